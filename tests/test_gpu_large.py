"""GPU rank at the full sizes of the large configurations, where the oracle
cannot run in seconds, checked through size-independent properties.

  * Planted corank (workloads.c3): rows 0..n-65 uniform, the last 64 rows random
    F_p combinations of the first 128, so the corank is exactly 64 unless the
    planted rows or the uniform block lose rank by chance (probability O(n/p)).
    Both update methods at n = 8192 and 16384, over 2^61-1 (8 int8 limbs) and
    2^31-1 (4 limbs); n = 32768 (8 GiB of u64 entries) over 2^61-1, where the
    32-bit tile, row and column counters reach 2^15 tiles per side.
  * The linear-complexity bound of MIXMAX-256 (skip 2): every row of a matrix of
    consecutive outputs is a length-n window of one sequence of linear
    complexity 65280 (README.md:107), and such windows span at most 65280
    dimensions, so the 65536 x 65536 matrix (32 GiB) has corank >= 256.  Its
    rank must also not depend on the panel width (the rank is unique).  The
    same generator at 510, 850 and 1326 (one, three and five row phases mod the
    generator's 255-output blocks) is checked against the oracle.
  * The device-built matrices equal the host-built ones (workloads.c3 against
    workloads.c3_device at a small side).
"""
import numpy as np
import pytest

from conftest import modlin, oracle, workloads

pytestmark = pytest.mark.gpu
M61 = (1 << 61) - 1
M31 = (1 << 31) - 1
METHODS = [modlin.METHOD_PLAIN, modlin.METHOD_MFMA]


def test_device_c3_equals_host(ctx):
    for n, p in [(300, M61), (200, M31)]:
        buf = workloads.c3_device(ctx, n, p)
        try:
            assert np.array_equal(buf.to_host().reshape(n, n), workloads.c3(n, p))
        finally:
            buf.free()


@pytest.mark.parametrize("p", [M61, M31])
@pytest.mark.parametrize("n", [8192, 16384])
def test_planted_corank_full_size(ctx, n, p):
    buf = workloads.c3_device(ctx, n, p)
    try:
        f = modlin.Field(p)
        for method in METHODS:
            print(f"planted n={n} p={p} method={method}", flush=True)
            assert ctx.rank_batch(f, [buf], method=method) == [n - 64], (n, p, method)
    finally:
        buf.free()


def test_planted_corank_32768(ctx):
    n = 32768
    buf = workloads.c3_device(ctx, n, M61)
    try:
        print(f"planted n={n} p=2^61-1 method=mfma", flush=True)
        assert ctx.rank_batch(modlin.Field(M61), [buf], method=modlin.METHOD_MFMA) == [n - 64]
    finally:
        buf.free()


@pytest.mark.parametrize("n", [510, 850, 1326])
def test_mixmax256_against_oracle(ctx, n):
    buf = workloads.mixmax256_matrix_device(ctx, n)
    try:
        a = buf.to_host().reshape(n, n)
        want = oracle.rank(M61, a, n)
        for method in METHODS:
            assert ctx.rank_batch(modlin.Field(M61), [buf], method=method) == [want], (n, method)
    finally:
        buf.free()


def test_mixmax256_65536(ctx):
    n = 65536
    buf = workloads.mixmax256_matrix_device(ctx, n)
    try:
        f = modlin.Field(M61)
        print(f"MIXMAX-256 n={n} method=mfma panel=64", flush=True)
        r64 = ctx.rank_batch(f, [buf], method=modlin.METHOD_MFMA)[0]
        print(f"MIXMAX-256 n={n} rank={r64}; panel=48", flush=True)
        r48 = ctx.rank_batch(f, [buf], method=modlin.METHOD_MFMA, panel=48)[0]
        assert r64 == r48
        assert n - r64 >= 256
    finally:
        buf.free()
