"""The allocation-free entry point modlin_linear_complexity_noalloc against the
existing entry point (modlin_lc_batch on the same device sequence) and the
oracle, bit-exact: two calls in a row on each kernel path, the second with a
smaller input than the first, so the shared scratch array holds the previous
call's state when the next one starts.

Paths: p = 2^61-1 long (the blocked cooperative kernel) then short (one
workgroup, blocks on the matrix cores); p = 2^63-25 long (the one-exchange-per-
step cooperative kernel) then short (one step per pass); p = 65521 long then
short.  Plus the size limit and the empty sequence, rejected before any launch.
"""
import numpy as np
import pytest

from conftest import modlin, oracle, workloads

pytestmark = pytest.mark.gpu
M61 = (1 << 61) - 1


def _lfsr(p, L, n, seed):
    rng = np.random.default_rng(seed)
    return modlin.lfsr(p, rng.integers(1, p, size=L, dtype=np.uint64), rng.integers(0, p, size=L, dtype=np.uint64), n)


def _uniform(p, n, seed):
    return np.random.default_rng(seed).integers(0, p, size=n, dtype=np.uint64)


CASES = [
    (M61, lambda: workloads.c5_bm(200_000), 65280, lambda: _uniform(M61, 5000, 1)),
    ((1 << 63) - 25, lambda: _lfsr((1 << 63) - 25, 500, 70_000, 2), 500, lambda: _lfsr((1 << 63) - 25, 40, 3000, 3)),
    (65521, lambda: _lfsr(65521, 300, 100_000, 4), 300, lambda: _uniform(65521, 1000, 5)),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_noalloc_equals_lc_batch(ctx, case):
    p, make_long, want_long, make_short = CASES[case]
    f = modlin.Field(p)
    long_s, short_s = make_long(), make_short()
    bufs = [ctx.to_device(long_s, square=False), ctx.to_device(short_s, square=False)]
    try:
        got = []
        for b in bufs:  # the long sequence first, then the shorter one
            print(f"noalloc p={p} n={b.nbytes // 8}", flush=True)
            got.append(ctx.linear_complexity_noalloc(f, b))
        want = [ctx.lc_batch(f, [b])[0] for b in bufs]
        assert got == want
        assert got[0] == want_long
        assert got[1] == oracle.lc(p, short_s)
        # and once more in the opposite order of sizes' predecessor: short after short
        assert ctx.linear_complexity_noalloc(f, bufs[1]) == want[1]
    finally:
        for b in bufs:
            b.free()


def test_noalloc_rejects_bad_sizes(ctx):
    f = modlin.Field(M61)
    big = ctx.prng_fill_device(modlin.PRNG_SPLITMIX64, 1, int(modlin.LC_NOALLOC_MAX_N) + 1, field=f, square=False)
    try:
        with pytest.raises(ValueError):
            ctx.linear_complexity_noalloc(f, big)
        empty = modlin.DeviceBuffer(ctx, big.ptr, 0, modlin.ELEM_U64, 0)
        empty.owner = False
        with pytest.raises(ValueError):
            ctx.linear_complexity_noalloc(f, empty)
        # the largest accepted length still runs
        at_max = big.view(0, 1)
        at_max.nbytes = int(modlin.LC_NOALLOC_MAX_N) * 8
        s = big.read(0, int(modlin.LC_NOALLOC_MAX_N))
        assert ctx.linear_complexity_noalloc(f, at_max) == ctx.lc_batch(f, [s])[0]
    finally:
        big.free()
