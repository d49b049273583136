"""GPU parity of the device generators (modlin_prng_fill_device) against the host
generators of the same library, which test_prng.py pins to the reference's
known-answer vectors (test_prng.rs:23-62, 87-98) and to the oracle.

Bar: every output word identical, for every generator kind, at offsets near the
start and far into the stream (reached by jump-ahead), raw and reduced mod p
(Mersenne, generic Barrett, 63-bit and u32 storage), with the keep/drop
decimation of the mixmax-lux feature, and with MIXMAX lanes forced to start from
earlier jump states.  The 500 x 500 MIXMAX-17 matrix of the reference's plain
rank case, filled on the device, must keep its corank of 432.
"""
import os

import numpy as np
import pytest

from conftest import modlin

pytestmark = pytest.mark.gpu
M61 = (1 << 61) - 1
ALL = [modlin.PRNG_XOROSHIRO128PP, modlin.PRNG_MIXMAX17, modlin.PRNG_MIXMAX240, modlin.PRNG_MIXMAX256,
       modlin.PRNG_SPLITMIX64, modlin.PRNG_XOSHIRO256SS]
MIXMAX = [modlin.PRNG_MIXMAX17, modlin.PRNG_MIXMAX240, modlin.PRNG_MIXMAX256]


def _dev(ctx, kind, seed, count, **kw):
    buf = ctx.prng_fill_device(kind, seed, count, square=False, **kw)
    try:
        return buf.to_host()
    finally:
        buf.free()


def _lux(raw, keep=5, drop=9):
    idx = np.arange(raw.size)
    return raw[(idx % (keep + drop)) < keep]


@pytest.mark.parametrize("kind", ALL)
def test_prefix_raw(ctx, kind):
    for seed in [1, 7, 1234567890123456789]:
        for count in [1, 15, 16, 17, 255, 1000, 70001]:
            got = _dev(ctx, kind, seed, count)
            want = modlin.prng_fill(kind, seed, count)
            assert np.array_equal(got, want), (kind, seed, count)


@pytest.mark.parametrize("kind", ALL)
def test_far_offsets(ctx, kind):
    for skip in [1, 16, 254, 255, 256, 1000003, 20000000]:
        got = _dev(ctx, kind, 11, 3000, skip=skip)
        want = modlin.prng_fill(kind, 11, 3000, skip=skip)
        assert np.array_equal(got, want), (kind, skip)


@pytest.mark.parametrize("kind", ALL)
def test_large_fill_windows(ctx, kind):
    # many lanes: 4M outputs, compared in full against the host stream
    count = 1 << 22
    got = _dev(ctx, kind, 5, count, skip=777)
    want = modlin.prng_fill(kind, 5, count, skip=777)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("p", [M61, (1 << 31) - 1, 65521, 1000000007, (1 << 63) - 25, 2])
def test_reduced(ctx, p):
    f = modlin.Field(p)
    for kind in ALL:
        want = modlin.prng_fill(kind, 3, 20000, skip=5000) % np.uint64(p)
        got = _dev(ctx, kind, 3, 20000, skip=5000, field=f)
        assert np.array_equal(got, want), (kind, p)
        if p < (1 << 32):
            got32 = _dev(ctx, kind, 3, 20000, skip=5000, field=f, elem=modlin.ELEM_U32)
            assert got32.dtype == np.uint32 and np.array_equal(got32.astype(np.uint64), want), (kind, p)


@pytest.mark.parametrize("kind", ALL)
def test_decimation(ctx, kind):
    for skip, count in [(0, 5000), (3, 4097), (123457, 9000)]:
        raw = modlin.prng_fill(kind, 9, (skip + count) * 3 + 64)
        want = _lux(raw)[skip:skip + count]
        got = _dev(ctx, kind, 9, count, skip=skip, keep=5, drop=9)
        assert np.array_equal(got, want), (kind, skip, count)


@pytest.mark.parametrize("kind", MIXMAX)
def test_mixmax_lanes_walk_back(ctx, kind):
    # A start bound above every word marks each jumped state unusable, so every
    # lane walks back to the seeded state and runs forward from there.
    old = os.environ.get("MODLIN_PRNG_CLEAN_BOUND")
    os.environ["MODLIN_PRNG_CLEAN_BOUND"] = str(1 << 62)
    try:
        for skip in [0, 500]:
            got = _dev(ctx, kind, 13, 30000, skip=skip)
            want = modlin.prng_fill(kind, 13, 30000, skip=skip)
            assert np.array_equal(got, want), (kind, skip)
    finally:
        if old is None:
            del os.environ["MODLIN_PRNG_CLEAN_BOUND"]
        else:
            os.environ["MODLIN_PRNG_CLEAN_BOUND"] = old


def test_errors(ctx):
    with pytest.raises(ValueError):
        ctx.prng_fill_device(modlin.PRNG_MIXMAX17, 0, 10)  # prng.rs:153-155
    with pytest.raises(ValueError):
        ctx.prng_fill_device(modlin.PRNG_SPLITMIX64, 1, 10, elem=modlin.ELEM_U32)  # u32 needs p < 2^32
    with pytest.raises(ValueError):
        ctx.prng_fill_device(modlin.PRNG_SPLITMIX64, 1, 10, field=modlin.Field(M61), elem=modlin.ELEM_U32)


def test_plain_rank_case_on_device(ctx):
    # main.rs:94-105 with MIXMAX-17, seed 1, p = 2^61 - 1: 500 x 500, corank 432.
    f = modlin.Field(M61)
    buf = ctx.prng_fill_device(modlin.PRNG_MIXMAX17, 1, 500 * 500, field=f)
    try:
        assert buf.n == 500
        assert ctx.rank_batch(f, [buf]) == [500 - 432]
    finally:
        buf.free()
