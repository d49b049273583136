"""Synthetic inputs of the benchmark configurations (fixed conventions).

Every entry is `field.reduce(next_u64())`, i.e. next_u64() mod p, as in the
reference's fill loops (main.rs:94-105, 153-155); matrices are row-major.

  c1   500 x 500 over 2^61-1 from MIXMAX-17 seed 1 (README: corank 432) and a
       control of the same shape from SplitMix64 seed 1 (corank 0).
  c2   64 matrices, sides n_i = round(256 * 16**u_i), u_i = (x >> 11) * 2^-53 from
       SplitMix64 seed 7.  Even i: xoshiro256** seed 7 (one stream, consumed in
       order); odd i: MIXMAX-17 seed 7 (one orbit, disjoint consecutive stretches).
  c3   n x n, rows 0..n-65 uniform from xoshiro256** seed 11, then each of the last
       64 rows = sum of the first 128 rows with coefficients drawn from the same
       stream (expected corank 64).
  c4   1024 sequences, lengths uniform in [1000, 20000] from SplitMix64 seed 3
       (1000 + x % 19001).  i % 3 == 0: xoshiro256** seed 3 (one stream);
       i % 3 == 1: order-L recurrence, L = 16 + x % 385 from the SplitMix64
       stream, nonzero coefficients and initial values from xoshiro256** seed
       1003; i % 3 == 2: MIXMAX-17 seed 3, disjoint stretches.
  c5   BM over MIXMAX-256 (skip 2) seed 1, n = 200000 (README: LC 65280), and
       256 matrices 2048 x 2048 over 65521, u32, from xoshiro256** seed 5.
"""
from __future__ import annotations

import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import modlin as M  # noqa: E402

M61 = (1 << 61) - 1
M31 = (1 << 31) - 1


def _mod(a: np.ndarray, p: int) -> np.ndarray:
    return (a % np.uint64(p)).astype(np.uint64)


def c1() -> tuple[np.ndarray, np.ndarray]:
    mm = _mod(M.prng_fill(M.PRNG_MIXMAX17, 1, 500 * 500), M61).reshape(500, 500)
    ctl = _mod(M.prng_fill(M.PRNG_SPLITMIX64, 1, 500 * 500), M61).reshape(500, 500)
    return mm, ctl


def c2_sides(count: int = 64) -> list[int]:
    x = M.prng_fill(M.PRNG_SPLITMIX64, 7, count)
    u = (x >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return [int(round(256.0 * 16.0 ** float(v))) for v in u]


def c2(count: int = 64, p: int = M61, max_side: int | None = None) -> list[np.ndarray]:
    sides = c2_sides(count)
    if max_side:
        sides = [min(s, max_side) for s in sides]
    even_total = sum(s * s for i, s in enumerate(sides) if i % 2 == 0)
    odd_total = sum(s * s for i, s in enumerate(sides) if i % 2 == 1)
    ev = M.prng_fill(M.PRNG_XOSHIRO256SS, 7, even_total)
    od = M.prng_fill(M.PRNG_MIXMAX17, 7, odd_total)
    mats, pe, po = [], 0, 0
    for i, s in enumerate(sides):
        if i % 2 == 0:
            mats.append(_mod(ev[pe:pe + s * s], p).reshape(s, s))
            pe += s * s
        else:
            mats.append(_mod(od[po:po + s * s], p).reshape(s, s))
            po += s * s
    return mats


def c3(n: int, p: int) -> np.ndarray:
    indep = n - 64
    draws = M.prng_fill(M.PRNG_XOSHIRO256SS, 11, indep * n + 64 * 128)
    m = np.empty((n, n), dtype=np.uint64)
    m[:indep] = _mod(draws[:indep * n], p).reshape(indep, n)
    coef = _mod(draws[indep * n:], p).reshape(64, 128)
    m[indep:] = M.combine_rows(p, m[:128], coef)
    return m


def c3_device(ctx: "M.Context", n: int, p: int) -> "M.DeviceBuffer":
    """The c3 matrix built in device memory (the same entries as c3(n, p)): the
    whole n*n prefix of the xoshiro256** seed-11 stream is generated on the
    device, then the last 64 rows -- whose stream positions hold the
    coefficients -- are replaced by the combinations of the first 128 rows."""
    indep = n - 64
    buf = ctx.prng_fill_device(M.PRNG_XOSHIRO256SS, 11, n * n, field=M.Field(p))
    coef = buf.read(indep * n, 64 * 128).reshape(64, 128)
    top = buf.read(0, 128 * n).reshape(128, n)
    buf.write(indep * n, M.combine_rows(p, top, coef))
    return buf


def mixmax256_matrix_device(ctx: "M.Context", n: int, p: int = M61, seed: int = 1) -> "M.DeviceBuffer":
    """n x n matrix of consecutive MIXMAX-256 (skip 2) outputs reduced mod p,
    row-major, built in device memory (main.rs:94-99 fill order).

    Rank bounds.  Every row is a length-n window of one sequence of linear
    complexity 65280 (README.md:107), and such windows span at most 65280
    dimensions, so at n = 65536 the corank is at least 256.  Each output block
    of the generator holds the 255 entries v[1..256) of its 256-word state, so
    when 255 divides n every row starts a block and rank <= 256 (at 510, 765,
    1020 the oracle gives exactly 256)."""
    return ctx.prng_fill_device(M.PRNG_MIXMAX256, seed, n * n, field=M.Field(p))


def c4_specs(count: int = 1024) -> list[tuple[int, int, int]]:
    """(kind, n, L) per sequence; kind 0 uniform, 1 recurrence of order L, 2 MIXMAX-17."""
    x = M.prng_fill(M.PRNG_SPLITMIX64, 3, 2 * count)
    specs = []
    for i in range(count):
        n = 1000 + int(x[2 * i] % np.uint64(19001))
        L = 16 + int(x[2 * i + 1] % np.uint64(385))
        specs.append((i % 3, n, L if i % 3 == 1 else 0))
    return specs


def c4(count: int = 1024, p: int = M61) -> tuple[list[np.ndarray], list[tuple[int, int, int]]]:
    specs = c4_specs(count)
    tot_u = sum(n for k, n, _ in specs if k == 0)
    tot_m = sum(n for k, n, _ in specs if k == 2)
    tot_l = sum(2 * L for k, _, L in specs if k == 1)
    uni = M.prng_fill(M.PRNG_XOSHIRO256SS, 3, tot_u)
    mmx = M.prng_fill(M.PRNG_MIXMAX17, 3, tot_m)
    lf = M.prng_fill(M.PRNG_XOSHIRO256SS, 1003, tot_l)
    seqs, pu, pm, pl = [], 0, 0, 0
    for kind, n, L in specs:
        if kind == 0:
            seqs.append(_mod(uni[pu:pu + n], p)); pu += n
        elif kind == 2:
            seqs.append(_mod(mmx[pm:pm + n], p)); pm += n
        else:
            coef = _mod(lf[pl:pl + L], p - 1) + np.uint64(1)  # nonzero coefficients
            init = _mod(lf[pl + L:pl + 2 * L], p)
            pl += 2 * L
            seqs.append(M.lfsr(p, coef, init, n))
    return seqs, specs


def c5_bm(n: int = 200_000) -> np.ndarray:
    return _mod(M.prng_fill(M.PRNG_MIXMAX256, 1, n), M61)


def c5_rank(count: int = 256, n: int = 2048, p: int = 65521) -> list[np.ndarray]:
    d = M.prng_fill(M.PRNG_XOSHIRO256SS, 5, count * n * n)
    return [(d[i * n * n:(i + 1) * n * n] % np.uint64(p)).astype(np.uint32).reshape(n, n) for i in range(count)]


C5_PLANTED_DEFICIENCY = (1, 2, 5, 17, 64, 200, 511, 1000)


def c5_planted(n: int = 2048, p: int = 65521, seed: int = 55) -> list[np.ndarray]:
    """Rank-deficient members for the small-prime u32 batch: matrix k has its last
    d_k rows (d_k from C5_PLANTED_DEFICIENCY) replaced by random F_p combinations
    of its first 32 rows, so its corank is d_k when its first n - d_k rows are
    independent (w.h.p.; the oracle confirms it for these seeds); entries from
    xoshiro256** seed + k, reduced mod p, stored as u32."""
    mats = []
    for k, d in enumerate(C5_PLANTED_DEFICIENCY):
        src = 32
        draws = M.prng_fill(M.PRNG_XOSHIRO256SS, seed + k, n * n + d * src)
        m = _mod(draws[:n * n], p).reshape(n, n)
        coef = _mod(draws[n * n:], p).reshape(d, src)
        m[n - d:] = M.combine_rows(p, m[:src], coef)
        mats.append(m.astype(np.uint32))
    return mats


def c5_rank_device(ctx: "M.Context", count: int = 256, n: int = 2048, p: int = 65521) -> "tuple[M.DeviceBuffer, list]":
    """The c5 rank batch generated in device memory as u32: one buffer holding the
    count matrices back to back (the same entries as c5_rank), and a view of each."""
    buf = ctx.prng_fill_device(M.PRNG_XOSHIRO256SS, 5, count * n * n, field=M.Field(p), elem=M.ELEM_U32,
                               square=False)
    return buf, [buf.view(i * n * n, n) for i in range(count)]
