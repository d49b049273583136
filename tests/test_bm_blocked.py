"""The blocked Berlekamp-Massey algebra of k_bm_block / k_bm_batch_blk, on the CPU.

The GPU kernels take K steps per pass: with C, B the connection polynomials and
m, l, bl, bb the state at the start of a block of steps i0 .. i0+K-1, they form
only the 2K sums
    Ec(r) = [x^(i0+r)] (C S),   Eb(r) = [x^(i0+r)] (x^m B S),   r < K,
(S = sum_j s_j x^j; coefficients of C above l and of B at or above bl read as
zero, s at negative indices as zero), replay the K steps of fp.rs:271-305 in
inversionless form on the vectors e(r), eb(r) and on the polynomials U, V (and
Ub, Vb at the last length change) with c_t = U C + V x^m B, and then form the new
C (and B) coefficient by coefficient.  This restatement follows the kernels'
bookkeeping -- coefficients owned in chunks of W, chunk q G + g to "workgroup" g,
sums taken per owner and added up -- and is compared with the oracle's
linear complexity (the reference's per-step algorithm, fp.rs:257-308) over random
sequences, recurrences, sparse, constant and zero sequences, several moduli,
block lengths K and owner counts G.  It pins the algebra the kernels implement,
not the kernels (the GPU tests compare those with the oracle directly).
"""
import random

import numpy as np
import pytest

from conftest import oracle


def _blocked_lc(p, s, K, G, W, drop_eb=False):
    n = len(s)
    C = [[0] * n, [0] * n]
    B = [[0] * n, [0] * n]
    C[0][0] = 1 % p
    B[0][0] = 1 % p
    cc = bc = 0
    l, bl, m, bb = 0, 1, 1, 1 % p
    nchunks = (n + W - 1) // W
    owned = [[q * G + g for q in range((nchunks - g + G - 1) // G)] if nchunks > g else [] for g in range(G)]
    i0 = 0
    while i0 < n:
        Kp = min(K, n - i0)
        # per-owner parts of Ec, Eb, then their sum (the exchange)
        Ec, Eb = [0] * K, [0] * K
        for g in range(G):
            for r in range(K):
                pc = pb = 0
                for ch in owned[g]:
                    for j in range(ch * W, min(n, ch * W + W)):
                        if j <= l and i0 + r - j >= 0 and i0 + r - j < n:
                            pc += C[cc][j] * s[i0 + r - j]
                        x = i0 + r - m - j
                        if j < bl and 0 <= x < n:
                            pb += B[bc][j] * s[x]
                Ec[r] = (Ec[r] + pc) % p
                Eb[r] = 0 if drop_eb else (Eb[r] + pb) % p
        # the replay (lane r of e, eb; lane a of U, V, U', V')
        e, eb = Ec[:], Eb[:]
        U, V, Up, Vp, Ub, Vb = [0] * K, [0] * K, [0] * K, [0] * K, [0] * K, [0] * K
        U[0], Vp[0] = 1 % p, 1 % p
        shift = lambda v: [0] + v[:-1]  # noqa: E731  (the DPP whole-wave shift)
        l0, bl0, m0, upd, bch = l, bl, m, False, False
        for t in range(Kp):
            i = i0 + t
            d = e[t]
            if d == 0:  # fp.rs:279-282
                m += 1
                eb, Up, Vp = shift(eb), shift(Up), shift(Vp)
                continue
            upd = True
            ne = [(bb * e[r] - d * eb[r]) % p for r in range(K)]
            nU = [(bb * U[a] - d * Up[a]) % p for a in range(K)]
            nV = [(bb * V[a] - d * Vp[a]) % p for a in range(K)]
            if 2 * l <= i:  # length change, fp.rs:287-299
                eb, Up, Vp, Ub, Vb = shift(e), shift(U), shift(V), U[:], V[:]
                bch = True
                old_l = l
                bl, bb, l, m = old_l + 1, d, i + 1 - old_l, 1
            else:
                eb, Up, Vp = shift(eb), shift(Up), shift(Vp)
                m += 1
            e, U, V = ne, nU, nV
        if upd:
            # new C (and B) per owned coefficient, from windows of the old ones
            getC = lambda j: C[cc][j] if 0 <= j <= l0 else 0  # noqa: E731
            getB = lambda j: B[bc][j] if 0 <= j < bl0 else 0  # noqa: E731
            for g in range(G):
                for ch in owned[g]:
                    for j in range(ch * W, min(n, ch * W + W)):
                        C[cc ^ 1][j] = sum(U[a] * getC(j - a) + V[a] * getB(j - m0 - a) for a in range(K)) % p
                        if bch:
                            B[bc ^ 1][j] = sum(Ub[a] * getC(j - a) + Vb[a] * getB(j - m0 - a) for a in range(K)) % p
            cc ^= 1
            if bch:
                bc ^= 1
        i0 += Kp
    return l


def _cases(seed, count):
    rng = random.Random(seed)
    for _ in range(count):
        p = rng.choice([2, 3, 5, 101, 65521, (1 << 31) - 1, (1 << 61) - 1])
        n = rng.randint(1, 80)
        kind = rng.randint(0, 4)
        if kind == 0:
            s = [rng.randrange(p) for _ in range(n)]
        elif kind == 1:
            L = rng.randint(1, 12)
            co = [rng.randrange(1, p) if p > 1 else 1 for _ in range(L)]
            s = [rng.randrange(p) for _ in range(L)]
            while len(s) < n:
                s.append(sum(co[k] * s[-1 - k] for k in range(L)) % p)
            s = s[:n]
        elif kind == 2:
            s = [0] * n
            s[rng.randrange(n)] = rng.randrange(1, p)
        elif kind == 3:
            s = [rng.choice([0, 0, 1]) % p for _ in range(n)]
        else:
            s = [1 % p] * n
        yield p, s, rng.choice([1, 2, 3, 8, 16, 64]), rng.choice([1, 2, 3, 5]), rng.choice([1, 4, 16])


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_blocked_algebra_matches_oracle(seed):
    for p, s, K, G, W in _cases(seed, 80):
        want = oracle.lc(p, np.array(s, dtype=np.uint64))
        assert _blocked_lc(p, s, K, G, W) == want, (p, len(s), K, G, W)


def test_planted_mistake_is_caught():
    """With the Eb sums (the x^m B part of the discrepancies) dropped, the same
    comparison fails on recurrences, so it can fail at all."""
    p = 65521
    rng = random.Random(9)
    differ = 0
    for L in (3, 6, 9):
        co = [rng.randrange(1, p) for _ in range(L)]
        s = [rng.randrange(p) for _ in range(L)]
        while len(s) < 60:
            s.append(sum(co[k] * s[-1 - k] for k in range(L)) % p)
        want = oracle.lc(p, np.array(s, dtype=np.uint64))
        assert _blocked_lc(p, s, 8, 2, 4) == want
        differ += _blocked_lc(p, s, 8, 2, 4, drop_eb=True) != want
    assert differ > 0
