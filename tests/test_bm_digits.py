"""The signed-digit matrix-core arithmetic of the batched BM for p = 2^61 - 1, on the CPU.

k_bm_batch_blk (modlin-rs_amd/csrc/bm_kernels.hip) forms both products of a block of
64 steps as int8 MFMAs:
  * digits: the balanced representative of x mod p as eight signed base-256 digits, read
    off as (bal(x) + c) xor c with c = 0x80 in every byte, and moved into digit planes
    by byte transposes of __byte_perm selectors (store_planes8, store_planes4);
  * E sums: Ec(r) = sum_j C[j] s[i0+r-j] and Eb(r) = sum_j D[j] s[i0+r-j] with
    D = x^m B, as sum_{u,v} 256^(u+v) T_v[u][r]; rounds of 512 coefficients (rounds
    holding no coefficient of C or D skipped), the int8 products accumulated over at
    most 8 rounds, then the class sums (c = u+v within a wave's four digits of C or D
    and all eight digits of s) offset by 2^28, rotated by 8c bits mod p, summed, minus
    the offsets;
  * new C: sum_{a<64} U[a] C[j-a] + V[a] B[j-m0-a] as sum_s 256^s Q_s[j] over 15
    classes, combined three at a time in int64 (|T| < 2^41), offset 2^41, five
    rotations in one 64-bit sum.
This file restates that arithmetic with numpy int64 (the values the int32 accumulators
hold) and Python integers and compares it with direct modular sums, over random residues
and residues with many extreme digits; it checks the bounds the kernels rely on
(|T| <= 2^26 at a fold, class sums <= 2^28 and <= 2^24, triples below 2^41, no 64-bit
overflow in the partial sums).  A planted mistake (one rotation off by 8 bits) must fail.  The GPU
tests compare the kernel itself with the oracle.
"""
import random

import numpy as np

P = (1 << 61) - 1
C80 = 0x8080808080808080
M64 = (1 << 64) - 1


def bal(x):
    """rank_common.hpp balanced(): x in [0, p) -> (-p/2, p/2]."""
    return x - P if x > (P - 1) // 2 else x


def digits_addxor(x):
    w = ((bal(x) + C80) & M64) ^ C80
    return [((w >> (8 * u)) & 0xFF) - (256 if (w >> (8 * u)) & 0x80 else 0) for u in range(8)]


def digits_peel(x):
    v, out = bal(x), []
    for _ in range(8):
        d = v & 0xFF
        d = d - 256 if d & 0x80 else d
        out.append(d)
        v = (v - d) >> 8
    assert v == 0
    return out


def rot(x, s):
    """rot_m61 in bm_kernels.hip: x * 2^s mod p, not reduced, for x < 2^42 (result < 2^61 + 2^42)."""
    assert 0 <= x < (1 << 42)
    r = x if s == 0 else ((x << s) & P) + (x >> (61 - s))
    assert r < (1 << 61) + (1 << 42) and r % P == (x << s) % P
    return r


def canon(x):
    x = (x & P) + (x >> 61)
    return x - P if x >= P else x


def byte_perm(x, y, s):
    """HIP __byte_perm: byte i of the result is byte (s >> 4i) & 7 of y:x."""
    b = x | (y << 32)
    return sum(((b >> (8 * ((s >> (4 * i)) & 7))) & 0xFF) << (8 * i) for i in range(4))


SELECTORS = (0x5140, 0x7362, 0x5410, 0x7632)  # store_planes8's __byte_perm selectors


def store_planes8(d, sel=SELECTORS):
    """store_planes8: eight digit words -> eight plane words (byte t of word u is byte u of d[t])."""
    lo_pair, hi_pair, lo_half, hi_half = sel
    q = [[[0] * 4 for _ in range(2)] for _ in range(2)]
    for h in range(2):
        for g in range(2):
            a, b, c, e = ((d[4 * g + k] >> (32 * h)) & 0xFFFFFFFF for k in range(4))
            t0, t1 = byte_perm(a, b, lo_pair), byte_perm(a, b, hi_pair)
            t2, t3 = byte_perm(c, e, lo_pair), byte_perm(c, e, hi_pair)
            q[h][g] = [byte_perm(t0, t2, lo_half), byte_perm(t0, t2, hi_half),
                       byte_perm(t1, t3, lo_half), byte_perm(t1, t3, hi_half)]
    return [q[u >> 2][0][u & 3] | (q[u >> 2][1][u & 3] << 32) for u in range(8)]


def store_planes4(d, sel=SELECTORS):
    """store_planes4: four digit words -> eight 4-byte plane words (byte t of word u is byte u of d[t])."""
    lo_pair, hi_pair, lo_half, hi_half = sel
    out = [0] * 8
    for h in range(2):
        a, b, c, e = ((d[k] >> (32 * h)) & 0xFFFFFFFF for k in range(4))
        t0, t1 = byte_perm(a, b, lo_pair), byte_perm(a, b, hi_pair)
        t2, t3 = byte_perm(c, e, lo_pair), byte_perm(c, e, hi_pair)
        out[4 * h:4 * h + 4] = [byte_perm(t0, t2, lo_half), byte_perm(t0, t2, hi_half),
                                byte_perm(t1, t3, lo_half), byte_perm(t1, t3, hi_half)]
    return out


def _is_transpose(words, planes):
    return all((planes[u] >> (8 * t)) & 0xFF == (words[t] >> (8 * u)) & 0xFF for u in range(8) for t in range(8))


def _residues(rng, count, extreme):
    out = []
    for _ in range(count):
        if extreme and rng.random() < 0.5:
            # balanced values made of digits -128 / 127 (the largest products)
            v = sum(rng.choice((-128, 127)) * 256 ** u for u in range(7)) + rng.randrange(-8, 8) * 256 ** 7
            out.append(v % P)
        else:
            out.append(rng.randrange(P))
    return out


def _planes(vals):
    return np.array([digits_addxor(x) for x in vals], dtype=np.int64).T  # [8][len]


def test_digits_and_transpose():
    rng = random.Random(1)
    vals = [0, 1, P - 1, (P - 1) // 2, (P + 1) // 2, 2 ** 60, 2 ** 60 - 1] + _residues(rng, 500, True)
    for x in vals:
        d = digits_addxor(x)
        assert d == digits_peel(x)
        assert all(-128 <= e <= 127 for e in d)
        assert sum(e * 256 ** u for u, e in enumerate(d)) == bal(x)
    for _ in range(200):
        words = [rng.getrandbits(64) for _ in range(8)]
        assert _is_transpose(words, store_planes8(words))
    for _ in range(200):
        words = [rng.getrandbits(64) for _ in range(4)]
        planes = store_planes4(words)
        assert all((planes[u] >> (8 * k)) & 0xFF == (words[k] >> (8 * u)) & 0xFF for u in range(8) for k in range(4))
    # a selector with two nibbles swapped breaks the transpose
    words = [rng.getrandbits(64) for _ in range(8)]
    for k, typo in enumerate((0x1540, 0x7326, 0x4510, 0x7623)):
        sel = list(SELECTORS)
        sel[k] = typo
        assert not _is_transpose(words, store_planes8(words, tuple(sel)))


def _esums_mfma(C, B, s, l, bl, m, i0, shift_mistake=False):
    """bm_esums_mfma_m61's arithmetic: returns (Ec, Eb), each 64 residues."""
    RE, SWE, n = 512, 592, len(s)
    top = max(l, m + bl - 1)
    # a wave holds the digits u in [4 uh, 4 uh + 4) of C (or D) against all eight digits
    # of s; its lane's classes are cr = ul + 4 h + i (u = 4 uh + ul, v = 4 h + i)
    koff = {uh: sum(canon(rot(1 << 28, (8 * (4 * uh + cr)) % 61)) for cr in range(11)) % P for uh in range(2)}
    ep = [[[0] * 64 for _ in range(2)] for _ in range(2)]  # [which][uh][r]
    acc = np.zeros((8, 16, 64), dtype=np.int64)  # [v][row][r]
    rounds = 0

    def fold():
        assert np.abs(acc).max() <= 1 << 26
        for which in range(2):
            for uh in range(2):
                for r in range(64):
                    w = 0
                    for cr in range(11):
                        cl = sum(int(acc[4 * h + i, 8 * which + 4 * uh + ul, r])
                                 for ul in range(4) for h in range(2) for i in range(4) if ul + 4 * h + i == cr)
                        assert abs(cl) <= 1 << 28
                        sh = (8 * (4 * uh + cr) + (8 if shift_mistake and cr == 3 else 0)) % 61
                        w = w + rot(cl + (1 << 28), sh)
                        assert w < 1 << 64
                        w = (w & P) + (w >> 61)
                    ep[which][uh][r] = (ep[which][uh][r] + canon(w) - koff[uh]) % P
        acc[...] = 0

    j0 = 0
    while j0 <= top:
        if j0 > l and j0 + RE <= m:
            j0 += RE
            continue
        cv = [C[j] if j <= l else 0 for j in range(j0, j0 + RE)]
        dv = [B[j - m] if m <= j < m + bl else 0 for j in range(j0, j0 + RE)]
        AP = np.concatenate([_planes(cv), _planes(dv)])  # [16][RE]
        sv = [s[x] if 0 <= x < n else 0 for x in (i0 - j0 + 63 - z for z in range(SWE))]
        SP = _planes(sv)  # [8][SWE]
        for v in range(8):
            # B[k][r] = SP[v][63 - r + k]
            tw = np.lib.stride_tricks.sliding_window_view(SP[v], RE)[:64][::-1]  # row r: SP[v][63 - r : 63 - r + RE]
            acc[v] += AP @ tw.T
        rounds += 1
        if rounds == 8:
            fold()
            rounds = 0
        j0 += RE
    fold()
    return [sum(ep[0][uh][r] for uh in range(2)) % P for r in range(64)], \
           [sum(ep[1][uh][r] for uh in range(2)) % P for r in range(64)]


def _esums_direct(C, B, s, l, bl, m, i0):
    n = len(s)
    ec = [sum(C[j] * s[i0 + r - j] for j in range(l + 1) if 0 <= i0 + r - j < n) % P for r in range(64)]
    eb = [sum(B[j] * s[i0 + r - m - j] for j in range(bl) if 0 <= i0 + r - m - j < n) % P for r in range(64)]
    return ec, eb


def _state(rng, n, l, bl, m, extreme):
    s = _residues(rng, n, extreme)
    C = _residues(rng, l + 1, extreme)
    B = _residues(rng, bl, extreme)
    return C, B, s


def test_esums_digit_classes():
    rng = random.Random(7)
    # (n, l, bl, m, i0): one round; rounds crossing 512; D far above C (skipped rounds);
    # more than 8 rounds (a fold in the middle); s windows running off both ends
    cases = [(300, 40, 30, 3, 128), (2000, 700, 650, 2, 1350), (3000, 20, 5, 2400, 2430),
             (9000, 4300, 4200, 5, 8600), (1200, 600, 1, 1, 1190)]
    for k, (n, l, bl, m, i0) in enumerate(cases):
        C, B, s = _state(rng, n, l, bl, m, extreme=k % 2 == 1)
        assert _esums_mfma(C, B, s, l, bl, m, i0) == _esums_direct(C, B, s, l, bl, m, i0), k


def test_esums_planted_mistake():
    rng = random.Random(8)
    C, B, s = _state(rng, 400, 60, 50, 2, extreme=False)
    assert _esums_mfma(C, B, s, 60, 50, 2, 200, shift_mistake=True) != _esums_direct(C, B, s, 60, 50, 2, 200)


def _apply_mfma(C, B, U, V, l0, bl0, m0, n, triple_mistake=False):
    """bm_apply_mfma_m61's arithmetic for new C: 15 classes combined three at a time,
    T_k = Q_3k + 2^8 Q_3k+1 + 2^16 Q_3k+2, T_k + 2^41 rotated by 24 k bits, summed."""
    koff = sum(canon(rot(1 << 41, (24 * k) % 61)) for k in range(5)) % P
    cw = [C[j] if 0 <= j <= l0 else 0 for j in range(n)]
    bw = [B[j] if 0 <= j < bl0 else 0 for j in range(n)]
    Cd, Bd = _planes(cw), _planes(bw)
    Ud, Vd = _planes(U), _planes(V)
    out = []
    for j in range(n):
        a = np.arange(64)
        jc, jb = j - a, j - m0 - a
        xc = np.where(jc >= 0, Cd[:, np.clip(jc, 0, n - 1)], 0)  # [8][64]
        xb = np.where(jb >= 0, Bd[:, np.clip(jb, 0, n - 1)], 0)
        prod = xc @ Ud.T + xb @ Vd.T  # [u][v] = sum_a
        Q = [int(sum(prod[u, q - u] for u in range(8) if 0 <= q - u < 8)) for q in range(15)]
        assert max(abs(x) for x in Q) <= 1 << 24
        w = 0
        for k in range(5):
            t = Q[3 * k] + (Q[3 * k + 1] << (16 if triple_mistake else 8)) + (Q[3 * k + 2] << 16)
            assert abs(t) < 1 << 41
            w += rot(t + (1 << 41), (24 * k) % 61)
        assert w < 1 << 64
        out.append((canon((w & P) + (w >> 61)) - koff) % P)
    return out


def test_application_digit_classes():
    rng = random.Random(9)
    for k, (n, l0, bl0, m0) in enumerate([(200, 150, 90, 3), (300, 299, 120, 40)]):
        C, B, _ = _state(rng, 1, l0, bl0, m0, extreme=k == 1)
        U = _residues(rng, 64, k == 1)
        V = _residues(rng, 64, k == 1)
        want = [sum(U[a] * (C[j - a] if 0 <= j - a <= l0 else 0) + V[a] * (B[j - m0 - a] if 0 <= j - m0 - a < bl0 else 0)
                    for a in range(64)) % P for j in range(n)]
        assert _apply_mfma(C, B, U, V, l0, bl0, m0, n) == want, k
        if k == 0:  # a wrong weight inside the triples is caught
            assert _apply_mfma(C, B, U, V, l0, bl0, m0, n, triple_mistake=True) != want
