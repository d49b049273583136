"""CPU restatement of the rank engine's block elimination (k_block_ge) and block
inverse (k_block_z), checked against exact integer arithmetic.

k_block_ge eliminates the top block T (Bt rows of the next logical rows x the w
panel columns) without moving rows: the pivot of column j is the first row, in
block order, that is not yet a pivot and has a nonzero in column j (any pivot
choice gives the rank, which is the reference's contract, fp.rs:142-178), and
the other non-pivot rows with a nonzero multiplier e take the fraction-free
step x <- pi x - e y (no inverse on the step).  It records per step the pivot
row q, the pivot pi, the pivot row's own-row coefficient and the multipliers e
(zero for the pivot rows), plus the pivot rows' entries in later pivot columns.

k_block_z replays the record on the identity to get F (F A = U, A = T[R, J] in
pivot order) and solves U Z = F column by column: Z = A^-1.  The tests check
A Z = I on random, sparse, rank-deficient and duplicate-row blocks over small
primes and 2^61 - 1, that the pivot count equals the rank of T, and that a
planted mistake in the replay (the pivot-row term dropped) is caught.
"""
import random

import pytest

M61 = (1 << 61) - 1


def block_ge(T, p):
    """The record of the no-swap fraction-free elimination of T (list of rows)."""
    Bt, w = len(T), len(T[0]) if T else 0
    U = [list(r) for r in T]
    own = [1 % p] * Bt
    piv = [False] * Bt
    rec = {"q": [], "pi": [], "own": [], "j": [], "E": []}
    for j in range(w):
        if len(rec["q"]) == Bt:
            break
        q = next((r for r in range(Bt) if not piv[r] and U[r][j] % p != 0), None)
        if q is None:
            continue
        pi = U[q][j]
        E = [0] * Bt
        for r in range(Bt):
            e = U[r][j]
            if piv[r] or r == q or e == 0:
                continue
            E[r] = e
            for c in range(j + 1, w):
                U[r][c] = (pi * U[r][c] - e * U[q][c]) % p
            own[r] = pi * own[r] % p
        rec["q"].append(q)
        rec["pi"].append(pi)
        rec["own"].append(own[q])
        rec["j"].append(j)
        rec["E"].append(E)
        piv[q] = True
    pb = len(rec["q"])
    # pivot rows' entries in later pivot columns (slot coordinates)
    rec["U"] = [[U[rec["q"][r]][rec["j"][s]] if s > r else 0 for s in range(pb)] for r in range(pb)]
    return rec


def block_z(rec, p, Bt, drop_pivot_term=False):
    """Z = A^-1 from the record: forward replay of F per slot, then U Z = F."""
    pb = len(rec["q"])
    Q, PI = rec["q"], rec["pi"]
    Z = [[0] * pb for _ in range(pb)]
    for s in range(pb):
        ev = [0] * Bt
        ev[Q[s]] = rec["own"][s]
        for k in range(s, pb):
            y = 0 if drop_pivot_term else ev[Q[k]]
            E = rec["E"][k]
            for r in range(Bt):
                if E[r] != 0:
                    ev[r] = (PI[k] * ev[r] - E[r] * y) % p
        f = [ev[Q[t]] if t >= s else 0 for t in range(pb)]
        z = [0] * pb
        for t in range(pb - 1, -1, -1):
            acc = sum(rec["U"][t][u] * z[u] for u in range(t + 1, pb)) % p
            z[t] = (f[t] - acc) * pow(PI[t], p - 2, p) % p
        for t in range(pb):
            Z[t][s] = z[t]
    return Z


def rank_mod(T, p):
    A = [list(r) for r in T]
    rk, rows = 0, len(A)
    cols = len(A[0]) if A else 0
    for c in range(cols):
        piv = next((r for r in range(rk, rows) if A[r][c] % p), None)
        if piv is None:
            continue
        A[rk], A[piv] = A[piv], A[rk]
        inv = pow(A[rk][c], p - 2, p)
        for r in range(rows):
            if r != rk and A[r][c] % p:
                f = A[r][c] * inv % p
                A[r] = [(a - f * b) % p for a, b in zip(A[r], A[rk])]
        rk += 1
    return rk


def check(T, p, drop=False):
    Bt = len(T)
    rec = block_ge(T, p)
    pb = len(rec["q"])
    assert pb == rank_mod(T, p)
    Z = block_z(rec, p, Bt, drop)
    A = [[T[rec["q"][r]][rec["j"][c]] for c in range(pb)] for r in range(pb)]
    for r in range(pb):
        for c in range(pb):
            v = sum(A[r][k] * Z[k][c] for k in range(pb)) % p
            if v != (1 if r == c else 0):
                return False
    return True


def blocks(rng, p, Bt, w):
    uni = [[rng.randrange(p) for _ in range(w)] for _ in range(Bt)]
    yield uni
    sparse = [[rng.randrange(p) if rng.random() < 0.15 else 0 for _ in range(w)] for _ in range(Bt)]
    yield sparse
    # rank-deficient: rows are combinations of a few
    base = [[rng.randrange(p) for _ in range(w)] for _ in range(3)]
    yield [[sum(rng.randrange(p) * b[c] for b in base) % p for c in range(w)] for _ in range(Bt)]
    # zero leading columns, duplicate rows, a zero row
    z = min(3, w)
    t = [[0] * z + [rng.randrange(p) for _ in range(w - z)] for _ in range(Bt)]
    if Bt > 2:
        t[1] = list(t[0])
        t[-1] = [0] * w
    yield t
    # entries in {0, 1, p-1}
    yield [[rng.choice([0, 1, p - 1]) for _ in range(w)] for _ in range(Bt)]


@pytest.mark.parametrize("p", [2, 3, 65521, (1 << 31) - 1, M61])
def test_block_inverse(p):
    rng = random.Random(p)
    for Bt, w in [(1, 1), (5, 3), (3, 5), (12, 12), (16, 9)]:
        for T in blocks(rng, p, Bt, w):
            assert check(T, p), (p, Bt, w, T)


def test_planted_mistake_is_caught():
    rng = random.Random(5)
    T = [[rng.randrange(M61) for _ in range(8)] for _ in range(8)]
    assert check(T, M61)
    assert not check(T, M61, drop=True)
