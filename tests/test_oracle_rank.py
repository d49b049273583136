"""Pins the oracle's blocked rank (oracle.c or_rank, a restatement of fp.rs:116-255).

* known answers of tests/test_rank.rs:74-97 and the README (corank 432 / 0);
* equality with the naive elimination of the reference's own test
  (test_rank.rs:44-72) over its whole family (test_rank.rs:99-149);
* an independent method: GF(2) bit-packed elimination for p = 2;
* brute force: for n <= 3 over small p, rank from all minors' determinants;
* invariants: rank(A) = rank(A^T) = rank(P A Q); planted corank.
"""
import itertools

import numpy as np
import pytest

import families
from conftest import modlin, oracle, workloads

M61 = (1 << 61) - 1


def test_known_matrices():
    ident = np.eye(3, dtype=np.uint64)
    assert oracle.rank(M61, ident, 3) == 3
    assert oracle.rank(M61, np.array([1, 2, 3, 4, 5, 6, 5, 7, 9], dtype=np.uint64), 3) == 2
    assert oracle.rank(M61, np.zeros(16, dtype=np.uint64), 4) == 0
    assert oracle.rank(2, np.array([1, 1, 0, 0, 1, 1, 1, 0, 1], dtype=np.uint64), 3) == 2


def test_blocked_matches_reference_family():
    count = 0
    for p, n, target, a in families.rank_family():
        expected = oracle.rank_ref(p, a, n)
        got = oracle.rank(p, a, n)
        assert got == expected, (p, n, target)
        assert expected <= target
        count += 1
    assert count == 150


def test_gf2_independent_method():
    rng = np.random.default_rng(5)
    for n in [1, 7, 64, 65, 130, 257]:
        for density in [0.5, 0.1]:
            a = (rng.random((n, n)) < density).astype(np.uint64)
            assert oracle.rank(2, a, n) == oracle.rank_gf2(a, n), (n, density)


def _det(m, p):
    k = len(m)
    if k == 1:
        return m[0][0] % p
    return sum((-1) ** c * m[0][c] * _det([r[:c] + r[c + 1:] for r in m[1:]], p) for c in range(k)) % p


def _rank_by_minors(m, n, p):
    for k in range(n, 0, -1):
        for rows in itertools.combinations(range(n), k):
            for cols in itertools.combinations(range(n), k):
                if _det([[m[r][c] for c in cols] for r in rows], p):
                    return k
    return 0


@pytest.mark.parametrize("p", [2, 3, 5])
def test_brute_force_small(p):
    rng = np.random.default_rng(p)
    for n in [1, 2, 3]:
        for _ in range(60):
            a = rng.integers(0, p, size=(n, n)).astype(np.uint64)
            assert oracle.rank(p, a, n) == _rank_by_minors(a.tolist(), n, p)


def test_invariants_transpose_and_permutation():
    rng = np.random.default_rng(11)
    for p in [97, M61]:
        for n, r in [(70, 33), (130, 129), (65, 1)]:
            b = rng.integers(0, p, size=(r, n)).astype(np.uint64)
            c = rng.integers(0, p, size=(n, r)).astype(np.uint64)
            # A = C B has rank r w.h.p.; build it with the exact helper
            a = modlin.combine_rows(p, b, c)
            ra = oracle.rank(p, a, n)
            assert ra == r
            assert oracle.rank(p, np.ascontiguousarray(a.T), n) == ra
            perm_r, perm_c = rng.permutation(n), rng.permutation(n)
            assert oracle.rank(p, np.ascontiguousarray(a[perm_r][:, perm_c]), n) == ra


def test_readme_c1():
    mm, ctl = workloads.c1()
    assert 500 - oracle.rank(M61, mm, 500) == 432   # README.md:71
    assert 500 - oracle.rank(M61, ctl, 500) == 0


def test_readme_xoroshiro_1000():
    s = oracle.xoroshiro128pp(0, 1000 * 1000)
    for p in [2, M61]:  # README.md:129, 149
        a = (s % np.uint64(p)).astype(np.uint64)
        assert oracle.rank(p, a, 1000) == 1000


def test_planted_corank():
    for p in [M61, (1 << 31) - 1]:
        a = workloads.c3(320, p)
        assert oracle.rank(p, a, 320) == 256


def test_work_count_full_rank():
    # W = sum_k (n-1-k)(n-1-j_k) = (n-1)n(2n-1)/6 when every column pivots.
    rng = np.random.default_rng(3)
    n = 90
    a = rng.integers(0, M61, size=(n, n)).astype(np.uint64)
    r, pcols, w = oracle.rank(M61, a, n, with_trace=True)
    assert r == n and pcols == list(range(n))
    assert w == (n - 1) * n * (2 * n - 1) // 6
