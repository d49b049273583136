"""Pins the oracle's Berlekamp-Massey (oracle.c or_lc, a restatement of fp.rs:257-308).

* known answers of tests/test_linear_complexity.rs:72-91;
* equality with the naive BM of the reference's test (test_linear_complexity.rs:37-70)
  over its family (93-133), including the LFSR bound LC <= k;
* README known answers: 272 (MIXMAX-17, n=1000), 503 (xoroshiro128++, p=2),
  5000 (xoroshiro128++, n=10000);
* invariants: LC of a recurrence of order L with nonzero last coefficient and
  random start is L; a uniform sequence has LC = ceil(n/2) w.h.p.; LC is
  unchanged by scaling the sequence by a nonzero constant.
"""
import numpy as np
import pytest

import families
from conftest import modlin, oracle

M61 = (1 << 61) - 1


def test_known_answers():
    p = M61
    s = [1, 2, 3]
    for i in range(3, 40):
        s.append((3 * s[i - 1] + 5 * s[i - 2] + 7 * s[i - 3]) % p)
    assert oracle.lc(p, np.array(s, dtype=np.uint64)) == 3
    assert oracle.lc(p, np.full(20, 4, dtype=np.uint64)) == 1
    assert oracle.lc(p, np.zeros(20, dtype=np.uint64)) == 0


def test_matches_reference_family():
    count = 0
    for p, label, s, k in families.lc_family():
        got = oracle.lc(p, s)
        assert got == oracle.lc_ref(p, s), (p, label)
        if k is not None:
            assert got <= k, (p, label)
        count += 1
    assert count == 60


def test_empty_rejected():
    with pytest.raises(ValueError):
        oracle.lc(M61, np.zeros(0, dtype=np.uint64))


def test_readme_known_answers():
    s = (oracle.mixmax(17, 1, 1000) % np.uint64(M61)).astype(np.uint64)
    assert oracle.lc(M61, s) == 272                      # README.md:88
    s = (oracle.xoroshiro128pp(0, 1000) % np.uint64(2)).astype(np.uint64)
    assert oracle.lc(2, s) == 503                        # README.md:137
    s = (oracle.xoroshiro128pp(0, 10000) % np.uint64(M61)).astype(np.uint64)
    assert oracle.lc(M61, s) == 5000                     # README.md:158


def test_recurrence_and_uniform_invariants():
    rng = np.random.default_rng(9)
    for p in [65521, M61]:
        for L in [1, 7, 40]:
            coef = rng.integers(1, p, size=L).astype(np.uint64)
            init = rng.integers(0, p, size=L).astype(np.uint64)
            s = modlin.lfsr(p, coef, init, 6 * L + 10)
            assert oracle.lc(p, s) == L
        for n in [1, 2, 9, 300]:
            s = rng.integers(0, p, size=n).astype(np.uint64)
            assert oracle.lc(p, s) == (n + 1) // 2
            k = int(rng.integers(1, p))
            scaled = np.array([(int(x) * k) % p for x in s], dtype=np.uint64)
            assert oracle.lc(p, scaled) == oracle.lc(p, s)


def test_byte_count_formula():
    # stats: bytes = sum_i 8 + 16 l_i + [d_i != 0] 24 jmax_i + [change] 16 (old_l + 1)
    s = np.zeros(10, dtype=np.uint64)
    _, st = oracle.lc(M61, s, with_stats=True)
    assert st["bytes"] == 80 and st["mulmods"] == 0
    s = np.array([1], dtype=np.uint64)   # d=1 at i=0: jmax = min(1, 0) = 0, change copies 1 word
    _, st = oracle.lc(M61, s, with_stats=True)
    assert st["bytes"] == 8 + 16 and st["changes"] == 1
