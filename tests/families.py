"""The reference's own test input families, restated with numpy.

rank:  tests/test_rank.rs:99-149 -- a counter-based 3-round mix (test_rank.rs:11-16)
       feeds, in order: `target` random rows, then for every remaining row one
       coefficient per random row (zero coefficients skipped), then a Fisher-Yates
       row shuffle.  The counter runs across the whole family.
lc:    tests/test_linear_complexity.rs:93-133 -- a 2-round mix counter
       (test_linear_complexity.rs:95-99) feeds random streams of several lengths
       and order-k LFSR streams of length 150, per prime.
"""
from __future__ import annotations

import numpy as np

import conftest

M61 = (1 << 61) - 1
MASK = (1 << 64) - 1

RANK_PRIMES = [2, 97, M61]                                   # test_rank.rs:151
RANK_SIDES = [1, 2, 5, 63, 64, 65, 129, 200, 257, 300]       # test_rank.rs:153
LC_PRIMES = [2, 3, 97, 65537, M61]                           # test_linear_complexity.rs:106
LC_LENGTHS = [1, 2, 3, 5, 16, 50, 100, 200]                  # test_linear_complexity.rs:109
LC_ORDERS = [1, 4, 10, 25]                                   # test_linear_complexity.rs:118


def mix3_np(c: np.ndarray) -> np.ndarray:
    z = c.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def mix2_int(x: int) -> int:
    z = (x * 0x9E3779B97F4A7C15) & MASK
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    return z ^ (z >> 31)


class _Counter:
    def __init__(self, mixer):
        self.c = 0
        self.mixer = mixer

    def take(self, k: int) -> np.ndarray:
        out = self.mixer(np.arange(self.c + 1, self.c + 1 + k, dtype=np.uint64))
        self.c += k
        return out


def rank_family(primes=RANK_PRIMES, sides=RANK_SIDES):
    """Yields (p, n, target, matrix) exactly as test_rank.rs:99-149 builds them."""
    nxt = _Counter(mix3_np)
    for p in primes:
        for n in sides:
            for target in (n, n // 2, n // 3, 1, 0):
                target = min(target, n)
                a = np.zeros((n, n), dtype=np.uint64)
                a[:target] = (nxt.take(target * n) % np.uint64(p)).reshape(target, n)
                if n > target and target > 0:
                    coef = (nxt.take((n - target) * target) % np.uint64(p)).reshape(n - target, target)
                    a[target:] = conftest.modlin.combine_rows(p, a[:target], coef)
                r_draws = nxt.take(max(0, n - 1))
                for idx, r in enumerate(range(n - 1, 0, -1)):
                    s = int(r_draws[idx]) % (r + 1)
                    if s != r:
                        a[[r, s]] = a[[s, r]]
                yield p, n, target, a


def lc_family(primes=LC_PRIMES):
    """Yields (p, label, sequence, max_lc) as test_linear_complexity.rs:93-133 builds them."""
    c = 0

    def nxt():
        nonlocal c
        c = (c + 1) & MASK
        return mix2_int(c)

    for p in primes:
        for n in LC_LENGTHS:
            s = np.array([nxt() % p for _ in range(n)], dtype=np.uint64)
            yield p, f"random n={n}", s, None
        for k in LC_ORDERS:
            coeffs = [nxt() % p for _ in range(k)]
            s = [nxt() % p for _ in range(k)]
            for i in range(k, 150):
                v = 0
                for j, cj in enumerate(coeffs):
                    v = (v + cj * s[i - 1 - j]) % p
                s.append(v)
            yield p, f"lfsr k={k}", np.array(s, dtype=np.uint64), k
