"""Pins the p-values (stats.rs) in both the oracle and the library's host code.

Known answers of tests/test_stats.rs:11-71 (exact where the reference asserts
equality, the same tolerances elsewhere), the printed strings of the README, the
pmf-sum identity of the LC CDF, and the Rust `{:?}` float formatting rules.
Beyond these values, bit-level equality with the Rust binary is by construction
(same libm functions and operation order) -- parity unpinned there.
"""
import math

import pytest

from conftest import modlin, oracle

M61 = (1 << 61) - 1
IMPLS = {
    "oracle": (oracle.corank_prob, oracle.corank_tail_pvalue, oracle.lc_left_tail_pvalue, oracle.pretty_p_value),
    "library": (modlin.corank_prob, modlin.corank_tail_pvalue, modlin.lc_left_tail_pvalue, modlin.pretty_p_value),
}


@pytest.fixture(params=list(IMPLS))
def impl(request):
    return IMPLS[request.param]


def test_corank_distribution_sane(impl):  # test_stats.rs:11-23
    prob, _, _, _ = impl
    p0 = prob(M61, 64, 0)
    assert abs(p0 - (1.0 - 1.0 / float(M61))) / p0 < 1e-12
    p1 = prob(M61, 64, 1)
    assert abs(p1 - 1.0 / float(M61)) / p1 < 1e-6
    assert abs(prob(2, 256, 0) - 0.2887880951) < 1e-6
    s = sum(prob(M61, 64, d) for d in range(65))
    assert abs(s - 1.0) < 1e-12


def test_single_observation_tails(impl):  # test_stats.rs:25-37
    _, tail, lc, _ = impl
    assert tail(M61, 305, 0) == 1.0
    p1 = tail(M61, 305, 1)
    assert abs(p1 - 1.0 / float(M61)) / p1 < 1e-6
    assert tail(M61, 305, 33) == 1e-307
    assert lc(M61, 10_000, 5000) == 1.0
    assert lc(M61, 10_000, 272) == 1e-307
    v = lc(2, 100, 40)
    assert 0.0 < v < 1.0


def test_lc_left_tail_exact(impl):  # test_stats.rs:39-71
    _, _, lc, _ = impl
    assert abs(lc(2, 2, 1) - 0.75) < 1e-12
    assert abs(lc(2, 5, 2) - 0.34375) < 1e-12
    assert abs(lc(2, 5, 3) - 0.84375) < 1e-12

    def pmf(q, n, l):
        if l == 0:
            return q ** (-n)
        if 2 * l <= n:
            return (q - 1.0) * q ** (2 * l - n - 1)
        return (q - 1.0) * q ** (n - 2 * l)

    for q in [2, 3, 5, 7]:
        for n in range(1, 25):
            cdf = 0.0
            for ell in range(n + 1):
                cdf += pmf(float(q), n, ell)
                got = lc(q, n, ell)
                assert abs(got - min(cdf, 1.0)) <= 1e-9 * cdf + 1e-15, (q, n, ell)


def test_readme_strings(impl):
    _, tail, lc, pretty = impl
    assert pretty(tail(M61, 500, 432)) == "1e-307"          # README.md:71
    assert pretty(tail(M61, 1000, 0)) == "1"                # README.md:129, 149
    assert pretty(lc(M61, 1000, 272)) == "1e-307"           # README.md:88
    assert pretty(lc(2, 1000, 503)) == "0.9947916666666666" # README.md:137
    assert pretty(lc(M61, 10000, 5000)) == "1"              # README.md:158
    assert pretty(lc(M61, 200000, 65280)) == "1e-307"       # README.md:107


@pytest.mark.parametrize("x,s", [
    (0.0, "0"), (1.0, "1"), (0.5, "0.5"), (2.5e-5, "2.5e-5"), (1e-4, "0.0001"), (1e16, "1e16"),
    (123.0, "123.0"), (1.5e300, "1.5e300"), (0.1, "0.1"), (1.0 / 3.0, "0.3333333333333333"),
    (4.336808689942018e-19, "4.336808689942018e-19"), (9.999999999999999e-5, "9.999999999999999e-5"),
    (5e-324, "5e-324"), (0.2887880950866029, "0.2887880950866029"), (1234567.0, "1234567.0"),
])
def test_rust_debug_format(impl, x, s):
    assert impl[3](x) == s


def test_oracle_and_library_agree_bitwise():
    for p in [2, 3, 97, 65521, (1 << 31) - 1, M61, (1 << 63) - 25]:
        for n in [1, 2, 5, 64, 305, 1000]:
            for c in sorted({0, 1, 2, n // 3, n // 2, n - 1, n}):
                a = oracle.corank_tail_pvalue(p, n, c)
                b = modlin.corank_tail_pvalue(p, n, c)
                assert a == b or (math.isnan(a) and math.isnan(b)), (p, n, c)
                assert oracle.lc_left_tail_pvalue(p, n, c) == modlin.lc_left_tail_pvalue(p, n, c)
                assert oracle.pretty_p_value(a) == modlin.pretty_p_value(b)
