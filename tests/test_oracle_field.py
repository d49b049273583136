"""Pins the oracle's (and the host library's) Field arithmetic.

Restates fp.rs:314-367 (barrett_matches_division): the nine primes, the corner
values and 256 mixed samples each for reduce_u128 and mul, checked against the
exact remainder of an integer division (the closed form Barrett must equal).
"""
import pytest

from conftest import modlin, oracle

PRIMES = [2, 3, 97, 65537, (1 << 31) - 1, (1 << 61) - 1, 2305843009213693951, (1 << 62) + 135, (1 << 63) - 25]
U128 = (1 << 128) - 1


def corners(p):
    c = [0, 1, p - 1, p, p + 1, U128, U128 - 1, 1 << 127, (1 << 127) | 1]
    for i in range(256):
        c.append((oracle.mix2(i) << 64) | oracle.mix2(i ^ 0xABCD))
    return c


@pytest.mark.parametrize("p", PRIMES)
def test_oracle_barrett_matches_division(p):
    for t in corners(p):
        assert oracle.reduce_u128(p, t) == t % p, (p, t)
    for x in [0, 1, p - 1, p, (1 << 64) - 1, (1 << 64) - 2]:
        assert oracle.reduce(p, x) == x % p
    for i in range(256):
        a = oracle.mix2(i) % p
        b = oracle.mix2(i ^ 0x1234) % p
        assert oracle.mul(p, a, b) == a * b % p


@pytest.mark.parametrize("p", PRIMES)
def test_oracle_field_constants(p):
    f = oracle.field(p)
    assert f["mu"] == U128 // p
    assert f["safe_batch"] == max(1, min(U128 // ((p - 1) ** 2) if p > 1 else U128, (1 << 64) - 1))


def test_oracle_safe_batch_known():
    # fp.rs:53-59 and the blocking notes: 64 at 2^61-1, 4 at 2^63-25.
    assert oracle.field((1 << 61) - 1)["safe_batch"] == 64
    assert oracle.field((1 << 63) - 25)["safe_batch"] == 4
    assert oracle.field((1 << 31) - 1)["safe_batch"] == (1 << 64) - 1


@pytest.mark.parametrize("p", [2, 3, 5, 7, 97, 65521])
def test_oracle_inverse_brute_force(p):
    for a in range(1, min(p, 200)):
        v = oracle.inv(p, a)
        assert a * v % p == 1


@pytest.mark.parametrize("p", PRIMES)
def test_library_host_field(p):
    """The library's host Field (modlin_field_init / reduce / reduce_u128)."""
    f = modlin.Field(p)
    assert f.mu == U128 // p
    assert f.safe_batch == oracle.field(p)["safe_batch"]
    for t in corners(p):
        assert f.reduce_u128(t) == t % p
    for x in [0, 1, p - 1, p, (1 << 64) - 1]:
        assert f.reduce(x) == x % p


@pytest.mark.parametrize("p", [0, 1, 1 << 63, (1 << 64) - 1])
def test_library_rejects_bad_modulus(p):
    with pytest.raises(ValueError):
        modlin.Field(p)
