"""Pins the input generators (prng.rs) in the oracle and in the library.

Known-answer vectors of tests/test_prng.rs: the first ten raw MIXMAX outputs
(next_u64() >> 3) for seed 1234567890123456789 and N = 17, 240, 256
(test_prng.rs:23-62), xoroshiro128++ seed 0 (test_prng.rs:87-98) and its raw
recurrence anchor (test_prng.rs:78-80); SplitMix64 and xoshiro256** against the
published reference outputs (first outputs of splitmix64.c from state 0 and of
xoshiro256starstar.c from state {1, 2, 3, 4}).
"""
import numpy as np
import pytest

from conftest import modlin, oracle

SEED = 1234567890123456789
MIXMAX = {
    17: [725153038902651861, 1378221679953707950, 1286954255613532202, 2183378240693018162, 480095223529858829,
         1028486193651132422, 1683776858464191402, 752698145398996242, 1196878046144236647, 1221112387665998636],
    240: [1894724080392768937, 2178403735632896525, 82974682673747095, 1362029167703491378, 175589691633294316,
          1981764336260928646, 1644676922156291251, 463159379824582110, 73102843550620337, 783499878333038877],
    256: [379984832372095102, 2262036764019001978, 952890261108261993, 1848869284066170746, 412703182074144958,
          1796627384723021327, 1959941208714421405, 1149684300799040022, 525039624349349232, 110320404022327900],
}
XORO = [8027914721839836897, 13805533416164201645, 5256508173613850168, 7973558954284022901, 8526501294691771125,
        6116102375994396471, 16028966417245382669, 12808598746819302742, 15824426267781808726, 5829521525559713354]
KIND = {17: modlin.PRNG_MIXMAX17, 240: modlin.PRNG_MIXMAX240, 256: modlin.PRNG_MIXMAX256}


@pytest.mark.parametrize("n", [17, 240, 256])
def test_mixmax_matches_root(n):
    assert [int(x) >> 3 for x in oracle.mixmax(n, SEED, 10)] == MIXMAX[n]
    assert [int(x) >> 3 for x in modlin.prng_fill(KIND[n], SEED, 10)] == MIXMAX[n]


def test_xoroshiro_reference():
    assert [int(x) for x in oracle.xoroshiro128pp(0, 10)] == XORO
    assert [int(x) for x in modlin.prng_fill(modlin.PRNG_XOROSHIRO128PP, 0, 10)] == XORO


def test_xoroshiro_raw_anchor():
    # From raw state [1, 0]: first output 131073, next state [562949955518465, 268435456].
    M = (1 << 64) - 1
    rotl = lambda x, k: ((x << k) | (x >> (64 - k))) & M
    s0, s1 = 1, 0
    out = (rotl((s0 + s1) & M, 17) + s0) & M
    s1 ^= s0
    s0, s1 = rotl(s0, 49) ^ s1 ^ ((s1 << 21) & M), rotl(s1, 28)
    assert out == 131073 and [s0, s1] == [562949955518465, 268435456]


def test_splitmix64_published():
    # splitmix64.c from x = 0: first outputs.
    want = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    assert [int(v) for v in oracle.splitmix64(0, 3)] == want
    assert [int(v) for v in modlin.prng_fill(modlin.PRNG_SPLITMIX64, 0, 3)] == want


def _xoshiro_from_state(s, count):
    M = (1 << 64) - 1
    rotl = lambda x, k: ((x << k) | (x >> (64 - k))) & M
    s = list(s)
    out = []
    for _ in range(count):
        out.append((rotl((s[1] * 5) & M, 7) * 9) & M)
        t = (s[1] << 17) & M
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]
        s[2] ^= t
        s[3] = rotl(s[3], 45)
    return out


def test_xoshiro256ss_published():
    # xoshiro256starstar.c from state {1, 2, 3, 4}: 11520, 0, 1509978240, 1215971899390074240.
    assert _xoshiro_from_state([1, 2, 3, 4], 4) == [11520, 0, 1509978240, 1215971899390074240]
    # seeding convention: four SplitMix64 draws of the seed
    st = [int(v) for v in oracle.splitmix64(42, 4)]
    want = _xoshiro_from_state(st, 16)
    assert [int(v) for v in oracle.xoshiro256ss(42, 16)] == want
    assert [int(v) for v in modlin.prng_fill(modlin.PRNG_XOSHIRO256SS, 42, 16)] == want


def test_mixmax_zero_seed_rejected():
    with pytest.raises(ValueError):
        oracle.mixmax(17, 0, 1)
    with pytest.raises(ValueError):
        modlin.prng_fill(modlin.PRNG_MIXMAX17, 0, 1)


@pytest.mark.parametrize("kind,n", [(modlin.PRNG_MIXMAX17, 17), (modlin.PRNG_MIXMAX240, 240),
                                    (modlin.PRNG_MIXMAX256, 256)])
def test_library_matches_oracle_long_streams(kind, n):
    for seed in [1, 7, 3]:
        a = modlin.prng_fill(kind, seed, 5000, skip=123)
        b = oracle.mixmax(n, seed, 5000, skip=123)
        assert np.array_equal(a, b)


def test_library_matches_oracle_xor_family():
    for seed in [0, 1, 5, 11]:
        assert np.array_equal(modlin.prng_fill(modlin.PRNG_XOROSHIRO128PP, seed, 4096), oracle.xoroshiro128pp(seed, 4096))
        assert np.array_equal(modlin.prng_fill(modlin.PRNG_XOSHIRO256SS, seed, 4096), oracle.xoshiro256ss(seed, 4096))
        assert np.array_equal(modlin.prng_fill(modlin.PRNG_SPLITMIX64, seed, 4096), oracle.splitmix64(seed, 4096))


def test_names():
    assert modlin.prng_name(modlin.PRNG_MIXMAX17) == "MIXMAX (TRandomMixMax17, N=17)"
    assert modlin.prng_name(modlin.PRNG_MIXMAX256) == "MIXMAX (TRandomMixMax256, N=256, skip=2)"
    assert modlin.prng_name(modlin.PRNG_XOROSHIRO128PP) == "xoroshiro128++"
