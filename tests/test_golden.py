"""The committed fixtures of the large workloads (tests/golden/*.json).

CPU checks: the workload generators still produce the inputs the fixtures were
written from (SHA-256 digests), the oracle reproduces the fixtures on the items
it finishes in seconds, and the fixtures agree with what the mathematics fixes
for every item: a uniform sequence of length n has linear complexity ceil(n/2),
an order-L recurrence with nonzero last coefficient has L, a MIXMAX-17 stream
272 (its state dimension times 16, README:88); MIXMAX-256 with skip 2 reaches
65280 (README:107); planted rows never raise the rank.
"""
import hashlib
import json
import os

import numpy as np

from conftest import oracle, workloads

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
M61 = (1 << 61) - 1


def _load(name):
    with open(os.path.join(HERE, name)) as fh:
        return json.load(fh)


def _digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:16]


def test_c2_fixture():
    fx = _load("c2_coranks.json")
    mats = workloads.c2()
    assert [m.shape[0] for m in mats] == fx["sides"]
    assert [_digest(m) for m in mats] == fx["sha256_16"]
    for i, m in enumerate(mats):
        n = m.shape[0]
        if i % 2 == 1:
            assert fx["coranks"][i] >= n - 272          # MIXMAX-17: rank <= 272
        if n <= 400:
            r, _, w = oracle.rank(M61, m, n, with_trace=True)
            assert n - r == fx["coranks"][i], i
            assert w == fx["work"][i], i


def test_c4_fixture():
    fx = _load("c4_lc.json")
    specs = workloads.c4_specs()
    assert [list(s) for s in specs] == fx["specs"]
    for (kind, n, L), lc in zip(specs, fx["lc"]):
        want = {0: (n + 1) // 2, 1: L, 2: 272}[kind]
        assert lc == want, (kind, n, L, lc)
    seqs, _ = workloads.c4(count=24)
    for i, s in enumerate(seqs):
        assert _digest(s) == fx["sha256_16"][i]
        lc, st = oracle.lc(M61, s, with_stats=True)
        assert lc == fx["lc"][i] and st["bytes"] == fx["bytes"][i] and st["mulmods"] == fx["mulmods"][i]


def test_c5_fixtures():
    bm = _load("c5_bm.json")
    assert bm["lc"] == 65280 and bm["n"] == 200_000
    assert _digest(workloads.c5_bm()) == bm["sha256_16"]
    rk = _load("c5_rank_coranks.json")
    assert rk["p"] == 65521 and rk["n"] == 2048 and len(rk["coranks"]) == 256
    mats = workloads.c5_rank(count=2)
    assert [_digest(m) for m in mats] == rk["sha256_16"][:2]


def test_c5_planted_members():
    """The planted-deficiency u32 members over 65521 have exactly the coranks the
    GPU batch test asserts (the oracle on every one, 2048 x 2048)."""
    mats = workloads.c5_planted()
    assert len(mats) == len(workloads.C5_PLANTED_DEFICIENCY)
    for d, m in zip(workloads.C5_PLANTED_DEFICIENCY, mats):
        assert m.dtype == np.uint32 and m.shape == (2048, 2048)
        assert 2048 - oracle.rank(65521, m.astype(np.uint64), 2048) == d


def test_mixmax256_block_rank():
    """Rows that all start an output block of MIXMAX-256 (255 divides n) have rank
    at most the state dimension 256; rank exactly 256 once there are 256 rows."""
    for n in (255, 510):
        a = (workloads.M.prng_fill(workloads.M.PRNG_MIXMAX256, 1, n * n) % np.uint64(M61)).astype(np.uint64)
        assert oracle.rank(M61, a, n) == min(n, 256)
