"""GPU parity of Berlekamp-Massey (libmodlin.so through its C ABI) against the oracle.

Bit-exact bar: the linear complexity of every sequence equals the oracle's.
Covers the reference's known answers and family (test_linear_complexity.rs), the
README answers (272, 503, 5000, 65280), batches of uneven lengths, recurrences
of known order, u32 inputs and the empty-sequence error.
"""
import numpy as np
import pytest

import families
from conftest import modlin, oracle, workloads

pytestmark = pytest.mark.gpu
M61 = (1 << 61) - 1


def test_known_answers(ctx):
    p = M61
    f = modlin.Field(p)
    s = [1, 2, 3]
    for i in range(3, 40):
        s.append((3 * s[i - 1] + 5 * s[i - 2] + 7 * s[i - 3]) % p)
    assert ctx.linear_complexity(f, np.array(s, dtype=np.uint64)) == 3
    assert ctx.linear_complexity(f, np.full(20, 4, dtype=np.uint64)) == 1
    assert ctx.linear_complexity(f, np.zeros(20, dtype=np.uint64)) == 0


def test_empty_sequence_rejected(ctx):
    with pytest.raises(ValueError):
        ctx.linear_complexity(modlin.Field(M61), np.zeros(0, dtype=np.uint64))


def test_reference_family(ctx):
    by_p = {}
    for p, label, s, k in families.lc_family():
        by_p.setdefault(p, []).append((label, s, k))
    for p, items in by_p.items():
        got = ctx.lc_batch(modlin.Field(p), [s for _, s, _ in items])
        for (label, s, k), g in zip(items, got):
            assert g == oracle.lc_ref(p, s), (p, label)
            if k is not None:
                assert g <= k


def test_readme_short(ctx):
    s = (modlin.prng_fill(modlin.PRNG_MIXMAX17, 1, 1000) % np.uint64(M61)).astype(np.uint64)
    assert ctx.linear_complexity(modlin.Field(M61), s) == 272
    s = (modlin.prng_fill(modlin.PRNG_XOROSHIRO128PP, 0, 1000) % np.uint64(2)).astype(np.uint64)
    assert ctx.linear_complexity(modlin.Field(2), s) == 503
    s = (modlin.prng_fill(modlin.PRNG_XOROSHIRO128PP, 0, 10000) % np.uint64(M61)).astype(np.uint64)
    assert ctx.linear_complexity(modlin.Field(M61), s) == 5000


def test_c4_subset_against_oracle(ctx):
    seqs, specs = workloads.c4(count=48)
    seqs = [s[:3000] for s in seqs]
    got = ctx.lc_batch(modlin.Field(M61), seqs)
    for i, (s, g) in enumerate(zip(seqs, got)):
        assert g == oracle.lc(M61, s), (i, specs[i])


def test_recurrence_orders(ctx):
    rng = np.random.default_rng(4)
    for p in [65521, (1 << 31) - 1, M61, (1 << 63) - 25]:
        seqs, orders = [], []
        for L in [1, 16, 100, 400]:
            coef = rng.integers(1, p, size=L, dtype=np.uint64)
            init = rng.integers(0, p, size=L, dtype=np.uint64)
            seqs.append(modlin.lfsr(p, coef, init, 3 * L + 50))
            orders.append(L)
        assert ctx.lc_batch(modlin.Field(p), seqs) == orders


def test_u32_inputs(ctx):
    p = 65521
    rng = np.random.default_rng(2)
    seqs = [rng.integers(0, p, size=n).astype(np.uint32) for n in [10, 333, 2000]]
    got = ctx.lc_batch(modlin.Field(p), seqs)
    assert got == [oracle.lc(p, s.astype(np.uint64)) for s in seqs]


def test_readme_mixmax256_long(ctx):
    s = workloads.c5_bm(200_000)
    print("BM n=200000 MIXMAX-256", flush=True)
    assert ctx.linear_complexity(modlin.Field(M61), s) == 65280   # README.md:107


@pytest.mark.parametrize("p", [2, 65521, (1 << 31) - 1, M61, (1 << 63) - 25])
def test_cooperative_matches_oracle(ctx, p):
    """One sequence on the whole GPU (forced down to short lengths with coop_min):
    uniform, recurrences of known order, constant and zero sequences, MIXMAX."""
    rng = np.random.default_rng(11)
    f = modlin.Field(p)
    cases = [("uniform", rng.integers(0, p, size=5000, dtype=np.uint64))]
    for L in [1, 37, 700]:
        coef = rng.integers(1, p, size=L, dtype=np.uint64)
        init = rng.integers(0, p, size=L, dtype=np.uint64)
        cases.append((f"lfsr{L}", modlin.lfsr(p, coef, init, 2 * L + 3000)))
    cases.append(("const", np.full(3000, 1 % p, dtype=np.uint64)))
    cases.append(("zero", np.zeros(2500, dtype=np.uint64)))
    cases.append(("mixmax17", (modlin.prng_fill(modlin.PRNG_MIXMAX17, 5, 4000) % np.uint64(p)).astype(np.uint64)))
    for label, s in cases:
        print("coop BM", p, label, s.size, flush=True)
        got = ctx.lc_batch(f, [s], coop_min=1000)
        assert ctx.stats()["bm_coop"] > 0, label
        assert got[0] == oracle.lc(p, s), (p, label)


def test_cooperative_equals_single_workgroup(ctx):
    s = (modlin.prng_fill(modlin.PRNG_XOROSHIRO128PP, 3, 20000) % np.uint64(M61)).astype(np.uint64)
    f = modlin.Field(M61)
    coop = ctx.lc_batch(f, [s], coop_min=1)
    assert ctx.stats()["bm_coop"] > 0
    single = ctx.lc_batch(f, [s], coop_min=1 << 31)
    assert ctx.stats()["bm_coop"] == 0
    assert coop == single == [oracle.lc(M61, s)]


@pytest.mark.parametrize("wgs", [1, 3, 40, 130])
def test_blocked_cooperative_workgroup_counts(ctx, wgs):
    """The blocked cooperative kernel with few workgroups: each one adds up several
    of the 128 exchanged values and owns up to 16 chunks of c and b (the count is
    raised to what the chunks need)."""
    rng = np.random.default_rng(5 + wgs)
    f = modlin.Field(M61)
    coef = rng.integers(1, M61, size=300, dtype=np.uint64)
    init = rng.integers(0, M61, size=300, dtype=np.uint64)
    cases = [rng.integers(0, M61, size=3000, dtype=np.uint64), modlin.lfsr(M61, coef, init, 2500),
             np.zeros(1500, dtype=np.uint64)]
    ctx.set_tuning(bm_coop_wgs=wgs)
    try:
        for s in cases:
            print("blocked BM", wgs, s.size, flush=True)
            got = ctx.lc_batch(f, [s], coop_min=1000)
            assert ctx.stats()["bm_coop"] >= wgs
            assert got[0] == oracle.lc(M61, s)
    finally:
        ctx.set_tuning()


@pytest.mark.parametrize("threads", [256, 512, 1024])
def test_batch_thread_counts(ctx, threads):
    """The batched kernel at every thread count it is built for, on a mixed batch."""
    rng = np.random.default_rng(threads)
    f = modlin.Field(M61)
    seqs = [rng.integers(0, M61, size=int(n), dtype=np.uint64) for n in rng.integers(1, 3000, size=12)]
    coef = rng.integers(1, M61, size=40, dtype=np.uint64)
    seqs.append(modlin.lfsr(M61, coef, rng.integers(0, M61, size=40, dtype=np.uint64), 2000))
    seqs.append(np.zeros(700, dtype=np.uint64))
    ctx.set_tuning(bm_threads=threads)
    try:
        print("batch BM threads", threads, flush=True)
        assert ctx.lc_batch(f, seqs, method=modlin.METHOD_PLAIN) == [oracle.lc(M61, s) for s in seqs]
    finally:
        ctx.set_tuning()


@pytest.mark.parametrize("p", [2, 3, 65521, (1 << 31) - 1, M61, (1 << 62) - 57, (1 << 63) - 25])
def test_blocked_batch_equals_plain(ctx, p):
    """Blocks of 64 steps per pass (the default for p < 2^62) and one step per pass
    give the oracle's linear complexities: sequences shorter than a block, lengths
    around block multiples, recurrences whose order crosses block boundaries,
    constant, zero and single-spike sequences, and long low-order runs."""
    rng = np.random.default_rng(p % 1000 + 17)
    f = modlin.Field(p)
    seqs = [rng.integers(0, p, size=int(n), dtype=np.uint64) for n in (1, 2, 63, 64, 65, 127, 128, 129, 700, 2500)]
    for L in (1, 5, 64, 65, 130):
        coef = rng.integers(1, p, size=L, dtype=np.uint64)
        seqs.append(modlin.lfsr(p, coef, rng.integers(0, p, size=L, dtype=np.uint64), 3 * L + 400))
    seqs.append(np.full(900, 1 % p, dtype=np.uint64))
    seqs.append(np.zeros(333, dtype=np.uint64))
    spike = np.zeros(600, dtype=np.uint64)
    spike[450] = 1 % p
    seqs.append(spike)
    # low order over a long run: x^m B moves thousands of coefficients above C, so
    # whole rounds of coefficients between the two hold neither
    coef = rng.integers(1, p, size=5, dtype=np.uint64)
    seqs.append(modlin.lfsr(p, coef, rng.integers(0, p, size=5, dtype=np.uint64), 3000))
    late = np.zeros(3000, dtype=np.uint64)
    late[2800] = 1 % p
    seqs.append(late)
    print("blocked vs plain BM", p, len(seqs), flush=True)
    want = [oracle.lc(p, s) for s in seqs]
    assert ctx.lc_batch(f, seqs) == want
    assert ctx.lc_batch(f, seqs, method=modlin.METHOD_PLAIN) == want


def test_blocked_batch_high_order_m61(ctx):
    """The batched blocked kernel on the matrix cores (2^61-1) at large l: dense
    recurrences of order about 1500 and 4600 (many coefficient rounds of the E sums
    and of the new C and B, with the E sums' periodic folds), and one recurrence
    whose coefficients and initial values are all p-1, (p-1)/2 or (p+1)/2 (the
    balanced-digit edges).  Their linear complexity is exactly the order (a
    nonzero last coefficient and a full-rank Hankel start, w.h.p.)."""
    p = M61
    rng = np.random.default_rng(1500)
    f = modlin.Field(p)
    seqs, orders = [], []
    for L in (1500, 4600):
        coef = rng.integers(1, p, size=L, dtype=np.uint64)
        init = rng.integers(0, p, size=L, dtype=np.uint64)
        seqs.append(modlin.lfsr(p, coef, init, 2 * L + 300))
        orders.append(L)
    edge = np.array([p - 1, (p - 1) // 2, (p + 1) // 2], dtype=np.uint64)
    L = 1200
    seqs.append(modlin.lfsr(p, edge[rng.integers(0, 3, size=L)], edge[rng.integers(0, 3, size=L)], 2 * L + 300))
    orders.append(L)
    print("blocked BM high order", orders, flush=True)
    want = [oracle.lc(p, s) for s in seqs]
    assert want == orders
    assert ctx.lc_batch(f, seqs) == want
    assert ctx.stats()["bm_coop"] == 0
    assert ctx.lc_batch(f, seqs, method=modlin.METHOD_PLAIN) == want


@pytest.mark.parametrize("n, coop", [(5000, False), (200_000, True)])
def test_lc_progress_in_steps(ctx, n, coop):
    """modlin_linear_complexity reports progress in steps (fp.rs:273): monotone,
    total n, the last report n -- on the one-workgroup blocked kernel and on the
    cooperative one (the README's MIXMAX-256 run)."""
    s = workloads.c5_bm(n) if coop else (modlin.prng_fill(modlin.PRNG_XOROSHIRO128PP, 2, n) % np.uint64(M61)).astype(
        np.uint64)
    seen = []
    lc = ctx.linear_complexity(modlin.Field(M61), s, progress=lambda done, total: seen.append((done, total)))
    assert lc == (65280 if coop else oracle.lc(M61, s))
    assert seen and all(t == n for _, t in seen)
    done = [d for d, _ in seen]
    assert done == sorted(done) and done[-1] == n
    if coop:
        assert len(done) >= 3   # reports while the kernel runs
