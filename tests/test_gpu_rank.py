"""GPU parity of rank (libmodlin.so through its C ABI) against the oracle.

Bit-exact bar: the rank (hence corank) of every matrix equals the oracle's.
Covers the reference's known answers and family (test_rank.rs), the README
cases, panel and top-block boundaries (sides 1..129, 63/64/65), pivots the top
block cannot see (zero top rows, permutation matrices), sparse and zero
matrices, the prime classes of the Barrett code paths (k <= 62, k = 63), u32
inputs, batches of mixed sizes, and the planted-corank matrices of the large
configurations (size-independent property: corank exactly 64).
"""
import numpy as np
import pytest

import families
from conftest import modlin, oracle, workloads

pytestmark = pytest.mark.gpu
M61 = (1 << 61) - 1
METHODS = [modlin.METHOD_PLAIN, modlin.METHOD_MFMA]


def _rank_all(ctx, p, mats, method=modlin.METHOD_PLAIN, panel=0):
    return ctx.rank_batch(modlin.Field(p), mats, method=method, panel=panel)


def test_known_matrices(ctx):
    f = modlin.Field(M61)
    assert ctx.rank(f, np.eye(3, dtype=np.uint64)) == 3
    assert ctx.rank(f, np.array([1, 2, 3, 4, 5, 6, 5, 7, 9], dtype=np.uint64)) == 2
    assert ctx.rank(f, np.zeros(16, dtype=np.uint64)) == 0
    assert ctx.rank(modlin.Field(2), np.array([1, 1, 0, 0, 1, 1, 1, 0, 1], dtype=np.uint64)) == 2


@pytest.mark.parametrize("method", METHODS)
def test_reference_family(ctx, method):
    by_p = {}
    for p, n, target, a in families.rank_family():
        by_p.setdefault(p, []).append((n, target, a))
    for p, items in by_p.items():
        got = _rank_all(ctx, p, [a for _, _, a in items], method)
        for (n, target, a), g in zip(items, got):
            print(f"family p={p} n={n} target={target}", flush=True)
            assert g == oracle.rank_ref(p, a, n), (p, n, target)


def test_readme_c1(ctx):
    mm, ctl = workloads.c1()
    r = _rank_all(ctx, M61, [mm, ctl])
    assert [500 - r[0], 500 - r[1]] == [432, 0]


def _hidden_pivot_cases(p, rng):
    cases = []
    for n in [5, 70, 130]:
        a = np.zeros((n, n), dtype=np.uint64)
        a[np.arange(n), n - 1 - np.arange(n)] = 1        # anti-diagonal permutation
        cases.append(a)
        b = rng.integers(0, p, size=(n, n)).astype(np.uint64)
        b[: min(n - 1, 64)] = 0                          # zero top block, pivots below
        cases.append(b)
        c = np.zeros((n, n), dtype=np.uint64)
        c[n - 1, :] = rng.integers(1, p, size=n)         # a single nonzero row at the bottom
        cases.append(c)
        d = rng.integers(0, p, size=(n, n)).astype(np.uint64)
        d[:, : n // 2] = 0                               # zero leading columns
        cases.append(d)
        e = (rng.random((n, n)) < 0.02).astype(np.uint64) * rng.integers(1, p, size=(n, n)).astype(np.uint64)
        cases.append(e)                                  # sparse
    return cases


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("p", [2, 97, M61])
def test_hidden_pivots_and_sparse(ctx, p, method):
    rng = np.random.default_rng(p % 1000)
    cases = _hidden_pivot_cases(p, rng)
    got = _rank_all(ctx, p, cases, method)
    for a, g in zip(cases, got):
        assert g == oracle.rank_ref(p, a, a.shape[0])


# one or more primes for every limb count L = 1..8 of the MFMA update, and both
# Barrett classes (bit length <= 62 and = 63)
PRIMES = [2, 3, 127, 131, 251, 257, 32749, 32771, 65521, 8388593, (1 << 31) - 1, 68719476731, 17592186044399,
          4503599627370449, M61, (1 << 62) + 135, (1 << 63) - 25]


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("p", PRIMES)
def test_prime_classes(ctx, p, method):
    rng = np.random.default_rng(7)
    mats = []
    for n in [1, 2, 63, 64, 65, 127, 128, 129, 200]:
        mats.append(rng.integers(0, p, size=(n, n), dtype=np.uint64))
        r = max(1, n // 3)
        b = rng.integers(0, p, size=(r, n), dtype=np.uint64)
        cf = rng.integers(0, p, size=(n, r), dtype=np.uint64)
        mats.append(modlin.combine_rows(p, b, cf))          # rank <= n/3
    # extreme residues: 0, 1, p-1 and (p-1)/2, (p+1)/2 (the balanced-digit edges)
    edge = np.array([0, 1, p - 1, (p - 1) // 2, (p + 1) // 2 % p], dtype=np.uint64)
    mats.append(edge[rng.integers(0, 5, size=(130, 130))])
    mats.append(np.full((70, 70), p - 1, dtype=np.uint64) - np.eye(70, dtype=np.uint64) * (p > 2))
    got = _rank_all(ctx, p, mats, method)
    for a, g in zip(mats, got):
        assert g == oracle.rank(p, a, a.shape[0]), (p, a.shape)


@pytest.mark.parametrize("method", METHODS)
def test_panel_widths(ctx, method):
    rng = np.random.default_rng(1)
    p = M61
    mats = [rng.integers(0, p, size=(n, n), dtype=np.uint64) for n in [50, 100, 150]]
    mats.append(workloads.c3(300, p))
    want = [oracle.rank(p, a, a.shape[0]) for a in mats]
    for b in [1, 2, 7, 16, 32, 64]:
        assert _rank_all(ctx, p, mats, method, panel=b) == want, b


@pytest.mark.parametrize("method", METHODS)
def test_u32_inputs(ctx, method):
    p = 65521
    mats = workloads.c5_rank(count=4, n=300, p=p)
    got = _rank_all(ctx, p, mats, method)
    for a, g in zip(mats, got):
        assert g == oracle.rank(p, a.astype(np.uint64), 300)
    # a u32 rank-deficient matrix
    b = modlin.combine_rows(p, mats[0][:10].astype(np.uint64), mats[1][:, :10].astype(np.uint64))
    assert _rank_all(ctx, p, [b.astype(np.uint32)], method) == [10]


@pytest.mark.parametrize("method", METHODS)
def test_mixed_batch_c2_subset(ctx, method):
    mats = workloads.c2(count=16, max_side=700)
    got = _rank_all(ctx, M61, mats, method)
    for i, (a, g) in enumerate(zip(mats, got)):
        assert g == oracle.rank(M61, a, a.shape[0]), i


def test_empty_and_tiny(ctx):
    f = modlin.Field(97)
    assert ctx.rank_batch(f, []) == []
    assert ctx.rank_batch(f, [np.zeros(0, dtype=np.uint64)]) == [0]
    assert ctx.rank_batch(f, [np.array([5], dtype=np.uint64), np.array([0], dtype=np.uint64)]) == [1, 0]


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("p", [M61, (1 << 31) - 1])
def test_planted_corank_large(ctx, p, method):
    n = 4096
    a = workloads.c3(n, p)
    print(f"planted n={n} p={p} method={method}", flush=True)
    assert _rank_all(ctx, p, [a], method) == [n - 64]


@pytest.mark.parametrize("method", METHODS)
def test_planted_corank_block_elimination_lds_form(ctx, method):
    """Batches with a side of 24576 or more take the LDS form of the block
    elimination: the planted-corank matrix at that side (built on the device,
    4.5 GiB) has corank exactly 64 (size-independent property)."""
    n = 24576
    buf = workloads.c3_device(ctx, n, M61)
    try:
        print(f"planted lds form n={n} method={method}", flush=True)
        assert ctx.rank_batch(modlin.Field(M61), [buf], method=method) == [n - 64]
    finally:
        buf.free()


def test_readme_c1_both_methods(ctx):
    mm, ctl = workloads.c1()
    for method in METHODS:
        r = _rank_all(ctx, M61, [mm, ctl], method)
        assert [500 - r[0], 500 - r[1]] == [432, 0]


@pytest.mark.parametrize("method", METHODS)
def test_graph_replay_matches_direct(ctx, method):
    """The captured hipGraph of the iteration block, replayed (and re-captured when
    the batch changes), gives the same ranks as direct launches and the oracle."""
    rng = np.random.default_rng(3)
    batches = [workloads.c2(count=8, max_side=600),
               [rng.integers(0, M61, size=(n, n), dtype=np.uint64) for n in (300, 129, 64)],
               workloads.c2(count=8, max_side=600)]
    f = modlin.Field(M61)
    for mats in batches:
        want = [oracle.rank(M61, a, a.shape[0]) for a in mats]
        assert ctx.rank_batch(f, mats, method=method, graph=False) == want
        assert ctx.stats()["graph"] == 0.0
        for _ in range(2):
            assert ctx.rank_batch(f, mats, method=method) == want
            assert ctx.stats()["graph"] == 1.0


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("tuning", [dict(far_per_cu=2, near_per_cu=1, check_every=4),
                                    dict(far_per_cu=8, near_per_cu=4, gather_wgs=64, check_every=16),
                                    dict(panel=48), dict(panel=16, far_per_cu=6)])
def test_launch_parameters_do_not_change_ranks(ctx, method, tuning):
    """Every tuning-table setting gives the oracle's ranks (mixed sizes, deficient
    MIXMAX members, a planted corank)."""
    mats = workloads.c2(count=12, max_side=600) + [workloads.c3(520, M61)]
    print("tuning", tuning, "method", method, flush=True)
    ctx.set_tuning(**tuning)
    try:
        got = _rank_all(ctx, M61, mats, method)
    finally:
        ctx.set_tuning()
    for i, (a, g) in enumerate(zip(mats, got)):
        assert g == oracle.rank(M61, a, a.shape[0]), (i, tuning)


def test_tuning_table_loaded_and_cleared(ctx, tmp_path):
    table = tmp_path / "tuning.txt"
    table.write_text("rank m61 1024 many 2 1 256 4 32 0 0\n")
    ctx.load_tuning(str(table))
    try:
        mats = workloads.c2(count=6, max_side=900)
        got = _rank_all(ctx, M61, mats, modlin.METHOD_MFMA)
        assert got == [oracle.rank(M61, a, a.shape[0]) for a in mats]
    finally:
        try:
            ctx.load_tuning(None)   # the library's own table, if one is installed
        except ValueError:
            ctx.load_tuning("")
    with pytest.raises(ValueError):
        ctx.load_tuning(str(tmp_path / "missing.txt"))


@pytest.mark.parametrize("n", [1000, 3000])
def test_rank_progress_in_columns(ctx, n):
    """modlin_rank reports progress in columns (fp.rs:251): monotone, total n, the
    last report n; a deficient MIXMAX-17 matrix still reports every column (its
    eliminated rest is searched in one pass, so it may end within the first host
    check)."""
    rng = np.random.default_rng(n)
    f = modlin.Field(M61)
    for a, full in ((rng.integers(0, M61, size=(n, n), dtype=np.uint64), True),
                    ((modlin.prng_fill(modlin.PRNG_MIXMAX17, 9, n * n) % np.uint64(M61)).astype(np.uint64)
                     .reshape(n, n), False)):
        seen = []
        r = ctx.rank(f, a, progress=lambda done, total: seen.append((done, total)))
        assert r == oracle.rank(M61, a, n)
        assert seen and all(t == n for _, t in seen)
        done = [d for d, _ in seen]
        assert done == sorted(done) and done[-1] == n and len(set(done)) == len(done)
        if n >= 3000 and full:
            assert len(done) >= 2   # intermediate reports (one host check per 8 panels)


@pytest.mark.parametrize("p", [M61, (1 << 31) - 1, 65521, 251, (1 << 62) + 135])
def test_deferred_far_update(ctx, p):
    """The MFMA method's paired panels (far columns updated once per two panels,
    up to 128 pivots per pass, the second panel's pivot rows corrected inside its X
    product) give the oracle's ranks, as does one far update per panel: mixed
    sizes, deficient and planted matrices, hidden pivots (missed-pivot restarts
    between pairs), and panel widths below 64."""
    rng = np.random.default_rng(p % 997)
    mats = [rng.integers(0, p, size=(n, n), dtype=np.uint64) for n in (130, 200, 257, 390, 700)]
    mats += [workloads.c3(520, p), workloads.c3(777, p)]
    r = 90
    mats.append(modlin.combine_rows(p, rng.integers(0, p, size=(r, 600), dtype=np.uint64),
                                    rng.integers(0, p, size=(600, r), dtype=np.uint64)))
    mats += _hidden_pivot_cases(p, rng)[5:]
    z = rng.integers(0, p, size=(400, 400), dtype=np.uint64)
    z[:, 150:260] = 0                                   # a run of zero columns in mid-matrix
    mats.append(z)
    want = [oracle.rank(p, a, a.shape[0]) for a in mats]
    for panel in (0, 64, 48, 17):
        print(f"deferred p={p} panel={panel}", flush=True)
        assert _rank_all(ctx, p, mats, modlin.METHOD_MFMA, panel=panel) == want, panel
    got = ctx.rank_batch(modlin.Field(p), mats, method=modlin.METHOD_MFMA, defer=False)
    assert got == want


@pytest.mark.parametrize("p", [65521, (1 << 31) - 1, 4294967291, 251])
def test_u32_working_copy(ctx, p):
    """For p < 2^32 the MFMA method keeps the working copy in u32 entries: u64 and
    u32 inputs, in host and device memory, give the oracle's ranks, as does the u64
    working copy (flag MODLIN_FLAG_U64_STORE)."""
    rng = np.random.default_rng(p % 1009)
    mats = [rng.integers(0, p, size=(n, n), dtype=np.uint64) for n in (64, 129, 300, 513)]
    mats += [workloads.c3(450, p)]
    r = 70
    mats.append(modlin.combine_rows(p, rng.integers(0, p, size=(r, 380), dtype=np.uint64),
                                    rng.integers(0, p, size=(380, r), dtype=np.uint64)))
    mats += _hidden_pivot_cases(p, rng)[5:]
    want = [oracle.rank(p, a, a.shape[0]) for a in mats]
    f = modlin.Field(p)
    m32 = [a.astype(np.uint32) for a in mats]
    print(f"u32 store p={p} host u64 / u32", flush=True)
    assert ctx.rank_batch(f, mats, method=modlin.METHOD_MFMA) == want
    assert ctx.rank_batch(f, m32, method=modlin.METHOD_MFMA) == want
    print(f"u32 store p={p} device u64 / u32 / mixed", flush=True)
    dev = [ctx.to_device(a) for a in mats]
    dev32 = [ctx.to_device(a) for a in m32]
    try:
        assert ctx.rank_batch(f, dev, method=modlin.METHOD_MFMA) == want
        assert ctx.rank_batch(f, dev32, method=modlin.METHOD_MFMA) == want
        mixed = [dev[i] if i % 3 == 0 else dev32[i] if i % 3 == 1 else m32[i] for i in range(len(mats))]
        assert ctx.rank_batch(f, mixed, method=modlin.METHOD_MFMA) == want
    finally:
        for d in dev + dev32:
            d.free()
    print(f"u32 store p={p} u64 working copy", flush=True)
    assert ctx.rank_batch(f, mats, method=modlin.METHOD_MFMA, u32_store=False) == want
    assert ctx.rank_batch(f, m32, method=modlin.METHOD_MFMA, u32_store=False) == want


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("p", [M61, (1 << 31) - 1, 65521])
def test_pivot_free_panels(ctx, p, method):
    """A panel whose top block has no pivot makes the next iteration search every
    column left for the first nonzero (the reference's next pivot column and row):
    low-rank matrices whose rest is eliminated to zero, runs of zero columns of any
    width and position, leading and trailing zero columns, a lone nonzero entry."""
    rng = np.random.default_rng(p % 4093)
    mats = []
    for n, r in ((300, 5), (700, 70), (700, 130), (450, 64), (520, 1)):
        mats.append(modlin.combine_rows(p, rng.integers(0, p, size=(r, n), dtype=np.uint64),
                                        rng.integers(0, p, size=(n, r), dtype=np.uint64)))
    for n, lo, hi in ((400, 64, 300), (400, 100, 164), (500, 300, 500), (390, 0, 128), (600, 70, 530)):
        a = rng.integers(0, p, size=(n, n), dtype=np.uint64)
        a[:, lo:hi] = 0
        mats.append(a)
    z = np.zeros((333, 333), dtype=np.uint64)
    z[170, 331] = 1                                      # a lone nonzero far right
    mats.append(z)
    y = np.zeros((260, 260), dtype=np.uint64)
    y[200:, 230:] = rng.integers(1, p, size=(60, 30))    # nonzeros only bottom right
    mats.append(y)
    want = [oracle.rank(p, a, a.shape[0]) for a in mats]
    print(f"pivot-free panels p={p} method={method}", flush=True)
    assert _rank_all(ctx, p, mats, method) == want
    assert _rank_all(ctx, p, mats, method, panel=17) == want


@pytest.mark.parametrize("tuning", [dict(far_depth=1), dict(far_depth=2, far_quarters=1), dict(far_quarters=1),
                                    dict(far_depth=1, far_quarters=1, panel=40)])
@pytest.mark.parametrize("p", [M61, (1 << 31) - 1, 65521])
def test_far_update_forms(ctx, p, tuning):
    """The far update's K depth (one or two panels per pass) and limb-class
    grouping (quarters per B fragment, for 4 digits or fewer) give the oracle's
    ranks: mixed sizes, planted and low-rank matrices, hidden pivots."""
    rng = np.random.default_rng(p % 7919)
    mats = [rng.integers(0, p, size=(n, n), dtype=np.uint64) for n in (200, 333, 600)]
    mats += [workloads.c3(700, p)]
    r = 100
    mats.append(modlin.combine_rows(p, rng.integers(0, p, size=(r, 500), dtype=np.uint64),
                                    rng.integers(0, p, size=(500, r), dtype=np.uint64)))
    mats += _hidden_pivot_cases(p, rng)[10:]
    want = [oracle.rank(p, a, a.shape[0]) for a in mats]
    print("far form", p, tuning, flush=True)
    ctx.set_tuning(**tuning)
    try:
        assert _rank_all(ctx, p, mats, modlin.METHOD_MFMA) == want
    finally:
        ctx.set_tuning()


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("p", [M61, (1 << 31) - 1])
def test_missed_pivots_in_large_matrices(ctx, p, method):
    """Missed pivots while long updates run: the top 200 rows of 2048^2 and
    3072^2 matrices are zero in their first 100 columns, so the first panels'
    top blocks find no pivot there and every such column is a missed-pivot
    restart (a row swap committed while the previous panel's far update may
    still run).  Ranks by construction (size-independent): full rank for the
    random matrices, n - 64 when the last 64 rows are planted combinations of
    the first 128 rows of the modified matrix."""
    rng = np.random.default_rng(p % 4099)
    mats, want = [], []
    for n in (2048, 3072):
        a = rng.integers(0, p, size=(n, n), dtype=np.uint64)
        a[:200, :100] = 0
        mats.append(a)
        want.append(n)
        b = a.copy()
        b[n - 64:] = modlin.combine_rows(p, b[:128], rng.integers(0, p, size=(64, 128), dtype=np.uint64))
        mats.append(b)
        want.append(n - 64)
    print(f"missed pivots p={p} method={method}", flush=True)
    assert _rank_all(ctx, p, mats, method) == want
