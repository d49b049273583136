"""Benchmark: F_p rank columns/s of the batched rank on one MI355X.

Workload (one "step"): rank of the 64-matrix batch of the batched-rank
configuration -- square sides drawn log-uniform in [256, 4096] from
SplitMix64 seed 7, p = 2^61-1, u64 entries; even matrices uniform from
xoshiro256** seed 7, odd matrices consecutive MIXMAX-17 seed-7 stretches
(modlin-rs_amd/workloads.py:c2).  Inputs are resident in HBM before timing;
each step copies them into the library's workspace (the reference destroys its
input too) and runs the whole batch through one launch sequence.

Prints one JSON line.  value = total matrix columns ranked per second
(sum of sides / step time).  `roofline` refers to the dominant kernel, the
far-column trailing update k_update_far, timed with HIP events on the stream
it is launched on; `traffic` is its HBM bytes per launch from the committed
rocprofv3 counter summary; `near_far_update` gives the same figures for the
near and far update launches together; `cpu_baseline` is the
oracle's OpenMP port of fp.rs on a bounded sample of the same batch on this
host.  Extra keys report the plain (VALU) path on the same batch, the single
8192 x 8192 planted-corank matrix over 2^61-1 on both paths, the batched
Berlekamp-Massey workload (1024 sequences, steps/s) and the per-item checks
against the CPU-computed fixtures in tests/golden/.

--copy-calibration additionally ranks one all-zero 4096 x 4096 u32 matrix held
in device memory: its set-up kernel k_setup widens it into the workspace (64
MiB read, 128 MiB written, plus 16 KiB of row order), a copy kernel of known
size that a counter run uses to scale FETCH_SIZE / WRITE_SIZE to bytes.
"""
import argparse
import json
import platform
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "modlin-rs_amd"))

import modlin as M  # noqa: E402
import workloads as W  # noqa: E402

M61 = (1 << 61) - 1
PEAK_INT8_TOPS = 5000.0   # MI355X dense int8 MFMA, TOP/s
PEAK_HBM_GBS = 8000.0     # MI355X HBM3E, GB/s


def limb_count(p: int) -> int:
    """Signed 8-bit digits needed for the balanced residues |x| <= (p-1)/2."""
    half, L = (p - 1) // 2, 1
    while 127 * (256 ** L - 1) // 255 < half:
        L += 1
    return L


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    return O


_NATIVE = []


def _native_oracle():
    """The oracle compiled for this host (-march=native), built once per run."""
    if not _NATIVE:
        import tempfile
        O = _oracle()
        _NATIVE.append(O.load(O.build_native(tempfile.mkdtemp(prefix="modlin_oracle_"))))
    return _NATIVE[0]


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def _rank_run(O, lib, mats, threads, over_matrices):
    """Seconds to rank mats with the oracle library lib: one after another with
    the update's row loop on `threads` OpenMP threads, or (over_matrices) the
    matrices spread over `threads` host threads, one OpenMP thread each."""
    import threading
    if not over_matrices:
        lib.or_set_num_threads(threads)
        t0 = time.perf_counter()
        for m in mats:
            O.rank(M61, m, m.shape[0], lib=lib)
        return time.perf_counter() - t0
    order = sorted(range(len(mats)), key=lambda i: -mats[i].shape[0])  # largest first
    lock, nxt = threading.Lock(), [0]

    def worker():
        lib.or_set_num_threads(1)  # this thread's OpenMP team
        while True:
            with lock:
                if nxt[0] >= len(order):
                    return
                i = order[nxt[0]]
                nxt[0] += 1
            O.rank(M61, mats[i], mats[i].shape[0], lib=lib)

    th = [threading.Thread(target=worker) for _ in range(threads)]
    t0 = time.perf_counter()
    for t in th:
        t.start()
    for t in th:
        t.join()
    return time.perf_counter() - t0


def cpu_baseline(mats, w_target=2e10, w_cap=6e9, full=False):
    """The oracle's port of fp::rank on the host cores, from a -march=native
    build compiled here (gcc -O3 -fopenmp).  The sample is the batch's matrices in
    batch order, each of work W <= w_cap, until their W reaches w_target (the
    whole batch with full=True), W being the oracle's work count (the dense
    Schur-complement multiply-accumulates of a matrix, from the batch's fixture):
      value           the reference's way: one matrix after another, the update's
                      row loop over all host threads (OpenMP, as rayon in
                      fp.rs:202-248);
      batch_parallel  the same matrices spread over the host threads, one each;
      single_thread   one matrix after another on one thread, on the part of the
                      sample up to w_target / 4;
      portable_build  the x86-64-v3 library the tests load, as value.
    Each leg projects the whole batch's time from its rate per unit of W; the
    batch's largest matrices parallelise better than the sample's, so for the
    multi-threaded legs the projection is an upper bound."""
    O = _oracle()
    fx = fixture("c2_coranks.json")
    work = fx["work"] if fx else [float(m.shape[0]) ** 3 / 3 for m in mats]
    w_all = float(sum(work))
    native = _native_oracle()
    portable = O.load()
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    cores = min(cores, native.or_num_threads())

    def pick(target):
        idx, acc = [], 0.0
        for i, w in enumerate(work):
            if full:
                idx.append(i)
            elif w <= w_cap and acc < target:
                idx.append(i)
                acc += w
        return idx

    def leg(lib, idx, threads, over):
        sub = [mats[i] for i in idx]
        if full:
            print(f"cpu baseline: {len(sub)} matrices, {threads} threads{' over matrices' if over else ''}",
                  file=sys.stderr, flush=True)
        dt = _rank_run(O, lib, sub, threads, over)
        cols, w = sum(m.shape[0] for m in sub), float(sum(work[i] for i in idx))
        return {"value": cols / dt, "unit": "columns/s", "cores": threads, "seconds": dt, "matrices": len(idx),
                "sample": f"{len(idx)} of {len(mats)} matrices (sum n = {cols}, W = {w:.3g} of {w_all:.3g})",
                "projected_batch_s": dt * w_all / w}

    big, small = pick(w_target), pick(w_target / 4)
    warm = np.random.default_rng(1).integers(0, M61, size=(256, 256), dtype=np.uint64)
    for lib in (native, portable):  # thread pools and code paged in before any timing
        lib.or_set_num_threads(cores)
        O.rank(M61, warm, 256, lib=lib)
    ref = leg(native, big, cores, False)
    out = {"value": ref["value"], "unit": "columns/s", "cores": cores, "kind": "port",
           "sample": ref["sample"] + " in batch order, each W <= %.3g; one after another, update rows on all "
                                     "threads" % w_cap,
           "cpu_model": cpu_model(), "build": "gcc -O3 -march=native -fopenmp (compiled on this host)",
           "seconds": ref["seconds"], "projected_batch_s": ref["projected_batch_s"],
           "batch_parallel": leg(native, big, cores, True),
           "single_thread": leg(native, small, 1, False),
           "portable_build": dict(leg(portable, big, cores, False), build="gcc -O3 -march=x86-64-v3 -fopenmp")}
    portable.or_set_num_threads(cores)
    native.or_set_num_threads(cores)
    return out


def cpu_bm(seqs, long_seq, budget_s=4.0, prefixes=(20000, 40000)):
    """The oracle's port of fp::linear_complexity (one thread, as the reference),
    from the -march=native build: on the batch's sequences in order until the time
    budget is spent, and on two prefixes of the long sequence.  A BM step costs
    about twice the current length l, and l grows as n / 2 until it reaches the
    sequence's complexity (65,280 here, at n = 130,560), so the full run's time
    is projected from the prefixes as t(n) = b n^2 up to 2 l_max, then linear."""
    O, lib = _oracle(), _native_oracle()
    lib.or_set_num_threads(1)
    t0, steps, k = time.perf_counter(), 0, 0
    while k < len(seqs) and time.perf_counter() - t0 < budget_s:
        O.lc(M61, seqs[k], lib=lib)
        steps += seqs[k].size
        k += 1
    dt = time.perf_counter() - t0
    tp = []
    for n in prefixes:
        t1 = time.perf_counter()
        O.lc(M61, long_seq[:n], lib=lib)
        tp.append(time.perf_counter() - t1)
    n0, n1 = prefixes
    b = (tp[1] - tp[0]) / (n1 * n1 - n0 * n0)
    knee, n = 2 * 65280, long_seq.size
    proj = b * knee * knee + (n - knee) * 2 * b * knee
    lib.or_set_num_threads(lib.or_num_threads())
    return {"bm_batch": {"value": steps / dt, "unit": "steps/s", "cores": 1, "kind": "port",
                         "sample": f"first {k} sequences of the batch ({steps} steps), {dt:.1f} s"},
            "bm_long": {"value": tp[1] * 1e6 / n1, "unit": "us/step", "cores": 1, "kind": "port",
                        "sample": f"first {n1} of the {n} steps, {tp[1]:.2f} s (first {n0}: {tp[0]:.2f} s)",
                        "steps_20k_40k_us_per_step": (tp[1] - tp[0]) * 1e6 / (n1 - n0),
                        "projected_full_s": proj, "projected_us_per_step": proj * 1e6 / n}}


def traffic_per_launch(key="far_traffic_bytes_per_launch"):
    """HBM bytes per launch of the MFMA trailing update on this workload, from
    the committed counter summary (FETCH_SIZE and WRITE_SIZE runs, scaled by the
    same counters on a copy kernel of known size), or None.  `key` selects the
    far-column update alone or the mean over near and far launches."""
    path = os.path.join(ROOT, "profiles", "update_kernels_batched_rank_64_p61_counters.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            return float(line.split(":", 1)[1])
    return None


def copy_calibration(ctx, n=4096):
    dev = ctx.to_device(np.zeros((n, n), dtype=np.uint32))
    r = ctx.rank_batch(M.Field(M61), [dev])
    ctx.synchronize()
    dev.free()
    return {"copy_kernel": "k_setup", "elements": n * n, "bytes_read": 4 * n * n, "bytes_written": 8 * n * n + 4 * n,
            "rank": r[0]}


def fixture(name):
    path = os.path.join(ROOT, "tests", "golden", name)
    return json.load(open(path)) if os.path.exists(path) else None


def single_large(ctx, n, p, steps, plain=True, warmup=1):
    """One n x n planted-corank matrix (built on the device): MFMA and plain
    paths, columns/s; each result must be corank 64."""
    buf = W.c3_device(ctx, n, p)
    field = M.Field(p)
    dev = [buf]
    out = {"n": n, "p": p}
    L = limb_count(p)
    methods = (("mfma", M.METHOD_MFMA), ("plain", M.METHOD_PLAIN)) if plain else (("mfma", M.METHOD_MFMA),)
    for name, method in methods:
        ranks, t, st, _ = run(ctx, field, dev, method, steps, warmup)
        out[f"{name}_columns_per_s"] = n / t
        out[f"{name}_ms"] = t * 1e3
        out[f"{name}_corank"] = n - ranks[0]
        out[f"{name}_update_ms"] = st["update_ms"]
        if name == "mfma" and st["update_ms"]:
            out["mfma_update_frac"] = 2.0 * L * L * st["macs"] / (st["update_ms"] / 1e3) / 1e12 / PEAK_INT8_TOPS
            if st["far_ms"]:
                out["mfma_far_update_frac"] = (2.0 * L * L * st["far_macs"] / (st["far_ms"] / 1e3) / 1e12
                                               / PEAK_INT8_TOPS)
    if plain:
        out["speedup_mfma_vs_plain"] = out["mfma_columns_per_s"] / out["plain_columns_per_s"]
    out["corank_ok"] = all(out[f"{name}_corank"] == 64 for name, _ in methods)
    buf.free()
    return out


def mixmax256_rank(ctx, n=65536):
    """The README's MIXMAX-256-detecting rank test (a couple of hours on the
    reference's CPU, README.md:110-113): an n x n matrix of consecutive
    MIXMAX-256 (skip 2) seed-1 outputs over 2^61-1, built on the device; the
    corank must be at least n - 65280 (every row is a window of a sequence of
    linear complexity 65280).  One timed run."""
    buf = W.mixmax256_matrix_device(ctx, n)
    field = M.Field(M61)
    ctx.synchronize()
    t0 = time.perf_counter()
    r = ctx.rank_batch(field, [buf], method=M.METHOD_MFMA)[0]
    ctx.synchronize()
    dt = time.perf_counter() - t0
    buf.free()
    return {"n": n, "bytes": 8 * n * n, "rank": r, "corank": n - r, "s": dt, "columns_per_s": n / dt,
            "corank_bound_ok": n - r >= n - 65280,
            "p_value": M.pretty_p_value(M.corank_tail_pvalue(M61, n, n - r))}


def c1_line(ctx, steps=20):
    """The README's 500 x 500 MIXMAX-17 rank (README.md:61-71: 55,700.74 columns/s
    on the reference's CPU, corank 432) on both GPU paths, with the oracle's port
    on the host beside it (one thread and all threads)."""
    mm, ctl = W.c1()
    field = M.Field(M61)
    dev = [ctx.to_device(mm)]
    out = {"n": 500, "readme_cpu_columns_per_s": 55700.74}
    for name, method in (("mfma", M.METHOD_MFMA), ("plain", M.METHOD_PLAIN)):
        ranks, t, _, _ = run(ctx, field, dev, method, steps, 2)
        out[f"{name}_columns_per_s"] = 500 / t
        out[f"{name}_corank"] = 500 - ranks[0]
    dev[0].free()
    O, lib = _oracle(), _native_oracle()
    cores = lib.or_num_threads()
    for label, th in (("cpu_port_1t", 1), ("cpu_port_all", cores)):
        lib.or_set_num_threads(th)
        t0 = time.perf_counter()
        reps = 0
        while time.perf_counter() - t0 < 1.0:
            O.rank(M61, mm, 500, lib=lib)
            reps += 1
        out[f"{label}_columns_per_s"] = 500 * reps / (time.perf_counter() - t0)
    lib.or_set_num_threads(cores)
    out["cpu_threads"] = cores
    out["cpu_build"] = "gcc -O3 -march=native -fopenmp (compiled on this host)"
    return out


def bm_batch(ctx, steps):
    """Batched Berlekamp-Massey: 1024 sequences, lengths uniform in [1000, 20000]."""
    seqs, _ = W.c4()
    field = M.Field(M61)
    dev = [ctx.to_device(s, square=False) for s in seqs]
    total = sum(s.size for s in seqs)
    ctx.lc_batch(field, dev)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        lcs = ctx.lc_batch(field, dev)
    ctx.synchronize()
    dt = (time.perf_counter() - t0) / steps
    ctx.lc_batch(field, dev, profile=True)
    st = ctx.stats()
    fx = fixture("c4_lc.json")
    out = {"sequences": len(seqs), "steps": total, "steps_per_s": total / dt, "ms": dt * 1e3,
           "kernel_ms": st["bm_ms"],
           "reference_algorithmic_GBps": st["bm_bytes"] / (st["bm_ms"] / 1e3) / 1e9 if st["bm_ms"] else None,
           "lc_match_fixture": (fx["lc"] == lcs) if fx else None}
    for d in dev:
        d.free()
    return out


def bm_long(ctx, steps):
    """Long Berlekamp-Massey: one 200,000-element MIXMAX-256 (skip 2) sequence, seed 1."""
    seq = W.c5_bm()
    field = M.Field(M61)
    dev = [ctx.to_device(seq, square=False)]
    t0 = time.perf_counter()
    for _ in range(steps):
        lc = ctx.lc_batch(field, dev)
    ctx.synchronize()
    dt = (time.perf_counter() - t0) / steps
    ctx.lc_batch(field, dev, profile=True)
    st = ctx.stats()
    fx = fixture("c5_bm.json")
    dev[0].free()
    return {"n": int(seq.size), "steps_per_s": seq.size / dt, "ms": dt * 1e3, "us_per_step": dt * 1e6 / seq.size,
            "kernel_ms": st["bm_ms"], "workgroups": st["bm_coop"], "lc": lc[0],
            "lc_match_fixture": (fx["lc"] == lc[0]) if fx else None}


def rank_batch_u32(ctx, steps):
    """256 matrices 2048 x 2048 over 65521, u32 storage (the small-prime batch),
    generated on the device."""
    buf, views = W.c5_rank_device(ctx)
    field = M.Field(65521)
    ranks, t, st, _ = run(ctx, field, views, M.METHOD_AUTO, steps, 1)
    fx = fixture("c5_rank_coranks.json")
    buf.free()
    L = st["limb_count"]
    return {"matrices": 256, "n": 2048, "p": 65521, "columns_per_s": 256 * 2048 / t, "ms": t * 1e3,
            "update_ms": st["update_ms"], "limbs": L,
            "update_frac": (2.0 * L * L * st["macs"] / (st["update_ms"] / 1e3) / 1e12 / PEAK_INT8_TOPS)
            if st["update_ms"] else None,
            "coranks_match_fixture": ([2048 - r for r in ranks] == fx["coranks"]) if fx else None}


def c5_mixed(steps=2):
    """The mixed configuration: the 200,000-step MIXMAX-256 BM and the 256 x 2048^2
    u32 rank batch over 65521 running at the same time, each on a context (stream)
    of its own, driven from two host threads; wall time of the pair, against the
    two run one after the other."""
    import threading
    cb, cr = M.Context(0), M.Context(0)
    seq = cb.to_device(W.c5_bm(), square=False)
    buf, views = W.c5_rank_device(cr)
    fb, fr = M.Field(M61), M.Field(65521)
    res = {}

    def bm():
        res["lc"] = cb.lc_batch(fb, [seq])[0]

    def rk():
        res["ranks"] = cr.rank_batch(fr, views)

    bm(); rk()  # warm-up (allocations, graph capture)
    cb.synchronize(); cr.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        bm(); rk()
    t_seq = (time.perf_counter() - t0) / steps
    t0 = time.perf_counter()
    for _ in range(steps):
        th = [threading.Thread(target=bm), threading.Thread(target=rk)]
        for x in th:
            x.start()
        for x in th:
            x.join()
    t_par = (time.perf_counter() - t0) / steps
    fx_r, fx_b = fixture("c5_rank_coranks.json"), fixture("c5_bm.json")
    out = {"ms_concurrent": t_par * 1e3, "ms_one_after_other": t_seq * 1e3,
           "bm_steps_per_s": 200_000 / t_par, "rank_columns_per_s": 256 * 2048 / t_par,
           "lc": res["lc"], "lc_match_fixture": (fx_b["lc"] == res["lc"]) if fx_b else None,
           "coranks_match_fixture": ([2048 - r for r in res["ranks"]] == fx_r["coranks"]) if fx_r else None}
    seq.free()
    buf.free()
    cb.close()
    cr.close()
    return out


def run(ctx, field, dev, method, steps, warmup):
    for _ in range(warmup):
        ctx.rank_batch(field, dev, method=method)
    ctx.synchronize()
    t0 = time.perf_counter()
    ranks = None
    for _ in range(steps):
        ranks = ctx.rank_batch(field, dev, method=method)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    prof_ranks = ctx.rank_batch(field, dev, method=method, profile=True)
    st = ctx.stats()
    return ranks, dt / steps, st, prof_ranks


EXTRA_LINES = ("c1_500_mixmax17", "single_8192_p61", "single_8192_p31", "single_16384_p61", "single_16384_p31",
               "bm_batch_1024", "bm_long_200k", "rank_batch_256x2048_u32_p65521", "c5_mixed_bm200k_rank256x2048")
FULL_LINES = ("single_32768_p61", "mixmax256_rank_65536")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--method", choices=["auto", "plain", "mfma"], default="auto")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-path comparison")
    ap.add_argument("--no-extras", action="store_true", help="skip the 8192^2 and Berlekamp-Massey lines")
    ap.add_argument("--copy-calibration", action="store_true", help="also run a 256 MiB copy kernel")
    ap.add_argument("--full", action="store_true",
                    help="also the 32768^2 planted matrix and the 65536^2 MIXMAX-256 rank")
    ap.add_argument("--cpu-full", action="store_true",
                    help="time the CPU baseline on the whole batch (minutes) instead of a leading sample")
    ap.add_argument("--only", default="",
                    help="comma-separated lines to run, for profiling one workload at a time: "
                         "c2 and/or names of the extra lines (the output then holds only those)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: measures one GPU only; --gpus above 1 is not supported "
              "(a batch would shard by matrix or sequence, one process per GPU)", file=sys.stderr)
        sys.exit(2)
    method = {"auto": M.METHOD_AUTO, "plain": M.METHOD_PLAIN, "mfma": M.METHOD_MFMA}[args.method]

    only = set(x for x in args.only.split(",") if x)
    unknown = only - set(EXTRA_LINES) - set(FULL_LINES) - {"c2"}
    if unknown:
        print(f"bench.py: unknown lines {sorted(unknown)}", file=sys.stderr)
        sys.exit(2)
    mats = W.c2()
    ctx = M.Context(0)
    out = {}
    if not only or "c2" in only:
        out = c2_line(ctx, mats, method, args)
    extras = {
        "c1_500_mixmax17": lambda: c1_line(ctx),
        "single_8192_p61": lambda: single_large(ctx, 8192, M61, 2),
        "single_8192_p31": lambda: single_large(ctx, 8192, (1 << 31) - 1, 2),
        "single_16384_p61": lambda: single_large(ctx, 16384, M61, 1),
        "single_16384_p31": lambda: single_large(ctx, 16384, (1 << 31) - 1, 1),
        "bm_batch_1024": lambda: bm_batch(ctx, 2),
        "bm_long_200k": lambda: bm_long(ctx, 1),
        "rank_batch_256x2048_u32_p65521": lambda: rank_batch_u32(ctx, 2),
        "c5_mixed_bm200k_rank256x2048": lambda: c5_mixed(),
    }
    full = {
        "single_32768_p61": lambda: single_large(ctx, 32768, M61, 1, plain=False, warmup=0),
        "mixmax256_rank_65536": lambda: mixmax256_rank(ctx),
    }
    assert tuple(extras) == EXTRA_LINES and tuple(full) == FULL_LINES
    if out:
        print(f"bench: c2 done, {out['ms_per_step']:.2f} ms/step", file=sys.stderr, flush=True)
    for name, fn in list(extras.items()) + list(full.items()):
        want = not args.no_extras if name in extras else args.full
        if (only and name in only) or (not only and want):
            t0 = time.perf_counter()
            out[name] = fn()
            print(f"bench: {name} done in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    if args.copy_calibration:
        out["copy_calibration"] = copy_calibration(ctx)
    # release the context (graphs, streams, events, pinned words, arena) while the
    # runtime and any attached profiler are fully up, not from interpreter teardown
    ctx.synchronize()
    ctx.close()
    if not args.no_cpu and (not only or "c2" in only):
        out["cpu_baseline"] = cpu_baseline(mats, full=args.cpu_full)
        if not args.no_extras:
            out["cpu_baseline_bm"] = cpu_bm(W.c4()[0], W.c5_bm())
    print(json.dumps(out))


def c2_line(ctx, mats, method, args):
    """The headline line: the 64-matrix batch over 2^61-1 from device memory."""
    sides = [m.shape[0] for m in mats]
    cols = sum(sides)
    field = M.Field(M61)
    dev = [ctx.to_device(m) for m in mats]

    ranks, t_step, st, _ = run(ctx, field, dev, method, args.steps, args.warmup)
    coranks = [n - r for n, r in zip(sides, ranks)]
    fx = fixture("c2_coranks.json")
    check = (fx["coranks"] == coranks) if fx else None

    L = limb_count(M61) if st["limb_count"] else 0
    upd_s = st["update_ms"] / 1e3
    near_far = None
    if L:
        # int8 operations per F_p multiply-accumulate: 2 per product of L x L digit pairs
        far_ops = 2.0 * L * L * st["far_macs"]
        achieved = far_ops / (st["far_ms"] / 1e3) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": PEAK_INT8_TOPS, "unit": "TOP/s",
                "frac": achieved / PEAK_INT8_TOPS, "traffic": traffic_per_launch(),
                "kernel": "k_update_far", "launches_per_step": st["far_launches"],
                "avg_us": st["far_ms"] * 1e3 / max(1, st["far_launches"])}
        all_ops = 2.0 * L * L * st["macs"]
        near_far = {"kernels": "k_update_near + k_update_far", "achieved": all_ops / upd_s / 1e12,
                    "unit": "TOP/s", "frac": all_ops / upd_s / 1e12 / PEAK_INT8_TOPS,
                    "traffic": traffic_per_launch("traffic_bytes_per_launch")}
    else:
        # VALU update: algorithmic HBM bytes = 16 per updated element (read + write)
        # + 8 per operand element of each tile (counted by the kernel)
        achieved = st["update_bytes"] / upd_s / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": achieved / PEAK_HBM_GBS, "traffic": None}

    out = {
        "metric": "Fp rank cols/s & BM steps/s: batched limb-MFMA vs plain GPU path, bit-exact",
        "value": cols / t_step,
        "unit": "columns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_step * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64 (F_p, p=2^61-1)" + (f"; update int8 x{L} limbs, int32 accumulate" if L else ""),
        "data": "synthetic (deterministic generators, see workloads.py)",
        "config": {"workload": "batched rank: 64 square matrices, sides log-uniform in [256, 4096] (seed 7), "
                               "p=2^61-1, half xoshiro256** uniform, half MIXMAX-17",
                   "matrices": len(mats), "sum_sides": cols, "max_side": max(sides), "method": args.method},
        "roofline": roof,
        "near_far_update": near_far,
        "update_ms_per_step": st["update_ms"],
        "update_launches_per_step": st["update_launches"],
        "update_avg_us": st["update_ms"] * 1e3 / max(1, st["update_launches"]),
        "update_macs_per_step": st["macs"],
        "coranks_match_fixture": check,
    }
    if not args.no_plain and method != M.METHOD_PLAIN:
        _, t_plain, st_p, _ = run(ctx, field, dev, M.METHOD_PLAIN, max(1, args.steps // 2), 0)
        out["plain_columns_per_s"] = cols / t_plain
        out["speedup_vs_plain"] = t_plain / t_step
    for d in dev:
        d.free()
    return out


if __name__ == "__main__":
    main()
