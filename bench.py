"""Benchmark: F_p rank columns/s of the batched rank on one MI355X.

Workload (one "step"): rank of the 64-matrix batch of the batched-rank
configuration -- square sides drawn log-uniform in [256, 4096] from
SplitMix64 seed 7, p = 2^61-1, u64 entries; even matrices uniform from
xoshiro256** seed 7, odd matrices consecutive MIXMAX-17 seed-7 stretches
(modlin-rs_amd/workloads.py:c2).  Inputs are resident in HBM before timing;
each step copies them into the library's workspace (the reference destroys its
input too) and runs the whole batch through one launch sequence.

Prints one JSON line.  value = total matrix columns ranked per second
(sum of sides / step time).  `roofline` refers to the trailing-update kernel,
timed with HIP events on the library's stream; `traffic` is its HBM bytes per
launch from the committed rocprofv3 counter summary; `cpu_baseline` is the
oracle's OpenMP port of fp.rs on a bounded sample of the same batch on this
host.  Extra keys report the plain (VALU) path on the same batch, the single
8192 x 8192 planted-corank matrix over 2^61-1 on both paths, the batched
Berlekamp-Massey workload (1024 sequences, steps/s) and the per-item checks
against the CPU-computed fixtures in tests/golden/.

--copy-calibration additionally ranks one all-zero 4096 x 4096 u32 matrix held
in device memory: its staging copy (64 MiB read, 64 MiB written) and widening
kernel k_widen_u32 (64 MiB read, 128 MiB written) are copy kernels of known
size that a counter run uses to scale FETCH_SIZE / WRITE_SIZE to bytes.
"""
import argparse
import json
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


def _rank_sample(O, mats, budget_n3):
    sample, acc = [], 0.0
    for m in mats:
        n = m.shape[0]
        if sample and acc + float(n) ** 3 > budget_n3:
            break
        sample.append(m)
        acc += float(n) ** 3
    t0 = time.perf_counter()
    for m in sample:
        O.rank(M61, m, m.shape[0])
    dt = time.perf_counter() - t0
    cols = sum(m.shape[0] for m in sample)
    return cols / dt, f"first {len(sample)} matrices of the batch (sum n = {cols}, sum n^3 = {acc:.3g}), {dt:.1f} s"


def cpu_baseline(mats, budget_n3=4.0e10, budget_n3_1t=1.0e10):
    """The oracle's port of fp::rank (OpenMP over rows, as rayon does) on the
    first matrices of the batch whose sum of n^3 stays within the budget, on all
    host threads; and a smaller sample on one thread and on all threads (columns/s
    depend on the sides, so the two thread counts are compared on one sample)."""
    O = _oracle()
    cores = O.num_threads()
    v, sample = _rank_sample(O, mats, budget_n3)
    va, _ = _rank_sample(O, mats, budget_n3_1t)
    O.set_num_threads(1)
    try:
        v1, sample1 = _rank_sample(O, mats, budget_n3_1t)
    finally:
        O.set_num_threads(cores)
    return {"value": v, "unit": "columns/s", "cores": cores, "kind": "port", "sample": sample,
            "single_thread": {"value": v1, "unit": "columns/s", "cores": 1, "sample": sample1,
                              "all_threads_same_sample": va}}


def cpu_bm(seqs, long_seq, budget_s=4.0, long_prefix=20000):
    """The oracle's port of fp::linear_complexity (one thread, as the reference) on
    the batch's sequences in order until the time budget is spent, and on the
    first long_prefix steps of the long sequence (the same first steps as the
    full run, with l still small, so a lower bound of its time per step)."""
    O = _oracle()
    t0, steps, k = time.perf_counter(), 0, 0
    while k < len(seqs) and time.perf_counter() - t0 < budget_s:
        O.lc(M61, seqs[k])
        steps += seqs[k].size
        k += 1
    dt = time.perf_counter() - t0
    t1 = time.perf_counter()
    O.lc(M61, long_seq[:long_prefix])
    dl = time.perf_counter() - t1
    return {"bm_batch": {"value": steps / dt, "unit": "steps/s", "cores": 1, "kind": "port",
                         "sample": f"first {k} sequences of the batch ({steps} steps), {dt:.1f} s"},
            "bm_long": {"value": dl * 1e6 / long_prefix, "unit": "us/step", "cores": 1, "kind": "port",
                        "sample": f"first {long_prefix} of the 200000 steps, {dl:.2f} s"}}


def traffic_per_launch():
    """HBM bytes per launch of the MFMA trailing update on this workload, from
    the committed counter summary (FETCH_SIZE and WRITE_SIZE runs, scaled by the
    same counters on a copy kernel of known size), or None."""
    path = os.path.join(ROOT, "profiles", "update_kernels_batched_rank_64_p61_counters.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def copy_calibration(ctx, n=4096):
    dev = ctx.to_device(np.zeros((n, n), dtype=np.uint32))
    r = ctx.rank_batch(M.Field(M61), [dev])
    ctx.synchronize()
    dev.free()
    return {"copy_kernel": "k_widen_u32", "elements": n * n, "bytes_read": 4 * n * n, "bytes_written": 8 * n * n,
            "rank": r[0]}


def fixture(name):
    path = os.path.join(ROOT, "tests", "golden", name)
    return json.load(open(path)) if os.path.exists(path) else None


def single_large(ctx, n, p, steps):
    """One n x n planted-corank matrix: MFMA and plain paths, columns/s."""
    a = W.c3(n, p)
    field = M.Field(p)
    dev = [ctx.to_device(a)]
    out = {"n": n, "p": p}
    for name, method in (("mfma", M.METHOD_MFMA), ("plain", M.METHOD_PLAIN)):
        ranks, t, st, _ = run(ctx, field, dev, method, steps, 1)
        out[f"{name}_columns_per_s"] = n / t
        out[f"{name}_ms"] = t * 1e3
        out[f"{name}_corank"] = n - ranks[0]
        out[f"{name}_update_ms"] = st["update_ms"]
    out["speedup_mfma_vs_plain"] = out["mfma_columns_per_s"] / out["plain_columns_per_s"]
    dev[0].free()
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
           "kernel_ms": st["bm_ms"], "hbm_GBps": st["bm_bytes"] / (st["bm_ms"] / 1e3) / 1e9 if st["bm_ms"] else None,
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
    """256 matrices 2048 x 2048 over 65521, u32 storage (the small-prime batch)."""
    mats = W.c5_rank(count=256)
    field = M.Field(65521)
    dev = [ctx.to_device(m) for m in mats]
    ranks, t, st, _ = run(ctx, field, dev, M.METHOD_AUTO, steps, 1)
    fx = fixture("c5_rank_coranks.json")
    for d in dev:
        d.free()
    return {"matrices": 256, "n": 2048, "p": 65521, "columns_per_s": 256 * 2048 / t, "ms": t * 1e3,
            "update_ms": st["update_ms"], "limbs": st["limb_count"],
            "coranks_match_fixture": ([2048 - r for r in ranks] == fx["coranks"]) if fx else None}


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
                    help="also the 8192^2 matrix over 2^31-1, the 16384^2 matrices and the 256 x 2048^2 u32 batch")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: measures one GPU only; --gpus above 1 is not supported "
              "(a batch would shard by matrix or sequence, one process per GPU)", file=sys.stderr)
        sys.exit(2)
    method = {"auto": M.METHOD_AUTO, "plain": M.METHOD_PLAIN, "mfma": M.METHOD_MFMA}[args.method]

    mats = W.c2()
    sides = [m.shape[0] for m in mats]
    cols = sum(sides)
    field = M.Field(M61)
    ctx = M.Context(0)
    dev = [ctx.to_device(m) for m in mats]

    ranks, t_step, st, _ = run(ctx, field, dev, method, args.steps, args.warmup)
    coranks = [n - r for n, r in zip(sides, ranks)]
    fx = fixture("c2_coranks.json")
    check = (fx["coranks"] == coranks) if fx else None

    L = limb_count(M61) if st["limb_count"] else 0
    upd_s = st["update_ms"] / 1e3
    if L:
        ops = 2.0 * L * L * st["macs"]
        achieved = ops / upd_s / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": PEAK_INT8_TOPS, "unit": "TOP/s",
                "frac": achieved / PEAK_INT8_TOPS, "traffic": traffic_per_launch()}
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
    if not args.no_extras:
        out["single_8192_p61"] = single_large(ctx, 8192, M61, 2)
        out["bm_batch_1024"] = bm_batch(ctx, 2)
        out["bm_long_200k"] = bm_long(ctx, 1)
    if args.full:
        out["single_8192_p31"] = single_large(ctx, 8192, (1 << 31) - 1, 2)
        out["single_16384_p61"] = single_large(ctx, 16384, M61, 1)
        out["single_16384_p31"] = single_large(ctx, 16384, (1 << 31) - 1, 1)
        out["rank_batch_256x2048_u32_p65521"] = rank_batch_u32(ctx, 2)
    if args.copy_calibration:
        out["copy_calibration"] = copy_calibration(ctx)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(mats)
        if not args.no_extras:
            out["cpu_baseline_bm"] = cpu_bm(W.c4()[0], W.c5_bm())
    print(json.dumps(out))


if __name__ == "__main__":
    main()
