"""Autotuner of the library's launch parameters (runs on a gfx950 GPU).

For each workload class -- a field mode (m61, m31, gL), a size bucket and a batch
size -- it times candidate launch parameters on a representative synthetic input
of that class, keeps the fastest, and writes the winners as a tuning table (the
format of modlin_ctx_load_tuning in include/modlin.h).  A context loads
tuning.txt from the library's directory when it is created.

Search: coordinate descent from the built-in defaults, two passes over the
parameters (far-update workgroups per CU, near-update workgroups per CU, gather
grid, panel iterations per host check, panel width, panels per far update (K
depth 64 or 128), and for p < 2^32 the far update's limb-class grouping
(quarters per B fragment); threads per sequence of the batched
Berlekamp-Massey, workgroups of the cooperative one).  Each candidate is timed as the best of `reps`
runs after one warm-up run.  Launch parameters never change a result: every
candidate's ranks (or linear complexity) must equal those of the untuned run,
or the tuner stops with an error.

    python modlin-rs_amd/autotune.py [--out PATH] [--quick]
"""
import argparse
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import modlin as M  # noqa: E402
import workloads as W  # noqa: E402

M61 = (1 << 61) - 1
P31 = (1 << 31) - 1

RANK_SPACE = {
    "far_per_cu": [2, 4, 6, 8],
    "near_per_cu": [1, 2, 4],
    "gather_wgs": [512, 1024, 2048],
    "check_every": [4, 8, 16],
    "panel": [48, 64],
    "far_depth": [1, 2],      # K per far pass: one panel (<= 64) or two (<= 128)
}
# fields (p < 2^32, 4 digits or fewer): quarters of the far update's column step
# computed per B fragment read -- how many limb-class accumulator sets are live
RANK_SPACE_SMALL = dict(RANK_SPACE, far_quarters=[1, 2])
BM_SPACE = {"bm_coop_wgs": [128, 256, 512]}
BM_BATCH_SPACE = {"bm_threads": [256, 512, 1024]}


def side_bucket(n: int) -> int:
    s = 256
    while s < n:
        s <<= 1
    return s


def mode_of(p: int) -> str:
    if p == M61:
        return "m61"
    if p == P31:
        return "m31"
    L, half = 1, (p - 1) // 2
    while 127 * (256 ** L - 1) // 255 < half:
        L += 1
    return f"g{L}"


class Job:
    """A workload class: a callable that runs it once and returns its results."""

    def __init__(self, ctx, kind, p, items, label, method=M.METHOD_AUTO):
        self.ctx, self.kind, self.p, self.items, self.label, self.method = ctx, kind, p, items, label, method
        self.field = M.Field(p)
        self.max_n = max(it.n for it in items)
        self.count = len(items)

    def run(self):
        if self.kind == "rank":
            return self.ctx.rank_batch(self.field, self.items)
        return self.ctx.lc_batch(self.field, self.items, method=self.method)

    def time(self, reps):
        self.run()
        self.ctx.synchronize()
        best, out = float("inf"), None
        for _ in range(reps):
            t0 = time.perf_counter()
            out = self.run()
            self.ctx.synchronize()
            best = min(best, time.perf_counter() - t0)
        return best, out

    def key(self):
        return f"{self.kind} {mode_of(self.p)} {side_bucket(self.max_n)} {'1' if self.count == 1 else 'many'}"


def tune(job, space, reps, log):
    ctx = job.ctx
    ctx.set_tuning()
    t_default, want = job.time(reps)
    current = {}
    best_t = t_default
    for _ in range(2):
        for name, values in space.items():
            for v in values:
                cand = dict(current, **{name: v})
                ctx.set_tuning(**cand)
                t, got = job.time(reps)
                if got != want:
                    raise RuntimeError(f"{job.label}: results changed under {cand}: {got} != {want}")
                if t < best_t * 0.995:
                    best_t, current = t, cand
    ctx.set_tuning()
    log(f"{job.label:40s} default {t_default * 1e3:9.2f} ms  tuned {best_t * 1e3:9.2f} ms  {current}")
    return current, t_default, best_t


def entry_line(job, params):
    fields = [params.get(k, 0) for k in M.TUNING_FIELDS]
    return job.key() + " " + " ".join(str(v) for v in fields)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "tuning.txt"))
    ap.add_argument("--quick", action="store_true", help="skip the 16384^2 and long-BM classes")
    ap.add_argument("--reps", type=int, default=2)
    args = ap.parse_args()

    ctx = M.Context(0)
    ctx.load_tuning("")  # tune from the built-in defaults
    log = lambda s: print(s, flush=True)  # noqa: E731
    jobs = []
    mats = W.c2()
    jobs.append((Job(ctx, "rank", M61, [ctx.to_device(m) for m in mats], "64-matrix batch, sides <= 4096, 2^61-1"),
                 RANK_SPACE))
    mm, _ = W.c1()
    jobs.append((Job(ctx, "rank", M61, [ctx.to_device(mm)], "500^2 MIXMAX-17 (rank 68), 2^61-1"), RANK_SPACE))
    for n, p in ((8192, M61), (8192, P31)) + (() if args.quick else ((16384, M61), (16384, P31))):
        jobs.append((Job(ctx, "rank", p, [ctx.to_device(W.c3(n, p))], f"{n}^2 planted corank, {mode_of(p)}"),
                     RANK_SPACE if p == M61 else RANK_SPACE_SMALL))
    small = W.c5_rank(count=256)
    jobs.append((Job(ctx, "rank", 65521, [ctx.to_device(m) for m in small], "256 x 2048^2 u32, p = 65521"),
                 RANK_SPACE_SMALL))
    seqs, _ = W.c4()
    # bm_threads sizes the one-step-per-pass kernel (method PLAIN); the default,
    # blocked form has no launch parameter to tune
    jobs.append((Job(ctx, "bm", M61, [ctx.to_device(s, square=False) for s in seqs],
                     "1024 BM sequences, n in [1000, 20000], one step per pass", M.METHOD_PLAIN), BM_BATCH_SPACE))
    if not args.quick:
        jobs.append((Job(ctx, "bm", M61, [ctx.to_device(W.c5_bm(), square=False)], "200000-step BM, MIXMAX-256"),
                     BM_SPACE))

    lines = ["# Launch parameters chosen by modlin-rs_amd/autotune.py on one MI355X.",
             "# kind mode side count far_per_cu near_per_cu gather_wgs check_every panel bm_coop_wgs bm_threads"
             " far_depth far_quarters"]
    for job, space in jobs:
        params, t0, t1 = tune(job, space, args.reps, log)
        lines.append(f"{entry_line(job, params)}   # {job.label}: {t0 * 1e3:.2f} -> {t1 * 1e3:.2f} ms")
        for it in job.items:
            it.free()
    with open(args.out, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    log(f"wrote {args.out}")


if __name__ == "__main__":
    main()
