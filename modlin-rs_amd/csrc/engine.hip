// Host orchestration of the rank engine: workspace layout in HBM, the panel
// iteration loop and the context.  The reference's driver loop is
// src/fp.rs:138-252 (one matrix); here every matrix of a batch advances one
// panel per iteration.  An iteration is: block elimination and inverse, plan,
// gather, near update, finish (main stream), and the far columns' X product and
// far update (far stream) -- the far update overlaps the next iteration's block
// elimination and plan: the next panel's columns lie in the near part, already
// updated.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>

#include "engine.hpp"
#include "rank_common.hpp"

namespace mdl {

// rank_kernels.hip, mfma_update.hip
void launch_block(hipStream_t st, RankMat *dmats, uint32_t nmat, const FieldParams &f, uint32_t b, uint32_t par,
                  uint32_t L, uint32_t defer, const RankLaunch *rlp, bool lds);
void launch_plan(hipStream_t st, RankMat *dmats, uint32_t nmat, const RankLaunch &rl, uint32_t par);
void launch_finish(hipStream_t st, RankMat *dmats, uint32_t nmat, uint32_t *active, uint32_t par);
void launch_collect(hipStream_t st, RankMat *dmats, uint32_t nmat, const uint32_t *active, uint64_t *out);
void launch_gather_plain(hipStream_t st, RankMat *dmats, uint32_t nmat, const FieldParams &f, const RankLaunch &rl,
                         uint32_t par);
void launch_update_plain(hipStream_t st, RankMat *dmats, uint32_t nmat, const FieldParams &f, const RankLaunch &rl,
                         uint32_t par, uint32_t far);
void launch_mfma(hipStream_t st, RankMat *dmats, uint32_t nmat, const FieldParams &f, const LimbParams &lp,
                 const RankLaunch &rl, uint32_t par, int what);

int hip_fail(modlin_ctx *ctx, hipError_t e, const char *what) {
    ctx->err = std::string(what) + ": " + hipGetErrorString(e);
    return MODLIN_EHIP;
}

int arena_reserve(modlin_ctx *ctx, size_t bytes) {
    if (bytes <= ctx->arena_bytes) return MODLIN_OK;
    if (ctx->arena) {
        hipStreamSynchronize(ctx->stream);
        hipFree(ctx->arena);
        ctx->arena = nullptr;
        ctx->arena_bytes = 0;
    }
    hipError_t e = hipMalloc(&ctx->arena, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        ctx->err = "device allocation of " + std::to_string(bytes) + " bytes failed";
        return MODLIN_ENOMEM;
    }
    ctx->arena_bytes = bytes;
    return MODLIN_OK;
}

static constexpr size_t PINNED_CHUNK = 64ull << 20;

static int ensure_pinned(modlin_ctx *ctx) {
    if (ctx->pinned[0]) return MODLIN_OK;
    for (int i = 0; i < 2; i++) {
        if (hipHostMalloc(&ctx->pinned[i], PINNED_CHUNK, hipHostMallocDefault) != hipSuccess ||
            hipEventCreateWithFlags(&ctx->pinned_ev[i], hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            ctx->err = "pinned staging allocation failed";
            return MODLIN_ENOMEM;
        }
    }
    ctx->pinned_bytes = PINNED_CHUNK;
    return MODLIN_OK;
}

void release_pinned(modlin_ctx *ctx) {
    for (int i = 0; i < 2; i++) {
        if (ctx->pinned_ev[i]) hipEventDestroy(ctx->pinned_ev[i]);
        if (ctx->pinned[i]) hipHostFree(ctx->pinned[i]);
        ctx->pinned_ev[i] = nullptr;
        ctx->pinned[i] = nullptr;
    }
    ctx->pinned_bytes = 0;
}

int copy_h2d(modlin_ctx *ctx, void *dst, const void *src, size_t bytes) {
    int rc = ensure_pinned(ctx);
    if (rc) return rc;
    size_t off = 0;
    int buf = 0;
    bool used[2] = {false, false};
    while (off < bytes) {
        const size_t chunk = std::min(bytes - off, ctx->pinned_bytes);
        if (used[buf]) MDL_HIP(hipEventSynchronize(ctx->pinned_ev[buf]));  // staging slot free again
        memcpy(ctx->pinned[buf], (const char *)src + off, chunk);
        MDL_HIP(hipMemcpyAsync((char *)dst + off, ctx->pinned[buf], chunk, hipMemcpyHostToDevice, ctx->stream));
        MDL_HIP(hipEventRecord(ctx->pinned_ev[buf], ctx->stream));
        used[buf] = true;
        off += chunk;
        buf ^= 1;
    }
    // both slots must be free before the caller's next staging use
    for (int i = 0; i < 2; i++)
        if (used[i]) MDL_HIP(hipEventSynchronize(ctx->pinned_ev[i]));
    return MODLIN_OK;
}

int copy_d2h(modlin_ctx *ctx, void *dst, const void *src, size_t bytes) {
    int rc = ensure_pinned(ctx);
    if (rc) return rc;
    size_t off = 0;
    while (off < bytes) {
        const size_t chunk = std::min(bytes - off, ctx->pinned_bytes);
        MDL_HIP(hipMemcpyAsync(ctx->pinned[0], (const char *)src + off, chunk, hipMemcpyDeviceToHost, ctx->stream));
        MDL_HIP(hipStreamSynchronize(ctx->stream));
        memcpy((char *)dst + off, ctx->pinned[0], chunk);
        off += chunk;
    }
    return MODLIN_OK;
}

// Set-up of one matrix of a batch: its working copy (from the caller's device
// matrix or the staged u32 copy; none when the host copy already landed in M)
// and the identity row order.
struct SetupJob {
    const void *src;
    uint64_t *M;
    uint32_t *perm;
    uint64_t elems;  // n * n
    uint32_t n;
    uint32_t form;   // 0: u64 in 16-byte pieces, 1: u64 one by one, 2: u32 widened to u64,
                     // 3: u32 kept (u32 working copy), 4: u64 narrowed to u32 (entries below 2^32)
};

// All matrices of a batch in one launch: workgroup (x, y) covers part x of job y.
__global__ void k_setup(const SetupJob *jobs) {
    const SetupJob J = jobs[blockIdx.y];
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t t0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (J.src) {
        if (J.form == 3) {
            const uint32_t *src = (const uint32_t *)J.src;
            uint32_t *dst = (uint32_t *)J.M;
            for (uint64_t i = t0; i < J.elems; i += stride) dst[i] = src[i];
        } else if (J.form == 4) {
            const uint64_t *src = (const uint64_t *)J.src;
            uint32_t *dst = (uint32_t *)J.M;
            for (uint64_t i = t0; i < J.elems; i += stride) dst[i] = (uint32_t)src[i];
        } else if (J.form == 2) {
            const uint32_t *src = (const uint32_t *)J.src;
            for (uint64_t i = t0; i < J.elems; i += stride) J.M[i] = src[i];
        } else if (J.form == 1) {
            const uint64_t *src = (const uint64_t *)J.src;
            for (uint64_t i = t0; i < J.elems; i += stride) J.M[i] = src[i];
        } else {
            const ulonglong2 *src = (const ulonglong2 *)J.src;
            ulonglong2 *dst = (ulonglong2 *)J.M;
            for (uint64_t i = t0; i < J.elems / 2; i += stride) dst[i] = src[i];
            if (t0 == 0 && (J.elems & 1)) J.M[J.elems - 1] = ((const uint64_t *)J.src)[J.elems - 1];
        }
    }
    for (uint64_t i = t0; i < J.n; i += stride) J.perm[i] = (uint32_t)i;
}

static size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

int rank_batch(modlin_ctx *ctx, uint64_t p, const modlin_mat_desc_t *mats, uint64_t count, const modlin_opts_t *opts,
               uint64_t *ranks, modlin_progress_fn progress, void *user) {
    const auto t_start = std::chrono::steady_clock::now();
    ctx->stats = modlin_stats_t{};
    if (count == 0) return MODLIN_OK;
    const FieldParams f = make_params(p);
    uint64_t maxn = 0;
    for (uint64_t i = 0; i < count; i++) maxn = std::max<uint64_t>(maxn, mats[i].n);
    const modlin_tuning_t tu = select_tuning(ctx->tune_table, &ctx->tune_override, p, maxn, count, false);
    uint32_t b = (opts && opts->panel) ? opts->panel : (tu.panel ? tu.panel : (uint32_t)BMAX);
    if (b > (uint32_t)BMAX) b = BMAX;
    const bool profile = opts && (opts->flags & MODLIN_FLAG_PROFILE);
    const uint32_t method = opts ? opts->method : (uint32_t)MODLIN_METHOD_AUTO;
    const bool use_mfma = method != MODLIN_METHOD_PLAIN;
    // far updates of two panels at once (MFMA method), unless the caller asks for
    // one per panel
    const bool defer = use_mfma && !(opts && (opts->flags & MODLIN_FLAG_NO_DEFER)) && tu.far_depth != 1;
    // entries below 2^32 are kept as u32 in the working copy (MFMA method): the
    // trailing updates, which read and write it, move half the bytes
    const bool m32 = use_mfma && p <= 0xffffffffull && !(opts && (opts->flags & MODLIN_FLAG_U64_STORE));
    const LimbParams lp = make_limb_params(p);
    const size_t KA = (size_t)lp.L * KSTEP;

    // Workspace layout (256-byte aligned regions).
    std::vector<size_t> offM(count), offP(count), offA(count), offX(count), offZ(count), offE(count), offS(count),
        offAl(count), offXl(count), offZl(count), offAlP(count), offXlP(count), offR(count);
    size_t cur = 0;
    const size_t A = 256;
    // the matrices' state, their set-up jobs and the active count, copied to the
    // device together
    cur = count * sizeof(RankMat);
    const size_t off_jobs = cur;
    cur += count * sizeof(SetupJob);
    const size_t off_active = cur;
    cur = align_up(cur + sizeof(uint32_t), A);
    // the work plan, one per iteration parity (a far update reads its iteration's
    // while the next k_plan writes the other)
    size_t off_gx[2], off_upd[2], off_far[2], off_xf[2], off_plan[2];
    for (int q = 0; q < 2; q++) {
        off_gx[q] = cur;
        cur = align_up(cur + (count + 1) * sizeof(uint32_t), A);
        off_upd[q] = cur;
        cur = align_up(cur + (count + 1) * sizeof(uint32_t), A);
        off_far[q] = cur;
        cur = align_up(cur + (count + 1) * sizeof(uint32_t), A);
        off_xf[q] = cur;
        cur = align_up(cur + (count + 1) * sizeof(uint32_t), A);
        off_plan[q] = cur;
        cur = align_up(cur + sizeof(RankPlan), A);
    }
    const size_t off_trash = cur;
    cur = align_up(cur + TRASH_WORDS * sizeof(uint64_t), A);
    const size_t off_out = cur;  // k_collect's output: active, then rank, macs, bytes, far macs per matrix
    cur = align_up(cur + (1 + 4 * count) * sizeof(uint64_t), A);
    for (uint64_t i = 0; i < count; i++) {
        const uint64_t n = mats[i].n;
        if (n > 0xffffffffull / 2 || (n && n > (((uint64_t)1) << 62) / n / 8)) {
            ctx->err = "matrix side too large";
            return MODLIN_EINVAL;
        }
        if (mats[i].elem != MODLIN_ELEM_U64 && mats[i].elem != MODLIN_ELEM_U32) {
            ctx->err = "unknown element type";
            return MODLIN_EINVAL;
        }
        if (n && !mats[i].data) {
            ctx->err = "null matrix pointer";
            return MODLIN_EINVAL;
        }
        offM[i] = cur; cur = align_up(cur + n * n * (m32 ? 4 : 8), A);
        offP[i] = cur; cur = align_up(cur + n * 4, A);
        offA[i] = cur; cur = align_up(cur + n * BMAX * 8, A);
        offX[i] = cur; cur = align_up(cur + BMAX * n * 8, A);
        offZ[i] = cur; cur = align_up(cur + BMAX * BMAX * 8, A);
        offE[i] = cur; cur = align_up(cur + ETA_WORDS * 8, A);
        // stage for host entries whose width differs from the working copy's
        offS[i] = 0;
        const bool in32 = mats[i].elem == MODLIN_ELEM_U32;
        if (in32 != m32 && mats[i].mem != MODLIN_MEM_DEVICE) { offS[i] = cur; cur = align_up(cur + n * n * (in32 ? 4 : 8), A); }
        offAl[i] = offXl[i] = offZl[i] = offAlP[i] = offXlP[i] = offR[i] = 0;
        if (use_mfma) {
            // digit columns are read in whole 64-column tiles: each X buffer is
            // followed by 64 columns' worth of slack
            const size_t npad = (n + UT - 1) / UT * UT;
            offR[i] = cur; cur = align_up(cur + n * 4, A);
            offAl[i] = cur; cur = align_up(cur + npad * KA, A);
            offXl[i] = cur; cur = align_up(cur + npad * KA, A);
            offZl[i] = cur; cur = align_up(cur + 2 * BMAX * KA, A);  // one per iteration parity
            if (defer) {
                offAlP[i] = cur; cur = align_up(cur + npad * KA, A);
                offXlP[i] = cur; cur = align_up(cur + (npad + UT) * KA, A);
            }
        }
    }
    int rc = arena_reserve(ctx, cur);
    if (rc) return rc;
    char *base = (char *)ctx->arena;
    hipStream_t st = ctx->stream;

    // host image of [RankMat x count | SetupJob x count | active]
    std::vector<char> himg(off_active + sizeof(uint32_t));
    RankMat *hm = (RankMat *)himg.data();
    SetupJob *jobs = (SetupJob *)(himg.data() + off_jobs);
    uint32_t active = 0;
    for (uint64_t i = 0; i < count; i++) {
        const uint64_t n = mats[i].n;
        RankMat &S = hm[i];
        memset(&S, 0, sizeof(S));
        S.M = (uint64_t *)(base + offM[i]);
        S.perm = (uint32_t *)(base + offP[i]);
        S.Ab = (uint64_t *)(base + offA[i]);
        S.X = (uint64_t *)(base + offX[i]);
        S.Z = (uint64_t *)(base + offZ[i]);
        S.Eta = (uint64_t *)(base + offE[i]);
        S.Al = use_mfma ? (int8_t *)(base + offAl[i]) : nullptr;
        S.Xl = use_mfma ? (int8_t *)(base + offXl[i]) : nullptr;
        S.Zl = use_mfma ? (int8_t *)(base + offZl[i]) : nullptr;
        S.AlP = defer ? (int8_t *)(base + offAlP[i]) : nullptr;
        S.XlP = defer ? (int8_t *)(base + offXlP[i]) : nullptr;
        S.rows = use_mfma ? (uint32_t *)(base + offR[i]) : nullptr;
        S.n = (uint32_t)n;
        S.ld = (uint32_t)n;
        S.m32 = m32 ? 1u : 0u;
        S.done = (n == 0) ? 1u : 0u;
        S.pend = 0xffffffffu;
        if (n == 0) continue;
        active++;
        const size_t elems = n * n;
        const bool dev = mats[i].mem == MODLIN_MEM_DEVICE;
        const bool u32 = mats[i].elem == MODLIN_ELEM_U32;
        SetupJob &J = jobs[i];
        J.M = S.M;
        J.perm = S.perm;
        J.elems = elems;
        J.n = (uint32_t)n;
        J.form = m32 ? (u32 ? 3u : 4u) : (u32 ? 2u : (((uintptr_t)mats[i].data & 15) ? 1u : 0u));
        J.src = mats[i].data;
        if (!dev) {
            // host entries: straight into M when the widths agree, else into the stage
            const bool same = u32 == m32;
            void *dst = same ? (void *)S.M : (void *)(base + offS[i]);
            rc = copy_h2d(ctx, dst, mats[i].data, elems * (u32 ? 4 : 8));
            if (rc) return rc;
            J.src = same ? nullptr : dst;
        }
    }
    RankMat *dmats = (RankMat *)(base + 0);
    RankLaunch rl;
    rl.trash = (uint64_t *)(base + off_trash);
    uint32_t *dactive = (uint32_t *)(base + off_active);
    *(uint32_t *)(himg.data() + off_active) = active;
    rc = copy_h2d(ctx, dmats, himg.data(), himg.size());
    if (rc) return rc;
    {
        const uint32_t gx = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(1024, 4096 / count));
        hipLaunchKernelGGL(k_setup, dim3(gx, (uint32_t)count), dim3(256), 0, st, (const SetupJob *)(base + off_jobs));
    }
    MDL_HIP(hipGetLastError());

    const uint32_t nmat = (uint32_t)count;
    rl.grid_gx = tu.gather_wgs ? tu.gather_wgs : 1024;
    rl.grid_near = (uint32_t)ctx->cu_count * (tu.near_per_cu ? tu.near_per_cu : 2);
    // far grid: a multiple of 8 (XCD numbering)
    rl.grid_far = ((uint32_t)ctx->cu_count * (tu.far_per_cu ? tu.far_per_cu : (use_mfma ? 4 : 8)) + 7) / 8 * 8;
    rl.far_items = use_mfma ? 1u : 0u;
    rl.far_qw = tu.far_quarters ? tu.far_quarters : 2u;
    RankLaunch rlp[2] = {rl, rl};  // by iteration parity: the plan buffers
    for (int q = 0; q < 2; q++) {
        rlp[q].gx_off = (uint32_t *)(base + off_gx[q]);
        rlp[q].near_off = (uint32_t *)(base + off_upd[q]);
        rlp[q].far_off = (uint32_t *)(base + off_far[q]);
        rlp[q].xf_off = (uint32_t *)(base + off_xf[q]);
        rlp[q].plan = (RankPlan *)(base + off_plan[q]);
    }
    const int check_every = tu.check_every ? (int)tu.check_every : 8;
    std::vector<hipEvent_t> evs;
    if (profile) {
        evs.resize(4 * check_every);
        for (auto &e : evs) MDL_HIP(hipEventCreate(&e));
    }
    double upd_ms = 0, upd_launches = 0, far_ms = 0, far_launches = 0;
    uint32_t h_active = active;
    uint64_t iters = 0;
    // One block of check_every iterations.  The panel chain runs on the main
    // stream: block elimination and inverse, plan, gather, near update, finish,
    // then the next block elimination.  The far part runs on the far stream: the
    // far columns' X product (MFMA method) from the plan on, beside the gather and
    // near update; the far update once the near update is done (its digit rows
    // are written, and the two do not share the GPU), overlapping the finish and
    // the next block elimination and plan.  The next gather (which rewrites the
    // digit rows the far update reads, and reads pivot rows in columns it
    // writes) waits for it.  MFMA method: the far update reaches its rows through
    // the gather's row snapshot and its work queue through its iteration's plan
    // buffer, so the next k_plan (which reorders perm and writes the other plan
    // buffer) does not wait; the plain method's far update reads perm, so there
    // the next plan waits.  (check_every is even, so iteration parities repeat
    // per block.)
    hipStream_t fs = ctx->far_stream;
    auto enqueue_block = [&](bool prof) -> hipError_t {
        for (int s = 0; s < check_every; s++) {
            const uint32_t par = (uint32_t)(s & 1);
            const RankLaunch &rq = rlp[par];
            hipError_t e;
            // MFMA method: the plan runs as one more workgroup of k_block_z
            launch_block(st, dmats, nmat, f, b, par, use_mfma ? lp.L : 0u, defer ? 1u : 0u, use_mfma ? &rq : nullptr,
                         maxn >= GE_LDS_MIN_SIDE);
            if (!use_mfma) {
                if ((e = hipStreamWaitEvent(st, ctx->ev_far, 0)) != hipSuccess) return e;
                launch_plan(st, dmats, nmat, rq, par);
            }
            if ((e = hipEventRecord(ctx->ev_plan, st)) != hipSuccess) return e;
            if ((e = hipStreamWaitEvent(fs, ctx->ev_plan, 0)) != hipSuccess) return e;
            // far X: reads the plan, Z's digits (this parity) and the pivot rows, none
            // of which the gather, near update or finish writes
            if (use_mfma) launch_mfma(fs, dmats, nmat, f, lp, rq, par, -2);
            if (use_mfma && (e = hipStreamWaitEvent(st, ctx->ev_far, 0)) != hipSuccess) return e;
            if (use_mfma)
                launch_mfma(st, dmats, nmat, f, lp, rq, par, -1);
            else
                launch_gather_plain(st, dmats, nmat, f, rq, par);
            if (prof) hipEventRecord(evs[4 * s], st);
            if (use_mfma)
                launch_mfma(st, dmats, nmat, f, lp, rq, par, 0);
            else
                launch_update_plain(st, dmats, nmat, f, rq, par, 0);
            if (prof) hipEventRecord(evs[4 * s + 1], st);
            if ((e = hipEventRecord(ctx->ev_gather, st)) != hipSuccess) return e;
            if ((e = hipStreamWaitEvent(fs, ctx->ev_gather, 0)) != hipSuccess) return e;
            launch_finish(st, dmats, nmat, dactive, par);
            // far update: columns the finish and the next block elimination do not touch
            if (prof) hipEventRecord(evs[4 * s + 2], fs);
            if (use_mfma)
                launch_mfma(fs, dmats, nmat, f, lp, rq, par, 1);
            else
                launch_update_plain(fs, dmats, nmat, f, rq, par, 1);
            if (prof) hipEventRecord(evs[4 * s + 3], fs);
            if ((e = hipEventRecord(ctx->ev_far, fs)) != hipSuccess) return e;
        }
        // the far stream joins the main stream (the block ends with nothing pending)
        hipError_t e = hipStreamWaitEvent(st, ctx->ev_far, 0);
        return e != hipSuccess ? e : hipGetLastError();
    };
    // hipGraph of the block (not while profiling: those events are timed per launch).
    bool use_graph = !profile && !(opts && (opts->flags & MODLIN_FLAG_NO_GRAPH));
    if (use_graph) {
        char key[256];
        snprintf(key, sizeof key, "%p/%u/%u/%d/%d/%llu/%u/%u/%u/%u/%d/%u", (void *)base, nmat, b, (int)use_mfma,
                 (int)defer, (unsigned long long)p, rl.grid_gx, rl.grid_near, rl.grid_far, (unsigned)lp.L,
                 check_every, rl.far_qw);
        if (!ctx->rank_graph || ctx->rank_graph_key != key) {
            if (ctx->rank_graph) {
                hipGraphExecDestroy(ctx->rank_graph);
                ctx->rank_graph = nullptr;
                ctx->rank_graph_key.clear();
            }
            hipGraph_t graph = nullptr;
            hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
            if (e == hipSuccess) {
                hipError_t e1 = hipEventRecord(ctx->ev_far, st);
                if (e1 == hipSuccess) e1 = enqueue_block(false);
                e = hipStreamEndCapture(st, &graph);
                if (e == hipSuccess) e = e1;
            }
            if (e == hipSuccess && graph) e = hipGraphInstantiate(&ctx->rank_graph, graph, nullptr, nullptr, 0);
            if (graph) hipGraphDestroy(graph);
            if (e == hipSuccess && ctx->rank_graph) {
                ctx->rank_graph_key = key;
            } else {
                (void)hipGetLastError();  // capture unavailable: launch the blocks directly
                ctx->rank_graph = nullptr;
                use_graph = false;
            }
        }
    }
    ctx->stats.graph = use_graph ? 1.0 : 0.0;
    // Progress in matrix columns, as the reference's update_with_count(jend - col)
    // per panel (fp.rs:251): at every host check, the columns below each matrix's
    // current panel (all n once it is done), never decreasing (a missed pivot moves
    // a panel back).
    uint64_t total_cols = 0, shown_cols = 0, max_n = 0;
    for (uint64_t i = 0; i < count; i++) {
        total_cols += mats[i].n;
        max_n = std::max<uint64_t>(max_n, mats[i].n);
    }
    // Without progress reports or per-launch timing, the blocks that a batch with
    // no missed pivot needs (one panel of b columns per iteration) are queued
    // before the first host check, so that such a batch makes one round trip: the
    // active count and the ranks come back in one copy (k_collect).
    int ahead = 1;
    if (!progress && !profile)
        ahead = (int)std::max<uint64_t>(1, ((max_n + b - 1) / b + check_every - 1) / check_every);
    uint64_t *dout = (uint64_t *)(base + off_out);
    std::vector<uint64_t> hout(1 + 4 * count);
    for (uint64_t i = 0; i < count; i++) hout[1 + 4 * i] = hm[i].rank;  // a batch with nothing active
    MDL_HIP(hipEventRecord(ctx->ev_far, st));  // set-up done
    while (h_active > 0) {
        for (int k = 0; k < ahead; k++) {
            if (use_graph)
                MDL_HIP(hipGraphLaunch(ctx->rank_graph, st));
            else
                MDL_HIP(enqueue_block(profile));
            iters += check_every;
        }
        ahead = 1;
        launch_collect(st, dmats, nmat, dactive, dout);
        MDL_HIP(hipGetLastError());
        rc = copy_d2h(ctx, hout.data(), dout, hout.size() * sizeof(uint64_t));
        if (rc) return rc;
        h_active = (uint32_t)hout[0];
        if (profile) {
            for (int s = 0; s < check_every; s++) {
                float ms0 = 0, ms1 = 0;
                MDL_HIP(hipEventElapsedTime(&ms0, evs[4 * s], evs[4 * s + 1]));
                MDL_HIP(hipEventElapsedTime(&ms1, evs[4 * s + 2], evs[4 * s + 3]));
                upd_ms += ms0 + ms1;
                upd_launches += 2;
                far_ms += ms1;
                far_launches += 1;
            }
        }
        if (progress && h_active > 0) {
            rc = copy_d2h(ctx, hm, dmats, count * sizeof(RankMat));
            if (rc) return rc;
            uint64_t cols = 0;
            for (uint64_t i = 0; i < count; i++) cols += hm[i].done ? hm[i].n : std::min(hm[i].col, hm[i].n);
            if (cols > shown_cols) {
                shown_cols = cols;
                progress(user, cols, total_cols);
            }
        }
    }
    if (progress) progress(user, total_cols, total_cols);
    MDL_HIP(hipStreamSynchronize(fs));
    for (auto &e : evs) hipEventDestroy(e);
    double macs = 0, ubytes = 0, fmacs = 0;
    for (uint64_t i = 0; i < count; i++) {
        ranks[i] = hout[1 + 4 * i];
        macs += (double)hout[2 + 4 * i];
        ubytes += (double)hout[3 + 4 * i];
        fmacs += (double)hout[4 + 4 * i];
    }
    ctx->stats.macs = macs;
    ctx->stats.far_macs = fmacs;
    ctx->stats.far_ms = far_ms;
    ctx->stats.far_launches = far_launches;
    ctx->stats.update_bytes = ubytes;
    ctx->stats.update_ms = upd_ms;
    ctx->stats.update_launches = upd_launches;
    ctx->stats.iterations = (double)iters;
    ctx->stats.limb_count = use_mfma ? (double)lp.L : 0.0;
    ctx->stats.total_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    return MODLIN_OK;
}

}  // namespace mdl
