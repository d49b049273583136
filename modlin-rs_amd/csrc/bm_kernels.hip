// Berlekamp-Massey over F_p on gfx950 (reference: src/fp.rs:257-308).
//
// k_bm_batch: one workgroup per sequence.  Each step i
//   * the discrepancy d = s[i] + sum_{j=1..l} c[j] s[i-j] (fp.rs:274-278) is a
//     block-wide dot product: every lane accumulates its strided terms lazily in
//     128 bits (one Barrett reduction per safe_batch products), then the partials
//     are added mod p across the wavefront with DPP/shuffles and across waves in
//     LDS.  Addition mod p is associative, so d equals the reference's ordered sum.
//   * d = 0 only advances m (fp.rs:279-282);
//   * otherwise c[m..m+jmax) -= (d / bb) b[0..jmax) is a coalesced streaming axpy
//     (fp.rs:292-294, 301-303); on a length change (2l <= i, fp.rs:287-299) the old
//     c[0..l] is first copied to t and b/t swap roles by pointer.
// The state (c, b, t) lives in a per-sequence slice of HBM; s is read-only.

#include <algorithm>

#include "engine.hpp"
#include "rank_common.hpp"

namespace mdl {

struct BmSeq {
    const uint64_t *s;
    uint64_t *c, *b, *t;
    uint32_t n;
    uint32_t lc;
    unsigned long long bytes;  // algorithmic bytes: sum_i 8 + 16 l_i + [d_i != 0] 24 jmax_i + [change] 16 (old_l + 1)
};

template <int NT>
__device__ __forceinline__ uint64_t block_sum_mod(uint64_t v, uint64_t p, uint64_t *red) {
    // wavefront reduction (64 lanes)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        uint64_t o = __shfl_xor(v, off, 64);
        v = add_mod(v, o, p);
    }
    constexpr int NW = NT / 64;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (NW > 1) {
        if (lane == 0) red[wid] = v;
        __syncthreads();
        v = 0;
#pragma unroll
        for (int w = 0; w < NW; w++) v = add_mod(v, red[w], p);
        __syncthreads();
    }
    return v;
}

template <int NT, int MODE>
__global__ __launch_bounds__(NT) void k_bm_batch(BmSeq *seqs, FieldParams f) {
    BmSeq &Q = seqs[blockIdx.x];
    const uint32_t n = Q.n;
    const uint64_t *s = Q.s;
    uint64_t *c = Q.c, *b = Q.b, *t = Q.t;
    const uint64_t p = f.p, sb = f.safe_batch;
    const uint32_t tid = threadIdx.x;
    __shared__ uint64_t red[NT / 64];
    __shared__ uint64_t sh_inv;

    for (uint32_t j = tid; j < n; j += NT) {
        c[j] = 0;
        b[j] = 0;
        t[j] = 0;
    }
    __syncthreads();
    if (tid == 0) {
        c[0] = 1 % p;
        b[0] = 1 % p;
    }
    __syncthreads();
    uint32_t l = 0, bl = 1, m = 1;
    uint64_t bb_inv = 1 % p;
    unsigned long long bytes = 0;  // kept in a register, stored once at the end
    for (uint32_t i = 0; i < n; i++) {
        uint64_t ah = 0, al = 0, part = 0, cnt = 0;
        for (uint32_t j = 1 + tid; j <= l; j += NT) {
            uint64_t h, lo;
            mul_wide(c[j], s[i - j], h, lo);
            add128(ah, al, h, lo);
            if (++cnt == sb) {
                part = add_mod(part, reduce_t<MODE>(f, ah, al), p);
                ah = al = 0;
                cnt = 0;
            }
        }
        if (cnt) part = add_mod(part, reduce_t<MODE>(f, ah, al), p);
        uint64_t d = add_mod(block_sum_mod<NT>(part, p, red), s[i], p);
        bytes += 8ull + 16ull * l;
        if (d == 0) {
            m++;
            continue;
        }
        const uint64_t coef = mul_mod_t<MODE>(f, d, bb_inv);
        const uint32_t jmax = min(bl, n - m);
        if (2 * l <= i) {
            const uint32_t old_l = l;
            bytes += 24ull * jmax + 16ull * (old_l + 1);
            for (uint32_t j = tid; j <= old_l; j += NT) t[j] = c[j];
            if (tid == 0) sh_inv = inv_mod_bin(d, p);
            __syncthreads();
            for (uint32_t j = tid; j < jmax; j += NT) c[j + m] = sub_mod(c[j + m], mul_mod_t<MODE>(f, coef, b[j]), p);
            uint64_t *tmp = b;
            b = t;
            t = tmp;
            bl = old_l + 1;
            bb_inv = sh_inv;
            l = i + 1 - old_l;
            m = 1;
        } else {
            bytes += 24ull * jmax;
            for (uint32_t j = tid; j < jmax; j += NT) c[j + m] = sub_mod(c[j + m], mul_mod_t<MODE>(f, coef, b[j]), p);
            m++;
        }
        __syncthreads();
    }
    if (tid == 0) {
        Q.lc = l;
        Q.bytes = bytes;
    }
}

// ---------------------------------------------------------------------------
// k_bm_coop: one long sequence, cooperatively on the whole GPU.
//
// G workgroups each own a contiguous slice [lo, hi) of the indices of c, b and t.
// Per step a workgroup forms the partial discrepancy sum_{j in slice, 1<=j<=l}
// c[j] s[i-j] and publishes it; one grid barrier later every workgroup adds the G
// partials (addition mod p is associative, so d is the reference's sum) and
// applies the axpy c[j] -= coef b[j-m] to its own slice.  The scalar state
// (l, bl, m, 1/bb) is replicated: every workgroup derives it from the same d.
//
// Memory: a slice of c is only touched by its owner and s is read-only, so both
// stay in the ordinary caches.  b and t cross slices (b is read shifted by m; t
// becomes b at a length change) and the partials cross workgroups: they are
// accessed with device-scope atomic loads and stores, which are coherent across
// the XCDs' L2s, so no L2 write-back or invalidation is needed at the barrier --
// only a wait for each wave's own stores before its workgroup arrives.
//
// The launch is cooperative, so all workgroups are resident and every barrier
// completes; each wait is still bounded (BM_SPIN_LIMIT polls), and a wait that
// runs out sets the error flag instead of hanging.
// ---------------------------------------------------------------------------
struct BmCoop {
    unsigned int arrive;       // grid barrier arrivals (monotonic)
    unsigned int timeout;      // set if a barrier wait ran out
    uint64_t part[2][1024];  // partial discrepancies, double-buffered by step parity
};
constexpr uint32_t BM_COOP_THREADS = 512;
constexpr uint32_t BM_COOP_MIN = 65536;  // default length from which one sequence runs cooperatively
constexpr uint32_t BM_SPIN_LIMIT = 1u << 26;

__device__ __forceinline__ uint64_t ld_dev(gptr<const uint64_t> a) {
    return __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_dev(gptr<uint64_t> a, uint64_t v) {
    __hip_atomic_store(a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void coop_barrier(BmCoop *cs, unsigned int target) {
    __builtin_amdgcn_s_waitcnt(0);  // this wave's device-scope stores are complete
    __syncthreads();
    if (threadIdx.x == 0) {
        __hip_atomic_fetch_add(&cs->arrive, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t spins = 0;
        while (__hip_atomic_load(&cs->arrive, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (++spins == BM_SPIN_LIMIT) {
                __hip_atomic_store(&cs->timeout, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    __syncthreads();
}

template <int MODE>
__global__ __launch_bounds__(BM_COOP_THREADS) void k_bm_coop(BmSeq *seq, FieldParams f, BmCoop *cs) {
    constexpr int NT = BM_COOP_THREADS;
    constexpr int PF = 4;  // prefetched b operands per thread
    BmSeq &Q = *seq;
    const uint32_t n = Q.n, G = gridDim.x, g = blockIdx.x, tid = threadIdx.x, lane = tid & 63;
    const gptr<const uint64_t> s = to_global(Q.s);
    gptr<uint64_t> c = to_global(Q.c), b = to_global(Q.b), t = to_global(Q.t);
    gptr<uint64_t> part = to_global(&cs->part[0][0]);
    const uint64_t p = f.p, sb = f.safe_batch;
    const uint32_t span = (n + G - 1) / G;
    const uint32_t lo = min(n, g * span), hi = min(n, lo + span);
    __shared__ uint64_t red[NT / 64];
    __shared__ uint64_t sh_d;
    __shared__ uint32_t sh_abort;
    unsigned int bars = 0;
    if (tid == 0) sh_abort = 0;
    const bool tagged = f.p <= P61;  // partials below 2^61 leave 3 bits for a step tag

    for (uint32_t j = lo + tid; j < hi; j += NT) {
        c[j] = j == 0 ? 1 % p : 0;
        st_dev(b + j, j == 0 ? 1 % p : 0);
        st_dev(t + j, 0);
    }
    coop_barrier(cs, ++bars * G);  // b is initialised everywhere
    // The reference's state (fp.rs:267-270) with the saved discrepancy bb itself in
    // place of its inverse: the update is the inversionless c <- bb c - d x^m b,
    // which scales c (c[0] included), b and every later discrepancy by nonzero
    // factors, so the zero tests, the length changes and l are the reference's.
    uint32_t l = 0, bl = 1, m = 1;
    uint64_t bb = 1 % p;
    bool fresh_b = false;  // b was rewritten at the previous step (no prefetch)
    unsigned long long bytes = 0;
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t jmax = min(bl, n - m);
        const uint32_t a0 = max(lo, m), a1 = min(hi, m + jmax);
        // this thread's indices are j = lo + tid + q NT; the axpy ones (j in [a0, a1))
        // start at q = q0, and the first PF of them are prefetched
        const uint32_t first = lo + tid;
        const uint32_t q0 = a0 > first ? (a0 - first + NT - 1) / NT : 0;
        uint64_t bpre[PF];
        if (!fresh_b) {
#pragma unroll
            for (int r = 0; r < PF; r++) {
                const uint32_t j = first + (q0 + r) * NT;
                bpre[r] = j < a1 ? ld_dev(b + (j - m)) : 0;
            }
        }
        // this slice's part of d = sum_{j=0..l} c[j] s[i-j] (fp.rs:274-278 with c[0],
        // which the inversionless update scales, in place of the reference's 1)
        uint64_t ah = 0, al = 0, mine = 0, cnt = 0;
        const uint32_t j1 = min(hi, l + 1);
        for (uint32_t j = lo + tid; j < j1; j += NT) {
            uint64_t h, lw;
            mul_wide(c[j], s[i - j], h, lw);
            add128(ah, al, h, lw);
            if (++cnt == sb) {
                mine = add_mod(mine, reduce_t<MODE>(f, ah, al), p);
                ah = al = 0;
                cnt = 0;
            }
        }
        if (cnt) mine = add_mod(mine, reduce_t<MODE>(f, ah, al), p);
        mine = block_sum_mod<NT>(mine, p, red);
        const gptr<uint64_t> slot = part + (i & 1) * 1024;
        if (tagged) {
            // Publish the partial with the step's tag in bits 61..63 (the slot of
            // this parity was last written two steps ago, with the previous tag;
            // tags start at 1, so the zeroed slots never match); wave 0 then polls
            // every slot until it carries this step's tag.  One round trip.
            const uint64_t tag = (uint64_t)(((i >> 1) + 1) & 7) << 61;
            __builtin_amdgcn_s_waitcnt(0);  // this wave's t stores are complete
            __syncthreads();
            if (tid == 0) st_dev(slot + g, mine | tag);
            if (tid < 64) {
                uint64_t v = 0;
                for (uint32_t w = lane; w < G; w += 64) {
                    uint64_t x = ld_dev(slot + w);
                    uint32_t spins = 0;
                    while ((x & ~P61) != tag) {
                        if (++spins == BM_SPIN_LIMIT) {
                            __hip_atomic_store(&cs->timeout, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            sh_abort = 1;
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
                        x = ld_dev(slot + w);
                    }
                    v = add_mod(v, x & P61, p);
                }
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) v = add_mod(v, __shfl_xor(v, off, 64), p);
                if (tid == 0) sh_d = v;
            }
        } else {
            if (tid == 0) st_dev(slot + g, mine);
            coop_barrier(cs, ++bars * G);
            if (tid < 64) {
                uint64_t v = 0;
                for (uint32_t w = lane; w < G; w += 64) v = add_mod(v, ld_dev(slot + w), p);
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) v = add_mod(v, __shfl_xor(v, off, 64), p);
                if (tid == 0) sh_d = v;
            }
        }
        __syncthreads();
        // a wait that ran out ends the kernel: the other workgroups then run out
        // too, at their next exchange, and the host reports the error
        if (sh_abort || __hip_atomic_load(&cs->timeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
        const uint64_t d = sh_d;
        bytes += 8ull + 16ull * l;
        if (d == 0) {  // fp.rs:279-282
            m++;
            fresh_b = false;
            continue;
        }
        if (fresh_b) {
#pragma unroll
            for (int r = 0; r < PF; r++) {
                const uint32_t j = first + (q0 + r) * NT;
                bpre[r] = j < a1 ? ld_dev(b + (j - m)) : 0;
            }
        }
        // c[j] <- bb c[j] - d b[j - m] on this slice (the b term for j in [m, m + jmax))
        const bool change = 2 * l <= i;  // length change (fp.rs:287-299): t <- c[0..l] first
        const uint32_t old_l = l;
        bytes += 24ull * jmax + (change ? 16ull * (old_l + 1) : 0);
        if (change)
            for (uint32_t j = lo + tid; j < min(hi, old_l + 1); j += NT) st_dev(t + j, c[j]);
        const uint32_t top = min(hi, max(l + 1, m + jmax));
        uint32_t q = 0;
        for (uint32_t j = first; j < top; j += NT, q++) {
            uint64_t y = 0;
            if (j >= a0 && j < a1) {
                const uint32_t k = q - q0;  // this thread's k-th axpy element
                if (k < (uint32_t)PF) {
#pragma unroll
                    for (int r = 0; r < PF; r++)
                        if (r == (int)k) y = bpre[r];
                } else {
                    y = ld_dev(b + (j - m));
                }
            }
            c[j] = ff_step_t<MODE>(f, bb, c[j], d, y);
        }
        if (change) {
            const gptr<uint64_t> tmp = b;
            b = t;
            t = tmp;
            bl = old_l + 1;
            bb = d;
            l = i + 1 - old_l;
            m = 1;
            fresh_b = true;
        } else {
            m++;
            fresh_b = false;
        }
        __syncthreads();
    }
    if (g == 0 && tid == 0) {
        Q.lc = l;
        Q.bytes = bytes;
    }
}

static size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

__global__ void k_widen_seq(const uint32_t *src, uint64_t *dst, size_t count) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

int lc_batch(modlin_ctx *ctx, uint64_t p, const modlin_seq_desc_t *seqs, uint64_t count, const modlin_opts_t *opts,
             uint64_t *lcs) {
    ctx->stats = modlin_stats_t{};
    if (count == 0) return MODLIN_OK;
    const FieldParams f = make_params(p);
    const bool profile = opts && (opts->flags & MODLIN_FLAG_PROFILE);
    std::vector<size_t> offS(count), offC(count);
    size_t cur = align_up(count * sizeof(BmSeq), 256);
    uint64_t maxn = 0;
    for (uint64_t i = 0; i < count; i++) {
        const uint64_t n = seqs[i].n;
        if (n == 0 || n > 0x7fffffffull) {
            ctx->err = n == 0 ? "empty sequence (the reference panics, fp.rs:265)" : "sequence too long";
            return MODLIN_EINVAL;
        }
        if (!seqs[i].data || (seqs[i].elem != MODLIN_ELEM_U64 && seqs[i].elem != MODLIN_ELEM_U32)) {
            ctx->err = "bad sequence descriptor";
            return MODLIN_EINVAL;
        }
        maxn = n > maxn ? n : maxn;
        offS[i] = cur;
        cur = align_up(cur + n * 8, 256);
        offC[i] = cur;
        cur = align_up(cur + 3 * n * 8, 256);
    }
    const size_t off_stage = cur;
    cur = align_up(cur + maxn * 4, 256);
    const size_t off_coop = cur;
    cur = align_up(cur + sizeof(BmCoop), 256);
    int rc = arena_reserve(ctx, cur);
    if (rc) return rc;
    char *base = (char *)ctx->arena;
    hipStream_t st = ctx->stream;
    std::vector<BmSeq> hs(count);
    for (uint64_t i = 0; i < count; i++) {
        const uint64_t n = seqs[i].n;
        uint64_t *sd = (uint64_t *)(base + offS[i]);
        uint64_t *cd = (uint64_t *)(base + offC[i]);
        const bool dev = seqs[i].mem == MODLIN_MEM_DEVICE;
        const size_t esz = seqs[i].elem == MODLIN_ELEM_U64 ? 8 : 4;
        void *dst = seqs[i].elem == MODLIN_ELEM_U64 ? (void *)sd : (void *)(base + off_stage);
        if (dev) {
            MDL_HIP(hipMemcpyAsync(dst, seqs[i].data, n * esz, hipMemcpyDeviceToDevice, st));
        } else {
            rc = copy_h2d(ctx, dst, seqs[i].data, n * esz);
            if (rc) return rc;
        }
        if (seqs[i].elem == MODLIN_ELEM_U32)
            hipLaunchKernelGGL(k_widen_seq, dim3(256), dim3(256), 0, st, (const uint32_t *)dst, sd, (size_t)n);
        hs[i].s = sd;
        hs[i].c = cd;
        hs[i].b = cd + n;
        hs[i].t = cd + 2 * n;
        hs[i].n = (uint32_t)n;
        hs[i].lc = 0;
        hs[i].bytes = 0;
    }
    BmSeq *ds = (BmSeq *)base;
    rc = copy_h2d(ctx, ds, hs.data(), count * sizeof(BmSeq));
    if (rc) return rc;
    if (profile) MDL_HIP(hipEventRecord(ctx->evA, st));
    // One long sequence: the cooperative kernel over the whole GPU (one workgroup
    // per CU, at most 1024 partials).
    const uint32_t coop_min = (opts && opts->coop_min) ? opts->coop_min : BM_COOP_MIN;
    if (count == 1 && maxn >= coop_min) {
        BmCoop *cs = (BmCoop *)(base + off_coop);
        MDL_HIP(hipMemsetAsync(cs, 0, sizeof(BmCoop), st));
        const modlin_tuning_t tu = select_tuning(ctx->tune_table, &ctx->tune_override, p, maxn, count, true);
        const uint32_t G = tu.bm_coop_wgs ? tu.bm_coop_wgs : (uint32_t)std::min(ctx->cu_count, 1024);
        void *args[] = {(void *)&ds, (void *)&f, (void *)&cs};
        hipError_t e;
        if (f.mode == MODE_M61)
            e = hipLaunchCooperativeKernel((const void *)k_bm_coop<MODE_M61>, dim3(G), dim3(BM_COOP_THREADS), args, 0, st);
        else if (f.mode == MODE_M31)
            e = hipLaunchCooperativeKernel((const void *)k_bm_coop<MODE_M31>, dim3(G), dim3(BM_COOP_THREADS), args, 0, st);
        else
            e = hipLaunchCooperativeKernel((const void *)k_bm_coop<MODE_GENERIC>, dim3(G), dim3(BM_COOP_THREADS), args, 0,
                                           st);
        if (e != hipSuccess) {
            // refused (another process may hold the device's cooperative-launch
            // resources): the one-workgroup kernel below gives the same result
            (void)hipGetLastError();
            ctx->stats.bm_coop = 0;
            goto single;
        }
        if (profile) MDL_HIP(hipEventRecord(ctx->evB, st));
        uint32_t timeout = 0;
        rc = copy_d2h(ctx, &timeout, &cs->timeout, sizeof(timeout));
        if (rc) return rc;
        if (timeout) {
            ctx->err = "cooperative Berlekamp-Massey: a grid barrier wait ran out";
            return MODLIN_EHIP;
        }
        rc = copy_d2h(ctx, hs.data(), ds, count * sizeof(BmSeq));
        if (rc) return rc;
        lcs[0] = hs[0].lc;
        ctx->stats.bm_bytes = (double)hs[0].bytes;
        ctx->stats.bm_coop = (double)G;
        if (profile) {
            float ms = 0;
            MDL_HIP(hipEventElapsedTime(&ms, ctx->evA, ctx->evB));
            ctx->stats.bm_ms = ms;
        }
        return MODLIN_OK;
    }
single:
    const bool wide = count == 1 && maxn >= 4096;
#define MDL_BM(NT, MODE) hipLaunchKernelGGL((k_bm_batch<NT, MODE>), dim3((uint32_t)count), dim3(NT), 0, st, ds, f)
    if (f.mode == MODE_M61) {
        if (wide) MDL_BM(1024, MODE_M61); else MDL_BM(256, MODE_M61);
    } else if (f.mode == MODE_M31) {
        if (wide) MDL_BM(1024, MODE_M31); else MDL_BM(256, MODE_M31);
    } else {
        if (wide) MDL_BM(1024, MODE_GENERIC); else MDL_BM(256, MODE_GENERIC);
    }
#undef MDL_BM
    if (profile) MDL_HIP(hipEventRecord(ctx->evB, st));
    MDL_HIP(hipGetLastError());
    rc = copy_d2h(ctx, hs.data(), ds, count * sizeof(BmSeq));
    if (rc) return rc;
    double bytes = 0;
    for (uint64_t i = 0; i < count; i++) {
        lcs[i] = hs[i].lc;
        bytes += (double)hs[i].bytes;
    }
    ctx->stats.bm_bytes = bytes;
    if (profile) {
        float ms = 0;
        MDL_HIP(hipEventElapsedTime(&ms, ctx->evA, ctx->evB));
        ctx->stats.bm_ms = ms;
    }
    return MODLIN_OK;
}

}  // namespace mdl
