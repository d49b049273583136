// Berlekamp-Massey over F_p on gfx950 (reference: src/fp.rs:257-308).
//
// Which kernel runs (lc_batch):
//   one sequence of length >= coop_min, p <= 2^61   k_bm_block   (whole GPU, 64 steps per exchange)
//   one such sequence, p > 2^61                      k_bm_coop    (whole GPU, one exchange per step)
//   batches (and the rest), p < 2^62                 k_bm_batch_blk (one workgroup per sequence,
//                                                                  64 steps per pass; matrix
//                                                                  cores for 2^61 - 1)
//   method PLAIN, or p >= 2^62                       k_bm_batch   (one workgroup per sequence,
//                                                                  one step per pass)
// All four keep the reference's per-step decisions (d = 0, length change, l) and
// return its linear complexity; the blocked ones reach each step's discrepancy
// through the algebra described at k_bm_block.
//
// k_bm_batch: one workgroup per sequence.  Thread tid owns the indices
// j = tid (mod NT) of c and t in both phases of a step, so it only ever reads c
// values it wrote itself and a step needs one workgroup barrier (the one inside the
// discrepancy reduction).  Each step i
//   * the discrepancy d = sum_{j=0..l} c[j] s[i-j] (fp.rs:274-278; c[0] is 1 in the
//     reference and a nonzero scale here, see below) is a block-wide dot product:
//     every lane accumulates its terms lazily (MacSum), four loads in flight at a
//     time, the partials are added mod p across the wavefront with shuffles and
//     across waves in LDS.  Addition mod p is associative, so the zero test sees
//     the reference's sum;
//   * d = 0 only advances m (fp.rs:279-282);
//   * otherwise c <- bb c - d x^m b over j in [0, max(l+1, m+jmax)) (fp.rs:292-294,
//     301-303 without the inverse of bb: the inversionless update scales c and every
//     later discrepancy by nonzero factors, so the zero tests, the length changes
//     and l are the reference's); on a length change (2l <= i, fp.rs:287-299) the
//     old c[0..l] is first copied to t, and b/t swap roles by pointer.  The update
//     loop also forms the thread's part of the next step's dot product from the
//     values it has just written, so after an update the next step skips its own
//     pass over c.  b is read shifted by m (other threads' indices): the barrier of
//     the next step orders those reads after the copy, and the copy of a later
//     length change after them.
// The state (c, b, t) lives in a per-sequence slice of HBM; s is read-only.

#include <string.h>

#include <algorithm>
#include <chrono>
#include <mutex>
#include <thread>

#include "engine.hpp"
#include "rank_common.hpp"

namespace mdl {

struct BmSeq {
    const uint64_t *s;
    uint64_t *c, *b, *t;
    uint64_t *u;  // fourth state buffer (the blocked kernels double-buffer c and b)
    uint32_t n;
    uint32_t lc;
    uint32_t timeout;  // k_bm_batch_blk: a follower's wait for a step record ran out
    uint32_t pad;
    uint64_t *prog;    // steps done, for the host's progress callback (pinned host memory), or null
    unsigned long long bytes;  // algorithmic bytes: sum_i 8 + 16 l_i + [d_i != 0] 24 jmax_i + [change] 16 (old_l + 1)
};

// Steps done so far, for the host's progress callback: a system-scope store to
// pinned host memory that the host reads while the kernel runs.
__device__ __forceinline__ void bm_progress(uint64_t *prog, uint64_t steps) {
    if (prog) __hip_atomic_store(prog, steps, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

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
    constexpr int NW = NT / 64;
    constexpr uint32_t U = 4;  // loads in flight per lane
    BmSeq &Q = seqs[blockIdx.x];
    const uint32_t n = Q.n;
    const gptr<const uint64_t> s = to_global(Q.s);
    gptr<uint64_t> c = to_global(Q.c), b = to_global(Q.b), t = to_global(Q.t);
    const uint64_t p = f.p;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wid = uni32(tid >> 6);
    __shared__ uint64_t red[2][NW];

    // c = 1, b = 1 (fp.rs:267-270); entries of b and t at or above their length
    // are never read, so only c needs clearing
    for (uint32_t j = tid; j < n; j += NT) c[j] = j == 0 ? 1 % p : 0;
    if (tid == 0) b[0] = 1 % p;
    uint32_t l = 0, bl = 1, m = 1;
    uint64_t bb = 1 % p;
    unsigned long long bytes = 0;  // kept in a register, stored once at the end
    MacSum<MODE> nxt;    // this thread's part of the next discrepancy, formed by an update
    bool fused = false;  // nxt holds it (the previous step updated c)
    for (uint32_t i = 0; i < n; i++) {
        if ((i & 255) == 0 && tid == 0) bm_progress(Q.prog, i);
        uint32_t j = tid;
        uint64_t part;
        if (fused) {
            part = nxt.value(f);
        } else {
            MacSum<MODE> acc;
            for (; j + (U - 1) * NT <= l; j += U * NT) {
                uint64_t cv[U], sv[U];
#pragma unroll
                for (uint32_t u = 0; u < U; u++) {
                    cv[u] = c[j + u * NT];
                    sv[u] = s[i - j - u * NT];
                }
#pragma unroll
                for (uint32_t u = 0; u < U; u += 2) acc.add2(f, cv[u], sv[u], cv[u + 1], sv[u + 1]);
            }
            for (; j + NT <= l; j += 2 * NT) acc.add2(f, c[j], s[i - j], c[j + NT], s[i - j - NT]);
            if (j <= l) acc.add2(f, c[j], s[i - j], 0, 0);
            part = acc.value(f);
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) part = add_mod(part, __shfl_xor(part, off, 64), p);
        // red[i & 1] is rewritten two steps later, after every thread has passed
        // the next step's barrier and so has read it
        if (lane == 0) red[i & 1][wid] = part;
        __syncthreads();
        uint64_t d = 0;
#pragma unroll
        for (int w = 0; w < NW; w++) d = add_mod(d, red[i & 1][w], p);
        bytes += 8ull + 16ull * l;
        if (d == 0) {
            m++;
            fused = false;
            continue;
        }
        const uint32_t jmax = min(bl, n - m);
        const bool change = 2 * l <= i;
        const uint32_t old_l = l;
        bytes += 24ull * jmax + (change ? 16ull * (old_l + 1) : 0);
        const uint32_t top = max(l + 1, m + jmax), a1 = m + jmax;
        // The update also forms this thread's part of the next discrepancy,
        // sum_j c'[j] s[i+1-j]: it covers every j < top, c' is zero above the next
        // l (and below top it is exactly zero there), and c is zero from top on.
        const uint32_t inext = i + 1 < n ? i + 1 : 0;  // s index bound (no next step: s[0..0], unused)
        nxt = MacSum<MODE>();
        j = tid;
        for (; j + (U - 1) * NT < top; j += U * NT) {
            uint64_t cv[U], yv[U], sv[U], nv[U];
#pragma unroll
            for (uint32_t u = 0; u < U; u++) {
                const uint32_t jj = j + u * NT;
                cv[u] = c[jj];
                yv[u] = (jj >= m && jj < a1) ? b[jj - m] : 0;
                sv[u] = jj <= inext ? s[inext - jj] : 0;
            }
#pragma unroll
            for (uint32_t u = 0; u < U; u++) {
                const uint32_t jj = j + u * NT;
                if (change && jj <= old_l) t[jj] = cv[u];
                nv[u] = ff_step_t<MODE>(f, bb, cv[u], d, yv[u]);
                c[jj] = nv[u];
            }
#pragma unroll
            for (uint32_t u = 0; u < U; u += 2) nxt.add2(f, nv[u], sv[u], nv[u + 1], sv[u + 1]);
        }
        for (; j < top; j += NT) {
            const uint64_t cv = c[j];
            const uint64_t y = (j >= m && j < a1) ? b[j - m] : 0;
            const uint64_t sv = j <= inext ? s[inext - j] : 0;
            if (change && j <= old_l) t[j] = cv;
            const uint64_t nv = ff_step_t<MODE>(f, bb, cv, d, y);
            c[j] = nv;
            nxt.add2(f, nv, sv, 0, 0);
        }
        fused = true;
        if (change) {
            const gptr<uint64_t> tmp = b;
            b = t;
            t = tmp;
            bl = old_l + 1;
            bb = d;
            l = i + 1 - old_l;
            m = 1;
        } else {
            m++;
        }
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

__device__ __forceinline__ uint64_t lane_bcast64(uint64_t v, uint32_t src) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, (int)src);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), (int)src);
    return ((uint64_t)hi << 32) | lo;
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
        if ((i & 255) == 0 && blockIdx.x == 0 && tid == 0) bm_progress(seq->prog, i);
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

// ---------------------------------------------------------------------------
// k_bm_block: one long sequence on the whole GPU, BM_BK steps per exchange.
//
// With C, B the connection polynomials and m, l, bl, bb the scalar state at the
// start of a block of steps i0 .. i0+K-1, every c and b of the block is a
// combination  c_t = U_t C + V_t x^m B  with polynomials U_t, V_t of degree < K,
// and the discrepancy of step i0+t is  d_t = [x^(i0+t)] (c_t S),  S = sum s_j x^j
// (the reference's d = sum_{j<=l} c[j] s[i-j], fp.rs:274-278, with c[0] scaled;
// c has no terms above l).  So a block needs only the 2K numbers
//     Ec(r) = [x^(i0+r)] (C S),   Eb(r) = [x^(i0+r)] (x^m B S),   r < K,
// sums over the coefficients of C and B that every workgroup forms for the
// indices it owns.  One exchange adds them up (reduce-scatter then all-gather
// through tagged words, as in k_bm_coop); then one wavefront replays the K steps
// exactly as the reference does (fp.rs:271-305, inversionless): lane r holds the
// discrepancy of step i0+r of the current c and of the current x^m b,
//     d = e(t);  e <- bb e - d eb;  eb <- x eb, or x e on a length change,
// and lane a holds coefficient a of U, V (and of the b combinations).  Last, each
// workgroup forms the new C = U C + V x^m B (and B, if it changed) for its own
// indices, from windows of the old C and B.  Addition mod p is associative and
// every product is exact, so d, the length changes and l are the reference's.
//
// Ownership: indices of c and b in chunks of BM_BW, chunk q G + g to workgroup g,
// kept in LDS (sCo, sBo) and mirrored to HBM for the other workgroups' windows.
// The HBM copies are double-buffered: a block reads one and writes the other, and
// a workgroup only writes after the exchange, i.e. after every workgroup has
// finished the previous block's reads.  Cross-workgroup data uses device-scope
// loads and stores; each wave waits for its stores before its workgroup
// publishes.  Needs p < 2^61 (three tag bits) and at most BM_BQ chunks per
// workgroup; k_bm_coop covers the rest.
// ---------------------------------------------------------------------------
constexpr uint32_t BM_BK = 64;  // steps per block (one wavefront replays them)
constexpr uint32_t BM_BW = 64;  // chunk width
constexpr uint32_t BM_BQ = 16;  // chunks per workgroup at most
constexpr uint32_t BM_BLOCK_THREADS = 512;
constexpr uint32_t BM_BWIN = BM_BW + BM_BK - 1;
constexpr uint32_t BM_BLOCK_SPIN_LIMIT = 1u << 22;  // polls of one wait (seconds)

struct BmBlockX {
    unsigned int timeout;  // set if a wait ran out
    unsigned int pad[63];
    uint64_t part[2][2 * BM_BK][1024];  // per block parity, value, workgroup
    uint64_t tot[2][2 * BM_BK];         // per block parity, value
};

// Lane r gets lane r-1's value, lane 0 gets 0 (DPP whole-wave shift right).
__device__ __forceinline__ uint64_t shift_up1(uint64_t v, uint32_t) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)v, 0x138, 0xf, 0xf, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(v >> 32), 0x138, 0xf, 0xf, false);
    return ((uint64_t)hi << 32) | lo;
}

// Polls a tagged word until it carries `tag`; false (and the error flag set) if
// the wait ran out.
__device__ __forceinline__ bool poll_tag(gptr<const uint64_t> a, uint64_t tag, uint64_t &val, BmBlockX *x) {
    uint64_t v = ld_dev(a);
    uint32_t spins = 0;
    while ((v & ~P61) != tag) {
        if (++spins == BM_BLOCK_SPIN_LIMIT) {
            __hip_atomic_store(&x->timeout, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
        v = ld_dev(a);
    }
    val = v & P61;
    return true;
}

// Records of the steps of a block, written by wavefront 0 of bm_replay_block: the
// step's discrepancy d (< p < 2^62) with its kind in bits 62-63 (0: d = 0, 1: update,
// 2: update with a length change).  *prog, written with release order every few
// steps, says how many records are there; followers wait for record t with
// acquire loads.  A follower holds the block's records one per lane in rv, reloaded
// (one LDS read for the wave) each time the count moves past what it holds, and
// takes record t from lane t: no LDS round trip on a step.  A wait that runs out
// sets *fail (LDS); the replay then finishes its bounded loop on records that
// are not final, and the kernel stops after it and reports the failure.
__device__ __forceinline__ uint64_t bm_record(const uint64_t *rec, uint32_t *prog, uint32_t t, uint32_t lane,
                                              uint32_t &avail, uint64_t &rv, uint32_t *fail) {
    if (t >= avail) {
        uint32_t spins = 0;
        for (; t >= avail && spins < BM_BLOCK_SPIN_LIMIT; spins++) {
            avail = uni32(__hip_atomic_load(prog, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP));
            if (t >= avail) __builtin_amdgcn_s_sleep(0);
        }
        if (t >= avail) {
            *fail = 1;
            avail = t + 1;  // no further waits in this block
        }
        rv = rec[lane];  // records below avail are final; the others are not read
    }
    return lane_bcast64(rv, t);
}

// Replays the steps of one block for one polynomial pair (P, P'): P <- bb P - d P',
// P' <- x P' (or x P on a length change, when Pb keeps the old P and bb becomes d).
// For 2^61 - 1 the steps are lazy (values in [0, p + 7]) and P, Pb are reduced once
// at the end.
template <int MODE>
__device__ __forceinline__ void bm_replay_poly(const FieldParams &f, uint32_t Kp, uint32_t lane, uint64_t bb,
                                               uint64_t P, uint64_t Pp, uint32_t *prog, const uint64_t *rec,
                                               uint64_t *outP, uint64_t *outPb, uint32_t *fail) {
    uint64_t Pb = 0, rv = 0;
    uint32_t avail = 0;
    for (uint32_t t = 0; t < Kp; t++) {
        const uint64_t r = bm_record(rec, prog, t, lane, avail, rv, fail);
        const uint32_t fl = (uint32_t)(r >> 62);
        if (fl == 0) {
            Pp = shift_up1(Pp, lane);
            continue;
        }
        const uint64_t d = r & ((1ull << 62) - 1);  // d < p < 2^62
        const uint64_t nP = MODE == MODE_M61 ? ff_step_m61_lazy(bb, P, d, Pp) : ff_step_t<MODE>(f, bb, P, d, Pp);
        if (fl == 2) {
            Pp = shift_up1(P, lane);
            Pb = P;
            bb = d;
        } else {
            Pp = shift_up1(Pp, lane);
        }
        P = nP;
    }
    if (MODE == MODE_M61) {
        P = P >= P61 ? P - P61 : P;
        Pb = Pb >= P61 ? Pb - P61 : Pb;
    }
    outP[lane] = P;
    outPb[lane] = Pb;
}

// The replay of one block of K steps (k_bm_block, k_bm_batch_blk).  Wavefront 0
// replays the discrepancies (lane r: step i0+r of the current c and of the current
// x^m b) with the length test and writes one record per step; wavefronts 1 and 2
// follow it with the coefficients of U (C's part) and V (B's part), wavefront 3
// with the rest of the reference's scalar state (m, bl) and the byte count.  The
// block-start state (l, bl, m, bb) is the same in every thread; the new state goes
// to sState: l, bl, m, bb, whether B changed, the byte count, whether c changed.
// A follower whose wait for a record ran out sets *sFail.
template <int MODE>
__device__ __forceinline__ void bm_replay_block(const FieldParams &f, uint32_t n, uint32_t i0, uint32_t Kp,
                                                uint32_t lane, uint32_t wid, uint32_t l, uint32_t bl, uint32_t m,
                                                uint64_t bb, unsigned long long &bytes, const uint64_t *sE,
                                                uint64_t (*sPoly)[BM_BK], uint64_t *sRec, uint64_t *sJunk,
                                                uint32_t *sProg_, uint64_t *sState, uint32_t *sFail) {
    constexpr uint32_t K = BM_BK;
    const uint64_t p = f.p;
    uint32_t &sProg = *sProg_;
    if (wid == 0) {
        // for 2^61 - 1, e and eb are kept in [0, p + 7] and only the broadcast d
        // is reduced (on the scalar unit)
        uint64_t e = sE[lane], eb = sE[K + lane];
        __asm__ volatile("" : "+v"(e), "+v"(eb));  // loaded before the loop: no LDS wait at its head
        uint64_t recv = 0;  // lane t: the record of step t, stored to sRec with each progress update
        for (uint32_t t = 0; t < Kp; t++) {
            const uint32_t i = i0 + t;
            uint64_t d = lane_bcast64(e, t);
            if (MODE == MODE_M61 && ((d + 1) >> 61) != 0) d -= P61;  // d >= p for d < 2^62, on the scalar unit
            uint64_t fl = 0;
            if (d == 0) {  // fp.rs:279-282
                eb = shift_up1(eb, lane);
            } else {
                const uint64_t ne = MODE == MODE_M61 ? ff_step_m61_lazy(bb, e, d, eb) : ff_step_t<MODE>(f, bb, e, d, eb);
                if (2 * l <= i) {  // length change, fp.rs:287-299
                    eb = shift_up1(e, lane);
                    l = i + 1 - l;
                    bb = d;
                    fl = 2;
                } else {
                    eb = shift_up1(eb, lane);
                    fl = 1;
                }
                e = ne;
            }
            recv = lane == t ? d | (fl << 62) : recv;
            if ((t & 3) == 3 || t + 1 == Kp) {
                // records 0 .. t, and then the count; lanes above t store values that
                // no follower reads before they are rewritten.  Lane 0 releases the
                // count, the other lanes store to their junk slots: no branch on the
                // lane, so the loop's control stays scalar
                sRec[lane] = recv;
                __hip_atomic_store(lane == 0 ? &sProg : (uint32_t *)&sJunk[64 + lane], t + 1, __ATOMIC_RELEASE,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
        if (lane == 0) sState[3] = bb;
    } else if (wid == 1) {
        bm_replay_poly<MODE>(f, Kp, lane, bb, lane == 0 ? 1 % p : 0, 0, &sProg, sRec, sPoly[0], sPoly[2], sFail);
    } else if (wid == 2) {
        bm_replay_poly<MODE>(f, Kp, lane, bb, 0, lane == 0 ? 1 % p : 0, &sProg, sRec, sPoly[1], sPoly[3], sFail);
    } else if (wid == 3) {
        uint32_t avail = 0, upd = 0, bch = 0;
        uint64_t rv = 0;
        for (uint32_t t = 0; t < Kp; t++) {
            const uint32_t fl = (uint32_t)(bm_record(sRec, &sProg, t, lane, avail, rv, sFail) >> 62);
            const uint32_t i = i0 + t;
            bytes += 8ull + 16ull * l;
            if (fl == 0) {
                m++;
                continue;
            }
            const uint32_t jmax = min(bl, n - m);
            bytes += 24ull * jmax + (fl == 2 ? 16ull * (l + 1) : 0);
            upd = 1;
            if (fl == 2) {
                const uint32_t old_l = l;
                bl = old_l + 1;
                l = i + 1 - old_l;
                m = 1;
                bch = 1;
            } else {
                m++;
            }
        }
        if (lane == 0) {
            sState[0] = l;
            sState[1] = bl;
            sState[2] = m;
            sState[4] = bch;
            sState[5] = bytes;
            sState[6] = upd;
        }
    }
}

template <int MODE>
__global__ __launch_bounds__(BM_BLOCK_THREADS) void k_bm_block(BmSeq *seq, FieldParams f, BmBlockX *x) {
    constexpr uint32_t NT = BM_BLOCK_THREADS, K = BM_BK, W = BM_BW, QM = BM_BQ, WIN = BM_BWIN, V2 = 2 * K;
    constexpr uint32_t NW = NT / 64;
    BmSeq &Q = *seq;
    const uint32_t n = Q.n, G = gridDim.x, g = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wid = uni32(tid >> 6);
    const gptr<const uint64_t> s = to_global(Q.s);
    const gptr<uint64_t> Cg[2] = {to_global(Q.c), to_global(Q.t)}, Bg[2] = {to_global(Q.b), to_global(Q.u)};
    const uint64_t p = f.p;
    // chunks q = 0 .. nq-1 of this workgroup start at (q G + g) W
    const uint32_t nchunks = (n + W - 1) / W;
    const uint32_t nq = nchunks > g ? (nchunks - g + G - 1) / G : 0;  // <= QM (host)
    __shared__ uint64_t sCo[QM * W], sBo[QM * W];
    __shared__ uint64_t stage[QM * 2 * WIN];
    __shared__ uint64_t sPoly[4][K];
    __shared__ uint64_t sRed[NT / V2][V2];
    __shared__ uint64_t sE[V2];
    __shared__ uint64_t sWave[NW];
    __shared__ uint64_t sRec[K];
    __shared__ uint64_t sJunk[2 * 64];
    __shared__ uint64_t sState[7];
    __shared__ uint32_t sAbort, sProg, sFail;

    for (uint32_t k = tid; k < QM * W; k += NT) {
        sCo[k] = 0;
        sBo[k] = 0;
    }
    if (tid == 0) sAbort = sFail = 0;
    __syncthreads();
    if (g == 0 && tid == 0) {  // c = 1, b = 1 (fp.rs:267-270)
        sCo[0] = 1 % p;
        sBo[0] = 1 % p;
        st_dev(Cg[0], 1 % p);
        st_dev(Bg[0], 1 % p);
    }
    uint32_t l = 0, bl = 1, m = 1, cc = 0, bc = 0;
    uint64_t bb = 1 % p;
    unsigned long long bytes = 0;
    for (uint32_t i0 = 0, blk = 0; i0 < n; i0 += K, blk++) {
        const uint32_t Kp = min(K, n - i0);
        if (g == 0 && tid == 0) bm_progress(Q.prog, i0);
        const uint32_t par = blk & 1;
        const uint64_t tag = (uint64_t)(((blk >> 1) + 1) & 7) << 61;
        // ---- 1. this workgroup's parts of Ec(r) and Eb(r)
        // chunks holding a coefficient of C (index <= l) or of B (index < bl)
        const uint32_t top_e = max(l, bl - 1);
        uint32_t nqa = 0;
        while (nqa < nq && (nqa * G + g) * W <= top_e) nqa++;
        // s windows: chunk q at c0 needs s[i0 + r - j], j in [c0, c0+W), r < K, for
        // Ec and the same shifted by m for Eb: u = r - (j - c0) + W - 1 in [0, WIN)
        for (uint32_t k = tid; k < nqa * 2 * WIN; k += NT) {
            const uint32_t q = k / (2 * WIN), rem = k - q * 2 * WIN, which = rem / WIN, u = rem - which * WIN;
            const uint32_t c0 = (q * G + g) * W;
            const int64_t xi = (int64_t)i0 - (int64_t)c0 - (int64_t)(W - 1) + (int64_t)u - (which ? (int64_t)m : 0);
            stage[k] = (xi >= 0 && xi < (int64_t)n) ? s[(uint32_t)xi] : 0;
        }
        __syncthreads();
        {
            const uint32_t v = tid & (V2 - 1), part = tid / V2, which = v / K, r = v & (K - 1);
            constexpr uint32_t OP = W / (NT / V2);  // offsets per part (even)
            const uint64_t *own = which ? sBo : sCo;
            MacSum<MODE> acc;
            for (uint32_t q = 0; q < nqa; q++) {
                const uint64_t *win = stage + q * 2 * WIN + which * WIN + r + W - 1;
                const uint64_t *ow = own + q * W;
#pragma unroll 4
                for (uint32_t o = part * OP; o < part * OP + OP; o += 2)
                    acc.add2(f, ow[o], win[-(int32_t)o], ow[o + 1], win[-(int32_t)o - 1]);
            }
            sRed[part][v] = acc.value(f);
        }
        // every wave's HBM stores of the previous block are complete before publishing
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (tid < V2) {
            uint64_t v = 0;
#pragma unroll
            for (uint32_t k = 0; k < NT / V2; k++) v = add_mod(v, sRed[k][tid], p);
            st_dev(to_global(&x->part[par][tid][g]), v | tag);
        }
        // ---- 2. reduce-scatter: workgroup g adds up values v = g, g + G, ...
        for (uint32_t v = g; v < V2; v += G) {
            uint64_t acc = 0;
            for (uint32_t w = tid; w < G; w += NT) {
                uint64_t y = 0;
                if (!poll_tag(to_global(&x->part[par][v][w]), tag, y, x)) sAbort = 1;
                acc = add_mod(acc, y, p);
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) acc = add_mod(acc, __shfl_xor(acc, off, 64), p);
            if (lane == 0) sWave[wid] = acc;
            __syncthreads();
            if (tid == 0) {
                uint64_t t2 = 0;
                for (uint32_t k = 0; k < NW; k++) t2 = add_mod(t2, sWave[k], p);
                st_dev(to_global(&x->tot[par][v]), t2 | tag);
            }
            __syncthreads();
        }
        // ---- all-gather
        if (tid < V2) {
            uint64_t y = 0;
            if (!poll_tag(to_global(&x->tot[par][tid]), tag, y, x)) sAbort = 1;
            sE[tid] = y;
        }
        if (tid == 0) sProg = 0;
        __syncthreads();
        if (uni32(sAbort) != 0 || uni32(__hip_atomic_load(&x->timeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) != 0)
            break;
        // ---- 3. the K steps (wavefronts 0-3)
        const uint32_t l0 = l, bl0 = bl, m0 = m;
        bm_replay_block<MODE>(f, n, i0, Kp, lane, wid, l, bl, m, bb, bytes, sE, sPoly, sRec, sJunk, &sProg, sState,
                              &sFail);
        __syncthreads();
        if (uni32(sFail) != 0) {
            if (tid == 0) __hip_atomic_store(&x->timeout, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
        }
        l = uni32((uint32_t)sState[0]);
        bl = uni32((uint32_t)sState[1]);
        m = uni32((uint32_t)sState[2]);
        bb = uni64(sState[3]);
        const bool bch = uni32((uint32_t)sState[4]) != 0;
        if (uni32((uint32_t)sState[6]) == 0) continue;  // no update in this block: C and B stand
        // ---- 4. new C (and B) on this workgroup's chunks
        const uint32_t top_a = bch ? max(l, bl - 1) : l;
        uint32_t nqb = 0;
        while (nqb < nq && (nqb * G + g) * W <= top_a) nqb++;
        // windows: old C over [c0 - K + 1, c0 + W) and old x^m0 B over the same range
        for (uint32_t k = tid; k < nqb * 2 * WIN; k += NT) {
            const uint32_t q = k / (2 * WIN), rem = k - q * 2 * WIN, which = rem / WIN, u = rem - which * WIN;
            const int64_t j = (int64_t)((q * G + g) * W) - (int64_t)(K - 1) + (int64_t)u;
            uint64_t y = 0;
            if (which == 0) {
                if (u >= K - 1)
                    y = sCo[q * W + (u - (K - 1))];  // own chunk (zero above l0)
                else if (j >= 0 && j <= (int64_t)l0)
                    y = ld_dev(Cg[cc] + (uint32_t)j);
            } else {
                const int64_t jb = j - (int64_t)m0;
                if (jb >= 0 && jb < (int64_t)bl0) y = ld_dev(Bg[bc] + (uint32_t)jb);
            }
            stage[k] = y;
        }
        __syncthreads();
        // work items (chunk, tap range): with fewer chunks than wavefronts each
        // chunk's K taps are split over wpc wavefronts and the parts added in LDS
        const uint32_t wpc = nqb >= 5 ? 1 : nqb >= 3 ? 2 : nqb == 2 ? 4 : 8;
        const uint32_t taps = K / wpc;
        uint64_t *sAcc = stage + 8 * 2 * WIN;  // free when wpc > 1 (nqb <= 4): [item][2][64]
        for (uint32_t it = wid; it < nqb * wpc; it += NW) {
            const uint32_t q = it / wpc, a0 = (it - q * wpc) * taps;
            const uint64_t *cw = stage + q * 2 * WIN + lane + K - 1;  // cw[-a] = C[j - a]
            const uint64_t *bw = cw + WIN;                            // bw[-a] = B[j - m0 - a]
            MacSum<MODE> cs, bs;
            for (uint32_t a = a0; a < a0 + taps; a++) {
                const uint64_t xc = cw[-(int32_t)a], xb = bw[-(int32_t)a];
                cs.add2(f, sPoly[0][a], xc, sPoly[1][a], xb);
                if (bch) bs.add2(f, sPoly[2][a], xc, sPoly[3][a], xb);
            }
            const uint64_t cv = cs.value(f), bv = bch ? bs.value(f) : 0;
            if (wpc == 1) {
                const uint32_t j = (q * G + g) * W + lane;
                sCo[q * W + lane] = cv;
                if (j < n) st_dev(Cg[cc ^ 1] + j, cv);
                if (bch) {
                    sBo[q * W + lane] = bv;
                    if (j < n) st_dev(Bg[bc ^ 1] + j, bv);
                }
            } else {
                sAcc[it * 128 + lane] = cv;
                sAcc[it * 128 + 64 + lane] = bv;
            }
        }
        if (wpc > 1) {
            __syncthreads();
            for (uint32_t q = wid; q < nqb; q += NW) {
                uint64_t cv = 0, bv = 0;
                for (uint32_t k = 0; k < wpc; k++) {
                    cv = add_mod(cv, sAcc[(q * wpc + k) * 128 + lane], p);
                    bv = add_mod(bv, sAcc[(q * wpc + k) * 128 + 64 + lane], p);
                }
                const uint32_t j = (q * G + g) * W + lane;
                sCo[q * W + lane] = cv;
                if (j < n) st_dev(Cg[cc ^ 1] + j, cv);
                if (bch) {
                    sBo[q * W + lane] = bv;
                    if (j < n) st_dev(Bg[bc ^ 1] + j, bv);
                }
            }
        }
        cc ^= 1;
        if (bch) bc ^= 1;
        __syncthreads();
    }
    if (g == 0 && tid == 0) {
        Q.lc = l;
        Q.bytes = (unsigned long long)sState[5];
    }
}

// ---------------------------------------------------------------------------
// k_bm_batch_blk: the blocked form of k_bm_block inside one workgroup per
// sequence (batches).  A block of K steps costs one pass over c and b for the E
// sums, the replay of the K steps (bm_replay_block) and one pass forming the new C
// and B, with no barrier between steps: about 6 products per coefficient and step
// where k_bm_batch needs 3, but on independent data, not a latency chain.  C and B
// live in HBM, double-buffered (one pass reads one copy and writes the other);
// both passes run over the coefficients in rounds, staged through LDS.  For
// p = 2^61 - 1 both passes are signed-digit int8 products on the matrix cores
// (bm_esums_mfma_m61, bm_apply_mfma_m61); other moduli use lazy vector sums
// (MacSum).  Needs p < 2^62 (the replay records hold the step's kind in bits 62-63).
// ---------------------------------------------------------------------------
// x 2^s mod 2^61 - 1 for x < 2^42, 0 <= s < 61: the bits shifted past 2^61 wrap
// around (2^61 == 1); result < 2^61 + 2^42, congruent, not reduced.
__device__ __forceinline__ uint64_t rot_m61(uint64_t x, uint32_t s) {
    return s == 0 ? x : ((x << s) & P61) + (x >> (61 - s));
}

constexpr uint32_t BM_RA = 1024;           // outputs per round of the matrix-core application
constexpr uint32_t BM_PW = BM_RA + 128;    // digit plane length of a window round
constexpr uint32_t BM_RP = 160;            // reversed polynomial digit plane length

// The eight signed digits (in [-128, 127], least significant first) of the balanced
// representative of x mod 2^61 - 1 as the bytes of one word: with c = 0x80 in every
// byte, balanced(x) + c has byte u equal to digit u + 128, so xor c gives digit u in
// two's complement.  Equal to peeling the low byte as an int8 eight times.
__device__ __forceinline__ uint64_t digits8_m61(uint64_t x) {
    constexpr uint64_t c = 0x8080808080808080ull;
    return ((uint64_t)balanced(x, P61) + c) ^ c;
}

// Eight values' digit words d[t] -> eight plane words: byte t of plane word u is byte u
// of d[t] (an 8x8 byte transpose, two 4x4 transposes per half), stored at
// dst[u * stride + k .. + 8] (k and stride multiples of 8).
__device__ __forceinline__ void store_planes8(int8_t *dst, uint32_t stride, uint32_t k, const uint64_t d[8]) {
    uint32_t q[2][2][4];  // [half of the digits][group of four values][digit within the half]
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int g = 0; g < 2; g++) {
            const uint32_t a = (uint32_t)(d[4 * g] >> (32 * h)), b = (uint32_t)(d[4 * g + 1] >> (32 * h)),
                           c = (uint32_t)(d[4 * g + 2] >> (32 * h)), e = (uint32_t)(d[4 * g + 3] >> (32 * h));
            const uint32_t t0 = __byte_perm(a, b, 0x5140), t1 = __byte_perm(a, b, 0x7362);  // a0 b0 a1 b1 | a2 b2 a3 b3
            const uint32_t t2 = __byte_perm(c, e, 0x5140), t3 = __byte_perm(c, e, 0x7362);
            q[h][g][0] = __byte_perm(t0, t2, 0x5410);  // a0 b0 c0 e0
            q[h][g][1] = __byte_perm(t0, t2, 0x7632);
            q[h][g][2] = __byte_perm(t1, t3, 0x5410);
            q[h][g][3] = __byte_perm(t1, t3, 0x7632);
        }
#pragma unroll
    for (int u = 0; u < 8; u++)
        *(uint64_t *)(dst + u * stride + k) = (uint64_t)q[u >> 2][0][u & 3] | ((uint64_t)q[u >> 2][1][u & 3] << 32);
}

// Four values' digit words d[t] -> eight plane words of four bytes (byte t of word u
// is byte u of d[t]; the 4x4 transposes of store_planes8), stored at
// dst[u * stride + k .. + 4] (k and stride multiples of 4).
__device__ __forceinline__ void store_planes4(int8_t *dst, uint32_t stride, uint32_t k, const uint64_t d[4]) {
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const uint32_t a = (uint32_t)(d[0] >> (32 * h)), b = (uint32_t)(d[1] >> (32 * h)),
                       c = (uint32_t)(d[2] >> (32 * h)), e = (uint32_t)(d[3] >> (32 * h));
        const uint32_t t0 = __byte_perm(a, b, 0x5140), t1 = __byte_perm(a, b, 0x7362);
        const uint32_t t2 = __byte_perm(c, e, 0x5140), t3 = __byte_perm(c, e, 0x7362);
        *(uint32_t *)(dst + (4 * h + 0) * stride + k) = __byte_perm(t0, t2, 0x5410);
        *(uint32_t *)(dst + (4 * h + 1) * stride + k) = __byte_perm(t0, t2, 0x7632);
        *(uint32_t *)(dst + (4 * h + 2) * stride + k) = __byte_perm(t1, t3, 0x5410);
        *(uint32_t *)(dst + (4 * h + 3) * stride + k) = __byte_perm(t1, t3, 0x7632);
    }
}

// The application of k_bm_batch_blk for p = 2^61 - 1 on the matrix cores.
// With balanced 8-digit expansions W^ = sum_u 256^u W_u of the window values and
// P^ = sum_v 256^v P_v of the polynomial coefficients,
//   c_new[j] = sum_{a<64} U[a] C[j-a] + V[a] B[j-m0-a] == sum_s 256^s Q_s[j],
//   Q_s[j] = sum_{u+v=s} sum_a C_u[j-a] U_v[a] + B_u[j-m0-a] V_v[a],
// and b_new alike with Ub, Vb.  A 16 x 16 output tile holds 256 consecutive
// outputs j = j0 + 256 t + 16 r + c; with x = c - a it is the product of the
// Hankel rows A[r][x] = W_u[j0 + 256 t + 16 r + x] (x in [-64, 64), two K chunks of
// 64, read as aligned 16-byte runs of a forward digit plane) and the Toeplitz
// columns B[x][c] = P_v[c - x] (16-byte runs of a reversed, zero-padded digit plane
// of the polynomial).  The MFMAs accumulate straight into the 15 classes s = u + v:
// |Q_s| <= 8 pairs * 2 windows * 64 taps * 2^14 = 2^24, exact in int32.  Epilogue:
// three classes at a time, T_k = Q_3k + 2^8 Q_3k+1 + 2^16 Q_3k+2 (|T_k| < 2^41, exact
// in int64), T_k + 2^41 rotated by 24 k bits mod p and summed (5 rotations, not 15);
// the offsets' total koff is subtracted.  Wave w forms tile w & 3 of c_new (w < 4)
// or of b_new (w >= 4).
__device__ void bm_apply_mfma_m61(uint32_t tid, uint32_t lane, uint32_t wid, uint32_t n, uint32_t top_a, uint32_t l0,
                                  uint32_t bl0, uint32_t m0, bool bch, gptr<const uint64_t> cold,
                                  gptr<const uint64_t> bold, gptr<uint64_t> cnew, gptr<uint64_t> bnew,
                                  const uint64_t (*sPoly)[BM_BK], int8_t (*sPl)[8][BM_PW],
                                  int8_t (*sRP)[8][BM_RP], uint64_t koff) {
    typedef int v4i __attribute__((ext_vector_type(4)));
    constexpr uint32_t NT = BM_BLOCK_THREADS;
    __asm__ volatile("" : "+v"(lane), "+v"(tid));  // lane constants per call (as in the E sums)
    // this thread's eight window values of a round (threads < 2 BM_PW / 8): digit
    // planes sPl[w][u][y .. y + 8] of C[j0 - 64 + y] (w = 0, zero above l0) or of
    // B[j0 - 64 + y - m0] (w = 1, zero at or above bl0); fetched a round ahead
    uint32_t st = tid;
    __asm__ volatile("" : "+v"(st));  // the staging indices per call, not hoisted out of the step loop
    const uint32_t sw = st < BM_PW / 8 ? 0 : st < 2 * BM_PW / 8 ? 1 : 2, sy = 8 * (st % (BM_PW / 8));
    auto fetch = [&](uint32_t j0, uint64_t raw[8]) {
#pragma unroll
        for (uint32_t t = 0; t < 8; t++) raw[t] = 0;
        const int64_t x0 = (int64_t)j0 - 64 + (int64_t)sy;
        if (sw == 0) {
#pragma unroll
            for (uint32_t t = 0; t < 8; t++) {
                const int64_t x = x0 + (int64_t)t;
                if (x >= 0 && x <= (int64_t)l0) raw[t] = cold[(uint32_t)x];
            }
        } else if (sw == 1) {
#pragma unroll
            for (uint32_t t = 0; t < 8; t++) {
                const int64_t x = x0 + (int64_t)t - (int64_t)m0;
                if (x >= 0 && x < (int64_t)bl0) raw[t] = bold[(uint32_t)x];
            }
        }
    };
    uint64_t raw[8];
    fetch(0, raw);
    // reversed polynomial digit planes: sRP[poly][v][79 - a] = digit v of poly[a]
    // (polys U, V, Ub, Vb), zero outside a in [0, 64); thread k < 32 takes a = 8 (k & 7) ..
    // + 8 of poly k >> 3, stored reversed at 72 - 8 (k & 7)
    for (uint32_t k = tid; k < 4 * 8 * BM_RP / 4; k += NT) ((int32_t *)&sRP[0][0][0])[k] = 0;
    __syncthreads();
    if (st < 4 * BM_BK / 8) {
        const uint32_t poly = st >> 3, a0 = 8 * (st & 7);
        uint64_t d[8];
#pragma unroll
        for (uint32_t t = 0; t < 8; t++) d[t] = digits8_m61(sPoly[poly][a0 + 7 - t]);
        store_planes8(&sRP[poly][0][0], BM_RP, 72 - a0, d);
    }
    const uint32_t col = lane & 15, kq = lane >> 4, tau = wid & 3, pp = wid >> 2;
    const v4i zero = {0, 0, 0, 0};
    for (uint32_t j0 = 0; j0 <= top_a; j0 += BM_RA) {
        if (sw < 2) {
            uint64_t d[8];
#pragma unroll
            for (uint32_t t = 0; t < 8; t++) d[t] = digits8_m61(raw[t]);
            store_planes8(&sPl[sw][0][0], BM_PW, sy, d);
        }
        __syncthreads();
        if (j0 + BM_RA <= top_a) fetch(j0 + BM_RA, raw);  // in flight during the products
        if (pp == 0 || bch) {
            v4i acc[15];
#pragma unroll
            for (int q = 0; q < 15; q++) acc[q] = zero;
            const uint32_t abase = 256 * tau + 16 * col + 16 * kq;  // A row col (= r), K bytes 16 kq ..
            const uint32_t bbase = 15 - col + 16 * kq;               // B column col
            const uint32_t bsh = bbase & 3;
            // [window][K chunk]; bytes bbase .. + 16 (bbase mod 4 is fixed per lane)
            // from five aligned dwords, shifted
            auto run = [&](const int8_t *row) {
                const uint32_t *q = (const uint32_t *)(row - bsh);
                const uint32_t d0 = q[0], d1 = q[1], d2 = q[2], d3 = q[3], d4 = q[4];
                v4i r;
                r[0] = (int)__builtin_amdgcn_alignbyte(d1, d0, bsh);
                r[1] = (int)__builtin_amdgcn_alignbyte(d2, d1, bsh);
                r[2] = (int)__builtin_amdgcn_alignbyte(d3, d2, bsh);
                r[3] = (int)__builtin_amdgcn_alignbyte(d4, d3, bsh);
                return r;
            };
            // the round's A runs (the windows' Hankel rows) of this lane, loaded once
            v4i a[8][4];  // [digit u][window * 2 + K chunk]
#pragma unroll
            for (int u = 0; u < 8; u++) {
                __builtin_memcpy(&a[u][0], &sPl[0][u][abase], 16);
                __builtin_memcpy(&a[u][1], &sPl[0][u][abase + 64], 16);
                __builtin_memcpy(&a[u][2], &sPl[1][u][abase], 16);
                __builtin_memcpy(&a[u][3], &sPl[1][u][abase + 64], 16);
            }
#pragma unroll
            for (int v = 0; v < 8; v++) {
                __asm__ volatile("" ::: "memory");  // one digit's B runs at a time
                v4i b[4];  // [window * 2 + K chunk]
                b[0] = run(&sRP[2 * pp][v][bbase]);
                b[1] = run(&sRP[2 * pp][v][bbase + 64]);
                b[2] = run(&sRP[2 * pp + 1][v][bbase]);
                b[3] = run(&sRP[2 * pp + 1][v][bbase + 64]);
#pragma unroll
                for (int u = 0; u < 8; u++)
#pragma unroll
                    for (int x = 0; x < 4; x++)
                        acc[u + v] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[u][x], b[x], acc[u + v], 0, 0, 0);
            }
            // register i of this lane: row 4 kq + i of the tile, column col
#pragma unroll
            for (int i = 0; i < 4; i++) {
                uint64_t w = 0;  // five terms below 2^61 + 2^42 each
#pragma unroll
                for (int k = 0; k < 5; k++) {
                    const int64_t t = (int64_t)acc[3 * k][i] + ((int64_t)acc[3 * k + 1][i] << 8) +
                                      ((int64_t)acc[3 * k + 2][i] << 16);
                    w += rot_m61((uint64_t)(t + (1ll << 41)), (24 * k) % 61);
                }
                const uint64_t val = sub_mod(canon_m61(fold_m61_64(w)), koff, P61);
                const uint32_t j = j0 + 256 * tau + 16 * (4 * kq + i) + col;
                if (j < n) {
                    if (pp == 0)
                        cnew[j] = val;
                    else
                        bnew[j] = val;
                }
            }
        }
        __syncthreads();
    }
}

constexpr uint32_t BM_RE = 512;   // coefficients per round of the matrix-core E sums
constexpr uint32_t BM_EKC = 9;    // K chunks of 64 per round: s positions k < 576 (511 + 16 lags)
constexpr uint32_t BM_ESF = 624;  // forward s digit plane length of a round (48 + 576)
constexpr uint32_t BM_ECR = 608;  // reversed C / D digit plane length (16 + 512 + zero tail)

// The E sums of k_bm_batch_blk for p = 2^61 - 1 on the matrix cores.  With
// D = x^m B (D[j] = B[j - m]) both sums read the same window of s:
//   Ec(r) = sum_j C[j] s[i0+r-j],   Eb(r) = sum_j D[j] s[i0+r-j],
// and with signed base-256 digits Ec(r) == sum_{u,v} 256^(u+v) T_uv(r).  A round
// takes the 512 coefficients j = j0 + jj; with x0 = i0 - j0 - 511 and the s position
// k = 511 + c - jj (k < 576, nine K chunks of 64), a lag r = 16 g + c is
//   T_uv(r) = sum_k S_v[x0 + 16 g + k] C_u[j0 + 511 + c - k].
// One MFMA per (digit u of C or D, group of four digits v of s, K chunk): rows
// 4 g + vv hold the Hankel rows S_(4 h + vv)[x0 + 16 g + k] of the forward s digit
// planes (aligned 16-byte runs), columns c the Toeplitz C_u[j0 + 511 + c - k] of a
// reversed, zero-padded digit plane of C (or D): per lane a 16-byte run at a
// lane-fixed byte offset mod 4, put together from five aligned dwords with byte
// shifts.  Output register i of lane (c, g) is lag 16 g + c, digit v = 4 h + i.
// Wave w takes C (w < 4) or D, the digits u in [4 ((w >> 1) & 1), + 4), both groups
// h, and the K chunks [0, 5) (w even) or [5, 9): each A run feeds four MFMAs and each
// B run two.  |T| grows by at most 2^23 a round; every 8 rounds (|T| <= 2^26) and at
// the end the T of equal u + v in a lane (four at most, |sum| <= 2^28) go into a
// residue: plus 2^28, rotated by 8 (u + v) bits mod p, summed, minus the offsets; the
// four waves of C (of D) meet in LDS.  Rounds holding no coefficient of C (j <= l)
// nor of D (m <= j < m + bl) are skipped, and the next round's values are fetched
// from HBM during the current round's products.
__device__ void bm_esums_mfma_m61(uint32_t tid, uint32_t lane, uint32_t wid, uint32_t n, uint32_t i0, uint32_t l,
                                  uint32_t bl, uint32_t m, gptr<const uint64_t> s, gptr<const uint64_t> cp,
                                  gptr<const uint64_t> bp, int8_t *pl, uint64_t *sE) {
    typedef int v4i __attribute__((ext_vector_type(4)));
    constexpr uint32_t RE = BM_RE, KC = BM_EKC, SF = BM_ESF, CR = BM_ECR;
    // the lane constants below are formed anew per call: hoisted out of the caller's
    // loop they would stay live through the application's products
    __asm__ volatile("" : "+v"(lane), "+v"(tid));
    int8_t *CRp = pl, *SFp = pl + 16 * CR;  // CRp[16][CR]: planes of C, then of D; SFp[8][SF]
    const uint32_t which = wid >> 2, uh = (wid >> 1) & 1, kh = wid & 1, c = lane & 15, g = lane >> 4;
    const uint32_t kc0 = kh ? 5 : 0, kc1 = kh ? KC : 5;
    // A: row lane & 15 = 4 g' + vv of the tile, K bytes 16 (lane >> 4) ..
    const uint32_t arow = 16 * ((lane & 15) >> 2) + 16 * (lane >> 4), avv = lane & 3;
    // B: column c, K bytes 16 (lane >> 4) ..: reversed plane index 16 + k - c
    const uint32_t bz = 16 + 16 * (lane >> 4) - c, sh = bz & 3;
    const v4i zero = {0, 0, 0, 0};
    // this lane's offsets per fold: 2^28 256^(4 uh + cr) for its 11 classes cr
    uint64_t koff = 0;
    for (uint32_t cr = 0; cr < 11; cr++) koff = add_mod(koff, canon_m61(rot_m61(1ull << 28, (8 * (4 * uh + cr)) % 61)), P61);
    uint64_t epart = 0;
    v4i acc[4][2];  // [digit u - 4 uh][group h of digits of s]
#pragma unroll
    for (int ul = 0; ul < 4; ul++) acc[ul][0] = acc[ul][1] = zero;
    uint32_t folds = 0;  // rounds in acc since the last fold (at most 8: |T| <= 2^26)
    auto fold = [&]() {  // output register i: digit v = 4 h + i; class cr = ul + 4 h + i
        int32_t cl[11];
#pragma unroll
        for (int cr = 0; cr < 11; cr++) cl[cr] = 0;
#pragma unroll
        for (int ul = 0; ul < 4; ul++)
#pragma unroll
            for (int h = 0; h < 2; h++)
#pragma unroll
                for (int i = 0; i < 4; i++) cl[ul + 4 * h + i] += acc[ul][h][i];
        uint64_t w = 0;
#pragma unroll
        for (uint32_t cr = 0; cr < 11; cr++)
            w = fold_m61_64(w + rot_m61((uint64_t)(uint32_t)(cl[cr] + (1 << 28)), (8 * (4 * uh + cr)) % 61));
        epart = add_mod(epart, sub_mod(canon_m61(w), koff, P61), P61);
#pragma unroll
        for (int ul = 0; ul < 4; ul++) acc[ul][0] = acc[ul][1] = zero;
    };
    const uint32_t top_e = max(l, m + bl - 1);
    // rounds holding no coefficient of C (j <= l) nor of D (m <= j < m + bl) are skipped
    auto next_round = [&](uint32_t j) {
        while (j <= top_e && j > l && j + RE <= m) j += RE;
        return j;
    };
    // this thread's four values of a round: reversed C (tid < 128) or D (tid < 256)
    // digits at z = 4 (tid & 127) .., value j = j0 + 511 - z; or forward s digits
    // (tid < 412) at y = 4 (tid - 256) .., value s[x0 + y]; fetched a round ahead
    uint32_t st = tid;
    __asm__ volatile("" : "+v"(st));
    const uint32_t sw = st < RE / 4 ? 0 : st < 2 * RE / 4 ? 1 : st < 2 * RE / 4 + SF / 4 ? 2 : 3;
    const uint32_t sk = sw < 2 ? 4 * (st % (RE / 4)) : 4 * (st - 2 * RE / 4);
    auto fetch = [&](uint32_t j0, uint64_t raw[4]) {
#pragma unroll
        for (uint32_t t = 0; t < 4; t++) raw[t] = 0;
        if (sw == 0) {
#pragma unroll
            for (uint32_t t = 0; t < 4; t++) {
                const uint32_t j = j0 + RE - 1 - (sk + t);
                if (j <= l) raw[t] = cp[j];
            }
        } else if (sw == 1) {
#pragma unroll
            for (uint32_t t = 0; t < 4; t++) {
                const uint32_t j = j0 + RE - 1 - (sk + t);
                if (j >= m && j - m < bl) raw[t] = bp[j - m];
            }
        } else if (sw == 2) {
            const int64_t x0 = (int64_t)i0 - (int64_t)j0 - (int64_t)(RE - 1) + (int64_t)sk;
#pragma unroll
            for (uint32_t t = 0; t < 4; t++) {
                const int64_t x = x0 + (int64_t)t;
                if (x >= 0 && x < (int64_t)n) raw[t] = s[(uint32_t)x];
            }
        }
    };
    // the reversed planes' zero margins: bytes [0, 16) and [16 + RE, CR) of each
    for (uint32_t k = tid; k < 16 * (CR - RE) / 4; k += BM_BLOCK_THREADS) {
        const uint32_t pl_ = k / ((CR - RE) / 4), q = k - pl_ * ((CR - RE) / 4);
        const uint32_t off = q < 4 ? 4 * q : RE + 4 * q;  // 16 front bytes, then the tail from 16 + RE
        *(uint32_t *)(CRp + pl_ * CR + off) = 0;
    }
    uint64_t raw[4];
    uint32_t j0 = next_round(0);
    fetch(j0, raw);
    while (j0 <= top_e) {
        if (sw < 3) {
            uint64_t d[4];
#pragma unroll
            for (uint32_t t = 0; t < 4; t++) d[t] = digits8_m61(raw[t]);
            if (sw < 2)
                store_planes4(CRp + sw * 8 * CR, CR, 16 + sk, d);
            else
                store_planes4(SFp, SF, sk, d);
        }
        __syncthreads();
        const uint32_t jn = next_round(j0 + RE);
        if (jn <= top_e) fetch(jn, raw);  // in flight during the products
        const int8_t *cplanes = CRp + (which * 8 + 4 * uh) * CR;
        // at most five chunks per wave: unrolled, so the loads of later chunks can
        // be in flight during earlier products
#pragma unroll
        for (uint32_t kk = 0; kk < 5; kk++) {
            const uint32_t kc = kc0 + kk;
            if (kc >= kc1) break;  // uniform
            v4i a[2];
#pragma unroll
            for (int h = 0; h < 2; h++) __builtin_memcpy(&a[h], SFp + (4 * h + avv) * SF + arow + 64 * kc, 16);
#pragma unroll
            for (uint32_t ul = 0; ul < 4; ul++) {
                // bytes 16 + 64 kc + 16 (lane >> 4) - c .. + 16 of plane u: five aligned dwords
                const uint32_t *q = (const uint32_t *)(cplanes + ul * CR + bz + 64 * kc - sh);
                const uint32_t d0 = q[0], d1 = q[1], d2 = q[2], d3 = q[3], d4 = q[4];
                v4i bf;
                bf[0] = (int)__builtin_amdgcn_alignbyte(d1, d0, sh);
                bf[1] = (int)__builtin_amdgcn_alignbyte(d2, d1, sh);
                bf[2] = (int)__builtin_amdgcn_alignbyte(d3, d2, sh);
                bf[3] = (int)__builtin_amdgcn_alignbyte(d4, d3, sh);
#pragma unroll
                for (int h = 0; h < 2; h++) acc[ul][h] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[h], bf, acc[ul][h], 0, 0, 0);
            }
        }
        if (++folds == 8) {
            fold();
            folds = 0;
        }
        __syncthreads();
        j0 = jn;
    }
    fold();
    // lane (c, g): lag 16 g + c of Ec (waves 0-3) or Eb (waves 4-7); the four waves of
    // each meet in LDS
    uint64_t *red = (uint64_t *)pl;  // the planes are free after the last round's barrier
    red[(wid & 3) * 2 * BM_BK + which * BM_BK + 16 * g + c] = epart;
    __syncthreads();
    if (tid < 2 * BM_BK)
        sE[tid] = add_mod(add_mod(red[tid], red[2 * BM_BK + tid], P61),
                          add_mod(red[4 * BM_BK + tid], red[6 * BM_BK + tid], P61), P61);
}

template <int MODE>
__global__ __launch_bounds__(BM_BLOCK_THREADS) void k_bm_batch_blk(BmSeq *seqs, FieldParams f) {
    constexpr uint32_t NT = BM_BLOCK_THREADS, K = BM_BK, V2 = 2 * K, R = NT, SW = R + K - 1, NP = NT / V2;
    BmSeq &Q = seqs[blockIdx.x];
    const uint32_t n = Q.n, tid = threadIdx.x, lane = tid & 63, wid = uni32(tid >> 6);
    const gptr<const uint64_t> s = to_global(Q.s);
    const gptr<uint64_t> Cg[2] = {to_global(Q.c), to_global(Q.t)}, Bg[2] = {to_global(Q.b), to_global(Q.u)};
    const uint64_t p = f.p;
    __shared__ uint64_t sCr[R], sBr[R];  // E pass: this round's coefficients of C and B
    __shared__ uint64_t sW[2][SW];       // E pass: s windows; application: windows of C and x^m B
    __shared__ uint64_t sPoly[4][K];
    __shared__ uint64_t sRed[NP][V2];
    __shared__ uint64_t sE[V2];
    __shared__ uint64_t sRec[K];
    __shared__ uint64_t sJunk[2 * 64];
    __shared__ uint64_t sState[7];
    __shared__ uint32_t sProg, sFail;
    // matrix-core application (2^61 - 1): digit planes of the windows and of U|V|Ub|Vb
    __shared__ __attribute__((aligned(16))) int8_t sPl[MODE == MODE_M61 ? 2 : 1][8][MODE == MODE_M61 ? BM_PW : 16];
    __shared__ __attribute__((aligned(16))) int8_t sRPl[MODE == MODE_M61 ? 4 : 1][8][MODE == MODE_M61 ? BM_RP : 16];
    uint64_t koff = 0;  // sum over class triples k < 5 of 2^41 2^(24 k) mod p (the epilogue's offsets)
    if constexpr (MODE == MODE_M61) {
        for (uint32_t k = 0; k < 5; k++) koff = add_mod(koff, canon_m61(rot_m61(1ull << 41, (24 * k) % 61)), P61);
    }

    if (tid == 0) {  // c = 1, b = 1 (fp.rs:267-270); coefficients above l (of b: at or above bl) are never read
        Cg[0][0] = 1 % p;
        Bg[0][0] = 1 % p;
        sFail = 0;
    }
    __syncthreads();
    uint32_t l = 0, bl = 1, m = 1, cc = 0, bc = 0;
    uint64_t bb = 1 % p;
    unsigned long long bytes = 0;
    for (uint32_t i0 = 0; i0 < n; i0 += K) {
        const uint32_t Kp = min(K, n - i0);
        if (tid == 0) bm_progress(Q.prog, i0);
        // ---- E sums: thread (v, part), v = value (Ec(r), then Eb(r)), part = a
        // quarter of each round's coefficients; the window index of coefficient o
        // of the round for lag r is r - o + R - 1.
        if constexpr (MODE == MODE_M61) {
            bm_esums_mfma_m61(tid, lane, wid, n, i0, l, bl, m, s, Cg[cc], Bg[bc], &sPl[0][0][0], sE);
        } else {
            const uint32_t v = tid & (V2 - 1), part = tid / V2, which = v / K, r = v & (K - 1);
            const uint64_t *own = which ? sBr : sCr;
            const uint64_t *win = sW[which] + r + R - 1;
            MacSum<MODE> acc;
            const uint32_t top_e = max(l, bl - 1);
            for (uint32_t j0 = 0; j0 <= top_e; j0 += R) {
                for (uint32_t k = tid; k < R; k += NT) {
                    const uint32_t j = j0 + k;
                    sCr[k] = j <= l ? Cg[cc][j] : 0;
                    sBr[k] = j < bl ? Bg[bc][j] : 0;
                }
                for (uint32_t k = tid; k < 2 * SW; k += NT) {
                    const uint32_t w = k / SW, u = k - w * SW;
                    const int64_t xi = (int64_t)i0 - (int64_t)j0 - (int64_t)(R - 1) + (int64_t)u - (w ? (int64_t)m : 0);
                    sW[w][u] = (xi >= 0 && xi < (int64_t)n) ? s[(uint32_t)xi] : 0;
                }
                __syncthreads();
#pragma unroll 4
                for (uint32_t o = part * (R / NP); o < (part + 1) * (R / NP); o += 2)
                    acc.add2(f, own[o], win[-(int32_t)o], own[o + 1], win[-(int32_t)o - 1]);
                __syncthreads();
            }
            sRed[part][v] = acc.value(f);
            __syncthreads();
            if (tid < V2) {
                uint64_t v = 0;
#pragma unroll
                for (uint32_t k = 0; k < NP; k++) v = add_mod(v, sRed[k][tid], p);
                sE[tid] = v;
            }
        }
        if (tid == 0) sProg = 0;
        __syncthreads();
        // ---- the K steps (wavefronts 0-3)
        const uint32_t l0 = l, bl0 = bl, m0 = m;
        bm_replay_block<MODE>(f, n, i0, Kp, lane, wid, l, bl, m, bb, bytes, sE, sPoly, sRec, sJunk, &sProg, sState,
                              &sFail);
        __syncthreads();
        if (uni32(sFail) != 0) {  // a record wait ran out: stop, the host reports it
            if (tid == 0) Q.timeout = 1;
            return;
        }
        l = uni32((uint32_t)sState[0]);
        bl = uni32((uint32_t)sState[1]);
        m = uni32((uint32_t)sState[2]);
        bb = uni64(sState[3]);
        const bool bch = uni32((uint32_t)sState[4]) != 0;
        if (uni32((uint32_t)sState[6]) == 0) continue;  // no update in this block: C and B stand
        // ---- new C (and B): thread tid forms coefficient j0 + tid of each round
        // from the windows C[j - a], B[j - m0 - a], a < K
        const uint32_t top_a = bch ? max(l, bl - 1) : l;
        if constexpr (MODE == MODE_M61) {
            bm_apply_mfma_m61(tid, lane, wid, n, top_a, l0, bl0, m0, bch, Cg[cc], Bg[bc], Cg[cc ^ 1], Bg[bc ^ 1], sPoly,
                              sPl, sRPl, koff);
        } else
        for (uint32_t j0 = 0; j0 <= top_a; j0 += R) {
            for (uint32_t k = tid; k < 2 * SW; k += NT) {
                const uint32_t w = k / SW, u = k - w * SW;
                const int64_t j = (int64_t)j0 - (int64_t)(K - 1) + (int64_t)u - (w ? (int64_t)m0 : 0);
                uint64_t y = 0;
                if (w == 0) {
                    if (j >= 0 && j <= (int64_t)l0) y = Cg[cc][(uint32_t)j];
                } else {
                    if (j >= 0 && j < (int64_t)bl0) y = Bg[bc][(uint32_t)j];
                }
                sW[w][u] = y;
            }
            __syncthreads();
            const uint64_t *cw = sW[0] + tid + K - 1, *bw = sW[1] + tid + K - 1;
            MacSum<MODE> cs, bs;
            for (uint32_t a = 0; a < K; a++) {
                const uint64_t xc = cw[-(int32_t)a], xb = bw[-(int32_t)a];
                cs.add2(f, sPoly[0][a], xc, sPoly[1][a], xb);
                if (bch) bs.add2(f, sPoly[2][a], xc, sPoly[3][a], xb);
            }
            const uint32_t j = j0 + tid;
            if (j < n) {
                Cg[cc ^ 1][j] = cs.value(f);
                if (bch) Bg[bc ^ 1][j] = bs.value(f);
            }
            __syncthreads();
        }
        cc ^= 1;
        if (bch) bc ^= 1;
    }
    if (tid == 0) {
        Q.lc = l;
        Q.bytes = (unsigned long long)sState[5];
    }
}

static size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

__global__ void k_widen_seq(const uint32_t *src, uint64_t *dst, size_t count) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// Waits for the context's stream while reporting the steps a running kernel
// has stored to the pinned progress word (monotone, at most `total`).
static int wait_progress(modlin_ctx *ctx, uint64_t total, modlin_progress_fn progress, void *user) {
    const volatile uint64_t *w = ctx->prog_host;
    uint64_t last = 0;
    for (;;) {
        const hipError_t e = hipStreamQuery(ctx->stream);
        if (e == hipSuccess) return MODLIN_OK;
        if (e != hipErrorNotReady) return hip_fail(ctx, e, "hipStreamQuery");
        const uint64_t v = *w;
        if (v > last && v <= total) {
            last = v;
            progress(user, v, total);
        }
        std::this_thread::sleep_for(std::chrono::microseconds(200));
    }
}

// Words of a device area set to zero (a fill kernel rather than a memset node,
// so that the same code is safe inside captured graphs).
__global__ void k_zero_words(uint64_t *p, size_t words) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < words; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// Runs the batch described by the device array ds (host copy hs, count items,
// the longest maxn; item k is the caller's sequence ord[k]) with the kernel the
// batch shape selects, and stores the linear complexities in lcs.  exch is the
// cooperative kernels' exchange area (count == 1); rd(dst, src, bytes) copies
// device memory to the host and waits for the stream.
template <class Rd>
static int run_bm(modlin_ctx *ctx, const FieldParams &f, uint64_t p, BmSeq *ds, BmSeq *hs, uint64_t count,
                  uint64_t maxn, char *exch, const modlin_opts_t *opts, modlin_progress_fn progress, void *user,
                  const uint64_t *ord, uint64_t *lcs, Rd rd) {
    hipStream_t st = ctx->stream;
    const bool profile = opts && (opts->flags & MODLIN_FLAG_PROFILE);
    int rc = MODLIN_OK;
    if (profile) MDL_HIP(hipEventRecord(ctx->evA, st));
    // One long sequence: the cooperative kernel over the whole GPU (one workgroup
    // per CU, at most 1024 partials).
    const uint32_t coop_min = (opts && opts->coop_min) ? opts->coop_min : BM_COOP_MIN;
    if (count == 1 && maxn >= coop_min) {
        const modlin_tuning_t tu = select_tuning(ctx->tune_table, &ctx->tune_override, p, maxn, count, true);
        const uint32_t G0 = tu.bm_coop_wgs ? tu.bm_coop_wgs : (uint32_t)std::min(ctx->cu_count, 1024);
        // blocked kernel: p < 2^61 (tagged words) and at most BM_BQ chunks per workgroup
        const uint64_t nchunks = (maxn + BM_BW - 1) / BM_BW;
        const uint32_t Gb = (uint32_t)std::max<uint64_t>(G0, (nchunks + BM_BQ - 1) / BM_BQ);
        if (p <= P61 && Gb <= 1024) {
            const void *kb = f.mode == MODE_M61   ? (const void *)k_bm_block<MODE_M61>
                             : f.mode == MODE_M31 ? (const void *)k_bm_block<MODE_M31>
                                                  : (const void *)k_bm_block<MODE_GENERIC>;
            int per_cu = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kb, BM_BLOCK_THREADS, 0) != hipSuccess) per_cu = 0;
            if ((uint64_t)Gb <= (uint64_t)per_cu * (uint64_t)std::max(ctx->cu_count, 0)) {
                BmBlockX *bx = (BmBlockX *)exch;
                hipLaunchKernelGGL(k_zero_words, dim3(256), dim3(256), 0, st, (uint64_t *)bx, sizeof(BmBlockX) / 8);
                void *bargs[] = {(void *)&ds, (void *)&f, (void *)&bx};
                if (hipLaunchCooperativeKernel(kb, dim3(Gb), dim3(BM_BLOCK_THREADS), bargs, 0, st) == hipSuccess) {
                    if (profile) MDL_HIP(hipEventRecord(ctx->evB, st));
                    if (progress && (rc = wait_progress(ctx, maxn, progress, user))) return rc;
                    uint32_t timeout = 0;
                    rc = rd(&timeout, &bx->timeout, sizeof(timeout));
                    if (rc) return rc;
                    if (timeout) {
                        ctx->err = "blocked cooperative Berlekamp-Massey: an exchange wait ran out";
                        return MODLIN_EHIP;
                    }
                    rc = rd(hs, ds, sizeof(BmSeq));
                    if (rc) return rc;
                    lcs[0] = hs[0].lc;
                    ctx->stats.bm_bytes = (double)hs[0].bytes;
                    ctx->stats.bm_coop = (double)Gb;
                    if (profile) {
                        float ms = 0;
                        MDL_HIP(hipEventElapsedTime(&ms, ctx->evA, ctx->evB));
                        ctx->stats.bm_ms = ms;
                    }
                    return MODLIN_OK;
                }
                (void)hipGetLastError();  // refused: the kernels below give the same result
            }
        }
        BmCoop *cs = (BmCoop *)exch;
        hipLaunchKernelGGL(k_zero_words, dim3(16), dim3(256), 0, st, (uint64_t *)cs, sizeof(BmCoop) / 8);
        const uint32_t G = G0;
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
        if (progress && (rc = wait_progress(ctx, maxn, progress, user))) return rc;
        uint32_t timeout = 0;
        rc = rd(&timeout, &cs->timeout, sizeof(timeout));
        if (rc) return rc;
        if (timeout) {
            ctx->err = "cooperative Berlekamp-Massey: a grid barrier wait ran out";
            return MODLIN_EHIP;
        }
        rc = rd(hs, ds, count * sizeof(BmSeq));
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
    // Blocks of K steps per pass (k_bm_batch_blk) unless the caller asks for the
    // one-step-per-pass form (MODLIN_METHOD_PLAIN) or p >= 2^62.
    if ((!opts || opts->method != MODLIN_METHOD_PLAIN) && p < (1ull << 62)) {
        if (f.mode == MODE_M61)
            hipLaunchKernelGGL((k_bm_batch_blk<MODE_M61>), dim3((uint32_t)count), dim3(BM_BLOCK_THREADS), 0, st, ds, f);
        else if (f.mode == MODE_M31)
            hipLaunchKernelGGL((k_bm_batch_blk<MODE_M31>), dim3((uint32_t)count), dim3(BM_BLOCK_THREADS), 0, st, ds, f);
        else
            hipLaunchKernelGGL((k_bm_batch_blk<MODE_GENERIC>), dim3((uint32_t)count), dim3(BM_BLOCK_THREADS), 0, st, ds,
                               f);
    } else {
    // threads per sequence: the tuning table's, else as many as keep the whole
    // batch resident (about 32 waves per CU), 256 at least
    const modlin_tuning_t tb = select_tuning(ctx->tune_table, &ctx->tune_override, p, maxn, count, true);
    const uint64_t cus = (uint64_t)std::max(ctx->cu_count, 1);
    const uint32_t nt = tb.bm_threads ? tb.bm_threads : count <= 2 * cus ? 1024 : count <= 4 * cus ? 512 : 256;
#define MDL_BM(NT, MODE) hipLaunchKernelGGL((k_bm_batch<NT, MODE>), dim3((uint32_t)count), dim3(NT), 0, st, ds, f)
#define MDL_BM_NT(MODE)              \
    if (nt == 1024)                   \
        MDL_BM(1024, MODE);           \
    else if (nt == 512)               \
        MDL_BM(512, MODE);            \
    else                              \
        MDL_BM(256, MODE);
    if (f.mode == MODE_M61) {
        MDL_BM_NT(MODE_M61)
    } else if (f.mode == MODE_M31) {
        MDL_BM_NT(MODE_M31)
    } else {
        MDL_BM_NT(MODE_GENERIC)
    }
#undef MDL_BM_NT
#undef MDL_BM
    }
    if (profile) MDL_HIP(hipEventRecord(ctx->evB, st));
    MDL_HIP(hipGetLastError());
    if (progress && (rc = wait_progress(ctx, maxn, progress, user))) return rc;
    rc = rd(hs, ds, count * sizeof(BmSeq));
    if (rc) return rc;
    for (uint64_t k = 0; k < count; k++)
        if (hs[k].timeout) {
            ctx->err = "batched Berlekamp-Massey: a wait for a step record ran out";
            return MODLIN_EHIP;
        }
    double bytes = 0;
    for (uint64_t k = 0; k < count; k++) {
        lcs[ord[k]] = hs[k].lc;
        bytes += (double)hs[k].bytes;
    }
    ctx->stats.bm_bytes = bytes;
    if (profile) {
        float ms = 0;
        MDL_HIP(hipEventElapsedTime(&ms, ctx->evA, ctx->evB));
        ctx->stats.bm_ms = ms;
    }
    return MODLIN_OK;
}

int lc_batch(modlin_ctx *ctx, uint64_t p, const modlin_seq_desc_t *seqs, uint64_t count, const modlin_opts_t *opts,
             uint64_t *lcs, modlin_progress_fn progress, void *user) {
    ctx->stats = modlin_stats_t{};
    if (count == 0) return MODLIN_OK;
    if (count != 1) progress = nullptr;
    if (progress && !ctx->prog_host) {
        void *h = nullptr;
        if (hipHostMalloc(&h, 256, hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess) {
            (void)hipGetLastError();
            ctx->err = "pinned progress word allocation failed";
            return MODLIN_ENOMEM;
        }
        ctx->prog_host = (uint64_t *)h;
    }
    uint64_t *prog_dev = nullptr;
    if (progress) {
        *(volatile uint64_t *)ctx->prog_host = 0;
        MDL_HIP(hipHostGetDevicePointer((void **)&prog_dev, ctx->prog_host, 0));
    }
    const FieldParams f = make_params(p);
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
        cur = align_up(cur + 4 * n * 8, 256);  // c, b, t and a fourth buffer (blocked kernels)
    }
    const size_t off_stage = cur;
    cur = align_up(cur + maxn * 4, 256);
    const size_t off_coop = cur;
    cur = align_up(cur + (count == 1 ? std::max(sizeof(BmCoop), sizeof(BmBlockX)) : sizeof(BmCoop)), 256);
    int rc = arena_reserve(ctx, cur);
    if (rc) return rc;
    char *base = (char *)ctx->arena;
    hipStream_t st = ctx->stream;
    // longest sequences first: workgroups are dispatched in index order, so when
    // a batch does not fit on the device at once the long ones start first
    std::vector<uint64_t> ord(count);
    for (uint64_t i = 0; i < count; i++) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](uint64_t a, uint64_t b2) { return seqs[a].n > seqs[b2].n; });
    std::vector<BmSeq> hs(count);
    for (uint64_t k = 0; k < count; k++) {
        const uint64_t i = ord[k];
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
        hs[k].s = sd;
        hs[k].c = cd;
        hs[k].b = cd + n;
        hs[k].t = cd + 2 * n;
        hs[k].u = cd + 3 * n;
        hs[k].n = (uint32_t)n;
        hs[k].lc = 0;
        hs[k].timeout = 0;
        hs[k].pad = 0;
        hs[k].prog = prog_dev;
        hs[k].bytes = 0;
    }
    BmSeq *ds = (BmSeq *)base;
    rc = copy_h2d(ctx, ds, hs.data(), count * sizeof(BmSeq));
    if (rc) return rc;
    return run_bm(ctx, f, p, ds, hs.data(), count, maxn, base + off_coop, opts, progress, user, ord.data(), lcs,
                  [ctx](void *dst, const void *src, size_t bytes) { return copy_d2h(ctx, dst, src, bytes); });
}

// ---------------------------------------------------------------------------
// The allocation-free entry point for one sequence in device memory
// (modlin_linear_complexity_noalloc).  The sequence descriptor, the cooperative
// kernels' exchange area and the four state buffers of up to
// MODLIN_LC_NOALLOC_MAX_N coefficients live in `scratch`, a file-scope device
// array that each GPU holds from module load to the end of the process.  Host
// code never uses the array's name as an address: noalloc_init (called by
// modlin_ctx_create, with the context's device current) takes each device's
// address of it once, and the entry point passes that address to the kernels.
// Every call on a GPU shares its array, so calls that overlap on one GPU race;
// what must start at zero (the descriptor's counters, the exchange area) is
// cleared by kernels at the start of every call.
// ---------------------------------------------------------------------------
constexpr size_t NA_SEQ = 256;                                                                  // BmSeq slot
constexpr size_t NA_X = ((sizeof(BmCoop) > sizeof(BmBlockX) ? sizeof(BmCoop) : sizeof(BmBlockX)) + 255) / 256 * 256;
constexpr size_t NA_STATE = 4ull * MODLIN_LC_NOALLOC_MAX_N * 8;                                  // c, b, t, u
constexpr size_t NA_BYTES = NA_SEQ + NA_X + NA_STATE;
static_assert(sizeof(BmSeq) <= NA_SEQ, "descriptor slot");
static __device__ __align__(256) uint64_t scratch[NA_BYTES / 8];

constexpr int NA_MAX_DEVICES = 64;
static std::mutex na_mu;
static char *na_ptr[NA_MAX_DEVICES];  // each device's address of scratch

int noalloc_init(modlin_ctx *ctx) {
    if (ctx->device < 0 || ctx->device >= NA_MAX_DEVICES) return MODLIN_ENODEV;
    std::lock_guard<std::mutex> g(na_mu);
    if (na_ptr[ctx->device]) return MODLIN_OK;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        (void)hipGetLastError();
        ctx->err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        return MODLIN_EHIP;
    }
    na_ptr[ctx->device] = (char *)ptr;
    return MODLIN_OK;
}

// The descriptor of the one sequence, written on the device.
__global__ void k_na_seq(BmSeq *q, const uint64_t *s, uint64_t *state, uint32_t n) {
    q->s = s;
    q->c = state;
    q->b = state + n;
    q->t = state + 2 * (size_t)n;
    q->u = state + 3 * (size_t)n;
    q->n = n;
    q->lc = 0;
    q->timeout = 0;
    q->pad = 0;
    q->prog = nullptr;
    q->bytes = 0;
}

int lc_noalloc(modlin_ctx *ctx, uint64_t p, const uint64_t *s, uint64_t n, uint64_t *lc) {
    ctx->stats = modlin_stats_t{};
    if (n == 0 || n > MODLIN_LC_NOALLOC_MAX_N) {
        ctx->err = n == 0 ? "empty sequence (the reference panics, fp.rs:265)"
                          : "sequence longer than MODLIN_LC_NOALLOC_MAX_N";
        return MODLIN_EINVAL;
    }
    char *base = nullptr;
    if (ctx->device >= 0 && ctx->device < NA_MAX_DEVICES) {
        std::lock_guard<std::mutex> g(na_mu);
        base = na_ptr[ctx->device];
    }
    if (!base || !ctx->na_host) {
        ctx->err = "allocation-free entry point: context set-up did not complete";
        return MODLIN_EHIP;
    }
    BmSeq *ds = (BmSeq *)base;
    char *exch = base + NA_SEQ;
    uint64_t *state = (uint64_t *)(base + NA_SEQ + NA_X);
    hipLaunchKernelGGL(k_na_seq, dim3(1), dim3(1), 0, ctx->stream, ds, s, state, (uint32_t)n);
    MDL_HIP(hipGetLastError());
    // results come back through the context's pinned word area (no staging allocation)
    auto rd = [ctx](void *dst, const void *src, size_t bytes) -> int {
        if (bytes > NA_HOST_BYTES) {
            ctx->err = "allocation-free read-back too large";
            return MODLIN_EHIP;
        }
        MDL_HIP(hipMemcpyAsync(ctx->na_host, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        MDL_HIP(hipStreamSynchronize(ctx->stream));
        memcpy(dst, ctx->na_host, bytes);
        return MODLIN_OK;
    };
    BmSeq hs;
    const uint64_t ord = 0;
    return run_bm(ctx, make_params(p), p, ds, &hs, 1, n, exch, nullptr, nullptr, nullptr, &ord, lc, rd);
}

}  // namespace mdl
