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
#include "engine.hpp"

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
