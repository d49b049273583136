// Trailing update C <- C - A X (mod p) on the int8 matrix cores of gfx950.
//
// The only dense contraction of the rank (fp.rs:202-248).  Operands are exact
// integers below p < 2^63; the product is computed exactly with int8 MFMA:
//
//   * every element is written in balanced signed digits, x^ = x - p [x > (p-1)/2]
//     = sum_l d_l 256^l with d_l in [-128, 127] and L = limbs(p) digits
//     (1 for p < 2^8, 3 for 65521, 4 for 2^31-1, 8 for 2^61-1);
//   * the digit shifts of A are folded into X:  A X = sum_la 256^la A_la X
//     == sum_la A_la X'_la (mod p) with X'_la = 256^la X mod p, and X'_la is
//     itself split into digits D_{la,lb}.  Hence
//         A X == sum_lb 256^lb P_lb (mod p),  P_lb = sum_la A_la D_{la,lb},
//     i.e. L GEMMs with K = L * 64 that share the A operand (A digits of all la
//     are laid out along K).  |P_lb| <= L * 64 * 2^14 < 2^23: exact in int32;
//   * the epilogue forms v = sum_lb 256^lb P_lb (|v| < 2^81), adds a multiple of
//     p >= 2^85 and reduces the positive 128-bit value with the reference's
//     Barrett reduction (fp.rs:71-79).  Every result is the canonical residue, so
//     the update equals the VALU path and the CPU reference bit for bit.
//
// k_gather_split  produces the digit planes: A (rows O x L*64, K-contiguous) and
//                 X' (L planes of columns x L*64, K-contiguous), fusing the X
//                 product (the reference's U solve) with the digit split.
// k_update_mfma   64 x 64 output tiles per workgroup, 4 waves of 32 x 32, built
//                 from v_mfma_i32_16x16x64_i8 (lane l holds row / column l % 16 and
//                 the 16 contiguous K bytes 16 (l / 16) .. +16 of a 64-deep step).
#include "rank_common.hpp"

namespace mdl {

typedef int v4i __attribute__((ext_vector_type(4)));

template <int L>
__global__ __launch_bounds__(256) void k_gather_split(RankMat *mats, uint32_t nmat, const uint32_t *gx_off,
                                                      const RankPlan *plan, FieldParams f, LimbParams lp) {
    __shared__ uint64_t Zs[BMAX][BMAX + 1];
    __shared__ uint64_t Ms[BMAX][XCOLS + 1];
    constexpr uint32_t KA = L * KSTEP;
    const uint32_t total = plan->gx_total;
    const uint32_t tid = threadIdx.x;
    const uint64_t p = f.p;
    for (uint32_t t = blockIdx.x; t < total; t += gridDim.x) {
        const uint32_t mi = find_mat(gx_off, nmat, t);
        RankMat &S = mats[mi];
        const uint32_t lt = t - gx_off[mi];
        const uint32_t O = S.n - S.rank, pb = S.pb, ld = S.ld;
        const uint32_t ng = cdiv(O, GATHER_ROWS);
        if (lt < ng) {
            // A digits: Al[i][la*64 + k] = digit la of M[perm[rank + i]][J[k]] (zero for k >= pb)
            const uint32_t r0 = lt * GATHER_ROWS;
            const uint32_t rows = min(GATHER_ROWS, O - r0);
            for (uint32_t idx = tid; idx < rows * KSTEP; idx += blockDim.x) {
                const uint32_t i = idx / KSTEP, kk = idx - i * KSTEP;
                uint64_t x = 0;
                if (kk < pb) x = S.M[(size_t)S.perm[S.rank + r0 + i] * ld + S.J[kk]];
                int64_t v = balanced(x, p);
                int8_t *dst = S.Al + (size_t)(r0 + i) * KA + kk;
#pragma unroll
                for (int la = 0; la < L; la++) {
                    const int8_t d = (int8_t)(v & 0xff);
                    dst[la * KSTEP] = d;
                    v = (v - d) >> 8;
                }
            }
        } else {
            // X'[la][k][c] digits from X[k][c] = sum_s Z[k][s] M[R[s]][c0 + c]
            const uint32_t cb = (lt - ng) * XCOLS;
            const uint32_t cols = min(XCOLS, S.cend - S.c0 - cb);
            __syncthreads();
            for (uint32_t idx = tid; idx < pb * pb; idx += blockDim.x) Zs[idx / pb][idx % pb] = S.Z[(idx / pb) * BMAX + idx % pb];
            for (uint32_t idx = tid; idx < pb * XCOLS; idx += blockDim.x) {
                const uint32_t s = idx / XCOLS, c = idx - s * XCOLS;
                Ms[s][c] = (c < cols) ? S.M[(size_t)S.R[s] * ld + S.c0 + cb + c] : 0;
            }
            __syncthreads();
            const uint64_t sb = f.safe_batch;
            const size_t plane = (size_t)cdiv(S.n, UT) * UT * KA;
            for (uint32_t o = tid; o < pb * cols; o += blockDim.x) {
                const uint32_t c = o / pb, kk = o - c * pb;
                uint64_t ah = 0, al = 0, tot = 0, cnt = 0;
                for (uint32_t s = 0; s < pb; s++) {
                    uint64_t h, l;
                    mul_wide(Zs[kk][s], Ms[s][c], h, l);
                    add128(ah, al, h, l);
                    if (++cnt == sb) {
                        tot = add_mod(tot, reduce128(f, ah, al), p);
                        ah = al = 0;
                        cnt = 0;
                    }
                }
                if (cnt) tot = add_mod(tot, reduce128(f, ah, al), p);
                int8_t *dst = S.Xl + (size_t)(cb + c) * KA + kk;
#pragma unroll
                for (int la = 0; la < L; la++) {
                    const uint64_t xl = la ? mul_mod(f, tot, lp.pow256[la]) : tot;
                    int64_t v = balanced(xl, p);
#pragma unroll
                    for (int lb = 0; lb < L; lb++) {
                        const int8_t d = (int8_t)(v & 0xff);
                        dst[lb * plane + la * KSTEP] = d;
                        v = (v - d) >> 8;
                    }
                }
            }
            __syncthreads();
        }
    }
}

template <int L>
__global__ __launch_bounds__(256) void k_update_mfma(RankMat *mats, uint32_t nmat, const uint32_t *upd_off,
                                                     const RankPlan *plan, FieldParams f, LimbParams lp) {
    constexpr uint32_t KA = L * KSTEP;
    const uint32_t total = plan->upd_total;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    const uint64_t p = f.p;
    for (uint32_t t = blockIdx.x; t < total; t += gridDim.x) {
        const uint32_t mi = find_mat(upd_off, nmat, t);
        RankMat &S = mats[mi];
        const uint32_t lt = t - upd_off[mi];
        const uint32_t O = S.n - S.rank, pb = S.pb, ld = S.ld;
        const uint32_t cols = S.cend - S.c0;
        const uint32_t nct = cdiv(cols, UT);
        const uint32_t tr = lt / nct, tc = lt - tr * nct;
        const uint32_t r0 = tr * UT, cb = tc * UT;
        const uint32_t nr = min(UT, O - r0), nc = min(UT, cols - cb);
        uint32_t nzcol = 0xffffffffu;
        if (pb == 0) {
            // detection only (no pivots): scan the tile for a nonzero panel column
            for (uint32_t e = tid; e < nr * nc; e += blockDim.x) {
                const uint32_t i = e / nc, c = e - i * nc;
                const uint32_t gcol = S.c0 + cb + c;
                if (gcol < S.jend && S.M[(size_t)S.perm[S.rank + r0 + i] * ld + gcol] != 0) nzcol = min(nzcol, gcol);
            }
            if (S.resid && nzcol != 0xffffffffu) atomicMin(&S.first_nz_col, nzcol);
            continue;
        }
        v4i acc[L][2][2];
#pragma unroll
        for (int lb = 0; lb < L; lb++)
#pragma unroll
            for (int a = 0; a < 2; a++)
#pragma unroll
                for (int b = 0; b < 2; b++) acc[lb][a][b] = v4i{0, 0, 0, 0};
        const size_t plane = (size_t)cdiv(S.n, UT) * UT * KA;
        const int8_t *Ap = S.Al + (size_t)(r0 + wr + (lane & 15)) * KA + 16 * (lane >> 4);
        const int8_t *Bp = S.Xl + (size_t)(cb + wc + (lane & 15)) * KA + 16 * (lane >> 4);
#pragma unroll
        for (int s = 0; s < L; s++) {
            const v4i a0 = *(const v4i *)(Ap + s * KSTEP);
            const v4i a1 = *(const v4i *)(Ap + 16 * KA + s * KSTEP);
#pragma unroll
            for (int lb = 0; lb < L; lb++) {
                const v4i b0 = *(const v4i *)(Bp + lb * plane + s * KSTEP);
                const v4i b1 = *(const v4i *)(Bp + lb * plane + 16 * KA + s * KSTEP);
                acc[lb][0][0] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a0, b0, acc[lb][0][0], 0, 0, 0);
                acc[lb][0][1] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a0, b1, acc[lb][0][1], 0, 0, 0);
                acc[lb][1][0] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a1, b0, acc[lb][1][0], 0, 0, 0);
                acc[lb][1][1] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a1, b1, acc[lb][1][1], 0, 0, 0);
            }
        }
        // epilogue: C - sum_lb 256^lb P_lb, reduced
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const uint32_t i = wr + 16 * a + 4 * (lane >> 4) + r;
                    const uint32_t c = wc + 16 * b + (lane & 15);
                    if (i >= nr || c >= nc) continue;
                    int64_t lo = 0, hi = 0;
#pragma unroll
                    for (int lb = 0; lb < L; lb++) {
                        const int64_t pv = (int64_t)acc[lb][a][b][r];
                        if (lb < 4) lo += pv << (8 * lb);
                        else hi += pv << (8 * (lb - 4));
                    }
                    // V = hi * 2^32 + lo as a two's-complement 128-bit value
                    const uint64_t hl = (uint64_t)hi << 32, hh = (uint64_t)(hi >> 32);
                    uint64_t vl = hl + (uint64_t)lo;
                    uint64_t vh = hh + (lo < 0 ? ~0ull : 0ull) + (vl < hl ? 1ull : 0ull);
                    uint64_t *mp = S.M + (size_t)S.perm[S.rank + r0 + i] * ld + S.c0 + cb + c;
                    const uint64_t cv = *mp;
                    // t = pbig + C - V  (positive, below 2^87)
                    uint64_t tl = lp.pbig_lo + cv;
                    uint64_t th = lp.pbig_hi + (tl < cv ? 1ull : 0ull);
                    const uint64_t nl = tl - vl;
                    th = th - vh - (tl < vl ? 1ull : 0ull);
                    tl = nl;
                    const uint64_t res = reduce128(f, th, tl);
                    *mp = res;
                    const uint32_t gcol = S.c0 + cb + c;
                    if (res != 0 && gcol < S.jend) nzcol = min(nzcol, gcol);
                }
        if (S.resid && nzcol != 0xffffffffu) atomicMin(&S.first_nz_col, nzcol);
        if (tid == 0) {
            atomicAdd(&S.macs, (unsigned long long)nr * nc * pb);
            atomicAdd(&S.upd_bytes, 16ull * nr * nc + 8ull * pb * (nr + nc));
        }
    }
}

// Launchers: one panel iteration with the MFMA update (the block, plan and
// finish kernels are those of the plain path).
void launch_block_plan(hipStream_t st, RankMat *dmats, uint32_t nmat, const FieldParams &f, uint32_t b,
                       uint32_t *gx_off, uint32_t *upd_off, RankPlan *plan);
void launch_finish(hipStream_t st, RankMat *dmats, uint32_t nmat, uint32_t *active);

template <int L>
static void launch_mfma_L(hipStream_t st, RankMat *dmats, uint32_t nmat, const FieldParams &f, const LimbParams &lp,
                          uint32_t *gx_off, uint32_t *upd_off, RankPlan *plan, uint32_t grid_gx, uint32_t grid_upd,
                          hipEvent_t e0, hipEvent_t e1) {
    hipLaunchKernelGGL(k_gather_split<L>, dim3(grid_gx), dim3(256), 0, st, dmats, nmat, gx_off, plan, f, lp);
    if (e0) hipEventRecord(e0, st);
    hipLaunchKernelGGL(k_update_mfma<L>, dim3(grid_upd), dim3(256), 0, st, dmats, nmat, upd_off, plan, f, lp);
    if (e1) hipEventRecord(e1, st);
}

void launch_rank_iteration_mfma(hipStream_t st, RankMat *dmats, uint32_t nmat, const FieldParams &f,
                                const LimbParams &lp, uint32_t b, uint32_t *gx_off, uint32_t *upd_off,
                                RankPlan *plan, uint32_t *active, uint32_t grid_gx, uint32_t grid_upd,
                                hipEvent_t e0, hipEvent_t e1) {
    launch_block_plan(st, dmats, nmat, f, b, gx_off, upd_off, plan);
    switch (lp.L) {
        case 1: launch_mfma_L<1>(st, dmats, nmat, f, lp, gx_off, upd_off, plan, grid_gx, grid_upd, e0, e1); break;
        case 2: launch_mfma_L<2>(st, dmats, nmat, f, lp, gx_off, upd_off, plan, grid_gx, grid_upd, e0, e1); break;
        case 3: launch_mfma_L<3>(st, dmats, nmat, f, lp, gx_off, upd_off, plan, grid_gx, grid_upd, e0, e1); break;
        case 4: launch_mfma_L<4>(st, dmats, nmat, f, lp, gx_off, upd_off, plan, grid_gx, grid_upd, e0, e1); break;
        case 5: launch_mfma_L<5>(st, dmats, nmat, f, lp, gx_off, upd_off, plan, grid_gx, grid_upd, e0, e1); break;
        case 6: launch_mfma_L<6>(st, dmats, nmat, f, lp, gx_off, upd_off, plan, grid_gx, grid_upd, e0, e1); break;
        case 7: launch_mfma_L<7>(st, dmats, nmat, f, lp, gx_off, upd_off, plan, grid_gx, grid_upd, e0, e1); break;
        default: launch_mfma_L<8>(st, dmats, nmat, f, lp, gx_off, upd_off, plan, grid_gx, grid_upd, e0, e1); break;
    }
    launch_finish(st, dmats, nmat, active);
}

}  // namespace mdl
