// Trailing update C <- C - A X (mod p) on the int8 matrix cores of gfx950.
//
// The only dense contraction of the rank (fp.rs:202-248).  Operands are exact
// integers below p < 2^63; the product is computed exactly with int8 MFMA:
//
//   * every element x in [0, p) is written in balanced signed digits,
//     x^ = x - p [x > (p-1)/2] = sum_l d_l 256^l, d_l in [-128, 127], with
//     L = limbs(p) digits (1 for p < 2^8, 3 for 65521, 4 for 2^31-1, 8 for 2^61-1);
//   * A X = sum_s 256^s P_s with P_s = sum_{la+lb=s} A_la X_lb, s < 2L-1: each
//     digit pair is one int8 GEMM with K = 64 (the panel), accumulated exactly in
//     int32 per shift class (|P_s| <= 8 * 64 * 2^14 = 2^23);
//   * the epilogue (limb.hpp) forms the residue of sum_s 256^s P_s.  For
//     2^61 - 1 every class starts at KOFF2 (non-negative 26-bit sums) and
//     classes c and c + 8, whose weights differ by 2^3, are added as one 32-bit
//     value before the Mersenne folds (limb_sub_m61_k2).  For the other moduli
//     one panel's classes are joined into exact 32-bit pair words -- the even
//     classes' accumulators start at an offset that keeps them non-negative --
//     added as two 32-bit word strings and reduced once (Mersenne folds for
//     2^31 - 1, or the weights 2^(32 i) mod p and the reference's Barrett
//     reduction, fp.rs:71-79); two panels' classes (the far update) go through
//     limb_residue_k2.  Every result is the canonical residue, so the update
//     equals the VALU path and the CPU reference bit for bit.
//
// k_gather_split  digit planes of A = M[O, J] (rows x L*64, K-contiguous) and of
//                 X = Z M[R, c0:] (columns x L*64, K-contiguous), fusing the X
//                 product (the reference's U-block solve) with the digit split.
// k_update_near   the near part: one 64 x 64 output tile per item, on the far
//                 update's machinery (k_update_far: row digits in registers,
//                 column digits staged in LDS, v_mfma_i32_16x16x64_i8 with lane l
//                 holding row or column l % 16 and K bytes 16 (l / 16) .. +16 of
//                 a 64-byte step).
// k_update_far    the far part, by row band and XCD group, one or two panels.
#include "rank_common.hpp"

namespace mdl {

typedef int v4i __attribute__((ext_vector_type(4)));

// Geometry of one 64 x 64 update tile.
struct TileMeta {
    uint32_t mi, r0, cb, nr, nc, pb;
};

__device__ __forceinline__ TileMeta tile_meta(RankMat *mats, uint32_t nmat, const uint32_t *upd_off, uint32_t t,
                                              uint32_t par, uint32_t far) {
    TileMeta m;
    m.mi = find_mat(upd_off, nmat, t);
    const RankMat &S = mats[m.mi];
    const IterState &I = S.it[par];
    const uint32_t lt = t - upd_off[m.mi];
    const uint32_t O = S.n - I.rank;
    uint32_t cbase, cols;
    update_range(I, far, cbase, cols);
    m.pb = I.pb;
    if (!tile_geom(O, cbase, cols, far, lt, m.r0, m.cb, m.nr, m.nc)) {
        m.r0 = m.cb = m.nr = m.nc = 0;  // padding tile: no work
        m.pb = 0;
    }
    return m;
}

// Class accumulators of 2^61 - 1 start at KOFF2 (limb_sub_m61_k2: 32-bit class
// pairs, about half the VALU work of the pair words); the other moduli at
// HOFF / 0 (limb_residue's pair words, cheaper there for 4 digits or more).
template <int MODE, int L>
constexpr bool kind_k2() {
    return MODE == MODE_M61 && L == 8;
}

// Residue of the product for element (b, r) of this lane's accumulators.
template <int MODE, int NC>
__device__ __forceinline__ uint64_t acc_residue(const FieldParams &f, const LimbParams &lp, v4i (*acc)[2], int b,
                                                int r) {
    int32_t P[NC];
#pragma unroll
    for (int s = 0; s < NC; s++) P[s] = acc[s][b][r];
    if constexpr (kind_k2<MODE, (NC + 1) / 2>()) {
        const uint64_t neg = limb_sub_m61_k2(0, lp.kc2, P);  // (0 - product) mod p
        return neg == 0 ? 0 : P61 - neg;
    }
    return sub_mod(limb_residue<MODE, NC>(f, lp, P), lp.k0, f.p);
}

// Balanced signed digits of four consecutive elements, packed per digit plane:
// out[l] holds digit l of x[0..3] in its bytes 0..3.
template <int L>
__device__ __forceinline__ void digits4(const uint64_t *x, uint64_t p, uint32_t *out) {
    int64_t v[4];
#pragma unroll
    for (int j = 0; j < 4; j++) v[j] = balanced(x[j], p);
#pragma unroll
    for (int l = 0; l < L; l++) {
        uint32_t w = 0;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int8_t d = (int8_t)(v[j] & 0xff);
            w |= (uint32_t)(uint8_t)d << (8 * j);
            v[j] = (v[j] - d) >> 8;
        }
        out[l] = w;
    }
}

// One 64 x 64 x 64 limb product on the matrix cores: acc[s][b] = init_s +
// sum_{la+lb=s} A_la B_lb for this wave's 16 x 32 block (rows wr.., columns wc..),
// with init_s = HOFF for the even classes and 0 for the odd ones (limb.hpp), or
// KOFF2 for every class (K2): the first product of each class takes its initial
// value as the MFMA's C operand.  As / Bs hold the digit rows (K-contiguous, L
// planes of 64 bytes, row stride KAP).
template <int L, uint32_t KAP, bool K2>
__device__ __forceinline__ void limb_tile_mma(const int8_t *As, const int8_t *Bs, uint32_t lane, uint32_t wr,
                                              uint32_t wc, v4i (*acc)[2]) {
    constexpr int H = K2 ? KOFF2 : (int)HOFF;
    const v4i hoff = {H, H, H, H}, zero = {K2 ? KOFF2 : 0, K2 ? KOFF2 : 0, K2 ? KOFF2 : 0, K2 ? KOFF2 : 0};
    v4i bf[L][2];
    const uint32_t koff = 16 * (lane >> 4);
#pragma unroll
    for (int lb = 0; lb < L; lb++)
#pragma unroll
        for (int b = 0; b < 2; b++) bf[lb][b] = *(const v4i *)(Bs + (wc + 16 * b + (lane & 15)) * KAP + lb * KSTEP + koff);
#pragma unroll
    for (int la = 0; la < L; la++) {
        const v4i af = *(const v4i *)(As + (wr + (lane & 15)) * KAP + la * KSTEP + koff);
#pragma unroll
        for (int lb = 0; lb < L; lb++) {
            const int s = la + lb;
            const bool first = (la == 0) || (lb == L - 1);  // loop order: la outer, lb inner
#pragma unroll
            for (int b = 0; b < 2; b++) {
                const v4i c = first ? ((s & 1) ? zero : hoff) : acc[s][b];
                acc[s][b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af, bf[lb][b], c, 0, 0, 0);
            }
        }
    }
}

// Digit planes of the two update operands.
//  gather task: Al[i][la*64 + k] = digit la of M[perm[rank + i]][J[k]] (zero for k >= pb);
//               for a PAIR iteration also AlP[i][la*64 + k] = digit la of
//               M[perm[rank + i]][J'[k]], J' the pending panel's pivot columns.
//  X task:      Xl[c][lb*64 + k] = digit lb of X[k][c], X = Z M[R, c0:] mod p (the
//               reference's U-block solve, fp.rs:180-200), computed as a limb product
//               on the matrix cores: Z and the pivot rows are split into digits in LDS.
//               A DEFER iteration writes XlP instead.  In a PAIR iteration the
//               pivot rows still lack the pending panel's update in the far
//               columns c >= dcol, so there X = Z Y with
//                 Y[s][c] = M[R[s]][c] - sum_a M[R[s]][J'[a]] X'[a][c]  (mod p),
//               X' the pending panel's X (digits in XlP): a first limb product
//               forms Y, whose digits then take the pivot rows' place.
template <int L, int MODE>
__global__ __launch_bounds__(512) void k_gather_split(RankMat *mats, uint32_t nmat, const uint32_t *gx_off,
                                                      const RankPlan *plan, FieldParams f, LimbParams lp, uint32_t par,
                                                      uint32_t part) {
    constexpr uint32_t KA = L * KSTEP;
    constexpr uint32_t KAP = KA + 16;
    constexpr int NC = 2 * L - 1;
    __shared__ __attribute__((aligned(16))) int8_t As[UT * KAP];
    __shared__ __attribute__((aligned(16))) int8_t Bs[UT * KAP];
    const uint32_t total = part ? plan->xf_total : plan->gx_total;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = uni32(tid >> 6);
    const uint32_t wr = (wave >> 1) * 16, wc = (wave & 1) * 32;
    const uint64_t p = f.p;
    for (uint32_t t = blockIdx.x; t < total; t += gridDim.x) {
        const uint32_t mi = find_mat(gx_off, nmat, t);
        RankMat &S = mats[mi];
        const IterState &I = S.it[par];
        const uint32_t O = S.n - I.rank, pb = I.pb, ld = S.ld;
        const uint32_t ng = cdiv(O, GATHER_ROWS);
        // part 0: row tasks, then the near X tasks; part 1: the far X tasks
        const uint32_t lt = t - gx_off[mi] + (part ? ng + x_near_tasks(I) : 0u);
        if (lt < ng) {
            const uint32_t r0 = lt * GATHER_ROWS;
            const uint32_t rows = min(GATHER_ROWS, O - r0);
            const bool pair = I.kind == IT_PAIR;
            // own panel's columns (when it has pivots), then the pending panel's (PAIR)
            for (uint32_t pass = pb > 0 ? 0u : 1u; pass < (pair ? 2u : 1u); pass++) {
                const uint32_t *Jc = pass ? I.dJ : I.J;
                const uint32_t kb = pass ? I.dpb : pb;
                int8_t *out = pass ? S.AlP : S.Al;
                // GATHER_ROWS x 16 units of 4 entries, two per thread: every load of
                // both (row order, then entries) is issued before any digit is formed
                static_assert(GATHER_ROWS * (KSTEP / 4) <= 2 * 512, "two units per thread");
                uint32_t iu[2], ku[2];
                bool ok[2];
                size_t mrow[2];
#pragma unroll
                for (int u = 0; u < 2; u++) {
                    const uint32_t idx = tid + 512 * u;
                    iu[u] = idx / (KSTEP / 4);
                    ku[u] = (idx - iu[u] * (KSTEP / 4)) * 4;
                    ok[u] = iu[u] < rows;
                    const uint32_t prow = S.perm[I.rank + r0 + (ok[u] ? iu[u] : 0)];
                    mrow[u] = (size_t)prow * ld;
                    // the row snapshot of this iteration, for the far update
                    if (ok[u] && ku[u] == 0 && pass == (pb > 0 ? 0u : 1u)) S.rows[r0 + iu[u]] = prow;
                }
                uint64_t x[2][4];
#pragma unroll
                for (int u = 0; u < 2; u++)
#pragma unroll
                    for (int j = 0; j < 4; j++) x[u][j] = (ok[u] && ku[u] + j < kb) ? m_get(S, mrow[u] + Jc[ku[u] + j]) : 0;
#pragma unroll
                for (int u = 0; u < 2; u++) {
                    if (!ok[u]) continue;
                    uint32_t w[L];
                    digits4<L>(x[u], p, w);
                    uint32_t *dst = (uint32_t *)(out + (size_t)(r0 + iu[u]) * KA + ku[u]);
#pragma unroll
                    for (int la = 0; la < L; la++) dst[la * (KSTEP / 4)] = w[la];
                }
            }
            continue;
        }
        const uint32_t cb = (lt - ng) * XCOLS;
        const uint32_t cols = min(XCOLS, I.cend - I.c0 - cb);
        // uniform: the task has columns the pending panel has not updated
        const bool corr = I.kind == IT_PAIR && I.c0 + cb + XCOLS > I.dcol;
        __syncthreads();
        if (corr) {
            // As[s][la*64 + a] = digit la of M[R[s]][J'[a]]; Bs[c][lb*64 + a] = digit lb of
            // X'[a][c] for the columns c0 + cb + c >= dcol (zero for the others)
            const uint32_t dpb = I.dpb;
            for (uint32_t idx = tid; idx < UT * (KSTEP / 4); idx += blockDim.x) {
                const uint32_t sl = idx / (KSTEP / 4), a4 = (idx - sl * (KSTEP / 4)) * 4;
                uint64_t x[4];
#pragma unroll
                for (int j = 0; j < 4; j++)
                    x[j] = (sl < pb && a4 + j < dpb) ? m_get(S, (size_t)I.R[sl] * ld + I.dJ[a4 + j]) : 0;
                uint32_t w[L];
                digits4<L>(x, p, w);
#pragma unroll
                for (int la = 0; la < L; la++) *(uint32_t *)(As + sl * KAP + la * KSTEP + a4) = w[la];
            }
            for (uint32_t idx = tid; idx < XCOLS * (KA / 16); idx += blockDim.x) {
                const uint32_t c = idx / (KA / 16), ch = idx - c * (KA / 16);
                const bool live = c < cols && I.c0 + cb + c >= I.dcol;
                *(v4i *)(Bs + c * KAP + ch * 16) =
                    live ? ((const v4i *)(S.XlP + (size_t)(cb + c + I.dxoff) * KA))[ch] : (v4i){0, 0, 0, 0};
            }
            __syncthreads();
            v4i acc[NC][2];
            limb_tile_mma<L, KAP, kind_k2<MODE, L>()>(As, Bs, lane, wr, wc, acc);
            // Y = M[R[s]][c] - product (mod p) for this lane's 4 slots and 2 columns
            uint64_t y[2][4];
#pragma unroll
            for (int b = 0; b < 2; b++) {
                const uint32_t c = wc + 16 * b + (lane & 15);
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const uint32_t sl = wr + 4 * (lane >> 4) + r;
                    const uint64_t pr = acc_residue<MODE, NC>(f, lp, acc, b, r);
                    y[b][r] = (c < cols && sl < pb) ? sub_mod(m_get(S, (size_t)I.R[sl] * ld + I.c0 + cb + c), pr, p) : 0;
                }
            }
            __syncthreads();  // every wave is done with As and Bs
            // Bs[c][lb*64 + s] = digit lb of Y[s][c]; As = digits of Z
#pragma unroll
            for (int b = 0; b < 2; b++) {
                const uint32_t c = wc + 16 * b + (lane & 15), s4 = wr + 4 * (lane >> 4);
                uint32_t w[L];
                digits4<L>(y[b], p, w);
#pragma unroll
                for (int lb = 0; lb < L; lb++) *(uint32_t *)(Bs + c * KAP + lb * KSTEP + s4) = w[lb];
            }
            for (uint32_t idx = tid; idx < UT * (KA / 16); idx += blockDim.x) {
                const uint32_t k = idx / (KA / 16), ch = idx - k * (KA / 16);
                *(v4i *)(As + k * KAP + ch * 16) = ((const v4i *)(S.Zl + ((size_t)par * BMAX + k) * KA))[ch];
            }
        } else {
            // As[k][la*64 + s] = digit la of Z[k][s] (split once by k_block_z);
            // Bs[c][lb*64 + s] = digit lb of M[R[s]][c0 + cb + c]
            for (uint32_t idx = tid; idx < UT * (KA / 16); idx += blockDim.x) {
                const uint32_t k = idx / (KA / 16), ch = idx - k * (KA / 16);
                *(v4i *)(As + k * KAP + ch * 16) = ((const v4i *)(S.Zl + ((size_t)par * BMAX + k) * KA))[ch];
            }
            for (uint32_t idx = tid; idx < XCOLS * (KSTEP / 4); idx += blockDim.x) {
                const uint32_t c = idx % XCOLS, s4 = (idx / XCOLS) * 4;
                uint64_t x[4];
#pragma unroll
                for (int j = 0; j < 4; j++)
                    x[j] = (c < cols && s4 + j < pb) ? m_get(S, (size_t)I.R[s4 + j] * ld + I.c0 + cb + c) : 0;
                uint32_t w[L];
                digits4<L>(x, p, w);
#pragma unroll
                for (int lb = 0; lb < L; lb++) *(uint32_t *)(Bs + c * KAP + lb * KSTEP + s4) = w[lb];
            }
        }
        __syncthreads();
        v4i acc[NC][2];
        limb_tile_mma<L, KAP, kind_k2<MODE, L>()>(As, Bs, lane, wr, wc, acc);
        // X = sum_s 256^s P_s mod p, then its digits (4 consecutive k per lane and column)
        int8_t *xl = own_xl(S, I);
#pragma unroll
        for (int b = 0; b < 2; b++) {
            const uint32_t c = wc + 16 * b + (lane & 15);
            uint64_t x[4];
#pragma unroll
            for (int r = 0; r < 4; r++) x[r] = acc_residue<MODE, NC>(f, lp, acc, b, r);
            if (c < cols) {
                uint32_t dw[L];
                digits4<L>(x, p, dw);
                const uint32_t k4 = wr + 4 * (lane >> 4);
#pragma unroll
                for (int lb = 0; lb < L; lb++) *(uint32_t *)(xl + (size_t)(cb + c) * KA + lb * KSTEP + k4) = dw[lb];
            }
        }
    }
}

// The same limb product for a 16 x 16 block with the roles of the operands
// swapped in the register file: the wave's L A digit fragments (its 16 rows) are
// held, and the B fragment of its 16 columns is read per digit plane lb.
template <int L, uint32_t KAP>
__device__ __forceinline__ void limb_tile_mma_ahold(const v4i *af, const int8_t *Bs, uint32_t lane, uint32_t wc,
                                                    v4i *acc) {
    constexpr int H = (int)HOFF;
    const v4i hoff = {H, H, H, H}, zero = {0, 0, 0, 0};
    const uint32_t koff = 16 * (lane >> 4);
#pragma unroll
    for (int lb = 0; lb < L; lb++) {
        const v4i bf = *(const v4i *)(Bs + (wc + (lane & 15)) * KAP + lb * KSTEP + koff);
#pragma unroll
        for (int la = 0; la < L; la++) {
            const int s = la + lb;
            const bool first = (lb == 0) || (la == L - 1);  // loop order: lb outer, la inner
            const v4i c = first ? ((s & 1) ? zero : hoff) : acc[s];
            acc[s] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[la], bf, c, 0, 0, 0);
        }
    }
}

// Far part of the update, by row band: work item (row band, XCD group x)
// updates the column tiles tc = x, x + 8, ... of that row band.  Items are
// numbered rows-major with the group fastest; with a grid that is a multiple of
// 8, item t runs on XCD t mod 8 = x, so an XCD's L2 holds its own column tiles.
//
// A PAIR iteration's far part carries two panels: the pending one (digits AlP
// of M[O, J'], XlP of its X, offset by dxoff columns) and its own (Al, Xl), up
// to 128 pivots accumulated together -- each far element is read and written
// once per two panels.  Every class accumulator starts at KOFF2 and the
// recombination is limb_residue_k2 / limb_sub_m61_k2 (exact for K <= 128), for
// single panels too.
//
// A workgroup is 4 waves; wave w owns rows 16 w .. 16 w + 15 of the band and
// keeps its A digit fragments (L per panel, read once per item straight from
// HBM) in registers.  The columns go by in steps of 32: the B digit planes of a
// step (one tile per panel) are staged in LDS, double-buffered with one barrier
// per step: the next step's tiles are fetched into registers at the start of a
// step and stored once its first products are issued.  Each wave
// computes QW 16 x 16 column quarters at a time, digit plane lb outer and la
// inner, so consecutive products go to different classes.  Two workgroups share
// a CU and synchronise independently, so the matrix-core work of one overlaps
// the epilogue (VALU) of the other.  C values of a quarter are fetched one step
// ahead, right after the quarter's own epilogue.
constexpr uint32_t FAR_THREADS = 256;
constexpr uint32_t FSTEP = 32;  // columns per step

// Everything an item of the far update needs, read once into registers: the
// stores go through global pointers that the compiler cannot tell apart from
// the RankMat array.
struct FarItem {
    uint64_t *M;  // the working copy (u64 or, when m32, u32 entries), offset by c0 in far_item
    uint32_t c0;
    gptr<const uint32_t> permg;    // row r0 + i of the item is physical row permg[rank + r0 + i]
    gptr<const v4i> Ag[2], Xg[2];  // panel h: 0 the pending one, 1 this iteration's own
    uint32_t x, r0, nr, rank, ld, cbase, width, C;
    uint32_t scan_end;  // NEAR: nonzero results in columns c0 + c < scan_end are reported
};

// One item with the panels H0 .. H1 (compile-time, so every product's
// accumulator operand is known) and entries of type T.  NEAR: the item is the
// single column tile x (the near part's 64 x 64 tiles), and the first column
// with a nonzero result below it.scan_end lowers nzcol (the missed-pivot check).
template <int L, int MODE, int QW, int H0, int H1, typename T, bool NEAR = false>
__device__ __forceinline__ void far_item(const FarItem &it, int8_t *Bs, const FieldParams &f, const LimbParams &lp,
                                         gptr<T> tr, uint32_t tid, uint32_t lane, uint32_t wr, uint32_t &nzcol) {
    constexpr uint32_t KA = L * KSTEP;
    constexpr uint32_t KAP = KA + 16;           // LDS row stride: consecutive columns start 4 banks apart
    constexpr int NC = 2 * L - 1;
    constexpr uint32_t CH = KA / 16;            // 16-byte chunks per column
    constexpr uint32_t NCH = FSTEP * CH;        // chunks per panel tile of a step
    constexpr uint32_t NV = (NCH + FAR_THREADS - 1) / FAR_THREADS;
    constexpr uint32_t TILE = FSTEP * KAP;      // LDS bytes of one panel tile
    constexpr uint32_t BUF = 2 * TILE;
    const uint32_t koff = 16 * (lane >> 4);
    const uint32_t cbase = it.cbase, width = it.width, x = it.x;
    // Steps: tile x + 8 (j / 2), half j % 2; the last tile's second half only if
    // it holds a column.
    const uint32_t ntile = NEAR ? 1u : (it.C - x + 7) / 8;
    const uint32_t last_cb = cbase + (x + 8 * (ntile - 1)) * UT;
    const uint32_t nstep = 2 * ntile - (last_cb + FSTEP < cbase + width ? 0u : 1u);
    auto step_cb = [&](uint32_t j) { return cbase + (x + 8 * (j >> 1)) * UT + FSTEP * (j & 1); };
    // This wave's A fragments: row wr + (lane & 15), K bytes koff.. of each plane.
    v4i af[H1 - H0 + 1][L];
#pragma unroll
    for (int h = H0; h <= H1; h++) {
        const gptr<const v4i> ga = it.Ag[h] + ((size_t)(it.r0 + wr + (lane & 15)) * KA + koff) / 16;
#pragma unroll
        for (int la = 0; la < L; la++) af[h - H0][la] = ga[la * (KSTEP / 16)];
    }
    // This lane's rows; a row outside the band loads and stores through `tr`.
    const gptr<T> Mg = to_global((T *)it.M) + it.c0;
    gptr<T> mrow[4];
    bool rok[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const uint32_t i = wr + 4 * (lane >> 4) + r;
        rok[r] = i < it.nr;
        mrow[r] = Mg + (size_t)it.permg[it.rank + it.r0 + min(i, it.nr - 1)] * it.ld;
    }
    v4i rb[H1 - H0 + 1][NV];
    auto fetch_b = [&](uint32_t j) {
#pragma unroll
        for (int h = H0; h <= H1; h++) {
            const gptr<const v4i> gb = it.Xg[h] + (size_t)step_cb(j) * (KA / 16);
#pragma unroll
            for (uint32_t k = 0; k < NV; k++)
                if (NCH % FAR_THREADS == 0 || tid + k * FAR_THREADS < NCH) rb[h - H0][k] = gb[tid + k * FAR_THREADS];
        }
    };
    auto stage_b = [&](uint32_t buf) {
#pragma unroll
        for (int h = H0; h <= H1; h++) {
#pragma unroll
            for (uint32_t k = 0; k < NV; k++) {
                const uint32_t v = tid + k * FAR_THREADS, row = v / CH, ch = v - row * CH;
                if (NCH % FAR_THREADS == 0 || v < NCH) *(v4i *)(Bs + buf * BUF + h * TILE + row * KAP + ch * 16) = rb[h - H0][k];
            }
        }
    };
    // C values of this lane's elements of step j: cv[q][r] = row r, column
    // 16 q + (lane & 15) of the step.
    uint64_t cv[2][4];
    auto fetch_c = [&](uint32_t j, int q) {
        const uint32_t cb = step_cb(j), c = 16 * q + (lane & 15);
        const bool cok = cb + c < cbase + width;
#pragma unroll
        for (int r = 0; r < 4; r++) cv[q][r] = *((rok[r] && cok) ? mrow[r] + cb + c : tr);
    };
    fetch_b(0);
    __syncthreads();  // the previous item is done with Bs
    stage_b(0);
    fetch_c(0, 0);
    fetch_c(0, 1);
    __syncthreads();
    for (uint32_t j = 0; j < nstep; j++) {
        const uint32_t buf = j & 1;
        // the next step's tiles: fetched now, staged once the first quarters'
        // products are issued (their buffer's last reader, step j - 1, is past the
        // barrier), so that the registers are free again for the epilogue.  The
        // last step fetches and stages its own tiles again into the idle buffer
        // (no branch, so the compiler keeps the stores after the products).
        const uint32_t jn = j + 1 < nstep ? j + 1 : j;
        fetch_b(jn);
        __builtin_amdgcn_sched_barrier(0);  // the loads go out before the products
        const uint32_t cb = step_cb(j);
        const int8_t *B0 = Bs + buf * BUF + (lane & 15) * KAP + koff;
#pragma unroll
        for (int q0 = 0; q0 < 2; q0 += QW) {
            // acc[q][s] = KOFF2 + sum over the panels of sum_{la+lb=s} A_la B_lb for
            // this wave's rows and quarter q0 + q's columns
            v4i acc[QW][NC];
#pragma unroll
            for (int h = H0; h <= H1; h++) {
#pragma unroll
                for (int lb = 0; lb < L; lb++) {
                    v4i bq[QW];
#pragma unroll
                    for (int q = 0; q < QW; q++) bq[q] = *(const v4i *)(B0 + h * TILE + (q0 + q) * 16 * KAP + lb * KSTEP);
#pragma unroll
                    for (int la = 0; la < L; la++) {
                        const int s = la + lb;
                        // first product of class s: lb outer, la inner, first panel
                        const bool first = h == H0 && ((lb == 0) || (la == L - 1));
#pragma unroll
                        for (int q = 0; q < QW; q++)
                            acc[q][s] = __builtin_amdgcn_mfma_i32_16x16x64_i8(
                                af[h - H0][la], bq[q], first ? (v4i){KOFF2, KOFF2, KOFF2, KOFF2} : acc[q][s], 0, 0, 0);
                    }
                }
            }
            if (q0 == 0) {
                __builtin_amdgcn_sched_barrier(0);
                stage_b(buf ^ 1);
                __builtin_amdgcn_sched_barrier(0);
            }
            // epilogue: C - sum_s 256^s P_s (mod p), then quarter q's C values of the next step
#pragma unroll
            for (int qq = 0; qq < QW; qq++) {
                const int q = q0 + qq;
                const uint32_t c = 16 * q + (lane & 15);
                const bool cok = cb + c < cbase + width;
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    int32_t P[NC];
#pragma unroll
                    for (int k = 0; k < NC; k++) P[k] = acc[qq][k][r];
                    uint64_t v;
                    if constexpr (MODE == MODE_M61 && NC == 15)
                        v = limb_sub_m61_k2(cv[q][r], lp.kc2, P);
                    else
                        v = sub_mod(add_mod(cv[q][r], lp.k2, f.p), limb_residue_k2<MODE, NC>(f, lp, P), f.p);
                    *((rok[r] && cok) ? mrow[r] + cb + c : tr) = (T)v;
                    if constexpr (NEAR) {
                        const uint32_t gcol = it.c0 + cb + c;
                        if (rok[r] && cok && v != 0 && gcol < it.scan_end) nzcol = min(nzcol, gcol);
                    }
                }
                fetch_c(jn, q);
            }
        }
        __syncthreads();  // the next buffer is staged; this one is free
    }
}

// QW: 16-column quarters computed together per B fragment read (the limb-class
// accumulators of QW quarters are live at once: 2 only for L <= 4).
template <int L, int MODE, int QW>
__global__ __launch_bounds__(FAR_THREADS, 2) void k_update_far(RankMat *mats, uint32_t nmat, const uint32_t *item_off,
                                                               const RankPlan *plan, FieldParams f, LimbParams lp,
                                                               uint32_t par, uint64_t *trash) {
    constexpr uint32_t KA = L * KSTEP;
    __shared__ __attribute__((aligned(16))) int8_t Bs[2 * 2 * FSTEP * (KA + 16)];  // 2 buffers x 2 panel tiles
    const uint32_t total = plan->far_total;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = uni32(tid >> 6);
    const gptr<uint64_t> tr = to_global(trash) + tid;
    const gptr<uint32_t> tr32 = to_global((uint32_t *)trash) + tid;
    uint32_t nz_unused = 0;  // far items do not report nonzero columns
    for (uint32_t item = blockIdx.x; item < total; item += gridDim.x) {
        const uint32_t mi = find_mat(item_off, nmat, item);
        const RankMat &S = mats[mi];
        const IterState &I = S.it[par];
        const uint32_t lt = item - item_off[mi];
        FarItem it;
        it.x = lt & 7;
        const uint32_t band = lt >> 3;
        const uint32_t O = S.n - I.rank, pb = I.pb;
        const bool pair = uni32(I.kind) == IT_PAIR;
        update_range(I, 1, it.cbase, it.width);
        it.C = cdiv(it.width, UT);
        if ((pb == 0 && !pair) || it.x >= it.C) continue;  // uniform: no pivots, or no tile
        it.M = S.M;
        it.c0 = I.c0;
        const bool m32 = uni32(S.m32) != 0;
        // rows through the gather's snapshot (rank + r -> rows[r]): the next k_plan
        // may reorder perm while this update runs
        it.permg = to_global((const uint32_t *)S.rows);
        it.Ag[0] = to_global((const v4i *)S.AlP);
        it.Ag[1] = to_global((const v4i *)S.Al);
        it.Xg[0] = to_global((const v4i *)S.XlP) + (size_t)(pair ? I.dxoff : 0u) * (KA / 16);
        it.Xg[1] = to_global((const v4i *)S.Xl);
        it.r0 = band * UT;
        it.nr = min(UT, O - it.r0);
        it.rank = 0;
        it.ld = S.ld;
        if (m32) {  // p < 2^32
            if (!pair)
                far_item<L, MODE, QW, 1, 1>(it, Bs, f, lp, tr32, tid, lane, wave * 16, nz_unused);
            else if (pb > 0)
                far_item<L, MODE, QW, 0, 1>(it, Bs, f, lp, tr32, tid, lane, wave * 16, nz_unused);
            else
                far_item<L, MODE, QW, 0, 0>(it, Bs, f, lp, tr32, tid, lane, wave * 16, nz_unused);
        } else {
            if (!pair)
                far_item<L, MODE, QW, 1, 1>(it, Bs, f, lp, tr, tid, lane, wave * 16, nz_unused);
            else if (pb > 0)
                far_item<L, MODE, QW, 0, 1>(it, Bs, f, lp, tr, tid, lane, wave * 16, nz_unused);
            else
                far_item<L, MODE, QW, 0, 0>(it, Bs, f, lp, tr, tid, lane, wave * 16, nz_unused);
        }
    }
}

// Near part of the update on the far update's machinery: one 64 x 64 tile per
// item (the near part is one or two tiles wide and O rows tall), the wave's row
// digits in registers, the tile's two 32-column halves of X digits staged in LDS
// (double-buffered), C values fetched a half ahead; nonzero results below
// scan_end are reported (the missed-pivot check).  A tile of an iteration
// without pivots is only scanned for a nonzero.
template <int L, int MODE, int QW>
__global__ __launch_bounds__(FAR_THREADS, 2) void k_update_near(RankMat *mats, uint32_t nmat, const uint32_t *upd_off,
                                                                const RankPlan *plan, FieldParams f, LimbParams lp,
                                                                uint32_t par, uint64_t *trash) {
    constexpr uint32_t KA = L * KSTEP;
    __shared__ __attribute__((aligned(16))) int8_t Bs[2 * 2 * FSTEP * (KA + 16)];
    const uint32_t total = plan->near_total;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = uni32(tid >> 6);
    const gptr<uint64_t> tr = to_global(trash) + tid;
    const gptr<uint32_t> tr32 = to_global((uint32_t *)trash) + tid;
    for (uint32_t item = blockIdx.x; item < total; item += gridDim.x) {
        const uint32_t mi = find_mat(upd_off, nmat, item);
        RankMat &S = mats[mi];
        const IterState &I = S.it[par];
        const uint32_t lt = item - upd_off[mi];
        const uint32_t O = S.n - I.rank;
        uint32_t cbase, width, r0, cb, nr, nc;
        update_range(I, 0, cbase, width);
        if (!tile_geom(O, cbase, width, 0, lt, r0, cb, nr, nc)) continue;  // uniform: padding tile
        uint32_t nzcol = 0xffffffffu;
        if (I.pb == 0) {
            // detection only: thread = column tid % 64, rows tid / 64 + 4 k, every
            // load issued before any test; a tile right of a nonzero column already
            // found has nothing to add
            const uint32_t known = __hip_atomic_load(&S.first_nz_col, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const uint32_t c = tid & 63, i0 = tid >> 6, gcol = I.c0 + cb + c;
            if (I.c0 + cb <= known && c < nc && gcol < I.scan_end) {
                constexpr uint32_t RK = UT / (FAR_THREADS / 64);
                size_t idx[RK];
#pragma unroll
                for (uint32_t k = 0; k < RK; k++) {
                    const uint32_t i = i0 + (FAR_THREADS / 64) * k;
                    idx[k] = (size_t)S.perm[I.rank + r0 + min(i, nr - 1)] * S.ld + gcol;
                }
                bool any = false;
                if (S.m32) {
                    uint32_t v[RK];
#pragma unroll
                    for (uint32_t k = 0; k < RK; k++) v[k] = ((const uint32_t *)S.M)[idx[k]];
#pragma unroll
                    for (uint32_t k = 0; k < RK; k++) any |= i0 + (FAR_THREADS / 64) * k < nr && v[k] != 0;
                } else {
                    uint64_t v[RK];
#pragma unroll
                    for (uint32_t k = 0; k < RK; k++) v[k] = S.M[idx[k]];
#pragma unroll
                    for (uint32_t k = 0; k < RK; k++) any |= i0 + (FAR_THREADS / 64) * k < nr && v[k] != 0;
                }
                if (any) nzcol = gcol;
            }
        } else {
            FarItem it;
            it.x = (cb - cbase) / UT;
            it.cbase = cbase;
            it.width = width;
            it.C = cdiv(width, UT);
            it.M = S.M;
            it.c0 = I.c0;
            it.permg = to_global((const uint32_t *)S.perm);
            it.Ag[0] = it.Ag[1] = to_global((const v4i *)S.Al);
            it.Xg[0] = it.Xg[1] = to_global((const v4i *)own_xl(S, I));
            it.r0 = r0;
            it.nr = nr;
            it.rank = I.rank;
            it.ld = S.ld;
            it.scan_end = I.scan_end;
            if (uni32(S.m32))
                far_item<L, MODE, QW, 1, 1, uint32_t, true>(it, Bs, f, lp, tr32, tid, lane, wave * 16, nzcol);
            else
                far_item<L, MODE, QW, 1, 1, uint64_t, true>(it, Bs, f, lp, tr, tid, lane, wave * 16, nzcol);
        }
        if (I.resid && nzcol != 0xffffffffu) atomicMin(&S.first_nz_col, nzcol);
    }
}

// Launchers: the digit gather and the two parts of the MFMA update.
template <int L, int MODE>
static void launch_mfma_L(hipStream_t st, RankMat *dmats, uint32_t nmat, const FieldParams &f, const LimbParams &lp,
                          const RankLaunch &rl, uint32_t par, int what) {
    if (what < 0)  // -1: rows and near X; -2: far X
        hipLaunchKernelGGL((k_gather_split<L, MODE>), dim3(rl.grid_gx), dim3(512), 0, st, dmats, nmat,
                           what == -1 ? rl.gx_off : rl.xf_off, rl.plan, f, lp, par, what == -1 ? 0u : 1u);
    else if (what == 0 && L <= 4 && rl.far_qw != 1)
        hipLaunchKernelGGL((k_update_near<L, MODE, (L <= 4 ? 2 : 1)>), dim3(rl.grid_near), dim3(FAR_THREADS), 0, st,
                           dmats, nmat, rl.near_off, rl.plan, f, lp, par, rl.trash);
    else if (what == 0)
        hipLaunchKernelGGL((k_update_near<L, MODE, 1>), dim3(rl.grid_near), dim3(FAR_THREADS), 0, st, dmats, nmat,
                           rl.near_off, rl.plan, f, lp, par, rl.trash);
    else if (L <= 4 && rl.far_qw != 1)
        hipLaunchKernelGGL((k_update_far<L, MODE, (L <= 4 ? 2 : 1)>), dim3(rl.grid_far), dim3(FAR_THREADS), 0, st, dmats,
                           nmat, rl.far_off, rl.plan, f, lp, par, rl.trash);
    else
        hipLaunchKernelGGL((k_update_far<L, MODE, 1>), dim3(rl.grid_far), dim3(FAR_THREADS), 0, st, dmats, nmat,
                           rl.far_off, rl.plan, f, lp, par, rl.trash);
}

// what = -1: digit gather and the near X product; -2: the far X product; 0: near
// update; 1: far update.
void launch_mfma(hipStream_t st, RankMat *dmats, uint32_t nmat, const FieldParams &f, const LimbParams &lp,
                 const RankLaunch &rl, uint32_t par, int what) {
#define MDL_L(LL) launch_mfma_L<LL, MODE_GENERIC>(st, dmats, nmat, f, lp, rl, par, what)
    if (f.mode == MODE_M61)
        launch_mfma_L<8, MODE_M61>(st, dmats, nmat, f, lp, rl, par, what);
    else if (f.mode == MODE_M31)
        launch_mfma_L<4, MODE_M31>(st, dmats, nmat, f, lp, rl, par, what);
    else {
        switch (lp.L) {
            case 1: MDL_L(1); break;
            case 2: MDL_L(2); break;
            case 3: MDL_L(3); break;
            case 4: MDL_L(4); break;
            case 5: MDL_L(5); break;
            case 6: MDL_L(6); break;
            case 7: MDL_L(7); break;
            default: MDL_L(8); break;
        }
    }
#undef MDL_L
}

}  // namespace mdl
