// Rank over F_p on gfx950: the blocked right-looking elimination of the
// reference (src/fp.rs:116-255) restructured for the GPU.
//
// One iteration handles one column panel [col, jend) of every active matrix:
//
//   k_block_ge   one workgroup per matrix.  Loads the top block (the next BMAX logical rows) of the panel into
//                registers and eliminates it there: a wave ballot + ffs finds the pivot
//                row (the first block row not yet a pivot with a nonzero in the
//                column), and the elimination is fraction-free (pi*x - e*y), so no
//                modular inverse sits on the per-column critical path.  It records
//                the row operations for k_block_z.
//   k_block_z    Z = A^-1 for A = M[R, J] from that record; pb independent
//                inverses (Fermat, fp.rs:111-113) off the critical chain.
//   k_plan       prefix sums of the per-matrix tile counts (the work queue).
//   k_gather_x   Ab = M[O, J] (gathered multipliers' source) and
//                X = Z M[R, c0:] (the reference's U-block solve, fp.rs:180-200).
//   k_update     S = M[O, c0:] - Ab X mod p, the trailing update (fp.rs:202-248),
//                with lazy 128-bit accumulation and one Barrett reduction per
//                safe_batch products, exactly as the reference accumulates.
//   k_finish     advances the panel; handles a pivot the top block missed.
//
// Exactness.  For any pivot rows R and columns J with A = M[R, J] invertible,
// rank(M) = |R| + rank(M[O,:] - M[O,J] A^-1 M[R,:]) (Schur complement).  The top
// block is a prefix of the logical row order, so whenever it has a nonzero in a
// column the reference's first nonzero row is inside it.  If the block has fewer
// than b pivots and rows exist below it, the update also covers the panel columns
// and any nonzero left there (a pivot the block could not see) sends the next
// panel back to that column, with the first nonzero row swapped to the top --
// the reference's own pivot for that column.
#include "rank_common.hpp"

namespace mdl {

// The value of lane `src` (wave-uniform) in every lane.
__device__ __forceinline__ uint64_t lane_bcast(uint64_t v, uint32_t src) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, (int)src);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), (int)src);
    return ((uint64_t)hi << 32) | lo;
}

// Panel steps keep 2^61 - 1 residues lazily in [0, p + 7] (one fold per step, no
// final compare); other moduli stay canonical.  Records are canonical.
template <int MODE>
__device__ __forceinline__ uint64_t canon_t(uint64_t v) {
    return MODE == MODE_M61 ? (v >= P61 ? v - P61 : v) : v;
}
// pi x - e y for canonical pi, e (x, y lazy for 2^61 - 1)
template <int MODE>
__device__ __forceinline__ uint64_t ff_lazy_t(const FieldParams &f, uint64_t pi, uint64_t x, uint64_t e, uint64_t y) {
    if constexpr (MODE == MODE_M61)
        return ff_step_m61_lazy(pi, x, e, y);
    else
        return ff_step_t<MODE>(f, pi, x, e, y);
}

// ---------------------------------------------------------------------------
// finish_iter: the end of an iteration (parity par) of matrix S, once its near
// update has run -- advance the panel, bring a missed pivot to the top of the
// block, retire a finished matrix.  It runs once per iteration (S.fin < S.iters:
// the last iteration's finish is pending).  Every thread of the workgroup calls
// it; it ends with a barrier.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void finish_iter(RankMat &S, uint32_t *active, uint32_t par, uint32_t &best) {
    const uint32_t tid = threadIdx.x;
    if (S.done || S.fin == S.iters) return;  // uniform: nothing written yet
    const IterState &I = S.it[par];
    const uint32_t n = S.n;
    const uint32_t jstar = S.first_nz_col;
    const bool miss = I.resid && jstar < I.scan_end;
    if (miss) {
        if (tid == 0) best = 0xffffffffu;
        __syncthreads();
        uint32_t mine = 0xffffffffu;
        for (uint32_t q = S.rank + tid; q < n; q += blockDim.x) {
            if (m_get(S, (size_t)S.perm[q] * S.ld + jstar) != 0) { mine = q; break; }
        }
        if (mine != 0xffffffffu) atomicMin(&best, mine);
        __syncthreads();
        // The far part of this iteration's update may still run (the plain method's
        // reads perm), so the swap to position rank is only recorded here; the next
        // k_block_ge reads through it and the next k_plan commits it.
        if (tid == 0 && best != 0xffffffffu && best != S.rank) S.pend = best;
    }
    __syncthreads();  // every thread has read the state
    if (tid == 0) {
        S.col = miss ? jstar : I.scan_end;  // (scan_end = n: no nonzero left, the rank is final)
        S.fin = S.iters;
        if (S.col >= n || S.rank >= n) {
            S.done = 1;
            atomicSub(active, 1u);
        }
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// k_block_ge: pivot search and fraction-free elimination of the panel part of
// the top block
// ---------------------------------------------------------------------------
// The block elimination runs on GE_WAVES waves, two per SIMD: the step loop is
// bound by VALU issue, every wave repeats the pivot search, and two waves per
// SIMD are what keeps a SIMD issuing.  Wave v keeps the panel columns v + 8 m.
constexpr uint32_t GE_WAVES = 8, GE_COLS = BMAX / GE_WAVES;

// One fraction-free step of a block row: x <- pi x - e y where the row takes it
// (act), x's residue where it does not.  For 2^61 - 1 the choice is folded into
// the coefficients, (b, d) = (pi, p - e) or (1, 0), in 31-bit halves per lane:
// lin2_m61_lazy_acc then needs no select, and with (1, 0) it returns x for
// x < 2^61 and x - p for x in [2^61, p + 7].  Results stay in [0, p + 7].
template <int MODE>
struct FFStep {
    uint64_t pi, e;
    bool act;
    uint32_t b0, b1, d0, d1;
    __device__ __forceinline__ FFStep(uint64_t pi_, uint64_t e_, bool act_)
        : pi(pi_), e(e_), act(act_), b0(0), b1(0), d0(0), d1(0) {
        if constexpr (MODE == MODE_M61) {
            const uint64_t ne = P61 - e;
            b0 = act ? ((uint32_t)pi & 0x7fffffffu) : 1u;
            b1 = act ? (uint32_t)(pi >> 31) : 0u;
            d0 = act ? ((uint32_t)ne & 0x7fffffffu) : 0u;
            d1 = act ? (uint32_t)(ne >> 31) : 0u;
        }
    }
    __device__ __forceinline__ uint64_t operator()(const FieldParams &f, uint64_t x, uint64_t y) const {
        if constexpr (MODE == MODE_M61)
            return lin2_m61_lazy_acc(x, y, b0, b1, d0, d1);
        else
            return act ? ff_step_t<MODE>(f, pi, x, e, y) : x;
    }
};

// A step on the wave's panel columns wave + 8 m for m >= M0 (those after the
// pivot column) and, with OWN, on the own column.  Their products are independent
// and interleave (for 2^61 - 1 written stage by stage across the columns, so the
// multiply-add chains interleave in the instruction stream too).
template <int MODE, int M0, bool OWN>
__device__ __forceinline__ void ge_columns(const FieldParams &f, const FFStep<MODE> &st,
                                           uint64_t (&x)[GE_COLS], uint64_t &own, uint32_t q) {
    constexpr int NC = GE_COLS, NS = OWN ? NC + 1 : NC;  // slot NC: the own column (y = 0)
    uint64_t y[NC];
#pragma unroll
    for (int m = M0; m < NC; m++) y[m] = lane_bcast(x[m], q);
    if constexpr (MODE == MODE_M61) {
        uint32_t x0[NS], x1[NS], y0[NS], y1[NS];
        uint64_t t[NS], s[NS];
#pragma unroll
        for (int m = M0; m < NS; m++) {
            const uint64_t xv = m < NC ? x[m] : own, yv = m < NC ? y[m] : 0;
            x0[m] = (uint32_t)xv & 0x7fffffffu;
            x1[m] = (uint32_t)(xv >> 31);
            y0[m] = (uint32_t)yv & 0x7fffffffu;
            y1[m] = (uint32_t)(yv >> 31);
        }
#pragma unroll
        for (int m = M0; m < NS; m++) t[m] = (uint64_t)x0[m] * st.b1;
#pragma unroll
        for (int m = M0; m < NS; m++) s[m] = (uint64_t)x0[m] * st.b0;
#pragma unroll
        for (int m = M0; m < NS; m++) t[m] += (uint64_t)x1[m] * st.b0;
#pragma unroll
        for (int m = M0; m < NS; m++) s[m] += (uint64_t)x1[m] * (2u * st.b1);
#pragma unroll
        for (int m = M0; m < NS; m++) t[m] += (uint64_t)y0[m] * st.d1;
#pragma unroll
        for (int m = M0; m < NS; m++) s[m] += (uint64_t)y0[m] * st.d0;
#pragma unroll
        for (int m = M0; m < NS; m++) t[m] += (uint64_t)y1[m] * st.d0;
#pragma unroll
        for (int m = M0; m < NS; m++) s[m] += (uint64_t)y1[m] * (2u * st.d1);
#pragma unroll
        for (int m = M0; m < NS; m++) {
            // s < 1.5 2^63, t < 2^63: the sum below is < 2^64 (lin2_m61_lazy_acc)
            const uint64_t v = s[m] + (t[m] >> 30) + ((t[m] & 0x3fffffffull) << 31);
            const uint64_t z = (v & P61) + (v >> 61);  // <= p + 7
            if (m < NC)
                x[m] = z;
            else
                own = z;
        }
    } else {
#pragma unroll
        for (int m = M0; m < NC; m++) x[m] = st(f, x[m], y[m]);
        if constexpr (OWN) own = st(f, own, 0);
    }
}

__device__ void ge_commit(RankMat &S, uint32_t par, uint32_t defer, uint32_t pb, uint64_t piv, const uint32_t *rowid,
                          const uint32_t *Jl, const uint32_t *Qk, uint32_t rank0, uint32_t col, uint32_t jend,
                          uint32_t w, uint32_t Bt, uint32_t dpend);

template <int MODE>
__global__ __launch_bounds__(64 * GE_WAVES) void k_block_ge(RankMat *mats, FieldParams f, uint32_t b, uint32_t par,
                                                            uint32_t defer) {
    RankMat &S = mats[blockIdx.x];
    // U: one padded row per block row: the block as loaded, then, after the
    // elimination, its final state (panel columns, and the own column COWN).
    constexpr uint32_t COWN = BMAX, UW = BMAX + 1;
    __shared__ uint64_t U[BMAX][UW];
    __shared__ uint64_t colb[2][BMAX];  // the column of the next pivot search, double-buffered
    __shared__ uint64_t piK[BMAX];
    __shared__ uint32_t rowid[BMAX], Jl[BMAX], Qk[BMAX];
    if (S.done) return;

    const uint32_t tid = threadIdx.x;
    const uint32_t n = S.n, ld = S.ld, rank0 = S.rank, col = S.col;
    const uint32_t dpend = S.dpend;  // read before the barriers below; thread 0 clears it at the end
    const uint32_t jend = min(col + b, n);
    const uint32_t w = jend - col;
    const uint32_t Bt = min((uint32_t)BMAX, n - rank0);

    // Rows of the top block; a missed pivot found by the last k_finish moves to
    // position rank0 (the swap itself is committed to perm by k_plan).
    const uint32_t pend = S.pend;
    if (tid < Bt) {
        const uint32_t pos = rank0 + tid;
        const uint32_t src = (pend == 0xffffffffu) ? pos : (pos == rank0 ? pend : (pos == pend ? rank0 : pos));
        rowid[tid] = S.perm[src];
    }
    __syncthreads();
    {
        // GE_WAVES lanes per block row, GE_COLS panel columns each: all loads issued first.
        const uint32_t q = tid / GE_WAVES, c0 = tid % GE_WAVES;
        const size_t src = (size_t)rowid[q < Bt ? q : 0] * ld + col;
        uint64_t v[GE_COLS];
#pragma unroll
        for (int m = 0; m < (int)GE_COLS; m++) {
            const uint32_t c = c0 + GE_WAVES * m;
            v[m] = (q < Bt && c < w) ? m_get(S, src + c) : 0;
        }
#pragma unroll
        for (int m = 0; m < (int)GE_COLS; m++) U[q][c0 + GE_WAVES * m] = v[m];
    }
    __syncthreads();

    // Rows stay where they are (lane = block row).  The pivot of column j is the
    // first block row that is not yet a pivot and has a nonzero in column j: any
    // pivot choice gives the rank (the reference takes the first nonzero row at or
    // below the current position and swaps it up, fp.rs:150-163).  Wave v keeps
    // the panel columns v + 8 m (m < 8) in registers for the whole elimination,
    // and wave 0 also the own column (the coefficient on the row's own original
    // row), so the waves' counts of columns after the pivot column differ by at
    // most one at every step.  The pivot row's entries come from its lane by
    // readlane; only the column of the next search goes through LDS: its owner
    // publishes it at the end of a step, and one barrier per step separates that
    // from the reads (two buffers, so the next column's write cannot meet a late
    // read of this one).
    const uint32_t lane = tid & 63, wave = uni32(tid >> 6);
    const bool inblk = lane < Bt;
    uint64_t x[GE_COLS];
#pragma unroll
    for (int m = 0; m < (int)GE_COLS; m++) x[m] = U[lane][wave + GE_WAVES * m];
    uint64_t own = inblk ? 1 % f.p : 0;
    if (wave == 0) colb[0][lane] = x[0];
    __syncthreads();  // (also: every read of the loaded block precedes the final state's writes)

    uint64_t *eta = S.Eta;
    uint64_t piv = 0;  // block rows that are pivots (the same in every lane)
    uint32_t k = 0;
    for (uint32_t j = 0; j < w && k < Bt; j++) {
        const uint64_t colv = colb[j & 1][lane];  // canonical (published so)
        const bool free_row = inblk && !((piv >> lane) & 1);
        const uint64_t mask = __ballot(free_row && colv != 0);
        if (mask != 0) {  // uniform (else no pivot in this column)
            const uint32_t q = __ffsll((unsigned long long)mask) - 1;
            // Fraction-free step on the other free rows with e != 0 (fp.rs:166-174
            // without the division): x <- pi x - e y on the panel columns after j,
            // own scaled by pi.  A pivot row takes no step from then on, so its
            // entries after its column, and its own, are final: the record reads
            // them at the end.
            const uint64_t pi = lane_bcast(colv, q);
            const uint64_t e = colv;
            const bool act = free_row && lane != q && e != 0;
            const FFStep<MODE> st(pi, e, act);
            // this wave's columns after j: wave + 8 m for m >= mlo
            const uint32_t mlo = j < wave ? 0u : (j - wave) / GE_WAVES + 1;
            if (wave == 0) {
                eta[ETA_E + k * BMAX + lane] = act ? e : 0;  // the step's multipliers
                if (lane == 0) {
                    Jl[k] = j;
                    piK[k] = pi;
                    Qk[k] = q;
                }
                switch (mlo) {
                    case 0: ge_columns<MODE, 0, true>(f, st, x, own, q); break;
                    case 1: ge_columns<MODE, 1, true>(f, st, x, own, q); break;
                    case 2: ge_columns<MODE, 2, true>(f, st, x, own, q); break;
                    case 3: ge_columns<MODE, 3, true>(f, st, x, own, q); break;
                    case 4: ge_columns<MODE, 4, true>(f, st, x, own, q); break;
                    case 5: ge_columns<MODE, 5, true>(f, st, x, own, q); break;
                    case 6: ge_columns<MODE, 6, true>(f, st, x, own, q); break;
                    case 7: ge_columns<MODE, 7, true>(f, st, x, own, q); break;
                    default: ge_columns<MODE, 8, true>(f, st, x, own, q); break;
                }
            } else {
                switch (mlo) {
                    case 0: ge_columns<MODE, 0, false>(f, st, x, own, q); break;
                    case 1: ge_columns<MODE, 1, false>(f, st, x, own, q); break;
                    case 2: ge_columns<MODE, 2, false>(f, st, x, own, q); break;
                    case 3: ge_columns<MODE, 3, false>(f, st, x, own, q); break;
                    case 4: ge_columns<MODE, 4, false>(f, st, x, own, q); break;
                    case 5: ge_columns<MODE, 5, false>(f, st, x, own, q); break;
                    case 6: ge_columns<MODE, 6, false>(f, st, x, own, q); break;
                    case 7: ge_columns<MODE, 7, false>(f, st, x, own, q); break;
                    default: break;
                }
            }
            piv |= 1ull << q;
            k++;
        }
        const uint32_t jn = j + 1;
        if (jn < w && wave == jn % GE_WAVES) {
            uint64_t v = x[0];
#pragma unroll
            for (int m = 1; m < (int)GE_COLS; m++)
                if (jn / GE_WAVES == (uint32_t)m) v = x[m];
            colb[jn & 1][lane] = canon_t<MODE>(v);
        }
        __syncthreads();
    }
    const uint32_t pb = k;
    // the final state, for the record
#pragma unroll
    for (int m = 0; m < (int)GE_COLS; m++) U[lane][wave + GE_WAVES * m] = x[m];
    if (wave == 0) U[lane][COWN] = own;
    __syncthreads();

    // The record for k_block_z, which builds A^-1 from it: the multipliers (above,
    // by step), the pivot rows' entries at the later pivot columns, U[r][s] =
    // entry of pivot row r at column J_s for s > r, and the pivot rows' own.
    for (uint32_t idx = tid; idx < pb * BMAX; idx += blockDim.x) {
        const uint32_t r = idx / BMAX, s = idx % BMAX;
        eta[ETA_U + idx] = (s > r && s < pb) ? canon_t<MODE>(U[Qk[r]][Jl[s]]) : 0;
    }
    if (tid < pb) {  // (the pivots' inverses are computed by k_block_z, off this kernel's path)
        eta[ETA_PI + tid] = piK[tid];
        eta[ETA_OWNK + tid] = canon_t<MODE>(U[Qk[tid]][COWN]);
        eta[ETA_Q + tid] = Qk[tid];
    }
    ge_commit(S, par, defer, pb, piv, rowid, Jl, Qk, rank0, col, jend, w, Bt, dpend);
}

// The block elimination in its LDS form, used for the largest matrices (sides of
// GE_LDS_MIN_SIDE and more): 16 waves, the block in LDS, each step's columns read
// from and written back to it, a barrier per step.  It is slower on its own than
// k_block_ge (about 90 against 80 us a panel on a 65536^2 matrix), but with it
// the far update that runs beside the panel chain takes 4.68 instead of 5.01 ms a
// launch there (the 65536^2 planted-corank rank: 4.95 against 5.38 s; measured,
// the cause not found), while on the 64-matrix batch and 8192^2 k_block_ge is
// faster overall (17.2 against 17.5 ms, 22.7 against 23.7 ms).  The record and
// the iteration state are the same.
template <int MODE>
__device__ __forceinline__ bool nz_t(uint64_t v) {
    return MODE == MODE_M61 ? (v != 0 && v != P61) : v != 0;
}

template <int MODE>
__global__ __launch_bounds__(1024) void k_block_ge_lds(RankMat *mats, FieldParams f, uint32_t b, uint32_t par,
                                                       uint32_t defer) {
    RankMat &S = mats[blockIdx.x];
    // One LDS row per block row: [0, BMAX) panel entries, then own (the
    // coefficient on the row's own original row).
    constexpr uint32_t COWN = BMAX, UW = BMAX + 1;
    __shared__ uint64_t U[BMAX][UW];
    __shared__ uint64_t Et[BMAX][BMAX];  // multiplier e of each block row at each step
    __shared__ uint64_t piK[BMAX], ownK[BMAX];
    __shared__ uint32_t rowid[BMAX], Jl[BMAX], Qk[BMAX];
    if (S.done) return;

    const uint32_t tid = threadIdx.x;
    const uint32_t n = S.n, ld = S.ld, rank0 = S.rank, col = S.col;
    const uint32_t dpend = S.dpend;  // read before the barriers below; thread 0 clears it at the end
    const uint32_t jend = min(col + b, n);
    const uint32_t w = jend - col;
    const uint32_t Bt = min((uint32_t)BMAX, n - rank0);
    const uint32_t pend = S.pend;  // (as in k_block_ge)
    if (tid < Bt) {
        const uint32_t pos = rank0 + tid;
        const uint32_t src = (pend == 0xffffffffu) ? pos : (pos == rank0 ? pend : (pos == pend ? rank0 : pos));
        rowid[tid] = S.perm[src];
    }
    for (uint32_t idx = tid; idx < BMAX * BMAX; idx += blockDim.x) Et[idx / BMAX][idx % BMAX] = 0;
    __syncthreads();
    {
        // 16 lanes per block row, 4 panel columns each: all loads issued first.
        const uint32_t q = tid >> 4, c0 = tid & 15;
        const size_t src = (size_t)rowid[q < Bt ? q : 0] * ld + col;
        uint64_t v[4];
#pragma unroll
        for (int m = 0; m < 4; m++) {
            const uint32_t c = c0 + 16 * m;
            v[m] = (q < Bt && c < w) ? m_get(S, src + c) : 0;
        }
#pragma unroll
        for (int m = 0; m < 4; m++) U[q][c0 + 16 * m] = v[m];
        if (c0 == 0) U[q][COWN] = q < Bt ? 1 % f.p : 0;
    }
    __syncthreads();

    // Wave v owns the panel columns v, v + 16, v + 32, v + 48 (and, at step j, the
    // own column if v = j mod 16), lane = block row; the pivot search and the
    // fraction-free step are those of k_block_ge.  A step reads column j and the
    // lane's own columns before the ballot; row q and column j are only read, so
    // no barrier separates the reads from the writes.
    const uint32_t lane = tid & 63, wave = uni32(tid >> 6);
    const bool inblk = lane < Bt;
    uint64_t piv = 0;
    uint32_t k = 0;
    for (uint32_t j = 0; j < w && k < Bt; j++) {
        const uint64_t colv = U[lane][j];
        const bool own = wave == (j & 15);
        uint64_t xv[5];
#pragma unroll
        for (int m = 0; m < 4; m++) {
            const uint32_t c = wave + 16 * m;
            if (c > j && c < w) xv[m] = U[lane][c];  // uniform over the wave
        }
        if (own) xv[4] = U[lane][COWN];
        const bool free_row = inblk && !((piv >> lane) & 1);
        const uint64_t mask = __ballot(free_row && nz_t<MODE>(colv));
        if (mask == 0) continue;  // uniform: no pivot in this column
        const uint32_t q = __ffsll((unsigned long long)mask) - 1;
        const uint64_t pi = canon_t<MODE>(lane_bcast(colv, q));
        const uint64_t e = canon_t<MODE>(colv);
        const bool act = free_row && lane != q && e != 0;
        if (wave == 0 && act) Et[k][lane] = e;
#pragma unroll
        for (int m = 0; m < 4; m++) {
            const uint32_t c = wave + 16 * m;
            if (c > j && c < w) {
                const uint64_t y = lane_bcast(xv[m], q);
                if (act) U[lane][c] = ff_lazy_t<MODE>(f, pi, xv[m], e, y);
            }
        }
        if (own && act) U[lane][COWN] = ff_lazy_t<MODE>(f, pi, xv[4], e, 0);
        if (tid == 0) {
            Jl[k] = j;
            piK[k] = pi;
            ownK[k] = canon_t<MODE>(U[q][COWN]);  // row q takes no step now
            Qk[k] = q;
        }
        __syncthreads();
        piv |= 1ull << q;
        k++;
    }
    const uint32_t pb = k;

    // The record (the layout of k_block_ge's): a pivot row is final once it is the
    // pivot, so U[r][s] = U[Q[r]][J_s] for s > r.
    uint64_t *eta = S.Eta;
    for (uint32_t idx = tid; idx < pb * BMAX; idx += blockDim.x) {
        const uint32_t r = idx / BMAX, s = idx % BMAX;
        eta[ETA_E + idx] = Et[r][s];
        eta[ETA_U + idx] = (s > r && s < pb) ? canon_t<MODE>(U[Qk[r]][Jl[s]]) : 0;
    }
    if (tid < pb) {
        eta[ETA_PI + tid] = piK[tid];
        eta[ETA_OWNK + tid] = ownK[tid];
        eta[ETA_Q + tid] = Qk[tid];
    }
    ge_commit(S, par, defer, pb, piv, rowid, Jl, Qk, rank0, col, jend, w, Bt, dpend);
}

// The end of a block elimination (either form), every thread of the workgroup:
// the pivots and the block's new row order into the iteration's state, and the
// iteration kind with its column ranges.
__device__ void ge_commit(RankMat &S, uint32_t par, uint32_t defer, uint32_t pb, uint64_t piv, const uint32_t *rowid,
                          const uint32_t *Jl, const uint32_t *Qk, uint32_t rank0, uint32_t col, uint32_t jend,
                          uint32_t w, uint32_t Bt, uint32_t dpend) {
    const uint32_t tid = threadIdx.x, n = S.n;
    // The new order of the block -- the pivot rows in pivot order, then the other
    // rows in block order -- is applied to perm by k_plan.
    IterState &I = S.it[par];
    if (tid < pb) {
        const uint32_t r = rowid[Qk[tid]];
        I.J[tid] = col + Jl[tid];
        I.R[tid] = r;
        I.top[tid] = r;
    }
    if (tid < Bt && !((piv >> tid) & 1)) I.top[pb + __popcll(~piv & ((1ull << tid) - 1))] = rowid[tid];
    // Iteration kind (MFMA method): a pending far contribution makes this a PAIR
    // iteration; otherwise a clean panel (pivots, no missed-pivot check) with far
    // columns beyond the next two panels and rows left below DEFERs its own.
    // Odd-numbered matrices of a batch start their pairs one iteration later, so
    // that every iteration of the batch has far work to overlap the next panel
    // eliminations.
    const bool below = (n - rank0) > Bt;
    const uint32_t resid = (pb < w && below) ? 1u : 0u;
    const bool stagger = S.iters == 0 && (blockIdx.x & 1);
    const uint32_t kind = dpend ? IT_PAIR
                          : (defer && !stagger && pb > 0 && !resid && jend + 2 * BMAX < n && rank0 + pb < n)
                              ? IT_DEFER
                              : IT_NORMAL;
    if (kind == IT_PAIR && tid < BMAX) I.dJ[tid] = S.dJ[tid];
    if (kind == IT_DEFER && tid < pb) S.dJ[tid] = col + Jl[tid];
    if (tid == 0) {
        I.rank0 = rank0;
        I.rank = rank0 + pb;
        I.Bt = Bt;
        I.jend = jend;
        I.pb = pb;
        I.resid = resid;
        I.kind = kind;
        I.scan_end = jend;
        if (kind == IT_PAIR) {
            // near part [c0, dcol): this panel only; far part [dcol, n): both panels
            I.split = S.dcol;
            I.c0 = pb > 0 ? (resid ? col : jend) : (resid ? col : S.dcol);
            I.cend = n;
            I.dpb = S.dpb;
            I.dcol = S.dcol;
            I.dxoff = I.c0 - S.dc0;
            S.dpend = 0;
        } else if (kind == IT_DEFER) {
            I.split = jend + 2 * BMAX;  // near part only; the far part is pending
            I.c0 = jend;
            I.cend = n;
            S.dpend = 1;
            S.dpb = pb;
            S.dcol = jend + 2 * BMAX;
            S.dc0 = jend;
        } else {
            I.split = jend + BMAX;
            if (pb > 0) {
                I.c0 = resid ? col : jend;
                I.cend = n;
            } else if (resid) {
                // No pivot in the top block and no update to apply: search every
                // column left (the rows below the block included), so that a run of
                // eliminated columns -- the whole rest of a rank-deficient matrix --
                // costs one pass instead of one iteration per panel.  The first
                // nonzero column is the reference's next pivot column, the first
                // nonzero row in it its pivot row (fp.rs:150-163).
                I.c0 = col;
                I.cend = n;
                I.split = n;
                I.scan_end = n;
            } else {
                I.c0 = col;
                I.cend = col;  // nothing at all
            }
        }
        S.rank = rank0 + pb;
        S.first_nz_col = 0xffffffffu;
        S.iters++;
    }
}

// ---------------------------------------------------------------------------
// k_block_z: Z = A^-1 for the pivot block A = M[R, J] of the last elimination.
//
// The elimination's row operations, applied to the identity, give F with
// F A = U (upper triangular in pivot-slot coordinates, diagonal = the pivots).
// F's column for slot s (the pivot row q_s's coefficient in every block row)
// starts at step s as own[s] in row q_s and is then carried through the later
// steps (x <- pi x - e y, y the step's pivot row).  Then Z = U^-1 F.  Every slot
// (a column of F and of Z) is independent: one wave per slot, lane = block row
// for the forward chain, with the step's pivot row broadcast from its lane -- no
// barriers.  One more wave per workgroup inverts the pivots (Fermat, in
// lockstep) while the slot waves run their forward chains.  The back
// substitution then runs on V = D^-1 U (unit diagonal) and g = D^-1 f, so that a
// step of its dependent chain is one product added to lazy sums:
// z_t = g_t - sum_{u > t} V[t][u] z_u.
// ---------------------------------------------------------------------------

// Balanced signed digits of Z[k][s] for the matrix-core X product (L > 0):
// Zl[k][la*64 + s] = digit la, zero outside the pb x pb block.
__device__ __forceinline__ void z_digits(int8_t *Zl, uint32_t L, uint32_t k, uint32_t s, uint64_t z, uint64_t p) {
    int64_t v = balanced(z, p);
    for (uint32_t la = 0; la < L; la++) {
        const int8_t d = (int8_t)(v & 0xff);
        Zl[(size_t)k * L * KSTEP + la * KSTEP + s] = d;
        v = (v - d) >> 8;
    }
}

__device__ void plan_body(RankMat *mats, uint32_t nmat, uint32_t *gx_off, uint32_t *near_off, uint32_t *far_off,
                          uint32_t *xf_off, RankPlan *plan, uint32_t par, uint32_t far_items);

// with_plan: one more workgroup (the last) computes the iteration's plan beside
// the slot workgroups -- the plan reads only k_block_ge's record and the matrices'
// state, none of which this kernel writes, and this kernel reads nothing it writes.
template <int MODE>
__global__ __launch_bounds__(64 * (ZSLOTS + 1)) void k_block_z(RankMat *mats, FieldParams f, uint32_t par,
                                                               uint32_t L, uint32_t nmat, RankLaunch rl,
                                                               uint32_t with_plan) {
    if (with_plan && blockIdx.x == nmat * ZG) {  // uniform
        plan_body(mats, nmat, rl.gx_off, rl.near_off, rl.far_off, rl.xf_off, rl.plan, par, rl.far_items);
        return;
    }
    RankMat &S = mats[blockIdx.x / ZG];
    if (S.done) return;
    const uint32_t pb = S.it[par].pb;
    const uint32_t s0 = (blockIdx.x % ZG) * ZSLOTS;
    if (s0 >= pb) {  // uniform over the workgroup: no slots, only zero digit columns
        if (L)
            for (uint32_t idx = threadIdx.x; idx < BMAX * ZSLOTS; idx += blockDim.x)
                z_digits(S.Zl + (size_t)par * BMAX * L * KSTEP, L, idx / ZSLOTS, s0 + idx % ZSLOTS, 0, f.p);
        return;
    }
    __shared__ uint64_t Ek[BMAX][BMAX];  // multipliers of steps s0.. (rows: step)
    __shared__ uint64_t Ur[BMAX][BMAX + 1];  // U[r][t], t > r, then V (padded: the back substitution reads columns)
    __shared__ uint64_t piK[BMAX], dinv[BMAX], ownK[BMAX];
    __shared__ uint64_t Fcol[ZSLOTS][BMAX];  // F's column of each slot wave, by block row
    __shared__ uint32_t qk[BMAX];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = uni32(tid >> 6);
    const uint64_t *eta = S.Eta;
    const uint64_t p = f.p;
    {
        // Ek rows s0 .. pb-1 and Ur rows 0 .. pb-1 in 16-byte pieces, all loads issued
        // before the LDS stores: at most (2 BMAX x BMAX / 2) / blockDim pieces each.
        constexpr uint32_t NT = 64 * (ZSLOTS + 1), HALF = BMAX * BMAX / 2, NP = (2 * HALF + NT - 1) / NT;
        const uint32_t ne = (pb - s0) * (BMAX / 2), nu = pb * (BMAX / 2);
        const ulonglong2 *ge = (const ulonglong2 *)(eta + ETA_E + s0 * BMAX);
        const ulonglong2 *gu = (const ulonglong2 *)(eta + ETA_U);
        ulonglong2 v[NP];
#pragma unroll
        for (uint32_t i = 0; i < NP; i++) {
            const uint32_t idx = tid + i * NT;
            v[i] = idx < ne ? ge[idx] : (idx >= HALF && idx - HALF < nu) ? gu[idx - HALF] : ulonglong2{0, 0};
        }
#pragma unroll
        for (uint32_t i = 0; i < NP; i++) {
            const uint32_t idx = tid + i * NT;
            if (idx < ne) {
                const uint32_t r = s0 + (2 * idx) / BMAX, c = (2 * idx) % BMAX;
                Ek[r][c] = v[i].x;
                Ek[r][c + 1] = v[i].y;
            } else if (idx >= HALF && idx - HALF < nu) {
                const uint32_t r = (2 * (idx - HALF)) / BMAX, c = (2 * (idx - HALF)) % BMAX;
                Ur[r][c] = v[i].x;
                Ur[r][c + 1] = v[i].y;
            }
        }
    }
    if (tid < BMAX) {
        piK[tid] = tid < pb ? eta[ETA_PI + tid] : 1;
        ownK[tid] = tid < pb ? eta[ETA_OWNK + tid] : 0;
        qk[tid] = tid < pb ? (uint32_t)eta[ETA_Q + tid] : 0;
    }
    __syncthreads();
    const uint32_t s = s0 + wave;
    if (wave == ZSLOTS) {
        // the inverse wave
        if constexpr (MODE == MODE_M61)
            dinv[lane] = lane < pb ? inv_m61(piK[lane]) : 0;
        else
            dinv[lane] = lane < pb ? inv_fermat_t<MODE>(f, piK[lane]) : 0;
    } else if (s < pb) {
        // Forward: ev = F[lane][q_s] for block row `lane`.  The step's pivot row,
        // pivot and multiplier come from lane registers (readlane) and one LDS read
        // issued a step ahead.
        const uint32_t my_q = qk[lane];
        const uint64_t my_pi = piK[lane];
        const uint32_t qs = (uint32_t)__builtin_amdgcn_readlane((int)my_q, (int)s);
        uint64_t ev = lane == qs ? ownK[s] : 0;
        uint64_t e_next = Ek[s][lane];
        for (uint32_t k = s; k < pb; k++) {
            const uint64_t e = e_next;
            if (k + 1 < pb) e_next = Ek[k + 1][lane];
            const uint32_t q = (uint32_t)__builtin_amdgcn_readlane((int)my_q, (int)k);
            const uint64_t y = lane_bcast(ev, q);
            const uint64_t pik = lane_bcast(my_pi, k);
            if (e != 0) ev = ff_lazy_t<MODE>(f, pik, ev, e, y);
        }
        Fcol[wave][lane] = canon_t<MODE>(ev);
    }
    __syncthreads();  // the inverses and F's columns are in LDS
    // V = D^-1 U (the rows of U scaled by the inverse pivots), in place
    for (uint32_t idx = tid; idx < BMAX * BMAX; idx += blockDim.x) {
        const uint32_t r = idx / BMAX, t = idx % BMAX;
        if (t > r && t < pb) Ur[r][t] = mul_mod_t<MODE>(f, dinv[r], Ur[r][t]);
    }
    __syncthreads();
    if (wave == ZSLOTS) return;
    if (s >= pb) {
        if (L) z_digits(S.Zl + (size_t)par * BMAX * L * KSTEP, L, lane, s, 0, p);
        return;
    }

    // Back substitution, column-oriented, lane = slot t: once z_t is known every
    // lane u < t adds V[u][t] z_t to its lazy sum; lane t-1 is then complete.
    const uint64_t fv = (lane >= s && lane < pb) ? Fcol[wave][qk[lane]] : 0;
    const uint64_t g = lane < pb ? mul_mod_t<MODE>(f, dinv[lane], fv) : 0;
    uint64_t acc = 0, z = 0;
    for (int t = (int)pb - 1; t >= 0; t--) {
        const uint64_t v = Ur[lane][t];
        if (lane == (uint32_t)t) z = sub_mod(g, canon_t<MODE>(acc), p);
        const uint64_t zt = lane_bcast(z, (uint32_t)t);
        if (lane < (uint32_t)t) {
            if constexpr (MODE == MODE_M61)
                acc = fold_m61_64(acc + fold_m61_64(mul_m61_lazy(v, (uint32_t)zt & 0x7fffffffu, (uint32_t)(zt >> 31))));
            else
                acc = add_mod(acc, mul_mod_t<MODE>(f, v, zt), p);
        }
    }
    if (lane < pb) S.Z[lane * BMAX + s] = z;
    if (L) z_digits(S.Zl + (size_t)par * BMAX * L * KSTEP, L, lane, s, lane < pb ? z : 0, p);
}

// ---------------------------------------------------------------------------
// k_plan: tile counts and prefix sums (single workgroup)
// ---------------------------------------------------------------------------

// Tile counts of matrix S in this iteration: gather tasks g, near tiles un, far
// tiles or items uf; adds the update kernels' work to its counters.
__device__ __forceinline__ void plan_counts(RankMat &S, uint32_t par, uint32_t far_items, uint32_t &g, uint32_t &un,
                                            uint32_t &uf, uint32_t &xf) {
    g = un = uf = xf = 0;
    if (S.done) return;
    const IterState &I = S.it[par];
    const uint32_t O = S.n - I.rank;
    const uint32_t cols = I.cend - I.c0;
    if (O == 0 || cols == 0) return;
    const bool pair = I.kind == IT_PAIR;
    // gather rows (own and, for a PAIR, the pending panel's columns) and X columns;
    // the MFMA method's far X columns go to a launch of their own beside the gather
    if (I.pb > 0) {
        const uint32_t nx = cdiv(cols, XCOLS), nxn = far_items ? x_near_tasks(I) : nx;
        g = cdiv(O, GATHER_ROWS) + nxn;
        xf = nx - nxn;
    } else if (pair) {
        g = cdiv(O, GATHER_ROWS);
    }
    uint32_t base, wn, wf;
    update_range(I, 0, base, wn);
    un = part_tiles(O, wn, 0);
    update_range(I, 1, base, wf);
    if (I.kind == IT_DEFER) wf = 0;  // pending: applied by the next iteration
    // far work: XCD-numbered tiles (VALU), or (row band, XCD group) items (MFMA)
    uf = far_items ? (wf > 0 ? cdiv(O, UT) * 8 : 0) : part_tiles(O, wf, 1);
    const uint32_t kfar = I.pb + (pair ? I.dpb : 0u);
    if (I.pb > 0 || pair) {
        // The update kernels' work: O x wn x pb products in the near part and
        // O x wf x kfar in the far part; per tile 16 B per element and 8 B per
        // operand element, summed over the tiles of both parts (R row tiles, C
        // column tiles).
        const unsigned long long R = cdiv(O, UT), C = cdiv(wn, UT) + cdiv(wf, UT);
        S.macs += (unsigned long long)O * wn * I.pb + (unsigned long long)O * wf * kfar;
        S.far_macs += (unsigned long long)O * wf * kfar;
        S.upd_bytes += 16ull * O * (wn + wf) + 8ull * kfar * (O * C + (unsigned long long)(wn + wf) * R);
    }
}

// Inclusive prefix sums of the 4 values v over the workgroup (at most 16 waves).
__device__ __forceinline__ void block_incl_scan4(uint32_t *v, uint32_t (*wsum)[4], uint32_t tid) {
    const uint32_t lane = tid & 63, w = tid >> 6;
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t t = __shfl_up(v[k], off, 64);
            if (lane >= off) v[k] += t;
        }
    }
    if (lane == 63)
#pragma unroll
        for (int k = 0; k < 4; k++) wsum[w][k] = v[k];
    __syncthreads();
    for (uint32_t j = 0; j < w; j++)
#pragma unroll
        for (int k = 0; k < 4; k++) v[k] += wsum[j][k];
}

constexpr uint32_t PLAN_THREADS = 1024;

// The plan of one iteration, by one workgroup of any size up to PLAN_THREADS.
__device__ void plan_body(RankMat *mats, uint32_t nmat, uint32_t *gx_off, uint32_t *near_off, uint32_t *far_off,
                          uint32_t *xf_off, RankPlan *plan, uint32_t par, uint32_t far_items) {
    __shared__ uint32_t wsum[PLAN_THREADS / 64][4];
    __shared__ uint32_t base[4];
    const uint32_t tid = threadIdx.x, nt = blockDim.x;
    // Commit the last iteration's missed-pivot swap, then the top blocks' new row
    // order (written by k_block_ge of this iteration, which saw the swap).  The
    // previous iteration's far update may still run: the plain method's waits
    // for it before this kernel (it reads perm); the MFMA method's reads its rows
    // through the gather's snapshot instead, and this iteration's plan buffer is
    // not the one it reads.
    for (uint32_t i = tid; i < nmat; i += blockDim.x) {
        RankMat &S = mats[i];
        const uint32_t pend = S.pend;
        if (pend != 0xffffffffu && !S.done) {
            const uint32_t r = S.it[par].rank0;
            const uint32_t t = S.perm[r];
            S.perm[r] = S.perm[pend];
            S.perm[pend] = t;
            S.pend = 0xffffffffu;
        }
    }
    __syncthreads();
    for (uint32_t idx = tid; idx < nmat * BMAX; idx += blockDim.x) {
        RankMat &S = mats[idx / BMAX];
        const IterState &I = S.it[par];
        const uint32_t q = idx % BMAX;
        if (!S.done && q < I.Bt) S.perm[I.rank0 + q] = I.top[q];
    }
    if (tid < 4) base[tid] = 0;
    __syncthreads();
    uint32_t *const offs[4] = {gx_off, near_off, far_off, xf_off};
    for (uint32_t i0 = 0; i0 < nmat; i0 += nt) {
        const uint32_t i = i0 + tid;
        uint32_t c[4] = {0, 0, 0, 0};
        if (i < nmat) plan_counts(mats[i], par, far_items, c[0], c[1], c[2], c[3]);
        uint32_t v[4] = {c[0], c[1], c[2], c[3]};
        block_incl_scan4(v, wsum, tid);
        if (i < nmat)
#pragma unroll
            for (int k = 0; k < 4; k++) offs[k][i] = base[k] + v[k] - c[k];
        __syncthreads();
        if (tid == nt - 1)
#pragma unroll
            for (int k = 0; k < 4; k++) base[k] += v[k];
        __syncthreads();
    }
    if (tid == 0) {
#pragma unroll
        for (int k = 0; k < 4; k++) offs[k][nmat] = base[k];
        plan->gx_total = base[0];
        plan->near_total = base[1];
        plan->far_total = base[2];
        plan->xf_total = base[3];
    }
}

__global__ __launch_bounds__(PLAN_THREADS) void k_plan(RankMat *mats, uint32_t nmat, uint32_t *gx_off,
                                                       uint32_t *near_off, uint32_t *far_off, uint32_t *xf_off,
                                                       RankPlan *plan, uint32_t par, uint32_t far_items) {
    plan_body(mats, nmat, gx_off, near_off, far_off, xf_off, plan, par, far_items);
}

// ---------------------------------------------------------------------------
// k_gather_x
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_gather_x(RankMat *mats, uint32_t nmat, const uint32_t *gx_off,
                                                  const RankPlan *plan, FieldParams f, uint32_t par) {
    __shared__ uint64_t Zs[BMAX][BMAX + 1];
    __shared__ uint64_t Ms[BMAX][XCOLS + 1];
    const uint32_t total = plan->gx_total;
    const uint32_t tid = threadIdx.x;
    for (uint32_t t = blockIdx.x; t < total; t += gridDim.x) {
        const uint32_t mi = find_mat(gx_off, nmat, t);
        RankMat &S = mats[mi];
        const IterState &I = S.it[par];
        const uint32_t lt = t - gx_off[mi];
        const uint32_t O = S.n - I.rank, pb = I.pb, ld = S.ld;
        const uint32_t ng = cdiv(O, GATHER_ROWS);
        if (lt < ng) {
            // Ab[i][k] = M[perm[rank + i]][J[k]]
            const uint32_t r0 = lt * GATHER_ROWS;
            const uint32_t rows = min(GATHER_ROWS, O - r0);
            for (uint32_t idx = tid; idx < rows * pb; idx += blockDim.x) {
                uint32_t i = idx / pb, kk = idx - i * pb;
                uint32_t r = S.perm[I.rank + r0 + i];
                S.Ab[(size_t)(r0 + i) * BMAX + kk] = S.M[(size_t)r * ld + I.J[kk]];
            }
        } else {
            // X[k][c] = sum_s Z[k][s] M[R[s]][c0 + c]
            const uint32_t cb = (lt - ng) * XCOLS;
            const uint32_t cols = min(XCOLS, I.cend - I.c0 - cb);
            __syncthreads();
            for (uint32_t idx = tid; idx < pb * pb; idx += blockDim.x) Zs[idx / pb][idx % pb] = S.Z[(idx / pb) * BMAX + idx % pb];
            for (uint32_t idx = tid; idx < pb * XCOLS; idx += blockDim.x) {
                uint32_t s = idx / XCOLS, c = idx - s * XCOLS;
                Ms[s][c] = (c < cols) ? S.M[(size_t)I.R[s] * ld + I.c0 + cb + c] : 0;
            }
            __syncthreads();
            const uint64_t sb = f.safe_batch;
            for (uint32_t o = tid; o < pb * cols; o += blockDim.x) {
                uint32_t kk = o / cols, c = o - kk * cols;
                uint64_t ah = 0, al = 0, tot = 0, cnt = 0;
                for (uint32_t s = 0; s < pb; s++) {
                    uint64_t h, l;
                    mul_wide(Zs[kk][s], Ms[s][c], h, l);
                    add128(ah, al, h, l);
                    if (++cnt == sb) {
                        tot = add_mod(tot, reduce128(f, ah, al), f.p);
                        ah = al = 0;
                        cnt = 0;
                    }
                }
                if (cnt) tot = add_mod(tot, reduce128(f, ah, al), f.p);
                S.X[(size_t)kk * S.n + cb + c] = tot;
            }
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------------------
// k_update (VALU Barrett): S = M[O, c0:cend] - Ab X
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_update_valu(RankMat *mats, uint32_t nmat, const uint32_t *upd_off,
                                                     const RankPlan *plan, FieldParams f, uint32_t par, uint32_t far) {
    __shared__ uint64_t As[UT][BMAX + 1];
    __shared__ uint64_t Xs[BMAX][UT + 1];
    const uint32_t total = far ? plan->far_total : plan->near_total;
    const uint32_t tid = threadIdx.x;
    const uint32_t ty = tid >> 4, tx = tid & 15;
    for (uint32_t t = blockIdx.x; t < total; t += gridDim.x) {
        const uint32_t mi = find_mat(upd_off, nmat, t);
        RankMat &S = mats[mi];
        const IterState &I = S.it[par];
        const uint32_t lt = t - upd_off[mi];
        const uint32_t O = S.n - I.rank, pb = I.pb, ld = S.ld;
        uint32_t cbase, cols, r0, cb, nr, nc;
        update_range(I, far, cbase, cols);
        if (!tile_geom(O, cbase, cols, far, lt, r0, cb, nr, nc)) continue;  // padding tile (uniform)
        // detection only, right of a nonzero column already found: nothing to add
        if (pb == 0 && I.c0 + cb > __hip_atomic_load(&S.first_nz_col, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            continue;
        __syncthreads();
        for (uint32_t idx = tid; idx < UT * pb; idx += blockDim.x) {
            uint32_t i = idx / pb, kk = idx - i * pb;
            As[i][kk] = (i < nr) ? S.Ab[(size_t)(r0 + i) * BMAX + kk] : 0;
        }
        for (uint32_t idx = tid; idx < pb * UT; idx += blockDim.x) {
            uint32_t kk = idx / UT, c = idx - kk * UT;
            Xs[kk][c] = (c < nc) ? S.X[(size_t)kk * S.n + cb + c] : 0;
        }
        __syncthreads();
        const uint64_t sb = f.safe_batch, p = f.p;
        uint64_t ah[4][4], al[4][4], tot[4][4];
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int c = 0; c < 4; c++) ah[a][c] = al[a][c] = tot[a][c] = 0;
        uint64_t cnt = 0;
        for (uint32_t kk = 0; kk < pb; kk++) {
            uint64_t av[4], xv[4];
#pragma unroll
            for (int a = 0; a < 4; a++) av[a] = As[ty * 4 + a][kk];
#pragma unroll
            for (int c = 0; c < 4; c++) xv[c] = Xs[kk][tx * 4 + c];
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    uint64_t h, l;
                    mul_wide(av[a], xv[c], h, l);
                    add128(ah[a][c], al[a][c], h, l);
                }
            if (++cnt == sb) {
#pragma unroll
                for (int a = 0; a < 4; a++)
#pragma unroll
                    for (int c = 0; c < 4; c++) {
                        tot[a][c] = add_mod(tot[a][c], reduce128(f, ah[a][c], al[a][c]), p);
                        ah[a][c] = al[a][c] = 0;
                    }
                cnt = 0;
            }
        }
        if (cnt) {
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int c = 0; c < 4; c++) tot[a][c] = add_mod(tot[a][c], reduce128(f, ah[a][c], al[a][c]), p);
        }
        uint32_t nzcol = 0xffffffffu;
#pragma unroll
        for (int a = 0; a < 4; a++) {
            const uint32_t i = ty * 4 + a;
            if (i >= nr) continue;
            const uint32_t r = S.perm[I.rank + r0 + i];
            uint64_t *mrow = S.M + (size_t)r * ld + I.c0 + cb;
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const uint32_t cc = tx * 4 + c;
                if (cc >= nc) continue;
                uint64_t v = sub_mod(mrow[cc], tot[a][c], p);
                if (pb > 0) mrow[cc] = v;
                const uint32_t gcol = I.c0 + cb + cc;
                if (v != 0 && gcol < I.scan_end) nzcol = min(nzcol, gcol);
            }
        }
        if (I.resid && nzcol != 0xffffffffu) atomicMin(&S.first_nz_col, nzcol);
    }
}

// ---------------------------------------------------------------------------
// k_finish: the end of the iteration (parity par), after its near update
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_finish(RankMat *mats, uint32_t *active, uint32_t par) {
    __shared__ uint32_t best;
    finish_iter(mats[blockIdx.x], active, par, best);
}

// ---------------------------------------------------------------------------
// k_collect: the active count and every matrix's rank and counters, packed for
// one copy to the host: out[0] = active, out[1 + 4 i ..] = rank, macs, bytes, far macs
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_collect(const RankMat *mats, uint32_t nmat, const uint32_t *active,
                                                 uint64_t *out) {
    const uint32_t tid = threadIdx.x;
    if (tid == 0) out[0] = *active;
    for (uint32_t i = tid; i < nmat; i += blockDim.x) {
        out[1 + 4 * (size_t)i] = mats[i].rank;
        out[2 + 4 * (size_t)i] = mats[i].macs;
        out[3 + 4 * (size_t)i] = mats[i].upd_bytes;
        out[4 + 4 * (size_t)i] = mats[i].far_macs;
    }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
// The block elimination (the LDS form with lds) and Z; with rlp, the iteration's
// plan too (one more workgroup of k_block_z).
template <int MODE>
static void launch_block_t(hipStream_t st, RankMat *dmats, uint32_t nmat, const FieldParams &f, uint32_t b,
                           uint32_t par, uint32_t L, uint32_t defer, const RankLaunch *rlp, bool lds) {
    if (lds)
        hipLaunchKernelGGL(k_block_ge_lds<MODE>, dim3(nmat), dim3(1024), 0, st, dmats, f, b, par, defer);
    else
        hipLaunchKernelGGL(k_block_ge<MODE>, dim3(nmat), dim3(64 * GE_WAVES), 0, st, dmats, f, b, par, defer);
    const RankLaunch rl = rlp ? *rlp : RankLaunch{};
    hipLaunchKernelGGL(k_block_z<MODE>, dim3(nmat * ZG + (rlp ? 1 : 0)), dim3(64 * (ZSLOTS + 1)), 0, st, dmats, f,
                       par, L, nmat, rl, rlp ? 1u : 0u);
}

void launch_block(hipStream_t st, RankMat *dmats, uint32_t nmat, const FieldParams &f, uint32_t b, uint32_t par,
                  uint32_t L, uint32_t defer, const RankLaunch *rlp, bool lds) {
    if (f.mode == MODE_M61)
        launch_block_t<MODE_M61>(st, dmats, nmat, f, b, par, L, defer, rlp, lds);
    else if (f.mode == MODE_M31)
        launch_block_t<MODE_M31>(st, dmats, nmat, f, b, par, L, defer, rlp, lds);
    else
        launch_block_t<MODE_GENERIC>(st, dmats, nmat, f, b, par, L, defer, rlp, lds);
}

void launch_plan(hipStream_t st, RankMat *dmats, uint32_t nmat, const RankLaunch &rl, uint32_t par) {
    hipLaunchKernelGGL(k_plan, dim3(1), dim3(PLAN_THREADS), 0, st, dmats, nmat, rl.gx_off, rl.near_off, rl.far_off,
                       rl.xf_off, rl.plan,
                       par, rl.far_items);
}

void launch_finish(hipStream_t st, RankMat *dmats, uint32_t nmat, uint32_t *active, uint32_t par) {
    hipLaunchKernelGGL(k_finish, dim3(nmat), dim3(1024), 0, st, dmats, active, par);
}

void launch_collect(hipStream_t st, RankMat *dmats, uint32_t nmat, const uint32_t *active, uint64_t *out) {
    hipLaunchKernelGGL(k_collect, dim3(1), dim3(256), 0, st, dmats, nmat, active, out);
}

void launch_gather_plain(hipStream_t st, RankMat *dmats, uint32_t nmat, const FieldParams &f, const RankLaunch &rl,
                         uint32_t par) {
    hipLaunchKernelGGL(k_gather_x, dim3(rl.grid_gx), dim3(256), 0, st, dmats, nmat, rl.gx_off, rl.plan, f, par);
}

void launch_update_plain(hipStream_t st, RankMat *dmats, uint32_t nmat, const FieldParams &f, const RankLaunch &rl,
                         uint32_t par, uint32_t far) {
    hipLaunchKernelGGL(k_update_valu, dim3(far ? rl.grid_far : rl.grid_near), dim3(256), 0, st, dmats, nmat,
                       far ? rl.far_off : rl.near_off, rl.plan, f, par, far);
}

}  // namespace mdl
