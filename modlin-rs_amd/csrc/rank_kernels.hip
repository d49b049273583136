// Rank over F_p on gfx950: the blocked right-looking elimination of the
// reference (src/fp.rs:116-255) restructured for the GPU.
//
// One iteration handles one column panel [col, jend) of every active matrix:
//
//   k_block_ge   one workgroup per matrix.  Loads the top block (the next BMAX
//                logical rows) of the panel into LDS and eliminates it with the
//                reference's first-nonzero pivot rule (fp.rs:150-163): a wave
//                ballot + ffs finds the pivot row, rows are swapped in LDS, and the
//                elimination is fraction-free (pi*x - e*y), so no modular inverse
//                sits on the per-column critical path.  The elimination is
//                Gauss-Jordan and tracks the row operations E, so that at the end
//                E A = D (diagonal) for A = M[R, J], and Z = A^-1 = D^-1 E needs
//                only pb independent inverses (Fermat, fp.rs:111-113).
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

// ---------------------------------------------------------------------------
// k_block_ge: pivot search and elimination of the panel part of the top block
// ---------------------------------------------------------------------------
template <int MODE>
__global__ __launch_bounds__(1024) void k_block_ge(RankMat *mats, FieldParams f, uint32_t b) {
    RankMat &S = mats[blockIdx.x];
    if (S.done) return;
    // One LDS row per block row: [0, BMAX) panel entries T, then own (coefficient
    // on the row's own original row) and D (its pivot once the row is a pivot
    // row, scaled like the row afterwards).
    constexpr uint32_t COWN = BMAX, CD = BMAX + 1, UW = BMAX + 2;
    __shared__ uint64_t U[BMAX][UW];
    __shared__ uint64_t Et[BMAX][BMAX];     // multiplier e of each block row at each step
    __shared__ uint64_t piK[BMAX], ownK[BMAX];
    __shared__ uint32_t rowid[BMAX], Jl[BMAX], swq[BMAX];

    const uint32_t tid = threadIdx.x;
    const uint32_t n = S.n, ld = S.ld, rank0 = S.rank, col = S.col;
    const uint32_t jend = min(col + b, n);
    const uint32_t w = jend - col;
    const uint32_t Bt = min((uint32_t)BMAX, n - rank0);
    const uint64_t p = f.p;

    for (uint32_t idx = tid; idx < BMAX * UW; idx += blockDim.x) {
        const uint32_t q = idx / UW, c = idx - q * UW;
        uint64_t v = 0;
        if (q < Bt) {
            if (c < w) v = S.M[(size_t)S.perm[rank0 + q] * ld + col + c];
            else if (c == COWN) v = 1 % p;
        }
        U[q][c] = v;
    }
    for (uint32_t idx = tid; idx < BMAX * BMAX; idx += blockDim.x) Et[idx / BMAX][idx % BMAX] = 0;
    if (tid < Bt) rowid[tid] = S.perm[rank0 + tid];
    __syncthreads();

    const uint32_t lane = tid & 63;
    const uint32_t row = tid >> 4, sub = tid & 15;
    uint32_t k = 0;
    for (uint32_t j = 0; j < w && k < Bt; j++) {
        // Pivot search: first block position q >= k with a nonzero in column j.
        // Every wave evaluates the same ballot, so no barrier is needed.
        bool nz = (lane >= k) && (lane < Bt) && (U[lane][j] != 0);
        uint64_t mask = __ballot(nz);
        if (mask == 0) continue;  // uniform: no pivot in this column
        uint32_t q = __ffsll((unsigned long long)mask) - 1;
        if (q != k) {  // swap block positions k and q (fp.rs:158-162)
            __syncthreads();  // every wave has read column j for its ballot before rows move
            for (uint32_t c = tid; c < UW; c += blockDim.x) {
                uint64_t t = U[k][c]; U[k][c] = U[q][c]; U[q][c] = t;
            }
            if (tid == 0) {
                uint32_t r = rowid[k]; rowid[k] = rowid[q]; rowid[q] = r;
            }
            __syncthreads();
        }
        // Fraction-free Gauss-Jordan step on every other row with e = U[row][j] != 0:
        // x <- pi x - e y on the panel columns after j; own and D scale by pi.
        // Row k is only read and column j is never written, so the step needs no
        // barrier between reading e and writing.
        const uint64_t pi = U[k][j];
        if (row < Bt && row != k) {
            const uint64_t e = U[row][j];
            if (sub == 0) Et[k][row] = e;
            if (e != 0) {
                const uint32_t nT = w - j - 1;
                for (uint32_t it = sub; it < nT + 2; it += 16) {
                    const uint32_t c = it < nT ? j + 1 + it : (it == nT ? COWN : CD);
                    const uint64_t y = it < nT ? U[k][c] : 0;
                    U[row][c] = ff_step_t<MODE>(f, pi, U[row][c], e, y);
                }
            }
        }
        if (tid == 0) {
            U[k][CD] = pi;
            Jl[k] = j;
            piK[k] = pi;
            ownK[k] = U[k][COWN];
            swq[k] = q;
        }
        __syncthreads();
        k++;
    }
    const uint32_t pb = k;

    // The record for k_block_z (which builds A^-1 from it) and 1 / D.
    uint64_t *eta = S.Eta;
    for (uint32_t idx = tid; idx < pb * BMAX; idx += blockDim.x) eta[ETA_E + idx] = Et[idx / BMAX][idx % BMAX];
    if (tid < pb) {
        eta[ETA_PI + tid] = piK[tid];
        eta[ETA_OWNK + tid] = ownK[tid];
        eta[ETA_OWNF + tid] = U[tid][COWN];
        eta[ETA_DINV + tid] = inv_mod_bin(U[tid][CD], p);
        eta[ETA_SWQ + tid] = swq[tid];
    }
    if (tid < Bt) S.perm[rank0 + tid] = rowid[tid];
    if (tid < pb) {
        S.J[tid] = col + Jl[tid];
        S.R[tid] = rowid[tid];
    }
    if (tid == 0) {
        S.jend = jend;
        S.pb = pb;
        S.rank = rank0 + pb;
        const bool below = (n - rank0) > Bt;
        const uint32_t resid = (pb < w && below) ? 1u : 0u;
        S.resid = resid;
        if (pb > 0) {
            S.c0 = resid ? col : jend;
            S.cend = n;
        } else {
            S.c0 = col;
            S.cend = resid ? jend : col;  // detection only, or nothing at all
        }
        S.first_nz_col = 0xffffffffu;
        S.iters++;
    }
}

// ---------------------------------------------------------------------------
// k_block_z: Z = A^-1 for the pivot block A = M[R, J] of the last elimination.
//
// The elimination applied row operations to the block; the same operations on
// the identity give rows E with E A = D (diagonal).  Only pivot slots < k are
// nonzero in a row's coefficients at step k, so E is kept per pivot slot s:
// slot s starts at step s as -e own[s] and is then carried through the later
// steps (x <- pi x - e y, with the recorded swaps).  Slots are independent, so
// ZG workgroups per matrix replay the record on BMAX / ZG slots each.  Finally
// Z = D^-1 E' with E'[r][s] = E[r][s] (s != r) and own[r] + E[r][r] (s == r):
// slot r of a pivot row and its own original row are the same row.
// ---------------------------------------------------------------------------
template <int MODE>
__global__ __launch_bounds__(1024) void k_block_z(RankMat *mats, FieldParams f) {
    constexpr uint32_t SG = BMAX / ZG;  // slots per workgroup
    RankMat &S = mats[blockIdx.x / ZG];
    if (S.done) return;
    const uint32_t pb = S.pb;
    const uint32_t s0 = (blockIdx.x % ZG) * SG;
    if (s0 >= pb) return;
    __shared__ uint64_t Ec[BMAX][SG + 1];
    __shared__ uint64_t Ek[BMAX][BMAX];          // the record's multipliers, steps s0..pb-1
    __shared__ uint64_t piK[BMAX], ownK[BMAX];
    __shared__ uint32_t swq[BMAX];
    const uint32_t tid = threadIdx.x, row = tid / SG, sl = tid % SG, s = s0 + sl;
    const uint64_t *eta = S.Eta;
    const uint64_t p = f.p;
    Ec[row][sl] = 0;
    for (uint32_t idx = s0 * BMAX + tid; idx < pb * BMAX; idx += blockDim.x) Ek[idx / BMAX][idx % BMAX] = eta[ETA_E + idx];
    if (tid < pb) {
        piK[tid] = eta[ETA_PI + tid];
        ownK[tid] = eta[ETA_OWNK + tid];
        swq[tid] = (uint32_t)eta[ETA_SWQ + tid];
    }
    __syncthreads();
    for (uint32_t k = s0; k < pb; k++) {
        const uint32_t q = swq[k];
        if (q != k) {
            if (row == k) {
                const uint64_t t = Ec[k][sl]; Ec[k][sl] = Ec[q][sl]; Ec[q][sl] = t;
            }
            __syncthreads();
        }
        const uint64_t e = Ek[k][row];
        if (s <= k && s < pb && row != k && e != 0) {
            const uint64_t x = (s == k) ? 0 : Ec[row][sl];
            const uint64_t y = (s == k) ? ownK[k] : Ec[k][sl];
            Ec[row][sl] = ff_step_t<MODE>(f, piK[k], x, e, y);
        }
        __syncthreads();
    }
    if (row < pb && s < pb) {
        const uint64_t v = (row == s) ? add_mod(eta[ETA_OWNF + row], Ec[row][sl], p) : Ec[row][sl];
        S.Z[row * BMAX + s] = mul_mod_t<MODE>(f, eta[ETA_DINV + row], v);
    }
}

// ---------------------------------------------------------------------------
// k_plan: tile counts and prefix sums (single workgroup)
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_plan(RankMat *mats, uint32_t nmat, uint32_t *gx_off, uint32_t *upd_off,
                                              RankPlan *plan) {
    __shared__ uint32_t sg[256], su[256];
    __shared__ uint32_t base_g, base_u;
    const uint32_t tid = threadIdx.x;
    if (tid == 0) base_g = base_u = 0;
    __syncthreads();
    for (uint32_t i0 = 0; i0 < nmat; i0 += 256) {
        const uint32_t i = i0 + tid;
        uint32_t g = 0, u = 0;
        if (i < nmat) {
            const RankMat &S = mats[i];
            if (!S.done) {
                const uint32_t O = S.n - S.rank;
                const uint32_t cols = S.cend - S.c0;
                if (O > 0 && cols > 0) {
                    if (S.pb > 0) g = cdiv(O, GATHER_ROWS) + cdiv(cols, XCOLS);
                    u = cdiv(O, UT) * cdiv(cols, UT);
                }
            }
        }
        sg[tid] = g;
        su[tid] = u;
        __syncthreads();
        // inclusive Hillis-Steele scan
        for (uint32_t off = 1; off < 256; off <<= 1) {
            const uint32_t ag = tid >= off ? sg[tid - off] : 0, au = tid >= off ? su[tid - off] : 0;
            __syncthreads();
            sg[tid] += ag;
            su[tid] += au;
            __syncthreads();
        }
        if (i < nmat) {
            gx_off[i] = base_g + sg[tid] - g;
            upd_off[i] = base_u + su[tid] - u;
        }
        __syncthreads();
        if (tid == 255) {
            base_g += sg[255];
            base_u += su[255];
        }
        __syncthreads();
    }
    if (tid == 0) {
        gx_off[nmat] = base_g;
        upd_off[nmat] = base_u;
        plan->gx_total = base_g;
        plan->upd_total = base_u;
    }
}

// ---------------------------------------------------------------------------
// k_gather_x
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_gather_x(RankMat *mats, uint32_t nmat, const uint32_t *gx_off,
                                                  const RankPlan *plan, FieldParams f) {
    __shared__ uint64_t Zs[BMAX][BMAX + 1];
    __shared__ uint64_t Ms[BMAX][XCOLS + 1];
    const uint32_t total = plan->gx_total;
    const uint32_t tid = threadIdx.x;
    for (uint32_t t = blockIdx.x; t < total; t += gridDim.x) {
        const uint32_t mi = find_mat(gx_off, nmat, t);
        RankMat &S = mats[mi];
        const uint32_t lt = t - gx_off[mi];
        const uint32_t O = S.n - S.rank, pb = S.pb, ld = S.ld;
        const uint32_t ng = cdiv(O, GATHER_ROWS);
        if (lt < ng) {
            // Ab[i][k] = M[perm[rank + i]][J[k]]
            const uint32_t r0 = lt * GATHER_ROWS;
            const uint32_t rows = min(GATHER_ROWS, O - r0);
            for (uint32_t idx = tid; idx < rows * pb; idx += blockDim.x) {
                uint32_t i = idx / pb, kk = idx - i * pb;
                uint32_t r = S.perm[S.rank + r0 + i];
                S.Ab[(size_t)(r0 + i) * BMAX + kk] = S.M[(size_t)r * ld + S.J[kk]];
            }
        } else {
            // X[k][c] = sum_s Z[k][s] M[R[s]][c0 + c]
            const uint32_t cb = (lt - ng) * XCOLS;
            const uint32_t cols = min(XCOLS, S.cend - S.c0 - cb);
            __syncthreads();
            for (uint32_t idx = tid; idx < pb * pb; idx += blockDim.x) Zs[idx / pb][idx % pb] = S.Z[(idx / pb) * BMAX + idx % pb];
            for (uint32_t idx = tid; idx < pb * XCOLS; idx += blockDim.x) {
                uint32_t s = idx / XCOLS, c = idx - s * XCOLS;
                Ms[s][c] = (c < cols) ? S.M[(size_t)S.R[s] * ld + S.c0 + cb + c] : 0;
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
                                                     const RankPlan *plan, FieldParams f) {
    __shared__ uint64_t As[UT][BMAX + 1];
    __shared__ uint64_t Xs[BMAX][UT + 1];
    const uint32_t total = plan->upd_total;
    const uint32_t tid = threadIdx.x;
    const uint32_t ty = tid >> 4, tx = tid & 15;
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
            const uint32_t r = S.perm[S.rank + r0 + i];
            uint64_t *mrow = S.M + (size_t)r * ld + S.c0 + cb;
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const uint32_t cc = tx * 4 + c;
                if (cc >= nc) continue;
                uint64_t v = sub_mod(mrow[cc], tot[a][c], p);
                if (pb > 0) mrow[cc] = v;
                const uint32_t gcol = S.c0 + cb + cc;
                if (v != 0 && gcol < S.jend) nzcol = min(nzcol, gcol);
            }
        }
        if (S.resid && nzcol != 0xffffffffu) atomicMin(&S.first_nz_col, nzcol);
        if (tid == 0 && pb > 0) {
            atomicAdd(&S.macs, (unsigned long long)nr * nc * pb);
            atomicAdd(&S.upd_bytes, 16ull * nr * nc + 8ull * pb * (nr + nc));
        }
    }
}

// ---------------------------------------------------------------------------
// k_finish: advance the panel; bring a missed pivot to the top of the block
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_finish(RankMat *mats, uint32_t *active) {
    RankMat &S = mats[blockIdx.x];
    if (S.done) return;
    __shared__ uint32_t best;
    const uint32_t tid = threadIdx.x;
    const uint32_t n = S.n;
    uint32_t jstar = S.first_nz_col;
    const bool miss = S.resid && jstar < S.jend;
    if (miss) {
        if (tid == 0) best = 0xffffffffu;
        __syncthreads();
        uint32_t mine = 0xffffffffu;
        for (uint32_t q = S.rank + tid; q < n; q += blockDim.x) {
            if (S.M[(size_t)S.perm[q] * S.ld + jstar] != 0) { mine = q; break; }
        }
        if (mine != 0xffffffffu) atomicMin(&best, mine);
        __syncthreads();
        if (tid == 0 && best != 0xffffffffu && best != S.rank) {
            uint32_t t = S.perm[S.rank];
            S.perm[S.rank] = S.perm[best];
            S.perm[best] = t;
        }
    }
    if (tid == 0) {
        S.col = miss ? jstar : S.jend;
        if (S.col >= n || S.rank >= n) {
            S.done = 1;
            atomicSub(active, 1u);
        }
    }
}

// ---------------------------------------------------------------------------
// Host launchers (one panel iteration of every matrix in the batch)
// ---------------------------------------------------------------------------
void launch_block_plan(hipStream_t st, RankMat *dmats, uint32_t nmat, const FieldParams &f, uint32_t b,
                       uint32_t *gx_off, uint32_t *upd_off, RankPlan *plan) {
    if (f.mode == MODE_M61) {
        hipLaunchKernelGGL(k_block_ge<MODE_M61>, dim3(nmat), dim3(1024), 0, st, dmats, f, b);
        hipLaunchKernelGGL(k_block_z<MODE_M61>, dim3(nmat * ZG), dim3(1024), 0, st, dmats, f);
    } else if (f.mode == MODE_M31) {
        hipLaunchKernelGGL(k_block_ge<MODE_M31>, dim3(nmat), dim3(1024), 0, st, dmats, f, b);
        hipLaunchKernelGGL(k_block_z<MODE_M31>, dim3(nmat * ZG), dim3(1024), 0, st, dmats, f);
    } else {
        hipLaunchKernelGGL(k_block_ge<MODE_GENERIC>, dim3(nmat), dim3(1024), 0, st, dmats, f, b);
        hipLaunchKernelGGL(k_block_z<MODE_GENERIC>, dim3(nmat * ZG), dim3(1024), 0, st, dmats, f);
    }
    hipLaunchKernelGGL(k_plan, dim3(1), dim3(256), 0, st, dmats, nmat, gx_off, upd_off, plan);
}

void launch_finish(hipStream_t st, RankMat *dmats, uint32_t nmat, uint32_t *active) {
    hipLaunchKernelGGL(k_finish, dim3(nmat), dim3(1024), 0, st, dmats, active);
}

void launch_rank_iteration(hipStream_t st, RankMat *dmats, uint32_t nmat, const FieldParams &f, uint32_t b,
                           uint32_t *gx_off, uint32_t *upd_off, RankPlan *plan, uint32_t *active, uint32_t grid_gx,
                           uint32_t grid_upd, hipEvent_t ev_upd0, hipEvent_t ev_upd1) {
    launch_block_plan(st, dmats, nmat, f, b, gx_off, upd_off, plan);
    hipLaunchKernelGGL(k_gather_x, dim3(grid_gx), dim3(256), 0, st, dmats, nmat, gx_off, plan, f);
    if (ev_upd0) hipEventRecord(ev_upd0, st);
    hipLaunchKernelGGL(k_update_valu, dim3(grid_upd), dim3(256), 0, st, dmats, nmat, upd_off, plan, f);
    if (ev_upd1) hipEventRecord(ev_upd1, st);
    launch_finish(st, dmats, nmat, active);
}

}  // namespace mdl
