// Shared declarations of the rank engine (device state layout, launch helpers).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fp_dev.hpp"
#include "limb.hpp"

namespace mdl {

// Panel width and height of the pivot-search block.  The trailing update is a
// rank-pb update with pb <= BMAX.
constexpr int BMAX = 64;

// Kinds of panel iteration (MFMA method).  A DEFER iteration updates only the
// near columns [c0, jend + 2 BMAX) -- what the next two panels need -- and leaves
// its contribution to the far columns pending; the next iteration (PAIR) applies
// it together with its own, as one update with up to 2 BMAX pivots, so each far
// element is read and written once per two panels.  Its own pivot rows, which
// the pending update has not reached, get it inside the X product:
// X' = Z (M[R, c] - M[R, J'] X'_pending[c]) for the far columns c.
enum : uint32_t { IT_NORMAL = 0, IT_DEFER = 1, IT_PAIR = 2 };

// Parameters of one panel iteration, written by k_block_ge and read by the
// kernels of the same iteration.  RankMat keeps two, selected by iteration
// parity, so that the next panel's elimination can run while the far columns
// of this panel's update are still being written.
struct IterState {
    uint32_t rank0;          // pivots before this panel: the top block starts here
    uint32_t rank;           // pivots after it: the rows O still to update start here
    uint32_t Bt;             // rows in the top block
    uint32_t jend, pb;       // panel end; pivots found in the panel
    uint32_t c0, cend;       // trailing-update column range
    uint32_t split;          // first column of the far part of the update (jend + BMAX)
    uint32_t resid;          // the panel needs a residual (missed-pivot) check
    uint32_t kind;           // IT_NORMAL, IT_DEFER or IT_PAIR
    uint32_t dpb;            // IT_PAIR: pivots of the pending panel
    uint32_t dcol;           // IT_PAIR: first column the pending panel has not updated (= split)
    uint32_t dxoff;          // IT_PAIR: c0 minus the pending panel's c0 (its X digit columns)
    uint32_t scan_end;       // resid: columns [col, scan_end) are searched for a missed pivot --
                             // the panel (jend), or, for a panel without pivots, every column left (n)
    uint32_t J[BMAX];        // pivot columns of the panel (absolute)
    uint32_t R[BMAX];        // physical pivot rows, in pivot order
    uint32_t top[BMAX];      // physical rows of the top block in their new order
    uint32_t dJ[BMAX];       // IT_PAIR: pivot columns of the pending panel (absolute)
};

// One matrix of a rank batch, resident in device memory.  The matrix is the
// library's working copy (row-major, leading dimension ld); rows are addressed
// through perm (logical position -> physical row) so that pivoting swaps
// indices instead of 8n-byte rows.
struct RankMat {
    uint64_t *M;       // working copy, n x n, row stride ld; u32 entries when m32
    uint32_t *perm;    // logical -> physical row
    uint64_t *Ab;      // gathered A operand M[O, J]: n x BMAX (row stride BMAX)
    uint64_t *X;       // A^-1 M[R, c0:], BMAX x n (row stride n)
    uint64_t *Z;       // A^-1, BMAX x BMAX
    uint64_t *Eta;     // row operations of the last block elimination (ETA_* layout)
    int8_t *Al;        // int8 limb planes of Ab (MFMA path)
    int8_t *Xl;        // int8 limb planes of the shifted X (MFMA path)
    int8_t *Zl;        // int8 limb planes of Z, rows K-contiguous (MFMA path), one block per iteration
                       // parity: the far X product of an iteration reads it while the next k_block_z writes
    int8_t *AlP;       // IT_PAIR: int8 limb planes of M[O, J'] for the pending panel's columns J'
    int8_t *XlP;       // IT_DEFER: int8 limb planes of its X (read by its near update and the next PAIR)
    uint32_t *rows;    // MFMA path: physical row of position rank + r, written by the gather's row
                       // tasks; the far update addresses its rows through it, so the next k_plan
                       // may reorder perm while it runs
    uint32_t n, ld;
    uint32_t m32;            // M holds u32 entries (p < 2^32, MFMA method): half the update traffic
    uint32_t rank, col;      // pivots found; first column not yet eliminated
    uint32_t first_nz_col;   // smallest panel column with a nonzero residual
    uint32_t done;
    uint32_t iters;
    uint32_t pend;           // missed pivot: position whose row moves to `rank`, applied by
                             // the next k_plan (0xffffffff: none); see finish_iter
    uint32_t fin;            // iterations whose finish (finish_iter) has run: the last one's
                             // is pending while fin < iters
    uint32_t dpend;          // a DEFER iteration's far contribution is pending
    uint32_t dpb, dcol, dc0; // its pivots, first column not updated, and base column c0
    uint32_t dJ[BMAX];       // its pivot columns
    IterState it[2];
    // Counters, on a cache line of their own, kept by k_plan (no atomics in the
    // update kernels).
    alignas(128) unsigned long long macs;  // F_p multiply-accumulates of the trailing updates
    unsigned long long upd_bytes;          // algorithmic HBM bytes of the trailing updates
    unsigned long long far_macs;           // the part of macs in the far parts
};

// Work plan of one iteration: per-matrix tile offsets (prefix sums) of the
// gather tasks and of the near (columns [c0, split)) and far ([split, cend))
// update tiles.
struct RankPlan {
    uint32_t gx_total, near_total, far_total, xf_total;
};

// Record of one block elimination, written by k_block_ge and replayed by
// k_block_z (u64 offsets into RankMat::Eta): the multiplier e of every block row
// at every step (zero for rows that take no step: the pivot rows and rows with a
// zero entry), the pivot rows' entries U[r][s] in the pivot columns (upper
// triangle, pivot slots), and per step the pivot, the pivot row's own-row
// coefficient and the pivot row's position in the block (rows do not move).
// Every value is canonical (in [0, p)).
constexpr uint32_t ETA_E = 0;                          // [step][block row]
constexpr uint32_t ETA_U = BMAX * BMAX;                // [pivot slot][pivot slot]
constexpr uint32_t ETA_PI = 2 * BMAX * BMAX;           // [step]
constexpr uint32_t ETA_OWNK = ETA_PI + BMAX;           // [step]
constexpr uint32_t ETA_Q = ETA_OWNK + BMAX;            // [step]
constexpr uint32_t ETA_WORDS = ETA_Q + BMAX;
constexpr uint32_t ZG = 16;                            // workgroups per matrix in k_block_z
constexpr uint64_t GE_LDS_MIN_SIDE = 24576;            // batches with a side this large: k_block_ge_lds
constexpr uint32_t ZSLOTS = BMAX / ZG;                 // slot waves per workgroup (plus one inverse wave)


// Device buffers and grid sizes of the per-iteration launches.
struct RankLaunch {
    uint32_t *gx_off, *near_off, *far_off;  // nmat + 1 offsets each
    uint32_t *xf_off;                       // MFMA: X tasks of the far columns, launched after k_finish
    RankPlan *plan;
    uint32_t grid_gx, grid_near, grid_far;
    uint32_t far_items;                     // far work counted as (row band, XCD group) items (MFMA)
    uint32_t far_qw;                        // MFMA far update: quarters per B fragment read (1, or 2 for L <= 4)
    uint64_t *trash;                        // TRASH_WORDS words that lanes with no element load and store
};

// Words of the per-launch trash area: a lane whose element lies outside the
// matrix loads from and stores to trash[tid] instead, so that every load and
// store of the update kernels is unconditional (no per-lane branches, and exact
// outstanding-memory counts for the compiler's waits).
constexpr uint32_t TRASH_WORDS = 1024;

constexpr uint32_t GATHER_ROWS = 64;   // rows per gather task
constexpr uint32_t XCOLS = 64;         // columns per X task
constexpr uint32_t UT = 64;            // trailing-update tile edge
constexpr uint32_t KSTEP = 64;         // int8 K per MFMA step (one digit plane of a panel)

// Pointers into global memory.  Loads and stores through them compile to global_*
// instructions; through a generic pointer they are flat_*, whose completion also
// holds up every wait for LDS results.
// A value every lane holds alike (read from LDS), as a scalar: branches on it stay
// uniform instead of running both sides under an exec mask.
__device__ __forceinline__ uint32_t uni32(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
__device__ __forceinline__ uint64_t uni64(uint64_t v) {
    return ((uint64_t)uni32((uint32_t)(v >> 32)) << 32) | uni32((uint32_t)v);
}

template <typename T>
using gptr = __attribute__((address_space(1))) T *;
template <typename T>
__device__ __forceinline__ gptr<T> to_global(T *p) {
    return (gptr<T>)p;
}

__device__ __forceinline__ uint32_t cdiv(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

__device__ __forceinline__ uint32_t pad8(uint32_t x) { return (x + 7) & ~7u; }

// Update tiles of one part of a matrix (O rows, `width` columns).  Near tiles are
// numbered row-major.  Far tiles are numbered for the XCDs: a persistent grid of
// a multiple of 8 workgroups walks tile t from workgroup t mod grid, and
// workgroups are dispatched round-robin over the 8 XCDs, so tile t runs on XCD
// t mod 8.  Column band tc goes to XCD tc mod 8 (bands padded to a multiple of 8,
// padding tiles are empty) and each XCD walks its bands fastest, row band by row
// band: an XCD's L2 then holds its own digit column tiles and the current digit
// row tiles instead of all of them.
__device__ __forceinline__ uint32_t part_tiles(uint32_t O, uint32_t width, uint32_t far) {
    const uint32_t C = cdiv(width, UT);
    return (C == 0 || O == 0) ? 0 : cdiv(O, UT) * (far ? pad8(C) : C);
}

// Geometry of tile lt of a part starting at column cbase (relative to c0);
// false for a padding tile.
__device__ __forceinline__ bool tile_geom(uint32_t O, uint32_t cbase, uint32_t width, uint32_t far, uint32_t lt,
                                          uint32_t &r0, uint32_t &cb, uint32_t &nr, uint32_t &nc) {
    const uint32_t C = cdiv(width, UT);
    uint32_t tr, tc;
    if (far) {
        const uint32_t Cg = pad8(C) >> 3, x = lt & 7, j = lt >> 3;
        tr = j / Cg;
        tc = x + 8 * (j - tr * Cg);
        if (tc >= C) return false;
    } else {
        tr = lt / C;
        tc = lt - tr * C;
    }
    r0 = tr * UT;
    cb = cbase + tc * UT;
    nr = min(UT, O - r0);
    nc = min(UT, cbase + width - cb);
    return true;
}

// Entry idx (row * ld + col) of the working copy, in either storage width.
__device__ __forceinline__ uint64_t m_get(const RankMat &S, size_t idx) {
    return S.m32 ? (uint64_t)((const uint32_t *)S.M)[idx] : S.M[idx];
}
__device__ __forceinline__ void m_set(const RankMat &S, size_t idx, uint64_t v) {
    if (S.m32)
        ((uint32_t *)S.M)[idx] = (uint32_t)v;
    else
        S.M[idx] = v;
}

// The X digit planes an iteration writes and its near update reads.
__device__ __forceinline__ int8_t *own_xl(const RankMat &S, const IterState &I) {
    return I.kind == IT_DEFER ? S.XlP : S.Xl;
}

// Column range of the near (far = 0) or far (far = 1) part of the update,
// relative to c0: [base, base + width).
__device__ __forceinline__ void update_range(const IterState &I, uint32_t far, uint32_t &base, uint32_t &width) {
    const uint32_t cols = I.cend - I.c0;
    const uint32_t nw = min(cols, I.split - I.c0);
    base = far ? nw : 0;
    width = far ? cols - nw : nw;
}

// MFMA method: the X column tasks (XCOLS wide, from c0) that the near update
// needs -- those that start below split; the rest are computed after k_finish,
// off the chain that the next panel waits for.
__device__ __forceinline__ uint32_t x_near_tasks(const IterState &I) {
    const uint32_t cols = I.cend - I.c0, nw = min(cols, I.split - I.c0);
    return I.pb > 0 ? cdiv(nw, XCOLS) : 0u;
}

__device__ __forceinline__ uint32_t find_mat(const uint32_t *off, uint32_t nmat, uint32_t t) {
    // largest i with off[i] <= t (off is non-decreasing, off[nmat] > t)
    uint32_t lo = 0, hi = nmat;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= t) lo = mid; else hi = mid;
    }
    return lo;
}

// Balanced signed digits of x in [0, p).
__device__ __forceinline__ int64_t balanced(uint64_t x, uint64_t p) {
    return x > (p - 1) / 2 ? (int64_t)x - (int64_t)p : (int64_t)x;
}

}  // namespace mdl
