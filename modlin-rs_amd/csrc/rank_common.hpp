// Shared declarations of the rank engine (device state layout, launch helpers).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fp_dev.hpp"

namespace mdl {

// Panel width and height of the pivot-search block.  The trailing update is a
// rank-pb update with pb <= BMAX.
constexpr int BMAX = 64;

// One matrix of a rank batch, resident in device memory.  The matrix is the
// library's working copy (row-major, leading dimension ld); rows are addressed
// through perm (logical position -> physical row) so that pivoting swaps
// indices instead of 8n-byte rows.
struct RankMat {
    uint64_t *M;       // working copy, n x n, row stride ld
    uint32_t *perm;    // logical -> physical row
    uint64_t *Ab;      // gathered A operand M[O, J]: n x BMAX (row stride BMAX)
    uint64_t *X;       // A^-1 M[R, c0:], BMAX x n (row stride n)
    uint64_t *Z;       // A^-1, BMAX x BMAX
    int8_t *Al;        // int8 limb planes of Ab (MFMA path)
    int8_t *Xl;        // int8 limb planes of the shifted X (MFMA path)
    uint32_t n, ld;
    uint32_t rank, col;      // pivots found; first column not yet eliminated
    uint32_t jend, pb;       // current panel end; pivots in the current panel
    uint32_t c0, cend;       // trailing-update column range of this iteration
    uint32_t resid;          // the panel needs a residual (missed-pivot) check
    uint32_t first_nz_col;   // smallest panel column with a nonzero residual
    uint32_t done;
    uint32_t iters;
    uint32_t J[BMAX];        // pivot columns of the panel (absolute)
    uint32_t R[BMAX];        // physical pivot rows, in pivot order
    unsigned long long macs; // F_p multiply-accumulates of the trailing updates
    unsigned long long upd_launch_macs;  // MACs of this iteration's update
};

// Work plan of one iteration: per-matrix tile offsets (prefix sums).
struct RankPlan {
    uint32_t gx_total, upd_total;
    uint32_t pad0, pad1;
};

}  // namespace mdl
