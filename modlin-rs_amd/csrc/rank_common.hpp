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
    uint64_t *Eta;     // row operations of the last block elimination (ETA_* layout)
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
    unsigned long long upd_bytes;        // algorithmic HBM bytes of the trailing updates
};

// Record of one block elimination, written by k_block_ge and replayed by
// k_block_z (u64 offsets into RankMat::Eta): the multiplier e of every block row
// at every step, and per step the pivot, the pivot row's own-row coefficient and
// the swapped position; per final row its own-row coefficient and 1 / D.
constexpr uint32_t ETA_E = 0;                          // [step][row]
constexpr uint32_t ETA_PI = BMAX * BMAX;               // [step]
constexpr uint32_t ETA_OWNK = ETA_PI + BMAX;           // [step]
constexpr uint32_t ETA_OWNF = ETA_OWNK + BMAX;         // [row]
constexpr uint32_t ETA_DINV = ETA_OWNF + BMAX;         // [row]
constexpr uint32_t ETA_SWQ = ETA_DINV + BMAX;          // [step]
constexpr uint32_t ETA_WORDS = ETA_SWQ + BMAX;
constexpr uint32_t ZG = 4;                             // workgroups per matrix in k_block_z

// Work plan of one iteration: per-matrix tile offsets (prefix sums).
struct RankPlan {
    uint32_t gx_total, upd_total;
    uint32_t pad0, pad1;
};

// Constants of the signed-digit limb representation used by the MFMA update.
// An element x in [0, p) is first mapped to its balanced residue
// x^ = x - p [x > (p-1)/2], |x^| <= (p-1)/2, then written with L signed 8-bit
// digits: x^ = sum_l d_l 256^l, d_l in [-128, 127].
struct LimbParams {
    uint64_t pow256[16];  // 256^l mod p
    uint64_t k0;          // 2^49 * sum_{i : 4i < 2L-1} 256^(4i) mod p (epilogue offset)
    uint32_t L;           // digits per element
    uint32_t pad;
};

constexpr uint32_t GATHER_ROWS = 256;  // rows per gather task
constexpr uint32_t XCOLS = 64;         // columns per X task
constexpr uint32_t UT = 64;            // trailing-update tile edge
constexpr uint32_t KSTEP = 64;         // int8 K per MFMA step (one digit plane of a panel)

__device__ __forceinline__ uint32_t cdiv(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

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
