// Host helpers that synthesise workloads (planted-deficiency matrices, linear
// recurrences) with exact 128-bit modular arithmetic.  They play the role of the
// reference's fill loops (main.rs:86-105) for inputs the generators alone do
// not produce.
#include <stdint.h>

#include "../../include/modlin.h"

namespace {
typedef unsigned __int128 u128;
bool valid_p(uint64_t p) { return p >= 2 && p < (1ull << 63); }
}  // namespace

extern "C" int modlin_util_combine_rows(uint64_t p, const uint64_t *rows, uint64_t nsrc, uint64_t n,
                                        const uint64_t *coef, uint64_t nout, uint64_t *out) {
    if (!valid_p(p) || (nout && n && (!rows || !coef || !out))) return MODLIN_EINVAL;
    for (uint64_t i = 0; i < nout; i++) {
        uint64_t *o = out + i * n;
        for (uint64_t c = 0; c < n; c++) o[c] = 0;
        for (uint64_t k = 0; k < nsrc; k++) {
            const uint64_t a = coef[i * nsrc + k] % p;
            if (!a) continue;
            const uint64_t *r = rows + k * n;
            for (uint64_t c = 0; c < n; c++) o[c] = (uint64_t)(((u128)a * (r[c] % p) + o[c]) % p);
        }
    }
    return MODLIN_OK;
}

extern "C" int modlin_util_lfsr(uint64_t p, const uint64_t *coef, uint64_t L, const uint64_t *init, uint64_t n,
                                uint64_t *out) {
    if (!valid_p(p) || (n && !out) || (L && (!coef || !init))) return MODLIN_EINVAL;
    // products of (p-1)^2 that fit a u128 before a reduction (at least 3 for p < 2^63)
    const u128 maxp = (u128)(p - 1) * (p - 1);
    u128 sb = maxp ? (~(u128)0 - p) / maxp : 64;
    const uint64_t batch = sb > 64 ? 64 : (uint64_t)sb;
    for (uint64_t i = 0; i < n && i < L; i++) out[i] = init[i] % p;
    for (uint64_t i = L; i < n; i++) {
        u128 acc = 0;
        uint64_t v = 0, cnt = 0;
        for (uint64_t j = 0; j < L; j++) {
            acc += (u128)(coef[j] % p) * out[i - 1 - j];
            if (++cnt == batch) {
                v = (uint64_t)((acc % p + v) % p);
                acc = 0;
                cnt = 0;
            }
        }
        out[i] = (uint64_t)((acc % p + v) % p);
    }
    return MODLIN_OK;
}
