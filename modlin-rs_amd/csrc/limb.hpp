// Signed-digit limb representation of F_p elements and the exact recombination
// of int8 limb products (the epilogue of the matrix-core trailing update).
//
// An element x in [0, p) is mapped to its balanced residue
// x^ = x - p [x > (p-1)/2], |x^| <= (p-1)/2, and written with L signed 8-bit
// digits x^ = sum_l d_l 256^l, d_l in [-128, 127].  A product of two K-deep
// digit panels (K <= 64) is A X = sum_s 256^s P_s over the NC = 2L-1 shift
// classes s = la + lb, P_s = sum_{la+lb=s} A_la X_lb, each accumulated exactly in
// int32 by the int8 MFMA: a class has at most 8 digit pairs and
// -16256 <= d_a d_b <= 16384, so -2^23 < P_s <= 2^23.
//
// Recombination.  Two neighbouring classes fit one 32-bit word:
//   P_2t + 256 P_2t+1  lies in [-HOFF, 2^32 - HOFF), HOFF = 8 * 64 * 16256 * 257,
// so with the offset HOFF added (it is the initial value of the even classes'
// accumulators, which then stay inside int32) the pair word
//   h_t = acc_2t + 256 acc_2t+1   (mod 2^32)
// is exact and non-negative.  The NH = L pair words sit 16 bits apart:
//   S = sum_t h_t 2^(16 t) = sum_s 256^s P_s + sum_t HOFF 2^(16 t),
// whose even words are disjoint 32-bit fields and whose odd words are the same
// shifted by 16, so S is a two-operand add of 32-bit word strings.  It is
// reduced mod p (Mersenne folds for 2^61 - 1 and 2^31 - 1, otherwise weights
// 2^(32 i) mod p and one Barrett reduction), and the constant
// k0 = sum_t HOFF 2^(16 t) mod p is subtracted by the caller.  Every step is
// integer-exact, so the result is the canonical residue of the product.
#pragma once
#include <stdint.h>

#include "fp_dev.hpp"

namespace mdl {

constexpr uint32_t HOFF = 2139029504u;  // 8 pairs * 64 terms * 16256 * 257

// Two panels accumulated together (K <= 128, the far update of a deferred
// pair of panels): a class sum then lies in [-2^24, 2^24] (8 pairs x 128 terms
// x 2^14), too wide for the pair words.  Every class accumulator instead starts
// at KOFF2 = 2^25 and stays in [2^24, 3 2^24], a non-negative 26-bit value, so
// S' = sum_s 256^s P'_s = sum_s 256^s P_s + k2 with k2 = KOFF2 sum_s 256^s mod p.
constexpr int32_t KOFF2 = 1 << 25;

struct LimbParams {
    uint64_t pow256[16];  // 256^l mod p
    uint64_t pow32[5];    // 2^(32 i) mod p: weights of the recombined 32-bit words
    uint64_t k0;          // sum_{t < L} HOFF 2^(16 t) mod p (the pair words' offsets)
    uint64_t k2;          // sum_{s < 2L-1} KOFF2 256^s mod p (the offsets of limb_residue_k2)
    uint64_t kc2;         // k2 + 2p (limb_sub_m61_k2)
    uint32_t L;           // digits per element
    uint32_t pad;
};

// Digits needed for the balanced residues: the smallest L with
// 127 (256^L - 1) / 255 >= (p - 1) / 2.
inline uint32_t limb_count(uint64_t p) {
    typedef unsigned __int128 u128;
    const u128 half = (p - 1) / 2;
    u128 cap = 127, pw = 1;
    uint32_t L = 1;
    while (cap < half) {
        pw *= 256;
        cap += 127 * pw;
        L++;
    }
    return L;
}

inline LimbParams make_limb_params(uint64_t p) {
    typedef unsigned __int128 u128;
    LimbParams lp{};
    lp.L = limb_count(p);
    u128 x = 1 % p;
    for (int l = 0; l < 16; l++) {
        lp.pow256[l] = (uint64_t)x;
        x = (x * 256) % p;
    }
    for (int i = 0; i < 4; i++) lp.pow32[i] = lp.pow256[4 * i];
    lp.pow32[4] = (uint64_t)(((u128)lp.pow256[12] * lp.pow256[4]) % p);  // 2^128
    u128 k0 = 0, w = 1 % p;  // w = 2^(16 t) mod p
    for (uint32_t t = 0; t < lp.L; t++) {
        k0 = (k0 + (u128)HOFF % p * w) % p;
        w = (w * 65536) % p;
    }
    lp.k0 = (uint64_t)k0;
    u128 k2 = 0;
    for (uint32_t c = 0; c + 1 < 2 * lp.L; c++) k2 = (k2 + (u128)KOFF2 % p * lp.pow256[c]) % p;
    lp.k2 = (uint64_t)k2;
    lp.kc2 = (uint64_t)(k2 + 2 * (u128)p);
    return lp;
}

// S = sum_t h_t 2^(16 t) as five 32-bit words (NH <= 8 pair words).
template <int NH>
MDL_HD void pair_words_sum(const uint32_t *h, uint32_t *s) {
    uint32_t e[5] = {0, 0, 0, 0, 0}, q[5] = {0, 0, 0, 0, 0};
#pragma unroll
    for (int t = 0; t < NH; t += 2) e[t / 2] = h[t];
#pragma unroll
    for (int t = 1; t < NH; t += 2) {  // h_t 2^(16 t) = (h_t << 16) 2^(32 (t-1)/2)
        q[(t - 1) / 2] |= h[t] << 16;
        q[(t + 1) / 2] |= h[t] >> 16;
    }
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < 5; i++) {
        const uint64_t v = (uint64_t)e[i] + q[i] + c;
        s[i] = (uint32_t)v;
        c = (uint32_t)(v >> 32);
    }
}

// Residue mod p of S = sum_s 256^s P_s + sum_t HOFF 2^(16 t), from the class
// accumulators acc[0..NC) of one element (even classes started at HOFF, odd
// ones at 0).  Returns the canonical residue in [0, p).
template <int MODE, int NC>
MDL_HD uint64_t limb_residue(const FieldParams &f, const LimbParams &lp, const int32_t *acc) {
    constexpr int NH = (NC + 1) / 2;
    uint32_t h[NH];
#pragma unroll
    for (int t = 0; t < NH; t++) {
        uint32_t v = (uint32_t)acc[2 * t];
        if (2 * t + 1 < NC) v += (uint32_t)acc[2 * t + 1] << 8;
        h[t] = v;
    }
    if constexpr (MODE == MODE_M31 && NH == 4) {
        // 2^32 == 2, 2^48 == 2^17 (mod 2^31 - 1): S == h0 + 2^16 h1 + 2 h2 + 2^17 h3 < 2^50
        const uint64_t x = (uint64_t)h[0] + ((uint64_t)h[1] << 16) + ((uint64_t)h[2] << 1) + ((uint64_t)h[3] << 17);
        return fold_m31(0, x);
    }
    uint32_t s[5];
    pair_words_sum<NH>(h, s);
    if constexpr (MODE == MODE_M61 && NH == 8) {
        // S = lo + hi 2^64 + s4 2^128 with 2^64 == 8 and 2^128 == 2^6 (mod 2^61 - 1)
        const uint64_t lo = ((uint64_t)s[1] << 32) | s[0], hi = ((uint64_t)s[3] << 32) | s[2];
        uint64_t x = (lo & P61) + (lo >> 61) + ((hi << 3) & P61) + (hi >> 58) + ((uint64_t)s[4] << 6);  // < 2^63
        x = (x & P61) + (x >> 61);                                                                       // <= p + 3
        return x >= P61 ? x - P61 : x;
    }
    // weights 2^(32 i) mod p; the sum is below 5 * 2^95 < 2^98
    constexpr int NS = (16 * (NH - 1) + 32) / 32 + 1;
    uint64_t th = 0, tl = s[0];
#pragma unroll
    for (int i = 1; i < NS; i++) {
        uint64_t ph, pl;
        mul_wide((uint64_t)s[i], lp.pow32[i], ph, pl);
        add128(th, tl, ph, pl);
    }
    return f.k >= 49 ? reduce_prod(f, th, tl) : reduce128(f, th, tl);
}

// Residue mod p of S' = sum_s 256^s P'_s from the class accumulators of two
// accumulated panels (every class started at KOFF2, so each P'_s lies in
// [0, 2^26)); the caller subtracts k2.  Returns the canonical residue.
template <int MODE, int NC>
MDL_HD uint64_t limb_residue_k2(const FieldParams &f, const LimbParams &lp, const int32_t *acc) {
    if constexpr (MODE == MODE_M31 && NC == 7) {
        // 256^s == 2^(8 s mod 31): exponents 0, 8, 16, 24, 1, 9, 17; the sum is below 2^53
        uint64_t x = 0;
#pragma unroll
        for (int c = 0; c < NC; c++) x += (uint64_t)(uint32_t)acc[c] << ((8 * c) % 31);
        return fold_m31(0, x);
    }
    if constexpr (NC <= 5) {
        // S' < 2^26 (1 + 2^8 + ... + 2^32) < 2^59 fits one word
        uint64_t x = 0;
#pragma unroll
        for (int c = 0; c < NC; c++) x += (uint64_t)(uint32_t)acc[c] << (8 * c);
        return reduce128(f, 0, x);
    }
    // weights 256^s mod p: each product is below 2^89, the sum below 2^93
    uint64_t th = 0, tl = 0;
#pragma unroll
    for (int c = 0; c < NC; c++) {
        uint64_t ph, pl;
        mul_wide((uint64_t)(uint32_t)acc[c], lp.pow256[c], ph, pl);
        add128(th, tl, ph, pl);
    }
    return reduce128(f, th, tl);
}

// Recombination of the 15 class sums P_s of 2^61 - 1 (L = 8), each in [0, 2^26)
// (two accumulated panels, every class started at KOFF2), into a 64-bit value
// congruent to S' = sum_s 2^(8 s mod 61) P_s.  Classes c and c + 8 (c < 7) have
// the weights 2^(8c) and 2^(8c + 3) (2^64 == 2^3), so t_c = P_c + 8 P_{c+8} <
// 2^29.2 carries the weight 2^(8c); class 7 (2^56) stands alone.  The parts:
//   lo = 2^8 t_1 + 2^16 t_2 + 2^24 t_3                                 < 2^53.2
//   hi = t_4 + 2^8 (t_5 mod 2^21) + 2^16 (t_6 mod 2^13) + 2^24 (P_7 mod 2^5)
//                                                      < 2^31.1, weight 2^32
//   w  = t_0 + (t_5 >> 21) + (t_6 >> 13) + (P_7 >> 5)                  < 2^30
// (t 2^(8c) for c = 5, 6 and P_7 2^56 are split at bit 61, whose weight is 1),
// and x = lo + 2^32 hi + w < 2^64.  Mostly 32-bit work: about half the
// instructions of rotating every class into a 64-bit sum.  The pairs can be
// added in any order (the update kernels add each as soon as both classes are
// complete).
struct M61Sum {
    uint64_t lo;
    uint32_t hi, w;
};

MDL_HD M61Sum m61_sum_init() { return M61Sum{0, 0, 0}; }

// Adds classes C and C + 8 (C < 7).
template <int C>
MDL_HD void m61_sum_pair(M61Sum &a, uint32_t pc, uint32_t pc8) {
    static_assert(C >= 0 && C < 7, "class pairs (0, 8) .. (6, 14)");
    const uint32_t t = pc + (pc8 << 3);
    if constexpr (C == 0) {
        a.w += t;
    } else if constexpr (C <= 3) {
        a.lo += (uint64_t)t << (8 * C);
    } else if constexpr (C == 4) {
        a.hi += t;
    } else {
        constexpr int S = 61 - 8 * C;  // bits of t below 2^61: 21 or 13
        a.hi += (t & ((1u << S) - 1u)) << (8 * C - 32);
        a.w += t >> S;
    }
}

// Adds class 7 (weight 2^56).
MDL_HD void m61_sum_class7(M61Sum &a, uint32_t p7) {
    a.hi += (p7 & 31u) << 24;
    a.w += p7 >> 5;
}

// (cv - (S' - k2)) mod 2^61 - 1, canonical, for cv in [0, p); kc2 = k2 + 2p.
MDL_HD uint64_t m61_sum_sub(uint64_t cv, uint64_t kc2, const M61Sum &a) {
    const uint64_t x = a.lo + (((uint64_t)a.hi << 32) | a.w);
    const uint64_t y = (x & P61) + (x >> 61);  // <= p + 7, congruent to S'
    // r = cv + k2 + 2p - y in [p - 7, 4p), congruent to cv - (S' - k2)
    uint64_t r = cv + kc2 - y;
    r = (r & P61) + (r >> 61);  // <= p + 3
    return r >= P61 ? r - P61 : r;
}

// (cv - sum_s 256^s P_s) mod 2^61 - 1 for cv in [0, p), from the 15 class
// accumulators of two accumulated panels of L = 8 digits (every class started
// at KOFF2, so each lies in [0, 2^26)); kc2 = k2 + 2p.
MDL_HD uint64_t limb_sub_m61_k2(uint64_t cv, uint64_t kc2, const int32_t *acc) {
    M61Sum a = m61_sum_init();
    m61_sum_pair<0>(a, (uint32_t)acc[0], (uint32_t)acc[8]);
    m61_sum_pair<1>(a, (uint32_t)acc[1], (uint32_t)acc[9]);
    m61_sum_pair<2>(a, (uint32_t)acc[2], (uint32_t)acc[10]);
    m61_sum_pair<3>(a, (uint32_t)acc[3], (uint32_t)acc[11]);
    m61_sum_pair<4>(a, (uint32_t)acc[4], (uint32_t)acc[12]);
    m61_sum_pair<5>(a, (uint32_t)acc[5], (uint32_t)acc[13]);
    m61_sum_pair<6>(a, (uint32_t)acc[6], (uint32_t)acc[14]);
    m61_sum_class7(a, (uint32_t)acc[7]);
    return m61_sum_sub(cv, kc2, a);
}

}  // namespace mdl
