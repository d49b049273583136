// Prime-field arithmetic for p < 2^63 on gfx950 (and on the host, for set-up).
//
// Contract: the reference's Field (src/fp.rs:34-114).  Every function returns the
// canonical residue in [0, p), so results are bit-identical to the reference's
// Barrett reduction whatever reduction is used internally.
//
//  * reduce128  -- the reference's own Barrett (fp.rs:71-79): q = mulhi128(t, mu)
//                  with mu = floor((2^128-1)/p), r = t - q p < 3p.  Exact for any
//                  128-bit t; used once per lazily accumulated batch.
//  * mulmod     -- HAC 14.42 Barrett for t = a b < p^2 with k = bitlen(p):
//                  q1 = t >> (k-1), q3 = (q1 mu_k) >> (k+1), mu_k = floor(2^2k / p),
//                  0 <= t - q3 p < 3p.  mu_k is stored pre-shifted by (63-k) so q3
//                  is a single high-word product.  For k <= 62 the remainder fits a
//                  64-bit word; k = 63 takes a 128-bit correction path.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define MDL_HD __host__ __device__ __forceinline__
#else
#define MDL_HD inline
#endif

namespace mdl {

// Reduction mode: generic Barrett, or shift-and-add folding for the Mersenne
// primes 2^61 - 1 and 2^31 - 1 (2^61 == 1, resp. 2^31 == 1).  All modes return
// the same canonical residues.
enum : uint32_t { MODE_GENERIC = 0, MODE_M61 = 1, MODE_M31 = 2 };
constexpr uint64_t P61 = 2305843009213693951ull;  // 2^61 - 1
constexpr uint64_t P31 = 2147483647ull;           // 2^31 - 1

struct FieldParams {
    uint64_t p;
    uint64_t mu_hi, mu_lo;  // floor((2^128 - 1) / p)                     fp.rs:52
    uint64_t mup;           // floor(2^2k / p) << (63 - k), at most 2^64-1
    uint64_t r64;           // 2^64 mod p
    uint64_t safe_batch;    // products of (p-1)^2 that fit in 128 bits    fp.rs:53-59
    uint32_t k;             // bit length of p
    uint32_t big;           // k == 63
    uint32_t mode;          // MODE_*
    uint32_t pad;
};

// Host set-up of the field parameters from p (Field::new, fp.rs:49-61).
inline FieldParams make_params(uint64_t p) {
    typedef unsigned __int128 u128;
    FieldParams f{};
    f.p = p;
    const u128 mu = (~(u128)0) / p;
    f.mu_hi = (uint64_t)(mu >> 64);
    f.mu_lo = (uint64_t)mu;
    const uint32_t k = 64 - __builtin_clzll(p);
    f.k = k;
    f.big = (k == 63) ? 1u : 0u;
    const u128 muk = (((u128)1) << (2 * k)) / p;
    const u128 mup = muk << (63 - k);
    f.mup = (mup >> 64) ? ~(uint64_t)0 : (uint64_t)mup;
    f.r64 = (uint64_t)((((u128)1) << 64) % p);
    const u128 maxp = (u128)(p - 1) * (u128)(p - 1);
    u128 sb = (~(u128)0) / (maxp ? maxp : 1);
    if (sb > (u128)UINT64_MAX) sb = UINT64_MAX;
    if (sb < 1) sb = 1;
    f.safe_batch = (uint64_t)sb;
    f.mode = p == P61 ? MODE_M61 : (p == P31 ? MODE_M31 : MODE_GENERIC);
    return f;
}

// t = th 2^64 + tl mod 2^61 - 1 (any 128-bit t): 2^64 == 8.
MDL_HD uint64_t fold_m61(uint64_t th, uint64_t tl) {
    uint64_t x = (tl & P61) + (tl >> 61) + ((th << 3) & P61) + (th >> 58);  // < 2^63
    x = (x & P61) + (x >> 61);                                               // <= p + 3
    return x >= P61 ? x - P61 : x;
}

// t = th 2^64 + tl mod 2^31 - 1 (any 128-bit t): 2^31 == 1, 2^64 == 4.
MDL_HD uint64_t fold_m31(uint64_t th, uint64_t tl) {
    const uint64_t a = (tl & P31) + ((tl >> 31) & P31) + (tl >> 62);  // == tl, < 2^33
    const uint64_t b = (th & P31) + ((th >> 31) & P31) + (th >> 62);  // == th, < 2^33
    uint64_t x = a + 4 * b;                                             // < 2^36
    x = (x & P31) + (x >> 31);                                          // < 2^31 + 32
    x = (x & P31) + (x >> 31);                                          // <= p + 1
    return x >= P31 ? x - P31 : x;
}

MDL_HD uint64_t umulhi64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul64hi(a, b);
#else
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

// 64x64 -> 128
MDL_HD void mul_wide(uint64_t a, uint64_t b, uint64_t &hi, uint64_t &lo) {
    lo = a * b;
    hi = umulhi64(a, b);
}

// 128-bit add in place: (h, l) += (bh, bl)
MDL_HD void add128(uint64_t &h, uint64_t &l, uint64_t bh, uint64_t bl) {
    uint64_t nl = l + bl;
    h += bh + (nl < l ? 1u : 0u);
    l = nl;
}

MDL_HD uint64_t add_mod(uint64_t a, uint64_t b, uint64_t p) {  // fp.rs:82-85
    uint64_t s = a + b;
    return s >= p ? s - p : s;
}
MDL_HD uint64_t sub_mod(uint64_t a, uint64_t b, uint64_t p) {  // fp.rs:88-90
    return a >= b ? a - b : a + p - b;
}

// t mod p for any 128-bit t (fp.rs:71-79).
MDL_HD uint64_t reduce128(const FieldParams &f, uint64_t th, uint64_t tl) {
    if (f.mode == MODE_M61) return fold_m61(th, tl);
    if (f.mode == MODE_M31) return fold_m31(th, tl);
    // q = high 128 bits of t * mu  (fp.rs:12-24 with four 64x64 partials)
    uint64_t a1 = th, a0 = tl, b1 = f.mu_hi, b0 = f.mu_lo;
    uint64_t lo_hi = umulhi64(a0, b0);
    uint64_t m1h, m1l, m2h, m2l, hh, hl;
    mul_wide(a1, b0, m1h, m1l);
    mul_wide(a0, b1, m2h, m2l);
    mul_wide(a1, b1, hh, hl);
    // carry = lo_hi + m1l + m2l (up to 66 bits)
    uint64_t c = lo_hi + m1l;
    uint64_t cc = (c < lo_hi) ? 1u : 0u;
    uint64_t c2 = c + m2l;
    cc += (c2 < c) ? 1u : 0u;
    // q = (hh, hl) + m1h + m2h + cc
    uint64_t qh = hh, ql = hl;
    add128(qh, ql, 0, m1h);
    add128(qh, ql, 0, m2h);
    add128(qh, ql, 0, cc);
    // r = t - q p   (low 128 bits suffice: 0 <= r < 3p)
    uint64_t qph, qpl;
    mul_wide(ql, f.p, qph, qpl);
    qph += qh * f.p;
    uint64_t rl = tl - qpl;
    uint64_t rh = th - qph - (tl < qpl ? 1u : 0u);
    // r < 3p < 2^65: subtract p while r >= p
    while (rh != 0 || rl >= f.p) {
        uint64_t nl = rl - f.p;
        rh -= (rl < f.p) ? 1u : 0u;
        rl = nl;
    }
    return rl;
}

// t mod p for t < 2^2k (in particular t = a b with a, b < p).
MDL_HD uint64_t reduce_prod(const FieldParams &f, uint64_t th, uint64_t tl) {
    if (f.mode == MODE_M61) return fold_m61(th, tl);
    if (f.mode == MODE_M31) return fold_m31(th, tl);
    const uint32_t k = f.k;
    uint64_t q1 = (th << (65 - k)) | (tl >> (k - 1));
    uint64_t q3 = umulhi64(q1, f.mup);
    if (!f.big) {
        uint64_t r = tl - q3 * f.p;  // exact: 0 <= r < 4p < 2^64
        if (r >= 2 * f.p) r -= 2 * f.p;
        if (r >= f.p) r -= f.p;
        return r;
    }
    uint64_t qph, qpl;
    mul_wide(q3, f.p, qph, qpl);
    uint64_t rl = tl - qpl;
    uint64_t rh = th - qph - (tl < qpl ? 1u : 0u);
    while (rh != 0 || rl >= f.p) {
        uint64_t nl = rl - f.p;
        rh -= (rl < f.p) ? 1u : 0u;
        rl = nl;
    }
    return rl;
}

MDL_HD uint64_t mul_mod(const FieldParams &f, uint64_t a, uint64_t b) {  // fp.rs:93-95
    uint64_t h, l;
    mul_wide(a, b, h, l);
    return reduce_prod(f, h, l);
}

// (pi x - e y) mod p, the fraction-free elimination step.
MDL_HD uint64_t ff_step(const FieldParams &f, uint64_t pi, uint64_t x, uint64_t e, uint64_t y) {
    if (f.mode != MODE_GENERIC) {
        // pi x + (p - e) y < 2 p^2 < 2^128: one folding reduction
        uint64_t h, l, h2, l2;
        mul_wide(pi, x, h, l);
        mul_wide(f.p - e, y, h2, l2);
        add128(h, l, h2, l2);
        return f.mode == MODE_M61 ? fold_m61(h, l) : fold_m31(h, l);
    }
    return sub_mod(mul_mod(f, pi, x), mul_mod(f, e, y), f.p);
}

// a b mod 2^61 - 1 up to a few multiples of p (result < 2^63 + 2^33), for
// a, b < 2^61, from 31-bit halves: with a = a1 2^31 + a0, b = b1 2^31 + b0,
//   a b = a1 b1 2^62 + (a0 b1 + a1 b0) 2^31 + a0 b0,   2^62 == 2,
//   t 2^31 = (t >> 30) 2^61 + (t mod 2^30) 2^31 == (t >> 30) + (t mod 2^30) 2^31,
// so four 32 x 32 -> 64 products and no 128-bit carries.  b is given split:
// b0 = b mod 2^31, b1 = b >> 31.
MDL_HD uint64_t mul_m61_lazy(uint64_t a, uint32_t b0, uint32_t b1) {
    const uint32_t a0 = (uint32_t)a & 0x7fffffffu, a1 = (uint32_t)(a >> 31);
    const uint64_t s = (uint64_t)a0 * b0 + (uint64_t)a1 * (2u * b1);  // < 2^62 + 2^61
    const uint64_t t = (uint64_t)a0 * b1 + (uint64_t)a1 * b0;         // < 2^62
    return s + (t >> 30) + ((t & 0x3fffffffull) << 31);
}
MDL_HD uint64_t fold_m61_64(uint64_t x) {  // x < 2^64 -> [0, 2^61 + 8)
    return (x & P61) + (x >> 61);
}
MDL_HD uint64_t canon_m61(uint64_t x) {    // x < 2^63 -> [0, p)
    x = (x & P61) + (x >> 61);               // <= p + 3
    return x >= P61 ? x - P61 : x;
}

// Compile-time-mode variants (MODE = MODE_*; MODE_GENERIC defers to f.mode).
template <int MODE>
MDL_HD uint64_t mul_mod_t(const FieldParams &f, uint64_t a, uint64_t b) {
    if (MODE == MODE_M61) return canon_m61(fold_m61_64(mul_m61_lazy(a, (uint32_t)b & 0x7fffffffu, (uint32_t)(b >> 31))));
    uint64_t h, l;
    mul_wide(a, b, h, l);
    if (MODE == MODE_M31) return fold_m31(h, l);
    return reduce_prod(f, h, l);
}
template <int MODE>
MDL_HD uint64_t reduce_t(const FieldParams &f, uint64_t th, uint64_t tl) {  // any 128-bit t
    if (MODE == MODE_M61) return fold_m61(th, tl);
    if (MODE == MODE_M31) return fold_m31(th, tl);
    return reduce128(f, th, tl);
}
// Lazy sums of products mod p.  add2(a, b, c, d) adds a b + c d (all < p); any
// number of calls stays exact, and value() is the canonical residue.
//   MODE_M61: with 31-bit halves (a = a1 2^31 + a0),
//     a b == a0 b0 + a1 (2 b1) + (a0 b1 + a1 b0) 2^31   (2^62 == 2 mod 2^61 - 1),
//   so s collects a0 b0 + 2 a1 b1 (< 1.5 2^62 per product) and t the cross terms
//   (< 2^62 per product): four 32 x 32 -> 64 products, no 128-bit carries.  Both
//   are folded to < 2^61 + 8 after every pair, so a pair never overflows 64 bits.
//   Other modes: 128-bit sums of full products, reduced every safe_batch products.
template <int MODE>
struct MacSum {
    uint64_t s = 0, t = 0, r = 0;
    uint64_t cnt = 0;
    MDL_HD void add2(const FieldParams &f, uint64_t a, uint64_t b, uint64_t c, uint64_t d) {
        if constexpr (MODE == MODE_M61) {
            const uint32_t a0 = (uint32_t)a & 0x7fffffffu, a1 = (uint32_t)(a >> 31);
            const uint32_t b0 = (uint32_t)b & 0x7fffffffu, b1 = (uint32_t)(b >> 31);
            const uint32_t c0 = (uint32_t)c & 0x7fffffffu, c1 = (uint32_t)(c >> 31);
            const uint32_t d0 = (uint32_t)d & 0x7fffffffu, d1 = (uint32_t)(d >> 31);
            s += (uint64_t)a0 * b0 + (uint64_t)a1 * (2u * b1) + (uint64_t)c0 * d0 + (uint64_t)c1 * (2u * d1);
            t += (uint64_t)a0 * b1 + (uint64_t)a1 * b0 + (uint64_t)c0 * d1 + (uint64_t)c1 * d0;
            s = (s & P61) + (s >> 61);
            t = (t & P61) + (t >> 61);
        } else {
            if (cnt + 2 > f.safe_batch) {  // safe_batch >= 4 for p < 2^63
                r = add_mod(r, reduce_t<MODE>(f, s, t), f.p);
                s = t = 0;
                cnt = 0;
            }
            uint64_t h, l;
            mul_wide(a, b, h, l);
            add128(s, t, h, l);
            mul_wide(c, d, h, l);
            add128(s, t, h, l);
            cnt += 2;
        }
    }
    MDL_HD uint64_t value(const FieldParams &f) const {
        if constexpr (MODE == MODE_M61) {
            // t 2^31 = (t >> 30) 2^61 + (t mod 2^30) 2^31 == (t >> 30) + (t mod 2^30) 2^31
            return canon_m61(s + (t >> 30) + ((t & 0x3fffffffull) << 31));  // < 2^63
        } else {
            return add_mod(r, reduce_t<MODE>(f, s, t), f.p);
        }
    }
};

// b x + d y mod 2^61 - 1 for x, y, b, d < 2^61, b and d given split into 31-bit
// halves (b0 = b mod 2^31, b1 = b >> 31).  Both products share the two sums
//   s = x0 b0 + x1 2 b1 + y0 d0 + y1 2 d1 < 3 2^62,   t = x0 b1 + x1 b0 + y0 d1 + y1 d0 < 2^63
// (2^62 == 2), so b x + d y == s + t 2^31 and a single fold reduces them.
// lin2_m61_lazy: the same, also for x, y < 2^61 + 8 (then x1, y1 <= 2^30, and s <
// 1.5 2^63, t <= 2^63 still hold), result in [0, p + 7], congruent, not reduced.
MDL_HD uint64_t lin2_m61_lazy(uint64_t x, uint64_t y, uint32_t b0, uint32_t b1, uint32_t d0, uint32_t d1) {
    const uint32_t x0 = (uint32_t)x & 0x7fffffffu, x1 = (uint32_t)(x >> 31);
    const uint32_t y0 = (uint32_t)y & 0x7fffffffu, y1 = (uint32_t)(y >> 31);
    // two independent pairs of multiply-adds per sum (not one chain of four): the
    // dependent depth is two products and an add
    const uint64_t s = ((uint64_t)x0 * b0 + (uint64_t)x1 * (2u * b1)) + ((uint64_t)y0 * d0 + (uint64_t)y1 * (2u * d1));
    const uint64_t t = ((uint64_t)x0 * b1 + (uint64_t)x1 * b0) + ((uint64_t)y0 * d1 + (uint64_t)y1 * d0);
    const uint64_t r = s + (t >> 30) + ((t & 0x3fffffffull) << 31);  // < 2^64
    return (r & P61) + (r >> 61);                                      // <= p + 7
}
MDL_HD uint64_t lin2_m61(uint64_t x, uint64_t y, uint32_t b0, uint32_t b1, uint32_t d0, uint32_t d1) {
    const uint64_t r = lin2_m61_lazy(x, y, b0, b1, d0, d1);
    return r >= P61 ? r - P61 : r;
}
// pi x - e y mod 2^61 - 1 for pi, e < p and x, y < 2^61 + 8, in [0, p + 7].
MDL_HD uint64_t ff_step_m61_lazy(uint64_t pi, uint64_t x, uint64_t e, uint64_t y) {
    const uint64_t ne = P61 - e;
    return lin2_m61_lazy(x, y, (uint32_t)pi & 0x7fffffffu, (uint32_t)(pi >> 31), (uint32_t)ne & 0x7fffffffu,
                         (uint32_t)(ne >> 31));
}
// lin2_m61_lazy for throughput: t first, then s accumulated by multiply-adds onto
// t's two parts, r = (t >> 30) + (t mod 2^30) 2^31 + s < 2^33 + 2^61 + 1.5 2^63 < 2^64.
// Eight multiply-adds and no separate 64-bit adds, at a dependent depth of eight:
// for code that interleaves several independent ones.
MDL_HD uint64_t lin2_m61_lazy_acc(uint64_t x, uint64_t y, uint32_t b0, uint32_t b1, uint32_t d0, uint32_t d1) {
    const uint32_t x0 = (uint32_t)x & 0x7fffffffu, x1 = (uint32_t)(x >> 31);
    const uint32_t y0 = (uint32_t)y & 0x7fffffffu, y1 = (uint32_t)(y >> 31);
    uint64_t t = (uint64_t)x0 * b1;
    t += (uint64_t)x1 * b0;
    t += (uint64_t)y0 * d1;
    t += (uint64_t)y1 * d0;
    uint64_t r = (t >> 30) + ((t & 0x3fffffffull) << 31);
    r += (uint64_t)x0 * b0;
    r += (uint64_t)x1 * (2u * b1);
    r += (uint64_t)y0 * d0;
    r += (uint64_t)y1 * (2u * d1);
    return (r & P61) + (r >> 61);  // <= p + 7
}

template <int MODE>
MDL_HD uint64_t ff_step_t(const FieldParams &f, uint64_t pi, uint64_t x, uint64_t e, uint64_t y) {
    if (MODE == MODE_M61) {
        const uint64_t ne = P61 - e;
        return lin2_m61(x, y, (uint32_t)pi & 0x7fffffffu, (uint32_t)(pi >> 31), (uint32_t)ne & 0x7fffffffu,
                        (uint32_t)(ne >> 31));
    }
    if (MODE == MODE_M31) {
        uint64_t h, l, h2, l2;
        mul_wide(pi, x, h, l);
        mul_wide(f.p - e, y, h2, l2);
        add128(h, l, h2, l2);
        return fold_m31(h, l);
    }
    return ff_step(f, pi, x, e, y);
}

MDL_HD uint64_t pow_mod(const FieldParams &f, uint64_t a, uint64_t e) {  // fp.rs:97-107
    uint64_t r = 1 % f.p;
    while (e) {
        if (e & 1) r = mul_mod(f, r, a);
        a = mul_mod(f, a, a);
        e >>= 1;
    }
    return r;
}

MDL_HD uint64_t inv_mod(const FieldParams &f, uint64_t a) { return pow_mod(f, a, f.p - 2); }  // fp.rs:111-113

// a^(p-2) = 1 / a for a in [1, p) (Fermat, fp.rs:111-113) by left-to-right
// square-and-multiply over the bits of p - 2.  The exponent is the same in every
// lane, so a wavefront inverts 64 values in lockstep (no divergence).
template <int MODE>
MDL_HD uint64_t inv_fermat_t(const FieldParams &f, uint64_t a) {
    const uint64_t e = f.p - 2;
    if (e == 0) return 1;  // p = 2
    int b = 63;
    while (!((e >> b) & 1)) b--;
    uint64_t r = a;
    for (b--; b >= 0; b--) {
        r = mul_mod_t<MODE>(f, r, r);
        if ((e >> b) & 1) r = mul_mod_t<MODE>(f, r, a);
    }
    return r;
}

// 1 / a mod 2^61 - 1 for a in [1, p): a^(p-2) (Fermat, fp.rs:111-113) by an
// addition chain instead of square-and-multiply.  With x_k = a^(2^k - 1):
// x_{j+k} = x_j^(2^k) x_k builds x_2, x_3, x_6, x_12, x_24, x_48, x_54, x_57,
// x_59, and p - 2 = 2^61 - 3 = 4 (2^59 - 1) + 1 gives the inverse as x_59^4 a:
// 60 squarings and 10 products (square-and-multiply takes 60 and 59), lazy in
// [0, 2^61 + 8) until the end.
MDL_HD uint64_t sq_m61_lazy(uint64_t x) {  // x < 2^61 + 8 -> x^2 in [0, 2^61 + 8)
    const uint32_t x0 = (uint32_t)x & 0x7fffffffu, x1 = (uint32_t)(x >> 31);
    const uint64_t s = (uint64_t)x0 * x0 + (uint64_t)x1 * (2u * x1);  // x^2 == s + 2 x0 x1 2^31
    const uint64_t t = ((uint64_t)x0 * x1) << 1;                         // < 2^62
    return fold_m61_64(s + (t >> 30) + ((t & 0x3fffffffull) << 31));
}
MDL_HD uint64_t sqn_m61_lazy(uint64_t x, int n) {
    for (int i = 0; i < n; i++) x = sq_m61_lazy(x);
    return x;
}
MDL_HD uint64_t mulf_m61_lazy(uint64_t x, uint64_t y) {
    return fold_m61_64(mul_m61_lazy(x, (uint32_t)y & 0x7fffffffu, (uint32_t)(y >> 31)));
}
MDL_HD uint64_t inv_m61(uint64_t a) {
    const uint64_t x2 = mulf_m61_lazy(sq_m61_lazy(a), a);
    const uint64_t x3 = mulf_m61_lazy(sq_m61_lazy(x2), a);
    const uint64_t x6 = mulf_m61_lazy(sqn_m61_lazy(x3, 3), x3);
    const uint64_t x12 = mulf_m61_lazy(sqn_m61_lazy(x6, 6), x6);
    const uint64_t x24 = mulf_m61_lazy(sqn_m61_lazy(x12, 12), x12);
    const uint64_t x48 = mulf_m61_lazy(sqn_m61_lazy(x24, 24), x24);
    const uint64_t x54 = mulf_m61_lazy(sqn_m61_lazy(x48, 6), x6);
    const uint64_t x57 = mulf_m61_lazy(sqn_m61_lazy(x54, 3), x3);
    const uint64_t x59 = mulf_m61_lazy(sqn_m61_lazy(x57, 2), x2);
    return canon_m61(mulf_m61_lazy(sqn_m61_lazy(x59, 2), a));
}

// Inverse of a in [1, p) by the binary extended Euclidean algorithm (Knuth,
// TAOCP vol. 2, 4.5.2): about 2 bitlen(p) short steps instead of the ~2 bitlen(p)
// dependent Barrett products of Fermat.  Any exact inverse is bit-identical to
// the reference's a^(p-2).  p = 2 has the single unit 1.
MDL_HD uint64_t inv_mod_bin(uint64_t a, uint64_t p) {
    if (p == 2) return 1;
    uint64_t u = a, v = p, x1 = 1, x2 = 0;
    while (u != 1 && v != 1) {
        while ((u & 1) == 0) {
            u >>= 1;
            x1 = (x1 & 1) ? (x1 + p) >> 1 : x1 >> 1;  // x1 + p < 2^64 since x1, p < 2^63
        }
        while ((v & 1) == 0) {
            v >>= 1;
            x2 = (x2 & 1) ? (x2 + p) >> 1 : x2 >> 1;
        }
        if (u >= v) {
            u -= v;
            x1 = sub_mod(x1, x2, p);
        } else {
            v -= u;
            x2 = sub_mod(x2, x1, p);
        }
    }
    return u == 1 ? x1 : x2;
}

}  // namespace mdl
