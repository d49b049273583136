// Null distributions of the corank and of the linear complexity, the one-sided
// p-values, and their printing (reference: src/stats.rs:14-137).
//
// Bit-level agreement with the reference follows from using the same
// operations in the same order: glibc log / log1p / exp (what Rust's f64::ln,
// ln_1p and exp call), an explicit fma for f64::mul_add (stats.rs:22), the
// run-time powi routine (binary exponentiation, reciprocal for negative
// exponents), round-to-nearest u64 -> f64 conversions, and no contraction of
// a*b+c (this file is compiled with -ffp-contract=off).
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>

#include "stats.hpp"

namespace modlin {
namespace stats {

namespace {

// Rust's f64::powi at run time: multiply on odd bits, square, halve the
// exponent; reciprocal at the end for a negative exponent.
double powi(double a, int b) {
    const bool recip = b < 0;
    double r = 1.0;
    for (;;) {
        if (b & 1) r *= a;
        b /= 2;
        if (b == 0) break;
        a *= a;
    }
    return recip ? 1.0 / r : r;
}

constexpr double LN_2 = 0.693147180559945309417232121458176568;

double ln_u64(uint64_t p) {  // stats.rs:19-23
    const int e = 63 - __builtin_clzll(p);
    const double mantissa = (double)p / (double)(1ull << e);
    return fma((double)e, LN_2, log(mantissa));
}

double sum_ln_1m(double p, uint64_t a, uint64_t b) {  // stats.rs:26-41
    if (a > b) return 0.0;
    const double inv = 1.0 / p;
    double pj = powi(inv, (int)a);
    double s = 0.0;
    for (uint64_t j = a; j <= b; j++) {
        if (pj < 1e-18) break;
        s += log1p(-pj);
        pj *= inv;
    }
    return s;
}

double clamp(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }

}  // namespace

double corank_prob(uint64_t p, uint64_t n, uint64_t d) {  // stats.rs:52-62
    if (d > n) return 0.0;
    const double pf = (double)p;
    const double lp = -powi((double)d, 2) * ln_u64(p) + 2.0 * sum_ln_1m(pf, d + 1, n) - sum_ln_1m(pf, 1, n - d);
    return lp < -745.0 ? 0.0 : exp(lp);
}

double corank_tail_pvalue(uint64_t p, uint64_t n, uint64_t c) {  // stats.rs:69-84
    if (c == 0) return 1.0;
    double s = 0.0;
    for (uint64_t d = c; d <= n; d++) {
        const double pr = corank_prob(p, n, d);
        if (pr == 0.0) break;
        s += pr;
    }
    return clamp(s, P_FLOOR, 1.0);
}

double lc_left_tail_pvalue(uint64_t p, uint64_t n, uint64_t ell) {  // stats.rs:101-126
    if (ell >= n) return 1.0;
    const double lnp = ln_u64(p);
    double pr;
    if (2 * ell <= n) {
        const double a = (double)(2 * ell + 1) * lnp;
        const double ln_num = a > 0.0 ? a + log1p(exp(-a)) : log1p(exp(a));
        const double ln_den = (double)n * lnp + log((double)p + 1.0);
        pr = exp(ln_num - ln_den);
    } else {
        const double hi = exp(((double)n - 2.0 * (double)ell) * lnp);
        const double lo = exp(-(double)n * lnp);
        pr = 1.0 - (hi - lo) / ((double)p + 1.0);
    }
    return clamp(pr, P_FLOOR, 1.0);
}

// Rust's `{:?}` for a finite f64: the shortest digit string that parses back to
// x, in decimal notation for 1e-4 <= |x| < 1e16 (an integer keeps ".0") and in
// "d.ddde-N" notation otherwise (no '+', no zero padding).  The digits are the
// correctly rounded ones at the smallest precision that round-trips.
std::string debug_f64(double x) {
    if (x == 0.0) return signbit(x) ? "-0.0" : "0.0";
    if (isnan(x)) return "NaN";
    if (isinf(x)) return x > 0 ? "inf" : "-inf";
    char buf[64];
    int prec = 0;
    for (; prec < 17; prec++) {
        snprintf(buf, sizeof buf, "%.*e", prec, x);
        if (strtod(buf, nullptr) == x) break;
    }
    snprintf(buf, sizeof buf, "%.*e", prec, x);
    std::string digits;
    const char *q = buf;
    const bool neg = *q == '-';
    if (neg) q++;
    for (; *q && *q != 'e'; q++)
        if (*q >= '0' && *q <= '9') digits.push_back(*q);
    const int e10 = atoi(q + 1);
    while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
    std::string out = neg ? "-" : "";
    const double ax = fabs(x);
    const int nd = (int)digits.size();
    if (ax >= 1e-4 && ax < 1e16) {
        if (e10 >= 0) {
            for (int i = 0; i <= e10; i++) out.push_back(i < nd ? digits[i] : '0');
            out.push_back('.');
            if (nd > e10 + 1)
                out.append(digits, e10 + 1, std::string::npos);
            else
                out.push_back('0');
        } else {
            out += "0.";
            out.append((size_t)(-e10 - 1), '0');
            out += digits;
        }
    } else {
        out.push_back(digits[0]);
        if (nd > 1) {
            out.push_back('.');
            out.append(digits, 1, std::string::npos);
        }
        out += "e" + std::to_string(e10);
    }
    return out;
}

std::string pretty_p_value(double p) {  // stats.rs:129-137
    if (p == 0.0) return "0";
    if (p == 1.0) return "1";
    return debug_f64(p);
}

}  // namespace stats
}  // namespace modlin
