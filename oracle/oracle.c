/*
 * CPU oracle for the modlin hot paths -- TEST INFRASTRUCTURE ONLY.
 *
 * This file restates, in plain C, the algorithms of the reference crate
 * (vigna/modlin-rs) that the GPU library must reproduce bit-exactly:
 *
 *   - the prime-field arithmetic with Barrett reduction   (src/fp.rs:12-114)
 *   - the blocked right-looking rank                      (src/fp.rs:116-255)
 *   - Berlekamp-Massey linear complexity                  (src/fp.rs:257-308)
 *   - the naive checkers of the reference's own tests     (tests/test_rank.rs:44-72,
 *                                                          tests/test_linear_complexity.rs:37-70)
 *   - the null distributions and p-value formatting       (src/stats.rs:14-137)
 *   - the input generators                                (src/prng.rs:46-343)
 *
 * Only the test suite, the smoke check and the CPU-baseline leg of the
 * benchmark load this code, and only as the checker.  The product library
 * never links it and never falls back to it.
 *
 * Pinning (what each function is checked against, in tests/):
 *   of_* Barrett        : exact u128 '%' over the reference test's primes, corners
 *                         and mixed samples (fp.rs:314-367).
 *   or_rank             : or_rank_ref (test_rank.rs:44-72) on the reference's
 *                         test family (test_rank.rs:99-149); known answers
 *                         (test_rank.rs:74-97, README corank 432 / 0); GF(2)
 *                         bit-packed elimination; exhaustive determinants n<=3.
 *   or_lc               : or_lc_ref (test_linear_complexity.rs:37-70) on the
 *                         reference's family; known answers (3, 1, 0; README 272,
 *                         503, 5000, 65280); LFSR order invariant.
 *   or_stats            : test_stats.rs known answers and the pmf-sum identity;
 *                         README printed strings.  Bit-level equality beyond those
 *                         values is by construction only ("parity unpinned" for
 *                         (p, n, c) outside them).
 *   or_prng             : test_prng.rs known-answer vectors; xoshiro256** and
 *                         SplitMix64 against the published reference outputs.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef unsigned __int128 u128;

/* ------------------------------------------------------------------------ */
/* Field arithmetic (fp.rs:12-114)                                           */
/* ------------------------------------------------------------------------ */

typedef struct {
    uint64_t p;
    u128 mu;             /* floor((2^128 - 1) / p)              fp.rs:52 */
    uint64_t safe_batch; /* products of (p-1)^2 that fit in u128 fp.rs:53-59 */
} ofield;

/* High 128 bits of the 256-bit product (fp.rs:12-24). */
static inline u128 of_mul_hi_128(u128 a, u128 b) {
    u128 a1 = a >> 64, a0 = (uint64_t)a;
    u128 b1 = b >> 64, b0 = (uint64_t)b;
    u128 lo = a0 * b0, m1 = a1 * b0, m2 = a0 * b1, hi = a1 * b1;
    u128 carry = (lo >> 64) + (uint64_t)m1 + (uint64_t)m2;
    return hi + (m1 >> 64) + (m2 >> 64) + (carry >> 64);
}

/* Field::new (fp.rs:49-61).  Returns 0 on success, -1 for p outside [2, 2^63). */
int of_init(ofield *f, uint64_t p) {
    if (p < 2 || p >= (1ull << 63)) return -1;
    f->p = p;
    f->mu = (~(u128)0) / p;
    u128 max_prod = (u128)(p - 1) * (u128)(p - 1);
    u128 sb = (~(u128)0) / max_prod;
    if (sb > (u128)UINT64_MAX) sb = UINT64_MAX;
    if (sb < 1) sb = 1;
    f->safe_batch = (uint64_t)sb;
    return 0;
}

/* Field::reduce_u128 (fp.rs:71-79). */
static inline uint64_t of_reduce_u128(const ofield *f, u128 t) {
    u128 q = of_mul_hi_128(t, f->mu);
    u128 p = f->p;
    u128 r = t - q * p;
    while (r >= p) r -= p;
    return (uint64_t)r;
}
static inline uint64_t of_add(const ofield *f, uint64_t a, uint64_t b) { /* fp.rs:82-85 */
    uint64_t s = a + b;
    return s >= f->p ? s - f->p : s;
}
static inline uint64_t of_sub(const ofield *f, uint64_t a, uint64_t b) { /* fp.rs:88-90 */
    return a >= b ? a - b : a + f->p - b;
}
static inline uint64_t of_mul(const ofield *f, uint64_t a, uint64_t b) { /* fp.rs:93-95 */
    return of_reduce_u128(f, (u128)a * b);
}
static uint64_t of_pow(const ofield *f, uint64_t a, uint64_t e) { /* fp.rs:97-107 */
    uint64_t r = 1;
    while (e) {
        if (e & 1) r = of_mul(f, r, a);
        a = of_mul(f, a, a);
        e >>= 1;
    }
    return r;
}
static inline uint64_t of_inv(const ofield *f, uint64_t a) { return of_pow(f, a, f->p - 2); } /* fp.rs:111-113 */

/* ctypes-facing wrappers (u128 split into hi/lo words). */
uint64_t or_field_mu_hi(uint64_t p) { ofield f; of_init(&f, p); return (uint64_t)(f.mu >> 64); }
uint64_t or_field_mu_lo(uint64_t p) { ofield f; of_init(&f, p); return (uint64_t)f.mu; }
uint64_t or_field_safe_batch(uint64_t p) { ofield f; of_init(&f, p); return f.safe_batch; }
uint64_t or_reduce_u128(uint64_t p, uint64_t hi, uint64_t lo) {
    ofield f; of_init(&f, p);
    return of_reduce_u128(&f, ((u128)hi << 64) | lo);
}
uint64_t or_reduce(uint64_t p, uint64_t x) { ofield f; of_init(&f, p); return of_reduce_u128(&f, x); }
uint64_t or_mul(uint64_t p, uint64_t a, uint64_t b) { ofield f; of_init(&f, p); return of_mul(&f, a, b); }
uint64_t or_inv(uint64_t p, uint64_t a) { ofield f; of_init(&f, p); return of_inv(&f, a); }
/* Exact remainders with the compiler's u128 division: the closed form Barrett must equal. */
uint64_t or_u128_mod(uint64_t hi, uint64_t lo, uint64_t p) { return (uint64_t)((((u128)hi << 64) | lo) % p); }

/* ------------------------------------------------------------------------ */
/* Blocked rank (fp.rs:116-255)                                              */
/* ------------------------------------------------------------------------ */

#define OR_PANEL 64  /* fp.rs:29 */
#define OR_TILE 256  /* fp.rs:32 */

/*
 * rank(field, m, n) -- the reference's blocked right-looking elimination,
 * statement for statement: panel factorisation with first-nonzero pivoting and
 * in-place multipliers (fp.rs:142-178), the U-block triangular solve
 * (fp.rs:180-200), and the tiled trailing update with lazy u128 accumulation
 * (fp.rs:202-248).  OpenMP replaces rayon on the two parallel row loops
 * (fp.rs:166, fp.rs:222).
 *
 * Extra outputs for the tests (may be NULL):
 *   pcols_out[k]  pivot column of pivot k (k < rank);
 *   work_out      W = sum_k (n-1-k)(n-1-j_k), the dense Schur-complement MAC count.
 */
size_t or_rank(uint64_t p, uint64_t *m, size_t n, uint64_t *pcols_out, double *work_out) {
    ofield fs;
    of_init(&fs, p);
    const ofield *f = &fs;
    uint64_t *u = (uint64_t *)calloc((size_t)OR_PANEL * n + 1, sizeof(uint64_t));
    uint64_t *utile = (uint64_t *)calloc((size_t)OR_PANEL * OR_TILE, sizeof(uint64_t));
    size_t pcols[OR_PANEL];
    size_t rank = 0, col = 0;
    double work = 0.0;
    while (col < n && rank < n) {
        size_t jend = col + OR_PANEL < n ? col + OR_PANEL : n;
        size_t p0 = rank, pb = 0;
        for (size_t j = col; j < jend; j++) {
            if (rank == n) break;
            size_t pr = n;
            for (size_t r = rank; r < n; r++)
                if (m[r * n + j] != 0) { pr = r; break; }
            if (pr == n) continue;
            if (pr != rank) {
                for (size_t c = 0; c < n; c++) {
                    uint64_t t = m[rank * n + c];
                    m[rank * n + c] = m[pr * n + c];
                    m[pr * n + c] = t;
                }
            }
            uint64_t inv = of_inv(f, m[rank * n + j]);
            const uint64_t *prow = m + rank * n;
            long long lo = (long long)rank + 1, hi = (long long)n;
#pragma omp parallel for schedule(dynamic, 16)
            for (long long r = lo; r < hi; r++) {
                uint64_t *rr = m + (size_t)r * n;
                uint64_t e = rr[j];
                if (e != 0) {
                    uint64_t fm = of_mul(f, e, inv);
                    rr[j] = fm;
                    for (size_t c = j + 1; c < jend; c++) rr[c] = of_sub(f, rr[c], of_mul(f, fm, prow[c]));
                }
            }
            if (pcols_out) pcols_out[rank] = j;
            work += (double)(n - 1 - rank) * (double)(n - 1 - j);
            pcols[pb++] = j;
            rank++;
        }
        if (pb > 0 && jend < n) {
            for (size_t k = 0; k < pb; k++)
                memcpy(u + k * n + jend, m + (p0 + k) * n + jend, (n - jend) * sizeof(uint64_t));
            for (size_t k = 1; k < pb; k++)
                for (size_t k2 = 0; k2 < k; k2++) {
                    uint64_t fm = m[(p0 + k) * n + pcols[k2]];
                    if (fm == 0) continue;
                    uint64_t *uk = u + k * n;
                    const uint64_t *uk2 = u + k2 * n;
                    for (size_t c = jend; c < n; c++) uk[c] = of_sub(f, uk[c], of_mul(f, fm, uk2[c]));
                }
            uint64_t sb = f->safe_batch;
            for (size_t c0 = jend; c0 < n; c0 += OR_TILE) {
                size_t c1 = c0 + OR_TILE < n ? c0 + OR_TILE : n, w = c1 - c0;
                for (size_t k = 0; k < pb; k++)
                    for (size_t ci = 0; ci < w; ci++) utile[ci * pb + k] = u[k * n + c0 + ci];
                long long lo = (long long)rank, hi = (long long)n;
#pragma omp parallel for schedule(dynamic, 16)
                for (long long r = lo; r < hi; r++) {
                    uint64_t *rr = m + (size_t)r * n;
                    uint64_t fsv[OR_PANEL];
                    for (size_t k = 0; k < pb; k++) fsv[k] = rr[pcols[k]];
                    for (size_t ci = 0; ci < w; ci++) {
                        const uint64_t *ut = utile + ci * pb;
                        uint64_t total = 0, cnt = 0;
                        u128 acc = 0;
                        for (size_t k = 0; k < pb; k++) {
                            acc += (u128)fsv[k] * ut[k];
                            if (++cnt == sb) {
                                total = of_add(f, total, of_reduce_u128(f, acc));
                                acc = 0;
                                cnt = 0;
                            }
                        }
                        if (cnt) total = of_add(f, total, of_reduce_u128(f, acc));
                        rr[c0 + ci] = of_sub(f, rr[c0 + ci], total);
                    }
                }
            }
        }
        col = jend;
    }
    free(u);
    free(utile);
    if (work_out) *work_out = work;
    return rank;
}

/* Naive unblocked elimination of the reference's test (test_rank.rs:44-72). */
size_t or_rank_ref(uint64_t p, uint64_t *m, size_t n) {
    size_t rank = 0;
    for (size_t col = 0; col < n; col++) {
        if (rank == n) break;
        size_t pr = n;
        for (size_t r = rank; r < n; r++)
            if (m[r * n + col] != 0) { pr = r; break; }
        if (pr == n) continue;
        if (pr != rank)
            for (size_t c = col; c < n; c++) {
                uint64_t t = m[pr * n + c];
                m[pr * n + c] = m[rank * n + c];
                m[rank * n + c] = t;
            }
        /* powmod / mulmod / submod with u128 '%' as in the test helpers */
        uint64_t a = m[rank * n + col] % p, e = p - 2, inv = 1 % p;
        u128 base = a;
        while (e) {
            if (e & 1) inv = (uint64_t)(((u128)inv * base) % p);
            base = (base * base) % p;
            e >>= 1;
        }
        for (size_t r = rank + 1; r < n; r++) {
            uint64_t head = m[r * n + col];
            if (!head) continue;
            uint64_t fm = (uint64_t)(((u128)head * inv) % p);
            for (size_t c = col; c < n; c++) {
                uint64_t prod = (uint64_t)(((u128)fm * m[rank * n + c]) % p);
                uint64_t x = m[r * n + c];
                m[r * n + c] = x >= prod ? x - prod : x + p - prod;
            }
        }
        rank++;
    }
    return rank;
}

/* Rank over GF(2) by bit-packed elimination (a textbook method, independent of
 * the modular code): entries are taken mod 2. */
size_t or_rank_gf2(const uint64_t *m, size_t n) {
    size_t words = (n + 63) / 64;
    uint64_t *b = (uint64_t *)calloc(n * words + 1, sizeof(uint64_t));
    for (size_t r = 0; r < n; r++)
        for (size_t c = 0; c < n; c++)
            if (m[r * n + c] & 1) b[r * words + c / 64] |= 1ull << (c % 64);
    size_t rank = 0;
    for (size_t c = 0; c < n && rank < n; c++) {
        size_t w = c / 64;
        uint64_t bit = 1ull << (c % 64);
        size_t pr = n;
        for (size_t r = rank; r < n; r++)
            if (b[r * words + w] & bit) { pr = r; break; }
        if (pr == n) continue;
        for (size_t k = 0; k < words; k++) {
            uint64_t t = b[pr * words + k];
            b[pr * words + k] = b[rank * words + k];
            b[rank * words + k] = t;
        }
        for (size_t r = 0; r < n; r++)
            if (r != rank && (b[r * words + w] & bit))
                for (size_t k = 0; k < words; k++) b[r * words + k] ^= b[rank * words + k];
        rank++;
    }
    free(b);
    return rank;
}

/* ------------------------------------------------------------------------ */
/* Berlekamp-Massey (fp.rs:257-308)                                          */
/* ------------------------------------------------------------------------ */

/*
 * linear_complexity(field, s) -- the reference's loop, step for step.
 * stats_out (may be NULL) receives [bytes, mulmods, length_changes] where
 * bytes = sum_i [8 + 16 l_i + (d_i != 0) 24 jmax_i + (change_i) 16 (old_l_i + 1)]
 * and mulmods = sum_i (l_i + (d_i != 0) jmax_i): the algorithmic traffic and
 * work of each step.  n = 0 is rejected (-1) where the reference panics
 * (fp.rs:265).
 */
long long or_lc(uint64_t p, const uint64_t *s, size_t n, double *stats_out) {
    if (n == 0) return -1;
    ofield fs;
    of_init(&fs, p);
    const ofield *f = &fs;
    uint64_t *c = (uint64_t *)calloc(n, sizeof(uint64_t));
    uint64_t *b = (uint64_t *)calloc(n, sizeof(uint64_t));
    uint64_t *t = (uint64_t *)calloc(n, sizeof(uint64_t));
    c[0] = 1;
    b[0] = 1;
    size_t l = 0, bl = 1, m = 1;
    uint64_t bb_inv = 1;
    double bytes = 0, mulmods = 0, changes = 0;
    for (size_t i = 0; i < n; i++) {
        uint64_t d = s[i];
        for (size_t j = 1; j <= l; j++) d = of_add(f, d, of_mul(f, c[j], s[i - j]));
        bytes += 8.0 + 16.0 * (double)l;
        mulmods += (double)l;
        if (d == 0) { m++; continue; }
        uint64_t coef = of_mul(f, d, bb_inv);
        size_t jmax = bl < n - m ? bl : n - m;
        bytes += 24.0 * (double)jmax;
        mulmods += (double)jmax;
        if (2 * l <= i) {
            size_t old_l = l;
            memcpy(t, c, (old_l + 1) * sizeof(uint64_t));
            for (size_t j = 0; j < jmax; j++) c[j + m] = of_sub(f, c[j + m], of_mul(f, coef, b[j]));
            uint64_t *tmp = b; b = t; t = tmp;
            bl = old_l + 1;
            bb_inv = of_inv(f, d);
            l = i + 1 - old_l;
            m = 1;
            bytes += 16.0 * (double)(old_l + 1);
            changes += 1;
        } else {
            for (size_t j = 0; j < jmax; j++) c[j + m] = of_sub(f, c[j + m], of_mul(f, coef, b[j]));
            m++;
        }
    }
    free(c); free(b); free(t);
    if (stats_out) { stats_out[0] = bytes; stats_out[1] = mulmods; stats_out[2] = changes; }
    return (long long)l;
}

/* Naive BM of the reference's test (test_linear_complexity.rs:37-70): full-range
 * update, fresh copy of c on every nonzero discrepancy. */
long long or_lc_ref(uint64_t p, const uint64_t *s, size_t n) {
    if (n == 0) return -1;
    uint64_t *c = (uint64_t *)calloc(n, sizeof(uint64_t));
    uint64_t *b = (uint64_t *)calloc(n, sizeof(uint64_t));
    uint64_t *t = (uint64_t *)calloc(n, sizeof(uint64_t));
    c[0] = 1; b[0] = 1;
    size_t l = 0, m = 1;
    uint64_t bb = 1;
    for (size_t i = 0; i < n; i++) {
        uint64_t d = s[i];
        for (size_t j = 1; j <= l; j++)
            d = (uint64_t)(((u128)d + (u128)(((u128)c[j] * s[i - j]) % p)) % p);
        if (d == 0) { m++; continue; }
        uint64_t e = p - 2, inv = 1 % p;
        u128 base = bb % p;
        while (e) {
            if (e & 1) inv = (uint64_t)(((u128)inv * base) % p);
            base = (base * base) % p;
            e >>= 1;
        }
        uint64_t coef = (uint64_t)(((u128)d * inv) % p);
        memcpy(t, c, n * sizeof(uint64_t));
        for (size_t j = 0; j + m < n; j++) {
            uint64_t prod = (uint64_t)(((u128)coef * b[j]) % p);
            uint64_t x = c[j + m];
            c[j + m] = x >= prod ? x - prod : x + p - prod;
        }
        if (2 * l <= i) {
            l = i + 1 - l;
            uint64_t *tmp = b; b = t; t = tmp;
            bb = d;
            m = 1;
        } else {
            m++;
        }
    }
    free(c); free(b); free(t);
    return (long long)l;
}

/* ------------------------------------------------------------------------ */
/* stats.rs:14-137 (f64, glibc libm; compile with -ffp-contract=off)         */
/* ------------------------------------------------------------------------ */

#define OR_P_FLOOR 1e-307 /* stats.rs:16 */

/* Binary exponentiation with a reciprocal for negative exponents: the
 * run-time routine that Rust's f64::powi lowers to (LLVM compiler-rt
 * __powidf2).  Multiply on odd bits, square, halve. */
static double or_powi(double a, int b) {
    int recip = b < 0;
    double r = 1.0;
    for (;;) {
        if (b & 1) r *= a;
        b /= 2;
        if (b == 0) break;
        a *= a;
    }
    return recip ? 1.0 / r : r;
}

static int or_ilog2(uint64_t x) { return 63 - __builtin_clzll(x); }

static double or_ln_u64(uint64_t p) { /* stats.rs:19-23 */
    int e = or_ilog2(p);
    double mant = (double)p / (double)(1ull << e);
    return fma((double)e, 0.693147180559945309417232121458176568 /* LN_2 */, log(mant));
}

static double or_sum_ln_1m(double p, size_t a, size_t b) { /* stats.rs:26-41 */
    if (a > b) return 0.0;
    double inv = 1.0 / p;
    double pj = or_powi(inv, (int)a);
    double s = 0.0;
    for (size_t j = a; j <= b; j++) {
        if (pj < 1e-18) break;
        s += log1p(-pj);
        pj *= inv;
    }
    return s;
}

double or_corank_prob(uint64_t p, size_t n, size_t d) { /* stats.rs:52-62 */
    if (d > n) return 0.0;
    double pf = (double)p;
    double dd = (double)d;
    double lp = -or_powi(dd, 2) * or_ln_u64(p) + 2.0 * or_sum_ln_1m(pf, d + 1, n) - or_sum_ln_1m(pf, 1, n - d);
    return lp < -745.0 ? 0.0 : exp(lp);
}

static double or_clamp(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }

double or_corank_tail_pvalue(uint64_t p, size_t n, size_t c) { /* stats.rs:69-84 */
    if (c == 0) return 1.0;
    double s = 0.0;
    for (size_t d = c; d <= n; d++) {
        double pr = or_corank_prob(p, n, d);
        if (pr == 0.0) break;
        s += pr;
    }
    return or_clamp(s, OR_P_FLOOR, 1.0);
}

double or_lc_left_tail_pvalue(uint64_t p, size_t n, size_t ell) { /* stats.rs:101-126 */
    if (ell >= n) return 1.0;
    double lnp = or_ln_u64(p);
    double pr;
    if (2 * ell <= n) {
        double a = (double)(2 * ell + 1) * lnp;
        double ln_num = a > 0.0 ? a + log1p(exp(-a)) : log1p(exp(a));
        double ln_den = (double)n * lnp + log((double)p + 1.0);
        pr = exp(ln_num - ln_den);
    } else {
        double hi = exp(((double)n - 2.0 * (double)ell) * lnp);
        double lo = exp(-(double)n * lnp);
        pr = 1.0 - (hi - lo) / ((double)p + 1.0);
    }
    return or_clamp(pr, OR_P_FLOOR, 1.0);
}

/* Shortest round-trip digits of a finite positive double, in the style of
 * Rust's `{:?}` (stats.rs:129-137): decimal notation for 1e-4 <= x < 1e16
 * (integers keep a ".0"), scientific "d.ddde-N" otherwise.  The digit string is
 * the correctly rounded one of the smallest precision that parses back to x. */
int or_pretty_p_value(double x, char *out, size_t cap) {
    char buf[64];
    if (x == 0.0) { snprintf(out, cap, "0"); return 0; }
    if (x == 1.0) { snprintf(out, cap, "1"); return 0; }
    int prec;
    for (prec = 0; prec < 17; prec++) {
        snprintf(buf, sizeof buf, "%.*e", prec, x);
        if (strtod(buf, NULL) == x) break;
    }
    snprintf(buf, sizeof buf, "%.*e", prec, x);
    /* buf = [-]D[.DDD]e[+-]XX */
    char digits[32];
    int nd = 0, neg = 0;
    const char *q = buf;
    if (*q == '-') { neg = 1; q++; }
    for (; *q && *q != 'e'; q++)
        if (*q >= '0' && *q <= '9') digits[nd++] = *q;
    digits[nd] = 0;
    int exp10 = atoi(q + 1);
    while (nd > 1 && digits[nd - 1] == '0') digits[--nd] = 0;
    char res[96];
    int pos = 0;
    if (neg) res[pos++] = '-';
    double ax = fabs(x);
    if (ax >= 1e-4 && ax < 1e16) {
        if (exp10 >= 0) {
            for (int i = 0; i <= exp10; i++) res[pos++] = i < nd ? digits[i] : '0';
            res[pos++] = '.';
            if (nd > exp10 + 1)
                for (int i = exp10 + 1; i < nd; i++) res[pos++] = digits[i];
            else
                res[pos++] = '0';
        } else {
            res[pos++] = '0';
            res[pos++] = '.';
            for (int i = 0; i < -exp10 - 1; i++) res[pos++] = '0';
            for (int i = 0; i < nd; i++) res[pos++] = digits[i];
        }
        res[pos] = 0;
    } else {
        res[pos++] = digits[0];
        if (nd > 1) {
            res[pos++] = '.';
            for (int i = 1; i < nd; i++) res[pos++] = digits[i];
        }
        pos += snprintf(res + pos, sizeof res - pos, "e%d", exp10);
    }
    snprintf(out, cap, "%s", res);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Generators (prng.rs:46-343) and the published ones the workloads use      */
/* ------------------------------------------------------------------------ */

static inline uint64_t or_rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

/* SplitMix64 (Steele, Lea, Flood; public-domain reference splitmix64.c):
 * state += golden gamma, then the three-round finaliser. */
static inline uint64_t or_splitmix_next(uint64_t *x) {
    uint64_t z = (*x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
void or_splitmix64_fill(uint64_t seed, uint64_t *out, size_t count) {
    uint64_t x = seed;
    for (size_t i = 0; i < count; i++) out[i] = or_splitmix_next(&x);
}

/* xoroshiro128++ seeded by two SplitMix64 draws (prng.rs:58-83). */
void or_xoroshiro128pp_fill(uint64_t seed, uint64_t *out, size_t count) {
    uint64_t z = seed;
    uint64_t s0 = or_splitmix_next(&z), s1 = or_splitmix_next(&z);
    for (size_t i = 0; i < count; i++) {
        uint64_t r = or_rotl(s0 + s1, 17) + s0;
        s1 ^= s0;
        s0 = or_rotl(s0, 49) ^ s1 ^ (s1 << 21);
        s1 = or_rotl(s1, 28);
        out[i] = r;
    }
}

/* xoshiro256** (Blackman, Vigna; public-domain xoshiro256starstar.c), state
 * from four SplitMix64 draws of the seed. */
void or_xoshiro256ss_fill(uint64_t seed, uint64_t *out, size_t count) {
    uint64_t z = seed, s[4];
    for (int i = 0; i < 4; i++) s[i] = or_splitmix_next(&z);
    for (size_t i = 0; i < count; i++) {
        uint64_t r = or_rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t;
        s[3] = or_rotl(s[3], 45);
        out[i] = r;
    }
}

/* MIXMAX over 2^61-1 (prng.rs:92-280: the engine of CERN ROOT's MixMaxEngine).
 * Variants: N=17 (SPECIALMUL 36, SPECIAL 0, SKIP 0), N=240 (51, 487013230256099140, 0),
 * N=256 (0, -1, 2). */
#define OR_MM_P 2305843009213693951ull
#define OR_MM_BITS 61
typedef struct {
    int n, specialmul, skip;
    uint64_t special; /* UINT64_MAX encodes the C value -1 */
    uint64_t v[256];
    uint64_t sumtot;
    int counter;
} or_mixmax;

static inline uint64_t or_mm_mod(uint64_t k) { return (k & OR_MM_P) + (k >> OR_MM_BITS); }
static inline uint64_t or_mm_add(uint64_t a, uint64_t b) { return or_mm_mod(a + b); }
static inline uint64_t or_mm_fmul(uint64_t cum, uint64_t a, uint64_t b) {
    u128 s = (u128)a * b + cum;
    uint64_t lo = (uint64_t)s, hi = (uint64_t)(s >> 64);
    return or_mm_mod((lo & OR_MM_P) + hi * 8 + (lo >> OR_MM_BITS));
}

static int or_mm_seed(or_mixmax *g, int n, int specialmul, uint64_t special, int skip, uint64_t seed) {
    if (seed == 0) return -1; /* prng.rs:153-155 */
    g->n = n; g->specialmul = specialmul; g->special = special; g->skip = skip;
    uint64_t l = seed, sum = 0, ovf = 0;
    for (int i = 0; i < n; i++) {
        l *= 6364136223846793005ull;
        l = (l << 32) ^ (l >> 32);
        g->v[i] = l & OR_MM_P;
        uint64_t s = sum + g->v[i];
        ovf += s < sum;
        sum = s;
    }
    g->sumtot = or_mm_mod(or_mm_mod(sum) + (ovf << 3));
    g->counter = n;
    return 0;
}

static void or_mm_iterate(or_mixmax *g) {
    int n = g->n, sm = g->specialmul;
    uint64_t temp2 = g->v[1];
    uint64_t tv = g->sumtot;
    g->v[0] = tv;
    uint64_t sum = tv, ovf = 0, tp = 0;
    for (int i = 1; i < n; i++) {
        if (sm) {
            uint64_t tpo = ((tp << sm) & OR_MM_P) | (tp >> (OR_MM_BITS - sm));
            tp = or_mm_add(tp, g->v[i]);
            tv = or_mm_mod(tv + tp + tpo);
        } else {
            tp = or_mm_add(tp, g->v[i]);
            tv = or_mm_add(tv, tp);
        }
        g->v[i] = tv;
        uint64_t s = sum + tv;
        ovf += s < sum;
        sum = s;
    }
    if (g->special) {
        uint64_t t2 = g->special == UINT64_MAX ? OR_MM_P - temp2 : or_mm_fmul(0, g->special, temp2);
        g->v[2] = or_mm_add(g->v[2], t2);
        uint64_t s = sum + t2;
        ovf += s < sum;
        sum = s;
    }
    g->sumtot = or_mm_mod(or_mm_mod(sum) + (ovf << 3));
}

static inline uint64_t or_mm_raw(or_mixmax *g) {
    int counter = g->counter;
    if (counter >= g->n)
        for (int k = 0; k < g->skip; k++) or_mm_iterate(g);
    g->counter = counter;
    int i = g->counter;
    if (i < g->n) {
        g->counter++;
        return g->v[i];
    }
    or_mm_iterate(g);
    g->counter = 2;
    return g->v[1];
}

/* Fills out[] with next_u64() (raw << 3, prng.rs:268) of MIXMAX-`variant`
 * (17, 240 or 256) after discarding `skip_outputs` outputs.  Returns -1 for a
 * zero seed or an unknown variant. */
int or_mixmax_fill(int variant, uint64_t seed, uint64_t skip_outputs, uint64_t *out, size_t count) {
    or_mixmax g;
    int rc;
    if (variant == 17) rc = or_mm_seed(&g, 17, 36, 0, 0, seed);
    else if (variant == 240) rc = or_mm_seed(&g, 240, 51, 487013230256099140ull, 0, seed);
    else if (variant == 256) rc = or_mm_seed(&g, 256, 0, UINT64_MAX, 2, seed);
    else return -1;
    if (rc) return rc;
    for (uint64_t i = 0; i < skip_outputs; i++) (void)or_mm_raw(&g);
    for (size_t i = 0; i < count; i++) out[i] = or_mm_raw(&g) << 3;
    return 0;
}

/* The reference tests' input mixers: 2 rounds (fp.rs:318-322,
 * test_linear_complexity.rs:95-99) and 3 rounds (test_rank.rs:11-16). */
uint64_t or_mix2(uint64_t x) {
    uint64_t z = x * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    return z ^ (z >> 31);
}
uint64_t or_mix3(uint64_t x) {
    uint64_t z = x * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

int or_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
void or_set_num_threads(int t) {
#ifdef _OPENMP
    omp_set_num_threads(t);
#else
    (void)t;
#endif
}
