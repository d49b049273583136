// Exactness of the limb recombination (modlin-rs_amd/csrc/limb.hpp) on the CPU.
//
// For arbitrary int8 digit panels -- uniform digits, and the extreme panels that
// drive every shift class to its upper or lower bound -- the class accumulators
// are formed exactly as the matrix cores form them (even classes start at HOFF,
// int32 accumulation, checked for overflow), then recombined with limb_residue
// and limb_sub_m61_k2.  The result must equal the product sum_k A_k X_k mod p with
// A_k = sum_l a_kl 256^l computed directly in 128-bit integers.
//
// Built and run by tests/test_limb.py; prints "bad=0" and exits 0 on success.
#include <stdint.h>
#include <stdio.h>

#include <random>

#include "limb.hpp"

using namespace mdl;
typedef __int128 i128;
typedef unsigned __int128 u128;

static std::mt19937_64 rng(12345);
static int bad = 0;

// Digit of panel kind `kind` at (k, l): 0 uniform, 1 all -128 (classes at their
// maximum), 2 -128 against 127 (minimum), 3 random extremes, 4 alternating.
static void digits(int kind, int l, int &da, int &dx) {
    switch (kind) {
        case 0: da = (int)(rng() % 256) - 128; dx = (int)(rng() % 256) - 128; break;
        case 1: da = -128; dx = -128; break;
        case 2: da = -128; dx = 127; break;
        case 3: da = (rng() & 1) ? -128 : 127; dx = (rng() & 1) ? -128 : 127; break;
        default: da = (l & 1) ? -128 : 127; dx = (l & 1) ? 127 : -128; break;
    }
}

static i128 value_mod(const int8_t *d, int L, uint64_t p) {
    i128 v = 0, w = 1;
    for (int l = 0; l < L; l++) {
        v += (i128)d[l] * w;
        w *= 256;
    }
    v %= (i128)p;
    return v < 0 ? v + (i128)p : v;
}

template <int MODE, int L>
static void run(uint64_t p, int trials, int kind) {
    constexpr int NC = 2 * L - 1;
    const FieldParams f = make_params(p);
    const LimbParams lp = make_limb_params(p);
    for (int tr = 0; tr < trials; tr++) {
        const int K = kind == 0 ? 1 + (int)(rng() % 64) : 64;
        int8_t a[64][8], x[64][8];
        for (int k = 0; k < K; k++)
            for (int l = 0; l < L; l++) {
                int da, dx;
                digits(kind, l, da, dx);
                a[k][l] = (int8_t)da;
                x[k][l] = (int8_t)dx;
            }
        int32_t acc[NC];
        for (int s = 0; s < NC; s++) acc[s] = (s & 1) ? 0 : (int32_t)HOFF;
        for (int k = 0; k < K; k++)
            for (int la = 0; la < L; la++)
                for (int lb = 0; lb < L; lb++) {
                    const int64_t v = (int64_t)acc[la + lb] + (int64_t)a[k][la] * x[k][lb];
                    if (v > INT32_MAX || v < INT32_MIN) {
                        printf("int32 overflow p=%llu L=%d kind=%d\n", (unsigned long long)p, L, kind);
                        bad++;
                        return;
                    }
                    acc[la + lb] = (int32_t)v;
                }
        i128 want = 0;
        for (int k = 0; k < K; k++)
            want = (want + (i128)(((u128)value_mod(a[k], L, p) * (u128)value_mod(x[k], L, p)) % p)) % (i128)p;
        uint64_t got = limb_residue<MODE, NC>(f, lp, acc);
        got = got >= lp.k0 ? got - lp.k0 : got + p - lp.k0;
        if ((uint64_t)want != got) {
            if (bad < 8)
                printf("limb_residue mismatch p=%llu L=%d kind=%d K=%d\n", (unsigned long long)p, L, kind, K);
            bad++;
        }
    }
}

// Two accumulated panels (K up to 128): classes started at KOFF2, recombined with
// limb_residue_k2 (and limb_sub_m61_k2 for 2^61 - 1).
template <int MODE, int L>
static void run_k2(uint64_t p, int trials, int kind) {
    constexpr int NC = 2 * L - 1;
    const FieldParams f = make_params(p);
    const LimbParams lp = make_limb_params(p);
    for (int tr = 0; tr < trials; tr++) {
        const int K = kind == 0 ? 1 + (int)(rng() % 128) : 128;
        static int8_t a[128][8], x[128][8];
        for (int k = 0; k < K; k++)
            for (int l = 0; l < L; l++) {
                int da, dx;
                digits(kind, l, da, dx);
                a[k][l] = (int8_t)da;
                x[k][l] = (int8_t)dx;
            }
        int32_t acc[NC];
        for (int s = 0; s < NC; s++) acc[s] = KOFF2;
        for (int k = 0; k < K; k++)
            for (int la = 0; la < L; la++)
                for (int lb = 0; lb < L; lb++) {
                    const int64_t v = (int64_t)acc[la + lb] + (int64_t)a[k][la] * x[k][lb];
                    if (v >= (1 << 26) || v < 0) {
                        printf("class outside [0, 2^26) p=%llu L=%d kind=%d\n", (unsigned long long)p, L, kind);
                        bad++;
                        return;
                    }
                    acc[la + lb] = (int32_t)v;
                }
        i128 want = 0;
        for (int k = 0; k < K; k++)
            want = (want + (i128)(((u128)value_mod(a[k], L, p) * (u128)value_mod(x[k], L, p)) % p)) % (i128)p;
        uint64_t got = limb_residue_k2<MODE, NC>(f, lp, acc);
        got = got >= lp.k2 ? got - lp.k2 : got + p - lp.k2;
        if ((uint64_t)want != got) {
            if (bad < 8)
                printf("limb_residue_k2 mismatch p=%llu L=%d kind=%d K=%d\n", (unsigned long long)p, L, kind, K);
            bad++;
        }
        if (MODE == MODE_M61 && L == 8) {
            for (uint64_t cv : {(uint64_t)(rng() % p), (uint64_t)0, p - 1}) {
                const uint64_t sub = limb_sub_m61_k2(cv, lp.kc2, acc);
                const uint64_t exp = cv >= (uint64_t)want ? cv - (uint64_t)want : cv + p - (uint64_t)want;
                if (sub != exp) {
                    if (bad < 8) printf("limb_sub_m61_k2 mismatch kind=%d K=%d\n", kind, K);
                    bad++;
                }
            }
        }
    }
}

// limb_sub_m61_k2 over its whole input domain: every class value in [0, 2^26)
// (uniform, all at the maximum 2^26 - 1, all zero, and each class alone at its
// maximum), against (cv - sum_s 2^(8 s) P_s + k2) mod p in 128-bit integers.
static void m61_k2_domain(int trials) {
    const LimbParams lp = make_limb_params(P61);
    auto check = [&](const int32_t *acc, uint64_t cv) {
        u128 s = 0;
        for (int c = 0; c < 15; c++) s = (s + (u128)(uint32_t)acc[c] * lp.pow256[c]) % P61;
        const uint64_t sv = (uint64_t)((s + P61 - lp.k2) % P61);  // S' - k2
        const uint64_t exp = cv >= sv ? cv - sv : cv + P61 - sv;
        const uint64_t got = limb_sub_m61_k2(cv, lp.kc2, acc);
        if (got != exp) {
            if (bad < 8) printf("limb_sub_m61_k2 domain mismatch cv=%llu\n", (unsigned long long)cv);
            bad++;
        }
    };
    int32_t acc[15];
    for (int tr = 0; tr < trials; tr++) {
        for (int c = 0; c < 15; c++) acc[c] = (int32_t)(rng() & ((1u << 26) - 1));
        check(acc, rng() % P61);
    }
    for (uint64_t cv : {(uint64_t)0, (uint64_t)1, P61 - 1, lp.k2, (uint64_t)(rng() % P61)}) {
        for (int c = 0; c < 15; c++) acc[c] = (1 << 26) - 1;
        check(acc, cv);
        for (int c = 0; c < 15; c++) acc[c] = 0;
        check(acc, cv);
        for (int one = 0; one < 15; one++) {
            for (int c = 0; c < 15; c++) acc[c] = c == one ? (1 << 26) - 1 : 0;
            check(acc, cv);
        }
    }
}

// The addition-chain inverse of 2^61 - 1 against the binary extended Euclid,
// and a * inv(a) == 1, on random and edge values.
static void m61_inverse(int trials) {
    auto check = [&](uint64_t a) {
        const uint64_t r = inv_m61(a);
        if (r != inv_mod_bin(a, P61) || (uint64_t)(((u128)a * r) % P61) != 1) {
            if (bad < 8) printf("inv_m61 mismatch a=%llu\n", (unsigned long long)a);
            bad++;
        }
    };
    const uint64_t edge[] = {1, 2, 3, P61 - 1, P61 - 2, (P61 + 1) / 2, (uint64_t)1 << 60, ((uint64_t)1 << 31) - 1,
                             (uint64_t)1 << 31, (uint64_t)0x7fffffff << 30};
    for (uint64_t a : edge) check(a);
    for (int i = 0; i < trials; i++) check(1 + rng() % (P61 - 1));
}

template <int L>
static void generic(uint64_t p, int trials) {
    for (int kind = 0; kind < 5; kind++) {
        run<MODE_GENERIC, L>(p, kind ? 20 : trials, kind);
        run_k2<MODE_GENERIC, L>(p, kind ? 20 : trials / 4, kind);
    }
}

int main() {
    const int n = 20000;
    for (int kind = 0; kind < 5; kind++) {
        run<MODE_M61, 8>(P61, kind ? 50 : 10 * n, kind);
        run<MODE_M31, 4>(P31, kind ? 50 : 10 * n, kind);
        run_k2<MODE_M61, 8>(P61, kind ? 50 : 2 * n, kind);
        run_k2<MODE_M31, 4>(P31, kind ? 50 : 2 * n, kind);
    }
    m61_k2_domain(20 * n);
    m61_inverse(n);
    const uint64_t primes[] = {2ull, 3ull, 5ull, 7ull, 127ull, 131ull, 251ull, 257ull, 65521ull, 65537ull,
                               16777213ull, 4294967291ull, 1099511627689ull, 281474976710597ull,
                               72057594037927931ull, P31, P61, 4611686018427387847ull, 9223372036854775783ull};
    for (uint64_t p : primes) {
        switch (limb_count(p)) {
            case 1: generic<1>(p, n); break;
            case 2: generic<2>(p, n); break;
            case 3: generic<3>(p, n); break;
            case 4: generic<4>(p, n); break;
            case 5: generic<5>(p, n); break;
            case 6: generic<6>(p, n); break;
            case 7: generic<7>(p, n); break;
            default: generic<8>(p, n); break;
        }
    }
    printf("bad=%d\n", bad);
    return bad != 0;
}
