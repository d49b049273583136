// The GPU field arithmetic (modlin-rs_amd/csrc/fp_dev.hpp) on the CPU, against
// exact 128-bit remainders: the fused Mersenne step lin2_m61 (b x + d y), the
// fraction-free step ff_step_t in every reduction mode (and ff_step_m61_lazy on
// operands up to p + 7), mul_mod_t, and the lazy
// sums of products MacSum (long runs of add2, every mode).  Inputs are
// random canonical residues plus the corners 0, 1, p - 1, 2^31 - 1, 2^31, 2^32.
// Built and run by tests/test_limb.py; prints "bad=0" and exits 0 on success.
#include <stdint.h>
#include <stdio.h>

#include <random>

#include "fp_dev.hpp"

using namespace mdl;
typedef unsigned __int128 u128;

static std::mt19937_64 rng(777);
static int bad = 0;

static uint64_t pick(uint64_t p, int i) {
    const uint64_t corners[] = {0, 1, p - 1, p - 2, 0x7fffffffull % p, 0x80000000ull % p, 0x100000000ull % p};
    return i < 7 ? corners[i] : rng() % p;
}

template <int MODE>
static void run(uint64_t p, int n) {
    const FieldParams f = make_params(p);
    for (int it = 0; it < n; it++) {
        const uint64_t pi = pick(p, it % 23), x = pick(p, (it / 23) % 11), e = pick(p, (it / 7) % 13), y = pick(p, it % 9);
        const uint64_t want = (uint64_t)(((u128)pi * x % p + (u128)(p - e) % p * y % p) % p);
        const uint64_t got = ff_step_t<MODE>(f, pi, x, e, y);
        if (got != want) {
            if (bad < 8) printf("ff_step mismatch p=%llu\n", (unsigned long long)p);
            bad++;
        }
        const uint64_t m = mul_mod_t<MODE>(f, x, y);
        if (m != (uint64_t)((u128)x * y % p)) {
            if (bad < 8) printf("mul_mod mismatch p=%llu\n", (unsigned long long)p);
            bad++;
        }
        if (MODE == MODE_M61) {
            // lazily reduced operands x + k p, y + k' p <= p + 7 (k, k' in {0, 1} when
            // x, y <= 7): result congruent and at most p + 7
            const uint64_t xl = x <= 7 ? x + P61 : x, yl = y <= 7 ? y + P61 : y;
            const uint64_t z = ff_step_m61_lazy(pi, xl, e, yl);
            if (z > P61 + 7 || z % P61 != want) {
                if (bad < 8) printf("ff_step_m61_lazy mismatch\n");
                bad++;
            }
            const uint64_t ne = P61 - e;
            const uint64_t za = lin2_m61_lazy_acc(xl, yl, (uint32_t)pi & 0x7fffffffu, (uint32_t)(pi >> 31),
                                                  (uint32_t)ne & 0x7fffffffu, (uint32_t)(ne >> 31));
            if (za > P61 + 7 || za % P61 != want) {
                if (bad < 8) printf("lin2_m61_lazy_acc mismatch\n");
                bad++;
            }
            const uint64_t l = lin2_m61(x, y, (uint32_t)pi & 0x7fffffffu, (uint32_t)(pi >> 31), (uint32_t)e & 0x7fffffffu,
                                        (uint32_t)(e >> 31));
            if (l != (uint64_t)(((u128)pi * x + (u128)e * y) % p)) {
                if (bad < 8) printf("lin2_m61 mismatch\n");
                bad++;
            }
        }
    }
}

// MacSum over runs of up to 300 pairs, every residue drawn from the corners and at
// random (long runs of p - 1 stress the 2^61 - 1 folds and the safe_batch cut).
template <int MODE>
static void run_sum(uint64_t p, int n) {
    const FieldParams f = make_params(p);
    for (int it = 0; it < n; it++) {
        MacSum<MODE> acc;
        u128 want = 0;
        const int len = it % 7 == 0 ? 300 : (int)(rng() % 40);
        const bool worst = it % 5 == 0;
        for (int k = 0; k < len; k++) {
            const uint64_t a = worst ? p - 1 : pick(p, (int)(rng() % 12)), b = worst ? p - 1 : pick(p, (int)(rng() % 12));
            const uint64_t c = pick(p, (int)(rng() % 12)), d = worst ? p - 1 : pick(p, (int)(rng() % 12));
            acc.add2(f, a, b, c, d);
            want = (want + (u128)a * b % p + (u128)c * d % p) % p;
        }
        if (acc.value(f) != (uint64_t)want) {
            if (bad < 8) printf("MacSum mismatch p=%llu len=%d\n", (unsigned long long)p, len);
            bad++;
        }
    }
}

// The lazy 2^61 - 1 steps at the ends of their operand ranges: x, y up to p + 7,
// pi up to p - 1, e = 0 (so d = p) and e up to p - 1; every pair of corners.
static void m61_lazy_corners() {
    const uint64_t xs[] = {0, 1, 7, P61 - 1, P61, P61 + 1, P61 + 7, 0x7fffffffull, 0x80000000ull,
                           (1ull << 60) + 12345, P61 - 0x7fffffffull};
    const uint64_t cs[] = {0, 1, 2, P61 - 1, P61 - 2, 0x7fffffffull, 0x80000000ull, 1ull << 60, (1ull << 61) - 0x80000000ull};
    for (uint64_t x : xs)
        for (uint64_t y : xs)
            for (uint64_t pi : cs)
                for (uint64_t e : cs) {
                    const uint64_t want =
                        (uint64_t)(((u128)pi * (x % P61) + (u128)(P61 - e) % P61 * (y % P61)) % P61);
                    const uint64_t z = ff_step_m61_lazy(pi, x, e, y);
                    const uint64_t ne = P61 - e;
                    const uint64_t za = lin2_m61_lazy_acc(x, y, (uint32_t)pi & 0x7fffffffu, (uint32_t)(pi >> 31),
                                                          (uint32_t)ne & 0x7fffffffu, (uint32_t)(ne >> 31));
                    if (z > P61 + 7 || z % P61 != want || za > P61 + 7 || za % P61 != want) {
                        if (bad < 8) printf("lazy 2^61-1 corner mismatch\n");
                        bad++;
                    }
                }
}

int main() {
    m61_lazy_corners();
    run_sum<MODE_M61>(P61, 20000);
    run_sum<MODE_M31>(P31, 5000);
    run_sum<MODE_GENERIC>(4611686018427387847ull, 5000);
    run_sum<MODE_GENERIC>(9223372036854775783ull, 5000);
    run_sum<MODE_GENERIC>(65521ull, 5000);
    run<MODE_M61>(P61, 2000000);
    run<MODE_M31>(P31, 500000);
    const uint64_t primes[] = {2ull, 3ull, 65521ull, 4294967291ull, 1099511627689ull, 4611686018427387847ull,
                               9223372036854775783ull, P61, P31};
    for (uint64_t p : primes) run<MODE_GENERIC>(p, 200000);
    printf("bad=%d\n", bad);
    return bad != 0;
}
