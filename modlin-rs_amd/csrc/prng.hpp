// Input generators of modlin (src/prng.rs) plus the two published generators the
// benchmark workloads use (SplitMix64, xoshiro256**).  The state transitions are
// __host__ __device__: the host generators below and the device fill kernels
// (prng_dev.hip) run the same operations word for word.
#pragma once
#include <stdint.h>

#include <array>
#include <vector>

#include "fp_dev.hpp"

#if defined(__clang__)
#define MDL_PRAGMA(x) _Pragma(#x)
#else
#define MDL_PRAGMA(x)
#endif

namespace modlin {
namespace prng {

MDL_HD uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

// SplitMix64 (public-domain splitmix64.c): add the golden gamma, then finalise.
struct SplitMix64 {
    uint64_t x;
    MDL_HD explicit SplitMix64(uint64_t seed) : x(seed) {}
    MDL_HD uint64_t next_u64() {
        uint64_t z = (x += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
};

// xoroshiro128++ with a SplitMix64-expanded seed (prng.rs:48-84).
struct Xoroshiro128pp {
    static constexpr const char *NAME = "xoroshiro128++";
    uint64_t s0, s1;
    MDL_HD Xoroshiro128pp() : s0(0), s1(0) {}
    MDL_HD explicit Xoroshiro128pp(uint64_t seed) {
        SplitMix64 z(seed);
        s0 = z.next_u64();
        s1 = z.next_u64();
    }
    MDL_HD uint64_t next_u64() {
        const uint64_t a = s0;
        uint64_t b = s1;
        const uint64_t r = rotl(a + b, 17) + a;
        b ^= a;
        s0 = rotl(a, 49) ^ b ^ (b << 21);
        s1 = rotl(b, 28);
        return r;
    }
};

// xoshiro256** (public-domain xoshiro256starstar.c), state from four SplitMix64 draws.
struct Xoshiro256ss {
    uint64_t s[4];
    MDL_HD Xoshiro256ss() : s{0, 0, 0, 0} {}
    MDL_HD explicit Xoshiro256ss(uint64_t seed) {
        SplitMix64 z(seed);
        for (int i = 0; i < 4; i++) s[i] = z.next_u64();
    }
    MDL_HD uint64_t next_u64() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
};

// The MIXMAX state transition over 2^61 - 1 (ROOT's iterate_raw_vec,
// prng.rs:191-222), on the words v[0], v[S], ..., v[(N-1) S].  N is the
// dimension, SPECIALMUL/SPECIAL the per-N matrix tweak (SPECIAL = UINT64_MAX
// stands for ROOT's -1).  Values are kept in the reference's folded form, which
// is canonical except that a few residues (0, 1, 2) may be held as p, p + 1 or
// p + 2; the device fill kernels rely on exactly this property when they start
// a stream from a jumped (canonical) state.
template <int N, int SPECIALMUL, uint64_t SPECIAL>
struct MixMaxMap {
    static constexpr uint64_t P = 2305843009213693951ull;  // 2^61 - 1
    static constexpr int BITS = 61;
    MDL_HD static uint64_t fold(uint64_t k) { return (k & P) + (k >> BITS); }
    MDL_HD static uint64_t addm(uint64_t a, uint64_t b) { return fold(a + b); }
    MDL_HD static uint64_t fmul(uint64_t a, uint64_t b) {  // fmodmul_m61(0, a, b)
        uint64_t hi, lo;
        mdl::mul_wide(a, b, hi, lo);
        return fold((lo & P) + hi * 8 + (lo >> BITS));
    }
    // Applies the matrix in place and returns the new sumtot.
    template <int S>
    MDL_HD static uint64_t iterate(uint64_t *v, uint64_t sumtot_in) {
        const uint64_t temp2 = v[S];
        uint64_t tv = sumtot_in;
        v[0] = tv;
        uint64_t sum = tv, ovf = 0, tp = 0;
        [[maybe_unused]] constexpr int UN = N <= 32 ? N : 4;
        MDL_PRAGMA(unroll UN)
        for (int i = 1; i < N; i++) {
            if (SPECIALMUL != 0) {
                const uint64_t tpo = ((tp << SPECIALMUL) & P) | (tp >> (BITS - SPECIALMUL));
                tp = addm(tp, v[i * S]);
                tv = fold(tv + tp + tpo);
            } else {
                tp = addm(tp, v[i * S]);
                tv = addm(tv, tp);
            }
            v[i * S] = tv;
            const uint64_t s = sum + tv;
            ovf += s < sum;
            sum = s;
        }
        if (SPECIAL != 0) {
            const uint64_t t2 = SPECIAL == UINT64_MAX ? P - temp2 : fmul(SPECIAL, temp2);
            v[2 * S] = addm(v[2 * S], t2);
            const uint64_t s = sum + t2;
            ovf += s < sum;
            sum = s;
        }
        return fold(fold(sum) + (ovf << 3));
    }
};

// The MIXMAX generator of CERN ROOT's MixMaxEngine, as in prng.rs:92-280.  SKIP
// is the number of extra iterations before each re-iteration.
template <int N, int SPECIALMUL, uint64_t SPECIAL, int SKIP>
class MixMax {
  public:
    using Map = MixMaxMap<N, SPECIALMUL, SPECIAL>;
    static constexpr uint64_t P = Map::P;
    static constexpr int DIM = N, SPECIAL_MUL = SPECIALMUL, SKIP_ITERS = SKIP;
    static constexpr uint64_t SPECIAL_ENTRY = SPECIAL;

    // Returns false for the zero seed (prng.rs:153-155).
    bool seed(uint64_t sd) {
        if (sd == 0) return false;
        uint64_t l = sd, sum = 0, ovf = 0;
        for (int i = 0; i < N; i++) {
            l *= 6364136223846793005ull;  // Knuth's LCG multiplier (seed_spbox)
            l = (l << 32) ^ (l >> 32);
            v_[i] = l & P;
            const uint64_t s = sum + v_[i];
            ovf += s < sum;
            sum = s;
        }
        sumtot_ = Map::fold(Map::fold(sum) + (ovf << 3));
        counter_ = N;
        return true;
    }
    // Raw 61-bit output (prng.rs:243-252).
    uint64_t next_raw() {
        const int counter = counter_;
        if (counter >= N)
            for (int k = 0; k < SKIP; k++) iterate();
        counter_ = counter;
        if (counter_ < N) return v_[counter_++];
        iterate();
        counter_ = 2;
        return v_[1];
    }
    uint64_t next_u64() { return next_raw() << 3; }  // prng.rs:268
    // The seeded state, before the first draw (words as the reference holds them).
    const uint64_t *state() const { return v_.data(); }
    uint64_t sumtot() const { return sumtot_; }

  private:
    void iterate() { sumtot_ = Map::template iterate<1>(v_.data(), sumtot_); }
    std::array<uint64_t, N> v_{};
    uint64_t sumtot_ = 0;
    int counter_ = N;
};

using MixMax17 = MixMax<17, 36, 0, 0>;                        // prng.rs:303-322
using MixMax240 = MixMax<240, 51, 487013230256099140ull, 0>;  // prng.rs:282-301
using MixMax256 = MixMax<256, 0, UINT64_MAX, 2>;              // prng.rs:324-343

}  // namespace prng
}  // namespace modlin
