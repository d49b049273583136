// Device-side generation of the modlin input streams (prng.rs:46-343) with exact
// jump-ahead: the matrices and sequences of main.rs:94-105 and main.rs:153-155
// (entries next_u64() mod p, row-major, successive outputs) are written straight
// into HBM, bit-identical to the host generators of prng.hpp.
//
//  * SplitMix64: output r is the finaliser of seed + (r + 1) gamma; one thread
//    per output, coalesced.
//  * xoroshiro128++ / xoshiro256**: the state transition T is linear over GF(2).
//    Lane j starts from T^(j per) s, s the state at the first wanted output,
//    applied as a product of the tabulated T^(2^b) (bit matrices in column form,
//    read uniformly by all lanes), then steps the generator sequentially.
//  * MIXMAX-17/240/256: the transition is a linear map A over F_q, q = 2^61 - 1.
//    Lane j starts at block kb0 + j bpl (a block is the SKIP + 1 iterations and
//    N - 1 outputs of one re-iteration, prng.rs:243-252).  Its state is M^j v
//    with M = A^(bpl (SKIP + 1)), produced for all lanes by doubling
//    (V[:, k..2k) = M^k V[:, 0..k)), and v = A^(kb0 (SKIP + 1)) v_seed from the
//    tabulated A^(2^b).  The reference keeps its words folded, in [0, q + 2],
//    and folding preserves residues, so a word differs from its canonical
//    residue only when that residue is 0, 1 or 2 (held as q, q + 1, q + 2), and
//    sumtot only when it is 0.  A jumped canonical state with no such word and a
//    nonzero sum therefore equals the reference's state bit for bit.  A lane
//    whose start state is not of that kind starts from an earlier lane's state
//    (the host keeps the base state clean, or exact: the seeded state itself),
//    and every lane then runs the reference's own transition
//    (MixMaxMap::iterate), so every output word equals the host generator's.
#include <stdlib.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "engine.hpp"
#include "prng.hpp"

namespace mdl {
namespace {

using namespace modlin::prng;

// Where the outputs go: output indices [skip, skip + count) of the (optionally
// decimated) stream, covered by raw indices [r_lo, r_hi).
struct FillOut {
    FieldParams f;
    void *out;
    uint64_t skip, count;
    uint64_t r_lo, r_hi;
    uint32_t keep, period;  // keep outputs, then drop period - keep; keep = 0: no decimation
    uint32_t u32, raw;      // store as u32; store next_u64() unreduced
};

__device__ __forceinline__ void store_out(const FillOut &o, uint64_t i, uint64_t x) {
    const uint64_t y = o.raw ? x : reduce128(o.f, 0, x);  // Field::reduce (main.rs:96, 154)
    if (o.u32)
        ((uint32_t *)o.out)[i] = (uint32_t)y;
    else
        ((uint64_t *)o.out)[i] = y;
}

// One lane's walk over consecutive raw outputs, with the keep/drop phase.
struct Emit {
    uint64_t r, q;
    uint32_t m;
    __device__ void init(const FillOut &o, uint64_t r0) {
        r = r0;
        q = o.keep ? r0 / o.period : 0;
        m = o.keep ? (uint32_t)(r0 % o.period) : 0;
    }
    __device__ __forceinline__ void put(const FillOut &o, uint64_t x) {
        bool kept = true;
        uint64_t e = r;
        if (o.keep) {
            kept = m < o.keep;
            e = q * o.keep + m;
            if (++m == o.period) {
                m = 0;
                ++q;
            }
        }
        if (kept && r >= o.r_lo && r < o.r_hi) store_out(o, e - o.skip, x);
        ++r;
    }
};

// ---------------------------------------------------------------------------
// SplitMix64: the state after r + 1 draws is seed + (r + 1) gamma.
__global__ __launch_bounds__(256) void k_splitmix_fill(uint64_t seed, FillOut o) {
    for (uint64_t t = blockIdx.x * 256ull + threadIdx.x; t < o.count; t += (uint64_t)gridDim.x * 256) {
        const uint64_t e = o.skip + t;
        const uint64_t r = o.keep ? (e / o.keep) * o.period + e % o.keep : e;
        SplitMix64 g(seed + r * 0x9e3779b97f4a7c15ull);
        store_out(o, t, g.next_u64());
    }
}

// ---------------------------------------------------------------------------
// GF(2)-linear generators.  tab[b] holds the 64 SW columns of T^(2^b), each SW
// words: column c is the image of state bit c (word c / 64, bit c % 64).
template <int SW>
struct Gf2Gen;
template <>
struct Gf2Gen<2> {
    Xoroshiro128pp g;
    __device__ void set(const uint64_t *s) {
        g.s0 = s[0];
        g.s1 = s[1];
    }
    __device__ uint64_t next() { return g.next_u64(); }
};
template <>
struct Gf2Gen<4> {
    Xoshiro256ss g;
    __device__ void set(const uint64_t *s) {
        for (int w = 0; w < 4; w++) g.s[w] = s[w];
    }
    __device__ uint64_t next() { return g.next_u64(); }
};

struct Gf2Args {
    uint64_t base[4];  // state at raw index r_lo
    uint64_t per, lanes;
    int nbits;  // bit length of (lanes - 1) per
};

template <int SW>
__global__ __launch_bounds__(256) void k_gf2_fill(const uint64_t *__restrict__ tab, Gf2Args a, FillOut o) {
    const uint64_t j = blockIdx.x * 256ull + threadIdx.x;
    if (j >= a.lanes) return;  // no barriers below
    uint64_t s[SW];
    for (int w = 0; w < SW; w++) s[w] = a.base[w];
    const uint64_t off = j * a.per;
    for (int b = 0; b < a.nbits; b++) {
        const uint64_t *T = tab + (uint64_t)b * 64 * SW * SW;
        uint64_t y[SW] = {};
        for (int c = 0; c < 64 * SW; c++) {
            const uint64_t mask = 0ull - ((s[c >> 6] >> (c & 63)) & 1ull);
#pragma unroll
            for (int w = 0; w < SW; w++) y[w] ^= T[c * SW + w] & mask;
        }
        if ((off >> b) & 1)
            for (int w = 0; w < SW; w++) s[w] = y[w];
    }
    Gf2Gen<SW> g;
    g.set(s);
    const uint64_t r0 = o.r_lo + off;
    const uint64_t n = std::min<uint64_t>(a.per, o.r_hi - r0);
    Emit em;
    em.init(o, r0);
    for (uint64_t k = 0; k < n; k++) em.put(o, g.next());
}

// ---------------------------------------------------------------------------
// Arithmetic mod q = 2^61 - 1 on canonical operands for the MIXMAX jumps.

// C = A B mod q; A is m x k, B is k x n, C is m x n, row-major with leading
// dimensions; C must not overlap A or B.  16 x 16 output tile per workgroup.
__global__ __launch_bounds__(256) void k_m61_gemm(const uint64_t *__restrict__ A, uint64_t lda,
                                                  const uint64_t *__restrict__ B, uint64_t ldb, uint64_t *__restrict__ C,
                                                  uint64_t ldc, int m, int n, int k) {
    __shared__ uint64_t As[16][17], Bs[16][17];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int row = blockIdx.y * 16 + ty, col = blockIdx.x * 16 + tx;
    uint64_t acc = 0;
    for (int k0 = 0; k0 < k; k0 += 16) {
        As[ty][tx] = (row < m && k0 + tx < k) ? A[(uint64_t)row * lda + k0 + tx] : 0;
        Bs[ty][tx] = (k0 + ty < k && col < n) ? B[(uint64_t)(k0 + ty) * ldb + col] : 0;
        __syncthreads();
#pragma unroll
        for (int t = 0; t < 16; t++) {
            const uint64_t b = Bs[t][tx];
            const uint64_t pr = fold_m61_64(mul_m61_lazy(As[ty][t], (uint32_t)b & 0x7fffffffu, (uint32_t)(b >> 31)));
            acc = fold_m61_64(acc + pr);  // <= q + 7 throughout
        }
        __syncthreads();
    }
    if (row < m && col < n) C[(uint64_t)row * ldc + col] = canon_m61(acc);
}

// v <- A^e v with A^(2^b) = tab[b] (n x n row-major), n <= 256; one workgroup.
// Writes the result to column 0 of V (leading dimension ldv) and to vout.
__global__ __launch_bounds__(256) void k_m61_powvec(const uint64_t *__restrict__ tab, int n, uint64_t e,
                                                    const uint64_t *__restrict__ vin, uint64_t *__restrict__ V,
                                                    uint64_t ldv, uint64_t *__restrict__ vout) {
    __shared__ uint64_t v[256];
    const int i = threadIdx.x;
    if (i < n) v[i] = vin[i];
    __syncthreads();
    for (int b = 0; b < 64; b++) {
        if (!((e >> b) & 1)) continue;  // uniform
        const uint64_t *T = tab + (uint64_t)b * n * n + (uint64_t)i * n;
        uint64_t acc = 0;
        if (i < n)
            for (int c = 0; c < n; c++) {
                const uint64_t x = v[c];
                acc = fold_m61_64(acc + fold_m61_64(mul_m61_lazy(T[c], (uint32_t)x & 0x7fffffffu, (uint32_t)(x >> 31))));
            }
        __syncthreads();
        if (i < n) v[i] = canon_m61(acc);
        __syncthreads();
    }
    if (i < n) {
        V[(uint64_t)i * ldv] = v[i];
        vout[i] = v[i];
    }
}

// ---------------------------------------------------------------------------
// MIXMAX fill.  Lane j covers blocks [kb0 + j bpl, min(kb0 + (j+1) bpl, kb_hi)).
struct MixArgs {
    const uint64_t *V;      // N x lanes start states (canonical), leading dimension ldv
    const uint64_t *exact;  // seeded state words and sumtot (N + 1), used when base_exact
    uint64_t ldv, kb0, bpl, kb_hi, lanes;
    uint64_t bound;  // a start word below bound makes the state unusable (3 in production)
    uint32_t base_exact, pad;
};

template <int N>
__device__ bool mix_clean(const MixArgs &a, uint64_t col) {
    uint64_t sum = 0;
    for (int i = 0; i < N; i++) {
        const uint64_t w = a.V[(uint64_t)i * a.ldv + col];
        if (w < a.bound) return false;
        sum = add_mod(sum, w, P61);
    }
    return sum != 0;
}

template <int N, int SM, uint64_t SP, int SKIP, int S>
__device__ void mix_lane(uint64_t *v, const MixArgs &a, const FillOut &o, uint64_t j) {
    using Map = MixMaxMap<N, SM, SP>;
    constexpr int UN = S == 1 ? N : 4;
    uint64_t s = j;
    while (s > 0 && !mix_clean<N>(a, s)) --s;
    uint64_t sumtot = 0;
    if (s == 0 && a.base_exact) {
        MDL_PRAGMA(unroll UN)
        for (int i = 0; i < N; i++) v[i * S] = a.exact[i];
        sumtot = a.exact[N];
    } else {
        MDL_PRAGMA(unroll UN)
        for (int i = 0; i < N; i++) {
            v[i * S] = a.V[(uint64_t)i * a.ldv + s];
            sumtot = add_mod(sumtot, v[i * S], P61);
        }
    }
    const uint64_t k_begin = a.kb0 + j * a.bpl;
    const uint64_t k_end = std::min<uint64_t>(k_begin + a.bpl, a.kb_hi);
    Emit em;
    em.init(o, k_begin * (N - 1));
    for (uint64_t k = a.kb0 + s * a.bpl; k < k_end; ++k) {
        for (int t = 0; t <= SKIP; t++) sumtot = Map::template iterate<S>(v, sumtot);
        if (k < k_begin) continue;
        MDL_PRAGMA(unroll UN)
        for (int i = 1; i < N; i++) em.put(o, v[i * S] << 3);  // next_u64 (prng.rs:268)
    }
}

template <int N, int SM, uint64_t SP, int SKIP>
__global__ __launch_bounds__(64) void k_mixmax_fill(MixArgs a, FillOut o) {
    const uint64_t j = blockIdx.x * 64ull + threadIdx.x;
    if constexpr (N <= 32) {
        if (j >= a.lanes) return;
        uint64_t v[N];
        mix_lane<N, SM, SP, SKIP, 1>(v, a, o, j);
    } else {
        // one column of N words per lane, word i of lane l at [i][l]
        __shared__ uint64_t lds[N * 64];
        if (j >= a.lanes) return;  // lanes never synchronise
        mix_lane<N, SM, SP, SKIP, 64>(lds + threadIdx.x, a, o, j);
    }
}

// ---------------------------------------------------------------------------
// Host side.

template <class G>
void gf2_step(uint64_t *s);
template <>
void gf2_step<Xoroshiro128pp>(uint64_t *s) {
    Xoroshiro128pp g;
    g.s0 = s[0];
    g.s1 = s[1];
    (void)g.next_u64();
    s[0] = g.s0;
    s[1] = g.s1;
}
template <>
void gf2_step<Xoshiro256ss>(uint64_t *s) {
    Xoshiro256ss g;
    for (int w = 0; w < 4; w++) g.s[w] = s[w];
    (void)g.next_u64();
    for (int w = 0; w < 4; w++) s[w] = g.s[w];
}

void gf2_apply(const uint64_t *cols, int sw, const uint64_t *x, uint64_t *y) {
    uint64_t t[4] = {0, 0, 0, 0};
    for (int c = 0; c < 64 * sw; c++)
        if ((x[c >> 6] >> (c & 63)) & 1)
            for (int w = 0; w < sw; w++) t[w] ^= cols[c * sw + w];
    for (int w = 0; w < sw; w++) y[w] = t[w];
}

// T^(2^b) for b < 64, column form; computed once per process.
template <class G, int SW>
const std::vector<uint64_t> &gf2_table() {
    static std::vector<uint64_t> tab;
    static std::once_flag once;
    std::call_once(once, [] {
        const size_t mat = (size_t)64 * SW * SW;
        tab.assign(64 * mat, 0);
        for (int c = 0; c < 64 * SW; c++) {
            uint64_t *col = &tab[(size_t)c * SW];
            col[c >> 6] = 1ull << (c & 63);
            gf2_step<G>(col);
        }
        for (int b = 1; b < 64; b++)
            for (int c = 0; c < 64 * SW; c++)
                gf2_apply(&tab[(b - 1) * mat], SW, &tab[(b - 1) * mat + (size_t)c * SW], &tab[b * mat + (size_t)c * SW]);
    });
    return tab;
}

// The MIXMAX matrix: column c is the canonical image of unit vector c under the
// map of MixMaxMap::iterate (v'_0 = sum v, tp_i = tp_{i-1} + v_i,
// v'_i = v'_{i-1} + tp_i [+ 2^SM tp_{i-1}], v'_2 += SPECIAL v_1).
template <class MM>
std::vector<uint64_t> mixmax_matrix() {
    constexpr int N = MM::DIM, SM = MM::SPECIAL_MUL;
    constexpr uint64_t SP = MM::SPECIAL_ENTRY;
    auto mulq = [](uint64_t a, uint64_t b) { return mul_mod_t<MODE_M61>(FieldParams{}, a, b); };
    std::vector<uint64_t> A((size_t)N * N, 0);
    for (int c = 0; c < N; c++) {
        std::vector<uint64_t> v(N, 0), w(N, 0);
        v[c] = 1;
        uint64_t tv = 0;
        for (int i = 0; i < N; i++) tv = add_mod(tv, v[i], P61);
        w[0] = tv;
        uint64_t tp = 0;
        for (int i = 1; i < N; i++) {
            const uint64_t tpo = SM ? mulq(tp, 1ull << SM) : 0;
            tp = add_mod(tp, v[i], P61);
            tv = add_mod(add_mod(tv, tp, P61), tpo, P61);
            w[i] = tv;
        }
        if (SP != 0) {
            const uint64_t t2 = SP == UINT64_MAX ? (P61 - v[1]) % P61 : mulq(SP % P61, v[1]);
            w[2] = add_mod(w[2], t2, P61);
        }
        for (int i = 0; i < N; i++) A[(size_t)i * N + c] = w[i];
    }
    return A;
}

int ensure_ws(modlin_ctx *ctx, size_t bytes) {
    if (ctx->gen_ws_bytes >= bytes) return MODLIN_OK;
    if (ctx->gen_ws) {
        MDL_HIP(hipStreamSynchronize(ctx->stream));
        MDL_HIP(hipFree(ctx->gen_ws));
        ctx->gen_ws = nullptr;
        ctx->gen_ws_bytes = 0;
    }
    if (hipMalloc(&ctx->gen_ws, bytes) != hipSuccess) {
        (void)hipGetLastError();
        ctx->gen_ws = nullptr;
        ctx->err = "device allocation of the generator workspace failed";
        return MODLIN_ENOMEM;
    }
    ctx->gen_ws_bytes = bytes;
    return MODLIN_OK;
}

template <class G, int SW>
int fill_gf2(modlin_ctx *ctx, uint32_t kind, const G &g0, FillOut o) {
    const auto &tab = gf2_table<G, SW>();
    if (!ctx->gen_tab[kind]) {
        const size_t bytes = tab.size() * 8;
        if (hipMalloc(&ctx->gen_tab[kind], bytes) != hipSuccess) {
            (void)hipGetLastError();
            ctx->gen_tab[kind] = nullptr;
            ctx->err = "device allocation of the generator table failed";
            return MODLIN_ENOMEM;
        }
        MDL_HIP(hipMemcpy(ctx->gen_tab[kind], tab.data(), bytes, hipMemcpyHostToDevice));
    }
    Gf2Args a{};
    uint64_t s[4] = {0, 0, 0, 0};
    if constexpr (SW == 2) {
        s[0] = g0.s0;
        s[1] = g0.s1;
    } else {
        for (int w = 0; w < 4; w++) s[w] = g0.s[w];
    }
    const size_t mat = (size_t)64 * SW * SW;
    for (int b = 0; b < 64; b++)
        if ((o.r_lo >> b) & 1) gf2_apply(&tab[b * mat], SW, s, s);
    for (int w = 0; w < 4; w++) a.base[w] = s[w];
    const uint64_t R = o.r_hi - o.r_lo;
    a.per = std::max<uint64_t>(256, (R + 65535) / 65536);
    a.lanes = (R + a.per - 1) / a.per;
    a.nbits = 0;
    while (a.nbits < 64 && ((a.lanes - 1) * a.per) >> a.nbits) a.nbits++;
    const unsigned grid = (unsigned)((a.lanes + 255) / 256);
    hipLaunchKernelGGL(k_gf2_fill<SW>, dim3(grid), dim3(256), 0, ctx->stream, (const uint64_t *)ctx->gen_tab[kind], a, o);
    MDL_HIP(hipGetLastError());
    return MODLIN_OK;
}

bool clean_host(const uint64_t *w, int n, uint64_t bound) {
    uint64_t sum = 0;
    for (int i = 0; i < n; i++) {
        if (w[i] < bound) return false;
        sum = add_mod(sum, w[i], P61);
    }
    return sum != 0;
}

void launch_gemm(modlin_ctx *ctx, const uint64_t *A, uint64_t lda, const uint64_t *B, uint64_t ldb, uint64_t *C,
                 uint64_t ldc, int m, int n, int k) {
    dim3 grid((unsigned)((n + 15) / 16), (unsigned)((m + 15) / 16));
    hipLaunchKernelGGL(k_m61_gemm, grid, dim3(256), 0, ctx->stream, A, lda, B, ldb, C, ldc, m, n, k);
}

template <class MM>
int fill_mixmax(modlin_ctx *ctx, uint32_t kind, const MM &g0, FillOut o) {
    constexpr int N = MM::DIM, SKIP = MM::SKIP_ITERS;
    const size_t NN = (size_t)N * N;
    // tabulated A^(2^b), b < 64, squared on the device once per context
    if (!ctx->gen_tab[kind]) {
        if (hipMalloc(&ctx->gen_tab[kind], 64 * NN * 8) != hipSuccess) {
            (void)hipGetLastError();
            ctx->gen_tab[kind] = nullptr;
            ctx->err = "device allocation of the generator table failed";
            return MODLIN_ENOMEM;
        }
        uint64_t *tab = (uint64_t *)ctx->gen_tab[kind];
        const std::vector<uint64_t> A = mixmax_matrix<MM>();
        MDL_HIP(hipMemcpyAsync(tab, A.data(), NN * 8, hipMemcpyHostToDevice, ctx->stream));
        for (int b = 1; b < 64; b++) launch_gemm(ctx, tab + (b - 1) * NN, N, tab + (b - 1) * NN, N, tab + b * NN, N, N, N, N);
        MDL_HIP(hipGetLastError());
        MDL_HIP(hipStreamSynchronize(ctx->stream));
    }
    const uint64_t *tab = (const uint64_t *)ctx->gen_tab[kind];
    const uint64_t lanes_max = N <= 32 ? 65536 : 64ull * (uint64_t)ctx->cu_count;
    // workspace: V (N x lanes_max), two N x N matrices, the base vector and the exact seed state
    const size_t ws_words = (size_t)N * lanes_max + 2 * NN + 2 * (N + 1);
    int rc = ensure_ws(ctx, ws_words * 8);
    if (rc != MODLIN_OK) return rc;
    uint64_t *V = (uint64_t *)ctx->gen_ws;
    uint64_t *M0 = V + (size_t)N * lanes_max, *M1 = M0 + NN;
    uint64_t *vbase = M1 + NN, *exact = vbase + (N + 1);

    // Start words below this bound disqualify a jumped state.  The residues 0, 1
    // and 2 are the ones the reference may hold in another form; a larger bound
    // (MODLIN_PRNG_CLEAN_BOUND, for tests) only makes lanes start further back.
    uint64_t bound = 3;
    if (const char *env = getenv("MODLIN_PRNG_CLEAN_BOUND")) bound = std::max<uint64_t>(3, strtoull(env, nullptr, 0));
    const uint64_t kb_lo = o.r_lo / (N - 1), kb_hi = (o.r_hi - 1) / (N - 1) + 1;
    // seeded state, exactly as the reference holds it, and canonically
    std::vector<uint64_t> ex(N + 1), canon(N);
    for (int i = 0; i < N; i++) {
        ex[i] = g0.state()[i];
        canon[i] = ex[i] % P61;
    }
    ex[N] = g0.sumtot();
    MDL_HIP(hipMemcpyAsync(exact, ex.data(), (N + 1) * 8, hipMemcpyHostToDevice, ctx->stream));
    // base block: the latest block <= kb_lo whose jumped state is clean (or block 0)
    uint64_t kb0 = kb_lo;
    uint32_t base_exact = 0;
    const uint64_t ldv = lanes_max;
    for (;;) {
        if (kb0 == 0) {
            base_exact = 1;
            MDL_HIP(hipMemcpy2DAsync(V, ldv * 8, canon.data(), 8, 8, N, hipMemcpyHostToDevice, ctx->stream));
            MDL_HIP(hipStreamSynchronize(ctx->stream));
            break;
        }
        MDL_HIP(hipMemcpyAsync(vbase, canon.data(), N * 8, hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(k_m61_powvec, dim3(1), dim3(256), 0, ctx->stream, tab, N, kb0 * (SKIP + 1),
                           (const uint64_t *)vbase, V, ldv, vbase);
        MDL_HIP(hipGetLastError());
        std::vector<uint64_t> b(N);
        MDL_HIP(hipMemcpyAsync(b.data(), vbase, N * 8, hipMemcpyDeviceToHost, ctx->stream));
        MDL_HIP(hipStreamSynchronize(ctx->stream));
        if (clean_host(b.data(), N, bound)) break;
        if (kb_lo - kb0 >= 64) {
            ctx->err = "no usable MIXMAX jump state within 64 blocks";
            return MODLIN_EUNSUPPORTED;
        }
        --kb0;
    }
    const uint64_t blocks = kb_hi - kb0;
    const uint64_t bpl = (blocks + lanes_max - 1) / lanes_max;
    const uint64_t lanes = (blocks + bpl - 1) / bpl;
    if (lanes > 1) {
        // M = A^(bpl (SKIP + 1)) from the table, then doubling over the lanes
        const uint64_t L = bpl * (SKIP + 1);
        uint64_t *Mk = M0, *Mt = M1;
        bool have = false;
        for (int b = 0; b < 64; b++) {
            if (!((L >> b) & 1)) continue;
            if (!have) {
                MDL_HIP(hipMemcpyAsync(Mk, tab + b * NN, NN * 8, hipMemcpyDeviceToDevice, ctx->stream));
                have = true;
            } else {
                launch_gemm(ctx, Mk, N, tab + b * NN, N, Mt, N, N, N, N);
                std::swap(Mk, Mt);
            }
        }
        for (uint64_t k = 1; k < lanes; k *= 2) {
            const uint64_t w = std::min<uint64_t>(k, lanes - k);
            launch_gemm(ctx, Mk, N, V, ldv, V + k, ldv, N, (int)w, N);
            if (2 * k < lanes) {
                launch_gemm(ctx, Mk, N, Mk, N, Mt, N, N, N, N);
                std::swap(Mk, Mt);
            }
        }
        MDL_HIP(hipGetLastError());
    }
    MixArgs a{};
    a.V = V;
    a.exact = exact;
    a.ldv = ldv;
    a.kb0 = kb0;
    a.bpl = bpl;
    a.kb_hi = kb_hi;
    a.lanes = lanes;
    a.bound = bound;
    a.base_exact = base_exact;
    hipLaunchKernelGGL((k_mixmax_fill<N, MM::SPECIAL_MUL, MM::SPECIAL_ENTRY, SKIP>), dim3((unsigned)((lanes + 63) / 64)),
                       dim3(64), 0, ctx->stream, a, o);
    MDL_HIP(hipGetLastError());
    return MODLIN_OK;
}

}  // namespace

int prng_fill_device(modlin_ctx *ctx, uint32_t kind, uint64_t seed, uint64_t skip, uint32_t keep, uint32_t drop,
                     const modlin_field_t *field, uint32_t elem, void *out, uint64_t count) {
    if (kind > MODLIN_PRNG_XOSHIRO256SS || elem > MODLIN_ELEM_U32 || (count && !out)) return MODLIN_EINVAL;
    if (keep == 0 && drop != 0) return MODLIN_EINVAL;
    if (elem == MODLIN_ELEM_U32 && (!field || field->p > 0xffffffffull)) return MODLIN_EINVAL;
    if (field && (field->p < 2 || field->p >= (1ull << 63))) return MODLIN_EINVAL;
    const bool mixmax = kind == MODLIN_PRNG_MIXMAX17 || kind == MODLIN_PRNG_MIXMAX240 || kind == MODLIN_PRNG_MIXMAX256;
    if (mixmax && seed == 0) {
        ctx->err = "the zero seed is not allowed for MIXMAX";  // prng.rs:153-155
        return MODLIN_EINVAL;
    }
    if (count == 0) return MODLIN_OK;
    FillOut o{};
    o.f = field ? make_params(field->p) : FieldParams{};
    o.raw = field ? 0 : 1;
    o.out = out;
    o.skip = skip;
    o.count = count;
    o.u32 = elem == MODLIN_ELEM_U32;
    o.keep = keep;
    o.period = keep + drop;
    const uint64_t last = skip + count - 1;
    if (last < skip) return MODLIN_EINVAL;
    if (keep) {
        const unsigned __int128 rl = (unsigned __int128)(last / keep) * o.period + last % keep;
        if (rl >= ((unsigned __int128)1 << 62)) return MODLIN_EINVAL;
        o.r_lo = (skip / keep) * o.period + skip % keep;
        o.r_hi = (uint64_t)rl + 1;
    } else {
        if (last >= (1ull << 62)) return MODLIN_EINVAL;
        o.r_lo = skip;
        o.r_hi = last + 1;
    }
    MDL_HIP(hipSetDevice(ctx->device));
    int rc = MODLIN_OK;
    switch (kind) {
        case MODLIN_PRNG_SPLITMIX64: {
            const unsigned grid = (unsigned)std::min<uint64_t>((count + 255) / 256, 64ull * (uint64_t)ctx->cu_count);
            hipLaunchKernelGGL(k_splitmix_fill, dim3(grid), dim3(256), 0, ctx->stream, seed, o);
            MDL_HIP(hipGetLastError());
            break;
        }
        case MODLIN_PRNG_XOROSHIRO128PP: rc = fill_gf2<Xoroshiro128pp, 2>(ctx, kind, Xoroshiro128pp(seed), o); break;
        case MODLIN_PRNG_XOSHIRO256SS: rc = fill_gf2<Xoshiro256ss, 4>(ctx, kind, Xoshiro256ss(seed), o); break;
        case MODLIN_PRNG_MIXMAX17: {
            MixMax17 g;
            g.seed(seed);
            rc = fill_mixmax(ctx, kind, g, o);
            break;
        }
        case MODLIN_PRNG_MIXMAX240: {
            MixMax240 g;
            g.seed(seed);
            rc = fill_mixmax(ctx, kind, g, o);
            break;
        }
        default: {
            MixMax256 g;
            g.seed(seed);
            rc = fill_mixmax(ctx, kind, g, o);
            break;
        }
    }
    if (rc != MODLIN_OK) return rc;
    MDL_HIP(hipStreamSynchronize(ctx->stream));
    return MODLIN_OK;
}

void release_generators(modlin_ctx *ctx) {
    for (auto &t : ctx->gen_tab)
        if (t) {
            hipFree(t);
            t = nullptr;
        }
    if (ctx->gen_ws) hipFree(ctx->gen_ws);
    ctx->gen_ws = nullptr;
    ctx->gen_ws_bytes = 0;
}

}  // namespace mdl
