// The C ABI (include/modlin.h): argument checking, status codes, context
// management, and thin forwarding to the engines.  No C++ exception crosses it.
#include <string.h>

#include <new>
#include <string>

#include "../../include/modlin.h"
#include "engine.hpp"
#include "prng.hpp"
#include "stats.hpp"

using mdl::FieldParams;

namespace {

bool valid_p(uint64_t p) { return p >= 2 && p < (1ull << 63); }

int fill_field(uint64_t p, modlin_field_t *out) {
    if (!out || !valid_p(p)) return MODLIN_EINVAL;
    const FieldParams f = mdl::make_params(p);
    out->p = p;
    out->mu_hi = f.mu_hi;
    out->mu_lo = f.mu_lo;
    out->safe_batch = f.safe_batch;
    return MODLIN_OK;
}

bool field_ok(const modlin_field_t *f) {
    if (!f || !valid_p(f->p)) return false;
    const FieldParams g = mdl::make_params(f->p);
    return g.mu_hi == f->mu_hi && g.mu_lo == f->mu_lo && g.safe_batch == f->safe_batch;
}

}  // namespace

extern "C" {

int modlin_field_init(uint64_t p, modlin_field_t *out) { return fill_field(p, out); }

// An invalid field gives UINT64_MAX, which no residue (< p < 2^63) equals.
uint64_t modlin_reduce(const modlin_field_t *f, uint64_t x) {
    if (!f || !valid_p(f->p)) return UINT64_MAX;
    const FieldParams g = mdl::make_params(f->p);
    return mdl::reduce128(g, 0, x);
}

uint64_t modlin_reduce_u128(const modlin_field_t *f, uint64_t hi, uint64_t lo) {
    if (!f || !valid_p(f->p)) return UINT64_MAX;
    const FieldParams g = mdl::make_params(f->p);
    return mdl::reduce128(g, hi, lo);
}

int modlin_ctx_create(int device, modlin_ctx_t **out) {
    if (!out) return MODLIN_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        return MODLIN_ENODEV;
    }
    if (device < 0 || device >= ndev) return MODLIN_ENODEV;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return MODLIN_ENODEV;
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) return MODLIN_ENODEV;
    modlin_ctx *ctx = new (std::nothrow) modlin_ctx();
    if (!ctx) return MODLIN_ENOMEM;
    ctx->device = device;
    ctx->cu_count = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess ||
        hipEventCreate(&ctx->evA) != hipSuccess || hipEventCreate(&ctx->evB) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->far_stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_plan, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_gather, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_far, hipEventDisableTiming) != hipSuccess ||
        hipHostMalloc(&ctx->na_host, mdl::NA_HOST_BYTES, hipHostMallocDefault) != hipSuccess ||
        mdl::noalloc_init(ctx) != MODLIN_OK) {
        (void)hipGetLastError();
        modlin_ctx_destroy(ctx);
        return MODLIN_EHIP;
    }
    // launch parameters from tuning.txt next to the library, if one was written
    std::vector<mdl::TuneEntry> table;
    std::string err;
    if (mdl::parse_tuning(mdl::default_tuning_path(), table, err)) ctx->tune_table.swap(table);
    *out = ctx;
    return MODLIN_OK;
}

int modlin_ctx_set_tuning(modlin_ctx_t *ctx, const modlin_tuning_t *t) {
    if (!ctx) return MODLIN_EINVAL;
    if (t && ((t->check_every & 1) || t->panel > 64 || t->bm_coop_wgs > 1024 || !mdl::bm_threads_ok(t->bm_threads) ||
              t->far_depth > 2 || t->far_quarters > 2)) {
        ctx->err = "tuning: check_every must be even, panel <= 64, bm_coop_wgs <= 1024, bm_threads 256, 512 or 1024, "
                   "far_depth and far_quarters 1 or 2";
        return MODLIN_EINVAL;
    }
    ctx->tune_override = t ? *t : modlin_tuning_t{};
    return MODLIN_OK;
}

int modlin_ctx_load_tuning(modlin_ctx_t *ctx, const char *path) {
    if (!ctx) return MODLIN_EINVAL;
    if (path && !*path) {
        ctx->tune_table.clear();
        return MODLIN_OK;
    }
    std::vector<mdl::TuneEntry> table;
    if (!mdl::parse_tuning(path ? std::string(path) : mdl::default_tuning_path(), table, ctx->err)) return MODLIN_EINVAL;
    ctx->tune_table.swap(table);
    return MODLIN_OK;
}

void modlin_ctx_destroy(modlin_ctx_t *ctx) {
    if (!ctx) return;
    hipSetDevice(ctx->device);
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    if (ctx->far_stream) hipStreamSynchronize(ctx->far_stream);
    if (ctx->arena) hipFree(ctx->arena);
    mdl::release_pinned(ctx);
    mdl::release_generators(ctx);
    if (ctx->prog_host) hipHostFree(ctx->prog_host);
    if (ctx->na_host) hipHostFree(ctx->na_host);
    if (ctx->ev0) hipEventDestroy(ctx->ev0);
    if (ctx->ev1) hipEventDestroy(ctx->ev1);
    if (ctx->evA) hipEventDestroy(ctx->evA);
    if (ctx->evB) hipEventDestroy(ctx->evB);
    if (ctx->rank_graph) hipGraphExecDestroy(ctx->rank_graph);
    if (ctx->ev_plan) hipEventDestroy(ctx->ev_plan);
    if (ctx->ev_gather) hipEventDestroy(ctx->ev_gather);
    if (ctx->ev_far) hipEventDestroy(ctx->ev_far);
    if (ctx->far_stream) hipStreamDestroy(ctx->far_stream);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *modlin_ctx_last_error(const modlin_ctx_t *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int modlin_ctx_last_stats(const modlin_ctx_t *ctx, modlin_stats_t *out) {
    if (!ctx || !out) return MODLIN_EINVAL;
    *out = ctx->stats;
    return MODLIN_OK;
}

int modlin_rank(modlin_ctx_t *ctx, const modlin_field_t *f, uint64_t *m, uint64_t n, modlin_progress_fn progress,
                void *user, uint64_t *rank) {
    if (!ctx || !rank) return MODLIN_EINVAL;
    if (!field_ok(f)) {
        ctx->err = "invalid field";
        return MODLIN_EINVAL;
    }
    if (n > 0 && !m) {
        ctx->err = "null matrix";
        return MODLIN_EINVAL;
    }
    hipSetDevice(ctx->device);
    modlin_mat_desc_t d{m, n, MODLIN_ELEM_U64, MODLIN_MEM_HOST};
    try {
        return mdl::rank_batch(ctx, f->p, &d, 1, nullptr, rank, progress, user);
    } catch (const std::bad_alloc &) {
        ctx->err = "host allocation failed";
        return MODLIN_ENOMEM;
    } catch (...) {
        ctx->err = "internal error";
        return MODLIN_EHIP;
    }
}

int modlin_rank_batch(modlin_ctx_t *ctx, const modlin_field_t *f, const modlin_mat_desc_t *mats, uint64_t count,
                      const modlin_opts_t *opts, uint64_t *ranks) {
    if (!ctx || (count && (!mats || !ranks))) return MODLIN_EINVAL;
    if (!field_ok(f)) {
        ctx->err = "invalid field";
        return MODLIN_EINVAL;
    }
    if (opts && (opts->method > MODLIN_METHOD_MFMA || opts->panel > 64)) {
        ctx->err = "invalid options";
        return MODLIN_EINVAL;
    }
    hipSetDevice(ctx->device);
    try {
        return mdl::rank_batch(ctx, f->p, mats, count, opts, ranks, nullptr, nullptr);
    } catch (const std::bad_alloc &) {
        ctx->err = "host allocation failed";
        return MODLIN_ENOMEM;
    } catch (...) {
        ctx->err = "internal error";
        return MODLIN_EHIP;
    }
}

int modlin_linear_complexity(modlin_ctx_t *ctx, const modlin_field_t *f, const uint64_t *s, uint64_t n,
                             modlin_progress_fn progress, void *user, uint64_t *lc) {
    if (!ctx || !lc) return MODLIN_EINVAL;
    if (!field_ok(f)) {
        ctx->err = "invalid field";
        return MODLIN_EINVAL;
    }
    if (n == 0 || !s) {
        ctx->err = "empty sequence (the reference panics, fp.rs:265)";
        return MODLIN_EINVAL;
    }
    hipSetDevice(ctx->device);
    modlin_seq_desc_t d{s, n, MODLIN_ELEM_U64, MODLIN_MEM_HOST};
    try {
        int rc = mdl::lc_batch(ctx, f->p, &d, 1, nullptr, lc, progress, user);
        if (rc == MODLIN_OK && progress) progress(user, n, n);
        return rc;
    } catch (const std::bad_alloc &) {
        ctx->err = "host allocation failed";
        return MODLIN_ENOMEM;
    } catch (...) {
        ctx->err = "internal error";
        return MODLIN_EHIP;
    }
}

int modlin_linear_complexity_noalloc(modlin_ctx_t *ctx, const modlin_field_t *f, const uint64_t *s, uint64_t n,
                                     uint64_t *lc) {
    if (!ctx || !lc) return MODLIN_EINVAL;
    if (!field_ok(f)) {
        ctx->err = "invalid field";
        return MODLIN_EINVAL;
    }
    if (!s) {
        ctx->err = "null sequence";
        return MODLIN_EINVAL;
    }
    hipSetDevice(ctx->device);
    return mdl::lc_noalloc(ctx, f->p, s, n, lc);
}

int modlin_lc_batch(modlin_ctx_t *ctx, const modlin_field_t *f, const modlin_seq_desc_t *seqs, uint64_t count,
                    const modlin_opts_t *opts, uint64_t *lcs) {
    if (!ctx || (count && (!seqs || !lcs))) return MODLIN_EINVAL;
    if (!field_ok(f)) {
        ctx->err = "invalid field";
        return MODLIN_EINVAL;
    }
    hipSetDevice(ctx->device);
    try {
        return mdl::lc_batch(ctx, f->p, seqs, count, opts, lcs);
    } catch (const std::bad_alloc &) {
        ctx->err = "host allocation failed";
        return MODLIN_ENOMEM;
    } catch (...) {
        ctx->err = "internal error";
        return MODLIN_EHIP;
    }
}

double modlin_corank_prob(uint64_t p, uint64_t n, uint64_t d) { return modlin::stats::corank_prob(p, n, d); }
double modlin_corank_tail_pvalue(uint64_t p, uint64_t n, uint64_t c) {
    return modlin::stats::corank_tail_pvalue(p, n, c);
}
double modlin_lc_left_tail_pvalue(uint64_t p, uint64_t n, uint64_t ell) {
    return modlin::stats::lc_left_tail_pvalue(p, n, ell);
}
int modlin_pretty_p_value(double x, char *buf, size_t cap) {
    const std::string s = modlin::stats::pretty_p_value(x);
    if (!buf || cap < s.size() + 1) return -1;
    memcpy(buf, s.c_str(), s.size() + 1);
    return (int)s.size();
}

const char *modlin_prng_name(uint32_t kind) {
    switch (kind) {
        case MODLIN_PRNG_XOROSHIRO128PP: return "xoroshiro128++";                            // prng.rs:56
        case MODLIN_PRNG_MIXMAX17: return "MIXMAX (TRandomMixMax17, N=17)";                 // prng.rs:312
        case MODLIN_PRNG_MIXMAX240: return "MIXMAX (TRandomMixMax, N=240, s=487013230256099140)";  // prng.rs:291
        case MODLIN_PRNG_MIXMAX256: return "MIXMAX (TRandomMixMax256, N=256, skip=2)";      // prng.rs:333
        case MODLIN_PRNG_SPLITMIX64: return "SplitMix64";
        case MODLIN_PRNG_XOSHIRO256SS: return "xoshiro256**";
        default: return nullptr;
    }
}

int modlin_prng_fill(uint32_t kind, uint64_t seed, uint64_t skip, uint64_t *out, uint64_t count) {
    using namespace modlin::prng;
    if (count && !out) return MODLIN_EINVAL;
    auto run = [&](auto &g) {
        for (uint64_t i = 0; i < skip; i++) (void)g.next_u64();
        for (uint64_t i = 0; i < count; i++) out[i] = g.next_u64();
    };
    switch (kind) {
        case MODLIN_PRNG_XOROSHIRO128PP: { Xoroshiro128pp g(seed); run(g); return MODLIN_OK; }
        case MODLIN_PRNG_SPLITMIX64: { SplitMix64 g(seed); run(g); return MODLIN_OK; }
        case MODLIN_PRNG_XOSHIRO256SS: { Xoshiro256ss g(seed); run(g); return MODLIN_OK; }
        case MODLIN_PRNG_MIXMAX17: { MixMax17 g; if (!g.seed(seed)) return MODLIN_EINVAL; run(g); return MODLIN_OK; }
        case MODLIN_PRNG_MIXMAX240: { MixMax240 g; if (!g.seed(seed)) return MODLIN_EINVAL; run(g); return MODLIN_OK; }
        case MODLIN_PRNG_MIXMAX256: { MixMax256 g; if (!g.seed(seed)) return MODLIN_EINVAL; run(g); return MODLIN_OK; }
        default: return MODLIN_EINVAL;
    }
}

int modlin_prng_fill_device(modlin_ctx_t *ctx, uint32_t kind, uint64_t seed, uint64_t skip, uint32_t keep,
                            uint32_t drop, const modlin_field_t *f, uint32_t elem, void *out, uint64_t count) {
    if (!ctx) return MODLIN_EINVAL;
    return mdl::prng_fill_device(ctx, kind, seed, skip, keep, drop, f, elem, out, count);
}

int modlin_dev_alloc(modlin_ctx_t *ctx, uint64_t bytes, void **out) {
    if (!ctx || !out) return MODLIN_EINVAL;
    hipSetDevice(ctx->device);
    if (hipMalloc(out, bytes ? bytes : 1) != hipSuccess) {
        (void)hipGetLastError();
        ctx->err = "device allocation failed";
        return MODLIN_ENOMEM;
    }
    return MODLIN_OK;
}
int modlin_dev_free(modlin_ctx_t *ctx, void *ptr) {
    if (!ctx) return MODLIN_EINVAL;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    return hipFree(ptr) == hipSuccess ? MODLIN_OK : MODLIN_EHIP;
}
int modlin_memcpy_h2d(modlin_ctx_t *ctx, void *dst, const void *src, uint64_t bytes) {
    if (!ctx) return MODLIN_EINVAL;
    hipSetDevice(ctx->device);
    return mdl::copy_h2d(ctx, dst, src, bytes);
}
int modlin_memcpy_d2h(modlin_ctx_t *ctx, void *dst, const void *src, uint64_t bytes) {
    if (!ctx) return MODLIN_EINVAL;
    hipSetDevice(ctx->device);
    return mdl::copy_d2h(ctx, dst, src, bytes);
}
int modlin_synchronize(modlin_ctx_t *ctx) {
    if (!ctx) return MODLIN_EINVAL;
    hipSetDevice(ctx->device);
    hipError_t e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? MODLIN_OK : mdl::hip_fail(ctx, e, "hipStreamSynchronize");
}

}  // extern "C"
