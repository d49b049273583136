// Internal host-side declarations shared by the engine translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/modlin.h"
#include "fp_dev.hpp"
#include "tuning.hpp"

struct modlin_ctx {
    int device = 0;
    // launch parameters: explicit override (nonzero fields) over the table
    modlin_tuning_t tune_override{};
    std::vector<mdl::TuneEntry> tune_table;
    hipStream_t stream = nullptr;
    std::string err;
    modlin_stats_t stats{};
    // grow-only device arena reused across calls
    void *arena = nullptr;
    size_t arena_bytes = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, evA = nullptr, evB = nullptr;
    // second stream of the rank loop (the far part of each update) and its
    // handshake events
    hipStream_t far_stream = nullptr;
    hipEvent_t ev_plan = nullptr, ev_gather = nullptr, ev_far = nullptr;
    // captured 8-iteration launch block of the rank loop, reused while its
    // arguments (arena, batch shape, field, method) are unchanged
    hipGraphExec_t rank_graph = nullptr;
    std::string rank_graph_key;
    int cu_count = 256;
    // pinned host staging buffers for host <-> device transfers (double-buffered)
    void *pinned[2] = {nullptr, nullptr};
    size_t pinned_bytes = 0;
    hipEvent_t pinned_ev[2] = {nullptr, nullptr};
    // one pinned, coherent host word that a running BM kernel updates with its
    // step count (the progress callback of modlin_linear_complexity)
    uint64_t *prog_host = nullptr;
    // pinned host words for the allocation-free entry point's read-backs,
    // allocated with the context
    void *na_host = nullptr;
    // device generators: tabulated jump powers per generator kind, and a workspace
    void *gen_tab[8] = {};
    void *gen_ws = nullptr;
    size_t gen_ws_bytes = 0;
};

namespace mdl {

// Records a HIP failure on the context and returns MODLIN_EHIP.
int hip_fail(modlin_ctx *ctx, hipError_t e, const char *what);

// Ensures the context arena holds at least `bytes`; returns MODLIN_OK or an error.
int arena_reserve(modlin_ctx *ctx, size_t bytes);

// Host -> device copy of `bytes` from pageable memory through the context's
// pinned staging buffers, ordered on ctx->stream.  The source may be reused as
// soon as the call returns.
int copy_h2d(modlin_ctx *ctx, void *dst, const void *src, size_t bytes);
// Device -> host copy through the staging buffers; synchronous.
int copy_d2h(modlin_ctx *ctx, void *dst, const void *src, size_t bytes);
void release_pinned(modlin_ctx *ctx);

#define MDL_HIP(call)                                              \
    do {                                                           \
        hipError_t e_ = (call);                                    \
        if (e_ != hipSuccess) return mdl::hip_fail(ctx, e_, #call); \
    } while (0)

int rank_batch(modlin_ctx *ctx, uint64_t p, const modlin_mat_desc_t *mats, uint64_t count, const modlin_opts_t *opts,
               uint64_t *ranks, modlin_progress_fn progress, void *user);
// Device fill of a generator stream (prng_dev.hip); see modlin_prng_fill_device.
int prng_fill_device(modlin_ctx *ctx, uint32_t kind, uint64_t seed, uint64_t skip, uint32_t keep, uint32_t drop,
                     const modlin_field_t *field, uint32_t elem, void *out, uint64_t count);
void release_generators(modlin_ctx *ctx);

// modlin_linear_complexity_noalloc: the scratch array's address on the context's
// device (once per device, at context creation), and the call itself.
int noalloc_init(modlin_ctx *ctx);
int lc_noalloc(modlin_ctx *ctx, uint64_t p, const uint64_t *s, uint64_t n, uint64_t *lc);
constexpr size_t NA_HOST_BYTES = 4096;

// progress (may be null) is called with the steps done of a single sequence
// (count == 1) while its kernel runs.
int lc_batch(modlin_ctx *ctx, uint64_t p, const modlin_seq_desc_t *seqs, uint64_t count, const modlin_opts_t *opts,
             uint64_t *lcs, modlin_progress_fn progress = nullptr, void *user = nullptr);

}  // namespace mdl
