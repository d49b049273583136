/*
 * modlin -- MI355X (gfx950) implementation of the hot paths of modlin-rs:
 * rank over F_p and Berlekamp-Massey linear complexity, with the reference's
 * Field contract and p-values.  C ABI: plain pointers, sizes and status codes;
 * no exceptions or aborts cross it.
 *
 * Reference interfaces replaced (paths relative to the modlin-rs tree):
 *   modlin_field_init         Field::new                     src/fp.rs:49-61
 *   modlin_reduce             Field::reduce                  src/fp.rs:63-67
 *   modlin_reduce_u128        Field::reduce_u128             src/fp.rs:69-79
 *   modlin_rank               fp::rank                       src/fp.rs:116-255
 *   modlin_linear_complexity  fp::linear_complexity          src/fp.rs:257-308
 *   modlin_rank_batch         main.rs's loop of fp::rank over reps (main.rs:82-127),
 *                             one launch sequence for all matrices
 *   modlin_lc_batch           main.rs's loop of linear_complexity (main.rs:152-175)
 *   modlin_corank_prob        stats::corank_prob             src/stats.rs:52-62
 *   modlin_corank_tail_pvalue stats::corank_tail_pvalue      src/stats.rs:69-84
 *   modlin_lc_left_tail_pvalue stats::lc_left_tail_pvalue    src/stats.rs:101-126
 *   modlin_pretty_p_value     stats::pretty_p_value          src/stats.rs:129-137
 *   modlin_prng_fill          Prng::try_new + next_u64       src/prng.rs:46-343
 *   modlin_prng_fill_device   the fill loops, on the device   main.rs:94-105, 153-155
 *   modlin_ctx_set_tuning, modlin_ctx_load_tuning, modlin_tuning_select
 *                             no counterpart: launch parameters of this library, chosen by
 *                             modlin-rs_amd/autotune.py (the reference's PANEL = 64 and
 *                             TILE = 256 constants, src/fp.rs:26-32, are fixed)
 *
 * Ownership: the caller owns every buffer.  modlin_rank leaves m unspecified
 * (the reference also destroys it, holding partial factors).  A context owns
 * device memory, the stream and cached plans; use one context per thread.
 */
#ifndef MODLIN_H
#define MODLIN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define MODLIN_OK 0
#define MODLIN_EINVAL 1      /* bad argument: p outside [2, 2^63), n = 0 for BM, overflow */
#define MODLIN_ENOMEM 2      /* host or device allocation failed */
#define MODLIN_EHIP 3        /* HIP runtime error; see modlin_ctx_last_error */
#define MODLIN_ENODEV 4      /* no usable gfx950 device */
#define MODLIN_EUNSUPPORTED 5

/* Field (src/fp.rs:34-47): p, Barrett reciprocal mu = floor((2^128-1)/p) split
 * in two words, and safe_batch = number of (p-1)^2 products that fit a u128. */
typedef struct modlin_field {
    uint64_t p;
    uint64_t mu_hi, mu_lo;
    uint64_t safe_batch;
} modlin_field_t;

typedef struct modlin_ctx modlin_ctx_t;

/* Progress callback, called on the calling thread while a call runs (NULL
 * disables it).  modlin_rank: done = matrix columns below the current panel
 * (the reference's update_with_count per panel, src/fp.rs:251), at every host
 * check of the iteration loop (every check_every panels), total = n; done never
 * decreases and the last call has done = total.  modlin_linear_complexity:
 * done = steps the kernel has passed (the reference's update() per step,
 * src/fp.rs:273, reported every 64 to 256 steps), total = n, last call
 * done = total. */
typedef void (*modlin_progress_fn)(void *user, uint64_t done, uint64_t total);

enum modlin_method {
    MODLIN_METHOD_AUTO = 0,  /* MFMA limb update where it pays, else plain */
    MODLIN_METHOD_PLAIN = 1, /* VALU Barrett trailing update */
    MODLIN_METHOD_MFMA = 2   /* signed-digit int8 limb update on MFMA */
};
enum modlin_elem { MODLIN_ELEM_U64 = 0, MODLIN_ELEM_U32 = 1 };
enum modlin_mem { MODLIN_MEM_HOST = 0, MODLIN_MEM_DEVICE = 1 };

/* One matrix of a batch: n x n row-major, entries in [0, p). */
typedef struct modlin_mat_desc {
    const void *data;
    uint64_t n;
    uint32_t elem; /* modlin_elem */
    uint32_t mem;  /* modlin_mem */
} modlin_mat_desc_t;

/* One sequence of a batch, elements in [0, p). */
typedef struct modlin_seq_desc {
    const void *data;
    uint64_t n;
    uint32_t elem;
    uint32_t mem;
} modlin_seq_desc_t;

#define MODLIN_FLAG_PROFILE 1u  /* time the dominant kernels with HIP events (direct launches) */
#define MODLIN_FLAG_NO_GRAPH 2u /* launch each iteration block directly instead of replaying the
                                   context's captured hipGraph of it */
#define MODLIN_FLAG_NO_DEFER 4u /* rank, MFMA method: update the far columns once per panel instead
                                   of once per two panels (the same ranks) */
#define MODLIN_FLAG_U64_STORE 8u /* rank, MFMA method, p < 2^32: keep the working copy in u64 entries
                                    instead of u32 (the same ranks) */

typedef struct modlin_opts {
    uint32_t method; /* modlin_method.  BM: PLAIN runs one step per pass over c; AUTO and MFMA
                        run blocks of 64 steps (p < 2^62).  The result is the same. */
    uint32_t panel;  /* panel width (0 = tuned default, else 1..64) */
    uint32_t flags;  /* MODLIN_FLAG_* */
    uint32_t coop_min; /* BM: a batch of one sequence at least this long runs cooperatively on
                          the whole GPU (0 = default, 65536) */
} modlin_opts_t;

/* Counters of the last rank / BM call on a context. */
typedef struct modlin_stats {
    double macs;            /* F_p multiply-accumulates in trailing updates */
    double update_ms;       /* summed duration of trailing-update kernels (PROFILE) */
    double update_launches; /* number of trailing-update launches (PROFILE) */
    double limb_count;      /* int8 limbs per element of the MFMA update (0 = VALU) */
    double iterations;      /* panel iterations */
    double bm_bytes;        /* algorithmic BM bytes */
    double bm_ms;           /* BM kernel time (PROFILE) */
    double total_ms;        /* wall time of the device part of the call */
    double update_bytes;    /* algorithmic HBM bytes of the trailing updates: 16 per updated
                               element (read + write) + 8 per operand element per tile */
    double graph;           /* 1 if the rank loop replayed a captured hipGraph */
    double bm_coop;         /* workgroups of the cooperative BM kernel (0: one per sequence) */
    double far_macs;        /* the part of macs in the far-column updates */
    double far_ms;          /* summed duration of the far-column update kernels (PROFILE) */
    double far_launches;    /* number of far-column update launches (PROFILE) */
} modlin_stats_t;

/* Launch parameters of the rank and BM engines.  A zero field means "from the
   context's tuning table, else the built-in default".  No value changes a result:
   rank and linear complexity do not depend on blocking or on launch shapes. */
typedef struct modlin_tuning {
    uint32_t far_per_cu;  /* far-update workgroups per CU (default 4 MFMA, 8 VALU) */
    uint32_t near_per_cu; /* near-update workgroups per CU (default 2) */
    uint32_t gather_wgs;  /* workgroups of the gather / X-product kernel (default 1024) */
    uint32_t check_every; /* panel iterations between host checks, even (default 8) */
    uint32_t panel;       /* panel width 1..64 (default 64; modlin_opts.panel overrides) */
    uint32_t bm_coop_wgs; /* workgroups of the cooperative BM, 1..1024 (default: one per CU) */
    uint32_t bm_threads;  /* threads per sequence of the batched BM: 256, 512 or 1024 (default:
                             the most that keep the whole batch resident) */
    uint32_t far_depth;   /* MFMA method: panels per far update, 1 or 2 (default 2: a clean panel
                             leaves its far columns to the next one, K up to 128 per pass) */
    uint32_t far_quarters; /* MFMA method, 4 digits or fewer (p < 2^32): 16-column quarters per
                              B fragment read in the far update, 1 or 2 (default 2) */
    uint32_t reserved[3];
} modlin_tuning_t;

/* --- Field (src/fp.rs:49-79) -------------------------------------------- */
/* modlin_reduce and modlin_reduce_u128 return UINT64_MAX for an invalid field
 * (NULL, or p outside [2, 2^63)): no residue equals it, since residues are < p. */
int modlin_field_init(uint64_t p, modlin_field_t *out);
uint64_t modlin_reduce(const modlin_field_t *f, uint64_t x);
uint64_t modlin_reduce_u128(const modlin_field_t *f, uint64_t hi, uint64_t lo);

/* --- Context --------------------------------------------------------------- */
int modlin_ctx_create(int device, modlin_ctx_t **out);
void modlin_ctx_destroy(modlin_ctx_t *ctx);
const char *modlin_ctx_last_error(const modlin_ctx_t *ctx);
int modlin_ctx_last_stats(const modlin_ctx_t *ctx, modlin_stats_t *out);

/* Tuning.  A table file has one entry per line,
 *     <rank|bm> <mode> <side> <count> far_per_cu near_per_cu gather_wgs check_every panel bm_coop_wgs bm_threads
 *         [far_depth far_quarters]
 * with mode m61 (p = 2^61-1), m31 (p = 2^31-1) or gL (other p, L int8 digits per
 * element), side the power of two at or above the largest n of a call (at least
 * 256), count 1 or many; '#' starts a comment.  A call uses the entry of its kind,
 * mode and count with the nearest side, else built-in defaults.
 * modlin_ctx_create loads tuning.txt from the library's directory if it exists.
 * modlin_ctx_load_tuning(ctx, NULL) reloads it, (ctx, "") clears the table, any
 * other path loads that file (MODLIN_EINVAL if it cannot be read or parsed).
 * modlin_ctx_set_tuning(ctx, &t) makes t's nonzero fields override the table for
 * every later call; NULL removes the override.
 * modlin_tuning_select applies the same selection to a table file without a
 * context (no GPU needed); bm != 0 selects a Berlekamp-Massey entry. */
int modlin_ctx_set_tuning(modlin_ctx_t *ctx, const modlin_tuning_t *t);
int modlin_ctx_load_tuning(modlin_ctx_t *ctx, const char *path);
int modlin_tuning_select(const char *path, uint64_t p, uint64_t max_n, uint64_t count, int bm, modlin_tuning_t *out);

/* --- Hot paths ------------------------------------------------------------- */
int modlin_rank(modlin_ctx_t *ctx, const modlin_field_t *f, uint64_t *m, uint64_t n, modlin_progress_fn progress,
                void *user, uint64_t *rank);
int modlin_linear_complexity(modlin_ctx_t *ctx, const modlin_field_t *f, const uint64_t *s, uint64_t n,
                             modlin_progress_fn progress, void *user, uint64_t *lc);
int modlin_rank_batch(modlin_ctx_t *ctx, const modlin_field_t *f, const modlin_mat_desc_t *mats, uint64_t count,
                      const modlin_opts_t *opts, uint64_t *ranks);
int modlin_lc_batch(modlin_ctx_t *ctx, const modlin_field_t *f, const modlin_seq_desc_t *seqs, uint64_t count,
                    const modlin_opts_t *opts, uint64_t *lcs);

/* Allocation-free Berlekamp-Massey of one sequence already in device memory.
 * No counterpart in the reference: for callers that cannot allocate device or
 * pinned host memory during a call (tight serving loops, harnesses that forbid
 * hipMalloc).  s: n u64 entries in [0, p) in device memory of the context's GPU,
 * 1 <= n <= MODLIN_LC_NOALLOC_MAX_N (MODLIN_EINVAL otherwise, before any launch);
 * *lc receives the linear complexity, the same as modlin_linear_complexity's.
 * The call allocates nothing: its state (about 34 MiB) is a fixed device array
 * that each GPU holds until the process ends, shared by every call on that GPU,
 * so calls that overlap on one GPU (several contexts or threads) race.
 * Synchronous on the context's stream. */
#define MODLIN_LC_NOALLOC_MAX_N 1048576u
int modlin_linear_complexity_noalloc(modlin_ctx_t *ctx, const modlin_field_t *f, const uint64_t *s, uint64_t n,
                                     uint64_t *lc);

/* --- Statistics (src/stats.rs) ------------------------------------------- */
double modlin_corank_prob(uint64_t p, uint64_t n, uint64_t d);
double modlin_corank_tail_pvalue(uint64_t p, uint64_t n, uint64_t c);
double modlin_lc_left_tail_pvalue(uint64_t p, uint64_t n, uint64_t ell);
/* Writes "0", "1" or Rust's {:?} of x; returns the length, or -1 if cap is too small. */
int modlin_pretty_p_value(double x, char *buf, size_t cap);

/* --- Generators (src/prng.rs) ---------------------------------------------- */
enum modlin_prng {
    MODLIN_PRNG_XOROSHIRO128PP = 0, /* prng.rs:46-84 */
    MODLIN_PRNG_MIXMAX17 = 1,       /* prng.rs:303-322 */
    MODLIN_PRNG_MIXMAX240 = 2,      /* prng.rs:282-301 */
    MODLIN_PRNG_MIXMAX256 = 3,      /* prng.rs:324-343 */
    MODLIN_PRNG_SPLITMIX64 = 4,     /* published SplitMix64 (workload generation) */
    MODLIN_PRNG_XOSHIRO256SS = 5    /* published xoshiro256**, SplitMix64-seeded */
};
/* Fills out[0..count) with next_u64() after discarding `skip` outputs.
 * Returns MODLIN_EINVAL for a zero MIXMAX seed (prng.rs:153-155). */
int modlin_prng_fill(uint32_t kind, uint64_t seed, uint64_t skip, uint64_t *out, uint64_t count);
const char *modlin_prng_name(uint32_t kind);
/* The fill loops of main.rs:94-105 and main.rs:153-155 on the device: writes
 * out[i] = Field::reduce(next_u64()) for the outputs skip .. skip+count-1 of the
 * stream (kind, seed), straight into device memory `out` (u64, or u32 when
 * p < 2^32), bit-identical to modlin_prng_fill.  f == NULL writes the raw
 * next_u64() words (u64 only).  keep > 0 decimates the stream first: keep
 * outputs, drop `drop` (the mixmax-lux feature uses 5 and 9, prng.rs:272-279).
 * Positions are reached by exact jump-ahead, so the cost does not grow with
 * skip.  Runs on the context's stream and returns when the fill is complete.
 * Returns MODLIN_EINVAL for a zero MIXMAX seed (prng.rs:153-155) or bad sizes. */
int modlin_prng_fill_device(modlin_ctx_t *ctx, uint32_t kind, uint64_t seed, uint64_t skip, uint32_t keep,
                            uint32_t drop, const modlin_field_t *f, uint32_t elem, void *out, uint64_t count);

/* --- Workload helpers (host; the fill side of main.rs:86-105) ---------------- */
/* out[i][c] = sum_k coef[i][k] * rows[k][c] mod p, for i < nout, k < nsrc, c < n
 * (row-major; rows and out may not overlap).  Builds planted-deficiency matrices. */
int modlin_util_combine_rows(uint64_t p, const uint64_t *rows, uint64_t nsrc, uint64_t n, const uint64_t *coef,
                             uint64_t nout, uint64_t *out);
/* s[i] = sum_{j<L} coef[j] * s[i-1-j] mod p for i >= L, with s[0..L) = init:
 * a sequence of linear complexity at most L. */
int modlin_util_lfsr(uint64_t p, const uint64_t *coef, uint64_t L, const uint64_t *init, uint64_t n, uint64_t *out);

/* --- Device buffers for zero-copy callers ---------------------------------- */
int modlin_dev_alloc(modlin_ctx_t *ctx, uint64_t bytes, void **out);
int modlin_dev_free(modlin_ctx_t *ctx, void *ptr);
int modlin_memcpy_h2d(modlin_ctx_t *ctx, void *dst, const void *src, uint64_t bytes);
int modlin_memcpy_d2h(modlin_ctx_t *ctx, void *dst, const void *src, uint64_t bytes);
int modlin_synchronize(modlin_ctx_t *ctx);

#ifdef __cplusplus
}
#endif

#endif /* MODLIN_H */
