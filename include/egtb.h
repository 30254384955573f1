/*
 * egtb.h -- C ABI of libegtb_hip.so, the MI355X (gfx950) engine for the
 * classic 2048 endgame-tablebase path: forward layer enumeration, sort/unique
 * dedup, backward success-rate solve, and batched lookups over solved tables.
 *
 * All functions are plain C: pointers and sizes, no C++ or torch types.
 * Functions returning int return 0 on success and a negative code on
 * failure; egtb_last_error() then gives the message (per thread).
 *
 * Board encoding: a uint64 of 16 nibbles, nibble k = cell (row 3-k/4,
 * column 3-k%4), value v = tile 2^v, 0 = empty.
 */
#ifndef EGTB_H
#define EGTB_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#if defined(__GNUC__)
#define EGTB_EXPORT __attribute__((visibility("default")))
#else
#define EGTB_EXPORT
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define EGTB_OK 0
#define EGTB_ERR_ARG (-1)
#define EGTB_ERR_HIP (-2)
#define EGTB_ERR_IO (-3)
#define EGTB_ERR_MEMORY (-4)
#define EGTB_ERR_INTERNAL (-5)

/* Pattern of a build.  Replaces PatternSpec
 * (reference native_core/include/FormationRuntime.h:17-22).
 * symm_mode: 0 identity, 1 full, 2 diagonal, 3 horizontal (SymmMode,
 * FormationRuntime.h:13; the variant-board modes 4-6 are rejected). */
typedef struct egtb_pattern {
    const uint64_t *masks;
    uint32_t n_masks;   /* <= 16 */
    const uint8_t *shifts;
    uint32_t n_shifts;  /* <= 16 */
    int32_t symm_mode;
} egtb_pattern;

/* Options of a build.  Replaces the classic-path fields of RunOptions
 * (FormationRuntime.h:24-42); target is log2 of the target tile. */
typedef struct egtb_options {
    int32_t target;
    int32_t steps;
    int32_t docheck_step;
    const char *pathname;            /* file prefix "<folder>/<pattern>_<tile>_" */
    int32_t is_free;                 /* accepted for parity; sizes are exact here */
    double spawn_rate4;              /* probability of a 4, reference default 0.1 */
    const char *success_rate_dtype;  /* "float32", "float64", "uint32", "uint64" */
    int32_t write_files;             /* 1: write <pathname><i>.book + stats.txt */
    int32_t keep_raw;                /* 1: keep raw layers for egtb_result_raw */
    int32_t device;                  /* HIP device ordinal */
    int32_t radix_bits;              /* 0 = autotuned; else 4, 6, 8 or 11 */
    int32_t reserved[7];
} egtb_options;

/* Build a tablebase end to end on the GPU and write the reference's file
 * set (<pathname><i>.book for every layer, <pathname>stats.txt).
 * Replaces formation_core.run_pattern_build
 * (reference native_core/src/bindings_formation.cpp:220-229 ->
 * run_pattern_build_cpp, BookSolver.cpp:823-832). */
EGTB_EXPORT int egtb_build(const uint64_t *seeds, size_t n_seeds, const egtb_pattern *pattern,
               const egtb_options *options);

/* Progress counters of the running build (relaxed atomics, any thread).
 * Replaces formation_core.get_build_progress
 * (bindings_formation.cpp:147-153, FormationRuntime.h:61-86). */
EGTB_EXPORT void egtb_get_build_progress(uint32_t *current, uint32_t *total);

/* Message of the last failure on this thread ("" when none).  The
 * reference maps C++ exceptions to Python exceptions through nanobind
 * (bindings_formation.cpp:220-229); this is the C equivalent. */
EGTB_EXPORT const char *egtb_last_error(void);

/* ---- in-memory results (tests, serving, batched builds) ------------------- */

typedef struct egtb_result egtb_result;

/* As egtb_build, but also keeps every solved layer resident in HBM and
 * returns a handle.  Same semantics as egtb_build otherwise. */
EGTB_EXPORT int egtb_build_ex(const uint64_t *seeds, size_t n_seeds, const egtb_pattern *pattern,
                  const egtb_options *options, egtb_result **out);

/* Per-layer counts: raw size N_i, live size N'_i, forward candidates
 * (spawn-2 emitted at step i, spawn-4 emitted at step i). */
EGTB_EXPORT int egtb_result_layer_info(const egtb_result *r, int layer, uint64_t *n_raw, uint64_t *n_live,
                           uint64_t *c2, uint64_t *c4);
EGTB_EXPORT int egtb_result_steps(const egtb_result *r);
/* Copy raw layer `layer` to host (needs options.keep_raw). */
EGTB_EXPORT int egtb_result_raw(const egtb_result *r, int layer, uint64_t *keys, size_t cap, size_t *n);
/* Copy solved layer `layer` to host: keys and rates (rate element size is
 * 4 for float32/uint32, 8 for float64/uint64). */
EGTB_EXPORT int egtb_result_book(const egtb_result *r, int layer, uint64_t *keys, void *rates, size_t cap,
                     size_t *n);
/* Timing of the last build: forward, backward and total milliseconds. */
EGTB_EXPORT int egtb_result_timing(const egtb_result *r, double *fwd_ms, double *bwd_ms, double *total_ms);
EGTB_EXPORT void egtb_result_free(egtb_result *r);

/* ---- dedup primitives: drop-in for the reference's bookgen_native.so ----- */

/* Ascending (or descending) sort of a host array on the GPU.  Same name and
 * signature as the reference plugin's export (native_core/xss_wrapper.cpp:
 * 14-31), resolved by BookGeneratorUtils.cpp:68-94. */
EGTB_EXPORT void sort_uint64(uint64_t *arr, size_t arrsize, bool descending);

/* Adjacent dedup of a sorted host array in place; returns the new length.
 * Same name and signature as native_core/src/unique_wrapper.cpp:12-14. */
EGTB_EXPORT size_t unique_sorted_u64_inplace(uint64_t *data, size_t size);

/* ---- kernel statistics (bench / profiling) ------------------------------- */

/* Accumulated HIP-event time and algorithmic bytes of one kernel family
 * ("expand", "sort", "solve", ...) since the last reset, when enabled. */
EGTB_EXPORT void egtb_stats_enable(int on);
EGTB_EXPORT void egtb_stats_reset(void);
EGTB_EXPORT int egtb_stats_get(const char *kernel, uint64_t *launches, double *total_ms, double *alg_bytes);

#ifdef __cplusplus
}
#endif

#endif /* EGTB_H */
