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
    int32_t sort_schedule;           /* 0 = autotuned, 1 = per-digit 3 kernels, 2 = single-pass look-back */
    int32_t chunk_mkeys;             /* candidate budget per chunk in Mi keys; 0 = from free HBM */
    int32_t host_tables;             /* 1: stream finished layers to host memory, free them in HBM */
    int32_t write_raw_files;         /* 1: also write raw layers <pathname><i> (N_i x u64 LE) as the
                                        forward pass finishes them; each is removed once solved;
                                        2: write them and keep them */
    double deletion_threshold;       /* RunOptions::deletion_threshold: after layer i is solved, book
                                        i+2 is rewritten with rates > threshold only (0 = off) */
    int32_t solve_hash_mb;           /* successor lookups of the solve: hash tables up to this many MiB
                                        per layer, bucket-indexed binary search above; 0 = default,
                                        -1 = always search */
    int32_t forward_hash_steps;      /* forward dedup: 0 = key sets (hashing) while they fit free HBM, then
                                        sort + unique; -1 = sort + unique at every step; k > 0 = key sets
                                        for steps below k only */
    int32_t solve_hash_layout;       /* successor hashes of the solve: 0 = 16-byte slots; 1 = compact
                                        8-byte slots in 64-byte buckets where the layer's key span allows
                                        (4-byte rates), else 16-byte slots */
    int32_t optimal_branch_only;     /* RunOptions::optimal_branch_only: after the solve, book i (i > 20)
                                        keeps only the records some state of layers i-1 / i-2 reaches as its
                                        best move (BookSolver.cpp:611-735) */
    int32_t compress;                /* RunOptions::compress: at the end of a build with files, every book
                                        above 2 MiB except layers 0-2 is replaced by its compressed form
                                        <pathname><i>.z (CompressionBridge.cpp:41-60) */
    int32_t own_streams;             /* 1: the build runs on two streams of its own, kept with its result, so
                                        builds from several host threads overlap on the device; 0: the
                                        device's shared pair (builds queue behind each other) */
} egtb_options;

/* Build a tablebase end to end on the GPU and write the reference's file
 * set (<pathname><i>.book for every layer, <pathname>stats.txt).
 * Resumes like the reference (BookGenerator.cpp:384-430,
 * BookSolver.cpp:476-511): empty or misaligned .book files are removed,
 * solved layers whose .book exists are skipped from the top down, the first
 * layer solved loads books i+1 and i+2 from their files, and leftover raw
 * layer files are removed as their layers are solved.  The forward pass is
 * recomputed on the device rather than read back from raw files.
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
/* Large-build accounting: peak HBM in use during the backward pass (host
 * tables mode), bytes and time of the layers streamed to the host, forward
 * steps that ran chunked and their chunk count. */
EGTB_EXPORT int egtb_result_memory(const egtb_result *r, uint64_t *peak_device_bytes, uint64_t *streamed_bytes,
                                   double *stream_ms, uint64_t *chunked_steps, uint64_t *chunks);
/* Forward dedup of the last build: steps whose candidates went into key
 * sets (the rest were sorted and uniqued) and how often a set was regrown
 * and its steps run again. */
EGTB_EXPORT int egtb_result_forward(const egtb_result *r, uint64_t *hash_steps, uint64_t *set_regrows);
EGTB_EXPORT void egtb_result_free(egtb_result *r);

/* ---- advanced (AD, large-tile masking) forward pass ----------------------- */

/* Pattern of an advanced build.  Replaces AdvancedPatternSpec
 * (reference native_core/include/FormationRuntime.h:88-96) as
 * engine_core/BookBuilder.py:90-108 fills it: the pattern's masks and
 * canonical mode, the 32768 tiles of the first seed that no mask pins
 * (n_free_32k), the nibble shifts of the pinned ones, SmallTileSumLimit and
 * the target (log2). */
typedef struct egtb_ad_pattern {
    const uint64_t *masks;
    uint32_t n_masks;                /* <= 16 */
    int32_t symm_mode;               /* 0 identity, 1 full, 2 diagonal, 3 horizontal */
    int32_t n_free_32k;
    const uint8_t *fixed_32k_shifts; /* nibble shifts 0, 4, ..., 60 */
    uint32_t n_fixed;                /* <= 16 */
    uint32_t small_tile_sum_limit;   /* reference default 96 */
    int32_t target;
} egtb_ad_pattern;

typedef struct egtb_ad_options {
    int32_t steps;           /* layers of the build (steps-1 raw layers are made) */
    const char *pathname;    /* NULL, or the prefix of raw layer files <pathname><i> */
    int32_t keep_layers;     /* 1: keep every raw layer in HBM for egtb_ad_result_layer */
    int32_t device;          /* HIP device ordinal */
    int32_t radix_bits;      /* 0 = autotuned; else 4, 6, 8 or 11 */
    int32_t sort_schedule;   /* 0 = autotuned, 1 = per-digit 3 kernels, 2 = single-pass look-back */
    int32_t chunk_mkeys;     /* boards per parent range in Mi keys (sort + unique steps); 0 = from free HBM */
    int32_t forward_dedup;   /* 0: key sets (up to 2^25 slots), sort + unique where a set would be larger or
                                overflowed; 1: sort + unique at every step; 2: key sets at 1/64 of their size
                                (exercises the overflow path) */
    int32_t reserved[2];
} egtb_ad_options;

typedef struct egtb_ad_result egtb_ad_result;

/* The advanced builder's forward pass on the GPU: raw layers 0 .. steps-2
 * of masked boards (layer 0 = the masked seeds as given, layer i sorted and
 * unique), validated every 32 steps, written as the reference writes them
 * when pathname is set.  Replaces generate_process_ad
 * (reference native_core/src/BookGeneratorAD.cpp:1256-1396, called from
 * run_pattern_build_ad_cpp :1421-1438, bound as
 * formation_core.run_pattern_build_ad, bindings_formation.cpp:231-235).  With
 * files it resumes as handle_restart_ad (:391-420) does: steps whose raw
 * layers (or books) are on disk are skipped and the first step run reads
 * layers i-1 and i back.  out may be NULL. */
EGTB_EXPORT int egtb_ad_generate(const uint64_t *seeds, size_t n_seeds, const egtb_ad_pattern *pattern,
                                 const egtb_ad_options *options, egtb_ad_result **out);
/* steps, the steps-1 layer sizes (sizes may be NULL), wall time, boards
 * emitted before dedup, the steps that ran over parent ranges, and the
 * steps found on disk and skipped. */
EGTB_EXPORT int egtb_ad_result_info(const egtb_ad_result *r, int *steps, uint64_t *sizes, double *fwd_ms,
                                    uint64_t *candidates, uint64_t *chunked_steps, uint64_t *skipped_steps);
/* Forward dedup of the pass: steps deduplicated in key sets, and key-set
 * steps whose sets overflowed and ran again through sort + unique. */
EGTB_EXPORT int egtb_ad_result_forward(const egtb_ad_result *r, uint64_t *hash_steps, uint64_t *set_overflows);
/* Copy raw layer `layer` to host (needs options.keep_layers; keys may be
 * NULL to query n). */
EGTB_EXPORT int egtb_ad_result_layer(const egtb_ad_result *r, int layer, uint64_t *keys, size_t cap, size_t *n);
EGTB_EXPORT void egtb_ad_result_free(egtb_ad_result *r);

/* Solver options of an advanced build (the RunOptions fields
 * run_pattern_solve_ad_cpp reads, reference native_core/src/BookSolverAD.cpp:
 * 2086-2284). */
typedef struct egtb_ad_solve_options {
    const char *success_rate_dtype; /* "float32" (NULL), "float64", "uint32" or "uint64" */
    double spawn_rate4;             /* probability of a 4 (reference default 0.1) */
    double deletion_threshold;      /* > 0: layer i+2's rows with no rate above it are dropped (and rewritten) */
    int32_t keep_books;             /* 1: keep every solved layer's buckets in HBM (egtb_ad_result_book) */
    int32_t reserved[3];
} egtb_ad_solve_options;

/* The advanced build: the forward pass (as egtb_ad_generate, every raw layer
 * kept in HBM) and then the solver, layers steps-3 .. 0.  Layer i's boards
 * are split into buckets by their large tiles (expand_ad, BookSolverAD.cpp:
 * 1243-1318); a board's row holds one rate per placement of its large tiles
 * (derive_size_for_bucket, :291-299, columns in the order of the permutation
 * table, BoardMaskerAD.cpp:291-358), computed from layers i+1 and i+2
 * (recalculate_ad, :1321-1505); rows with no rate above zero are dropped
 * (remove_died_ad, :495-536).  With options.pathname it writes the reference's
 * files: <pathname><i>b/<key>.i (sorted boards) and <key>.b (rates) per
 * bucket (dict_tofile, :320-352), stats.txt lines, and removes each raw
 * layer once solved; a layer whose <i>b folder exists is skipped and the
 * layers it needs are read back (recalculate_process_ad_impl, :2086-2242).
 * Replaces run_pattern_build_ad_cpp (BookGeneratorAD.cpp:1421-1438, bound as
 * formation_core.run_pattern_build_ad, bindings_formation.cpp:231-235).
 * out may be NULL. */
EGTB_EXPORT int egtb_ad_build(const uint64_t *seeds, size_t n_seeds, const egtb_ad_pattern *pattern,
                              const egtb_ad_options *options, const egtb_ad_solve_options *solve,
                              egtb_ad_result **out);
/* Solver figures: wall time of the solve, rates computed (rows x columns,
 * before dropped rows) and layers found solved on disk. */
EGTB_EXPORT int egtb_ad_result_solve_info(const egtb_ad_result *r, double *solve_ms, uint64_t *values,
                                          uint64_t *skipped_layers);
/* Bucket `key` (-16 .. 31) of solved layer `layer` (needs
 * solve.keep_books): its rows (boards, sorted) and cols; boards (rows) and
 * rates (rows x cols of the rate type) are copied when not NULL. */
EGTB_EXPORT int egtb_ad_result_book(const egtb_ad_result *r, int layer, int key, uint64_t *boards, void *rates,
                                    size_t cap_rows, size_t *rows, int *cols);

/* ---- batched builds: many (pattern, target) jobs through one pipeline ------ */

/* One job of a batched build: the arguments of egtb_build for one table.
 * Every mask of the pattern must pin nibbles 0 and 1 (all catalogue
 * patterns but the free* ones do); rates are float32. */
typedef struct egtb_batch_job {
    const uint64_t *seeds;
    size_t n_seeds;
    egtb_pattern pattern;
    int32_t target;
    int32_t steps;
    int32_t docheck_step;
    int32_t reserved0;
    double spawn_rate4;
} egtb_batch_job;

typedef struct egtb_batch_options {
    int32_t device;
    int32_t radix_bits;          /* 0 = autotuned */
    int32_t sort_schedule;       /* 0 = autotuned */
    int32_t keep_raw;
    int32_t forward_hash_steps;  /* as egtb_options::forward_hash_steps */
    int32_t reserved[3];
} egtb_batch_options;

typedef struct egtb_batch_result egtb_batch_result;

/* Build every job through the plain build's schedule over job keys
 * ((board >> 8) | job << 56): one set of kernel launches per forward step
 * and per backward layer covers all jobs.  Each job's layers equal its
 * egtb_build layers.  No reference counterpart (the reference builds one
 * table per run_pattern_build call). */
EGTB_EXPORT int egtb_batch_build(const egtb_batch_job *jobs, int n_jobs, const egtb_batch_options *options,
                                 egtb_batch_result **out);
EGTB_EXPORT int egtb_batch_timing(const egtb_batch_result *b, double *fwd_ms, double *bwd_ms, double *total_ms);
/* Sizes of layer `layer` summed over all jobs. */
EGTB_EXPORT int egtb_batch_layer_info(const egtb_batch_result *b, int layer, uint64_t *n_raw_all,
                                      uint64_t *n_live_all);
/* Solved layer `layer` of job `job`: keys (plain encoding) and float32 rates. */
EGTB_EXPORT int egtb_batch_book(const egtb_batch_result *b, int job, int layer, uint64_t *keys, float *rates,
                                size_t cap, size_t *n);
EGTB_EXPORT void egtb_batch_free(egtb_batch_result *b);

/* ---- solved tables resident in HBM and batched lookups ------------------- */

typedef struct egtb_table egtb_table;

#define EGTB_LOOKUP_OK 0        /* answered from the table */
#define EGTB_LOOKUP_FALLBACK 1  /* off-table: depth-2 expectimax fallback values */
#define EGTB_LOOKUP_ERROR (-1)  /* serving only: the query's batch could not run (its table was
                                   detached, or a HIP call failed); rates are NaN */

/* Load every <pathname><i>.book (i < steps) of a build into HBM.  A missing
 * layer file reads as off-table, an empty one as all-zero rates, as in the
 * reference's reader (ReaderRuntime.cpp:739-776).  tile_sum is the seed tile
 * sum (minus the catalogue's nums_adjust).  Replaces the per-lookup on-disk
 * binary search of find_success_entry_in_file (ReaderRuntime.cpp:279-316). */
EGTB_EXPORT int egtb_table_open(const char *pathname, int steps, const egtb_pattern *pattern, int target,
                                int64_t tile_sum, const char *success_rate_dtype, double spawn_rate4,
                                int device, egtb_table **out);
/* Serve the solved layers of an in-memory build (no copy; the result must
 * outlive the table). */
EGTB_EXPORT int egtb_table_from_result(const egtb_result *result, const egtb_pattern *pattern, int target,
                                       int64_t tile_sum, double spawn_rate4, egtb_table **out);
EGTB_EXPORT int egtb_table_info(const egtb_table *t, int *steps, uint64_t *records, uint64_t *bytes);
/* For each board: four directional rates in (left, right, up, down) order,
 * NaN for an invalid move, normalised to [0, 1] as the reference reader
 * does (ReaderRuntime.cpp:271-277) -- i.e. ClassicBookReader::move_on_dic
 * (ReaderRuntime.cpp:739-840) with a stateless symmetry order 0..7.
 * status[i] is EGTB_LOOKUP_OK or EGTB_LOOKUP_FALLBACK.  Host buffers. */
EGTB_EXPORT int egtb_lookup_batch(egtb_table *t, const uint64_t *boards, size_t n, float *out4, int32_t *status);
/* As egtb_lookup_batch, but allocates nothing: no hipMalloc, hipFree or
 * allocating library call.  No counterpart in the reference; for callers
 * that cannot allocate device memory (tight serving loops, harnesses that
 * forbid it).  The device staging of boards, rates and status lives in a
 * file-scope device array of 28 MiB per GPU, held until the process ends:
 * at most 2^20 boards per call (more: EGTB_ERR_ARG).  Calls on one GPU share
 * that array and would race; the library serialises them with a per-device
 * lock.  Same answers as egtb_lookup_batch.  Host buffers. */
EGTB_EXPORT int egtb_lookup_batch_noalloc(egtb_table *t, const uint64_t *boards, size_t n, float *out4,
                                          int32_t *status);
/* Same in float64, also returning the symmetry op that answered (-1 none). */
EGTB_EXPORT int egtb_lookup_batch_f64(egtb_table *t, const uint64_t *boards, size_t n, double *out4,
                                      int32_t *status, int32_t *op_used);
EGTB_EXPORT void egtb_table_free(egtb_table *t);

/* Answer from the symmetry op whose best rate is largest instead of the
 * first op that answers; a board whose rates are all 0 then counts as not
 * answered (status EGTB_LOOKUP_FALLBACK).  The reference reader does this
 * for 4442f, 4442ff and 4tiler (prefer_max_result_,
 * ReaderRuntime.cpp:827-832,1053). */
EGTB_EXPORT int egtb_table_set_prefer_max(egtb_table *t, int32_t on);

/* Moves and canonical forms of n boards on the device.  out4 (optional,
 * n x 4): the boards after left, right, up, down, an impossible move giving
 * the board unchanged (MoverEngine::move_all_dir, CommonMover.h:136-139).
 * canon (optional, n): each board's canonical form under symm_mode 0..3
 * (Calculator.h:80-98,192-203).  Replaces mover_runtime.std.move_all_dir and
 * mover_runtime.canonical_* as used by start_build's free-pattern seeds
 * (engine_core/BookBuilder.py:146-181). */
EGTB_EXPORT int egtb_board_moves(const uint64_t *boards, size_t n, int32_t symm_mode, uint64_t *out4,
                                 uint64_t *canon, int device);

/* ---- heuristic expectimax player (the dispatcher's AI) ------------------- */

/* AIPlayer(board).start_search(depth) for n boards at once on the device
 * (reference native_core/src/AIPlayer.cpp:893-941, bound as
 * ai_core.AIPlayer in bindings_ai.cpp:20-56 and called by the AI dispatcher,
 * engine_core/AIPlayer.py:575-1010): apply_dynamic_mask, then
 * search_ai_player(masked board, depth), then the search again on the
 * unmasked board when the reference's condition holds.  depth 1..3;
 * do_check / prune / spawn_rate4 as the player's members of those names.
 * best (n): the best operation, 1..4 = left, right, up, down, 0 when no
 * move is valid; scores (n x 4): top_scores, -dead_score for invalid moves.
 * Equal to the reference's single-thread search up to depth 2 (its value
 * cache matters there only on a signature collision); at depth 3 each board
 * is searched at its own depth, where the reference may reuse a deeper
 * cached value. */
EGTB_EXPORT int egtb_ai_search(const uint64_t *boards, size_t n, int depth, int do_check, int prune,
                               double spawn_rate4, int device, int32_t *best, int32_t *scores);
/* AIPlayer::evaluate (AIPlayer.cpp:645-667) of n boards on the device, with
 * table variant `variant` (see egtb_ai_table), do_check and the masked count
 * that selects check_corner's shapes (AIPlayer.cpp:470-643). */
EGTB_EXPORT int egtb_ai_evaluate(const uint64_t *boards, size_t n, int variant, int do_check, int masked_count,
                                 int device, int32_t *out);
/* The player's evaluation table (host only, no device): variant 0 is
 * init_evaluate_tables' table (AIPlayer.cpp:103-127); variants 1..15 carry
 * the adjustments update_specific_scores makes (AIPlayer.cpp:254-336) for
 * flags (bit 0 flag1 .. bit 3 flag4) when the 2x2 corner shape is present.
 * out: 65536 x {d1, d2}. */
EGTB_EXPORT int egtb_ai_table(int16_t *out, int variant);

/* ---- serving: continuous batching over resident tables -------------------- */

/* A request loop that keeps tables in HBM and answers position queries.
 * A worker owns `slots` batch slots (own stream + pinned buffers each);
 * whenever a slot is free and queries wait, it takes up to max_batch of
 * them (oldest table first) and launches the lookup, so queries arriving
 * while other batches are on the GPU are admitted at once and batches
 * overlap on the device.  Tables are admitted against memory_budget bytes
 * (0 = unlimited).  Replaces the reference's per-call on-disk reader for
 * serving use (ReaderRuntime.cpp:1055-1068); no reference counterpart for
 * the batching itself. */
typedef struct egtb_server egtb_server;
typedef struct egtb_server_options {
    int32_t device;
    int32_t slots;       /* batches in flight (default 4) */
    uint32_t max_batch;  /* queries per batch (default 4096) */
    int32_t reserved0;
    int64_t memory_budget;
    int32_t reserved[4];
} egtb_server_options;

EGTB_EXPORT int egtb_server_create(const egtb_server_options *options, egtb_server **out);
/* Admit a table (fails with EGTB_ERR_MEMORY past the budget). */
EGTB_EXPORT int egtb_server_attach(egtb_server *s, egtb_table *table, int32_t *table_id);
/* Release a table once its queued and in-flight queries are done. */
EGTB_EXPORT int egtb_server_detach(egtb_server *s, int32_t table_id);
/* Queue n boards for table_id; tickets[i] identifies each answer. */
EGTB_EXPORT int egtb_server_submit(egtb_server *s, int32_t table_id, const uint64_t *boards, size_t n,
                                   uint64_t *tickets);
/* Take up to max finished answers (4 floats + status each, as
 * egtb_lookup_batch, or EGTB_LOOKUP_ERROR), waiting up to wait_us for the
 * first.  Once the worker has stopped on a failure, submit fails and poll
 * fails when no answer is left. */
EGTB_EXPORT int egtb_server_poll(egtb_server *s, size_t max, uint64_t *tickets, float *out4, int32_t *status,
                                 size_t *n_out, int32_t wait_us);
EGTB_EXPORT int egtb_server_stats(const egtb_server *s, uint64_t *batches, uint64_t *queries, uint64_t *max_batch);
EGTB_EXPORT void egtb_server_destroy(egtb_server *s);

/* Poisson load: queries at rate qps for duration_s seconds, each board
 * drawn from pool[] by a seeded uniform index (n_games is the number of
 * concurrent game sources the rate is split over, for reporting).  The
 * generator owns the server's completion queue while it runs. */
typedef struct egtb_loadgen_options {
    double qps;
    double duration_s;
    uint64_t seed;
    int32_t n_games;
    int32_t table_id;
    const uint64_t *pool;
    size_t pool_size;
} egtb_loadgen_options;

typedef struct egtb_loadgen_report {
    double qps_offered, qps_achieved;
    double p50_ms, p90_ms, p99_ms, p999_ms, max_ms, mean_ms;
    double submit_seconds;
    uint64_t completed, batches, max_batch, fallbacks;
} egtb_loadgen_report;

EGTB_EXPORT int egtb_server_loadgen(egtb_server *s, const egtb_loadgen_options *options, egtb_loadgen_report *report);

/* ---- compressed books: the reference's .z layout ------------------------- */

/* Records of the compressed form of a book: book_path is the .book path
 * (<pathname><i>.book); its compressed form is the directory <pathname><i>.z
 * holding the trie index (<stem>i, <stem>ii), the block table (<stem>s) and
 * the xz blocks (<stem>z).  With keys and rates NULL only *n is set; else
 * up to cap records are copied (keys ascending, rates of rate_size 4 or 8
 * bytes, raw as stored).  Host work, no device.  Replaces reading a .z
 * layer per probe (trie_search_with_context, ReaderRuntime.cpp:318-476);
 * egtb_table_open reads compressed layers through it. */
EGTB_EXPORT int egtb_zbook_read(const char *book_path, int rate_size, uint64_t *keys, void *rates, size_t cap,
                                size_t *n);
/* Writes the compressed form of a book of n strictly ascending records (the
 * .book itself is left alone).  Same layout as trie_compress_progress_native
 * (TrieCompression.cpp:124-551): blocks of about 32768 records, xz preset 1
 * with CRC64, compressed on host threads.  Host work, no device. */
EGTB_EXPORT int egtb_zbook_write(const char *book_path, const uint64_t *keys, const void *rates, size_t n,
                                 int rate_size);

/* ---- dedup primitives: drop-in for the reference's bookgen_native.so ----- */

/* Ascending (or descending) sort of a host array on the GPU.  Same name and
 * signature as the reference plugin's export (native_core/xss_wrapper.cpp:
 * 14-31), resolved by BookGeneratorUtils.cpp:68-94. */
EGTB_EXPORT void sort_uint64(uint64_t *arr, size_t arrsize, bool descending);

/* Adjacent dedup of a sorted host array in place; returns the new length.
 * Same name and signature as native_core/src/unique_wrapper.cpp:12-14. */
EGTB_EXPORT size_t unique_sorted_u64_inplace(uint64_t *data, size_t size);

/* ---- tuning ---------------------------------------------------------------- */

/* Sorted unique copy of keys[0..n) into out (room for n), count in *n_out,
 * with the dedup the builder uses for its layers: schedule 0 = tuned,
 * 1/2 = radix sort (per-digit / look-back) + unique, 3 = sample sort with
 * per-bucket sort and unique in LDS (n <= 2^21; larger n use radix).
 * Replaces sort_array + parallel_unique (BookGeneratorUtils.cpp:138-206). */
EGTB_EXPORT int egtb_sort_unique_u64(const uint64_t *keys, size_t n, int schedule, uint64_t *out, size_t *n_out);

/* Install the sort tuning table used when options ask for autotuned values
 * (radix_bits = 0, sort_schedule = 0): a sort of n keys uses the first row
 * with n <= max_n[i] (rows ascending; the last row also covers larger n).
 * rows = 0 restores the built-in defaults.  No counterpart in the reference
 * (its sort has no tunables); written by the package's autotuner. */
EGTB_EXPORT int egtb_set_sort_tuning(const uint64_t *max_n, const int32_t *bits, const int32_t *schedule,
                                     int rows);

/* Install a tile tuning table: kind 0 = keys per thread of the per-digit
 * radix schedule's tiles (4, 8 or 16), kind 1 = threads per block of the
 * lookup kernel (64..1024, whole wavefronts).  A launch over n items uses
 * the first row with n <= max_n[i] (rows ascending; the last row also covers
 * larger n); a value of 0, or rows = 0, keeps the built-in rule.  No
 * counterpart in the reference; written by the package's autotuner. */
EGTB_EXPORT int egtb_set_tile_tuning(int kind, const uint64_t *max_n, const int32_t *value, int rows);

/* Sizing of the forward pass's key sets.  slots_per_key: slots per
 * predicted key (default 2: a set is at most about half full when the
 * growth estimate holds; valid 0 < v <= 64).  max_slots: the largest set
 * (default 2^25 slots, 256 MiB: random inserts pay while the sets being
 * filled stay near the 256 MiB last-level cache); a step that would need a larger one
 * and the rest of the forward pass sort and unique their candidates instead.
 * 0 restores a default.  A set that fills up is regrown and its steps run
 * again, so small slots_per_key values only cost time.  No counterpart in
 * the reference (its forward pass sorts every step). */
EGTB_EXPORT int egtb_set_key_set_sizing(double slots_per_key, uint64_t max_slots);

/* Time sort + unique of n host keys on device 0 with a given digit width
 * (4, 6, 8, 11) and schedule (1 per-digit kernels, 2 look-back); average
 * milliseconds over `reps` runs after one warm-up.  For the autotuner. */
EGTB_EXPORT int egtb_bench_sort(const uint64_t *keys, size_t n, int bits, int schedule, int reps,
                                double *ms_per_sort);

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
