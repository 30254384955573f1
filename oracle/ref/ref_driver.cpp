// Test infrastructure only: a C-ABI shim around the reference's own classic
// builder (`run_pattern_build_cpp`, native_core/src/BookSolver.cpp:823-832),
// compiled from the reference's sources by oracle/ref/Makefile into
// oracle/_ref/libegtb_ref.so.  It is loaded by tests (golden-vector
// generation) and by bench.py's CPU-baseline leg, never by the product.
//
// The shim fills the reference's PatternSpec / RunOptions
// (native_core/include/FormationRuntime.h:17-42) the way
// engine_core/BookBuilder.py:44-92 does, and turns C++ exceptions into a
// return code plus message.  It must be loaded from a process with an
// initialised Python interpreter (e.g. through ctypes), because the
// reference's debug_log (BookGenerator.cpp:45-51) acquires the GIL.

#include <cstdint>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "BookSolver.h"
#include "FormationRuntime.h"

// deletion_threshold, optimal_branch_only and compress are the RunOptions
// fields of the same names (FormationRuntime.h:24-42): the post-solve
// pruning of BookSolver.cpp:589-593, the optimal-branch pass :611-735 and the
// .z compression of finished books (CompressionBridge.cpp:41-60).
extern "C" int egtb_ref_build_ex(const uint64_t *seeds, size_t n_seeds,
                                 const uint64_t *masks, size_t n_masks,
                                 const uint8_t *shifts, size_t n_shifts,
                                 int symm_mode, int target, int steps,
                                 int docheck_step, const char *pathname,
                                 int is_free, double spawn_rate4,
                                 const char *dtype, int num_threads,
                                 double deletion_threshold,
                                 int optimal_branch_only, int compress,
                                 char *err, size_t err_len) {
    try {
        PatternSpec spec;
        spec.name = "ref";
        spec.pattern_masks.assign(masks, masks + n_masks);
        spec.success_shifts.assign(shifts, shifts + n_shifts);
        spec.symm_mode = symm_mode;

        RunOptions opt;
        opt.target = target;
        opt.steps = steps;
        opt.docheck_step = docheck_step;
        opt.pathname = pathname;
        opt.is_free = is_free != 0;
        opt.is_variant = false;
        opt.spawn_rate4 = spawn_rate4;
        opt.success_rate_dtype = dtype;
        opt.deletion_threshold = deletion_threshold;
        opt.compress = compress != 0;
        opt.compress_temp_files = false;
        opt.optimal_branch_only = optimal_branch_only != 0;
        opt.chunked_solve = false;
        opt.num_threads = num_threads;
        opt.direct_io = false;

        std::vector<uint64_t> init(seeds, seeds + n_seeds);
        run_pattern_build_cpp(init, spec, opt);
        return 0;
    } catch (const std::exception &e) {
        if (err && err_len) {
            std::strncpy(err, e.what(), err_len - 1);
            err[err_len - 1] = '\0';
        }
        return -1;
    } catch (...) {
        if (err && err_len) {
            std::strncpy(err, "unknown exception", err_len - 1);
            err[err_len - 1] = '\0';
        }
        return -2;
    }
}

extern "C" int egtb_ref_build(const uint64_t *seeds, size_t n_seeds,
                              const uint64_t *masks, size_t n_masks,
                              const uint8_t *shifts, size_t n_shifts,
                              int symm_mode, int target, int steps,
                              int docheck_step, const char *pathname,
                              int is_free, double spawn_rate4,
                              const char *dtype, int num_threads, char *err,
                              size_t err_len) {
    return egtb_ref_build_ex(seeds, n_seeds, masks, n_masks, shifts, n_shifts,
                             symm_mode, target, steps, docheck_step, pathname,
                             is_free, spawn_rate4, dtype, num_threads, 0.0, 0,
                             0, err, err_len);
}
