// Test infrastructure only: a C-ABI shim around the reference's advanced
// ("AD", large-tile masking) builder, run_pattern_build_ad_cpp
// (native_core/src/BookGeneratorAD.cpp:1421-1438), compiled from the
// reference's sources by oracle/ref/Makefile into oracle/_ref/libegtb_ref.so.
// Loaded by tests/golden scripts, never by the product.
//
// The shim fills AdvancedPatternSpec / RunOptions
// (native_core/include/FormationRuntime.h:24-42,88-96) the way
// engine_core/BookBuilder.py:90-108,211-236 does.  With stop_after_forward
// set it leaves the forward pass's raw layer files in place: a regular file
// at <pathname><steps-1>b makes the solver's first create_directories
// (BookSolverAD.cpp:2113-2114) fail, so the build stops right after
// generate_process_ad wrote raw layers 0 .. steps-2, before the solver reads
// and deletes them.

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <exception>
#include <filesystem>
#include <string>
#include <vector>

#include "BookSolver.h"
#include "FormationRuntime.h"

extern "C" int egtb_ref_build_ad(const uint64_t *seeds, size_t n_seeds, const uint64_t *masks, size_t n_masks,
                                 int symm_mode, int num_free_32k, const uint8_t *fixed_shifts, size_t n_fixed,
                                 int small_tile_sum_limit, int target, int steps, int docheck_step,
                                 const char *pathname, int is_free, double spawn_rate4, const char *dtype,
                                 int num_threads, int stop_after_forward, double deletion_threshold,
                                 char *err, size_t err_len) {
    const std::string blocker = std::string(pathname) + std::to_string(steps - 1) + "b";
    try {
        AdvancedPatternSpec spec;
        spec.name = "ref";
        spec.pattern_masks.assign(masks, masks + n_masks);
        spec.symm_mode = symm_mode;
        spec.num_free_32k = (uint8_t)num_free_32k;
        spec.fixed_32k_shifts.assign(fixed_shifts, fixed_shifts + n_fixed);
        spec.small_tile_sum_limit = (uint32_t)small_tile_sum_limit;
        spec.target = (uint8_t)target;

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
        opt.compress = false;
        opt.compress_temp_files = false;
        opt.optimal_branch_only = false;
        opt.chunked_solve = false;
        opt.num_threads = num_threads;
        opt.direct_io = false;

        if (stop_after_forward) {
            FILE *f = std::fopen(blocker.c_str(), "wb");
            if (!f) throw std::runtime_error("cannot create the forward-pass stop file");
            std::fclose(f);
        }
        std::vector<uint64_t> init(seeds, seeds + n_seeds);
        try {
            run_pattern_build_ad_cpp(init, spec, opt);
        } catch (const std::filesystem::filesystem_error &e) {
            // only the stop file's own error ends the run quietly
            if (!stop_after_forward || e.path1() != std::filesystem::path(blocker)) throw;
        }
        if (stop_after_forward) std::remove(blocker.c_str());
        return 0;
    } catch (const std::exception &e) {
        if (stop_after_forward) std::remove(blocker.c_str());
        if (err && err_len) {
            std::strncpy(err, e.what(), err_len - 1);
            err[err_len - 1] = '\0';
        }
        return -1;
    }
}
