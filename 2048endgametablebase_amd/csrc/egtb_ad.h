// Shared pieces of the advanced ("AD", large-tile masking) forward pass
// (egtb_ad.hip) and solver (egtb_ad_solve.hip): the masked pattern passed to
// every kernel, the per-device tables, the result handle and host helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "egtb.h"
#include "egtb_device.h"
#include "egtb_runtime.h"

namespace egtb {

constexpr int kCombStride = 11;    // length, then up to 10 tiles (log2, ascending)
constexpr int kCombEntries = 2560; // index (sum >> 6) + (count << 8), count < 10
constexpr int kAdThreads = 1024;
// bucket keys of the solver's books (bucket_key_min/max, FormationRuntime.h:145-151)
constexpr int kAdKeyMin = -16, kAdKeyMax = 31, kAdKeys = kAdKeyMax - kAdKeyMin + 1;

// The masked pattern (AdvancedPatternSpec + AdvancedMaskParam,
// FormationRuntime.h:88-104) by value in every launch.
struct AdSpecDev {
    uint64_t masks[kMaxMasks];
    uint64_t fixed_mask;  // 0xF at every pinned 32768 cell
    uint32_t n_masks;
    uint32_t symm;
    int32_t n_free, n_fixed;
    uint32_t limit;  // small_tile_sum_limit
    int32_t target;  // log2 of the target tile
};

__device__ __forceinline__ bool ad_is_pattern(uint64_t b, const AdSpecDev &s) {
    if (s.n_masks == 0) return true;
    bool ok = false;
    for (uint32_t i = 0; i < s.n_masks; ++i) ok |= (b & s.masks[i]) == s.masks[i];
    return ok;
}

// Number of masked (0xF) cells.  A masked tile never merges and a merge
// reaching 64 makes one, so a move made a new masked tile exactly when the
// moved board holds more of them -- the flag BoardMoverAD's row bitmap
// gives (BoardMoverAD.cpp:100-117).
__device__ __forceinline__ int f_count(uint64_t b) {
    return __popcll(b & (b >> 1) & (b >> 2) & (b >> 3) & 0x1111111111111111ull);
}

// Per-device tables, made on a device's first advanced pass and kept for the
// process's lifetime: the masked row table and the large-tile combination
// table (init_masker, BoardMaskerAD.cpp:514-533).
struct AdDevice {
    uint16_t *lut = nullptr;
    uint8_t *comb = nullptr;
};
AdDevice &ad_device(int dev);

// The pattern's device form, after checking its fields.
AdSpecDev make_ad_spec(const egtb_ad_pattern &pat);

// One solved bucket of a layer: boards (sorted) and rows x cols rates.
struct AdBucket {
    DBuf ind, rates;
    uint64_t rows = 0;
    uint32_t cols = 0;
    // range index over the bits below the boards' shared prefix: the boards
    // with window value v are ind[idx[v] .. idx[v+1]) (idx empty: none)
    DBuf idx;
    uint64_t pre_mask = 0, pre_val = 0;
    uint32_t shift = 0, bits = 0;
};
struct AdBookLayer {
    AdBucket b[kAdKeys];
    // the solver's device table of the 48 buckets (and its host copy, which
    // outlives the copy to the device)
    DBuf desc;
    std::vector<uint8_t> desc_host;
};

struct AdResult {
    int steps = 0;
    int dev = 0;
    std::vector<uint64_t> n;     // raw layer sizes, layers 0 .. steps-2
    std::vector<DBuf> layers;    // kept layers (device)
    double fwd_ms = 0;
    uint64_t candidates = 0;     // boards emitted over every step
    uint64_t chunked_steps = 0, chunks = 0;
    uint64_t skipped_steps = 0;  // steps found on disk (resume)
    uint64_t hash_steps = 0;     // steps deduplicated in key sets
    uint64_t set_overflows = 0;  // key-set steps redone by sort + unique
    bool started = false;        // a forward step ran (not all found on disk)
    // solver
    bool solved = false;
    int rate_size = 4;
    double solve_ms = 0;
    uint64_t solve_values = 0;    // rates computed (rows x cols, before pruning)
    uint64_t solve_skipped = 0;   // solved layers found on disk (resume)
    std::vector<AdBookLayer> books;  // kept solved layers (keep_books)
};

void ad_generate(const uint64_t *seeds, size_t n_seeds, const egtb_ad_pattern &pat, const egtb_ad_options &o,
                 AdResult &res);
void ad_solve(const uint64_t *seeds, size_t n_seeds, const egtb_ad_pattern &pat, const egtb_ad_options &o,
              const egtb_ad_solve_options &so, AdResult &res);

uint32_t seed_sum(uint64_t b);
uint64_t mask_board(uint64_t b);
bool file_exists(const std::string &p);
uint64_t file_size(const std::string &p);
// read_binary_vector_direct (FileIOUtils.h:190-215): a missing file reads as
// empty, a size that is not whole boards is an error
std::vector<uint64_t> read_raw(const std::string &p);
uint32_t grid_for(uint64_t n, uint32_t threads, uint32_t cap);

template <class F>
int guarded_ad(F &&f) {
    try {
        g_last_error.clear();
        f();
        return EGTB_OK;
    } catch (const Error &e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        g_last_error = "host out of memory";
        return EGTB_ERR_MEMORY;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return EGTB_ERR_INTERNAL;
    }
}

}  // namespace egtb
