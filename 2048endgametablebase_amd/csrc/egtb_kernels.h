// Host-visible declarations of the kernel launch wrappers.
#pragma once

#include <utility>
#include <vector>

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "egtb_device.h"

namespace egtb {

// A sorted, zero-compacted layer as the solve kernel sees it: SoA keys and
// rates plus a bucket index idx[0..2^bits] over the top `bits` key bits.
template <typename T>
struct TableDev {
    const uint64_t *keys;
    const T *rates;
    const uint32_t *idx;
    const uint64_t *n_ptr;  // live size in device memory (read when non-null)
    uint64_t n;             // live size when n_ptr is null
    uint32_t bits;
    uint32_t present;       // 1 when the layer exists (reader: its file exists)
};

// Exact-match lookup of `key` in a sorted table of n records: true and the
// rate when present (reference search_arr, BookSolver.cpp:317-327, whose
// miss reads as the rate type's zero value).
template <typename T>
__device__ __forceinline__ bool table_find(const TableDev<T> &t, uint64_t n, uint64_t key, T &rate) {
    if (n == 0) return false;
    const uint64_t bk = key >> (64 - t.bits);
    uint64_t lo = t.idx[bk], hi = t.idx[bk + 1];
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (t.keys[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    if (lo < n && t.keys[lo] == key) {
        rate = t.rates[lo];
        return true;
    }
    return false;
}

// As table_find, also giving the record's position.
template <typename T>
__device__ __forceinline__ bool table_find_pos(const TableDev<T> &t, uint64_t n, uint64_t key, T &rate,
                                               uint64_t &pos) {
    if (n == 0) return false;
    const uint64_t bk = key >> (64 - t.bits);
    uint64_t lo = t.idx[bk], hi = t.idx[bk + 1];
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (t.keys[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    if (lo < n && t.keys[lo] == key) {
        rate = t.rates[lo];
        pos = lo;
        return true;
    }
    return false;
}

// Order-preserving unsigned image of a rate (floats of either sign order as
// these keys; unsigned rates are their own), for the max-rate statistic.
template <typename T>
__host__ __device__ __forceinline__ uint64_t rate_order_key(T v) {
    if constexpr (sizeof(T) == 4) {
        const uint32_t b = __builtin_bit_cast(uint32_t, v);
        if constexpr (T(-1) < T(0)) return (b >> 31) ? ~b : (b | 0x80000000u);
        else return b;
    } else {
        const uint64_t b = __builtin_bit_cast(uint64_t, v);
        if constexpr (T(-1) < T(0)) return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
        else return b;
    }
}

template <typename T>
__host__ __device__ __forceinline__ T rate_from_order_key(uint64_t k) {
    if constexpr (sizeof(T) == 4) {
        uint32_t b = (uint32_t)k;
        if constexpr (T(-1) < T(0)) b = (b >> 31) ? (b & 0x7fffffffu) : ~b;
        return __builtin_bit_cast(T, b);
    } else {
        uint64_t b = k;
        if constexpr (T(-1) < T(0)) b = (b >> 63) ? (b & 0x7fffffffffffffffull) : ~b;
        return __builtin_bit_cast(T, b);
    }
}

// Hash of one solved layer, the structure the backward solve looks
// successors up in.  Two layouts:
//
// wide (slots): 16-byte slots {key, rate bits}, key 0 marks an empty slot
//   (every stored board holds tiles, so no key is 0), linear probing from
//   hash_mix(key); the capacity is a power of two at least four times the
//   layer's size bound, so every probe run ends at an empty slot.
//
// compact (cslots, 4-byte rates only): 8-byte slots in buckets of 8 (64
//   bytes).  A layer's keys agree outside one span of s bits (the varying
//   bits of the layer, from the OR / AND of its keys), so a key is its span
//   value x plus the constant bits; a key with other constant bits is not in
//   the layer.  x goes through a bijective mix on s bits; the top cb bits
//   pick the home bucket and the low s - cb <= 26 bits are stored.  A slot
//   holds [valid:1][displacement:5][stored:26][rate bits:32], 0 = empty; a
//   record sits in bucket home + d (d < 32) with d in its slot, so (bucket,
//   d, stored) names the key exactly.  Buckets fill front to back and slots
//   only go from empty to full, so a lookup stops at the first empty slot.
//   Tables are 4-8x smaller than the wide layout, which keeps the solve's
//   successor tables in the caches.
struct HashDev {
    const ulonglong2 *slots = nullptr;  // wide layout
    uint64_t mask = 0;                  // wide: capacity - 1; compact: buckets - 1
    const uint64_t *cslots = nullptr;   // compact layout
    uint64_t cconst = 0;                // compact: the key bits outside the span
    uint64_t cspan = 0;                 // compact: 2^s - 1
    uint32_t clo = 0, cs = 0, cb = 0;   // compact: span shift, span bits s, bucket bits
    uint32_t *err = nullptr;            // fills: bit 0 set when a compact insert found no slot
    __host__ __device__ bool present() const { return slots != nullptr || cslots != nullptr; }
    __host__ __device__ uint64_t bytes() const {
        return cslots ? (mask + 1) * 64 : slots ? (mask + 1) * 16 : 0;
    }
    __host__ __device__ void *table() const { return cslots ? (void *)cslots : (void *)slots; }
};

constexpr uint32_t kCompactStoredBits = 26;
constexpr uint32_t kCompactMaxDisp = 32;

// Home bucket and stored bits of `key` in a compact hash; false when the
// key's constant bits differ from the layer's (then it is not in the layer).
__device__ __forceinline__ bool chash_split(const HashDev &h, uint64_t key, uint64_t &bucket, uint32_t &stored) {
    if ((key & ~(h.cspan << h.clo)) != h.cconst) return false;
    uint64_t x = (key >> h.clo) & h.cspan;
    if (h.cs) {
        // odd multipliers and right xorshifts are bijections on s bits
        const uint32_t r = (h.cs + 1) >> 1;
        x = (x * 0x9e3779b97f4a7c15ull) & h.cspan;
        x ^= x >> r;
        x = (x * 0xbf58476d1ce4e5b9ull) & h.cspan;
        x ^= x >> r;
    }
    const uint32_t sb = h.cs - h.cb;
    bucket = x >> sb;
    stored = (uint32_t)(x & ((1ull << sb) - 1));
    return true;
}

// the upper word of a compact slot holding (d, stored)
__device__ __forceinline__ uint32_t chash_tag(uint32_t d, uint32_t stored) {
    return 0x80000000u | (d << kCompactStoredBits) | stored;
}

// murmur3's 64-bit finaliser: board keys share most of their bits
__device__ __forceinline__ uint64_t hash_mix(uint64_t x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33;
    return x;
}

// Insert into an open-addressing key set (slots hold the key, 0 = empty;
// linear probing from hash_mix). A slot only goes from 0 to a key within a
// launch, so a plain load that sees a key is exact and one that sees 0 is
// settled by the CAS. false: kSetProbeLimit slots in a row were taken (the
// key was not inserted; the caller raises its overflow flag).
constexpr uint32_t kSetProbeLimit = 256;

__device__ __forceinline__ bool set_insert(uint64_t *__restrict__ tab, uint64_t mask, uint64_t k) {
    uint64_t h = hash_mix(k) & mask;
    for (uint32_t q = 0; q < kSetProbeLimit; ++q) {
        const uint64_t cur = tab[h];
        if (cur == k) return true;
        if (cur == 0) {
            const uint64_t old = atomicCAS((unsigned long long *)&tab[h], 0ull, (unsigned long long)k);
            if (old == 0 || old == k) return true;
        }
        h = (h + 1) & mask;
    }
    return false;
}

template <typename T>
__device__ __forceinline__ T rate_from_bits(uint64_t v) {
    if constexpr (sizeof(T) == 4) return __builtin_bit_cast(T, (uint32_t)v);
    else return __builtin_bit_cast(T, v);
}

uint64_t hash_capacity(uint64_t bound);
// Insert records [0, n) of a layer (n = *n_ptr when non-null, else bound)
// into a cleared-here table.
template <typename T>
void launch_hash_build(hipStream_t s, const uint64_t *keys, const T *rates, uint64_t bound, const uint64_t *n_ptr,
                       const HashDev &h);

// Sample-sort dedup (egtb_bucket.hip): sorted unique keys of up to
// kBucketMaxKeys candidates in six launches.  BucketWork holds its device
// scratch: splitters (1023 x u64), per-bucket counts, starts, cursors and
// unique counts (1024 x u32 each) and each key's bucket.
constexpr uint64_t kBucketMaxKeys = 1ull << 21;
struct BucketWork {
    uint64_t *spl;
    uint32_t *counts, *starts, *cursor, *ucount;
    uint16_t *ids;  // bucket of each key (n entries)
};
uint32_t bucket_count_for(uint64_t n);
void configure_bucket_kernels();
// keys is read only; tmp holds n keys; scratch holds 2n keys (used only by a
// bucket too large for LDS).
void launch_bucket_dedup(hipStream_t s, const uint64_t *keys, uint64_t *tmp, uint64_t *scratch, uint64_t n,
                         BucketWork &bw, uint64_t *out, uint64_t *count_out);

struct LaunchShape {
    uint32_t blocks;   // persistent grid size (upper bound)
    uint32_t threads;  // threads per block
};

// The job table of a batched build (tab == nullptr: a plain build, whose
// pattern and docheck flag the launch passes).  err gets bit 1 when a
// canonical child's low byte is not 0xFF (it cannot be stored as a job key)
// and bit 2 when a child's key is 0 (the empty slot of key sets and hashes).
struct JobArg {
    const JobDev *tab = nullptr;
    uint32_t *err = nullptr;
    int32_t step = 0;  // forward: the step (parents in layer step-1); solve: the layer
};

void configure_kernels();
void launch_build_lut(hipStream_t s, uint16_t *lut);
// Sizes: *_bound is a host-side upper bound (launch geometry); when the
// matching n_ptr is non-null the kernel reads the true size from device
// memory, so a launch can be queued before the producer has finished.
void launch_expand(hipStream_t s, bool write, const LaunchShape &ls, const uint64_t *par, uint64_t n_bound,
                   const uint64_t *n_ptr, const uint16_t *lut, const PatternDev &p, int do_check,
                   uint64_t *counts, const uint64_t *offs, uint64_t *out2, uint64_t *out4, uint64_t *or_and2,
                   uint64_t *or_and4, uint64_t offs_base = 0, uint64_t cap2 = ~0ull, uint64_t cap4 = ~0ull,
                   const JobArg &ja = JobArg{});
// One-launch forward step (any candidate order): *tot (zero before the
// launch) gets c2 | c4 << 32; children at or past cap2 / cap4 are counted,
// not written.
void launch_expand_alloc(hipStream_t s, const LaunchShape &ls, const uint64_t *par, uint64_t n_bound,
                         const uint64_t *n_ptr, const uint16_t *lut, const PatternDev &p, int do_check,
                         uint64_t *tot, uint64_t *out2, uint64_t *out4, uint64_t *or_and2, uint64_t *or_and4,
                         uint64_t cap2, uint64_t cap4, const JobArg &ja = JobArg{});
// Forward step into key sets (open addressing, 0 = empty, capacities are
// powers of two): spawn-2 children into set2, spawn-4 children into set4
// (either may be null: counted, not inserted). *tot gets c2 | c4 << 32;
// meta2 / meta4 point at the LayerDev of layers step and step+1: their
// or_all / and_all fold the children's keys and pad[0] becomes nonzero when
// an insert was dropped (the set was too full).
void launch_expand_hash(hipStream_t s, const LaunchShape &ls, const uint64_t *par, uint64_t n_bound,
                        const uint64_t *n_ptr, const uint16_t *lut, const PatternDev &p, int do_check, uint64_t *tot,
                        uint64_t *set2, uint64_t cap2, uint64_t *set4, uint64_t cap4, uint64_t *meta2,
                        uint64_t *meta4, const JobArg &ja = JobArg{});
// keys[0..n) (nonzero) into a set of cap > n slots (a power of two); cannot
// overflow.
void launch_set_insert(hipStream_t s, const uint64_t *keys, uint64_t n, uint64_t *set, uint64_t cap);
// Keys of a set to out[0..*count) in any order; the set's slots are cleared.
// *count must be zero before; out must hold cap keys.
void launch_set_extract(hipStream_t s, uint64_t *set, uint64_t cap, uint64_t *out, uint64_t *count);
void launch_gather_u64(hipStream_t s, const uint64_t *src, const uint64_t *idx, uint32_t m, uint64_t *out);
void launch_copy_u64(hipStream_t s, const uint64_t *src, uint64_t *dst);

uint64_t scan_partials_needed(uint64_t n);
void launch_scan_u64(hipStream_t s, const uint64_t *in, uint64_t n_bound, const uint64_t *n_ptr, uint64_t *out,
                     uint64_t *partial, uint64_t *total_out);
void launch_scan_u32(hipStream_t s, const uint32_t *in, uint64_t n, uint32_t *out, uint32_t *partial);

// Device state of the single-pass (decoupled look-back) kernels of one
// stream: status words tagged with a pass epoch (never cleared), cumulative
// tile counters (never reset: each launch claims ids from `claimed` on),
// double-buffered digit histograms, and a stall flag.
constexpr int kHistWords = 16 * 2048;
struct LookbackWork {
    uint64_t *status_sort = nullptr;  // [tiles * 2048]
    uint64_t *status_uniq = nullptr;  // [tiles] of unique and compaction launches
    unsigned long long *counters = nullptr;  // [0] sort tiles, [1] unique tiles
    uint32_t *hist = nullptr;         // [2 * kHistWords]
    uint32_t *err = nullptr;
    unsigned long long claimed = 0, claimed_uniq = 0;
    uint32_t epoch = 0;
    int hist_parity = 0;
    int dirty[2] = {0, 0};  // histogram words left non-zero in each buffer
    uint32_t next_epoch() {
        epoch = epoch + 1 >= (1u << 20) ? 1u : epoch + 1;
        return epoch;
    }
};

// keys per radix tile for a sort of n keys with the given schedule (1 or 2)
uint32_t radix_tile(int bits, uint64_t n, int schedule);

// Tile tunings installed by the autotuner (egtb_set_tile_tuning): per kind,
// ascending rows of (max_n, value); tuned_tile gives the value of the first
// row with n <= max_n (the last row also covers larger n), 0 when the kind
// has no rows (the built-in rule applies).
enum TileKind { kTileRadixItems = 0, kTileLookupThreads = 1, kTileKinds = 2 };
void set_tile_tuning(int kind, std::vector<std::pair<uint64_t, int>> rows);
int tuned_tile(int kind, uint64_t n);
// Classic schedule: per digit a tile-histogram kernel, a per-digit scan of
// the [digit][tile] counts and a scatter kernel.  counts holds
// tiles << bits words, offsets tiles << bits words plus 2^bits digit totals
// (tiles = n / radix_tile(bits, n, 1), rounded up).
uint64_t *launch_classic_sort(hipStream_t s, int bits, uint64_t *keys, uint64_t *tmp, uint64_t n, int lo, int hi,
                              uint32_t *counts, uint32_t *offsets);
// tiles of the chained scan of n words (status words needed)
// Stable LSD sort of keys[0..n) on bits [lo, hi); returns keys or tmp,
// whichever holds the result.
uint64_t *launch_onesweep_sort(hipStream_t s, int bits, uint64_t *keys, uint64_t *tmp, uint64_t n, int lo, int hi,
                               LookbackWork &lw);

uint64_t unique_tiles(uint64_t n);
void launch_unique(hipStream_t s, const uint64_t *a, uint64_t n, uint64_t *out, uint64_t *count_out,
                   LookbackWork &lw);

void launch_bucket_index(hipStream_t s, const uint64_t *keys, uint64_t n_bound, const uint64_t *n_ptr,
                         uint32_t bits, uint32_t *idx);

template <typename T>
// clear: a hash table the same launch zeroes (a later layer's, whose
// slots no probe of this launch reads), so its fill needs no launch of its own.
// own: this layer's own solve hash (cleared before the launch), which gets
// every state with a rate above zero.
void launch_solve(hipStream_t s, const LaunchShape &ls, const uint64_t *keys, uint64_t n, const uint16_t *lut,
                  const PatternDev &p, int do_check, const HashDev &h1, const HashDev &h2,
                  const TableDev<T> &t1, const TableDev<T> &t2, double w2, double w4, T max_scale, T zero,
                  T *rates, const HashDev &clear = HashDev{}, const HashDev &own = HashDev{},
                  const JobArg &ja = JobArg{});

// Final layers: success -> max_scale, else zero.
template <typename T>
void launch_final_rates(hipStream_t s, const uint64_t *keys, uint64_t n, const PatternDev &p, T max_scale, T zero,
                        T *rates, const JobArg &ja = JobArg{});

// Records with rate > threshold (or, with `keep`, those whose keep byte is
// set), in order, in one look-back launch; it shares the unique kernel's
// status words (ensure_lookback with compact_tiles(n) unique tiles first).
// With a hash `fill`, the table is cleared (unless `zeroed`) and the kept
// records hashed into it too.
uint64_t compact_tiles(uint64_t n);
template <typename T>
void launch_compact(hipStream_t s, const uint64_t *keys, const T *rates, uint64_t n, uint64_t *out_keys,
                    T *out_rates, unsigned long long *max_bits, uint64_t *count_out, LookbackWork &lw,
                    T threshold = T(0), const HashDev &fill = HashDev{}, bool zeroed = false,
                    const uint8_t *keep = nullptr);

// Marks (mark[pos] = 1) the records of table t that the n parents, with
// `spawn` on each empty cell, reach as their best move; t has a bucket index.
template <typename T>
void launch_opt_mark(hipStream_t s, const LaunchShape &ls, const uint64_t *parents, uint64_t n, const uint16_t *lut,
                     const PatternDev &p, uint64_t spawn, const TableDev<T> &t, T zero, uint8_t *mark);

template <typename T>
void launch_pack_records(hipStream_t s, const uint64_t *keys, const T *rates, uint64_t n, uint32_t *out);

// Packed {u64 board; rate} records (rate_size 4 or 8 bytes) -> SoA.
void launch_unpack_records(hipStream_t s, const uint32_t *recs, uint64_t n, int rate_size, uint64_t *keys,
                           void *rates);

}  // namespace egtb
