// Host-side runtime types shared by the build, batched and serving modules.
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "egtb.h"
#include "egtb_kernels.h"

namespace egtb {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

extern thread_local std::string g_last_error;

void hip_check(hipError_t e, const char *what);
#define HIPCHK(x) ::egtb::hip_check((x), #x)

struct DeviceCtx {
    int dev = 0;
    int cus = 0;
    uint16_t *lut = nullptr;  // 65536-entry left-slide row table
    hipStream_t stream = nullptr;
    LaunchShape shape{};
};

DeviceCtx &device_ctx(int dev);

// Move-only device buffer, allocated and freed in stream order.
struct DBuf {
    void *p = nullptr;
    size_t bytes = 0;
    hipStream_t stream = nullptr;

    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept { *this = std::move(o); }
    DBuf &operator=(DBuf &&o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            bytes = o.bytes;
            stream = o.stream;
            o.p = nullptr;
            o.bytes = 0;
        }
        return *this;
    }
    ~DBuf() { release(); }

    void alloc(size_t nbytes, hipStream_t s);
    void ensure(size_t nbytes, hipStream_t s);  // grow-only
    void release();
};

// Pinned host staging buffer (grow-only).
struct HostStage {
    void *p = nullptr;
    size_t bytes = 0;
    HostStage() = default;
    HostStage(const HostStage &) = delete;
    HostStage &operator=(const HostStage &) = delete;
    ~HostStage();
    void ensure(size_t n);
};

// Scratch of the dedup / compaction pipelines of one stream.
struct SortWork {
    DBuf counts, offsets, partial, tile_counts, tile_offs, maxbits;
    DBuf bk_u64, bk_u32, bk_scratch, bk_ids;  // sample-sort dedup
    DBuf lb_status_sort, lb_status_uniq, lb_status_scan, lb_misc;
    LookbackWork lw;
    // grow the look-back state for a sort of `sort_tiles` tiles of `bins`
    // digits and a unique of `uniq_tiles` tiles (0 = no change)
    void ensure_lookback(hipStream_t s, uint64_t sort_tiles, int bins, uint64_t uniq_tiles, uint64_t scan_tiles = 0);
    // stall flag of the look-back kernels (synchronises)
    uint32_t read_error(hipStream_t s);
};

// HIP-event timing of one launch group, accumulated per kernel family when
// statistics are enabled (egtb_stats_enable); a no-op otherwise.
class Timer {
  public:
    Timer(hipStream_t s, const char *name, double alg_bytes);
    ~Timer();
    Timer(const Timer &) = delete;
    Timer &operator=(const Timer &) = delete;

  private:
    hipStream_t s_;
    const char *name_;
    double bytes_;
    bool on_ = false;
    hipEvent_t a_ = nullptr, b_ = nullptr;
};

// Bump allocator over large device chunks: layers are carved out in build
// order, the newest allocation can be trimmed once its true size is known,
// and everything is released at once. With a stream set, chunks come from
// the stream-ordered pool and the chunks above a consumed allocation can be
// handed back while the stream still runs.
class Arena {
  public:
    explicit Arena(size_t min_chunk = (size_t)1 << 30) : min_chunk_(min_chunk) {}
    Arena(const Arena &) = delete;
    Arena &operator=(const Arena &) = delete;
    ~Arena() { release(); }
    void set_stream(hipStream_t s) { stream_ = s; }
    void *alloc(size_t bytes);
    void trim_last(void *p, size_t bytes);  // shrink the newest allocation
    void release_after(const void *p);      // free every chunk newer than the one holding p
    void release();
    size_t used() const;

  private:
    struct Chunk {
        char *p;
        size_t cap, used;
    };
    void free_chunk(Chunk &c);
    std::vector<Chunk> chunks_;
    size_t min_chunk_;
    hipStream_t stream_ = nullptr;
};

// Per-layer device record: sizes, key span and stats of layer i.
struct LayerDev {
    uint64_t n_raw;    // unique raw size N_i
    uint64_t n_live;   // records with rate > 0
    uint64_t or_all;   // OR of every candidate key of layer i
    uint64_t and_all;  // AND of every candidate key of layer i
    uint64_t maxbits;  // bits of the largest rate
    uint64_t tot;      // candidates of forward step i: c2 | c4 << 32
    uint64_t pad[2];
};

// A solved layer resident in HBM: SoA keys/rates (arena) plus its bucket index.
struct Book {
    uint64_t *keys = nullptr;
    void *rates = nullptr;
    uint32_t *idx = nullptr;
    uint64_t *n_dev = nullptr;  // &LayerDev::n_live
    uint64_t n = 0;             // host copy, valid once the build finished
    uint32_t bits = 1;
    const ulonglong2 *hslots = nullptr;  // solve-side hash (HashDev) while the layer has readers
    uint64_t hmask = 0;
};

struct LayerInfo {
    uint64_t n_raw = 0, n_live = 0, c2 = 0, c4 = 0;
};

// Page-aligned host memory mapped on first write (no zero-fill pass), with
// transparent huge pages requested so that filling it costs few page faults.
struct HostArray {
    char *p = nullptr;
    size_t bytes = 0;
    HostArray() = default;
    HostArray(const HostArray &) = delete;
    HostArray &operator=(const HostArray &) = delete;
    HostArray(HostArray &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr, o.bytes = 0; }
    HostArray &operator=(HostArray &&o) noexcept {
        std::swap(p, o.p);
        std::swap(bytes, o.bytes);
        return *this;
    }
    ~HostArray() { reset(); }
    void allocate(size_t n);
    void reset();
};

struct Result {
    int steps = 0;
    int rate_size = 4;
    bool rate_is_float = true;
    int dev = 0;
    bool keep_tables = true;
    Arena raw_arena, book_arena;
    DBuf meta;  // LayerDev[steps + 1]
    std::vector<uint64_t *> raw;
    std::vector<Book> books;
    std::vector<LayerInfo> info;
    double fwd_ms = 0, bwd_ms = 0, total_ms = 0;
    // host-tables mode: books live in per-layer device buffers until they are
    // no longer needed, then in host memory
    bool host_tables = false;
    std::vector<DBuf> book_bufs;  // 3 per layer: keys, rates, index
    std::vector<DBuf> hash_bufs;  // per layer, freed after the layer's last solve
    double hash_max_bytes = 0;    // larger layers are searched, not hashed
    std::vector<HostArray> host_keys, host_rates;
    uint64_t streamed_bytes = 0, chunked_steps = 0, chunks = 0, peak_device_bytes = 0;
    double stream_ms = 0;
    LayerDev *meta_dev() const { return (LayerDev *)meta.p; }
};

PatternDev make_pattern(const egtb_pattern *pat, int target);
int choose_radix_bits(int span, int requested);
int choose_schedule(uint64_t n, int requested);
uint64_t *radix_sort(hipStream_t s, uint64_t *keys, uint64_t *tmp, uint64_t n, uint64_t or_all, uint64_t and_all,
                     int bits_req, SortWork &w, int schedule);
void unique_into(hipStream_t s, const uint64_t *sorted, uint64_t n, uint64_t *out, uint64_t *count_out,
                 SortWork &w, int schedule);
// Sorted unique keys of keys[0..n) into out (size to count_out, on the
// device).  keys and tmp (n keys) are clobbered; out is a third buffer.
void dedup_into(hipStream_t s, uint64_t *keys, uint64_t *tmp, uint64_t n, uint64_t or_all, uint64_t and_all,
                uint64_t *out, uint64_t *count_out, int bits_req, SortWork &w, int schedule);
void write_file_atomic(const std::string &path, const void *data, size_t bytes);
void progress_set(uint32_t cur, uint32_t total);
void build_dispatch(const uint64_t *seeds, size_t n_seeds, const egtb_pattern *pattern, const egtb_options *o,
                    Result &res);

}  // namespace egtb
