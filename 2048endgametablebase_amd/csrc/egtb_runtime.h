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
    hipStream_t aux = nullptr;  // second stream: the backward pass's compactions
    LaunchShape shape{};
};

DeviceCtx &device_ctx(int dev);
// Device blocks from the per-device cache (idle blocks of the request's size
// class are reused after the taking stream waits for their release; new
// blocks come from hipMalloc, and when the device is out of memory the idle
// blocks are freed and the request tried once more).  The flag is unused.
hipError_t device_alloc(void **p, size_t nbytes, hipStream_t s, bool async = true);
// Back to the cache, idle once the work queued on `s` so far is done (s ==
// nullptr: once the device is idle).
void device_release(void *p, hipStream_t s);
// Bytes of the current device's idle cached blocks: free for this library.
size_t device_idle_bytes(int dev);

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
    DBuf counts, offsets, partial;
    DBuf bk_u64, bk_u32, bk_scratch, bk_ids;  // sample-sort dedup
    DBuf lb_status_sort, lb_status_uniq, lb_misc;
    LookbackWork lw;
    // grow the look-back state for a sort of `sort_tiles` tiles of `bins`
    // digits and a unique of `uniq_tiles` tiles (0 = no change)
    void ensure_lookback(hipStream_t s, uint64_t sort_tiles, int bins, uint64_t uniq_tiles);
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

// Device memory of the books in host-tables mode: large chunks handed out
// front to back, a chunk whose books have all gone to the host being reused
// for later books. A build frees a few dozen chunks at its end instead of
// three stream-ordered frees per layer, which cost seconds on the pool
// (most take microseconds, one in a few hundred stalls for ~100 ms).
// Build thread only.
class BookRing {
  public:
    explicit BookRing(size_t chunk = (size_t)2 << 30) : chunk_(chunk) {}
    BookRing(const BookRing &) = delete;
    BookRing &operator=(const BookRing &) = delete;
    ~BookRing() { free_all(); }
    void reset(hipStream_t s, int layers);
    void *alloc(int layer, size_t bytes);
    void release(int layer);  // every array of the layer is no longer read
    void free_all();

  private:
    struct Chunk {
        char *p;
        size_t cap, used;
        int live;  // arrays handed out and not yet released
    };
    std::vector<Chunk> chunks_;
    std::vector<int> spare_;                  // chunks with no live array
    std::vector<std::vector<int>> of_layer_;  // chunk of each array, per layer
    int cur_ = -1;
    size_t chunk_;
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
    uint64_t pad[2];   // pad[0]: layer i's key set overflowed
    // at least the candidates of forward step i of each spawn value, summed
    // in 64 bits (reruns add again): below 2^32, tot's halves are exact
    uint64_t c2x, c4x;
};

// A solved layer resident in HBM: SoA keys/rates (arena) plus its bucket index.
struct Book {
    uint64_t *keys = nullptr;
    void *rates = nullptr;
    uint32_t *idx = nullptr;
    uint64_t *n_dev = nullptr;  // &LayerDev::n_live
    uint64_t n = 0;             // host copy, valid once the build finished
    uint32_t bits = 1;
    HashDev h{};  // solve-side hash while the layer has readers
};

struct LayerInfo {
    uint64_t n_raw = 0, n_live = 0, c2 = 0, c4 = 0;
    // OR / AND of the layer's keys (from its forward step); the defaults
    // mean "unknown" and give the full 64-bit span
    uint64_t key_or = ~0ull, key_and = 0;
};

// Host memory that streamed books are copied into: anonymous pages mapped
// on first write, with transparent huge pages requested so that faulting
// them in costs few faults. `used` is the part handed out.
struct HostSlab {
    char *p = nullptr;
    size_t bytes = 0, used = 0;
    HostSlab() = default;
    HostSlab(const HostSlab &) = delete;
    HostSlab &operator=(const HostSlab &) = delete;
    HostSlab(HostSlab &&o) noexcept { *this = std::move(o); }
    HostSlab &operator=(HostSlab &&o) noexcept {
        std::swap(p, o.p);
        std::swap(bytes, o.bytes);
        std::swap(used, o.used);
        return *this;
    }
    ~HostSlab();
    void map(size_t n);  // throws Error(EGTB_ERR_MEMORY)
};

// One streamed array (a range of a HostSlab).
struct HostView {
    char *p = nullptr;
    size_t bytes = 0;
};

// Streams a build owns (options.own_streams): made for the build and kept
// while the result's buffers may still be released on them; the destructor
// waits for both and destroys them.
struct OwnStreams {
    hipStream_t s = nullptr, s2 = nullptr;
    OwnStreams() = default;
    OwnStreams(const OwnStreams &) = delete;
    OwnStreams &operator=(const OwnStreams &) = delete;
    ~OwnStreams();
    void create();  // on the current device
};

struct Result {
    OwnStreams streams;  // first member: destroyed after every buffer that may use them
    int steps = 0;
    int rate_size = 4;
    bool rate_is_float = true;
    double rate_zero = 0.0;  // the rate type's zero value: -1 for the 1-float types
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
    BookRing book_ring;           // books' device memory until they are on the host
    std::vector<DBuf> hash_bufs;  // ring of solve hashes: layer i's in slot i % hash_ring
    int hash_ring = 3;
    double hash_max_bytes = 0;    // larger layers are searched, not hashed
    bool compact_hashes = true;   // compact solve-hash layout where a layer's key span allows
    DBuf hash_err;                // set by a compact insert that found no slot
    bool book_index = true;       // books get the bucket index of the solve's search path
    std::vector<HostSlab> host_slabs;            // own the memory of host_keys / host_rates
    std::vector<HostView> host_keys, host_rates;  // per layer
    uint64_t streamed_bytes = 0, chunked_steps = 0, chunks = 0, peak_device_bytes = 0;
    uint64_t hash_steps = 0, hash_regrows = 0;  // forward steps deduplicated in key sets; set regrowths
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
// A batched build through the plain schedule's stages (float32 rates, books
// kept in HBM): keys[0..n_keys) are the seeds of every job as job keys, job
// by job; o.steps is the largest job's layer count.  Kernels raise bits of
// *job_err (device memory, zero before) for children that cannot be stored
// as job keys.
void run_build_jobs(const uint64_t *keys, size_t n_keys, const egtb_options &o, const JobDev *jobs,
                    uint32_t *job_err, Result &res);

}  // namespace egtb
