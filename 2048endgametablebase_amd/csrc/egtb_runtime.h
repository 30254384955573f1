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

// Scratch of the dedup / compaction pipelines.
struct SortWork {
    DBuf counts, offsets, partial, tile_counts, tile_offs, maxbits;
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

// A solved layer resident in HBM: SoA keys/rates plus its bucket index.
struct Book {
    DBuf keys, rates, idx;
    uint64_t n = 0;
    uint32_t bits = 0;
    uint32_t rates_size = 4;
};

struct LayerInfo {
    uint64_t n_raw = 0, n_live = 0, c2 = 0, c4 = 0;
};

struct Result {
    int steps = 0;
    int rate_size = 4;
    bool keep_tables = true;
    std::vector<DBuf> raw;
    std::vector<Book> books;
    std::vector<LayerInfo> info;
    double fwd_ms = 0, bwd_ms = 0, total_ms = 0;
};

PatternDev make_pattern(const egtb_pattern *pat, int target);
int choose_radix_bits(int span, int requested);
uint64_t *radix_sort(hipStream_t s, uint64_t *keys, uint64_t *tmp, uint64_t n, uint64_t or_all, uint64_t and_all,
                     int bits_req, SortWork &w);
uint64_t sort_unique(hipStream_t s, DBuf &keys, DBuf &tmp, uint64_t n, uint64_t or_all, uint64_t and_all,
                     int bits_req, SortWork &w, DBuf &out);
void write_file_atomic(const std::string &path, const void *data, size_t bytes);
void progress_set(uint32_t cur, uint32_t total);
void build_dispatch(const uint64_t *seeds, size_t n_seeds, const egtb_pattern *pattern, const egtb_options *o,
                    Result &res);

}  // namespace egtb
