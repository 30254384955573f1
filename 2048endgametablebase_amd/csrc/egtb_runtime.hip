// Host runtime of libegtb_hip.so: device context, stream-ordered buffers,
// the dedup pipeline (LSD radix sort + unique), the plain per-layer build
// schedule (forward, final layers, backward), reference-layout file output
// and the C ABI declared in include/egtb.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>
#include <condition_variable>
#include <deque>
#include <thread>
#include <type_traits>
#include <vector>

#include "egtb.h"
#include "egtb_kernels.h"
#include "egtb_runtime.h"
#include "egtb_zbook.h"

namespace egtb {

// ---------------------------------------------------------------------------
// Errors
// ---------------------------------------------------------------------------

thread_local std::string g_last_error;

void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess)
        throw Error(EGTB_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// ---------------------------------------------------------------------------
// Device context
// ---------------------------------------------------------------------------

static std::mutex g_ctx_mu;
static std::map<int, std::unique_ptr<DeviceCtx>> g_ctx;

DeviceCtx &device_ctx(int dev) {
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    auto it = g_ctx.find(dev);
    if (it != g_ctx.end()) {
        HIPCHK(hipSetDevice(dev));
        return *it->second;
    }
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (dev < 0 || dev >= ndev) throw Error(EGTB_ERR_ARG, "no HIP device " + std::to_string(dev));
    HIPCHK(hipSetDevice(dev));
    auto c = std::make_unique<DeviceCtx>();
    c->dev = dev;
    HIPCHK(hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, dev));
    int lds = 0;
    HIPCHK(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    if (lds < kLutEntries * 2)
        throw Error(EGTB_ERR_HIP, "device LDS per block (" + std::to_string(lds) +
                                      " B) cannot hold the 128 KiB row table");
    configure_kernels();
    configure_bucket_kernels();
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking));
    HIPCHK(hipMalloc(&c->lut, kLutEntries * sizeof(uint16_t)));
    launch_build_lut(c->stream, c->lut);
    HIPCHK(hipGetLastError());
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
        uint64_t thr = UINT64_MAX;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    // one 1024-thread block per CU: the 128 KiB LDS table allows no more
    c->shape.blocks = (uint32_t)c->cus;
    c->shape.threads = 1024;
    DeviceCtx &ref = *c;
    g_ctx[dev] = std::move(c);
    return ref;
}

// ---------------------------------------------------------------------------
// Stream-ordered device buffers
// ---------------------------------------------------------------------------

// Device memory: a per-device cache of blocks.  A freed block is kept,
// idle, for the next request of its size class instead of going back to
// the driver: after a build frees tens of GB, fresh allocations stalled for
// seconds (single hipMalloc / hipMallocAsync calls of 1-3 s, with the pool
// holding no memory and 284 GB free), which points at the driver clearing
// the freed memory before handing it out again.  Classes are powers of two
// up to 1 MiB and eight steps per power of two above (at most 12.5% over the
// request).  A freed block carries an event recorded on the stream that freed
// it, and the stream that takes it next waits for that event, so reuse keeps
// stream order.  A block freed without a stream waits for the whole device
// first, as hipFree would.  When the device runs out of memory the idle
// blocks are freed (after a device synchronisation), the stream-ordered pool
// is trimmed, and the request is tried once more.
static void stats_add_host(const char *name, double ms);

namespace {

struct IdleBlock {
    void *p;
    hipEvent_t ev;
};

struct DevCache {
    std::map<size_t, std::vector<IdleBlock>> idle;  // by class size
    size_t idle_bytes = 0;
};

// never destroyed: buffers held by other static objects are released at
// process exit, in any order
struct CacheState {
    std::mutex mu;
    std::map<int, DevCache> devs;
    std::map<void *, std::pair<int, size_t>> live;  // block -> (device, class size)
    std::vector<hipEvent_t> free_events;
};
CacheState &cache_state() {
    static CacheState *c = new CacheState();
    return *c;
}

size_t size_class(size_t n) {
    if (n <= 256) return 256;
    if (n <= ((size_t)1 << 20)) {
        size_t c = 256;
        while (c < n) c <<= 1;
        return c;
    }
    size_t p2 = (size_t)1 << (63 - __builtin_clzll((unsigned long long)n));
    const size_t step = p2 / 8;
    return (n + step - 1) / step * step;
}

// frees every idle block of device `dev` (caller holds the cache mutex)
void flush_idle_locked(int dev) {
    CacheState &C = cache_state();
    (void)hipDeviceSynchronize();
    DevCache &dc = C.devs[dev];
    for (auto &kv : dc.idle)
        for (IdleBlock &b : kv.second) {
            (void)hipFree(b.p);
            C.free_events.push_back(b.ev);
        }
    dc.idle.clear();
    dc.idle_bytes = 0;
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) (void)hipMemPoolTrimTo(pool, 0);
}

}  // namespace

size_t device_idle_bytes(int dev) {
    CacheState &C = cache_state();
    std::lock_guard<std::mutex> lk(C.mu);
    auto it = C.devs.find(dev);
    return it == C.devs.end() ? 0 : it->second.idle_bytes;
}

hipError_t device_alloc(void **p, size_t nbytes, hipStream_t s, bool) {
    const auto t0 = std::chrono::steady_clock::now();
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    const size_t cls = size_class(nbytes);
    CacheState &C = cache_state();
    std::unique_lock<std::mutex> lk(C.mu);
    DevCache &dc = C.devs[dev];
    auto it = dc.idle.find(cls);
    if (it != dc.idle.end() && !it->second.empty()) {
        const IdleBlock b = it->second.back();
        it->second.pop_back();
        dc.idle_bytes -= cls;
        C.live[b.p] = {dev, cls};
        e = hipStreamWaitEvent(s, b.ev, 0);
        C.free_events.push_back(b.ev);
        *p = b.p;
        stats_add_host("host_alloc_reused", 0.0);
        return e;
    }
    e = hipMalloc(p, cls);
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) {
        (void)hipGetLastError();
        flush_idle_locked(dev);
        stats_add_host("host_cache_flush", 0.0);
        e = hipMalloc(p, cls);
    }
    if (e == hipSuccess) C.live[*p] = {dev, cls};
    lk.unlock();
    stats_add_host("host_alloc",
                   std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    return e;
}

void device_release(void *p, hipStream_t s) {
    if (!p) return;
    CacheState &C = cache_state();
    std::lock_guard<std::mutex> lk(C.mu);
    auto it = C.live.find(p);
    if (it == C.live.end()) return;
    const int dev = it->second.first;
    const size_t cls = it->second.second;
    C.live.erase(it);
    if (!s) (void)hipDeviceSynchronize();
    hipEvent_t ev = nullptr;
    if (!C.free_events.empty()) {
        ev = C.free_events.back();
        C.free_events.pop_back();
    } else if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
        (void)hipDeviceSynchronize();  // no event: the block is idle once the device is
        (void)hipFree(p);
        return;
    }
    (void)hipEventRecord(ev, s);
    DevCache &dc = C.devs[dev];
    dc.idle[cls].push_back(IdleBlock{p, ev});
    dc.idle_bytes += cls;
}

void DBuf::alloc(size_t nbytes, hipStream_t s) {
    release();
    stream = s;
    bytes = nbytes;
    if (nbytes == 0) return;
    hipError_t e = device_alloc(&p, nbytes, s, true);
    if (e != hipSuccess) {
        p = nullptr;
        bytes = 0;
        throw Error(EGTB_ERR_MEMORY, "device allocation of " + std::to_string(nbytes) + " bytes failed: " +
                                         hipGetErrorString(e));
    }
}

void DBuf::release() {
    if (p) device_release(p, stream);
    p = nullptr;
    bytes = 0;
}

// Grows by at least half, from 4 MiB: the forward pass's scratch follows
// layers that grow a little every step from a few keys, and each
// reallocation (tens of microseconds of host time) sits between a step's
// sizes and its first launch.
constexpr size_t kEnsureMin = (size_t)4 << 20;

void DBuf::ensure(size_t nbytes, hipStream_t s) {
    if (nbytes <= bytes && p) return;
    alloc(std::max({nbytes + nbytes / 4, bytes + bytes / 2, kEnsureMin}) + 256, s);
}

// ---------------------------------------------------------------------------
// Kernel statistics (HIP events around launch groups on the build stream)
// ---------------------------------------------------------------------------

static std::atomic<int> g_stats_on{0};
// slots per predicted key of the forward pass's key sets (2: 442@512's sets
// of 2-4 MiB then stay about within an XCD's L2; 3 was 1-3% slower)
constexpr double kSetSlotsPerKey = 2.0;
static std::atomic<double> g_set_slots_per_key{kSetSlotsPerKey};
// largest forward key set (256 MiB); steps that need more sort their
// candidates. 444@2048's forward pass, with the sorts on the second
// stream: 2.78 s at a limit of 2^24 slots (176 steps on sets), 1.78 s at
// 2^25 (every step), 4.8 s at 2^26 (inserts at HBM speed)
constexpr uint64_t kSetMaxSlots = (uint64_t)1 << 25;
static std::atomic<uint64_t> g_set_max_slots{kSetMaxSlots};
static std::mutex g_stats_mu;
struct StatAcc {
    uint64_t launches = 0;
    double ms = 0, bytes = 0;
};
struct PendingStat {
    const char *name;
    hipEvent_t a, b;
    double bytes;
};
static std::map<std::string, StatAcc> g_stats;
static std::vector<PendingStat> g_pending;
static std::vector<hipEvent_t> g_event_pool;

static hipEvent_t pooled_event() {
    if (!g_event_pool.empty()) {
        hipEvent_t e = g_event_pool.back();
        g_event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    return e;
}

// Resolve recorded event pairs into the accumulators (waits for them).
static void resolve_pending_locked() {
    for (auto &pe : g_pending) {
        (void)hipEventSynchronize(pe.b);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, pe.a, pe.b);
        auto &acc = g_stats[pe.name];
        acc.launches++;
        acc.ms += ms;
        acc.bytes += pe.bytes;
        g_event_pool.push_back(pe.a);
        g_event_pool.push_back(pe.b);
    }
    g_pending.clear();
}

// Host time of a host-side operation (device allocations), per family.
static void stats_add_host(const char *name, double ms) {
    if (!g_stats_on.load(std::memory_order_relaxed)) return;
    std::lock_guard<std::mutex> lk(g_stats_mu);
    auto &acc = g_stats[name];
    acc.launches++;
    acc.ms += ms;
}

// Add algorithmic bytes known only after a build to a kernel family.
static void stats_add_bytes(const char *name, double bytes) {
    if (!g_stats_on.load(std::memory_order_relaxed)) return;
    std::lock_guard<std::mutex> lk(g_stats_mu);
    resolve_pending_locked();
    g_stats[name].bytes += bytes;
}

Timer::Timer(hipStream_t s, const char *name, double bytes) : s_(s), name_(name), bytes_(bytes) {
    if (!g_stats_on.load(std::memory_order_relaxed)) return;
    on_ = true;
    std::lock_guard<std::mutex> lk(g_stats_mu);
    a_ = pooled_event();
    b_ = pooled_event();
    HIPCHK(hipEventRecord(a_, s_));
}

Timer::~Timer() {
    if (!on_) return;
    (void)hipEventRecord(b_, s_);
    std::lock_guard<std::mutex> lk(g_stats_mu);
    g_pending.push_back(PendingStat{name_, a_, b_, bytes_});
    if (g_pending.size() > 4096) resolve_pending_locked();
}

// ---------------------------------------------------------------------------
// Dedup pipeline: LSD radix sort over the varying-bit span + unique
// ---------------------------------------------------------------------------

// Sort tuning table: for keys counts up to max_n[i] (ascending) use
// bits[i] / schedule[i]; installed by the autotuner (egtb_set_sort_tuning).
static std::mutex g_tune_mu;
struct TuneRow {
    uint64_t max_n;
    int bits, schedule;
};
static std::vector<TuneRow> g_tune;

static const TuneRow *tune_row(uint64_t n) {
    for (const TuneRow &r : g_tune)
        if (n <= r.max_n) return &r;
    return g_tune.empty() ? nullptr : &g_tune.back();
}

int choose_radix_bits(int span, int requested) {
    if (requested == 4 || requested == 6 || requested == 8 || requested == 11) return requested;
    if (span <= 0) return 8;
    // fewest passes; on a tie the narrower digit (cheaper histograms)
    const int p8 = (span + 7) / 8, p11 = (span + 10) / 11;
    return p11 < p8 ? 11 : 8;
}

void SortWork::ensure_lookback(hipStream_t s, uint64_t sort_tiles, int bins, uint64_t uniq_tiles) {
    if (!lb_misc.p) {
        // counters (2 x u64), error flag, double histogram
        const size_t bytes = 64 + (size_t)2 * kHistWords * 4;
        lb_misc.alloc(bytes, s);
        HIPCHK(hipMemsetAsync(lb_misc.p, 0, bytes, s));
        // counters[0], [1] at bytes 0-15, error flag at 16
        lw.counters = (unsigned long long *)lb_misc.p;
        lw.err = (uint32_t *)((char *)lb_misc.p + 16);
        lw.hist = (uint32_t *)((char *)lb_misc.p + 64);
        lw.claimed = lw.claimed_uniq = 0;
    }
    if (lw.epoch > (1u << 20) - 64) {
        // status words are told apart by a 20-bit epoch: clear them before
        // an epoch value could come round again
        if (lb_status_sort.p) HIPCHK(hipMemsetAsync(lb_status_sort.p, 0, lb_status_sort.bytes, s));
        if (lb_status_uniq.p) HIPCHK(hipMemsetAsync(lb_status_uniq.p, 0, lb_status_uniq.bytes, s));
        lw.epoch = 0;
    }
    const size_t want_sort = (size_t)sort_tiles * (size_t)bins * 8;
    if (want_sort > lb_status_sort.bytes) {
        lb_status_sort.alloc(want_sort + want_sort / 2, s);
        HIPCHK(hipMemsetAsync(lb_status_sort.p, 0, lb_status_sort.bytes, s));
        lw.status_sort = (uint64_t *)lb_status_sort.p;
    }
    const size_t want_uniq = (size_t)uniq_tiles * 8;
    if (want_uniq > lb_status_uniq.bytes) {
        lb_status_uniq.alloc(want_uniq + want_uniq / 2 + 64, s);
        HIPCHK(hipMemsetAsync(lb_status_uniq.p, 0, lb_status_uniq.bytes, s));
        lw.status_uniq = (uint64_t *)lb_status_uniq.p;
    }
}

uint32_t SortWork::read_error(hipStream_t s) {
    if (!lw.err) return 0;
    uint32_t e = 0;
    HIPCHK(hipMemcpyAsync(&e, lw.err, 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return e;
}

// 1: per-digit radix (3 kernels a digit), 2: single-pass look-back radix,
// 3: sample-sort dedup (layers up to kBucketMaxKeys; else the table's radix).
int choose_schedule(uint64_t n, int requested) {
    if (requested == 1 || requested == 2) return requested;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    const TuneRow *r = tune_row(n);
    const int sc = requested == 3 ? 3 : (r ? r->schedule : 1);
    if (sc == 3) return n <= kBucketMaxKeys ? 3 : (r && (r->schedule == 1 || r->schedule == 2) ? r->schedule : 1);
    return sc == 2 ? 2 : 1;
}

static int radix_schedule(uint64_t n, int requested) {
    const int sc = choose_schedule(n, requested);
    if (sc != 3) return sc;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    const TuneRow *r = tune_row(n);
    return r && r->schedule == 2 ? 2 : 1;
}

static int tuned_bits(uint64_t n, int span, int requested) {
    if (requested) return choose_radix_bits(span, requested);
    std::lock_guard<std::mutex> lk(g_tune_mu);
    const TuneRow *r = tune_row(n);
    return choose_radix_bits(span, r ? r->bits : 0);
}

uint64_t *radix_sort(hipStream_t s, uint64_t *keys, uint64_t *tmp, uint64_t n, uint64_t or_all,
                     uint64_t and_all, int bits_req, SortWork &w, int schedule) {
    const uint64_t diff = or_all ^ and_all;
    if (n < 2 || diff == 0) return keys;
    if (n >= (1ull << 32)) throw Error(EGTB_ERR_ARG, "a single sort is limited to 2^32 keys");
    const int lo = __builtin_ctzll(diff);
    const int hi = 64 - __builtin_clzll(diff);
    const int bits = tuned_bits(n, hi - lo, bits_req);
    const int sched = radix_schedule(n, schedule);
    const uint64_t tile = radix_tile(bits, n, sched);
    const uint64_t ntiles = (n + tile - 1) / tile;
    uint64_t *out;
    if (sched == 2) {
        w.ensure_lookback(s, ntiles, 1 << bits, 0);
        out = launch_onesweep_sort(s, bits, keys, tmp, n, lo, hi, w.lw);
    } else {
        const uint64_t nc = ntiles << bits;
        w.counts.ensure(nc * 4, s);
        w.offsets.ensure((nc + (1ull << bits)) * 4, s);  // + the digit totals
        out = launch_classic_sort(s, bits, keys, tmp, n, lo, hi, (uint32_t *)w.counts.p, (uint32_t *)w.offsets.p);
    }
    HIPCHK(hipGetLastError());
    return out;
}

// One launch with decoupled look-back for every schedule: the three-launch
// form (tile counts, scan, write) cost ~25 us a layer at 1-4 Mi keys
// against ~5 us, all of it latency.
void unique_into(hipStream_t s, const uint64_t *sorted, uint64_t n, uint64_t *out, uint64_t *count_out,
                 SortWork &w, int schedule) {
    (void)schedule;
    w.ensure_lookback(s, 0, 0, unique_tiles(n));
    launch_unique(s, sorted, n, out, count_out, w.lw);
}

void dedup_into(hipStream_t s, uint64_t *keys, uint64_t *tmp, uint64_t n, uint64_t or_all, uint64_t and_all,
                uint64_t *out, uint64_t *count_out, int bits_req, SortWork &w, int schedule) {
    if (n >= 2 && (or_all ^ and_all) != 0 && choose_schedule(n, schedule) == 3) {
        Timer t(s, "sort", 16.0 * (double)n);
        w.bk_u64.ensure(1024 * 8, s);
        w.bk_u32.ensure(4 * 1024 * 4, s);
        w.bk_scratch.ensure(2 * n * 8, s);
        w.bk_ids.ensure(n * 2 + 64, s);
        BucketWork bw;
        bw.spl = (uint64_t *)w.bk_u64.p;
        bw.counts = (uint32_t *)w.bk_u32.p;
        bw.starts = bw.counts + 1024;
        bw.cursor = bw.counts + 2048;
        bw.ucount = bw.counts + 3072;
        bw.ids = (uint16_t *)w.bk_ids.p;
        launch_bucket_dedup(s, keys, tmp, (uint64_t *)w.bk_scratch.p, n, bw, out, count_out);
        HIPCHK(hipGetLastError());
        return;
    }
    uint64_t *sorted;
    {
        Timer t(s, "sort", 16.0 * (double)n);
        sorted = radix_sort(s, keys, tmp, n, or_all, and_all, bits_req, w, schedule);
    }
    Timer t(s, "unique", 16.0 * (double)n);
    unique_into(s, sorted, n, out, count_out, w, schedule);
}

// ---------------------------------------------------------------------------
// Files (reference FileIOUtils write path: <path>.tmp then rename,
// FileIOUtils.cpp:730-830)
// ---------------------------------------------------------------------------

void write_file_atomic(const std::string &path, const void *data, size_t bytes) {
    const std::string tmp = path + ".tmp";
    FILE *f = std::fopen(tmp.c_str(), "wb");
    if (!f) throw Error(EGTB_ERR_IO, "cannot open " + tmp + " for writing");
    size_t done = bytes ? std::fwrite(data, 1, bytes, f) : 0;
    const bool ok = done == bytes && std::fflush(f) == 0;
    std::fclose(f);
    if (!ok) {
        std::remove(tmp.c_str());
        throw Error(EGTB_ERR_IO, "short write to " + tmp);
    }
    if (std::rename(tmp.c_str(), path.c_str()) != 0) throw Error(EGTB_ERR_IO, "cannot rename " + tmp);
}

static bool file_exists(const std::string &path) {
    struct stat st;
    return ::stat(path.c_str(), &st) == 0;
}

static std::string now_string() {
    std::time_t t = std::time(nullptr);
    std::tm tmv{};
    localtime_r(&t, &tmv);
    char buf[64];
    std::strftime(buf, sizeof buf, "%Y-%m-%d %H:%M:%S", &tmv);
    return buf;
}

static double round_to_2(double v) { return std::round(v * 100.0) / 100.0; }

// ---------------------------------------------------------------------------
// Build progress (reference FormationProgress, FormationRuntime.h:61-86)
// ---------------------------------------------------------------------------

static std::atomic<uint32_t> g_prog_cur{0}, g_prog_total{0};

void progress_set(uint32_t cur, uint32_t total) {
    g_prog_total.store(total, std::memory_order_relaxed);
    g_prog_cur.store(cur, std::memory_order_relaxed);
}

// ---------------------------------------------------------------------------
// Rate type traits
// ---------------------------------------------------------------------------

template <typename T>
struct RateTraits;
template <>
struct RateTraits<float> {
    static float max_scale() { return 1.0f; }
    static double to_unit(float v) { return (double)v; }
};
template <>
struct RateTraits<double> {
    static double max_scale() { return 1.0; }
    static double to_unit(double v) { return v; }
};
template <>
struct RateTraits<uint32_t> {
    static uint32_t max_scale() { return 4000000000u; }
    static double to_unit(uint32_t v) { return (double)v / 4000000000.0; }
};
template <>
struct RateTraits<uint64_t> {
    static uint64_t max_scale() { return 1600000000000000000ull; }
    static double to_unit(uint64_t v) { return (double)v / 1.6e18; }
};

// ---------------------------------------------------------------------------
// The plain per-layer build
// ---------------------------------------------------------------------------

PatternDev make_pattern(const egtb_pattern *pat, int target) {
    if (!pat) throw Error(EGTB_ERR_ARG, "pattern is NULL");
    if (pat->n_masks > (uint32_t)kMaxMasks) throw Error(EGTB_ERR_ARG, "more than 16 pattern masks");
    if (pat->n_shifts > 16) throw Error(EGTB_ERR_ARG, "more than 16 success shifts");
    if (pat->symm_mode < 0 || pat->symm_mode > 3)
        throw Error(EGTB_ERR_ARG, "symm_mode " + std::to_string(pat->symm_mode) +
                                      " is not supported (identity, full, diagonal, horizontal)");
    if (target < 1 || target > 15) throw Error(EGTB_ERR_ARG, "target must be in 1..15 (log2 of the tile)");
    PatternDev p;
    std::memset(&p, 0, sizeof p);
    p.n_masks = pat->n_masks;
    for (uint32_t i = 0; i < pat->n_masks; ++i) p.masks[i] = pat->masks[i];
    p.symm = (uint32_t)pat->symm_mode;
    p.any_success = pat->n_shifts == 0 ? 1u : 0u;
    for (uint32_t i = 0; i < pat->n_shifts; ++i) {
        const uint32_t sh = pat->shifts[i];
        if (sh > 60 || sh % 4) throw Error(EGTB_ERR_ARG, "success shift must be a nibble offset 0..60");
        p.success_mask |= 0xFull << sh;
    }
    p.target_stacked = (uint64_t)target * 0x1111111111111111ull;
    return p;
}

// ---------------------------------------------------------------------------
// Pinned host staging
// ---------------------------------------------------------------------------

void HostStage::ensure(size_t n) {
    if (n <= bytes && p) return;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
    const size_t want = n + n / 4 + 4096;
    if (hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess) {
        p = nullptr;
        throw Error(EGTB_ERR_MEMORY, "pinned host allocation of " + std::to_string(want) + " bytes failed");
    }
    bytes = want;
}

HostStage::~HostStage() {
    if (p) (void)hipHostFree(p);
}

// ---------------------------------------------------------------------------
// Arena
// ---------------------------------------------------------------------------

void *Arena::alloc(size_t bytes) {
    bytes = (bytes + 255) & ~(size_t)255;
    if (chunks_.empty() || chunks_.back().cap - chunks_.back().used < bytes) {
        size_t cap = bytes > min_chunk_ ? bytes : min_chunk_;
        void *p = nullptr;
        hipError_t e = device_alloc(&p, cap, stream_, stream_ != nullptr);
        if (e != hipSuccess)
            throw Error(EGTB_ERR_MEMORY, "device allocation of " + std::to_string(cap) + " bytes failed: " +
                                             hipGetErrorString(e));
        chunks_.push_back(Chunk{(char *)p, cap, 0});
    }
    Chunk &c = chunks_.back();
    void *p = c.p + c.used;
    c.used += bytes;
    return p;
}

void Arena::trim_last(void *p, size_t bytes) {
    if (chunks_.empty() || !p) return;
    Chunk &c = chunks_.back();
    const size_t off = (size_t)((char *)p - c.p);
    if (off > c.used) return;  // not in the newest chunk
    const size_t end = off + ((bytes + 255) & ~(size_t)255);
    if (end < c.used) c.used = end;
}

void Arena::free_chunk(Chunk &c) { device_release(c.p, stream_); }

void Arena::release_after(const void *p) {
    for (size_t k = 0; k < chunks_.size(); ++k) {
        const Chunk &c = chunks_[k];
        if ((const char *)p >= c.p && (const char *)p < c.p + c.cap) {
            while (chunks_.size() > k + 1) {
                free_chunk(chunks_.back());
                chunks_.pop_back();
            }
            return;
        }
    }
}

void Arena::release() {
    for (auto &c : chunks_) free_chunk(c);
    chunks_.clear();
}

// ---------------------------------------------------------------------------
// BookRing
// ---------------------------------------------------------------------------

void BookRing::reset(hipStream_t s, int layers) {
    free_all();
    stream_ = s;
    of_layer_.assign((size_t)layers, {});
}

void *BookRing::alloc(int layer, size_t bytes) {
    bytes = (bytes + 255) & ~(size_t)255;
    if (cur_ < 0 || chunks_[cur_].cap - chunks_[cur_].used < bytes) {
        // leave the current chunk; it is spare already if its arrays are gone
        if (cur_ >= 0 && chunks_[cur_].live == 0) {
            chunks_[cur_].used = 0;
            spare_.push_back(cur_);
        }
        cur_ = -1;
        for (size_t k = 0; k < spare_.size(); ++k)
            if (chunks_[spare_[k]].cap >= bytes) {
                cur_ = spare_[k];
                spare_.erase(spare_.begin() + (std::ptrdiff_t)k);
                break;
            }
        if (cur_ < 0) {
            const size_t cap = std::max(bytes, chunk_);
            void *p = nullptr;
            const hipError_t e = device_alloc(&p, cap, stream_, true);
            if (e != hipSuccess)
                throw Error(EGTB_ERR_MEMORY, "device allocation of " + std::to_string(cap) +
                                                 " bytes for streamed books failed: " + hipGetErrorString(e));
            chunks_.push_back(Chunk{(char *)p, cap, 0, 0});
            cur_ = (int)chunks_.size() - 1;
        }
    }
    Chunk &c = chunks_[cur_];
    void *p = c.p + c.used;
    c.used += bytes;
    ++c.live;
    of_layer_[(size_t)layer].push_back(cur_);
    return p;
}

void BookRing::release(int layer) {
    for (int k : of_layer_[(size_t)layer]) {
        Chunk &c = chunks_[k];
        if (--c.live == 0 && k != cur_) {
            c.used = 0;
            spare_.push_back(k);
        }
    }
    of_layer_[(size_t)layer].clear();
}

void BookRing::free_all() {
    for (auto &c : chunks_) device_release(c.p, stream_);
    chunks_.clear();
    spare_.clear();
    cur_ = -1;
    for (auto &v : of_layer_) v.clear();
}

size_t Arena::used() const {
    size_t u = 0;
    for (auto &c : chunks_) u += c.used;
    return u;
}

static uint32_t index_bits_for(uint64_t n) {
    uint32_t b = 1;
    while (b < 24 && (1ull << (b + 3)) < n) ++b;
    return b;
}

// The solve-side hash of book `layer` (at most `bound` live records, count
// in device memory).
// Layers whose hash would exceed this many MiB are searched instead: a
// hash is built with one random atomic per record into a table four times
// the layer's size, which outgrows the caches on the largest layers.
// per-layer rates kept for the second stream's compactions, at most
constexpr double kTwoStreamRatesMax = 16.0 * 1024 * 1024 * 1024;
constexpr int kDefaultSolveHashMB = 1024;

// Keys per bucket (of 8 slots) at a compact table's size bound: buckets are
// the power of two at or above bound / 4, so a table is 25-50% full at its
// bound and less with the zero-rate states left out.  (At 6 keys a bucket,
// 38-75% full, 444@64's largest layers displaced records by up to 15
// buckets; at 4, by at most 5.)
constexpr double kCompactKeysPerBucket = 4.0;

// Layout of book `layer`'s solve hash for at most `bound` records (no table
// yet): compact when the rates are 4 bytes and the layer's key span leaves at
// most 26 stored bits at a size no larger than the wide table's, else wide;
// not present when the table would exceed the hash budget (the solve
// searches that layer instead).
static HashDev hash_layout(const Result &res, int layer, uint64_t bound) {
    HashDev h{};
    const uint64_t wide_cap = hash_capacity(bound);
    const double wide_bytes = (double)wide_cap * sizeof(ulonglong2);
    const LayerInfo &li = res.info[layer];
    const uint64_t diff = li.key_or ^ li.key_and;
    if (res.compact_hashes && res.rate_size == 4 && (li.key_and & ~li.key_or) == 0) {
        const uint32_t lo = diff ? (uint32_t)__builtin_ctzll(diff) : 0;
        const uint32_t span = diff ? (uint32_t)(64 - __builtin_clzll(diff)) - lo : 0;
        uint32_t cb = 0;
        while ((double)(1ull << cb) * kCompactKeysPerBucket < (double)bound && cb < 40) ++cb;
        if (span > kCompactStoredBits && cb < span - kCompactStoredBits) cb = span - kCompactStoredBits;
        if (cb > span) cb = span;
        const double bytes = (double)(1ull << cb) * 64.0;
        if (span < 64 && span - cb <= kCompactStoredBits && bytes <= std::max(wide_bytes, 65536.0) &&
            bytes <= res.hash_max_bytes) {
            h.mask = (1ull << cb) - 1;
            h.clo = lo;
            h.cs = span;
            h.cb = cb;
            h.cspan = span ? ((1ull << span) - 1) : 0;
            h.cconst = li.key_and & ~(h.cspan << lo);
            h.cslots = reinterpret_cast<const uint64_t *>(uintptr_t(1));  // marks the layout; set with the table
            return h;
        }
    }
    if (wide_bytes > res.hash_max_bytes) return h;
    h.mask = wide_cap - 1;
    h.slots = reinterpret_cast<const ulonglong2 *>(uintptr_t(1));
    return h;
}

// The table of book `layer`'s solve hash (the hash goes on the book), or a
// hash that is not present when the layer is too large to hash and the
// solve searches it instead.
static HashDev hash_table_for(hipStream_t s, Result &res, int layer, uint64_t bound) {
    Book &bk = res.books[layer];
    HashDev h = hash_layout(res, layer, bound);
    if (!h.present()) {
        bk.h = HashDev{};
        return bk.h;
    }
    // with a ring of 3, at most three hashes are alive (layers i, i+1, i+2
    // while layer i is solved): layer i takes slot i % 3, whose previous
    // table (layer i+3) was last read by the solve of layer i+1, queued
    // before. A ring of 4 lets the solve of layer i+1 clear layer i's slot.
    DBuf &buf = res.hash_bufs[layer % res.hash_ring];
    buf.ensure(h.bytes(), s);
    if (h.cslots) h.cslots = (const uint64_t *)buf.p;
    else h.slots = (const ulonglong2 *)buf.p;
    h.err = (uint32_t *)res.hash_err.p;
    bk.h = h;
    return h;
}

// The solve hash of a book that is already in place (one loaded from its
// file); solved layers get theirs from the compaction that makes them.
template <typename T>
static void build_hash(hipStream_t s, Result &res, int layer, uint64_t bound) {
    const HashDev h = hash_table_for(s, res, layer, bound);
    if (!h.present()) return;
    const Book &bk = res.books[layer];
    Timer t(s, "hash", (double)bound * (8 + sizeof(T)) + (double)h.bytes());
    launch_hash_build<T>(s, bk.keys, (const T *)bk.rates, bound, bk.n_dev, h);
}

template <typename T>
static TableDev<T> table_of(const Book &bk) {
    TableDev<T> t;
    t.keys = bk.keys;
    t.rates = (const T *)bk.rates;
    t.idx = bk.idx;
    t.n_ptr = bk.n_dev;
    t.n = bk.n;
    t.bits = bk.bits;
    t.present = 1;
    return t;
}

static HashDev hash_of(const Book &bk) { return bk.h; }

static void drop_hash(Result &res, int layer) { res.books[layer].h = HashDev{}; }

// Carve book storage for a layer of at most `bound` live records, compact
// (keys, rates) into it (stable, rate > 0) and index it.  Sizes stay on the
// device; nothing is synchronised.
// Book storage for a layer of at most `bound` live records (host-side
// carving; a new arena chunk is allocated on the build stream).
static void *book_bytes(Result &res, int layer, size_t bytes) {
    if (!res.host_tables) return res.book_arena.alloc(bytes);
    return res.book_ring.alloc(layer, bytes + 256);
}

template <typename T>
static void carve_book(Result &res, int layer, uint64_t bound) {
    Book &bk = res.books[layer];
    bk.n_dev = &res.meta_dev()[layer].n_live;
    bk.keys = (uint64_t *)book_bytes(res, layer, bound * 8);
    bk.rates = book_bytes(res, layer, bound * sizeof(T));
}

// Compact a carved book: (keys, rates) -> book storage (stable, rate > 0)
// and index it. Sizes stay on the device; nothing is synchronised.
template <typename T>
static void compact_book(hipStream_t s, Result &res, int layer, const uint64_t *keys, const T *rates, uint64_t bound,
                         SortWork &w, bool index, T zero, bool hash_zeroed = false) {
    Book &bk = res.books[layer];
    LayerDev *md = res.meta_dev() + layer;
    if (bound) w.ensure_lookback(s, 0, 0, compact_tiles(bound));
    // the compaction also fills the layer's solve hash (read by the solves
    // of layers i-1 and i-2; layer 0 is read by none)
    // (with hash_zeroed the table is in place already, cleared by the solve)
    const HashDev fill = !index ? HashDev{} : hash_zeroed ? bk.h : hash_table_for(s, res, layer, bound);
    {
        Timer t(s, "compact", (double)bound * (sizeof(T) + 8) * 2 + (double)fill.bytes());
        launch_compact<T>(s, keys, rates, bound, bk.keys, (T *)bk.rates, (unsigned long long *)&md->maxbits,
                          &md->n_live, w.lw, zero, fill, hash_zeroed);
    }
    if (index && res.book_index) {
        bk.bits = index_bits_for(bound);
        bk.idx = (uint32_t *)book_bytes(res, layer, ((1ull << bk.bits) + 1) * 4);
        launch_bucket_index(s, bk.keys, bound, bk.n_dev, bk.bits, bk.idx);
    }
}

// Carve and compact (the single-stream path).
template <typename T>
static void make_book(hipStream_t s, Result &res, int layer, const uint64_t *keys, const T *rates, uint64_t bound,
                      SortWork &w, bool index, T zero, bool hash_zeroed = false) {
    carve_book<T>(res, layer, bound);
    compact_book<T>(s, res, layer, keys, rates, bound, w, index, zero, hash_zeroed);
}

void HostSlab::map(size_t n) {
    void *q = mmap(nullptr, n, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
    if (q == MAP_FAILED)
        throw Error(EGTB_ERR_MEMORY, "host memory for streamed books (" + std::to_string(n) + " B)");
    (void)madvise(q, n, MADV_HUGEPAGE);
    p = (char *)q;
    bytes = n;
}

HostSlab::~HostSlab() {
    if (p) (void)munmap(p, bytes);
}

// memcpy split across a fixed set of worker threads plus the caller.
class CopyPool {
  public:
    explicit CopyPool(int workers) {
        for (int w = 0; w < workers; ++w) th_.emplace_back([this, w] { loop(w + 1); });
    }
    ~CopyPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : th_) t.join();
    }
    // src == nullptr writes one byte per 4 KiB page (faults the range in)
    void copy(char *dst, const char *src, size_t bytes) {
        const int parts = (int)th_.size() + 1;
        {
            std::lock_guard<std::mutex> lk(mu_);
            dst_ = dst, src_ = src, bytes_ = bytes, left_ = parts - 1;
            ++gen_;
        }
        cv_.notify_all();
        part(0, parts);
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [this] { return left_ == 0; });
    }

  private:
    void part(int k, int parts) {
        const size_t per = ((bytes_ + parts - 1) / parts + 4095) & ~(size_t)4095;
        const size_t a = std::min(bytes_, per * (size_t)k), b = std::min(bytes_, a + per);
        if (b <= a) return;
        if (src_) {
            std::memcpy(dst_ + a, src_ + a, b - a);
        } else {
            for (size_t off = a; off < b; off += 4096) dst_[off] = 0;
        }
    }
    void loop(int k) {
        uint64_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
            }
            part(k, (int)th_.size() + 1);
            std::lock_guard<std::mutex> lk(mu_);
            if (--left_ == 0) done_cv_.notify_one();
        }
    }
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    char *dst_ = nullptr;
    const char *src_ = nullptr;
    size_t bytes_ = 0;
    int left_ = 0;
    uint64_t gen_ = 0;
    bool stop_ = false;
};

// Host-tables mode: every finished book (keys + rates) is copied to host
// memory once the event that marks its last use on the device has passed.
//  - The forward pass reserves each layer's upper bound (its raw size); a
//    fault thread maps host slabs for it and faults their pages in while the
//    device is still busy, so the copies land in resident pages.
//  - Retiring a book stores its live size into host-mapped memory from the
//    build stream. The issue thread waits for the retire event, takes that
//    many records from the slabs and queues the book's arrays, in pieces, as
//    device-to-host copies into a ring of page-locked staging slots on a
//    stream of its own; it only waits when every slot is still being
//    drained. (A size read by a copy of its own queues behind the copies
//    already in flight and stalls the pipeline at every book.)
//  - The drain thread takes the pieces in order, waits for each copy and
//    moves the slot into the book's host memory with a thread pool.
//  - The device memory of a book that has arrived goes back to the build
//    thread (reclaim()).
// Page-locking the destinations instead costs the device time while the
// pages are mapped for it, and the system caps how much may be locked.
class BookStreamer {
  public:
    BookStreamer(Result &res, int rate_size, int steps)
        : res_(res), rate_size_(rate_size), copy_pool_(kCopyWorkers), fault_pool_(kFaultWorkers) {
        HIPCHK(hipStreamCreateWithFlags(&cs_, hipStreamNonBlocking));
        HIPCHK(hipHostMalloc((void **)&sizes_, (size_t)steps * 8, hipHostMallocMapped | hipHostMallocCoherent));
        HIPCHK(hipHostGetDevicePointer((void **)&sizes_dev_, sizes_, 0));
        for (int k = 0; k < kSlots; ++k) {
            HIPCHK(hipHostMalloc(&stage_[k], kStage, hipHostMallocDefault));
            HIPCHK(hipEventCreateWithFlags(&slot_ev_[k], hipEventDisableTiming));
        }
        issue_th_ = std::thread([this] { issue_loop(); });
        drain_th_ = std::thread([this] { drain_loop(); });
        fault_th_ = std::thread([this] { fault_loop(); });
    }
    ~BookStreamer() {
        try {
            finish();
        } catch (...) {
        }
    }

    // a book of at most `records` records will be streamed
    void reserve(uint64_t records) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            // 1/16 more for the slab tails that are too short for a book's next array
            const uint64_t b = records * (8 + (uint64_t)rate_size_);
            want_ += b + b / 16 + 2 * kAlign;
        }
        fault_cv_.notify_one();
    }

    // book `layer`'s device buffers are read by work queued on s so far
    void retire(int layer, hipStream_t s) {
        hipEvent_t ev;
        launch_copy_u64(s, res_.books[layer].n_dev, sizes_dev_ + layer);
        HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        HIPCHK(hipEventRecord(ev, s));
        {
            std::lock_guard<std::mutex> lk(mu_);
            jobs_.push_back({layer, ev});
        }
        cv_.notify_all();
    }

    // frees the device memory of the books that are on the host (build thread)
    void reclaim() {
        std::vector<int> done;
        {
            std::lock_guard<std::mutex> lk(mu_);
            done.swap(arrived_);
        }
        for (int layer : done) {
            res_.book_ring.release(layer);
            Book &bk = res_.books[layer];
            bk.keys = nullptr;
            bk.rates = nullptr;
            bk.idx = nullptr;
        }
    }

    // waits for every retired book; slabs no book reached are unmapped
    void finish() {
        if (!issue_th_.joinable()) return;
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        issue_th_.join();
        {
            std::lock_guard<std::mutex> lk(mu_);
            issue_done_ = true;
        }
        cv_.notify_all();
        drain_th_.join();
        {
            std::lock_guard<std::mutex> lk(mu_);
            fault_stop_ = true;
        }
        fault_cv_.notify_all();
        fault_th_.join();
        reclaim();
        auto &sl = res_.host_slabs;
        sl.erase(std::remove_if(sl.begin(), sl.end(), [](const HostSlab &x) { return x.used == 0; }), sl.end());
        for (int k = 0; k < kSlots; ++k) {
            (void)hipHostFree(stage_[k]);
            (void)hipEventDestroy(slot_ev_[k]);
        }
        (void)hipHostFree(sizes_);
        (void)hipStreamDestroy(cs_);
        if (!err_.empty()) throw Error(EGTB_ERR_HIP, "streaming a layer to the host failed: " + err_);
    }

  private:
    static constexpr size_t kAlign = 256;
    static constexpr size_t kSlabMin = (size_t)64 << 20, kSlabMax = (size_t)1 << 30;
    static constexpr size_t kStage = (size_t)64 << 20;
    static constexpr int kSlots = 8;
    static constexpr int kCopyWorkers = 7;
    static constexpr int kFaultWorkers = 5;
    struct Job {
        int layer;
        hipEvent_t ev;
    };
    // one device-to-host copy into staging slot `slot` (-1: a book with no
    // records); `last` closes the book
    struct Piece {
        int slot;
        char *dst;
        size_t bytes;
        int layer;
        hipEvent_t retired;
        size_t book_bytes;
        bool last;
    };

    void fail(const std::string &what) {
        std::lock_guard<std::mutex> lk(mu_);
        if (err_.empty()) err_ = what;
        failed_ = true;
    }

    // `bytes` of host memory from the slabs whose pages are in (waiting for
    // the fault thread while it still has reserved memory to map), else from
    // a slab mapped here and faulted in by the copies
    char *take(size_t bytes) {
        bytes = (bytes + kAlign - 1) & ~(kAlign - 1);
        std::unique_lock<std::mutex> lk(mu_);
        auto &v = res_.host_slabs;
        for (;;) {
            // first fit over the faulted-in slabs that still have room;
            // slabs with less than kAlign * 4096 left are closed
            for (size_t k = cur_; k < ready_; ++k) {
                HostSlab &sl = v[k];
                if (sl.bytes - sl.used >= bytes) {
                    char *p = sl.p + sl.used;
                    sl.used += bytes;
                    while (cur_ < ready_ && v[cur_].bytes - v[cur_].used < kAlign * 4096) ++cur_;
                    return p;
                }
            }
            if (bytes > kSlabMax || !(fault_busy_ || mapped_ < want_)) break;
            slab_cv_.wait(lk);
        }
        // past the reservation: a slab of its own, after the faulted-in ones
        HostSlab own;
        own.map(std::max(bytes, kSlabMin));
        own.used = bytes;
        char *p = own.p;
        v.push_back(std::move(own));
        return p;
    }

    // maps and faults in slabs until the reservation is covered
    void fault_loop() {
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            fault_cv_.wait(lk, [this] { return fault_stop_ || mapped_ < want_; });
            if (fault_stop_) return;
            const size_t sz = want_ >= kSlabMax ? kSlabMax
                                                : std::max(kSlabMin, (want_ - mapped_ + ((size_t)2 << 20) - 1) &
                                                                         ~(((size_t)2 << 20) - 1));
            mapped_ += sz;
            fault_busy_ = true;
            lk.unlock();
            HostSlab sl;
            bool ok = true;
            try {
                sl.map(sz);
                fault_pool_.copy(sl.p, nullptr, sl.bytes);
            } catch (const std::exception &e) {
                ok = false;
                fail(e.what());
            }
            lk.lock();
            fault_busy_ = false;
            if (ok) {
                // slabs [0, ready_) are faulted in; take() maps its own slabs
                // only while this thread is idle, so they come after
                auto &v = res_.host_slabs;
                v.insert(v.begin() + (std::ptrdiff_t)ready_, std::move(sl));
                ++ready_;
            } else {
                want_ = mapped_;  // no more mapping ahead
            }
            slab_cv_.notify_all();
        }
    }

    void issue_loop() {
        (void)hipSetDevice(res_.dev);
        for (;;) {
            Job j;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [this] { return stop_ || !jobs_.empty(); });
                if (jobs_.empty()) return;
                j = jobs_.front();
                jobs_.pop_front();
            }
            try {
                issue(j);
            } catch (const std::exception &e) {
                fail(e.what());
                // the drain thread still closes the book (and its event)
                push_piece({-1, nullptr, 0, j.layer, j.ev, 0, true});
            }
        }
    }

    void push_piece(const Piece &p) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            pieces_.push_back(p);
        }
        cv_.notify_all();
    }

    void issue(const Job &j) {
        if (failed_) throw std::runtime_error("an earlier book failed");
        Book &bk = res_.books[j.layer];
        HIPCHK(hipEventSynchronize(j.ev));
        const uint64_t n = ((volatile const uint64_t *)sizes_)[j.layer];
        const size_t kb = n * 8, rb = n * (size_t)rate_size_;
        char *hk = kb ? take(kb) : nullptr, *hr = rb ? take(rb) : nullptr;
        res_.host_keys[j.layer] = {hk, kb};
        res_.host_rates[j.layer] = {hr, rb};
        struct Part {
            const char *src;
            char *dst;
            size_t bytes;
        };
        std::vector<Part> parts;
        for (size_t off = 0; off < kb; off += kStage)
            parts.push_back({(const char *)bk.keys + off, hk + off, std::min(kStage, kb - off)});
        for (size_t off = 0; off < rb; off += kStage)
            parts.push_back({(const char *)bk.rates + off, hr + off, std::min(kStage, rb - off)});
        if (parts.empty()) {
            push_piece({-1, nullptr, 0, j.layer, j.ev, 0, true});
            return;
        }
        for (size_t k = 0; k < parts.size(); ++k) {
            int slot;
            {
                // the slot's previous piece must have been drained
                std::unique_lock<std::mutex> lk(mu_);
                slot_cv_.wait(lk, [this] { return issued_ - drained_ < (uint64_t)kSlots; });
                slot = (int)(issued_ % kSlots);
                ++issued_;
            }
            try {
                HIPCHK(hipMemcpyAsync(stage_[slot], parts[k].src, parts[k].bytes, hipMemcpyDeviceToHost, cs_));
                HIPCHK(hipEventRecord(slot_ev_[slot], cs_));
            } catch (...) {
                push_piece({slot, nullptr, 0, j.layer, nullptr, 0, false});  // hands the slot back
                throw;
            }
            push_piece({slot, parts[k].dst, parts[k].bytes, j.layer, j.ev, kb + rb, k + 1 == parts.size()});
        }
    }

    void drain_loop() {
        (void)hipSetDevice(res_.dev);
        auto busy_from = std::chrono::steady_clock::now();
        bool busy = false;
        for (;;) {
            Piece p;
            {
                std::unique_lock<std::mutex> lk(mu_);
                if (pieces_.empty() && busy) {
                    res_.stream_ms +=
                        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - busy_from).count();
                    busy = false;
                }
                cv_.wait(lk, [this] { return issue_done_ || !pieces_.empty(); });
                if (pieces_.empty()) return;
                p = pieces_.front();
                pieces_.pop_front();
            }
            if (!busy) {
                busy = true;
                busy_from = std::chrono::steady_clock::now();
            }
            if (p.slot >= 0) {
                const hipError_t e = p.bytes ? hipEventSynchronize(slot_ev_[p.slot]) : hipSuccess;
                if (e == hipSuccess) {
                    copy_pool_.copy(p.dst, (const char *)stage_[p.slot], p.bytes);
                } else {
                    fail(std::string("device-to-host copy: ") + hipGetErrorString(e));
                }
                {
                    std::lock_guard<std::mutex> lk(mu_);
                    ++drained_;
                }
                slot_cv_.notify_all();
            }
            if (p.last) {
                (void)hipEventDestroy(p.retired);
                res_.streamed_bytes += p.book_bytes;
                std::lock_guard<std::mutex> lk(mu_);
                arrived_.push_back(p.layer);
            }
        }
    }

    Result &res_;
    int rate_size_;
    CopyPool copy_pool_, fault_pool_;
    hipStream_t cs_ = nullptr;
    uint64_t *sizes_ = nullptr, *sizes_dev_ = nullptr;  // live size per layer, host-mapped
    void *stage_[kSlots] = {};
    hipEvent_t slot_ev_[kSlots] = {};
    std::thread issue_th_, drain_th_, fault_th_;
    std::mutex mu_;  // guards everything below and res_.host_slabs
    std::condition_variable cv_, slot_cv_, fault_cv_, slab_cv_;
    std::deque<Job> jobs_;
    std::deque<Piece> pieces_;
    std::vector<int> arrived_;
    uint64_t issued_ = 0, drained_ = 0;
    size_t want_ = 0, mapped_ = 0, cur_ = 0, ready_ = 0;
    bool stop_ = false, issue_done_ = false, fault_stop_ = false, fault_busy_ = false;
    std::string err_;
    std::atomic<bool> failed_{false};
};

// Raw layer files <pathname><i> (N_i x u64, little endian; reference
// write_binary_vector_direct, FileIOUtils.h:444-451) written by a thread of
// their own while the forward pass goes on: device-to-host copies on a
// private stream into pinned slots, fwrite of one slot while the next fills.
// A layer is queued with its producer's stream: the copies wait for what that
// stream had queued (the layer's sort) at that point.
class RawFileWriter {
  public:
    explicit RawFileWriter(int dev) : dev_(dev) {
        HIPCHK(hipStreamCreateWithFlags(&cs_, hipStreamNonBlocking));
        for (int k = 0; k < kSlots; ++k) {
            HIPCHK(hipHostMalloc(&slot_[k], kStage, hipHostMallocDefault));
            HIPCHK(hipEventCreateWithFlags(&ev_[k], hipEventDisableTiming));
        }
        th_ = std::thread([this] { run(); });
    }
    ~RawFileWriter() {
        try {
            finish();
        } catch (...) {
        }
    }
    void write(const std::string &path, const uint64_t *keys, uint64_t n, hipStream_t producer) {
        hipEvent_t ready;
        HIPCHK(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
        const hipError_t e = hipEventRecord(ready, producer);
        if (e != hipSuccess) {
            (void)hipEventDestroy(ready);
            HIPCHK(e);
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            q_.push_back({path, keys, n, ready});
        }
        cv_.notify_one();
    }
    // waits until every queued file is on disk
    void finish() {
        if (!th_.joinable()) return;
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_one();
        th_.join();
        for (int k = 0; k < kSlots; ++k) {
            (void)hipHostFree(slot_[k]);
            (void)hipEventDestroy(ev_[k]);
        }
        (void)hipStreamDestroy(cs_);
        if (!err_.empty()) throw Error(EGTB_ERR_IO, err_);
    }

  private:
    static constexpr size_t kStage = (size_t)16 << 20;
    static constexpr int kSlots = 3;
    struct Job {
        std::string path;
        const uint64_t *keys;
        uint64_t n;
        hipEvent_t ready;
    };
    void write_one(const Job &j) {
        HIPCHK(hipStreamWaitEvent(cs_, j.ready, 0));
        const std::string tmp = j.path + ".tmp";
        FILE *f = std::fopen(tmp.c_str(), "wb");
        if (!f) throw Error(EGTB_ERR_IO, "cannot open " + tmp + " for writing");
        const size_t bytes = j.n * 8, pieces = (bytes + kStage - 1) / kStage;
        bool ok = true;
        size_t issued = 0, done = 0;
        try {
            while (done < pieces) {
                while (issued < pieces && issued - done < (size_t)kSlots) {
                    const int k = (int)(issued % kSlots);
                    const size_t off = issued * kStage, m = std::min(kStage, bytes - off);
                    HIPCHK(hipMemcpyAsync(slot_[k], (const char *)j.keys + off, m, hipMemcpyDeviceToHost, cs_));
                    HIPCHK(hipEventRecord(ev_[k], cs_));
                    ++issued;
                }
                const int k = (int)(done % kSlots);
                const size_t off = done * kStage, m = std::min(kStage, bytes - off);
                HIPCHK(hipEventSynchronize(ev_[k]));
                ok = ok && std::fwrite(slot_[k], 1, m, f) == m;
                ++done;
            }
        } catch (...) {
            (void)hipStreamSynchronize(cs_);
            std::fclose(f);
            std::remove(tmp.c_str());
            throw;
        }
        ok = ok && std::fflush(f) == 0;
        std::fclose(f);
        if (!ok) {
            std::remove(tmp.c_str());
            throw Error(EGTB_ERR_IO, "short write to " + tmp);
        }
        if (std::rename(tmp.c_str(), j.path.c_str()) != 0) throw Error(EGTB_ERR_IO, "cannot rename " + tmp);
    }
    void run() {
        (void)hipSetDevice(dev_);
        for (;;) {
            Job j;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [this] { return stop_ || !q_.empty(); });
                if (q_.empty()) return;
                j = std::move(q_.front());
                q_.pop_front();
            }
            if (err_.empty()) {  // after a failure the remaining files are dropped
                try {
                    write_one(j);
                } catch (const std::exception &e) {
                    err_ = e.what();
                }
            }
            (void)hipEventSynchronize(j.ready);
            (void)hipEventDestroy(j.ready);
        }
    }
    int dev_;
    hipStream_t cs_ = nullptr;
    void *slot_[kSlots] = {};
    hipEvent_t ev_[kSlots] = {};
    std::thread th_;
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<Job> q_;
    bool stop_ = false;
    std::string err_;
};

// Restart-file checks (reference is_valid_restart_file /
// remove_invalid_restart_file, BookGenerator.cpp:71-99): a file counts when
// it is a non-empty regular file whose size is a multiple of the record
// size; one that does not is removed.
static bool valid_restart_file(const std::string &path, uint64_t align) {
    struct stat st;
    if (::stat(path.c_str(), &st) != 0 || !S_ISREG(st.st_mode) || st.st_size <= 0) return false;
    return align == 0 || (uint64_t)st.st_size % align == 0;
}

static void drop_invalid_restart_file(const std::string &path, uint64_t align) {
    struct stat st;
    if (::stat(path.c_str(), &st) != 0 || valid_restart_file(path, align)) return;
    if (std::remove(path.c_str()) != 0) throw Error(EGTB_ERR_IO, "failed to remove invalid restart file: " + path);
}

static void remove_if_exists(const std::string &path) {
    if (file_exists(path) && std::remove(path.c_str()) != 0) throw Error(EGTB_ERR_IO, "cannot remove " + path);
}

// Resume: load a solved layer from its .book file into a device book
// (reference read_split_layer_file on restart, BookSolver.cpp:500-507).
template <typename T>
static void load_book_file(hipStream_t s, Result &res, int layer, const std::string &path, HostStage &stage,
                           DBuf &packed, LayerDev *H, bool hash = true) {
    constexpr size_t rec = 8 + sizeof(T);
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) throw Error(EGTB_ERR_IO, "cannot open " + path);
    std::fseek(f, 0, SEEK_END);
    const long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (sz < 0 || (size_t)sz % rec) {
        std::fclose(f);
        throw Error(EGTB_ERR_IO, path + ": size is not a multiple of the record size");
    }
    const uint64_t n = (uint64_t)sz / rec;
    stage.ensure((size_t)sz + 16);
    const size_t got = n ? std::fread(stage.p, 1, (size_t)sz, f) : 0;
    std::fclose(f);
    if (got != (size_t)sz) throw Error(EGTB_ERR_IO, "short read of " + path);
    Book &bk = res.books[layer];
    LayerDev *md = res.meta_dev() + layer;
    bk.n_dev = &md->n_live;
    bk.keys = (uint64_t *)res.book_arena.alloc(n * 8 + 8);
    bk.rates = res.book_arena.alloc(n * sizeof(T) + 8);
    bk.bits = index_bits_for(n);
    bk.idx = (uint32_t *)res.book_arena.alloc(((1ull << bk.bits) + 1) * 4);
    H[layer].n_live = n;
    HIPCHK(hipMemcpyAsync(&md->n_live, &H[layer].n_live, 8, hipMemcpyHostToDevice, s));
    if (n) {
        packed.ensure((size_t)sz, s);
        HIPCHK(hipMemcpyAsync(packed.p, stage.p, (size_t)sz, hipMemcpyHostToDevice, s));
        launch_unpack_records(s, (const uint32_t *)packed.p, n, (int)sizeof(T), bk.keys, bk.rates);
    }
    launch_bucket_index(s, bk.keys, n, bk.n_dev, bk.bits, bk.idx);
    if (hash) build_hash<T>(s, res, layer, n);
    HIPCHK(hipStreamSynchronize(s));  // the staging buffers are reused
}

template <typename T>
static void write_book_file(hipStream_t s, const std::string &path, const Book &bk, uint64_t n, HostStage &stage,
                            DBuf &packed) {
    constexpr size_t rec = 8 + sizeof(T);
    const size_t bytes = n * rec;
    if (n) {
        packed.ensure(bytes, s);
        launch_pack_records<T>(s, bk.keys, (const T *)bk.rates, n, (uint32_t *)packed.p);
        stage.ensure(bytes);
        HIPCHK(hipMemcpyAsync(stage.p, packed.p, bytes, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    write_file_atomic(path, stage.p, bytes);
}

// The stats row's max success rate from the largest kept rate's order key:
// (max - zero) / (max_scale - zero), as the reference writes it
// (BookSolver.cpp:570-581).
template <typename T>
static double rate_unit(uint64_t key, T zero, T max_scale) {
    const T v = rate_from_order_key<T>(key);
    return static_cast<double>(v - zero) / static_cast<double>(max_scale - zero);
}

// Copy layer records [lo, hi] of the device metadata into the pinned mirror
// and wait for them.
static void read_meta(hipStream_t s, Result &res, LayerDev *host, int lo, int hi) {
    HIPCHK(hipMemcpyAsync(host + lo, res.meta_dev() + lo, (size_t)(hi - lo + 1) * sizeof(LayerDev),
                          hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
}

// Destroys an event on the way out of a scope.
struct EventGuard {
    hipEvent_t e;
    ~EventGuard() { (void)hipEventDestroy(e); }
};

// What the stages of one build share: the device, options and pattern,
// the rate type's scale and zero, the per-layer records (pinned host mirror
// H, device M), the resume state and the scratch one stage hands the next.
OwnStreams::~OwnStreams() {
    if (s) {
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
    }
    if (s2) {
        (void)hipStreamSynchronize(s2);
        (void)hipStreamDestroy(s2);
    }
}

void OwnStreams::create() {
    if (s) return;
    HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
}

// The build stream and second stream of a build: the device's shared pair,
// or (options.own_streams) a pair of the result's own, so that builds from
// several host threads run side by side.
static hipStream_t build_stream(DeviceCtx &C, const egtb_options &o, Result &res, bool second) {
    if (!o.own_streams) return second ? C.aux : C.stream;
    HIPCHK(hipSetDevice(C.dev));
    res.streams.create();
    return second ? res.streams.s2 : res.streams.s;
}

template <typename T>
struct BuildCtx {
    DeviceCtx &C;
    hipStream_t s, s2;
    const egtb_options &o;
    const PatternDev &p;
    Result &res;
    const int steps;
    const bool files;
    const std::string prefix;
    // the rate type's scale and zero (FormationRuntime.h:261-303): the
    // 1-float types store rate - 1, so success is 0 and a miss is -1
    const T max_scale, zero;
    const double w4, w2;
    const uint32_t prog_total;
    std::unique_ptr<BookStreamer> streamer;
    SortWork w;
    HostStage stage, meta_host_buf;
    DBuf rates, packed;
    LayerDev *H = nullptr, *M = nullptr;
    std::vector<char> have_book;
    std::unique_ptr<RawFileWriter> raw_writer;
    std::chrono::steady_clock::time_point t_begin, t_fwd;
    bool two_stream = false;  // the backward pass's schedule
    uint64_t max_n = 0;       // largest layer
    // batched builds: the device job table (keys are job keys; the pattern,
    // docheck step and spawn weights of the options are not used)
    const JobDev *jobs = nullptr;
    uint32_t *job_err = nullptr;
    JobArg ja(int step) const { return JobArg{jobs, job_err, step}; }

    BuildCtx(const egtb_options &o_, const PatternDev &p_, Result &res_, bool one_minus)
        : C(device_ctx(o_.device)), s(build_stream(C, o_, res_, false)), s2(build_stream(C, o_, res_, true)), o(o_),
          p(p_), res(res_), steps(o_.steps),
          files(o_.write_files != 0), prefix(o_.pathname ? o_.pathname : ""),
          max_scale(one_minus ? T(0) : RateTraits<T>::max_scale()), zero(one_minus ? T(-1) : T(0)),
          w4(o_.spawn_rate4), w2(1.0 - o_.spawn_rate4), prog_total((uint32_t)(2 * o_.steps)), have_book(o_.steps, 0) {}

    // raw layer file of layer L (write_raw_files), after `producer`'s work
    void queue_raw(int L, uint64_t n, hipStream_t producer = nullptr) {
        if (raw_writer && !have_book[L])
            raw_writer->write(prefix + std::to_string(L), res.raw[L], n, producer ? producer : s);
    }
};

// The plain per-layer schedule.  Forward: one host synchronisation per step
// (after the count pass, to size the candidate streams); the spawn-4 stream
// of step i lands at the front of the candidate buffer of layer i+1 and the
// spawn-2 stream of step i+1 is appended behind it, so layer i+1 is one sort
// + unique of that buffer.  Backward: sizes stay on the device, no sync.
template <typename T>
static void forward_pass(BuildCtx<T> &b, const uint64_t *seeds, size_t n_seeds) {
    DeviceCtx &C = b.C;
    hipStream_t s = b.s;
    const egtb_options &o = b.o;
    const PatternDev &p = b.p;
    Result &res = b.res;
    const int steps = b.steps;
    const uint32_t prog_total = b.prog_total;
    std::unique_ptr<BookStreamer> &streamer = b.streamer;
    SortWork &w = b.w;
    LayerDev *const H = b.H;
    LayerDev *const M = b.M;
    auto queue_raw = [&](int L, uint64_t n, hipStream_t producer = nullptr) { b.queue_raw(L, n, producer); };
    DBuf cand[2], sort_tmp, counts, offs;

    // ---- forward (reference generate_process, BookGenerator.cpp:963-1121)
    res.raw[0] = (uint64_t *)res.raw_arena.alloc(n_seeds * 8);
    if (n_seeds) HIPCHK(hipMemcpyAsync(res.raw[0], seeds, n_seeds * 8, hipMemcpyHostToDevice, s));
    res.info[0].n_raw = n_seeds;

    auto launch_count = [&](int step, uint64_t bound) {
        // count pass of forward step `step` over layer step-1 (size on device)
        counts.ensure(bound * 8 + 8, s);
        offs.ensure(bound * 8 + 8, s);
        w.partial.ensure(scan_partials_needed(bound) * 8 + 64, s);
        {
            Timer t(s, "expand_count", 8.0 * (double)bound);
            launch_expand(s, false, C.shape, res.raw[step - 1], bound, &M[step - 1].n_raw, C.lut, p,
                          step > o.docheck_step ? 1 : 0, (uint64_t *)counts.p, nullptr, nullptr, nullptr,
                          &M[step].or_all, &M[step + 1].or_all, 0, ~0ull, ~0ull, b.ja(step));
        }
        launch_scan_u64(s, (const uint64_t *)counts.p, bound, &M[step - 1].n_raw, (uint64_t *)offs.p,
                        (uint64_t *)w.partial.p, &M[step].tot);
    };
    // one-pass expand of forward step `step` over layer step-1 (size on the
    // device) into out2 / out4, up to cap2 / cap4 candidates; the totals it
    // leaves in M[step].tot are exact even where writes were dropped
    uint64_t spec_cap2 = 0, spec_cap4 = 0;  // capacities the last one-pass expand ran with
    // (the per-layer records start zeroed, so only a rerun clears the counter)
    auto launch_alloc = [&](int step, uint64_t bound, uint64_t *out2, uint64_t *out4, uint64_t cap2,
                            uint64_t cap4, bool rerun) {
        if (rerun) HIPCHK(hipMemsetAsync(&M[step].tot, 0, 8, s));
        spec_cap2 = cap2;
        spec_cap4 = cap4;
        Timer t(s, "expand", 0.0);  // its bytes are counted once per step, below
        launch_expand_alloc(s, C.shape, res.raw[step - 1], bound, &M[step - 1].n_raw, C.lut, p,
                            step > o.docheck_step ? 1 : 0, &M[step].tot, out2, out4, &M[step].or_all,
                            &M[step + 1].or_all, cap2, cap4, b.ja(step));
    };

    // ---- chunked steps: runs (sorted unique pieces) of the layer being
    // assembled and of the next one; used once a step's candidates exceed
    // the budget, and from then on for the rest of the forward pass
    bool chunked = false;
    std::vector<DBuf> runs_cur, runs_next;
    std::vector<uint64_t> runs_cur_n, runs_next_n;
    DBuf chunk2, chunk4, run_tmp, run_out, cnt_dev, gidx, gval;
    cnt_dev.alloc(64, s);
    // free HBM is read (a driver call) at the first step and again only
    // when a step's candidates come within half of the last budget
    uint64_t budget_cache = 0;
    auto budget_keys = [&](uint64_t need) -> uint64_t {
        if (o.chunk_mkeys > 0) return (uint64_t)o.chunk_mkeys << 20;
        if (budget_cache == 0 || need > budget_cache / 2) {
            size_t fr = 0, tot = 0;
            HIPCHK(hipMemGetInfo(&fr, &tot));
            fr += device_idle_bytes(o.device);  // the cache's idle blocks are free to this build
            // a chunk's two streams, the sort scratch and the unique output
            budget_cache = std::max<uint64_t>(1ull << 20, (uint64_t)(fr / 8 / 6));
        }
        return budget_cache;
    };
    // sort + unique keys[0..n) into an exact-size run (synchronises)
    auto to_run = [&](uint64_t *keys, uint64_t n, uint64_t or_all, uint64_t and_all, std::vector<DBuf> &runs,
                      std::vector<uint64_t> &runs_n) {
        if (n == 0) return;
        run_tmp.ensure(n * 8, s);
        run_out.ensure(n * 8, s);
        uint64_t *u = (uint64_t *)run_out.p;
        dedup_into(s, keys, (uint64_t *)run_tmp.p, n, or_all, and_all, u, (uint64_t *)cnt_dev.p, o.radix_bits, w,
                   o.sort_schedule);
        uint64_t m = 0;
        HIPCHK(hipMemcpyAsync(&m, cnt_dev.p, 8, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        DBuf r;
        r.alloc(m * 8, s);
        if (m) HIPCHK(hipMemcpyAsync(r.p, u, m * 8, hipMemcpyDeviceToDevice, s));
        runs.push_back(std::move(r));
        runs_n.push_back(m);
    };
    // layer `L` = sort + unique of the concatenation of its runs
    auto merge_runs = [&](int L, std::vector<DBuf> &runs, std::vector<uint64_t> &runs_n) -> uint64_t {
        uint64_t tot = 0;
        for (uint64_t v : runs_n) tot += v;
        DBuf cat;
        cat.alloc(tot * 8 + 8, s);
        uint64_t off = 0;
        for (size_t k = 0; k < runs.size(); ++k) {
            if (runs_n[k])
                HIPCHK(hipMemcpyAsync((uint64_t *)cat.p + off, runs[k].p, runs_n[k] * 8, hipMemcpyDeviceToDevice, s));
            off += runs_n[k];
        }
        runs.clear();
        runs_n.clear();
        run_tmp.ensure(tot * 8, s);
        res.raw[L] = (uint64_t *)res.raw_arena.alloc(tot * 8);
        dedup_into(s, (uint64_t *)cat.p, (uint64_t *)run_tmp.p, tot, H[L].or_all, H[L].and_all, res.raw[L],
                   &M[L].n_raw, o.radix_bits, w, o.sort_schedule);
        return tot;
    };

    // Step i's one-pass expand is queued at the end of step i-1 into the
    // candidate buffers as they are; its totals (copied to H, marked by
    // meta_ev) tell whether every write landed. A step whose totals outgrew
    // the buffers grows them and runs again. Chunked steps use the count
    // pass and per-parent offsets instead.
    hipEvent_t meta_ev;
    HIPCHK(hipEventCreateWithFlags(&meta_ev, hipEventDisableTiming));
    EventGuard meta_ev_guard{meta_ev};
    auto queue_meta = [&](int lo, int hi) {
        HIPCHK(hipMemcpyAsync(H + lo, res.meta_dev() + lo, (size_t)(hi - lo + 1) * sizeof(LayerDev),
                              hipMemcpyDeviceToHost, s));
        HIPCHK(hipEventRecord(meta_ev, s));
    };
    uint64_t c4_prev = 0;

    // ---- hash-set forward: step i inserts its spawn-2 children into layer
    // i's key set and its spawn-4 children into layer i+1's; once step i has
    // run, layer i's set is extracted (unsorted, unique) and feeds step i+1
    // with its size on the device, and is sorted into the raw arena after
    // step i+1 is queued, so the host's wait for a layer's size overlaps the
    // next step on the device. Sets are sized ahead from the growth of the
    // known layers; a set that overflows is regrown and the steps feeding it
    // run again from the sorted layers. When a set would not fit free HBM
    // (or forward_hash_steps says so) the sort loop below takes over.
    struct KeySet {
        DBuf buf;
        uint64_t cap = 0;  // slots in use (a power of two); the rest of buf is zero
    };
    KeySet sets[3];
    // unsorted unique layers, layer L in ubuf[L % kURing]: a layer's sort
    // runs on the second stream a few steps after its extraction
    constexpr int kURing = 8;
    DBuf ubuf[kURing];
    hipEvent_t hev[2];
    HIPCHK(hipEventCreateWithFlags(&hev[0], hipEventDisableTiming));
    EventGuard hev0_guard{hev[0]};
    HIPCHK(hipEventCreateWithFlags(&hev[1], hipEventDisableTiming));
    EventGuard hev1_guard{hev[1]};
    static_assert(offsetof(LayerDev, pad) - offsetof(LayerDev, or_all) == 4 * sizeof(uint64_t),
                  "k_expand_hash raises the overflow flag at or_all + 4 words");
    auto set_of = [&](int L) -> KeySet & { return sets[L % 3]; };
    auto set_prepare = [&](KeySet &k, uint64_t cap) {
        if (k.buf.bytes < cap * 8) {
            k.buf.alloc(cap * 8, s);
            HIPCHK(hipMemsetAsync(k.buf.p, 0, k.buf.bytes, s));
        }
        k.cap = cap;
    };
    // slots for layer L from the sizes of layers up to `known`: the larger
    // of the last two growth ratios, compounded, times the slots per key
    const double slots_per_key = g_set_slots_per_key.load();
    const uint64_t max_slots = g_set_max_slots.load();
    auto cap_for = [&](int L, int known) -> uint64_t {
        const double n_k = (double)res.info[known].n_raw;
        double g = 2.0;
        if (known >= 1 && res.info[known - 1].n_raw) g = n_k / (double)res.info[known - 1].n_raw;
        if (known >= 2 && res.info[known - 2].n_raw)
            g = std::max(g, (double)res.info[known - 1].n_raw / (double)res.info[known - 2].n_raw);
        g = std::min(std::max(g, 1.0), 8.0);
        const double pred = n_k * std::pow(g, (double)(L - known)) * slots_per_key + 1024.0;
        uint64_t c = (uint64_t)1 << 15;
        while ((double)c < pred) c <<= 1;
        return c;
    };
    auto hash_step = [&](int step, const uint64_t *parents, uint64_t bound, KeySet *s2, KeySet *s4, uint64_t *tot) {
        Timer t(s, "expand", 0.0);  // its bytes are counted once per step, from its totals
        launch_expand_hash(s, C.shape, parents, bound, &M[step - 1].n_raw, C.lut, p, step > o.docheck_step ? 1 : 0,
                           tot, s2 ? (uint64_t *)s2->buf.p : nullptr, s2 ? s2->cap : 0,
                           s4 ? (uint64_t *)s4->buf.p : nullptr, s4 ? s4->cap : 0, &M[step].or_all,
                           &M[step + 1].or_all, b.ja(step));
    };
    // Sorts of unique layers on the second stream, in batches of kSortBatch
    // layers: one fork (the second stream waits for everything queued on
    // the build stream, which includes the expand that reads each layer) per
    // batch, and the build stream waits for a batch only before reusing its
    // layers' buffers. Explicit sort schedules sort in line instead.
    // 442@512: 65-67 ms with a fork per layer, 67-70 per 2, 72.7 per 4, 78
    // per 8 layers: a sort started early matters more than the ~18 us of
    // host time of an event pair.
    constexpr size_t kSortBatch = 1;
    const bool defer = o.sort_schedule == 0;
    hipStream_t s2 = b.s2;
    SortWork w_side;  // the second stream's sort state
    std::vector<int> pending;
    struct Batch {
        int last;
        hipEvent_t ev;
    };
    std::deque<Batch> batches;
    int s_waited = 0;  // layers up to this one are sorted before anything the build stream runs now
    hipEvent_t fork2;
    HIPCHK(hipEventCreateWithFlags(&fork2, hipEventDisableTiming));
    EventGuard fork2_guard{fork2};
    struct BatchGuard {
        std::deque<Batch> &b;
        ~BatchGuard() {
            for (Batch &x : b) (void)hipEventDestroy(x.ev);
        }
    } batch_guard{batches};
    struct SideSync {  // on the way out: sorts still running read buffers freed on the build stream
        hipStream_t side;
        ~SideSync() { (void)hipStreamSynchronize(side); }
    } sort_sync{s2};
    auto submit = [&]() {
        if (pending.empty()) return;
        HIPCHK(hipEventRecord(fork2, s));
        HIPCHK(hipStreamWaitEvent(s2, fork2, 0));
        for (int L : pending) {
            const uint64_t n = res.info[L].n_raw;
            uint64_t *u = (uint64_t *)ubuf[L % kURing].p;
            {
                Timer t(s2, "sort", 16.0 * (double)n);
                uint64_t *out = radix_sort(s2, u, res.raw[L], n, H[L].or_all, H[L].and_all, o.radix_bits, w_side, 0);
                if (out != res.raw[L] && n)
                    HIPCHK(hipMemcpyAsync(res.raw[L], out, n * 8, hipMemcpyDeviceToDevice, s2));
            }
            if (L <= steps - 2) queue_raw(L, n, s2);
        }
        hipEvent_t ev;
        HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        batches.push_back(Batch{pending.back(), ev});
        HIPCHK(hipEventRecord(ev, s2));
        pending.clear();
    };
    // the build stream waits for the sort of layer L (already submitted)
    auto wait_sorted = [&](int L) {
        if (L <= s_waited) return;
        while (batches.size() > 1 && batches.front().last < L) {  // a later batch implies this one
            (void)hipEventDestroy(batches.front().ev);
            batches.pop_front();
        }
        if (batches.empty() || batches.front().last < L)
            throw Error(EGTB_ERR_INTERNAL, "layer " + std::to_string(L) + " was not queued for its sort");
        HIPCHK(hipStreamWaitEvent(s, batches.front().ev, 0));
        s_waited = batches.front().last;
    };
    // every sort queued so far finishes before the build stream goes on
    auto join_sorts = [&]() {
        submit();
        if (!batches.empty()) wait_sorted(batches.back().last);
    };
    auto extract = [&](int L) {
        KeySet &k = set_of(L);
        if (defer && L - kURing >= 1) wait_sorted(L - kURing);  // the slot's previous layer
        ubuf[L % kURing].ensure(k.cap * 8, s);
        {
            Timer t(s, "extract", 16.0 * (double)k.cap);
            launch_set_extract(s, (uint64_t *)k.buf.p, k.cap, (uint64_t *)ubuf[L % kURing].p, &M[L].n_raw);
        }
        HIPCHK(hipMemcpyAsync(H + L, M + L, sizeof(LayerDev), hipMemcpyDeviceToHost, s));
        HIPCHK(hipEventRecord(hev[L & 1], s));
    };
    auto wait_meta = [&](int L) {
        for (hipError_t q; (q = hipEventQuery(hev[L & 1])) != hipSuccess;)
            if (q != hipErrorNotReady) HIPCHK(q);
    };
    // step i's candidates are counted as c2 | c4 << 32 (one atomic gives a
    // block both write positions), exact while each half stays below 2^32
    auto check_step = [&](int i) {
        if (H[i].c2x >= (1ull << 32) || H[i].c4x >= (1ull << 32))
            throw Error(EGTB_ERR_ARG, "forward step " + std::to_string(i) +
                                          " has 2^32 or more candidates of one spawn value");
    };
    // layer L (size known, sets drained) sorted into the raw arena
    auto sort_layer = [&](int L, uint64_t n) {
        res.raw[L] = (uint64_t *)res.raw_arena.alloc(n * 8);  // before the batch's fork
        if (defer) {
            pending.push_back(L);
            return;
        }
        uint64_t *u = (uint64_t *)ubuf[L % kURing].p;
        if (o.sort_schedule == 0) {
            Timer t(s, "sort", 16.0 * (double)n);
            uint64_t *out = radix_sort(s, u, res.raw[L], n, H[L].or_all, H[L].and_all, o.radix_bits, w, 0);
            if (out != res.raw[L] && n)
                HIPCHK(hipMemcpyAsync(res.raw[L], out, n * 8, hipMemcpyDeviceToDevice, s));
        } else {  // an explicit schedule sorts (and uniques) as it would the candidates
            sort_tmp.ensure(n * 8, s);
            dedup_into(s, u, (uint64_t *)sort_tmp.p, n, H[L].or_all, H[L].and_all, res.raw[L], &M[L].n_raw,
                       o.radix_bits, w, o.sort_schedule);
        }
    };
    // layer L's set dropped keys: regrow it, empty every set, run the steps
    // that feed it again from the sorted layers below, extract it again
    // (synchronous; rare). Leaves the state of a step whose set L+1 holds
    // exactly the spawn-4 children of step L.
    auto recover = [&](int L) {
        uint64_t cap = set_of(L).cap;
        submit();  // the steps run again from the sorted layers below
        HIPCHK(hipStreamSynchronize(s2));
        while (H[L].pad[0]) {
            HIPCHK(hipStreamSynchronize(s));
            cap *= 4;
            res.hash_regrows++;
            for (KeySet &k : sets)
                if (k.buf.p) HIPCHK(hipMemsetAsync(k.buf.p, 0, k.buf.bytes, s));
            set_prepare(set_of(L), cap);
            if (L + 1 <= steps - 1 && !set_of(L + 1).buf.p) set_prepare(set_of(L + 1), cap);
            for (int q = L; q <= std::min(L + 2, steps); ++q) {
                HIPCHK(hipMemsetAsync(&M[q].pad[0], 0, 8, s));
                if (q > L) HIPCHK(hipMemsetAsync(&M[q].tot, 0, 8, s));
            }
            HIPCHK(hipMemsetAsync(&M[L].n_raw, 0, 8, s));
            if (L + 1 <= steps) HIPCHK(hipMemsetAsync(&M[L + 1].n_raw, 0, 8, s));
            uint64_t *sink = (uint64_t *)cnt_dev.p;  // the totals of these steps are already known
            if (L >= 2) hash_step(L - 1, res.raw[L - 2], res.info[L - 2].n_raw, nullptr, &set_of(L), sink);
            if (L <= steps - 2)
                hash_step(L, res.raw[L - 1], res.info[L - 1].n_raw, &set_of(L), &set_of(L + 1), sink);
            extract(L);
            HIPCHK(hipEventSynchronize(hev[L & 1]));
        }
    };

    int first_sort_step = 1;
    bool hash_done = false;
    const bool hash_fwd = steps >= 3 && o.chunk_mkeys == 0 && o.forward_hash_steps >= 0 &&
                          o.forward_hash_steps != 1;
    if (hash_fwd) {
        res.info[0].n_raw = n_seeds;
        const uint64_t c1 = cap_for(1, 0), c2 = cap_for(2, 0);
        if (c1 + c2 <= budget_keys(c1 + c2) && std::max(c1, c2) <= max_slots) {
            queue_raw(0, n_seeds);
            if (streamer) streamer->reserve(n_seeds);
            set_prepare(set_of(1), c1);
            set_prepare(set_of(2), c2);
            hash_step(1, res.raw[0], n_seeds, &set_of(1), &set_of(2), &M[1].tot);
            extract(1);
            first_sort_step = 0;
            for (int i = 1; i <= steps - 2; ++i) {
                progress_set((uint32_t)i, prog_total);
                bool handover = false;
                bool queued = false;  // step i+1 (or the last layer's extraction) is queued
                if (i + 1 <= steps - 2) {
                    const uint64_t c = cap_for(i + 2, i - 1);
                    if ((o.forward_hash_steps > 0 && i + 1 >= o.forward_hash_steps) || c > max_slots ||
                        5 * c > budget_keys(5 * c)) {
                        handover = true;
                    } else {
                        set_prepare(set_of(i + 2), c);
                        hash_step(i + 1, (const uint64_t *)ubuf[i % kURing].p, ubuf[i % kURing].bytes / 8,
                                  &set_of(i + 1), &set_of(i + 2), &M[i + 1].tot);
                        extract(i + 1);
                        queued = true;
                    }
                } else {
                    extract(steps - 1);  // the last layer: spawn-4 children of step steps-2 only
                    queued = true;
                }
                wait_meta(i);
                if (H[i].pad[0]) {
                    recover(i);
                    if (queued) {  // queue again what the recovery undid
                        if (i + 1 <= steps - 2) {
                            hash_step(i + 1, (const uint64_t *)ubuf[i % kURing].p, ubuf[i % kURing].bytes / 8,
                                      &set_of(i + 1), &set_of(i + 2), &M[i + 1].tot);
                            extract(i + 1);
                        } else {
                            extract(steps - 1);
                        }
                    }
                }
                const uint64_t n = H[i].n_raw;
                res.hash_steps++;
                res.info[i].n_raw = n;
                res.info[i].c2 = H[i].tot & 0xFFFFFFFFull;
                res.info[i].c4 = H[i].tot >> 32;
                stats_add_bytes("expand", 8.0 * (double)res.info[i - 1].n_raw +
                                              8.0 * (double)(res.info[i].c2 + res.info[i].c4));
                check_step(i);
                sort_layer(i, n);
                if (handover) {
                    join_sorts();  // the sort loop reads the sorted layers on the build stream
                    // the sort loop continues at step i+1. Its candidate
                    // buffer starts with the spawn-4 children of step i:
                    // written here again by the one-pass expand from the
                    // sorted layer i-1 (set i+1 may have dropped some; its
                    // overflow is only looked at when that layer is due)
                    const uint64_t c4 = H[i].tot >> 32, c2 = H[i].tot & 0xFFFFFFFFull;
                    DBuf &cur = cand[(i + 1) & 1], &nxt = cand[i & 1];
                    const uint64_t room = std::max<uint64_t>(c4 + 2 * (c2 + c4) + 64, kEnsureMin / 8);
                    cur.alloc(room * 8 + 256, s);
                    nxt.alloc(room * 8 + 256, s);
                    for (KeySet &q : sets) q.buf.release();
                    for (DBuf &u : ubuf) u.release();
                    HIPCHK(hipMemsetAsync(cnt_dev.p, 0, 8, s));  // its write positions come from this total
                    {
                        Timer t(s, "expand", 0.0);
                        launch_expand_alloc(s, C.shape, res.raw[i - 1], res.info[i - 1].n_raw, &M[i - 1].n_raw,
                                            C.lut, p, i > o.docheck_step ? 1 : 0, (uint64_t *)cnt_dev.p, nullptr,
                                            (uint64_t *)cur.p, &M[i].or_all, &M[i + 1].or_all, 0, c4, b.ja(i));
                    }
                    c4_prev = c4;
                    const uint64_t cap2 = cur.bytes / 8 - c4, cap4 = nxt.bytes / 8;
                    launch_alloc(i + 1, n, (uint64_t *)cur.p + c4, (uint64_t *)nxt.p, cap2, cap4, false);
                    queue_meta(i, i + 2);
                    first_sort_step = i + 1;
                    break;
                }
                if (!defer) queue_raw(i, n);  // (a deferred sort queues it after sorting)
                if (streamer) streamer->reserve(n);
                if (defer && pending.size() >= kSortBatch) submit();
                HIPCHK(hipGetLastError());
            }
            if (first_sort_step == 0) {  // the last layer
                const int L = steps - 1;
                wait_meta(L);
                if (H[L].pad[0]) recover(L);
                res.info[L].n_raw = H[L].n_raw;
                sort_layer(L, H[L].n_raw);
                join_sorts();
                if (streamer) streamer->reserve(H[L].n_raw);
                for (KeySet &q : sets) q.buf.release();
                for (DBuf &u : ubuf) u.release();
                hash_done = true;
            }
        }
    }
    if (first_sort_step == 1 && steps >= 3) {
        launch_alloc(1, n_seeds, nullptr, nullptr, 0, 0, false);  // no buffers yet: totals and spans only
        queue_meta(0, 2);
    }
    for (int i = std::max(first_sort_step, 1); !hash_done && i <= steps - 2; ++i) {
        progress_set((uint32_t)i, prog_total);
        // poll: the host must wake as soon as the sizes land, while the
        // device is still busy with the speculative write
        for (hipError_t q; (q = hipEventQuery(meta_ev)) != hipSuccess;)
            if (q != hipErrorNotReady) HIPCHK(q);
        const uint64_t n_par = H[i - 1].n_raw;
        if (i > 1) res.raw_arena.trim_last(res.raw[i - 1], n_par * 8);
        res.info[i - 1].n_raw = n_par;
        queue_raw(i - 1, n_par);
        if (streamer) streamer->reserve(n_par);
        check_step(i);
        const uint64_t c2 = H[i].tot & 0xFFFFFFFFull, c4 = H[i].tot >> 32;
        res.info[i].c2 = c2;
        res.info[i].c4 = c4;
        const uint64_t n_c = c4_prev + c2;
        const int do_check = i > o.docheck_step ? 1 : 0;
        const uint64_t budget = budget_keys(std::max(n_c, c4));
        if (!chunked && (n_c > budget || c4 > budget)) {
            // switch to runs: the spawn-4 prefix already in `cur` becomes a run
            chunked = true;
            if (c4_prev)
                to_run((uint64_t *)cand[i & 1].p, c4_prev, H[i].or_all, H[i].and_all, runs_cur, runs_cur_n);
            cand[0].release();
            cand[1].release();
            sort_tmp.release();
            launch_count(i, n_par);  // per-parent offsets of this step for its chunk ranges
        }
        uint64_t bound_next;
        if (!chunked) {
            DBuf &cur = cand[i & 1], &nxt = cand[(i + 1) & 1];
            bool write = c2 > spec_cap2 || c4 > spec_cap4;  // the one-pass expand dropped writes
            // candidate buffers keep a quarter of headroom over this step so
            // the next speculative write usually fits, and grow by at least
            // half their size: a reallocation (with its copy) sits between
            // this step's sizes and its sort, so it must stay rare while the
            // layers grow a little every step
            const uint64_t room = n_c + n_c / 4 + 64;
            auto grown = [&](const DBuf &b) {
                return std::max<uint64_t>({room * 8 + 256, b.bytes + b.bytes / 2, kEnsureMin});
            };
            if (cur.bytes < n_c * 8) {  // grow, keeping the spawn-4 prefix
                DBuf g;
                g.alloc(grown(cur), s);
                if (c4_prev) HIPCHK(hipMemcpyAsync(g.p, cur.p, c4_prev * 8, hipMemcpyDeviceToDevice, s));
                cur = std::move(g);
                write = true;
            }
            if (nxt.bytes < c4 * 8) {
                nxt.alloc(std::max<uint64_t>(c4 * 8 + 256, grown(nxt)), s);
                write = true;
            } else if (nxt.bytes < room * 8) {  // headroom for the next step, keeping what is written
                DBuf g;
                g.alloc(grown(nxt), s);
                if (!write && c4) HIPCHK(hipMemcpyAsync(g.p, nxt.p, c4 * 8, hipMemcpyDeviceToDevice, s));
                nxt = std::move(g);
            }
            stats_add_bytes("expand", 8.0 * (double)n_par + 8.0 * (double)(c2 + c4));
            if (write && n_par && (c2 || c4))
                launch_alloc(i, n_par, (uint64_t *)cur.p + c4_prev, (uint64_t *)nxt.p, c2, c4, true);  // fits now
            // (the sort's second buffer stays out of the raw arena: a layer
            // must remain the arena's newest allocation to be trimmed to its
            // size, and a larger block per step would leave the arena's 1 GiB
            // chunks mostly unused on large layers)
            sort_tmp.ensure(n_c * 8, s);
            res.raw[i] = (uint64_t *)res.raw_arena.alloc(n_c * 8);
            dedup_into(s, (uint64_t *)cur.p, (uint64_t *)sort_tmp.p, n_c, H[i].or_all, H[i].and_all, res.raw[i],
                       &M[i].n_raw, o.radix_bits, w, o.sort_schedule);
            bound_next = n_c;
        } else {
            // parent ranges whose candidates fit the budget, from the prefix
            res.chunked_steps++;
            const JobArg ja_i = b.ja(i);
            uint32_t K =(uint32_t)std::max<uint64_t>(1, (std::max(c2, c4) + budget - 1) / budget);
            std::vector<uint64_t> bnd, ofs;
            for (;;) {
                bnd.assign(K + 1, 0);
                for (uint32_t k = 0; k <= K; ++k) bnd[k] = n_par * k / K;
                ofs.assign(K + 1, 0);
                gidx.ensure((K + 1) * 8, s);
                gval.ensure((K + 1) * 8, s);
                HIPCHK(hipMemcpyAsync(gidx.p, bnd.data(), K * 8, hipMemcpyHostToDevice, s));
                launch_gather_u64(s, (const uint64_t *)offs.p, (const uint64_t *)gidx.p, K, (uint64_t *)gval.p);
                HIPCHK(hipMemcpyAsync(ofs.data(), gval.p, K * 8, hipMemcpyDeviceToHost, s));
                HIPCHK(hipStreamSynchronize(s));
                ofs[K] = H[i].tot;
                bool fits = true;
                for (uint32_t k = 0; k < K; ++k) {
                    const uint64_t d = ofs[k + 1] - ofs[k];
                    fits &= (d & 0xFFFFFFFFull) <= budget && (d >> 32) <= budget;
                }
                if (fits || K >= n_par) break;
                K *= 2;
            }
            for (uint32_t k = 0; k < K; ++k) {
                const uint64_t a = bnd[k], b = bnd[k + 1];
                const uint64_t d = ofs[k + 1] - ofs[k];
                const uint64_t k2 = d & 0xFFFFFFFFull, k4 = d >> 32;
                if (a == b || (k2 == 0 && k4 == 0)) continue;
                res.chunks++;
                chunk2.ensure(k2 * 8 + 8, s);
                chunk4.ensure(k4 * 8 + 8, s);
                {
                    Timer t(s, "expand", 8.0 * (double)(b - a) + 8.0 * (double)(k2 + k4));
                    launch_expand(s, true, C.shape, res.raw[i - 1] + a, b - a, nullptr, C.lut, p, do_check, nullptr,
                                  (const uint64_t *)offs.p + a, (uint64_t *)chunk2.p, (uint64_t *)chunk4.p, nullptr,
                                  nullptr, ofs[k], ~0ull, ~0ull, ja_i);
                }
                to_run((uint64_t *)chunk2.p, k2, H[i].or_all, H[i].and_all, runs_cur, runs_cur_n);
                to_run((uint64_t *)chunk4.p, k4, H[i + 1].or_all, H[i + 1].and_all, runs_next, runs_next_n);
            }
            bound_next = merge_runs(i, runs_cur, runs_cur_n);
            runs_cur = std::move(runs_next);
            runs_cur_n = std::move(runs_next_n);
            runs_next.clear();
            runs_next_n.clear();
        }
        if (i + 1 <= steps - 2) {
            if (!chunked) {
                // step i+1 into the buffers as they are; it runs again if
                // its totals outgrow them
                const DBuf &cur1 = cand[(i + 1) & 1], &nxt1 = cand[i & 1];
                const uint64_t cap2 = cur1.bytes / 8 > c4 ? cur1.bytes / 8 - c4 : 0, cap4 = nxt1.bytes / 8;
                launch_alloc(i + 1, bound_next, cap2 ? (uint64_t *)cur1.p + c4 : nullptr, (uint64_t *)nxt1.p,
                             cap2, cap4, false);
            } else {
                launch_count(i + 1, bound_next);
            }
            queue_meta(i, i + 2);
        }
        HIPCHK(hipGetLastError());
        c4_prev = c4;
    }
    // layer steps-1: only the spawn-4 children of layer steps-3
    if (!hash_done) {
        const int L = steps - 1;
        read_meta(s, res, H, L - 1, L);
        res.raw_arena.trim_last(res.raw[L - 1], H[L - 1].n_raw * 8);
        res.info[L - 1].n_raw = H[L - 1].n_raw;
        queue_raw(L - 1, H[L - 1].n_raw);
        if (streamer) streamer->reserve(H[L - 1].n_raw);
        if (!chunked) {
            const uint64_t n_c = c4_prev;
            DBuf &cur = cand[L & 1];
            sort_tmp.ensure(n_c * 8, s);
            res.raw[L] = (uint64_t *)res.raw_arena.alloc(n_c * 8);
            dedup_into(s, (uint64_t *)cur.p, (uint64_t *)sort_tmp.p, n_c, H[L].or_all, H[L].and_all, res.raw[L],
                       &M[L].n_raw, o.radix_bits, w, o.sort_schedule);
        } else {
            merge_runs(L, runs_cur, runs_cur_n);
        }
        read_meta(s, res, H, L, L);
        res.raw_arena.trim_last(res.raw[L], H[L].n_raw * 8);
        res.info[L].n_raw = H[L].n_raw;
        if (streamer) streamer->reserve(H[L].n_raw);
    }
    chunk2.release();
    chunk4.release();
    sort_tmp.release();
    run_tmp.release();
    run_out.release();
    cand[0].release();
    cand[1].release();
    counts.release();
    offs.release();
    // every layer's key span (the OR / AND its forward steps folded) sizes
    // its compact solve hash; layer 0 (the seeds) gets no hash
    if (res.compact_hashes) {
        read_meta(s, res, H, 0, steps - 1);
        for (int i = 1; i < steps; ++i) {
            res.info[i].key_or = H[i].or_all;
            res.info[i].key_and = H[i].and_all;
        }
    }
    // the second stream's sorts carry their own look-back error flag (the
    // small unique layers are sorted there with the single-pass look-back):
    // a raised flag means a partial prefix was used
    if (w_side.read_error(s2)) throw Error(EGTB_ERR_INTERNAL, "a look-back kernel exceeded its spin limit");
}

// Final layers and the stats header of a build with files.
template <typename T>
static void final_layers(BuildCtx<T> &b) {
    hipStream_t s = b.s;
    const egtb_options &o = b.o;
    const PatternDev &p = b.p;
    Result &res = b.res;
    const int steps = b.steps;
    const bool files = b.files;
    const std::string &prefix = b.prefix;
    const T max_scale = b.max_scale;
    const T zero = b.zero;
    SortWork &w = b.w;
    HostStage &stage = b.stage;
    DBuf &rates = b.rates;
    DBuf &packed = b.packed;
    LayerDev *const H = b.H;
    std::vector<char> &have_book = b.have_book;
    std::unique_ptr<RawFileWriter> &raw_writer = b.raw_writer;
    // ---- final layers (reference final_steps, BookSolver.cpp:237-284)

    uint64_t &max_n = b.max_n;
    max_n = 0;
    for (int i = 0; i < steps; ++i) max_n = std::max(max_n, res.info[i].n_raw);
    rates.ensure(max_n * sizeof(T) + 8, s);
    for (int i = steps - 2; i <= steps - 1; ++i) {
        const uint64_t n = res.info[i].n_raw;
        launch_final_rates<T>(s, res.raw[i], n, p, max_scale, zero, (T *)rates.p, b.ja(i));
        make_book<T>(s, res, i, res.raw[i], (const T *)rates.p, n, w, true, zero);
    }
    if (raw_writer) raw_writer->finish();
    if (files) {
        read_meta(s, res, H, steps - 2, steps - 1);
        for (int i = steps - 2; i <= steps - 1; ++i)
            if (!have_book[i])
                write_book_file<T>(s, prefix + std::to_string(i) + ".book", res.books[i], H[i].n_live, stage,
                                   packed);
        if (o.write_raw_files != 2) remove_if_exists(prefix + std::to_string(steps - 2));
        if (!file_exists(prefix + "stats.txt")) {
            FILE *f = std::fopen((prefix + "stats.txt").c_str(), "a");
            if (!f) throw Error(EGTB_ERR_IO, "cannot open " + prefix + "stats.txt");
            std::fputs("layer,length,max success rate,speed,deletion_threshold,time\n", f);
            std::fclose(f);
        }
    }

}

// The backward pass, on two streams or one (the schedule b.two_stream
// names), then the checks of the flags its kernels raise.
template <typename T>
static void backward_pass(BuildCtx<T> &b) {
    DeviceCtx &C = b.C;
    hipStream_t s = b.s;
    const egtb_options &o = b.o;
    const PatternDev &p = b.p;
    Result &res = b.res;
    const int steps = b.steps;
    const bool files = b.files;
    const std::string &prefix = b.prefix;
    const T max_scale = b.max_scale;
    const T zero = b.zero;
    const double w4 = b.w4;
    const double w2 = b.w2;
    const uint32_t prog_total = b.prog_total;
    std::unique_ptr<BookStreamer> &streamer = b.streamer;
    SortWork &w = b.w;
    HostStage &stage = b.stage;
    DBuf &rates = b.rates;
    DBuf &packed = b.packed;
    LayerDev *const H = b.H;
    std::vector<char> &have_book = b.have_book;
    const bool two_stream = b.two_stream;
    const uint64_t max_n = b.max_n;
    // ---- backward (reference recalculate_process_impl, BookSolver.cpp:513-605)
    if (two_stream) {
        // layer i: the solve (build stream) reads the hashes of layers i+1
        // and i+2, fills layer i's (cleared by the solve of layer i+1) and
        // clears layer i-1's; then the compaction that makes book i runs on
        // the second stream, from this layer's own rates
        hipStream_t s2 = b.s2;
        std::vector<uint64_t> roff(steps + 1, 0);
        for (int i = 0; i < steps; ++i)
            roff[i + 1] = roff[i] + ((res.info[i].n_raw * sizeof(T) + 255) & ~(uint64_t)255);
        DBuf rates_all;
        rates_all.alloc(roff[steps] + 256, s);
        struct SideSync {  // on the way out: the second stream's compactions read rates_all
            hipStream_t side;
            ~SideSync() { (void)hipStreamSynchronize(side); }
        } side_sync{s2};
        // the look-back words at their largest size now: the second stream
        // must never see them reallocated
        w.ensure_lookback(s, 0, 0, compact_tiles(max_n ? max_n : 1));
        hipEvent_t fork_ev, join_ev;
        HIPCHK(hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming));
        EventGuard fork_guard{fork_ev};
        HIPCHK(hipEventCreateWithFlags(&join_ev, hipEventDisableTiming));
        EventGuard join_guard{join_ev};
        const int i0 = steps - 3;
        HashDev own = i0 > 0 ? hash_table_for(s, res, i0, res.info[i0].n_raw) : HashDev{};
        if (own.present()) HIPCHK(hipMemsetAsync(own.table(), 0, own.bytes(), s));
        for (int i = i0; i >= 0; --i) {
            progress_set(prog_total - (uint32_t)i - 2u, prog_total);
            const uint64_t n = res.info[i].n_raw;
            const Book &b1 = res.books[i + 1], &b2 = res.books[i + 2];
            const HashDev next = i - 1 > 0 ? hash_table_for(s, res, i - 1, res.info[i - 1].n_raw) : HashDev{};
            T *ri = (T *)((char *)rates_all.p + roff[i]);
            {
                Timer t(s, "solve", (double)n * (8 + sizeof(T)));
                launch_solve<T>(s, C.shape, res.raw[i], n, C.lut, p, i > o.docheck_step ? 1 : 0, hash_of(b1),
                                hash_of(b2), table_of<T>(b1), table_of<T>(b2), w2, w4, max_scale, zero, ri, next,
                                own, b.ja(i));
            }
            carve_book<T>(res, i, n);  // before the fork: a new arena chunk is allocated on the build stream
            HIPCHK(hipEventRecord(fork_ev, s));
            HIPCHK(hipStreamWaitEvent(s2, fork_ev, 0));
            compact_book<T>(s2, res, i, res.raw[i], ri, n, w, false, zero);
            drop_hash(res, i + 2);
            HIPCHK(hipGetLastError());
            own = next;
        }
        HIPCHK(hipEventRecord(join_ev, s2));
        HIPCHK(hipStreamWaitEvent(s, join_ev, 0));
        rates_all.release();
    }
    bool started = false;
    DBuf thr_keys, thr_rates, thr_cnt;
    for (int i = two_stream ? -1 : steps - 3; i >= 0; --i) {
        progress_set(prog_total - (uint32_t)i - 2u, prog_total);
        if (files && !started && have_book[i]) continue;  // solved by an earlier run
        if (!started) {
            started = true;
            if (files && i != steps - 3)
                for (int j = i + 1; j <= i + 2; ++j)
                    load_book_file<T>(s, res, j, prefix + std::to_string(j) + ".book", stage, packed, H);
        }
        const auto l0 = std::chrono::steady_clock::now();
        const uint64_t n = res.info[i].n_raw;
        const Book &b1 = res.books[i + 1], &b2 = res.books[i + 2];
        // layer i's solve hash (ring slot i % 3; its previous table, layer
        // i+3's, was last read by the solve of layer i+1) is cleared by this
        // solve and filled by the compaction that makes book i
        const HashDev hs = i > 0 ? hash_table_for(s, res, i, n) : HashDev{};
        {
            Timer t(s, "solve", (double)n * (8 + sizeof(T)));
            launch_solve<T>(s, C.shape, res.raw[i], n, C.lut, p, i > o.docheck_step ? 1 : 0, hash_of(b1),
                            hash_of(b2), table_of<T>(b1), table_of<T>(b2), w2, w4, max_scale, zero, (T *)rates.p,
                            hs, HashDev{}, b.ja(i));
        }
        make_book<T>(s, res, i, res.raw[i], (const T *)rates.p, n, w, i > 0, zero, hs.present());
        drop_hash(res, i + 2);  // its last reader was this layer's solve
        HIPCHK(hipGetLastError());
        if (streamer) {
            // book i+2 was last read by this layer's solve; raw layers above
            // i are consumed
            streamer->retire(i + 2, s);
            streamer->reclaim();
            if (!o.keep_raw) res.raw_arena.release_after(res.raw[i]);
            size_t fr = 0, tot = 0;
            if (hipMemGetInfo(&fr, &tot) == hipSuccess)
                res.peak_device_bytes =
                    std::max<uint64_t>(res.peak_device_bytes, tot - fr - device_idle_bytes(o.device));
        }
        if (files) {
            read_meta(s, res, H, i, i);
            const auto l1 = std::chrono::steady_clock::now();
            if (H[i].n_live) {
                const double dt = std::chrono::duration<double>(l1 - l0).count();
                char row[256];
                std::snprintf(row, sizeof row, "%d,%llu,%g,%g mbps,%g,%s\n", i, (unsigned long long)n,
                              rate_unit<T>(H[i].maxbits, zero, max_scale),
                              round_to_2((double)n / std::max(dt, 0.01) / 1e6),
                              o.deletion_threshold, now_string().c_str());
                FILE *f = std::fopen((prefix + "stats.txt").c_str(), "a");
                if (f) {
                    std::fputs(row, f);
                    std::fclose(f);
                }
            }
            write_book_file<T>(s, prefix + std::to_string(i) + ".book", res.books[i], H[i].n_live, stage, packed);
            if (o.write_raw_files != 2) remove_if_exists(prefix + std::to_string(i));
            if (o.deletion_threshold > 0.0) {
                // book i+2 has had its last reader: its file keeps rates > threshold
                // (reference BookSolver.cpp:584-588)
                const int j = i + 2;
                const uint64_t nj = H[j].n_live;
                // the threshold in the rate type's own scale (BookSolver.cpp:590)
                const T thr = static_cast<T>(o.deletion_threshold * static_cast<double>(max_scale - zero) +
                                             static_cast<double>(zero));
                thr_keys.ensure(nj * 8 + 8, s);
                thr_rates.ensure(nj * sizeof(T) + 8, s);
                thr_cnt.ensure(16, s);
                w.ensure_lookback(s, 0, 0, compact_tiles(nj ? nj : 1));
                launch_compact<T>(s, res.books[j].keys, (const T *)res.books[j].rates, nj, (uint64_t *)thr_keys.p,
                                  (T *)thr_rates.p, (unsigned long long *)thr_cnt.p + 1, (uint64_t *)thr_cnt.p, w.lw,
                                  thr);
                uint64_t kept = 0;
                HIPCHK(hipMemcpyAsync(&H[steps].n_live, thr_cnt.p, 8, hipMemcpyDeviceToHost, s));
                HIPCHK(hipStreamSynchronize(s));
                kept = H[steps].n_live;
                Book tb;
                tb.keys = (uint64_t *)thr_keys.p;
                tb.rates = thr_rates.p;
                write_book_file<T>(s, prefix + std::to_string(j) + ".book", tb, kept, stage, packed);
            }
        }
    }
    if (steps > 1) drop_hash(res, 1);
    for (DBuf &h : res.hash_bufs) h.release();  // the solve is done with them
    if (streamer) {
        streamer->retire(1, s);
        streamer->retire(0, s);
        streamer->finish();
        // the build's chunks go back to the device cache (idle, for later
        // builds)
        res.book_ring.free_all();
        if (!o.keep_raw) res.raw_arena.release();
        HIPCHK(hipStreamSynchronize(s));
    }
    // the build stream's look-back error flag (the second stream's is read
    // at the end of the forward pass): a raised flag means a partial prefix
    // was used
    if (w.read_error(s)) throw Error(EGTB_ERR_INTERNAL, "a look-back kernel exceeded its spin limit");
    {
        uint32_t herr = 0;
        HIPCHK(hipMemcpyAsync(&herr, res.hash_err.p, 4, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (herr)
            throw Error(EGTB_ERR_INTERNAL, "a compact solve hash had no free slot within 32 buckets (flags " +
                                               std::to_string(herr) +
                                               "); build again with solve_hash_layout = 0 (wide tables)");
    }
    read_meta(s, res, H, 0, steps - 1);

}

template <typename T>
static void optimal_branch_pass(BuildCtx<T> &b) {
    DeviceCtx &C = b.C;
    hipStream_t s = b.s;
    const egtb_options &o = b.o;
    const PatternDev &p = b.p;
    Result &res = b.res;
    const int steps = b.steps;
    const bool files = b.files;
    const std::string &prefix = b.prefix;
    const T zero = b.zero;
    SortWork &w = b.w;
    HostStage &stage = b.stage;
    DBuf &packed = b.packed;
    LayerDev *const H = b.H;
    // ---- optimal branches only (reference keep_only_optimal_branches_impl,
    // BookSolver.cpp:687-735): from layer 21 up, book i keeps the records
    // that some state of book i-2 (spawning a 4) or of book i-1 (spawning a
    // 2) reaches as its best move, the parents being the already filtered
    // books.  With files the pass reads the books from their files (so a
    // deletion threshold applies first, as in the reference) and keeps its
    // place in <pathname>optlayer for a resumed run; without files it filters
    // the books in HBM.
    if (o.optimal_branch_only && steps > 21) {
        const std::string optpath = prefix + "optlayer";
        int done = -1;  // layers up to here were filtered by an earlier run
        if (files) {
            if (FILE *f = std::fopen(optpath.c_str(), "r")) {
                int v = 0;
                if (std::fscanf(f, "%d", &v) == 1) done = v;
                std::fclose(f);
            }
        }
        DBuf mark, cnt;
        cnt.alloc(64, s);
        bool started = false;
        for (int i = 21; i < steps; ++i) {
            if (!started) {
                if (done >= i) continue;
                if (files) {
                    const std::string p0 = prefix + std::to_string(i - 2) + ".book";
                    const std::string p1 = prefix + std::to_string(i - 1) + ".book";
                    if (!file_exists(p0) || !file_exists(p1)) continue;
                    load_book_file<T>(s, res, i - 2, p0, stage, packed, H, false);
                    load_book_file<T>(s, res, i - 1, p1, stage, packed, H, false);
                }
                started = true;
            }
            if (files) {
                FILE *f = std::fopen(optpath.c_str(), "w");
                if (!f) throw Error(EGTB_ERR_IO, "cannot write " + optpath);
                std::fprintf(f, "%d", i - 1);
                std::fclose(f);
                load_book_file<T>(s, res, i, prefix + std::to_string(i) + ".book", stage, packed, H, false);
            }
            Book &bt = res.books[i];
            const uint64_t nt = H[i].n_live;
            if (!bt.idx) {  // in-HBM books of hashed builds carry no bucket index
                bt.bits = index_bits_for(nt);
                bt.idx = (uint32_t *)res.book_arena.alloc(((1ull << bt.bits) + 1) * 4);
                launch_bucket_index(s, bt.keys, nt, nullptr, bt.bits, bt.idx);
            }
            mark.ensure(nt + 64, s);
            HIPCHK(hipMemsetAsync(mark.p, 0, nt + 64, s));
            TableDev<T> t{bt.keys, (const T *)bt.rates, bt.idx, nullptr, nt, bt.bits, 1};
            launch_opt_mark<T>(s, C.shape, res.books[i - 2].keys, H[i - 2].n_live, C.lut, p, 2, t, zero,
                               (uint8_t *)mark.p);
            launch_opt_mark<T>(s, C.shape, res.books[i - 1].keys, H[i - 1].n_live, C.lut, p, 1, t, zero,
                               (uint8_t *)mark.p);
            uint64_t *keys2 = (uint64_t *)res.book_arena.alloc(nt * 8 + 8);
            T *rates2 = (T *)res.book_arena.alloc(nt * sizeof(T) + 8);
            w.ensure_lookback(s, 0, 0, compact_tiles(nt ? nt : 1));
            launch_compact<T>(s, bt.keys, (const T *)bt.rates, nt, keys2, rates2, (unsigned long long *)cnt.p + 1,
                              &res.meta_dev()[i].n_live, w.lw, zero, HashDev{}, false, (const uint8_t *)mark.p);
            HIPCHK(hipGetLastError());
            read_meta(s, res, H, i, i);
            bt.keys = keys2;
            bt.rates = rates2;
            bt.n_dev = &res.meta_dev()[i].n_live;
            bt.idx = nullptr;  // the index was over the unfiltered keys
            if (files)
                write_book_file<T>(s, prefix + std::to_string(i) + ".book", bt, H[i].n_live, stage, packed);
        }
        if (files) remove_if_exists(optpath);
        if (w.read_error(s)) throw Error(EGTB_ERR_INTERNAL, "a look-back kernel exceeded its spin limit");
    }
}

template <typename T>
static void compress_pass(BuildCtx<T> &b) {
    const egtb_options &o = b.o;
    const int steps = b.steps;
    const bool files = b.files;
    const std::string &prefix = b.prefix;
    HostStage &stage = b.stage;
    // ---- compressed books (reference maybe_do_compress_classic,
    // CompressionBridge.cpp:41-60): every final .book above 2 MiB, except
    // layers 0-2, becomes a <pathname><i>.z directory (trie index + xz
    // blocks) and the .book is removed
    if (files && o.compress) {
        constexpr size_t rec = 8 + sizeof(T);
        for (int i = 3; i < steps; ++i) {
            const std::string bp = prefix + std::to_string(i) + ".book";
            struct stat st;
            if (stat(bp.c_str(), &st) != 0 || (uint64_t)st.st_size <= 2097152ull || st.st_size % rec) continue;
            const size_t n = (size_t)st.st_size / rec;
            stage.ensure((size_t)st.st_size + 16);
            FILE *f = std::fopen(bp.c_str(), "rb");
            if (!f) throw Error(EGTB_ERR_IO, "cannot open " + bp);
            const size_t got = std::fread(stage.p, 1, (size_t)st.st_size, f);
            std::fclose(f);
            if (got != (size_t)st.st_size) throw Error(EGTB_ERR_IO, "short read of " + bp);
            std::vector<uint64_t> keys(n);
            std::vector<uint8_t> rts(n * sizeof(T));
            for (size_t r = 0; r < n; ++r) {
                std::memcpy(&keys[r], (const char *)stage.p + r * rec, 8);
                std::memcpy(rts.data() + r * sizeof(T), (const char *)stage.p + r * rec + 8, sizeof(T));
            }
            zbook_write(bp, keys.data(), rts.data(), n, (int)sizeof(T));
            remove_if_exists(bp);
        }
    }
}

// One build: set-up and resume scan, the forward pass, the choice of the
// backward schedule, the final layers, the backward pass, the optional
// post-solve passes (optimal branches, compression) and the timings.
template <typename T>
void run_build(const uint64_t *seeds, size_t n_seeds, const PatternDev &p, const egtb_options &o, Result &res,
               bool one_minus, const JobDev *jobs = nullptr, uint32_t *job_err = nullptr) {
    BuildCtx<T> b(o, p, res, one_minus);
    b.jobs = jobs;
    b.job_err = job_err;
    if (jobs && (o.write_files || o.host_tables || o.optimal_branch_only || o.deletion_threshold > 0.0 ||
                 o.compress || o.write_raw_files))
        throw Error(EGTB_ERR_ARG, "batched builds keep their books in HBM: no files, host tables or pruning");
    hipStream_t s = b.s;
    const int steps = b.steps;
    const bool files = b.files;
    const T zero = b.zero;
    progress_set(0, b.prog_total);

    res.dev = o.device;
    res.steps = steps;
    res.rate_size = sizeof(T);
    res.rate_is_float = std::is_floating_point<T>::value;
    res.raw.assign(steps, nullptr);
    res.books.assign(steps, Book{});
    res.info.assign(steps, LayerInfo{});
    // the arenas' chunks come from the stream-ordered pool, which keeps the
    // memory earlier builds freed: a plain hipMalloc of a 1 GiB chunk after a
    // build that freed tens of GB into the pool stalled for seconds (3.2 s
    // and 1.6 s in single calls) while the driver reclaimed that memory
    res.raw_arena.set_stream(s);
    res.book_arena.set_stream(s);
    res.host_tables = o.host_tables != 0;
    if (res.host_tables) {
        if (files) throw Error(EGTB_ERR_ARG, "host_tables and write_files cannot be combined");
        res.book_ring.reset(s, steps);
        res.host_keys.clear();
        res.host_rates.clear();
        res.host_keys.resize(steps);
        res.host_rates.resize(steps);
        b.streamer = std::make_unique<BookStreamer>(res, (int)sizeof(T), steps);
    }
    res.hash_bufs.clear();
    res.hash_bufs.resize(4);
    res.hash_ring = 3;
    res.hash_max_bytes = o.solve_hash_mb < 0 ? 0.0
                         : (double)(o.solve_hash_mb ? o.solve_hash_mb : kDefaultSolveHashMB) * (1 << 20);
    res.rate_size = (int)sizeof(T);
    res.rate_zero = static_cast<double>(zero);
    if (o.optimal_branch_only && o.host_tables)
        throw Error(EGTB_ERR_ARG, "optimal_branch_only needs the books in HBM (host_tables = 0)");
    // compact tables are 4-8x smaller but take more probe rounds per
    // lookup: 442@512 k_solve 101.6 us a launch against 72 us wide
    res.compact_hashes = o.solve_hash_layout == 1;
    res.hash_err.alloc(256, s);
    HIPCHK(hipMemsetAsync(res.hash_err.p, 0, 256, s));

    b.t_begin = std::chrono::steady_clock::now();

    // per-layer device records: zero sizes, OR = 0, AND = all ones
    b.meta_host_buf.ensure((size_t)(steps + 1) * sizeof(LayerDev));
    LayerDev *H = b.H = (LayerDev *)b.meta_host_buf.p;
    for (int i = 0; i <= steps; ++i) {
        std::memset(&H[i], 0, sizeof(LayerDev));
        H[i].and_all = ~0ull;
    }
    H[0].n_raw = n_seeds;
    res.meta.alloc((size_t)(steps + 1) * sizeof(LayerDev), s);
    HIPCHK(hipMemcpyAsync(res.meta.p, H, (size_t)(steps + 1) * sizeof(LayerDev), hipMemcpyHostToDevice, s));

    // resume (reference handle_restart / handle_restart_recalculate): drop
    // invalid leftovers, note which layers are already solved on disk
    if (files) {
        for (int i = 0; i < steps; ++i) {
            const std::string bp = b.prefix + std::to_string(i) + ".book";
            drop_invalid_restart_file(bp, 8 + sizeof(T));
            drop_invalid_restart_file(b.prefix + std::to_string(i), 8);
            b.have_book[i] = valid_restart_file(bp, 8 + sizeof(T)) ? 1 : 0;
        }
        if (o.write_raw_files) b.raw_writer = std::make_unique<RawFileWriter>(o.device);
    }
    b.M = res.meta_dev();

    forward_pass(b, seeds, n_seeds);
    b.t_fwd = std::chrono::steady_clock::now();

    // when every layer gets a solve hash the solve never searches, and the
    // books go without the bucket index (a table set over the result makes
    // its own)
    {
        bool all_hashed = res.hash_max_bytes > 0;
        for (int i = 1; i < steps && all_hashed; ++i) all_hashed = hash_layout(res, i, res.info[i].n_raw).present();
        res.book_index = !all_hashed;
    }

    // Two streams for the backward pass of an in-HBM build: each solve
    // fills its own layer's hash, so the compaction that makes the book
    // leaves the chain of solves and runs on the second stream; every layer
    // keeps its own rates until its compaction has read them.
    uint64_t rates_all_bytes = 0;
    for (int i = 0; i < steps; ++i) rates_all_bytes += (res.info[i].n_raw * sizeof(T) + 255) & ~(uint64_t)255;
    // the per-layer rates must also fit in half the free HBM, so a build that
    // fits with one stream never fails for want of the second's buffers
    size_t hbm_free = 0, hbm_total = 0;
    HIPCHK(hipMemGetInfo(&hbm_free, &hbm_total));
    hbm_free += device_idle_bytes(o.device);
    b.two_stream = steps >= 3 && !files && !b.streamer && !res.book_index && o.deletion_threshold <= 0.0 &&
                   (double)rates_all_bytes <= kTwoStreamRatesMax && rates_all_bytes + 256 <= (uint64_t)hbm_free / 2;
    if (b.two_stream) res.hash_ring = 4;

    final_layers(b);
    backward_pass(b);
    optimal_branch_pass(b);
    compress_pass(b);
    for (int i = 0; i < steps; ++i) {
        res.info[i].n_live = H[i].n_live;
        res.books[i].n = H[i].n_live;
    }
    {
        // each solve launch streams the two tables it looks up in once
        double table_bytes = 0;
        for (int i = 0; i <= steps - 3; ++i)
            table_bytes += (double)(H[i + 1].n_live + H[i + 2].n_live) * (8 + sizeof(T));
        stats_add_bytes("solve", table_bytes);
    }
    const auto t_end = std::chrono::steady_clock::now();
    if (!o.keep_raw) {
        res.raw_arena.release();
        res.raw.assign(steps, nullptr);
    }
    if (!res.keep_tables) res.book_arena.release();
    res.fwd_ms = std::chrono::duration<double, std::milli>(b.t_fwd - b.t_begin).count();
    res.bwd_ms = std::chrono::duration<double, std::milli>(t_end - b.t_fwd).count();
    res.total_ms = std::chrono::duration<double, std::milli>(t_end - b.t_begin).count();
    progress_set(b.prog_total, b.prog_total);
}

void run_build_jobs(const uint64_t *keys, size_t n_keys, const egtb_options &o, const JobDev *jobs,
                    uint32_t *job_err, Result &res) {
    PatternDev unused{};
    unused.any_success = 1;
    run_build<float>(keys, n_keys, unused, o, res, false, jobs, job_err);
}

void build_dispatch(const uint64_t *seeds, size_t n_seeds, const egtb_pattern *pattern, const egtb_options *o,
                    Result &res) {
    if (!o) throw Error(EGTB_ERR_ARG, "options is NULL");
    if (!seeds && n_seeds) throw Error(EGTB_ERR_ARG, "seeds is NULL");
    if (o->steps < 3) throw Error(EGTB_ERR_ARG, "steps must be at least 3");
    if (o->write_files && !o->pathname) throw Error(EGTB_ERR_ARG, "pathname is NULL");
    if (!(o->spawn_rate4 >= 0.0 && o->spawn_rate4 <= 1.0)) throw Error(EGTB_ERR_ARG, "spawn_rate4 out of [0, 1]");
    const PatternDev p = make_pattern(pattern, o->target);
    const std::string dt = o->success_rate_dtype ? o->success_rate_dtype : "float32";
    // the reference's rate types (success_rate_kind_from_name,
    // FormationRuntime.h:256-270); 1-float32 / 1-float64 store rate - 1
    if (dt == "float32") run_build<float>(seeds, n_seeds, p, *o, res, false);
    else if (dt == "float64") run_build<double>(seeds, n_seeds, p, *o, res, false);
    else if (dt == "uint32") run_build<uint32_t>(seeds, n_seeds, p, *o, res, false);
    else if (dt == "uint64") run_build<uint64_t>(seeds, n_seeds, p, *o, res, false);
    else if (dt == "1-float32") run_build<float>(seeds, n_seeds, p, *o, res, true);
    else if (dt == "1-float64") run_build<double>(seeds, n_seeds, p, *o, res, true);
    else throw Error(EGTB_ERR_ARG, "success_rate_dtype '" + dt + "' is not supported");
}

}  // namespace egtb

// ===========================================================================
// C ABI
// ===========================================================================

using namespace egtb;

struct egtb_result : Result {};

template <typename F>
static int guarded(F &&f) {
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

extern "C" {

const char *egtb_last_error(void) { return g_last_error.c_str(); }

void egtb_get_build_progress(uint32_t *current, uint32_t *total) {
    uint32_t t = g_prog_total.load(std::memory_order_relaxed);
    uint32_t c = g_prog_cur.load(std::memory_order_relaxed);
    if (c > t) t = c;
    if (current) *current = c;
    if (total) *total = t;
}

int egtb_build(const uint64_t *seeds, size_t n_seeds, const egtb_pattern *pattern, const egtb_options *options) {
    return guarded([&] {
        Result r;
        r.keep_tables = false;
        build_dispatch(seeds, n_seeds, pattern, options, r);
    });
}

int egtb_build_ex(const uint64_t *seeds, size_t n_seeds, const egtb_pattern *pattern, const egtb_options *options,
                  egtb_result **out) {
    if (!out) {
        g_last_error = "out is NULL";
        return EGTB_ERR_ARG;
    }
    *out = nullptr;
    auto r = std::make_unique<egtb_result>();
    r->keep_tables = true;
    const int rc = guarded([&] { build_dispatch(seeds, n_seeds, pattern, options, *r); });
    if (rc == EGTB_OK) *out = r.release();
    return rc;
}

int egtb_result_steps(const egtb_result *r) { return r ? r->steps : 0; }

int egtb_result_layer_info(const egtb_result *r, int layer, uint64_t *n_raw, uint64_t *n_live, uint64_t *c2,
                           uint64_t *c4) {
    if (!r || layer < 0 || layer >= r->steps) {
        g_last_error = "layer out of range";
        return EGTB_ERR_ARG;
    }
    const LayerInfo &li = r->info[layer];
    if (n_raw) *n_raw = li.n_raw;
    if (n_live) *n_live = li.n_live;
    if (c2) *c2 = li.c2;
    if (c4) *c4 = li.c4;
    return EGTB_OK;
}

int egtb_result_raw(const egtb_result *r, int layer, uint64_t *keys, size_t cap, size_t *n) {
    return guarded([&] {
        if (!r || layer < 0 || layer >= r->steps) throw Error(EGTB_ERR_ARG, "layer out of range");
        const uint64_t m = r->info[layer].n_raw;
        if (n) *n = m;
        if (m && !r->raw[layer]) throw Error(EGTB_ERR_ARG, "raw layers were not kept (options.keep_raw)");
        if (!keys) return;
        if (cap < m) throw Error(EGTB_ERR_ARG, "buffer too small");
        if (m) HIPCHK(hipMemcpy(keys, r->raw[layer], m * 8, hipMemcpyDeviceToHost));
    });
}

int egtb_result_book(const egtb_result *r, int layer, uint64_t *keys, void *rates, size_t cap, size_t *n) {
    return guarded([&] {
        if (!r || layer < 0 || layer >= r->steps) throw Error(EGTB_ERR_ARG, "layer out of range");
        const Book &b = r->books[layer];
        if (n) *n = b.n;
        if (!keys && !rates) return;
        if (cap < b.n) throw Error(EGTB_ERR_ARG, "buffer too small");
        if (r->host_tables) {
            if (b.n && keys) std::memcpy(keys, r->host_keys[layer].p, b.n * 8);
            if (b.n && rates) std::memcpy(rates, r->host_rates[layer].p, b.n * r->rate_size);
            return;
        }
        if (b.n && keys) HIPCHK(hipMemcpy(keys, b.keys, b.n * 8, hipMemcpyDeviceToHost));
        if (b.n && rates) HIPCHK(hipMemcpy(rates, b.rates, b.n * r->rate_size, hipMemcpyDeviceToHost));
    });
}

int egtb_result_memory(const egtb_result *r, uint64_t *peak_device_bytes, uint64_t *streamed_bytes,
                       double *stream_ms, uint64_t *chunked_steps, uint64_t *chunks) {
    if (!r) return EGTB_ERR_ARG;
    if (peak_device_bytes) *peak_device_bytes = r->peak_device_bytes;
    if (streamed_bytes) *streamed_bytes = r->streamed_bytes;
    if (stream_ms) *stream_ms = r->stream_ms;
    if (chunked_steps) *chunked_steps = r->chunked_steps;
    if (chunks) *chunks = r->chunks;
    return EGTB_OK;
}

int egtb_result_forward(const egtb_result *r, uint64_t *hash_steps, uint64_t *set_regrows) {
    if (!r) return EGTB_ERR_ARG;
    if (hash_steps) *hash_steps = r->hash_steps;
    if (set_regrows) *set_regrows = r->hash_regrows;
    return EGTB_OK;
}

int egtb_result_timing(const egtb_result *r, double *fwd_ms, double *bwd_ms, double *total_ms) {
    if (!r) return EGTB_ERR_ARG;
    if (fwd_ms) *fwd_ms = r->fwd_ms;
    if (bwd_ms) *bwd_ms = r->bwd_ms;
    if (total_ms) *total_ms = r->total_ms;
    return EGTB_OK;
}

void egtb_result_free(egtb_result *r) {
    if (!r) return;
    try {
        (void)hipSetDevice(r->dev);
        (void)hipDeviceSynchronize();
        delete r;
    } catch (...) {
    }
}

void sort_uint64(uint64_t *arr, size_t arrsize, bool descending) {
    if (!arr || arrsize < 2) return;
    const int rc = guarded([&] {
        DeviceCtx &C = device_ctx(0);
        hipStream_t s = C.stream;
        DBuf a, t;
        a.alloc(arrsize * 8, s);
        t.alloc(arrsize * 8, s);
        HIPCHK(hipMemcpyAsync(a.p, arr, arrsize * 8, hipMemcpyHostToDevice, s));
        SortWork w;
        // full 64-bit span: no key-range assumption for arbitrary input
        uint64_t *out = radix_sort(s, (uint64_t *)a.p, (uint64_t *)t.p, arrsize, ~0ull, 0ull, 0, w, 0);
        HIPCHK(hipMemcpyAsync(arr, out, arrsize * 8, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (w.read_error(s)) throw Error(EGTB_ERR_INTERNAL, "a look-back kernel exceeded its spin limit");
    });
    if (rc != EGTB_OK) {
        std::fprintf(stderr, "sort_uint64: %s\n", g_last_error.c_str());
        std::abort();
    }
    if (descending)
        for (size_t i = 0, j = arrsize - 1; i < j; ++i, --j) std::swap(arr[i], arr[j]);
}

size_t unique_sorted_u64_inplace(uint64_t *data, size_t size) {
    if (!data || size < 2) return size;
    size_t m = 0;
    const int rc = guarded([&] {
        DeviceCtx &C = device_ctx(0);
        hipStream_t s = C.stream;
        DBuf a, out;
        a.alloc(size * 8, s);
        HIPCHK(hipMemcpyAsync(a.p, data, size * 8, hipMemcpyHostToDevice, s));
        SortWork w;
        out.alloc(size * 8, s);
        DBuf cnt;
        cnt.alloc(8, s);
        unique_into(s, (const uint64_t *)a.p, size, (uint64_t *)out.p, (uint64_t *)cnt.p, w, 0);
        uint64_t mm = 0;
        HIPCHK(hipMemcpyAsync(&mm, cnt.p, 8, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (mm) HIPCHK(hipMemcpyAsync(data, out.p, mm * 8, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        m = mm;
    });
    if (rc != EGTB_OK) {
        std::fprintf(stderr, "unique_sorted_u64_inplace: %s\n", g_last_error.c_str());
        std::abort();
    }
    return m;
}

int egtb_bench_sort(const uint64_t *keys, size_t n, int bits, int schedule, int reps, double *ms_per_sort) {
    return guarded([&] {
        if (!keys || n == 0 || reps <= 0 || !ms_per_sort) throw Error(EGTB_ERR_ARG, "invalid sort benchmark");
        DeviceCtx &C = device_ctx(0);
        hipStream_t s = C.stream;
        DBuf src, a, t, u, cnt;
        src.alloc(n * 8, s);
        a.alloc(n * 8, s);
        t.alloc(n * 8, s);
        u.alloc(n * 8, s);
        cnt.alloc(8, s);
        HIPCHK(hipMemcpyAsync(src.p, keys, n * 8, hipMemcpyHostToDevice, s));
        uint64_t o = 0, an = ~0ull;
        for (size_t i = 0; i < n; ++i) {
            o |= keys[i];
            an &= keys[i];
        }
        SortWork w;
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        double total = 0;
        for (int r = 0; r <= reps; ++r) {  // first round warms up
            HIPCHK(hipMemcpyAsync(a.p, src.p, n * 8, hipMemcpyDeviceToDevice, s));
            HIPCHK(hipEventRecord(e0, s));
            dedup_into(s, (uint64_t *)a.p, (uint64_t *)t.p, n, o, an, (uint64_t *)u.p, (uint64_t *)cnt.p, bits, w,
                       schedule);
            HIPCHK(hipEventRecord(e1, s));
            HIPCHK(hipEventSynchronize(e1));
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            if (r) total += ms;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (w.read_error(s)) throw Error(EGTB_ERR_INTERNAL, "a look-back kernel exceeded its spin limit");
        *ms_per_sort = total / reps;
    });
}

int egtb_zbook_read(const char *book_path, int rate_size, uint64_t *keys, void *rates, size_t cap, size_t *n) {
    return guarded([&] {
        if (!book_path || !n) throw Error(EGTB_ERR_ARG, "book_path and n must not be NULL");
        if (rate_size != 4 && rate_size != 8) throw Error(EGTB_ERR_ARG, "rate_size must be 4 or 8");
        const std::string bp = book_path;
        if (bp.size() < 5 || bp.compare(bp.size() - 5, 5, ".book") != 0)
            throw Error(EGTB_ERR_ARG, "book_path must end in .book");
        ZBook z = zbook_read(bp, rate_size);
        *n = z.keys.size();
        if (!keys && !rates) return;
        if (cap < z.keys.size()) throw Error(EGTB_ERR_ARG, "buffers hold fewer records than the book");
        if (keys && !z.keys.empty()) std::memcpy(keys, z.keys.data(), z.keys.size() * 8);
        if (rates && !z.rates.empty()) std::memcpy(rates, z.rates.data(), z.rates.size());
    });
}

int egtb_zbook_write(const char *book_path, const uint64_t *keys, const void *rates, size_t n, int rate_size) {
    return guarded([&] {
        if (!book_path || !keys || !rates) throw Error(EGTB_ERR_ARG, "book_path, keys and rates must not be NULL");
        if (rate_size != 4 && rate_size != 8) throw Error(EGTB_ERR_ARG, "rate_size must be 4 or 8");
        const std::string bp = book_path;
        if (bp.size() < 5 || bp.compare(bp.size() - 5, 5, ".book") != 0)
            throw Error(EGTB_ERR_ARG, "book_path must end in .book");
        for (size_t r = 1; r < n; ++r)
            if (keys[r] <= keys[r - 1]) throw Error(EGTB_ERR_ARG, "compressed books hold strictly ascending keys");
        zbook_write(bp, keys, rates, n, rate_size);
    });
}

int egtb_sort_unique_u64(const uint64_t *keys, size_t n, int schedule, uint64_t *out, size_t *n_out) {
    return guarded([&] {
        if ((n && (!keys || !out)) || !n_out) throw Error(EGTB_ERR_ARG, "invalid dedup arguments");
        if (schedule < 0 || schedule > 3) throw Error(EGTB_ERR_ARG, "schedule must be 0..3");
        *n_out = 0;
        if (n == 0) return;
        DeviceCtx &C = device_ctx(0);
        hipStream_t s = C.stream;
        DBuf a, t, u, cnt;
        a.alloc(n * 8, s);
        t.alloc(n * 8, s);
        u.alloc(n * 8, s);
        cnt.alloc(8, s);
        HIPCHK(hipMemcpyAsync(a.p, keys, n * 8, hipMemcpyHostToDevice, s));
        uint64_t o = 0, an = ~0ull;
        for (size_t i = 0; i < n; ++i) {
            o |= keys[i];
            an &= keys[i];
        }
        SortWork w;
        dedup_into(s, (uint64_t *)a.p, (uint64_t *)t.p, n, o, an, (uint64_t *)u.p, (uint64_t *)cnt.p, 0, w, schedule);
        uint64_t m = 0;
        HIPCHK(hipMemcpyAsync(&m, cnt.p, 8, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (m) HIPCHK(hipMemcpy(out, u.p, m * 8, hipMemcpyDeviceToHost));
        if (w.read_error(s)) throw Error(EGTB_ERR_INTERNAL, "a look-back kernel exceeded its spin limit");
        *n_out = m;
    });
}

int egtb_set_sort_tuning(const uint64_t *max_n, const int32_t *bits, const int32_t *schedule, int rows) {
    if (rows < 0 || (rows > 0 && (!max_n || !bits || !schedule))) {
        g_last_error = "invalid tuning table";
        return EGTB_ERR_ARG;
    }
    std::vector<TuneRow> t;
    for (int i = 0; i < rows; ++i) {
        if (i && max_n[i] <= max_n[i - 1]) {
            g_last_error = "tuning rows must have ascending max_n";
            return EGTB_ERR_ARG;
        }
        t.push_back(TuneRow{max_n[i], bits[i], schedule[i]});
    }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune = std::move(t);
    return EGTB_OK;
}

int egtb_set_tile_tuning(int kind, const uint64_t *max_n, const int32_t *value, int rows) {
    if (kind < 0 || kind >= kTileKinds || rows < 0 || (rows > 0 && (!max_n || !value))) {
        g_last_error = "invalid tile tuning table";
        return EGTB_ERR_ARG;
    }
    std::vector<std::pair<uint64_t, int>> t;
    for (int i = 0; i < rows; ++i) {
        if (i && max_n[i] <= max_n[i - 1]) {
            g_last_error = "tuning rows must have ascending max_n";
            return EGTB_ERR_ARG;
        }
        const int v = value[i];
        const bool ok = kind == kTileRadixItems ? (v == 0 || v == 4 || v == 8 || v == 16)
                                                : (v == 0 || (v >= 64 && v <= 1024 && v % 64 == 0));
        if (!ok) {
            g_last_error = kind == kTileRadixItems ? "radix keys per thread must be 4, 8 or 16 (0 = rule)"
                                                   : "lookup threads per block must be a multiple of 64 in "
                                                     "64..1024 (0 = rule)";
            return EGTB_ERR_ARG;
        }
        t.emplace_back(max_n[i], v);
    }
    set_tile_tuning(kind, std::move(t));
    return EGTB_OK;
}

int egtb_set_key_set_sizing(double slots_per_key, uint64_t max_slots) {
    if (!(slots_per_key >= 0.0 && slots_per_key <= 64.0)) {
        g_last_error = "slots per key must lie in (0, 64], or 0 for the default";
        return EGTB_ERR_ARG;
    }
    if (max_slots > ((uint64_t)1 << 34)) {
        g_last_error = "max_slots must be at most 2^34, or 0 for the default";
        return EGTB_ERR_ARG;
    }
    g_set_slots_per_key.store(slots_per_key > 0.0 ? slots_per_key : kSetSlotsPerKey);
    g_set_max_slots.store(max_slots ? max_slots : kSetMaxSlots);
    return EGTB_OK;
}

void egtb_stats_enable(int on) { g_stats_on.store(on ? 1 : 0); }

void egtb_stats_reset(void) {
    std::lock_guard<std::mutex> lk(g_stats_mu);
    resolve_pending_locked();
    g_stats.clear();
}

int egtb_stats_get(const char *kernel, uint64_t *launches, double *total_ms, double *alg_bytes) {
    std::lock_guard<std::mutex> lk(g_stats_mu);
    resolve_pending_locked();
    auto it = g_stats.find(kernel ? kernel : "");
    if (it == g_stats.end()) {
        if (launches) *launches = 0;
        if (total_ms) *total_ms = 0;
        if (alg_bytes) *alg_bytes = 0;
        return EGTB_ERR_ARG;
    }
    if (launches) *launches = it->second.launches;
    if (total_ms) *total_ms = it->second.ms;
    if (alg_bytes) *alg_bytes = it->second.bytes;
    return EGTB_OK;
}

}  // extern "C"
