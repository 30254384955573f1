// Solved tables resident in HBM and batched position lookups.
//
// A table set holds every solved layer of one build (pattern, target) as
// sorted keys + rates + bucket index.  A lookup answers the reference's
// ClassicBookReader::move_on_dic (ReaderRuntime.cpp:739-840) for a batch of
// boards: layer = (tile sum - seed tile sum) / 2; the first symmetry op
// (0..7) under which the board is in the pattern and has a valid move gives
// the four directional rates (NaN = invalid move), remapped to absolute
// directions.  Boards that no op answers get the heuristic expectimax
// player's scores instead (status EGTB_LOOKUP_FALLBACK): the reference hands
// such positions to its AIPlayer (engine_core/AIPlayer.py:273-291 ->
// AIPlayer::start_search); here its search at depth 1 -- two moves deep:
// move, spawn, move, evaluate -- runs in the lookup's own thread.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "egtb.h"
#include "egtb_ai.h"
#include "egtb_device.h"
#include "egtb_kernels.h"
#include "egtb_runtime.h"
#include "egtb_serving.h"
#include "egtb_zbook.h"

namespace egtb {

// absolute direction (left, right, up, down) of move d of the board under
// symmetry op `op` (reference adjust_direction, ReaderRuntime.cpp:149-169)
__constant__ uint8_t c_absdir[8][4] = {{0, 1, 2, 3}, {2, 3, 1, 0}, {1, 0, 3, 2}, {3, 2, 0, 1},
                                       {1, 0, 2, 3}, {3, 2, 1, 0}, {0, 1, 3, 2}, {2, 3, 0, 1}};

// The reader's board operations (apply_operation, ReaderRuntime.cpp:127-147)
// on the nibble encoding: 1 rotate left, 2 rotate 180, 3 rotate right,
// 4 flip horizontal, 5-7 flip horizontal after 1-3.
__device__ __forceinline__ uint64_t apply_op(uint64_t b, int op) {
    switch (op) {
        case 1: return mirror_ud(transpose(b));
        case 2: return mirror_ud(mirror_lr(b));
        case 3: return mirror_lr(transpose(b));
        case 4: return mirror_lr(b);
        case 5: return mirror_lr(mirror_ud(transpose(b)));
        case 6: return mirror_ud(b);
        case 7: return transpose(b);
        default: return b;
    }
}

__device__ __forceinline__ uint32_t tile_sum(uint64_t b) {
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const uint32_t v = (b >> (4 * k)) & 0xF;
        s += v ? (1u << v) : 0u;
    }
    return s;
}

template <typename T, typename OutT>
__global__ __launch_bounds__(1024) void k_lookup(const uint64_t *__restrict__ boards, uint64_t n,
                                                 const TableDev<T> *__restrict__ tabs, int steps,
                                                 int64_t nums_adjust, PatternDev p, double scale, double zero,
                                                 double p4,
                                                 int prefer_max, int round9, const uint16_t *__restrict__ glut,
                                                 AiTables ai, int64_t w4, OutT *__restrict__ out4,
                                                 int32_t *__restrict__ status, int32_t *__restrict__ op_out) {
    extern __shared__ uint16_t lut[];
    stage_lut(lut, glut);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) {
        const uint64_t b = boards[q];
        const OutT nanv = (OutT)NAN;
        OutT v[4] = {nanv, nanv, nanv, nanv};
        int st = EGTB_LOOKUP_FALLBACK, used = -1;
        const int64_t nums = ((int64_t)tile_sum(b) + nums_adjust) / 2;
        if (nums >= 0 && nums < steps && tabs[nums].present) {
            const TableDev<T> &t = tabs[nums];
            double best = 0.0;
            for (int op = 0; op < 8 && (prefer_max || used < 0); ++op) {
                const uint64_t x = apply_op(b, op);
                if (!is_pattern(x, p)) continue;
                uint64_t mv[4];
                move4(x, lut, mv);
                bool any = false;
                double top = -1.0;
                OutT r[4] = {nanv, nanv, nanv, nanv};
#pragma unroll
                for (int d = 0; d < 4; ++d) {
                    if (mv[d] != x && is_pattern(mv[d], p)) {
                        // a present key reads as its normalised rate, an absent
                        // one as the type's zero (find_success_entry_in_file,
                        // ReaderRuntime.cpp:270-316)
                        T raw;
                        const double val = table_find(t, t.n, canonical(mv[d], p.symm), raw) ? (double)raw / scale
                                                                                                  : zero;
                        r[c_absdir[op][d]] = (OutT)val;
                        // compared as the reader reports them: 4-byte rates
                        // rounded to 1e-9 (maybe_round_value, ReaderRuntime.cpp:208-214,
                        // whose release build multiplies by 1e-9 rather than divides)
                        top = fmax(top, (round9 && fabs(val) > 1e-7) ? round(val * 1e9) * 1e-9 : val);
                        any = true;
                    }
                }
                if (prefer_max ? (any && top > best) : any) {
                    best = top;
#pragma unroll
                    for (int d = 0; d < 4; ++d) v[d] = r[d];
                    used = op;
                    st = EGTB_LOOKUP_OK;
                }
            }
        }
        if (used < 0) {
            // the player's four top scores; NaN where the move does not
            // change the queried board
            st = EGTB_LOOKUP_FALLBACK;
            int32_t sc[4];
            ai_start_search<1>(b, ai, 0, 0, w4, sc);
            uint64_t mv[4];
            move4(b, lut, mv);
#pragma unroll
            for (int d = 0; d < 4; ++d)
                if (mv[d] != b) v[d] = (OutT)sc[d];
        }
#pragma unroll
        for (int d = 0; d < 4; ++d) out4[q * 4 + d] = v[d];
        if (status) status[q] = st;
        if (op_out) op_out[q] = used;
    }
}

// Moves and symmetry reduction of a batch of boards: host-side preparation
// (free-pattern seeds) and tests use the same device code as the build.
__global__ __launch_bounds__(256) void k_board_ops(const uint64_t *__restrict__ boards, uint64_t n,
                                                   const uint16_t *__restrict__ lut, uint32_t symm,
                                                   uint64_t *__restrict__ out4, uint64_t *__restrict__ canon) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t b = boards[i];
    if (out4) {
        uint64_t m[4];
        move4(b, lut, m);
#pragma unroll
        for (int d = 0; d < 4; ++d) out4[i * 4 + d] = m[d];
    }
    if (canon) canon[i] = canonical(b, symm);
}

void configure_serving_kernels() {
    const int lds = kLutEntries * sizeof(uint16_t);
    (void)hipFuncSetAttribute((const void *)k_lookup<float, float>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    (void)hipFuncSetAttribute((const void *)k_lookup<float, double>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    (void)hipFuncSetAttribute((const void *)k_lookup<double, float>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    (void)hipFuncSetAttribute((const void *)k_lookup<double, double>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              lds);
    (void)hipFuncSetAttribute((const void *)k_lookup<uint32_t, float>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              lds);
    (void)hipFuncSetAttribute((const void *)k_lookup<uint32_t, double>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    (void)hipFuncSetAttribute((const void *)k_lookup<uint64_t, float>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              lds);
    (void)hipFuncSetAttribute((const void *)k_lookup<uint64_t, double>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, lds);
}

template <typename T, typename OutT>
static void launch_lookup_t(hipStream_t s, const TableSet &ts, const DeviceCtx &C, const uint64_t *boards,
                            uint64_t n, OutT *out4, int32_t *status, int32_t *op_out) {
    // the 128 KiB LDS row table allows one block per CU: about n / CUs
    // threads each (whole wavefronts, up to 1024), so a small batch still
    // spreads over every CU, unless the autotuner installed block sizes
    const uint64_t per = (n + C.cus - 1) / (uint64_t)C.cus;
    const int tuned = tuned_tile(kTileLookupThreads, n);
    const uint32_t threads = tuned ? (uint32_t)tuned
                                   : (uint32_t)std::min<uint64_t>(1024, std::max<uint64_t>(64, (per + 63) / 64 * 64));
    uint64_t want = (n + threads - 1) / threads;
    const uint64_t cap = (uint64_t)C.cus;
    const uint32_t grid = (uint32_t)(want < cap ? want : cap);
    // the player's spawn weight of a 4, truncated to 1/65536 (AIPlayer.cpp:636-641)
    const int64_t w4 = (int64_t)(ts.p4 * 65536.0);
    hipLaunchKernelGGL((k_lookup<T, OutT>), dim3(grid ? grid : 1), dim3(threads), kLutEntries * sizeof(uint16_t), s,
                       boards, n, (const TableDev<T> *)ts.tabs.p, ts.steps, ts.nums_adjust, ts.pattern, ts.scale,
                       ts.zero, ts.p4, ts.prefer_max, ts.rate_size == 4 ? 1 : 0, C.lut, ts.ai, w4, out4, status,
                       op_out);
}

template <typename OutT>
void launch_lookup(hipStream_t s, const TableSet &ts, const uint64_t *boards, uint64_t n, OutT *out4,
                   int32_t *status, int32_t *op_out) {
    if (n == 0) return;
    const DeviceCtx &C = device_ctx(ts.dev);
    switch (ts.rate_size * 10 + (ts.is_float ? 1 : 0)) {
        case 41: launch_lookup_t<float, OutT>(s, ts, C, boards, n, out4, status, op_out); break;
        case 81: launch_lookup_t<double, OutT>(s, ts, C, boards, n, out4, status, op_out); break;
        case 40: launch_lookup_t<uint32_t, OutT>(s, ts, C, boards, n, out4, status, op_out); break;
        default: launch_lookup_t<uint64_t, OutT>(s, ts, C, boards, n, out4, status, op_out); break;
    }
}

template void launch_lookup<float>(hipStream_t, const TableSet &, const uint64_t *, uint64_t, float *, int32_t *,
                                   int32_t *);
template void launch_lookup<double>(hipStream_t, const TableSet &, const uint64_t *, uint64_t, double *,
                                    int32_t *, int32_t *);

// ---------------------------------------------------------------------------
// Table sets
// ---------------------------------------------------------------------------

struct DTypeDesc {
    int size;
    bool is_float;
    double scale;  // reader normalisation (ReaderRuntime.cpp:183-206,271-277)
    double zero;   // what a missing key reads as (dtype_info_for_name's zero_value)
};

static DTypeDesc dtype_desc(const std::string &dt) {
    if (dt == "float32") return {4, true, 1.0, 0.0};
    if (dt == "float64") return {8, true, 1.0, 0.0};
    if (dt == "uint32") return {4, false, 4e9, 0.0};
    if (dt == "uint64") return {8, false, 1.6e18, 0.0};
    if (dt == "1-float32") return {4, true, 1.0, -1.0};
    if (dt == "1-float64") return {8, true, 1.0, -1.0};
    throw Error(EGTB_ERR_ARG, "success_rate_dtype '" + dt + "' is not supported");
}

static void finish_tables(TableSet &ts, hipStream_t s, const std::vector<TableDev<char>> &host_tabs) {
    ts.tabs.alloc(host_tabs.size() * sizeof(TableDev<char>), s);
    HIPCHK(hipMemcpyAsync(ts.tabs.p, host_tabs.data(), host_tabs.size() * sizeof(TableDev<char>),
                          hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
}

static uint32_t table_index_bits(uint64_t n) {
    uint32_t b = 1;
    while (b < 24 && (1ull << (b + 3)) < n) ++b;
    return b;
}

std::unique_ptr<TableSet> table_from_files(const char *prefix, int steps, const egtb_pattern *pat, int target,
                                           int64_t tile_sum, const char *dtype, double p4, int device) {
    if (!prefix) throw Error(EGTB_ERR_ARG, "prefix is NULL");
    if (steps < 1) throw Error(EGTB_ERR_ARG, "steps must be positive");
    const DTypeDesc dd = dtype_desc(dtype ? dtype : "float32");
    DeviceCtx &C = device_ctx(device);
    configure_serving_kernels();
    hipStream_t s = C.stream;
    auto ts = std::make_unique<TableSet>();
    ts->dev = device;
    ts->ai = ai_tables(device);
    ts->steps = steps;
    ts->pattern = make_pattern(pat, target);
    ts->nums_adjust = -tile_sum;
    ts->rate_size = dd.size;
    ts->is_float = dd.is_float;
    ts->scale = dd.scale;
    ts->zero = dd.zero;
    ts->p4 = p4;
    ts->n.assign(steps, 0);
    const size_t rec = 8 + dd.size;
    std::vector<TableDev<char>> tabs(steps);
    HostStage stage;
    DBuf recs;
    for (int i = 0; i < steps; ++i) {
        TableDev<char> &t = tabs[i];
        std::memset(&t, 0, sizeof t);
        t.bits = 1;
        const std::string path = std::string(prefix) + std::to_string(i) + ".book";
        FILE *f = std::fopen(path.c_str(), "rb");
        if (!f && zbook_exists(path)) {
            // a compressed layer (<pathname><i>.z): decoded on the host, then
            // resident in HBM like a .book (ReaderRuntime.cpp:318-476 reads
            // it block by block per probe instead)
            t.present = 1;
            ZBook z = zbook_read(path, dd.size);
            const uint64_t n = z.keys.size();
            ts->n[i] = n;
            if (n == 0) continue;
            uint64_t *keys = (uint64_t *)ts->storage.alloc(n * 8);
            void *rates = ts->storage.alloc(n * dd.size);
            HIPCHK(hipMemcpyAsync(keys, z.keys.data(), n * 8, hipMemcpyHostToDevice, s));
            HIPCHK(hipMemcpyAsync(rates, z.rates.data(), n * dd.size, hipMemcpyHostToDevice, s));
            const uint32_t bits = table_index_bits(n);
            uint32_t *idx = (uint32_t *)ts->storage.alloc(((1ull << bits) + 1) * 4);
            launch_bucket_index(s, keys, n, nullptr, bits, idx);
            t.keys = keys;
            t.rates = (const char *)rates;
            t.idx = idx;
            t.n_ptr = nullptr;
            t.n = n;
            t.bits = bits;
            HIPCHK(hipStreamSynchronize(s));  // the host vectors go out of scope
            continue;
        }
        if (!f) continue;  // absent layer: lookups there read as off-table
        t.present = 1;
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
        ts->n[i] = n;
        if (n == 0) continue;
        recs.ensure((size_t)sz, s);
        HIPCHK(hipMemcpyAsync(recs.p, stage.p, (size_t)sz, hipMemcpyHostToDevice, s));
        uint64_t *keys = (uint64_t *)ts->storage.alloc(n * 8);
        void *rates = ts->storage.alloc(n * dd.size);
        launch_unpack_records(s, (const uint32_t *)recs.p, n, dd.size, keys, rates);
        const uint32_t bits = table_index_bits(n);
        uint32_t *idx = (uint32_t *)ts->storage.alloc(((1ull << bits) + 1) * 4);
        launch_bucket_index(s, keys, n, nullptr, bits, idx);
        t.keys = keys;
        t.rates = (const char *)rates;
        t.idx = idx;
        t.n_ptr = nullptr;
        t.n = n;
        t.bits = bits;
        HIPCHK(hipStreamSynchronize(s));  // the staging buffer is reused
    }
    finish_tables(*ts, s, tabs);
    ts->bytes = ts->storage.used();
    return ts;
}

std::unique_ptr<TableSet> table_from_result(const Result &r, const egtb_pattern *pat, int target, int64_t tile_sum,
                                            double p4) {
    DeviceCtx &C = device_ctx(r.dev);
    configure_serving_kernels();
    auto ts = std::make_unique<TableSet>();
    ts->dev = r.dev;
    ts->ai = ai_tables(r.dev);
    ts->steps = r.steps;
    ts->pattern = make_pattern(pat, target);
    ts->nums_adjust = -tile_sum;
    ts->rate_size = r.rate_size;
    ts->is_float = r.rate_is_float;
    ts->scale = r.rate_is_float ? 1.0 : (r.rate_size == 4 ? 4e9 : 1.6e18);
    ts->zero = r.rate_zero;
    ts->p4 = p4;
    ts->borrowed = &r;
    ts->n.assign(r.steps, 0);
    std::vector<TableDev<char>> tabs(r.steps);
    for (int i = 0; i < r.steps; ++i) {
        const Book &bk = r.books[i];
        TableDev<char> &t = tabs[i];
        std::memset(&t, 0, sizeof t);
        t.bits = 1;
        ts->n[i] = bk.n;
        t.present = 1;
        if (bk.n == 0) continue;
        t.keys = bk.keys;
        t.rates = (const char *)bk.rates;
        if (bk.idx) {
            t.idx = bk.idx;
            t.bits = bk.bits;
        } else {  // layer 0 has no index in a build; give it one
            const uint32_t bits = table_index_bits(bk.n);
            uint32_t *idx = (uint32_t *)ts->storage.alloc(((1ull << bits) + 1) * 4);
            launch_bucket_index(C.stream, bk.keys, bk.n, nullptr, bits, idx);
            t.idx = idx;
            t.bits = bits;
        }
        t.n = bk.n;
        ts->bytes += bk.n * (8 + r.rate_size);
    }
    finish_tables(*ts, C.stream, tabs);
    ts->bytes += ts->storage.used();
    return ts;
}

// ---------------------------------------------------------------------------
// Allocation-free lookups: staging in a file-scope device array
// ---------------------------------------------------------------------------

// The staging of egtb_lookup_batch_noalloc: boards (8 B), four float rates
// (16 B) and a status word (4 B) per query, for at most kNoAllocMaxBoards
// queries (28 MiB).  Every GPU the library runs on holds its own copy until
// the process ends.  In host code the array's name is not a device address:
// kernels and copies use the address hipGetSymbolAddress gives per device.
constexpr size_t kNoAllocMaxBoards = (size_t)1 << 20;
constexpr size_t kNoAllocBoardsOff = 0;
constexpr size_t kNoAllocOutOff = kNoAllocMaxBoards * 8;
constexpr size_t kNoAllocStatusOff = kNoAllocOutOff + kNoAllocMaxBoards * 16;
constexpr size_t kNoAllocBytes = kNoAllocStatusOff + kNoAllocMaxBoards * 4;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

struct NoAllocStage {
    std::mutex mu;            // calls on one device run one at a time
    unsigned char *ptr = nullptr;  // device address of `scratch` on this device
};
static std::mutex g_noalloc_mu;
static std::map<int, std::unique_ptr<NoAllocStage>> g_noalloc;

// Set-up, once per device, with that device current: look up the array's
// device address.  Allocates nothing on the device.
static NoAllocStage &noalloc_stage(int dev) {
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    auto it = g_noalloc.find(dev);
    if (it != g_noalloc.end()) return *it->second;
    HIPCHK(hipSetDevice(dev));
    auto st = std::make_unique<NoAllocStage>();
    void *ptr = nullptr;
    HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) throw Error(EGTB_ERR_HIP, "hipGetSymbolAddress gave no address for the lookup staging");
    st->ptr = (unsigned char *)ptr;
    NoAllocStage &ref = *st;
    g_noalloc[dev] = std::move(st);
    return ref;
}

}  // namespace egtb

// ===========================================================================
// C ABI
// ===========================================================================

using namespace egtb;

struct egtb_table : TableSet {};

template <typename F>
static int guarded_s(F &&f) {
    try {
        g_last_error.clear();
        f();
        return EGTB_OK;
    } catch (const Error &e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return EGTB_ERR_INTERNAL;
    }
}

extern "C" {

int egtb_table_open(const char *pathname, int steps, const egtb_pattern *pattern, int target, int64_t tile_sum,
                    const char *success_rate_dtype, double spawn_rate4, int device, egtb_table **out) {
    if (!out) return EGTB_ERR_ARG;
    *out = nullptr;
    return guarded_s([&] {
        auto ts = table_from_files(pathname, steps, pattern, target, tile_sum, success_rate_dtype, spawn_rate4,
                                   device);
        *out = static_cast<egtb_table *>(ts.release());
    });
}

int egtb_table_from_result(const egtb_result *result, const egtb_pattern *pattern, int target, int64_t tile_sum,
                           double spawn_rate4, egtb_table **out) {
    if (!out || !result) return EGTB_ERR_ARG;
    *out = nullptr;
    return guarded_s([&] {
        auto ts = table_from_result(*reinterpret_cast<const Result *>(result), pattern, target, tile_sum,
                                    spawn_rate4);
        *out = static_cast<egtb_table *>(ts.release());
    });
}

int egtb_table_info(const egtb_table *t, int *steps, uint64_t *records, uint64_t *bytes) {
    if (!t) return EGTB_ERR_ARG;
    if (steps) *steps = t->steps;
    uint64_t r = 0;
    for (uint64_t v : t->n) r += v;
    if (records) *records = r;
    if (bytes) *bytes = t->bytes;
    return EGTB_OK;
}

int egtb_lookup_batch(egtb_table *t, const uint64_t *boards, size_t n, float *out4, int32_t *status) {
    return guarded_s([&] {
        if (!t) throw Error(EGTB_ERR_ARG, "table is NULL");
        if (n == 0) return;
        DeviceCtx &C = device_ctx(t->dev);
        hipStream_t s = C.stream;
        DBuf b, o, st;
        b.alloc(n * 8, s);
        o.alloc(n * 16, s);
        st.alloc(n * 4, s);
        HIPCHK(hipMemcpyAsync(b.p, boards, n * 8, hipMemcpyHostToDevice, s));
        launch_lookup<float>(s, *t, (const uint64_t *)b.p, n, (float *)o.p, (int32_t *)st.p, nullptr);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(out4, o.p, n * 16, hipMemcpyDeviceToHost, s));
        if (status) HIPCHK(hipMemcpyAsync(status, st.p, n * 4, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    });
}

int egtb_lookup_batch_noalloc(egtb_table *t, const uint64_t *boards, size_t n, float *out4, int32_t *status) {
    return guarded_s([&] {
        if (!t) throw Error(EGTB_ERR_ARG, "table is NULL");
        if (n == 0) return;
        if (!boards || !out4) throw Error(EGTB_ERR_ARG, "boards and out4 must not be NULL");
        if (n > kNoAllocMaxBoards)
            throw Error(EGTB_ERR_ARG, "egtb_lookup_batch_noalloc takes at most " + std::to_string(kNoAllocMaxBoards) +
                                          " boards per call (got " + std::to_string(n) + ")");
        DeviceCtx &C = device_ctx(t->dev);  // the table made it: no allocation here
        NoAllocStage &st = noalloc_stage(t->dev);
        std::lock_guard<std::mutex> lk(st.mu);
        hipStream_t s = C.stream;
        uint64_t *d_boards = (uint64_t *)(st.ptr + kNoAllocBoardsOff);
        float *d_out = (float *)(st.ptr + kNoAllocOutOff);
        int32_t *d_status = (int32_t *)(st.ptr + kNoAllocStatusOff);
        // every query's four rates and status are written by the kernel, so
        // nothing left by the previous call needs clearing
        HIPCHK(hipMemcpyAsync(d_boards, boards, n * 8, hipMemcpyHostToDevice, s));
        launch_lookup<float>(s, *t, d_boards, n, d_out, d_status, nullptr);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(out4, d_out, n * 16, hipMemcpyDeviceToHost, s));
        if (status) HIPCHK(hipMemcpyAsync(status, d_status, n * 4, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    });
}

int egtb_lookup_batch_f64(egtb_table *t, const uint64_t *boards, size_t n, double *out4, int32_t *status,
                          int32_t *op_used) {
    return guarded_s([&] {
        if (!t) throw Error(EGTB_ERR_ARG, "table is NULL");
        if (n == 0) return;
        DeviceCtx &C = device_ctx(t->dev);
        hipStream_t s = C.stream;
        DBuf b, o, st, op;
        b.alloc(n * 8, s);
        o.alloc(n * 32, s);
        st.alloc(n * 4, s);
        op.alloc(n * 4, s);
        HIPCHK(hipMemcpyAsync(b.p, boards, n * 8, hipMemcpyHostToDevice, s));
        launch_lookup<double>(s, *t, (const uint64_t *)b.p, n, (double *)o.p, (int32_t *)st.p, (int32_t *)op.p);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(out4, o.p, n * 32, hipMemcpyDeviceToHost, s));
        if (status) HIPCHK(hipMemcpyAsync(status, st.p, n * 4, hipMemcpyDeviceToHost, s));
        if (op_used) HIPCHK(hipMemcpyAsync(op_used, op.p, n * 4, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    });
}

int egtb_table_set_prefer_max(egtb_table *t, int32_t on) {
    if (!t) return EGTB_ERR_ARG;
    t->prefer_max = on ? 1 : 0;
    return EGTB_OK;
}

int egtb_board_moves(const uint64_t *boards, size_t n, int32_t symm_mode, uint64_t *out4, uint64_t *canon,
                     int device) {
    return guarded_s([&] {
        if (n && !boards) throw Error(EGTB_ERR_ARG, "boards is NULL");
        if (symm_mode < 0 || symm_mode > 3) throw Error(EGTB_ERR_ARG, "symm_mode must be 0..3");
        if (n == 0 || (!out4 && !canon)) return;
        DeviceCtx &C = device_ctx(device);
        hipStream_t s = C.stream;
        DBuf b, o, c;
        b.alloc(n * 8, s);
        if (out4) o.alloc(n * 32, s);
        if (canon) c.alloc(n * 8, s);
        HIPCHK(hipMemcpyAsync(b.p, boards, n * 8, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_board_ops, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (const uint64_t *)b.p,
                           (uint64_t)n, C.lut, (uint32_t)symm_mode, (uint64_t *)o.p, (uint64_t *)c.p);
        HIPCHK(hipGetLastError());
        if (out4) HIPCHK(hipMemcpyAsync(out4, o.p, n * 32, hipMemcpyDeviceToHost, s));
        if (canon) HIPCHK(hipMemcpyAsync(canon, c.p, n * 8, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    });
}

void egtb_table_free(egtb_table *t) {
    if (!t) return;
    (void)hipSetDevice(t->dev);
    (void)hipDeviceSynchronize();
    delete t;
}

}  // extern "C"
