// Advanced ("AD") forward pass: layer-by-layer enumeration of masked boards.
//
// The reference's advanced builder (native_core/src/BookGeneratorAD.cpp)
// stores every tile of 64 or more as the masked value 0xF, so one stored
// board stands for all placements of its large tiles.  A move whose merge
// reaches 64 makes a new masked tile; such a child goes through derive
// (BookGeneratorAD.cpp:435-509), which drops boards whose large tiles could
// not sum to the layer's total or already hold the target, and spells out
// equal pairs (or three 64s) of large tiles so that their merges stay
// possible.  Every 32 steps both layers at hand are validated
// (BoardMaskerAD.cpp:422-456).  The layers written here are the reference's
// raw layer files <pathname><i> byte for byte.
//
// Layout and schedule, per forward step i (parents = layer i-1, sorted;
// layer i = spawn-2 boards of step i with the spawn-4 boards of step i-1,
// which wait for step i+1 in its key set or as a unique list):
//   key sets (default): layer i's open-addressing set holds the spawn-4
//      boards of step i-1 (the set carried from that step, or their list
//      inserted) and takes the step's spawn-2 boards; layer i+1's set takes
//      its spawn-4 boards and is carried to step i+1 (k_ad_expand<kHash>:
//      one thread per parent, the masked row table in LDS; exact totals and
//      OR/AND spans on the side). Layer i's set is extracted (unique, any
//      order) and sorted. Sets are sized from the
//      last step's growth at two slots per key, up to 2^25 slots; a step
//      whose set would be larger, or overflowed, takes the sort path
//      (a carried set is drained into the list first; validation steps
//      drain it too):
//   sort + unique: k_ad_expand<kCount> (per parent the number of spawn-2 and
//      spawn-4 boards, packed c2 | c4 << 32), an exclusive scan,
//      k_ad_expand<kWrite> (the boards at their offsets), then the classic
//      path's radix sort + look-back unique. Steps whose boards exceed the
//      candidate budget run over parent ranges, each range deduplicated into
//      a run, and the runs merged by one more sort + unique.
// On validation steps k_ad_validate + scan + k_ad_scatter filter both lists.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <sys/stat.h>
#include <string>
#include <vector>

#include "egtb_ad.h"
#include "egtb_kernels.h"

namespace egtb {
namespace {

// Masked row slide toward the high nibble (merge_line,
// BoardMoverAD.cpp:47-86): equal neighbours merge once from the move side,
// 0xF never merges, a merge of 64 or more becomes 0xF.
__global__ void k_build_ad_lut(uint16_t *lut) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= (uint32_t)kLutEntries) return;
    int out[4] = {0, 0, 0, 0}, m = 0, pend = 0;
    for (int i = 0; i < 4; ++i) {
        const int x = (int)((r >> (4 * (3 - i))) & 0xF);
        if (!x) continue;
        if (pend && pend == x && x != 0xF) {
            out[m++] = x + 1 >= 6 ? 0xF : x + 1;
            pend = 0;
        } else {
            if (pend) out[m++] = pend;
            pend = x;
        }
    }
    if (pend) out[m++] = pend;
    lut[r] = (uint16_t)((out[0] << 12) | (out[1] << 8) | (out[2] << 4) | out[3]);
}

// derive (BookGeneratorAD.cpp:435-509) on a canonical child that made a new
// masked tile, with tile_sum_and_32k_count3's statistics
// (BoardMaskerAD.cpp:190-216): cells above 32 off the pinned ones are the
// large cells, the others sum to `total`.  put(x) takes the boards it keeps:
// the child itself when nothing needs spelling out, else the spelled-out
// boards that still match the pattern, canonicalised.
template <class Put>
__device__ __forceinline__ void derive_emit(uint64_t b, uint32_t board_sum, const AdSpecDev &s,
                                            const uint8_t *__restrict__ comb, Put &&put) {
    uint32_t total = 0;
    int count = 0, n64 = 0, pos64 = 0;
    uint64_t large = 0;  // 0xF at every large cell
    for (int k = 0; k < 16; ++k) {
        const int v = (int)((b >> (4 * k)) & 0xF);
        if (v > 5) {
            if (((s.fixed_mask >> (4 * k)) & 0xF) == 0) {
                ++count;
                large |= 0xFull << (4 * k);
                if (v == 6) {
                    ++n64;
                    pos64 = 4 * k;
                }
            }
        } else if (v > 0) {
            total += 1u << v;
        }
    }
    if (total >= s.limit + 64u) return;
    const uint32_t lsum = board_sum - total - ((uint32_t)(s.n_free + s.n_fixed) << 15);
    const uint32_t idx = (uint32_t)(uint8_t)(lsum >> 6) + ((uint32_t)(uint8_t)(count - s.n_free) << 8);
    if (idx >= (uint32_t)kCombEntries) return;
    const uint8_t *c = comb + idx * kCombStride;
    const int n = c[0];
    if (n == 0 || c[n] == s.target) return;
    if (count - s.n_free < 2) {
        put(b);
        return;
    }
    const uint64_t c0 = c[1];
    auto spelled = [&](uint64_t x) {
        if (ad_is_pattern(x, s)) put(canonical(x, s.symm));
    };
    if (c[1] == c[2]) {
        if (total >= s.limit) return;
        if (n > 2 && c[1] == c[3]) {
            if (total >= s.limit - 64u) return;
            if (n64 == 0) return;
            // derive_3x64 (:421-433): each large cell not already 64 set to 64
            for (uint64_t m = large; m;) {
                const int p = __builtin_ctzll(m) & ~3;
                m &= ~(0xFull << p);
                const uint64_t v = (b >> p) & 0xF;
                if (v != 6) spelled(b ^ ((v ^ 6ull) << p));
            }
            return;
        }
        // two equal large tiles: every pair of large cells set to that tile
        for (uint64_t mp = large; mp;) {
            const int p = __builtin_ctzll(mp) & ~3;
            mp &= ~(0xFull << p);
            const uint64_t dp = (((b >> p) & 0xF) ^ c0) << p;
            for (uint64_t mq = mp; mq;) {
                const int q = __builtin_ctzll(mq) & ~3;
                mq &= ~(0xFull << q);
                spelled(b ^ dp ^ ((((b >> q) & 0xF) ^ c0) << q));
            }
        }
        return;
    }
    if (n64 == 1) {
        spelled(b | (0xFull << pos64));
        return;
    }
    put(b);
}

// validate (BoardMaskerAD.cpp:422-456) with tile_sum_and_32k_count2's
// statistics (:165-187): pinned cells cleared first.
__device__ __forceinline__ bool ad_validate(uint64_t b, uint32_t board_sum, const AdSpecDev &s,
                                            const uint8_t *__restrict__ comb) {
    b &= ~s.fixed_mask;
    uint32_t total = 0;
    int count = 0;
    for (int k = 0; k < 16; ++k) {
        const int v = (int)((b >> (4 * k)) & 0xF);
        if (v > 5) ++count;
        else if (v > 0) total += 1u << v;
    }
    if (total >= s.limit + 64u) return false;
    if (count == s.n_free) return true;
    const uint32_t lsum = board_sum - total - ((uint32_t)(s.n_free + s.n_fixed) << 15);
    const uint32_t idx = (uint32_t)(uint8_t)(lsum >> 6) + ((uint32_t)(uint8_t)(count - s.n_free) << 8);
    if (idx >= (uint32_t)kCombEntries) return false;
    const uint8_t *c = comb + idx * kCombStride;
    const int n = c[0];
    if (n == 0 || c[n] == s.target) return false;
    if (count < 2) return true;
    if (n > 2 && c[1] == c[3]) {
        if (total >= s.limit - 64u) return false;
    } else if (n > 1 && c[1] == c[2]) {
        if (total >= s.limit) return false;
    }
    return true;
}

// One forward step over parents [0, n) (gen_boards_ad,
// BookGeneratorAD.cpp:785-960): for every empty cell, spawn 2 then 4, the
// four moves in the reference's order; a moved board that changed and
// matches the pattern is canonicalised; without a new masked tile it is
// kept, with one it goes through derive.
//   kCount: counts[i] gets c2 | c4 << 32, acc[0..5] the OR/AND spans
//           (atomic) and exact totals;
//   kWrite: counts[i] holds the exclusive scan and the boards are written
//           to out2 / out4 at counts[i] - offs_base;
//   kHash:  as kCount (without counts), the boards inserted into the key
//           sets set2 / set4 (masks mask2 / mask4); acc[6] becomes nonzero
//           when an insert found its probe run full (the step is redone).
enum AdMode { kCount = 0, kWrite = 1, kHash = 2 };

template <int MODE>
__global__ __launch_bounds__(kAdThreads) void k_ad_expand(const uint64_t *__restrict__ par, uint64_t n,
                                                          const uint16_t *__restrict__ glut, AdSpecDev sp,
                                                          const uint8_t *__restrict__ comb, uint32_t board_sum,
                                                          uint64_t *__restrict__ counts, uint64_t offs_base,
                                                          uint64_t *__restrict__ out2, uint64_t *__restrict__ out4,
                                                          uint64_t mask2, uint64_t mask4,
                                                          unsigned long long *__restrict__ acc) {
    constexpr bool WRITE = MODE == kWrite;
    extern __shared__ uint16_t lut[];
    __shared__ uint64_t red[6][kAdThreads / 64];
    __shared__ uint32_t s_fail;
    if (MODE == kHash && threadIdx.x == 0) s_fail = 0;
    stage_lut(lut, glut);
    uint64_t or2 = 0, and2 = ~0ull, or4 = 0, and4 = ~0ull, x2 = 0, x4 = 0;
    bool fail = false;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t b = par[i];
        uint64_t w2 = 0, w4 = 0;
        if (WRITE) {
            const uint64_t o = counts[i] - offs_base;
            w2 = o & 0xFFFFFFFFull;
            w4 = o >> 32;
        }
        uint32_t c2 = 0, c4 = 0;
        auto put2 = [&](uint64_t key) {
            if (WRITE) {
                out2[w2 + c2] = key;
            } else {
                or2 |= key;
                and2 &= key;
                if (MODE == kHash && !set_insert(out2, mask2, key)) fail = true;
            }
            ++c2;
        };
        auto put4 = [&](uint64_t key) {
            if (WRITE) {
                out4[w4 + c4] = key;
            } else {
                or4 |= key;
                and4 &= key;
                if (MODE == kHash && !set_insert(out4, mask4, key)) fail = true;
            }
            ++c4;
        };
        uint32_t em = empty_cells(b);
        const SpawnMoves smv(b, lut);
        const int f0 = f_count(b);
        while (em) {
            const int k = __builtin_ctz(em);
            em &= em - 1;
            for (int s = 1; s <= 2; ++s) {
                const uint64_t spb = b | ((uint64_t)s << (4 * k));
                uint64_t mv[4];
                smv.of(k, (uint64_t)s, lut, mv);
                for (int d = 0; d < 4; ++d) {
                    if (mv[d] == spb || !ad_is_pattern(mv[d], sp)) continue;
                    const uint64_t canon = canonical(mv[d], sp.symm);
                    const bool mnt = f_count(mv[d]) > f0;
                    if (s == 1) {
                        if (mnt) derive_emit(canon, board_sum + 2u, sp, comb, put2);
                        else put2(canon);
                    } else {
                        if (mnt) derive_emit(canon, board_sum + 4u, sp, comb, put4);
                        else put4(canon);
                    }
                }
            }
        }
        if (!WRITE) {
            if (MODE == kCount) counts[i] = (uint64_t)c2 | ((uint64_t)c4 << 32);
            x2 += c2;
            x4 += c4;
        }
    }
    if (WRITE) return;
    if (MODE == kHash && fail) s_fail = 1;  // benign race: any writer sets it
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        or2 |= __shfl_xor(or2, off, 64);
        and2 &= __shfl_xor(and2, off, 64);
        or4 |= __shfl_xor(or4, off, 64);
        and4 &= __shfl_xor(and4, off, 64);
        x2 += __shfl_xor(x2, off, 64);
        x4 += __shfl_xor(x4, off, 64);
    }
    const uint32_t wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0][wave] = or2;
        red[1][wave] = and2;
        red[2][wave] = or4;
        red[3][wave] = and4;
        red[4][wave] = x2;
        red[5][wave] = x4;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t w = 1; w < (blockDim.x >> 6); ++w) {
            or2 |= red[0][w];
            and2 &= red[1][w];
            or4 |= red[2][w];
            and4 &= red[3][w];
            x2 += red[4][w];
            x4 += red[5][w];
        }
        atomicOr(&acc[0], (unsigned long long)or2);
        atomicAnd(&acc[1], (unsigned long long)and2);
        atomicOr(&acc[2], (unsigned long long)or4);
        atomicAnd(&acc[3], (unsigned long long)and4);
        if (x2) atomicAdd(&acc[4], (unsigned long long)x2);
        if (x4) atomicAdd(&acc[5], (unsigned long long)x4);
        if (MODE == kHash && s_fail) atomicOr(&acc[6], 1ull);
    }
}

__global__ void k_ad_acc_init(unsigned long long *acc) {
    if (threadIdx.x < 7) acc[threadIdx.x] = (threadIdx.x == 1 || threadIdx.x == 3) ? ~0ull : 0ull;
}

__global__ void k_ad_validate(const uint64_t *__restrict__ keys, uint64_t n, uint32_t board_sum, AdSpecDev sp,
                              const uint8_t *__restrict__ comb, uint32_t *__restrict__ keep) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        keep[i] = ad_validate(keys[i], board_sum, sp, comb) ? 1u : 0u;
}

// Stable compaction by an exclusive scan of the keep flags; *count gets the
// number kept.
__global__ void k_ad_scatter(const uint64_t *__restrict__ keys, uint64_t n, const uint32_t *__restrict__ keep,
                             const uint32_t *__restrict__ offs, uint64_t *__restrict__ out,
                             uint64_t *__restrict__ count) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        if (keep[i]) out[offs[i]] = keys[i];
        if (i == n - 1) *count = (uint64_t)offs[i] + keep[i];
    }
}

// masked_tiles_combinations (BoardMaskerAD.cpp:256-285) for
// (remaining_sum, remaining_count), written ascending as init_masker keeps
// it (:517-526); returns the length, 0 when there is none.
int comb_entry(uint32_t sum, int count, uint8_t *out) {
    static const uint32_t tiles[8] = {8192, 4096, 2048, 1024, 512, 256, 128, 64};
    uint8_t tmp[16];
    int n = 0;
    for (int k = 0; k < 8; ++k) {
        const uint32_t t = tiles[k];
        const uint8_t lg = (uint8_t)(13 - k);
        if (t > sum) continue;
        if (t == sum) {
            if (count == 1) {
                tmp[n++] = lg;
                for (int i = 0; i < n; ++i) out[i] = tmp[n - 1 - i];
                return n;
            }
            continue;
        }
        if (2 * t == sum && count == 2) {
            tmp[n++] = lg;
            tmp[n++] = lg;
        } else if (t != 64 && sum == 192 && count == 3) {
            tmp[n++] = 6;
            tmp[n++] = 6;
            tmp[n++] = 6;
        } else {
            sum -= t;
            --count;
            tmp[n++] = lg;
            continue;
        }
        for (int i = 0; i < n; ++i) out[i] = tmp[n - 1 - i];
        return n;
    }
    return 0;
}

}  // namespace

// Per-device tables: the masked row table and the combination table, made on
// a device's first advanced pass and kept for the process's lifetime (as the
// device context keeps the classic row table).
AdDevice &ad_device(int dev) {
    static std::mutex mu;
    static std::vector<std::unique_ptr<AdDevice>> devs;
    std::lock_guard<std::mutex> lock(mu);
    if ((int)devs.size() <= dev) devs.resize(dev + 1);
    if (!devs[dev]) {
        DeviceCtx &C = device_ctx(dev);
        auto d = std::make_unique<AdDevice>();
        HIPCHK(hipSetDevice(dev));
        HIPCHK(hipMalloc(&d->lut, kLutEntries * sizeof(uint16_t)));
        hipLaunchKernelGGL(k_build_ad_lut, dim3(kLutEntries / 256), dim3(256), 0, C.stream, d->lut);
        HIPCHK(hipGetLastError());
        std::vector<uint8_t> host(kCombEntries * kCombStride, 0);
        for (int a = 1; a < 256; ++a)
            for (int c = 1; c < 10; ++c) {
                uint8_t *e = &host[(size_t)(a + (c << 8)) * kCombStride];
                e[0] = (uint8_t)comb_entry((uint32_t)a << 6, c, e + 1);
            }
        HIPCHK(hipMalloc(&d->comb, host.size()));
        HIPCHK(hipMemcpyAsync(d->comb, host.data(), host.size(), hipMemcpyHostToDevice, C.stream));
        HIPCHK(hipStreamSynchronize(C.stream));
        devs[dev] = std::move(d);
    }
    return *devs[dev];
}

uint32_t grid_for(uint64_t n, uint32_t threads, uint32_t cap) {
    const uint64_t b = (n + threads - 1) / threads;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(b, cap));
}

namespace {

struct AdRun {
    hipStream_t s;
    int dev, cus;
    const AdDevice &D;
    AdSpecDev spec;
    const egtb_ad_options &o;
    SortWork w;
    DBuf counts, partial, acc, cnt_dev;
    HostStage stage;

    AdRun(hipStream_t s_, int dev_, int cus_, const AdDevice &D_, const AdSpecDev &sp, const egtb_ad_options &o_)
        : s(s_), dev(dev_), cus(cus_), D(D_), spec(sp), o(o_) {}

    uint64_t read_u64(const void *p) {
        uint64_t v = 0;
        HIPCHK(hipMemcpyAsync(&v, p, 8, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        return v;
    }

    // count pass over par[0..n): totals and spans into acc[0..5]
    void count(const uint64_t *par, uint64_t n, uint32_t bsum) {
        counts.ensure(n * 8 + 8, s);
        acc.ensure(64, s);
        hipLaunchKernelGGL(k_ad_acc_init, dim3(1), dim3(64), 0, s, (unsigned long long *)acc.p);
        const size_t lds = kLutEntries * sizeof(uint16_t);
        hipLaunchKernelGGL(k_ad_expand<kCount>, dim3(grid_for(n, kAdThreads, cus)), dim3(kAdThreads), lds, s, par,
                           n, D.lut, spec, D.comb, bsum, (uint64_t *)counts.p, 0ull, (uint64_t *)nullptr,
                           (uint64_t *)nullptr, 0ull, 0ull, (unsigned long long *)acc.p);
        HIPCHK(hipGetLastError());
    }

    // write pass after count(): exclusive scan of the packed counts, then the
    // boards into out2 / out4
    void write(const uint64_t *par, uint64_t n, uint32_t bsum, uint64_t *out2, uint64_t *out4) {
        partial.ensure(scan_partials_needed(n) * 8 + 64, s);
        launch_scan_u64(s, (const uint64_t *)counts.p, n, nullptr, (uint64_t *)counts.p, (uint64_t *)partial.p,
                        nullptr);
        const size_t lds = kLutEntries * sizeof(uint16_t);
        hipLaunchKernelGGL(k_ad_expand<kWrite>, dim3(grid_for(n, kAdThreads, cus)), dim3(kAdThreads), lds, s, par,
                           n, D.lut, spec, D.comb, bsum, (uint64_t *)counts.p, 0ull, out2, out4, 0ull, 0ull,
                           (unsigned long long *)nullptr);
        HIPCHK(hipGetLastError());
    }

    // one step into key sets of cap2 / cap4 slots (powers of two, zero);
    // totals, spans and the overflow flag into acc[0..6]
    void hash(const uint64_t *par, uint64_t n, uint32_t bsum, uint64_t *set2, uint64_t cap2, uint64_t *set4,
              uint64_t cap4) {
        acc.ensure(64, s);
        hipLaunchKernelGGL(k_ad_acc_init, dim3(1), dim3(64), 0, s, (unsigned long long *)acc.p);
        if (n == 0) return;
        const size_t lds = kLutEntries * sizeof(uint16_t);
        hipLaunchKernelGGL(k_ad_expand<kHash>, dim3(grid_for(n, kAdThreads, cus)), dim3(kAdThreads), lds, s, par,
                           n, D.lut, spec, D.comb, bsum, (uint64_t *)nullptr, 0ull, set2, set4, cap2 - 1, cap4 - 1,
                           (unsigned long long *)acc.p);
        HIPCHK(hipGetLastError());
    }

    // sorted unique keys of keys[0..n) (clobbered, with tmp) into a new
    // buffer; the count lands in device slot `slot` (read with counts2)
    DBuf dedup_async(uint64_t *keys, uint64_t *tmp, uint64_t n, uint64_t or_all, uint64_t and_all, int slot) {
        DBuf out;
        out.alloc(n * 8 + 8, s);
        cnt_dev.ensure(64, s);
        uint64_t *cnt = (uint64_t *)cnt_dev.p + slot;
        if (n == 0) HIPCHK(hipMemsetAsync(cnt, 0, 8, s));
        else dedup_into(s, keys, tmp, n, or_all, and_all, (uint64_t *)out.p, cnt, o.radix_bits, w, o.sort_schedule);
        return out;
    }

    void counts2(uint64_t &a, uint64_t &b) {
        uint64_t v[2] = {0, 0};
        HIPCHK(hipMemcpyAsync(v, cnt_dev.p, 16, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        a = v[0];
        b = v[1];
    }

    // keys[0..n) (unique) sorted; the buffer that holds them
    DBuf sort_keys(DBuf &u, uint64_t n, uint64_t or_all, uint64_t and_all) {
        if (n < 2) return std::move(u);
        DBuf t;
        t.alloc(n * 8 + 8, s);
        const int sched = o.sort_schedule == 1 || o.sort_schedule == 2 ? o.sort_schedule : 0;
        uint64_t *out = radix_sort(s, (uint64_t *)u.p, (uint64_t *)t.p, n, or_all, and_all, o.radix_bits, w, sched);
        return out == (uint64_t *)u.p ? std::move(u) : std::move(t);
    }

    DBuf dedup(uint64_t *keys, uint64_t *tmp, uint64_t n, uint64_t or_all, uint64_t and_all, uint64_t &m) {
        DBuf out = dedup_async(keys, tmp, n, or_all, and_all, 0);
        m = read_u64(cnt_dev.p);
        return out;
    }

    // keys[0..n) keeping those validate() accepts for board sum bsum
    uint64_t validate(DBuf &keys, uint64_t n, uint32_t bsum) {
        if (n == 0) return 0;
        DBuf keep, offs, part, out;
        keep.alloc(n * 4 + 4, s);
        offs.alloc(n * 4 + 4, s);
        part.alloc(scan_partials_needed(n) * 4 + 64, s);
        out.alloc(n * 8 + 8, s);
        const uint32_t g = grid_for(n, 256, (uint32_t)cus * 16);
        hipLaunchKernelGGL(k_ad_validate, dim3(g), dim3(256), 0, s, (const uint64_t *)keys.p, n, bsum, spec, D.comb,
                           (uint32_t *)keep.p);
        launch_scan_u32(s, (const uint32_t *)keep.p, n, (uint32_t *)offs.p, (uint32_t *)part.p);
        cnt_dev.ensure(64, s);
        hipLaunchKernelGGL(k_ad_scatter, dim3(g), dim3(256), 0, s, (const uint64_t *)keys.p, n,
                           (const uint32_t *)keep.p, (const uint32_t *)offs.p, (uint64_t *)out.p,
                           (uint64_t *)cnt_dev.p);
        HIPCHK(hipGetLastError());
        const uint64_t m = read_u64(cnt_dev.p);
        keys = std::move(out);
        return m;
    }

    void write_raw(const std::string &path, const DBuf &keys, uint64_t n) {
        stage.ensure(n * 8 + 8);
        if (n) HIPCHK(hipMemcpyAsync(stage.p, keys.p, n * 8, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        write_file_atomic(path, stage.p, n * 8);
    }
};

}  // namespace

bool file_exists(const std::string &p) {
    struct stat st;
    return ::stat(p.c_str(), &st) == 0;
}

uint64_t file_size(const std::string &p) {
    struct stat st;
    return ::stat(p.c_str(), &st) == 0 ? (uint64_t)st.st_size : 0;
}

std::vector<uint64_t> read_raw(const std::string &p) {
    FILE *f = std::fopen(p.c_str(), "rb");
    if (!f) return {};
    std::fseek(f, 0, SEEK_END);
    const long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (sz < 0 || sz % 8) {
        std::fclose(f);
        throw Error(EGTB_ERR_IO, "misaligned binary file size: " + p);
    }
    std::vector<uint64_t> v((size_t)sz / 8);
    const size_t got = v.empty() ? 0 : std::fread(v.data(), 8, v.size(), f);
    std::fclose(f);
    if (got != v.size()) throw Error(EGTB_ERR_IO, "short read of " + p);
    return v;
}

uint32_t seed_sum(uint64_t b) {
    uint32_t t = 0;
    for (int k = 0; k < 16; ++k) {
        const uint32_t v = (uint32_t)((b >> (4 * k)) & 0xF);
        if (v) t += 1u << v;
    }
    return t;
}

uint64_t mask_board(uint64_t b) {
    uint64_t r = b;
    for (int k = 0; k < 16; ++k)
        if (((b >> (4 * k)) & 0xF) >= 6) r |= 0xFull << (4 * k);
    return r;
}

AdSpecDev make_ad_spec(const egtb_ad_pattern &pat) {
    if (pat.n_masks > (uint32_t)kMaxMasks) throw Error(EGTB_ERR_ARG, "at most 16 masks");
    if (pat.n_masks && !pat.masks) throw Error(EGTB_ERR_ARG, "masks is NULL");
    if (pat.symm_mode < 0 || pat.symm_mode > 3)
        throw Error(EGTB_ERR_ARG, "symm_mode must be 0-3 (identity, full, diagonal, horizontal)");
    if (pat.n_fixed > 16 || (pat.n_fixed && !pat.fixed_32k_shifts))
        throw Error(EGTB_ERR_ARG, "fixed_32k_shifts: at most 16 shifts");
    if (pat.n_free_32k < 0 || pat.n_free_32k > 16) throw Error(EGTB_ERR_ARG, "n_free_32k must be 0-16");
    if (pat.target < 1 || pat.target > 15) throw Error(EGTB_ERR_ARG, "target must be log2 of a tile (1-15)");
    AdSpecDev sp{};
    for (uint32_t i = 0; i < pat.n_masks; ++i) sp.masks[i] = pat.masks[i];
    sp.n_masks = pat.n_masks;
    sp.symm = (uint32_t)pat.symm_mode;
    for (uint32_t i = 0; i < pat.n_fixed; ++i) {
        if (pat.fixed_32k_shifts[i] > 60 || pat.fixed_32k_shifts[i] % 4)
            throw Error(EGTB_ERR_ARG, "fixed_32k_shifts must be nibble shifts 0, 4, ..., 60");
        sp.fixed_mask += 0xFull << pat.fixed_32k_shifts[i];
    }
    sp.n_fixed = (int32_t)pat.n_fixed;
    sp.n_free = pat.n_free_32k;
    sp.limit = pat.small_tile_sum_limit;
    sp.target = pat.target;
    return sp;
}

// generate_process_ad (BookGeneratorAD.cpp:1256-1396) for a fresh build.
void ad_generate(const uint64_t *seeds, size_t n_seeds, const egtb_ad_pattern &pat, const egtb_ad_options &o,
                 AdResult &res) {
    if (!seeds && n_seeds) throw Error(EGTB_ERR_ARG, "seeds is NULL");
    const AdSpecDev sp = make_ad_spec(pat);
    if (o.steps < 2) throw Error(EGTB_ERR_ARG, "steps must be at least 2");
    if (o.radix_bits != 0 && o.radix_bits != 4 && o.radix_bits != 6 && o.radix_bits != 8 && o.radix_bits != 11)
        throw Error(EGTB_ERR_ARG, "radix_bits must be 0, 4, 6, 8 or 11");
    if (o.sort_schedule < 0 || o.sort_schedule > 3) throw Error(EGTB_ERR_ARG, "sort_schedule must be 0-3");
    if (o.chunk_mkeys < 0) throw Error(EGTB_ERR_ARG, "chunk_mkeys must be >= 0");
    if (o.forward_dedup < 0 || o.forward_dedup > 2) throw Error(EGTB_ERR_ARG, "forward_dedup must be 0, 1 or 2");

    const int dev = o.device;
    DeviceCtx &C = device_ctx(dev);
    HIPCHK(hipSetDevice(dev));
    const AdDevice &D = ad_device(dev);
    hipStream_t s = C.stream;
    AdRun run(s, dev, C.cus, D, sp, o);
    res.steps = o.steps;
    res.dev = dev;
    res.n.assign((size_t)std::max(o.steps - 1, 1), 0);
    if (o.keep_layers) res.layers.resize(res.n.size());
    const std::string prefix = o.pathname ? o.pathname : "";
    const bool files = o.pathname != nullptr;

    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t ini = n_seeds ? seed_sum(seeds[0]) : 0u;
    DBuf d0, d1;
    uint64_t n0 = 0, n1 = 0;
    uint64_t d1_or = 0, d1_and = ~0ull;
    // keep a copy of layer i (device keys[0..n)) in the result
    auto keep = [&](int i, const DBuf &keys, uint64_t n) {
        if (!o.keep_layers) return;
        res.layers[i].alloc(n * 8 + 8, s);
        if (n) HIPCHK(hipMemcpyAsync(res.layers[i].p, keys.p, n * 8, hipMemcpyDeviceToDevice, s));
    };
    // a raw layer file into a new device buffer
    auto load = [&](const std::string &path, DBuf &out, uint64_t &n, uint64_t *or_all, uint64_t *and_all) {
        const std::vector<uint64_t> h = read_raw(path);
        n = h.size();
        out.alloc(n * 8 + 8, s);
        if (n) HIPCHK(hipMemcpyAsync(out.p, h.data(), n * 8, hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));  // h leaves scope
        if (or_all) {
            *or_all = 0;
            *and_all = ~0ull;
            for (uint64_t v : h) {
                *or_all |= v;
                *and_all &= v;
            }
        }
    };
    auto exists = [&](const std::string &suffix) { return file_exists(prefix + suffix); };
    bool started = false;
    size_t free_b = 0, total_b = 0;
    HIPCHK(hipMemGetInfo(&free_b, &total_b));
    free_b += device_idle_bytes(dev);
    // candidate keys per parent range: three buffers of them (keys, sort
    // scratch, output) plus the layers at hand must fit
    const uint64_t budget = o.chunk_mkeys ? (uint64_t)o.chunk_mkeys << 20
                                          : std::max<uint64_t>(1ull << 22, (uint64_t)(free_b / 8 / 8));
    const uint64_t kMaxSort = (1ull << 32) - 1;
    // key sets: two slots per predicted key, sets up to 2^25 slots (256 MiB,
    // the size of the last-level cache the random inserts live in)
    const bool use_sets = o.forward_dedup != 1 && o.chunk_mkeys == 0;
    const double slots_per_key = o.forward_dedup == 2 ? 2.0 / 64.0 : 2.0;
    const uint64_t min_slots = o.forward_dedup == 2 ? 1024 : (1ull << 15), max_slots = 1ull << 25;
    DBuf set2, set4;
    double r2 = 0;  // last step's layer / parents
    // carry: the spawn-4 boards of the last step are in set2 (carry_cap
    // slots), which the next step fills with its spawn-2 boards, instead of
    // being the list d1
    bool carry = false;
    uint64_t carry_cap = 0;
    auto cap_for = [&](double keys) {
        uint64_t c = min_slots;
        while ((double)c < keys * slots_per_key) c <<= 1;
        return c;
    };
    auto set_ready = [&](DBuf &b, uint64_t cap) {
        if (b.bytes < cap * 8) {
            b.alloc(cap * 8, s);
            HIPCHK(hipMemsetAsync(b.p, 0, b.bytes, s));
        }
    };
    // the carried set's keys into the list d1 (the set is left empty)
    auto drain = [&]() {
        DBuf u;
        u.alloc(carry_cap * 8, s);
        run.cnt_dev.ensure(64, s);
        HIPCHK(hipMemsetAsync(run.cnt_dev.p, 0, 8, s));
        launch_set_extract(s, (uint64_t *)set2.p, carry_cap, (uint64_t *)u.p, (uint64_t *)run.cnt_dev.p);
        n1 = run.read_u64(run.cnt_dev.p);
        d1 = std::move(u);
        carry = false;
    };

    for (int i = 1; i < o.steps - 1; ++i) {
        const uint32_t bsum = (uint32_t)(2 * i) + ini - 2u;
        if (files) {
            // handle_restart_ad (BookGeneratorAD.cpp:391-420): a step whose
            // layer (and the next) is on disk already, as a raw file or as
            // books, is skipped; the first step run after skipped ones reads
            // layers i-1 and i back
            const std::string a = std::to_string(i), b = std::to_string(i + 1);
            const bool have_i = exists(a);
            if ((exists(b) && have_i) || (exists(b + "b") && have_i) || (exists(b + ".z") && have_i) ||
                exists(a + "b") || exists(a + ".z") || exists(a + "b.7z") || exists(a + ".7z")) {
                res.n[i] = have_i ? file_size(prefix + a) / 8 : 0;
                if (o.keep_layers && have_i) {
                    DBuf t;
                    uint64_t m = 0;
                    load(prefix + a, t, m, nullptr, nullptr);
                    keep(i, t, m);
                }
                ++res.skipped_steps;
                continue;
            }
        }
        if (i == 1) {
            // layer 0: the masked seeds as given (:409-411)
            std::vector<uint64_t> h0(n_seeds);
            for (size_t k = 0; k < n_seeds; ++k) h0[k] = mask_board(seeds[k]);
            n0 = n_seeds;
            d0.alloc(n0 * 8 + 8, s);
            if (n0) HIPCHK(hipMemcpyAsync(d0.p, h0.data(), n0 * 8, hipMemcpyHostToDevice, s));
            HIPCHK(hipStreamSynchronize(s));
            res.n[0] = n0;
            if (files) run.write_raw(prefix + "0", d0, n0);
            keep(0, d0, n0);
            started = true;
        } else if (!started) {
            load(prefix + std::to_string(i - 1), d0, n0, nullptr, nullptr);
            load(prefix + std::to_string(i), d1, n1, &d1_or, &d1_and);
            if (o.keep_layers && !res.layers[0].p) {
                DBuf t;
                uint64_t m = 0;
                load(prefix + "0", t, m, nullptr, nullptr);
                res.n[0] = m;
                keep(0, t, m);
            }
            if (res.n[0] == 0 && file_exists(prefix + "0")) res.n[0] = file_size(prefix + "0") / 8;
            started = true;
        }
        uint64_t m0 = 0, m1 = 0;
        uint64_t or4 = 0, and4 = ~0ull;
        DBuf layer, next;
        bool done = false;
        if (use_sets && n0) {
            // the step into key sets: layer i's set holds the spawn-4 boards
            // of step i-1 (carried, or inserted from the list d1) and takes
            // the spawn-2 boards; layer i+1's set takes the spawn-4 boards
            // and is carried to the next step. Layer i's set is extracted
            // (unique, any order) and sorted. Layer i+1's set is sized for
            // the whole of layer i+1 from the last growth ratio.
            const double pn = (double)n0;
            const uint64_t cap2 = carry ? carry_cap : cap_for(std::max((double)n1, r2 * pn) * 1.3 + 1024.0);
            const uint64_t cap4 = cap_for(r2 * r2 * pn * 1.3 + 1024.0);
            if (cap2 <= max_slots && cap4 <= max_slots) {
                if (!carry) set_ready(set2, cap2);
                set_ready(set4, cap4);
                if (!carry && n1) launch_set_insert(s, (const uint64_t *)d1.p, n1, (uint64_t *)set2.p, cap2);
                run.hash((const uint64_t *)d0.p, n0, bsum, (uint64_t *)set2.p, cap2, (uint64_t *)set4.p, cap4);
                uint64_t a[7];
                HIPCHK(hipMemcpyAsync(a, run.acc.p, sizeof a, hipMemcpyDeviceToHost, s));
                HIPCHK(hipStreamSynchronize(s));
                if (a[6] == 0) {
                    DBuf u2;
                    u2.alloc(cap2 * 8, s);
                    run.cnt_dev.ensure(64, s);
                    HIPCHK(hipMemsetAsync(run.cnt_dev.p, 0, 8, s));
                    launch_set_extract(s, (uint64_t *)set2.p, cap2, (uint64_t *)u2.p, (uint64_t *)run.cnt_dev.p);
                    m0 = run.read_u64(run.cnt_dev.p);
                    layer = run.sort_keys(u2, m0, d1_or | a[0], d1_and & a[1]);
                    std::swap(set2, set4);  // the emptied set takes the next step's spawn-4 boards
                    d1.release();           // the list (if any) is in the set now
                    n1 = 0;
                    carry = true;
                    carry_cap = cap4;
                    or4 = a[2];
                    and4 = a[3];
                    res.candidates += a[4] + a[5];
                    ++res.hash_steps;
                    done = true;
                } else {
                    // an insert found its probe run full: layer i's set (with
                    // the spawn-4 boards of step i-1 in it) becomes the list
                    // d1, layer i+1's is emptied, and the step runs again
                    // through sort + unique
                    if (carry) {
                        drain();
                    } else {
                        HIPCHK(hipMemsetAsync(set2.p, 0, cap2 * 8, s));
                    }
                    HIPCHK(hipMemsetAsync(set4.p, 0, cap4 * 8, s));
                    ++res.set_overflows;
                }
            }
        }
        if (!done && carry) drain();  // sort + unique takes the list
        if (!done) {
            std::vector<DBuf> runs2, runs4;
            std::vector<uint64_t> runs2_n, runs4_n;
            uint64_t or2 = d1_or, and2 = d1_and;
            // parent ranges: the whole layer, halved until a range's boards fit
            std::vector<std::pair<uint64_t, uint64_t>> todo{{0, n0}};
            bool whole = true;
            DBuf c2buf, c4buf, tmp;
            uint64_t t2_whole = 0;
            while (!todo.empty()) {
                const auto [lo, hi] = todo.back();
                todo.pop_back();
                const uint64_t np = hi - lo;
                if (np == 0) continue;
                const uint64_t *par = (const uint64_t *)d0.p + lo;
                run.count(par, np, bsum);
                uint64_t a[6];
                HIPCHK(hipMemcpyAsync(a, run.acc.p, sizeof a, hipMemcpyDeviceToHost, s));
                HIPCHK(hipStreamSynchronize(s));
                const uint64_t t2 = a[4], t4 = a[5];
                const uint64_t extra = whole ? n1 : 0;  // a whole step also takes d1 into the spawn-2 buffer
                if ((t2 + extra > budget || t4 > budget || t2 + extra > kMaxSort || t4 > kMaxSort ||
                     t2 >= (1ull << 32) || t4 >= (1ull << 32)) && np > 1) {
                    whole = false;
                    todo.push_back({lo + np / 2, hi});
                    todo.push_back({lo, lo + np / 2});
                    continue;
                }
                if (t2 + extra > kMaxSort || t4 > kMaxSort)
                    throw Error(EGTB_ERR_MEMORY, "one parent's boards exceed a sort");
                res.candidates += t2 + t4;
                or2 |= a[0];
                and2 &= a[1];
                or4 |= a[2];
                and4 &= a[3];
                c2buf.ensure((t2 + extra) * 8 + 8, s);
                c4buf.ensure(t4 * 8 + 8, s);
                run.write(par, np, bsum, (uint64_t *)c2buf.p, (uint64_t *)c4buf.p);
                tmp.ensure(std::max(t2 + extra, t4) * 8 + 8, s);
                if (whole) {
                    if (n1) HIPCHK(hipMemcpyAsync((uint64_t *)c2buf.p + t2, d1.p, n1 * 8, hipMemcpyDeviceToDevice, s));
                    t2_whole = t2 + n1;
                } else {
                    ++res.chunks;
                }
                uint64_t m2 = 0, m4 = 0;
                const uint64_t n2 = whole ? t2_whole : t2;
                runs2.push_back(run.dedup_async((uint64_t *)c2buf.p, (uint64_t *)tmp.p, n2, whole ? or2 : (a[0]),
                                                whole ? and2 : (a[1]), 0));
                runs4.push_back(run.dedup_async((uint64_t *)c4buf.p, (uint64_t *)tmp.p, t4, a[2], a[3], 1));
                run.counts2(m2, m4);
                runs2_n.push_back(m2);
                runs4_n.push_back(m4);
            }
            c2buf.release();
            c4buf.release();
            tmp.release();
            // merge runs: layer i = spawn-2 runs (+ d1 when the step was split)
            auto merge = [&](std::vector<DBuf> &rs, std::vector<uint64_t> &rn, const DBuf *extra_buf,
                             uint64_t extra_n, uint64_t or_all, uint64_t and_all, uint64_t &m) -> DBuf {
                if (rs.size() == 1 && !extra_buf) {
                    m = rn[0];
                    return std::move(rs[0]);
                }
                uint64_t tot = extra_n;
                for (uint64_t v : rn) tot += v;
                DBuf cat, tbuf;
                cat.alloc(tot * 8 + 8, s);
                uint64_t off = 0;
                for (size_t k = 0; k < rs.size(); ++k) {
                    if (rn[k]) HIPCHK(hipMemcpyAsync((uint64_t *)cat.p + off, rs[k].p, rn[k] * 8, hipMemcpyDeviceToDevice, s));
                    off += rn[k];
                    rs[k].release();
                }
                if (extra_n) HIPCHK(hipMemcpyAsync((uint64_t *)cat.p + off, extra_buf->p, extra_n * 8, hipMemcpyDeviceToDevice, s));
                tbuf.alloc(tot * 8 + 8, s);
                return run.dedup((uint64_t *)cat.p, (uint64_t *)tbuf.p, tot, or_all, and_all, m);
            };
            if (runs2.empty()) {
                // no parents: the layer is d1 as it stands
                layer.alloc(n1 * 8 + 8, s);
                if (n1) HIPCHK(hipMemcpyAsync(layer.p, d1.p, n1 * 8, hipMemcpyDeviceToDevice, s));
                m0 = n1;
            } else {
                layer = merge(runs2, runs2_n, whole ? nullptr : &d1, whole ? 0 : n1, or2, and2, m0);
            }
            if (runs4.empty()) {
                next.alloc(8, s);
                m1 = 0;
            } else {
                next = merge(runs4, runs4_n, nullptr, 0, or4, and4, m1);
            }
            if (!whole) ++res.chunked_steps;
        }  // sort + unique
        if (n0) r2 = (double)m0 / (double)n0;
        if (!carry) {
            d1 = std::move(next);
            n1 = m1;
        }
        d1_or = or4;
        d1_and = and4;
        // validation every 32 steps (BookGeneratorAD.cpp:1383-1387)
        if (((i + (int)(ini % 64u / 2u)) % 32) == (int)((sp.limit / 2u) % 32u) + 1) {
            m0 = run.validate(layer, m0, bsum + 2u);
            if (carry) drain();
            n1 = run.validate(d1, n1, bsum + 4u);
        }
        res.n[i] = m0;
        if (files) run.write_raw(prefix + std::to_string(i), layer, m0);
        keep(i, layer, m0);
        d0 = std::move(layer);
        n0 = m0;
        progress_set((uint32_t)i, (uint32_t)(o.steps - 2));
    }
    res.started = started;
    // the look-back kernels' stall flag is sticky: one read covers every sort
    if (run.w.read_error(s)) throw Error(EGTB_ERR_INTERNAL, "a look-back sort stalled");
    HIPCHK(hipStreamSynchronize(s));
    res.fwd_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace egtb

struct egtb_ad_result : egtb::AdResult {};

extern "C" {

int egtb_ad_generate(const uint64_t *seeds, size_t n_seeds, const egtb_ad_pattern *pattern,
                     const egtb_ad_options *options, egtb_ad_result **out) {
    if (out) *out = nullptr;
    if (!pattern || !options) {
        egtb::g_last_error = "pattern and options are required";
        return EGTB_ERR_ARG;
    }
    auto r = std::make_unique<egtb_ad_result>();
    const int rc = egtb::guarded_ad([&] { egtb::ad_generate(seeds, n_seeds, *pattern, *options, *r); });
    if (rc == EGTB_OK && out) *out = r.release();
    return rc;
}

int egtb_ad_result_info(const egtb_ad_result *r, int *steps, uint64_t *sizes, double *fwd_ms, uint64_t *candidates,
                        uint64_t *chunked_steps, uint64_t *skipped_steps) {
    if (!r) {
        egtb::g_last_error = "result is NULL";
        return EGTB_ERR_ARG;
    }
    if (steps) *steps = r->steps;
    if (sizes) std::copy(r->n.begin(), r->n.end(), sizes);
    if (fwd_ms) *fwd_ms = r->fwd_ms;
    if (candidates) *candidates = r->candidates;
    if (chunked_steps) *chunked_steps = r->chunked_steps;
    if (skipped_steps) *skipped_steps = r->skipped_steps;
    return EGTB_OK;
}

int egtb_ad_result_forward(const egtb_ad_result *r, uint64_t *hash_steps, uint64_t *set_overflows) {
    if (!r) {
        egtb::g_last_error = "result is NULL";
        return EGTB_ERR_ARG;
    }
    if (hash_steps) *hash_steps = r->hash_steps;
    if (set_overflows) *set_overflows = r->set_overflows;
    return EGTB_OK;
}

int egtb_ad_result_layer(const egtb_ad_result *r, int layer, uint64_t *keys, size_t cap, size_t *n) {
    return egtb::guarded_ad([&] {
        if (!r || layer < 0 || layer >= (int)r->n.size()) throw egtb::Error(EGTB_ERR_ARG, "layer out of range");
        const uint64_t m = r->n[layer];
        if (n) *n = m;
        if (!keys) return;
        if (r->layers.empty()) throw egtb::Error(EGTB_ERR_ARG, "layers were not kept (options.keep_layers)");
        if (cap < m) throw egtb::Error(EGTB_ERR_ARG, "buffer too small");
        HIPCHK(hipSetDevice(r->dev));
        if (m) HIPCHK(hipMemcpy(keys, r->layers[layer].p, m * 8, hipMemcpyDeviceToHost));
    });
}

void egtb_ad_result_free(egtb_ad_result *r) {
    if (!r) return;
    (void)hipSetDevice(r->dev);
    delete r;
}

}  // extern "C"
