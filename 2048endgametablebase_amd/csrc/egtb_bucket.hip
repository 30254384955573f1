// Sample-sort dedup of one layer's candidates: sorted unique keys in six
// launches, for layers of up to kBucketMaxKeys candidates, where the radix
// pass chain (three to five launches per 8-bit digit) is bound by launch
// latency rather than bandwidth.
//
//   1. k_bk_sample:  one block sorts a strided sample of 8 keys per bucket
//                    and takes every 8th as a splitter (B buckets, B a power
//                    of two, about 2048 keys each).
//   2. k_bk_count:   bucket of every key (branch-free search of the
//                    splitters in LDS), block histograms added to global
//                    counts.
//   3. k_bk_scan:    bucket starts.
//   4. k_bk_scatter: keys to their bucket; each block reserves its run per
//                    bucket with one atomic (order inside a bucket is free).
//   5. k_bk_sort_unique: one block per bucket: bitonic sort with 8 keys per
//                    thread in registers (distances below 8 in registers, up
//                    to 256 by wave shuffles, the rest through LDS), then
//                    adjacent unique, written back at the bucket's start.  A
//                    bucket above 8192 keys (with 8 samples per bucket the
//                    odds are about 1e-9 a bucket) takes a plain bitonic sort
//                    in global scratch instead.
//   6. k_bk_gather:  unique runs concatenated in bucket order; the last block
//                    writes the layer size.
// Buckets are value ranges in increasing order, so the concatenation is
// sorted and a key's duplicates all land in one bucket.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "egtb_kernels.h"
#include "egtb_wave.h"

namespace egtb {

namespace {

constexpr int kBkThreads = 1024;
constexpr int kBkWaves = kBkThreads / 64;
constexpr int kBkItems = 8;
constexpr uint32_t kBkCap = kBkThreads * kBkItems;  // keys one block sorts in LDS
constexpr int kScatterThreads = 512;
constexpr int kScatterItems = 16;
constexpr uint32_t kScatterTile = kScatterThreads * kScatterItems;
constexpr uint32_t kBkMaxBuckets = 1024;
constexpr uint32_t kBkOversample = 8;
constexpr uint32_t kBkTarget = 2048;  // mean keys per bucket

struct BkLds {
    uint64_t a[kBkCap];
    uint32_t scratch[kBkWaves];
};

// Bitonic sort of a[0..P) ascending, P a power of two, by the whole block;
// `a` is LDS or (for oversized buckets) global memory private to the block.
__device__ void bitonic_sort(uint64_t *a, uint64_t P) {
    for (uint64_t k = 2; k <= P; k <<= 1) {
        for (uint64_t j = k >> 1; j > 0; j >>= 1) {
            for (uint64_t t = threadIdx.x; t < P / 2; t += kBkThreads) {
                const uint64_t i = ((t & ~(j - 1)) << 1) | (t & (j - 1));  // bit j of i clear
                const uint64_t l = i | j;
                const uint64_t x = a[i], y = a[l];
                const bool up = (i & k) == 0;
                if ((x > y) == up) {
                    a[i] = y;
                    a[l] = x;
                }
            }
            __threadfence_block();
            __syncthreads();
        }
    }
}

// One compare-exchange stage at distance J < 8 between a thread's own keys
// (elements 8t+e); ascending where bit k of the element index is clear.
template <int J>
__device__ __forceinline__ void reg_stage(uint64_t (&v)[kBkItems], uint32_t k, uint32_t t) {
#pragma unroll
    for (int e = 0; e < kBkItems; ++e) {
        if (e & J) continue;
        const bool asc = ((8u * t + (uint32_t)e) & k) == 0;
        const uint64_t x = v[e], y = v[e | J];
        if ((x > y) == asc) {
            v[e] = y;
            v[e | J] = x;
        }
    }
}

// Bitonic sort of P keys (a power of two, 8 <= P <= kBkCap) held 8 per
// thread, thread t holding elements 8t..8t+7; all threads of the block call
// it.  Threads with 8t >= P hold +inf and only ever meet each other.
__device__ void block_bitonic(uint64_t (&v)[kBkItems], uint32_t P, uint64_t *lds) {
    const uint32_t t = threadIdx.x;
    for (uint32_t k = 2; k <= P; k <<= 1) {
        for (uint32_t j = k >> 1; j >= 8; j >>= 1) {
            const uint32_t d = j >> 3;  // partner thread t ^ d holds element e at i ^ j
            const bool keep_min = ((t & d) == 0) == (((8u * t) & k) == 0);
            if (d < 64) {
#pragma unroll
                for (int e = 0; e < kBkItems; ++e) {
                    const uint64_t p = __shfl_xor(v[e], (int)d, 64);
                    v[e] = keep_min ? (v[e] < p ? v[e] : p) : (v[e] < p ? p : v[e]);
                }
            } else {
#pragma unroll
                for (int e = 0; e < kBkItems; ++e) lds[8 * t + e] = v[e];
                __syncthreads();
#pragma unroll
                for (int e = 0; e < kBkItems; ++e) {
                    const uint64_t p = lds[(8 * t + e) ^ j];
                    v[e] = keep_min ? (v[e] < p ? v[e] : p) : (v[e] < p ? p : v[e]);
                }
                __syncthreads();
            }
        }
        if (k >= 8) reg_stage<4>(v, k, t);
        if (k >= 4) reg_stage<2>(v, k, t);
        reg_stage<1>(v, k, t);
    }
}

__device__ __forceinline__ uint64_t pow2_at_least(uint64_t m) {
    uint64_t P = 1;
    while (P < m) P <<= 1;
    return P;
}

// number of splitters <= k, for B a power of two and B-1 sorted splitters
__device__ __forceinline__ uint32_t bucket_of(const uint64_t *spl, uint32_t B, uint64_t k) {
    uint32_t pos = 0;
    for (uint32_t step = B >> 1; step; step >>= 1)
        if (spl[pos + step - 1] <= k) pos += step;
    return pos;
}

__global__ __launch_bounds__(kBkThreads) void k_bk_sample(const uint64_t *__restrict__ keys, uint64_t n,
                                                          uint32_t B, uint64_t *__restrict__ spl,
                                                          uint32_t *__restrict__ counts,
                                                          uint32_t *__restrict__ starts) {
    extern __shared__ uint64_t bk_lds[];
    BkLds &L = *reinterpret_cast<BkLds *>(bk_lds);
    for (uint32_t b = threadIdx.x; b < B; b += kBkThreads) counts[b] = 0;
    if (B == 1) {
        if (threadIdx.x == 0) {
            counts[0] = (uint32_t)n;
            starts[0] = 0;
        }
        return;
    }
    // B >= 2 means n > 1024 * B, so the sample has exactly 8B keys, a power
    // of two no larger than kBkCap
    const uint32_t s = kBkOversample * B, t = threadIdx.x;
    uint64_t v[kBkItems];
#pragma unroll
    for (int e = 0; e < kBkItems; ++e) {
        const uint32_t i = 8 * t + e;
        v[e] = i < s ? keys[(uint64_t)i * n / s] : ~0ull;
    }
    block_bitonic(v, s, L.a);
    // splitter j is sample element 8(j+1): element 0 of thread j+1
    if (t >= 1 && t < B) spl[t - 1] = v[0];
}

__global__ __launch_bounds__(kScatterThreads) void k_bk_count(const uint64_t *__restrict__ keys, uint64_t n,
                                                              const uint64_t *__restrict__ spl, uint32_t B,
                                                              uint32_t *__restrict__ counts,
                                                              uint16_t *__restrict__ ids) {
    __shared__ uint64_t sspl[kBkMaxBuckets];
    __shared__ uint32_t hist[kBkMaxBuckets];
    for (uint32_t i = threadIdx.x; i + 1 < B; i += kScatterThreads) sspl[i] = spl[i];
    for (uint32_t b = threadIdx.x; b < B; b += kScatterThreads) hist[b] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * kScatterTile;
    uint64_t key[kScatterItems];
#pragma unroll
    for (int j = 0; j < kScatterItems; ++j) {
        const uint64_t idx = base + (uint64_t)j * kScatterThreads + threadIdx.x;
        key[j] = idx < n ? keys[idx] : 0ull;
    }
#pragma unroll
    for (int j = 0; j < kScatterItems; ++j) {
        const uint64_t idx = base + (uint64_t)j * kScatterThreads + threadIdx.x;
        if (idx < n) {
            const uint32_t b = bucket_of(sspl, B, key[j]);
            ids[idx] = (uint16_t)b;
            atomicAdd(&hist[b], 1u);
        }
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < B; b += kScatterThreads)
        if (hist[b]) atomicAdd(&counts[b], hist[b]);
}

__global__ __launch_bounds__(1024) void k_bk_scan(const uint32_t *__restrict__ counts, uint32_t B,
                                                  uint32_t *__restrict__ starts, uint32_t *__restrict__ cursor) {
    __shared__ uint32_t scratch[16];
    const uint32_t v = threadIdx.x < B ? counts[threadIdx.x] : 0u;
    uint32_t total;
    const uint32_t ex = block_exclusive_sum(v, scratch, total);
    if (threadIdx.x < B) {
        starts[threadIdx.x] = ex;
        cursor[threadIdx.x] = ex;
    }
}

__global__ __launch_bounds__(kScatterThreads) void k_bk_scatter(const uint64_t *__restrict__ keys, uint64_t n,
                                                                const uint16_t *__restrict__ ids, uint32_t B,
                                                                uint32_t *__restrict__ cursor,
                                                                uint64_t *__restrict__ out) {
    __shared__ uint32_t hist[kBkMaxBuckets];
    __shared__ uint32_t bbase[kBkMaxBuckets];
    for (uint32_t b = threadIdx.x; b < B; b += kScatterThreads) hist[b] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * kScatterTile;
    uint64_t key[kScatterItems];
    uint32_t bk[kScatterItems], r[kScatterItems];
#pragma unroll
    for (int j = 0; j < kScatterItems; ++j) {
        const uint64_t idx = base + (uint64_t)j * kScatterThreads + threadIdx.x;
        key[j] = idx < n ? keys[idx] : 0ull;
        bk[j] = idx < n ? ids[idx] : 0u;
    }
#pragma unroll
    for (int j = 0; j < kScatterItems; ++j) {
        const uint64_t idx = base + (uint64_t)j * kScatterThreads + threadIdx.x;
        if (idx < n) r[j] = atomicAdd(&hist[bk[j]], 1u);
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < B; b += kScatterThreads)
        if (hist[b]) bbase[b] = atomicAdd(&cursor[b], hist[b]);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kScatterItems; ++j) {
        const uint64_t idx = base + (uint64_t)j * kScatterThreads + threadIdx.x;
        if (idx < n) out[(uint64_t)bbase[bk[j]] + r[j]] = key[j];
    }
}

__global__ __launch_bounds__(kBkThreads) void k_bk_sort_unique(const uint64_t *__restrict__ src,
                                                               uint64_t *__restrict__ runs,
                                                               uint64_t *__restrict__ scratch,
                                                               const uint32_t *__restrict__ counts,
                                                               const uint32_t *__restrict__ starts,
                                                               uint32_t *__restrict__ ucount) {
    extern __shared__ uint64_t bk_lds[];
    BkLds &L = *reinterpret_cast<BkLds *>(bk_lds);
    const uint32_t m = counts[blockIdx.x];
    const uint64_t start = starts[blockIdx.x];
    if (m == 0) {
        if (threadIdx.x == 0) ucount[blockIdx.x] = 0;
        return;
    }
    if (m <= kBkCap) {
        const uint32_t P = (uint32_t)pow2_at_least(m < 8 ? 8 : m), t = threadIdx.x;
        uint64_t v[kBkItems];
#pragma unroll
        for (int e = 0; e < kBkItems; ++e) {
            const uint32_t i = 8 * t + e;
            v[e] = i < m ? src[start + i] : ~0ull;
        }
        block_bitonic(v, P, L.a);
#pragma unroll
        for (int e = 0; e < kBkItems; ++e) L.a[8 * t + e] = v[e];
        __syncthreads();
        // adjacent unique; thread t owns keys [t*kBkItems, (t+1)*kBkItems)
        const uint32_t first = threadIdx.x * kBkItems;
        uint32_t keep = 0, cnt = 0;
#pragma unroll
        for (int q = 0; q < kBkItems; ++q) {
            const uint32_t i = first + q;
            const bool kq = i < m && (i == 0 || L.a[i] != L.a[i - 1]);
            keep |= (kq ? 1u : 0u) << q;
            cnt += kq ? 1u : 0u;
        }
        uint32_t total;
        uint32_t pos = block_exclusive_sum(cnt, L.scratch, total);
#pragma unroll
        for (int q = 0; q < kBkItems; ++q)
            if ((keep >> q) & 1u) runs[start + pos++] = L.a[first + q];
        if (threadIdx.x == 0) ucount[blockIdx.x] = total;
        return;
    }
    // oversized bucket: the same sort on a padded copy in global scratch.
    // Bucket b owns scratch[2*start, 2*start + 2*m), which holds its padded
    // size and does not meet the next bucket's region.
    const uint64_t P = pow2_at_least(m);
    uint64_t *sc = scratch + 2 * start;
    for (uint64_t i = threadIdx.x; i < P; i += kBkThreads) sc[i] = i < m ? src[start + i] : ~0ull;
    __threadfence_block();
    __syncthreads();
    bitonic_sort(sc, P);
    uint32_t run = 0;
    for (uint64_t c = 0; c < m; c += kBkThreads) {
        const uint64_t i = c + threadIdx.x;
        const bool kq = i < m && (i == 0 || sc[i] != sc[i - 1]);
        uint32_t total;
        const uint32_t ex = block_exclusive_sum(kq ? 1u : 0u, L.scratch, total);
        if (kq) runs[start + run + ex] = sc[i];
        run += total;
    }
    if (threadIdx.x == 0) ucount[blockIdx.x] = run;
}

__global__ __launch_bounds__(kBkThreads) void k_bk_gather(const uint64_t *__restrict__ runs,
                                                          const uint32_t *__restrict__ starts,
                                                          const uint32_t *__restrict__ ucount, uint32_t B,
                                                          uint64_t *__restrict__ out,
                                                          uint64_t *__restrict__ count_out) {
    __shared__ uint64_t scratch[kBkWaves];
    uint64_t part = 0;
    for (uint32_t b = threadIdx.x; b < blockIdx.x; b += kBkThreads) part += ucount[b];
    uint64_t prefix;
    (void)block_exclusive_sum(part, scratch, prefix);
    const uint32_t u = ucount[blockIdx.x];
    const uint64_t st = starts[blockIdx.x];
    for (uint32_t i = threadIdx.x; i < u; i += kBkThreads) out[prefix + i] = runs[st + i];
    if (blockIdx.x == B - 1 && threadIdx.x == 0) *count_out = prefix + u;
}

}  // namespace

uint32_t bucket_count_for(uint64_t n) {
    uint32_t B = 1;
    while (B < kBkMaxBuckets && (uint64_t)B * kBkTarget < n) B <<= 1;
    return B;
}

void configure_bucket_kernels() {
    const int lds = (int)sizeof(BkLds);
    (void)hipFuncSetAttribute((const void *)k_bk_sample, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    (void)hipFuncSetAttribute((const void *)k_bk_sort_unique, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
}

void launch_bucket_dedup(hipStream_t s, const uint64_t *keys, uint64_t *tmp, uint64_t *scratch, uint64_t n,
                         BucketWork &bw, uint64_t *out, uint64_t *count_out) {
    if (n == 0) {
        (void)hipMemsetAsync(count_out, 0, 8, s);
        return;
    }
    const uint32_t B = bucket_count_for(n);
    const size_t lds = sizeof(BkLds);
    hipLaunchKernelGGL(k_bk_sample, dim3(1), dim3(kBkThreads), lds, s, keys, n, B, bw.spl, bw.counts, bw.starts);
    const uint64_t *src = keys;
    if (B > 1) {
        const uint32_t grid = (uint32_t)((n + kScatterTile - 1) / kScatterTile);
        hipLaunchKernelGGL(k_bk_count, dim3(grid), dim3(kScatterThreads), 0, s, keys, n, bw.spl, B, bw.counts,
                           bw.ids);
        hipLaunchKernelGGL(k_bk_scan, dim3(1), dim3(1024), 0, s, bw.counts, B, bw.starts, bw.cursor);
        hipLaunchKernelGGL(k_bk_scatter, dim3(grid), dim3(kScatterThreads), 0, s, keys, n, bw.ids, B, bw.cursor,
                           tmp);
        src = tmp;
    }
    hipLaunchKernelGGL(k_bk_sort_unique, dim3(B), dim3(kBkThreads), lds, s, src, tmp, scratch, bw.counts,
                       bw.starts, bw.ucount);
    hipLaunchKernelGGL(k_bk_gather, dim3(B), dim3(kBkThreads), 0, s, tmp, bw.starts, bw.ucount, B, out,
                       count_out);
}

}  // namespace egtb
