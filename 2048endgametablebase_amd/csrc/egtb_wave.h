// Wave64 and block-level helpers shared by the sort, unique and dedup
// kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace egtb {

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

__device__ __forceinline__ uint64_t lanes_below(uint64_t m) { return m & __lanemask_lt(); }

template <typename T>
__device__ __forceinline__ T wave_inclusive_sum(T v) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        T y = __shfl_up(v, off, 64);
        if (lane >= (uint32_t)off) v += y;
    }
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_reduce_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Exclusive scan of one value per thread over the block; `total` receives
// the block sum.  `scratch` needs blockDim/64 entries.  Ends synchronised.
template <typename T>
__device__ __forceinline__ T block_exclusive_sum(T v, T *scratch, T &total) {
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
    const T inc = wave_inclusive_sum(v);
    if (lane == 63) scratch[wave] = inc;
    __syncthreads();
    T before = 0, sum = 0;
    for (uint32_t w = 0; w < nwaves; ++w) {
        const T t = scratch[w];
        if (w < wave) before += t;
        sum += t;
    }
    __syncthreads();
    total = sum;
    return before + inc - v;
}

// Lanes of this wave whose digit equals mine, among the lanes in `valid`.
template <int BITS>
__device__ __forceinline__ uint64_t match_digit(uint32_t digit, uint64_t valid) {
    uint64_t m = valid;
#pragma unroll
    for (int b = 0; b < BITS; ++b) {
        const bool bit = (digit >> b) & 1u;
        const uint64_t bb = __ballot(bit);
        m &= bit ? bb : ~bb;
    }
    return m;
}

}  // namespace egtb
