// HIP kernels for the classic tablebase build on gfx950 (wave64).
//
//   k_build_lut        65536-row left-slide table (global; staged to LDS)
//   k_expand<WRITE>    forward expansion: count pass / write pass
//   k_scan_*           device-wide exclusive scan (reduce / partials / down)
//   k_radix_hist/scatter  LSD radix sort pass with wave-ballot multisplit
//   k_unique_*         adjacent dedup of a sorted array
//   k_bucket_index     top-bit range index over a sorted layer
//   k_solve<T>         backward success-rate recurrence
//   k_final_rates<T>   final layers: success -> max_scale else zero
//   k_compact_*<T>     keep rate > zero (stable)
//   k_pack_records<T>  SoA -> packed {u64 board; T rate} records
//
// Launch wrappers at the bottom are the only entry points used by the host.
#include <hip/hip_runtime.h>

#include <mutex>
#include <stddef.h>
#include <stdint.h>

#include <utility>

#include "egtb_device.h"
#include "egtb_kernels.h"
#include "egtb_runtime.h"
#include "egtb_wave.h"

namespace egtb {

// ---------------------------------------------------------------------------
// Row table
// ---------------------------------------------------------------------------

// Slide a 4-tile row toward its high nibble: compact the non-empty tiles,
// merge equal neighbours once from the move side, 0xF (32768) never merges
// (reference StandardMergePolicy::_internal_merge, BoardMover.h:7-55).
__global__ void k_build_lut(uint16_t *lut) {
    const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= kLutEntries) return;
    uint32_t tiles[4], n = 0;
    for (int c = 0; c < 4; ++c) {
        const uint32_t v = (row >> (12 - 4 * c)) & 0xFu;
        if (v) tiles[n++] = v;
    }
    uint32_t out[4] = {0, 0, 0, 0}, m = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (i + 1 < n && tiles[i] == tiles[i + 1] && tiles[i] != 0xFu) {
            out[m++] = tiles[i] + 1;
            ++i;
        } else {
            out[m++] = tiles[i];
        }
    }
    lut[row] = (uint16_t)((out[0] << 12) | (out[1] << 8) | (out[2] << 4) | out[3]);
}

// ---------------------------------------------------------------------------
// Forward expansion (reference gen_boards_* , BookGenerator.cpp:473-760)
// ---------------------------------------------------------------------------

// Word offsets, from a layer record's or_all (the pointer the expand
// kernels get for the layer of their spawn-2 children), of the step's
// exact 64-bit candidate totals and of its key set's overflow flag.
constexpr int kExactC2 = (int)((offsetof(LayerDev, c2x) - offsetof(LayerDev, or_all)) / sizeof(uint64_t));
constexpr int kExactC4 = (int)((offsetof(LayerDev, c4x) - offsetof(LayerDev, or_all)) / sizeof(uint64_t));
constexpr int kSetOverflow = (int)((offsetof(LayerDev, pad) - offsetof(LayerDev, or_all)) / sizeof(uint64_t));

// How a stored key is expanded.  Plain builds store boards, and one
// pattern and docheck flag (launch arguments) hold for the whole launch.
// Batched builds store job keys and take both from the job table; a job
// whose last forward step is past expands nothing, so its layers end at its
// own layer count.
template <bool JOBS>
struct ParentView;

template <>
struct ParentView<false> {
    uint64_t board;
    bool live;
    __device__ __forceinline__ ParentView(uint64_t key, const PatternDev &p, int do_check, const JobArg &)
        : board(key), live(!(do_check && is_success(key, p))) {}
    __device__ __forceinline__ const PatternDev &pattern(const PatternDev &p, const JobArg &) const { return p; }
    __device__ __forceinline__ uint64_t child_key(uint64_t c, uint32_t &) const { return c; }
};

template <>
struct ParentView<true> {
    uint64_t board;
    uint32_t job;
    bool live;
    __device__ __forceinline__ ParentView(uint64_t key, const PatternDev &, int, const JobArg &ja)
        : board(job_board(key)), job(job_of(key)) {
        const JobDev &J = ja.tab[job];
        live = ja.step <= J.steps - 2 && !(ja.step > J.docheck && is_success(board, J.p));
    }
    __device__ __forceinline__ const PatternDev &pattern(const PatternDev &, const JobArg &ja) const {
        return ja.tab[job].p;
    }
    // a child's key; bit 1 of bad: the child leaves the job's pinned low
    // byte, bit 2: its key is 0
    __device__ __forceinline__ uint64_t child_key(uint64_t c, uint32_t &bad) const {
        const uint64_t k = job_key(c, job);
        bad |= ((c & 0xFF) != 0xFF ? 2u : 0u) | (k == 0 ? 4u : 0u);
        return k;
    }
};

// One thread per parent.  For each empty nibble k (ascending) and spawn
// value 1 (a 2) / 2 (a 4): the four moves of the spawned board; a move is
// kept when it changed the board and the result is in the pattern; the
// canonical form is emitted into the spawn-2 (next layer) or spawn-4
// (layer after) stream.  Parents are skipped when do_check and they are
// success boards (BookGenerator.cpp:634,991).
//
// WRITE = false: store packed counts (c2 | c4 << 32) per parent and fold the
//   OR / AND of every child into or_and2 / or_and4 (varying-bit span used to
//   size the radix passes).
// WRITE = true: write children at the exclusive prefix `offs` (packed the
//   same way) into out2 / out4; indices at or past cap2 / cap4 are dropped
//   (a write queued before its totals were known; the host then grows the
//   buffers and writes again).
template <bool WRITE, bool JOBS>
__global__ __launch_bounds__(1024) void k_expand(const uint64_t *__restrict__ par, uint64_t n_bound,
                                                 const uint64_t *__restrict__ n_ptr,
                                                 const uint16_t *__restrict__ glut, PatternDev p,
                                                 int do_check, uint64_t *__restrict__ counts,
                                                 const uint64_t *__restrict__ offs, uint64_t offs_base,
                                                 uint64_t *__restrict__ out2,
                                                 uint64_t *__restrict__ out4,
                                                 uint64_t *__restrict__ or_and2,
                                                 uint64_t *__restrict__ or_and4, uint64_t cap2,
                                                 uint64_t cap4, JobArg ja) {
    extern __shared__ uint16_t lut[];
    __shared__ uint64_t red[6][16];
    stage_lut(lut, glut);
    const uint64_t n = n_ptr ? (*n_ptr < n_bound ? *n_ptr : n_bound) : n_bound;
    uint64_t or2 = 0, and2 = ~0ull, or4 = 0, and4 = ~0ull, x2 = 0, x4 = 0;
    uint32_t bad = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const ParentView<JOBS> pv(par[i], p, do_check, ja);
        const uint64_t b = pv.board;
        const PatternDev &P = pv.pattern(p, ja);
        uint32_t c2 = 0, c4 = 0;
        uint64_t w2 = 0, w4 = 0;
        if (WRITE) {
            const uint64_t o = offs[i] - offs_base;  // both halves stay >= 0
            w2 = o & 0xFFFFFFFFull;
            w4 = o >> 32;
        }
        if (pv.live) {
            uint32_t em = empty_cells(b);
            const SpawnMoves smv(b, lut);
            while (em) {
                const int k = __builtin_ctz(em);
                em &= em - 1;
#pragma unroll
                for (int s = 1; s <= 2; ++s) {
                    const uint64_t sp = b | ((uint64_t)s << (4 * k));
                    uint64_t mv[4];
                    smv.of(k, (uint64_t)s, lut, mv);
#pragma unroll
                    for (int d = 0; d < 4; ++d) {
                        if (mv[d] != sp && is_pattern(mv[d], P)) {
                            const uint64_t c = pv.child_key(canonical(mv[d], P.symm), bad);
                            if (s == 1) {
                                if (WRITE && w2 + c2 < cap2) out2[w2 + c2] = c;
                                else { or2 |= c; and2 &= c; }
                                ++c2;
                            } else {
                                if (WRITE && w4 + c4 < cap4) out4[w4 + c4] = c;
                                else { or4 |= c; and4 &= c; }
                                ++c4;
                            }
                        }
                    }
                }
            }
        }
        if (!WRITE) {
            counts[i] = (uint64_t)c2 | ((uint64_t)c4 << 32);
            x2 += c2;
            x4 += c4;
        }
    }
    if (JOBS && bad) atomicOr(ja.err, bad);
    if (!WRITE) {
        // fold the span accumulators and the exact totals: wave, then one
        // set of atomics per block
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
        if (lane_id() == 0) {
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
            atomicOr((unsigned long long *)&or_and2[0], (unsigned long long)or2);
            atomicAnd((unsigned long long *)&or_and2[1], (unsigned long long)and2);
            atomicOr((unsigned long long *)&or_and4[0], (unsigned long long)or4);
            atomicAnd((unsigned long long *)&or_and4[1], (unsigned long long)and4);
            if (x2) atomicAdd((unsigned long long *)&or_and2[kExactC2], (unsigned long long)x2);
            if (x4) atomicAdd((unsigned long long *)&or_and2[kExactC4], (unsigned long long)x4);
        }
    }
}

// A forward step in one launch, for candidates whose order does not matter
// (a sort follows). Per block iteration (blockDim parents) each thread
// counts its children, one atomic on *tot (c2 | c4 << 32) reserves the
// block's run in out2 and in out4, and the children are generated again and
// written. Writes at or past cap2 / cap4 are dropped while the counter still
// counts them, so the totals are exact either way. The key spans of both
// child streams are folded into or_and2 / or_and4.
template <bool JOBS>
__global__ __launch_bounds__(1024) void k_expand_alloc(const uint64_t *__restrict__ par, uint64_t n_bound,
                                                       const uint64_t *__restrict__ n_ptr,
                                                       const uint16_t *__restrict__ glut, PatternDev p,
                                                       int do_check, unsigned long long *__restrict__ tot,
                                                       uint64_t *__restrict__ out2, uint64_t *__restrict__ out4,
                                                       uint64_t *__restrict__ or_and2,
                                                       uint64_t *__restrict__ or_and4, uint64_t cap2,
                                                       uint64_t cap4, JobArg ja) {
    extern __shared__ uint16_t lut[];
    __shared__ uint64_t red[6][16];
    __shared__ uint64_t scratch[16];
    __shared__ uint64_t s_base;
    stage_lut(lut, glut);
    const uint64_t n = n_ptr ? (*n_ptr < n_bound ? *n_ptr : n_bound) : n_bound;
    uint64_t or2 = 0, and2 = ~0ull, or4 = 0, and4 = ~0ull, x2 = 0, x4 = 0;
    uint32_t bad = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x; i0 < n; i0 += stride) {
        const uint64_t i = i0 + threadIdx.x;
        const ParentView<JOBS> pv(i < n ? par[i] : 0, p, do_check, ja);
        const uint64_t b = pv.board;
        const PatternDev &P = pv.pattern(p, ja);
        const bool live = i < n && pv.live;
        const SpawnMoves smv(b, lut);  // for both passes
        uint32_t c2 = 0, c4 = 0;
        if (live) {
            uint32_t em = empty_cells(b);
            while (em) {
                const int k = __builtin_ctz(em);
                em &= em - 1;
#pragma unroll
                for (int s = 1; s <= 2; ++s) {
                    const uint64_t sp = b | ((uint64_t)s << (4 * k));
                    uint64_t mv[4];
                    smv.of(k, (uint64_t)s, lut, mv);
#pragma unroll
                    for (int d = 0; d < 4; ++d) {
                        if (mv[d] != sp && is_pattern(mv[d], P)) {
                            const uint64_t c = pv.child_key(canonical(mv[d], P.symm), bad);
                            if (s == 1) {
                                or2 |= c;
                                and2 &= c;
                                ++c2;
                            } else {
                                or4 |= c;
                                and4 &= c;
                                ++c4;
                            }
                        }
                    }
                }
            }
        }
        uint64_t block_total;
        const uint64_t ex = block_exclusive_sum((uint64_t)c2 | ((uint64_t)c4 << 32), scratch, block_total);
        if (threadIdx.x == 0) {
            s_base = atomicAdd(tot, (unsigned long long)block_total);
            x2 += block_total & 0xFFFFFFFFull;  // a block's halves stay far below 2^32
            x4 += block_total >> 32;
        }
        __syncthreads();
        const uint64_t at = s_base + ex;  // both halves: c2 and c4 of a step stay below 2^32
        uint64_t w2 = at & 0xFFFFFFFFull, w4 = at >> 32;
        __syncthreads();  // s_base is rewritten by the next iteration
        if (live && (c2 | c4)) {
            uint32_t em = empty_cells(b);
            while (em) {
                const int k = __builtin_ctz(em);
                em &= em - 1;
#pragma unroll
                for (int s = 1; s <= 2; ++s) {
                    const uint64_t sp = b | ((uint64_t)s << (4 * k));
                    uint64_t mv[4];
                    smv.of(k, (uint64_t)s, lut, mv);
#pragma unroll
                    for (int d = 0; d < 4; ++d) {
                        if (mv[d] != sp && is_pattern(mv[d], P)) {
                            const uint64_t c = pv.child_key(canonical(mv[d], P.symm), bad);
                            if (s == 1) {
                                if (w2 < cap2) out2[w2] = c;
                                ++w2;
                            } else {
                                if (w4 < cap4) out4[w4] = c;
                                ++w4;
                            }
                        }
                    }
                }
            }
        }
    }
    if (JOBS && bad) atomicOr(ja.err, bad);
    // fold the span accumulators: wave, then one atomic per block
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        or2 |= __shfl_xor(or2, off, 64);
        and2 &= __shfl_xor(and2, off, 64);
        or4 |= __shfl_xor(or4, off, 64);
        and4 &= __shfl_xor(and4, off, 64);
    }
    const uint32_t wave = threadIdx.x >> 6;
    if (lane_id() == 0) {
        red[0][wave] = or2;
        red[1][wave] = and2;
        red[2][wave] = or4;
        red[3][wave] = and4;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t w = 1; w < (blockDim.x >> 6); ++w) {
            or2 |= red[0][w];
            and2 &= red[1][w];
            or4 |= red[2][w];
            and4 &= red[3][w];
        }
        atomicOr((unsigned long long *)&or_and2[0], (unsigned long long)or2);
        atomicAnd((unsigned long long *)&or_and2[1], (unsigned long long)and2);
        atomicOr((unsigned long long *)&or_and4[0], (unsigned long long)or4);
        atomicAnd((unsigned long long *)&or_and4[1], (unsigned long long)and4);
        if (x2) atomicAdd((unsigned long long *)&or_and2[kExactC2], (unsigned long long)x2);
        if (x4) atomicAdd((unsigned long long *)&or_and2[kExactC4], (unsigned long long)x4);
    }
}

// ---------------------------------------------------------------------------
// Forward step into hash sets: the children of a step are inserted into the
// open-addressing key sets of their layers (spawn 2: layer step, spawn 4:
// layer step+1) instead of being written out for a sort. Slots hold the key
// itself, 0 = empty (a spawned board is never 0); linear probing from
// hash_mix. A slot only ever goes from 0 to a key within a launch, so a
// plain load that sees a key is exact and one that sees 0 is settled by the
// CAS. An insert that finds kSetProbeLimit slots in a row taken raises the
// layer's overflow flag and drops the key: the host then regrows the set and
// runs the step again.
// ---------------------------------------------------------------------------

template <bool JOBS>
__global__ __launch_bounds__(1024) void k_expand_hash(const uint64_t *__restrict__ par, uint64_t n_bound,
                                                      const uint64_t *__restrict__ n_ptr,
                                                      const uint16_t *__restrict__ glut, PatternDev p, int do_check,
                                                      unsigned long long *__restrict__ tot, uint64_t *__restrict__ set2,
                                                      uint64_t mask2, uint64_t *__restrict__ set4, uint64_t mask4,
                                                      uint64_t *__restrict__ meta2, uint64_t *__restrict__ meta4,
                                                      JobArg ja) {
    extern __shared__ uint16_t lut[];
    __shared__ uint64_t red[4][16];
    __shared__ uint64_t wtot[2][16];
    __shared__ uint32_t s_fail;
    stage_lut(lut, glut);
    if (threadIdx.x == 0) s_fail = 0;
    __syncthreads();
    const uint64_t n = n_ptr ? (*n_ptr < n_bound ? *n_ptr : n_bound) : n_bound;
    uint64_t or2 = 0, and2 = ~0ull, or4 = 0, and4 = ~0ull, x2 = 0, x4 = 0;
    uint32_t fail = 0, bad = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const ParentView<JOBS> pv(par[i], p, do_check, ja);
        if (!pv.live) continue;
        const uint64_t b = pv.board;
        const PatternDev &P = pv.pattern(p, ja);
        const SpawnMoves smv(b, lut);
        // the last child per (spawn value, move): spawns in neighbouring
        // cells often slide to the same board, which needs no second insert
        uint64_t last[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        uint32_t em = empty_cells(b);
        while (em) {
            const int k = __builtin_ctz(em);
            em &= em - 1;
#pragma unroll
            for (int sv = 1; sv <= 2; ++sv) {
                const uint64_t sp = b | ((uint64_t)sv << (4 * k));
                uint64_t mv[4];
                smv.of(k, (uint64_t)sv, lut, mv);
#pragma unroll
                for (int d = 0; d < 4; ++d) {
                    if (mv[d] != sp && is_pattern(mv[d], P)) {
                        const uint64_t c = pv.child_key(canonical(mv[d], P.symm), bad);
                        const int j = (sv - 1) * 4 + d;
                        const bool again = c == last[j];
                        last[j] = c;
                        if (sv == 1) {
                            or2 |= c;
                            and2 &= c;
                            ++x2;
                            if (set2 && !again && !set_insert(set2, mask2, c)) fail |= 1u;
                        } else {
                            or4 |= c;
                            and4 &= c;
                            ++x4;
                            if (set4 && !again && !set_insert(set4, mask4, c)) fail |= 2u;
                        }
                    }
                }
            }
        }
    }
    // fold: wave shuffles, then one set of atomics per block
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
    if (fail) atomicOr(&s_fail, fail);
    if (JOBS && bad) atomicOr(ja.err, bad);
    if (lane_id() == 0) {
        red[0][wave] = or2;
        red[1][wave] = and2;
        red[2][wave] = or4;
        red[3][wave] = and4;
        wtot[0][wave] = x2;
        wtot[1][wave] = x4;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t w = 1; w < (blockDim.x >> 6); ++w) {
            or2 |= red[0][w];
            and2 &= red[1][w];
            or4 |= red[2][w];
            and4 &= red[3][w];
            x2 += wtot[0][w];
            x4 += wtot[1][w];
        }
        // the packed total (c2 | c4 << 32) and the exact 64-bit halves: a
        // block's halves stay far below 2^32, the packed sum over the launch
        // is exact while the step's totals are
        if (x2 | x4) atomicAdd(tot, (unsigned long long)(x2 | (x4 << 32)));
        if (x2) atomicAdd((unsigned long long *)&meta2[kExactC2], (unsigned long long)x2);
        if (x4) atomicAdd((unsigned long long *)&meta2[kExactC4], (unsigned long long)x4);
        atomicOr((unsigned long long *)&meta2[0], (unsigned long long)or2);
        atomicAnd((unsigned long long *)&meta2[1], (unsigned long long)and2);
        atomicOr((unsigned long long *)&meta4[0], (unsigned long long)or4);
        atomicAnd((unsigned long long *)&meta4[1], (unsigned long long)and4);
        if (s_fail & 1u) atomicOr((unsigned long long *)&meta2[kSetOverflow], 1ull);
        if (s_fail & 2u) atomicOr((unsigned long long *)&meta4[kSetOverflow], 1ull);
    }
}

// Inserts keys[0..n) (nonzero) into a set of more than n slots: probing
// ends at the first empty slot, which a set that cannot fill always has.
__global__ __launch_bounds__(256) void k_set_insert(const uint64_t *__restrict__ keys, uint64_t n,
                                                    uint64_t *__restrict__ set, uint64_t mask) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t k = keys[i];
        uint64_t h = hash_mix(k) & mask;
        for (uint64_t q = 0; q <= mask; ++q) {
            const uint64_t cur = set[h];
            if (cur == k) break;
            if (cur == 0) {
                const uint64_t old = atomicCAS((unsigned long long *)&set[h], 0ull, (unsigned long long)k);
                if (old == 0 || old == k) break;
            }
            h = (h + 1) & mask;
        }
    }
}

// Moves the keys of a set to out[] (any order) and clears their slots, so
// the set is empty again for its next layer; *count (zero before) gets the
// number of keys.
constexpr int kExtractThreads = 256;
constexpr int kExtractItems = 8;
constexpr int kExtractTile = kExtractThreads * kExtractItems;

__global__ __launch_bounds__(kExtractThreads) void k_set_extract(uint64_t *__restrict__ set, uint64_t cap,
                                                                 uint64_t *__restrict__ out,
                                                                 unsigned long long *__restrict__ count) {
    __shared__ uint64_t stage[kExtractTile];
    __shared__ uint32_t scratch[kExtractThreads / 64];
    __shared__ unsigned long long s_base;
    const uint64_t base = (uint64_t)blockIdx.x * kExtractTile;
    uint64_t v[kExtractItems];
    uint32_t c = 0;
#pragma unroll
    for (int j = 0; j < kExtractItems; ++j) {
        const uint64_t idx = base + (uint64_t)j * kExtractThreads + threadIdx.x;
        v[j] = idx < cap ? set[idx] : 0ull;
        if (v[j]) {
            ++c;
            set[idx] = 0;
        }
    }
    uint32_t total;
    uint32_t at = block_exclusive_sum(c, scratch, total);
    if (threadIdx.x == 0) s_base = total ? atomicAdd(count, (unsigned long long)total) : 0ull;
#pragma unroll
    for (int j = 0; j < kExtractItems; ++j)
        if (v[j]) stage[at++] = v[j];
    __syncthreads();
    const unsigned long long b0 = s_base;
    for (uint32_t q = threadIdx.x; q < total; q += kExtractThreads) out[b0 + q] = stage[q];
}

// ---------------------------------------------------------------------------
// Exclusive scan (out[0..n) exclusive prefix, out[n] = total)
// ---------------------------------------------------------------------------

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ __forceinline__ uint64_t dev_n(uint64_t bound, const uint64_t *n_ptr) {
    if (!n_ptr) return bound;
    const uint64_t v = *n_ptr;
    return v < bound ? v : bound;
}

template <typename TI, typename TO>
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const TI *__restrict__ in, uint64_t n_bound,
                                                              const uint64_t *__restrict__ n_ptr,
                                                              TO *__restrict__ partial) {
    __shared__ TO scratch[kScanThreads / 64];
    const uint64_t n = dev_n(n_bound, n_ptr);
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    TO s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const uint64_t idx = base + (uint64_t)j * kScanThreads + threadIdx.x;
        if (idx < n) s += (TO)in[idx];
    }
    s = wave_reduce_sum(s);
    if (lane_id() == 0) scratch[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        TO t = 0;
        for (int w = 0; w < kScanThreads / 64; ++w) t += scratch[w];
        partial[blockIdx.x] = t;
    }
}

// Single block: exclusive scan of the per-tile partials in place.
template <typename TO>
__global__ __launch_bounds__(kScanThreads) void k_scan_partials(TO *__restrict__ partial, uint64_t m) {
    __shared__ TO scratch[kScanThreads / 64];
    TO carry = 0;
    for (uint64_t base = 0; base < m; base += kScanThreads) {
        const uint64_t idx = base + threadIdx.x;
        const TO v = idx < m ? partial[idx] : (TO)0;
        TO total;
        const TO ex = block_exclusive_sum(v, scratch, total);
        if (idx < m) partial[idx] = carry + ex;
        carry += total;
    }
}

template <typename TI, typename TO>
__global__ __launch_bounds__(kScanThreads) void k_scan_down(const TI *__restrict__ in, uint64_t n_bound,
                                                            const uint64_t *__restrict__ n_ptr,
                                                            TO *__restrict__ out,
                                                            const TO *__restrict__ partial,
                                                            TO *__restrict__ total_out) {
    __shared__ TO tile[kScanTile];
    __shared__ TO scratch[kScanThreads / 64];
    const uint64_t n = dev_n(n_bound, n_ptr);
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const uint64_t idx = base + (uint64_t)j * kScanThreads + threadIdx.x;
        tile[j * kScanThreads + threadIdx.x] = idx < n ? (TO)in[idx] : (TO)0;
    }
    __syncthreads();
    TO local[kScanItems];
    TO s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        local[j] = s;
        s += tile[threadIdx.x * kScanItems + j];
    }
    TO total;
    const TO ex = block_exclusive_sum(s, scratch, total) + partial[blockIdx.x];
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) tile[threadIdx.x * kScanItems + j] = ex + local[j];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const uint64_t idx = base + (uint64_t)j * kScanThreads + threadIdx.x;
        if (idx < n) out[idx] = tile[j * kScanThreads + threadIdx.x];
    }
    if (total_out && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *total_out = partial[blockIdx.x] + total;
}

// ---------------------------------------------------------------------------
// LSD radix sort (stable), onesweep schedule, wave64 ballot multisplit
//
//   k_os_hist     one read of the keys: digit histograms of every pass
//   k_os_scan     exclusive scan of each pass's histogram (digit bases)
//   k_os_scatter  one pass: tile ranks by wave multisplit, tile offsets by a
//                 decoupled look-back over per-(tile, digit) status words,
//                 digit runs written contiguously through LDS
//
// Status words: [epoch:20 | flag:2 | value:42]; flag 1 = tile aggregate,
// 2 = inclusive prefix.  The epoch changes every pass, so the array is
// never cleared; a word from another pass reads as "not ready".  Spins are
// bounded: past the limit the kernel raises *err and carries on.
// ---------------------------------------------------------------------------

constexpr int kSortThreads = 256;
constexpr int kSortWaves = kSortThreads / 64;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr int kMaxPasses = 16;
constexpr uint32_t kSpinLimit = 1u << 22;

__device__ __forceinline__ uint64_t lb_pack(uint32_t epoch, uint32_t flag, uint64_t v) {
    return ((uint64_t)epoch << 44) | ((uint64_t)flag << 42) | (v & ((1ull << 42) - 1));
}

__device__ __forceinline__ void lb_store(uint64_t *p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t lb_load(const uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sum of the values of tiles [0, tile) from their status words, reading up
// to 16 predecessors per round so their loads are in flight together; waits
// for aggregates that are not yet published, stops at the first inclusive
// prefix (tile 0 always publishes one).
__device__ __forceinline__ uint64_t lb_lookback(const uint64_t *status, uint64_t stride, uint32_t tile,
                                                uint32_t epoch, uint32_t *err) {
    constexpr int B = 16;
    uint64_t prefix = 0;
    int64_t j = (int64_t)tile - 1;
    uint32_t spins = 0;
    while (j >= 0) {
        uint64_t w[B];
#pragma unroll
        for (int q = 0; q < B; ++q) w[q] = (j - q >= 0) ? lb_load(status + (uint64_t)(j - q) * stride) : 0;
        int q = 0;
        bool stall = false;
#pragma unroll
        for (int r = 0; r < B; ++r) {
            if (stall || j - r < 0) continue;
            const uint32_t f = (uint32_t)(w[r] >> 42) & 3u;
            if ((uint32_t)(w[r] >> 44) != epoch || f == 0) {
                stall = true;
                continue;
            }
            prefix += w[r] & ((1ull << 42) - 1);
            if (f == 2) return prefix;
            ++q;
        }
        j -= q;
        if (stall) {
            if (++spins > kSpinLimit) {
                atomicOr(err, 1u);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    return prefix;
}

template <int BITS>
__global__ __launch_bounds__(kSortThreads) void k_os_hist(const uint64_t *__restrict__ keys, uint64_t n, int lo,
                                                          int passes, uint32_t *__restrict__ hist) {
    constexpr int BINS = 1 << BITS;
    __shared__ uint32_t h[kMaxPasses * BINS > 12288 ? 12288 : kMaxPasses * BINS];
    for (int i = threadIdx.x; i < passes * BINS; i += kSortThreads) h[i] = 0;
    __syncthreads();
    const uint64_t stride = (uint64_t)gridDim.x * kSortThreads;
    for (uint64_t i = (uint64_t)blockIdx.x * kSortThreads + threadIdx.x; i < n; i += stride) {
        const uint64_t k = keys[i] >> lo;
        for (int q = 0; q < passes; ++q) atomicAdd(&h[q * BINS + ((k >> (q * BITS)) & (BINS - 1))], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < passes * BINS; i += kSortThreads)
        if (h[i]) atomicAdd(&hist[i], h[i]);
}

// One block: hist[q][*] -> exclusive digit bases, in place; also clears
// the other histogram buffer for the next sort (its last reader finished
// before this sort's histogram kernel started).
template <int BITS>
__global__ __launch_bounds__(kSortThreads) void k_os_scan(uint32_t *__restrict__ hist, int passes,
                                                          uint32_t *__restrict__ next_hist, int next_words) {
    constexpr int BINS = 1 << BITS;
    constexpr int PER = BINS >= kSortThreads ? BINS / kSortThreads : 1;
    __shared__ uint32_t scratch[kSortWaves];
    for (int i = threadIdx.x; i < next_words; i += kSortThreads) next_hist[i] = 0;
    for (int q = 0; q < passes; ++q) {
        uint32_t *hq = hist + q * BINS;
        const int first = threadIdx.x * PER;
        uint32_t local[PER];
        uint32_t sum = 0;
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            local[j] = sum;
            sum += first + j < BINS ? hq[first + j] : 0u;
        }
        uint32_t total;
        const uint32_t ex = block_exclusive_sum(sum, scratch, total);
#pragma unroll
        for (int j = 0; j < PER; ++j)
            if (first + j < BINS) hq[first + j] = ex + local[j];
        __syncthreads();
    }
}

template <int BITS>
__global__ __launch_bounds__(kSortThreads) void k_os_scatter(const uint64_t *__restrict__ src,
                                                             uint64_t *__restrict__ dst, uint64_t n, int shift,
                                                             const uint32_t *__restrict__ digit_base,
                                                             uint64_t *__restrict__ status,
                                                             unsigned long long *__restrict__ tile_counter,
                                                             unsigned long long tile_base, uint32_t epoch,
                                                             uint32_t *__restrict__ err) {
    constexpr int BINS = 1 << BITS;
    constexpr int ITEMS = kSortItems;
    constexpr int TILE = kSortTile;
    __shared__ uint64_t stage[TILE];
    __shared__ uint32_t whist[kSortWaves][BINS];
    __shared__ uint32_t binbase[BINS];
    __shared__ uint32_t gexcl[BINS];
    __shared__ uint32_t scratch[kSortWaves];
    __shared__ uint32_t s_tile;

    if (threadIdx.x == 0) s_tile = (uint32_t)(atomicAdd(tile_counter, 1ull) - tile_base);
    for (int i = threadIdx.x; i < kSortWaves * BINS; i += kSortThreads) (&whist[0][0])[i] = 0;
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t wave = threadIdx.x >> 6, lane = lane_id();
    const uint64_t tile0 = (uint64_t)tile * TILE;
    const uint64_t base = tile0 + (uint64_t)wave * ITEMS * 64;
    uint64_t key[ITEMS];
    uint32_t rank[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const uint64_t idx = base + (uint64_t)j * 64 + lane;
        key[j] = idx < n ? src[idx] : 0ull;
    }
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const uint64_t idx = base + (uint64_t)j * 64 + lane;
        const bool valid = idx < n;
        const uint32_t d = (uint32_t)((key[j] >> shift) & (BINS - 1));
        const uint64_t m = match_digit<BITS>(d, __ballot(valid));
        const uint32_t below = (uint32_t)__popcll(lanes_below(m));
        uint32_t prev = 0;
        if (valid) prev = whist[wave][d];
        rank[j] = prev + below;
        if (valid && below == 0) whist[wave][d] = prev + (uint32_t)__popcll(m);
    }
    __syncthreads();
    // per digit: wave-exclusive prefix, tile count; publish, then look back
    for (int bin = threadIdx.x; bin < BINS; bin += kSortThreads) {
        uint32_t c = 0;
#pragma unroll
        for (int w = 0; w < kSortWaves; ++w) {
            const uint32_t t = whist[w][bin];
            whist[w][bin] = c;
            c += t;
        }
        binbase[bin] = c;
        uint64_t *st = status + (uint64_t)tile * BINS + bin;
        if (tile == 0) {
            lb_store(st, lb_pack(epoch, 2, c));
            gexcl[bin] = 0;
        } else {
            lb_store(st, lb_pack(epoch, 1, c));
        }
    }
    if (tile > 0) {
        for (int bin = threadIdx.x; bin < BINS; bin += kSortThreads) {
            const uint64_t pre = lb_lookback(status + bin, BINS, tile, epoch, err);
            lb_store(status + (uint64_t)tile * BINS + bin, lb_pack(epoch, 2, pre + binbase[bin]));
            gexcl[bin] = (uint32_t)pre;
        }
    }
    __syncthreads();
    {
        constexpr int PER = BINS >= kSortThreads ? BINS / kSortThreads : 1;
        const int first = threadIdx.x * PER;
        uint32_t local[PER];
        uint32_t sum = 0;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            local[q] = sum;
            sum += (first + q < BINS) ? binbase[first + q] : 0u;
        }
        uint32_t total;
        const uint32_t ex = block_exclusive_sum(sum, scratch, total);
#pragma unroll
        for (int q = 0; q < PER; ++q)
            if (first + q < BINS) binbase[first + q] = ex + local[q];
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const uint64_t idx = base + (uint64_t)j * 64 + lane;
        if (idx < n) {
            const uint32_t d = (uint32_t)((key[j] >> shift) & (BINS - 1));
            stage[binbase[d] + whist[wave][d] + rank[j]] = key[j];
        }
    }
    __syncthreads();
    const uint64_t m = (n - tile0) < (uint64_t)TILE ? (n - tile0) : (uint64_t)TILE;
    for (uint32_t i = threadIdx.x; i < m; i += kSortThreads) {
        const uint64_t k = stage[i];
        const uint32_t d = (uint32_t)((k >> shift) & (BINS - 1));
        dst[(uint64_t)digit_base[d] + gexcl[d] + (i - binbase[d])] = k;
    }
}

// Classic three-kernel pass (tile histogram, scan, scatter): no look-back.
// Per-tile digit histogram -> counts[bin * ntiles + tile].
template <int BITS, int ITEMS>
__global__ __launch_bounds__(kSortThreads) void k_radix_hist(const uint64_t *__restrict__ src, uint64_t n,
                                                             int shift, uint32_t *__restrict__ counts,
                                                             uint32_t ntiles) {
    constexpr int BINS = 1 << BITS;
    constexpr int TILE = kSortThreads * ITEMS;
    __shared__ uint32_t hist[BINS];
    for (int i = threadIdx.x; i < BINS; i += kSortThreads) hist[i] = 0;
    __syncthreads();
    const uint32_t wave = threadIdx.x >> 6, lane = lane_id();
    const uint64_t base = (uint64_t)blockIdx.x * TILE + (uint64_t)wave * ITEMS * 64;
#pragma unroll 4
    for (int j = 0; j < ITEMS; ++j) {
        const uint64_t idx = base + (uint64_t)j * 64 + lane;
        const bool valid = idx < n;
        const uint32_t d = valid ? (uint32_t)((src[idx] >> shift) & (BINS - 1)) : 0u;
        const uint64_t m = match_digit<BITS>(d, __ballot(valid));
        if (valid && lanes_below(m) == 0) atomicAdd(&hist[d], (uint32_t)__popcll(m));
    }
    __syncthreads();
    for (int i = threadIdx.x; i < BINS; i += kSortThreads) counts[(uint64_t)i * ntiles + blockIdx.x] = hist[i];
}

// Per-digit scan of one pass's [digit][tile] counts: block d owns digit d
// and writes, for every tile, the keys of digit d in the tiles before it
// (offsets) and the digit's total (totals[d]). All digits in one launch and
// no block waits on another.
template <int BITS>
__global__ __launch_bounds__(kSortThreads) void k_radix_binscan(const uint32_t *__restrict__ counts,
                                                                uint32_t ntiles, uint32_t *__restrict__ offsets,
                                                                uint32_t *__restrict__ totals) {
    constexpr int PER = 4;
    __shared__ uint32_t scratch[kSortWaves];
    const uint32_t *c = counts + (uint64_t)blockIdx.x * ntiles;
    uint32_t *o = offsets + (uint64_t)blockIdx.x * ntiles;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < ntiles; base += kSortThreads * PER) {
        const uint32_t t0 = base + threadIdx.x * PER;
        uint32_t v[PER], sum = 0;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            v[q] = t0 + q < ntiles ? c[t0 + q] : 0u;
            sum += v[q];
        }
        uint32_t total;
        uint32_t run = carry + block_exclusive_sum(sum, scratch, total);
#pragma unroll
        for (int q = 0; q < PER; ++q)
            if (t0 + q < ntiles) {
                o[t0 + q] = run;
                run += v[q];
            }
        carry += total;
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

// Scatter one tile: stable local ranks by wave multisplit, reorder through
// LDS by digit, then write each digit run contiguously at its global offset:
// the keys of smaller digits (exclusive scan of totals, done here) plus the
// keys of the same digit in earlier tiles (offsets).
template <int BITS, int ITEMS>
__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(const uint64_t *__restrict__ src,
                                                                uint64_t *__restrict__ dst, uint64_t n,
                                                                int shift,
                                                                const uint32_t *__restrict__ offsets,
                                                                const uint32_t *__restrict__ totals,
                                                                uint32_t ntiles) {
    constexpr int BINS = 1 << BITS;
    constexpr int TILE = kSortThreads * ITEMS;
    __shared__ uint64_t stage[TILE];
    __shared__ uint32_t whist[kSortWaves][BINS];
    __shared__ uint32_t binbase[BINS];
    __shared__ uint32_t gbase[BINS];
    __shared__ uint32_t scratch[kSortWaves];

    for (int i = threadIdx.x; i < kSortWaves * BINS; i += kSortThreads) (&whist[0][0])[i] = 0;
    __syncthreads();

    const uint32_t wave = threadIdx.x >> 6, lane = lane_id();
    const uint64_t tile0 = (uint64_t)blockIdx.x * TILE;
    const uint64_t base = tile0 + (uint64_t)wave * ITEMS * 64;
    uint64_t key[ITEMS];
    uint32_t rank[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const uint64_t idx = base + (uint64_t)j * 64 + lane;
        key[j] = idx < n ? src[idx] : 0ull;
    }
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const uint64_t idx = base + (uint64_t)j * 64 + lane;
        const bool valid = idx < n;
        const uint32_t d = (uint32_t)((key[j] >> shift) & (BINS - 1));
        const uint64_t m = match_digit<BITS>(d, __ballot(valid));
        const uint32_t below = (uint32_t)__popcll(lanes_below(m));
        uint32_t prev = 0;
        if (valid) prev = whist[wave][d];
        rank[j] = prev + below;
        if (valid && below == 0) whist[wave][d] = prev + (uint32_t)__popcll(m);
    }
    __syncthreads();
    // wave-exclusive prefix per bin, block total per bin
    for (int bin = threadIdx.x; bin < BINS; bin += kSortThreads) {
        uint32_t s = 0;
#pragma unroll
        for (int w = 0; w < kSortWaves; ++w) {
            const uint32_t t = whist[w][bin];
            whist[w][bin] = s;
            s += t;
        }
        binbase[bin] = s;
    }
    __syncthreads();
    // exclusive scans over the bins (each thread owns a run of bins): of this
    // tile's per-bin totals, and of the digit totals of the whole pass
    {
        constexpr int PER = BINS >= kSortThreads ? BINS / kSortThreads : 1;
        const int first = threadIdx.x * PER;
        uint32_t local[PER], glocal[PER];
        uint32_t s = 0, g = 0;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            local[q] = s;
            glocal[q] = g;
            s += (first + q < BINS) ? binbase[first + q] : 0u;
            g += (first + q < BINS) ? totals[first + q] : 0u;
        }
        uint32_t total;
        const uint32_t ex = block_exclusive_sum(s, scratch, total);
        const uint32_t gex = block_exclusive_sum(g, scratch, total);
#pragma unroll
        for (int q = 0; q < PER; ++q)
            if (first + q < BINS) {
                binbase[first + q] = ex + local[q];
                gbase[first + q] = gex + glocal[q];
            }
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const uint64_t idx = base + (uint64_t)j * 64 + lane;
        if (idx < n) {
            const uint32_t d = (uint32_t)((key[j] >> shift) & (BINS - 1));
            stage[binbase[d] + whist[wave][d] + rank[j]] = key[j];
        }
    }
    __syncthreads();
    const uint64_t m = (n - tile0) < (uint64_t)TILE ? (n - tile0) : (uint64_t)TILE;
    for (uint32_t i = threadIdx.x; i < m; i += kSortThreads) {
        const uint64_t k = stage[i];
        const uint32_t d = (uint32_t)((k >> shift) & (BINS - 1));
        dst[(uint64_t)gbase[d] + offsets[(uint64_t)d * ntiles + blockIdx.x] + (i - binbase[d])] = k;
    }
}

// ---------------------------------------------------------------------------
// Unique of a sorted array (reference unique_sorted_u64_inplace,
// UniqueUtils.h:218-256): keep a[i] when i == 0 or a[i] != a[i-1].
// ---------------------------------------------------------------------------

// 32 keys a thread (8192 a tile, the most the 32-bit keep mask holds): a
// look-back launch's tiles all start together and each sums predecessor
// counts until it meets a published prefix, so fewer, larger tiles finish
// sooner. 442@512 (1-2 Mi-key layers): 108-110 ms a build at 8 keys a
// thread, 104 ms at 32.
constexpr int kUniqThreads = 256;
constexpr int kUniqItems = 32;
constexpr int kUniqTile = kUniqThreads * kUniqItems;

// One pass: each tile counts its kept elements, publishes the count, looks
// back for its output offset, writes; the last tile stores the total.
__global__ __launch_bounds__(kUniqThreads) void k_unique_lb(const uint64_t *__restrict__ a, uint64_t n,
                                                            uint64_t *__restrict__ out,
                                                            uint64_t *__restrict__ count_out,
                                                            uint64_t *__restrict__ status,
                                                            unsigned long long *__restrict__ tile_counter,
                                                            unsigned long long tile_base, uint32_t epoch,
                                                            uint32_t *__restrict__ err) {
    __shared__ uint64_t scratch[kUniqThreads / 64];
    __shared__ uint32_t s_tile;
    __shared__ uint64_t s_off;
    if (threadIdx.x == 0) s_tile = (uint32_t)(atomicAdd(tile_counter, 1ull) - tile_base);
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint64_t base = (uint64_t)tile * kUniqTile + (uint64_t)threadIdx.x * kUniqItems;
    uint64_t v[kUniqItems];
    uint32_t f = 0;
    uint64_t cnt = 0;
#pragma unroll
    for (int j = 0; j < kUniqItems; ++j) {
        const uint64_t idx = base + j;
        v[j] = idx < n ? a[idx] : 0;
        const bool keep = idx < n && (idx == 0 || v[j] != (j ? v[j - 1] : a[idx - 1]));
        f |= (keep ? 1u : 0u) << j;
        cnt += keep;
    }
    uint64_t total;
    const uint64_t ex = block_exclusive_sum(cnt, scratch, total);
    if (threadIdx.x == 0) {
        if (tile == 0) {
            lb_store(status, lb_pack(epoch, 2, total));
            s_off = 0;
        } else {
            lb_store(status + tile, lb_pack(epoch, 1, total));
            const uint64_t pre = lb_lookback(status, 1, tile, epoch, err);
            lb_store(status + tile, lb_pack(epoch, 2, pre + total));
            s_off = pre;
        }
    }
    __syncthreads();
    uint64_t pos = s_off + ex;
    if (count_out && (uint64_t)(tile + 1) * kUniqTile >= n && threadIdx.x == 0) *count_out = s_off + total;
#pragma unroll
    for (int j = 0; j < kUniqItems; ++j)
        if ((f >> j) & 1u) out[pos++] = v[j];
}

// ---------------------------------------------------------------------------
// Bucket index over a sorted layer: idx[v] = lower_bound(keys, v << shift),
// idx[2^bits] = n.  Narrows every lookup to one bucket (the reference's
// AdaptiveIndex level 1 over the top key bits, AdaptiveIndex.cpp:89-91).
// ---------------------------------------------------------------------------

__global__ void k_bucket_index(const uint64_t *__restrict__ keys, uint64_t n_bound,
                               const uint64_t *__restrict__ n_ptr, uint32_t bits, uint32_t *__restrict__ idx) {
    const uint64_t n = dev_n(n_bound, n_ptr);
    const uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t nb = 1ull << bits;
    if (v > nb) return;
    if (v == nb) {
        idx[v] = (uint32_t)n;
        return;
    }
    const uint64_t target = v << (64 - bits);
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (keys[mid] < target) lo = mid + 1;
        else hi = mid;
    }
    idx[v] = (uint32_t)lo;
}

// ---------------------------------------------------------------------------
// Backward solve (reference recalculate_layer, BookSolver.cpp:413-470)
// ---------------------------------------------------------------------------

// rate(b) = max_scale when do_check and b is a success board; otherwise
// with E empty nibbles visited in ascending order,
//   sp = fma(best2_k, 1 - p4, sp); sp = fma(best4_k, p4, sp)   (double)
//   rate = (T)(sp / E)   (E = 0 -> zero)
// where best_x_k is the T-typed max, starting from zero, over the valid
// moves of the spawned board of the next (spawn 2) / after-next (spawn 4)
// layer's rate of the canonical moved board (absent -> zero).  zero and
// max_scale are the rate type's (0 and 1 for probabilities, -1 and 0 for the
// 1-float types, 0 and 4e9 / 1.6e18 for the integer ones).  The fused update is what the
// reference's -Ofast -march=x86-64-v3 build computes (CMakeLists.txt:82-83).
template <typename T>
__device__ __forceinline__ uint64_t rate_bits(T v) {
    if constexpr (sizeof(T) == 4) return (uint64_t)__builtin_bit_cast(uint32_t, v);
    else return __builtin_bit_cast(uint64_t, v);
}

// Record (k, r) in a layer's solve hash.  Wide: keys of a layer are unique
// and the table is at most a quarter full, so the claim ends at an empty
// slot.  Compact: the first empty slot of the home bucket or of the next 31
// (one 64-bit CAS writes key and rate together); none free sets the error
// flag.
template <typename T>
__device__ __forceinline__ void hash_insert(const HashDev &h, uint64_t k, T r) {
    if (h.cslots) {
        if constexpr (sizeof(T) == 4) {
            uint64_t bk;
            uint32_t st;
            if (!chash_split(h, k, bk, st)) {  // a key outside its layer's span: cannot happen
                if (h.err) atomicOr(h.err, 2u);
                return;
            }
            unsigned long long *sl = (unsigned long long *)h.cslots;
            for (uint32_t d = 0; d < kCompactMaxDisp; ++d) {
                const uint64_t base = ((bk + d) & h.mask) * 8;
                const unsigned long long v = ((unsigned long long)chash_tag(d, st) << 32) | rate_bits(r);
                for (int j = 0; j < 8; ++j) {
                    if (__hip_atomic_load(&sl[base + j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0ull) continue;
                    if (atomicCAS(&sl[base + j], 0ull, v) == 0ull) return;
                }
            }
            if (h.err) atomicOr(h.err, 1u);
        }
        return;
    }
    ulonglong2 *slots = (ulonglong2 *)h.slots;
    uint64_t sl = hash_mix(k) & h.mask;
    while (atomicCAS((unsigned long long *)&slots[sl].x, 0ull, (unsigned long long)k) != 0ull) sl = (sl + 1) & h.mask;
    slots[sl].y = rate_bits(r);
}

// Batched builds (JOBS): a key takes its pattern, docheck step and spawn
// weights from its job, a key in one of its job's two final layers gets the
// final rate (success -> max_scale, else zero), and successors are looked up
// under the same job's keys.
template <typename T, bool JOBS>
__global__ __launch_bounds__(1024) void k_solve(const uint64_t *__restrict__ keys, uint64_t n,
                                                const uint16_t *__restrict__ glut, PatternDev p,
                                                int do_check, HashDev h1, HashDev h2, TableDev<T> t1,
                                                TableDev<T> t2, double w2, double w4, T max_scale, T zero,
                                                T *__restrict__ rates, ulonglong2 *__restrict__ clear,
                                                uint64_t clear_n, HashDev own, JobArg ja) {
    extern __shared__ uint16_t lut[];
    // the table filled after this launch: disjoint from h1 / h2
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < clear_n;
         j += (uint64_t)gridDim.x * blockDim.x)
        clear[j] = make_ulonglong2(0ull, 0ull);
    stage_lut(lut, glut);
    // hash probes when both successor layers have a hash, else binary
    // searches narrowed by the bucket index (uniform over the launch)
    const bool use_hash = h1.present() && h2.present();
    // (the sizes of the successor books are read only by the searches: with
    // hashes their compaction may still be running on another stream)
    const uint64_t n1 = use_hash ? 0 : (t1.n_ptr ? *t1.n_ptr : t1.n);
    const uint64_t n2 = use_hash ? 0 : (t2.n_ptr ? *t2.n_ptr : t2.n);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t key = keys[i];
        uint64_t b = key;
        uint32_t job = 0;
        bool chk = do_check != 0;
        double W2 = w2, W4 = w4;
        if constexpr (JOBS) {
            b = job_board(key);
            job = job_of(key);
            const JobDev &J = ja.tab[job];
            if (ja.step >= J.steps - 2) {  // the job's final layers
                const T rv = is_success(b, J.p) ? max_scale : zero;
                rates[i] = rv;
                if (own.present() && rv > zero) hash_insert(own, key, rv);
                continue;
            }
            chk = ja.step > J.docheck;
            W2 = J.w2;
            W4 = J.w4;
        }
        const PatternDev &P = JOBS ? ja.tab[job].p : p;
        if (chk && is_success(b, P)) {
            rates[i] = max_scale;
            if (own.present()) hash_insert(own, key, max_scale);
            continue;
        }
        double sp = 0.0;
        uint32_t em = empty_cells(b);
        const int e = __popc(em);
        const SpawnMoves smv(b, lut);
        while (em) {
            const int k = __builtin_ctz(em);
            em &= em - 1;
            // the 8 successors of this cell: spawn 2 (j < 4) and spawn 4
            // (j >= 4) times the four moves; their probes advance in
            // lockstep so the loads of one round are all in flight together
            uint64_t skey[8], slot[8];
            uint32_t active = 0;
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const uint64_t sp_board = b | ((uint64_t)(s + 1) << (4 * k));
                uint64_t mv[4];
                smv.of(k, (uint64_t)(s + 1), lut, mv);
#pragma unroll
                for (int d = 0; d < 4; ++d) {
                    const int j = s * 4 + d;
                    const uint64_t c = canonical(mv[d], P.symm);
                    skey[j] = JOBS ? job_key(c, job) : c;
                    if (mv[d] != sp_board && is_pattern(mv[d], P)) active |= 1u << j;
                }
            }
            T r[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) r[j] = zero;
            if (!use_hash) {
                // bucket-narrowed binary searches, advanced in lockstep
                uint32_t lo[8], hi[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    lo[j] = hi[j] = 0;
                    const TableDev<T> &t = j < 4 ? t1 : t2;
                    if (((active >> j) & 1u) && (j < 4 ? n1 : n2)) {
                        const uint64_t bk = skey[j] >> (64 - t.bits);
                        lo[j] = t.idx[bk];
                        hi[j] = t.idx[bk + 1];
                    }
                }
                uint32_t open = active;
                while (open) {
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        if (lo[j] < hi[j]) {
                            const uint32_t mid = (lo[j] + hi[j]) >> 1;
                            if ((j < 4 ? t1.keys : t2.keys)[mid] < skey[j]) lo[j] = mid + 1;
                            else hi[j] = mid;
                        }
                    }
                    open = 0;
#pragma unroll
                    for (int j = 0; j < 8; ++j) open |= (lo[j] < hi[j] ? 1u : 0u) << j;
                }
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const TableDev<T> &t = j < 4 ? t1 : t2;
                    if (((active >> j) & 1u) && lo[j] < (j < 4 ? n1 : n2) && t.keys[lo[j]] == skey[j])
                        r[j] = t.rates[lo[j]];
                }
                active = 0;
            }
            // wide: slot[j] is the probed slot; compact: slot[j] is the home
            // bucket, step[j] = 4 * displacement + pair within the bucket and
            // want[j] the stored bits
            uint32_t step[8], want[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const HashDev &h = j < 4 ? h1 : h2;
                step[j] = 0;
                want[j] = 0;
                if (h.cslots) {
                    if (((active >> j) & 1u) && !chash_split(h, skey[j], slot[j], want[j]))
                        active &= ~(1u << j);  // constant bits differ: not in the layer
                } else {
                    slot[j] = hash_mix(skey[j]) & h.mask;
                }
            }
            while (active) {
                ulonglong2 ent[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    if ((active >> j) & 1u) {
                        const HashDev &h = j < 4 ? h1 : h2;
                        ent[j] = h.cslots ? ((const ulonglong2 *)h.cslots)[((slot[j] + (step[j] >> 2)) & h.mask) * 4 +
                                                                           (step[j] & 3)]
                                          : h.slots[slot[j]];
                    }
                }
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    if ((active >> j) & 1u) {
                        const HashDev &h = j < 4 ? h1 : h2;
                        if (h.cslots) {
                            if constexpr (sizeof(T) == 4) {
                                // two slots a round; an empty slot ends the run
                                const uint32_t tag = chash_tag(step[j] >> 2, want[j]);
                                if ((uint32_t)(ent[j].x >> 32) == tag) {
                                    r[j] = rate_from_bits<T>(ent[j].x & 0xffffffffull);
                                    active &= ~(1u << j);
                                } else if (ent[j].x == 0) {
                                    active &= ~(1u << j);
                                } else if ((uint32_t)(ent[j].y >> 32) == tag) {
                                    r[j] = rate_from_bits<T>(ent[j].y & 0xffffffffull);
                                    active &= ~(1u << j);
                                } else if (ent[j].y == 0 || ++step[j] >= 4 * kCompactMaxDisp) {
                                    active &= ~(1u << j);
                                }
                            }
                        } else if (ent[j].x == skey[j]) {
                            r[j] = rate_from_bits<T>(ent[j].y);
                            active &= ~(1u << j);
                        } else if (ent[j].x == 0) {
                            active &= ~(1u << j);  // absent: rate 0
                        } else {
                            slot[j] = (slot[j] + 1) & h.mask;
                        }
                    }
                }
            }
            T best[2] = {zero, zero};
#pragma unroll
            for (int j = 0; j < 8; ++j) best[j >> 2] = best[j >> 2] < r[j] ? r[j] : best[j >> 2];
            sp = __fma_rn((double)best[0], W2, sp);
            sp = __fma_rn((double)best[1], W4, sp);
        }
        const T rv = e > 0 ? (T)(sp / (double)e) : zero;
        rates[i] = rv;
        if (own.present() && rv > zero) hash_insert(own, key, rv);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_hash_build(const uint64_t *__restrict__ keys, const T *__restrict__ rates,
                                                    uint64_t bound, const uint64_t *__restrict__ n_ptr, HashDev h) {
    const uint64_t n = n_ptr ? *n_ptr : bound;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t k = keys[i];
        if (k == 0) continue;
        hash_insert(h, k, rates[i]);
    }
}

// Final layers (reference final_situation_process, BookSolver.cpp:237-261).
template <typename T, bool JOBS>
__global__ void k_final_rates(const uint64_t *__restrict__ keys, uint64_t n, PatternDev p, T max_scale, T zero,
                              T *__restrict__ rates, JobArg ja) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t k = keys[i];
    if constexpr (JOBS) rates[i] = is_success(job_board(k), ja.tab[job_of(k)].p) ? max_scale : zero;
    else rates[i] = is_success(k, p) ? max_scale : zero;
}

// ---------------------------------------------------------------------------
// Optimal-branch marking (reference find_optimal_branches,
// BookSolver.cpp:611-646): for every parent and every empty cell, the parent
// with `spawn` (1 = a 2, 2 = a 4) on that cell has its four moves looked up
// in the target layer; the first move (left, right, up, down) with the
// strictly largest rate above zero marks its record.  A spawn with no such
// move marks record 0, as the reference's best_index = 0 does.
// ---------------------------------------------------------------------------

template <typename T>
__global__ __launch_bounds__(1024) void k_opt_mark(const uint64_t *__restrict__ parents, uint64_t n,
                                                   const uint16_t *__restrict__ glut, PatternDev p, uint64_t spawn,
                                                   TableDev<T> t, T zero, uint8_t *__restrict__ mark) {
    extern __shared__ uint16_t lut[];
    stage_lut(lut, glut);
    if (t.n == 0) return;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t b = parents[i];
        uint32_t em = empty_cells(b);
        const SpawnMoves smv(b, lut);
        while (em) {
            const int k = __builtin_ctz(em);
            em &= em - 1;
            const uint64_t sp = b | (spawn << (4 * k));
            uint64_t mv[4];
            smv.of(k, spawn, lut, mv);
            T best = zero;
            uint64_t best_idx = 0;
#pragma unroll
            for (int d = 0; d < 4; ++d) {
                if (mv[d] == sp || !is_pattern(mv[d], p)) continue;
                uint64_t pos;
                T r;
                if (table_find_pos(t, t.n, canonical(mv[d], p.symm), r, pos) && r > best) {
                    best = r;
                    best_idx = pos;
                }
            }
            mark[best_idx] = 1;
        }
    }
}

// ---------------------------------------------------------------------------
// Stable compaction of rate > 0 (reference compact_live_entries,
// BookSolver.cpp:380-394), SoA in -> SoA out; also max rate for stats.
// ---------------------------------------------------------------------------

constexpr int kCompThreads = 256;
constexpr int kCompItems = 8;
constexpr int kCompTile = kCompThreads * kCompItems;

// Keep the records with rate > thr, in order, in one launch: tiles are
// claimed in order by a ticket, count their kept records, publish the count
// and take their output offset by decoupled look-back. max_bits gets the
// bits of the largest rate; count_out the number kept. With hslots, every
// kept record also goes into that cleared open-addressing table (the solve
// hash of the layer: keys are unique, so a claim ends at the first empty
// slot of its run).
template <typename T>
__global__ __launch_bounds__(kCompThreads) void k_compact_lb(const uint64_t *__restrict__ keys,
                                                             const T *__restrict__ rates, uint64_t n, T thr,
                                                             uint64_t *__restrict__ out_keys,
                                                             T *__restrict__ out_rates,
                                                             unsigned long long *__restrict__ max_bits,
                                                             uint64_t *__restrict__ count_out,
                                                             uint64_t *__restrict__ status,
                                                             unsigned long long *__restrict__ tile_counter,
                                                             unsigned long long tile_base, uint32_t epoch,
                                                             uint32_t *__restrict__ err, HashDev fill,
                                                             const uint8_t *__restrict__ keep) {
    __shared__ uint64_t scratch[kCompThreads / 64];
    __shared__ uint32_t s_tile;
    __shared__ uint64_t s_off;
    if (threadIdx.x == 0) s_tile = (uint32_t)(atomicAdd(tile_counter, 1ull) - tile_base);
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint64_t base = (uint64_t)tile * kCompTile + (uint64_t)threadIdx.x * kCompItems;
    T r[kCompItems];
    uint64_t cnt = 0;
    T mx = thr;
    uint32_t kept = 0;  // bit j: item j is kept (rate > thr, or its keep byte is set)
#pragma unroll
    for (int j = 0; j < kCompItems; ++j) {
        const uint64_t idx = base + j;
        r[j] = idx < n ? rates[idx] : thr;
        const bool k = idx < n && (keep ? keep[idx] != 0 : r[j] > thr);
        kept |= (uint32_t)k << j;
        cnt += k;
        if (k && !keep) mx = mx < r[j] ? r[j] : mx;
    }
    uint64_t total;
    const uint64_t ex = block_exclusive_sum(cnt, scratch, total);
    if (threadIdx.x == 0) {
        if (tile == 0) {
            lb_store(status, lb_pack(epoch, 2, total));
            s_off = 0;
        } else {
            lb_store(status + tile, lb_pack(epoch, 1, total));
            const uint64_t pre = lb_lookback(status, 1, tile, epoch, err);
            lb_store(status + tile, lb_pack(epoch, 2, pre + total));
            s_off = pre;
        }
    }
    __syncthreads();
    uint64_t pos = s_off + ex;
    if (count_out && (uint64_t)(tile + 1) * kCompTile >= n && threadIdx.x == 0) *count_out = s_off + total;
#pragma unroll
    for (int j = 0; j < kCompItems; ++j) {
        if ((kept >> j) & 1u) {
            const uint64_t k = keys[base + j];
            out_keys[pos] = k;
            out_rates[pos] = r[j];
            ++pos;
            if (fill.present()) hash_insert(fill, k, r[j]);
        }
    }
    // the largest kept rate, as its order-preserving key (0 = none kept)
    if (mx > thr) atomicMax(max_bits, (unsigned long long)rate_order_key(mx));
}

// SoA -> packed records {u64 board; T rate} (SuccessEntry<T>,
// FormationRuntime.h:190-200), as 32-bit words.
template <typename T>
__global__ void k_pack_records(const uint64_t *__restrict__ keys, const T *__restrict__ rates, uint64_t n,
                               uint32_t *__restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    constexpr int W = (8 + sizeof(T)) / 4;
    uint32_t *rec = out + i * W;
    const uint64_t k = keys[i];
    rec[0] = (uint32_t)k;
    rec[1] = (uint32_t)(k >> 32);
    const uint64_t rb = rate_bits(rates[i]);
    rec[2] = (uint32_t)rb;
    if constexpr (W == 4) rec[3] = (uint32_t)(rb >> 32);
}

// out[k] = src[idx[k]] for k < m (offsets at chunk boundaries).
__global__ void k_gather_u64(const uint64_t *__restrict__ src, const uint64_t *__restrict__ idx, uint32_t m,
                             uint64_t *__restrict__ out) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < m) out[k] = src[idx[k]];
}

// *dst = *src, one thread: a device-side count into host-mapped memory, so
// the host reads it after an event without a copy of its own.
__global__ void k_copy_u64(const uint64_t *__restrict__ src, uint64_t *__restrict__ dst) { *dst = *src; }

// Packed records -> SoA (the reader's table load).
__global__ void k_unpack_records(const uint32_t *__restrict__ recs, uint64_t n, int words, uint64_t *__restrict__ keys,
                                 uint32_t *__restrict__ rates) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t *r = recs + i * words;
    keys[i] = (uint64_t)r[0] | ((uint64_t)r[1] << 32);
    rates[i * (words - 2)] = r[2];
    if (words == 4) rates[i * 2 + 1] = r[3];
}

// ===========================================================================
// Launch wrappers
// ===========================================================================

static inline uint32_t blocks_for(uint64_t n, uint64_t per) { return (uint32_t)((n + per - 1) / per); }

template <typename T>
static void configure_solve(int lds) {
    (void)hipFuncSetAttribute((const void *)k_solve<T, false>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    (void)hipFuncSetAttribute((const void *)k_solve<T, true>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
}

void configure_kernels() {
    const int lds = kLutEntries * sizeof(uint16_t);
    (void)hipFuncSetAttribute((const void *)k_expand<true, false>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    (void)hipFuncSetAttribute((const void *)k_expand<false, false>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    (void)hipFuncSetAttribute((const void *)k_expand<true, true>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    (void)hipFuncSetAttribute((const void *)k_expand<false, true>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    (void)hipFuncSetAttribute((const void *)k_expand_alloc<false>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    (void)hipFuncSetAttribute((const void *)k_expand_alloc<true>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    (void)hipFuncSetAttribute((const void *)k_expand_hash<false>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    (void)hipFuncSetAttribute((const void *)k_expand_hash<true>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    configure_solve<float>(lds);
    configure_solve<double>(lds);
    configure_solve<uint32_t>(lds);
    configure_solve<uint64_t>(lds);
    (void)hipFuncSetAttribute((const void *)k_opt_mark<float>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    (void)hipFuncSetAttribute((const void *)k_opt_mark<double>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    (void)hipFuncSetAttribute((const void *)k_opt_mark<uint32_t>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    (void)hipFuncSetAttribute((const void *)k_opt_mark<uint64_t>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
}

void launch_build_lut(hipStream_t s, uint16_t *lut) {
    hipLaunchKernelGGL(k_build_lut, dim3(kLutEntries / 256), dim3(256), 0, s, lut);
}

void launch_expand(hipStream_t s, bool write, const LaunchShape &ls, const uint64_t *par, uint64_t n_bound,
                   const uint64_t *n_ptr, const uint16_t *lut, const PatternDev &p, int do_check,
                   uint64_t *counts, const uint64_t *offs, uint64_t *out2, uint64_t *out4, uint64_t *or_and2,
                   uint64_t *or_and4, uint64_t offs_base, uint64_t cap2, uint64_t cap4, const JobArg &ja) {
    if (n_bound == 0) return;
    const uint32_t want = blocks_for(n_bound, ls.threads);
    const uint32_t grid = want < ls.blocks ? want : ls.blocks;
    const size_t lds = kLutEntries * sizeof(uint16_t);
    if (!write) {
        offs_base = 0;
        cap2 = cap4 = 0;
    }
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(grid), dim3(ls.threads), lds, s, par, n_bound, n_ptr, lut, p, do_check, counts,
                           offs, offs_base, out2, out4, or_and2, or_and4, cap2, cap4, ja);
    };
    if (ja.tab) write ? go(k_expand<true, true>) : go(k_expand<false, true>);
    else write ? go(k_expand<true, false>) : go(k_expand<false, false>);
}

void launch_expand_alloc(hipStream_t s, const LaunchShape &ls, const uint64_t *par, uint64_t n_bound,
                         const uint64_t *n_ptr, const uint16_t *lut, const PatternDev &p, int do_check,
                         uint64_t *tot, uint64_t *out2, uint64_t *out4, uint64_t *or_and2, uint64_t *or_and4,
                         uint64_t cap2, uint64_t cap4, const JobArg &ja) {
    if (n_bound == 0) return;
    const uint32_t want = blocks_for(n_bound, ls.threads);
    const uint32_t grid = want < ls.blocks ? want : ls.blocks;
    const size_t lds = kLutEntries * sizeof(uint16_t);
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(grid), dim3(ls.threads), lds, s, par, n_bound, n_ptr, lut, p, do_check,
                           (unsigned long long *)tot, out2, out4, or_and2, or_and4, cap2, cap4, ja);
    };
    if (ja.tab) go(k_expand_alloc<true>);
    else go(k_expand_alloc<false>);
}

void launch_expand_hash(hipStream_t s, const LaunchShape &ls, const uint64_t *par, uint64_t n_bound,
                        const uint64_t *n_ptr, const uint16_t *lut, const PatternDev &p, int do_check, uint64_t *tot,
                        uint64_t *set2, uint64_t cap2, uint64_t *set4, uint64_t cap4, uint64_t *meta2,
                        uint64_t *meta4, const JobArg &ja) {
    if (n_bound == 0) return;
    const uint32_t want = blocks_for(n_bound, ls.threads);
    const uint32_t grid = want < ls.blocks ? want : ls.blocks;
    const size_t lds = kLutEntries * sizeof(uint16_t);
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(grid), dim3(ls.threads), lds, s, par, n_bound, n_ptr, lut, p, do_check,
                           (unsigned long long *)tot, set2, set2 ? cap2 - 1 : 0ull, set4, set4 ? cap4 - 1 : 0ull,
                           meta2, meta4, ja);
    };
    if (ja.tab) go(k_expand_hash<true>);
    else go(k_expand_hash<false>);
}

void launch_set_insert(hipStream_t s, const uint64_t *keys, uint64_t n, uint64_t *set, uint64_t cap) {
    if (n == 0) return;
    const uint64_t want = blocks_for(n, 256);
    hipLaunchKernelGGL(k_set_insert, dim3(want < 65536 ? want : 65536), dim3(256), 0, s, keys, n, set, cap - 1);
}

void launch_set_extract(hipStream_t s, uint64_t *set, uint64_t cap, uint64_t *out, uint64_t *count) {
    if (cap == 0) return;
    hipLaunchKernelGGL(k_set_extract, dim3(blocks_for(cap, kExtractTile)), dim3(kExtractThreads), 0, s, set, cap, out,
                       (unsigned long long *)count);
}

template <typename TI, typename TO>
static void scan_impl(hipStream_t s, const TI *in, uint64_t n_bound, const uint64_t *n_ptr, TO *out, TO *partial,
                      TO *total_out) {
    const uint32_t tiles = blocks_for(n_bound ? n_bound : 1, kScanTile);
    hipLaunchKernelGGL((k_scan_reduce<TI, TO>), dim3(tiles), dim3(kScanThreads), 0, s, in, n_bound, n_ptr,
                       partial);
    hipLaunchKernelGGL((k_scan_partials<TO>), dim3(1), dim3(kScanThreads), 0, s, partial, (uint64_t)tiles);
    hipLaunchKernelGGL((k_scan_down<TI, TO>), dim3(tiles), dim3(kScanThreads), 0, s, in, n_bound, n_ptr, out,
                       partial, total_out);
}

uint64_t scan_partials_needed(uint64_t n) { return blocks_for(n ? n : 1, kScanTile); }

void launch_scan_u64(hipStream_t s, const uint64_t *in, uint64_t n_bound, const uint64_t *n_ptr, uint64_t *out,
                     uint64_t *partial, uint64_t *total_out) {
    scan_impl<uint64_t, uint64_t>(s, in, n_bound, n_ptr, out, partial, total_out);
}

void launch_scan_u32(hipStream_t s, const uint32_t *in, uint64_t n, uint32_t *out, uint32_t *partial) {
    scan_impl<uint32_t, uint32_t>(s, in, n, nullptr, out, partial, nullptr);
}

// Keys per thread of the classic schedule's tiles: small sorts get more,
// smaller tiles, so that each CU has several workgroups to hide latency
// with; 11-bit digits keep the large tile (their per-tile bin work grows
// with the tile count). Measured on gfx950 with 8-bit digits, ms per sort:
// 256 Ki keys 0.106 / 0.124 / 0.161 (4 / 8 / 16 items), 1 Mi 0.178 / 0.172 /
// 0.189, 4 Mi 0.367 / 0.325 / 0.323, 16 Mi 1.086 / 0.92 / 0.84.
static int classic_items(int bits, uint64_t n) {
    if (bits >= 11) return 16;
    const int t = tuned_tile(kTileRadixItems, n);
    if (t == 4 || t == 8 || t == 16) return t;
    return n <= (1u << 19) ? 4 : n <= (1u << 22) ? 8 : 16;
}

static std::mutex g_tile_mu;
static std::vector<std::pair<uint64_t, int>> g_tile_rows[kTileKinds];

void set_tile_tuning(int kind, std::vector<std::pair<uint64_t, int>> rows) {
    std::lock_guard<std::mutex> lk(g_tile_mu);
    g_tile_rows[kind] = std::move(rows);
}

int tuned_tile(int kind, uint64_t n) {
    std::lock_guard<std::mutex> lk(g_tile_mu);
    const auto &rows = g_tile_rows[kind];
    for (const auto &r : rows)
        if (n <= r.first) return r.second;
    return rows.empty() ? 0 : rows.back().second;
}

uint32_t radix_tile(int bits, uint64_t n, int schedule) {
    return schedule == 2 ? (uint32_t)kSortTile : (uint32_t)(kSortThreads * classic_items(bits, n));
}

uint64_t unique_tiles(uint64_t n) { return blocks_for(n ? n : 1, kUniqTile); }

template <int BITS>
static uint64_t *onesweep_impl(hipStream_t s, uint64_t *keys, uint64_t *tmp, uint64_t n, int lo, int passes,
                               LookbackWork &lw) {
    constexpr int BINS = 1 << BITS;
    const uint32_t ntiles = blocks_for(n, kSortTile);
    const int cur = lw.hist_parity, other = cur ^ 1;
    uint32_t *hist = lw.hist + (cur ? kHistWords : 0);
    uint32_t *next = lw.hist + (other ? kHistWords : 0);
    lw.hist_parity = other;
    const uint32_t hgrid = blocks_for(n, 8192) < 256 ? blocks_for(n, 8192) : 256;
    hipLaunchKernelGGL((k_os_hist<BITS>), dim3(hgrid), dim3(kSortThreads), 0, s, keys, n, lo, passes, hist);
    hipLaunchKernelGGL((k_os_scan<BITS>), dim3(1), dim3(kSortThreads), 0, s, hist, passes, next, lw.dirty[other]);
    lw.dirty[other] = 0;
    lw.dirty[cur] = passes * BINS;
    uint64_t *src = keys, *dst = tmp;
    for (int q = 0; q < passes; ++q) {
        const uint32_t epoch = lw.next_epoch();
        hipLaunchKernelGGL((k_os_scatter<BITS>), dim3(ntiles), dim3(kSortThreads), 0, s, src, dst, n,
                           lo + q * BITS, hist + q * BINS, lw.status_sort, lw.counters, lw.claimed, epoch, lw.err);
        lw.claimed += ntiles;
        std::swap(src, dst);
    }
    return src;
}

template <int BITS, int ITEMS>
static void classic_pass_impl(hipStream_t s, const uint64_t *src, uint64_t *dst, uint64_t n, int shift,
                              uint32_t *counts, uint32_t *offsets) {
    const uint32_t ntiles = blocks_for(n, kSortThreads * ITEMS);
    uint32_t *totals = offsets + ((uint64_t)ntiles << BITS);
    {
        Timer t(s, "radix_hist", 8.0 * (double)n);  // reads every key once
        hipLaunchKernelGGL((k_radix_hist<BITS, ITEMS>), dim3(ntiles), dim3(kSortThreads), 0, s, src, n, shift,
                           counts, ntiles);
    }
    hipLaunchKernelGGL((k_radix_binscan<BITS>), dim3(1u << BITS), dim3(kSortThreads), 0, s, counts, ntiles, offsets,
                       totals);
    {
        Timer t(s, "radix_scatter", 16.0 * (double)n);  // reads and writes every key once
        hipLaunchKernelGGL((k_radix_scatter<BITS, ITEMS>), dim3(ntiles), dim3(kSortThreads), 0, s, src, dst, n,
                           shift, offsets, totals, ntiles);
    }
}

template <int BITS>
static void classic_pass(hipStream_t s, const uint64_t *src, uint64_t *dst, uint64_t n, int shift,
                         uint32_t *counts, uint32_t *offsets) {
    switch (classic_items(BITS, n)) {
        case 4: classic_pass_impl<BITS, 4>(s, src, dst, n, shift, counts, offsets); break;
        case 8: classic_pass_impl<BITS, 8>(s, src, dst, n, shift, counts, offsets); break;
        default: classic_pass_impl<BITS, 16>(s, src, dst, n, shift, counts, offsets); break;
    }
}

uint64_t *launch_classic_sort(hipStream_t s, int bits, uint64_t *keys, uint64_t *tmp, uint64_t n, int lo, int hi,
                              uint32_t *counts, uint32_t *offsets) {
    if (n < 2 || hi <= lo) return keys;
    uint64_t *src = keys, *dst = tmp;
    for (int shift = lo; shift < hi; shift += bits) {
        switch (bits) {
            case 4: classic_pass<4>(s, src, dst, n, shift, counts, offsets); break;
            case 6: classic_pass<6>(s, src, dst, n, shift, counts, offsets); break;
            case 11: classic_pass_impl<11, 16>(s, src, dst, n, shift, counts, offsets); break;
            default: classic_pass<8>(s, src, dst, n, shift, counts, offsets); break;
        }
        std::swap(src, dst);
    }
    return src;
}

uint64_t *launch_onesweep_sort(hipStream_t s, int bits, uint64_t *keys, uint64_t *tmp, uint64_t n, int lo, int hi,
                               LookbackWork &lw) {
    if (n < 2 || hi <= lo) return keys;
    const int passes = (hi - lo + bits - 1) / bits;
    switch (bits) {
        case 4: return onesweep_impl<4>(s, keys, tmp, n, lo, passes, lw);
        case 6: return onesweep_impl<6>(s, keys, tmp, n, lo, passes, lw);
        case 11: return onesweep_impl<11>(s, keys, tmp, n, lo, passes, lw);
        default: return onesweep_impl<8>(s, keys, tmp, n, lo, passes, lw);
    }
}

void launch_unique(hipStream_t s, const uint64_t *a, uint64_t n, uint64_t *out, uint64_t *count_out,
                   LookbackWork &lw) {
    if (n == 0) {
        if (count_out) (void)hipMemsetAsync(count_out, 0, 8, s);
        return;
    }
    const uint32_t tiles = (uint32_t)unique_tiles(n);
    const uint32_t epoch = lw.next_epoch();
    hipLaunchKernelGGL(k_unique_lb, dim3(tiles), dim3(kUniqThreads), 0, s, a, n, out, count_out, lw.status_uniq,
                       lw.counters + 1, lw.claimed_uniq, epoch, lw.err);
    lw.claimed_uniq += tiles;
}

void launch_bucket_index(hipStream_t s, const uint64_t *keys, uint64_t n_bound, const uint64_t *n_ptr,
                         uint32_t bits, uint32_t *idx) {
    const uint64_t nb = (1ull << bits) + 1;
    hipLaunchKernelGGL(k_bucket_index, dim3(blocks_for(nb, 256)), dim3(256), 0, s, keys, n_bound, n_ptr, bits, idx);
}

template <typename T>
void launch_solve(hipStream_t s, const LaunchShape &ls, const uint64_t *keys, uint64_t n, const uint16_t *lut,
                  const PatternDev &p, int do_check, const HashDev &h1, const HashDev &h2,
                  const TableDev<T> &t1, const TableDev<T> &t2, double w2, double w4, T max_scale, T zero,
                  T *rates, const HashDev &clear, const HashDev &own, const JobArg &ja) {
    const uint64_t clear_n = clear.bytes() / sizeof(ulonglong2);
    if (n == 0) {
        if (clear_n) (void)hipMemsetAsync(clear.table(), 0, clear.bytes(), s);
        return;
    }
    const uint32_t want = blocks_for(n > clear_n ? n : clear_n, ls.threads);
    const uint32_t grid = want < ls.blocks ? want : ls.blocks;
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(grid), dim3(ls.threads), kLutEntries * sizeof(uint16_t), s, keys, n, lut, p,
                           do_check, h1, h2, t1, t2, w2, w4, max_scale, zero, rates, (ulonglong2 *)clear.table(),
                           clear_n, own, ja);
    };
    if (ja.tab) go(k_solve<T, true>);
    else go(k_solve<T, false>);
}

// Slots for a layer of at most `bound` records: a power of two, at most a
// quarter full. Short probe runs matter more than bytes here: the solve's
// probes advance in lockstep, so a round waits for the longest run among a
// wave's 512 probes. 442@512 backward pass, one MI355X: 35.2 ms at 2x the
// bound, 32.4 ms at 4x, 33.5 ms at 8x.
uint64_t hash_capacity(uint64_t bound) {
    uint64_t c = 64;
    while (c < 4 * bound) c <<= 1;
    return c;
}

template <typename T>
void launch_hash_build(hipStream_t s, const uint64_t *keys, const T *rates, uint64_t bound, const uint64_t *n_ptr,
                       const HashDev &h) {
    (void)hipMemsetAsync(h.table(), 0, h.bytes(), s);
    if (bound == 0) return;
    const uint32_t grid = blocks_for(bound, 256) < 8192 ? blocks_for(bound, 256) : 8192;
    hipLaunchKernelGGL((k_hash_build<T>), dim3(grid), dim3(256), 0, s, keys, rates, bound, n_ptr, h);
}

template <typename T>
void launch_final_rates(hipStream_t s, const uint64_t *keys, uint64_t n, const PatternDev &p, T max_scale, T zero,
                        T *rates, const JobArg &ja) {
    if (n == 0) return;
    if (ja.tab)
        hipLaunchKernelGGL((k_final_rates<T, true>), dim3(blocks_for(n, 256)), dim3(256), 0, s, keys, n, p, max_scale,
                           zero, rates, ja);
    else
        hipLaunchKernelGGL((k_final_rates<T, false>), dim3(blocks_for(n, 256)), dim3(256), 0, s, keys, n, p,
                           max_scale, zero, rates, ja);
}

uint64_t compact_tiles(uint64_t n) { return blocks_for(n ? n : 1, kCompTile); }

template <typename T>
void launch_opt_mark(hipStream_t s, const LaunchShape &ls, const uint64_t *parents, uint64_t n, const uint16_t *lut,
                     const PatternDev &p, uint64_t spawn, const TableDev<T> &t, T zero, uint8_t *mark) {
    if (n == 0 || t.n == 0) return;
    const uint32_t want = blocks_for(n, ls.threads);
    const uint32_t grid = want < ls.blocks ? want : ls.blocks;
    hipLaunchKernelGGL((k_opt_mark<T>), dim3(grid), dim3(ls.threads), kLutEntries * sizeof(uint16_t), s, parents, n,
                       lut, p, spawn, t, zero, mark);
}

template <typename T>
void launch_compact(hipStream_t s, const uint64_t *keys, const T *rates, uint64_t n, uint64_t *out_keys,
                    T *out_rates, unsigned long long *max_bits, uint64_t *count_out, LookbackWork &lw, T threshold,
                    const HashDev &fill, bool zeroed, const uint8_t *keep) {
    if (fill.present() && !zeroed) (void)hipMemsetAsync(fill.table(), 0, fill.bytes(), s);
    if (n == 0) {
        if (count_out) (void)hipMemsetAsync(count_out, 0, 8, s);
        return;
    }
    const uint32_t tiles = (uint32_t)compact_tiles(n);
    const uint32_t epoch = lw.next_epoch();
    hipLaunchKernelGGL((k_compact_lb<T>), dim3(tiles), dim3(kCompThreads), 0, s, keys, rates, n, threshold,
                       out_keys, out_rates, max_bits, count_out, lw.status_uniq, lw.counters + 1, lw.claimed_uniq,
                       epoch, lw.err, fill, keep);
    lw.claimed_uniq += tiles;
}

template <typename T>
void launch_pack_records(hipStream_t s, const uint64_t *keys, const T *rates, uint64_t n, uint32_t *out) {
    if (n == 0) return;
    hipLaunchKernelGGL((k_pack_records<T>), dim3(blocks_for(n, 256)), dim3(256), 0, s, keys, rates, n, out);
}

void launch_gather_u64(hipStream_t s, const uint64_t *src, const uint64_t *idx, uint32_t m, uint64_t *out) {
    if (m) hipLaunchKernelGGL(k_gather_u64, dim3((m + 255) / 256), dim3(256), 0, s, src, idx, m, out);
}

void launch_copy_u64(hipStream_t s, const uint64_t *src, uint64_t *dst) {
    hipLaunchKernelGGL(k_copy_u64, dim3(1), dim3(1), 0, s, src, dst);
}

void launch_unpack_records(hipStream_t s, const uint32_t *recs, uint64_t n, int rate_size, uint64_t *keys,
                           void *rates) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_unpack_records, dim3(blocks_for(n, 256)), dim3(256), 0, s, recs, n, (8 + rate_size) / 4,
                       keys, (uint32_t *)rates);
}

#define EGTB_INSTANTIATE(T)                                                                                   \
    template void launch_solve<T>(hipStream_t, const LaunchShape &, const uint64_t *, uint64_t,              \
                                  const uint16_t *, const PatternDev &, int, const HashDev &,                 \
                                  const HashDev &, const TableDev<T> &, const TableDev<T> &, double, double,  \
                                  T, T, T *, const HashDev &, const HashDev &, const JobArg &);               \
    template void launch_hash_build<T>(hipStream_t, const uint64_t *, const T *, uint64_t, const uint64_t *,  \
                                       const HashDev &);                                                      \
    template void launch_final_rates<T>(hipStream_t, const uint64_t *, uint64_t, const PatternDev &, T, T, T *, \
                                        const JobArg &);                                                      \
    template void launch_compact<T>(hipStream_t, const uint64_t *, const T *, uint64_t, uint64_t *, T *,      \
                                    unsigned long long *, uint64_t *, LookbackWork &, T, const HashDev &,     \
                                    bool, const uint8_t *);                                                   \
    template void launch_opt_mark<T>(hipStream_t, const LaunchShape &, const uint64_t *, uint64_t,            \
                                     const uint16_t *, const PatternDev &, uint64_t, const TableDev<T> &, T,  \
                                     uint8_t *);                                                              \
    template void launch_pack_records<T>(hipStream_t, const uint64_t *, const T *, uint64_t, uint32_t *);

EGTB_INSTANTIATE(float)
EGTB_INSTANTIATE(double)
EGTB_INSTANTIATE(uint32_t)
EGTB_INSTANTIATE(uint64_t)

}  // namespace egtb
