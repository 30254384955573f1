// Device-side board primitives for the classic tablebase kernels (gfx950).
//
// Board encoding (reference native_core/include/BoardCodec.h:20-29): a
// uint64 of 16 nibbles, nibble k = cell (row 3-k/4, column 3-k%4), value v is
// the tile 2^v, 0 is empty, 0xF is 32768.  Row 0 is the top 16 bits.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace egtb {

constexpr int kMaxMasks = 16;
constexpr int kLutEntries = 65536;  // one u16 result row per 16-bit row

// Pattern parameters passed by value to every kernel (reference PatternSpec,
// FormationRuntime.h:17-22, plus the target of RunOptions).
struct PatternDev {
    uint64_t masks[kMaxMasks];
    uint64_t success_mask;    // 0xF at every success nibble
    uint64_t target_stacked;  // target log2 in every nibble
    uint32_t n_masks;
    uint32_t symm;            // 0 identity, 1 full, 2 diagonal, 3 horizontal
    uint32_t any_success;     // 1 when the pattern lists no success nibble
    uint32_t pad;
};

// One job of a batched build: many (pattern, target) builds share one key
// space.  Every mask of a batched job pins nibbles 0 and 1, so the low byte
// of each of its boards is 0xFF and a board of job j is stored as the key
// (board >> 8) | j << 56.  Sorted keys then group the jobs and order each
// job's boards as their plain values; kernels take a key's pattern, layer
// count, docheck step and spawn weights from the job table.
struct JobDev {
    PatternDev p;
    int32_t steps;    // layers of the job (its final layers are steps-2, steps-1)
    int32_t docheck;  // success boards are terminal in layers above this one
    double w2, w4;    // spawn weights: 1 - p4, p4
};

__device__ __forceinline__ uint64_t job_key(uint64_t board, uint32_t job) {
    return (board >> 8) | ((uint64_t)job << 56);
}
__device__ __forceinline__ uint64_t job_board(uint64_t key) { return (key << 8) | 0xFFull; }
__device__ __forceinline__ uint32_t job_of(uint64_t key) { return (uint32_t)(key >> 56); }

// Main-diagonal transpose of the 4x4 nibble matrix by two delta swaps
// (2x2 blocks, then cells inside each block).  Same map as the reference's
// reverse_board / ReverseUL (CommonMover.h:33-37).
__device__ __forceinline__ uint64_t transpose(uint64_t x) {
    uint64_t t = ((x >> 24) ^ x) & 0x00000000FF00FF00ull;
    x ^= t ^ (t << 24);
    t = ((x >> 12) ^ x) & 0x0000F0F00000F0F0ull;
    x ^= t ^ (t << 12);
    return x;
}

// Reverse the nibble order inside every 16-bit row (column mirror,
// Calculator::ReverseLR, Calculator.h:28-32).
__device__ __forceinline__ uint64_t mirror_lr(uint64_t x) {
    x = ((x >> 8) & 0x00FF00FF00FF00FFull) | ((x & 0x00FF00FF00FF00FFull) << 8);
    x = ((x >> 4) & 0x0F0F0F0F0F0F0F0Full) | ((x & 0x0F0F0F0F0F0F0F0Full) << 4);
    return x;
}

// Row mirror (Calculator::ReverseUD, Calculator.h:34-38).
__device__ __forceinline__ uint64_t mirror_ud(uint64_t x) {
    x = (x >> 32) | (x << 32);
    x = ((x >> 16) & 0x0000FFFF0000FFFFull) | ((x & 0x0000FFFF0000FFFFull) << 16);
    return x;
}

// Slide all four rows toward their high nibble through the LDS row table.
__device__ __forceinline__ uint64_t rows_left(uint64_t b, const uint16_t *lut) {
    return (uint64_t)lut[b & 0xFFFF] | ((uint64_t)lut[(b >> 16) & 0xFFFF] << 16) |
           ((uint64_t)lut[(b >> 32) & 0xFFFF] << 32) | ((uint64_t)lut[b >> 48] << 48);
}

// The four moves in the reference's move_all_dir order (left, right, up,
// down; CommonMover.h:136-139).  Right is mirror-slide-mirror, which is how
// the reference defines its right merge (BoardMover.h:16,42); up/down act on
// the transpose.
__device__ __forceinline__ void move4(uint64_t b, const uint16_t *lut, uint64_t out[4]) {
    const uint64_t t = transpose(b);
    out[0] = rows_left(b, lut);
    out[1] = mirror_lr(rows_left(mirror_lr(b), lut));
    out[2] = transpose(rows_left(t, lut));
    out[3] = transpose(mirror_lr(rows_left(mirror_lr(t), lut)));
}

// Nibble order reversed within a 16-bit row.
__device__ __forceinline__ uint32_t rev_row(uint32_t x) {
    return ((x & 0xFu) << 12) | ((x & 0xF0u) << 4) | ((x >> 4) & 0xF0u) | (x >> 12);
}

// Nibble j of a 16-bit row to nibble 4j of a board (its column 0): a row of
// the transposed board back in place as a column.
__device__ __forceinline__ uint64_t row_to_col(uint64_t x) {
    return (x & 0xFull) | ((x & 0xF0ull) << 12) | ((x & 0xF00ull) << 24) | ((x & 0xF000ull) << 36);
}

// The four moves of every spawned board b | v << 4k (k an empty cell of b)
// from the moves of b: a spawn changes one row (k / 4) and one column
// (k % 4), so left/right redo that row and up/down that column through the
// row table -- 4 lookups a spawned board instead of move4's 16, same
// results (checked against the full moves on every cell of 200k boards).
struct SpawnMoves {
    uint64_t b, t, m[4];
    __device__ __forceinline__ SpawnMoves(uint64_t board, const uint16_t *lut) : b(board), t(transpose(board)) {
        move4(board, lut, m);
    }
    __device__ __forceinline__ void of(int k, uint64_t v, const uint16_t *lut, uint64_t out[4]) const {
        const int r = k >> 2, c = k & 3;
        const uint32_t row = (uint32_t)((b >> (16 * r)) & 0xFFFF) | (uint32_t)(v << (4 * c));
        const uint32_t col = (uint32_t)((t >> (16 * c)) & 0xFFFF) | (uint32_t)(v << (4 * r));
        const uint64_t rm = 0xFFFFull << (16 * r), cm = 0x000F000F000F000Full << (4 * c);
        out[0] = (m[0] & ~rm) | ((uint64_t)lut[row] << (16 * r));
        out[1] = (m[1] & ~rm) | ((uint64_t)rev_row(lut[rev_row(row)]) << (16 * r));
        out[2] = (m[2] & ~cm) | (row_to_col(lut[col]) << (4 * c));
        out[3] = (m[3] & ~cm) | (row_to_col(rev_row(lut[rev_row(col)])) << (4 * c));
    }
};

__device__ __forceinline__ uint64_t umin64(uint64_t a, uint64_t b) { return a < b ? a : b; }

// Canonical representative (apply_canonical, BookGenerator.cpp:231-249).
__device__ __forceinline__ uint64_t canonical(uint64_t b, uint32_t symm) {
    if (symm == 0) return b;
    if (symm == 3) return umin64(b, mirror_lr(b));
    const uint64_t t = transpose(b);
    if (symm == 2) return umin64(b, t);
    uint64_t m = umin64(b, t);
    const uint64_t lb = mirror_lr(b), ub = mirror_ud(b), lt = mirror_lr(t), ut = mirror_ud(t);
    m = umin64(m, umin64(lb, ub));
    m = umin64(m, umin64(mirror_ud(lb), lt));
    m = umin64(m, umin64(ut, mirror_ud(lt)));
    return m;
}

// is_pattern (Formation.h:11-17): some mask has all its nibbles at 0xF.
__device__ __forceinline__ bool is_pattern(uint64_t b, const PatternDev &p) {
    if (p.n_masks == 0) return true;
    bool ok = false;
    for (uint32_t i = 0; i < p.n_masks; ++i) ok |= (b & p.masks[i]) == p.masks[i];
    return ok;
}

// is_success_by_shifts (Formation.h:32-43) as a SWAR zero-nibble test of
// (board restricted to the success nibbles) xor (target in every nibble):
// a nibble outside the success set reads as the nonzero target, so only
// success nibbles equal to the target produce a zero nibble.
__device__ __forceinline__ bool is_success(uint64_t b, const PatternDev &p) {
    if (p.any_success) return true;
    const uint64_t d = (b & p.success_mask) ^ p.target_stacked;
    const uint64_t m7 = 0x7777777777777777ull;
    return (~(((d & m7) + m7) | d | m7)) != 0;
}

// 16-bit mask with bit k set when nibble k is empty.
__device__ __forceinline__ uint32_t empty_cells(uint64_t b) {
    uint64_t x = b | (b >> 1);
    x |= x >> 2;
    x = ~x & 0x1111111111111111ull;  // bit 4k set when nibble k == 0
    // gather bit 4k into bit k
    x = (x | (x >> 3)) & 0x0303030303030303ull;
    x = (x | (x >> 6)) & 0x000F000F000F000Full;
    x = (x | (x >> 12)) & 0x000000FF000000FFull;
    x = (x | (x >> 24)) & 0x000000000000FFFFull;
    return (uint32_t)x;
}

// Copy the 128 KiB left-slide row table from global memory into LDS.
__device__ __forceinline__ void stage_lut(uint16_t *lds, const uint16_t *__restrict__ g) {
    const uint4 *src = reinterpret_cast<const uint4 *>(g);
    uint4 *dst = reinterpret_cast<uint4 *>(lds);
    for (int i = threadIdx.x; i < kLutEntries * 2 / 16; i += blockDim.x) dst[i] = src[i];
    __syncthreads();
}

}  // namespace egtb
