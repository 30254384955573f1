// Device side of the heuristic expectimax player (the reference's AIPlayer,
// native_core/src/AIPlayer.cpp): evaluation, moves with merge scores, the
// dynamic mask and the fixed-depth search, shared by the batched search
// kernels (egtb_ai.hip) and the serving lookup's off-table fallback
// (egtb_serving.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "egtb_device.h"

namespace egtb {

constexpr int kAiVariants = 16;

struct AiTables {
    const int16_t *eval;   // [kAiVariants][65536][2]
    const uint32_t *rows;  // [65536] row move | score << 16
};

// Per-search state chosen by the dynamic mask.
struct AiRoot {
    uint64_t masked;
    uint64_t fixed_mask;
    uint32_t board_sum;
    int32_t dead;
    int32_t masked_count;
    int32_t variant;
};

struct AiEnv {
    const int16_t *eval;  // the variant's table
    const uint32_t *rows;
    int32_t dead;
    int32_t masked_count;
    int32_t do_check;
    int64_t w2, w4;
};

__device__ __forceinline__ uint32_t nib(uint64_t b, int k) { return (uint32_t)(b >> (4 * k)) & 0xFu; }

// tiles at or above `threshold` become 0xF (mask, AIPlayer.cpp:19-28)
__device__ __forceinline__ uint64_t ai_mask(uint64_t b, uint32_t threshold) {
    uint64_t out = b;
    for (int k = 0; k < 16; ++k)
        if (nib(b, k) >= threshold) out |= 0xFull << (4 * k);
    return out;
}

__device__ __forceinline__ uint64_t first_match(uint64_t b, const uint64_t *m, int n) {
    for (int i = 0; i < n; ++i)
        if ((b & m[i]) == m[i]) return m[i];
    return 0;
}

// corner shapes (get_23_mask .. get_20_mask, AIPlayer.cpp:152-252)
static __constant__ uint64_t c_m23[8] = {0xff00fff0ULL, 0xfff00ff00000000ULL, 0xf00ff00ffULL, 0xff00ff00f0000000ULL,
                                  0xfff0ff0000000000ULL, 0xff00ff000f0000ULL, 0xf000ff00ff00ULL, 0xff0fffULL};
static __constant__ uint64_t c_m22[4] = {0xff00ff0000000000ULL, 0x00ff00ff00000000ULL, 0x00000000ff00ff00ULL,
                                  0x0000000000ff00ffULL};
static __constant__ uint64_t c_m21[4] = {0xff00f00000000000ULL, 0x00ff000f00000000ULL, 0x00000000f000ff00ULL,
                                  0x00000000000f00ffULL};
static __constant__ uint64_t c_m20[8] = {0xff00ULL, 0xff000000000000ULL, 0xf000fULL, 0xf000f00000000000ULL,
                                  0xff00000000000000ULL, 0xf000f00000000ULL, 0xf000f000ULL, 0xffULL};

// check_corner (AIPlayer.cpp:470-643): shapes of masked tiles by masked
// count, each group with its penalty, tried in the reference's order.
struct CornerGroup {
    int8_t mc_lo, mc_hi;  // masked counts the group applies to
    int16_t n;
    int32_t penalty;
    uint64_t m[8];
};
static __constant__ CornerGroup c_corner[10] = {
    {6, 6, 8, -600, {0xf000fff0ffULL, 0xff0fff000f000000ULL, 0xff0fff0000000fULL, 0xf0000000fff0ff00ULL,
                     0xf0ff00ff00f00000ULL, 0xf00000fff00ffULL, 0xff00fff00000f000ULL, 0xf00ff00ff0fULL}},
    {6, 6, 8, -500, {0xf0000000ff0fff00ULL, 0xfff0ff0000000fULL, 0xf0000000fff0ffULL, 0xff0fff0000000f00ULL,
                     0xff00ff0f0000f000ULL, 0xf0ff00ff000000f0ULL, 0xf000000ff00ff0fULL, 0xf0000f0ff00ffULL}},
    {6, 6, 8, -1600, {0xf00000fff0ffULL, 0xff0fff00000f0000ULL, 0xff00ff00000f0fULL, 0xf0f00000ff00ff00ULL,
                      0xf0ff00fff0000000ULL, 0xf0f000000ff00ffULL, 0xff00ff000000f0f0ULL, 0xfff00ff0fULL}},
    {6, 6, 8, -2400, {0xff0fff0fULL, 0xf0fff0ff00000000ULL, 0xff000000ff00ffULL, 0xff00ff000000ff00ULL,
                      0xff0fff0f00000000ULL, 0xff00ff000000ffULL, 0xff000000ff00ff00ULL, 0xf0fff0ffULL}},
    {6, 6, 8, 600, {0xffff0ffULL, 0xff0ffff000000000ULL, 0xff00ff00f0000fULL, 0xf0000f00ff00ff00ULL,
                    0xf0ff0fff00000000ULL, 0xf00f000ff00ffULL, 0xff00ff000f00f000ULL, 0xfff0ff0fULL}},
    {4, 4, 8, 3000, {0xff00f00fULL, 0xff00f00000000fULL, 0xfff00fULL, 0xff000f000000f000ULL, 0xf00000000f00ff00ULL,
                     0xf00f00ff00000000ULL, 0xf00fff0000000000ULL, 0xf000000f000ffULL}},
    {4, 4, 8, 800, {0xf000f000f00fULL, 0xf000ff00fULL, 0xfff000000000f000ULL, 0xf00000000000fff0ULL,
                    0xf00f000f000f0000ULL, 0xf00ff000f0000000ULL, 0xf000000000fffULL, 0xfff00000000000fULL}},
    {4, 4, 8, 1600, {0xff000000f000f000ULL, 0xf000f000000ffULL, 0xf000f0ffULL, 0xff0f000f00000000ULL,
                     0xf000f0000000ff00ULL, 0xf0fff00000000000ULL, 0xfff0fULL, 0xff0000000f000fULL}},
    {3, 4, 4, 3000, {0xf00000000f00fULL, 0xf00f00000000000fULL, 0xf00000000000f00fULL, 0xf00f00000000f000ULL}},
    {3, 4, 8, 2000, {0xf0000000000f00fULL, 0xf000000000f00fULL, 0xf0000000000ff000ULL, 0xf000000f0000f000ULL,
                     0xf00f0000000000f0ULL, 0xf00f000000000f00ULL, 0xf0000f000000fULL, 0xff0000000000fULL}},
};
static __constant__ uint64_t c_m3_1000[8] = {0xff00fULL, 0xf000f00fULL, 0xf00000000000ff00ULL, 0xff0000000000f000ULL,
                                      0xf00ff00000000000ULL, 0xf00f000f00000000ULL, 0xff00000000000fULL,
                                      0xf0000000000ffULL};

__device__ int32_t check_corner(uint64_t b, int32_t mc) {
    for (int g = 0; g < 10; ++g) {
        const CornerGroup &G = c_corner[g];
        if (mc < G.mc_lo || mc > G.mc_hi) continue;
        for (int i = 0; i < G.n; ++i)
            if ((b & G.m[i]) == G.m[i]) return G.penalty;
    }
    if (mc == 3)
        for (int i = 0; i < 8; ++i)
            if ((b & c_m3_1000[i]) == c_m3_1000[i]) return 1000;
    return 0;
}

__device__ __forceinline__ int32_t ai_evaluate(uint64_t s, const AiEnv &E) {
    const uint64_t t = transpose(s);
    int32_t x1 = 0, x2 = 0, y1 = 0, y2 = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const uint32_t px = *reinterpret_cast<const uint32_t *>(E.eval + 2 * ((s >> (16 * i)) & 0xFFFF));
        const uint32_t py = *reinterpret_cast<const uint32_t *>(E.eval + 2 * ((t >> (16 * i)) & 0xFFFF));
        x1 += (int16_t)(px & 0xFFFF);
        x2 += (int16_t)(px >> 16);
        y1 += (int16_t)(py & 0xFFFF);
        y2 += (int16_t)(py >> 16);
    }
    int32_t r = max(x1, x2) + max(y1, y2);
    if (E.do_check) r -= check_corner(s, E.masked_count);
    return r;
}

// process_score (AIPlayer.cpp:688-699)
__device__ __forceinline__ int32_t process_score(uint32_t score) {
    const int32_t s = (int32_t)score;
    if (s < 200) return max(0, (s >> 2) - 10);
    if (s < 500) return (s >> 1) - 12;
    if (s < 1000) return (s >> 1) + 144;
    if (s < 2000) return s + 600;
    return 3000;
}

// left / right slides of the four rows with the summed merge score
__device__ __forceinline__ uint64_t slide_rows(uint64_t b, const uint32_t *rows, bool right, uint32_t &score) {
    if (right) b = mirror_lr(b);
    uint64_t out = 0;
    score = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const uint32_t e = rows[(b >> (16 * i)) & 0xFFFF];
        out |= (uint64_t)(e & 0xFFFF) << (16 * i);
        score += e >> 16;
    }
    return right ? mirror_lr(out) : out;
}

// s_move_board_all (CommonMover.h:181-193): moves in the order left, right,
// up, down with their merge scores; a move is valid when it changes the board
__device__ __forceinline__ void ai_moves(uint64_t b, const uint32_t *rows, uint64_t mv[4], uint32_t sc[4]) {
    const uint64_t t = transpose(b);
    mv[0] = slide_rows(b, rows, false, sc[0]);
    mv[1] = slide_rows(b, rows, true, sc[1]);
    mv[2] = transpose(slide_rows(t, rows, false, sc[2]));
    mv[3] = transpose(slide_rows(t, rows, true, sc[3]));
}

__device__ __forceinline__ uint64_t empty_nibbles(uint64_t b) {
    uint64_t x = b | (b >> 1);
    x |= x >> 2;
    return ~x & 0x1111111111111111ull;
}

// search0 (AIPlayer.cpp:669-686)
__device__ int32_t ai_search0(uint64_t b, const AiEnv &E) {
    int32_t best = -E.dead;
    uint64_t mv[4];
    uint32_t sc[4];
    ai_moves(b, E.rows, mv, sc);
    for (int i = 0; i < 4; ++i) {
        if (mv[i] == b) continue;
        const int32_t v = ai_evaluate(mv[i], E) + process_score(sc[i]);
        if (v > best) best = v;
    }
    return best;
}

template <int D>
__device__ int32_t ai_player(uint64_t b, const AiEnv &E);

// search_branch at depth D >= 1 (AIPlayer.cpp:703-791; no pruning at these
// depths): the mean over empty cells, in int64 with an arithmetic shift and
// a division that truncates, as the reference computes it
template <int D>
__device__ int32_t ai_branch(uint64_t t, const AiEnv &E) {
    uint64_t em = empty_nibbles(t);
    const int e = __popcll(em);
    int64_t acc = 0;
    while (em) {
        const int bit = __builtin_ctzll(em);
        em &= em - 1;
        const int32_t s4 = ai_player<D - 1>(t | (2ull << bit), E);
        const int32_t s2 = ai_player<D - 1>(t | (1ull << bit), E);
        acc += ((int64_t)s2 * E.w2 + (int64_t)s4 * E.w4) >> 16;
    }
    return e ? (int32_t)(acc / e) : 0;
}

// search_ai_player at depth D (AIPlayer.cpp:793-891), the task form every
// level of a search of depth <= 3 takes: best over valid moves of
// search_branch + process_score + 1
template <int D>
__device__ int32_t ai_player(uint64_t b, const AiEnv &E) {
    if constexpr (D <= 0) {
        return ai_search0(b, E);
    } else {
        int32_t best = -E.dead;
        uint64_t mv[4];
        uint32_t sc[4];
        ai_moves(b, E.rows, mv, sc);
        for (int i = 0; i < 4; ++i) {
            if (mv[i] == b) continue;
            const int32_t v = ai_branch<D>(mv[i], E) + process_score(sc[i]) + 1;
            if (v > best) best = v;
        }
        return best;
    }
}

// apply_dynamic_mask (AIPlayer.cpp:943-1063)
__device__ AiRoot dynamic_mask(uint64_t board, int prune, int do_check) {
    uint32_t counts[16];
    for (int i = 0; i < 16; ++i) counts[i] = 0;
    uint32_t bs = 0, small = 0;
    for (int k = 0; k < 16; ++k) {
        const uint32_t v = nib(board, k);
        counts[v]++;
        if (v) {
            bs += 1u << v;
            if (v < 9) small += 1u << v;
        }
    }
    uint32_t gt128 = 0, gt256 = 0;
    bool distinct = true;
    for (int i = 8; i < 16; ++i) {
        gt128 += counts[i];
        if (i >= 9) {
            gt256 += counts[i];
            if (counts[i] > 1 && i < 15) distinct = false;
        }
    }
    // large_tile_count(8, counts) (AIPlayer.cpp:129-142)
    int32_t large = 0;
    {
        int i = 8;
        while (i < 16 && counts[i] != 0) ++i;
        for (++i; i < 16; ++i) {
            if (counts[i] == 2 && i < 15) {
                large = 0;
                goto large_done;
            }
            large += counts[i];
        }
    }
large_done:
    (void)prune;  // the pruning threshold only acts at depths above 3
    AiRoot R;
    R.board_sum = bs;
    R.masked = board;
    R.fixed_mask = 0;
    R.masked_count = 0;
    R.variant = 0;
    const uint32_t m1024 = bs % 1024;
    if (gt128 <= 4) R.dead = 262144;
    else if (m1024 < 8 || m1024 > 1008) R.dead = 131072;
    else if (gt128 == 5) R.dead = 131072;
    else if (gt256 > 5 && distinct) {
        if (gt128 == 8 && (counts[7] == 1 || counts[6] >= 1)) R.dead = 4096;
        else if (gt256 >= 6 && 960 < m1024) R.dead = 32768;
        else R.dead = 48000;
    } else {
        R.dead = 65536;
    }
    if (!distinct) return R;
    const bool a1 = (m1024 >= 48 || (m1024 > 12 && small == m1024)) && m1024 <= 512 && counts[9] == 0;
    const bool a2 = m1024 >= 512 || m1024 < 6;
    const uint64_t mb = (a1 || a2) ? ai_mask(board, m1024 > 1000 ? 11 : 9) : ai_mask(board, 12);
    uint64_t x = mb & (mb >> 1);
    x &= x >> 2;
    const int32_t mc = __popcll(x & 0x1111111111111111ull);
    uint32_t min_masked = 0;
    if (mc > 0) {
        int32_t acc = 0;
        for (int i = 15; i >= 0; --i) {
            acc += counts[i];
            if (acc >= mc) {
                min_masked = (uint32_t)i;
                break;
            }
        }
    }
    uint32_t gt512 = 0;
    for (int i = 10; i < 16; ++i) gt512 += counts[i];
    const bool f1 = gt512 == 5 && (counts[9] == 1 || counts[8] > 0) && bs % 512 < 72;
    const bool f2 = ((gt512 == 6 || (gt512 == 5 && counts[9] == 1)) && m1024 < 72) ||
                    (gt512 == 5 && counts[9] == 0 && (counts[8] == 1 || counts[7] > 0) && bs % 256 < 72);
    const bool f3 = gt512 >= 4 && mc >= 5 && (bs % 256 > 132 && bs % 256 < 234);
    const bool f4 = gt512 == 4 && counts[9] == 1 && counts[10] == 1 && bs % 256 < 60;
    const bool corner = first_match(mb, c_m22, 4) != 0;
    R.variant = corner ? (int)f1 | ((int)f2 << 1) | ((int)f3 << 2) | ((int)f4 << 3) : 0;
    R.masked = mb;
    R.masked_count = mc;
    // is_5tiler (AIPlayer.cpp:144-150)
    uint32_t s11 = 0;
    for (int i = 11; i < 16; ++i) s11 += counts[i];
    const bool five = ((bs > 62000 && bs < 65520) || (s11 == 4 && counts[10] == 0)) && (m1024 < 24 || m1024 > 996);
    const uint32_t m256 = bs % 256;
    uint64_t fm = 0;
    if (mc >= 5 && counts[0] >= 3) {
        fm = first_match(mb, c_m22, 4) | first_match(mb, c_m21, 4);
    } else if ((min_masked == 12 && mc == 4) || (min_masked == 11 && mc == 5) || (min_masked == 10 && mc == 6)) {
        fm = first_match(mb, c_m23, 8) | first_match(mb, c_m22, 4);
    } else if (mc >= 5 && (large > 5 || (m256 > 48 && m256 < 224) || five)) {
        fm = first_match(mb, c_m23, 8) | first_match(mb, c_m22, 4);
    } else if (do_check || (prune == 0 && (large > 2 || large == 0)) ||
               ((m256 > 246 || m256 < 24) && ((large < 5 && large > 2) || large == 0)) || counts[7] > 1 ||
               (counts[6] > 1 && counts[7] == 1)) {
        fm = 0;
    } else if (mc >= 4) {
        fm = first_match(mb, c_m22, 4) | first_match(mb, c_m21, 4);
    } else if (mc == 3 && gt512 == 3 && m1024 > 24) {
        fm = first_match(mb, c_m21, 4) | first_match(mb, c_m20, 8);
    } else if (mc >= 2 && gt512 == 2 && bs % 512 > 16) {
        fm = first_match(mb, c_m20, 8);
    }
    R.fixed_mask = fm;
    return R;
}

// The whole search of one board on one thread (AIPlayer::start_search at
// depth D): the four top scores and the best operation (1..4, 0 = none).
template <int D>
__device__ int ai_start_search(uint64_t board, const AiTables &tab, int do_check, int prune, int64_t w4,
                               int32_t out[4]) {
    const AiRoot R = dynamic_mask(board, prune, do_check);
    AiEnv E;
    E.eval = tab.eval + (size_t)R.variant * 65536 * 2;
    E.rows = tab.rows;
    E.dead = R.dead;
    E.masked_count = R.masked_count;
    E.do_check = do_check;
    E.w4 = w4;
    E.w2 = 65536 - w4;
    int best = 0;
    for (int pass = 0; pass < 2; ++pass) {
        const uint64_t root = pass ? board : R.masked;
        uint64_t mv[4];
        uint32_t sc[4];
        ai_moves(root, E.rows, mv, sc);
        int32_t bv = -E.dead;
        for (int i = 0; i < 4; ++i) {
            out[i] = mv[i] == root ? -E.dead : ai_branch<D>(mv[i], E) + process_score(sc[i]) + 1;
            if (out[i] > bv) {
                bv = out[i];
                best = i + 1;
            }
        }
        // the search again on the unmasked board (AIPlayer.cpp:921-940)
        if (pass || !(R.fixed_mask && best > 0 &&
                      (bv < min(-2000, -(R.dead >> 2)) ||
                       (R.masked_count < 5 && bv < 480 * R.masked_count && R.board_sum % 512 < 160))))
            break;
        E.masked_count = 0;
    }
    return best;
}

// The tables of device `dev` (built on first use).
AiTables ai_tables(int dev);

}  // namespace egtb
