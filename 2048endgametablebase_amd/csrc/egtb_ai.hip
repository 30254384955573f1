// Heuristic expectimax player as a batched GPU search: the reference's
// AIPlayer (native_core/src/AIPlayer.cpp) at a fixed depth of 1 to 3, one
// wavefront per board.
//
// What it computes, per board b (AIPlayer::start_search, :893-941):
//   1. apply_dynamic_mask (:943-1063): tile counts and sum pick the dead
//      score, a masked board (tiles at or above 9, 11 or 12 become 0xF), the
//      masked count, the adjustment of 16 entries of the evaluation table
//      (update_specific_scores, :254-336) and a fixed mask;
//   2. search_ai_player(masked board, depth) (:793-891): for each valid move
//      (left, right, up, down), search_branch (:703-791) = the integer mean
//      over empty cells of (score after a 2 * w2 + score after a 4 * w4) >> 16,
//      plus process_score of the move's merge score plus 1; the move score
//      below depth 0 is search0 (:669-686): the best evaluate + process_score;
//   3. when the fixed mask is set and the best score is poor, the search again
//      on the unmasked board with masked count 0 (:921-940).
// evaluate (:645-667) sums a 65536-row table over the rows of the board and
// of its transpose; with do_check it subtracts check_corner (:470-643).
//
// Depths 1-3 never reach the reference's pruning (depth < max_d - 2) or its
// per-move depth cut (serial levels), so the search is the plain expectimax
// above.  The reference also caches values per board in a shared table with
// 11-bit signatures; up to depth 2 every board it caches is searched at one
// depth only, so its cache changes nothing but on a signature collision, and
// this search equals the reference's; at depth 3 a board reachable at two
// depths can take its deeper cached value in the reference, while this
// search evaluates each occurrence at its own depth.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "egtb.h"
#include "egtb_device.h"
#include "egtb_runtime.h"

namespace egtb {

// ---------------------------------------------------------------------------
// The evaluation table (init_evaluate_tables / diffs_evaluation_func,
// AIPlayer.cpp:40-127): row r (4 nibbles, nibble j = tile exponent of
// position j) maps each tile to a weight and scores the line of weights in
// nibble order (d1) and reversed (d2).
// ---------------------------------------------------------------------------

constexpr int32_t kTileWeight[16] = {0, 2, 4, 8, 16, 32, 64, 128, 248, 388, 488, 518, 519, 519, 519, 520};

static int32_t line_score(const int32_t w[4]) {
    // descending-run score
    int32_t dpdf = w[0];
    for (int x = 0; x < 3; ++x) {
        if (w[x] < w[x + 1]) {
            if (w[x] > 400) {
                dpdf += (w[x] << 1) + (w[x + 1] - w[x]) * x;
            } else if (w[x] > 300 && x == 1 && w[0] > w[1]) {
                dpdf += w[x] << 1;
            } else {
                dpdf -= (w[x + 1] - w[x]) << 3;
                dpdf -= w[x + 1] * 3;
                if (x < 2 && w[x + 2] < w[x + 1] && w[x + 1] > 30) dpdf -= std::max(80, w[x + 1]);
            }
        } else if (x < 2) {
            dpdf += w[x + 1] + w[x];
        } else {
            dpdf += (int32_t)((double)(w[x + 1] + w[x]) * 0.5);
        }
    }
    if (w[0] > 400 && w[1] > 300 && w[2] > 200 && w[2] > w[3] && w[3] < 300) dpdf += w[3] >> 2;
    // both-ends score
    int32_t t;
    const int32_t lo12 = std::min(w[1], w[2]), hi12 = std::max(w[1], w[2]);
    if (std::min(w[0], w[3]) < 32) {
        t = -16384;
    } else if ((w[0] < w[1] && w[0] < 400) || (w[3] < w[2] && w[3] < 400)) {
        t = -(hi12 * 10);
    } else {
        t = (int32_t)(((double)w[0] * 1.8 + (double)w[3] * 1.8) + (double)hi12 * 1.5 +
                      (double)std::min(160, lo12) * 2.5);
        if (lo12 < 8) t -= 60;
    }
    int zeros = 0;
    for (int k = 0; k < 4; ++k) zeros += w[k] == 0;
    const int32_t s123 = w[1] + w[2] + w[3];
    const int32_t penalty = (w[0] > 100 && ((zeros > 1 && s123 < 32) || s123 < 12)) ? 4 : 0;
    return std::max(dpdf, t) / 4 - penalty;
}

// out[2 r] = d1, out[2 r + 1] = d2 for every row r
static void base_table(int16_t *out) {
    for (int r = 0; r < 65536; ++r) {
        int32_t w[4], wr[4];
        for (int j = 0; j < 4; ++j) {
            w[j] = kTileWeight[(r >> (4 * j)) & 0xF];
            wr[3 - j] = w[j];
        }
        out[2 * r] = (int16_t)line_score(w);
        out[2 * r + 1] = (int16_t)line_score(wr);
    }
}

// The rows a search may adjust (update_specific_scores): d1 of `row` and
// d2 of its nibble mirror `mirror` get the line score of `row` plus a
// bonus chosen by the board's flags.
struct AdjustedRow {
    uint16_t row, mirror;
    int16_t bonus_a, bonus_b;  // base rows: bonus_a with flag3; dynamic rows: flag1 / flag2 (flag4: bonus_a >> 1)
    bool dynamic;
};
constexpr AdjustedRow kAdjusted[8] = {
    {0x78ff, 0xff87, 300, 0, false},  {0x7fff, 0xfff7, 320, 0, false},   {0x8fff, 0xfff8, 360, 0, false},
    {0x2fff, 0xfff2, 4, -4, true},    {0x3fff, 0xfff3, 12, -10, true},   {0x4fff, 0xfff4, 18, -16, true},
    {0x5fff, 0xfff5, 24, -20, true},  {0x6fff, 0xfff6, 25, -8, true},
};

// Table variant v (bit 0 flag1, 1 flag2, 2 flag3, 3 flag4, with the 2x2
// corner present; v = 0 is the unadjusted table).
static void table_variant(const int16_t *base, int v, int16_t *out) {
    std::memcpy(out, base, 65536 * 2 * sizeof(int16_t));
    const bool f1 = v & 1, f2 = v & 2, f3 = v & 4, f4 = v & 8;
    for (const AdjustedRow &a : kAdjusted) {
        int32_t w[4];
        for (int j = 0; j < 4; ++j) w[j] = kTileWeight[(a.row >> (4 * j)) & 0xF];
        int32_t s = line_score(w);
        if (!a.dynamic) {
            if (f3) s += a.bonus_a;
        } else if (f1) {
            s += a.bonus_a;
        } else if (f2) {
            s += a.bonus_b;
        } else if (f4) {
            s += a.bonus_a >> 1;
        }
        out[2 * a.row] = (int16_t)s;
        out[2 * a.mirror + 1] = (int16_t)s;
    }
}

// Row moves with merge scores (StandardMergePolicy::_internal_merge,
// BoardMover.h:7-55): low 16 bits the row slid toward its high nibble, high
// 16 bits the sum of merged tile values, at most 65535.  Sliding the other
// way merges the same pairs within each run of equal tiles, so the score of
// a move right equals the score of a move left of the same row, as the
// reference's table gives it (CommonMover.h:87-94,150-158).
static void row_table(uint32_t *out) {
    for (uint32_t r = 0; r < 65536; ++r) {
        uint32_t t[4], n = 0;
        for (int c = 0; c < 4; ++c) {
            const uint32_t v = (r >> (12 - 4 * c)) & 0xF;
            if (v) t[n++] = v;
        }
        uint32_t o[4] = {0, 0, 0, 0}, m = 0, score = 0;
        for (uint32_t i = 0; i < n; ++i) {
            if (i + 1 < n && t[i] == t[i + 1] && t[i] != 0xF) {
                o[m++] = t[i] + 1;
                score += 2u << t[i];
                ++i;
            } else {
                o[m++] = t[i];
            }
        }
        out[r] = ((o[0] << 12) | (o[1] << 8) | (o[2] << 4) | o[3]) | (std::min(score, 65535u) << 16);
    }
}

// ---------------------------------------------------------------------------
// Device side
// ---------------------------------------------------------------------------

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
__constant__ uint64_t c_m23[8] = {0xff00fff0ULL, 0xfff00ff00000000ULL, 0xf00ff00ffULL, 0xff00ff00f0000000ULL,
                                  0xfff0ff0000000000ULL, 0xff00ff000f0000ULL, 0xf000ff00ff00ULL, 0xff0fffULL};
__constant__ uint64_t c_m22[4] = {0xff00ff0000000000ULL, 0x00ff00ff00000000ULL, 0x00000000ff00ff00ULL,
                                  0x0000000000ff00ffULL};
__constant__ uint64_t c_m21[4] = {0xff00f00000000000ULL, 0x00ff000f00000000ULL, 0x00000000f000ff00ULL,
                                  0x00000000000f00ffULL};
__constant__ uint64_t c_m20[8] = {0xff00ULL, 0xff000000000000ULL, 0xf000fULL, 0xf000f00000000000ULL,
                                  0xff00000000000000ULL, 0xf000f00000000ULL, 0xf000f000ULL, 0xffULL};

// check_corner (AIPlayer.cpp:470-643): shapes of masked tiles by masked
// count, each group with its penalty, tried in the reference's order.
struct CornerGroup {
    int8_t mc_lo, mc_hi;  // masked counts the group applies to
    int16_t n;
    int32_t penalty;
    uint64_t m[8];
};
__constant__ CornerGroup c_corner[10] = {
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
__constant__ uint64_t c_m3_1000[8] = {0xff00fULL, 0xf000f00fULL, 0xf00000000000ff00ULL, 0xff0000000000f000ULL,
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

constexpr int kAiItems = 128;  // (root move, empty cell, spawn) of one board

// The four root move scores of search_ai_player(root, D) on one wavefront:
// lane work items are (root move i, empty cell, spawn value); lanes 0..3
// then form search_branch + process_score + 1 per valid root move.
template <int D>
__device__ void wave_search(uint64_t root, const AiEnv &E, int32_t *vals, int32_t out[4]) {
    const int lane = threadIdx.x & 63;
    uint64_t mv[4];
    uint32_t sc[4];
    ai_moves(root, E.rows, mv, sc);
    uint32_t ne[4], base[4], total = 0;
    for (int i = 0; i < 4; ++i) {
        ne[i] = mv[i] == root ? 0u : (uint32_t)__popcll(empty_nibbles(mv[i]));
        base[i] = total;
        total += 2 * ne[i];
    }
    for (uint32_t it = lane; it < total; it += 64) {
        int i = 0;
        while (i < 3 && it >= base[i + 1]) ++i;
        const uint32_t r = it - base[i];
        const uint32_t cell = r >> 1, spawn = (r & 1) ? 2u : 1u;
        uint64_t em = empty_nibbles(mv[i]);
        for (uint32_t c = 0; c < cell; ++c) em &= em - 1;
        const int bit = __builtin_ctzll(em);
        vals[it] = ai_player<D - 1>(mv[i] | ((uint64_t)spawn << bit), E);
    }
    __syncthreads();  // one wavefront per block
    if (lane < 4) {
        const int i = lane;
        int32_t v = -E.dead;
        if (ne[i]) {
            int64_t acc = 0;
            for (uint32_t c = 0; c < ne[i]; ++c) {
                const int32_t s2 = vals[base[i] + 2 * c], s4 = vals[base[i] + 2 * c + 1];
                acc += ((int64_t)s2 * E.w2 + (int64_t)s4 * E.w4) >> 16;
            }
            v = (int32_t)(acc / (int64_t)ne[i]) + process_score(sc[i]) + 1;
        }
        vals[kAiItems + i] = v;
    }
    __syncthreads();
    for (int i = 0; i < 4; ++i) out[i] = vals[kAiItems + i];
    __syncthreads();  // vals is reused by a second search
}

// One board per 64-thread block (one wavefront).
template <int D>
__global__ __launch_bounds__(64) void k_ai_search(const uint64_t *__restrict__ boards, uint64_t n, AiTables tab,
                                                  int do_check, int prune, int64_t w4, int32_t *__restrict__ best_out,
                                                  int32_t *__restrict__ scores_out) {
    __shared__ int32_t vals[kAiItems + 4];
    for (uint64_t q = blockIdx.x; q < n; q += gridDim.x) {
        const uint64_t board = boards[q];
        const AiRoot R = dynamic_mask(board, prune, do_check);
        AiEnv E;
        E.eval = tab.eval + (size_t)R.variant * 65536 * 2;
        E.rows = tab.rows;
        E.dead = R.dead;
        E.masked_count = R.masked_count;
        E.do_check = do_check;
        E.w4 = w4;
        E.w2 = 65536 - w4;
        int32_t s[4];
        wave_search<D>(R.masked, E, vals, s);
        int32_t best = 0, bv = -E.dead;
        for (int i = 0; i < 4; ++i)
            if (s[i] > bv) {
                bv = s[i];
                best = i + 1;
            }
        // the search again on the unmasked board (AIPlayer.cpp:921-940)
        if (R.fixed_mask && best > 0 &&
            (bv < min(-2000, -(R.dead >> 2)) ||
             (R.masked_count < 5 && bv < 480 * R.masked_count && R.board_sum % 512 < 160))) {
            E.masked_count = 0;
            wave_search<D>(board, E, vals, s);
            bv = -E.dead;
            for (int i = 0; i < 4; ++i)
                if (s[i] > bv) {
                    bv = s[i];
                    best = i + 1;
                }
        }
        if (threadIdx.x == 0) {
            best_out[q] = best;
            for (int i = 0; i < 4; ++i) scores_out[4 * q + i] = s[i];
        }
    }
}

// evaluate(board) of a player with the given masked count, table variant
// and do_check (one thread per board).
__global__ void k_ai_evaluate(const uint64_t *__restrict__ boards, uint64_t n, AiTables tab, int variant,
                              int do_check, int masked_count, int32_t *__restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    AiEnv E;
    E.eval = tab.eval + (size_t)variant * 65536 * 2;
    E.rows = tab.rows;
    E.dead = 0;
    E.masked_count = masked_count;
    E.do_check = do_check;
    E.w2 = E.w4 = 0;
    out[i] = ai_evaluate(boards[i], E);
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

namespace {

struct AiDevice {
    DBuf eval, rows;
};

std::mutex g_ai_mu;
std::map<int, std::unique_ptr<AiDevice>> g_ai;

AiDevice &ai_device(DeviceCtx &C) {
    std::lock_guard<std::mutex> lk(g_ai_mu);
    auto &slot = g_ai[C.dev];
    if (!slot) {
        auto d = std::make_unique<AiDevice>();
        std::vector<int16_t> base(65536 * 2), all((size_t)kAiVariants * 65536 * 2);
        base_table(base.data());
        for (int v = 0; v < kAiVariants; ++v) table_variant(base.data(), v, all.data() + (size_t)v * 65536 * 2);
        std::vector<uint32_t> rows(65536);
        row_table(rows.data());
        d->eval.alloc(all.size() * sizeof(int16_t), C.stream);
        d->rows.alloc(rows.size() * sizeof(uint32_t), C.stream);
        HIPCHK(hipMemcpyAsync(d->eval.p, all.data(), all.size() * sizeof(int16_t), hipMemcpyHostToDevice, C.stream));
        HIPCHK(hipMemcpyAsync(d->rows.p, rows.data(), rows.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                              C.stream));
        HIPCHK(hipStreamSynchronize(C.stream));
        slot = std::move(d);
    }
    return *slot;
}

}  // namespace

void ai_search(const uint64_t *boards, size_t n, int depth, int do_check, int prune, double spawn_rate4, int dev,
               int32_t *best, int32_t *scores) {
    if (depth < 1 || depth > 3) throw Error(EGTB_ERR_ARG, "depth must be 1, 2 or 3");
    if (!(spawn_rate4 >= 0.0 && spawn_rate4 <= 1.0)) throw Error(EGTB_ERR_ARG, "spawn_rate4 out of [0, 1]");
    if (n && (!boards || !best || !scores)) throw Error(EGTB_ERR_ARG, "null buffer");
    if (n == 0) return;
    DeviceCtx &C = device_ctx(dev);
    AiDevice &A = ai_device(C);
    hipStream_t s = C.stream;
    DBuf d_boards, d_best, d_scores;
    d_boards.alloc(n * 8, s);
    d_best.alloc(n * 4, s);
    d_scores.alloc(n * 16, s);
    HIPCHK(hipMemcpyAsync(d_boards.p, boards, n * 8, hipMemcpyHostToDevice, s));
    const AiTables tab{(const int16_t *)A.eval.p, (const uint32_t *)A.rows.p};
    // update_spawn_rate (AIPlayer.cpp:636-641): the weight of a 4 truncated to 1/65536
    const int64_t w4 = (int64_t)(spawn_rate4 * 65536.0);
    const uint32_t grid = (uint32_t)std::min<uint64_t>(n, 65536);
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(grid), dim3(64), 0, s, (const uint64_t *)d_boards.p, (uint64_t)n, tab, do_check,
                           prune, w4, (int32_t *)d_best.p, (int32_t *)d_scores.p);
    };
    if (depth == 1) go(k_ai_search<1>);
    else if (depth == 2) go(k_ai_search<2>);
    else go(k_ai_search<3>);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(best, d_best.p, n * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(scores, d_scores.p, n * 16, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
}

void ai_evaluate_batch(const uint64_t *boards, size_t n, int variant, int do_check, int masked_count, int dev,
                       int32_t *out) {
    if (variant < 0 || variant >= kAiVariants) throw Error(EGTB_ERR_ARG, "variant must be 0..15");
    if (n && (!boards || !out)) throw Error(EGTB_ERR_ARG, "null buffer");
    if (n == 0) return;
    DeviceCtx &C = device_ctx(dev);
    AiDevice &A = ai_device(C);
    hipStream_t s = C.stream;
    DBuf d_boards, d_out;
    d_boards.alloc(n * 8, s);
    d_out.alloc(n * 4, s);
    HIPCHK(hipMemcpyAsync(d_boards.p, boards, n * 8, hipMemcpyHostToDevice, s));
    const AiTables tab{(const int16_t *)A.eval.p, (const uint32_t *)A.rows.p};
    hipLaunchKernelGGL(k_ai_evaluate, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s,
                       (const uint64_t *)d_boards.p, (uint64_t)n, tab, variant, do_check, masked_count,
                       (int32_t *)d_out.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, d_out.p, n * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
}

}  // namespace egtb

using namespace egtb;

template <typename F>
static int guarded_ai(F &&f) {
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

int egtb_ai_table(int16_t *out, int variant) {
    if (!out || variant < 0 || variant >= kAiVariants) return EGTB_ERR_ARG;
    std::vector<int16_t> base(65536 * 2);
    base_table(base.data());
    table_variant(base.data(), variant, out);
    return EGTB_OK;
}

int egtb_ai_search(const uint64_t *boards, size_t n, int depth, int do_check, int prune, double spawn_rate4,
                   int device, int32_t *best, int32_t *scores) {
    return guarded_ai([&] { ai_search(boards, n, depth, do_check, prune, spawn_rate4, device, best, scores); });
}

int egtb_ai_evaluate(const uint64_t *boards, size_t n, int variant, int do_check, int masked_count, int device,
                     int32_t *out) {
    return guarded_ai([&] { ai_evaluate_batch(boards, n, variant, do_check, masked_count, device, out); });
}

}  // extern "C"
