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
#include "egtb_ai.h"
#include "egtb_device.h"
#include "egtb_runtime.h"

namespace egtb {

// ---------------------------------------------------------------------------
// The evaluation table (init_evaluate_tables / diffs_evaluation_func,
// AIPlayer.cpp:40-127): row r (4 nibbles, nibble j = tile exponent of
// position j) maps each tile to a weight and scores the line of weights in
// nibble order (d1) and reversed (d2).
// ---------------------------------------------------------------------------

static constexpr int32_t kTileWeight[16] = {0, 2, 4, 8, 16, 32, 64, 128, 248, 388, 488, 518, 519, 519, 519, 520};

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

AiTables ai_tables(int dev) {
    DeviceCtx &C = device_ctx(dev);
    AiDevice &A = ai_device(C);
    return AiTables{(const int16_t *)A.eval.p, (const uint32_t *)A.rows.p};
}

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
