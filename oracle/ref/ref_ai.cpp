// Test infrastructure only: a C-ABI shim around the reference's own
// expectimax player (AIPlayer, native_core/src/AIPlayer.cpp:645-1063),
// compiled from the reference's sources by oracle/ref/Makefile into
// oracle/_ref/libegtb_ref.so.  Loaded by tests/golden/make_ai_golden.py to
// write search fixtures and by the CPU tests that check the evaluation
// table, never by the product.
//
// The reference keeps its evaluation table (_diffs_merged) in a global that
// every search's apply_dynamic_mask may rewrite at 16 entries
// (update_specific_scores, AIPlayer.cpp:254-336); a search whose board has
// repeated tiles above 256 skips that rewrite and inherits the previous
// search's entries.  egtb_ref_ai_search restores the unadjusted entries
// before each search (update_specific_scores with every flag false, which
// rewrites them to the values init_evaluate_tables gave them), so each
// fixture depends on its own board only.

#include <cstdint>
#include <cstring>

#include "AIPlayer.h"

// defined in AIPlayer.cpp with external linkage, declared there only
void update_specific_scores(uint64_t board, bool flag1, bool flag2, bool flag3, bool flag4);

extern "C" {

// The evaluation table as init_evaluate_tables left it: out[2*i] = d1,
// out[2*i+1] = d2 for row i (65536 rows).
void egtb_ref_ai_table(int16_t *out) {
    update_specific_scores(0, false, false, false, false);
    for (int i = 0; i < 65536; ++i) {
        out[2 * i] = _diffs_merged[i].d1;
        out[2 * i + 1] = _diffs_merged[i].d2;
    }
}

// The table after update_specific_scores(board, flag1..flag4): the
// adjustments apply when `board` holds a 2x2 corner of masked tiles.  The
// unadjusted entries are restored afterwards.
void egtb_ref_ai_table_adjusted(uint64_t board, int f1, int f2, int f3, int f4, int16_t *out) {
    update_specific_scores(board, f1 != 0, f2 != 0, f3 != 0, f4 != 0);
    for (int i = 0; i < 65536; ++i) {
        out[2 * i] = _diffs_merged[i].d1;
        out[2 * i + 1] = _diffs_merged[i].d2;
    }
    update_specific_scores(0, false, false, false, false);
}

// AIPlayer::evaluate(board) of a player whose masked_count is given (it
// selects check_corner's masks) with do_check on or off, on the unadjusted
// table.
int32_t egtb_ref_ai_evaluate(uint64_t board, int do_check, int masked_count) {
    update_specific_scores(0, false, false, false, false);
    AIPlayer p(board);
    p.do_check = (uint8_t)(do_check != 0);
    p.masked_count = (uint8_t)masked_count;
    return p.evaluate(board);
}

// One AIPlayer(board).start_search(depth) on one thread (max_threads 1, as
// the reference itself uses below depth 5).  Outputs: the best operation
// (1..4 = left, right, up, down; 0 none), the four top scores, and the
// dead score, masked count and fixed mask apply_dynamic_mask chose.
void egtb_ref_ai_search(uint64_t board, int depth, int do_check, int prune, double spawn_rate4, int32_t *best,
                        int32_t *scores, int32_t *dead_score, int32_t *masked_count, uint64_t *fixed_mask) {
    update_specific_scores(0, false, false, false, false);
    AIPlayer p(board);
    p.max_threads = 1;
    p.do_check = (uint8_t)(do_check != 0);
    p.prune = (uint8_t)(prune != 0);
    p.update_spawn_rate(spawn_rate4);
    p.start_search(depth);
    *best = p.best_operation;
    for (int i = 0; i < 4; ++i) scores[i] = p.top_scores[i];
    *dead_score = p.dead_score;
    *masked_count = p.masked_count;
    *fixed_mask = p.fixed_mask;
}

}  // extern "C"
