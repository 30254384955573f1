"""Generate expectimax-player fixtures from the reference's own AIPlayer.

Runs oracle/_ref/libegtb_ref.so (the reference's AIPlayer.cpp compiled from
its sources by oracle/ref/Makefile, in this container only) and writes
tests/golden/ai_search.json:

  * ``evaluate``: boards with AIPlayer::evaluate (AIPlayer.cpp:645-667) on
    the unadjusted table, for do_check on with masked counts 0, 3, 4 and 6
    (check_corner's shapes, :470-643) and for do_check off;
  * ``table_sha256``: SHA-256 of the evaluation table (65536 x {d1, d2}
    int16, little endian) as init_evaluate_tables leaves it and as
    update_specific_scores leaves it for each of the 15 flag combinations
    with the 2x2 corner shape present (variants 1..15, bit 0 = flag1);
  * ``search``: AIPlayer(board).start_search(depth) on one thread
    (:893-941) -- best operation, the four top scores, dead score, masked
    count, fixed mask -- at depths 1, 2 and 3 with do_check / prune off and
    on and spawn rates 0.1 and 0.2.

Positions are seeded endgame-like boards: a few distinct large tiles (256 and
up, sometimes a repeated one or two 32768s), small tiles and empty cells, so
that the dead score, the masking thresholds, the table adjustments, the
fixed mask and the second search on the unmasked board all come up; the
evaluate boards also put 0xF tiles on every corner shape check_corner tests.
Edge positions: a full board without a move, the empty board, single tiles,
a row of 32768s, a board of 2s.

Each reference call resets the evaluation table first, so each case depends
on its own board only.

Usage:  python tests/golden/make_ai_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, ROOT)

from oracle.ref import ref  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "ai_search.json")

# check_corner's shapes (AIPlayer.cpp:473-640), as boards to put 0xF tiles on
CORNER_SHAPES = [
    0xf000fff0ff, 0xff0fff000f000000, 0xff0fff0000000f, 0xf0000000fff0ff00, 0xf0ff00ff00f00000, 0xf00000fff00ff,
    0xff00fff00000f000, 0xf00ff00ff0f, 0xf0000000ff0fff00, 0xfff0ff0000000f, 0xf0000000fff0ff, 0xff0fff0000000f00,
    0xff00ff0f0000f000, 0xf0ff00ff000000f0, 0xf000000ff00ff0f, 0xf0000f0ff00ff, 0xf00000fff0ff, 0xff0fff00000f0000,
    0xff00ff00000f0f, 0xf0f00000ff00ff00, 0xf0ff00fff0000000, 0xf0f000000ff00ff, 0xff00ff000000f0f0, 0xfff00ff0f,
    0xff0fff0f, 0xf0fff0ff00000000, 0xff000000ff00ff, 0xff00ff000000ff00, 0xff0fff0f00000000, 0xff00ff000000ff,
    0xff000000ff00ff00, 0xf0fff0ff, 0xffff0ff, 0xff0ffff000000000, 0xff00ff00f0000f, 0xf0000f00ff00ff00,
    0xf0ff0fff00000000, 0xf00f000ff00ff, 0xff00ff000f00f000, 0xfff0ff0f, 0xff00f00f, 0xff00f00000000f, 0xfff00f,
    0xff000f000000f000, 0xf00000000f00ff00, 0xf00f00ff00000000, 0xf00fff0000000000, 0xf000000f000ff,
    0xf000f000f00f, 0xf000ff00f, 0xfff000000000f000, 0xf00000000000fff0, 0xf00f000f000f0000, 0xf00ff000f0000000,
    0xf000000000fff, 0xfff00000000000f, 0xff000000f000f000, 0xf000f000000ff, 0xf000f0ff, 0xff0f000f00000000,
    0xf000f0000000ff00, 0xf0fff00000000000, 0xfff0f, 0xff0000000f000f, 0xf00000000f00f, 0xf00f00000000000f,
    0xf00000000000f00f, 0xf00f00000000f000, 0xf0000000000f00f, 0xf000000000f00f, 0xf0000000000ff000,
    0xf000000f0000f000, 0xf00f0000000000f0, 0xf00f000000000f00, 0xf0000f000000f, 0xff0000000000f, 0xff00f, 0xf000f00f,
    0xf00000000000ff00, 0xff0000000000f000, 0xf00ff00000000000, 0xf00f000f00000000, 0xff00000000000f, 0xf0000000000ff,
]


def endgame_board(rng):
    """A board with 3-8 large tiles (mostly distinct), small tiles and gaps."""
    cells = rng.permutation(16)
    n_large = int(rng.integers(3, 9))
    large = list(rng.choice(np.arange(8, 16), size=n_large, replace=False))
    if rng.random() < 0.2:  # a repeated large tile
        large[-1] = large[0] if large[0] < 15 else 14
    if rng.random() < 0.1:  # two 32768s
        large[:2] = [15, 15]
    b = 0
    for k, v in zip(cells[:n_large], large):
        b |= int(v) << (4 * int(k))
    n_small = int(rng.integers(0, 16 - n_large + 1))
    for k in cells[n_large:n_large + n_small]:
        b |= int(rng.integers(1, 8)) << (4 * int(k))
    return b


def corner_board(rng, shape):
    b = shape
    for k in range(16):
        if (shape >> (4 * k)) & 0xF:
            continue
        r = rng.random()
        if r < 0.5:
            b |= int(rng.integers(1, 9)) << (4 * k)
        elif r < 0.6:
            b |= 0xF << (4 * k)
    return b


def main():
    rng = np.random.default_rng(2048)
    evaluate = []
    boards = [corner_board(rng, s) for s in CORNER_SHAPES for _ in range(3)]
    boards += [endgame_board(rng) for _ in range(200)]
    for b in boards:
        row = [f"{b:016x}", ref.ai_evaluate(b, False, 0)]
        row += [ref.ai_evaluate(b, True, mc) for mc in (0, 3, 4, 6)]
        evaluate.append(row)
    search = []
    settings = [(1, 0, 0, 0.1, 300), (2, 0, 0, 0.1, 300), (2, 1, 0, 0.1, 150), (2, 0, 1, 0.1, 100),
                (2, 0, 0, 0.2, 100), (1, 1, 1, 0.1, 100), (3, 0, 0, 0.1, 80)]
    for depth, do_check, prune, p4, count in settings:
        for _ in range(count):
            b = endgame_board(rng)
            best, scores, dead, mc, fm = ref.ai_search(b, depth, do_check, prune, p4)
            search.append([f"{b:016x}", depth, do_check, prune, p4, best, scores, dead, mc, f"{fm:016x}"])
    # a full board without a move, the empty board, single tiles, a row of
    # 32768s, a board of 2s
    for b in (0x123456789ABC1234, 0, 0x1, 0x2, 0xFFFF000000000000, 0x1111111111111111):
        for depth in (1, 2):
            best, scores, dead, mc, fm = ref.ai_search(b, depth, 0, 0, 0.1)
            search.append([f"{b:016x}", depth, 0, 0, 0.1, best, scores, dead, mc, f"{fm:016x}"])
    tables = [hashlib.sha256(ref.ai_table().astype("<i2").tobytes()).hexdigest()]
    for v in range(1, 16):
        t = ref.ai_table_adjusted(0xff00ff0000000000, v & 1, (v >> 1) & 1, (v >> 2) & 1, (v >> 3) & 1)
        tables.append(hashlib.sha256(t.astype("<i2").tobytes()).hexdigest())
    with open(OUT, "w") as fh:
        json.dump({"source": "reference AIPlayer (native_core/src/AIPlayer.cpp) via oracle/_ref, one thread",
                   "table_sha256": tables,
                   "evaluate_columns": ["board", "no check", "check mc 0", "check mc 3", "check mc 4",
                                        "check mc 6"],
                   "evaluate": evaluate,
                   "search_columns": ["board", "depth", "do_check", "prune", "spawn_rate4", "best", "scores",
                                      "dead_score", "masked_count", "fixed_mask"],
                   "search": search}, fh, separators=(",", ":"))
    print(f"wrote {OUT}: {len(evaluate)} evaluate rows, {len(search)} searches")


if __name__ == "__main__":
    main()
