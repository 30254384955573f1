"""Multi-table dispatch for play: which resident table answers a position,
and the GPU expectimax player when none does.

``DispatcherCommon`` restates the reference's dispatcher
(engine_core/AIPlayer.py:34-291): the tables of the configuration's
``filepath_map`` are grouped by endgame level (32768 tiles of the pattern +
log2 target) and by the pattern's free 32768 tiles; for a position, the
endgame levels its large tiles allow are tried in three tiers
(``get_endgame_lvls``), each table is looked up on the position with the n
largest tiles masked to 32768 (``mask``), a table whose best move has a rate
above zero answers, and a table that is about to finish its target hands
over to the player for 20 moves (``check_table``'s cooldown).  ``dispatcher``
returns a direction name or "AI".  Lookups go through ``BookReader.move_on_dic``
(the reference's classic reader; AD tables are not part of this build).

``dispatch_batch`` makes the same decisions for many games at once: each
round looks every undecided game up in its next candidate table, one device
lookup per table (``Table.move_on_dic_batch`` over tables resident in HBM),
and the games left to the player are searched together by ``ai.search``.
"""
from __future__ import annotations

import os
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import ai
from .BookReader import BookReader, encode_board
from .config import SingletonConfig
from .patterns import pattern_catalog

# the zero of each rate type (Config.py:41-78): a looked-up rate plus this is
# the success rate the dispatcher compares
ZERO_OF_DTYPE = {"uint32": 0, "uint64": 0, "float32": 0.0, "float64": 0.0, "1-float32": -1.0, "1-float64": -1.0}

# the board a position gets when its small tiles sum below 24: no table
# holds it (engine_core/AIPlayer.py:63-74)
UNFINDABLE = np.array([[32768, 32768, 32768, 32768], [0, 32768, 32768, 0], [0, 32768, 32768, 0],
                       [32768, 32768, 32768, 32768]], dtype=np.int32)

COOLDOWN_STEPS = 20


def pattern_32k_tiles(name: str) -> Tuple[int, int, Tuple[int, ...]]:
    """(32768 tiles of the first seed, of those the ones no mask pins, the
    nibble offsets every mask pins) -- Config.py:250-256."""
    meta = pattern_catalog[name]
    seed = meta.seed_boards[0]
    count = sum(1 for k in range(16) if (seed >> (4 * k)) & 0xF == 0xF)
    inter = 0xFFFFFFFFFFFFFFFF
    for m in meta.pattern_masks:
        inter &= m
    if not meta.pattern_masks:
        inter = 0
    fixed = tuple(4 * k for k in range(16) if (inter >> (4 * k)) & 0xF == 0xF)
    return count, count - len(fixed), fixed


def has_pattern_files(table: str, path_list) -> bool:
    """Config.check_pattern_file (Config.py:588-620): some folder of the
    table holds a file named <table>_... ending in .book, .z or b."""
    prefix = f"{table}_"
    for path, _dtype in path_list or []:
        if not path or not os.path.isdir(path):
            continue
        for item in os.listdir(path):
            if item.startswith(prefix) and (item.endswith(".book") or item.endswith(".z") or item.endswith("b")):
                return True
    return False


def decode(board_encoded: int) -> np.ndarray:
    """4x4 tile values of an encoded board (row 0 = the top 16 bits)."""
    m = np.zeros((4, 4), dtype=np.int32)
    for r in range(4):
        for c in range(4):
            e = (board_encoded >> (4 * ((3 - r) * 4 + (3 - c)))) & 0xF
            m[r, c] = 0 if e == 0 else 1 << e
    return m


class DispatcherCommon:
    """engine_core/AIPlayer.py:34-291 (BaseDispatcher + DispatcherCommon)."""

    def __init__(self, board, board_encoded):
        self.board_encoded = int(board_encoded) & 0xFFFFFFFFFFFFFFFF
        self.board = np.asarray(board)
        self.counts = self.frequency_count()
        self.last_operator = 0
        self.current_table = "AI"
        self.readers: Dict[Tuple[int, int], list] = {}
        self._table_cooldowns: Dict[str, int] = {}
        self.init_bookreader()

    def reset(self, board, board_encoded):
        self.board_encoded = int(board_encoded) & 0xFFFFFFFFFFFFFFFF
        self.board = np.asarray(board)
        self.counts = self.frequency_count()
        self._table_cooldowns.clear()

    def frequency_count(self) -> np.ndarray:
        counts = np.zeros(16, dtype=np.uint32)
        for i in range(16):
            counts[(self.board_encoded >> (4 * i)) & 0xF] += 1
        return counts

    def mask(self, n: int) -> np.ndarray:
        """The n largest tiles set to 32768 (np.argpartition's choice among
        equal tiles, as the reference makes it)."""
        flat = self.board.flatten().copy()
        idx = np.argpartition(flat, -n)[-n:]
        flat[idx] = 32768
        masked = flat.reshape(4, 4)
        if masked.sum() - n * 32768 < 24:
            masked = UNFINDABLE.copy()
        return masked

    def init_bookreader(self):
        cfg = SingletonConfig().config
        rate = float(cfg["4_spawn_rate"])
        for i, (table, spawn_rate4) in enumerate(cfg["filepath_map"].keys()):
            if abs(rate - float(spawn_rate4)) >= 0.01:
                continue
            if not has_pattern_files(table, cfg["filepath_map"][(table, spawn_rate4)]):
                continue
            parts = table.split("_")
            if len(parts) > 2:
                continue
            pattern, target_str = parts[0], parts[1]
            if pattern not in pattern_catalog:
                continue
            target = int(np.log2(int(target_str)))
            c32k, free32k, _fixed = pattern_32k_tiles(pattern)
            lvl = c32k + target
            self.readers.setdefault((lvl, c32k), []).append(
                (lvl, c32k, free32k, pattern, target, target_str, table, i + 1))
        for key in self.readers:
            self.readers[key].sort(key=lambda x: x[2], reverse=True)

    def _decide(self, table_param, table_type: int, r1, dtype) -> Optional[str]:
        """check_table's decision from the lookup's (dict, dtype)."""
        _lvl, _c32k, _free, _pattern, target, _ts, table, i = table_param
        zero = ZERO_OF_DTYPE.get(dtype, 0)
        r1 = {k: (v + zero if isinstance(v, (int, float, np.integer, np.floating)) else v) for k, v in r1.items()}
        move = list(r1.keys())[0]
        rate = r1[move]
        if isinstance(rate, (float, np.floating)):
            target_val = 1 << target
            rem = (int(np.sum(self.board[self.board < target_val])) if table_type == 1
                   else int(np.sum(self.board)) % target_val)
            if ((table_type == 1 and rate > 0.9999999 and rem < 24)
                    or (table_type == 2 and (rem < 32 or int((self.board == target_val // 2).sum()) > 1))
                    or (table_type == 3 and rate > 0.9999999 and (rem > (1 << target) - 4 or rem < 24))):
                self._table_cooldowns[table] = COOLDOWN_STEPS
                self.last_operator = 0
                self.current_table = "AI"
                return "AI"
            if rate > 0:
                self.last_operator = i
                self.current_table = table
                return move
        return None

    def check_table(self, table_param, table_type: int) -> Optional[str]:
        table = table_param[6]
        if table in self._table_cooldowns:
            return None
        masked = self.mask(table_param[1])
        r1, dtype = BookReader.move_on_dic(masked, table_param[3], table_param[5], table)
        return self._decide(table_param, table_type, r1, dtype)

    def is_unfree_endgame(self, i: int) -> bool:
        for j in range(i + 1, 14):
            if self.counts[j] == 0:
                return int(sum(self.counts[j:])) >= 4
        return False

    def get_endgame_lvls(self):
        lv1, lv2, lv3 = [], [], []
        large = 0
        current_large = int(np.sum(self.counts[9:]))
        ntiler_after = 1
        for tt in range(15, 6, -1):
            if self.counts[tt] > 1 and tt != 15:
                break
            if self.counts[tt] == 0:
                ntiler_after = int(np.sum(self.counts[tt + 1:])) + 1
            large += int(self.counts[tt])
            lvl = large + tt
            if lvl < 12 or (lvl == 12 and tt < 8):
                continue
            if self.counts[tt] > 0:
                if tt <= 12 and self.is_unfree_endgame(tt):
                    readers = self.readers.get((lvl, large), [])
                    lv3.extend([r for r in readers if (r[2] > 4 and current_large > 4) or ntiler_after > 4])
                    lv1.extend([r for r in readers if (r[2] <= 4 or current_large < 4) and ntiler_after <= 4])
                else:
                    lv1.extend(self.readers.get((lvl, large), []))
                if self.counts[tt - 1] < 2:
                    (lv2 if ntiler_after < 5 else lv3).extend(self.readers.get((lvl + 1, large), []))
                    if lvl > 12:
                        lv3.extend([r for r in self.readers.get((lvl + 2, large), []) if r[2] <= 4])
            elif self.counts[tt] == 0:
                lv3.extend(self.readers.get((lvl, large), []))
        lv1.sort(key=lambda x: x[2], reverse=True)
        lv2.sort(key=lambda x: 16 - x[1], reverse=True)
        lv3.sort(key=lambda x: 16 - x[1], reverse=True)
        return lv1, lv2, lv3

    def tick_cooldowns(self):
        expired = [t for t, c in self._table_cooldowns.items() if c <= 1]
        for t in expired:
            del self._table_cooldowns[t]
        for t in self._table_cooldowns:
            self._table_cooldowns[t] -= 1

    def candidates(self) -> List[Tuple[tuple, int]]:
        lv = self.get_endgame_lvls()
        return [(tp, ty) for tables, ty in zip(lv, (1, 2, 3)) for tp in tables]

    def dispatcher(self) -> str:
        self.tick_cooldowns()
        for tp, ty in self.candidates():
            result = self.check_table(tp, ty)
            if result == "AI":
                return "AI"
            if result:
                return result
        self.last_operator = 0
        self.current_table = "AI"
        return "AI"


def dispatch_batch(games: Sequence[DispatcherCommon], tables: Dict[str, object]) -> List[str]:
    """``dispatcher()`` for every game, with the lookups of one round grouped
    per table into one device call.  tables: table name -> resident
    ``Table`` (tables missing here are skipped, as a table whose files are
    gone would answer nothing)."""
    out: List[Optional[str]] = [None] * len(games)
    cands = []
    for g in games:
        g.tick_cooldowns()
        cands.append(g.candidates())
    pos = [0] * len(games)
    live = list(range(len(games)))
    while live:
        groups: Dict[str, List[int]] = {}
        nxt = []
        for gi in live:
            g = games[gi]
            while pos[gi] < len(cands[gi]):
                tp, _ty = cands[gi][pos[gi]]
                if tp[6] in g._table_cooldowns or tp[6] not in tables:
                    pos[gi] += 1
                    continue
                groups.setdefault(tp[6], []).append(gi)
                break
            else:
                g.last_operator = 0
                g.current_table = "AI"
                out[gi] = "AI"
        for name, members in groups.items():
            boards = [encode_board(games[gi].mask(cands[gi][pos[gi]][0][1])) for gi in members]
            results = tables[name].move_on_dic_batch(boards)
            for gi, (r1, dtype) in zip(members, results):
                tp, ty = cands[gi][pos[gi]]
                d = games[gi]._decide(tp, ty, r1, dtype) if dtype else None
                if d is not None:
                    out[gi] = d
                else:
                    pos[gi] += 1
                    nxt.append(gi)
        live = nxt
    return [o if o is not None else "AI" for o in out]


def best_moves(games: Sequence[DispatcherCommon], tables: Dict[str, object], depth: int = 2,
               device: int = 0) -> List[str]:
    """The move of every game: the table's answer, or the GPU player's best
    move (one batched search) where the dispatcher says "AI"; "" when the
    player finds no valid move."""
    moves = dispatch_batch(games, tables)
    ai_idx = [i for i, m in enumerate(moves) if m == "AI"]
    if ai_idx:
        cfg = SingletonConfig().config
        best, _scores = ai.search([games[i].board_encoded for i in ai_idx], depth,
                                  spawn_rate4=float(cfg["4_spawn_rate"]), device=device)
        for i, b in zip(ai_idx, best):
            moves[i] = ai.MOVES[b - 1] if b > 0 else ""
    return moves
