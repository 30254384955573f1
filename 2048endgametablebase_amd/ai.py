"""The dispatcher's heuristic expectimax player, searched on the GPU.

``search`` runs the reference's AIPlayer search (native_core/src/AIPlayer.cpp:
893-941) for many boards at once through ``egtb_ai_search``: one wavefront
per board, depths 1-3.  ``AIPlayer`` keeps the surface of the reference's
``ai_core.AIPlayer`` binding (native_core/src/bindings_ai.cpp:20-56) over it:
construct with a board, set ``do_check`` / ``prune`` / ``spawn_rate4``, call
``start_search(depth)``, read ``best_operation`` and ``scores``.

Up to depth 2 the results equal the reference's single-thread search; at
depth 3 each board is searched at its own depth where the reference may reuse
a deeper value from its cache.  Depths above 3 (the reference's pruned,
time-limited levels) are not offered.
"""
from __future__ import annotations

import ctypes
from typing import Tuple

import numpy as np

from . import native

MOVES = ("left", "right", "up", "down")
i32p = ctypes.POINTER(ctypes.c_int32)


def search(boards, depth: int = 2, do_check: bool = False, prune: bool = False, spawn_rate4: float = 0.1,
           device: int = 0) -> Tuple[np.ndarray, np.ndarray]:
    """best (n,) int32: 1..4 = left, right, up, down, 0 = no valid move;
    scores (n, 4) int32: the top scores, -dead_score for invalid moves."""
    b = np.ascontiguousarray(np.atleast_1d(np.asarray(boards, dtype=np.uint64)))
    best = np.zeros(b.size, dtype=np.int32)
    scores = np.zeros((b.size, 4), dtype=np.int32)
    native.check(native.lib().egtb_ai_search(b.ctypes.data_as(native.u64p), b.size, int(depth), int(bool(do_check)),
                                             int(bool(prune)), float(spawn_rate4), int(device),
                                             best.ctypes.data_as(i32p), scores.ctypes.data_as(i32p)))
    return best, scores


def evaluate(boards, variant: int = 0, do_check: bool = False, masked_count: int = 0, device: int = 0) -> np.ndarray:
    """AIPlayer::evaluate on the device (table variant 0 = unadjusted)."""
    b = np.ascontiguousarray(np.atleast_1d(np.asarray(boards, dtype=np.uint64)))
    out = np.zeros(b.size, dtype=np.int32)
    native.check(native.lib().egtb_ai_evaluate(b.ctypes.data_as(native.u64p), b.size, int(variant),
                                               int(bool(do_check)), int(masked_count), int(device),
                                               out.ctypes.data_as(i32p)))
    return out


def table(variant: int = 0) -> np.ndarray:
    """The evaluation table (65536, 2) int16 of (d1, d2), computed on the host."""
    out = np.zeros((65536, 2), dtype=np.int16)
    native.check(native.lib().egtb_ai_table(out.ctypes.data_as(ctypes.POINTER(ctypes.c_int16)), int(variant)))
    return out


class AIPlayer:
    """ai_core.AIPlayer's surface (bindings_ai.cpp:20-56) over the GPU search."""

    def __init__(self, board: int):
        self.board = int(board) & 0xFFFFFFFFFFFFFFFF
        self.max_d = 3
        self.best_operation = 0
        self.max_threads = 1  # accepted for parity; the device search is batched
        self.do_check = 0
        self.spawn_rate4 = 0.1
        self.prune = 0
        self.device = 0
        self._scores = (0, 0, 0, 0)

    def reset_board(self, new_board: int) -> None:
        self.best_operation = 0
        self.board = int(new_board) & 0xFFFFFFFFFFFFFFFF

    def clear_cache(self) -> None:
        """No value cache is kept between searches."""

    def update_spawn_rate(self, new_rate: float) -> None:
        self.spawn_rate4 = float(new_rate)

    def evaluate(self, board: int) -> int:
        return int(evaluate([board], do_check=bool(self.do_check), device=self.device)[0])

    def start_search(self, depth: int = 3) -> None:
        self.max_d = int(depth)
        best, scores = search([self.board], depth, bool(self.do_check), bool(self.prune), self.spawn_rate4,
                              self.device)
        self.best_operation = int(best[0])
        self._scores = tuple(int(x) for x in scores[0])

    @property
    def scores(self):
        return self._scores
