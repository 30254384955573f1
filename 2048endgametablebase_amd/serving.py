"""Serving: continuous batching of position lookups over tables resident in
HBM (the C++ request loop of egtb_server_*), plus the synthetic query load
of the serving benchmark."""
from __future__ import annotations

import ctypes
from typing import Dict, Tuple

import numpy as np

from . import native
from .patterns import pattern_catalog

i32p = ctypes.POINTER(ctypes.c_int32)


class Server:
    def __init__(self, slots: int = 4, max_batch: int = 4096, memory_budget: int = 0, device: int = 0):
        o = native.ServerOptions()
        o.device = device
        o.slots = slots
        o.max_batch = max_batch
        o.memory_budget = memory_budget
        h = ctypes.c_void_p()
        native.check(native.lib().egtb_server_create(ctypes.byref(o), ctypes.byref(h)))
        self._h = h
        self._tables = {}

    def attach(self, table) -> int:
        tid = ctypes.c_int32()
        native.check(native.lib().egtb_server_attach(self._h, table.handle, ctypes.byref(tid)))
        self._tables[tid.value] = table
        return tid.value

    def detach(self, table_id: int) -> None:
        native.check(native.lib().egtb_server_detach(self._h, table_id))
        self._tables.pop(table_id, None)

    def submit(self, table_id: int, boards) -> np.ndarray:
        b = np.ascontiguousarray(np.asarray(boards, dtype=np.uint64).reshape(-1))
        t = np.empty(b.size, dtype=np.uint64)
        native.check(native.lib().egtb_server_submit(self._h, table_id, b.ctypes.data_as(native.u64p), b.size,
                                                     t.ctypes.data_as(native.u64p)))
        return t

    def poll(self, max_n: int = 4096, wait_us: int = 1000):
        t = np.empty(max_n, dtype=np.uint64)
        o = np.empty((max_n, 4), dtype=np.float32)
        s = np.empty(max_n, dtype=np.int32)
        n = ctypes.c_size_t()
        native.check(native.lib().egtb_server_poll(self._h, max_n, t.ctypes.data_as(native.u64p),
                                                   o.ctypes.data_as(native.f32p), s.ctypes.data_as(i32p),
                                                   ctypes.byref(n), wait_us))
        k = n.value
        return t[:k], o[:k], s[:k]

    def collect(self, tickets, timeout_polls: int = 20000) -> Dict[int, Tuple[np.ndarray, int]]:
        want = set(int(x) for x in tickets)
        got = {}
        polls = 0
        while want - got.keys() and polls < timeout_polls:
            t, o, s = self.poll()
            for i in range(t.size):
                got[int(t[i])] = (o[i].copy(), int(s[i]))
            polls += 1
        return got

    def stats(self):
        b, q, m = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        native.check(native.lib().egtb_server_stats(self._h, ctypes.byref(b), ctypes.byref(q), ctypes.byref(m)))
        return {"batches": b.value, "queries": q.value, "max_batch": m.value}

    def loadgen(self, table_id: int, pool, qps: float, duration_s: float, seed: int = 0,
                n_games: int = 4096) -> dict:
        p = np.ascontiguousarray(pool, dtype=np.uint64)
        o = native.LoadgenOptions()
        o.qps = qps
        o.duration_s = duration_s
        o.seed = seed
        o.n_games = n_games
        o.table_id = table_id
        o.pool = p.ctypes.data_as(native.u64p)
        o.pool_size = p.size
        r = native.LoadgenReport()
        native.check(native.lib().egtb_server_loadgen(self._h, ctypes.byref(o), ctypes.byref(r)))
        return {name: getattr(r, name) for name, _ in native.LoadgenReport._fields_}

    def close(self):
        if self._h is not None and self._h.value:
            native.lib().egtb_server_destroy(self._h)
        self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def spawn_random(boards: np.ndarray, rng, p4: float = 0.1) -> np.ndarray:
    """Put a 2 (prob 1-p4) or a 4 on a uniformly chosen empty cell of each
    board (the position the player then moves from; gen_new_num,
    CommonMover.h:59-67).  Boards without an empty cell are unchanged."""
    b = np.ascontiguousarray(boards, dtype=np.uint64)
    shifts = (np.arange(16, dtype=np.uint64) * np.uint64(4))
    nib = (b[:, None] >> shifts[None, :]) & np.uint64(0xF)
    empty = nib == 0
    cnt = empty.sum(axis=1)
    r = np.floor(rng.random(b.size) * np.maximum(cnt, 1)).astype(np.int64)
    rank = np.cumsum(empty, axis=1) - 1
    pick = np.argmax(empty & (rank == r[:, None]), axis=1)
    val = np.where(rng.random(b.size) < p4, 2, 1).astype(np.uint64)
    spawned = b | (val << (pick.astype(np.uint64) * np.uint64(4)))
    return np.where(cnt > 0, spawned, b)


def make_query_pool(result, pattern: str, n: int = 1 << 18, off_table_frac: float = 0.2, seed: int = 0,
                    p4: float = 0.1) -> np.ndarray:
    """Serving queries: (1 - off_table_frac) are table states drawn by a
    seeded uniform index over every live record of the build, with a random
    spawn applied; the rest are such boards with one pinned cell set to a
    random small tile, so no symmetry of them is in the pattern and they
    take the fallback (none for a pattern without pinned cells)."""
    rng = np.random.default_rng(seed)
    sizes = np.array([result.layer_info(i)[1] for i in range(result.steps)], dtype=np.int64)
    cum = np.concatenate([[0], np.cumsum(sizes)])
    idx = np.sort(rng.integers(0, cum[-1], n))
    layer = np.searchsorted(cum, idx, side="right") - 1
    boards = np.empty(n, dtype=np.uint64)
    for L in np.unique(layer):
        keys, _ = result.book(int(L))
        sel = np.nonzero(layer == L)[0]
        boards[sel] = keys[idx[sel] - cum[L]]
    rng.shuffle(boards)
    boards = spawn_random(boards, rng, p4)
    meta = pattern_catalog[pattern]
    pinned = np.array(meta.fixed_positions, dtype=np.uint64)
    # a pattern without pinned cells (free*) has no off-pattern boards
    n_off = int(round(off_table_frac * n)) if pinned.size else 0
    sh = pinned[rng.integers(0, pinned.size, n_off)]
    v = rng.integers(1, 10, n_off).astype(np.uint64)
    boards[:n_off] = (boards[:n_off] & ~(np.uint64(0xF) << sh)) | (v << sh)
    perm = rng.permutation(n)
    return boards[perm]
