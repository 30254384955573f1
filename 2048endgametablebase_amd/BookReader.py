"""Classic table reader: the reference's BookReader.move_on_dic surface
(engine_core/BookReader.py:65-99) answered from tables resident in HBM.

The reference binary-searches each .book file on disk per probe
(ReaderRuntime.cpp:279-316).  Here every layer of a build is loaded into
HBM once (egtb_table_open) and the whole move_on_dic evaluation -- symmetry
ops, moves, pattern checks, canonical lookups, direction remap -- runs in
one GPU kernel; Python only formats the result exactly as the reference's
binding does (bindings_formation.cpp:30-40).

Deviation (documented): the reference tries the symmetry op that answered
the previous call first (a per-reader ``last_operation_index_``,
ReaderRuntime.cpp:171-182,826); this reader is stateless and always tries
ops 0..7.  Results differ only for boards that match the pattern under more
than one op with different answers.
"""
from __future__ import annotations

import ctypes
import math
import os
import re
import threading
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import native
from .config import SingletonConfig
from .patterns import pattern_catalog

_ORDERED_KEYS = ("down", "right", "left", "up")   # kOrderedResultKeys
# readers that keep the best op's answer (ReaderRuntime.cpp:1053)
_PREFER_MAX_PATTERNS = ("4442ff", "4442f", "4tiler")
_ABS_NAMES = ("left", "right", "up", "down")     # lookup output order
_DIR_NAMES = ("up", "right", "down", "left")     # kDirectionNames


def adjust_direction(op: int, direction: str) -> str:
    """ReaderRuntime.cpp:149-169."""
    d = direction
    if op >= 4:
        d = {"left": "right", "right": "left"}.get(d, d)
    idx = {"up": 0, "right": 1, "down": 2, "left": 3}[d]
    return _DIR_NAMES[(idx + op % 4) % 4]


def round_half_away(x: float) -> float:
    """std::round: nearest integer, halves away from zero."""
    t = float(math.trunc(x))
    if abs(x - t) >= 0.5:
        t += math.copysign(1.0, x)
    return t


def maybe_round_value(value: float, dtype: str) -> float:
    """ReaderRuntime.cpp:208-214: 32-bit dtypes are rounded to 1e-9.  The
    reference's release build (-Ofast, CMakeLists.txt:79-95) compiles the
    division by 10^9 as a multiplication by the double nearest 1e-9
    (-freciprocal-math), so that is the value its reader returns: e.g.
    0.23655372900000002 where a division would give 0.236553729."""
    if "32" not in dtype or abs(value) <= 1e-7:
        return value
    return round_half_away(value * 1e9) * 1e-9


def blank_entries() -> Dict[str, object]:
    return {"down": "", "right": "", "left": "", "up": ""}


def encode_board(board) -> int:
    """encode_board_matrix (BoardCodec.h:20-41) of a 4x4 array of tile values."""
    b = 0
    a = np.asarray(board).reshape(4, 4)
    for r in range(4):
        for c in range(4):
            v = int(a[r, c])
            e = v.bit_length() - 1 if v > 0 else 0
            b |= (e & 0xF) << (4 * ((3 - r) * 4 + (3 - c)))
    return b


def format_result(values4, op: int, status: int, dtype: str) -> Tuple[Dict[str, object], str]:
    """Turn one lookup row into the reference's (dict, dtype) pair."""
    if status != 0:
        return blank_entries(), ""
    by_abs = {name: values4[i] for i, name in enumerate(_ABS_NAMES)}
    # the entry under key k of the transformed board lands on adjust(op, k);
    # the lookup already returns absolute directions
    entries = []
    for key in _ORDERED_KEYS:
        name = adjust_direction(op, key)
        v = by_abs[name]
        entries.append((name, None if np.isnan(v) else maybe_round_value(float(v), dtype)))
    numeric = [e for e in entries if e[1] is not None]
    others = [e for e in entries if e[1] is None]
    numeric.sort(key=lambda e: -e[1])  # stable, like the reference's small-range sort
    return dict(numeric + others), dtype


class Table:
    """One build's solved layers resident in HBM (egtb_table)."""

    def __init__(self, handle, dtype: str):
        self._h = ctypes.c_void_p(handle)
        self.dtype = dtype

    @classmethod
    def open(cls, prefix: str, pattern: str, target: int, dtype: str, steps: Optional[int] = None,
             spawn_rate4: float = 0.1, device: int = 0, prefer_max: Optional[bool] = None) -> "Table":
        meta = pattern_catalog[pattern]
        if steps is None:
            folder, base = os.path.split(prefix)
            rx = re.compile(re.escape(base) + r"(\d+)\.(?:book|z)$")
            found = [int(m.group(1)) for f in os.listdir(folder or ".") if (m := rx.match(f))]
            steps = max(found) + 1 if found else 0
        masks = np.asarray(meta.pattern_masks or (0,), dtype=np.uint64)
        shifts = np.asarray(meta.success_shifts or (0,), dtype=np.uint8)
        pat = native.Pattern(masks.ctypes.data_as(native.u64p), len(meta.pattern_masks),
                             shifts.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), len(meta.success_shifts),
                             meta.symm_mode)
        h = ctypes.c_void_p()
        native.check(native.lib().egtb_table_open(prefix.encode(), int(steps), ctypes.byref(pat), int(target),
                                                  int(meta.tile_sum), dtype.encode(), float(spawn_rate4),
                                                  int(device), ctypes.byref(h)))
        t = cls(h.value, dtype)
        if prefer_max is None:
            prefer_max = pattern in _PREFER_MAX_PATTERNS
        if prefer_max:
            t.set_prefer_max(True)
        return t

    @classmethod
    def from_result(cls, result, pattern: str, target: int, spawn_rate4: float = 0.1) -> "Table":
        meta = pattern_catalog[pattern]
        masks = np.asarray(meta.pattern_masks or (0,), dtype=np.uint64)
        shifts = np.asarray(meta.success_shifts or (0,), dtype=np.uint8)
        pat = native.Pattern(masks.ctypes.data_as(native.u64p), len(meta.pattern_masks),
                             shifts.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), len(meta.success_shifts),
                             meta.symm_mode)
        h = ctypes.c_void_p()
        native.check(native.lib().egtb_table_from_result(result.handle, ctypes.byref(pat), int(target),
                                                         int(meta.tile_sum), float(spawn_rate4), ctypes.byref(h)))
        t = cls(h.value, result.dtype)
        t._result = result  # keep the build alive
        return t

    def set_prefer_max(self, on: bool = True):
        """Answer from the symmetry op with the largest best rate
        (egtb_table_set_prefer_max)."""
        native.check(native.lib().egtb_table_set_prefer_max(self._h, int(bool(on))))

    @property
    def handle(self):
        return self._h

    def info(self):
        steps, recs, by = ctypes.c_int(), ctypes.c_uint64(), ctypes.c_uint64()
        native.check(native.lib().egtb_table_info(self._h, ctypes.byref(steps), ctypes.byref(recs),
                                                  ctypes.byref(by)))
        return steps.value, recs.value, by.value

    def lookup(self, boards) -> Tuple[np.ndarray, np.ndarray]:
        """float32 [n, 4] rates (left, right, up, down; NaN invalid), status [n]."""
        b = np.ascontiguousarray(np.asarray(boards, dtype=np.uint64).reshape(-1))
        out = np.empty((b.size, 4), dtype=np.float32)
        st = np.empty(b.size, dtype=np.int32)
        native.check(native.lib().egtb_lookup_batch(self._h, b.ctypes.data_as(native.u64p), b.size,
                                                    out.ctypes.data_as(native.f32p),
                                                    st.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return out, st

    def lookup_f64(self, boards):
        b = np.ascontiguousarray(np.asarray(boards, dtype=np.uint64).reshape(-1))
        out = np.empty((b.size, 4), dtype=np.float64)
        st = np.empty(b.size, dtype=np.int32)
        op = np.empty(b.size, dtype=np.int32)
        i32p = ctypes.POINTER(ctypes.c_int32)
        native.check(native.lib().egtb_lookup_batch_f64(self._h, b.ctypes.data_as(native.u64p), b.size,
                                                        out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                        st.ctypes.data_as(i32p), op.ctypes.data_as(i32p)))
        return out, st, op

    def move_on_dic_batch(self, boards) -> List[Tuple[Dict[str, object], str]]:
        vals, st, op = self.lookup_f64(boards)
        return [format_result(vals[i], int(op[i]), int(st[i]), self.dtype) for i in range(vals.shape[0])]

    def close(self):
        if self._h is not None and self._h.value:
            native.lib().egtb_table_free(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# record layout, raw-to-rate scale and the value of a missing key per dtype
# (FormationRuntime.h:197-200; ReaderRuntime.cpp:183-206,270-277: integer
# rates are divided by their max scale, the 1-float types store rate - 1 and
# read a missing key as -1)
_REC = {
    "float32": (np.dtype([("board", "<u8"), ("rate", "<f4")]), 1.0, 0.0),
    "float64": (np.dtype([("board", "<u8"), ("rate", "<f8")]), 1.0, 0.0),
    "uint32": (np.dtype([("board", "<u8"), ("rate", "<u4")]), 4e9, 0.0),
    "uint64": (np.dtype([("board", "<u8"), ("rate", "<u8")]), 1.6e18, 0.0),
    "1-float32": (np.dtype([("board", "<u8"), ("rate", "<f4")]), 1.0, -1.0),
    "1-float64": (np.dtype([("board", "<u8"), ("rate", "<f8")]), 1.0, -1.0),
}


def _record_layout(dtype: str):
    try:
        rec, scale, _ = _REC[dtype]
        return rec, scale
    except KeyError:
        raise ValueError(f"unsupported success_rate_dtype {dtype!r}") from None


def find_classic_value(pathname: str, filename: str, search_key: int, success_rate_dtype: str = "uint32"):
    """One on-disk probe of a .book file (formation_core.find_classic_value,
    bindings_formation.cpp:181-195 -> find_success_entry_in_file,
    ReaderRuntime.cpp:279-316).  None when the file cannot be opened; the
    key's rate when found (integer dtypes scaled to [0, 1]); 0.0 otherwise.
    The bisection is the reference's own (inclusive bounds, lower mid), so
    answers agree also on layer 0, which is stored in seed order."""
    rec, scale = _record_layout(success_rate_dtype)
    path = os.path.join(pathname, filename)
    try:
        size = os.path.getsize(path)
        if not os.path.isfile(path):
            return None
    except OSError:
        return None
    n = size // rec.itemsize
    if n == 0:
        return _REC[success_rate_dtype][2]
    zero = _REC[success_rate_dtype][2]
    recs = np.memmap(path, dtype=rec, mode="r", shape=(n,))
    key = int(search_key) & 0xFFFFFFFFFFFFFFFF
    left, right = 0, n - 1
    while left <= right:
        mid = (left + right) // 2
        board = int(recs[mid]["board"])
        if board == key:
            v = float(recs[mid]["rate"])
            return v / scale if scale > 1.0 else v
        if board < key:
            left = mid + 1
        else:
            if mid == 0:
                break
            right = mid - 1
    return zero


def spawn_random_tile(board: int, spawn_rate4: float, rng) -> int:
    """A 2 (or a 4 with probability spawn_rate4) on a uniformly chosen empty
    cell (gen_new_num, CommonMover.h:59-67); a full board is returned as is."""
    empty = [k for k in range(16) if (board >> (4 * k)) & 0xF == 0]
    if not empty:
        return board
    pos = empty[int(rng.integers(len(empty)))]
    val = 2 if rng.random() < spawn_rate4 else 1
    return board | (val << (4 * pos))


class BookReader:
    """engine_core/BookReader.py: move_on_dic / get_random_state."""

    _tables: Dict[Tuple[str, str, str], Table] = {}
    _lock = threading.Lock()

    @classmethod
    def _table(cls, folder: str, pattern_full: str, dtype: str) -> Optional[Table]:
        key = (folder, pattern_full, dtype)
        with cls._lock:
            t = cls._tables.get(key)
            if t is None:
                pattern, tile = pattern_full.split("_", 1)[0], pattern_full.split("_")[1]
                target = int(tile).bit_length() - 1
                cfg = SingletonConfig().config
                t = Table.open(os.path.join(folder, pattern_full + "_"), pattern, target, dtype,
                               spawn_rate4=float(cfg["4_spawn_rate"]), device=int(cfg.get("device", 0)))
                cls._tables[key] = t
            return t

    @classmethod
    def move_on_dic(cls, board, pattern: str, target: str, pattern_full: str):
        del target
        meta = pattern_catalog.get(pattern)
        cfg = SingletonConfig().config
        path_list = cfg["filepath_map"].get((pattern_full, cfg["4_spawn_rate"]), [])
        if meta is None:
            return {"?": "?"}, ""
        if not path_list:
            return {"?": "?"}, ""
        b = encode_board(board)
        v = int(np.asarray(board).sum()) - meta.tile_sum
        nums = abs(v) // 2 if v >= 0 else -(abs(v) // 2)  # C++ truncation toward zero
        if nums < 0:
            return blank_entries(), ""
        for folder, dtype in path_list:
            if not os.path.exists(folder):
                continue
            # the layer as a .book or in its compressed form (a .z directory;
            # the same check as ReaderRuntime.cpp:482-498,765)
            stem = os.path.join(folder, f"{pattern_full}_{nums}")
            if not os.path.exists(stem + ".book") and not os.path.isdir(stem + ".z"):
                continue
            res = cls._table(folder, pattern_full, dtype).move_on_dic_batch([b])[0]
            if res[1]:
                return res
        return blank_entries(), ""

    @classmethod
    def get_random_state(cls, path_list: list, pattern_full: str, rng=None) -> np.uint64:
        """A random stored state of one of the books 0..9 with one tile
        spawned on it (engine_core/BookReader.py:87-99 ->
        sample_classic_book_state, ReaderRuntime.cpp:944-990): per folder,
        books are tried in random order, a record is drawn uniformly, and
        the first non-zero board is returned after gen_new_num.  0 when no
        book qualifies or the pattern is unknown."""
        if pattern_full.split("_", 1)[0] not in pattern_catalog:
            return np.uint64(0)
        rng = rng if rng is not None else np.random.default_rng()
        p4 = float(SingletonConfig().config["4_spawn_rate"])
        for folder, dtype in path_list:
            rec, _ = _record_layout(dtype)
            books = list(range(10))
            while books:
                book_id = books.pop(int(rng.integers(len(books))))
                path = os.path.join(folder, f"{pattern_full}_{book_id}.book")
                if not os.path.isfile(path):
                    continue
                n = os.path.getsize(path) // rec.itemsize
                if n == 0:
                    continue
                r = int(rng.integers(n))
                state = int(np.memmap(path, dtype=rec, mode="r", shape=(n,))[r]["board"])
                if state != 0:
                    return np.uint64(spawn_random_tile(state, p4, rng))
        return np.uint64(0)

    @classmethod
    def clear(cls):
        with cls._lock:
            for t in cls._tables.values():
                t.close()
            cls._tables.clear()
