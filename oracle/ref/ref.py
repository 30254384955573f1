"""ctypes loader for oracle/_ref/libegtb_ref.so: the reference's own classic
builder, classic reader, advanced (AD) builder and expectimax player compiled
from their sources by oracle/ref/Makefile.

TEST INFRASTRUCTURE ONLY (golden-vector generation, CPU baseline timing).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.abspath(os.path.join(_HERE, "..", "_ref", "libegtb_ref.so"))
REFERENCE_SRC = "/root/reference/native_core"


def available() -> bool:
    return os.path.exists(LIB_PATH)


def build() -> bool:
    """Compile the reference library when its sources are present."""
    if not os.path.isdir(REFERENCE_SRC):
        return False
    subprocess.run(["make", "-s", "-C", _HERE, "-j8"], check=True)
    return True


_lib = None


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(LIB_PATH)
        f = L.egtb_ref_build
        f.restype = ctypes.c_int
        f.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
                      ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                      ctypes.c_int, ctypes.c_char_p, ctypes.c_int, ctypes.c_double,
                      ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t]
        f = L.egtb_ref_build_ex
        f.restype = ctypes.c_int
        f.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
                      ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                      ctypes.c_int, ctypes.c_char_p, ctypes.c_int, ctypes.c_double,
                      ctypes.c_char_p, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                      ctypes.c_char_p, ctypes.c_size_t]
        res = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p,
               ctypes.c_char_p, ctypes.c_char_p, ctypes.c_size_t]
        f = L.egtb_ref_move_on_dic
        f.restype = ctypes.c_int
        f.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int64,
                      ctypes.c_char_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int] + res
        L.egtb_ref_reader_new.restype = ctypes.c_void_p
        L.egtb_ref_reader_new.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        f = L.egtb_ref_reader_move_on_dic
        f.restype = ctypes.c_int
        f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                      ctypes.c_int64] + res
        L.egtb_ref_reader_free.restype = None
        L.egtb_ref_reader_free.argtypes = [ctypes.c_void_p]
        L.egtb_ref_trie_compress.restype = ctypes.c_int
        L.egtb_ref_trie_compress.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
        L.egtb_ref_trie_search.restype = ctypes.c_int
        L.egtb_ref_trie_search.argtypes = [ctypes.c_char_p, ctypes.c_uint64, ctypes.c_char_p,
                                           ctypes.POINTER(ctypes.c_double)]
        f = L.egtb_ref_find_classic_value
        f.restype = ctypes.c_int
        f.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_uint64, ctypes.c_char_p,
                      ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32)]
        L.egtb_ref_ai_table.restype = None
        L.egtb_ref_ai_table.argtypes = [ctypes.c_void_p]
        L.egtb_ref_ai_table_adjusted.restype = None
        L.egtb_ref_ai_table_adjusted.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                 ctypes.c_int, ctypes.c_void_p]
        L.egtb_ref_ai_evaluate.restype = ctypes.c_int32
        L.egtb_ref_ai_evaluate.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int]
        f = L.egtb_ref_build_ad
        f.restype = ctypes.c_int
        P, Z, I = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
        f.argtypes = [P, Z, P, Z, I, I, P, Z, I, I, I, I, ctypes.c_char_p, I, ctypes.c_double, ctypes.c_char_p,
                      I, I, ctypes.c_double, ctypes.c_char_p, Z]
        L.egtb_ref_ai_search.restype = None
        L.egtb_ref_ai_search.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                         ctypes.POINTER(ctypes.c_int32), ctypes.c_void_p,
                                         ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                         ctypes.POINTER(ctypes.c_uint64)]
        _lib = L
    return _lib


def run_build(seeds, masks, shifts, symm, target, steps, docheck, pathname, is_free=False,
              p4=0.1, dtype="float32", threads=8, deletion_threshold=0.0,
              optimal_branch_only=False, compress=False) -> None:
    """Run the reference's run_pattern_build_cpp; raises on its exceptions."""
    s = np.ascontiguousarray(seeds, dtype=np.uint64)
    m = np.ascontiguousarray(masks if len(masks) else [0], dtype=np.uint64)
    sh = np.ascontiguousarray(shifts if len(shifts) else [0], dtype=np.uint8)
    # engine_core/BookBuilder.py:237-261: the capacity error is retried once,
    # resuming from the layer files already written.
    for attempt in range(2):
        err = ctypes.create_string_buffer(1024)
        rc = lib().egtb_ref_build_ex(s.ctypes.data, s.size, m.ctypes.data, len(masks), sh.ctypes.data,
                                     len(shifts), symm, target, steps, docheck, pathname.encode(),
                                     int(is_free), p4, dtype.encode(), threads,
                                     float(deletion_threshold), int(bool(optimal_branch_only)),
                                     int(bool(compress)), err, 1024)
        if rc == 0:
            return
        msg = err.value.decode(errors="replace")
        if attempt or not msg.startswith("The length multiplier"):
            raise RuntimeError(msg)


def run_build_ad(seeds, masks, symm, n_free_32k, fixed_shifts, target, steps, docheck, pathname,
                 is_free=False, small_tile_sum_limit=96, p4=0.1, dtype="float32", threads=8,
                 stop_after_forward=True, deletion_threshold=0.0) -> None:
    """Run the reference's advanced builder run_pattern_build_ad_cpp; with
    stop_after_forward its raw layer files <pathname>0 .. <pathname>steps-2
    are left in place and the solver does not run."""
    s = np.ascontiguousarray(seeds, dtype=np.uint64)
    m = np.ascontiguousarray(masks if len(masks) else [0], dtype=np.uint64)
    fx = np.ascontiguousarray(fixed_shifts if len(fixed_shifts) else [0], dtype=np.uint8)
    for attempt in range(2):
        err = ctypes.create_string_buffer(1024)
        rc = lib().egtb_ref_build_ad(s.ctypes.data, s.size, m.ctypes.data, len(masks), symm, n_free_32k,
                                     fx.ctypes.data, len(fixed_shifts), small_tile_sum_limit, target, steps,
                                     docheck, pathname.encode(), int(is_free), p4, dtype.encode(), threads,
                                     int(bool(stop_after_forward)), float(deletion_threshold), err, 1024)
        if rc == 0:
            return
        msg = err.value.decode(errors="replace")
        if attempt or not msg.startswith("The length multiplier"):
            raise RuntimeError(msg)


def _result(keys, kinds, values, texts, n, dt):
    out = []
    for i in range(n.value):
        k = keys.raw[16 * i:16 * i + 16].split(b"\0", 1)[0].decode()
        kind = kinds[i]
        if kind == 1:
            v = float(values[i])
        elif kind == 2:
            v = texts.raw[16 * i:16 * i + 16].split(b"\0", 1)[0].decode()
        else:
            v = None
        out.append((k, v))
    return dict(out), dt.value.decode()


def _buffers():
    return (ctypes.create_string_buffer(64), (ctypes.c_int32 * 4)(), (ctypes.c_double * 4)(),
            ctypes.create_string_buffer(64), ctypes.c_int32(), ctypes.create_string_buffer(32),
            ctypes.create_string_buffer(512))


def _board16(board):
    b = np.ascontiguousarray(np.asarray(board, dtype=np.int32).reshape(16))
    return b


def move_on_dic(board, folder, dtype, pattern_full, nums_adjust, name, masks, symm):
    """ClassicBookReader(spec).move_on_dic(board, [(folder, dtype)], pattern_full,
    nums_adjust) with a fresh reader (symmetry ops tried 0..7).  Returns the
    reference binding's (dict, dtype) pair: numbers as float, "" strings,
    None for invalid directions."""
    b = _board16(board)
    m = np.ascontiguousarray(masks if len(masks) else [0], dtype=np.uint64)
    keys, kinds, values, texts, n, dt, err = _buffers()
    rc = lib().egtb_ref_move_on_dic(b.ctypes.data, folder.encode() if folder else b"", dtype.encode(),
                                    pattern_full.encode(), int(nums_adjust), name.encode(), m.ctypes.data,
                                    len(masks), int(symm), keys, kinds, values, texts, ctypes.byref(n), dt,
                                    err, 512)
    if rc:
        raise RuntimeError(err.value.decode(errors="replace"))
    return _result(keys, kinds, values, texts, n, dt)


class Reader:
    """One ClassicBookReader kept across calls (the reference's stateful
    symmetry order: the op that answered last is tried first)."""

    def __init__(self, name, masks, symm):
        m = np.ascontiguousarray(masks if len(masks) else [0], dtype=np.uint64)
        self._h = lib().egtb_ref_reader_new(name.encode(), m.ctypes.data, len(masks), int(symm))
        if not self._h:
            raise RuntimeError("reader construction failed")

    def move_on_dic(self, board, folder, dtype, pattern_full, nums_adjust):
        b = _board16(board)
        keys, kinds, values, texts, n, dt, err = _buffers()
        rc = lib().egtb_ref_reader_move_on_dic(self._h, b.ctypes.data, folder.encode(), dtype.encode(),
                                               pattern_full.encode(), int(nums_adjust), keys, kinds, values,
                                               texts, ctypes.byref(n), dt, err, 512)
        if rc:
            raise RuntimeError(err.value.decode(errors="replace"))
        return _result(keys, kinds, values, texts, n, dt)

    def close(self):
        if self._h:
            lib().egtb_ref_reader_free(self._h)
            self._h = None


def find_classic_value(pathname, filename, key, dtype):
    """find_classic_value_native: (value, found)."""
    v = ctypes.c_double()
    f = ctypes.c_int32()
    if lib().egtb_ref_find_classic_value(pathname.encode(), filename.encode(), int(key), dtype.encode(),
                                         ctypes.byref(v), ctypes.byref(f)):
        raise RuntimeError("find_classic_value_native raised")
    return v.value, bool(f.value)


def trie_compress(book_path, dtype):
    """trie_compress_progress_native: writes <book without .book>.z/ beside it."""
    rc = lib().egtb_ref_trie_compress(book_path.encode(), dtype.encode())
    if rc:
        raise RuntimeError(f"trie_compress_progress_native failed ({rc})")


def trie_search(path_prefix, board, dtype):
    """trie_decompress_search_native: the normalised rate, or None."""
    v = ctypes.c_double()
    rc = lib().egtb_ref_trie_search(path_prefix.encode(), int(board), dtype.encode(), ctypes.byref(v))
    if rc < 0:
        raise RuntimeError("trie_decompress_search_native raised")
    return v.value if rc else None


def ai_table():
    """The reference's evaluation table (_diffs_merged after
    init_evaluate_tables): int16 array [65536, 2] of (d1, d2)."""
    out = np.zeros((65536, 2), dtype=np.int16)
    lib().egtb_ref_ai_table(out.ctypes.data)
    return out


def ai_table_adjusted(board, f1, f2, f3, f4):
    """The table as update_specific_scores(board, f1, f2, f3, f4) leaves it."""
    out = np.zeros((65536, 2), dtype=np.int16)
    lib().egtb_ref_ai_table_adjusted(int(board), int(f1), int(f2), int(f3), int(f4), out.ctypes.data)
    return out


def ai_evaluate(board, do_check=False, masked_count=0):
    """AIPlayer::evaluate(board) with the unadjusted table."""
    return int(lib().egtb_ref_ai_evaluate(int(board), int(bool(do_check)), int(masked_count)))


def ai_search(board, depth, do_check=False, prune=False, spawn_rate4=0.1):
    """AIPlayer(board).start_search(depth) on one thread: (best operation
    1..4 or 0, [4 top scores], dead_score, masked_count, fixed_mask)."""
    best, dead, mc, fm = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_uint64()
    scores = np.zeros(4, dtype=np.int32)
    lib().egtb_ref_ai_search(int(board), int(depth), int(bool(do_check)), int(bool(prune)), float(spawn_rate4),
                             ctypes.byref(best), scores.ctypes.data, ctypes.byref(dead), ctypes.byref(mc),
                             ctypes.byref(fm))
    return int(best.value), [int(x) for x in scores], int(dead.value), int(mc.value), int(fm.value)
