"""ctypes wrapper of the CPU oracle (oracle/egtb_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libegtb_oracle.so")
_lib = None

u64p = ctypes.POINTER(ctypes.c_uint64)
f32p = ctypes.POINTER(ctypes.c_float)


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE, "libegtb_oracle.so", "libegtb_ad_oracle.so"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_init.restype = None
        L.orc_row_left.restype = ctypes.c_uint16
        L.orc_row_left.argtypes = [ctypes.c_uint16]
        L.orc_row_right.restype = ctypes.c_uint16
        L.orc_row_right.argtypes = [ctypes.c_uint16]
        for name in ("orc_transpose", "orc_mirror_lr", "orc_mirror_ud"):
            getattr(L, name).restype = ctypes.c_uint64
            getattr(L, name).argtypes = [ctypes.c_uint64]
        L.orc_move.restype = ctypes.c_uint64
        L.orc_move.argtypes = [ctypes.c_uint64, ctypes.c_int]
        L.orc_canonical.restype = ctypes.c_uint64
        L.orc_canonical.argtypes = [ctypes.c_uint64, ctypes.c_int]
        L.orc_is_pattern.restype = ctypes.c_int
        L.orc_is_pattern.argtypes = [ctypes.c_uint64, u64p, ctypes.c_int]
        L.orc_is_success.restype = ctypes.c_int
        L.orc_is_success.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
        L.orc_sort_unique.restype = ctypes.c_size_t
        L.orc_sort_unique.argtypes = [u64p, ctypes.c_size_t, ctypes.c_int]
        L.orc_build.restype = ctypes.c_void_p
        L.orc_build.argtypes = [u64p, ctypes.c_size_t, u64p, ctypes.c_int, ctypes.c_char_p,
                                ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                ctypes.c_int]
        L.orc_free.restype = None
        L.orc_free.argtypes = [ctypes.c_void_p]
        L.orc_raw_layer.restype = ctypes.c_size_t
        L.orc_raw_layer.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(u64p)]
        L.orc_book_layer.restype = ctypes.c_size_t
        L.orc_book_layer.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(u64p),
                                     ctypes.POINTER(f32p)]
        L.orc_candidates.restype = None
        L.orc_candidates.argtypes = [ctypes.c_void_p, ctypes.c_int, u64p, u64p]
        L.orc_solve_layer.restype = None
        L.orc_solve_layer.argtypes = [u64p, ctypes.c_size_t, f32p, u64p, f32p, ctypes.c_size_t,
                                      u64p, f32p, ctypes.c_size_t, u64p, ctypes.c_int,
                                      ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.orc_init()
        _lib = L
    return _lib


def _u64(a):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.uint64))
    return a, a.ctypes.data_as(u64p)


def move(board: int, direction: int) -> int:
    return int(lib().orc_move(board, direction))


def canonical(board: int, mode: int) -> int:
    return int(lib().orc_canonical(board, mode))


def sort_unique(a, threads: int = 8) -> np.ndarray:
    a = np.array(a, dtype=np.uint64, copy=True)
    n = lib().orc_sort_unique(a.ctypes.data_as(u64p), a.size, threads)
    return a[:n]


class OracleBuild:
    """A complete in-memory build: raw layers, candidate counts, solved books."""

    def __init__(self, seeds, masks, shifts, symm, target, steps, docheck, p4=0.1,
                 fma_mode=1, solve=True, threads=8):
        L = lib()
        s, sp = _u64(seeds)
        m, mp = _u64(masks if len(masks) else [0])
        sh = bytes(bytearray(int(x) for x in shifts))
        self.steps = steps
        self._h = L.orc_build(sp, s.size, mp, len(masks), sh, len(shifts), symm, target, steps,
                              docheck, p4, fma_mode, int(solve), threads)
        if not self._h:
            raise ValueError("oracle build rejected its parameters")

    def raw(self, i: int) -> np.ndarray:
        p = u64p()
        n = lib().orc_raw_layer(self._h, i, ctypes.byref(p))
        if n == 0:
            return np.zeros(0, dtype=np.uint64)
        return np.ctypeslib.as_array(p, shape=(n,)).copy()

    def book(self, i: int):
        k = u64p()
        r = f32p()
        n = lib().orc_book_layer(self._h, i, ctypes.byref(k), ctypes.byref(r))
        if n == 0:
            return np.zeros(0, dtype=np.uint64), np.zeros(0, dtype=np.float32)
        return (np.ctypeslib.as_array(k, shape=(n,)).copy(),
                np.ctypeslib.as_array(r, shape=(n,)).copy())

    def candidates(self, i: int):
        a = ctypes.c_uint64()
        b = ctypes.c_uint64()
        lib().orc_candidates(self._h, i, ctypes.byref(a), ctypes.byref(b))
        return int(a.value), int(b.value)

    def close(self):
        if self._h:
            lib().orc_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve_layer(keys, book1, book2, masks, shifts, symm, target, do_check, p4=0.1,
                fma_mode=1, threads=8) -> np.ndarray:
    L = lib()
    k, kp = _u64(keys)
    out = np.zeros(k.size, dtype=np.float32)
    k1, k1p = _u64(book1[0])
    r1 = np.ascontiguousarray(book1[1], dtype=np.float32)
    k2, k2p = _u64(book2[0])
    r2 = np.ascontiguousarray(book2[1], dtype=np.float32)
    m, mp = _u64(masks if len(masks) else [0])
    sh = bytes(bytearray(int(x) for x in shifts))
    L.orc_solve_layer(kp, k.size, out.ctypes.data_as(f32p), k1p, r1.ctypes.data_as(f32p), k1.size,
                      k2p, r2.ctypes.data_as(f32p), k2.size, mp, len(masks), sh, len(shifts),
                      symm, target, int(do_check), p4, fma_mode, threads)
    return out


def read_book(path: str):
    """Read a reference-layout .book file: packed {u64 board; f32 rate}."""
    rec = np.dtype([("board", "<u8"), ("rate", "<f4")])
    a = np.fromfile(path, dtype=rec)
    return a["board"].copy(), a["rate"].copy()
