"""ctypes wrapper of oracle/libegtb_ad_oracle.so, the CPU restatement of the
reference's advanced (AD, large-tile masking) forward pass.

TEST INFRASTRUCTURE ONLY: loaded by tests/ as the checker, never by the
product path.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libegtb_ad_oracle.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-s", "-C", _HERE], check=True)
        L = ctypes.CDLL(LIB_PATH)
        P, I = ctypes.c_void_p, ctypes.c_int
        L.orc_ad_forward.restype = I
        L.orc_ad_forward.argtypes = [P, ctypes.c_size_t, P, I, I, I, P, I, I, I, I, ctypes.c_char_p, P]
        L.orc_ad_row_left.restype = ctypes.c_uint16
        L.orc_ad_row_left.argtypes = [ctypes.c_uint16]
        L.orc_ad_move4.restype = None
        L.orc_ad_move4.argtypes = [ctypes.c_uint64, P, P]
        L.orc_ad_comb_table.restype = None
        L.orc_ad_comb_table.argtypes = [P]
        L.orc_ad_derive.restype = I
        L.orc_ad_derive.argtypes = [ctypes.c_uint64, ctypes.c_uint32, P, I, I, I, P, I, I, I, P, P]
        L.orc_ad_validate.restype = I
        L.orc_ad_validate.argtypes = [ctypes.c_uint64, ctypes.c_uint32, P, I, I, I, P, I, I, I]
        L.orc_ad_build.restype = I
        L.orc_ad_build.argtypes = [P, ctypes.c_size_t, P, I, I, I, P, I, I, I, I, ctypes.c_char_p,
                                   ctypes.c_double, I]
        _lib = L
    return _lib


def _spec(p):
    masks = np.ascontiguousarray(p["masks"] if len(p["masks"]) else [0], dtype=np.uint64)
    fixed = np.ascontiguousarray(p["fixed_32k_shifts"] if len(p["fixed_32k_shifts"]) else [0], dtype=np.uint8)
    return masks, len(p["masks"]), fixed, len(p["fixed_32k_shifts"])


def forward(p, pathname: str, steps: int | None = None) -> np.ndarray:
    """Raw layers <pathname>0 .. <pathname>steps-2 of the advanced forward
    pass for the parameters of patterns.ad_params; returns the layer sizes.
    A smaller `steps` than p["steps"] gives the same first layers."""
    steps = int(p["steps"] if steps is None else steps)
    seeds = np.ascontiguousarray(p["seeds"], dtype=np.uint64)
    masks, nm, fixed, nf = _spec(p)
    sizes = np.zeros(max(steps, 1), dtype=np.uint64)
    rc = lib().orc_ad_forward(seeds.ctypes.data, seeds.size, masks.ctypes.data, nm, int(p["symm_mode"]),
                              int(p["n_free_32k"]), fixed.ctypes.data, nf, int(p["small_tile_sum_limit"]),
                              int(p["target"]), steps, pathname.encode(), sizes.ctypes.data)
    if rc != 0:
        raise OSError(f"oracle AD forward pass could not write under {pathname}")
    return sizes[:max(steps - 1, 0)]


RATE_TYPES = {"float32": 0, "float64": 1, "uint32": 2}


def build(p, pathname: str, p4: float = 0.1, dtype: str = "float32") -> None:
    """The whole advanced build (forward pass, then the solver of layers
    steps-3 .. 0) as bucket files <pathname><i>b/<key>.i|.b, rates of dtype
    float32, float64 or uint32."""
    seeds = np.ascontiguousarray(p["seeds"], dtype=np.uint64)
    masks, nm, fixed, nf = _spec(p)
    rc = lib().orc_ad_build(seeds.ctypes.data, seeds.size, masks.ctypes.data, nm, int(p["symm_mode"]),
                            int(p["n_free_32k"]), fixed.ctypes.data, nf, int(p["small_tile_sum_limit"]),
                            int(p["target"]), int(p["steps"]), pathname.encode(), float(p4),
                            RATE_TYPES[dtype])
    if rc != 0:
        raise OSError(f"oracle AD build could not read or write under {pathname}")


def row_left(row: int) -> int:
    return int(lib().orc_ad_row_left(row))


def move4(board: int):
    out = np.zeros(4, dtype=np.uint64)
    mnt = np.zeros(4, dtype=np.int32)
    lib().orc_ad_move4(board, out.ctypes.data, mnt.ctypes.data)
    return [int(x) for x in out], [bool(x) for x in mnt]


def comb_table() -> np.ndarray:
    """[2560, 11] uint8: length, then the tiles (log2, ascending)."""
    out = np.zeros((2560, 11), dtype=np.uint8)
    lib().orc_ad_comb_table(out.ctypes.data)
    return out


def derive(board: int, board_sum: int, p):
    """(-1 invalid | 0 kept as is | 1 derived, boards)"""
    masks, nm, fixed, nf = _spec(p)
    out = np.zeros(128, dtype=np.uint64)
    m = ctypes.c_int32()
    r = lib().orc_ad_derive(board, board_sum, masks.ctypes.data, nm, int(p["symm_mode"]), int(p["n_free_32k"]),
                            fixed.ctypes.data, nf, int(p["small_tile_sum_limit"]), int(p["target"]),
                            out.ctypes.data, ctypes.byref(m))
    return r, [int(x) for x in out[:m.value]]


def validate(board: int, board_sum: int, p) -> bool:
    masks, nm, fixed, nf = _spec(p)
    return bool(lib().orc_ad_validate(board, board_sum, masks.ctypes.data, nm, int(p["symm_mode"]),
                                      int(p["n_free_32k"]), fixed.ctypes.data, nf,
                                      int(p["small_tile_sum_limit"]), int(p["target"])))
