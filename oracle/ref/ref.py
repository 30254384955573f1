"""ctypes loader for oracle/_ref/libegtb_ref.so: the reference's own classic
builder compiled from its sources by oracle/ref/Makefile.

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
        _lib = L
    return _lib


def run_build(seeds, masks, shifts, symm, target, steps, docheck, pathname, is_free=False,
              p4=0.1, dtype="float32", threads=8) -> None:
    """Run the reference's run_pattern_build_cpp; raises on its exceptions."""
    s = np.ascontiguousarray(seeds, dtype=np.uint64)
    m = np.ascontiguousarray(masks if len(masks) else [0], dtype=np.uint64)
    sh = np.ascontiguousarray(shifts if len(shifts) else [0], dtype=np.uint8)
    # engine_core/BookBuilder.py:237-261: the capacity error is retried once,
    # resuming from the layer files already written.
    for attempt in range(2):
        err = ctypes.create_string_buffer(1024)
        rc = lib().egtb_ref_build(s.ctypes.data, s.size, m.ctypes.data, len(masks), sh.ctypes.data,
                                  len(shifts), symm, target, steps, docheck, pathname.encode(),
                                  int(is_free), p4, dtype.encode(), threads, err, 1024)
        if rc == 0:
            return
        msg = err.value.decode(errors="replace")
        if attempt or not msg.startswith("The length multiplier"):
            raise RuntimeError(msg)
