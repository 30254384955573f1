"""Batched builds: many (pattern, target) jobs through one layer pipeline
(egtb_batch_build).  Each job's solved layers equal its plain build's."""
from __future__ import annotations

import ctypes
import time
from typing import List, Sequence, Tuple

import numpy as np

from . import native
from .patterns import build_params

# The batched workload: 24 jobs of L3, 442 and t (10 free cells) at
# targets 64..1024 and the 12-free-cell 444-style patterns 444, LL and 4431
# at 32..128, plus 4432 / 4441 at 32 and 4442 at 16 -- layers from ~1e1 to
# ~1e7 states, 3.7e9 raw states in all.  The 12-free-cell patterns stop at
# 128: a batch keeps every job's raw layers and books in HBM at once (the
# 10-free-cell dozen peaks at 71 GB, 444@256 adds 34 GB), 444@512 alone has
# 1.9e9 states and 444@1024 4.2e9.
DEFAULT_JOBS: List[Tuple[str, int]] = [
    ("L3", 6), ("L3", 7), ("L3", 8), ("L3", 9), ("L3", 10),
    ("442", 6), ("442", 7), ("442", 8), ("442", 9), ("442", 10),
    ("t", 7), ("t", 8), ("t", 9), ("t", 10),
    ("444", 5), ("444", 6), ("444", 7),
    ("LL", 5), ("LL", 6),
    ("4431", 5), ("4431", 6),
    ("4432", 5), ("4441", 5), ("4442", 4),
]

# 24 jobs at targets 16..64 of every 10- and 12-free-cell pattern family:
# merged layers of ~1e7 states, past the key-set limit.
SMALL_JOBS: List[Tuple[str, int]] = [(p, t) for p in ("L3", "442", "t", "LL", "444", "4431")
                                     for t in (4, 5, 6, 6)]

# 24 tiny 10-free-cell jobs at targets 8..32 (layers of 1e3..1e5 states):
# the regime where a plain build is its fixed per-layer cost (L3@8: 52
# layers in ~3.7 ms) and batching shares that cost across all jobs.
TINY_JOBS: List[Tuple[str, int]] = [(p, t) for p in ("L3", "442", "t") for t in (3, 3, 4, 4, 5, 5, 3, 4)]


class BatchResult:
    def __init__(self, handle, jobs):
        self._h = ctypes.c_void_p(handle)
        self.jobs = list(jobs)

    def timing(self):
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        native.check(native.lib().egtb_batch_timing(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def book(self, job: int, layer: int):
        L = native.lib()
        n = ctypes.c_size_t()
        native.check(L.egtb_batch_book(self._h, job, layer, None, None, 0, ctypes.byref(n)))
        keys = np.empty(n.value, dtype=np.uint64)
        rates = np.empty(n.value, dtype=np.float32)
        native.check(L.egtb_batch_book(self._h, job, layer, keys.ctypes.data_as(native.u64p),
                                       rates.ctypes.data_as(native.f32p), keys.size, ctypes.byref(n)))
        return keys, rates

    def layer_totals(self, layer: int):
        a, b = ctypes.c_uint64(), ctypes.c_uint64()
        native.check(native.lib().egtb_batch_layer_info(self._h, layer, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def close(self):
        if self._h is not None and self._h.value:
            native.lib().egtb_batch_free(self._h)
        self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def compare_with_plain(jobs: Sequence[Tuple[str, int]], reps: int = 1) -> dict:
    """Wall time of the batched build vs the same jobs built one by one
    with the plain schedule (both with inputs resident, no file I/O), each
    the best of `reps` runs (a process's first large build also pays for
    the device memory it takes)."""
    from .BookBuilder import build_in_memory

    plain_ms = None
    states = 0
    for _ in range(reps):
        total = 0.0
        states = 0
        for p, t in jobs:
            t0 = time.perf_counter()
            r = build_in_memory(p, t)
            total += (time.perf_counter() - t0) * 1000.0
            states += sum(r.layer_info(i)[0] for i in range(r.steps))
            r.close()
        plain_ms = total if plain_ms is None else min(plain_ms, total)
    best = None
    for _ in range(reps):
        t0 = time.perf_counter()
        br = build_batch(jobs)
        ms = (time.perf_counter() - t0) * 1000.0
        br.close()
        best = ms if best is None else min(best, ms)
    return {"jobs": len(jobs), "states": states, "plain_sequential_ms": round(plain_ms, 2),
            "batched_ms": round(best, 2), "speedup": round(plain_ms / best, 3),
            "batched_states_per_s": round(states / (best / 1000.0), 1)}


def build_batch(jobs: Sequence[Tuple[str, int]] = DEFAULT_JOBS, spawn_rate4: float = 0.1, device: int = 0,
                radix_bits: int = 0, sort_schedule: int = 0, forward_hash_steps: int = 0) -> BatchResult:
    params = [build_params(p, t) for p, t in jobs]
    keep = []
    arr = (native.BatchJob * len(params))()
    for i, p in enumerate(params):
        seeds = np.ascontiguousarray(p["seeds"], dtype=np.uint64)
        masks = np.ascontiguousarray(p["masks"] or [0], dtype=np.uint64)
        shifts = np.ascontiguousarray(p["shifts"] or [0], dtype=np.uint8)
        keep += [seeds, masks, shifts]
        j = arr[i]
        j.seeds = seeds.ctypes.data_as(native.u64p)
        j.n_seeds = seeds.size
        j.pattern = native.Pattern(masks.ctypes.data_as(native.u64p), len(p["masks"]),
                                   shifts.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), len(p["shifts"]),
                                   p["symm_mode"])
        j.target = p["target"]
        j.steps = p["steps"]
        j.docheck_step = p["docheck_step"]
        j.spawn_rate4 = spawn_rate4
    o = native.BatchOptions()
    o.device = device
    o.radix_bits = radix_bits
    o.sort_schedule = sort_schedule
    o.forward_hash_steps = forward_hash_steps
    h = ctypes.c_void_p()
    native.check(native.lib().egtb_batch_build(arr, len(params), ctypes.byref(o), ctypes.byref(h)))
    return BatchResult(h.value, jobs)
