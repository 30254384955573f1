"""Batched builds: many (pattern, target) jobs through one layer pipeline
(egtb_batch_build), or as plan_batch groups them (merged groups of small
jobs, large jobs alone).  Each job's solved layers equal its plain
build's."""
from __future__ import annotations

import ctypes
import json
import os
import time
from typing import Dict, List, Sequence, Tuple

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


# Per-job sizes of plain builds (peak raw layer, total states), written by
# the autotuner (autotune.tune_job_sizes) for the catalogue jobs above.
JOB_SIZES_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning", "job_sizes.json")
# The peaks of a merged group's jobs (reached at different layers) may sum
# to this many keys. Merging pays only while the group's key sets stay small
# enough to live in the caches, as each job's own do: measured on the 24-job
# mixes, a 1 Mi cap keeps the planned batch at 1.00x / 1.04x the one-by-one
# time, 2 Mi gave 0.96x / 1.05x, 4 Mi 0.90x / 0.98x and 12 Mi (merged sets up
# to the key-set limit) 0.86x / 0.85x.
MAX_GROUP_PEAK = 1 << 20


def job_sizes() -> Dict[str, dict]:
    try:
        with open(JOB_SIZES_FILE) as fh:
            return json.load(fh)["jobs"]
    except (OSError, ValueError, KeyError):
        return {}


def plan_batch(jobs: Sequence[Tuple[str, int]], max_group_peak: int = MAX_GROUP_PEAK):
    """(groups, solo): groups of job indices to merge, and the jobs to build
    alone with the plain schedule.  Jobs whose peak raw layer is known merge
    first-fit by decreasing peak into groups whose peaks sum to at most
    max_group_peak; a job above it, or of unknown size, is built alone (a
    merged layer past the key-set limit is deduplicated by sorting, which
    costs more per state than the key sets its jobs use alone)."""
    sizes = job_sizes()
    solo, sized = [], []
    for j, (p, t) in enumerate(jobs):
        e = sizes.get(f"{p}@{2 ** int(t)}")
        if e is None or e["peak"] > max_group_peak:
            solo.append(j)
        else:
            sized.append((int(e["peak"]), j))
    sized.sort(key=lambda x: (-x[0], x[1]))
    bins: List[list] = []
    for peak, j in sized:
        for b in bins:
            if b[0] + peak <= max_group_peak:
                b[0] += peak
                b[1].append(j)
                break
        else:
            bins.append([peak, [j]])
    groups = []
    for _, members in bins:
        if len(members) == 1:
            solo.append(members[0])
        else:
            groups.append(sorted(members))
    return groups, sorted(solo)


class PlannedBatch:
    """The result of a planned batch: merged groups and plain builds, with
    BatchResult's book(job, layer)."""

    def __init__(self, jobs, parts, wall_ms):
        self.jobs = list(jobs)
        self._parts = parts  # [(result, {job: index in the result, or None for a plain build})]
        self._where = {}
        for r, idx in parts:
            for j, k in idx.items():
                self._where[j] = (r, k)
        self.wall_ms = wall_ms

    def book(self, job: int, layer: int):
        r, k = self._where[job]
        return r.book(layer) if k is None else r.book(k, layer)

    def close(self):
        for r, _ in self._parts:
            r.close()
        self._parts = []
        self._where = {}

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


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


def compare_with_plain(jobs: Sequence[Tuple[str, int]], reps: int = 1, plan: str = "merge") -> dict:
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
        br = build_batch(jobs) if plan == "merge" else build_planned(jobs)
        ms = (time.perf_counter() - t0) * 1000.0
        br.close()
        best = ms if best is None else min(best, ms)
    out = {"jobs": len(jobs), "states": states, "plain_sequential_ms": round(plain_ms, 2),
           "batched_ms": round(best, 2), "speedup": round(plain_ms / best, 3),
           "batched_states_per_s": round(states / (best / 1000.0), 1), "plan": plan}
    if plan != "merge":
        groups, solo = plan_batch(jobs)
        out["groups"] = [[f"{jobs[j][0]}@{2 ** jobs[j][1]}" for j in g] for g in groups]
        out["alone"] = [f"{jobs[j][0]}@{2 ** jobs[j][1]}" for j in solo]
    return out


def build_planned(jobs: Sequence[Tuple[str, int]] = DEFAULT_JOBS, device: int = 0,
                  max_group_peak: int = MAX_GROUP_PEAK) -> PlannedBatch:
    """The jobs as plan_batch groups them: each group in one batched build,
    the others built alone (float32 rates, books in HBM)."""
    from .BookBuilder import build_in_memory

    groups, solo = plan_batch(jobs, max_group_peak)
    parts = []
    t0 = time.perf_counter()
    try:
        for g in groups:
            br = build_batch([jobs[j] for j in g], device=device)
            parts.append((br, {j: k for k, j in enumerate(g)}))
        for j in solo:
            p, t = jobs[j]
            parts.append((build_in_memory(p, t, device=device), {j: None}))
    except Exception:
        for r, _ in parts:
            r.close()
        raise
    return PlannedBatch(jobs, parts, (time.perf_counter() - t0) * 1000.0)


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
