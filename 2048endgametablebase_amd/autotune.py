"""Autotuner of the dedup sort: digit width (4, 6, 8, 11 bits) and schedule
(per-digit three-kernel passes, single-pass decoupled look-back, or sample
sort with per-bucket sort + unique in LDS up to 2^21 keys), chosen per
key-count bucket by timing sort + unique of real layer keys on the GPU.

    python -m ... autotune   (see main below)

The winners are written to tuning/sort_gfx950.json, which native.lib()
installs on load; options with radix_bits = 0 / sort_schedule = 0 use it.
"""
from __future__ import annotations

import ctypes
import json
import os
import time

import numpy as np

from . import native
from .BookBuilder import build_in_memory

BUCKETS = [1 << 12, 1 << 14, 1 << 16, 1 << 18, 1 << 20, 1 << 22, 1 << 24]
CANDIDATES = [(b, s) for s in (1, 2) for b in (4, 6, 8, 11)]
SAMPLE_SORT_MAX = 1 << 21  # kBucketMaxKeys of the sample-sort dedup


def sample_keys(pattern: str = "442", target: int = 8, seed: int = 0):
    """A pool of real layer keys (with the duplicates a candidate stream has)."""
    res = build_in_memory(pattern, target, keep_raw=True)
    try:
        mid = res.steps // 2
        pool = np.concatenate([res.raw(i) for i in range(max(1, mid - 8), mid + 8)])
    finally:
        res.close()
    return pool, np.random.default_rng(seed)


def time_sort(keys: np.ndarray, bits: int, schedule: int, reps: int = 3) -> float:
    k = np.ascontiguousarray(keys, dtype=np.uint64)
    ms = ctypes.c_double()
    native.check(native.lib().egtb_bench_sort(k.ctypes.data_as(native.u64p), k.size, bits, schedule, reps,
                                              ctypes.byref(ms)))
    return ms.value


def tune(pattern: str = "442", target: int = 8, buckets=BUCKETS, reps: int = 3, log=print):
    pool, rng = sample_keys(pattern, target)
    rows, table = [], []
    for n in buckets:
        keys = pool[rng.integers(0, pool.size, n)]
        timings = {}
        for bits, sched in CANDIDATES + ([(8, 3)] if n <= SAMPLE_SORT_MAX else []):
            timings[(bits, sched)] = time_sort(keys, bits, sched, reps)
        (bits, sched), best = min(timings.items(), key=lambda kv: kv[1])
        rows.append({"max_n": int(n), "bits": bits, "schedule": sched})
        table.append({"n": int(n), "ms": {f"{b}/{s}": round(v, 4) for (b, s), v in timings.items()}})
        log(f"n={n:>9}: best {bits}-bit schedule {sched} {best:.4f} ms  "
            + " ".join(f"{b}/{s}:{v:.3f}" for (b, s), v in sorted(timings.items())))
    return rows, table


def main():
    rows, table = tune()
    os.makedirs(os.path.dirname(native.TUNING_FILE), exist_ok=True)
    with open(native.TUNING_FILE, "w") as fh:
        json.dump({"device": "gfx950", "pool": "442@256 layers", "created": time.strftime("%Y-%m-%d"),
                   "rows": rows, "timings": table}, fh, indent=1)
    native.install_sort_tuning(rows)
    print(json.dumps(rows))


if __name__ == "__main__":
    main()
