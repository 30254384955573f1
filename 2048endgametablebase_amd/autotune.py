"""Autotuner of the dedup sort and the kernels' tile sizes, per size bucket,
by timing on the GPU:

* the sort's digit width (4, 6, 8, 11 bits) and schedule (per-digit
  three-kernel passes, single-pass decoupled look-back, or sample sort with
  per-bucket sort + unique in LDS up to 2^21 keys), and for the per-digit
  schedule its tile size (4, 8 or 16 keys per thread), timing sort + unique
  of real layer keys;
* the lookup kernel's threads per block (64 to 1024), timing batched
  lookups of a real table's states;
* the sizes of the batched mixes' jobs (peak raw layer, total states), by
  which the batch planner groups jobs (tuning/job_sizes.json).

    python -m 2048endgametablebase_amd.autotune

The winners are written to tuning/sort_gfx950.json, which native.lib()
installs on load; options with radix_bits = 0 / sort_schedule = 0 use them,
and the tile sizes apply to every sort and lookup.
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


RADIX_ITEMS = (4, 8, 16)
LOOKUP_THREADS = (64, 128, 256, 512, 1024)
LOOKUP_BUCKETS = [1 << 10, 1 << 12, 1 << 14, 1 << 16, 1 << 18, 1 << 20]
_ALL = (1 << 64) - 1


def _one_tile(kind: int, value: int) -> None:
    native.install_tile_tuning(kind, [{"max_n": _ALL, "value": value}] if value else [])


def tune(pattern: str = "442", target: int = 8, buckets=BUCKETS, reps: int = 3, log=print):
    """Sort rows {max_n, bits, schedule} and radix tile rows {max_n, value}
    (keys per thread of the per-digit schedule, 0 when another schedule
    wins), with every timing."""
    pool, rng = sample_keys(pattern, target)
    rows, tiles, table = [], [], []
    try:
        for n in buckets:
            keys = pool[rng.integers(0, pool.size, n)]
            timings = {}
            for bits, sched in CANDIDATES + ([(8, 3)] if n <= SAMPLE_SORT_MAX else []):
                for items in (RADIX_ITEMS if sched == 1 and bits < 11 else (0,)):
                    _one_tile(0, items)
                    timings[(bits, sched, items)] = time_sort(keys, bits, sched, reps)
            (bits, sched, items), best = min(timings.items(), key=lambda kv: kv[1])
            rows.append({"max_n": int(n), "bits": bits, "schedule": sched})
            tiles.append({"max_n": int(n), "value": int(items)})
            table.append({"n": int(n), "ms": {f"{b}/{s}/{i}": round(v, 4) for (b, s, i), v in timings.items()}})
            log(f"n={n:>9}: best {bits}-bit schedule {sched} items {items} {best:.4f} ms  "
                + " ".join(f"{b}/{s}/{i}:{v:.3f}" for (b, s, i), v in sorted(timings.items())))
    finally:
        _one_tile(0, 0)
    return rows, tiles, table


def tune_lookup(pattern: str = "442", target: int = 8, buckets=LOOKUP_BUCKETS, reps: int = 5, log=print):
    """Lookup threads-per-block rows {max_n, value}: batches of n queries
    drawn like the serving workload (80% table states, 20% off-table),
    median wall time of a batched lookup per candidate block size."""
    from .BookReader import Table
    from .serving import make_query_pool

    res = build_in_memory(pattern, target)
    table = Table.from_result(res, pattern, target)
    rows, timings_all = [], []
    try:
        pool = make_query_pool(res, pattern, n=max(buckets), off_table_frac=0.2, seed=1)
        for n in buckets:
            q = np.ascontiguousarray(pool[:n])
            timings = {}
            for th in LOOKUP_THREADS:
                _one_tile(1, th)
                table.lookup(q)  # warm-up
                ts = []
                for _ in range(reps):
                    t0 = time.perf_counter()
                    table.lookup(q)
                    ts.append(time.perf_counter() - t0)
                timings[th] = float(np.median(ts)) * 1000.0
            th, best = min(timings.items(), key=lambda kv: kv[1])
            rows.append({"max_n": int(n), "value": int(th)})
            timings_all.append({"n": int(n), "ms": {str(k): round(v, 4) for k, v in timings.items()}})
            log(f"lookup n={n:>8}: best {th} threads {best:.4f} ms  "
                + " ".join(f"{k}:{v:.3f}" for k, v in sorted(timings.items())))
    finally:
        _one_tile(1, 0)
        table.close()
        res.close()
    return rows, timings_all


def tune_job_sizes(jobs=None, log=print):
    """Peak raw layer, total states, layers and best-of-two plain-build time
    of each job (the mixes of batched.py by default): the table the batch
    planner groups jobs by."""
    from . import batched

    jobs = sorted(set(jobs if jobs is not None else batched.DEFAULT_JOBS + batched.SMALL_JOBS + batched.TINY_JOBS))
    build_in_memory("442", 7).close()  # warm-up
    out = {}
    for p, t in jobs:
        best = None
        for _ in range(2):
            t0 = time.perf_counter()
            r = build_in_memory(p, t)
            dt = time.perf_counter() - t0
            info = [r.layer_info(i)[0] for i in range(r.steps)]
            r.close()
            best = dt if best is None else min(best, dt)
        out[f"{p}@{2 ** t}"] = {"peak": int(max(info)), "states": int(sum(info)), "layers": len(info),
                                "plain_ms": round(best * 1000.0, 2)}
        log(f"{p}@{2 ** t}: {out[f'{p}@{2 ** t}']}")
    return out


def main():
    rows, tiles, table = tune()
    lrows, ltable = tune_lookup()
    os.makedirs(os.path.dirname(native.TUNING_FILE), exist_ok=True)
    with open(native.TUNING_FILE, "w") as fh:
        json.dump({"device": "gfx950", "pool": "442@256 layers; lookups of a 442@256 table",
                   "created": time.strftime("%Y-%m-%d"), "rows": rows,
                   "tiles": {"radix_items": tiles, "lookup_threads": lrows},
                   "timings": table, "lookup_timings": ltable}, fh, indent=1)
    native.install_sort_tuning(rows)
    native.install_tile_tuning(0, tiles)
    native.install_tile_tuning(1, lrows)
    from . import batched

    with open(batched.JOB_SIZES_FILE, "w") as fh:
        json.dump({"device": "gfx950", "created": time.strftime("%Y-%m-%d"),
                   "note": "peak raw layer, total raw states, layers and plain-build ms of catalogue jobs (the "
                           "sizes are properties of the pattern and target; the times are this GPU's)",
                   "jobs": tune_job_sizes()}, fh, indent=1)
    print(json.dumps({"rows": rows, "tiles": {"radix_items": tiles, "lookup_threads": lrows}}))


if __name__ == "__main__":
    main()
