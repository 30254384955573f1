"""Generate golden vectors of the advanced (AD) builder's solved books from the
reference's own builder.

Runs oracle/_ref/libegtb_ref.so (the reference C++ compiled from its sources
by oracle/ref/Makefile, in this container only): run_pattern_build_ad_cpp
(native_core/src/BookGeneratorAD.cpp:1421-1438) through its solver
(recalculate_process_ad_impl, native_core/src/BookSolverAD.cpp:2086-2242),
and writes tests/golden/ad_book_<pattern>_<tile>[_<dtype>].json holding, per
solved layer <pathname><i>b (i = 0 .. steps-3) and bucket key k (the files
<k>.i, sorted uint64 boards, and <k>.b, rows x cols rates, BookSolverAD.cpp:
320-352):

  * ``rows`` / ``cols`` -- boards in the bucket and rates per board;
  * ``i_sha`` / ``b_sha`` -- SHA-256 of the two files' bytes;
  * ``sample`` -- up to 2 evenly spaced rows: board (hex) and its rates.

The spec is engine_core/BookBuilder.py:90-108's (SmallTileSumLimit 96) and
the spawn rate of 4s 0.1.

Usage:  python tests/golden/make_ad_book_golden.py [tag ...]
"""
from __future__ import annotations

import hashlib
import importlib
import json
import os
import shutil
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, ROOT)

from oracle.ref import ref  # noqa: E402

patterns = importlib.import_module("2048endgametablebase_amd.patterns")

NP_TYPES = {"float32": "<f4", "float64": "<f8", "uint32": "<u4", "uint64": "<u8"}

# (pattern, log2 target, dtype[, deletion threshold]): one-rate buckets only
# (442@64, 442@128 with its spelled-out pairs), rows of two rates (442@256),
# free 32768 tiles (t, rows of three), the diagonal and horizontal forms (LL,
# 444), nine free 32768 tiles (free8), the integer and double rate types, and
# rows pruned below a deletion threshold (layer i+2 rewritten after layer i)
RUNS = [("442", 6, "float32"), ("442", 7, "float32"), ("442", 8, "float32"), ("t", 6, "float32"),
        ("t", 7, "float32"), ("L3", 7, "float32"), ("LL", 6, "float32"), ("444", 6, "float32"),
        ("442", 7, "uint32"), ("442", 7, "float64"), ("free8", 6, "float32"), ("442", 8, "float32", 0.3),
        ("t", 7, "uint32", 0.05)]


def tag_of(name, t, dtype, thr=0.0):
    return (f"ad_book_{name}_{2 ** t}" + ("" if dtype == "float32" else f"_{dtype}") +
            (f"_del{int(round(thr * 100)):02d}" if thr else ""))


def digest_run(name, t, dtype, workdir, thr=0.0):
    p = patterns.ad_params(name, t, 96)
    prefix = os.path.join(workdir, f"{name}_{2 ** t}_")
    t0 = time.time()
    ref.run_build_ad(p["seeds"], p["masks"], p["symm_mode"], p["n_free_32k"], p["fixed_32k_shifts"], t,
                     p["steps"], p["docheck_step"], prefix, is_free=p["is_free"], dtype=dtype, threads=8,
                     stop_after_forward=False, deletion_threshold=thr)
    secs = time.time() - t0
    layers = []
    for i in range(p["steps"] - 2):
        folder = prefix + f"{i}b"
        buckets = []
        keys = sorted(int(f[:-2]) for f in os.listdir(folder) if f.endswith(".i"))
        for k in keys:
            ib = open(os.path.join(folder, f"{k}.i"), "rb").read()
            bb = open(os.path.join(folder, f"{k}.b"), "rb").read()
            idx = np.frombuffer(ib, dtype="<u8")
            vals = np.frombuffer(bb, dtype=NP_TYPES[dtype])
            rows = int(idx.size)
            cols = int(vals.size // rows) if rows else 0
            pick = np.unique(np.linspace(0, max(rows - 1, 0), num=min(2, rows)).astype(np.int64))
            sample = [[f"{int(idx[j]):016x}", [float(v) for v in vals[j * cols:(j + 1) * cols]]] for j in pick]
            buckets.append({"key": k, "rows": rows, "cols": cols, "i_sha": hashlib.sha256(ib).hexdigest(),
                            "b_sha": hashlib.sha256(bb).hexdigest(), "sample": sample})
        layers.append({"i": i, "buckets": buckets})
    return {
        "pattern": name, "target": t, "steps": p["steps"], "dtype": dtype, "spawn_rate4": 0.1,
        "deletion_threshold": thr,
        "seeds": [f"{s:016x}" for s in p["seeds"]], "masks": [f"{m:x}" for m in p["masks"]],
        "symm": p["symm_mode"], "n_free_32k": p["n_free_32k"], "fixed_32k_shifts": p["fixed_32k_shifts"],
        "small_tile_sum_limit": 96,
        "source": "reference run_pattern_build_ad_cpp (forward pass and solver), g++ -Ofast -march=x86-64-v3 "
        "-flto (oracle/ref/Makefile)", "reference_seconds_8_threads": round(secs, 1), "layers": layers,
    }


def main(argv):
    if not ref.available() and not ref.build():
        raise SystemExit("reference sources not present; cannot make golden vectors")
    want = set(argv)
    for run in RUNS:
        name, t, dtype = run[:3]
        thr = run[3] if len(run) > 3 else 0.0
        tag = tag_of(name, t, dtype, thr)
        if want and tag not in want:
            continue
        work = tempfile.mkdtemp(prefix="egtb_adbook_", dir="/dev/shm")
        try:
            out = digest_run(name, t, dtype, work, thr)
        finally:
            shutil.rmtree(work, ignore_errors=True)
        with open(os.path.join(os.path.dirname(__file__), f"{tag}.json"), "w") as fh:
            json.dump(out, fh, separators=(",", ":"))
        n = sum(b["rows"] for l in out["layers"] for b in l["buckets"])
        print(tag, "layers", len(out["layers"]), "rows", n, "s", out["reference_seconds_8_threads"], flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
