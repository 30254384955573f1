"""Generate golden vectors from the reference's own classic builder.

Runs oracle/_ref/libegtb_ref.so (the reference C++ compiled from its sources
by oracle/ref/Makefile, in this container only) on a set of patterns and
writes tests/golden/<pattern>_<tile>.json holding, per layer:

  * ``n``      -- raw layer size N_i from the reference's stats.txt
                  (solved layers) -- absent for the two final layers;
  * ``live``   -- number of records in <i>.book (rate > 0);
  * ``sha256`` -- SHA-256 of the .book bytes ({u64 board; rate} records,
                  the rate in the run's success_rate_dtype);
  * ``sample`` -- up to 16 evenly spaced records as [board_hex, rate_bits].

Runs in a dtype other than float32 are written to <pattern>_<tile>_<dtype>.json;
runs with post-solve pruning add "_dt<threshold>" (RunOptions::deletion_threshold)
and "_opt" (RunOptions::optimal_branch_only) to the tag, and record the
options under "options".

Usage:  python tests/golden/make_golden.py [name ...]
"""
from __future__ import annotations

import csv
import hashlib
import importlib
import json
import os
import shutil
import sys
import tempfile

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, ROOT)

from oracle.ref import ref  # noqa: E402

patterns = importlib.import_module("2048endgametablebase_amd.patterns")

# (pattern, log2 target[, dtype]); the full-symmetry case uses the free8
# spec below.  uint32 is the reference's default success_rate_dtype.
RUNS = [
    ("L3", 7), ("L3", 8), ("442", 7), ("442", 8), ("442", 9), ("t", 8), ("LL", 7),
    ("4431", 7), ("LL", 6), ("444", 6), ("4432", 6), ("free8", 6), ("t", 6),
    ("442", 7, "uint32"), ("L3", 7, "uint32"), ("LL", 6, "uint32"), ("442", 7, "float64"),
    ("442", 7, "uint64"), ("442", 7, "1-float32"), ("442", 7, "1-float64"), ("LL", 6, "1-float32"),
    ("442", 7, "float32", {"deletion_threshold": 0.3}),
    ("442", 7, "uint32", {"deletion_threshold": 0.05}),
    ("442", 7, "float32", {"optimal_branch_only": True}),
    ("442", 7, "uint32", {"optimal_branch_only": True, "deletion_threshold": 0.3}),
    ("LL", 6, "float32", {"optimal_branch_only": True}),
    ("444", 9),
]
RATE_FORMAT = {"float32": "<f4", "uint32": "<u4", "float64": "<f8", "uint64": "<u8", "1-float32": "<f4",
               "1-float64": "<f8"}


def run_tag(name, t, dtype="float32", opts=None):
    tag = f"{name}_{2 ** t}" + ("" if dtype == "float32" else f"_{dtype}")
    opts = opts or {}
    if opts.get("deletion_threshold"):
        tag += f"_dt{opts['deletion_threshold']}"
    if opts.get("optimal_branch_only"):
        tag += "_opt"
    return tag

# free8 (docs_and_configs/patterns_config.json): no masks, every nibble a
# success position, full symmetry, single seed.  Built from its seed as is.
FREE8 = dict(masks=[], shifts=list(range(0, 64, 4)), symm=1, seeds=[0x01111111FFFFFFFF],
             extra=36)


def params(name, t):
    if name == "free8":
        ts = patterns.tile_sum_of(FREE8["seeds"][0])
        steps, doc = patterns.steps_and_docheck(ts, t, FREE8["extra"])
        return dict(seeds=FREE8["seeds"], masks=FREE8["masks"], shifts=FREE8["shifts"],
                    symm=FREE8["symm"], steps=steps, docheck=doc, is_free=True)
    p = patterns.build_params(name, t)
    return dict(seeds=p["seeds"], masks=p["masks"], shifts=p["shifts"], symm=p["symm_mode"],
                steps=p["steps"], docheck=p["docheck_step"], is_free=p["is_free"])


def digest_run(name, t, workdir, dtype="float32", opts=None):
    p = params(name, t)
    tile = 2 ** t
    opts = opts or {}
    prefix = os.path.join(workdir, f"{name}_{tile}_")
    ref.run_build(p["seeds"], p["masks"], p["shifts"], p["symm"], t, p["steps"], p["docheck"],
                  prefix, is_free=p["is_free"], dtype=dtype, threads=8, **opts)
    lengths = {}
    with open(prefix + "stats.txt") as fh:
        for row in list(csv.reader(fh))[1:]:
            lengths[int(row[0])] = int(row[1])
    layers = []
    rec = np.dtype([("board", "<u8"), ("rate", RATE_FORMAT[dtype])])
    bits = np.uint32 if rec["rate"].itemsize == 4 else np.uint64
    for i in range(p["steps"]):
        path = prefix + f"{i}.book"
        raw = open(path, "rb").read() if os.path.exists(path) else b""
        a = np.frombuffer(raw, dtype=rec)
        idx = np.unique(np.linspace(0, max(a.size - 1, 0), num=min(16, a.size)).astype(np.int64))
        layers.append({
            "i": i,
            "n": lengths.get(i),
            "live": int(a.size),
            "sha256": hashlib.sha256(raw).hexdigest(),
            "sample": [[f"{int(a['board'][j]):016x}", int(a["rate"][j:j + 1].view(bits)[0])]
                       for j in idx] if a.size else [],
        })
    return {
        "pattern": name, "target": t, "steps": p["steps"], "docheck": p["docheck"],
        "seeds": [f"{s:016x}" for s in p["seeds"]], "masks": [f"{m:x}" for m in p["masks"]],
        "shifts": list(p["shifts"]), "symm": p["symm"], "spawn_rate4": 0.1,
        "dtype": dtype, "options": opts, "source": "reference run_pattern_build_cpp, g++ -Ofast "
        "-march=x86-64-v3 -flto (oracle/ref/Makefile)", "layers": layers,
    }


def main(argv):
    if not ref.available() and not ref.build():
        raise SystemExit("reference sources not present; cannot make golden vectors")
    want = set(argv)
    for run in RUNS:
        name, t = run[0], run[1]
        dtype = run[2] if len(run) > 2 else "float32"
        opts = run[3] if len(run) > 3 else {}
        tag = run_tag(name, t, dtype, opts)
        if want and tag not in want and name not in want:
            continue
        work = tempfile.mkdtemp(prefix="egtb_golden_", dir="/tmp")
        try:
            out = digest_run(name, t, work, dtype, opts)
        finally:
            shutil.rmtree(work, ignore_errors=True)
        with open(os.path.join(os.path.dirname(__file__), f"{tag}.json"), "w") as fh:
            json.dump(out, fh, separators=(",", ":"))
        print(tag, "layers", out["steps"], "live", sum(l["live"] for l in out["layers"]))


if __name__ == "__main__":
    main(sys.argv[1:])
