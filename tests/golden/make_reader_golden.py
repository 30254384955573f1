"""Generate reader fixtures from the reference's own classic reader.

Builds a table with oracle/_ref/libegtb_ref.so (the reference's classic
builder, compiled from its sources by oracle/ref/Makefile, in this container
only), then asks the reference's ClassicBookReader::move_on_dic
(ReaderRuntime.cpp:1055-1068 -> :739-840) about a seeded set of positions and
writes tests/golden/reader_<tag>.json:

  * the case: pattern, target, dtype, the reader's pattern name (which decides
    prefer_max, ReaderRuntime.cpp:1053), nums_adjust, pattern_full;
  * ``cases``: [board as 16 tile values row-major, result, dtype], where
    result is the binding's ordered dict as a list of [key, value] pairs
    (value: a float, "" or null) and dtype is the returned dtype string.

Each query uses a fresh reader, so the symmetry ops are tried in the order
0..7.  Positions: stored states with one spawned tile, under every symmetry
op; boards with a pinned cell broken; boards of negative layers and of layers
past the last book; full boards without a move; the empty board.

Usage:  python tests/golden/make_reader_golden.py [tag ...]
"""
from __future__ import annotations

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

# tag: (pattern, log2 target, dtype, reader name, positions)
CASES = {
    "442_128": ("442", 7, "float32", "442", 1200),
    "442_128_uint32": ("442", 7, "uint32", "442", 600),
    # the reference keys prefer_max on the reader's pattern name only; the
    # 442 tables read under the name 4442f exercise that rule
    "442_128_prefer_max": ("442", 7, "float32", "4442f", 600),
    "LL_64": ("LL", 6, "float32", "LL", 600),
    "444_64": ("444", 6, "float32", "444", 600),
    "442_512": ("442", 9, "float32", "442", 2400),
    # rate - 1 stored, a missing key reads as -1
    "442_128_1-float32": ("442", 7, "1-float32", "442", 400),
}


def decode(b):
    m = np.zeros((4, 4), dtype=np.int64)
    for r in range(4):
        for c in range(4):
            e = (b >> (4 * ((3 - r) * 4 + (3 - c)))) & 0xF
            m[r, c] = 0 if e == 0 else 1 << e
    return m


def read_keys(path, rec_size):
    raw = np.fromfile(path, dtype=np.uint8)
    n = raw.size // rec_size
    return raw[: n * rec_size].reshape(n, rec_size)[:, :8].copy().view("<u8").reshape(n)


def positions(prefix, steps, n, seed, rec_size, tile_sum):
    rng = np.random.default_rng(seed)
    out = []
    layers = [i for i in range(steps) if os.path.exists(prefix + f"{i}.book")
              and os.path.getsize(prefix + f"{i}.book") >= rec_size]
    while len(out) < n:
        i = layers[int(rng.integers(0, len(layers)))]
        keys = read_keys(prefix + f"{i}.book", rec_size)
        b = int(keys[rng.integers(0, keys.size)])
        empt = [k for k in range(16) if (b >> (4 * k)) & 0xF == 0]
        if empt and rng.random() < 0.9:
            k = empt[rng.integers(0, len(empt))]
            b |= (1 if rng.random() < 0.9 else 2) << (4 * k)
        m = decode(b)
        kind = int(rng.integers(0, 10))
        if kind <= 5:  # any of the 8 symmetry ops
            m = np.rot90(m, int(rng.integers(0, 4)))
            if rng.random() < 0.5:
                m = np.fliplr(m)
        elif kind == 6:  # a pinned cell broken
            r, c = np.argwhere(m == 32768)[int(rng.integers(0, int((m == 32768).sum())))]
            m = m.copy()
            m[r, c] = int(2 ** rng.integers(1, 7))
        elif kind == 7:  # a layer past the last book (one more large tile)
            z = np.argwhere(m == 0)
            if z.size:
                m = m.copy()
                r, c = z[int(rng.integers(0, len(z)))]
                m[r, c] = 4096
        elif kind == 8:  # a negative layer: small tiles only
            m = np.where(m == 32768, 0, m)
        out.append(np.ascontiguousarray(m))
    # full boards without a move, in and out of the pattern
    out.append(np.array([[2, 4, 2, 4], [4, 2, 4, 2], [2, 4, 2, 4], [4, 2, 4, 2]]))
    out.append(np.array([[2, 4, 2, 4], [4, 2, 4, 2], [32768, 32768, 32768, 32768],
                         [32768, 32768, 32768, 32768]]))
    out.append(np.zeros((4, 4), dtype=np.int64))
    return out


def make(tag, work):
    pattern, t, dtype, name, n = CASES[tag]
    p = patterns.build_params(pattern, t)
    meta = patterns.pattern_catalog[pattern]
    pf = f"{pattern}_{2 ** t}"
    prefix = os.path.join(work, pf + "_")
    ref.run_build(p["seeds"], p["masks"], p["shifts"], p["symm_mode"], t, p["steps"], p["docheck_step"],
                  prefix, is_free=p["is_free"], dtype=dtype, threads=8)
    rec_size = 16 if dtype.endswith("64") else 12
    boards = positions(prefix, p["steps"], n, 1000 + t, rec_size, meta.tile_sum)
    cases = []
    for m in boards:
        res, dt = ref.move_on_dic(m, work, dtype, pf, meta.nums_adjust, name, list(meta.pattern_masks),
                                  meta.symm_mode)
        cases.append([[int(x) for x in m.reshape(16)], [[k, v] for k, v in res.items()], dt])
    numeric = sum(1 for c in cases if c[2])
    return {
        "pattern": pattern, "target": t, "dtype": dtype, "reader_name": name, "pattern_full": pf,
        "nums_adjust": meta.nums_adjust, "steps": p["steps"], "symm": meta.symm_mode,
        "masks": [f"{x:x}" for x in meta.pattern_masks],
        "source": "reference ClassicBookReader::move_on_dic over reference-built books, g++ -Ofast "
                  "-march=x86-64-v3 -flto (oracle/ref/Makefile)",
        "answered": numeric, "cases": cases,
    }


def main(argv):
    if not ref.available() and not ref.build():
        raise SystemExit("reference sources not present; cannot make reader fixtures")
    for tag in CASES:
        if argv and tag not in argv:
            continue
        work = tempfile.mkdtemp(prefix="egtb_reader_golden_", dir="/tmp")
        try:
            out = make(tag, work)
        finally:
            shutil.rmtree(work, ignore_errors=True)
        with open(os.path.join(os.path.dirname(__file__), f"reader_{tag}.json"), "w") as fh:
            json.dump(out, fh, separators=(",", ":"))
        print(f"reader_{tag}", "cases", len(out["cases"]), "answered", out["answered"], flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
