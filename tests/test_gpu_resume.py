"""GPU tests of the file-set contract of egtb_build: resume of an interrupted
build (BookGenerator.cpp:384-430, BookSolver.cpp:476-511), raw layer files
(FileIOUtils.h:444-451) and the deletion threshold (BookSolver.cpp:584-588).
Every build goes through the C ABI; the expected bytes are the reference's
golden digests or the same build's own in-memory layers."""
import hashlib
import os

import numpy as np
import pytest

from conftest import book_bytes, book_sha, golden_params, load_golden

pytestmark = pytest.mark.gpu


def _files_build(pkg, gp, prefix, **kw):
    pkg.run_pattern_build(gp["seeds"], gp["masks"], gp["shifts"], gp["symm"], gp["target"], gp["steps"],
                          gp["docheck"], prefix, success_rate_dtype="float32", write_files=True, **kw)


def _book_sha_file(path):
    return hashlib.sha256(open(path, "rb").read()).hexdigest()


def _check_against_golden(g, prefix):
    bad = [L["i"] for L in g["layers"] if _book_sha_file(f"{prefix}{L['i']}.book") != L["sha256"]]
    assert not bad, f"layers differing from the reference: {bad[:10]}"


def _leftovers(folder):
    return sorted(f for f in os.listdir(folder) if f.endswith(".tmp") or not (f.endswith(".book") or
                                                                              f.endswith(".txt")))


def test_resume_skips_solved_layers(pkg, tmp_path):
    g = load_golden("442_128")
    gp = golden_params(g)
    prefix = str(tmp_path / "442_128_")
    print("case resume 442_128", flush=True)
    _files_build(pkg, gp, prefix)
    _check_against_golden(g, prefix)
    steps = gp["steps"]
    cut = steps // 2
    # an interrupted backward pass: layers below `cut` unsolved, one truncated
    # book, a leftover raw layer and an empty raw file
    for i in range(cut):
        os.remove(f"{prefix}{i}.book")
    with open(f"{prefix}{cut - 1}.book", "wb") as fh:
        fh.write(b"\x00" * 13)  # not a multiple of 12: removed as invalid
    np.arange(1000, dtype=np.uint64).tofile(f"{prefix}{cut - 3}")
    open(f"{prefix}{cut - 5}", "wb").close()
    kept = {i: os.stat(f"{prefix}{i}.book").st_mtime_ns for i in range(cut, steps)}
    rows_before = len(open(prefix + "stats.txt").read().splitlines())
    _files_build(pkg, gp, prefix)
    _check_against_golden(g, prefix)
    assert {i: os.stat(f"{prefix}{i}.book").st_mtime_ns for i in range(cut, steps)} == kept, \
        "solved layers were written again"
    rows = open(prefix + "stats.txt").read().splitlines()
    new_layers = [int(r.split(",")[0]) for r in rows[rows_before:]]
    assert new_layers and max(new_layers) < cut and new_layers == sorted(new_layers, reverse=True)
    assert _leftovers(tmp_path) == []


def test_resume_from_final_layers(pkg, tmp_path):
    """Only the two final books left: every other layer is solved again,
    starting from books loaded from their files."""
    g = load_golden("L3_128")
    gp = golden_params(g)
    prefix = str(tmp_path / "L3_128_")
    print("case resume L3_128", flush=True)
    _files_build(pkg, gp, prefix)
    steps = gp["steps"]
    for i in range(steps - 3):
        os.remove(f"{prefix}{i}.book")
    os.remove(f"{prefix}{steps - 3}.book")
    _files_build(pkg, gp, prefix)
    _check_against_golden(g, prefix)
    assert _leftovers(tmp_path) == []


def test_resume_loads_books_from_files(pkg, tmp_path):
    """The first layer solved reads books i+1 and i+2 from disk: planting
    modified records there changes that layer, which proves the files are
    the source (the reference reads them the same way)."""
    g = load_golden("442_128")
    gp = golden_params(g)
    prefix = str(tmp_path / "442_128_")
    print("case resume-from-files 442_128", flush=True)
    _files_build(pkg, gp, prefix)
    steps = gp["steps"]
    cut = steps // 2
    for i in range(cut):
        os.remove(f"{prefix}{i}.book")
    rec = np.dtype([("board", "<u8"), ("rate", "<f4")])
    for j in (cut, cut + 1):
        r = np.fromfile(f"{prefix}{j}.book", dtype=rec)
        r["rate"] = np.float32(1.0)
        r.tofile(f"{prefix}{j}.book")
    _files_build(pkg, gp, prefix)
    want = {L["i"]: L["sha256"] for L in g["layers"]}
    assert _book_sha_file(f"{prefix}{cut - 1}.book") != want[cut - 1]


def test_raw_layer_files_match_layers(pkg, orc, tmp_path):
    """write_raw_files=2 keeps <pathname><i>: N_i little-endian u64, layer 0
    the seeds in their given order, the rest the oracle's sorted layers."""
    p = pkg.patterns.build_params("L3", 7)
    gp = dict(seeds=p["seeds"], masks=p["masks"], shifts=p["shifts"], symm=p["symm_mode"], target=7,
              steps=p["steps"], docheck=p["docheck_step"])
    prefix = str(tmp_path / "L3_128_")
    print("case raw files L3_128", flush=True)
    _files_build(pkg, gp, prefix, write_raw_files=2)
    ob = orc.OracleBuild(gp["seeds"], gp["masks"], gp["shifts"], gp["symm"], 7, gp["steps"], gp["docheck"])
    try:
        for i in range(gp["steps"] - 1):
            raw = np.fromfile(f"{prefix}{i}", dtype="<u8")
            assert np.array_equal(raw, ob.raw(i)), f"raw layer file {i}"
        assert not os.path.exists(f"{prefix}{gp['steps'] - 1}")  # the partial last layer is never written
    finally:
        ob.close()
    # mode 1 removes every raw file once its layer is solved
    prefix1 = str(tmp_path / "again_")
    _files_build(pkg, gp, prefix1, write_raw_files=1)
    assert not [f for f in os.listdir(tmp_path) if f.startswith("again_") and f[len("again_"):].isdigit()]


def test_deletion_threshold(pkg, tmp_path):
    """Books 2..steps-1 keep only rates > threshold on disk; books 0 and 1
    stay whole; the stats rows carry the threshold."""
    g = load_golden("442_128")
    gp = golden_params(g)
    thr = 0.5
    prefix = str(tmp_path / "442_128_")
    print("case deletion threshold 442_128", flush=True)
    _files_build(pkg, gp, prefix, deletion_threshold=thr)
    res = pkg.run_pattern_build(gp["seeds"], gp["masks"], gp["shifts"], gp["symm"], gp["target"], gp["steps"],
                                gp["docheck"], "", write_files=False, keep=True)
    try:
        dropped = 0
        for i in range(gp["steps"]):
            keys, rates = res.book(i)
            data = open(f"{prefix}{i}.book", "rb").read() if os.path.exists(f"{prefix}{i}.book") else b""
            if i >= 2:
                m = rates > np.float32(thr)
                dropped += int((~m).sum())
                assert data == book_bytes(keys[m], rates[m]), f"layer {i}"
            else:
                assert hashlib.sha256(data).hexdigest() == book_sha(keys, rates), f"layer {i}"
        assert dropped > 0
    finally:
        res.close()
    rows = open(prefix + "stats.txt").read().splitlines()[1:]
    assert rows and all(r.split(",")[4] == "0.5" for r in rows)
