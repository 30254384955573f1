"""The advanced (AD) solver on the GPU against the reference's own advanced
builder: every solved layer's bucket files (<i>b/<key>.i, the sorted boards,
and <key>.b, a row of rates per board) of each build in
tests/golden/ad_book_*.json (written by tests/golden/make_ad_book_golden.py
from the reference compiled in oracle/_ref) must have the reference's rows,
columns and SHA-256.  The builds cover one-rate buckets (442@64, 442@128
with spelled-out pairs), rows of two and three rates (442@256, t@128 with
free 32768 tiles), the diagonal and horizontal forms (LL, 444), the
uint32 and float64 rate types and two deletion thresholds (layer i+2's rows
pruned and rewritten after layer i).  Also: the books kept in HBM equal the
files, and a rerun over partly solved folders resumes and finishes them as
the reference's recalculate_process_ad_impl does."""
import hashlib
import importlib
import os
import shutil
import tempfile

import numpy as np
import pytest

from conftest import ad_book_tags, load_ad_book_golden

pytestmark = pytest.mark.gpu


def _adv():
    return importlib.import_module("2048endgametablebase_amd.advanced")


def _sha(path):
    with open(path, "rb") as fh:
        return hashlib.sha256(fh.read()).hexdigest()


def _check_files(prefix, g, tag):
    """Every solved layer of golden g against the files under prefix."""
    for layer in g["layers"]:
        folder = f"{prefix}{layer['i']}b"
        assert os.path.isdir(folder), (tag, folder)
        have = sorted(int(f[:-2]) for f in os.listdir(folder) if f.endswith(".i"))
        want = [b["key"] for b in layer["buckets"]]
        assert have == want, (tag, layer["i"], have, want)
        for b in layer["buckets"]:
            base = os.path.join(folder, str(b["key"]))
            rows = os.path.getsize(base + ".i") // 8
            assert rows == b["rows"], (tag, layer["i"], b["key"], rows, b["rows"])
            assert _sha(base + ".i") == b["i_sha"], (tag, layer["i"], b["key"], "boards")
            assert _sha(base + ".b") == b["b_sha"], (tag, layer["i"], b["key"], "rates")


def _build(g, prefix, **kw):
    adv = _adv()
    spec = adv.ad_spec(g["pattern"], g["target"])
    return adv.run_pattern_build_ad(spec["seeds"], spec, spec["steps"], prefix,
                                    success_rate_dtype=g["dtype"], spawn_rate4=g["spawn_rate4"],
                                    deletion_threshold=g.get("deletion_threshold", 0.0), **kw)


@pytest.mark.parametrize("tag", ad_book_tags())
def test_books_match_reference(pkg, tag):
    g = load_ad_book_golden(tag)
    rows = sum(b["rows"] for l in g["layers"] for b in l["buckets"])
    print(f"case advanced solve {tag}: {len(g['layers'])} layers, {rows} rows", flush=True)
    work = tempfile.mkdtemp(prefix="egtb_adsolve_", dir="/dev/shm" if os.path.isdir("/dev/shm") else None)
    try:
        prefix = os.path.join(work, f"{g['pattern']}_{2 ** g['target']}_")
        with _build(g, prefix) as res:
            info = res.solve_info()
            assert info["skipped_layers"] == 0
        _check_files(prefix, g, tag)
        # the raw layers were removed as the reference removes them
        assert not any(os.path.exists(f"{prefix}{i}") for i in range(g["steps"]))
    finally:
        shutil.rmtree(work, ignore_errors=True)


def test_kept_books_equal_files(pkg):
    g = load_ad_book_golden("442_256")
    print("case advanced solve kept books 442_256", flush=True)
    with _build(g, None, keep_books=True) as res:
        for layer in g["layers"][::7]:
            got = res.buckets(layer["i"])
            assert sorted(got) == [b["key"] for b in layer["buckets"]], layer["i"]
            for b in layer["buckets"]:
                boards, rates = res.book(layer["i"], b["key"])
                assert rates.shape == (b["rows"], b["cols"])
                assert hashlib.sha256(boards.astype("<u8").tobytes()).hexdigest() == b["i_sha"]
                assert hashlib.sha256(rates.astype("<f4").tobytes()).hexdigest() == b["b_sha"]


def test_solve_resumes_from_folders(pkg):
    """Solved layers 0..39 removed after a full build: the rerun regenerates
    their raw layers, skips the layers still solved on disk, reads layers 40
    and 41 back and solves 39..0 to the reference's bytes."""
    g = load_ad_book_golden("442_128")
    print("case advanced solve resume 442_128 from layer 40", flush=True)
    work = tempfile.mkdtemp(prefix="egtb_adsolve_", dir="/dev/shm" if os.path.isdir("/dev/shm") else None)
    try:
        prefix = os.path.join(work, "442_128_")
        with _build(g, prefix):
            pass
        for i in range(40):
            shutil.rmtree(f"{prefix}{i}b")
        with _build(g, prefix) as res:
            info = res.solve_info()
        assert info["skipped_layers"] == g["steps"] - 2 - 40
        _check_files(prefix, g, "442_128 resumed")
    finally:
        shutil.rmtree(work, ignore_errors=True)


def test_start_build_advanced_writes_reference_books(pkg):
    """start_build with advanced_algo runs the whole advanced build, as the
    reference's start_build -> _run_advanced_build does."""
    g = load_ad_book_golden("442_64")
    print("case advanced start_build 442_64", flush=True)
    cfg = pkg.config.SingletonConfig().config
    saved = {k: cfg.get(k) for k in ("advanced_algo", "success_rate_dtype", "deletion_threshold")}
    work = tempfile.mkdtemp(prefix="egtb_adsolve_", dir="/dev/shm" if os.path.isdir("/dev/shm") else None)
    try:
        cfg.update(advanced_algo=True, success_rate_dtype="float32", deletion_threshold=0.0)
        prefix = os.path.join(work, "442_64_")
        assert pkg.BookBuilder.start_build("442", 6, prefix) is True
        _check_files(prefix, g, "start_build 442_64")
        assert os.path.exists(prefix + "config.txt")
    finally:
        for k, v in saved.items():
            if v is None:
                cfg.pop(k, None)
            else:
                cfg[k] = v
        shutil.rmtree(work, ignore_errors=True)
