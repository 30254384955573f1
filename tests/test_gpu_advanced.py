"""The advanced (AD, large-tile masking) forward pass on the GPU against
the reference's own AD builder: every raw layer of twelve builds (two of
them with other SmallTileSumLimit values)
(tests/golden/ad_*.json, written by tests/golden/make_ad_golden.py from the
reference compiled in oracle/_ref) must have the reference's size and
SHA-256.  Also: the raw files written under a path, steps split into parent
ranges, and the single-pass sort schedule give the same bytes."""
import hashlib
import importlib
import os
import shutil
import tempfile

import numpy as np
import pytest

from conftest import ad_tags, load_ad_golden

pytestmark = pytest.mark.gpu


def _adv():
    return importlib.import_module("2048endgametablebase_amd.advanced")


def _check_layers(res, g, tag):
    assert res.steps == g["steps"]
    for layer in g["layers"]:
        a = res.layer(layer["i"])
        assert a.size == layer["n"], (tag, layer["i"], a.size, layer["n"])
        assert hashlib.sha256(a.astype("<u8").tobytes()).hexdigest() == layer["sha256"], (tag, layer["i"])


@pytest.mark.parametrize("tag", ad_tags())
def test_layers_match_reference(pkg, tag):
    g = load_ad_golden(tag)
    print(f"case advanced forward {tag}: {sum(l['n'] for l in g['layers'])} states", flush=True)
    adv = _adv()
    spec = dict(adv.ad_spec(g["pattern"], g["target"]), small_tile_sum_limit=g["small_tile_sum_limit"])
    with adv.run_pattern_generate_ad(spec["seeds"], spec, spec["steps"], keep_layers=True) as res:
        _check_layers(res, g, tag)
        print(f"{tag}: {res.fwd_ms:.1f} ms, {res.candidates} boards before dedup")


def test_raw_files_match_reference(pkg):
    g = load_ad_golden("442_64")
    work = tempfile.mkdtemp(prefix="egtb_ad_")
    try:
        prefix = os.path.join(work, "442_64_")
        print("case advanced raw files 442@64", flush=True)
        _adv().generate_ad("442", 6, prefix).close()
        for layer in g["layers"]:
            raw = open(prefix + str(layer["i"]), "rb").read()
            assert hashlib.sha256(raw).hexdigest() == layer["sha256"], layer["i"]
        assert not os.path.exists(prefix + str(len(g["layers"])))
    finally:
        shutil.rmtree(work, ignore_errors=True)


def test_parent_ranges_and_sort_schedule(pkg):
    g = load_ad_golden("t_128")
    print("case advanced t@128 with 1 Mi-board parent ranges", flush=True)
    with _adv().generate_ad("t", 7, keep_layers=True, chunk_mkeys=1) as res:
        assert res.chunked_steps > 0
        _check_layers(res, g, "t_128 chunked")
    print("case advanced t@128 with the single-pass sort schedule", flush=True)
    with _adv().generate_ad("t", 7, keep_layers=True, sort_schedule=2) as res:
        _check_layers(res, g, "t_128 schedule 2")


@pytest.mark.parametrize("tag", ["t_128", "LL_64"])
def test_forward_dedup_paths(pkg, tag):
    """Key sets (the default), sort + unique at every step, and key sets
    sized at 1/64 (steps whose sets overflow run again through sort +
    unique) give the same layers."""
    g = load_ad_golden(tag)
    adv = _adv()
    with adv.generate_ad(g["pattern"], g["target"], keep_layers=True) as res:
        assert res.hash_steps > 0.9 * (res.steps - 2) and res.set_overflows == 0, (res.hash_steps,
                                                                                  res.set_overflows)
    for mode in (1, 2):
        print(f"case advanced {tag} forward_dedup {mode}", flush=True)
        with adv.generate_ad(g["pattern"], g["target"], keep_layers=True, forward_dedup=mode) as res:
            _check_layers(res, g, f"{tag} forward_dedup {mode}")
            if mode == 1:
                assert res.hash_steps == 0
            else:
                assert res.set_overflows > 0 and res.hash_steps > 0, (res.hash_steps, res.set_overflows)


def test_resume_from_raw_files(pkg):
    """An interrupted pass resumes as the reference's handle_restart_ad does:
    steps whose raw layers are on disk are skipped, the first step run reads
    layers i-1 and i back, and every file ends up with the reference's bytes;
    a raw file that is not whole boards is an error."""
    g = load_ad_golden("442_128")
    adv = _adv()
    work = tempfile.mkdtemp(prefix="egtb_ad_resume_")
    try:
        prefix = os.path.join(work, "442_128_")
        print("case advanced resume 442@128: full pass", flush=True)
        adv.generate_ad("442", 7, prefix).close()
        n = len(g["layers"])
        for i in range(41, n):
            os.remove(prefix + str(i))
        print("case advanced resume 442@128: from layer 40", flush=True)
        with adv.generate_ad("442", 7, prefix, keep_layers=True) as res:
            assert res.skipped_steps == 39
            _check_layers(res, g, "442_128 resumed")
        for layer in g["layers"]:
            raw = open(prefix + str(layer["i"]), "rb").read()
            assert hashlib.sha256(raw).hexdigest() == layer["sha256"], layer["i"]
        for i in range(41, n):
            os.remove(prefix + str(i))
        with open(prefix + "40", "r+b") as fh:
            fh.truncate(8 * (g["layers"][40]["n"] - 1) + 3)
        print("case advanced resume 442@128: misaligned layer 40", flush=True)
        with pytest.raises(pkg.native.EgtbError, match="misaligned"):
            adv.generate_ad("442", 7, prefix)
    finally:
        shutil.rmtree(work, ignore_errors=True)


def test_full_size_4442_64(pkg):
    """4442@64 at full size (2.67e9 states over 79 raw layers, the largest
    94.6e6 boards, most steps past the key-set limit): every layer's size
    and SHA-256 equal the reference's."""
    g = load_ad_golden("4442_64")
    print("case advanced 4442@64 full size", flush=True)
    with _adv().generate_ad("4442", 6, keep_layers=True) as res:
        print(f"4442@64: {res.fwd_ms:.0f} ms, {res.hash_steps} key-set steps", flush=True)
        assert res.hash_steps > 0 and res.hash_steps < res.steps - 2
        _check_layers(res, g, "4442_64")


def test_edge_cases(pkg, tmp_path):
    """No seeds: every layer is empty. steps = 2: the reference's loop runs
    no step, so no raw file is written (handle_restart_ad writes layer 0 only
    when step 1 runs)."""
    adv = _adv()
    spec = adv.ad_spec("442", 7)
    print("case advanced edge: no seeds", flush=True)
    with adv.run_pattern_generate_ad([], spec, 20, keep_layers=True) as res:
        assert res.sizes.size == 19 and not res.sizes.any()
        assert res.layer(5).size == 0
    print("case advanced edge: steps = 2", flush=True)
    prefix = str(tmp_path / "x_")
    with adv.run_pattern_generate_ad(spec["seeds"], spec, 2, prefix) as res:
        assert res.steps == 2 and not res.sizes.any()
    assert not os.path.exists(prefix + "0")
