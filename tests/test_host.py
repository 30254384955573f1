"""Host-side logic and the C ABI boundary (CPU only: no GPU compute)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))


@pytest.mark.parametrize("pattern,target,steps,docheck", [
    # the reference's steps / docheck formulas (engine_core/BookBuilder.py:124-128) worked out per config
    ("L3", 7, 112, 60), ("L3", 8, 176, 124), ("L3", 9, 304, 252), ("L3", 10, 560, 508),
    ("442", 9, 304, 252), ("444", 8, 188, 127), ("444", 11, 1084, 1023),
    ("4442", 11, 1072, 1020), ("LL", 9, 304, 249), ("4431", 9, 304, 249),
])
def test_steps_and_docheck(pkg, pattern, target, steps, docheck):
    p = pkg.patterns.build_params(pattern, target)
    assert (p["steps"], p["docheck_step"]) == (steps, docheck)


def test_catalogue_entries(pkg):
    cat = pkg.patterns.pattern_catalog
    assert cat["L3"].success_shifts == tuple(range(0, 64, 4))
    assert cat["442"].success_shifts == (24,)
    assert cat["444"].success_shifts == (20, 24)
    assert cat["444"].symm_mode == 3 and cat["LL"].symm_mode == 2 and cat["442"].symm_mode == 0
    assert cat["L3"].tile_sum == 196616 and cat["444"].tile_sum == 131074
    assert cat["442"].free_cells == 10 and cat["444"].free_cells == 12
    assert pkg.patterns.build_params("442", 9)["is_free"] is False


def _header_symbols():
    text = open(os.path.join(ROOT, "include", "egtb.h")).read()
    return re.findall(r"EGTB_EXPORT\s+[\w\s\*]+?\b(\w+)\s*\(", text)


def test_library_exports_every_header_symbol(pkg):
    syms = _header_symbols()
    assert "egtb_build" in syms and "sort_uint64" in syms and len(syms) >= 15
    lib = ctypes.CDLL(pkg.native.LIB_PATH)
    for name in syms:
        assert hasattr(lib, name), name
    assert set(syms) == set(pkg.native.SIGNATURES), "ctypes binding out of step with the header"


def test_invalid_arguments_fail_before_device_use(pkg):
    """Argument errors come back as codes + messages (no GPU touched)."""
    with pytest.raises(pkg.native.EgtbError, match="symm_mode"):
        pkg.run_pattern_build([1], [0xFFFF], [24], 5, 9, 304, 252, "", write_files=False)
    with pytest.raises(pkg.native.EgtbError, match="target"):
        pkg.run_pattern_build([1], [0xFFFF], [24], 0, 0, 304, 252, "", write_files=False)
    with pytest.raises(pkg.native.EgtbError, match="steps"):
        pkg.run_pattern_build([1], [0xFFFF], [24], 0, 9, 2, 0, "", write_files=False)
    with pytest.raises(pkg.native.EgtbError, match="16 pattern masks"):
        pkg.run_pattern_build([1], list(range(17)), [24], 0, 9, 304, 252, "", write_files=False)
    with pytest.raises(pkg.native.EgtbError, match="success_rate_dtype"):
        pkg.run_pattern_build([1], [0xFFFF], [24], 0, 9, 304, 252, "", write_files=False,
                              success_rate_dtype="2-float32")
    assert "success_rate_dtype" in pkg.native.last_error()


def test_save_config_to_txt(pkg, tmp_path):
    path = tmp_path / "x_config.txt"
    pkg.BookBuilder.save_config_to_txt(str(path))
    lines = path.read_text().splitlines()
    assert lines[0] == "compress: False"
    assert "4_spawn_rate: 0.1" in lines and "success_rate_dtype: uint32" in lines
    assert len(lines) == 10


def test_unknown_pattern_raises(pkg, tmp_path):
    with pytest.raises(KeyError):
        pkg.start_build("nope", 9, str(tmp_path / "x_"))


def test_key_set_sizing_and_forward_query_check_arguments(pkg):
    """The forward key-set tuning and the forward-dedup query validate their
    arguments on the host (no GPU touched)."""
    L = pkg.native.lib()
    assert L.egtb_set_key_set_sizing(-1.0, 0) != 0
    assert "slots per key" in pkg.native.last_error()
    assert L.egtb_set_key_set_sizing(64.5, 0) != 0
    assert L.egtb_set_key_set_sizing(2.0, (1 << 34) + 1) != 0
    assert "max_slots" in pkg.native.last_error()
    assert L.egtb_set_key_set_sizing(2.0, 1 << 20) == 0
    assert L.egtb_set_key_set_sizing(0.0, 0) == 0  # back to the defaults
    a, b = ctypes.c_uint64(), ctypes.c_uint64()
    assert L.egtb_result_forward(None, ctypes.byref(a), ctypes.byref(b)) != 0


def test_every_sort_state_error_flag_is_read():
    """A build keeps one sort state per stream (the build context's and the
    forward pass's second-stream one), each with its own look-back error
    flag; every one of them must be read before the build returns, or a
    sort that hit its spin limit would pass unnoticed."""
    import re

    src = open(os.path.join(os.path.dirname(__file__), "..", "2048endgametablebase_amd", "csrc",
                            "egtb_runtime.hip")).read()
    start = src.index("struct BuildCtx {")
    end = src.index("\nvoid build_dispatch(", start)
    body = src[start:end]
    states = re.findall(r"\bSortWork\s+(\w+)\s*;", body)
    assert len(states) >= 2, states
    for name in states:
        assert f"{name}.read_error(" in body, name


def test_advanced_forward_arguments_fail_before_device_use(pkg):
    """The advanced forward pass checks its pattern and options on the host."""
    import importlib

    adv = importlib.import_module("2048endgametablebase_amd.advanced")
    s = adv.ad_spec("442", 7)
    for bad, msg in ((dict(symm_mode=5), "symm_mode"), (dict(target=0), "target"),
                     (dict(num_free_32k=-1), "n_free_32k"), (dict(fixed_32k_shifts=[3]), "nibble shifts"),
                     (dict(pattern_masks=list(range(1, 18))), "16 masks")):
        with pytest.raises(pkg.native.EgtbError, match=msg):
            adv.run_pattern_generate_ad(s["seeds"], {**s, **bad}, s["steps"])
    with pytest.raises(pkg.native.EgtbError, match="steps"):
        adv.run_pattern_generate_ad(s["seeds"], s, 1)
    with pytest.raises(pkg.native.EgtbError, match="radix_bits"):
        adv.run_pattern_generate_ad(s["seeds"], s, s["steps"], radix_bits=3)
    p = pkg.patterns.ad_params("t", 7)
    assert p["n_free_32k"] == 2 and p["fixed_32k_shifts"] == [0, 4, 16, 20]


def test_advanced_build_arguments_fail_before_device_use(pkg):
    """The advanced build (forward pass + solver) checks its solver options
    on the host before the forward pass starts."""
    import importlib

    adv = importlib.import_module("2048endgametablebase_amd.advanced")
    s = adv.ad_spec("442", 7)
    with pytest.raises(ValueError, match="success_rate_dtype"):
        adv.run_pattern_build_ad(s["seeds"], s, s["steps"], success_rate_dtype="float16")
    with pytest.raises(pkg.native.EgtbError, match="spawn_rate4"):
        adv.run_pattern_build_ad(s["seeds"], s, s["steps"], spawn_rate4=1.5)
    with pytest.raises(pkg.native.EgtbError, match="deletion_threshold"):
        adv.run_pattern_build_ad(s["seeds"], s, s["steps"], deletion_threshold=-0.1)
    with pytest.raises(pkg.native.EgtbError, match="3 steps"):
        adv.run_pattern_build_ad(s["seeds"], s, 2)
    with pytest.raises(pkg.native.EgtbError, match="symm_mode"):
        adv.run_pattern_build_ad(s["seeds"], {**s, "symm_mode": 7}, s["steps"])
    # the C entry point names an unknown rate type itself
    L = pkg.native.lib()
    seeds, keep, pat, opt = adv._ad_args(s["seeds"], s, s["steps"], None, False, 0, 0, 0, 0, 0)
    so = pkg.native.AdSolveOptions()
    so.success_rate_dtype = b"1-float32"
    so.spawn_rate4 = 0.1
    h = ctypes.c_void_p()
    assert L.egtb_ad_build(seeds.ctypes.data_as(pkg.native.u64p), seeds.size, ctypes.byref(pat), ctypes.byref(opt),
                           ctypes.byref(so), ctypes.byref(h)) != 0
    assert "success_rate_dtype" in pkg.native.last_error()
    assert L.egtb_ad_build(seeds.ctypes.data_as(pkg.native.u64p), seeds.size, ctypes.byref(pat), ctypes.byref(opt),
                           None, ctypes.byref(h)) != 0
    assert L.egtb_ad_result_book(None, 0, 0, None, None, 0, None, None) != 0


def test_tile_tuning_checks_arguments(pkg):
    """Tile tunings are validated on the host (no GPU touched)."""
    L = pkg.native.lib()
    u64 = ctypes.c_uint64 * 2
    i32 = ctypes.c_int32 * 2
    assert L.egtb_set_tile_tuning(2, u64(1, 2), i32(4, 8), 2) != 0
    assert L.egtb_set_tile_tuning(0, u64(1, 2), i32(4, 5), 2) != 0
    assert "keys per thread" in pkg.native.last_error()
    assert L.egtb_set_tile_tuning(1, u64(1, 2), i32(64, 100), 2) != 0
    assert "multiple of 64" in pkg.native.last_error()
    assert L.egtb_set_tile_tuning(0, u64(2, 1), i32(4, 8), 2) != 0
    assert "ascending" in pkg.native.last_error()
    assert L.egtb_set_tile_tuning(0, u64(1 << 20, 1 << 30), i32(4, 16), 2) == 0
    pkg.native._install_cached_tuning(L)  # back to the committed tables


def test_free_patterns_pin_no_cells(pkg):
    """A pattern without masks pins no cell (get_nibble_intersection of no
    masks is 0, Config.py:163-166): all its 32768 tiles are free ones."""
    cat = pkg.patterns.pattern_catalog
    assert cat["free8"].fixed_positions == () and cat["free8"].free_cells == 16
    p = pkg.patterns.ad_params("free8", 7)
    assert p["n_free_32k"] == 8 and p["fixed_32k_shifts"] == [] and p["is_free"]
    assert pkg.patterns.ad_params("free10", 7)["n_free_32k"] == 6
