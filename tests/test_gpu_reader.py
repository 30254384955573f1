"""GPU reader / lookup parity: BookReader.move_on_dic and batched table
lookups over HBM tables against the reference's own compiled reader (the
fixtures in tests/golden/reader_*.json) and against the CPU restatement of
that reader (on-disk binary search, itself pinned to the same fixtures);
the off-table fallback against the expectimax player, and table sources
agreeing."""
import importlib
import math
import os

import numpy as np
import pytest

from conftest import load_reader_fixture, reader_tags
from oracle import oracle as orc
from oracle import reader_oracle as ro

pytestmark = pytest.mark.gpu

PATTERN, TARGET = "442", 7


def decode_values(b):
    m = np.zeros((4, 4), dtype=np.int64)
    for r in range(4):
        for c in range(4):
            e = (b >> (4 * ((3 - r) * 4 + (3 - c)))) & 0xF
            m[r, c] = 0 if e == 0 else 1 << e
    return m


@pytest.fixture(scope="module")
def built(pkg, tmp_path_factory):
    cfg = pkg.config.SingletonConfig().config
    old = cfg["success_rate_dtype"]
    cfg["success_rate_dtype"] = "float32"
    folder = tmp_path_factory.mktemp("tables")
    prefix = str(folder / f"{PATTERN}_{2**TARGET}_")
    try:
        pkg.start_build(PATTERN, TARGET, prefix)
    finally:
        cfg["success_rate_dtype"] = old
    cfg["filepath_map"][(f"{PATTERN}_{2**TARGET}", cfg["4_spawn_rate"])] = [(str(folder), "float32")]
    yield str(folder), prefix
    pkg.BookReader.clear()
    cfg["filepath_map"].clear()


def sample_boards(prefix, steps, n, seed):
    """Positions to move from: stored states plus one spawned tile, some
    rotated/mirrored (other symmetry ops), some with a pinned cell broken."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        i = int(rng.integers(1, steps - 2))
        keys, _ = orc.read_book(prefix + f"{i}.book")
        if keys.size == 0:
            continue
        b = int(keys[rng.integers(0, keys.size)])
        empt = [k for k in range(16) if (b >> (4 * k)) & 0xF == 0]
        if empt:
            k = empt[rng.integers(0, len(empt))]
            b |= (1 if rng.random() < 0.9 else 2) << (4 * k)
        m = decode_values(b)
        kind = rng.integers(0, 4)
        if kind == 1:
            m = np.rot90(m, int(rng.integers(1, 4)))
        elif kind == 2:
            m = np.fliplr(m)
        elif kind == 3:
            m = m.copy()
            m[3, 3] = 2  # break a pinned 32768
        out.append(np.ascontiguousarray(m))
    return out


def test_move_on_dic_matches_reference_reader(pkg, built):
    folder, prefix = built
    meta = pkg.patterns.pattern_catalog[PATTERN]
    boards = sample_boards(prefix, pkg.patterns.build_params(PATTERN, TARGET)["steps"], 600, 1)
    n_numeric = 0
    for m in boards:
        want = ro.move_on_dic(m, folder, f"{PATTERN}_{2**TARGET}", meta.nums_adjust, list(meta.pattern_masks),
                              meta.symm_mode)
        got = pkg.BookReader.move_on_dic(m, PATTERN, str(2**TARGET), f"{PATTERN}_{2**TARGET}")
        assert got == want, (m.tolist(), got, want)
        assert list(got[0].keys()) == list(want[0].keys())
        n_numeric += want[1] != ""
    assert n_numeric > 300


def test_unknown_paths_and_negative_layers(pkg, built):
    m = np.zeros((4, 4), dtype=np.int64)
    assert pkg.BookReader.move_on_dic(m, PATTERN, "128", "442_128") == (
        {"down": "", "right": "", "left": "", "up": ""}, "")
    assert pkg.BookReader.move_on_dic(m, PATTERN, "128", "442_999") == ({"?": "?"}, "")


def test_fallback_and_table_sources_agree(pkg, built):
    folder, prefix = built
    res = pkg.build_in_memory(PATTERN, TARGET)
    t_mem = pkg.Table.from_result(res, PATTERN, TARGET)
    t_file = pkg.Table.open(prefix, PATTERN, TARGET, "float32")
    rng = np.random.default_rng(7)
    boards = rng.integers(0, 2**63, size=4000, dtype=np.uint64)
    boards &= np.uint64(0x7777777777777777)  # small tiles: mostly off-pattern
    keys, _ = res.book(40)
    boards[:2000] = keys[rng.integers(0, keys.size, 2000)] | np.uint64(1 << 60)
    a, sa, oa = t_mem.lookup_f64(boards)
    b, sb, ob = t_file.lookup_f64(boards)
    assert np.array_equal(sa, sb) and np.array_equal(oa, ob)
    assert np.array_equal(np.nan_to_num(a, nan=-1), np.nan_to_num(b, nan=-1))
    f32, s32 = t_mem.lookup(boards)
    assert np.array_equal(s32, sa)
    assert np.array_equal(np.nan_to_num(f32, nan=-1), np.nan_to_num(a.astype(np.float32), nan=-1))
    # off-table boards carry the expectimax player's depth-1 scores (the
    # batched search, which equals the reference's AIPlayer), NaN where the
    # move does not change the board
    fb = np.nonzero(sa == 1)[0]
    assert fb.size > 100
    ai = importlib.import_module("2048endgametablebase_amd.ai")
    _best, scores = ai.search(boards[fb], 1)
    for j, i in enumerate(fb):
        b = int(boards[i])
        want = [math.nan if orc.move(b, d) == b else float(scores[j][d]) for d in range(4)]
        assert np.array_equal(np.nan_to_num(a[i], nan=-1e18), np.nan_to_num(want, nan=-1e18)), hex(b)
    t_mem.close()
    t_file.close()
    res.close()


@pytest.mark.parametrize("threads", [64, 256, 1024])
def test_lookup_block_sizes_agree(pkg, built, threads):
    """The lookup's threads per block are an autotuner knob: every size gives
    the same answers (on-table rows, off-table player scores, statuses)."""
    res = pkg.build_in_memory(PATTERN, TARGET)
    t = pkg.Table.from_result(res, PATTERN, TARGET)
    try:
        rng = np.random.default_rng(11)
        keys, _ = res.book(30)
        boards = np.concatenate([keys[rng.integers(0, keys.size, 3000)] | np.uint64(1 << 60),
                                 rng.integers(0, 2**63, size=1000, dtype=np.uint64) & np.uint64(0x7777777777777777)])
        a, sa, oa = t.lookup_f64(boards)
        pkg.native.install_tile_tuning(1, [{"max_n": (1 << 64) - 1, "value": threads}])
        try:
            b, sb, ob = t.lookup_f64(boards)
        finally:
            pkg.native._install_cached_tuning(pkg.native.lib())
        assert np.array_equal(sa, sb) and np.array_equal(oa, ob)
        assert np.array_equal(np.nan_to_num(a, nan=-1), np.nan_to_num(b, nan=-1))
    finally:
        t.close()
        res.close()


def test_prefer_max_matches_reference_reader(pkg, built):
    """prefer_max (the reference reader's rule for 4442f/4442ff/4tiler): all
    ops are tried and the one with the largest rounded best rate answers."""
    folder, prefix = built
    meta = pkg.patterns.pattern_catalog[PATTERN]
    t = pkg.Table.open(prefix, PATTERN, TARGET, "float32", prefer_max=True)
    boards = sample_boards(prefix, pkg.patterns.build_params(PATTERN, TARGET)["steps"], 400, 5)
    reader_mod = importlib.import_module("2048endgametablebase_amd.BookReader")
    enc = [reader_mod.encode_board(m) for m in boards]
    got = t.move_on_dic_batch(enc)
    differs_from_first = 0
    for m, g in zip(boards, got):
        want = ro.move_on_dic(m, folder, f"{PATTERN}_{2**TARGET}", meta.nums_adjust, list(meta.pattern_masks),
                              meta.symm_mode, prefer_max=True)
        assert g == want, (m.tolist(), g, want)
        first = ro.move_on_dic(m, folder, f"{PATTERN}_{2**TARGET}", meta.nums_adjust, list(meta.pattern_masks),
                               meta.symm_mode)
        differs_from_first += first != want
    assert differs_from_first > 0  # the rule is exercised
    t.close()


@pytest.fixture(scope="module")
def built_u32(pkg, tmp_path_factory):
    """The same tables written by start_build with the default configuration
    (uint32 rates, the reference's default)."""
    cfg = pkg.config.SingletonConfig().config
    assert cfg["success_rate_dtype"] == "uint32"
    folder = tmp_path_factory.mktemp("tables_u32")
    prefix = str(folder / f"{PATTERN}_{2**TARGET}_")
    pkg.start_build(PATTERN, TARGET, prefix)
    yield str(folder), prefix


def test_move_on_dic_uint32_tables(pkg, built_u32):
    """Readers over uint32 files: rates scaled by 4e9 and rounded as the
    reference reader does, both through BookReader and find_classic_value."""
    folder, prefix = built_u32
    meta = pkg.patterns.pattern_catalog[PATTERN]
    t = pkg.Table.open(prefix, PATTERN, TARGET, "uint32")
    reader_mod = importlib.import_module("2048endgametablebase_amd.BookReader")
    boards = sample_boards(prefix, pkg.patterns.build_params(PATTERN, TARGET)["steps"], 300, 9)
    got = t.move_on_dic_batch([reader_mod.encode_board(m) for m in boards])
    n_numeric = 0
    for m, g in zip(boards, got):
        want = ro.move_on_dic(m, folder, f"{PATTERN}_{2**TARGET}", meta.nums_adjust, list(meta.pattern_masks),
                              meta.symm_mode, dtype="uint32")
        assert g == want, (m.tolist(), g, want)
        n_numeric += want[1] != ""
    assert n_numeric > 150
    t.close()
    keys, _ = orc.read_book(prefix + "40.book")
    for k in keys[:: max(1, keys.size // 50)].tolist():
        want = ro.find_in_file(prefix + "40.book", k, *ro.REC_FORMAT["uint32"])
        assert pkg.find_classic_value(folder, os.path.basename(prefix) + "40.book", k, "uint32") == want


# ---- against the reference's own reader (tests/golden/reader_*.json) --------

PREFER_MAX_NAMES = ("4442ff", "4442f", "4tiler")


def same_answers(fx, got):
    bad = []
    for (m, want), g in zip(fx["expected"], got):
        if g != want or list(g[0].keys()) != list(want[0].keys()):
            bad.append((m.tolist(), g, want))
    return bad


@pytest.mark.parametrize("tag", reader_tags())
def test_lookup_matches_reference_reader_fixture(pkg, tag):
    """Every fixture position through k_lookup on a table built on the GPU:
    the same ordered dict and dtype as the reference's reader gave."""
    fx = load_reader_fixture(tag)
    print(f"reader fixture {tag}: {fx['pattern']}@{2 ** fx['target']} {fx['dtype']}", flush=True)
    res = pkg.build_in_memory(fx["pattern"], fx["target"], success_rate_dtype=fx["dtype"])
    t = pkg.Table.from_result(res, fx["pattern"], fx["target"])
    try:
        if fx["reader_name"] in PREFER_MAX_NAMES:
            t.set_prefer_max(True)
        reader_mod = importlib.import_module("2048endgametablebase_amd.BookReader")
        got = t.move_on_dic_batch([reader_mod.encode_board(m) for m, _ in fx["expected"]])
        bad = same_answers(fx, got)
        assert not bad, (len(bad), bad[:3])
    finally:
        t.close()
        res.close()


def test_book_reader_matches_reference_reader_fixture(pkg, built):
    """BookReader.move_on_dic over the files start_build wrote (float32)."""
    fx = load_reader_fixture("442_128")
    got = [pkg.BookReader.move_on_dic(m, PATTERN, str(2 ** TARGET), fx["pattern_full"]) for m, _ in fx["expected"]]
    bad = same_answers(fx, got)
    assert not bad, (len(bad), bad[:3])


def test_book_reader_uint32_matches_reference_reader_fixture(pkg, built_u32):
    """The same over the default configuration's uint32 files."""
    folder, prefix = built_u32
    fx = load_reader_fixture("442_128_uint32")
    cfg = pkg.config.SingletonConfig().config
    key = (fx["pattern_full"], cfg["4_spawn_rate"])
    old = cfg["filepath_map"].get(key)
    cfg["filepath_map"][key] = [(folder, "uint32")]
    try:
        got = [pkg.BookReader.move_on_dic(m, PATTERN, str(2 ** TARGET), fx["pattern_full"])
               for m, _ in fx["expected"]]
    finally:
        if old is None:
            cfg["filepath_map"].pop(key, None)
        else:
            cfg["filepath_map"][key] = old
    bad = same_answers(fx, got)
    assert not bad, (len(bad), bad[:3])


def test_compressed_tables_serve_the_reference_answers(pkg, tmp_path):
    """A 442@512 build with compress on: every book above 2 MiB (but layers
    0-2) is replaced by its .z directory, each layer decodes to the
    reference's .book bytes, and a table opened over the compressed files
    answers the reference reader's 442@512 fixture positions exactly."""
    import ctypes

    from conftest import book_sha, load_golden

    g = load_golden("442_512")
    p = pkg.patterns.build_params("442", 9)
    prefix = str(tmp_path / "442_512_")
    print("case compressed 442@512", flush=True)
    pkg.run_pattern_build(p["seeds"], p["masks"], p["shifts"], p["symm_mode"], 9, p["steps"], p["docheck_step"],
                          prefix, success_rate_dtype="float32", write_files=True, compress=True)
    L = pkg.native.lib()
    n_z = 0
    for layer in g["layers"]:
        i = layer["i"]
        book, zdir = prefix + f"{i}.book", prefix + f"{i}.z"
        if os.path.isdir(zdir):
            assert not os.path.exists(book) and i > 2
            n = ctypes.c_size_t()
            pkg.native.check(L.egtb_zbook_read(book.encode(), 4, None, None, 0, ctypes.byref(n)))
            keys = np.empty(n.value, dtype=np.uint64)
            rates = np.empty(n.value, dtype=np.float32)
            pkg.native.check(L.egtb_zbook_read(book.encode(), 4, keys.ctypes.data_as(pkg.native.u64p),
                                               rates.ctypes.data, n.value, ctypes.byref(n)))
            assert book_sha(keys, rates) == layer["sha256"], i
            n_z += 1
        else:
            assert os.path.getsize(book) <= 2 * 1024 * 1024 or i <= 2, i
    assert n_z > 100
    fx = load_reader_fixture("442_512")
    t = pkg.Table.open(prefix, "442", 9, "float32")
    try:
        got = t.move_on_dic_batch([importlib.import_module("2048endgametablebase_amd.BookReader").encode_board(m)
                                   for m, _ in fx["expected"]])
        bad = same_answers(fx, got)
        assert not bad, (len(bad), bad[:3])
    finally:
        t.close()
