"""GPU parity of the plain build: the HIP path (through the C ABI) against
the CPU oracle (bit-exact raw layers, solved keys and float32 rate bits) and
against golden digests of the reference's own compiled builder."""
import ctypes
import hashlib
import os

import numpy as np
import pytest

from conftest import book_sha, golden_params, golden_tags, load_golden

pytestmark = pytest.mark.gpu


def _build(pkg, gp, dtype="float32", keep_raw=False, radix_bits=0):
    return pkg.run_pattern_build(gp["seeds"], gp["masks"], gp["shifts"], gp["symm"], gp["target"], gp["steps"],
                                 gp["docheck"], "", success_rate_dtype=dtype, write_files=False, keep=True,
                                 keep_raw=keep_raw, radix_bits=radix_bits)


@pytest.mark.parametrize("pattern,target", [("L3", 7), ("442", 7), ("LL", 6), ("444", 6), ("t", 6)])
def test_layers_match_oracle(pkg, orc, pattern, target):
    p = pkg.patterns.build_params(pattern, target)
    gp = dict(seeds=p["seeds"], masks=p["masks"], shifts=p["shifts"], symm=p["symm_mode"], target=target,
              steps=p["steps"], docheck=p["docheck_step"])
    print(f"case {pattern}@{2**target}", flush=True)
    res = _build(pkg, gp, keep_raw=True)
    ob = orc.OracleBuild(gp["seeds"], gp["masks"], gp["shifts"], gp["symm"], target, gp["steps"], gp["docheck"])
    try:
        for i in range(gp["steps"]):
            assert np.array_equal(res.raw(i), ob.raw(i)), f"raw layer {i}"
            n_raw, n_live, c2, c4 = res.layer_info(i)
            oc2, oc4 = ob.candidates(i)
            if 0 < i <= gp["steps"] - 2:
                assert (c2, c4) == (oc2, oc4), f"candidate counts of step {i}"
            gk, gr = res.book(i)
            ok, orr = ob.book(i)
            assert np.array_equal(gk, ok), f"book keys {i}"
            assert np.array_equal(gr.view(np.uint32), orr.view(np.uint32)), f"book rates {i}"
            assert n_live == gk.size
    finally:
        res.close()
        ob.close()


def file_build(pkg, gp, prefix, **extra):
    pkg.run_pattern_build(gp["seeds"], gp["masks"], gp["shifts"], gp["symm"], gp["target"], gp["steps"],
                          gp["docheck"], prefix, success_rate_dtype=gp["dtype"], write_files=True,
                          deletion_threshold=gp["options"].get("deletion_threshold", 0.0),
                          optimal_branch_only=gp["options"].get("optimal_branch_only", False), **extra)


def file_digest_mismatches(g, prefix):
    bad = []
    for layer in g["layers"]:
        path = prefix + f"{layer['i']}.book"
        raw = open(path, "rb").read() if os.path.exists(path) else b""
        if hashlib.sha256(raw).hexdigest() != layer["sha256"]:
            bad.append(layer["i"])
    return bad


@pytest.mark.parametrize("tag", golden_tags())
def test_matches_reference_golden(pkg, tag, tmp_path):
    """Every solved layer's packed bytes hash to the reference's .book file.
    Runs with post-solve pruning (deletion_threshold, optimal_branch_only)
    act on the files, so those are compared as written; an optimal-branch
    run without a threshold is also checked on the books kept in HBM."""
    g = load_golden(tag)
    gp = golden_params(g)
    print(f"case golden {tag}", flush=True)
    if gp["options"]:
        prefix = str(tmp_path / f"{tag}_")
        file_build(pkg, gp, prefix)
        bad = file_digest_mismatches(g, prefix)
        assert not bad, f"layers differing from the reference: {bad[:10]}"
        assert not os.path.exists(prefix + "optlayer")
        if gp["options"].get("deletion_threshold"):
            return
        res = pkg.run_pattern_build(gp["seeds"], gp["masks"], gp["shifts"], gp["symm"], gp["target"], gp["steps"],
                                    gp["docheck"], "", success_rate_dtype=gp["dtype"], write_files=False,
                                    keep=True, optimal_branch_only=True)
    else:
        res = _build(pkg, gp, dtype=gp["dtype"])
    try:
        bad = []
        for layer in g["layers"]:
            keys, rates = res.book(layer["i"])
            if keys.size != layer["live"] or book_sha(keys, rates) != layer["sha256"]:
                bad.append(layer["i"])
            n_raw = res.layer_info(layer["i"])[0]
            if layer["n"] is not None:
                assert n_raw == layer["n"], f"raw size of layer {layer['i']}"
        assert not bad, f"layers differing from the reference: {bad[:10]}"
    finally:
        res.close()


@pytest.mark.parametrize("bits,schedule", [(b, s) for s in (1, 2) for b in (4, 6, 8, 11)] + [(0, 3)])
def test_sort_tunings_agree(pkg, bits, schedule):
    """Digit width and sort schedule are tuning knobs: every choice gives the
    reference's table."""
    p = pkg.patterns.build_params("442", 7)
    gp = dict(seeds=p["seeds"], masks=p["masks"], shifts=p["shifts"], symm=p["symm_mode"], target=7,
              steps=p["steps"], docheck=p["docheck_step"])
    g = load_golden("442_128")
    res = pkg.run_pattern_build(gp["seeds"], gp["masks"], gp["shifts"], gp["symm"], gp["target"], gp["steps"],
                                gp["docheck"], "", write_files=False, keep=True, radix_bits=bits,
                                sort_schedule=schedule)
    try:
        for layer in g["layers"]:
            keys, rates = res.book(layer["i"])
            assert book_sha(keys, rates) == layer["sha256"], f"layer {layer['i']}, {bits}-bit, schedule {schedule}"
    finally:
        res.close()


@pytest.mark.parametrize("items", [4, 8, 16])
def test_radix_tile_sizes_agree(pkg, items):
    """The per-digit schedule's tile size (keys per thread) is an autotuner
    knob: every size gives the reference's table."""
    p = pkg.patterns.build_params("442", 7)
    g = load_golden("442_128")
    pkg.native.install_tile_tuning(0, [{"max_n": (1 << 64) - 1, "value": items}])
    try:
        res = pkg.run_pattern_build(p["seeds"], p["masks"], p["shifts"], p["symm_mode"], 7, p["steps"],
                                    p["docheck_step"], "", write_files=False, keep=True, radix_bits=8,
                                    sort_schedule=1, forward_hash_steps=-1)
        try:
            for layer in g["layers"]:
                keys, rates = res.book(layer["i"])
                assert book_sha(keys, rates) == layer["sha256"], f"layer {layer['i']}, {items} keys per thread"
        finally:
            res.close()
    finally:
        pkg.native._install_cached_tuning(pkg.native.lib())


def test_start_build_writes_reference_layout(pkg, tmp_path):
    """start_build writes config.txt, stats.txt and every <i>.book with the
    reference's bytes (442 at 128, float32)."""
    cfg = pkg.config.SingletonConfig().config
    old = cfg["success_rate_dtype"]
    cfg["success_rate_dtype"] = "float32"
    try:
        prefix = str(tmp_path / "442_128_")
        assert pkg.start_build("442", 7, prefix) is True
    finally:
        cfg["success_rate_dtype"] = old
    g = load_golden("442_128")
    for layer in g["layers"]:
        data = open(prefix + f"{layer['i']}.book", "rb").read()
        assert hashlib.sha256(data).hexdigest() == layer["sha256"], layer["i"]
    assert not [f for f in os.listdir(tmp_path) if f.endswith(".tmp")]
    stats = open(prefix + "stats.txt").read().splitlines()
    assert stats[0] == "layer,length,max success rate,speed,deletion_threshold,time"
    rows = [r.split(",") for r in stats[1:]]
    assert [int(r[0]) for r in rows] == list(range(g["steps"] - 3, -1, -1))[: len(rows)]
    by_layer = {int(r[0]): int(r[1]) for r in rows}
    for layer in g["layers"]:
        if layer["n"] is not None and layer["live"]:
            assert by_layer[layer["i"]] == layer["n"]
    cfg_txt = open(prefix + "config.txt").read()
    assert "success_rate_dtype: float32" in cfg_txt and "4_spawn_rate: 0.1" in cfg_txt


def test_start_build_default_dtype_is_uint32(pkg, tmp_path):
    """With the default configuration (success_rate_dtype uint32, the
    reference's default) start_build writes the reference's uint32 books."""
    cfg = pkg.config.SingletonConfig().config
    assert cfg["success_rate_dtype"] == "uint32"
    prefix = str(tmp_path / "442_128_")
    assert pkg.start_build("442", 7, prefix) is True
    g = load_golden("442_128_uint32")
    for layer in g["layers"]:
        data = open(prefix + f"{layer['i']}.book", "rb").read()
        assert hashlib.sha256(data).hexdigest() == layer["sha256"], layer["i"]


def test_dedup_primitives_match_numpy(pkg):
    """sort_uint64 / unique_sorted_u64_inplace (bookgen_native.so names)."""
    L = pkg.native.lib()
    rng = np.random.default_rng(0)
    for n in [0, 1, 2, 1000, 4097, 300001]:
        a = rng.integers(0, 2**63, size=n, dtype=np.uint64) | np.uint64(1 << 63) * (rng.integers(0, 2, n) > 0)
        a = np.concatenate([a, a[: n // 3]])
        want = np.sort(a)
        L.sort_uint64(a.ctypes.data_as(pkg.native.u64p), a.size, False)
        assert np.array_equal(a, want)
        m = L.unique_sorted_u64_inplace(a.ctypes.data_as(pkg.native.u64p), a.size)
        assert np.array_equal(a[:m], np.unique(want))
    b = rng.integers(0, 2**40, size=5000, dtype=np.uint64)
    L.sort_uint64(b.ctypes.data_as(pkg.native.u64p), b.size, True)
    assert np.array_equal(b, np.sort(b)[::-1])


def _dedup_cases():
    rng = np.random.default_rng(5)

    def boards(n):
        # nibble-structured keys like real layers: small tile values, pinned 0xF cells
        v = rng.integers(0, 10, size=(n, 10), dtype=np.uint64)
        k = np.full(n, 0xFFFFFF, dtype=np.uint64)
        for j in range(10):
            k |= v[:, j] << np.uint64(24 + 4 * j)
        return k

    out = {}
    for n in [1, 2, 3, 100, 2047, 2048, 2049, 5000, 8193, 100_000]:
        a = rng.integers(0, 2**63, size=n, dtype=np.uint64)
        out[f"random_{n}"] = np.concatenate([a, a[: n // 3]])
    out["boards_1M"] = boards(1_000_000)
    out["boards_max"] = boards(1 << 21)          # the largest sample-sort input
    out["boards_over"] = boards((1 << 21) + 7)   # falls back to the radix sort
    heavy = boards(60_000)
    heavy[:30_000] = heavy[0]                    # one bucket far past the LDS capacity
    out["heavy_duplicates"] = heavy
    out["all_equal"] = np.full(5000, 0x1234, dtype=np.uint64)
    skew = boards(300_000)
    skew[::2] = skew[::2] & np.uint64(0xFFFFFFFFFF)  # half the keys in a narrow range
    out["skewed"] = skew
    return out


@pytest.mark.parametrize("schedule", [0, 1, 2, 3])
def test_sort_unique_matches_numpy(pkg, schedule):
    """egtb_sort_unique_u64 (the builder's layer dedup) under every schedule."""
    L = pkg.native.lib()
    for name, keys in _dedup_cases().items():
        print(f"case dedup {name} schedule {schedule}", flush=True)
        keys = np.ascontiguousarray(keys)
        out = np.empty(keys.size, dtype=np.uint64)
        m = ctypes.c_size_t()
        pkg.native.check(L.egtb_sort_unique_u64(keys.ctypes.data_as(pkg.native.u64p), keys.size, schedule,
                                                out.ctypes.data_as(pkg.native.u64p), ctypes.byref(m)))
        want = np.unique(keys)
        assert m.value == want.size, name
        assert np.array_equal(out[: m.value], want), name


@pytest.mark.parametrize("solve_hash_mb", [-1, 1, 100000])
def test_solve_lookup_paths_agree(pkg, solve_hash_mb):
    """Successor lookups by hash, by bucket-indexed binary search, or mixed
    per layer (a 1 MiB limit) give the reference's table."""
    g = load_golden("442_128")
    gp = golden_params(g)
    res = pkg.run_pattern_build(gp["seeds"], gp["masks"], gp["shifts"], gp["symm"], gp["target"], gp["steps"],
                                gp["docheck"], "", write_files=False, keep=True, solve_hash_mb=solve_hash_mb)
    try:
        for layer in g["layers"]:
            keys, rates = res.book(layer["i"])
            assert book_sha(keys, rates) == layer["sha256"], f"layer {layer['i']}, solve_hash_mb {solve_hash_mb}"
    finally:
        res.close()


@pytest.mark.parametrize("tag,layout", [("442_256", 1), ("442_256", 0), ("LL_64", 1), ("444_64", 1),
                                        ("L3_128_uint32", 1), ("t_64", 1)])
def test_solve_hash_layouts(pkg, tag, layout):
    """Compact (8-byte slots keyed by each layer's key span) and wide
    (16-byte slots) successor hashes give the reference's tables, in the
    two-stream in-HBM backward pass and, with files, the single-stream one;
    the compact layout covers 4-byte rates on 10- and 12-free-cell patterns
    with identity, diagonal and horizontal symmetry."""
    g = load_golden(tag)
    gp = golden_params(g)
    res = pkg.run_pattern_build(gp["seeds"], gp["masks"], gp["shifts"], gp["symm"], gp["target"], gp["steps"],
                                gp["docheck"], "", success_rate_dtype=gp["dtype"], write_files=False, keep=True,
                                solve_hash_layout=layout)
    try:
        for layer in g["layers"]:
            keys, rates = res.book(layer["i"])
            assert book_sha(keys, rates) == layer["sha256"], f"{tag} layer {layer['i']}, layout {layout}"
    finally:
        res.close()


def test_optimal_branch_pass_resumes(pkg, tmp_path):
    """An optimal-branch pass cut short (its place in <pathname>optlayer, the
    books above it still unfiltered) finishes on a rerun with the
    reference's files; an unfinished pass of a threshold build likewise."""
    for tag in ("442_128_opt", "442_128_uint32_dt0.3_opt"):
        g = load_golden(tag)
        gp = golden_params(g)
        plain = dict(gp, options={"deletion_threshold": gp["options"].get("deletion_threshold", 0.0)})
        a = str(tmp_path / f"plain_{tag}_")
        b = str(tmp_path / f"opt_{tag}_")
        file_build(pkg, plain, a)
        file_build(pkg, gp, b)
        assert not file_digest_mismatches(g, b)
        # interrupt after layer 40: books 41.. as the solve left them
        for i in range(41, gp["steps"]):
            with open(a + f"{i}.book", "rb") as src, open(b + f"{i}.book", "wb") as dst:
                dst.write(src.read())
        with open(b + "optlayer", "w") as fh:
            fh.write("40")
        assert file_digest_mismatches(g, b)
        file_build(pkg, gp, b)
        assert not file_digest_mismatches(g, b), tag
        assert not os.path.exists(b + "optlayer")


def test_concurrent_builds_on_own_streams(pkg):
    """Builds from two host threads, each on streams of its own, overlap on
    the device and each gives the reference's table."""
    import threading

    runs = [("442", 7, "442_128"), ("L3", 7, "L3_128"), ("442", 7, "442_128"), ("LL", 6, "LL_64")]
    errors = []

    def work(pattern, target, tag):
        try:
            res = pkg.build_in_memory(pattern, target, own_streams=True)
            try:
                for layer in load_golden(tag)["layers"]:
                    keys, rates = res.book(layer["i"])
                    assert book_sha(keys, rates) == layer["sha256"], (tag, layer["i"])
            finally:
                res.close()
        except Exception as exc:  # reported by the main thread
            errors.append(exc)

    print("case concurrent builds 442@128, L3@128, 442@128, LL@64 on own streams", flush=True)
    threads = [threading.Thread(target=work, args=r) for r in runs]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
