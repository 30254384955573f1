"""Large-build machinery on small cases: forced tiny chunk budgets and
host-streamed tables give the reference's tables."""
import pytest

from conftest import book_sha, golden_params, load_golden

pytestmark = pytest.mark.gpu


def _check(res, tag):
    g = load_golden(tag)
    for layer in g["layers"]:
        keys, rates = res.book(layer["i"])
        assert book_sha(keys, rates) == layer["sha256"], (tag, layer["i"])


@pytest.mark.parametrize("tag,mkeys", [("442_128", 1), ("L3_128", 1), ("LL_64", 2)])
def test_chunked_forward_matches_golden(pkg, tag, mkeys):
    gp = golden_params(load_golden(tag))
    res = pkg.run_pattern_build(gp["seeds"], gp["masks"], gp["shifts"], gp["symm"], gp["target"], gp["steps"],
                                gp["docheck"], "", write_files=False, keep=True, chunk_mkeys=mkeys)
    try:
        mem = res.memory()
        print(f"case {tag} chunk {mkeys} Mi keys: {mem}", flush=True)
        assert mem["chunked_steps"] > 0
        if tag == "L3_128":
            assert mem["chunks"] > mem["chunked_steps"]  # some steps split into several chunks
        _check(res, tag)
    finally:
        res.close()


@pytest.mark.parametrize("tag", ["442_256", "444_64", "442_128_float64", "442_128_uint32"])
def test_host_tables_match_golden(pkg, tag):
    """Books streamed to host memory give the reference's bytes, for 4- and
    8-byte rates, and exactly the live records are copied."""
    gp = golden_params(load_golden(tag))
    res = pkg.run_pattern_build(gp["seeds"], gp["masks"], gp["shifts"], gp["symm"], gp["target"], gp["steps"],
                                gp["docheck"], "", write_files=False, keep=True, host_tables=True,
                                success_rate_dtype=gp["dtype"])
    try:
        mem = res.memory()
        print(f"case host tables {tag}: {mem}", flush=True)
        rec = 8 + {"float32": 4, "uint32": 4, "float64": 8, "uint64": 8}[gp["dtype"]]
        assert mem["streamed_bytes"] == rec * sum(res.layer_info(i)[1] for i in range(res.steps))
        _check(res, tag)
    finally:
        res.close()


def _tile_sum_table():
    import numpy as np

    t = np.zeros(1 << 16, dtype=np.int64)
    for x in range(1 << 16):
        t[x] = sum(1 << v for v in ((x >> 0) & 15, (x >> 4) & 15, (x >> 8) & 15, (x >> 12) & 15) if v)
    return t


def _mirror_lr(k):
    import numpy as np

    return (((k & np.uint64(0xF000F000F000F000)) >> np.uint64(12)) | ((k & np.uint64(0x0F000F000F000F00)) >> np.uint64(4))
            | ((k & np.uint64(0x00F000F000F000F0)) << np.uint64(4)) | ((k & np.uint64(0x000F000F000F000F)) << np.uint64(12)))


def test_444_2048_in_hbm_and_host_tables_agree(pkg, orc):
    """The 12-free-cell 444@2048 build at full size on one GPU (7.1e9
    states, 1084 layers).  Every layer of the build kept in HBM holds: keys
    strictly ascending and in horizontal canonical form (layer 0: the seeds
    as given), in the pattern (bottom row pinned), of tile sum tile_sum + 2i (on
    every 8th record), and float32 rates in (0, 1] with every success board
    of a layer past docheck at 1.  Five layers' live records are solved
    again by the CPU oracle from the stored books i+1 and i+2 and must match
    bit for bit.  Streamed to host memory the tables hold the same records
    (five layers compared in full, every layer's counts)."""
    import numpy as np

    p = pkg.patterns.build_params("444", 11)
    tile_sum = pkg.patterns.pattern_catalog["444"].tile_sum
    tsum = _tile_sum_table()
    picks = [40, 300, 540, 800, 1060]
    r = pkg.build_in_memory("444", 11)
    try:
        info = [r.layer_info(i) for i in range(r.steps)]
        print(f"case 444@2048 in HBM: {r.steps} layers, {sum(x[0] for x in info)} states", flush=True)
        sha = {}
        solve_picks = []
        for i in range(r.steps):
            keys, rates = r.book(i)
            assert keys.size == info[i][1], i
            if keys.size == 0:
                continue
            if i > 0:
                assert bool(np.all(keys[1:] > keys[:-1])), f"layer {i} not strictly ascending"
            assert bool(np.all((keys & np.uint64(0xFFFF)) == np.uint64(0xFFFF))), f"layer {i}: off the pattern"
            sub = keys[::8]
            if i > 0:  # layer 0 holds the seeds as given
                assert bool(np.all(sub <= _mirror_lr(sub))), f"layer {i}: not in horizontal canonical form"
            ts = sum(tsum[((sub >> np.uint64(16 * q)) & np.uint64(0xFFFF)).astype(np.int64)] for q in range(4))
            assert bool(np.all(ts == tile_sum + 2 * i)), f"layer {i}: tile sums"
            assert bool(np.all((rates > 0) & (rates <= 1))), f"layer {i}: rates outside (0, 1]"
            if i > p["docheck_step"]:
                succ = (((keys >> np.uint64(20)) & np.uint64(15)) == 11) | (((keys >> np.uint64(24)) & np.uint64(15)) == 11)
                assert bool(np.all(rates[succ] == 1.0)), f"layer {i}: success boards below 1"
            if i in picks:
                sha[i] = book_sha(keys, rates)
            if 100_000 < keys.size < 2_000_000 and i < r.steps - 2:
                solve_picks.append(i)
        solve_picks = [solve_picks[j] for j in np.linspace(0, len(solve_picks) - 1, 5).astype(int)]
        print(f"case 444@2048 local recompute of layers {solve_picks}", flush=True)
        assert len(set(solve_picks)) == 5
        for i in solve_picks:
            keys, rates = r.book(i)
            got = orc.solve_layer(keys, r.book(i + 1), r.book(i + 2), p["masks"], p["shifts"], p["symm_mode"], 11,
                                  i > p["docheck_step"], threads=16)
            assert np.array_equal(got.view(np.uint32), rates.view(np.uint32)), f"layer {i}: recomputed rates"
    finally:
        r.close()
    h = pkg.build_in_memory("444", 11, host_tables=True)
    try:
        print(f"case 444@2048 host tables: {h.memory()}", flush=True)
        assert [h.layer_info(i) for i in range(h.steps)] == info
        for i in picks:
            assert book_sha(*h.book(i)) == sha[i], i
        assert h.memory()["streamed_bytes"] == 12 * sum(x[1] for x in info)
    finally:
        h.close()
