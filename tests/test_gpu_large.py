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


def test_444_2048_in_hbm_and_host_tables_agree(pkg):
    """The 12-free-cell 444@2048 build at full size on one GPU: its layers
    come out sorted and unique, and streamed to host memory they hold the
    same records as kept in HBM. (Large layers cross the raw arena's 1 GiB
    chunks here, which no smaller case does.)"""
    import numpy as np
    picks = [40, 300, 540, 800, 1060]
    r = pkg.build_in_memory("444", 11)
    try:
        info = [r.layer_info(i) for i in range(r.steps)]
        print(f"case 444@2048 in HBM: {r.steps} layers, {sum(x[0] for x in info)} states", flush=True)
        sha = {}
        for i in picks:
            keys, rates = r.book(i)
            assert keys.size > 1 and bool(np.all(keys[1:] > keys[:-1])), i
            sha[i] = book_sha(keys, rates)
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
