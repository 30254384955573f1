"""Forward dedup in key sets against the sort + unique forward pass: the
same raw layers, layer sizes, candidate counts and books, also when the
sets overflow and are regrown, when the sort loop takes over part way, and
under an explicit sort schedule; and the reference's books on a free
pattern."""
import numpy as np
import pytest

from conftest import book_sha, golden_params, load_golden

pytestmark = pytest.mark.gpu


def _build(pkg, gp, **kw):
    return pkg.run_pattern_build(gp["seeds"], gp["masks"], gp["shifts"], gp["symm"], gp["target"], gp["steps"],
                                 gp["docheck"], "", write_files=False, keep=True, keep_raw=True,
                                 success_rate_dtype=gp["dtype"], **kw)


def _same(a, b, tag):
    assert a.steps == b.steps
    for i in range(a.steps):
        assert a.layer_info(i) == b.layer_info(i), (tag, i)
        assert np.array_equal(a.raw(i), b.raw(i)), (tag, i)
        ka, ra = a.book(i)
        kb, rb = b.book(i)
        assert np.array_equal(ka, kb) and ra.tobytes() == rb.tobytes(), (tag, i)


@pytest.mark.parametrize("tag,kw,max_slots,expect", [
    ("442_128", {}, 0, "all"),
    ("LL_64", {}, 1 << 26, "all"),  # layers of up to 4 Mi states: sets above the default limit
    ("442_128", {"forward_hash_steps": 30}, 0, 29),
    ("L3_128", {"sort_schedule": 3}, 0, "all"),
    ("444_64", {"sort_schedule": 2, "radix_bits": 4}, 1 << 26, "all"),
])
def test_key_sets_match_sorted_forward(pkg, tag, kw, max_slots, expect):
    gp = golden_params(load_golden(tag))
    print(f"case {tag} {kw} max_slots {max_slots}", flush=True)
    L = pkg.native.lib()
    ref = _build(pkg, gp, forward_hash_steps=-1, **{k: v for k, v in kw.items() if k != "forward_hash_steps"})
    assert L.egtb_set_key_set_sizing(0.0, max_slots) == 0
    try:
        got = _build(pkg, gp, **kw)
    finally:
        assert L.egtb_set_key_set_sizing(0.0, 0) == 0
    try:
        assert ref.forward()["hash_steps"] == 0
        fw = got.forward()
        assert fw["hash_steps"] == (gp["steps"] - 2 if expect == "all" else expect), fw
        _same(ref, got, tag)
    finally:
        ref.close()
        got.close()


@pytest.mark.parametrize("tag", ["442_256", "free8_64"])
def test_regrown_key_sets(pkg, tag):
    """Sets sized far too small overflow: they are regrown, their steps run
    again, and the result is still the reference's books."""
    L = pkg.native.lib()
    g = load_golden(tag)
    gp = golden_params(g)
    print(f"case {tag} regrow", flush=True)
    assert L.egtb_set_key_set_sizing(0.02, 0) == 0
    try:
        res = _build(pkg, gp)
    finally:
        assert L.egtb_set_key_set_sizing(0.0, 0) == 0
    try:
        fw = res.forward()
        assert fw["set_regrows"] > 0 and fw["hash_steps"] == gp["steps"] - 2, fw
        for layer in g["layers"]:
            keys, rates = res.book(layer["i"])
            assert book_sha(keys, rates) == layer["sha256"], (tag, layer["i"])
    finally:
        res.close()


def test_small_set_limit_hands_over_to_sorting(pkg):
    """A set limit below what the middle layers need: the early steps use
    key sets, the sort loop takes the rest, and the books are the
    reference's."""
    L = pkg.native.lib()
    g = load_golden("442_256")
    gp = golden_params(g)
    print("case 442_256 set limit 2^16", flush=True)
    assert L.egtb_set_key_set_sizing(0.0, 1 << 16) == 0
    try:
        res = _build(pkg, gp)
    finally:
        assert L.egtb_set_key_set_sizing(0.0, 0) == 0
    try:
        fw = res.forward()
        assert 0 < fw["hash_steps"] < gp["steps"] - 2, fw
        for layer in g["layers"]:
            keys, rates = res.book(layer["i"])
            assert book_sha(keys, rates) == layer["sha256"], ("442_256", layer["i"])
    finally:
        res.close()


def test_key_set_sizing_rejects_bad_values(pkg):
    L = pkg.native.lib()
    assert L.egtb_set_key_set_sizing(-1.0, 0) != 0
    assert L.egtb_set_key_set_sizing(65.0, 0) != 0
    assert L.egtb_set_key_set_sizing(1.0, (1 << 34) + 1) != 0
    assert L.egtb_set_key_set_sizing(0.0, 0) == 0


def test_hand_over_after_overflow_free12(pkg):
    """free12 from its generated layer 0: the sets of layers 2 and 3
    overflow and the sort loop takes over at step 3 (its layers grow about
    eightfold per step). Every raw layer and book equals the sorted forward
    pass's."""
    p = pkg.patterns.build_params("free12", 4)
    seeds = pkg.BookBuilder.resolve_seeds("free12", p)
    gp = dict(seeds=seeds, masks=p["masks"], shifts=p["shifts"], symm=p["symm_mode"], target=4, steps=6,
              docheck=2, dtype="float32")
    print("case free12 hand-over after overflow", flush=True)
    ref = _build(pkg, gp, forward_hash_steps=-1)
    got = _build(pkg, gp)
    try:
        fw = got.forward()
        assert fw["set_regrows"] >= 1 and 0 < fw["hash_steps"] < gp["steps"] - 2, fw
        _same(ref, got, "free12")
    finally:
        ref.close()
        got.close()


def test_hand_over_after_overflow_442_256(pkg):
    """Undersized sets (regrown at nearly every step) and then the sort loop
    from step 100 on."""
    L = pkg.native.lib()
    gp = golden_params(load_golden("442_256"))
    print("case 442_256 sizing 0.02, sort loop from step 100", flush=True)
    ref = _build(pkg, gp, forward_hash_steps=-1)
    assert L.egtb_set_key_set_sizing(0.02, 0) == 0
    try:
        got = _build(pkg, gp, forward_hash_steps=100)
    finally:
        assert L.egtb_set_key_set_sizing(0.0, 0) == 0
    try:
        fw = got.forward()
        assert fw["set_regrows"] >= 1 and fw["hash_steps"] == 99, fw
        _same(ref, got, "442_256")
    finally:
        ref.close()
        got.close()
