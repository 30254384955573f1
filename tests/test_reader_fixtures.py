"""The reader oracle against the reference's own compiled reader.

tests/golden/reader_*.json hold what ClassicBookReader::move_on_dic
(ReaderRuntime.cpp:739-840, built from the reference's sources with its
release flags) answered for seeded positions over reference-built books.
Here the C oracle builds the same tables on the CPU (its books equal the
reference's byte for byte on the builder goldens), writes them in the
reference layout, and oracle/reader_oracle.py must give the same ordered
dict and dtype for every position.  Planted mistakes in the reader oracle
(division instead of the reference's reciprocal multiply, a swapped
direction remap, first-op instead of prefer-max) must each fail the
fixtures.  CPU only."""
import os

import numpy as np
import pytest

from conftest import load_reader_fixture, reader_tags
from oracle import reader_oracle as ro

PREFER_MAX_NAMES = ("4442ff", "4442f", "4tiler")
# float32 tables the C oracle builds on the CPU (442@128 in seconds, the
# 12-free-cell LL@64 / 444@64 in under a minute each)
CPU_TAGS = [t for t in ("442_128", "442_128_prefer_max", "LL_64", "444_64") if t in reader_tags()]


def write_books(orc, pkg, pattern, target, folder):
    p = pkg.patterns.build_params(pattern, target)
    ob = orc.OracleBuild(p["seeds"], p["masks"], p["shifts"], p["symm_mode"], target, p["steps"],
                         p["docheck_step"], fma_mode=1)
    rec = np.dtype([("board", "<u8"), ("rate", "<f4")])
    try:
        for i in range(p["steps"]):
            k, r = ob.book(i)
            a = np.empty(k.size, dtype=rec)
            a["board"], a["rate"] = k, r
            a.tofile(os.path.join(folder, f"{pattern}_{2 ** target}_{i}.book"))
    finally:
        ob.close()


@pytest.fixture(scope="module")
def tables(orc, pkg, tmp_path_factory):
    made = {}

    def get(pattern, target):
        key = (pattern, target)
        if key not in made:
            folder = str(tmp_path_factory.mktemp(f"reader_{pattern}_{target}"))
            write_books(orc, pkg, pattern, target, folder)
            made[key] = folder
        return made[key]

    return get


def oracle_answers(fx, folder, **planted):
    prefer_max = fx["reader_name"] in PREFER_MAX_NAMES
    masks = [int(m, 16) for m in fx["masks"]]
    out = []
    for m, _ in fx["expected"]:
        out.append(ro.move_on_dic(m, folder, fx["pattern_full"], fx["nums_adjust"], masks, fx["symm"],
                                  dtype=fx["dtype"], prefer_max=planted.get("prefer_max", prefer_max)))
    return out


def mismatches(fx, got):
    bad = []
    for (m, want), g in zip(fx["expected"], got):
        if g != want or list(g[0].keys()) != list(want[0].keys()):
            bad.append((m.tolist(), g, want))
    return bad


@pytest.mark.parametrize("tag", CPU_TAGS)
def test_reader_oracle_matches_reference_reader(tables, tag):
    fx = load_reader_fixture(tag)
    folder = tables(fx["pattern"], fx["target"])
    bad = mismatches(fx, oracle_answers(fx, folder))
    assert not bad, bad[:3]
    answered = sum(1 for _, w in fx["expected"] if w[1])
    blank = sum(1 for _, w in fx["expected"] if not w[1])
    assert answered > len(fx["expected"]) // 2 and blank > 20


def test_fixtures_cover_the_edge_cases():
    """Every fixture holds answered positions under several symmetry ops,
    positions with an invalid move, blank answers and a full board."""
    for tag in reader_tags():
        fx = load_reader_fixture(tag)
        results = [w for _, w in fx["expected"]]
        assert any(None in d.values() for d, dt in results if dt), tag
        assert any(dt == "" and set(d.values()) == {""} for d, dt in results), tag
        assert sum(1 for d, dt in results if dt) > 250, tag
        assert all(dt in ("", fx["dtype"]) for _, dt in results), tag
        for d, dt in results:
            if not dt:
                continue
            # answered: the numbers first, in descending order, then the invalid moves
            vals = list(d.values())
            nums = [v for v in vals if v is not None]
            assert vals == nums + [None] * (4 - len(nums)) and nums == sorted(nums, reverse=True), (tag, d)


def test_planted_reader_mistakes_fail_the_fixtures(tables, monkeypatch):
    fx = load_reader_fixture("442_128")
    folder = tables(fx["pattern"], fx["target"])
    # (1) the textbook division instead of the reference build's reciprocal multiply
    monkeypatch.setattr(ro, "round9", lambda x: ro.std_round(x * 1e9) / 1e9)
    got = oracle_answers(fx, folder)
    monkeypatch.undo()
    assert mismatches(fx, got)
    # (2) a direction remap with left/right not swapped under the flips
    orig = ro.adjust_direction
    monkeypatch.setattr(ro, "adjust_direction", lambda op, d: orig(op % 4, d))
    got = oracle_answers(fx, folder)
    monkeypatch.undo()
    assert mismatches(fx, got)
    # (3) the first answering op where the reference keeps the best one
    fx = load_reader_fixture("442_128_prefer_max")
    assert mismatches(fx, oracle_answers(fx, folder, prefer_max=False))
