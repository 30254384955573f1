"""GPU tests of the free* pattern path: device moves/canonical forms
(egtb_board_moves) against the oracle, layer 0 from generate_free_inits
(engine_core/BookBuilder.py:146-181) against the same procedure run on the
oracle's primitives, and a short build from those seeds against the oracle."""
import itertools

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rand_boards(n, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, 2**63, size=n, dtype=np.uint64) | (rng.integers(0, 2, size=n, dtype=np.uint64) << 63)


def test_board_moves_match_oracle(pkg, orc):
    boards = np.concatenate([_rand_boards(4000, 11), np.array([0, 0x1111, 0xF00F00F00F00F00F], np.uint64)])
    print("case board moves", flush=True)
    for mode in range(4):
        moved, canon = pkg.BookBuilder.board_moves(boards, mode)
        for k, b in enumerate(boards.tolist()):
            assert [int(x) for x in moved[k]] == [orc.move(b, d) for d in range(4)], hex(b)
            assert int(canon[k]) == orc.canonical(b, mode), (hex(b), mode)


def _free_inits_on_oracle(orc, t32ks, t2s):
    """The reference's generate_free_inits loop, step by step, with the
    oracle's move and canonical functions."""
    generated = []
    for pos32 in itertools.combinations(range(16), t32ks):
        base = sum(15 << (4 * p) for p in pos32)
        remain = sorted(set(range(16)) - set(pos32))
        for pos2 in itertools.combinations(remain, t2s):
            generated.append(base | sum(1 << (4 * p) for p in pos2))
    gen = np.unique(np.asarray(generated, dtype=np.uint64))
    out = []
    for b in gen.tolist():
        for d in range(4):
            m = orc.move(b, d)
            if m == orc.canonical(m, 1):
                out.append(m)
        if len(out) >= gen.size - 3:
            break
    return np.unique(np.asarray(out, dtype=np.uint64))


@pytest.mark.parametrize("t32ks,t2s", [(4, 11), (6, 9)])
def test_generate_free_inits_matches_oracle_loop(pkg, orc, t32ks, t2s):
    print(f"case free inits {t32ks} {t2s}", flush=True)
    got = pkg.BookBuilder.generate_free_inits(t32ks, t2s)
    want = _free_inits_on_oracle(orc, t32ks, t2s)
    assert np.array_equal(got, want)
    assert got.size > 1 and bool(np.all(got[1:] > got[:-1]))
    tile_sum = 32768 * t32ks + 2 * t2s
    assert {pkg.patterns.tile_sum_of(int(b)) for b in got[:500]} == {tile_sum}


def test_free_pattern_seeds_and_short_build(pkg, orc):
    """free12's layer 0 comes from generate_free_inits; six layers built
    from it match the oracle bit for bit (states and float32 rates)."""
    p = pkg.patterns.build_params("free12", 4)
    seeds = pkg.BookBuilder.resolve_seeds("free12", p)
    assert seeds.size > 1 and p["is_free"]
    print(f"case free12 short build, {seeds.size} seeds", flush=True)
    steps, docheck = 6, 2
    res = pkg.run_pattern_build(seeds, p["masks"], p["shifts"], p["symm_mode"], 4, steps, docheck, "",
                                write_files=False, keep=True, keep_raw=True)
    ob = orc.OracleBuild(seeds, p["masks"], p["shifts"], p["symm_mode"], 4, steps, docheck)
    try:
        for i in range(steps):
            assert np.array_equal(res.raw(i), ob.raw(i)), f"raw layer {i}"
            gk, gr = res.book(i)
            ok, orr = ob.book(i)
            assert np.array_equal(gk, ok), f"book keys {i}"
            assert np.array_equal(gr.view(np.uint32), orr.view(np.uint32)), f"book rates {i}"
    finally:
        res.close()
        ob.close()
