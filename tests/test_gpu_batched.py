"""Batched builds: every job's solved layers equal the reference golden
digests / the plain GPU build of that job alone."""
import numpy as np
import pytest

from conftest import book_sha, load_golden

pytestmark = pytest.mark.gpu


def test_batch_matches_plain_and_golden(pkg):
    from importlib import import_module

    batched = import_module("2048endgametablebase_amd.batched")
    jobs = [("442", 7), ("L3", 7), ("LL", 6), ("444", 6), ("t", 6), ("442", 5), ("4442", 4), ("4431", 5)]
    goldens = {("442", 7): "442_128", ("L3", 7): "L3_128", ("LL", 6): "LL_64", ("444", 6): "444_64"}
    with batched.build_batch(jobs) as br:
        for j, (pat, t) in enumerate(jobs):
            print(f"case job {j} {pat}@{2**t}", flush=True)
            if (pat, t) in goldens:
                g = load_golden(goldens[(pat, t)])
                for layer in g["layers"]:
                    keys, rates = br.book(j, layer["i"])
                    assert book_sha(keys, rates) == layer["sha256"], (pat, t, layer["i"])
            else:
                res = pkg.build_in_memory(pat, t)
                try:
                    for i in range(res.steps):
                        pk, pr = res.book(i)
                        bk, brr = br.book(j, i)
                        assert np.array_equal(pk, bk) and np.array_equal(pr.view(np.uint32), brr.view(np.uint32)), \
                            (pat, t, i)
                finally:
                    res.close()


def test_batch_rejects_unpinned_patterns(pkg):
    from importlib import import_module

    native = pkg.native
    job = native.BatchJob()
    seeds = np.array([0x01111111FFFFFFFF], dtype=np.uint64)
    job.seeds = seeds.ctypes.data_as(native.u64p)
    job.n_seeds = 1
    job.pattern = native.Pattern(None, 0, None, 0, 1)
    job.target, job.steps, job.docheck_step, job.spawn_rate4 = 7, 100, 50, 0.1
    o = native.BatchOptions()
    import ctypes

    h = ctypes.c_void_p()
    rc = native.lib().egtb_batch_build(ctypes.byref(job), 1, ctypes.byref(o), ctypes.byref(h))
    assert rc != 0 and "pin nibbles" in native.last_error()


@pytest.mark.parametrize("variant", [dict(sort_schedule=1), dict(forward_hash_steps=-1),
                                     dict(forward_hash_steps=-1, sort_schedule=2), dict(forward_hash_steps=40)])
def test_batch_key_sets_match_sorted_dedup(pkg, variant):
    """The batched forward pass deduplicates in key sets by default; with
    forward_hash_steps = -1 it sorts and uniques every step's candidates, and
    with k > 0 it hands over to sort + unique at step k.  Every variant gives
    every job the same books."""
    from importlib import import_module

    batched = import_module("2048endgametablebase_amd.batched")
    jobs = [("442", 7), ("LL", 6), ("t", 6)]
    print(f"case batched key sets vs {variant}", flush=True)
    with batched.build_batch(jobs) as a, batched.build_batch(jobs, **variant) as b:
        for j, (pat, t) in enumerate(jobs):
            for i in range(pkg.patterns.build_params(pat, t)["steps"]):
                ka, ra = a.book(j, i)
                kb, rb = b.book(j, i)
                assert np.array_equal(ka, kb) and ra.tobytes() == rb.tobytes(), (pat, t, i)


def test_default_mix_matches_goldens_and_plain(pkg):
    """The benchmark's 24-job mix (L3 / 442 / t at 64-1024, the 12-free-cell
    patterns at 32-128) in one batched build: every job whose table has a
    reference golden equals it on every layer; every other job equals its
    plain build on three layers spread over its range."""
    from importlib import import_module

    from conftest import golden_tags

    batched = import_module("2048endgametablebase_amd.batched")
    jobs = batched.DEFAULT_JOBS
    assert len(jobs) == 24
    have = set(golden_tags())
    print(f"case batched default mix: {len(jobs)} jobs", flush=True)
    with batched.build_batch(jobs) as br:
        checked = 0
        for j, (pat, t) in enumerate(jobs):
            tag = f"{pat}_{2 ** t}"
            if tag in have:
                print(f"case job {j} {tag}: golden", flush=True)
                for layer in load_golden(tag)["layers"]:
                    keys, rates = br.book(j, layer["i"])
                    assert book_sha(keys, rates) == layer["sha256"], (tag, layer["i"])
                checked += 1
                continue
            print(f"case job {j} {tag}: plain build", flush=True)
            res = pkg.build_in_memory(pat, t)
            try:
                for i in (res.steps // 4, res.steps // 2, 3 * res.steps // 4):
                    pk, pr = res.book(i)
                    bk, brr = br.book(j, i)
                    assert np.array_equal(pk, bk) and np.array_equal(pr.view(np.uint32), brr.view(np.uint32)), \
                        (tag, i)
            finally:
                res.close()
        assert checked >= 8


def test_planned_batch_matches_goldens_and_plain(pkg):
    """plan_batch merges small jobs into groups whose peak layers sum to at
    most MAX_GROUP_PEAK keys and builds larger ones alone; every job of the
    planned mix equals its golden or its plain build."""
    from importlib import import_module

    from conftest import golden_tags

    batched = import_module("2048endgametablebase_amd.batched")
    jobs = batched.SMALL_JOBS[:12] + [("4442", 4), ("L3", 7), ("442", 7), ("L3", 3), ("442", 3)]
    groups, solo = batched.plan_batch(jobs)
    assert groups and solo and sorted(sum(groups, []) + solo) == list(range(len(jobs)))
    assert jobs.index(("4442", 4)) in solo  # its 4.7e7-state layers would overflow a merged set
    have = set(golden_tags())
    print(f"case planned batch: {len(groups)} groups, {len(solo)} alone", flush=True)
    with batched.build_planned(jobs) as br:
        for j, (pat, t) in enumerate(jobs):
            tag = f"{pat}_{2 ** t}"
            if tag in have:
                for layer in load_golden(tag)["layers"]:
                    keys, rates = br.book(j, layer["i"])
                    assert book_sha(keys, rates) == layer["sha256"], (tag, layer["i"])
                continue
            res = pkg.build_in_memory(pat, t)
            try:
                for i in (1, res.steps // 2, res.steps - 3):
                    pk, pr = res.book(i)
                    bk, brr = br.book(j, i)
                    assert np.array_equal(pk, bk) and np.array_equal(pr.view(np.uint32), brr.view(np.uint32)), \
                        (tag, i)
            finally:
                res.close()
