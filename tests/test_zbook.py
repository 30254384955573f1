"""Compressed books (the reference's .z layout: trie index + xz blocks),
host code of libegtb_hip.so, on the CPU.

Round trips on synthetic books of 4- and 8-byte rates whose prefix groups
and blocks cover the layout's cases; and, where the reference's library is
built (oracle/_ref), the files written here equal the reference's
trie_compress_progress_native output byte for byte on oracle-built 442@128
layers, the reference-written files decode to the same records, and the
reference's own per-probe search finds every record in the files written
here."""
import ctypes
import filecmp
import os

import numpy as np
import pytest

from oracle.ref import ref


def zwrite(pkg, path, keys, rates):
    L = pkg.native.lib()
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    rates = np.ascontiguousarray(rates)
    rc = L.egtb_zbook_write(path.encode(), keys.ctypes.data_as(pkg.native.u64p), rates.ctypes.data, keys.size,
                            rates.itemsize)
    assert rc == 0, pkg.native.last_error()


def zread(pkg, path, dtype):
    L = pkg.native.lib()
    size = np.dtype(dtype).itemsize
    n = ctypes.c_size_t()
    assert L.egtb_zbook_read(path.encode(), size, None, None, 0, ctypes.byref(n)) == 0, pkg.native.last_error()
    keys = np.empty(n.value, dtype=np.uint64)
    rates = np.empty(n.value, dtype=dtype)
    assert L.egtb_zbook_read(path.encode(), size, keys.ctypes.data_as(pkg.native.u64p), rates.ctypes.data,
                             n.value, ctypes.byref(n)) == 0, pkg.native.last_error()
    return keys, rates


def synthetic(seed, n, dtype):
    """Sorted unique keys mixing one huge prefix group (several blocks),
    many small groups and lone records, with random rates."""
    rng = np.random.default_rng(seed)
    big = (np.uint64(0x0123456700000000) | rng.integers(0, 2 ** 32, n // 2, dtype=np.uint64))
    small = rng.integers(0, 2 ** 63, n - n // 2, dtype=np.uint64) & np.uint64(0xFFFF0FFF0FFFFFFF)
    keys = np.unique(np.concatenate([big, small]))
    if np.dtype(dtype).kind == "f":
        rates = rng.random(keys.size).astype(dtype)
    else:
        rates = rng.integers(1, 2 ** 31, keys.size).astype(dtype)
    return keys, rates


@pytest.mark.parametrize("dtype,n", [(np.float32, 150_000), (np.uint64, 90_000), (np.float32, 5)])
def test_round_trip(pkg, tmp_path, dtype, n):
    keys, rates = synthetic(n, n, dtype)
    path = str(tmp_path / "p_64_7.book")
    zwrite(pkg, path, keys, rates)
    assert os.path.isdir(str(tmp_path / "p_64_7.z"))
    assert sorted(os.listdir(str(tmp_path / "p_64_7.z"))) == ["p_64_7.i", "p_64_7.ii", "p_64_7.s", "p_64_7.z"]
    k2, r2 = zread(pkg, path, dtype)
    assert np.array_equal(k2, keys) and r2.tobytes() == rates.tobytes()


def test_rejects_unsorted_and_bad_arguments(pkg, tmp_path):
    L = pkg.native.lib()
    keys = np.array([5, 3], dtype=np.uint64)
    rates = np.zeros(2, dtype=np.float32)
    rc = L.egtb_zbook_write(str(tmp_path / "x_1.book").encode(), keys.ctypes.data_as(pkg.native.u64p),
                            rates.ctypes.data, 2, 4)
    assert rc != 0 and "ascending" in pkg.native.last_error()
    n = ctypes.c_size_t()
    assert L.egtb_zbook_read(str(tmp_path / "missing_3.book").encode(), 4, None, None, 0, ctypes.byref(n)) != 0
    assert L.egtb_zbook_read(b"no_suffix", 4, None, None, 0, ctypes.byref(n)) != 0


@pytest.fixture(scope="module")
def oracle_books(orc, pkg, tmp_path_factory):
    p = pkg.patterns.build_params("442", 7)
    ob = orc.OracleBuild(p["seeds"], p["masks"], p["shifts"], p["symm_mode"], 7, p["steps"], p["docheck_step"])
    try:
        layers = {i: ob.book(i) for i in (30, 60, 90)}
    finally:
        ob.close()
    return layers, tmp_path_factory.mktemp("zbooks")


@pytest.mark.skipif(not ref.available(), reason="the reference's library is built only where its sources are")
def test_files_equal_the_reference_compressor(pkg, oracle_books):
    layers, folder = oracle_books
    rec = np.dtype([("board", "<u8"), ("rate", "<f4")])
    for i, (keys, rates) in layers.items():
        ours = folder / "ours"
        theirs = folder / "theirs"
        ours.mkdir(exist_ok=True)
        theirs.mkdir(exist_ok=True)
        a = np.empty(keys.size, dtype=rec)
        a["board"], a["rate"] = keys, rates
        a.tofile(str(theirs / f"442_128_{i}.book"))
        ref.trie_compress(str(theirs / f"442_128_{i}.book"), "float32")
        zwrite(pkg, str(ours / f"442_128_{i}.book"), keys, rates)
        for suffix in ("i", "ii", "s", "z"):
            f1 = str(ours / f"442_128_{i}.z" / f"442_128_{i}.{suffix}")
            f2 = str(theirs / f"442_128_{i}.z" / f"442_128_{i}.{suffix}")
            assert filecmp.cmp(f1, f2, shallow=False), (i, suffix)
        k2, r2 = zread(pkg, str(theirs / f"442_128_{i}.book"), np.float32)
        assert np.array_equal(k2, keys) and np.array_equal(r2.view(np.uint32), rates.view(np.uint32))
        prefix = str(ours / f"442_128_{i}.z" / f"442_128_{i}.")
        step = max(1, keys.size // 500)
        for j in range(0, keys.size, step):
            assert ref.trie_search(prefix, int(keys[j]), "float32") == float(rates[j]), (i, j)
