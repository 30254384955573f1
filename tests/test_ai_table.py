"""The expectimax player's evaluation table (host code of libegtb_hip.so, no
device): every variant equals the reference's AIPlayer table, by the
SHA-256 digests tests/golden/make_ai_golden.py took of the reference's own
table, and row by row against oracle/_ref where that library is built."""
import ctypes
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN


def _fixture():
    with open(os.path.join(GOLDEN, "ai_search.json")) as fh:
        return json.load(fh)


def _table(pkg, variant):
    out = np.zeros((65536, 2), dtype=np.int16)
    assert pkg.native.lib().egtb_ai_table(out.ctypes.data_as(ctypes.POINTER(ctypes.c_int16)), variant) == 0
    return out


def test_table_variants_match_reference_digests(pkg):
    shas = _fixture()["table_sha256"]
    assert len(shas) == 16
    for v in range(16):
        got = hashlib.sha256(_table(pkg, v).astype("<i2").tobytes()).hexdigest()
        assert got == shas[v], v


def test_table_rejects_bad_variant(pkg):
    out = np.zeros((65536, 2), dtype=np.int16)
    assert pkg.native.lib().egtb_ai_table(out.ctypes.data_as(ctypes.POINTER(ctypes.c_int16)), 16) != 0


def test_planted_table_mistake_is_caught(pkg):
    """A single changed entry (one row's d2 off by one) fails the digest."""
    t = _table(pkg, 0)
    t[0x7fff, 1] += 1
    assert hashlib.sha256(t.astype("<i2").tobytes()).hexdigest() != _fixture()["table_sha256"][0]


def test_table_rows_match_compiled_reference(pkg):
    from oracle.ref import ref

    if not ref.available():
        pytest.skip("oracle/_ref is not built (needs the reference sources)")
    assert np.array_equal(_table(pkg, 0), ref.ai_table())
    for v in (1, 2, 4, 8, 15):
        theirs = ref.ai_table_adjusted(0xff00ff0000000000, v & 1, (v >> 1) & 1, (v >> 2) & 1, (v >> 3) & 1)
        assert np.array_equal(_table(pkg, v), theirs), v
