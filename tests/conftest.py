import hashlib
import importlib
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and libegtb_hip.so")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def pkg():
    return importlib.import_module("2048endgametablebase_amd")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle

    oracle.lib()
    return oracle


def load_golden(tag):
    with open(os.path.join(GOLDEN, f"{tag}.json")) as fh:
        return json.load(fh)


def golden_tags():
    """Builder goldens (per-layer book digests); reader fixtures are reader_*,
    the expectimax player's ai_*, the advanced forward pass's ad_*."""
    return sorted(f[:-5] for f in os.listdir(GOLDEN)
                  if f.endswith(".json") and not f.startswith("reader_") and not f.startswith("ai_")
                  and not f.startswith("ad_"))


def ad_tags():
    """Advanced (masked) forward-pass goldens: per raw layer file digests."""
    return sorted(f[3:-5] for f in os.listdir(GOLDEN)
                  if f.startswith("ad_") and not f.startswith("ad_book_") and f.endswith(".json"))


def ad_book_tags():
    """Advanced solver goldens: per solved layer and bucket, file digests."""
    return sorted(f[8:-5] for f in os.listdir(GOLDEN) if f.startswith("ad_book_") and f.endswith(".json"))


def load_ad_book_golden(tag):
    with open(os.path.join(GOLDEN, f"ad_book_{tag}.json")) as fh:
        return json.load(fh)


def load_ad_golden(tag):
    with open(os.path.join(GOLDEN, f"ad_{tag}.json")) as fh:
        return json.load(fh)


def reader_tags():
    return sorted(f[7:-5] for f in os.listdir(GOLDEN) if f.startswith("reader_") and f.endswith(".json"))


def load_reader_fixture(tag):
    """Reader fixture: cases of (4x4 tile values, expected (dict, dtype))."""
    with open(os.path.join(GOLDEN, f"reader_{tag}.json")) as fh:
        fx = json.load(fh)
    fx["expected"] = [(np.array(b, dtype=np.int64).reshape(4, 4), (dict((k, v) for k, v in res), dt))
                      for b, res, dt in fx["cases"]]
    return fx


def book_bytes(keys, rates):
    """Packed reference records {u64 board; rate} in the rates' dtype (12 or
    16 bytes each)."""
    rec = np.empty(keys.size, dtype=np.dtype([("board", "<u8"), ("rate", rates.dtype.newbyteorder("<"))]))
    rec["board"] = keys
    rec["rate"] = rates
    return rec.tobytes()


def book_sha(keys, rates):
    return hashlib.sha256(book_bytes(keys, rates)).hexdigest()


def golden_params(g):
    return dict(seeds=[int(s, 16) for s in g["seeds"]], masks=[int(m, 16) for m in g["masks"]],
                shifts=list(g["shifts"]), symm=g["symm"], target=g["target"], steps=g["steps"],
                docheck=g["docheck"], dtype=g.get("dtype", "float32"), options=g.get("options", {}))
