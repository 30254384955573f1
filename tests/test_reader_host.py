"""Host-side reader helpers on synthetic .book files (no GPU):
find_classic_value (ReaderRuntime.cpp:279-316) and get_random_state
(ReaderRuntime.cpp:944-990)."""
import numpy as np


def _write_book(path, boards, rates, dtype="float32"):
    rdt = {"float32": "<f4", "float64": "<f8", "uint32": "<u4", "uint64": "<u8"}[dtype]
    rec = np.empty(len(boards), dtype=np.dtype([("board", "<u8"), ("rate", rdt)]))
    rec["board"] = boards
    rec["rate"] = rates
    rec.tofile(path)


def test_find_classic_value_sorted_book(pkg, tmp_path):
    rng = np.random.default_rng(3)
    boards = np.unique(rng.integers(1, 2**63, size=5000, dtype=np.uint64))
    rates = rng.random(boards.size).astype(np.float32)
    _write_book(tmp_path / "442_512_7.book", boards, rates)
    f = pkg.find_classic_value
    for k in [0, 1, 2499, boards.size - 1]:
        assert f(str(tmp_path), "442_512_7.book", int(boards[k]), "float32") == float(rates[k])
    absent = int(boards[10]) + 1
    assert absent not in set(boards.tolist())
    assert f(str(tmp_path), "442_512_7.book", absent, "float32") == 0.0
    assert f(str(tmp_path), "442_512_7.book", 0, "float32") == 0.0
    assert f(str(tmp_path), "442_512_7.book", 2**64 - 1, "float32") == 0.0
    assert f(str(tmp_path), "missing.book", 5, "float32") is None


def test_find_classic_value_integer_scale_and_empty(pkg, tmp_path):
    _write_book(tmp_path / "a.book", np.array([3, 8, 20], np.uint64), np.array([0, 2e9, 4e9], np.uint32), "uint32")
    f = pkg.find_classic_value
    assert f(str(tmp_path), "a.book", 8, "uint32") == 0.5
    assert f(str(tmp_path), "a.book", 20, "uint32") == 1.0
    _write_book(tmp_path / "b.book", np.array([3, 8], np.uint64), np.array([0.25, 0.75]), "float64")
    assert f(str(tmp_path), "b.book", 3, "float64") == 0.25
    open(tmp_path / "empty.book", "wb").close()
    assert f(str(tmp_path), "empty.book", 3, "uint32") == 0.0


def test_find_classic_value_on_seed_order_layer(pkg, tmp_path):
    """Layer 0 keeps the seeds' order; the reference's bisection finds the
    middle record and misses a key on the wrong side of it."""
    _write_book(tmp_path / "s.book", np.array([50, 10, 90], np.uint64), np.array([0.1, 0.2, 0.3], np.float32))
    f = pkg.find_classic_value
    assert f(str(tmp_path), "s.book", 10, "float32") == np.float32(0.2)
    assert f(str(tmp_path), "s.book", 90, "float32") == np.float32(0.3)
    assert f(str(tmp_path), "s.book", 50, "float32") == 0.0


def test_get_random_state(pkg, tmp_path):
    stored = np.array([0x1000000021ffffff, 0x0000000112ffffff, 0x0000001122ffffff], np.uint64)
    for book_id in (2, 5):
        _write_book(tmp_path / f"442_512_{book_id}.book", stored, np.ones(3, np.float32))
    open(tmp_path / "442_512_7.book", "wb").close()  # empty books are passed over
    reader = pkg.BookReader
    rng = np.random.default_rng(0)
    seen_vals = set()
    for _ in range(200):
        s = int(reader.get_random_state([(str(tmp_path), "float32")], "442_512", rng=rng))
        parents = [int(b) for b in stored if (s & int(b)) == int(b) and s != int(b)]
        assert len(parents) == 1, hex(s)
        diff = s ^ parents[0]
        k = (diff.bit_length() - 1) // 4
        assert diff == ((diff >> (4 * k)) & 0xF) << (4 * k)  # exactly one new nibble
        seen_vals.add((diff >> (4 * k)) & 0xF)
    assert seen_vals == {1, 2}
    assert reader.get_random_state([(str(tmp_path / "nowhere"), "float32")], "442_512") == 0
    assert reader.get_random_state([(str(tmp_path), "float32")], "nosuch_512") == 0
