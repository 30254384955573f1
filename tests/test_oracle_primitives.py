"""Pin the oracle's primitives against independent restatements: the 2048
merge rule on actual tile values, numpy matrix transforms, orbit minima."""
import itertools

import numpy as np
import pytest


def decode(board):
    """4x4 matrix of exponents; row 0 = top (nibbles 15..12)."""
    m = np.zeros((4, 4), dtype=np.int64)
    for r in range(4):
        for c in range(4):
            m[r, c] = (board >> (4 * ((3 - r) * 4 + (3 - c)))) & 0xF
    return m


def encode(m):
    b = 0
    for r in range(4):
        for c in range(4):
            b |= int(m[r, c]) << (4 * ((3 - r) * 4 + (3 - c)))
    return b


def slide_values_left(values):
    """Textbook 2048 row move on tile values (0 empty): tiles slide left,
    equal neighbours merge once (leftmost pair first); 32768 never merges."""
    tiles = [v for v in values if v]
    out = []
    i = 0
    while i < len(tiles):
        if i + 1 < len(tiles) and tiles[i] == tiles[i + 1] and tiles[i] != 32768:
            out.append(tiles[i] * 2)
            i += 2
        else:
            out.append(tiles[i])
            i += 1
    return out + [0] * (4 - len(out))


def row_exps_to_int(exps):
    return (exps[0] << 12) | (exps[1] << 8) | (exps[2] << 4) | exps[3]


def test_row_table_exhaustive(orc):
    L = orc.lib()
    for r in range(65536):
        exps = [(r >> (12 - 4 * c)) & 0xF for c in range(4)]
        vals = [0 if e == 0 else 1 << e for e in exps]
        left = [0 if v == 0 else v.bit_length() - 1 for v in slide_values_left(vals)]
        right = [0 if v == 0 else v.bit_length() - 1 for v in slide_values_left(vals[::-1])][::-1]
        assert L.orc_row_left(r) == row_exps_to_int(left), hex(r)
        assert L.orc_row_right(r) == row_exps_to_int(right), hex(r)


def test_row_table_known_answers(orc):
    L = orc.lib()
    # [2,2,2,2] -> [4,4,0,0]; [2,2,4,0] -> [4,4,0,0]; [32768,32768,..] stays
    assert L.orc_row_left(0x1111) == 0x2200
    assert L.orc_row_left(0x1120) == 0x2200
    assert L.orc_row_left(0xFF00) == 0xFF00
    assert L.orc_row_left(0x0101) == 0x2000
    assert L.orc_row_right(0x1111) == 0x0022
    assert L.orc_row_left(0xEE00) == 0xF000  # 16384 + 16384 -> 32768


def _rand_boards(n, seed=0):
    rng = np.random.default_rng(seed)
    for _ in range(n):
        yield int(rng.integers(0, 2**63, dtype=np.uint64)) | (int(rng.integers(0, 2)) << 63)


def _move_matrix(m, direction):
    """Moves on the matrix by rotating so the move is a left slide."""
    k = {0: 0, 1: 2, 2: 1, 3: 3}[direction]  # left, right, up, down
    rot = np.rot90(m, k)
    out = np.zeros_like(rot)
    for r in range(4):
        vals = [0 if e == 0 else 1 << int(e) for e in rot[r]]
        out[r] = [0 if v == 0 else v.bit_length() - 1 for v in slide_values_left(vals)]
    return np.rot90(out, -k)


def test_moves_match_matrix_restatement(orc):
    for b in itertools.islice(_rand_boards(3000, 1), 3000):
        m = decode(b)
        for d in range(4):
            assert orc.move(b, d) == encode(_move_matrix(m, d)), (hex(b), d)


def test_transforms_match_numpy(orc):
    L = orc.lib()
    for b in itertools.islice(_rand_boards(2000, 2), 2000):
        m = decode(b)
        assert L.orc_transpose(b) == encode(m.T)
        assert L.orc_mirror_lr(b) == encode(np.fliplr(m))
        assert L.orc_mirror_ud(b) == encode(np.flipud(m))


def _orbit(m):
    for k in range(4):
        r = np.rot90(m, k)
        yield r
        yield np.fliplr(r)


def test_canonical_modes(orc):
    for b in itertools.islice(_rand_boards(2000, 3), 2000):
        m = decode(b)
        assert orc.canonical(b, 0) == b
        assert orc.canonical(b, 1) == min(encode(x) for x in _orbit(m))
        assert orc.canonical(b, 2) == min(b, encode(m.T))
        assert orc.canonical(b, 3) == min(b, encode(np.fliplr(m)))
        for mode in range(4):
            c = orc.canonical(b, mode)
            assert orc.canonical(c, mode) == c


def test_predicates(orc):
    L = orc.lib()
    import ctypes

    masks = (ctypes.c_uint64 * 2)(0xFFFFFF, 0xF0FFF0FF)
    assert L.orc_is_pattern(0x1000000021FFFFFF, masks, 1)
    assert not L.orc_is_pattern(0x1000000021FFFFFE, masks, 1)
    assert L.orc_is_pattern(0x00000000F0FFF0FF, masks, 2)
    assert L.orc_is_pattern(0x1234, masks, 0)  # no masks: everything matches
    assert L.orc_is_success(0x0000000009000000, 9, bytes([24]), 1)
    assert not L.orc_is_success(0x0000000090000000, 9, bytes([24]), 1)
    assert L.orc_is_success(0x0000000090000000, 9, bytes([24, 28]), 2)
    assert L.orc_is_success(0x5, 9, b"", 0)  # no shifts: always success


def test_sort_unique_matches_numpy(orc):
    rng = np.random.default_rng(5)
    for n in [0, 1, 7, 1000, 65537]:
        a = rng.integers(0, 2**64 - 1, size=n, dtype=np.uint64)
        a = np.concatenate([a, a[: n // 2]])
        assert np.array_equal(orc.sort_unique(a), np.unique(a))
