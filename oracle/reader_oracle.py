"""CPU restatement of the reference's classic reader (test infrastructure).

move_on_dic follows ReaderRuntime.cpp:739-840 on 4x4 numpy matrices:
apply_operation (:127-147) with the reference's rotate/flip definitions
(:89-125), moves and canonical forms from the C oracle, and the on-disk
binary search of find_success_entry_in_file (:279-316) with one seek + read
per probe.  The symmetry order is 0..7 (stateless; see the product reader).
"""
from __future__ import annotations

import math
import os
import struct

import numpy as np

from . import oracle as orc

ORDERED_KEYS = ("down", "right", "left", "up")
DIR_NAMES = ("up", "right", "down", "left")


def rotate_left(b):
    r = np.zeros_like(b)
    for row in range(4):
        for col in range(4):
            r[3 - col][row] = b[row][col]
    return r


def rotate_right(b):
    r = np.zeros_like(b)
    for row in range(4):
        for col in range(4):
            r[col][3 - row] = b[row][col]
    return r


def rotate_180(b):
    r = np.zeros_like(b)
    for row in range(4):
        for col in range(4):
            r[3 - row][3 - col] = b[row][col]
    return r


def flip_horizontal(b):
    return np.array([list(row)[::-1] for row in b])


def apply_operation(b, op):
    return [b, rotate_left(b), rotate_180(b), rotate_right(b), flip_horizontal(b),
            flip_horizontal(rotate_left(b)), flip_horizontal(rotate_180(b)),
            flip_horizontal(rotate_right(b))][op]


def encode(b):
    out = 0
    for r in range(4):
        for c in range(4):
            v = int(b[r][c])
            e = v.bit_length() - 1 if v > 0 else 0
            out |= (e & 0xF) << (4 * ((3 - r) * 4 + (3 - c)))
    return out


def adjust_direction(op, d):
    if op >= 4:
        d = {"left": "right", "right": "left"}.get(d, d)
    idx = {"up": 0, "right": 1, "down": 2, "left": 3}[d]
    return DIR_NAMES[(idx + op % 4) % 4]


def std_round(x):
    t = float(math.trunc(x))
    if abs(x - t) >= 0.5:
        t += math.copysign(1.0, x)
    return t


def round9(x):
    """maybe_round_value (ReaderRuntime.cpp:208-214) as the reference's
    release flags compile it: -Ofast implies -freciprocal-math, so the
    division by 10^9 becomes a multiplication by the double nearest 1e-9."""
    return std_round(x * 1e9) * 1e-9


# record layout and max scale per dtype (FormationRuntime.h:197-200,
# ReaderRuntime.cpp:183-206: integer rates are divided by their max scale)
REC_FORMAT = {"float32": ("<Qf", 1.0), "float64": ("<Qd", 1.0), "uint32": ("<QI", 4e9),
              "uint64": ("<QQ", 1.6e18), "1-float32": ("<Qf", 1.0, -1.0), "1-float64": ("<Qd", 1.0, -1.0)}


def find_in_file(path, key, rec_fmt="<Qf", scale=1.0, zero=0.0):
    size = struct.calcsize(rec_fmt)
    with open(path, "rb") as fh:
        fh.seek(0, os.SEEK_END)
        n = fh.tell() // size
        left, right = 0, n - 1 if n else 0
        while left <= right and n:
            mid = (left + right) // 2
            fh.seek(mid * size)
            board, rate = struct.unpack(rec_fmt, fh.read(size))
            if board == key:
                return float(rate) / scale if scale > 1.0 else float(rate)
            if board < key:
                left = mid + 1
            else:
                if mid == 0:
                    break
                right = mid - 1
    return zero


def is_pattern(b, masks):
    return not masks or any((b & m) == m for m in masks)


def move_on_dic(board, folder, pattern_full, nums_adjust, masks, symm, dtype="float32", prefer_max=False):
    """evaluate_classic_result_candidates (ReaderRuntime.cpp:739-840) for one
    folder.  prefer_max: every op is tried and the op whose best (rounded)
    rate is strictly largest, starting from 0, answers (:827-832)."""
    board = np.asarray(board, dtype=np.int64).reshape(4, 4)
    v = int(board.sum()) + nums_adjust
    nums = abs(v) // 2 if v >= 0 else -(abs(v) // 2)
    blank = {"down": "", "right": "", "left": "", "up": ""}
    if nums < 0:
        return blank, ""
    path = os.path.join(folder, f"{pattern_full}_{nums}.book")
    if not os.path.exists(path):
        return blank, ""
    best, found = 0.0, None
    for op in range(8):
        enc = encode(apply_operation(board, op))
        if not is_pattern(enc, masks):
            continue
        vals = [None, None, None, None]  # down, right, left, up (result_values)
        for index in range(4):
            moved = orc.move(enc, index)
            if moved == enc or not is_pattern(moved, masks):
                continue
            val = find_in_file(path, orc.canonical(moved, symm), *REC_FORMAT[dtype])
            if "32" in dtype and abs(val) > 1e-7:
                val = round9(val)
            vals[{0: 2, 1: 1, 2: 3, 3: 0}[index]] = val
        if all(x is None for x in vals):
            continue
        entries = [(adjust_direction(op, k), vals[i]) for i, k in enumerate(ORDERED_KEYS)]
        num = [e for e in entries if e[1] is not None]
        oth = [e for e in entries if e[1] is None]
        num.sort(key=lambda e: -e[1])
        if not prefer_max:
            return dict(num + oth), dtype
        if num[0][1] > best:
            best, found = num[0][1], dict(num + oth)
    if prefer_max and found is not None:
        return found, dtype
    return blank, ""
