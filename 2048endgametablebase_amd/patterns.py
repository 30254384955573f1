"""Pattern catalogue and build-parameter formulas for the classic builder.

The entries restate, as plain data, the patterns of the reference's
``docs_and_configs/patterns_config.json`` that the classic path is run on,
parsed the way ``Config.load_patterns_from_file`` does (Config.py:179-282):

* ``pattern_masks``  -- the "valid pattern" masks; a board belongs to the
  pattern when some mask's nibbles are all 0xF on it (Formation.h:11-17);
* ``success_shifts`` -- bit offsets of the 0xF nibbles of "target pos"
  (``find_f_nibble_positions``, Config.py:154-160);
* ``canonical_mode`` -- symmetry reduction applied before storage;
* ``seed_boards``    -- layer 0, in catalogue order;
* ``nums_adjust``    -- minus the tile sum of the first seed;
* ``extra_steps``    -- layers past target/2 (default 36).

``steps_and_docheck`` is ``_steps_and_docheck`` (engine_core/BookBuilder.py:
124-128).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Tuple

# Symmetry mode numbering of the reference's SymmMode enum
# (native_core/include/FormationRuntime.h:13).
SYMM_MODES = {
    "identity": 0,
    "full": 1,
    "diagonal": 2,
    "horizontal": 3,
    "min33": 4,
    "min24": 5,
    "min34": 6,
}


def _f_nibble_shifts(mask: int) -> Tuple[int, ...]:
    return tuple(4 * i for i in range(16) if (mask >> (4 * i)) & 0xF == 0xF)


def tile_sum_of(board: int) -> int:
    """Sum of tile values of an encoded board (BoardCodec.h:20-74)."""
    total = 0
    for k in range(16):
        v = (board >> (4 * k)) & 0xF
        if v:
            total += 1 << v
    return total


@dataclass(frozen=True)
class PatternMeta:
    name: str
    category: str
    pattern_masks: Tuple[int, ...]
    success_shifts: Tuple[int, ...]
    canonical_mode: str
    seed_boards: Tuple[int, ...]
    extra_steps: int = 36

    @property
    def symm_mode(self) -> int:
        return SYMM_MODES[self.canonical_mode]

    @property
    def tile_sum(self) -> int:
        return tile_sum_of(self.seed_boards[0])

    @property
    def nums_adjust(self) -> int:
        return -self.tile_sum

    @property
    def mask_intersection(self) -> int:
        """Cells every mask pins (get_nibble_intersection, Config.py:163-166:
        0 when the pattern has no masks)."""
        if not self.pattern_masks:
            return 0
        inter = 0xFFFFFFFFFFFFFFFF
        for m in self.pattern_masks:
            inter &= m
        return inter

    @property
    def free_cells(self) -> int:
        inter = self.mask_intersection
        return 16 - sum(1 for k in range(16) if (inter >> (4 * k)) & 0xF == 0xF)

    @property
    def fixed_positions(self) -> Tuple[int, ...]:
        return _f_nibble_shifts(self.mask_intersection)


def _entry(name, category, masks, target_pos, mode, seeds, extra):
    mask_ints = tuple(int(m, 16) for m in masks)
    seed_ints = tuple(int(s, 16) for s in seeds)
    # Config.py:243-248 keeps only seeds with the first seed's tile sum.
    first = tile_sum_of(seed_ints[0])
    seed_ints = tuple(s for s in seed_ints if tile_sum_of(s) == first)
    return PatternMeta(
        name=name,
        category=category,
        pattern_masks=mask_ints,
        success_shifts=_f_nibble_shifts(int(target_pos, 16)),
        canonical_mode=mode,
        seed_boards=seed_ints,
        extra_steps=extra,
    )


# Values from docs_and_configs/patterns_config.json (the runtime file,
# Config.py:182), for the classic patterns that the tablebase path builds.
pattern_catalog: Dict[str, PatternMeta] = {
    m.name: m
    for m in (
        _entry("L3", "10 space", ["0xfff0fff"], "0xffffffffffffffff", "identity",
               ["0x100000001fff2fff", "0x000000012fff1fff"], 48),
        _entry("442", "10 space", ["0xffffff"], "0xf000000", "identity",
               ["0x1000000021ffffff", "0x0000000112ffffff"], 48),
        _entry("t", "10 space", ["0xf0fff0ff", "0xf000f0ff00ff", "0xf000f00000ff00ff"],
               "0xff0f000f00", "identity",
               ["0x10000000f1fff2ff", "0x00000001f2fff1ff"], 48),
        _entry("LL", "12 space", ["0xff00ff"], "0xff0f000f00", "diagonal",
               ["0x1000000021ff12ff", "0x0000000112ff21ff"], 48),
        _entry("4431", "12 space", ["0xf0fff"], "0xf00000", "identity",
               ["0x10000000121f2fff", "0x00000001121f2fff"], 48),
        _entry("444", "12 space", ["0xffff"], "0xff00000", "horizontal",
               ["0x100000000000ffff", "0x000000010000ffff"], 60),
        _entry("4442", "others", ["0xff"], "0xff0f00", "identity",
               ["0x00001000000012ff", "0x00010000000021ff"], 48),
        _entry("4432", "others", ["0xf00ff"], "0xf00000", "diagonal",
               ["0x00001000121f21ff", "0x00010000121f21ff"], 48),
        _entry("4441", "others", ["0xfff"], "0xfff000", "identity",
               ["0x0000100000001fff", "0x0001000000001fff"], 48),
        _entry("44", "others", ["0xffffffff"], "0xffff00000000", "identity",
               ["0x00001000ffffffff"], 48),
        _entry("L3f", "10 space",
               ["0xfff0fff", "0xf0fff0ff0", "0xf000f0ff00ff0", "0xf00fff0ff0", "0xf000fff0ff0", "0xff0ff00ff0",
                "0xff00ff00ff0", "0xf000fff00ff", "0xf00fff00ff", "0xff0fff0f00", "f00f00ff00ff0",
                "fff0fff0000", "ff00fff0f00", "f00fff0f0f"],
               "0xffffffffffffffff", "identity", ["0x100000001fff2fff", "0x000000012fff1fff"], 48),
        _entry("442t", "10 space",
               ["0xffffff", "0xf0ff0fff", "0xffff0ff", "0xffff00ff", "0xf00fffff0", "0xff0ff0ff0",
                "0xf00000ff0fff", "0xf0000fff00ff", "0xf00f00ff0ff0", "0xf00f0fff00f0"],
               "0xff0ff000000", "identity", ["0x1000000021ffffff", "0x0000000112ffffff"], 48),
        _entry("4432f", "12 space", ["0xf00ff"], "0xf00000", "diagonal",
               ["0x00001000121f2fff", "0x00010000121ff2ff", "0x0001000012ff21ff"], 48),
        _entry("4442ff", "12 space", ["0xff"], "0xff0fff0f00", "identity",
               ["0x0000100012ff21ff", "0x000100001212ffff", "0x00ff0012102100ff"], 48),
        _entry("4421", "others", ["0xff0fff"], "0xf000000", "identity",
               ["0x1000000012ff2fff", "0x1000000121ff1fff"], 48),
        _entry("4442f", "others", ["0xff"], "0xffff00", "identity",
               ["0x00001000121f21ff", "0x0001000012112fff"], 48),
        _entry("L1", "others", ["ff0fff0fff"], "0xf0000000000", "diagonal", ["0x11202ff2fff1fff"], 12),
        _entry("free8", "free", [], "0xffffffffffffffff", "full", ["0x01111111ffffffff"], 36),
        _entry("free9", "free", [], "0xffffffffffffffff", "full", ["0x011111111fffffff"], 36),
        _entry("free10", "free", [], "0xffffffffffffffff", "full", ["0x0111111111ffffff"], 36),
        _entry("free11", "free", [], "0xffffffffffffffff", "full", ["0x01111111111fffff"], 36),
        _entry("free12", "free", [], "0xffffffffffffffff", "full", ["0x011111111111ffff"], 36),
        # variant boards (min24/min33/min34 symmetry): listed so that a build
        # request fails with a clear message rather than "unknown pattern"
        _entry("2x4", "variant", [], "0xffffffff0000", "min24", ["0xffff00000000ffff"], 48),
        _entry("3x3", "variant", [], "0xfff0fff0fff00000", "min33", ["0x000f000f000fffff"], 60),
        _entry("3x4", "variant", [], "0xffffffffffff0000", "min34", ["0x000000000000ffff"], 60),
        _entry("3x405", "variant", [], "0xffffffffffff0000", "min34", ["0x0001218912abffff"], -1900),
    )
}


def steps_and_docheck(tile_sum: int, target: int, extra_steps: int) -> Tuple[int, int]:
    """Number of layers and the success-check threshold for a build
    (engine_core/BookBuilder.py:124-128).  ``target`` is log2 of the tile."""
    target_tile = 2 ** int(target)
    steps = int(target_tile / 2 + extra_steps)
    docheck_step = int(target_tile / 2) - tile_sum % target_tile // 2
    return steps, docheck_step


def build_params(pattern: str, target: int) -> Dict[str, object]:
    """Everything a classic build of ``pattern`` at 2**target needs.  The
    seeds are the catalogue's; start_build widens those of the free*
    patterns with generate_free_inits (BookBuilder.py:272-281)."""
    meta = pattern_catalog[pattern]
    if meta.category == "variant":
        raise NotImplementedError(f"pattern {pattern} uses variant boards ({meta.canonical_mode} symmetry), "
                                  "which this engine does not build")
    steps, docheck = steps_and_docheck(meta.tile_sum, target, meta.extra_steps)
    # BookBuilder.py:272-285: free patterns are always is_free; otherwise
    # fewer than 4 fixed cells and a small sum.
    is_free = pattern.startswith("free") or (len(meta.fixed_positions) < 4 and meta.tile_sum < 180000)
    return {
        "pattern": pattern,
        "target": int(target),
        "steps": steps,
        "docheck_step": docheck,
        "masks": list(meta.pattern_masks),
        "shifts": list(meta.success_shifts),
        "symm_mode": meta.symm_mode,
        "seeds": list(meta.seed_boards),
        "tile_sum": meta.tile_sum,
        "is_free": is_free,
    }


def ad_params(pattern: str, target: int, small_tile_sum_limit: int = 96) -> Dict[str, object]:
    """build_params plus what the advanced (masked) builder needs
    (engine_core/BookBuilder.py:90-108 with Config.py:250-256): the number of
    32768 tiles of the first seed that no mask pins (``n_free_32k``), the
    shifts of the pinned ones (``fixed_32k_shifts``, cells that every mask
    holds at 0xF) and the small-tile sum limit (SmallTileSumLimit, default
    96)."""
    p = build_params(pattern, target)
    meta = pattern_catalog[pattern]
    seed = int(meta.seed_boards[0])
    count = sum(1 for k in range(16) if (seed >> (4 * k)) & 0xF == 0xF)
    fixed = list(meta.fixed_positions)
    p.update(n_free_32k=count - len(fixed), fixed_32k_shifts=fixed,
             small_tile_sum_limit=int(small_tile_sum_limit))
    return p
