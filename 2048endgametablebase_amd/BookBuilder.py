"""Classic tablebase builder: the reference's Python build surface
(engine_core/BookBuilder.py) over the HIP engine.

``start_build(pattern, target, pathname)`` keeps the reference's signature,
file set and error behaviour: it writes ``pathname + "config.txt"``, runs the
whole forward/backward build on the GPU, writes ``<pathname><i>.book`` for
every layer plus ``<pathname>stats.txt``, returns True, and raises on
failure.
"""
from __future__ import annotations

import ctypes
from itertools import combinations
from typing import Iterable, Optional, Sequence

import numpy as np

from . import native
from .config import SingletonConfig
from .patterns import build_params, pattern_catalog, steps_and_docheck

# success_rate_dtype -> element type of the rates (the 1-float types store
# rate - 1 in a float; FormationRuntime.h:256-312)
_RATE_DTYPES = {"float32": np.float32, "float64": np.float64, "uint32": np.uint32, "uint64": np.uint64,
                "1-float32": np.float32, "1-float64": np.float64}


def save_config_to_txt(output_path: str) -> None:
    """engine_core/BookBuilder.py:130-146: the option snapshot of a build."""
    keys = [
        "compress",
        "optimal_branch_only",
        "compress_temp_files",
        "advanced_algo",
        "direct_io",
        "direct_io_queue_depth",
        "direct_io_chunk_mib",
        "deletion_threshold",
        "4_spawn_rate",
        "success_rate_dtype",
    ]
    cfg = SingletonConfig().config
    with open(output_path, "w", encoding="utf-8") as fh:
        for key in keys:
            fh.write(f"{key}: {str(cfg.get(key, '?'))}\n")


class _Args:
    """ctypes argument block kept alive for the duration of a native call."""

    def __init__(self, seeds, masks, shifts, symm, target, steps, docheck, pathname, is_free,
                 spawn_rate4, dtype, write_files, keep_raw, device, radix_bits, sort_schedule=0,
                 chunk_mkeys=0, host_tables=False, write_raw_files=False, deletion_threshold=0.0,
                 solve_hash_mb=0, forward_hash_steps=0, solve_hash_layout=0, optimal_branch_only=False,
                 compress=False, own_streams=False):
        self.seeds = np.ascontiguousarray(np.asarray(seeds, dtype=np.uint64))
        self.masks = np.ascontiguousarray(np.asarray(list(masks) or [0], dtype=np.uint64))
        self.shifts = np.ascontiguousarray(np.asarray(list(shifts) or [0], dtype=np.uint8))
        self.path = (pathname or "").encode()
        self.dtype = str(dtype).encode()
        self.pattern = native.Pattern(
            self.masks.ctypes.data_as(native.u64p), len(masks),
            self.shifts.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), len(shifts), int(symm))
        self.options = native.Options()
        o = self.options
        o.target = int(target)
        o.steps = int(steps)
        o.docheck_step = int(docheck)
        o.pathname = self.path
        o.is_free = int(bool(is_free))
        o.spawn_rate4 = float(spawn_rate4)
        o.success_rate_dtype = self.dtype
        o.write_files = int(bool(write_files))
        o.keep_raw = int(bool(keep_raw))
        o.device = int(device)
        o.radix_bits = int(radix_bits)
        o.sort_schedule = int(sort_schedule)
        o.chunk_mkeys = int(chunk_mkeys)
        o.host_tables = int(bool(host_tables))
        o.write_raw_files = int(write_raw_files)
        o.deletion_threshold = float(deletion_threshold)
        o.solve_hash_mb = int(solve_hash_mb)
        o.forward_hash_steps = int(forward_hash_steps)
        o.solve_hash_layout = int(solve_hash_layout)
        o.optimal_branch_only = int(bool(optimal_branch_only))
        o.compress = int(bool(compress))
        o.own_streams = int(bool(own_streams))


class BuildResult:
    """Solved layers of a build kept resident in HBM (egtb_result)."""

    def __init__(self, handle: int, dtype: str):
        self._h = ctypes.c_void_p(handle)
        self.dtype = dtype
        self.steps = native.lib().egtb_result_steps(self._h)

    @property
    def handle(self):
        return self._h

    def layer_info(self, i: int):
        L = native.lib()
        v = [ctypes.c_uint64() for _ in range(4)]
        native.check(L.egtb_result_layer_info(self._h, i, *[ctypes.byref(x) for x in v]))
        return tuple(int(x.value) for x in v)  # (n_raw, n_live, c2, c4)

    def raw(self, i: int) -> np.ndarray:
        L = native.lib()
        n = ctypes.c_size_t()
        native.check(L.egtb_result_raw(self._h, i, None, 0, ctypes.byref(n)))
        out = np.empty(n.value, dtype=np.uint64)
        native.check(L.egtb_result_raw(self._h, i, out.ctypes.data_as(native.u64p), out.size,
                                       ctypes.byref(n)))
        return out

    def book(self, i: int):
        L = native.lib()
        n = ctypes.c_size_t()
        native.check(L.egtb_result_book(self._h, i, None, None, 0, ctypes.byref(n)))
        keys = np.empty(n.value, dtype=np.uint64)
        rates = np.empty(n.value, dtype=_RATE_DTYPES[self.dtype])
        native.check(L.egtb_result_book(self._h, i, keys.ctypes.data_as(native.u64p),
                                        rates.ctypes.data_as(ctypes.c_void_p), keys.size, ctypes.byref(n)))
        return keys, rates

    def memory(self):
        """(peak HBM bytes, streamed bytes, stream ms, chunked steps, chunks)."""
        a, b, d, e = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        c = ctypes.c_double()
        native.check(native.lib().egtb_result_memory(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                                     ctypes.byref(d), ctypes.byref(e)))
        return {"peak_device_bytes": a.value, "streamed_bytes": b.value, "stream_ms": c.value,
                "chunked_steps": d.value, "chunks": e.value}

    def forward(self):
        """(forward steps deduplicated in key sets, key-set regrowths)."""
        a, b = ctypes.c_uint64(), ctypes.c_uint64()
        native.check(native.lib().egtb_result_forward(self._h, ctypes.byref(a), ctypes.byref(b)))
        return {"hash_steps": a.value, "set_regrows": b.value}

    def timing(self):
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        native.check(native.lib().egtb_result_timing(self._h, ctypes.byref(a), ctypes.byref(b),
                                                     ctypes.byref(c)))
        return a.value, b.value, c.value

    def close(self):
        if self._h is not None and self._h.value:
            native.lib().egtb_result_free(self._h)
        self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_pattern_build(arr_init, masks: Sequence[int], shifts: Sequence[int], symm_mode: int, target: int,
                      steps: int, docheck_step: int, pathname: str = "", *, is_free: bool = False,
                      spawn_rate4: float = 0.1, success_rate_dtype: str = "float32",
                      write_files: bool = True, keep: bool = False, keep_raw: bool = False,
                      device: int = 0, radix_bits: int = 0, sort_schedule: int = 0, chunk_mkeys: int = 0,
                      host_tables: bool = False, write_raw_files: int = 0,
                      deletion_threshold: float = 0.0, solve_hash_mb: int = 0,
                      forward_hash_steps: int = 0, solve_hash_layout: int = 0,
                      optimal_branch_only: bool = False, compress: bool = False,
                      own_streams: bool = False) -> Optional[BuildResult]:
    """The native build call (reference formation_core.run_pattern_build,
    native_core/src/bindings_formation.cpp:220-229) with the PatternSpec and
    RunOptions fields spelled out.  ``keep=True`` returns the solved layers
    resident in HBM as a BuildResult.  With files, an interrupted build in
    the same folder is resumed: layers whose .book exists are not solved
    again (BookSolver.cpp:476-511)."""
    a = _Args(arr_init, masks, shifts, symm_mode, target, steps, docheck_step, pathname, is_free,
              spawn_rate4, success_rate_dtype, write_files, keep_raw, device, radix_bits, sort_schedule,
              chunk_mkeys, host_tables, write_raw_files, deletion_threshold, solve_hash_mb, forward_hash_steps,
              solve_hash_layout, optimal_branch_only, compress, own_streams)
    L = native.lib()
    seeds_p = a.seeds.ctypes.data_as(native.u64p)
    if keep or keep_raw:
        h = ctypes.c_void_p()
        native.check(L.egtb_build_ex(seeds_p, a.seeds.size, ctypes.byref(a.pattern), ctypes.byref(a.options),
                                     ctypes.byref(h)))
        return BuildResult(h.value, success_rate_dtype)
    native.check(L.egtb_build(seeds_p, a.seeds.size, ctypes.byref(a.pattern), ctypes.byref(a.options)))
    return None


def get_build_progress():
    """(current, total) of the running build (bindings_formation.cpp:147-153)."""
    cur, tot = ctypes.c_uint32(), ctypes.c_uint32()
    native.lib().egtb_get_build_progress(ctypes.byref(cur), ctypes.byref(tot))
    return int(cur.value), int(tot.value)


def _should_retry_build_resume(exc: Exception) -> bool:
    message = str(exc).strip()
    return message == "The length multiplier is not big enough. Please restart." or \
        message.startswith("length multiplier ")


def _run_with_single_resume_retry(build_fn) -> None:
    """engine_core/BookBuilder.py:246-261.  The GPU build sizes every buffer
    exactly, so the capacity error it retries never occurs here; the retry
    is kept for the same error contract."""
    retried = False
    while True:
        try:
            build_fn()
            return
        except Exception as exc:
            if retried or not _should_retry_build_resume(exc):
                raise
            retried = True


def board_moves(boards, symm_mode: int = 0, device: int = 0):
    """(moves, canonical) of many boards on the device: moves is n x 4 in
    move_all_dir order (left, right, up, down), canonical the boards'
    canonical forms under ``symm_mode`` (egtb_board_moves)."""
    b = np.ascontiguousarray(np.asarray(boards, dtype=np.uint64).ravel())
    out4 = np.empty((b.size, 4), dtype=np.uint64)
    canon = np.empty(b.size, dtype=np.uint64)
    native.check(native.lib().egtb_board_moves(b.ctypes.data_as(native.u64p), b.size, int(symm_mode),
                                               out4.ctypes.data_as(native.u64p),
                                               canon.ctypes.data_as(native.u64p), int(device)))
    return out4, canon


def generate_free_inits(t32ks: int, t2s: int, device: int = 0) -> np.ndarray:
    """Layer 0 of a free* pattern (engine_core/BookBuilder.py:146-181):
    every board with ``t32ks`` 32768-tiles and ``t2s`` 2-tiles (at most 1e6,
    enumerated in the same combination order), deduplicated, then each
    board's four moves (an impossible move keeps the board) that are their
    own full-symmetry canonical form, stopping after the board that brings
    the count within 3 of the number of boards; sorted and unique."""
    max_estimated = 1_000_000
    generated = []
    for pos32 in combinations(range(16), t32ks):
        base = 0
        for pos in pos32:
            base |= 15 << (4 * pos)
        remain = sorted(set(range(16)) - set(pos32))
        for pos2 in combinations(remain, t2s):
            if len(generated) >= max_estimated:
                break
            board = base
            for pos in pos2:
                board |= 1 << (4 * pos)
            generated.append(board)
    gen = np.unique(np.asarray(generated, dtype=np.uint64))
    if gen.size == 0:
        return gen
    moved, _ = board_moves(gen, 0, device)
    _, canon = board_moves(moved.ravel(), 1, device)
    keep = (moved.ravel() == canon).reshape(moved.shape)
    per_board = keep.sum(axis=1)
    # the reference's capacity stop: after the first board whose running
    # count reaches len(gen) - 3
    running = np.cumsum(per_board)
    stop = np.flatnonzero(running >= gen.size - 3)
    last = int(stop[0]) if stop.size else gen.size - 1
    return np.unique(moved[: last + 1][keep[: last + 1]])


def resolve_seeds(pattern: str, p: dict, device: int = 0) -> np.ndarray:
    """Layer 0 as start_build computes it (engine_core/BookBuilder.py:
    272-281): free* patterns whose small tiles fit the free cells as 2s get
    generate_free_inits; every other pattern keeps its catalogue seeds."""
    seeds = np.asarray(p["seeds"], dtype=np.uint64)
    if not pattern.startswith("free"):
        return seeds
    seed0 = int(seeds[0])
    num_32k = sum(1 for k in range(16) if (seed0 >> (4 * k)) & 0xF == 0xF)
    extra_tile_sum = int(p["tile_sum"]) - 32768 * num_32k
    if extra_tile_sum <= (15 - num_32k) * 2:
        return generate_free_inits(num_32k, extra_tile_sum // 2, device)
    return seeds


def start_build(pattern: str, target: int, pathname: str) -> bool:
    """Build ``pattern`` at tile 2**target into files prefixed ``pathname``
    (engine_core/BookBuilder.py:264-317, classic path)."""
    cfg = SingletonConfig().config
    if pattern not in pattern_catalog:
        raise KeyError(f"Unknown pattern: {pattern}")
    if cfg.get("advanced_algo", False):
        # _run_advanced_build (engine_core/BookBuilder.py:211-236): the
        # advanced forward pass and solver into bucket folders <i>b
        from .advanced import build_ad

        save_config_to_txt(pathname + "config.txt")
        with build_ad(pattern, target, pathname, success_rate_dtype=str(cfg.get("success_rate_dtype", "uint32")),
                      spawn_rate4=float(cfg["4_spawn_rate"]), deletion_threshold=float(cfg.get("deletion_threshold", 0.0)),
                      device=int(cfg.get("device", 0))):
            pass
        return True
    p = build_params(pattern, target)
    save_config_to_txt(pathname + "config.txt")
    dtype = str(cfg.get("success_rate_dtype", "uint32"))
    seeds = resolve_seeds(pattern, p, int(cfg.get("device", 0)))
    _run_with_single_resume_retry(lambda: run_pattern_build(
        seeds, p["masks"], p["shifts"], p["symm_mode"], p["target"], p["steps"], p["docheck_step"],
        pathname, is_free=p["is_free"], spawn_rate4=float(cfg["4_spawn_rate"]), success_rate_dtype=dtype,
        write_files=True, device=int(cfg.get("device", 0)),
        write_raw_files=bool(cfg.get("write_raw_files", False)),
        deletion_threshold=float(cfg.get("deletion_threshold", 0.0)),
        optimal_branch_only=bool(cfg.get("optimal_branch_only", False)),
        compress=bool(cfg.get("compress", False))))
    return True


def build_in_memory(pattern: str, target: int, *, success_rate_dtype: str = "float32",
                    spawn_rate4: float = 0.1, keep_raw: bool = False, device: int = 0,
                    radix_bits: int = 0, sort_schedule: int = 0, chunk_mkeys: int = 0,
                    host_tables: bool = False, solve_hash_mb: int = 0,
                    forward_hash_steps: int = 0, solve_hash_layout: int = 0,
                    own_streams: bool = False) -> BuildResult:
    """Build a catalogue pattern (layer 0 as start_build makes it) and keep
    every solved layer in HBM."""
    p = build_params(pattern, target)
    return run_pattern_build(resolve_seeds(pattern, p, device), p["masks"], p["shifts"], p["symm_mode"], p["target"], p["steps"],
                             p["docheck_step"], "", is_free=p["is_free"], spawn_rate4=spawn_rate4,
                             success_rate_dtype=success_rate_dtype, write_files=False, keep=True,
                             keep_raw=keep_raw, device=device, radix_bits=radix_bits,
                             sort_schedule=sort_schedule, chunk_mkeys=chunk_mkeys, host_tables=host_tables,
                             solve_hash_mb=solve_hash_mb, forward_hash_steps=forward_hash_steps,
                             solve_hash_layout=solve_hash_layout, own_streams=own_streams)
