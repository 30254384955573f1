"""The advanced (AD, large-tile masking) builder on the GPU: forward pass and
solver.

``generate_ad(pattern, target, pathname)`` runs what the reference's
``formation_core.run_pattern_build_ad`` does before its solver
(native_core/src/BookGeneratorAD.cpp:1256-1396, called from
run_pattern_build_ad_cpp :1421-1438): raw layer files
``<pathname>0 .. <pathname><steps-2>`` of masked boards (tiles of 64 or more
stored as 0xF), byte-equal to the reference's; with a pathname it resumes
from the raw files already there, as the reference's handle_restart_ad does.  ``run_pattern_generate_ad``
takes the reference binding's arguments (seeds, an AdvancedPatternSpec-like
dict, steps, pathname).

``build_ad(pattern, target, pathname)`` / ``run_pattern_build_ad`` run the
whole advanced build as formation_core.run_pattern_build_ad does
(BookGeneratorAD.cpp:1421-1438): the forward pass, then the solver
(BookSolverAD.cpp:2086-2242), which writes each solved layer as bucket files
``<pathname><i>b/<key>.i`` (sorted boards) and ``<key>.b`` (a row of rates
per board, one per placement of its large tiles).
"""
from __future__ import annotations

import ctypes
from typing import Dict, Optional, Sequence

import numpy as np

from . import native, patterns
from .config import SingletonConfig


class AdResult:
    """Layers of one advanced forward pass (held in HBM when kept)."""

    rate_dtype = np.float32

    def __init__(self, handle: int):
        self._h = ctypes.c_void_p(handle)
        L = native.lib()
        steps = ctypes.c_int()
        ms = ctypes.c_double()
        cand = ctypes.c_uint64()
        chunked = ctypes.c_uint64()
        skipped = ctypes.c_uint64()
        native.check(L.egtb_ad_result_info(self._h, ctypes.byref(steps), None, ctypes.byref(ms),
                                           ctypes.byref(cand), ctypes.byref(chunked), ctypes.byref(skipped)))
        self.steps = steps.value
        sizes = np.zeros(max(self.steps - 1, 1), dtype=np.uint64)
        native.check(L.egtb_ad_result_info(self._h, None, sizes.ctypes.data_as(native.u64p), None, None, None,
                                           None))
        self.sizes = sizes
        self.fwd_ms = ms.value
        self.candidates = cand.value
        self.chunked_steps = chunked.value
        self.skipped_steps = skipped.value
        hs, ov = ctypes.c_uint64(), ctypes.c_uint64()
        native.check(L.egtb_ad_result_forward(self._h, ctypes.byref(hs), ctypes.byref(ov)))
        self.hash_steps = hs.value
        self.set_overflows = ov.value

    # the solver's figures (egtb_ad_build)
    def solve_info(self):
        ms, vals, skipped = ctypes.c_double(), ctypes.c_uint64(), ctypes.c_uint64()
        native.check(native.lib().egtb_ad_result_solve_info(self._h, ctypes.byref(ms), ctypes.byref(vals),
                                                            ctypes.byref(skipped)))
        return {"solve_ms": ms.value, "values": vals.value, "skipped_layers": skipped.value}

    def book(self, layer: int, key: int):
        """Bucket ``key`` of solved layer ``layer`` (kept with keep_books):
        (boards, rates of shape (rows, cols))."""
        L = native.lib()
        rows, cols = ctypes.c_size_t(), ctypes.c_int()
        native.check(L.egtb_ad_result_book(self._h, int(layer), int(key), None, None, 0, ctypes.byref(rows),
                                           ctypes.byref(cols)))
        boards = np.zeros(rows.value, dtype=np.uint64)
        rates = np.zeros((rows.value, max(cols.value, 0)), dtype=self.rate_dtype)
        if rows.value:
            native.check(L.egtb_ad_result_book(self._h, int(layer), int(key), boards.ctypes.data_as(native.u64p),
                                               rates.ctypes.data_as(ctypes.c_void_p), rows.value,
                                               ctypes.byref(rows), ctypes.byref(cols)))
        return boards, rates

    def buckets(self, layer: int):
        """{key: (rows, cols)} of the non-empty buckets of a kept solved layer."""
        L = native.lib()
        out = {}
        for k in range(-16, 32):
            rows, cols = ctypes.c_size_t(), ctypes.c_int()
            native.check(L.egtb_ad_result_book(self._h, int(layer), k, None, None, 0, ctypes.byref(rows),
                                               ctypes.byref(cols)))
            if rows.value:
                out[k] = (rows.value, cols.value)
        return out

    def layer(self, i: int) -> np.ndarray:
        L = native.lib()
        n = ctypes.c_size_t()
        native.check(L.egtb_ad_result_layer(self._h, int(i), None, 0, ctypes.byref(n)))
        out = np.zeros(n.value, dtype=np.uint64)
        native.check(L.egtb_ad_result_layer(self._h, int(i), out.ctypes.data_as(native.u64p), out.size,
                                            ctypes.byref(n)))
        return out

    def close(self) -> None:
        if self._h:
            native.lib().egtb_ad_result_free(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_pattern_generate_ad(arr_init: Sequence[int], spec: Dict[str, object], steps: int,
                            pathname: Optional[str] = None, *, keep_layers: bool = False, device: int = 0,
                            radix_bits: int = 0, sort_schedule: int = 0, chunk_mkeys: int = 0,
                            forward_dedup: int = 0) -> AdResult:
    """The forward pass of formation_core.run_pattern_build_ad
    (bindings_formation.cpp:231-235) for ``spec`` with the fields of
    AdvancedPatternSpec (FormationRuntime.h:88-96): pattern_masks,
    symm_mode, num_free_32k, fixed_32k_shifts, small_tile_sum_limit and
    target (log2)."""
    seeds, keep, pat, opt = _ad_args(arr_init, spec, steps, pathname, keep_layers, device, radix_bits,
                                     sort_schedule, chunk_mkeys, forward_dedup)
    h = ctypes.c_void_p()
    native.check(native.lib().egtb_ad_generate(seeds.ctypes.data_as(native.u64p), seeds.size, ctypes.byref(pat),
                                               ctypes.byref(opt), ctypes.byref(h)))
    return AdResult(h.value)


RATE_DTYPES = {"float32": np.float32, "float64": np.float64, "uint32": np.uint32, "uint64": np.uint64}


def _ad_args(arr_init, spec, steps, pathname, keep_layers, device, radix_bits, sort_schedule, chunk_mkeys,
             forward_dedup):
    seeds = np.ascontiguousarray(np.asarray(arr_init, dtype=np.uint64))
    masks = np.ascontiguousarray(np.asarray(list(spec["pattern_masks"]) or [0], dtype=np.uint64))
    fixed = np.ascontiguousarray(np.asarray(list(spec["fixed_32k_shifts"]) or [0], dtype=np.uint8))
    pat = native.AdPattern()
    pat.masks = masks.ctypes.data_as(native.u64p)
    pat.n_masks = len(spec["pattern_masks"])
    pat.symm_mode = int(spec["symm_mode"])
    pat.n_free_32k = int(spec["num_free_32k"])
    pat.fixed_32k_shifts = fixed.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
    pat.n_fixed = len(spec["fixed_32k_shifts"])
    pat.small_tile_sum_limit = int(spec.get("small_tile_sum_limit", 96))
    pat.target = int(spec["target"])
    opt = native.AdOptions()
    opt.steps = int(steps)
    opt.pathname = pathname.encode() if pathname is not None else None
    opt.keep_layers = int(bool(keep_layers))
    opt.device = int(device)
    opt.radix_bits = int(radix_bits)
    opt.sort_schedule = int(sort_schedule)
    opt.chunk_mkeys = int(chunk_mkeys)
    opt.forward_dedup = int(forward_dedup)
    return seeds, (masks, fixed), pat, opt


def run_pattern_build_ad(arr_init: Sequence[int], spec: Dict[str, object], steps: int,
                         pathname: Optional[str] = None, *, success_rate_dtype: str = "float32",
                         spawn_rate4: float = 0.1, deletion_threshold: float = 0.0, keep_books: bool = False,
                         device: int = 0, radix_bits: int = 0, sort_schedule: int = 0, chunk_mkeys: int = 0,
                         forward_dedup: int = 0) -> AdResult:
    """formation_core.run_pattern_build_ad (bindings_formation.cpp:231-235):
    the forward pass and the solver for ``spec`` (AdvancedPatternSpec's
    fields, as run_pattern_generate_ad takes them).  With a pathname the
    solved layers are written as the reference writes them and a rerun
    resumes; keep_books holds every solved layer in HBM (AdResult.book)."""
    if success_rate_dtype not in RATE_DTYPES:
        raise ValueError(f"success_rate_dtype must be one of {sorted(RATE_DTYPES)}")
    seeds, keep, pat, opt = _ad_args(arr_init, spec, steps, pathname, False, device, radix_bits, sort_schedule,
                                     chunk_mkeys, forward_dedup)
    so = native.AdSolveOptions()
    so.success_rate_dtype = success_rate_dtype.encode()
    so.spawn_rate4 = float(spawn_rate4)
    so.deletion_threshold = float(deletion_threshold)
    so.keep_books = int(bool(keep_books))
    h = ctypes.c_void_p()
    native.check(native.lib().egtb_ad_build(seeds.ctypes.data_as(native.u64p), seeds.size, ctypes.byref(pat),
                                            ctypes.byref(opt), ctypes.byref(so), ctypes.byref(h)))
    r = AdResult(h.value)
    r.rate_dtype = RATE_DTYPES[success_rate_dtype]
    return r


def build_ad(pattern: str, target: int, pathname: Optional[str] = None, **kw) -> AdResult:
    """The whole advanced build of ``pattern`` at 2**target (start_build's
    layer 0, as generate_ad), solved books as bucket files under
    ``pathname`` and/or kept in HBM (keep_books=True)."""
    from .BookBuilder import resolve_seeds

    s = ad_spec(pattern, target)
    seeds = resolve_seeds(pattern, patterns.build_params(pattern, target), int(kw.get("device", 0)))
    return run_pattern_build_ad(seeds, s, s["steps"], pathname, **kw)


def ad_spec(pattern: str, target: int) -> Dict[str, object]:
    """The AdvancedPatternSpec engine_core/BookBuilder.py:90-108 builds,
    with SmallTileSumLimit from the configuration."""
    limit = int(SingletonConfig().config.get("SmallTileSumLimit", 96))
    p = patterns.ad_params(pattern, target, limit)
    return {"pattern_masks": p["masks"], "symm_mode": p["symm_mode"], "num_free_32k": p["n_free_32k"],
            "fixed_32k_shifts": p["fixed_32k_shifts"], "small_tile_sum_limit": limit, "target": int(target),
            "steps": p["steps"], "seeds": p["seeds"]}


def generate_ad(pattern: str, target: int, pathname: Optional[str] = None, **kw) -> AdResult:
    """Forward pass of the advanced build of ``pattern`` at 2**target
    (raw layer files under ``pathname`` when given), from the layer 0
    start_build uses (engine_core/BookBuilder.py:272-285: generate_free_inits
    for free* patterns whose small tiles fit the free cells as 2s)."""
    from .BookBuilder import resolve_seeds

    s = ad_spec(pattern, target)
    seeds = resolve_seeds(pattern, patterns.build_params(pattern, target), int(kw.get("device", 0)))
    return run_pattern_generate_ad(seeds, s, s["steps"], pathname, **kw)


def _main(argv=None) -> None:
    """``python -m 2048endgametablebase_amd.advanced PATTERN LOG2_TARGET [--reps N] [--solve]``:
    time the forward pass (with --solve the whole build: forward pass and
    solver), layers and books in HBM, no files, and print one JSON line."""
    import argparse
    import json
    import time

    ap = argparse.ArgumentParser()
    ap.add_argument("pattern")
    ap.add_argument("target", type=int)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--solve", action="store_true")
    a = ap.parse_args(argv)
    times = []
    extra = {}
    for _ in range(a.reps):
        t0 = time.perf_counter()
        r = build_ad(a.pattern, a.target) if a.solve else generate_ad(a.pattern, a.target)
        times.append(time.perf_counter() - t0)
        n, c = int(r.sizes.sum()), int(r.candidates)
        if a.solve:
            extra = {"fwd_ms": round(r.fwd_ms, 2), "solve_ms": round(r.solve_info()["solve_ms"], 2),
                     "rates": r.solve_info()["values"]}
        r.close()
    what = "advanced build (forward pass and solver)" if a.solve else "advanced forward pass"
    print(json.dumps(dict({"workload": f"{a.pattern}@{2 ** a.target} {what}", "states": n,
                           "boards_before_dedup": c, "ms": [round(t * 1000.0, 2) for t in times],
                           "states_per_s": round(n / min(times), 1)}, **extra)))


if __name__ == "__main__":
    _main()
