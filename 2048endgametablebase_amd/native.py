"""ctypes binding of libegtb_hip.so (the C ABI in include/egtb.h).

The library is the product: every build and lookup runs through it on the
GPU.  There is no CPU fallback -- if the library is missing or fails to
load, ``lib()`` raises.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libegtb_hip.so")
CSRC = os.path.join(_HERE, "csrc")

u64p = ctypes.POINTER(ctypes.c_uint64)
u32p = ctypes.POINTER(ctypes.c_uint32)
f32p = ctypes.POINTER(ctypes.c_float)


class EgtbError(RuntimeError):
    """A failure reported by libegtb_hip.so (message from egtb_last_error)."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


class Pattern(ctypes.Structure):
    _fields_ = [
        ("masks", u64p),
        ("n_masks", ctypes.c_uint32),
        ("shifts", ctypes.POINTER(ctypes.c_uint8)),
        ("n_shifts", ctypes.c_uint32),
        ("symm_mode", ctypes.c_int32),
    ]


class Options(ctypes.Structure):
    _fields_ = [
        ("target", ctypes.c_int32),
        ("steps", ctypes.c_int32),
        ("docheck_step", ctypes.c_int32),
        ("pathname", ctypes.c_char_p),
        ("is_free", ctypes.c_int32),
        ("spawn_rate4", ctypes.c_double),
        ("success_rate_dtype", ctypes.c_char_p),
        ("write_files", ctypes.c_int32),
        ("keep_raw", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("radix_bits", ctypes.c_int32),
        ("sort_schedule", ctypes.c_int32),
        ("chunk_mkeys", ctypes.c_int32),
        ("host_tables", ctypes.c_int32),
        ("write_raw_files", ctypes.c_int32),
        ("deletion_threshold", ctypes.c_double),
        ("solve_hash_mb", ctypes.c_int32),
        ("forward_hash_steps", ctypes.c_int32),
        ("solve_hash_layout", ctypes.c_int32),
        ("optimal_branch_only", ctypes.c_int32),
        ("compress", ctypes.c_int32),
        ("own_streams", ctypes.c_int32),
    ]


class BatchJob(ctypes.Structure):
    _fields_ = [
        ("seeds", u64p),
        ("n_seeds", ctypes.c_size_t),
        ("pattern", Pattern),
        ("target", ctypes.c_int32),
        ("steps", ctypes.c_int32),
        ("docheck_step", ctypes.c_int32),
        ("reserved0", ctypes.c_int32),
        ("spawn_rate4", ctypes.c_double),
    ]


class BatchOptions(ctypes.Structure):
    _fields_ = [
        ("device", ctypes.c_int32),
        ("radix_bits", ctypes.c_int32),
        ("sort_schedule", ctypes.c_int32),
        ("keep_raw", ctypes.c_int32),
        ("forward_hash_steps", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 3),
    ]


class ServerOptions(ctypes.Structure):
    _fields_ = [
        ("device", ctypes.c_int32),
        ("slots", ctypes.c_int32),
        ("max_batch", ctypes.c_uint32),
        ("reserved0", ctypes.c_int32),
        ("memory_budget", ctypes.c_int64),
        ("reserved", ctypes.c_int32 * 4),
    ]


class LoadgenOptions(ctypes.Structure):
    _fields_ = [
        ("qps", ctypes.c_double),
        ("duration_s", ctypes.c_double),
        ("seed", ctypes.c_uint64),
        ("n_games", ctypes.c_int32),
        ("table_id", ctypes.c_int32),
        ("pool", ctypes.POINTER(ctypes.c_uint64)),
        ("pool_size", ctypes.c_size_t),
    ]


class LoadgenReport(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("qps_offered", "qps_achieved", "p50_ms", "p90_ms", "p99_ms",
                                               "p999_ms", "max_ms", "mean_ms", "submit_seconds")] + \
               [(n, ctypes.c_uint64) for n in ("completed", "batches", "max_batch", "fallbacks")]


class AdPattern(ctypes.Structure):
    _fields_ = [
        ("masks", u64p),
        ("n_masks", ctypes.c_uint32),
        ("symm_mode", ctypes.c_int32),
        ("n_free_32k", ctypes.c_int32),
        ("fixed_32k_shifts", ctypes.POINTER(ctypes.c_uint8)),
        ("n_fixed", ctypes.c_uint32),
        ("small_tile_sum_limit", ctypes.c_uint32),
        ("target", ctypes.c_int32),
    ]


class AdOptions(ctypes.Structure):
    _fields_ = [
        ("steps", ctypes.c_int32),
        ("pathname", ctypes.c_char_p),
        ("keep_layers", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("radix_bits", ctypes.c_int32),
        ("sort_schedule", ctypes.c_int32),
        ("chunk_mkeys", ctypes.c_int32),
        ("forward_dedup", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 2),
    ]


class AdSolveOptions(ctypes.Structure):
    _fields_ = [
        ("success_rate_dtype", ctypes.c_char_p),
        ("spawn_rate4", ctypes.c_double),
        ("deletion_threshold", ctypes.c_double),
        ("keep_books", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 3),
    ]


# Every symbol the header declares, with its ctypes signature.
SIGNATURES = {
    "egtb_build": (ctypes.c_int, [u64p, ctypes.c_size_t, ctypes.POINTER(Pattern), ctypes.POINTER(Options)]),
    "egtb_get_build_progress": (None, [u32p, u32p]),
    "egtb_last_error": (ctypes.c_char_p, []),
    "egtb_build_ex": (ctypes.c_int, [u64p, ctypes.c_size_t, ctypes.POINTER(Pattern), ctypes.POINTER(Options),
                                     ctypes.POINTER(ctypes.c_void_p)]),
    "egtb_result_layer_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, u64p, u64p, u64p, u64p]),
    "egtb_result_steps": (ctypes.c_int, [ctypes.c_void_p]),
    "egtb_result_raw": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, u64p, ctypes.c_size_t,
                                       ctypes.POINTER(ctypes.c_size_t)]),
    "egtb_result_book": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, u64p, ctypes.c_void_p, ctypes.c_size_t,
                                        ctypes.POINTER(ctypes.c_size_t)]),
    "egtb_result_timing": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    "egtb_result_memory": (ctypes.c_int, [ctypes.c_void_p, u64p, u64p, ctypes.POINTER(ctypes.c_double), u64p, u64p]),
    "egtb_result_forward": (ctypes.c_int, [ctypes.c_void_p, u64p, u64p]),
    "egtb_result_free": (None, [ctypes.c_void_p]),
    "egtb_ad_generate": (ctypes.c_int, [u64p, ctypes.c_size_t, ctypes.POINTER(AdPattern), ctypes.POINTER(AdOptions),
                                        ctypes.POINTER(ctypes.c_void_p)]),
    "egtb_ad_result_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), u64p,
                                           ctypes.POINTER(ctypes.c_double), u64p, u64p, u64p]),
    "egtb_ad_result_layer": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, u64p, ctypes.c_size_t,
                                            ctypes.POINTER(ctypes.c_size_t)]),
    "egtb_ad_result_forward": (ctypes.c_int, [ctypes.c_void_p, u64p, u64p]),
    "egtb_ad_result_free": (None, [ctypes.c_void_p]),
    "egtb_ad_build": (ctypes.c_int, [u64p, ctypes.c_size_t, ctypes.POINTER(AdPattern), ctypes.POINTER(AdOptions),
                                     ctypes.POINTER(AdSolveOptions), ctypes.POINTER(ctypes.c_void_p)]),
    "egtb_ad_result_solve_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), u64p, u64p]),
    "egtb_ad_result_book": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, u64p, ctypes.c_void_p,
                                           ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t),
                                           ctypes.POINTER(ctypes.c_int)]),
    "egtb_batch_build": (ctypes.c_int, [ctypes.POINTER(BatchJob), ctypes.c_int, ctypes.POINTER(BatchOptions),
                                        ctypes.POINTER(ctypes.c_void_p)]),
    "egtb_batch_timing": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    "egtb_batch_layer_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, u64p, u64p]),
    "egtb_batch_book": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, u64p, f32p, ctypes.c_size_t,
                                       ctypes.POINTER(ctypes.c_size_t)]),
    "egtb_batch_free": (None, [ctypes.c_void_p]),
    "egtb_table_open": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(Pattern), ctypes.c_int,
                                       ctypes.c_int64, ctypes.c_char_p, ctypes.c_double, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_void_p)]),
    "egtb_table_from_result": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Pattern), ctypes.c_int,
                                              ctypes.c_int64, ctypes.c_double, ctypes.POINTER(ctypes.c_void_p)]),
    "egtb_table_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), u64p, u64p]),
    "egtb_lookup_batch": (ctypes.c_int, [ctypes.c_void_p, u64p, ctypes.c_size_t, f32p,
                                         ctypes.POINTER(ctypes.c_int32)]),
    "egtb_lookup_batch_noalloc": (ctypes.c_int, [ctypes.c_void_p, u64p, ctypes.c_size_t, f32p,
                                                 ctypes.POINTER(ctypes.c_int32)]),
    "egtb_lookup_batch_f64": (ctypes.c_int, [ctypes.c_void_p, u64p, ctypes.c_size_t,
                                             ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32),
                                             ctypes.POINTER(ctypes.c_int32)]),
    "egtb_table_free": (None, [ctypes.c_void_p]),
    "egtb_table_set_prefer_max": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "egtb_sort_unique_u64": (ctypes.c_int, [u64p, ctypes.c_size_t, ctypes.c_int, u64p,
                                             ctypes.POINTER(ctypes.c_size_t)]),
    "egtb_zbook_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, u64p, ctypes.c_void_p, ctypes.c_size_t,
                                       ctypes.POINTER(ctypes.c_size_t)]),
    "egtb_zbook_write": (ctypes.c_int, [ctypes.c_char_p, u64p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]),
    "egtb_board_moves": (ctypes.c_int, [u64p, ctypes.c_size_t, ctypes.c_int32, u64p, u64p, ctypes.c_int]),
    "egtb_ai_search": (ctypes.c_int, [u64p, ctypes.c_size_t, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_int32),
                                      ctypes.POINTER(ctypes.c_int32)]),
    "egtb_ai_table": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int16), ctypes.c_int]),
    "egtb_ai_evaluate": (ctypes.c_int, [u64p, ctypes.c_size_t, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_int32)]),
    "egtb_server_create": (ctypes.c_int, [ctypes.POINTER(ServerOptions), ctypes.POINTER(ctypes.c_void_p)]),
    "egtb_server_attach": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]),
    "egtb_server_detach": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "egtb_server_submit": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, u64p, ctypes.c_size_t, u64p]),
    "egtb_server_poll": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, u64p, f32p, ctypes.POINTER(ctypes.c_int32),
                                        ctypes.POINTER(ctypes.c_size_t), ctypes.c_int32]),
    "egtb_server_stats": (ctypes.c_int, [ctypes.c_void_p, u64p, u64p, u64p]),
    "egtb_server_destroy": (None, [ctypes.c_void_p]),
    "egtb_server_loadgen": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(LoadgenOptions),
                                           ctypes.POINTER(LoadgenReport)]),
    "sort_uint64": (None, [u64p, ctypes.c_size_t, ctypes.c_bool]),
    "unique_sorted_u64_inplace": (ctypes.c_size_t, [u64p, ctypes.c_size_t]),
    "egtb_set_key_set_sizing": (ctypes.c_int, [ctypes.c_double, ctypes.c_uint64]),
    "egtb_set_tile_tuning": (ctypes.c_int, [ctypes.c_int, u64p, ctypes.POINTER(ctypes.c_int32), ctypes.c_int]),
    "egtb_set_sort_tuning": (ctypes.c_int, [u64p, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                            ctypes.c_int]),
    "egtb_bench_sort": (ctypes.c_int, [u64p, ctypes.c_size_t, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_double)]),
    "egtb_stats_enable": (None, [ctypes.c_int]),
    "egtb_stats_reset": (None, []),
    "egtb_stats_get": (ctypes.c_int, [ctypes.c_char_p, u64p, ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_double)]),
}

_lib = None
_lock = threading.Lock()


def build_library(jobs: int = 8) -> str:
    """Compile libegtb_hip.so for gfx950 in-tree (hipcc cross-compiles
    without a GPU)."""
    subprocess.run(["make", "-s", "-C", CSRC, f"-j{jobs}"], check=True)
    if not os.path.exists(LIB_PATH):
        raise EgtbError(-5, f"build did not produce {LIB_PATH}")
    return LIB_PATH


def lib():
    """Load the HIP library (once).  Raises when it is missing."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise EgtbError(-5, f"{LIB_PATH} is not built; run __graft_entry__.build() "
                                    "or make -C 2048endgametablebase_amd/csrc")
            L = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            _lib = L
            _install_cached_tuning(L)
        return _lib


TUNING_FILE = os.path.join(_HERE, "tuning", "sort_gfx950.json")


def _install_cached_tuning(L) -> None:
    """Install the autotuner's committed sort table, when there is one."""
    import json

    try:
        with open(TUNING_FILE) as fh:
            data = json.load(fh)
        rows = data["rows"]
    except (OSError, ValueError, KeyError):
        return
    install_sort_tuning(rows, L)
    for kind, name in enumerate(TILE_KINDS):
        install_tile_tuning(kind, data.get("tiles", {}).get(name, []), L)


# kinds of egtb_set_tile_tuning
TILE_KINDS = ("radix_items", "lookup_threads")


def install_tile_tuning(kind: int, rows, L=None) -> None:
    """rows: [{"max_n": int, "value": int}, ...] ascending (value 0 = rule)."""
    L = L or lib()
    n = len(rows)
    max_n = (ctypes.c_uint64 * max(n, 1))(*[int(r["max_n"]) for r in rows])
    val = (ctypes.c_int32 * max(n, 1))(*[int(r["value"]) for r in rows])
    rc = L.egtb_set_tile_tuning(int(kind), max_n, val, n)
    if rc != 0:
        raise EgtbError(rc, last_error())


def install_sort_tuning(rows, L=None) -> None:
    """rows: [{"max_n": int, "bits": int, "schedule": int}, ...] ascending."""
    L = L or lib()
    n = len(rows)
    max_n = (ctypes.c_uint64 * max(n, 1))(*[int(r["max_n"]) for r in rows])
    bits = (ctypes.c_int32 * max(n, 1))(*[int(r["bits"]) for r in rows])
    sched = (ctypes.c_int32 * max(n, 1))(*[int(r["schedule"]) for r in rows])
    rc = L.egtb_set_sort_tuning(max_n, bits, sched, n)
    if rc != 0:
        raise EgtbError(rc, last_error())


def last_error() -> str:
    msg = lib().egtb_last_error()
    return msg.decode(errors="replace") if msg else ""


def check(rc: int) -> None:
    if rc != 0:
        raise EgtbError(rc, last_error() or f"libegtb_hip error {rc}")


def gpu_available() -> bool:
    """True when a HIP device is visible (does not load the library)."""
    try:
        import torch

        return bool(torch.cuda.is_available())
    except Exception:
        return False
