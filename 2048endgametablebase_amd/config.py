"""Build configuration: the subset of the reference's SingletonConfig that
the classic build path reads (Config.py:423-528 defaults;
engine_core/BookBuilder.py:44-73 maps them to RunOptions)."""
from __future__ import annotations

import threading

DEFAULTS = {
    "compress": False,
    "optimal_branch_only": False,
    "compress_temp_files": False,
    "advanced_algo": False,
    "direct_io": True,
    "direct_io_queue_depth": 16,
    "direct_io_chunk_mib": 8,
    "deletion_threshold": 0.0,
    "SmallTileSumLimit": 96,
    "4_spawn_rate": 0.1,
    "success_rate_dtype": "uint32",
    "filepath_map": {},
    "device": 0,
}


class SingletonConfig:
    """Process-wide configuration dict (same keys as the reference)."""

    _instance = None
    _lock = threading.Lock()

    def __new__(cls):
        with cls._lock:
            if cls._instance is None:
                inst = super().__new__(cls)
                inst.config = dict(DEFAULTS)
                inst.config["filepath_map"] = {}
                cls._instance = inst
            return cls._instance
