"""MI355X-native engine for the classic 2048 endgame-tablebase path.

Import with ``importlib.import_module("2048endgametablebase_amd")`` (the
package name starts with a digit).  Submodules:

* ``patterns``    -- pattern catalogue, steps / docheck formulas
* ``BookBuilder`` -- ``start_build(pattern, target, pathname)`` and the
                     in-memory build API over libegtb_hip.so
* ``native``      -- ctypes binding of the C ABI (include/egtb.h)
"""
from . import config, native, patterns  # noqa: F401
from .BookReader import BookReader, Table, find_classic_value  # noqa: F401
from .BookBuilder import (  # noqa: F401
    BuildResult,
    build_in_memory,
    get_build_progress,
    run_pattern_build,
    start_build,
)
