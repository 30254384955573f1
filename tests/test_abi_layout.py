"""The ctypes mirrors in native.py lay their structs out exactly as
include/egtb.h does: a small C program built with gcc against the header
prints sizeof and every field's offset, and ctypes must agree."""
import ctypes
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

# ctypes class name in native.py -> C struct type in egtb.h
STRUCTS = {
    "Pattern": "egtb_pattern",
    "Options": "egtb_options",
    "BatchJob": "egtb_batch_job",
    "BatchOptions": "egtb_batch_options",
    "AdPattern": "egtb_ad_pattern",
    "AdOptions": "egtb_ad_options",
    "AdSolveOptions": "egtb_ad_solve_options",
    "ServerOptions": "egtb_server_options",
    "LoadgenOptions": "egtb_loadgen_options",
    "LoadgenReport": "egtb_loadgen_report",
}


def _c_layout(tmp_path, native):
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "egtb.h"', "int main(void) {"]
    for py, c in STRUCTS.items():
        lines.append(f'    printf("{py} sizeof %zu\\n", sizeof({c}));')
        for name, *_ in getattr(native, py)._fields_:
            lines.append(f'    printf("{py} {name} %zu\\n", offsetof({c}, {name}));')
    lines += ["    return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", INCLUDE, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return {tuple(l.split()[:2]): int(l.split()[2]) for l in out.splitlines()}


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_ctypes_structs_match_header(pkg, tmp_path):
    native = pkg.native
    c = _c_layout(tmp_path, native)
    for py in STRUCTS:
        cls = getattr(native, py)
        assert c[(py, "sizeof")] == ctypes.sizeof(cls), py
        for name, *_ in cls._fields_:
            assert c[(py, name)] == getattr(cls, name).offset, (py, name)
