"""The C-ABI library loads and exports every entry point include/chili_gpu.h
declares, and the header compiles as C (no GPU calls: these run anywhere)."""
import ctypes
import os
import re
import subprocess

import pytest

from chili_amd import lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "chili_gpu.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(chili_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared()
    assert "chili_mc_iterate" in names and "chili_flow_train" in names
    assert len(names) >= 40


def test_library_exports_every_declared_symbol():
    L = lib.load()
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, f"declared but not exported: {missing}"
    # the binding's list agrees with the header
    assert sorted(set(lib.EXPORTS)) == declared()


def test_header_compiles_as_c(tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include "chili_gpu.h"\nint main(void) { chili_stats s; (void)s; return 0; }\n')
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.dirname(HEADER), "-c", str(src), "-o",
                        str(tmp_path / "t.o")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_errors_cross_the_boundary_as_codes():
    # host-only entry points: bad arguments return a code and set the message
    L = lib.load()
    rc = L.chili_grid_adapt(0, 0, None, 1.0, None)
    assert rc != 0
    assert len(L.chili_last_error()) > 0


def test_specialized_kernel_source_compiles():
    # the run-time specialisation compiles for gfx950 (hipRTC, no GPU needed)
    from chili_amd import processes as P
    ch = lib.Channels(P.gluon_process(3))
    src = ch.kernel_source()
    assert "chili_prog_0" in src and "chili_k_main" in src
    secs = ch.compile_check()
    assert secs > 0
