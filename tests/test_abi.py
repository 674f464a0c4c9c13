"""The C-ABI library loads and exports every entry point include/d3fgpu.h
declares (no compute calls: this runs without a GPU)."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "d3fgpu.h")
LIBRARY = os.path.join(ROOT, "d3-force-wasm_amd", "libd3fgpu.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = "\n".join(l for l in text.splitlines() if not l.strip().startswith("typedef"))
    names = re.findall(r"^\s*(?:[A-Za-z_][\w\s\*]*?)\b([a-z_][a-z0-9_]*)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("typedef",)))


def test_header_declares_the_reference_exports():
    names = set(declared())
    # src/core/root.zig exports (the wasm alloc/free are replaced by caller-owned buffers)
    for n in ["sim_create", "sim_destroy", "sim_set_node_count", "sim_ptr", "force_create", "force_destroy",
              "force_set_count", "force_ptr", "force_set_scalar", "force_get_scalar", "force_run",
              "sim_integrate", "tree_build", "tree_info", "tree_order", "sim_snapshot", "core_set_random",
              "sim_sync"]:
        assert n in names, n


@pytest.mark.skipif(not os.path.exists(LIBRARY), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIBRARY)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_facade_fails_loudly_without_the_library(tmp_path):
    import d3force_amd.core as core
    with pytest.raises(OSError):
        core.Core(str(tmp_path / "missing.so"))
