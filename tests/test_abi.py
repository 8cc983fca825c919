"""C-ABI library checks that need no GPU: it loads, exports every function
declared in include/plato_hip.h, the ctypes table covers them all, and a
context on a machine without a GPU fails with a status code instead of a crash."""
import ctypes
import os
import re
import subprocess

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "plato_hip.h")
LIB = os.path.join(ROOT, "plato_amd", "libplato_hip.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pe_[a-z0-9_]+)\s*\(", text)))


def test_header_parses():
    names = declared_functions()
    assert "pe_ctx_create" in names and "pe_bfs" in names and len(names) > 40


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build libplato_hip.so first (python -m plato_amd.build)"
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (pe_[a-z0-9_]+)\b", out))
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, f"declared but not exported: {missing}"


def test_ctypes_table_covers_header():
    from plato_amd import _lib
    missing = [n for n in declared_functions() if n not in _lib.SIGNATURES]
    assert not missing, f"no ctypes signature for {missing}"
    lib = _lib.load()
    for n in declared_functions():
        assert getattr(lib, n) is not None


def test_header_compiles_as_c():
    src = '#include "plato_hip.h"\nint main(void){return 0;}\n'
    p = subprocess.run(["gcc", "-x", "c", "-std=c99", "-Wall", "-Werror", "-I",
                        os.path.dirname(HEADER), "-", "-o", os.devnull], input=src, text=True,
                       capture_output=True)
    assert p.returncode == 0, p.stderr


def test_null_arguments_are_rejected():
    from plato_amd import _lib
    lib = _lib.load()
    assert lib.pe_ctx_create(0, None) == 1
    assert lib.pe_graph_get_info(None, None) == 1
    assert lib.pe_vec_destroy(None) == 1
