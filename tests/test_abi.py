"""The C-ABI library loads and exports every entry point include/baas_hip.h
declares (no GPU needed: nothing is called)."""
import ctypes
import os

import pytest

from baas_amd import _lib


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libbaas_hip.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    declared = _lib.declared_symbols()
    assert len(declared) >= 30
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing
    # every symbol the ctypes layer binds is declared in the header
    assert set(_lib._SIGS) <= set(declared), set(_lib._SIGS) - set(declared)


def test_status_strings_without_gpu():
    lib = _lib.load()
    assert lib.baas_status_string(_lib.TIMEOUT) == b"ready-flag wait timed out"
    assert lib.baas_status_string(12345) == b"unknown status"


def test_product_does_not_import_the_oracle():
    """The product package never loads the checker."""
    import ast
    root = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "baas_amd")
    for dirpath, _, files in os.walk(root):
        for f in files:
            p = os.path.join(dirpath, f)
            if f.endswith(".py"):
                tree = ast.parse(open(p).read())
                for node in ast.walk(tree):
                    if isinstance(node, ast.Import):
                        assert all(not a.name.startswith("oracle") for a in node.names), p
                    if isinstance(node, ast.ImportFrom):
                        assert not (node.module or "").startswith("oracle"), p
            if f.endswith((".hip", ".cpp", ".h")):
                assert "oracle" not in open(p).read().lower(), p
