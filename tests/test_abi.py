"""The C-ABI library loads on the CPU and exports every symbol the headers declare."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(inc, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?\w[\w\s\*]*?\b(coffee_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_symbols()
    for n in ("coffee_ctx_create", "coffee_ez_step", "coffee_last_error", "coffee_total_energy"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from coffeegpu_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libcoffee_hip.so is not built")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in sorted(declared_symbols()) if not hasattr(lib, n)]
    assert not missing, missing
    # ctypes signatures cover every declared symbol
    assert declared_symbols() <= set(_lib.SIGNATURES)


def test_abi_version_and_create_error_without_gpu():
    from coffeegpu_amd import _lib

    lib = _lib.load()
    assert lib.coffee_abi_version() == 1
    # bad scalar size is rejected before any device call
    g = _lib.CoffeeGrid()
    p = _lib.CoffeeParams()
    ctx = ctypes.c_void_p()
    rc = lib.coffee_ctx_create(0, 3, ctypes.byref(g), ctypes.byref(p), ctypes.byref(ctx))
    assert rc == -1
    assert b"scalar_bytes" in lib.coffee_last_error(None)
