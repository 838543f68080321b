"""The C-ABI library loads on the CPU and exports every symbol the headers declare."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# headers of libcoffee_h5.so; every other include/*.h is libcoffee_hip.so's
H5_HEADERS = ("coffee_h5.h",)


def declared_symbols(headers=None):
    """coffee_* functions declared by include/*.h (all but the HDF5 layer's
    headers), or by the named headers."""
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if not fn.endswith(".h"):
            continue
        if (headers is None and fn in H5_HEADERS) or (headers is not None and fn not in headers):
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


def test_h5_library_exports_every_declared_symbol():
    from coffeegpu_amd import _h5

    if not os.path.exists(_h5.LIB_PATH):
        pytest.skip("libcoffee_h5.so not built (no HDF5 at build time)")
    lib = ctypes.CDLL(_h5.LIB_PATH)
    names = declared_symbols(H5_HEADERS)
    assert "coffee_h5_create" in names
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert names == set(_h5.SIGNATURES)


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


def _flat_grid_params(n=(16, 16, 16), guard=3):
    from coffeegpu_amd import _lib

    g = _lib.CoffeeGrid()
    p = _lib.CoffeeParams()
    for a in range(3):
        g.dims[a] = n[a] + 2 * guard
        g.guard[a] = guard
        g.delta[a] = 1.0 / n[a]
        g.sizes[a] = 1.0
        p.periodic[a] = 1
        p.size[a] = 1.0
    p.dt = 0.4 / n[0]
    p.tau = 2.0
    return g, p


def test_shift_ghost_rejected_before_any_device_call():
    """sim_params::shift_ghost > 0 (guard layers the reference also updates,
    field_solver_EZ.cpp:128-136) is not supported: the context is refused
    instead of running with frozen guards."""
    from coffeegpu_amd import _lib

    lib = _lib.load()
    g, p = _flat_grid_params()
    p.shift_ghost = 1
    ctx = ctypes.c_void_p()
    for solver in (_lib.COFFEE_SOLVER_EZ, _lib.COFFEE_SOLVER_GR_EZ):
        rc = lib.coffee_ctx_create(solver, 4, ctypes.byref(g), ctypes.byref(p), ctypes.byref(ctx))
        assert rc == -1
        assert b"shift_ghost" in lib.coffee_last_error(None)
        assert not ctx.value


def test_shift_ghost_rejected_by_python_mirror_params():
    from coffeegpu_amd.params import SimParams
    from coffeegpu_amd.solver import coffee_params

    p = coffee_params(SimParams(shift_ghost=2))
    assert p.shift_ghost == 2  # carried to the ABI, which refuses it


def test_oversized_grid_returns_err_size():
    """A grid of 2^31 or more cells per array does not fit the kernels' 32-bit
    cell index: COFFEE_ERR_SIZE, before any device call."""
    from coffeegpu_amd import _lib

    lib = _lib.load()
    g, p = _flat_grid_params(n=(1290, 1290, 1290))
    ctx = ctypes.c_void_p()
    rc = lib.coffee_ctx_create(0, 4, ctypes.byref(g), ctypes.byref(p), ctypes.byref(ctx))
    assert rc == -3
    assert b"2^31" in lib.coffee_last_error(None)
    assert not ctx.value


def test_batch_entries_refuse_null_context():
    from coffeegpu_amd import _lib

    lib = _lib.load()
    assert lib.coffee_sph_batch_step(None, 1, None) == -4
    assert lib.coffee_sph_batch_set_tile(None, 0) == -4
    assert lib.coffee_sph_batch_use_graph(None, 1) == -4
    assert lib.coffee_sph_batch_autotune(None, 1, None, None, None) == -4

