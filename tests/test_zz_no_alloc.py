"""coffee_ez_step_noalloc (the allocation-free flat step) against coffee_ez_step.

Both run the same kernels, so the states must agree bitwise (guards
included) after every step.  Two inputs in a row on the same process, the
second smaller than the first, so the second runs on a partner array that
still holds the first input's last step:
- the 48^3 FP32 Alfven-packet collision (periodic, cleaning, E.B damping);
- the 32x16x16 FP64 pulsar (problem 1, absorbing layer toward Bbg,
  non-periodic frozen guards, star boundary).
A grid whose partner state exceeds COFFEE_NOALLOC_SCRATCH_BYTES is refused
with COFFEE_ERR_SIZE, and the step after the refusal still matches.
"""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover - CPU containers
    pytest.skip("no GPU", allow_module_level=True)

from coffeegpu_amd import _lib  # noqa: E402
from coffeegpu_amd.solver import (FieldSolverEZ, SimData, SimEnvironment,  # noqa: E402
                                  coffee_grid, coffee_params)
from tests.golden.cases import case_alfven, case_pulsar  # noqa: E402

ERR_SIZE = -3


def _pair(sp, U0, dtype):
    """Two solvers on copies of one start: one steps through its context, the
    other through coffee_ez_step_noalloc with its own caller arrays."""
    out = []
    for _ in range(2):
        env = SimEnvironment(sp, dtype, "cuda:0")
        data = SimData(env)
        data.set_state(U0)
        out.append((env, data, FieldSolverEZ(data, env)))
    return out


def _noalloc_step(lib, solver, env, t):
    g = coffee_grid(env.grid())
    p = coffee_params(env.params(), solver.mask_lanes)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    return lib.coffee_ez_step_noalloc(ctypes.byref(g), ctypes.byref(p), env.dtype.itemsize,
                                      ctypes.byref(solver._fields), ctypes.c_double(t), s)


def _run_and_compare(lib, sp, U0, dtype, steps):
    (env_a, data_a, sol_a), (env_b, data_b, sol_b) = _pair(sp, U0, dtype)
    t = 0.0
    for n in range(steps):
        sol_a.evolve_fields(t)
        rc = _noalloc_step(lib, sol_b, env_b, t)
        assert rc == 0, lib.coffee_last_error(None)
        torch.cuda.synchronize()
        a, b = data_a.get_state().copy(), data_b.get_state().copy()
        assert np.isfinite(a).all()
        assert np.array_equal(a, b), f"step {n}: max |diff| {np.abs(a - b).max()}"
        t += sp.dt
    sol_a.close()
    sol_b.close()


def test_noalloc_step_matches_context_step_two_inputs():
    lib = _lib.load()
    print("[noalloc] alfven 48^3 fp32, 3 steps", flush=True)
    sp, U0 = case_alfven(np.float32, n=(48, 48, 48), edotb=True)
    _run_and_compare(lib, sp, U0, np.dtype(np.float32), 3)
    print("[noalloc] pulsar 32x16x16 fp64, 3 steps (smaller input)", flush=True)
    sp, U0 = case_pulsar(np.float64, n=(32, 16, 16))
    assert U0[0].nbytes < 54 ** 3 * 4
    _run_and_compare(lib, sp, U0, np.dtype(np.float64), 3)


def test_noalloc_step_refuses_oversized_grid_then_still_steps():
    lib = _lib.load()
    sp, U0 = case_alfven(np.float32, n=(16, 16, 16))
    env = SimEnvironment(sp, np.float32, "cuda:0")
    data = SimData(env)
    data.set_state(U0)
    sol = FieldSolverEZ(data, env)
    g = coffee_grid(env.grid())
    # 256^3 interior: 7 x 262^3 x 4 bytes = 503 MB > 80 MiB
    big = _lib.CoffeeGrid()
    ctypes.memmove(ctypes.byref(big), ctypes.byref(g), ctypes.sizeof(g))
    for a in range(3):
        big.dims[a] = 256 + 2 * g.guard[a]
    p = coffee_params(env.params(), sol.mask_lanes)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = lib.coffee_ez_step_noalloc(ctypes.byref(big), ctypes.byref(p), 4,
                                    ctypes.byref(sol._fields), ctypes.c_double(0.0), s)
    assert rc == ERR_SIZE
    assert b"COFFEE_NOALLOC_SCRATCH_BYTES" in lib.coffee_last_error(None)
    sol.close()
    _run_and_compare(lib, sp, U0, np.dtype(np.float32), 2)
