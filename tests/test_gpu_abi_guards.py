"""Regression tests of the batch entry points' argument checks and of the
graph cache.

- coffee_sph_batch_* return COFFEE_ERR_STATE on contexts that are not
  batches (a Kerr-Schild context, a single-run spherical context, a flat
  context) and launch nothing;
- a graph captured before coffee_sph_batch_set_tile rebuilds the work list is
  dropped with it: graph on, step, set_tile(same tile), step gives bitwise the
  result of the same steps with the graph off.
"""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

from coffeegpu_amd import _lib  # noqa: E402
from coffeegpu_amd.gr import FieldSolverGREZ  # noqa: E402
from coffeegpu_amd.solver import FieldSolverEZ, SimData, SimEnvironment  # noqa: E402
from coffeegpu_amd.spherical import FieldSolverEZSpherical, SphBatch  # noqa: E402
from tests.golden.cases import case_alfven  # noqa: E402
from tests.golden.cases_gr import case_gr  # noqa: E402
from tests.golden.cases_sph import case_sph  # noqa: E402

ERR_STATE = -4


def _batch_calls_all_refused(lib, ctx):
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    best = ctypes.c_int(-1)
    radius = (ctypes.c_double * 1)(2.0)
    out = (ctypes.c_double * 1)(0.0)
    act = (ctypes.c_int * 1)(0)
    t = (ctypes.c_double * 1)(0.0)
    st = (ctypes.c_int * 1)(0)
    rcs = {
        "step": lib.coffee_sph_batch_step(ctx, 1, s),
        "set_tile": lib.coffee_sph_batch_set_tile(ctx, 0),
        "autotune": lib.coffee_sph_batch_autotune(ctx, 1, s, ctypes.byref(best), None),
        "use_graph": lib.coffee_sph_batch_use_graph(ctx, 1),
        "status": lib.coffee_sph_batch_status(ctx, act, t, st),
        "luminosity": lib.coffee_sph_batch_luminosity(ctx, radius, out, s),
        "destroy": lib.coffee_sph_batch_destroy(ctx),
    }
    torch.cuda.synchronize()
    bad = {k: v for k, v in rcs.items() if v != ERR_STATE}
    assert not bad, bad
    assert b"not a spherical batch" in lib.coffee_last_error(ctx)
    assert best.value == -1


def test_batch_entries_refuse_gr_context():
    sp, U0 = case_gr(np.float32, a=0.9, n=(16, 16, 16))
    env = SimEnvironment(sp, np.float32, "cuda:0")
    data = SimData(env)
    data.set_state(U0)
    solver = FieldSolverGREZ(data, env)
    _batch_calls_all_refused(solver._lib, solver._ctx)
    # the context still works after the refused calls
    solver.evolve_fields(0.0)
    torch.cuda.synchronize()
    assert np.isfinite(data.get_state()).all()
    solver.close()


def test_batch_entries_refuse_single_run_spherical_context():
    sp, U0 = case_sph(np.float32, n=(32, 16))
    env = SimEnvironment(sp, np.float32, "cuda:0")
    data = SimData(env)
    data.set_state(U0)
    solver = FieldSolverEZSpherical(data, env)
    # before any step: its run descriptor has no arrays yet
    _batch_calls_all_refused(solver._lib, solver._ctx)
    solver.evolve_fields(0.0)
    _batch_calls_all_refused(solver._lib, solver._ctx)
    solver.close()


def test_batch_entries_refuse_flat_context():
    sp, U0 = case_alfven(np.float32, n=(16, 16, 16))
    env = SimEnvironment(sp, np.float32, "cuda:0")
    data = SimData(env)
    data.set_state(U0)
    solver = FieldSolverEZ(data, env)
    _batch_calls_all_refused(solver._lib, solver._ctx)
    solver.close()


SPECS = [dict(n=(64, 32), omega=0.3), dict(n=(48, 24), omega=0.25, pml=0, KO_geometry=False)]


def _run_batch(graph, reset_tile_between):
    dtype = np.float32
    envs, datas = [], []
    for kw in SPECS:
        sp, U0 = case_sph(dtype, **kw)
        env = SimEnvironment(sp, dtype, "cuda:0")
        d = SimData(env)
        d.set_state(U0)
        envs.append(env)
        datas.append(d)
    b = SphBatch(envs, datas, nsteps=[0, 0])
    b.set_tile(10)
    b.use_graph(graph)
    b.step(2)
    if reset_tile_between:
        # the same tile: the work list is rebuilt (new allocation), the
        # captured graph must not be replayed with the freed one
        b.set_tile(10)
        # allocate and free device memory so the freed work list's address is
        # unlikely to come back for the rebuilt one
        junk = torch.empty(1 << 20, device="cuda:0")
        del junk
        b.set_tile(10)
    b.step(3)
    torch.cuda.synchronize()
    out = [d.get_state().copy() for d in datas]
    b.close()
    return out


def test_graph_dropped_when_tile_reset():
    ref = _run_batch(graph=False, reset_tile_between=False)
    got = _run_batch(graph=True, reset_tile_between=True)
    for a, r in zip(got, ref):
        assert np.array_equal(a, r)


@pytest.mark.parametrize("words,unroll,blocks", [(1, 1, 1), (1000, 4, 7), (4097, 8, 3),
                                                 (65536, 2, 1024), (1 << 20, 4, 16384)])
def test_stream_copy_copies_every_word(words, unroll, blocks):
    """The bandwidth probe's copy (one contiguous chunk per workgroup) covers
    every 16-byte word whatever the grid, including more workgroups than
    words and chunks that end mid-row."""
    lib = _lib.load()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    src = torch.arange(words * 4, dtype=torch.float32, device="cuda")
    dst = torch.full_like(src, -1.0)
    assert lib.coffee_stream_copy(ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(dst.data_ptr()),
                                  words * 16, unroll, blocks, s) == 0
    torch.cuda.synchronize()
    assert torch.equal(src, dst)
