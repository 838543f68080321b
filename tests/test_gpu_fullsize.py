"""Properties of the HIP path at the benchmark sizes.

The oracle comparisons run at sizes the CPU oracle finishes in seconds.  These
tests carry the parity to the full benchmark shapes through properties that
hold bitwise at any size:

- schedule invariance: the tiled kernels (the ones the benchmarks time) equal
  the simple one-thread-per-cell kernels, which the oracle tests pin at small
  sizes, bitwise (flat 512^3 Alfven collision FP32 and FP64, Kerr-Schild Wald
  384^3 FP32);
- shift equivariance of the periodic flat box: a state shifted by a whole
  number of cells steps to the shifted result, bitwise, so every tile, XCD
  and wavefront seam sees other cells than in the unshifted run.  The
  cleaning term's grad P carries a tanh mask of the absolute position
  (field_solver_EZ.cpp:222-230), so this is checked with cleaning off;
- tile, graph and batch invariance of the batched 2D sweep on the sweep's
  grid sizes (512x256, 1024x512, 2048x1024).
"""

import gc

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

import bench  # noqa: E402
from coffeegpu_amd import initial_data as ic  # noqa: E402
from coffeegpu_amd.gr import FieldSolverGREZ  # noqa: E402
from coffeegpu_amd.solver import FieldSolverEZ, SimData, SimEnvironment  # noqa: E402
from coffeegpu_amd.spherical import FieldSolverEZSpherical, SphBatch  # noqa: E402

FLAT_N, FLAT_STEPS = 512, 3
GR_N, GR_STEPS = 384, 2
SPH_SHAPES = [(512, 256), (1024, 512), (2048, 1024)]


def _load(data, fields):
    for name, arr in zip(ic.FIELDS, data.state_arrays()):
        arr.set(fields[name])


def _release():
    gc.collect()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def _flat_run(sp, fields, variant, shift=None, dtype=np.float32, steps=FLAT_STEPS):
    """Steps the flat solver from `fields` (guards refilled periodically, the
    interior optionally rolled by `shift` = (dz, dy, dx) cells first) and
    returns device copies of the 7 state arrays."""
    env = SimEnvironment(sp, dtype, "cuda:0")
    data = SimData(env)
    _load(data, fields)
    solver = FieldSolverEZ(data, env)
    solver.set_schedule(variant, 0)
    it = env.grid().interior()
    if shift is not None:
        for a in data.state_arrays():
            a.dev[it] = torch.roll(a.dev[it], shifts=shift, dims=(0, 1, 2))
    solver.fill_guards()
    t = 0.0
    for _ in range(steps):
        solver.evolve_fields(t)
        t += sp.dt
    torch.cuda.synchronize()
    out = [a.dev.clone() for a in data.state_arrays()]
    solver.close()
    del solver, data, env
    _release()
    return out, it


def test_flat_512_schedules_and_shift_bitwise():
    sp = bench.alfven_params(FLAT_N, np.float32)
    grid = SimEnvironment(sp, np.float32, "cuda:0").grid()
    fields, _ = ic.alfven_packets_seeded(grid, 0, dtype=np.float32)
    print(f"flat {FLAT_N}^3 fp32 alfven, {FLAT_STEPS} steps: tiled (default)", flush=True)
    tiled, it = _flat_run(sp, fields, 0)
    assert all(bool(torch.isfinite(a).all()) for a in tiled)
    print("flat: simple one-thread-per-cell schedule", flush=True)
    simple, _ = _flat_run(sp, fields, 1)
    for f, (a, b) in enumerate(zip(tiled, simple)):
        assert torch.equal(a, b), f"field {f}: tiled and simple schedules differ"
    # the step changed the state (the test would pass trivially otherwise)
    start = torch.from_numpy(np.ascontiguousarray(fields["By"])).to("cuda:0")
    assert not torch.equal(tiled[4], start)
    del simple, tiled, start
    _release()
    # shift equivariance, cleaning off (no position-dependent term left); the
    # shift is not a multiple of the 128 x 8 tile, the 64-lane wavefront or
    # the XCD count in any axis
    sp.divB_clean = False
    shift = (211, 37, 101)
    print(f"flat: cleaning off, unshifted and shifted {shift} starts", flush=True)
    base, _ = _flat_run(sp, fields, 0)
    shifted, _ = _flat_run(sp, fields, 0, shift=shift)
    for f, (a, b) in enumerate(zip(base, shifted)):
        want = torch.roll(a[it], shifts=shift, dims=(0, 1, 2))
        assert torch.equal(b[it], want), f"field {f}: shifted run is not the shifted result"


def test_flat_512_fp64_schedules_bitwise():
    """The FP64 build at the benchmark size: the tiled kernels (default 64 x 8
    tile, fields addressed by 32-bit byte offsets) equal the simple
    one-thread-per-cell kernels (64-bit indexing) bitwise after 2 steps."""
    sp = bench.alfven_params(FLAT_N, np.float64)
    grid = SimEnvironment(sp, np.float64, "cuda:0").grid()
    fields, _ = ic.alfven_packets_seeded(grid, 0, dtype=np.float64)
    print(f"flat {FLAT_N}^3 fp64 alfven, 2 steps: tiled (default)", flush=True)
    tiled, _ = _flat_run(sp, fields, 0, dtype=np.float64, steps=2)
    assert all(bool(torch.isfinite(a).all()) for a in tiled)
    print("flat fp64: simple one-thread-per-cell schedule", flush=True)
    simple, _ = _flat_run(sp, fields, 1, dtype=np.float64, steps=2)
    for f, (a, b) in enumerate(zip(tiled, simple)):
        assert torch.equal(a, b), f"field {f}: tiled and simple schedules differ"
    start = torch.from_numpy(np.ascontiguousarray(fields["By"])).to("cuda:0")
    assert not torch.equal(tiled[4], start)
    del simple, tiled, start, fields
    _release()


def _gr_run(sp, fields, variant):
    env = SimEnvironment(sp, np.float32, "cuda:0")
    data = SimData(env)
    _load(data, fields)
    solver = FieldSolverGREZ(data, env)
    solver.set_schedule(variant, 0)
    for s in range(GR_STEPS):
        solver.evolve_fields(s * sp.dt)
    torch.cuda.synchronize()
    out = [a.dev.clone() for a in data.state_arrays()]
    solver.close()
    del solver, data, env
    _release()
    return out


@pytest.mark.parametrize("a", [0.5, 0.9, 0.99])
def test_gr_384_schedules_bitwise(a):
    sp = bench.gr_wald_params(GR_N, a)
    grid = SimEnvironment(sp, np.float32, "cuda:0").grid()
    fields = ic.wald_uniform(grid, sp.a, noise=1e-3, seed=0, dtype=np.float32)
    print(f"gr {GR_N}^3 fp32 wald a={a}, {GR_STEPS} steps: tiled (default)", flush=True)
    tiled = _gr_run(sp, fields, 0)
    assert all(bool(torch.isfinite(a).all()) for a in tiled)
    print("gr: point sweep + stage sweep schedule", flush=True)
    simple = _gr_run(sp, fields, 1)
    for f, (a, b) in enumerate(zip(tiled, simple)):
        assert torch.equal(a, b), f"field {f}: tiled and two-sweep schedules differ"


def _sph_setup():
    rng = np.random.default_rng(0)
    envs, datas, sps, starts = [], [], [], []
    for nr, nth in SPH_SHAPES:
        sp, _ = bench.sph_sweep_params(rng, nr, nth)
        env = SimEnvironment(sp, np.float32, "cuda:0")
        d = SimData(env)
        f = ic.pulsar_dipole_sph(env.grid(), sp, dtype=np.float32)
        _load(d, f)
        envs.append(env)
        datas.append(d)
        sps.append(sp)
        starts.append(d.get_state())
    return envs, datas, sps, starts


def test_sph_sweep_sizes_tiles_graph_batch_bitwise():
    steps = 3
    ref = None
    # march tiles (the autotuner's usual picks), an LDS tile and the
    # global-memory tile; graphs on for every other tile
    for n, tile in enumerate((16, 13, 5, 0)):
        envs, datas, sps, starts = _sph_setup()
        print(f"sph sweep sizes {SPH_SHAPES}: tile {tile}, graph {n % 2 == 0}", flush=True)
        b = SphBatch(envs, datas, nsteps=[0] * len(datas))
        b.set_tile(tile)
        b.use_graph(n % 2 == 0)
        b.step(steps)
        torch.cuda.synchronize()
        outs = [d.get_state() for d in datas]
        b.close()
        assert all(np.isfinite(o).all() for o in outs)
        if ref is None:
            ref = outs
            assert any(not np.array_equal(o, s) for o, s in zip(outs, starts))
        else:
            for r, (a, want) in enumerate(zip(outs, ref)):
                assert np.array_equal(a, want), f"run {r}: tile {tile} differs from tile 16"
    # each run advanced alone by the single-run solver
    envs, datas, sps, starts = _sph_setup()
    for r, (env, d, sp) in enumerate(zip(envs, datas, sps)):
        solver = FieldSolverEZSpherical(d, env)
        t = np.float32(0.0)
        for _ in range(steps):
            solver.evolve_fields(float(t))
            t = np.float32(t + np.float32(sp.dt))
        torch.cuda.synchronize()
        assert np.array_equal(d.get_state(), ref[r]), f"run {r}: batch differs from single run"
        solver.close()
