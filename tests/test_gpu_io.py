"""Restart snapshots and field output on the GPU (data_exporter).

Kill and resume: a run that stops after k steps, saves a snapshot, ends
(every solver, context and device array released) and is resumed by a new
process state from that snapshot reaches bitwise the state of the
uninterrupted run after n steps: E, B and P, guards included; the snapshot's
step and time are the ones saved.  Covered for the flat solver (FP32
periodic Alfven packets, FP64 pulsar with absorbing layer), the spherical
solver and the Kerr-Schild solver, through the Python mirror, and for the
flat and spherical solvers through the C++ surface (coffee_restart driver,
fresh process).

Field output: fld.%05d.h5 holds every array sampled on the GPU
(coffee_downsample), bitwise the oracle's sampling of the host copy (the
oracle equals the reference build's multi_array::downsample on the
committed fixtures), for downsample 1 and 2, 3D and 2D grids; the _cum
energy-accounting arrays are zero after output.
"""

import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover - CPU containers
    pytest.skip("no GPU", allow_module_level=True)

import oracle as O  # noqa: E402
from coffeegpu_amd import _h5  # noqa: E402
from coffeegpu_amd.data_exporter import (STAGGER_B, STAGGER_CENTRE, DataExporter,  # noqa: E402
                                         _output_members)
from coffeegpu_amd.gr import FieldSolverGREZ  # noqa: E402
from coffeegpu_amd.solver import FieldSolverEZ, SimData, SimEnvironment  # noqa: E402
from coffeegpu_amd.spherical import FieldSolverEZSpherical  # noqa: E402
from tests.golden.cases import case_alfven, case_pulsar  # noqa: E402
from tests.golden.cases_gr import case_gr  # noqa: E402
from tests.golden.cases_sph import case_sph  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SOLVERS = {"ez": FieldSolverEZ, "sph": FieldSolverEZSpherical, "gr": FieldSolverGREZ}


def _case(kind, dtype):
    if kind == "alfven":
        return "ez", case_alfven(dtype, n=(32, 32, 32), edotb=True)
    if kind == "pulsar":
        return "ez", case_pulsar(dtype)
    if kind == "sph":
        return "sph", case_sph(dtype, omega=0.3)
    return "gr", case_gr(dtype, a=0.9)


def _advance(solver_cls, sp, data, env, steps, t):
    """`steps` steps from time t; time accumulates in the Scalar type, as the
    reference's drivers keep it (Scalar time; time += dt)."""
    T = env.dtype.type
    t = T(t)
    solver = solver_cls(data, env)
    for _ in range(steps):
        solver.evolve_fields(float(t))
        t = T(t + T(sp.dt))
    torch.cuda.synchronize()
    solver.close()
    return float(t)


@pytest.mark.parametrize("kind,dtype", [("alfven", np.float32), ("pulsar", np.float64),
                                        ("sph", np.float32), ("gr", np.float64)])
def test_kill_and_resume_is_bitwise(tmp_path, kind, dtype):
    solver_kind, (sp, U0) = _case(kind, dtype)
    cls = SOLVERS[solver_kind]
    n, k = 6, 2
    print(f"[io] kill-and-resume {kind} {np.dtype(dtype).name} n={n} k={k}", flush=True)
    env = SimEnvironment(sp, dtype, "cuda:0")
    ref = SimData(env)
    ref.set_state(U0)
    _advance(cls, sp, ref, env, n, 0.0)
    want = ref.get_state().copy()
    del ref

    snap = str(tmp_path / "snapshot.h5")
    env1 = SimEnvironment(sp, dtype, "cuda:0")
    d1 = SimData(env1)
    d1.set_state(U0)
    t = _advance(cls, sp, d1, env1, k, 0.0)
    DataExporter(env1, 0, str(tmp_path / "Data"), copy_config=False).save_snapshot(snap, d1, k, t)
    del d1, env1
    torch.cuda.empty_cache()

    env2 = SimEnvironment(sp, dtype, "cuda:0")
    d2 = SimData(env2)
    step, t2 = DataExporter(env2, 0, str(tmp_path / "Data"), copy_config=False).load_snapshot(snap, d2)
    assert step == k and t2 == float(t)
    _advance(cls, sp, d2, env2, n - step, t2)
    got = d2.get_state()
    assert np.array_equal(got, want), f"max |diff| {np.abs(got - want).max()}"


def test_load_snapshot_missing_file_is_reported(tmp_path, capsys):
    sp, U0 = case_alfven(np.float32, n=(16, 16, 16))
    env = SimEnvironment(sp, np.float32, "cuda:0")
    d = SimData(env)
    assert DataExporter(env, 0, str(tmp_path / "D"), copy_config=False).load_snapshot(
        str(tmp_path / "none.h5"), d) is None
    assert "Can't find restart file" in capsys.readouterr().out


@pytest.mark.parametrize("kind,dtype,ds", [("pulsar", np.float64, 1), ("alfven", np.float32, 2),
                                           ("sph", np.float32, 2), ("sph", np.float64, 1)])
def test_field_output_matches_oracle_sampling(tmp_path, kind, dtype, ds):
    solver_kind, (sp, U0) = _case(kind, dtype)
    sp.downsample = ds
    sp.data_interval = 2
    print(f"[io] field output {kind} {np.dtype(dtype).name} downsample {ds}", flush=True)
    env = SimEnvironment(sp, dtype, "cuda:0")
    data = SimData(env)
    data.set_state(U0)
    solver = SOLVERS[solver_kind](data, env)
    if solver_kind == "ez":
        solver.set_diagnostics(True)  # J, divE, divB and the energy accounting
    for s in range(3):
        solver.evolve_fields(s * sp.dt)
    torch.cuda.synchronize()
    g = env.grid()
    host = {}
    for name, arr, st in _output_members(data):
        if arr is not None:
            host[name] = (arr.sync_to_host().copy(), st)
    ex = DataExporter(env, 0, str(tmp_path / "Data"), copy_config=False)
    path = ex.write_field_output(data, 4, 4 * sp.dt)
    assert os.path.basename(path) == "fld.00002.h5"
    two_d = g.dims[2] == 1
    with _h5.H5File(path) as f:
        for name, (a, st) in host.items():
            want = O.downsample(a.reshape(g.shape), g.guard, st, ds)
            shape = want.shape[1:] if two_d else want.shape
            got = f.read(name, shape, np.float32)
            assert np.array_equal(got, want.reshape(shape)), name
    assert STAGGER_B[0] == 0b001 and STAGGER_CENTRE == 0b111
    if solver_kind == "ez":
        assert "dU_KO_cum" in host and np.abs(host["dU_KO_cum"][0]).max() > 0
        for name in ("dU_KO_cum", "dU_Epar_cum", "dU_EgtB_cum"):
            assert not getattr(data, name).sync_to_host().any()
    solver.close()


def _cpp_restart(tmp_path, kind, dtype, sp, U0, n, k):
    exe = os.path.join(ROOT, "coffeegpu_amd", "bin",
                       "coffee_restart_f64" if dtype == np.float64 else "coffee_restart_f32")
    assert os.path.exists(exe), "C++ restart driver not built"
    (tmp_path / "config.toml").write_text(sp.to_toml())
    O.write_cfld(str(tmp_path / "in.cfld"), U0)
    r = subprocess.run([exe, "in.cfld", "out.cfld", str(n), str(k), "snap.h5", "Data", kind],
                       cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert f"resumed step {k}" in r.stdout
    return O.read_cfld(str(tmp_path / "out.cfld"))


@pytest.mark.parametrize("kind,dtype", [("pulsar", np.float64), ("alfven", np.float32),
                                        ("sph", np.float64)])
def test_cpp_restart_driver_matches_uninterrupted_run(tmp_path, kind, dtype):
    """The C++ surface (data_exporter.h) saves, ends the run, loads and
    continues: bitwise the uninterrupted run of the C++ driver over the same
    surface (coffee_drive, n steps in one go)."""
    solver_kind, (sp, U0) = _case(kind, dtype)
    n, k = 5, 2
    print(f"[io] C++ restart {kind} {np.dtype(dtype).name}", flush=True)
    got = _cpp_restart(tmp_path, solver_kind, dtype, sp, U0, n, k)
    exe = os.path.join(ROOT, "coffeegpu_amd", "bin",
                       "coffee_drive_f64" if dtype == np.float64 else "coffee_drive_f32")
    r = subprocess.run([exe, "in.cfld", "ref.cfld", str(n), solver_kind], cwd=str(tmp_path),
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    want = O.read_cfld(str(tmp_path / "ref.cfld"))
    assert np.array_equal(got, want), f"max |diff| {np.abs(got - want).max()}"
    fld = tmp_path / "Data" / f"fld.{k // sp.data_interval:05d}.h5"
    assert fld.exists()
