"""The C++ host surface (include/coffee/*.h): its config.toml parser agrees
with the reference's key set and types (and with coffeegpu_amd.params), and
the C++ driver built on it reproduces the oracle on the GPU."""

import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from coffeegpu_amd.params import SimParams, parse_config

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

EXAMPLE = """# reference-style config (config.toml.example layout)
dt = 0.001
max_steps = 100000
data_interval = 500
periodic_boundary = [ false, true, false ]
N = [ 3840, 2048, 1 ]
guard = [ 5, 5, 0 ]
lower = [ -0.2, 0.0, 0.0 ]
size = [ 5.7, 3.141592653589793, 0.0 ]
nodes = [ 1, 1, 1 ]
pml = [ 15, 15,
        0 ]
pmllen = 1
sigpml = 10.0
divB_clean = true
ch2 = 1
tau = 2.0
KOeps = 0.05
problem = 2.5   # a float for an int key keeps the default
radius = 1
omega = 0.0
b0 = 1e3
use_edotb_damping = false
pert_type = 2
dw0 = 0.4
"""


def test_cpp_parse_config_matches_python(tmp_path):
    exe = tmp_path / "pc"
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "cpp", "parse_config_check.cpp"), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    cfg = tmp_path / "config.toml"
    cfg.write_text(EXAMPLE)
    out = subprocess.run([str(exe), str(cfg)], capture_output=True, text=True, check=True).stdout
    got = dict(line.split() for line in out.strip().splitlines())
    p = parse_config(str(cfg))
    assert float(got["dt"]) == pytest.approx(p.dt, rel=1e-7)
    assert int(got["max_steps"]) == p.max_steps == 100000
    assert int(got["problem"]) == p.problem == 0  # float for an int key: default kept
    assert int(got["pmllen"]) == p.pmllen == 1
    for i in range(3):
        assert int(got[f"N{i}"]) == p.N[i]
        assert int(got[f"guard{i}"]) == p.guard[i]
        assert float(got[f"lower{i}"]) == pytest.approx(p.lower[i], rel=1e-7, abs=1e-12)
        assert float(got[f"size{i}"]) == pytest.approx(p.size[i], rel=1e-7, abs=1e-12)
        assert int(got[f"periodic{i}"]) == int(p.periodic_boundary[i])
        assert int(got[f"pml{i}"]) == p.pml[i]
        assert int(got[f"nodes{i}"]) == p.nodes[i]
    assert got["periodic1"] == "1" and got["pml2"] == "0"
    assert float(got["ch2"]) == 1.0 and float(got["radius"]) == 1.0  # ints read as floats
    assert int(got["use_edotb_damping"]) == 0 and int(got["divB_clean"]) == 1
    assert int(got["pert_type"]) == 2
    # untouched keys keep the reference defaults
    assert float(got["a"]) == pytest.approx(0.99, rel=1e-7)
    assert float(got["KOeps"]) == pytest.approx(0.05, rel=1e-7)


def test_python_defaults_match_reference_traps():
    p = SimParams()
    assert p.pml == [60, 60, 60] and p.use_edotb_damping and p.KOeps == 0.5
    assert p.KO_geometry and p.radius == 10.0 and p.omega == 0.001 and p.b0 == 1e3
    assert p.tau == 2.0 and p.pmllen == 10 and p.sigpml == 0.1


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_cpp_driver_matches_oracle(dtype):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import oracle as O
    from tests.golden.cases import case_pulsar

    sp, U0 = case_pulsar(dtype)
    exe = os.path.join(ROOT, "coffeegpu_amd", "bin",
                       "coffee_drive_f64" if dtype == np.float64 else "coffee_drive_f32")
    assert os.path.exists(exe), "C++ driver not built"
    tmp = tempfile.mkdtemp(prefix="coffee_drive_")
    try:
        with open(os.path.join(tmp, "config.toml"), "w") as f:
            f.write(sp.to_toml())
        O.write_cfld(os.path.join(tmp, "in.cfld"), U0)
        r = subprocess.run([exe, "in.cfld", "out.cfld", "5"], cwd=tmp, capture_output=True,
                           text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        got = O.read_cfld(os.path.join(tmp, "out.cfld"))
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    # the driver computes Bbg itself (host-side multipole), as the reference ctor does
    bbg = O.flat_bbg(sp, dtype)
    st = O.FlatState(U0, dtype, bbg)
    O.flat_evolve(sp, st, 5, mask_lanes=16 if dtype == np.float32 else 8)
    err = O.rel_l2(got, st.U)
    print(f"C++ driver vs oracle ({np.dtype(dtype).name}, pulsar, 5 steps): {err:.2e}")
    assert err <= (1e-12 if dtype == np.float64 else 1e-5)


def _run_driver(dtype, sp, U0, nsteps, kind):
    import oracle as O

    exe = os.path.join(ROOT, "coffeegpu_amd", "bin",
                       "coffee_drive_f64" if dtype == np.float64 else "coffee_drive_f32")
    assert os.path.exists(exe), "C++ driver not built"
    tmp = tempfile.mkdtemp(prefix="coffee_drive_")
    try:
        with open(os.path.join(tmp, "config.toml"), "w") as f:
            f.write(sp.to_toml())
        O.write_cfld(os.path.join(tmp, "in.cfld"), U0)
        r = subprocess.run([exe, "in.cfld", "out.cfld", str(nsteps), kind], cwd=tmp,
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        return O.read_cfld(os.path.join(tmp, "out.cfld")), r.stdout
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_cpp_spherical_solver_matches_oracle(dtype):
    """field_solver_EZ_spherical through the C++ surface (config.toml ->
    sim_environment -> sim_data -> solver.evolve_fields) against the oracle."""
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import oracle as O
    from tests.golden.cases_sph import case_sph

    sp, U0 = case_sph(dtype, omega=0.3)
    got, out = _run_driver(dtype, sp, U0, 5, "sph")
    st = O.SphState(U0, dtype)
    O.sph_evolve(sp, st, 5)
    err = O.rel_l2(got, st.U)
    print(f"C++ spherical solver vs oracle ({np.dtype(dtype).name}, 5 steps): {err:.2e}\n{out}")
    assert err <= (1e-12 if dtype == np.float64 else 1e-5)
    assert "nonfinite 0" in out


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_cpp_gr_solver_matches_oracle(dtype):
    """field_solver_gr_EZ through the C++ surface against the FP64 oracle
    (FP32: within 1e-5 outside the horizon, the bar of the GR FP32 tests)."""
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import oracle as O
    from coffeegpu_amd import initial_data as ic
    from coffeegpu_amd.params import Grid
    from tests.golden.cases_gr import case_gr

    sp, U0 = case_gr(np.float64, a=0.9, noise=0.0 if dtype == np.float32 else 1e-3)
    got, out = _run_driver(dtype, sp, U0.astype(dtype), 5, "gr")
    st = O.GrState(U0)
    O.gr_evolve(sp, st, 5)
    if dtype == np.float64:
        err = O.rel_l2(got, st.U)
    else:
        g = Grid.from_params(sp, np.float64)
        X, Y, Z = (np.array(v) for v in np.broadcast_arrays(*g.coords()))
        outside = ic.cks_metric_np(sp.a, X, Y, Z)[0] >= 1 + np.sqrt(1 - sp.a ** 2)
        d = got.astype(np.float64) - st.U
        err = float(np.sqrt((d[:, outside] ** 2).sum() / (st.U[:, outside] ** 2).sum()))
    print(f"C++ GR solver vs oracle ({np.dtype(dtype).name}, 5 steps): {err:.2e}\n{out}")
    assert err <= (1e-12 if dtype == np.float64 else 1e-5)
    assert "nonfinite 0" in out
