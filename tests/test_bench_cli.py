"""bench.py host logic (no GPU): the one-GPU refusal and the workload
parameters the benchmarks are quoted on."""

import json
import math
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["n_gpus"] == 2 and "one GPU" in out["error"]


def test_alfven_collision_parameters():
    """512^3 periodic unit box, guard 3, dt = 0.4 delta, cleaning (ch2 1,
    tau 2), KO 0.1, both projections, no absorbing layer."""
    sp = bench.alfven_params(512, np.float32)
    assert list(sp.N) == [512, 512, 512] and list(sp.guard) == [3, 3, 3]
    assert list(sp.periodic_boundary) == [True] * 3 and list(sp.pml) == [0, 0, 0]
    assert math.isclose(sp.dt, 0.4 / 512)
    assert sp.divB_clean and sp.ch2 == 1.0 and sp.tau == 2.0 and sp.KOeps == 0.1
    assert sp.clean_ep and sp.check_egb and sp.calc_current and not sp.use_edotb_damping


def test_gr_wald_parameters():
    """384^3 box [-20, 20]^3, absorbing layer of 32 cells, dt = 0.4 delta."""
    sp = bench.gr_wald_params(384, 0.9)
    delta = 40.0 / 384
    assert list(sp.N) == [384] * 3 and list(sp.lower) == [-20.0] * 3 and list(sp.size) == [40.0] * 3
    assert list(sp.pml) == [32] * 3 and list(sp.periodic_boundary) == [False] * 3
    assert math.isclose(sp.dt, 0.4 * delta) and sp.a == 0.9


def test_sph_sweep_parameters():
    """x = ln r up to ln 60, theta in [0, pi], absorbing layer from r = 48,
    dt = 0.4 min(dx, dtheta), R_LC ~ U[3, 20]."""
    rng = np.random.default_rng(0)
    for nr, nth in [(512, 256), (1024, 512), (2048, 1024)]:
        sp, rlc = bench.sph_sweep_params(rng, nr, nth)
        dx, dth = sp.size[0] / nr, sp.size[1] / nth
        assert math.isclose(sp.lower[0] + sp.size[0], math.log(60.0))
        assert math.isclose(sp.size[1], math.pi)
        assert sp.pml[0] == math.ceil((math.log(60.0) - math.log(48.0)) / dx)
        assert math.isclose(sp.dt, 0.4 * min(dx, dth))
        assert 3.0 <= rlc <= 20.0 and math.isclose(sp.omega, 1.0 / rlc)


def test_gr_un_fraction_matches_cell_by_cell_count():
    """The share of cells where the Kerr-Schild stages keep Un (outside the
    absorbing layer's inner bounds on any axis, or r < 0.7 r_H), vectorised,
    against a plain loop over the cells of a small grid with the same
    positions lower + i delta; and the 384^3 benchmark's 42 %."""
    sp = bench.gr_wald_params(24, 0.9)
    sp.pml = [4, 4, 4]
    n, d = 24, 40.0 / 24
    lo, hi = -20.0 + 4 * d, 20.0 - 4 * d
    a = 0.9
    r1 = 0.7 * (1.0 + math.sqrt(1.0 - a * a))
    count = 0
    for k in range(n):
        for j in range(n):
            for i in range(n):
                x, y, z = -20.0 + d * i, -20.0 + d * j, -20.0 + d * k
                w = x * x + y * y + z * z - a * a
                r = math.sqrt((w + math.sqrt(max(w * w + 4 * (a * z) ** 2, 1e-11))) / 2)
                outer = any(v > hi or v < lo for v in (x, y, z))
                count += outer or r < r1
    assert bench.gr_un_fraction(sp) == count / n ** 3
    sp = bench.gr_wald_params(192, 0.9)
    sp.pml = [16, 16, 16]  # the 384^3 layer's thickness in box units
    assert 0.41 < bench.gr_un_fraction(sp) < 0.43
