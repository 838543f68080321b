"""Named parity cases shared by the golden-fixture script and the tests.

Each case gives SimParams and a deterministic initial state (7, k, j, i) for a
Scalar dtype.  Grids keep the interior x-extent a multiple of 16 so the
reference's AVX-512 loops (16 floats / 8 doubles per vector,
src/utils/simd.h:27-51) never run into guard cells.
"""

from __future__ import annotations

import math

import numpy as np

from coffeegpu_amd import initial_data as ic
from coffeegpu_amd.params import Grid, SimParams


def _periodic_box(n, dt_frac=0.4, **kw):
    base = dict(dt=dt_frac / n[0], N=list(n), guard=[3, 3, 3], lower=[0.0, 0.0, 0.0],
                size=[1.0, n[1] / n[0], n[2] / n[0]], nodes=[1, 1, 1],
                periodic_boundary=[True, True, True], pml=[0, 0, 0], divB_clean=True, ch2=1.0,
                tau=2.0, KOeps=0.1, use_edotb_damping=False, problem=0, calc_current=True,
                clean_ep=True, check_egb=True, KO_geometry=False)
    base.update(kw)
    return SimParams(**base)


def case_wave(dtype, n=(16, 16, 16), seed=1):
    sp = _periodic_box(n, size=[1.0, 1.0, 1.0], dt=0.4 / n[0])
    g = Grid.from_params(sp, dtype)
    f, _ = ic.em_wave_seeded(g, seed, dtype=dtype)
    return sp, ic.stack(f)


def case_alfven(dtype, n=(16, 16, 16), seed=3, edotb=False):
    sp = _periodic_box(n, size=[1.0, 1.0, 1.0], dt=0.4 / n[0], use_edotb_damping=edotb,
                       damp_gamma=0.2)
    g = Grid.from_params(sp, dtype)
    f, _ = ic.alfven_packets_seeded(g, seed, dtype=dtype)
    return sp, ic.stack(f)


def case_noise(dtype, n=(16, 8, 8), seed=7, edotb=True):
    """Smooth background plus seeded white noise large enough that E > B and
    E.B != 0 occur: exercises J, the E.B-damped current and both projections."""
    sp = _periodic_box(n, use_edotb_damping=edotb, damp_gamma=0.2, KOeps=0.3, tau=3.0, ch2=0.7)
    g = Grid.from_params(sp, dtype)
    rng = np.random.default_rng(seed)
    U = 0.3 * rng.standard_normal((7,) + g.shape)
    U[5] += 1.0
    U[6] *= 0.1
    # guard cells consistent with the periodic wrap (as after an exchange)
    U = ic.periodic_fill(U, g, sp.periodic_boundary)
    return sp, U.astype(dtype)


def case_pulsar(dtype, n=(32, 16, 16)):
    """problem 1: rotating multipole star, non-periodic box with an absorbing
    layer (frozen guards, damping toward Bbg)."""
    sp = SimParams(dt=0.04, N=list(n), guard=[3, 3, 3], lower=[-2.0625, -1.0625, -1.0625],
                   size=[4.0, 2.0, 2.0], nodes=[1, 1, 1], periodic_boundary=[False] * 3,
                   pml=[4, 3, 3], pmllen=2, sigpml=0.5, divB_clean=True, ch2=1.0, tau=2.0,
                   KOeps=0.2, use_edotb_damping=False, problem=1, radius=0.55, omega=0.3,
                   b0=1.0, p1=0.1, p2=0.0, p3=1.0, q11=0.05, q12=0.02, q13=0.01, q22=-0.03,
                   q23=0.02, q_offset_x=0.05, q_offset_y=0.0, q_offset_z=0.02,
                   KO_geometry=False)
    g = Grid.from_params(sp, dtype)
    f = ic.pulsar_dipole(g, sp, dtype=dtype)
    return sp, ic.stack(f)


def case_alfven_boundary(dtype, n=(32, 16, 16), pert_type=2, **kw):
    """problem 2: dipole star with a time-dependent twist (pert_type 2), an
    axisymmetric angular-velocity pulse between the field lines through
    rpert1 and rpert2 (pert_type 0, wpert, field_solver_EZ.cpp:644-661) or the
    same pulse confined to ph1 pi < phi < ph2 pi (pert_type 1, wpert3d,
    .cpp:663-684)."""
    sp = SimParams(dt=0.04, N=list(n), guard=[3, 3, 3], lower=[-2.0625, -1.0625, -1.0625],
                   size=[4.0, 2.0, 2.0], nodes=[1, 1, 1], periodic_boundary=[False] * 3,
                   pml=[0, 0, 0], pmllen=1, sigpml=10.0, divB_clean=True, ch2=1.0, tau=2.0,
                   KOeps=0.1, use_edotb_damping=False, problem=2, radius=0.6, omega=0.0,
                   b0=1.0, p1=0.0, p2=0.0, p3=1.0, tp_start=0.0, tp_end=1.0, dw0=0.4, nT=2.0,
                   theta0=0.3, drpert=0.5, rpert1=1.2, rpert2=3.0, pert_type=pert_type,
                   KO_geometry=False, ph1=0.2, ph2=0.8, dph=0.1)
    sp = sp.replace(**kw) if kw else sp
    g = Grid.from_params(sp, dtype)
    f = ic.pulsar_dipole(g, sp, dtype=dtype)
    return sp, ic.stack(f)


CASES = {
    "wave": case_wave,
    "alfven": case_alfven,
    "noise": case_noise,
    "pulsar": case_pulsar,
    "alfven_bc": case_alfven_boundary,
    "alfven_bc0": lambda dtype, **kw: case_alfven_boundary(dtype, pert_type=0, **kw),
    "alfven_bc1": lambda dtype, **kw: case_alfven_boundary(dtype, pert_type=1, **kw),
}

# steps at which golden states are stored
DUMPS = (1, 10)
