"""Log-spherical parity cases: a rotating dipole star with optional twist
pulse, absorbing layer and E.B-damped current."""

from __future__ import annotations

import math

import numpy as np

from coffeegpu_amd import initial_data as ic
from coffeegpu_amd.params import Grid, SimParams


def sph_params(n=(64, 32), r_out=20.0, omega=0.3, dw0=0.0, pml=8, KO_geometry=True,
               edotb=False, dt=0.01, b0=1.0, **kw):
    base = dict(dt=dt, N=[n[0], n[1], 1], guard=[4, 4, 0], lower=[-0.2, 0.0, 0.0],
                size=[math.log(r_out) + 0.2, math.pi, 0.0], nodes=[1, 1, 1],
                periodic_boundary=[False, False, False], pml=[pml, 0, 0], pmllen=2,
                sigpml=0.5, divB_clean=True, ch2=1.0, tau=2.0, KOeps=0.1,
                KO_geometry=KO_geometry, use_edotb_damping=edotb, damp_gamma=0.2,
                calc_current=True, clean_ep=True, check_egb=True, problem=0, radius=1.0,
                omega=omega, b0=b0, tp_start=0.0, tp_end=0.5, dw0=dw0, nT=1.0, thp1=0.3,
                thp2=0.6, tp_start1=1.0, tp_end1=1.5, dw1=0.0)
    base.update(kw)
    return SimParams(**base)


def case_sph(dtype, **kw):
    sp = sph_params(**kw)
    g = Grid.from_params(sp, dtype)
    return sp, ic.stack(ic.pulsar_dipole_sph(g, sp, dtype=dtype))


SPH_CASES = {
    "rotator": dict(omega=0.3),
    "twist": dict(omega=0.2, dw0=0.5, edotb=True, KO_geometry=False),
    "vacuum": dict(omega=0.0, pml=0),
}
