"""CPU checks that pin the log-spherical oracle (oracle/ffe_sph*.c).

The reference has no CPU build of this path, so nothing from the reference
pins it directly ("parity unpinned" against the reference); these checks pin
it to closed forms instead."""

import ctypes
import math

import numpy as np
import pytest

import oracle as O
from coffeegpu_amd.params import Grid
from tests.golden.cases_sph import case_sph, sph_params


def _ptrs(U):
    P = ctypes.POINTER(ctypes.c_double)
    return (P * len(U))(*[u.ctypes.data_as(P) for u in U])


def test_static_vacuum_dipole_is_steady():
    """Omega = 0, no pulse: the dipole is a vacuum solution; the scheme keeps
    it to truncation error and E stays tiny.  (With KO_geometry the KO term is
    divided by sqrt(gamma), which vanishes on the axis, so the reference's
    default dissipates strongly next to the axis; the flat 2D KO is used.)"""
    sp, U0 = case_sph(np.float64, n=(96, 48), omega=0.0, pml=0, KO_geometry=False)
    st = O.SphState(U0, np.float64)
    O.sph_evolve(sp, st, 20)
    g = Grid.from_params(sp, np.float64)
    it = (slice(None),) + g.interior()
    B0, B1 = U0[3:6][it], st.U[3:6][it]
    relB = np.sqrt(((B1 - B0) ** 2).sum()) / np.sqrt((B0 ** 2).sum())
    relE = np.sqrt((st.U[0:3][it] ** 2).sum()) / np.sqrt((B0 ** 2).sum())
    assert relB < 5e-3, relB
    assert relE < 5e-3, relE


def test_axis_mirror_parity():
    sp, U0 = case_sph(np.float64)
    st = O.SphState(U0, np.float64)
    O.sph_evolve(sp, st, 3)
    g = Grid.from_params(sp, np.float64)
    U = st.U[:, 0]  # (7, j, i)
    j0 = g.guard[1]
    jpi = g.dims[1] - g.guard[1]
    ii = slice(g.guard[0], g.dims[0] - g.guard[0])
    sign = [1, -1, -1, 1, -1, -1]
    for l in range(1, 4):
        for c in range(6):
            assert np.array_equal(U[c, j0 - l, ii], sign[c] * U[c, j0 + l, ii])
            assert np.array_equal(U[c, jpi + l, ii], sign[c] * U[c, jpi - l, ii])
    # on-axis: E_theta, E_phi, B_theta, B_phi vanish
    for c in (1, 2, 4, 5):
        assert np.all(U[c, j0, ii] == 0) and np.all(U[c, jpi, ii] == 0)


def test_dipole_energy_matches_closed_form():
    """total_energy of the dipole: sum of gamma_ii B^i B^i sqrt(gamma) dx dtheta
    over r >= R approximates (4/3)(R^-3 - r_out^-3) b0^2."""
    sp = sph_params(n=(256, 128), omega=0.0, pml=0, r_out=20.0)
    g = Grid.from_params(sp, np.float64)
    from coffeegpu_amd import initial_data as ic
    U = ic.stack(ic.pulsar_dipole_sph(g, sp))
    W = O.sph_energy(sp, np.ascontiguousarray(U[3:6]))
    x_last = g.lower[0] + g.delta[0] * (g.reduced_dim(0) - 1)
    want = 4.0 / 3.0 * (1.0 - math.exp(-3 * x_last)) * sp.b0 ** 2
    assert abs(W - want) / want < 0.03, (W, want)


def test_ko_theta_nyquist_decay():
    """Flat 2D KO (KO_geometry off) multiplies a theta-Nyquist mode that is
    uniform in r by (1 - KOeps)."""
    sp = sph_params(KO_geometry=False, omega=0.0, pml=0)
    g = Grid.from_params(sp, np.float64)
    U = np.zeros((7,) + g.shape)
    j = np.arange(g.dims[1])
    U[6] = ((-1.0) ** j)[None, :, None]
    tmp = np.zeros_like(U)
    O.lib().ffe_sph_ko_f64(ctypes.byref(O.make_params(sp)), _ptrs(U), _ptrs(tmp))
    it = g.interior()
    want = (1 - sp.KOeps) * np.broadcast_to(((-1.0) ** j)[None, :, None], g.shape)
    assert np.allclose(U[6][it], want[it], atol=1e-13)


def test_sph_oracle_fp32_tracks_fp64():
    sp, U64 = case_sph(np.float64, omega=0.3)
    a = O.SphState(U64.astype(np.float32), np.float32)
    b = O.SphState(U64.astype(np.float32).astype(np.float64), np.float64)
    O.sph_evolve(sp, a, 10)
    O.sph_evolve(sp, b, 10)
    g = Grid.from_params(sp, np.float64)
    it = (slice(None),) + g.interior()
    e = np.sqrt(((a.U[3:6][it] - b.U[3:6][it]) ** 2).sum()) / np.sqrt((b.U[3:6][it] ** 2).sum())
    assert e < 1e-5
