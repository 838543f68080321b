"""Kerr-Schild GR cases: the reference's Wald start (user_wald2.hpp) with
seeded noise, absorbing layer and horizon excision; and the exact vacuum Wald
solution, built from its vector potential by complex-step differentiation."""

from __future__ import annotations

import numpy as np

from coffeegpu_amd import initial_data as ic
from coffeegpu_amd.params import Grid, SimParams


def gr_params(n=(24, 24, 24), lower=(-4.0, -4.0, -4.0), size=(8.0, 8.0, 8.0), a=0.9,
              pml=(4, 4, 4), pmllen=4, sigpml=0.5, cfl=0.4, periodic=(False, False, False),
              **kw):
    delta = size[0] / n[0]
    base = dict(dt=cfl * delta, N=list(n), guard=[3, 3, 3], lower=list(lower), size=list(size),
                nodes=[1, 1, 1], periodic_boundary=list(periodic), pml=list(pml), pmllen=pmllen,
                sigpml=sigpml, divB_clean=True, ch2=1.0, tau=2.0, KOeps=0.1,
                use_edotb_damping=False, calc_current=True, clean_ep=True, check_egb=True,
                problem=0, a=a)
    base.update(kw)
    return SimParams(**base)


def case_gr(dtype, noise=1e-3, seed=7, **kw):
    sp = gr_params(**kw)
    g = Grid.from_params(sp, dtype)
    return sp, ic.stack(ic.wald_uniform(g, sp.a, noise=noise, seed=seed, dtype=dtype))


GR_CASES = {
    "wald09": dict(a=0.9),
    "wald099": dict(a=0.99),
    "wald05_ragged": dict(a=0.5, n=(28, 20, 24), lower=(-4.0, -3.0, -4.5),
                          size=(7.0, 5.0, 6.0), pml=(3, 2, 4)),
    "wald_a0_nopml": dict(a=0.0, pml=(0, 0, 0), noise=1e-2),
    "periodic_z": dict(a=0.7, periodic=(False, False, True), pml=(4, 4, 0)),
}


# ---- exact vacuum Wald field ------------------------------------------------


def _wald_potential(a, x, y, z):
    """A = (psi + 2 a eta) / 2 with psi, eta the 1-forms of the Killing vectors
    d/dphi = -y d/dx + x d/dy and d/dt of Kerr-Schild (g = eta + 2H l l,
    l_mu = (1, l_x, l_y, l_z)); asymptotically uniform B = z-hat.  Works on
    complex arguments."""
    R2 = x * x + y * y + z * z
    w = R2 - a * a
    r = np.sqrt((w + np.sqrt(w * w + 4.0 * a * a * z * z)) / 2.0)
    H = r ** 3 / (r ** 4 + a * a * z * z)
    A = r * r + a * a
    lx, ly, lz = (r * x + a * y) / A, (r * y - a * x) / A, z / r
    lphi = -y * lx + x * ly
    At = 0.5 * (2.0 * H * lphi + 2.0 * a * (-1.0 + 2.0 * H))
    Ax = 0.5 * (-y + 2.0 * H * lphi * lx + 4.0 * a * H * lx)
    Ay = 0.5 * (x + 2.0 * H * lphi * ly + 4.0 * a * H * ly)
    Az = 0.5 * (2.0 * H * lphi * lz + 4.0 * a * H * lz)
    return At, Ax, Ay, Az


def _dpot(a, x, y, z, axis, h=1e-30):
    xs = [x.astype(np.complex128), y.astype(np.complex128), z.astype(np.complex128)]
    xs[axis] = xs[axis] + 1j * h
    return [np.imag(v) / h for v in _wald_potential(a, *xs)]


def wald_exact(grid: Grid, a: float):
    """(7, k, j, i) FP64 state D, B, P = 0 of the exact Wald solution at every
    node (E_i = d_i A_t, sqrt(g) B^i = curl A, D from E = alpha D_d + sqrt(g) beta x B)."""
    X, Y, Z = (np.array(v) for v in np.broadcast_arrays(
        *Grid(grid.dims, grid.guard, grid.delta, grid.inv_delta, grid.lower, grid.sizes,
              np.float64).coords()))
    dX, dY, dZ = (_dpot(a, X, Y, Z, ax) for ax in range(3))
    # d_j A_k: dX = d/dx of (At, Ax, Ay, Az), ...
    E = np.stack([dX[0], dY[0], dZ[0]])
    curl = np.stack([dY[3] - dZ[2], dZ[1] - dX[3], dX[2] - dY[1]])
    _, s, lx, ly, lz = ic.cks_metric_np(a, X, Y, Z)
    sg = np.sqrt(1.0 + s)
    alpha = 1.0 / sg
    B = curl / sg
    l = np.stack([lx, ly, lz])
    beta = (s / (1.0 + s)) * l
    bxB = np.stack([beta[1] * B[2] - beta[2] * B[1], beta[2] * B[0] - beta[0] * B[2],
                    beta[0] * B[1] - beta[1] * B[0]])
    Dd = (E - sg * bxB) / alpha  # gamma_ij D^j
    lDd = (l * Dd).sum(axis=0)
    D = Dd - (s / (1.0 + s)) * l * lDd  # gamma^ij (D_d)_j
    return np.concatenate([D, B, np.zeros((1,) + X.shape)])
