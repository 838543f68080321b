"""Pins the force-free parts of the log-spherical oracle (oracle/ffe_sph*.c).

The reference's spherical solver is CUDA-only
(src/algorithms/field_solver_EZ_spherical.cu), so no reference build pins
the oracle.  These checks pin it with what the mathematics fixes, stated here
independently of the oracle's own formulas:

- one RK stage's dU/dt (As = 0) on smooth manufactured fields with E.B != 0,
  div E != 0, P != 0 and E < B converges at 4th order to the continuum
  right-hand side of the force-free equations, written in orthonormal
  spherical components with the textbook curl / divergence / gradient
  (axisymmetric, d/dphi = 0) and evaluated by complex-step differentiation;
  both branches of the current (.cu:250-290) are covered;
- the Michel split monopole (B_r = 1/r^2, B_phi = E_theta = -Omega sin(theta)/r),
  an exact stationary force-free solution with J != 0, is stationary (to
  rounding: the truncation errors of rho and curl B cancel);
- closed forms of single cells: the metric-normed projections (.cu:361-442),
  the radial absorbing layer (.cu:750-802) and the star interior and surface
  values with the twist pulse envelope (.cu:576-595, 617-688);
- every planted mistake (oracle.PLANTS) makes at least one of these checks
  fail (mutation check of the checks themselves).
"""

import math

import numpy as np
import pytest

import oracle as O
from coffeegpu_amd.params import Grid, SimParams

H = 1e-30  # complex step


def d_r(f, r, th):
    return np.imag(f(r + 1j * H, th)) / H


def d_th(f, r, th):
    return np.imag(f(r, th + 1j * H)) / H


# ---- manufactured fields (orthonormal r, theta, phi components) ----------------

def mf_fields(r, th):
    Br = 1.0 + 0.3 * np.cos(th) * np.sin(2 * r)
    Bt = 0.4 * np.sin(th) * np.cos(r)
    Bp = 0.5 + 0.2 * np.sin(2 * th) * np.cos(1.3 * r)
    Er = 0.2 * np.sin(th) * np.sin(r)
    Et = 0.3 * np.cos(th) * np.cos(0.7 * r)
    Ep = 0.1 * np.sin(2 * th) * np.sin(1.1 * r)
    P = 0.05 * np.sin(r) * np.cos(th)
    return [Er, Et, Ep, Br, Bt, Bp, P]


OMEGA_M = 0.5


def monopole_fields(r, th):
    z = 0 * r * th
    Bp = -OMEGA_M * np.sin(th) / r + z
    return [z, Bp, z, 1.0 / r ** 2 + z, z, Bp, z]


def comp(fn, c):
    return lambda r, th: fn(r, th)[c]


def curl(fn, o, r, th):
    """Orthonormal curl of the vector fn[o:o+3] (d/dphi = 0)."""
    Fr, Ft, Fp = comp(fn, o), comp(fn, o + 1), comp(fn, o + 2)
    s = np.sin(th)
    cr = d_th(lambda r_, t_: np.sin(t_) * Fp(r_, t_), r, th) / (r * s)
    ct = -d_r(lambda r_, t_: r_ * Fp(r_, t_), r, th) / r
    cp = (d_r(lambda r_, t_: r_ * Ft(r_, t_), r, th) - d_th(Fr, r, th)) / r
    return np.stack([cr, ct, cp])


def div(fn, o, r, th):
    Fr, Ft = comp(fn, o), comp(fn, o + 1)
    return (d_r(lambda r_, t_: r_ ** 2 * Fr(r_, t_), r, th) / r ** 2 +
            d_th(lambda r_, t_: np.sin(t_) * Ft(r_, t_), r, th) / (r * np.sin(th)))


def continuum_rhs(fn, r, th, sp):
    """d/dt of (E, B, P) in orthonormal components: dE/dt = curl B - J,
    dB/dt = -curl E - grad P, dP/dt = -ch2 div B - P / (tau dt) (tau counted
    in steps, field_solver_EZ_spherical.cu:326-328), with the force-free
    current J = [rho E x B + (B.curl B - E.curl E) B] / B^2 or its E.B-damped
    form (E0^2 + B^2 in the drift denominator, damp_gamma E.B / dt added to
    the parallel part)."""
    F = [np.real(v) for v in fn(r, th)]
    E, B, P = np.stack(F[0:3]), np.stack(F[3:6]), F[6]
    cB, cE = curl(fn, 3, r, th), curl(fn, 0, r, th)
    rho, divB = div(fn, 0, r, th), div(fn, 3, r, th)
    B2, E2, EB = (B * B).sum(0), (E * E).sum(0), (E * B).sum(0)
    ExB = np.cross(E, B, axis=0)
    par = (B * cB).sum(0) - (E * cE).sum(0)
    if sp.use_edotb_damping:
        chi2 = B2 - E2
        E02 = 0.5 * (np.sqrt(chi2 ** 2 + 4 * EB ** 2) - chi2)
        J = rho * ExB / (E02 + B2) + (par + sp.damp_gamma * EB / sp.dt) * B / B2
    else:
        J = (rho * ExB + par * B) / B2
    gP = np.stack([d_r(comp(fn, 6), r, th), d_th(comp(fn, 6), r, th) / r, 0 * r])
    dE = cB - J
    dB = -cE - gP
    dP = -sp.ch2 * divB - P / (sp.tau * sp.dt)
    return dE, dB, dP


def to_code(v, r, th):
    """Orthonormal -> stored upper components in (x = ln r, theta, phi):
    F^x = F_r / r, F^theta = F_theta / r, F^phi = F_phi / (r sin theta)."""
    return np.stack([v[0] / r, v[1] / r, v[2] / (r * np.sin(th))])


def sph_box(n, r0=1.5, r1=3.5, th0=0.5, th1=2.6, **kw):
    base = dict(dt=0.01, N=[n, n, 1], guard=[3, 4, 0], lower=[math.log(r0), th0, 0.0],
                size=[math.log(r1 / r0), th1 - th0, 0.0], nodes=[1, 1, 1],
                periodic_boundary=[False] * 3, pml=[0, 0, 0], pmllen=2, sigpml=0.5,
                divB_clean=True, ch2=0.7, tau=3.0, KOeps=0.1, KO_geometry=False,
                use_edotb_damping=False, damp_gamma=0.2, calc_current=True, clean_ep=True,
                check_egb=True, problem=0, radius=1.0, omega=0.0, b0=1.0)
    base.update(kw)
    return SimParams(**base)


def node_rth(g):
    x = g.pos(0, np.arange(g.dims[0]))
    th = g.pos(1, np.arange(g.dims[1]))
    r = np.exp(x)[None, None, :] * np.ones((1, g.dims[1], 1))
    TH = th[None, :, None] * np.ones((1, 1, g.dims[0]))
    return r, TH


def state_from(fn, g):
    r, th = node_rth(g)
    F = [np.real(v) * np.ones(g.shape) for v in fn(r, th)]
    U = np.zeros((7,) + g.shape)
    U[0:3] = to_code(np.stack(F[0:3]), r, th)
    U[3:6] = to_code(np.stack(F[3:6]), r, th)
    U[6] = F[6]
    return U


def rhs_errors(fn, n, **kw):
    """Max |discrete - continuum| / max |continuum| of dE/dt, dB/dt, dP/dt
    over interior cells."""
    sp = sph_box(n, **kw)
    g = Grid.from_params(sp, np.float64)
    U = state_from(fn, g)
    got = O.sph_stage_rhs(sp, U)
    r, th = node_rth(g)
    dE, dB, dP = continuum_rhs(fn, r, th, sp)
    want = np.concatenate([to_code(dE, r, th), to_code(dB, r, th), dP[None] * np.ones((1,) + g.shape)])
    it = (slice(None),) + g.interior()
    d, w = got[it], want[it]
    out = []
    for sl in (slice(0, 3), slice(3, 6), slice(6, 7)):
        out.append(float(np.abs(d[sl] - w[sl]).max() / max(np.abs(w[sl]).max(), 1e-300)))
    return np.array(out)


def manufactured_converges(edotb=False):
    e1 = rhs_errors(mf_fields, 24, use_edotb_damping=edotb)
    e2 = rhs_errors(mf_fields, 48, use_edotb_damping=edotb)
    ratio = e1 / e2
    ok = bool(np.all(ratio > 11.0) and np.all(ratio < 22.0) and np.all(e2 < 2e-5))
    return ok, e1, e2


@pytest.mark.parametrize("edotb", [False, True])
def test_stage_rhs_converges_to_continuum_at_4th_order(edotb):
    ok, e1, e2 = manufactured_converges(edotb)
    print(f"edotb={edotb}: E,B,P rel err {e1} -> {e2}, ratio {e1 / e2}")
    assert ok, (e1, e2)


def monopole_residual(n):
    """Max |dU/dt| of the stationary Michel monopole over interior cells,
    relative to the size of curl B (2 Omega cos(theta)/r^2)."""
    sp = sph_box(n, divB_clean=True)
    g = Grid.from_params(sp, np.float64)
    got = O.sph_stage_rhs(sp, state_from(monopole_fields, g))
    r, th = node_rth(g)
    it = (slice(None),) + g.interior()
    phys = np.concatenate([got[0:3] * np.stack([r, r, r * np.sin(th)]),
                           got[3:6] * np.stack([r, r, r * np.sin(th)])])
    scale = (2 * OMEGA_M / r[g.interior()] ** 2).max()
    return float(np.abs(phys[it]).max() / scale)


def monopole_stationary():
    """The scheme keeps the monopole stationary to rounding, not only to
    4th order: rho = div E and (curl B)_r both reduce to the theta stencil of
    sin^2(theta), whose truncation error then cancels in curl B - J."""
    a, b = monopole_residual(24), monopole_residual(48)
    return bool(a < 1e-12 and b < 1e-12), a, b


def test_michel_monopole_is_stationary():
    ok, a, b = monopole_stationary()
    print(f"monopole |dU/dt| / |curl B| {a:.3e} (24^2), {b:.3e} (48^2)")
    assert ok


# ---- single-cell closed forms --------------------------------------------------

def _rand_state(g, seed, egtb_frac=0.3):
    rng = np.random.default_rng(seed)
    r, th = node_rth(g)
    Ep = rng.standard_normal((3,) + g.shape) * 0.5
    Bp = rng.standard_normal((3,) + g.shape)
    big = rng.random(g.shape) < egtb_frac
    Ep[:, big] *= 4.0
    U = np.zeros((7,) + g.shape)
    U[0:3] = to_code(Ep, r, th)
    U[3:6] = to_code(Bp, r, th)
    U[6] = rng.standard_normal(g.shape)
    return U, Ep, Bp


def projection_closed_form():
    """After the projections, in orthonormal components: E -> E - (E.B)B/B^2,
    then scaled down to |E| = |B| where it exceeds B; B and P unchanged.  The
    result has E.B = 0 and E^2 <= B^2 in the metric norm, and a second
    application changes nothing."""
    sp = sph_box(16)
    g = Grid.from_params(sp, np.float64)
    U, Ep, Bp = _rand_state(g, 5)
    out = O.sph_project(sp, U)
    r, th = node_rth(g)
    B2 = (Bp * Bp).sum(0)
    Eperp = Ep - (Ep * Bp).sum(0) / B2 * Bp
    E2 = (Eperp * Eperp).sum(0)
    sc = np.where(E2 > B2, np.sqrt(B2 / np.maximum(E2, 1e-300)), 1.0)
    want = to_code(Eperp * sc, r, th)
    it = (slice(None),) + g.interior()
    e_form = np.abs(out[0:3][it] - want[it]).max() / np.abs(want[it]).max()
    lo = lambda V: V * np.stack([r * r, r * r, (r * np.sin(th)) ** 2])  # noqa: E731
    EB = (lo(out[0:3]) * out[3:6]).sum(0)[g.interior()]
    E2o = (lo(out[0:3]) * out[0:3]).sum(0)[g.interior()]
    B2o = (lo(out[3:6]) * out[3:6]).sum(0)[g.interior()]
    twice = O.sph_project(sp, out)
    ok = (e_form < 1e-13 and np.abs(EB).max() < 1e-13 * B2o.max() and
          np.all(E2o <= B2o * (1 + 1e-13)) and np.abs(twice - out).max() < 1e-14 and
          np.array_equal(out[3:7], U[3:7]))
    return bool(ok), e_form


def test_projection_closed_form():
    ok, e = projection_closed_form()
    assert ok, e


def absorbing_closed_form():
    """Radial layer (x > x_h = lower + size - pml dx): with
    sigma = sigpml ((r - r_h) / pmllen)^3, r_h = e^{x_h} (pmllen counted in r
    units, .cu:784-785), E and B become e^-sigma Un + (1 - e^-sigma)/sigma (U - Un);
    P and every cell with x <= x_h are untouched."""
    sp = sph_box(32, r0=1.0, r1=20.0, pml=[8, 0, 0], pmllen=3, sigpml=0.7)
    g = Grid.from_params(sp, np.float64)
    rng = np.random.default_rng(11)
    U = rng.standard_normal((7,) + g.shape)
    Un = rng.standard_normal((7,) + g.shape)
    out = O.sph_absorb(sp, U, Un)
    x = g.pos(0, np.arange(g.dims[0]))
    xh = sp.lower[0] + sp.size[0] - sp.pml[0] * g.delta[0]
    rh = math.exp(xh)
    sig = np.where(x > xh, sp.sigpml * ((np.exp(x) - rh) / sp.pmllen) ** 3, 0.0)
    sig = np.broadcast_to(sig[None, None, :], g.shape)
    e = np.exp(-sig)
    f = np.where(sig > 1e-11, (1.0 - np.exp(-sig)) / np.where(sig > 0, sig, 1.0), 1.0)
    want = U.copy()
    want[0:6] = np.where(sig > 1e-11, e * Un[0:6] + f * (U[0:6] - Un[0:6]), U[0:6])
    it = (slice(None),) + g.interior()
    err = np.abs(out[it] - want[it]).max()
    layer = (sig[g.interior()] > 0.1).sum()
    return bool(err < 1e-13 and layer > 20), err


def test_absorbing_layer_closed_form():
    ok, e = absorbing_closed_form()
    assert ok, e


def star_closed_form():
    """Inside the star (r < R): B = b0 (2 cos(theta) r^-3, sin(theta) r^-3, 0)
    and the corotation field E = -(Omega x r) x B, i.e.
    E_r = Omega r sin(theta) B_theta, E_theta = -Omega r sin(theta) B_r, E_phi = 0;
    on the surface row (|r - R| < dx/2) only B^r, E^theta, E^phi are set.
    The twist pulse adds dw0 cos^2((theta - mu) pi/(th2 - th1)) sin^2(pi s)
    sin(2 pi nT s) to Omega, s = (t - t0)/(t1 - t0), inside its window only."""
    kw = dict(r0=0.5, r1=3.0, th0=0.2, th1=2.9, omega=0.37, b0=1.3, dw0=0.5, tp_start=1.0,
              tp_end=3.0, nT=0.75, thp1=0.8, thp2=1.6, dw1=0.0, tp_start1=9.0, tp_end1=9.5)
    g = Grid.from_params(sph_box(40, **kw), np.float64)
    # the surface row test compares r - R with dx/2 (dx in ln r): put R next to a node
    sp = sph_box(40, radius=float(np.exp(g.pos(0, g.guard[0] + 17))) + 1e-3, **kw)
    rng = np.random.default_rng(2)
    U = rng.standard_normal((7,) + g.shape)
    r, th = node_rth(g)
    ok = True
    worst = 0.0
    for t in (0.5, 2.0):
        out = O.sph_boundary(sp, U, t)
        s = (t - sp.tp_start) / (sp.tp_end - sp.tp_start)
        mu = 0.5 * (sp.thp1 + sp.thp2)
        inside = (th >= sp.thp1) & (th <= sp.thp2) & (0 <= s <= 1)
        w = np.where(inside, sp.dw0 * np.cos((th - mu) * math.pi / (sp.thp2 - sp.thp1)) ** 2 *
                     math.sin(math.pi * s) ** 2 * math.sin(2 * math.pi * sp.nT * s), 0.0)
        Om = sp.omega + w
        Br = sp.b0 * 2 * np.cos(th) / r ** 3
        Bt = sp.b0 * np.sin(th) / r ** 3
        v = Om * r * np.sin(th)
        want_in = np.concatenate([to_code(np.stack([v * Bt, -v * Br, 0 * r]), r, th),
                                  to_code(np.stack([Br, Bt, 0 * r]), r, th)])
        it = g.interior()
        core = (r < sp.radius - 1e-11) & (np.abs(r - sp.radius) >= g.delta[0] / 2)
        surf = np.abs(r - sp.radius) < g.delta[0] / 2
        rest = ~core & ~surf
        for c in range(6):
            a, b = out[c][it], want_in[c][it]
            worst = max(worst, np.abs(a - b)[core[it]].max())
        for c in (3, 1, 2):
            worst = max(worst, np.abs(out[c][it] - want_in[c][it])[surf[it]].max())
        for c in (0, 4, 5, 6):
            ok &= np.array_equal(out[c][it][surf[it]], U[c][it][surf[it]])
        ok &= np.array_equal(out[:, 0][:, rest[0]], U[:, 0][:, rest[0]])
        ok &= bool(core[it].sum() > 50 and surf[it].sum() > 10)
        if t == 2.0:
            ok &= bool(np.abs(w).max() > 0.2)  # the pulse is on at t = 2
    return bool(ok and worst < 1e-13), worst


def test_star_interior_and_surface_closed_form():
    ok, e = star_closed_form()
    assert ok, e


# ---- planted mistakes ------------------------------------------------------------

CHECKS = {
    "manufactured": lambda: manufactured_converges(False)[0],
    "manufactured_edotb": lambda: manufactured_converges(True)[0],
    "monopole": lambda: monopole_stationary()[0],
    "projection": lambda: projection_closed_form()[0],
    "absorbing": lambda: absorbing_closed_form()[0],
    "star": lambda: star_closed_form()[0],
}

# mistakes of the restatement that exist in the spherical path
SPH_PLANTS = ["J_SIGN", "J_NO_SQRTG", "GRADP_LOWER", "TAU_DT", "SIGMA_SQUARED", "PROJ_FLAT",
              "ROT_NO_SQRTG", "DIV_CENTRE", "STAR_SIGN"]


def test_checks_pass_without_plant():
    with O.planted(None):
        assert all(f() for f in CHECKS.values())


@pytest.mark.parametrize("mistake", SPH_PLANTS)
def test_planted_mistake_fails_a_check(mistake):
    with O.planted(mistake):
        failed = [name for name, f in CHECKS.items() if not f()]
    print(f"{mistake}: caught by {failed}")
    assert failed, f"no check catches {mistake}"


def test_geometric_ko_closed_form():
    """KO_2d6 (.cu:150-173, the KO_geometry default) damps sqrt(gamma) f:
    for f = (-1)^j / sqrt(gamma), sqrt(gamma) f is constant in x and a theta
    checkerboard, whose 6th difference is -64 times itself, so one
    application gives (1 - KOeps) f exactly; for f = c / sqrt(gamma) (c
    linear in x and theta) nothing changes."""
    sp = sph_box(24, KO_geometry=True, KOeps=0.15)
    g = Grid.from_params(sp, np.float64)
    r, th = node_rth(g)
    sg = r ** 3 * np.sin(th)
    j = np.arange(g.dims[1])[None, :, None]
    f = ((-1.0) ** j) / sg
    U = np.stack([f * (c + 1) for c in range(7)])
    out = O.sph_ko(sp, U)
    it = (slice(None),) + g.interior()
    assert np.abs(out[it] - (1 - sp.KOeps) * U[it]).max() <= 1e-13 * np.abs(U[it]).max()
    x = np.log(r)
    lin = np.stack([(0.5 * x - 0.3 * th + c) / sg for c in range(7)])
    assert np.abs(O.sph_ko(sp, lin)[it] - lin[it]).max() <= 1e-13 * np.abs(lin[it]).max()
