"""Pins the force-free parts of the Kerr-Schild oracle (oracle/ffe_gr.c).

The reference's GR solver is CUDA-only (src/algorithms/field_solver_gr_EZ.cu),
so no reference build pins the oracle.  These checks state the physics
independently of the oracle's formulas:

- the 3+1 split of the Kerr-Schild 4-metric g = eta + 2H l l (M = 1) is
  computed here from its definition (gamma_ij = g_ij, beta_i = g_0i,
  gamma^ij by matrix inversion, alpha^2 = beta_i beta^i - g_00,
  sqrt(gamma) = sqrt(det gamma_ij)), not from the oracle's closed forms;
- one RK stage's dU/dt (As = 0) on smooth manufactured D, B, P with
  E.B != 0, div D != 0 and P != 0 converges at 4th order to the continuum
  right-hand side of the 3+1 force-free equations (Komissarov 2004, MNRAS
  350, 427): E = alpha D + beta x B, H = alpha B - beta x D (covariant cross
  products with sqrt(gamma)), dD/dt = curl H - J, dB/dt = -curl E,
  J = [rho E x B + (B.curl H - D.curl E) B] / B^2, plus the reference's
  cleaning terms (.cu:281-332: (-alpha grad^i P + beta^i div B) times
  1/2 (1 - tanh(r - 5)) in dB/dt, dP/dt = -alpha div B + beta^i d_i P -
  alpha P / (tau dt)), derivatives by complex step;
- closed forms of single cells: the metric-normed projections
  (.cu:440-535), the outer absorbing layer (boundary.cu:7-79, sigma from
  the x spacing on every axis) and the horizon excision (.cu:610-701);
- every planted mistake (oracle.PLANTS) that exists in this path makes at
  least one of these checks fail.
"""

import numpy as np
import pytest

import oracle as O
from coffeegpu_amd.params import Grid
from tests.golden.cases_gr import gr_params

H = 1e-30


def ks_split(a, x, y, z):
    """3+1 quantities of Cartesian Kerr-Schild (M = 1) from the 4-metric
    g_mu_nu = eta_mu_nu + 2 H l_mu l_nu, H = r^3 / (r^4 + a^2 z^2),
    l = (1, (r x + a y)/(r^2 + a^2), (r y - a x)/(r^2 + a^2), z / r); works
    on complex coordinates.  Returns gd (..,3,3), gu, beta^i (..,3), alpha,
    sqrt(gamma), r."""
    R2 = x * x + y * y + z * z
    w = R2 - a * a
    r = np.sqrt((w + np.sqrt(w * w + 4.0 * a * a * z * z)) / 2.0)
    Hh = r ** 3 / (r ** 4 + a * a * z * z)
    l = np.stack([(r * x + a * y) / (r * r + a * a), (r * y - a * x) / (r * r + a * a), z / r],
                 axis=-1)
    gd = np.eye(3) + 2.0 * Hh[..., None, None] * l[..., :, None] * l[..., None, :]
    beta_d = 2.0 * Hh[..., None] * l
    gu = np.linalg.inv(gd)
    beta_u = np.einsum("...ij,...j->...i", gu, beta_d)
    g00 = -1.0 + 2.0 * Hh
    alpha = np.sqrt((beta_d * beta_u).sum(-1) - g00)
    sg = np.sqrt(np.linalg.det(gd))
    return gd, gu, beta_u, alpha, sg, r


def lower(gd, v):
    return np.einsum("...ij,...j->...i", gd, v)


def cross_cov(sg, a, b):
    """sqrt(gamma) [ijk] a^j b^k (lower result from upper a, b), or
    [ijk] a_j b_k / sqrt(gamma) (upper result from lower a, b) with 1/sg."""
    return sg[..., None] * np.cross(a, b)


# manufactured D, B (upper components) and P
def mf(x, y, z):
    B = np.stack([0.3 * np.sin(y + 0.2 * z), 0.2 * np.cos(z - 0.3 * x),
                  1.0 + 0.2 * np.sin(x) * np.cos(0.5 * y)], axis=-1)
    D = np.stack([0.2 * np.sin(z + 0.4 * x), 0.1 * np.cos(x + y), 0.15 * np.sin(y) + 0.05],
                 axis=-1)
    P = 0.05 * np.cos(x) * np.sin(z + 0.3)
    return D, B, P


def eh(a, x, y, z):
    """Lower E and H from the constitutive relations."""
    gd, gu, bu, al, sg, r = ks_split(a, x, y, z)
    D, B, P = mf(x, y, z)
    E = al[..., None] * lower(gd, D) + cross_cov(sg, bu, B)
    Hl = al[..., None] * lower(gd, B) - cross_cov(sg, bu, D)
    return E, Hl


def partial(f, a, x, y, z, axis):
    e = [0, 0, 0]
    e[axis] = 1j * H
    return np.imag(f(a, x + e[0], y + e[1], z + e[2])) / H


def curl_up(fl, a, x, y, z, sg):
    """(curl F)^i = [ijk] d_j F_k / sqrt(gamma) of a lower-index field fl."""
    d = [partial(fl, a, x, y, z, ax) for ax in range(3)]  # d[j][..., k]
    c = np.stack([d[1][..., 2] - d[2][..., 1], d[2][..., 0] - d[0][..., 2],
                  d[0][..., 1] - d[1][..., 0]], axis=-1)
    return c / sg[..., None]


def div_up(which, a, x, y, z, sg):
    """d_i (sqrt(gamma) F^i) / sqrt(gamma) of D (0) or B (1)."""
    def f(a_, x_, y_, z_):
        s = ks_split(a_, x_, y_, z_)[4]
        return s[..., None] * mf(x_, y_, z_)[which]
    return sum(partial(f, a, x, y, z, ax)[..., ax] for ax in range(3)) / sg


def continuum_rhs(sp, x, y, z):
    a = sp.a
    gd, gu, bu, al, sg, r = (np.real(v) for v in ks_split(a, x, y, z))
    D, B, P = (np.real(v) for v in mf(x, y, z))
    E, Hl = (np.real(v) for v in eh(a, x, y, z))
    cE = curl_up(lambda *q: eh(*q)[0], a, x, y, z, sg)
    cH = curl_up(lambda *q: eh(*q)[1], a, x, y, z, sg)
    rho = div_up(0, a, x, y, z, sg)
    divB = div_up(1, a, x, y, z, sg)
    Bl, Dl = lower(gd, B), lower(gd, D)
    B2 = (B * Bl).sum(-1)
    ExB = np.cross(E, Bl) / sg[..., None]
    J = (rho[..., None] * ExB + ((Bl * cH).sum(-1) - (Dl * cE).sum(-1))[..., None] * B) / B2[..., None]
    dP = np.stack([partial(lambda a_, x_, y_, z_: mf(x_, y_, z_)[2], a, x, y, z, ax)
                   for ax in range(3)], axis=-1)
    gradPu = np.einsum("...ij,...j->...i", gu, dP)
    shape = 0.5 * (1.0 - np.tanh(r - 5.0))
    dDdt = cH - J
    dBdt = -cE + (-al[..., None] * gradPu + bu * divB[..., None]) * shape[..., None]
    dPdt = -al * divB + (bu * dP).sum(-1) - al * P / (sp.tau * sp.dt)
    return dDdt, dBdt, dPdt


def node_xyz(g):
    xs = [g.pos(ax, np.arange(g.dims[ax])) for ax in range(3)]
    Z, Y, X = np.meshgrid(xs[2], xs[1], xs[0], indexing="ij")
    return X, Y, Z


def gr_box(n, a=0.7, **kw):
    base = dict(n=(n, n, n), lower=(3.5, -1.0, -1.0), size=(2.0, 2.0, 2.0), a=a,
                pml=(0, 0, 0), clean_ep=False, check_egb=False, ch2=1.0, tau=3.0)
    base.update(kw)
    return gr_params(**base)


def rhs_errors(n, a=0.7):
    sp = gr_box(n, a)
    g = Grid.from_params(sp, np.float64)
    X, Y, Z = node_xyz(g)
    D, B, P = mf(X, Y, Z)
    U = np.concatenate([np.moveaxis(D, -1, 0), np.moveaxis(B, -1, 0), P[None]])
    got = O.gr_stage_rhs(sp, U)
    dD, dB, dPt = continuum_rhs(sp, X, Y, Z)
    want = np.concatenate([np.moveaxis(dD, -1, 0), np.moveaxis(dB, -1, 0), dPt[None]])
    it = (slice(None),) + g.interior()
    out = []
    for sl in (slice(0, 3), slice(3, 6), slice(6, 7)):
        w = want[it][sl]
        out.append(float(np.abs(got[it][sl] - w).max() / np.abs(w).max()))
    return np.array(out)


def manufactured_converges(a=0.7):
    e1, e2 = rhs_errors(12, a), rhs_errors(24, a)
    ratio = e1 / e2
    # at least 4th order (the P row converges faster: its right-hand side is
    # dominated by the exact damping term)
    return bool(np.all(ratio > 11.0) and np.all(e2 < 1e-4)), e1, e2


@pytest.mark.parametrize("a", [0.0, 0.7, 0.99])
def test_stage_rhs_converges_to_continuum_at_4th_order(a):
    ok, e1, e2 = manufactured_converges(a)
    print(f"a={a}: D,B,P rel err {e1} -> {e2}, ratio {e1 / e2}")
    assert ok, (e1, e2)


def test_cleaning_weight_varies_over_the_box():
    """The box spans r ~ 3.4 .. 5.7, where 1/2 (1 - tanh(r - 5)) runs from
    0.98 down to 0.2, so the manufactured check sees the weight's shape."""
    g = Grid.from_params(gr_box(12), np.float64)
    r = ks_split(0.7, *node_xyz(g))[5]
    s = 0.5 * (1 - np.tanh(r - 5.0))
    assert s.max() > 0.95 and s.min() < 0.25


# ---- single-cell closed forms --------------------------------------------------

def projection_closed_form(a=0.9):
    """D -> D - (D.B)B/B^2 (metric dot products), then scaled to |D| = |B|
    where it exceeds B; B, P unchanged.  Afterwards D_i B^i = 0 and
    D.D <= B.B in the metric norm, and a second application changes nothing."""
    sp = gr_params(n=(10, 10, 10), lower=(-4.0, 1.5, -3.0), size=(5.0, 5.0, 5.0), a=a,
                   pml=(0, 0, 0))
    g = Grid.from_params(sp, np.float64)
    X, Y, Z = node_xyz(g)
    gd = ks_split(a, X, Y, Z)[0]
    rng = np.random.default_rng(4)
    D = rng.standard_normal(X.shape + (3,)) * 0.6
    B = rng.standard_normal(X.shape + (3,))
    D[rng.random(X.shape) < 0.3] *= 4.0
    U = np.concatenate([np.moveaxis(D, -1, 0), np.moveaxis(B, -1, 0),
                        rng.standard_normal((1,) + X.shape)])
    out = O.gr_project(sp, U)
    Bl = lower(gd, B)
    B2 = (B * Bl).sum(-1)
    Dp = D - ((D * Bl).sum(-1) / B2)[..., None] * B
    D2 = (Dp * lower(gd, Dp)).sum(-1)
    want = Dp * np.where(D2 > B2, np.sqrt(B2 / D2), 1.0)[..., None]
    it = g.interior()
    Do = np.moveaxis(out[0:3], 0, -1)
    err = np.abs(Do[it] - want[it]).max() / np.abs(want[it]).max()
    DB = (Do * Bl).sum(-1)[it]
    DD = (Do * lower(gd, Do)).sum(-1)[it]
    twice = O.gr_project(sp, out)
    ok = (err < 1e-13 and np.abs(DB).max() < 1e-12 and np.all(DD <= B2[it] * (1 + 1e-13)) and
          np.abs(twice - out).max() < 1e-14 and np.array_equal(out[3:7], U[3:7]))
    return bool(ok), err


def test_projection_closed_form():
    ok, e = projection_closed_form()
    assert ok, e


def absorbing_closed_form(a=0.9):
    """Outer layer: outside [lower + pml dx, lower + size - pml dx] on any
    axis, sigma = sigpml sum_axes (depth / (pmllen dx))^3 (the x spacing on
    every axis); horizon: r < 0.7 r_H adds sigma = sigpml ((0.7 r_H - r) /
    (0.2 r_H))^3 dt and r < 0.01 r_H is zeroed (r_H = 1 + sqrt(1 - a^2));
    D, B -> e^-sigma Un + (1 - e^-sigma)/sigma (U - Un) per stage of the two;
    P untouched."""
    sp = gr_params(n=(20, 16, 18), lower=(-3.0, -2.0, -2.5), size=(6.0, 4.4, 5.0), a=a,
                   pml=(4, 3, 5), pmllen=3, sigpml=0.8)
    g = Grid.from_params(sp, np.float64)
    X, Y, Z = node_xyz(g)
    # r of the quartic, with the reference's floor of TINY under the inner
    # square root (metric_cks_impl.hpp:17-24): it moves r only on the ring
    # singularity (x^2 + y^2 = a^2, z = 0), which this grid has a node on
    w = X * X + Y * Y + Z * Z - a * a
    r = np.sqrt((w + np.sqrt(np.maximum(w * w + 4 * a * a * Z * Z, 1e-11))) / 2)
    rng = np.random.default_rng(9)
    U = rng.standard_normal((7,) + X.shape)
    Un = rng.standard_normal((7,) + X.shape)
    out = O.gr_absorb(sp, U, Un)

    def blend(V, sig):
        f = np.where(sig > 1e-11, (1.0 - np.exp(-sig)) / np.where(sig > 0, sig, 1.0), 1.0)
        return np.where(sig > 1e-11, np.exp(-sig) * Un[0:6] + f * (V - Un[0:6]), V)

    sig = np.zeros(X.shape)
    scale = sp.pmllen * g.delta[0]
    for ax, C in enumerate((X, Y, Z)):
        lo = sp.lower[ax] + sp.pml[ax] * g.delta[ax]
        hi = sp.lower[ax] + sp.size[ax] - sp.pml[ax] * g.delta[ax]
        sig += np.where(C > hi, sp.sigpml * ((C - hi) / scale) ** 3, 0.0)
        sig += np.where(C < lo, sp.sigpml * ((lo - C) / scale) ** 3, 0.0)
    want = U.copy()
    want[0:6] = blend(U[0:6], sig)
    rH = 1 + np.sqrt(1 - a * a)
    sig_h = np.where(r < 0.7 * rH, sp.sigpml * ((0.7 * rH - r) / (0.2 * rH)) ** 3 * sp.dt, 0.0)
    want[0:6] = blend(want[0:6], sig_h)
    want[0:6] = np.where(r < 0.01 * rH, 0.0, want[0:6])
    it = (slice(None),) + g.interior()
    err = np.abs(out[it] - want[it]).max()
    n_layer = (sig[g.interior()] > 0.05).sum()
    n_hole = (sig_h[g.interior()] > 1e-6).sum()
    return bool(err < 1e-13 and n_layer > 100 and n_hole > 3), err


def test_absorbing_and_excision_closed_form():
    ok, e = absorbing_closed_form()
    assert ok, e


CHECKS = {
    "manufactured": lambda: manufactured_converges(0.7)[0],
    "projection": lambda: projection_closed_form()[0],
    "absorbing": lambda: absorbing_closed_form()[0],
}

GR_PLANTS = ["J_SIGN", "J_NO_SQRTG", "GRADP_LOWER", "TAU_DT", "SIGMA_SQUARED", "PROJ_FLAT",
             "ROT_NO_SQRTG", "DIV_CENTRE", "CLEAN_SHAPE", "EXCISION_NODT"]


def test_checks_pass_without_plant():
    with O.planted(None):
        assert all(f() for f in CHECKS.values())


@pytest.mark.parametrize("mistake", GR_PLANTS)
def test_planted_mistake_fails_a_check(mistake):
    with O.planted(mistake):
        failed = [name for name, f in CHECKS.items() if not f()]
    print(f"{mistake}: caught by {failed}")
    assert failed, f"no check catches {mistake}"


def test_ko_nyquist_decay_on_every_field():
    """Flat 6th-order KO (.cu:537-608) on D, B and P: the checkerboard
    (-1)^(i+j+k) has a 6th difference of -64 times itself along each axis,
    so one application multiplies it by 1 - 3 KOeps; a field linear in x, y, z
    is left untouched."""
    sp = gr_params(n=(10, 10, 10), lower=(2.0, 2.0, 2.0), size=(3.0, 3.0, 3.0), a=0.9,
                   pml=(0, 0, 0), KOeps=0.07)
    g = Grid.from_params(sp, np.float64)
    k, j, i = np.meshgrid(*[np.arange(d) for d in g.shape], indexing="ij")
    chk = (-1.0) ** (i + j + k)
    U = np.stack([chk * (c + 1) for c in range(7)])
    out = O.gr_ko(sp, U)
    it = (slice(None),) + g.interior()
    assert np.abs(out[it] - (1 - 3 * sp.KOeps) * U[it]).max() < 1e-13
    X, Y, Z = node_xyz(g)
    lin = np.stack([0.3 * X - Y + 2 * Z + c for c in range(7)])
    assert np.abs(O.gr_ko(sp, lin)[it] - lin[it]).max() < 1e-13
