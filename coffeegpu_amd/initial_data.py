"""Initial data for the solver configurations, evaluated on collocated nodes
including guard cells, as vector_field::initialize does
(src/data/fields_impl.hpp:8-30).

Each generator returns a dict of numpy arrays shaped (k, j, i) (Fortran order
in memory when flattened) in the requested dtype: Ex, Ey, Ez, Bx, By, Bz, P.
"""

from __future__ import annotations

import math
from typing import Dict

import numpy as np

from .params import Grid, SimParams

FIELDS = ("Ex", "Ey", "Ez", "Bx", "By", "Bz", "P")

_M64 = (1 << 64) - 1


def splitmix64(seed: int, n: int):
    """n uniform doubles in [0, 1) from the splitmix64 generator (Steele et
    al. 2014), seeded by `seed`; 53 high bits per draw."""
    out = []
    s = seed & _M64
    for _ in range(n):
        s = (s + 0x9E3779B97F4A7C15) & _M64
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        z = z ^ (z >> 31)
        out.append((z >> 11) * (1.0 / (1 << 53)))
    return out


def _coords64(grid: Grid):
    g64 = Grid(grid.dims, grid.guard, grid.delta, grid.inv_delta, grid.lower, grid.sizes, np.float64)
    return g64.coords()


def _pack(grid: Grid, dtype, **comp) -> Dict[str, np.ndarray]:
    out = {}
    for name in FIELDS:
        v = comp.get(name, 0.0)
        a = np.empty(grid.shape, dtype=dtype)
        a[...] = v
        out[name] = a
    return out


def em_wave(grid: Grid, amp: float = 0.1, phase: float = 0.0, dtype=np.float64):
    """Plane-polarised transverse wave on a uniform B0 = z:
    E = (-amp sin(2 pi z + phase), 0, 0), B = (0, amp sin(2 pi z + phase), 1).
    An exact force-free solution (E.B = 0, J = 0) travelling along -z."""
    x, y, z = _coords64(grid)
    s = amp * np.sin(2.0 * math.pi * z + phase)
    return _pack(grid, dtype, Ex=-s, By=s, Bz=1.0)


def em_wave_seeded(grid: Grid, seed: int, amp: float = 0.1, dtype=np.float64):
    phase = 2.0 * math.pi * splitmix64(seed, 1)[0]
    return em_wave(grid, amp, phase, dtype), phase


def alfven_packets(grid: Grid, a1: float, a2: float, width: float = 0.1, dtype=np.float64):
    """Two counter-propagating Alfven packets on B0 = z in a unit periodic box.

    Packet 1 (centre z=0.25, moving +z): dB = a1 g1 sin(2 pi x) y, dE = -z x dB.
    Packet 2 (centre z=0.75, moving -z): dB = a2 g2 sin(2 pi y) x, dE = +z x dB.
    g = exp(-d^2 / (2 width^2)) with d the periodic (minimum-image) distance
    in z.  Each dB is divergence-free."""
    x, y, z = _coords64(grid)
    L = grid.sizes[2] if grid.sizes[2] else 1.0

    def gauss(zc):
        d = (z - zc) / L
        d = (d + 0.5) - np.floor(d + 0.5) - 0.5
        return np.exp(-(d * L) ** 2 / (2.0 * width ** 2))

    sx = np.sin(2.0 * math.pi * x / (grid.sizes[0] or 1.0))
    sy = np.sin(2.0 * math.pi * y / (grid.sizes[1] or 1.0))
    b1 = a1 * gauss(0.25 * L) * sx  # y-component, packet 1
    b2 = a2 * gauss(0.75 * L) * sy  # x-component, packet 2
    # dE1 = -z x (b1 y) = b1 x ; dE2 = +z x (b2 x) = b2 y
    return _pack(grid, dtype, Ex=b1 + 0 * y, Ey=b2 + 0 * x, Bx=b2 + 0 * x, By=b1 + 0 * y, Bz=1.0)


def alfven_packets_seeded(grid: Grid, seed: int, dtype=np.float64):
    u = splitmix64(seed, 2)
    a1 = 0.1 + 0.4 * u[0]
    a2 = 0.1 + 0.4 * u[1]
    return alfven_packets(grid, a1, a2, dtype=dtype), (a1, a2)


def dipole2_np(x, y, z, p1, p2, p3, phase, n, tiny):
    """dipole2 (src/algorithms/pulsar_impl.hpp:57-80) in float64."""
    c, s = math.cos(phase), math.sin(phase)
    p1t = p1 * c - p2 * s
    p2t = p1 * s + p2 * c
    p3t = p3
    r2 = x * x + y * y + z * z
    den = np.sqrt(r2 * r2 * r2) + tiny
    pr = p1t * x + p2t * y + p3t * z
    comp = (x, y, z)[n]
    pt = (p1t, p2t, p3t)[n]
    return (3.0 * comp * pr / (r2 + tiny) - pt) / den


def quadrupole_np(x, y, z, p: SimParams, phase, n, tiny):
    """quadrupole (pulsar_impl.hpp:82-119) in float64."""
    c, s = math.cos(phase), math.sin(phase)
    q11, q12, q13, q22, q23 = p.q11, p.q12, p.q13, p.q22, p.q23
    q11t = q11 * c * c - 2.0 * q12 * c * s + q22 * s * s
    q12t = q12 * math.cos(2.0 * phase) + (q11 - q22) * c * s
    q22t = q22 * c * c + q11 * s * s + 2.0 * q12 * c * s
    q13t = q13 * c - q23 * s
    q23t = q23 * c + q13 * s
    q33t = -q11 - q22
    ox = p.q_offset_x * c - p.q_offset_y * s
    oy = p.q_offset_x * s + p.q_offset_y * c
    x1, y1, z1 = x - ox, y - oy, z - p.q_offset_z
    r2 = x1 * x1 + y1 * y1 + z1 * z1
    r7 = r2 * r2 * r2 * np.sqrt(r2)
    xqx = (q11t * x1 * x1 + 2.0 * q12t * x1 * y1 + q22t * y1 * y1
           + 2.0 * (q13t * x1 + q23t * y1) * z1 + q33t * z1 * z1)
    row = ((q11t, q12t, q13t), (q12t, q22t, q23t), (q13t, q23t, q33t))[n]
    comp = (x1, y1, z1)[n]
    return (-2.0 * (row[0] * x1 + row[1] * y1 + row[2] * z1) * r2 + 5.0 * comp * xqx) / (r7 + tiny)


def multipole_field(grid: Grid, p: SimParams, phase: float = 0.0, dtype=np.float64):
    """b0 * quadru_dipole(x, y, z, ..., phase, n) for n = 0, 1, 2
    (pulsar_impl.hpp:121-129)."""
    tiny = 1e-7 if np.dtype(dtype) == np.float32 else 1e-11
    x, y, z = _coords64(grid)
    out = []
    for n in range(3):
        v = p.b0 * (dipole2_np(x, y, z, p.p1, p.p2, p.p3, phase, n, tiny)
                    + quadrupole_np(x, y, z, p, phase, n, tiny))
        out.append(np.broadcast_to(v, grid.shape).astype(dtype))
    return out


def pulsar_dipole(grid: Grid, p: SimParams, dtype=np.float64):
    """user_pulsar3d_EZ.hpp: B = b0 * quadru_dipole(phase 0), E = 0, P = 0."""
    bx, by, bz = multipole_field(grid, p, 0.0, dtype)
    return _pack(grid, dtype, Bx=bx, By=by, Bz=bz)


def pulsar_dipole_sph(grid: Grid, p: SimParams, dtype=np.float64):
    """Aligned dipole on the log-spherical grid (x = ln r, y = theta), as
    user_pulsar_sph.hpp sets it: upper components
    B^x = b0 2 cos(th) / r^3 / sqrt(g11), B^y = b0 sin(th) / r^3 / sqrt(g22),
    B^z = 0; E = 0, P = 0 (dipole_sph_2d, pulsar_impl.hpp:131-139)."""
    tiny = 1e-7 if np.dtype(dtype) == np.float32 else 1e-11
    x, y, _ = _coords64(grid)
    r = np.exp(x)
    g11 = np.maximum(np.sqrt(r * r), tiny)
    g22 = np.maximum(np.sqrt(r * r), tiny)
    bx = p.b0 * 2.0 * np.cos(y) / r ** 3 / g11
    by = p.b0 * np.sin(y) / r ** 3 / g22
    shape = grid.shape
    return _pack(grid, dtype, Bx=np.broadcast_to(bx, shape), By=np.broadcast_to(by, shape))


def periodic_fill(U: np.ndarray, grid: Grid, periodic) -> np.ndarray:
    """Fill the guard cells of periodic axes with the interior one period away,
    x then y then z, as send_guard_cells does on one periodic rank
    (src/sim_env.cpp:365-438).  U is (..., k, j, i); returns a copy."""
    U = np.array(U, copy=True)
    for ax in range(3):
        if not periodic[ax]:
            continue
        g = grid.guard[ax]
        n = grid.reduced_dim(ax)
        if g == 0:
            continue
        nd = U.ndim
        dim = nd - 1 - ax  # numpy axis of grid axis `ax`
        idx_lo = [slice(None)] * nd
        src_lo = [slice(None)] * nd
        idx_lo[dim] = slice(0, g)
        src_lo[dim] = slice(n, n + g)
        U[tuple(idx_lo)] = U[tuple(src_lo)]
        idx_hi = [slice(None)] * nd
        src_hi = [slice(None)] * nd
        idx_hi[dim] = slice(n + g, n + 2 * g)
        src_hi[dim] = slice(g, 2 * g)
        U[tuple(idx_hi)] = U[tuple(src_hi)]
    return U


def stack(fields: Dict[str, np.ndarray]) -> np.ndarray:
    """(7, k, j, i) array in the field order Ex..Bz, P."""
    return np.stack([fields[n] for n in FIELDS])


def cks_metric_np(a: float, x, y, z):
    """Cartesian Kerr-Schild (M = 1) in FP64: r, s = 2H and l on numpy arrays
    (the quantities of src/algorithms/metric_cks_impl.hpp; gamma_ij =
    delta_ij + s l_i l_j, sqrt(gamma) = sqrt(1 + s))."""
    tiny = 1e-11
    x, y, z = (np.asarray(v, dtype=np.float64) for v in (x, y, z))
    w = x * x + y * y + z * z - a * a
    r = np.sqrt((w + np.sqrt(np.maximum(w * w + 4.0 * (a * z) ** 2, tiny))) / 2.0)
    Q = np.maximum(r ** 4 + (a * z) ** 2, tiny)
    A = np.maximum(a * a + r * r, tiny)
    s = 2.0 * r ** 3 / Q
    with np.errstate(divide="ignore", invalid="ignore"):
        lz = np.where(r > 0, z / np.where(r > 0, r, 1.0), 0.0)
    return r, s, (r * x + a * y) / A, (r * y - a * x) / A, lz


def wald_uniform(grid: Grid, a: float, noise: float = 0.0, seed: int = 0, dtype=np.float64):
    """Initial data of src/user_wald2.hpp:13-49: B = (0, 0, 1/sqrt(gamma)),
    D = 0, P = 0 on every node including guards, plus `noise` times seeded
    standard-normal white noise on the three B components."""
    X, Y, Z = _coords64(grid)
    _, s, _, _, _ = cks_metric_np(a, X, Y, Z)
    shape = grid.shape
    zero = np.zeros(shape)
    Bz = 1.0 / np.sqrt(1.0 + s)
    comp = {"Ex": zero, "Ey": zero, "Ez": zero, "Bx": zero.copy(), "By": zero.copy(), "Bz": Bz,
            "P": zero}
    if noise:
        rng = np.random.default_rng(seed)
        for key in ("Bx", "By", "Bz"):
            comp[key] = comp[key] + noise * rng.standard_normal(shape)
    return _pack(grid, dtype, **comp)
