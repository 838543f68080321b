"""Simulation parameters and grid geometry.

`SimParams` mirrors `Coffee::sim_params` (src/sim_params.h:15-124): the same
field names and the same defaults, including the reference's trap defaults
(pml=60, use_edotb_damping=True, KOeps=0.5, KO_geometry=True, radius=10,
omega=0.001, b0=1e3).  `parse_config` follows src/sim_params.cpp:7-146
(missing keys keep the defaults; integer arrays for `nodes` are read as
floats, as cpptoml's get_array_of<double> does).  `Grid` follows the set-up
in sim_environment (src/sim_env_impl.hpp:52-60) and Grid::pos
(src/data/grid.h:116-134).
"""

from __future__ import annotations

import dataclasses
import math
from dataclasses import dataclass, field
from typing import Any, Dict, List

import numpy as np


def _l3(v):
    return field(default_factory=lambda: list(v))


@dataclass
class SimParams:
    dt: float = 0.01
    max_steps: int = 10000
    data_interval: int = 100
    snapshot_interval: int = 1000
    periodic_boundary: List[bool] = _l3([False, False, False])
    downsample: int = 1
    slice_x: bool = False
    slice_x_pos: float = 0.0
    slice_y: bool = False
    slice_y_pos: float = 0.0
    slice_z: bool = False
    slice_z_pos: float = 0.0
    slice_xy: bool = False
    slice_interval: int = 100
    downsample_2d: int = 1
    N: List[int] = _l3([1, 0, 0])
    guard: List[int] = _l3([0, 0, 0])
    lower: List[float] = _l3([0.0, 0.0, 0.0])
    size: List[float] = _l3([1.0, 0.0, 0.0])
    nodes: List[int] = _l3([1, 0, 0])
    shift_ghost: int = 0
    kn: List[int] = _l3([1, 1, 0])
    a: float = 0.99
    calc_current: bool = True
    clean_ep: bool = True
    check_egb: bool = True
    pml: List[int] = _l3([60, 60, 60])
    pmllen: int = 10
    sigpml: float = 0.1
    divB_clean: bool = True
    ch2: float = 1.0
    tau: float = 2.0
    KOeps: float = 0.50
    KO_geometry: bool = True
    use_edotb_damping: bool = True
    damp_gamma: float = 0.2
    problem: int = 0
    radius: float = 10.0
    omega: float = 0.001
    b0: float = 1e3
    p1: float = 0.0
    p2: float = 0.0
    p3: float = 1.0
    q11: float = 0.0
    q12: float = 0.0
    q13: float = 0.0
    q22: float = 0.0
    q23: float = 0.0
    q_offset_x: float = 0.0
    q_offset_y: float = 0.0
    q_offset_z: float = 0.0
    tp_start: float = 0.0
    tp_end: float = 0.5
    rpert1: float = 5.0
    rpert2: float = 10.0
    thp1: float = 0.3
    thp2: float = 0.4
    dw0: float = 0.0
    nT: float = 1.0
    shear: float = 0.0
    ph1: float = 3.0 / 16.0
    ph2: float = 5.0 / 16.0
    dph: float = 1.0 / 64.0
    theta0: float = 0.2
    drpert: float = 0.1
    tp_start1: float = 1.0
    tp_end1: float = 1.5
    rpert11: float = 5.0
    rpert21: float = 10.0
    thp11: float = 0.3
    thp21: float = 0.4
    dw1: float = 0.0
    nT1: float = 1.0
    shear1: float = 0.0
    ph11: float = 3.0 * math.pi / 16.0
    ph21: float = 5.0 * math.pi / 16.0
    dph1: float = math.pi / 64.0
    pert_type: int = 0
    skymap_Nth: int = 256
    skymap_Nph: int = 512

    def replace(self, **kw) -> "SimParams":
        return dataclasses.replace(self, **kw)

    def to_dict(self) -> Dict[str, Any]:
        return dataclasses.asdict(self)

    def to_toml(self) -> str:
        """Serialise in the key/type layout parse_config reads."""
        out = []
        for f in dataclasses.fields(self):
            v = getattr(self, f.name)
            if f.name == "kn":
                continue
            if isinstance(v, list):
                if f.name in ("lower", "size", "nodes"):
                    out.append(f"{f.name} = [ " + ", ".join(repr(float(x)) for x in v) + " ]")
                elif f.name == "periodic_boundary":
                    out.append(f"{f.name} = [ " + ", ".join("true" if x else "false" for x in v) + " ]")
                else:
                    out.append(f"{f.name} = [ " + ", ".join(str(int(x)) for x in v) + " ]")
            elif isinstance(v, bool):
                out.append(f"{f.name} = {'true' if v else 'false'}")
            elif isinstance(v, int):
                out.append(f"{f.name} = {v}")
            else:
                out.append(f"{f.name} = {float(v)!r}")
        return "\n".join(out) + "\n"


_INT_KEYS = {"max_steps", "data_interval", "snapshot_interval", "downsample", "slice_interval",
             "downsample_2d", "shift_ghost", "pmllen", "problem", "pert_type"}


def parse_config(path: str) -> SimParams:
    """Read a config.toml into SimParams (src/sim_params.cpp:7-146)."""
    import tomli

    with open(path, "rb") as f:
        cfg = tomli.load(f)
    p = SimParams()
    for f_ in dataclasses.fields(p):
        if f_.name not in cfg or f_.name in ("kn", "skymap_Nth", "skymap_Nph", "thp1", "thp2",
                                              "thp11", "thp21", "shear", "shear1"):
            # parse_config never reads kn/skymap; thp*/shear* are read but unused.
            if f_.name in cfg and f_.name in ("thp1", "thp2", "thp11", "thp21", "shear", "shear1"):
                setattr(p, f_.name, float(cfg[f_.name]))
            continue
        v = cfg[f_.name]
        cur = getattr(p, f_.name)
        if isinstance(cur, list):
            vals = list(v)[:3]
            if f_.name == "periodic_boundary":
                vals = [bool(x) for x in vals]
            elif f_.name in ("lower", "size"):
                vals = [float(x) for x in vals]
            elif f_.name == "nodes":
                vals = [int(float(x)) for x in vals]
            else:
                vals = [int(x) for x in vals]
            setattr(p, f_.name, vals)
        elif isinstance(cur, bool):
            setattr(p, f_.name, bool(v))
        elif f_.name in _INT_KEYS:
            # cpptoml's get_as<int> yields nothing for a float value: default kept
            if isinstance(v, int) and not isinstance(v, bool):
                setattr(p, f_.name, int(v))
        else:
            setattr(p, f_.name, float(v))
    return p


@dataclass
class Grid:
    """Geometry of one (single-rank) domain; positions of collocated nodes."""

    dims: List[int]
    guard: List[int]
    delta: List[float]
    inv_delta: List[float]
    lower: List[float]
    sizes: List[float]
    dtype: Any = np.float64

    @staticmethod
    def from_params(p: SimParams, dtype=np.float64) -> "Grid":
        dt = np.dtype(dtype).type
        dims, guard, delta, inv, lower, sizes = [], [], [], [], [], []
        for i in range(3):
            n = int(p.N[i]) if p.N[i] else 1
            g = int(p.guard[i])
            dims.append(n + 2 * g)
            guard.append(g)
            size = dt(p.size[i])
            d = dt(size / dt(n)) if n else dt(1.0)
            delta.append(float(d))
            inv.append(float(dt(1.0 / float(d))) if float(d) != 0.0 else math.inf)
            lower.append(float(dt(p.lower[i])))
            sizes.append(float(size))
        return Grid(dims, guard, delta, inv, lower, sizes, np.dtype(dtype).type)

    @property
    def shape(self):
        """numpy shape (k, j, i) of one Fortran-order array."""
        return (self.dims[2], self.dims[1], self.dims[0])

    @property
    def size(self) -> int:
        return self.dims[0] * self.dims[1] * self.dims[2]

    def reduced_dim(self, i: int) -> int:
        return self.dims[i] - 2 * self.guard[i]

    def interior_cells(self) -> int:
        return self.reduced_dim(0) * self.reduced_dim(1) * self.reduced_dim(2)

    def pos(self, axis: int, n):
        """Grid::pos(axis, n, stagger=1) in the grid's Scalar precision."""
        t = self.dtype
        n = np.asarray(n)
        return t(self.lower[axis]) + t(self.delta[axis]) * ((n - self.guard[axis]).astype(t) + t(0.0))

    def coords(self):
        """Broadcastable node coordinates x[i], y[j], z[k] shaped for (k, j, i) arrays."""
        x = self.pos(0, np.arange(self.dims[0]))[None, None, :]
        y = self.pos(1, np.arange(self.dims[1]))[None, :, None]
        z = self.pos(2, np.arange(self.dims[2]))[:, None, None]
        return x, y, z

    def interior(self):
        """Slices selecting the interior of a (k, j, i) array."""
        g = self.guard
        return (slice(g[2], self.dims[2] - g[2]), slice(g[1], self.dims[1] - g[1]),
                slice(g[0], self.dims[0] - g[0]))


def reference_simd_lanes(dtype) -> int:
    """SIMD width of the reference's AVX-512 CPU build for this Scalar
    (src/utils/simd.h:27-51): 16 floats or 8 doubles per vector."""
    return 16 if np.dtype(dtype) == np.float32 else 8
