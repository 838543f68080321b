"""Host-side mirror of field_solver_EZ_spherical
(src/algorithms/field_solver_EZ_spherical.h:6-32) and the batched runner.

FieldSolverEZSpherical keeps the reference's constructor/evolve_fields/
total_energy surface on the same SimData as the flat solver (2D grids:
N[2] = 1, guard[2] = 0).  SphBatch advances many independent runs together
(no reference counterpart).  Both call libcoffee_hip.so only.
"""

from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from .solver import MultiArray, SimData, SimEnvironment, VectorField, coffee_grid, coffee_params


def _torch():
    import torch

    return torch


class FieldSolverEZSpherical:
    def __init__(self, data: SimData, env: SimEnvironment, stream=None):
        g, dt, dev = env.grid(), env.dtype, env.device
        self.m_data = data
        self.m_env = env
        self.dE = VectorField(g, dt, dev)
        self.dB = VectorField(g, dt, dev)
        self.Etmp = VectorField(g, dt, dev)
        self.Btmp = VectorField(g, dt, dev)
        self.dP = MultiArray(g, dt, dev)
        self.Ptmp = MultiArray(g, dt, dev)
        self._lib = _lib.load()
        self._ctx = ctypes.c_void_p()
        gg = coffee_grid(g)
        pp = coffee_params(env.params())
        _lib.check(self._lib.coffee_ctx_create(_lib.COFFEE_SOLVER_EZ_SPHERICAL, dt.itemsize,
                                               ctypes.byref(gg), ctypes.byref(pp),
                                               ctypes.byref(self._ctx)))
        self._stream = stream
        f = _lib.CoffeeFields()
        for i in range(3):
            f.E[i] = data.E.dev_ptr(i)
            f.B[i] = data.B.dev_ptr(i)
            f.dE[i] = self.dE.dev_ptr(i)
            f.dB[i] = self.dB.dev_ptr(i)
            f.Etmp[i] = self.Etmp.dev_ptr(i)
            f.Btmp[i] = self.Btmp.dev_ptr(i)
            f.J[i] = data.B0.dev_ptr(i)
        f.P = data.P.dev_ptr()
        f.dP = self.dP.dev_ptr()
        f.Ptmp = self.Ptmp.dev_ptr()
        f.divE = data.divE.dev_ptr()
        f.divB = data.divB.dev_ptr()
        self._fields = f

    def _s(self):
        s = self._stream if self._stream is not None else _torch().cuda.current_stream()
        return ctypes.c_void_p(s.cuda_stream)

    def evolve_fields(self, time: float):
        _lib.check(self._lib.coffee_ez_step(self._ctx, ctypes.byref(self._fields),
                                            ctypes.c_double(time), self._s()), self._ctx)

    def total_energy(self, f: VectorField) -> float:
        out = ctypes.c_double()
        comp = (ctypes.c_void_p * 3)(f.dev_ptr(0), f.dev_ptr(1), f.dev_ptr(2))
        _lib.check(self._lib.coffee_total_energy(self._ctx, comp, ctypes.byref(out), self._s()),
                   self._ctx)
        return out.value

    def set_diagnostics(self, on: bool):
        _lib.check(self._lib.coffee_set_diagnostics(self._ctx, int(on)), self._ctx)

    def residuals(self):
        """Residual monitor of the state (metric dot products, sqrt(gamma)-
        weighted div B): energy, max_divB, max_EdotB, max_E2mB2, nonfinite."""
        out = (ctypes.c_double * 5)()
        _lib.check(self._lib.coffee_residuals(self._ctx, ctypes.byref(self._fields), out,
                                              self._s()), self._ctx)
        return _residual_dict(out, 0)

    def set_timing(self, on: bool):
        _lib.check(self._lib.coffee_set_timing(self._ctx, int(on)), self._ctx)

    def close(self):
        if self._ctx:
            self._lib.coffee_ctx_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _residual_dict(out, o):
    return {"energy": out[o], "max_divB": out[o + 1], "max_EdotB": out[o + 2],
            "max_E2mB2": out[o + 3], "nonfinite": out[o + 4]}


def tile_shapes() -> List[tuple]:
    lib = _lib.load()
    out = []
    for t in range(lib.coffee_sph_num_tiles()):
        tx, ty = ctypes.c_int(), ctypes.c_int()
        _lib.check(lib.coffee_sph_tile_shape(t, ctypes.byref(tx), ctypes.byref(ty)))
        out.append((tx.value, ty.value))
    return out


class SphBatch:
    """Independent log-spherical runs advanced together on one GPU.

    envs: one SimEnvironment per run (its own grid and parameters); states:
    SimData per run holding E, B, P (caller-owned, as in a single run);
    nsteps: per-run step targets (runs stop advancing when reached)."""

    def __init__(self, envs: Sequence[SimEnvironment], datas: Sequence[SimData],
                 nsteps: Sequence[int], t0: Optional[Sequence[float]] = None, stream=None):
        n = len(envs)
        dt = envs[0].dtype
        self._lib = _lib.load()
        grids = (_lib.CoffeeGrid * n)(*[coffee_grid(e.grid()) for e in envs])
        params = (_lib.CoffeeParams * n)(*[coffee_params(e.params()) for e in envs])
        fields = (_lib.CoffeeFields * n)()
        for r, d in enumerate(datas):
            for i in range(3):
                fields[r].E[i] = d.E.dev_ptr(i)
                fields[r].B[i] = d.B.dev_ptr(i)
            fields[r].P = d.P.dev_ptr()
        ns = (ctypes.c_int * n)(*[int(v) for v in nsteps])
        tt = (ctypes.c_double * n)(*([0.0] * n if t0 is None else [float(v) for v in t0]))
        self._ctx = ctypes.c_void_p()
        _lib.check(self._lib.coffee_sph_batch_create(
            dt.itemsize, n, ctypes.cast(grids, ctypes.c_void_p), ctypes.cast(params, ctypes.c_void_p),
            ctypes.cast(fields, ctypes.c_void_p), ctypes.cast(ns, ctypes.c_void_p),
            ctypes.cast(tt, ctypes.c_void_p), ctypes.byref(self._ctx)))
        self.n = n
        self.datas = list(datas)
        self._stream = stream
        self._keep = (grids, params, fields, ns, tt)

    def _s(self):
        s = self._stream if self._stream is not None else _torch().cuda.current_stream()
        return ctypes.c_void_p(s.cuda_stream)

    def step(self, nsteps: int = 1):
        _lib.check(self._lib.coffee_sph_batch_step(self._ctx, int(nsteps), self._s()), self._ctx)

    def set_tile(self, index: int):
        _lib.check(self._lib.coffee_sph_batch_set_tile(self._ctx, int(index)), self._ctx)

    def autotune(self, trials: int = 5):
        best = ctypes.c_int()
        ms = (ctypes.c_double * len(tile_shapes()))()
        _lib.check(self._lib.coffee_sph_batch_autotune(self._ctx, int(trials), self._s(),
                                                       ctypes.byref(best),
                                                       ctypes.cast(ms, ctypes.c_void_p)),
                   self._ctx)
        return best.value, list(ms)

    def use_graph(self, on: bool = True):
        _lib.check(self._lib.coffee_sph_batch_use_graph(self._ctx, int(on)), self._ctx)

    def set_timing(self, on: bool):
        _lib.check(self._lib.coffee_set_timing(self._ctx, int(on)), self._ctx)

    def get_timing(self):
        n = ctypes.c_int()
        names = ((ctypes.c_char * 64) * 32)()
        ms = (ctypes.c_double * 32)()
        cnt = (ctypes.c_int * 32)()
        _lib.check(self._lib.coffee_get_timing(self._ctx, 32, names, ms, cnt, ctypes.byref(n)),
                   self._ctx)
        return {names[i].value.decode(): (cnt[i], ms[i]) for i in range(n.value)}

    def luminosity(self, radius) -> np.ndarray:
        """Spin-down luminosity of every run through r = radius (scalar or one
        per run)."""
        r = np.broadcast_to(np.asarray(radius, dtype=np.float64), (self.n,)).copy()
        out = np.zeros(self.n, dtype=np.float64)
        _lib.check(self._lib.coffee_sph_batch_luminosity(
            self._ctx, r.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p),
            self._s()), self._ctx)
        return out

    def residuals(self) -> List[dict]:
        """Residual monitor of every run (see FieldSolverEZSpherical.residuals)."""
        n = self.n
        out = (ctypes.c_double * (5 * n))()
        _lib.check(self._lib.coffee_sph_batch_residuals(self._ctx, ctypes.cast(out, ctypes.c_void_p),
                                                        self._s()), self._ctx)
        return [_residual_dict(out, 5 * r) for r in range(n)]

    def status(self):
        act = (ctypes.c_int * self.n)()
        t = (ctypes.c_double * self.n)()
        st = (ctypes.c_int * self.n)()
        _lib.check(self._lib.coffee_sph_batch_status(self._ctx, ctypes.cast(act, ctypes.c_void_p),
                                                     ctypes.cast(t, ctypes.c_void_p),
                                                     ctypes.cast(st, ctypes.c_void_p)), self._ctx)
        return list(act), list(t), list(st)

    def close(self):
        if self._ctx:
            self._lib.coffee_sph_batch_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
