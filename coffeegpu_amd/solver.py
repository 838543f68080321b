"""Host-side mirror of the reference's solver surface for the flat EZ path.

Same names and roles as the C++ classes of the reference:
  sim_environment  (src/sim_env.h:17-87)       -> SimEnvironment
  multi_array<T>   (src/data/multi_array.h)    -> MultiArray
  vector_field<T>  (src/data/fields.h)         -> VectorField
  sim_data         (src/data/sim_data.h)       -> SimData
  field_solver_EZ  (src/algorithms/field_solver_EZ.h:6-33) -> FieldSolverEZ
Device memory is held in torch tensors (plumbing only); every solver call goes
through libcoffee_hip.so.  Host copies exist only after sync_to_host(), as in
the reference.
"""

from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np

from . import _lib
from .params import Grid, SimParams, reference_simd_lanes


def _torch():
    import torch

    return torch


def _dtype_torch(dtype):
    torch = _torch()
    return torch.float64 if np.dtype(dtype) == np.float64 else torch.float32


class SimEnvironment:
    """Single-GPU environment: parameters and the (one-rank) grid.  The MPI
    decomposition of the reference collapses to one domain; the guard exchange
    is the periodic fill inside the solver."""

    def __init__(self, params: SimParams, dtype=np.float64, device: str = "cuda:0"):
        self._params = params
        self.dtype = np.dtype(dtype)
        self._grid = Grid.from_params(params, dtype)
        self.device = device

    def params(self) -> SimParams:
        return self._params

    def grid(self) -> Grid:
        return self._grid

    def send_guard_cells(self, data: "SimData", stream=None):
        """sim_environment::send_guard_cells(sim_data&) on one rank
        (src/sim_env_impl.hpp:179-195, src/sim_env.cpp:365-438): periodic
        axes take the interior one period away (x, then y, then z) in E, B
        and P; other axes keep their guard cells."""
        torch = _torch()
        g = coffee_grid(self._grid)
        per = (ctypes.c_int * 3)(*[int(bool(v)) for v in self._params.periodic_boundary])
        arrs = (ctypes.c_void_p * 7)(*[a.dev_ptr() for a in data.state_arrays()])
        s = stream if stream is not None else torch.cuda.current_stream()
        lib = _lib.load()
        _lib.check(lib.coffee_fill_guards_arrays(ctypes.byref(g), per, arrs, 7,
                                                 self.dtype.itemsize, ctypes.c_void_p(s.cuda_stream)))

    def rank(self) -> int:
        return 0

    def size(self) -> int:
        return 1


class MultiArray:
    """Fortran-order 3D array (i fastest) with a device buffer and an optional
    host mirror."""

    def __init__(self, grid: Grid, dtype, device: str):
        torch = _torch()
        self.grid = grid
        self.dtype = np.dtype(dtype)
        self.dev = torch.zeros(grid.shape, dtype=_dtype_torch(dtype), device=device)
        self.host: Optional[np.ndarray] = None
        self._pinned = None  # page-locked host mirror, allocated on first sync_to_host

    def dev_ptr(self) -> int:
        return self.dev.data_ptr()

    def assign(self, value: float):
        self.dev.fill_(value)

    def copy_from(self, other: "MultiArray"):
        self.dev.copy_(other.dev)

    def sync_to_host(self) -> np.ndarray:
        """Copies the device buffer into the host mirror and returns a numpy
        VIEW of it (no copy).  The mirror is one page-locked buffer kept for
        the array's lifetime (as multi_array keeps its host copy): the next
        sync_to_host of this array overwrites the data every earlier returned
        view shows, so a caller that keeps snapshots must `.copy()` them.
        Page-locked copies run at the link rate (≈57 GB/s on MI355X against
        ≈3 GB/s into freshly allocated pageable memory)."""
        torch = _torch()
        if self._pinned is None:
            self._pinned = torch.empty(self.dev.shape, dtype=self.dev.dtype,
                                       pin_memory=self.dev.is_cuda)
        self._pinned.copy_(self.dev)
        self.host = self._pinned.numpy()
        return self.host

    def sync_to_device(self):
        torch = _torch()
        self.dev.copy_(torch.from_numpy(np.ascontiguousarray(self.host, dtype=self.dtype)))

    def set(self, arr: np.ndarray):
        self.host = np.ascontiguousarray(arr, dtype=self.dtype).reshape(self.grid.shape)
        self.sync_to_device()


class VectorField:
    def __init__(self, grid: Grid, dtype, device: str):
        self.grid = grid
        self._data = [MultiArray(grid, dtype, device) for _ in range(3)]

    def data(self, n: int) -> MultiArray:
        return self._data[n]

    def dev_ptr(self, n: int) -> int:
        return self._data[n].dev_ptr()

    def initialize(self):
        for a in self._data:
            a.assign(0.0)

    def assign(self, value: float):
        for a in self._data:
            a.assign(value)

    def copy_from(self, other: "VectorField"):
        for a, b in zip(self._data, other._data):
            a.copy_from(b)

    def sync_to_host(self):
        return [a.sync_to_host() for a in self._data]

    def sync_to_device(self):
        for a in self._data:
            a.sync_to_device()


class SimData:
    """sim_data (src/data/sim_data.h:11-24): E, B, B0 (holds J), P, divB, divE."""

    def __init__(self, env: SimEnvironment):
        g, dt, dev = env.grid(), env.dtype, env.device
        self.env = env
        self.E = VectorField(g, dt, dev)
        self.B = VectorField(g, dt, dev)
        self.B0 = VectorField(g, dt, dev)
        self.P = MultiArray(g, dt, dev)
        self.divB = MultiArray(g, dt, dev)
        self.divE = MultiArray(g, dt, dev)
        # energy accounting (sim_data.h:15-16): allocated by
        # alloc_energy_accounting(), i.e. when a solver turns diagnostics on
        self.dU_KO = self.dU_Epar = self.dU_EgtB = None
        self.dU_KO_cum = self.dU_Epar_cum = self.dU_EgtB_cum = None

    ACCOUNTING = ("dU_KO", "dU_Epar", "dU_EgtB", "dU_KO_cum", "dU_Epar_cum", "dU_EgtB_cum")

    def alloc_energy_accounting(self):
        g, dt, dev = self.env.grid(), self.env.dtype, self.env.device
        for name in self.ACCOUNTING:
            if getattr(self, name) is None:
                setattr(self, name, MultiArray(g, dt, dev))

    def state_arrays(self):
        return [self.E.data(0), self.E.data(1), self.E.data(2), self.B.data(0), self.B.data(1),
                self.B.data(2), self.P]

    def set_state(self, U: np.ndarray):
        """Set E, B, P from a (7, k, j, i) array."""
        for a, u in zip(self.state_arrays(), U):
            a.set(u)

    def get_state(self) -> np.ndarray:
        return np.stack([a.sync_to_host() for a in self.state_arrays()])

    def sync_to_host(self):
        for a in self.state_arrays():
            a.sync_to_host()

    def sync_to_device(self):
        for a in self.state_arrays():
            a.sync_to_device()


def coffee_grid(grid: Grid) -> _lib.CoffeeGrid:
    g = _lib.CoffeeGrid()
    for i in range(3):
        g.dims[i] = grid.dims[i]
        g.guard[i] = grid.guard[i]
        g.delta[i] = grid.delta[i]
        g.lower[i] = grid.lower[i]
        g.sizes[i] = grid.sizes[i]
    return g


def coffee_params(sp: SimParams, mask_lanes: int = 1) -> _lib.CoffeeParams:
    p = _lib.CoffeeParams()
    for i in range(3):
        p.periodic[i] = int(bool(sp.periodic_boundary[i]))
        p.pml[i] = int(sp.pml[i])
        p.lower[i] = float(sp.lower[i])
        p.size[i] = float(sp.size[i])
    for name, ctype in _lib.CoffeeParams._fields_:
        if name in ("periodic", "pml", "lower", "size", "mask_lanes"):
            continue
        v = getattr(sp, name)
        setattr(p, name, int(v) if ctype is ctypes.c_int else float(v))
    p.mask_lanes = int(mask_lanes)
    return p


class FieldSolverEZ:
    """field_solver_EZ (src/algorithms/field_solver_EZ.h:6-33) on one GPU.

    Public scratch members keep the reference names: Etmp, Btmp, dE, dB, Bbg,
    dP, Ptmp.  `mask_lanes` defaults to the reference CPU build's SIMD width
    for this Scalar (parity with the normative CPU semantics)."""

    _energy_accounting = True  # the flat solver fills sim_data's dU_* arrays

    def __init__(self, data: SimData, env: SimEnvironment, mask_lanes: Optional[int] = None,
                 stream=None):
        torch = _torch()
        self.m_data = data
        self.m_env = env
        g, dt, dev = env.grid(), env.dtype, env.device
        self.dE = VectorField(g, dt, dev)
        self.dB = VectorField(g, dt, dev)
        self.Etmp = VectorField(g, dt, dev)
        self.Btmp = VectorField(g, dt, dev)
        self.Bbg = VectorField(g, dt, dev)
        self.dP = MultiArray(g, dt, dev)
        self.Ptmp = MultiArray(g, dt, dev)
        sp = env.params()
        if any(int(v) > 0 for v in sp.pml):
            from .initial_data import multipole_field

            for n, arr in enumerate(multipole_field(g, sp, 0.0, dt)):
                self.Bbg.data(n).set(arr)
        self.mask_lanes = reference_simd_lanes(dt) if mask_lanes is None else int(mask_lanes)
        self._lib = _lib.load()
        self._ctx = ctypes.c_void_p()
        gg = coffee_grid(g)
        pp = coffee_params(sp, self.mask_lanes)
        _lib.check(self._lib.coffee_ctx_create(_lib.COFFEE_SOLVER_EZ, dt.itemsize, ctypes.byref(gg),
                                               ctypes.byref(pp), ctypes.byref(self._ctx)))
        self._stream = stream
        self._fields = self._make_fields()
        self._diag = False

    def _stream_ptr(self):
        torch = _torch()
        s = self._stream if self._stream is not None else torch.cuda.current_stream()
        return ctypes.c_void_p(s.cuda_stream)

    def _make_fields(self) -> _lib.CoffeeFields:
        f = _lib.CoffeeFields()
        d = self.m_data
        for i in range(3):
            f.E[i] = d.E.dev_ptr(i)
            f.B[i] = d.B.dev_ptr(i)
            f.dE[i] = self.dE.dev_ptr(i)
            f.dB[i] = self.dB.dev_ptr(i)
            f.Etmp[i] = self.Etmp.dev_ptr(i)
            f.Btmp[i] = self.Btmp.dev_ptr(i)
            f.Bbg[i] = self.Bbg.dev_ptr(i)
            f.J[i] = d.B0.dev_ptr(i)
        f.P = d.P.dev_ptr()
        f.dP = self.dP.dev_ptr()
        f.Ptmp = self.Ptmp.dev_ptr()
        f.divE = d.divE.dev_ptr()
        f.divB = d.divB.dev_ptr()
        for name in SimData.ACCOUNTING:
            arr = getattr(d, name, None)
            setattr(f, name, arr.dev_ptr() if arr is not None else None)
        return f

    def set_diagnostics(self, on: bool):
        """Diagnostics mode: J, divE, divB every stage and (flat solver) the
        energy accounting arrays of sim_data, allocated on first use."""
        self._diag = bool(on)
        if on and self._energy_accounting and hasattr(self.m_data, "alloc_energy_accounting"):
            self.m_data.alloc_energy_accounting()
            self._fields = self._make_fields()
        _lib.check(self._lib.coffee_set_diagnostics(self._ctx, int(on)), self._ctx)

    def evolve_fields(self, time: float):
        """One time step (field_solver_EZ::evolve_fields)."""
        _lib.check(self._lib.coffee_ez_step(self._ctx, ctypes.byref(self._fields),
                                            ctypes.c_double(time), self._stream_ptr()), self._ctx)

    def fill_guards(self):
        _lib.check(self._lib.coffee_fill_guards(self._ctx, ctypes.byref(self._fields),
                                                self._stream_ptr()), self._ctx)

    def total_energy(self, f: VectorField) -> float:
        out = ctypes.c_double()
        comp = (ctypes.c_void_p * 3)(f.dev_ptr(0), f.dev_ptr(1), f.dev_ptr(2))
        _lib.check(self._lib.coffee_total_energy(self._ctx, comp, ctypes.byref(out),
                                                 self._stream_ptr()), self._ctx)
        return out.value

    def residuals(self):
        out = (ctypes.c_double * 5)()
        _lib.check(self._lib.coffee_residuals(self._ctx, ctypes.byref(self._fields), out,
                                              self._stream_ptr()), self._ctx)
        return {"energy": out[0], "max_divB": out[1], "max_EdotB": out[2], "max_E2mB2": out[3],
                "nonfinite": out[4]}

    def set_schedule(self, variant: int = 0, zchunk: int = 0):
        """0 = LDS-tiled z-marching kernels (default tile), 1 = one thread per
        cell, 2/3/4 = 64-wide tiles of 4/8/16 rows, 5/6 = 128/256-wide tiles,
        7 = 128 x 4.
        Every schedule gives bitwise the same results."""
        _lib.check(self._lib.coffee_set_schedule(self._ctx, int(variant), int(zchunk)), self._ctx)

    def set_timing(self, on: bool):
        _lib.check(self._lib.coffee_set_timing(self._ctx, int(on)), self._ctx)

    def reset_timing(self):
        _lib.check(self._lib.coffee_reset_timing(self._ctx), self._ctx)

    def get_timing(self):
        n = ctypes.c_int()
        names = ((ctypes.c_char * 64) * 32)()
        ms = (ctypes.c_double * 32)()
        cnt = (ctypes.c_int * 32)()
        _lib.check(self._lib.coffee_get_timing(self._ctx, 32, names, ms, cnt, ctypes.byref(n)),
                   self._ctx)
        return {names[i].value.decode(): (cnt[i], ms[i]) for i in range(n.value)}

    def close(self):
        if self._ctx:
            self._lib.coffee_ctx_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
