"""field_solver_gr_EZ (src/algorithms/field_solver_gr_EZ.h:6-32) on one GPU.

Cartesian Kerr-Schild black hole (M = 1, spin `a` from the parameters).  The
sim_data "E" arrays hold D (upper components), "B" holds B, as in the
reference.  Public scratch members keep the reference names: Dtmp, Btmp, dD,
dB, Ed, Hd, dP, Ptmp (Ed, Hd are recomputed on the device every stage and
are not kept in host-visible arrays: their members exist for interface
parity and stay zero).
"""

from __future__ import annotations

import ctypes

from . import _lib
from .solver import FieldSolverEZ, MultiArray, SimData, SimEnvironment, VectorField, coffee_grid, \
    coffee_params


class FieldSolverGREZ(FieldSolverEZ):
    _energy_accounting = False

    def __init__(self, data: SimData, env: SimEnvironment, stream=None):
        self.m_data = data
        self.m_env = env
        g, dt, dev = env.grid(), env.dtype, env.device
        self.dD = VectorField(g, dt, dev)
        self.dB = VectorField(g, dt, dev)
        self.Dtmp = VectorField(g, dt, dev)
        self.Btmp = VectorField(g, dt, dev)
        self.Ed = VectorField(g, dt, dev)
        self.Hd = VectorField(g, dt, dev)
        self.dP = MultiArray(g, dt, dev)
        self.Ptmp = MultiArray(g, dt, dev)
        # aliases used by the shared field table
        self.dE, self.Etmp = self.dD, self.Dtmp
        self.Bbg = self.Btmp
        self.mask_lanes = 1
        self._lib = _lib.load()
        self._ctx = ctypes.c_void_p()
        gg = coffee_grid(g)
        pp = coffee_params(env.params(), 1)
        _lib.check(self._lib.coffee_ctx_create(_lib.COFFEE_SOLVER_GR_EZ, dt.itemsize,
                                               ctypes.byref(gg), ctypes.byref(pp),
                                               ctypes.byref(self._ctx)))
        self._stream = stream
        self._fields = self._make_fields()
        self._diag = False

    def set_schedule(self, variant: int = 0, zchunk: int = 0):
        """0 = LDS-tiled stage kernel computing E, H on the fly, the halo
        points' metric of stages 2-5 read from a table filled once (default);
        1 = point sweep (E, H arrays) + stage sweep; 2 = tiled with every
        metric point computed on the fly.  All give bitwise the same results."""
        _lib.check(self._lib.coffee_set_schedule(self._ctx, int(variant), int(zchunk)), self._ctx)
