"""data_exporter (src/utils/data_exporter.h:15-100) on one GPU: restart
snapshots and field output in HDF5 files laid out as the reference writes
them.

- save_snapshot / load_snapshot (data_exporter.cpp:134-259): the full
  guarded arrays Ex, Ey, Ez, Bx, By, Bz, B0x, B0y, B0z, P, divE, divB (the
  whole domain is one rank's block with both neighbours NULL, so the
  reference's extent is dims = N + 2 guard on every axis), plus "step"
  (uint32) and "time" (Scalar).  Loading copies the arrays to the device and
  fills the guard cells (send_guard_cells).  The LSRK(5,4) register dU is
  reset by the first stage of every step (A_1 = 0) and the absorbing layer's
  step-start copy is rewritten by it, so E, B, P, step and time are the whole
  state: a resumed run continues bitwise as the uninterrupted one.
- write_output / write_field_output (data_exporter.cpp:120-131,593-643):
  fld.%05d.h5 (number = timestep / data_interval) in the output directory,
  each array sampled every `downsample` cells from the first interior cell
  and interpolated from its stagger to the cell centre, as float32, on the
  GPU (coffee_downsample); the _cum energy-accounting arrays are reset to 0
  after they are written.
"""

from __future__ import annotations

import ctypes
import os
import shutil
from typing import Optional, Tuple

import numpy as np

from . import _lib
from ._h5 import H5File

# Stagger bits of sim_data's fields (src/data/sim_data.cpp:19-25): bit a set
# = cell-centred along axis a.
STAGGER_E = (0b110, 0b101, 0b011)
STAGGER_B = (0b001, 0b010, 0b100)
STAGGER_CENTRE = 0b111

SNAPSHOT_ARRAYS = ("Ex", "Ey", "Ez", "Bx", "By", "Bz", "B0x", "B0y", "B0z", "P", "divE", "divB")


def _torch():
    import torch

    return torch


def _snapshot_members(data):
    """(name, MultiArray) in the reference's snapshot order."""
    return [("Ex", data.E.data(0)), ("Ey", data.E.data(1)), ("Ez", data.E.data(2)),
            ("Bx", data.B.data(0)), ("By", data.B.data(1)), ("Bz", data.B.data(2)),
            ("B0x", data.B0.data(0)), ("B0y", data.B0.data(1)), ("B0z", data.B0.data(2)),
            ("P", data.P), ("divE", data.divE), ("divB", data.divB)]


def _output_members(data):
    """(name, MultiArray or None, stagger) of write_field_output
    (data_exporter.cpp:612-631): B0 holds J."""
    out = [("Ex", data.E.data(0), STAGGER_E[0]), ("Ey", data.E.data(1), STAGGER_E[1]),
           ("Ez", data.E.data(2), STAGGER_E[2]), ("Bx", data.B.data(0), STAGGER_B[0]),
           ("By", data.B.data(1), STAGGER_B[1]), ("Bz", data.B.data(2), STAGGER_B[2]),
           ("Jx", data.B0.data(0), STAGGER_B[0]), ("Jy", data.B0.data(1), STAGGER_B[1]),
           ("Jz", data.B0.data(2), STAGGER_B[2]), ("P", data.P, STAGGER_CENTRE),
           ("divB", data.divB, STAGGER_CENTRE), ("divE", data.divE, STAGGER_CENTRE)]
    for name in ("dU_EgtB", "dU_Epar", "dU_KO", "dU_EgtB_cum", "dU_Epar_cum", "dU_KO_cum"):
        out.append((name, getattr(data, name, None), STAGGER_CENTRE))
    return out


def _file_shape(grid):
    """numpy (C-order) shape of a Fortran-order array with the grid's dims:
    (dims2, dims1, dims0), or (dims1, dims0) for a 2D grid."""
    dims = tuple(int(d) for d in grid.dims)
    if len(dims) == 2 or dims[2] == 1:
        return (dims[1], dims[0])
    return (dims[2], dims[1], dims[0])


def downsample_extent(grid, d: int) -> Tuple[int, int, int]:
    """Sampled extent: N / d where N > d, else N (data_exporter.cpp:70-79)."""
    ext = []
    for a in range(3):
        dims = int(grid.dims[a]) if a < len(grid.dims) else 1
        guard = int(grid.guard[a]) if a < len(grid.guard) else 0
        n = dims - 2 * guard
        ext.append(n // d if n > d else n)
    return tuple(ext)


class DataExporter:
    """data_exporter(sim_environment&, uint32_t& timestep).  `output_dir`
    defaults to the reference's "./Data/"; config.toml in the working
    directory, when present, is copied there (data_exporter.cpp:105-116)."""

    def __init__(self, env, timestep: int = 0, output_dir: str = "./Data/",
                 copy_config: bool = True):
        self.m_env = env
        self.outputDirectory = output_dir if output_dir.endswith("/") else output_dir + "/"
        os.makedirs(self.outputDirectory, exist_ok=True)
        if copy_config and os.path.exists("config.toml"):
            shutil.copyfile("config.toml", os.path.join(self.outputDirectory, "config.toml"))
        self._lib = _lib.load()
        self._tmp = None  # device buffer of the sampled array (tmp_grid_data)

    def output_directory(self) -> str:
        return self.outputDirectory

    # -- restart snapshots --------------------------------------------------

    def save_snapshot(self, filename: str, data, step: int, time: float):
        """data_exporter::save_snapshot (data_exporter.cpp:134-197)."""
        torch = _torch()
        torch.cuda.synchronize()
        shape = _file_shape(self.m_env.grid())
        dtype = self.m_env.dtype
        with H5File(filename, "w") as f:
            for name, arr in _snapshot_members(data):
                f.write(name, arr.sync_to_host().reshape(shape))
            f.write_scalar("step", int(step), np.uint32)
            f.write_scalar("time", float(time), dtype)

    def load_snapshot(self, filename: str, data) -> Optional[Tuple[int, float]]:
        """data_exporter::load_snapshot (data_exporter.cpp:199-259): returns
        (step, time), or None (with the reference's message) when the file
        does not exist."""
        if not os.path.exists(filename):
            print("Can't find restart file, proceeding without loading it!", flush=True)
            return None
        grid = self.m_env.grid()
        shape = _file_shape(grid)
        dtype = self.m_env.dtype
        with H5File(filename, "r") as f:
            for name, arr in _snapshot_members(data):
                arr.set(f.read(name, shape, dtype))
            step = int(f.read_scalar("step", np.uint32))
            time = float(f.read_scalar("time", dtype))
        self.m_env.send_guard_cells(data)
        return step, time

    # -- field output -------------------------------------------------------

    def write_output(self, data, timestep: int, time: float):
        """data_exporter::write_output (data_exporter.cpp:120-131)."""
        self.write_field_output(data, timestep, time)

    def _sample(self, arr, stagger: int, d: int) -> np.ndarray:
        torch = _torch()
        grid = self.m_env.grid()
        ext = downsample_extent(grid, d)
        n = ext[0] * ext[1] * ext[2]
        if self._tmp is None or self._tmp.numel() < n:
            self._tmp = torch.empty(n, dtype=torch.float32, device=self.m_env.device)
        from .solver import coffee_grid

        g = coffee_grid(grid)
        s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        _lib.check(self._lib.coffee_downsample(ctypes.byref(g), ctypes.c_void_p(arr.dev_ptr()),
                                               self.m_env.dtype.itemsize, int(stagger), int(d),
                                               ctypes.c_void_p(self._tmp.data_ptr()), s))
        out = self._tmp[:n].cpu().numpy()
        if ext[2] == 1:
            return out.reshape(ext[1], ext[0])
        return out.reshape(ext[2], ext[1], ext[0])

    def write_field_output(self, data, timestep: int, time: float) -> str:
        """data_exporter::write_field_output (data_exporter.cpp:593-643):
        writes fld.%05d.h5 and returns its path."""
        sp = self.m_env.params()
        num = int(timestep) // int(sp.data_interval)
        path = os.path.join(self.outputDirectory, f"fld.{num:05d}.h5")
        d = max(1, int(sp.downsample))
        with H5File(path, "w") as f:
            for name, arr, st in _output_members(data):
                if arr is None:
                    continue
                f.write(name, self._sample(arr, st, d))
        for name in ("dU_EgtB_cum", "dU_Epar_cum", "dU_KO_cum"):
            arr = getattr(data, name, None)
            if arr is not None:
                arr.assign(0.0)
        return path
