"""ctypes access to the CPU oracle and the reference-build harness.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product path never imports
this module.
"""

from __future__ import annotations

import ctypes
import os
import shutil
import subprocess
import tempfile
from typing import Dict, List, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")


class FFEParams(ctypes.Structure):
    _fields_ = [
        ("N", ctypes.c_int * 3), ("guard", ctypes.c_int * 3),
        ("lower", ctypes.c_double * 3), ("size", ctypes.c_double * 3),
        ("periodic", ctypes.c_int * 3), ("dt", ctypes.c_double),
        ("pml", ctypes.c_int * 3), ("pmllen", ctypes.c_int), ("sigpml", ctypes.c_double),
        ("divB_clean", ctypes.c_int), ("ch2", ctypes.c_double), ("tau", ctypes.c_double),
        ("KOeps", ctypes.c_double), ("KO_geometry", ctypes.c_int),
        ("use_edotb_damping", ctypes.c_int), ("damp_gamma", ctypes.c_double),
        ("calc_current", ctypes.c_int), ("clean_ep", ctypes.c_int), ("check_egb", ctypes.c_int),
        ("problem", ctypes.c_int),
        ("radius", ctypes.c_double), ("omega", ctypes.c_double), ("b0", ctypes.c_double),
        ("p1", ctypes.c_double), ("p2", ctypes.c_double), ("p3", ctypes.c_double),
        ("q11", ctypes.c_double), ("q12", ctypes.c_double), ("q13", ctypes.c_double),
        ("q22", ctypes.c_double), ("q23", ctypes.c_double),
        ("q_offset_x", ctypes.c_double), ("q_offset_y", ctypes.c_double),
        ("q_offset_z", ctypes.c_double),
        ("tp_start", ctypes.c_double), ("tp_end", ctypes.c_double), ("rpert1", ctypes.c_double),
        ("rpert2", ctypes.c_double), ("dw0", ctypes.c_double), ("nT", ctypes.c_double),
        ("ph1", ctypes.c_double), ("ph2", ctypes.c_double), ("dph", ctypes.c_double),
        ("theta0", ctypes.c_double), ("drpert", ctypes.c_double),
        ("tp_start1", ctypes.c_double), ("tp_end1", ctypes.c_double),
        ("rpert11", ctypes.c_double), ("rpert21", ctypes.c_double), ("dw1", ctypes.c_double),
        ("nT1", ctypes.c_double), ("ph11", ctypes.c_double), ("ph21", ctypes.c_double),
        ("dph1", ctypes.c_double),
        ("pert_type", ctypes.c_int), ("a", ctypes.c_double),
        ("thp1", ctypes.c_double), ("thp2", ctypes.c_double), ("thp11", ctypes.c_double),
        ("thp21", ctypes.c_double), ("mask_lanes", ctypes.c_int),
    ]


def make_params(sp, mask_lanes: int = 1) -> FFEParams:
    """FFEParams from a coffeegpu_amd.params.SimParams."""
    p = FFEParams()
    for i in range(3):
        p.N[i] = int(sp.N[i])
        p.guard[i] = int(sp.guard[i])
        p.lower[i] = float(sp.lower[i])
        p.size[i] = float(sp.size[i])
        p.periodic[i] = int(bool(sp.periodic_boundary[i]))
        p.pml[i] = int(sp.pml[i])
    for name, _ in FFEParams._fields_:
        if name in ("N", "guard", "lower", "size", "periodic", "pml", "mask_lanes"):
            continue
        v = getattr(sp, name)
        setattr(p, name, int(v) if isinstance(v, (bool, int)) and not isinstance(v, float) else v)
    p.mask_lanes = int(mask_lanes)
    return p


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library not built: {LIB_PATH} (run make -C oracle)")
        _lib = ctypes.CDLL(LIB_PATH)
    return _lib


def _ptrs(arrs: Sequence[np.ndarray], ctype):
    for a in arrs:
        assert a.flags["C_CONTIGUOUS"]
    P = ctypes.POINTER(ctype)
    return (P * len(arrs))(*[a.ctypes.data_as(P) for a in arrs])


class FlatState:
    """Host state of the flat solver: U, dU, Un (7 each), Bbg (3)."""

    def __init__(self, U: np.ndarray, dtype, bbg: Optional[np.ndarray] = None):
        self.dtype = np.dtype(dtype)
        self.U = np.ascontiguousarray(U, dtype=self.dtype).copy()
        self.dU = np.zeros_like(self.U)
        self.Un = np.zeros_like(self.U)
        self.Bbg = None if bbg is None else np.ascontiguousarray(bbg, dtype=self.dtype).copy()


def flat_evolve(sp, state: FlatState, nsteps: int, time: float = 0.0, mask_lanes: int = 1,
                diag: Optional[Dict[str, np.ndarray]] = None):
    L = lib()
    p = make_params(sp, mask_lanes)
    ct = ctypes.c_double if state.dtype == np.float64 else ctypes.c_float
    U = _ptrs(list(state.U), ct)
    dU = _ptrs(list(state.dU), ct)
    Un = _ptrs(list(state.Un), ct)
    bbg = _ptrs(list(state.Bbg), ct) if state.Bbg is not None else None
    fn = L.ffe_flat_evolve_f64 if state.dtype == np.float64 else L.ffe_flat_evolve_f32
    fn.restype = ctypes.c_int
    dg = None
    if diag is not None:
        dg = _make_diag(diag, ct, state.dtype)
    rc = fn(ctypes.byref(p), U, dU, Un, bbg, ctypes.byref(dg) if dg is not None else None,
            ctypes.c_double(time), ctypes.c_int(nsteps))
    if rc != 0:
        raise RuntimeError(f"oracle flat_evolve failed rc={rc}")
    return state


def _make_diag(d, ct, dtype):
    P = ctypes.POINTER(ct)

    class Diag(ctypes.Structure):
        _fields_ = [("J", P * 3), ("divE", P), ("divB", P), ("dU_KO", P), ("dU_Epar", P),
                    ("dU_EgtB", P), ("dU_KO_cum", P), ("dU_Epar_cum", P), ("dU_EgtB_cum", P)]

    g = Diag()
    if "J" in d:
        for i in range(3):
            g.J[i] = d["J"][i].ctypes.data_as(P)
    for k in ("divE", "divB", "dU_KO", "dU_Epar", "dU_EgtB", "dU_KO_cum", "dU_Epar_cum",
              "dU_EgtB_cum"):
        if k in d:
            setattr(g, k, d[k].ctypes.data_as(P))
    return g


def flat_bbg(sp, dtype) -> np.ndarray:
    from coffeegpu_amd.params import Grid

    g = Grid.from_params(sp, dtype)
    out = np.zeros((3,) + g.shape, dtype=dtype)
    ct = ctypes.c_double if np.dtype(dtype) == np.float64 else ctypes.c_float
    fn = lib().ffe_flat_bbg_f64 if np.dtype(dtype) == np.float64 else lib().ffe_flat_bbg_f32
    fn(ctypes.byref(make_params(sp)), _ptrs(list(out), ct))
    return out


# ---- CFLD files and the reference-build harness ------------------------------

def write_cfld(path: str, U: np.ndarray):
    U = np.ascontiguousarray(U)
    nf, nk, nj, ni = U.shape
    with open(path, "wb") as f:
        f.write(b"CFLD")
        np.array([1, U.dtype.itemsize, ni, nj, nk, nf], dtype="<i4").tofile(f)
        U.astype(U.dtype.newbyteorder("<")).tofile(f)


def read_cfld(path: str) -> np.ndarray:
    with open(path, "rb") as f:
        assert f.read(4) == b"CFLD"
        hdr = np.fromfile(f, dtype="<i4", count=6)
        _, isz, ni, nj, nk, nf = [int(v) for v in hdr]
        dt = np.dtype("<f8") if isz == 8 else np.dtype("<f4")
        data = np.fromfile(f, dtype=dt, count=nf * nk * nj * ni)
    return data.reshape(nf, nk, nj, ni).astype(dt.newbyteorder("="))


def reference_available(dtype=np.float64) -> bool:
    name = "ref_ez_f64" if np.dtype(dtype) == np.float64 else "ref_ez_f32"
    return os.path.exists(os.path.join(REF_DIR, name))


MPIEXEC = "/opt/conda/bin/mpiexec"


def mpiexec_available() -> bool:
    return os.path.exists(MPIEXEC)


def run_reference(sp, U0: np.ndarray, nsteps: int, dumps: Sequence[int] = (),
                  norms: bool = False, timing_warmup: int = -1, env=None, ranks: int = 1):
    """Run the reference CPU build (oracle/_ref) from U0 for nsteps.

    ranks > 1 runs it under mpiexec with the z axis split over the ranks
    (nodes = [1, 1, ranks]; timing only: no dumps or norms).
    Returns (dict step -> (7,k,j,i) array, norms array or None, stderr text)."""
    dtype = U0.dtype
    exe = os.path.join(REF_DIR, "ref_ez_f64" if dtype == np.float64 else "ref_ez_f32")
    if not os.path.exists(exe):
        raise RuntimeError(f"reference harness not built: {exe}")
    tmp = tempfile.mkdtemp(prefix="coffee_ref_")
    try:
        if ranks > 1:
            sp = sp.replace(nodes=[1, 1, int(ranks)])
        with open(os.path.join(tmp, "config.toml"), "w") as f:
            f.write(sp.to_toml())
        write_cfld(os.path.join(tmp, "in.cfld"), U0)
        args = [exe, "in.cfld", "out", str(nsteps),
                ",".join(str(s) for s in dumps) if dumps else "-",
                "norms.txt" if norms else "-", str(timing_warmup)]
        if ranks > 1:
            args = [MPIEXEC, "-launcher", "fork", "-n", str(int(ranks))] + args
        r = subprocess.run(args, cwd=tmp, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE,
                           text=True, env=env)
        if r.returncode != 0:
            raise RuntimeError(f"reference run failed ({r.returncode}): {r.stderr[-2000:]}")
        out = {s: read_cfld(os.path.join(tmp, f"out_s{s}.cfld")) for s in dumps}
        # energy bookkeeping (dU_KO, dU_Epar, dU_EgtB, then the _cum arrays)
        for s in dumps:
            out[("dg", s)] = read_cfld(os.path.join(tmp, f"out_dg_s{s}.cfld"))
        nrm = np.loadtxt(os.path.join(tmp, "norms.txt"), ndmin=2) if norms else None
        return out, nrm, r.stderr
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def rel_l2(a: np.ndarray, b: np.ndarray) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.sqrt(np.sum(b * b))
    return float(np.sqrt(np.sum((a - b) ** 2)) / (den if den > 0 else 1.0))


# ---- log-spherical oracle ------------------------------------------------------

class SphState:
    """Host state of the spherical solver: U, dU, Un (7 each), L (6)."""

    def __init__(self, U: np.ndarray, dtype):
        self.dtype = np.dtype(dtype)
        self.U = np.ascontiguousarray(U, dtype=self.dtype).copy()
        self.dU = np.zeros_like(self.U)
        self.Un = np.zeros_like(self.U)
        self.L = np.zeros((6,) + self.U.shape[1:], dtype=self.dtype)


def sph_evolve(sp, state: SphState, nsteps: int, time: float = 0.0):
    L = lib()
    p = make_params(sp)
    ct = ctypes.c_double if state.dtype == np.float64 else ctypes.c_float
    fn = L.ffe_sph_evolve_f64 if state.dtype == np.float64 else L.ffe_sph_evolve_f32
    fn.restype = ctypes.c_int
    rc = fn(ctypes.byref(p), _ptrs(list(state.U), ct), _ptrs(list(state.dU), ct),
            _ptrs(list(state.Un), ct), _ptrs(list(state.L), ct), ctypes.c_double(time),
            ctypes.c_int(nsteps))
    if rc != 0:
        raise RuntimeError(f"oracle sph_evolve failed rc={rc}")
    return state


def sph_energy(sp, F: np.ndarray) -> float:
    L = lib()
    ct = ctypes.c_double if F.dtype == np.float64 else ctypes.c_float
    fn = L.ffe_sph_energy_f64 if F.dtype == np.float64 else L.ffe_sph_energy_f32
    fn.restype = ctypes.c_double
    return fn(ctypes.byref(make_params(sp)), _ptrs(list(np.ascontiguousarray(F)), ct))


class GrState:
    """Host state of the Kerr-Schild solver (FP64): U = D, B, P; dU; Un; EH."""

    def __init__(self, U: np.ndarray):
        self.U = np.ascontiguousarray(U, dtype=np.float64).copy()
        self.dU = np.zeros_like(self.U)
        self.Un = np.zeros_like(self.U)
        self.EH = np.zeros((6,) + self.U.shape[1:], dtype=np.float64)


def _gr_diag(diag):
    if diag is None:
        return None
    return ctypes.byref(_make_diag(diag, ctypes.c_double, np.float64))


def gr_evolve(sp, state: GrState, nsteps: int, diag=None):
    L = lib()
    p = make_params(sp)
    ct = ctypes.c_double
    fn = L.ffe_gr_evolve_f64
    fn.restype = ctypes.c_int
    rc = fn(ctypes.byref(p), _ptrs(list(state.U), ct), _ptrs(list(state.dU), ct),
            _ptrs(list(state.Un), ct), _ptrs(list(state.EH), ct), _gr_diag(diag),
            ctypes.c_double(0.0), ctypes.c_int(nsteps))
    if rc != 0:
        raise RuntimeError(f"oracle gr_evolve failed rc={rc}")
    return state


def gr_rk_stage(sp, state: GrState, As: float, Bs: float, diag=None):
    L = lib()
    ct = ctypes.c_double
    fn = L.ffe_gr_rk_stage_f64
    fn.restype = ctypes.c_int
    rc = fn(ctypes.byref(make_params(sp)), _ptrs(list(state.U), ct), _ptrs(list(state.dU), ct),
            _ptrs(list(state.EH), ct), _gr_diag(diag), ctypes.c_double(As), ctypes.c_double(Bs))
    if rc != 0:
        raise RuntimeError(f"oracle gr_rk_stage failed rc={rc}")
    return state


def gr_eh(sp, U: np.ndarray) -> np.ndarray:
    L = lib()
    ct = ctypes.c_double
    U = np.ascontiguousarray(U, dtype=np.float64)
    EH = np.zeros((6,) + U.shape[1:], dtype=np.float64)
    L.ffe_gr_eh_f64.restype = ctypes.c_int
    L.ffe_gr_eh_f64(ctypes.byref(make_params(sp)), _ptrs(list(U), ct), _ptrs(list(EH), ct))
    return EH


GR_METRIC_NAMES = ("gd11", "gd12", "gd13", "gd22", "gd23", "gd33", "gu11", "gu12", "gu13",
                   "gu22", "gu23", "gu33", "bd1", "bd2", "bd3", "bu1", "bu2", "bu3", "alpha",
                   "sg", "r")


def gr_metric(a: float, x: float, y: float, z: float) -> dict:
    L = lib()
    out = (ctypes.c_double * 21)()
    L.ffe_gr_metric_f64.restype = None
    L.ffe_gr_metric_f64(ctypes.c_double(a), ctypes.c_double(x), ctypes.c_double(y),
                        ctypes.c_double(z), out)
    return dict(zip(GR_METRIC_NAMES, list(out)))


# ---- building blocks and planted mistakes (pinning tests) ------------------------

PLANTS = {
    "J_SIGN": 1, "J_NO_SQRTG": 2, "GRADP_LOWER": 3, "TAU_DT": 4, "SIGMA_SQUARED": 5,
    "PROJ_FLAT": 6, "ROT_NO_SQRTG": 7, "DIV_CENTRE": 8, "STAR_SIGN": 9, "CLEAN_SHAPE": 10,
    "EXCISION_NODT": 11,
}


class planted:
    """Context manager switching one deliberate mistake (PLANTS) into the
    spherical / Kerr-Schild restatement, for mutation checks of the tests."""

    def __init__(self, name: Optional[str]):
        self.code = PLANTS[name] if name else 0

    def __enter__(self):
        lib().ffe_set_plant(ctypes.c_int(self.code))
        return self

    def __exit__(self, *exc):
        lib().ffe_set_plant(ctypes.c_int(0))
        return False


def _f64_ptrs(U):
    return _ptrs([np.ascontiguousarray(u) for u in U], ctypes.c_double)


def sph_stage_rhs(sp, U: np.ndarray) -> np.ndarray:
    """dU/dt of one stage (As = 0, Bs = 0: U is left unchanged), FP64."""
    U = np.ascontiguousarray(U, dtype=np.float64).copy()
    dU = np.zeros_like(U)
    Lw = np.zeros((6,) + U.shape[1:])
    rc = lib().ffe_sph_rk_stage_f64(ctypes.byref(make_params(sp)), _ptrs(list(U), ctypes.c_double),
                                    _ptrs(list(dU), ctypes.c_double),
                                    _ptrs(list(Lw), ctypes.c_double), ctypes.c_double(0.0),
                                    ctypes.c_double(0.0))
    assert rc == 0
    return dU / sp.dt


def sph_project(sp, U: np.ndarray) -> np.ndarray:
    U = np.ascontiguousarray(U, dtype=np.float64).copy()
    lib().ffe_sph_project_f64(ctypes.byref(make_params(sp)), _ptrs(list(U), ctypes.c_double))
    return U


def sph_boundary(sp, U: np.ndarray, t: float) -> np.ndarray:
    U = np.ascontiguousarray(U, dtype=np.float64).copy()
    lib().ffe_sph_boundary_f64(ctypes.byref(make_params(sp)), _ptrs(list(U), ctypes.c_double),
                               ctypes.c_double(t))
    return U


def sph_absorb(sp, U: np.ndarray, Un: np.ndarray) -> np.ndarray:
    U = np.ascontiguousarray(U, dtype=np.float64).copy()
    Un = np.ascontiguousarray(Un, dtype=np.float64).copy()
    lib().ffe_sph_absorb_f64(ctypes.byref(make_params(sp)), _ptrs(list(U), ctypes.c_double),
                             _ptrs(list(Un), ctypes.c_double))
    return U


def gr_stage_rhs(sp, U: np.ndarray) -> np.ndarray:
    st = GrState(U)
    gr_rk_stage(sp, st, 0.0, 0.0)
    return st.dU / sp.dt


def gr_project(sp, U: np.ndarray) -> np.ndarray:
    U = np.ascontiguousarray(U, dtype=np.float64).copy()
    lib().ffe_gr_project_f64(ctypes.byref(make_params(sp)), _ptrs(list(U), ctypes.c_double))
    return U


def gr_ko(sp, U: np.ndarray) -> np.ndarray:
    U = np.ascontiguousarray(U, dtype=np.float64).copy()
    tmp = np.zeros_like(U)
    lib().ffe_gr_ko_f64(ctypes.byref(make_params(sp)), _ptrs(list(U), ctypes.c_double),
                        _ptrs(list(tmp), ctypes.c_double))
    return U


def sph_ko(sp, U: np.ndarray) -> np.ndarray:
    U = np.ascontiguousarray(U, dtype=np.float64).copy()
    tmp = np.zeros_like(U)
    lib().ffe_sph_ko_f64(ctypes.byref(make_params(sp)), _ptrs(list(U), ctypes.c_double),
                         _ptrs(list(tmp), ctypes.c_double))
    return U


def gr_absorb(sp, U: np.ndarray, Un: np.ndarray) -> np.ndarray:
    U = np.ascontiguousarray(U, dtype=np.float64).copy()
    Un = np.ascontiguousarray(Un, dtype=np.float64).copy()
    lib().ffe_gr_absorb_f64(ctypes.byref(make_params(sp)), _ptrs(list(U), ctypes.c_double),
                            _ptrs(list(Un), ctypes.c_double))
    return U


# ---- field-output sampling -------------------------------------------------

def downsample_extent(dims, guard, d: int):
    """Extent of the sampled output: N / d where N > d, else N
    (data_exporter.cpp:70-79)."""
    out = []
    for n, g in zip(dims, guard):
        m = n - 2 * g
        out.append(m // d if m > d else m)
    return tuple(out)


def downsample(arr: np.ndarray, guard, stagger: int, d: int) -> np.ndarray:
    """Point sample every d-th interior cell, interpolated from the array's
    stagger to the cell centre (Stagger 0b111), in the array's Scalar type,
    returned as float32 of the sampled extent (Fortran order flattened as
    arr is).  3D arrays: multi_array::downsample (src/data/multi_array.cpp:
    56-73) with interpolate (src/algorithms/interpolation.h:9-34); 2D arrays
    (dims[2] == 1): the device path's downsample2d with interpolate2d
    (src/cuda/data/multi_array.cu:73-92, interpolation.h:36-50).

    `arr` has shape (dims[2], dims[1], dims[0]) (C order of a Fortran-order
    array).  With out = 0b111 the minus offset of an axis is -1 where the
    stagger bit is 0 and 0 where it is 1; the plus offset is always 0.

    An axis with 1 < N <= d keeps its N cells, sampled with stride 1: the
    reference keeps the extent N there but still steps by d, reading beyond
    the array (multi_array.cpp:62-66)."""
    T = arr.dtype.type
    dz, dy, dx = arr.shape
    half = T(0.5)
    ex = downsample_extent((dx, dy, dz), guard, d)
    sd = [d if n - 2 * g > d else 1 for n, g in zip((dx, dy, dz), guard)]
    i = np.arange(ex[0]) * sd[0] + guard[0]
    j = np.arange(ex[1]) * sd[1] + guard[1]
    mi = -1 if not (stagger & 1) else 0
    mj = -1 if not (stagger & 2) else 0
    if dz == 1:
        kk = np.zeros(1, dtype=int)
        J, I = np.meshgrid(j, i, indexing="ij")
        p = arr[0]
        f1 = half * (p[J, I] + p[J + mj, I])
        f0 = half * (p[J, I + mi] + p[J + mj, I + mi])
        return (half * (f1 + f0)).astype(np.float32)[None]
    k = np.arange(ex[2]) * sd[2] + guard[2]
    mk = -1 if not (stagger & 4) else 0
    K, J, I = np.meshgrid(k, j, i, indexing="ij")
    f11 = half * (arr[K + mk, J, I] + arr[K, J, I])
    f10 = half * (arr[K + mk, J + mj, I] + arr[K, J + mj, I])
    f01 = half * (arr[K + mk, J, I + mi] + arr[K, J, I + mi])
    f00 = half * (arr[K + mk, J + mj, I + mi] + arr[K, J + mj, I + mi])
    f1 = half * (f11 + f10)
    f0 = half * (f01 + f00)
    return (half * (f1 + f0)).astype(np.float32)


def reference_downsample(arr: np.ndarray, guard, stagger: int, d: int) -> np.ndarray:
    """The reference build's own multi_array::downsample (3D arrays), run in
    oracle/_ref/ref_downsample_f32|f64 on `arr` (shape (dz, dy, dx))."""
    exe = os.path.join(REF_DIR, "ref_downsample_f64" if arr.dtype == np.float64
                       else "ref_downsample_f32")
    if not os.path.exists(exe):
        raise FileNotFoundError(exe)
    dz, dy, dx = arr.shape
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            np.array([arr.dtype.itemsize, dx, dy, dz, *guard, stagger, d],
                     dtype="<i4").tofile(f)
            np.ascontiguousarray(arr).tofile(f)
        subprocess.run([exe, fin, fout], check=True)
        raw = np.fromfile(fout, dtype="<i4", count=3)
        out = np.fromfile(fout, dtype="<f4", offset=12)
    ex = tuple(int(v) for v in raw)
    return out.reshape(ex[2], ex[1], ex[0])
