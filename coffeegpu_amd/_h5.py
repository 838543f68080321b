"""ctypes binding of libcoffee_h5.so (include/coffee_h5.h): the HDF5 snapshot
and field-output files."""

from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcoffee_h5.so")

F32, F64, U32, I32 = 1, 2, 3, 4
_TYPES = {np.dtype(np.float32): F32, np.dtype(np.float64): F64, np.dtype(np.uint32): U32,
          np.dtype(np.int32): I32}
_CTYPES = {F32: ctypes.c_float, F64: ctypes.c_double, U32: ctypes.c_uint32, I32: ctypes.c_int32}

_U3 = ctypes.c_ulonglong * 3
_V, _P, _I, _C = ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_char_p

# every symbol include/coffee_h5.h declares, with its ctypes signature
SIGNATURES = {
    "coffee_h5_create": (_I, [_C, _P]),
    "coffee_h5_open": (_I, [_C, _I, _P]),
    "coffee_h5_close": (_I, [_V]),
    "coffee_h5_write_array": (_I, [_V, _C, _V, _I, _I, _U3]),
    "coffee_h5_read_array": (_I, [_V, _C, _V, _I, _I, _U3]),
    "coffee_h5_write_scalar": (_I, [_V, _C, _I, _V]),
    "coffee_h5_read_scalar": (_I, [_V, _C, _I, _V]),
    "coffee_h5_has": (_I, [_V, _C]),
    "coffee_h5_last_error": (_C, []),
}

_lib = None


def load():
    """Load libcoffee_h5.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: it is built with the HDF5 C library "
                               "(python -c 'import __graft_entry__ as g; g.build()')")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


class H5Error(RuntimeError):
    pass


def _check(rc: int):
    if rc != 0:
        raise H5Error(f"coffee_h5 error {rc}: {load().coffee_h5_last_error().decode()}")


class H5File:
    """One HDF5 file: H5File / hdf_create of the reference
    (src/utils/hdf_wrapper.h).  Arrays are numpy arrays in C order of a
    Fortran-order field, i.e. shape (z, y, x) for extent {x, y, z}; the file
    holds them as datasets [z][y][x], as the reference writes them."""

    def __init__(self, path: str, mode: str = "r"):
        lib = load()
        self._f = ctypes.c_void_p()
        if mode == "w":
            _check(lib.coffee_h5_create(path.encode(), ctypes.byref(self._f)))
        elif mode in ("r", "r+"):
            _check(lib.coffee_h5_open(path.encode(), int(mode == "r+"), ctypes.byref(self._f)))
        else:
            raise ValueError("mode must be 'w', 'r' or 'r+'")

    def close(self):
        if self._f:
            f, self._f = self._f, ctypes.c_void_p()
            _check(load().coffee_h5_close(f))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _ext(shape):
        ext = _U3(0, 0, 0)
        for i, n in enumerate(reversed(shape)):
            ext[i] = n
        return ext

    def write(self, name: str, arr: np.ndarray):
        a = np.ascontiguousarray(arr)
        t = _TYPES.get(a.dtype)
        if t is None:
            raise TypeError(f"unsupported dtype {a.dtype}")
        if not 1 <= a.ndim <= 3:
            raise ValueError("arrays of 1 to 3 dimensions")
        _check(load().coffee_h5_write_array(self._f, name.encode(), a.ctypes.data_as(_V), t,
                                            a.ndim, self._ext(a.shape)))

    def read(self, name: str, shape, dtype) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        t = _TYPES[out.dtype]
        _check(load().coffee_h5_read_array(self._f, name.encode(), out.ctypes.data_as(_V), t,
                                           out.ndim, self._ext(out.shape)))
        return out

    def write_scalar(self, name: str, value, dtype):
        t = _TYPES[np.dtype(dtype)]
        v = _CTYPES[t](value)
        _check(load().coffee_h5_write_scalar(self._f, name.encode(), t, ctypes.byref(v)))

    def read_scalar(self, name: str, dtype):
        t = _TYPES[np.dtype(dtype)]
        v = _CTYPES[t]()
        _check(load().coffee_h5_read_scalar(self._f, name.encode(), t, ctypes.byref(v)))
        return np.dtype(dtype).type(v.value)

    def has(self, name: str) -> bool:
        rc = load().coffee_h5_has(self._f, name.encode())
        if rc < 0:
            _check(rc)
        return rc == 1
