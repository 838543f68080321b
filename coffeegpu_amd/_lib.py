"""ctypes binding of the C ABI in include/coffee_hip.h (libcoffee_hip.so).

The library is the product path: nothing here falls back to CPU code.  If the
shared library is missing, `load()` raises.
"""

from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcoffee_hip.so")

COFFEE_SOLVER_EZ = 0
COFFEE_SOLVER_EZ_SPHERICAL = 1
COFFEE_SOLVER_GR_EZ = 2
# static partner-state array of coffee_ez_step_noalloc (COFFEE_NOALLOC_SCRATCH_BYTES)
NOALLOC_SCRATCH_BYTES = 80 << 20

_D3 = ctypes.c_double * 3
_I3 = ctypes.c_int * 3
_P3 = ctypes.c_void_p * 3


class CoffeeGrid(ctypes.Structure):
    _fields_ = [("dims", _I3), ("guard", _I3), ("delta", _D3), ("lower", _D3), ("sizes", _D3)]


class CoffeeParams(ctypes.Structure):
    _fields_ = [
        ("dt", ctypes.c_double), ("periodic", _I3), ("pml", _I3), ("pmllen", ctypes.c_int),
        ("sigpml", ctypes.c_double), ("divB_clean", ctypes.c_int), ("ch2", ctypes.c_double),
        ("tau", ctypes.c_double), ("KOeps", ctypes.c_double), ("KO_geometry", ctypes.c_int),
        ("use_edotb_damping", ctypes.c_int), ("damp_gamma", ctypes.c_double),
        ("calc_current", ctypes.c_int), ("clean_ep", ctypes.c_int), ("check_egb", ctypes.c_int),
        ("problem", ctypes.c_int), ("radius", ctypes.c_double), ("omega", ctypes.c_double),
        ("b0", ctypes.c_double), ("p1", ctypes.c_double), ("p2", ctypes.c_double),
        ("p3", ctypes.c_double), ("q11", ctypes.c_double), ("q12", ctypes.c_double),
        ("q13", ctypes.c_double), ("q22", ctypes.c_double), ("q23", ctypes.c_double),
        ("q_offset_x", ctypes.c_double), ("q_offset_y", ctypes.c_double),
        ("q_offset_z", ctypes.c_double), ("tp_start", ctypes.c_double),
        ("tp_end", ctypes.c_double), ("rpert1", ctypes.c_double), ("rpert2", ctypes.c_double),
        ("dw0", ctypes.c_double), ("nT", ctypes.c_double), ("ph1", ctypes.c_double),
        ("ph2", ctypes.c_double), ("dph", ctypes.c_double), ("theta0", ctypes.c_double),
        ("drpert", ctypes.c_double), ("tp_start1", ctypes.c_double),
        ("tp_end1", ctypes.c_double), ("rpert11", ctypes.c_double),
        ("rpert21", ctypes.c_double), ("dw1", ctypes.c_double), ("nT1", ctypes.c_double),
        ("ph11", ctypes.c_double), ("ph21", ctypes.c_double), ("dph1", ctypes.c_double),
        ("pert_type", ctypes.c_int), ("a", ctypes.c_double), ("thp1", ctypes.c_double),
        ("thp2", ctypes.c_double), ("thp11", ctypes.c_double), ("thp21", ctypes.c_double),
        ("lower", _D3), ("size", _D3),
        ("mask_lanes", ctypes.c_int), ("shift_ghost", ctypes.c_int),
    ]


class CoffeeFields(ctypes.Structure):
    _fields_ = [("E", _P3), ("B", _P3), ("P", ctypes.c_void_p), ("dE", _P3), ("dB", _P3),
                ("dP", ctypes.c_void_p), ("Etmp", _P3), ("Btmp", _P3), ("Ptmp", ctypes.c_void_p),
                ("Bbg", _P3), ("J", _P3), ("divE", ctypes.c_void_p), ("divB", ctypes.c_void_p),
                ("dU_KO", ctypes.c_void_p), ("dU_Epar", ctypes.c_void_p),
                ("dU_EgtB", ctypes.c_void_p), ("dU_KO_cum", ctypes.c_void_p),
                ("dU_Epar_cum", ctypes.c_void_p), ("dU_EgtB_cum", ctypes.c_void_p)]


_V, _I, _L = ctypes.c_void_p, ctypes.c_int, ctypes.c_long
_F, _D = ctypes.c_float, ctypes.c_double

# every symbol include/*.h declares, with its ctypes signature
SIGNATURES = {
    "coffee_abi_version": (ctypes.c_int, []),
    "coffee_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "coffee_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.POINTER(CoffeeGrid),
                                         ctypes.POINTER(CoffeeParams),
                                         ctypes.POINTER(ctypes.c_void_p)]),
    "coffee_ctx_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "coffee_set_params": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(CoffeeParams)]),
    "coffee_ez_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(CoffeeFields),
                                      ctypes.c_double, ctypes.c_void_p]),
    "coffee_ez_step_noalloc": (ctypes.c_int, [ctypes.POINTER(CoffeeGrid),
                                              ctypes.POINTER(CoffeeParams), ctypes.c_int,
                                              ctypes.POINTER(CoffeeFields), ctypes.c_double,
                                              ctypes.c_void_p]),
    "coffee_fill_guards": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(CoffeeFields),
                                          ctypes.c_void_p]),
    "coffee_total_energy": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p * 3,
                                           ctypes.POINTER(ctypes.c_double), ctypes.c_void_p]),
    "coffee_residuals": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(CoffeeFields),
                                        ctypes.c_double * 5, ctypes.c_void_p]),
    "coffee_set_diagnostics": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "coffee_set_timing": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "coffee_set_schedule": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "coffee_get_timing": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.POINTER(ctypes.c_int)]),
    "coffee_reset_timing": (ctypes.c_int, [ctypes.c_void_p]),
    "coffee_sph_batch_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                               ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                               ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]),
    "coffee_sph_batch_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "coffee_sph_batch_residuals": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p,
                                                  ctypes.c_void_p]),
    "coffee_sph_batch_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]),
    "coffee_sph_num_tiles": (ctypes.c_int, []),
    "coffee_sph_tile_shape": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                             ctypes.POINTER(ctypes.c_int)]),
    "coffee_sph_batch_set_tile": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "coffee_sph_batch_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                                 ctypes.POINTER(ctypes.c_int), ctypes.c_void_p]),
    "coffee_sph_batch_use_graph": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "coffee_sph_batch_luminosity": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p,
                                                    ctypes.c_void_p, ctypes.c_void_p]),
    "coffee_sph_batch_status": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                               ctypes.c_void_p]),
    # include/coffee_unet.h
    "coffee_unet_conv3x3": (ctypes.c_int, [_V, _V, _V, _V, _V, _I, _I, _I, _I, _I, _I, _I, _V]),
    "coffee_unet_conv3x3_wgrad": (ctypes.c_int, [_V, _V, _V, _I, _V, _V, _I, _I, _I, _I, _I, _V]),
    "coffee_unet_combine": (ctypes.c_int, [_V, _I, _I, _V, _I, _I, _V, _I, _I, _V, _I, _I, _I, _L,
                                           _V]),
    "coffee_unet_pool2": (ctypes.c_int, [_V, _V, _I, _I, _I, _I, _V]),
    "coffee_unet_pool2_bwd": (ctypes.c_int, [_V, _V, _I, _I, _I, _I, _V]),
    "coffee_unet_up2": (ctypes.c_int, [_V, _V, _I, _I, _I, _I, _V]),
    "coffee_unet_up2_bwd": (ctypes.c_int, [_V, _V, _I, _I, _I, _I, _V]),
    "coffee_unet_mse": (ctypes.c_int, [_V, _I, _V, _I, _L, _V, ctypes.POINTER(ctypes.c_double),
                                       _V]),
    "coffee_unet_adam": (ctypes.c_int, [_V, _V, _V, _V, _L, _F, _F, _F, _F, _I, _V, _V, _I, _I,
                                        _V]),
    "coffee_unet_weights_bf16": (ctypes.c_int, [_V, _V, _V, _I, _I, _V]),
    "coffee_unet_to_bf16": (ctypes.c_int, [_V, _V, _L, _V]),
    "coffee_unet_from_bf16": (ctypes.c_int, [_V, _V, _L, _V]),
    "coffee_unet_pack": (ctypes.c_int, [_V * 6, _I, _I, _I, _I, _I, _I, _D, _D, _D, _D, _D, _V, _I,
                                        _I, _I, _V]),
    "coffee_unet_unpack": (ctypes.c_int, [_V, _I, _I, _V * 6, _I, _I, _I, _I, _I, _I, _D, _D, _D,
                                          _D, _V]),
    "coffee_unet_corotate": (ctypes.c_int, [_V * 6, _I, _I, _I, _I, _I, _I, _D, _D, _D, _D, _V]),
    "coffee_fill": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_ulonglong, ctypes.c_int,
                                   ctypes.c_double, ctypes.c_void_p]),
    "coffee_stream_copy": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_ulonglong,
                                          ctypes.c_int, ctypes.c_int, ctypes.c_void_p]),
    "coffee_downsample": (ctypes.c_int, [ctypes.POINTER(CoffeeGrid), ctypes.c_void_p, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                         ctypes.c_void_p]),
    "coffee_fill_guards_arrays": (ctypes.c_int, [ctypes.POINTER(CoffeeGrid), ctypes.c_int * 3,
                                                 ctypes.POINTER(ctypes.c_void_p), ctypes.c_int,
                                                 ctypes.c_int, ctypes.c_void_p]),
}

_lib = None


def load():
    """Load libcoffee_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; "
                "g.build()')")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


class CoffeeError(RuntimeError):
    pass


def check(rc: int, ctx=None):
    if rc != 0:
        msg = load().coffee_last_error(ctx)
        raise CoffeeError(f"coffee error {rc}: {msg.decode() if msg else ''}")
