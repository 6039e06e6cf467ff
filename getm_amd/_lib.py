"""ctypes binding of the C-ABI library ``libgetm_amd.so`` (include/getm_amd.h).

The library must have been built (``python -m getm_amd.build`` or
``__graft_entry__.build()``); loading fails loudly otherwise -- there is no
CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgetm_amd.so")

GETM_OK = 0
# tracer precision of a transport object (include/getm_amd.h)
GETM_F64, GETM_F32, GETM_F32C = 1, 2, 3
ERRORS = {
    1: "invalid argument or shape mismatch",
    2: "invalid advection scheme",
    3: "vertical diffusion applied before prepare",
    4: "HIP allocation or launch failure",
    5: "size exceeds what the entry point supports",
}

_vp = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double

# name -> (restype, argtypes); mirrors include/getm_amd.h
_SIGNATURES = {
    "create_grid": (_vp, [_i] * 8),
    "grid_finalize": (None, [_vp]),
    "grid_set_array": (_i, [_vp, ctypes.c_char_p, _vp]),
    "advection_create": (_vp, [_i, _vp]),
    "advection_finalize": (None, [_vp]),
    "advection_uv_calculate": (_i, [_i, _i, _vp, _vp, _vp, _vp, _vp, _d, _vp, _vp, _vp, _vp]),
    "advection_w_calculate": (_i, [_vp, _vp, _vp, _vp, _d, _vp, _vp, _vp]),
    "advection_apply_3d": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _d, _vp, _vp, _i, _vp]),
    "advection_get_stats": (_i, [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i), _i, _vp]),
    "vertical_diffusion_create": (_vp, [_vp]),
    "vertical_diffusion_finalize": (None, [_vp]),
    "c_vertical_diffusion_prepare": (_i, [_vp, _i, _i, _i, _d, _vp, _d, _d, _vp, _vp, _vp, _vp]),
    "c_vertical_diffusion_apply": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "transport_create": (_vp, [_vp, _vp, _vp, _i, _d, _i, _i]),
    "transport_finalize": (None, [_vp]),
    "transport_step": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _d, _d, _vp, _vp]),
    "transport_step_mixed": (_i, [_vp, _i, _vp, _vp, _vp, _vp, ctypes.POINTER(_vp), _vp, _vp, _vp, _d,
                                  ctypes.POINTER(_d), _vp, _vp]),
    "transport_get_stats": (_i, [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i), _i, _vp]),
    "transport_set_tuning": (_i, [_vp, _i, _i, _i, _i]),
    "c_w_momentum_3d": (_i, [_i, _i, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _d, _vp, _vp]),
    "c_vertical_advection_to_sources": (_i, [_i, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp]),
    "c_elevation2depth": (_i, [_i, _vp, _vp, _d, _vp, _vp, _vp]),
    "transport_set_sources": (_i, [_vp, _vp]),
    "c_velocities_from_transports": (_i, [ctypes.c_longlong, _vp, _vp, _vp, _vp]),
    "c_alpha": (_i, [_i, _vp, _d, _d, _vp, _vp, _vp]),
    "c_update_gvc": (_i, [_i, _i, _i, _d, _vp, _d, _i, _vp, _vp, _vp, _vp]),
    "c_update_sigma": (_i, [_i, _i, _i, _vp, _vp, _vp, _vp, _vp]),
    "grid_interp_x": (_i, [_i, _i, _i, _vp, _vp, _i, _vp]),
    "grid_interp_y": (_i, [_i, _i, _i, _vp, _vp, _i, _vp]),
    "grid_interp_z": (_i, [_i, _i, _i, _i, _vp, _vp, _i, _vp]),
    "grid_interp_xy": (_i, [_i, _i, _i, _i, _i, _vp, _vp, _i, _i, _vp]),
    "c_reconstruct_transport_change": (_i, [_i, _i, _i, _i, _i, _i, _i, _vp, _vp, _vp, _d, _vp]),
    "c_thickness2vertical_coordinates": (_i, [_i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp]),
    "transport_set_solver": (_i, [_vp, _i]),
    "transport_set_rounding": (_i, [_vp, _i, _vp]),
    "transport_inventory": (_i, [_vp, _i, _vp, _vp, _vp, _vp]),
    "transport_conserve": (_i, [_vp, _i, _vp, _vp, _vp]),
    "transport_conserve_noalloc": (_i, [_vp, _i, _vp, _vp, _vp]),
    "transport_check_finite": (_i, [_vp, _i, _vp, _vp, _vp]),
    "ensemble_create": (_vp, [_i, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                               ctypes.POINTER(_i), _i, _d, _i]),
    "ensemble_finalize": (None, [_vp]),
    "ensemble_bind": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp, _d, _d]),
    "ensemble_set_tuning": (_i, [_vp, _i, _i, _i, _i]),
    "ensemble_set_tile_x": (_i, [_vp, _i]),
    "ensemble_set_solver": (_i, [_vp, _i]),
    "ensemble_set_rounding": (_i, [_vp, _i, ctypes.POINTER(_vp)]),
    "ensemble_commit": (_i, [_vp, _vp]),
    "ensemble_step": (_i, [_vp, _vp]),
    "ensemble_thickness": (_vp, [_vp, _i]),
    "getm_profile_enable": (_i, [_i]),
    "getm_profile_read": (_i, [_i, ctypes.POINTER(_d), ctypes.POINTER(ctypes.c_long), _i]),
}

# kernel identifiers of getm_profile_read (include/getm_amd.h)
KERNELS = {
    "adv_uv_f64": 0,
    "adv_w_f64": 1,
    "vdiff_prepare_f64": 2,
    "vdiff_apply_f64": 3,
    "hpair_vu": 4,
    "hpair_uv": 5,
    "wcol": 6,
    "vdiff_f32": 7,
    "inventory": 8,
    "fix": 9,
}


def profile_enable(on: bool) -> None:
    check(load().getm_profile_enable(int(on)), "getm_profile_enable")


def profile_read(kernel: str, reset: bool = True):
    """(total milliseconds, launches) of one kernel since the last reset."""
    ms, n = ctypes.c_double(0.0), ctypes.c_long(0)
    check(load().getm_profile_read(KERNELS[kernel], ctypes.byref(ms), ctypes.byref(n), int(reset)),
          "getm_profile_read")
    return ms.value, n.value

_lib = None


class GetmError(RuntimeError):
    pass


def load():
    """Load the library once and attach the C signatures."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m getm_amd.build` "
                "(the tracer-transport step has no CPU fallback)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def declared_symbols():
    return list(_SIGNATURES)


def check(rc: int, what: str) -> None:
    if rc != GETM_OK:
        raise GetmError(f"{what}: {ERRORS.get(rc, 'error')} (status {rc})")


def ptr(t):
    """Device pointer of a torch tensor (or None -> NULL)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def stream_handle(device=None):
    import torch

    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
