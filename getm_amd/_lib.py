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
}

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
