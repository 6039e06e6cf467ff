"""Depth helpers of Simulation.update_depth (python/pygetm/simulation.py:1498-1600)
on the device.

``elevation2depth`` and ``thickness2vertical_coordinates`` mirror
pygetm._pygetm's functions of the same names (python/_pygetm.pyx:350-367) over
the C-ABI entries ``c_elevation2depth`` and ``c_thickness2vertical_coordinates``,
bit-identical to the reference's Fortran (python/src/wrapper.F90).  They keep the
water depth D and the layer coordinates zc/zf of a coupled run on the GPU next
to the thicknesses the tracer step reads.
"""
from __future__ import annotations

from . import _lib
from .core import CENTERS, INTERFACES, Array


def _f64(a: Array, name: str, shape) -> None:
    t = a.all_values
    if t.dtype.itemsize != 8 or not t.is_floating_point() or not t.is_contiguous() or not t.is_cuda:
        raise ValueError(f"{name}: expected a contiguous float64 device tensor")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name}: expected shape {tuple(shape)}, got {tuple(t.shape)}")


def elevation2depth(z: Array, H: Array, Dmin: float, D: Array) -> None:
    """D = max(H + z, Dmin) where the grid's mask is non-zero; D elsewhere is
    left as it is (python/_pygetm.pyx:362-367)."""
    grid = z.grid
    if H.grid is not grid or D.grid is not grid:
        raise ValueError("z, H and D must be on the same grid")
    shape = z.all_values.shape
    for name, a in (("z", z), ("H", H), ("D", D)):
        _f64(a, name, shape)
    mask = grid.mask.all_values
    if tuple(mask.shape) != tuple(shape):
        raise ValueError("z must be a 2D field of its grid")
    rc = _lib.load().c_elevation2depth(z.all_values.numel(), _lib.ptr(z.all_values), _lib.ptr(H.all_values),
                                       float(Dmin), _lib.ptr(mask), _lib.ptr(D.all_values), _lib.stream_handle())
    _lib.check(rc, "c_elevation2depth")


def thickness2vertical_coordinates(mask: Array, H: Array, h: Array, zc: Array, zf: Array) -> None:
    """Layer-centre (zc) and interface (zf) depths from the layer thicknesses h
    above the bottom at -H, where mask is non-zero (python/_pygetm.pyx:350-355)."""
    grid = mask.grid
    if h.grid is not grid or h.z != CENTERS:
        raise ValueError("h must be a layer field of the mask's grid")
    if H.grid is not grid or H.z:
        raise ValueError("H must be a 2D field of the mask's grid")
    if zc.grid is not grid or zc.z != CENTERS:
        raise ValueError("zc must be a layer field of the mask's grid")
    if zf.grid is not grid or zf.z != INTERFACES:
        raise ValueError("zf must be an interface field of the mask's grid")
    nz, ny_, nx_ = grid.nz_, grid.ny_, grid.nx_
    _f64(H, "H", (ny_, nx_))
    _f64(h, "h", (nz, ny_, nx_))
    _f64(zc, "zc", (nz, ny_, nx_))
    _f64(zf, "zf", (nz + 1, ny_, nx_))
    if tuple(mask.all_values.shape) != (ny_, nx_):
        raise ValueError("mask must be a 2D field")
    rc = _lib.load().c_thickness2vertical_coordinates(nx_, ny_, nz, _lib.ptr(mask.all_values),
                                                      _lib.ptr(H.all_values), _lib.ptr(h.all_values),
                                                      _lib.ptr(zc.all_values), _lib.ptr(zf.all_values),
                                                      _lib.stream_handle())
    _lib.check(rc, "c_thickness2vertical_coordinates")
