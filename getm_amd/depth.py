"""Simulation.update_depth (python/pygetm/simulation.py:1498-1600) on the device.

``elevation2depth``, ``alpha`` and ``thickness2vertical_coordinates`` mirror
pygetm._pygetm's functions of the same names (python/_pygetm.pyx:350-367) over
the C-ABI entries ``c_elevation2depth``, ``c_alpha`` and
``c_thickness2vertical_coordinates``, bit-identical to the reference's Fortran
(python/src/wrapper.F90).  ``DepthUpdate`` is the whole of update_depth: water
depths on the T, U, V and X grids and the advection grids, the shallow-water
damping, and (each 3D step) the layer thicknesses ho/hn through a
vertical-coordinate scheme (getm_amd.vertical_coordinates) and the layer
coordinates zc/zf.  It keeps a coupled run's thicknesses on the GPU next to the
tracer step that reads them.
"""
from __future__ import annotations

import torch

from . import _lib
from .core import CENTERS, INTERFACES, Array, Grid


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


def alpha(D: Array, Dmin: float, Dcrit: float, out: Array) -> None:
    """out = max(0, min(1, (D - Dmin) / (Dcrit - Dmin))) where the grid's mask
    is non-zero (python/_pygetm.pyx:357-360)."""
    grid = D.grid
    if out.grid is not grid:
        raise ValueError("D and out must be on the same grid")
    shape = D.all_values.shape
    _f64(D, "D", shape)
    _f64(out, "out", shape)
    mask = grid.mask.all_values
    if tuple(mask.shape) != tuple(shape):
        raise ValueError("D must be a 2D field of its grid")
    rc = _lib.load().c_alpha(D.all_values.numel(), _lib.ptr(D.all_values), float(Dmin), float(Dcrit),
                             _lib.ptr(mask), _lib.ptr(out.all_values), _lib.stream_handle())
    _lib.check(rc, "c_alpha")


def _field(grid: Grid, name: str, z=None) -> Array:
    """The grid's field ``name``, made (NaN-filled) when the grid has none."""
    arr = grid._fields.get(name)
    if arr is None:
        arr = grid.array(z, fill=float("nan"), name=name)
        grid._fields[name] = arr
    return arr


class DepthUpdate:
    """Simulation.update_depth for the grids T, U, V and X (T.ugrid/T.vgrid are U
    and V; U.ugrid/U.vgrid and V.ugrid/V.vgrid the velocity advection grids UU,
    UV, VU, VV).  Fields are read by the reference's names: T.z, T.zo (and
    T.zin, T.zio for the 3D step), H on T, U, V and X, and (3D) hn on T, U and V.
    D, T.Dclip, U.alpha, V.alpha, ho, zc and zf are made where absent.  X has one
    point more than T in each direction, as pygetm's X grid.  The domain is one
    undivided subdomain without open-boundary mirrors, so the halo exchanges and
    ``mirror`` calls of the reference do nothing (core.py:723-728)."""

    def __init__(self, T: Grid, U: Grid, V: Grid, X: Grid, Dmin: float, Dcrit: float, vertical_coordinates=None):
        if (X.ny_, X.nx_) != (T.ny_ + 1, T.nx_ + 1):
            raise ValueError("the X grid must have one point more than T in each direction")
        self.T, self.U, self.V, self.X = T, U, V, X
        self.Dmin, self.Dcrit = float(Dmin), float(Dcrit)
        self.vertical_coordinates = vertical_coordinates
        self._work = {g.name: g.array(fill=float("nan")) for g in (T, U, V, X)}
        for g in (T, U, V, X):
            _field(g, "D")
        _field(T, "Dclip")
        _field(U, "alpha")
        _field(V, "alpha")
        for g in (U.ugrid, U.vgrid, V.ugrid, V.vgrid):
            if g is not None:
                _field(g, "D")
        if vertical_coordinates is not None:
            for g in (T, U, V):
                if "ho" not in g._fields:
                    g.set_field("ho", g.hn.all_values.clone())
                _field(g, "zc", CENTERS)
                _field(g, "zf", INTERFACES)
            # U and V depths stand in for their clipped depths (core.py:374)
            vertical_coordinates.initialize(T, U, V)

    def __call__(self, _3d: bool = False, timestep: float = 0.0) -> None:
        T, U, V, X = self.T, self.U, self.V, self.X
        if _3d:
            T.zio.all_values.copy_(T.zin.all_values)
            T.zin.all_values.copy_(T.z.all_values)
            z_T, zo_T = T.zin, T.zio
        else:
            z_T, zo_T = T.z, T.zo
        z_U, z_V, z_X, z_T_half = (self._work[g.name] for g in (U, V, X, T))

        # total water depth on T, raw and clipped (simulation.py:1530-1535)
        elevation2depth(z_T, T.H, -1000.0, T.D)
        torch.clamp(T.D.all_values, min=self.Dmin, out=T.Dclip.all_values)  # np.maximum(D, Dmin)
        # elevations half a step behind T, then depths on U, V and X (:1537-1555)
        torch.add(zo_T.all_values, z_T.all_values, out=z_T_half.all_values)
        z_T_half.all_values.mul_(0.5)
        z_T_half.interp(z_U)
        elevation2depth(z_U, U.H, self.Dmin, U.D)
        z_T_half.interp(z_V)
        elevation2depth(z_V, V.H, self.Dmin, V.D)
        z_T_half.interp(z_X)
        elevation2depth(z_X, X.H, self.Dmin, X.D)
        # shallow-water damping (:1568-1570)
        alpha(U.D, 2 * self.Dmin, self.Dcrit, U.alpha)
        alpha(V.D, 2 * self.Dmin, self.Dcrit, V.alpha)
        # depths on the advection grids, half a step behind T (:1572-1580); the
        # half-step elevation becomes the half-step depth in place
        D_T_half = z_T_half
        elevation2depth(z_T_half, T.H, self.Dmin, D_T_half)
        if U.ugrid is not None:
            U.ugrid.D.all_values[:, :-1].copy_(D_T_half.all_values[:, 1:])
        if V.vgrid is not None:
            V.vgrid.D.all_values[:-1, :].copy_(D_T_half.all_values[1:, :])
        if U.vgrid is not None:
            U.vgrid.D.all_values.copy_(X.D.all_values[1:, 1:])
            if V.ugrid is not None:
                V.ugrid.D.all_values.copy_(U.vgrid.D.all_values)

        if _3d:
            for g in (T, U, V):
                g.ho.all_values.copy_(g.hn.all_values)
            self.vertical_coordinates.update(timestep)
            for g in (T, U, V):
                thickness2vertical_coordinates(g.mask, g.H, g.hn, g.zc, g.zf)
