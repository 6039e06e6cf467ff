"""Drop-in GPU counterparts of pygetm.operators (python/pygetm/operators.py:1-230).

Class names, constructor and method signatures, enums and semantics are the
reference's; arrays are :class:`getm_amd.core.Array` device fields and all
work runs as HIP kernels of ``libgetm_amd.so`` on the current torch stream.
Halo exchanges are no-ops (one undivided domain).
"""
from __future__ import annotations

import enum
from typing import Callable, Iterable, Optional

import torch

from . import _lib
from .core import CENTERS, INTERFACES, Array, Grid


class AdvectionScheme(enum.IntEnum):
    """operators.py:11-18"""
    HSIMT = 1
    MUSCL = 2
    P2_PDM = 3
    SPLMAX13 = 4
    SUPERBEE = 5
    UPSTREAM = 6
    DEFAULT = SUPERBEE


class AdvectionSplit(enum.Enum):
    """operators.py:21-24"""
    FULL = 1
    HALF = 2
    HALF_ALWAYS = 3


def _stream():
    return _lib.stream_handle()


class Advection:
    """Split TVD advection (operators.py:27-193, Cython _pygetm.pyx:198-261).

    ``h`` is the FP64 work array of layer thicknesses that evolves through the
    stages (shape ``(max(nz, 1), ny_, nx_)``); ``D = h[0]`` serves 2D advection."""

    def __init__(self, grid: Grid, scheme: AdvectionScheme = AdvectionScheme.DEFAULT,
                 split_2d: AdvectionSplit = AdvectionSplit.HALF):
        self.grid = grid
        self.ugrid = grid.ugrid
        self.vgrid = grid.vgrid
        self.scheme = AdvectionScheme(int(scheme))
        lib = _lib.load()
        self.p = lib.advection_create(int(scheme), grid.p)
        if not self.p:
            raise ValueError(f"invalid advection scheme {scheme!r}")
        self.h = torch.empty((max(grid.nz_, 1), grid.ny_, grid.nx_), dtype=torch.float64, device=grid.device)
        self.D = self.h[0]
        self.ufirst = False
        self.split_2d = split_2d

    def __del__(self):
        p = getattr(self, "p", None)
        if p and _lib._lib is not None:
            _lib._lib.advection_finalize(p)
            self.p = None

    @property
    def ufirst(self) -> bool:
        return self._ufirst

    @ufirst.setter
    def ufirst(self, value: bool):
        self._ufirst = bool(value)

    # -- single sweeps (_pygetm.pyx:236-261) -----------------------------------
    def _uv(self, direction: int, nk: int, grid: Grid, vel: Array, Ah: Optional[Array], timestep: float,
            h: torch.Tensor, huv: torch.Tensor, var: Array) -> None:
        rc = _lib.load().advection_uv_calculate(
            direction, nk, self.p, self.grid.p, grid.p, _lib.ptr(vel.all_values),
            _lib.ptr(Ah.all_values) if Ah is not None else None, float(timestep), _lib.ptr(h), _lib.ptr(huv),
            _lib.ptr(var.all_values), _stream())
        _lib.check(rc, "advection_uv_calculate")

    def u_2d(self, u: Array, timestep: float, var: Array, Ah: Optional[Array] = None) -> None:
        self._uv(1, 1, self.ugrid, u, Ah, timestep, self.D, self.ugrid.D.all_values, var)

    def v_2d(self, v: Array, timestep: float, var: Array, Ah: Optional[Array] = None) -> None:
        self._uv(2, 1, self.vgrid, v, Ah, timestep, self.D, self.vgrid.D.all_values, var)

    def u_3d(self, u: Array, timestep: float, var: Array, Ah: Optional[Array] = None) -> None:
        self._uv(1, var.all_values.shape[0], self.ugrid, u, Ah, timestep, self.h, self.ugrid.hn.all_values, var)

    def v_3d(self, v: Array, timestep: float, var: Array, Ah: Optional[Array] = None) -> None:
        self._uv(2, var.all_values.shape[0], self.vgrid, v, Ah, timestep, self.h, self.vgrid.hn.all_values, var)

    def w_3d(self, w: Array, w_var: Array, timestep: float, var: Array) -> None:
        rc = _lib.load().advection_w_calculate(self.p, self.grid.p, _lib.ptr(w.all_values),
                                               _lib.ptr(w_var.all_values), float(timestep), _lib.ptr(self.h),
                                               _lib.ptr(var.all_values), _stream())
        _lib.check(rc, "advection_w_calculate")

    # -- 2D split advection (operators.py:52-93) -------------------------------
    def __call__(self, u: Array, v: Array, timestep: float, var: Array, Ah_u: Optional[Array] = None,
                 Ah_v: Optional[Array] = None, skip_initial_halo_exchange: bool = False):
        assert u.grid is self.ugrid and not u.z
        assert v.grid is self.vgrid and not v.z
        assert var.grid is self.grid and not var.z
        adv1 = lambda dt: self.u_2d(u, dt, var, Ah_u)
        adv2 = lambda dt: self.v_2d(v, dt, var, Ah_v)
        if not self._ufirst:
            adv1, adv2 = adv2, adv1
        self.D.copy_(self.grid.D.all_values)
        if self.split_2d == AdvectionSplit.FULL:
            adv1(timestep)
            adv2(timestep)
        elif self.split_2d == AdvectionSplit.HALF:
            adv1(0.5 * timestep)
            adv2(timestep)
            adv1(0.5 * timestep)
        else:
            adv1(0.5 * timestep)
            adv2(0.5 * timestep)
            adv2(0.5 * timestep)
            adv1(0.5 * timestep)

    # -- 3D (operators.py:95-193) ----------------------------------------------
    def apply_3d(self, u: Array, v: Array, w: Array, timestep: float, var: Array, Ah_u: Optional[Array] = None,
                 Ah_v: Optional[Array] = None, new_h: bool = False, skip_initial_halo_exchange: bool = False,
                 w_var: Optional[Array] = None):
        if w_var is None:
            w_var = w
        assert u.grid is self.ugrid and u.z == CENTERS
        assert v.grid is self.vgrid and v.z == CENTERS
        assert w.grid is self.grid and w.z == INTERFACES
        assert w_var.grid is self.grid and w_var.z == INTERFACES
        assert var.grid is self.grid and var.z == CENTERS
        self.h.copy_((self.grid.hn if new_h else self.grid.ho).all_values)
        rc = _lib.load().advection_apply_3d(
            self.p, self.grid.p, self.ugrid.p, self.vgrid.p, _lib.ptr(u.all_values), _lib.ptr(v.all_values),
            _lib.ptr(w.all_values), _lib.ptr(w_var.all_values),
            _lib.ptr(Ah_u.all_values) if Ah_u is not None else None,
            _lib.ptr(Ah_v.all_values) if Ah_v is not None else None, float(timestep), _lib.ptr(self.h),
            _lib.ptr(var.all_values), int(self._ufirst), _stream())
        _lib.check(rc, "advection_apply_3d")

    def apply_3d_batch(self, u: Array, v: Array, w: Array, timestep: float, vars: Iterable[Array],
                       Ah_u: Optional[Array] = None, Ah_v: Optional[Array] = None, new_h: bool = False,
                       skip_initial_halo_exchange: bool = False,
                       get_w: Optional[Callable[[Array], Array]] = None):
        """Stage-outer, tracer-inner loop of operators.py:135-193: every tracer
        starts each stage from the thickness the previous stage left, which for
        the w stage is the last tracer's (its sub-step count follows that
        tracer's w_var, and so do the roundings of h).  When every tracer has the
        same vertical velocity, the thickness sequence is the same for all of
        them, and each tracer runs its five stages in one ``advection_apply_3d``
        call from the starting thickness instead (stage pairs on the GPU, no
        copies of h between stages): the same bits, with ``h`` left at the final
        thickness."""
        if get_w is None:
            get_w = lambda tracer: w
        vars = list(vars)
        assert u.grid is self.ugrid and u.z == CENTERS
        assert v.grid is self.vgrid and v.z == CENTERS
        assert w.grid is self.grid and w.z == INTERFACES
        for var in vars:
            assert var.grid is self.grid and var.z == CENTERS
        w_vars = [get_w(var) for var in vars]
        for w_var in w_vars:
            assert w_var.grid is self.grid and w_var.z == INTERFACES
        if all(w_var is w_vars[0] for w_var in w_vars):
            for var in vars:
                self.apply_3d(u, v, w, timestep, var, Ah_u, Ah_v, new_h=new_h, w_var=w_vars[0])
            return
        adv1 = lambda var: self.u_3d(u, 0.5 * timestep, var, Ah_u)
        adv2 = lambda var: self.v_3d(v, 0.5 * timestep, var, Ah_v)
        if not self._ufirst:
            adv1, adv2 = adv2, adv1
        current_h = (self.grid.hn if new_h else self.grid.ho).all_values.clone()
        for stage in range(5):
            for var, w_var in zip(vars, w_vars):
                self.h.copy_(current_h)
                if stage in (0, 4):
                    adv1(var)
                elif stage in (1, 3):
                    adv2(var)
                else:
                    self.w_3d(w, w_var, timestep, var)
            current_h.copy_(self.h)

    def stats(self, reset: bool = True):
        """(largest z sub-iteration count, columns capped at 200) since the last reset."""
        import ctypes

        mi, cap = ctypes.c_int(0), ctypes.c_int(0)
        _lib.check(_lib.load().advection_get_stats(self.p, ctypes.byref(mi), ctypes.byref(cap), int(reset),
                                                   _stream()), "advection_get_stats")
        return mi.value, cap.value


class VerticalAdvection:
    """operators.py:196-215"""

    def __init__(self, grid: Grid, scheme: AdvectionScheme = AdvectionScheme.DEFAULT):
        self.grid = grid
        self.p = _lib.load().advection_create(int(scheme), grid.p)
        if not self.p:
            raise ValueError(f"invalid advection scheme {scheme!r}")
        self.h = torch.empty((grid.nz_, grid.ny_, grid.nx_), dtype=torch.float64, device=grid.device)

    def __del__(self):
        p = getattr(self, "p", None)
        if p and _lib._lib is not None:
            _lib._lib.advection_finalize(p)
            self.p = None

    def w_3d(self, w: Array, w_var: Array, timestep: float, var: Array) -> None:
        rc = _lib.load().advection_w_calculate(self.p, self.grid.p, _lib.ptr(w.all_values),
                                               _lib.ptr(w_var.all_values), float(timestep), _lib.ptr(self.h),
                                               _lib.ptr(var.all_values), _stream())
        _lib.check(rc, "advection_w_calculate")

    def __call__(self, w: Array, w_var: Array, timestep: float, var: Array, new_h: bool = False):
        assert w.grid is self.grid and w.z == INTERFACES
        assert w_var.grid is self.grid and w_var.z == INTERFACES
        self.h.copy_((self.grid.hn if new_h else self.grid.ho).all_values)
        self.w_3d(w, w_var, timestep, var)


class VerticalDiffusion:
    """theta-method vertical diffusion (operators.py:218-230, _pygetm.pyx:288-316)."""

    def __init__(self, grid: Grid, cnpar: float = 1.0):
        self.grid = grid
        self.cnpar = float(cnpar)
        self.p = _lib.load().vertical_diffusion_create(grid.p)
        if not self.p:
            raise _lib.GetmError("vertical_diffusion_create failed")

    def __del__(self):
        p = getattr(self, "p", None)
        if p and _lib._lib is not None:
            _lib._lib.vertical_diffusion_finalize(p)
            self.p = None

    def prepare(self, nuh: Array, timestep: float, molecular: float = 0.0, use_ho: bool = False) -> None:
        g = nuh.grid
        hn = g.hn.all_values
        ho = g.ho.all_values if use_ho else hn
        rc = _lib.load().c_vertical_diffusion_prepare(self.p, g.nx_, g.ny_, g.nz_, float(molecular),
                                                      _lib.ptr(nuh.all_values), float(timestep), self.cnpar,
                                                      _lib.ptr(g.mask.all_values), _lib.ptr(ho), _lib.ptr(hn),
                                                      _stream())
        _lib.check(rc, "c_vertical_diffusion_prepare")

    def apply(self, var: Array, ea2: Optional[Array] = None, ea4: Optional[Array] = None,
              use_ho: bool = False) -> None:
        g = var.grid
        hn = g.hn.all_values
        ho = g.ho.all_values if use_ho else hn
        rc = _lib.load().c_vertical_diffusion_apply(
            self.p, g.nx_, g.ny_, g.nz_, _lib.ptr(g.mask.all_values), _lib.ptr(ho), _lib.ptr(hn),
            _lib.ptr(var.all_values), _lib.ptr(ea2.all_values) if ea2 is not None else None,
            _lib.ptr(ea4.all_values) if ea4 is not None else None, _stream())
        _lib.check(rc, "c_vertical_diffusion_apply")

    def __call__(self, nuh: Array, timestep: float, var: Array, molecular: float = 0.0,
                 ea2: Optional[Array] = None, ea4: Optional[Array] = None, use_ho: bool = False):
        self.prepare(nuh, timestep, molecular, use_ho)
        self.apply(var, ea2, ea4, use_ho)


def vertical_advection_to_sources(halox: int, haloy: int, mask: torch.Tensor, c: torch.Tensor, w: torch.Tensor,
                                  h: torch.Tensor, s: torch.Tensor) -> None:
    """pygetm._pygetm.vertical_advection_to_sources (python/_pygetm.pyx:574-575,
    python/src/wrapper.F90:447-486): first-order upstream transfer of the vertical
    velocity ``w`` (sinking < 0, floating > 0) of concentration ``c`` into the
    sources ``s`` of the layers it moves material between, in place.  All arrays
    are contiguous (nz, ny_, nx_) device tensors with layer 0 at the surface (the
    order pygetm's FABM coupling uses), ``mask`` int32, the others float64."""
    shape = tuple(c.shape)
    if len(shape) != 3:
        raise ValueError("arrays must be (nz, ny_, nx_)")
    for name, a, dt in (("mask", mask, torch.int32), ("c", c, torch.float64), ("w", w, torch.float64),
                        ("h", h, torch.float64), ("s", s, torch.float64)):
        if tuple(a.shape) != shape or a.dtype != dt or not a.is_contiguous() or not a.is_cuda:
            raise ValueError(f"{name}: expected a contiguous {dt} device tensor of shape {shape}")
    rc = _lib.load().c_vertical_advection_to_sources(shape[2], shape[1], shape[0], int(halox), int(haloy),
                                                     _lib.ptr(mask), _lib.ptr(c), _lib.ptr(w), _lib.ptr(h),
                                                     _lib.ptr(s), _stream())
    _lib.check(rc, "c_vertical_advection_to_sources")
