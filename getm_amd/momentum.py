"""Input producers of the transport step that live in pygetm's momentum module.

``w_momentum_3d`` is pygetm._pygetm.w_momentum_3d (python/_pygetm.pyx:443-455):
vertical velocity at layer interfaces from the layer transports and the
thickness change ho -> hn of the T grid, on the device (C-ABI
``c_w_momentum_3d``, bit-identical to python/src/momentum.F90:191-216).
``velocities_from_transports`` is the division that follows it in
Momentum.update_diagnostics (python/pygetm/momentum.py:885-892).  Keeping both
on the GPU spares coupled runs the per-step host -> device copy of u, v and w.

``VelocityAdvection`` is the momentum self-advection that reuses the tracer
advection on the velocity points' own grids (momentum.py:578-583, 918-961 in 3D,
1081-1097 in 2D): velocities are interpolated onto the UU/UV and VU/VV face
grids (``Array.interp``, the reference's grid_interp_* routines), advected with
``new_h=True``, and turned into transport changes by
``reconstruct_transport_change`` (C-ABI ``c_reconstruct_transport_change``,
bit-identical to momentum.F90:323-340).
"""
from __future__ import annotations

import torch

from . import _lib
from .core import CENTERS, INTERFACES, Array, Grid
from .operators import Advection, AdvectionScheme, AdvectionSplit


def w_momentum_3d(pk: Array, qk: Array, timestep: float, w: Array) -> None:
    """Update ``w`` (T grid, interfaces) in place for interfaces 1 .. nz-1 of wet
    points, over the interior plus one column/row on the high side (the
    reference's range, so w is valid around every U and V point).  w at the
    bottom interface is taken as given (0 in pygetm)."""
    grid = w.grid
    if w.z != INTERFACES:
        raise ValueError("w must be defined on layer interfaces")
    if pk.grid is not grid.ugrid or pk.z != CENTERS:
        raise ValueError("pk must be a layer field of the U grid")
    if qk.grid is not grid.vgrid or qk.z != CENTERS:
        raise ValueError("qk must be a layer field of the V grid")
    ptr = _lib.ptr
    rc = _lib.load().c_w_momentum_3d(
        grid.nx_, grid.ny_, grid.nz_, grid.halox + 1, grid.halox + grid.nx + 1, grid.haloy + 1,
        grid.haloy + grid.ny + 1, ptr(grid.mask.all_values), ptr(pk.grid.dy.all_values), ptr(qk.grid.dx.all_values),
        ptr(grid.iarea.all_values), ptr(grid.ho.all_values), ptr(grid.hn.all_values), ptr(pk.all_values),
        ptr(qk.all_values), float(timestep), ptr(w.all_values), _lib.stream_handle())
    _lib.check(rc, "c_w_momentum_3d")


def velocities_from_transports(pk: Array, qk: Array, uk: Array, vk: Array) -> None:
    """uk = pk / hn (U grid), vk = qk / hn (V grid) over every point,
    momentum.py:885-892 (C-ABI ``c_velocities_from_transports``)."""
    for p, u in ((pk, uk), (qk, vk)):
        hn = p.grid.hn.all_values
        for name, t in (("transport", p.all_values), ("hn", hn), ("velocity", u.all_values)):
            if t.dtype != torch.float64 or not t.is_contiguous() or not t.is_cuda:
                raise ValueError(f"{name}: expected a contiguous float64 device tensor")
        if not (p.all_values.shape == hn.shape == u.all_values.shape):
            raise ValueError("transport, thickness and velocity must have one shape")
        rc = _lib.load().c_velocities_from_transports(p.all_values.numel(), _lib.ptr(p.all_values), _lib.ptr(hn),
                                                      _lib.ptr(u.all_values), _lib.stream_handle())
        _lib.check(rc, "c_velocities_from_transports")


def reconstruct_transport_change(un: Array, hn: torch.Tensor, Uo: Array, timestep: float) -> None:
    """pygetm._pygetm.reconstruct_transport_change (python/_pygetm.pyx:506-510):
    un = (un * hn - Uo) / timestep in place, over every column and the interior
    rows of each layer (momentum.F90:323-340 loops i over the full extent).
    ``hn`` is the thickness the advection ended with (``Advection.h`` or ``.D``)."""
    grid = Uo.grid
    nz = 1 if Uo.all_values.dim() == 2 else grid.nz_
    n = nz * grid.ny_ * grid.nx_
    for name, t in (("un", un.all_values), ("hn", hn), ("Uo", Uo.all_values)):
        if t.dtype != torch.float64 or not t.is_contiguous() or not t.is_cuda or t.numel() != n:
            raise ValueError(f"{name}: expected a contiguous float64 device tensor of {n} values")
    rc = _lib.load().c_reconstruct_transport_change(
        grid.nx_, grid.ny_, nz, grid.halox + 1, grid.halox + grid.nx, grid.haloy + 1, grid.haloy + grid.ny,
        _lib.ptr(un.all_values), _lib.ptr(hn), _lib.ptr(Uo.all_values), float(timestep), _lib.stream_handle())
    _lib.check(rc, "c_reconstruct_transport_change")


class VelocityAdvection:
    """Advection of the velocities by themselves on their own advection grids.

    ``ugrid`` and ``vgrid`` are the U and V grids; ``ugrid.ugrid``/``ugrid.vgrid``
    (UU, UV) and ``vgrid.ugrid``/``vgrid.vgrid`` (VU, VV) carry the face masks,
    metrics and thicknesses (``hn`` in 3D, ``D`` in 2D), as pygetm's domain builds
    them (domain.py:732-741).  The work arrays on the face grids start as NaN, as
    in momentum.py:453-456; the interface column that the T -> U (V)
    interpolation of ww does not reach holds 0."""

    def __init__(self, ugrid: Grid, vgrid: Grid, scheme: AdvectionScheme = AdvectionScheme.DEFAULT,
                 split_2d: AdvectionSplit = AdvectionSplit.HALF):
        self.uadv = Advection(ugrid, scheme=scheme, split_2d=split_2d)
        self.vadv = Advection(vgrid, scheme=scheme, split_2d=split_2d)
        nan = float("nan")
        self.uua3d = ugrid.ugrid.array(CENTERS, fill=nan)
        self.uva3d = ugrid.vgrid.array(CENTERS, fill=nan)
        self.vua3d = vgrid.ugrid.array(CENTERS, fill=nan)
        self.vva3d = vgrid.vgrid.array(CENTERS, fill=nan)
        self.ww_u = ugrid.array(INTERFACES, fill=0.0)
        self.ww_v = vgrid.array(INTERFACES, fill=0.0)
        self.uua = ugrid.ugrid.array(fill=nan)
        self.uva = ugrid.vgrid.array(fill=nan)
        self.vua = vgrid.ugrid.array(fill=nan)
        self.vva = vgrid.vgrid.array(fill=nan)

    def advect_3d(self, uk: Array, vk: Array, ww: Array, pk: Array, qk: Array, timestep: float, advpk: Array,
                  advqk: Array) -> None:
        """advpk, advqk = change of the layer transports pk, qk over ``timestep``
        from advecting uk and vk with velocities interpolated to their advection
        grids and ww interpolated to U and V points (momentum.py:928-961)."""
        advpk.all_values.copy_(uk.all_values)
        self.uadv.apply_3d(uk.interp(self.uua3d), vk.interp(self.uva3d), ww.interp(self.ww_u), timestep, advpk,
                           new_h=True, skip_initial_halo_exchange=True)
        reconstruct_transport_change(advpk, self.uadv.h, pk, timestep)
        advqk.all_values.copy_(vk.all_values)
        self.vadv.apply_3d(uk.interp(self.vua3d), vk.interp(self.vva3d), ww.interp(self.ww_v), timestep, advqk,
                           new_h=True, skip_initial_halo_exchange=True)
        reconstruct_transport_change(advqk, self.vadv.h, qk, timestep)

    def advect_2d(self, U: Array, V: Array, u1: Array, v1: Array, timestep: float, advU: Array,
                  advV: Array) -> None:
        """advU, advV = change of the depth-integrated transports U, V from the 2D
        advection of u1, v1 by U/D and V/D on the advection grids
        (momentum.py:1081-1097)."""
        U.interp(self.uua)
        V.interp(self.uva)
        self.uua.all_values.div_(self.uua.grid.D.all_values)
        self.uva.all_values.div_(self.uva.grid.D.all_values)
        advU.all_values.copy_(u1.all_values)
        self.uadv(self.uua, self.uva, timestep, advU, skip_initial_halo_exchange=True)
        reconstruct_transport_change(advU, self.uadv.D, U, timestep)
        U.interp(self.vua)
        V.interp(self.vva)
        self.vua.all_values.div_(self.vua.grid.D.all_values)
        self.vva.all_values.div_(self.vva.grid.D.all_values)
        advV.all_values.copy_(v1.all_values)
        self.vadv(self.vua, self.vva, timestep, advV, skip_initial_halo_exchange=True)
        reconstruct_transport_change(advV, self.vadv.D, V, timestep)
