"""Input producers of the transport step that live in pygetm's momentum module.

``w_momentum_3d`` is pygetm._pygetm.w_momentum_3d (python/_pygetm.pyx:443-455):
vertical velocity at layer interfaces from the layer transports and the
thickness change ho -> hn of the T grid, on the device (C-ABI
``c_w_momentum_3d``, bit-identical to python/src/momentum.F90:191-216).
``velocities_from_transports`` is the division that follows it in
Momentum.update_diagnostics (python/pygetm/momentum.py:885-892).  Keeping both
on the GPU spares coupled runs the per-step host -> device copy of u, v and w.
"""
from __future__ import annotations

import torch

from . import _lib
from .core import CENTERS, INTERFACES, Array


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
    """uk = pk / hn (U grid), vk = qk / hn (V grid), momentum.py:885-892."""
    torch.div(pk.all_values, pk.grid.hn.all_values, out=uk.all_values)
    torch.div(qk.all_values, qk.grid.hn.all_values, out=vk.all_values)
