"""Vertical coordinates: layer thicknesses hn from the water depth, on the device.

Mirrors pygetm.vertical_coordinates (python/pygetm/vertical_coordinates.py):
``calculate_sigma`` (the zoomed sigma fractions, a set-up computation on nz
values done with numpy exactly as the reference does), ``Sigma`` (h = dga D,
C-ABI ``c_update_sigma``) and ``GVC`` (generalized vertical coordinates, C-ABI
``c_update_gvc``, bit-identical to python/src/wrapper.F90:567-601).  ``update``
refreshes hn of every grid given to ``initialize`` from its clipped depth where
the grid's mask is non-zero, as PerGrid.update does (vertical_coordinates.py:33-48).
The reference's ``Adaptive`` class is a placeholder there and is not mirrored.
"""
from __future__ import annotations

from typing import Optional

import numpy as np
import torch

from . import _lib


def calculate_sigma(nz: int, ddl: float = 0.0, ddu: float = 0.0) -> np.ndarray:
    """Sigma fractions of the nz layers, bottom first, zoomed towards the bottom
    (ddl) and surface (ddu) (vertical_coordinates.py:69-90)."""
    if ddl <= 0.0 and ddu <= 0.0:
        return np.broadcast_to(1.0 / nz, (nz,))
    ddl, ddu = max(ddl, 0.0), max(ddu, 0.0)
    ga = np.linspace(0.0, 1.0, nz + 1)
    ga[1:-1] = np.tanh((ddl + ddu) * ga[1:-1] - ddl) + np.tanh(ddl)
    ga[1:-1] /= np.tanh(ddl) + np.tanh(ddu)
    dga = np.diff(ga)
    if not (dga > 0.0).all():
        raise ValueError(f"ga not monotonically increasing: {ga}")
    return dga


def _check2(D: torch.Tensor, mask: torch.Tensor, h: torch.Tensor, nz: int):
    if D.dim() != 2 or tuple(mask.shape) != tuple(D.shape) or tuple(h.shape) != (nz,) + tuple(D.shape):
        raise ValueError("expected D and mask (ny, nx) and h (nz, ny, nx)")
    for name, t, dt in (("D", D, torch.float64), ("mask", mask, torch.int32), ("h", h, torch.float64)):
        if t.dtype != dt or not t.is_contiguous() or not t.is_cuda:
            raise ValueError(f"{name}: expected a contiguous {dt} device tensor")


def update_gvc(dsigma: float, dbeta: torch.Tensor, Dgamma: float, kk: int, D: torch.Tensor, mask: torch.Tensor,
               h: torch.Tensor) -> None:
    """pygetm._pygetm.update_gvc (python/_pygetm.pyx:577-585): h from D where
    mask != 0; kk is the reference layer, negative counting from the top."""
    nz = dbeta.numel()
    _check2(D, mask, h, nz)
    if dbeta.dtype != torch.float64 or not dbeta.is_cuda or not dbeta.is_contiguous():
        raise ValueError("dbeta: expected a contiguous float64 device tensor")
    if kk < 0:
        kk = nz + kk
    rc = _lib.load().c_update_gvc(D.shape[1], D.shape[0], nz, float(dsigma), _lib.ptr(dbeta), float(Dgamma), 1 + kk,
                                  _lib.ptr(D), _lib.ptr(mask), _lib.ptr(h), _lib.stream_handle())
    _lib.check(rc, "c_update_gvc")


class Base:
    """Updates the layer thicknesses ``hn`` of the grids given to ``initialize``."""

    def __init__(self, nz: int):
        if nz <= 0:
            raise ValueError("Number of layers nz must be a positive number")
        self.nz = nz

    def initialize(self, *grids, logger=None):
        self.logger = logger

    def update(self, timestep: float) -> None:
        raise NotImplementedError


class PerGrid(Base):
    """The same operation on every grid, from its clipped depth ``Dclip`` (the
    depth ``D`` on grids without one) into its ``hn`` where its mask is non-zero."""

    def initialize(self, *grids, logger=None):
        super().initialize(*grids, logger=logger)
        self.grid_info = [self.prepare_update_args(g) for g in grids]

    def prepare_update_args(self, grid):
        D = grid.Dclip if "Dclip" in grid._fields else grid.D
        return (D.all_values, grid.hn.all_values, grid.mask.all_values)

    def update(self, timestep: float) -> None:
        for gi in self.grid_info:
            self(*gi)

    def __call__(self, D: torch.Tensor, out: torch.Tensor, where: torch.Tensor) -> torch.Tensor:
        raise NotImplementedError


class Sigma(PerGrid):
    """Sigma coordinates with optional zooming towards bottom and surface."""

    def __init__(self, nz: int, *, ddl: float = 0.0, ddu: float = 0.0):
        super().__init__(nz)
        self.dga_host = np.ascontiguousarray(calculate_sigma(nz, ddl, ddu), dtype=np.float64)
        self._dga = {}

    def _coeffs(self, device) -> torch.Tensor:
        t = self._dga.get(device)
        if t is None:
            t = self._dga[device] = torch.from_numpy(self.dga_host.copy()).to(device)
        return t

    def __call__(self, D: torch.Tensor, out: torch.Tensor, where: torch.Tensor) -> torch.Tensor:
        _check2(D, where, out, self.nz)
        rc = _lib.load().c_update_sigma(D.shape[1], D.shape[0], self.nz, _lib.ptr(self._coeffs(D.device)),
                                        _lib.ptr(D), _lib.ptr(where), _lib.ptr(out), _lib.stream_handle())
        _lib.check(rc, "c_update_sigma")
        return out


class GVC(PerGrid):
    """Generalized vertical coordinates (Burchard & Petersen 1997): the reference
    layer (surface if ``gamma_surf``, else bottom) keeps the thickness Dgamma/nz
    except in water shallower than Dgamma, where all layers are equal
    (vertical_coordinates.py:116-187)."""

    def __init__(self, nz: int, *, ddl: float = 0.0, ddu: float = 0.0, gamma_surf: bool = True,
                 Dgamma: float = 0.0):
        if ddl <= 0.0 and ddu <= 0.0:
            raise ValueError("ddl and/or ddu must be a positive number")
        if Dgamma <= 0.0:
            raise ValueError("Dgamma must be a positive number")
        super().__init__(nz)
        self.dsigma = 1.0 / nz
        self.dbeta = np.ascontiguousarray(calculate_sigma(nz, ddl, ddu), dtype=np.float64)
        self.k_ref = -1 if gamma_surf else 0
        if self.dbeta[self.k_ref] >= self.dsigma:
            raise ValueError("This GVC parameterization would always result in equidistant layers."
                             " If this is desired, use Sigma instead.")
        self.Dgamma = Dgamma
        alpha_lim = -self.dbeta / (self.dsigma - self.dbeta)
        alpha_min = alpha_lim[alpha_lim < 0.0].max()
        denom = alpha_min * self.dsigma + (1.0 - alpha_min) * self.dbeta[self.k_ref]
        self.D_max = np.inf if abs(denom) < 1e-15 else (Dgamma * self.dsigma) / denom
        self._dbeta = {}

    def __call__(self, D: torch.Tensor, out: torch.Tensor, where: torch.Tensor) -> torch.Tensor:
        t = self._dbeta.get(D.device)
        if t is None:
            t = self._dbeta[D.device] = torch.from_numpy(self.dbeta.copy()).to(D.device)
        update_gvc(self.dsigma, t, self.Dgamma, self.k_ref, D, where, out)
        return out
