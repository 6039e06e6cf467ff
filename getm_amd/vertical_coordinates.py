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
    """Layer fractions (bottom first) of the zoomed sigma grid of
    vertical_coordinates.py:70-90: nz equal layers without zooming, else the
    interfaces g in [0, 1] mapped through tanh zooms of strength ddl (bottom)
    and ddu (surface).  The tanh steps are numpy's, in the reference's order,
    so the fractions are the reference's bit for bit."""
    if max(ddl, 0.0) == 0.0 and max(ddu, 0.0) == 0.0:
        return np.full(nz, 1.0 / nz)
    lo, hi = max(ddl, 0.0), max(ddu, 0.0)
    g = np.linspace(0.0, 1.0, nz + 1)
    g[1:-1] = np.tanh((lo + hi) * g[1:-1] - lo) + np.tanh(lo)
    g[1:-1] /= np.tanh(lo) + np.tanh(hi)
    frac = np.diff(g)
    if np.any(frac <= 0.0):
        raise ValueError(f"zoomed sigma interfaces are not strictly increasing: {g}")
    return frac


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
    """A vertical coordinate scheme: ``initialize`` binds the grids whose layer
    thicknesses it maintains, ``update`` recomputes them (pygetm's interface:
    vertical_coordinates.py:12-66).  A scheme here is a per-column formula
    h_k = f_k(D) applied to every grid from its clipped depth (``Dclip``, or
    ``D`` where a grid has none) at the points its mask marks, which subclasses
    give as ``__call__(D, out, where)``."""

    def __init__(self, nz: int):
        if nz < 1:
            raise ValueError(f"a vertical coordinate needs at least one layer, got nz={nz}")
        self.nz = nz
        self._targets = []

    def initialize(self, *grids, logger=None):
        self.logger = logger
        self._targets = []
        for g in grids:
            depth = g.Dclip if "Dclip" in g._fields else g.D
            self._targets.append((depth.all_values, g.hn.all_values, g.mask.all_values))

    def update(self, timestep: float) -> None:
        for depth, h, mask in self._targets:
            self(depth, h, mask)

    def __call__(self, D: torch.Tensor, out: torch.Tensor, where: torch.Tensor) -> torch.Tensor:
        raise NotImplementedError


# pygetm names the per-grid base PerGrid; every scheme here is per grid
PerGrid = Base


def _device_copy(cache: dict, host: np.ndarray, device) -> torch.Tensor:
    t = cache.get(device)
    if t is None:
        t = cache[device] = torch.from_numpy(host.copy()).to(device)
    return t


class Sigma(Base):
    """Terrain-following layers, each a fixed fraction of the water depth
    (calculate_sigma with bottom/surface zooming ddl/ddu): h = dga D on the
    device (``c_update_sigma``)."""

    def __init__(self, nz: int, *, ddl: float = 0.0, ddu: float = 0.0):
        super().__init__(nz)
        self.dga_host = np.ascontiguousarray(calculate_sigma(nz, ddl, ddu), dtype=np.float64)
        self._dga = {}

    def __call__(self, D: torch.Tensor, out: torch.Tensor, where: torch.Tensor) -> torch.Tensor:
        _check2(D, where, out, self.nz)
        dga = _device_copy(self._dga, self.dga_host, D.device)
        rc = _lib.load().c_update_sigma(D.shape[1], D.shape[0], self.nz, _lib.ptr(dga), _lib.ptr(D),
                                        _lib.ptr(where), _lib.ptr(out), _lib.stream_handle())
        _lib.check(rc, "c_update_sigma")
        return out


class GVC(Base):
    """Generalized vertical coordinates (Burchard & Petersen 1997;
    vertical_coordinates.py:116-194): each layer's share of the depth blends
    the equal share dsigma = 1/nz and the zoomed share dbeta_k,
    h_k = D (a dsigma + (1 - a) dbeta_k), with the blend a chosen per column so
    that the reference layer (the surface layer if ``gamma_surf``, else the
    bottom one) is Dgamma/nz thick; columns shallower than Dgamma get equal
    layers.  The formula runs on the device (``c_update_gvc``)."""

    def __init__(self, nz: int, *, ddl: float = 0.0, ddu: float = 0.0, gamma_surf: bool = True,
                 Dgamma: float = 0.0):
        if max(ddl, 0.0) == 0.0 and max(ddu, 0.0) == 0.0:
            raise ValueError("GVC needs zooming: give ddl and/or ddu > 0")
        if not Dgamma > 0.0:
            raise ValueError(f"GVC needs a positive reference depth Dgamma, got {Dgamma}")
        super().__init__(nz)
        self.dsigma = 1.0 / nz
        self.dbeta = np.ascontiguousarray(calculate_sigma(nz, ddl, ddu), dtype=np.float64)
        self.k_ref = -1 if gamma_surf else 0
        if not self.dbeta[self.k_ref] < self.dsigma:
            # the blend would then keep every column at equal layers: that is Sigma
            raise ValueError("with these zoom factors the reference layer is not thinner than 1/nz, so GVC "
                             "reduces to equal layers everywhere; use Sigma for that")
        self.Dgamma = Dgamma
        self.D_max = self.max_depth()
        self._dbeta = {}

    def max_depth(self) -> float:
        """Deepest water the blend can serve.  As the depth grows past Dgamma the
        blend a falls below 0 (the reference layer keeps Dgamma dsigma while D
        grows), and the share dbeta_k + a (dsigma - dbeta_k) of every layer with
        dbeta_k < dsigma shrinks; it reaches 0 at a = -dbeta_k / (dsigma -
        dbeta_k).  The first layer to vanish fixes the smallest admissible a and,
        through the reference layer's thickness D (a dsigma + (1 - a)
        dbeta_ref) = Dgamma dsigma, the depth; infinite when no share vanishes."""
        gap = self.dsigma - self.dbeta
        vanish = [-b / g for b, g in zip(self.dbeta, gap) if g > 0.0 and -b / g < 0.0]
        if not vanish:
            return float("inf")
        a = max(vanish)
        ref_share = a * self.dsigma + (1.0 - a) * self.dbeta[self.k_ref]
        return float("inf") if abs(ref_share) < 1e-15 else (self.Dgamma * self.dsigma) / ref_share

    def initialize(self, *grids, logger=None):
        super().initialize(*grids, logger=logger)
        if logger is not None:
            logger.info(f"GVC layers serve water depths up to {self.D_max:.3f} m")

    def __call__(self, D: torch.Tensor, out: torch.Tensor, where: torch.Tensor) -> torch.Tensor:
        update_gvc(self.dsigma, _device_copy(self._dbeta, self.dbeta, D.device), self.Dgamma, self.k_ref, D, where,
                   out)
        return out
