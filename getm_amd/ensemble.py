"""Ensembles of independent domains stepped together.

Each domain has its own grid (nx, ny, nz <= 64, mask, metrics, thicknesses),
velocities, w, nuh and 1..T FP32 tracers; one ``step()`` is
TracerCollection.advance (python/pygetm/tracer.py:193-276) of every domain, run
as four launches that cover all domains (``ensemble_step`` in the C-ABI).
Per-domain results do not depend on the other domains or on the launch shape
and equal the single-domain FP32 batched step bit for bit.

``graph_chunk(n)`` captures n steps in a hipGraph (via torch.cuda.CUDAGraph on
the current device) for replay without per-launch host overhead.  ``snapshot``
/ ``restore`` copy the ensemble's state (bound tracers and rounding step
counters) out and back in place; graph_chunk and the autotuner use them so
that their warm-up and timing steps leave no trace in the results.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence

import torch

from . import _lib
from .core import Array, Grid

_ROUNDING = {"nearest": 0, "stochastic": 1}


class Ensemble:
    def __init__(self, grids: Sequence[Grid], ntracers: Sequence[int], scheme: int = 5, cnpar: float = 1.0,
                 ufirst: bool = False):
        if len(grids) != len(ntracers) or not grids:
            raise ValueError("one tracer count per grid is needed")
        self.grids: List[Grid] = list(grids)
        self.ntracers = [int(n) for n in ntracers]
        n = len(grids)
        arr = ctypes.c_void_p * n
        tg = arr(*[g.p for g in grids])
        ug = arr(*[g.ugrid.p for g in grids])
        vg = arr(*[g.vgrid.p for g in grids])
        nt = (ctypes.c_int * n)(*self.ntracers)
        self.p = _lib.load().ensemble_create(n, tg, ug, vg, nt, int(scheme), float(cnpar), int(ufirst))
        if not self.p:
            raise ValueError("ensemble_create failed (scheme, grids, nz <= 64 and ntracers >= 1 are required)")
        self._keep = [None] * n  # tensors bound to the library, kept alive here
        self.tracers: List[Optional[torch.Tensor]] = [None] * n
        self.tuning = (0, 0, 16, 16)
        self.tile_x = 64
        self.pcr_max_nz = 0
        self.rounding = "stochastic"
        self._counters = None
        # the ensemble's own per-domain step counters, held here so that they can be
        # snapshotted with the tracers (a caller's counters replace them)
        dev = self.grids[0].device
        self._own_counters = [torch.zeros(1, dtype=torch.int64, device=dev) for _ in range(n)]
        self.set_rounding("stochastic")

    def __del__(self):
        p = getattr(self, "p", None)
        if p and _lib._lib is not None:
            _lib._lib.ensemble_finalize(p)
            self.p = None

    def __len__(self) -> int:
        return len(self.grids)

    def bind(self, d: int, tracers: torch.Tensor, u: Array, v: Array, w: Array, nuh: Array, timestep: float,
             molecular: float = 0.0) -> None:
        """Fields of domain d; tracers is a (ntracers[d], nz, ny_, nx_) float32 device tensor
        that is stepped in place."""
        g = self.grids[d]
        if tracers.dtype != torch.float32 or not tracers.is_contiguous():
            raise ValueError("ensemble tracers must be contiguous float32")
        if tuple(tracers.shape) != (self.ntracers[d], g.nz_, g.ny_, g.nx_):
            raise ValueError(f"domain {d}: expected {(self.ntracers[d], g.nz_, g.ny_, g.nx_)}, "
                             f"got {tuple(tracers.shape)}")
        rc = _lib.load().ensemble_bind(self.p, d, _lib.ptr(tracers), _lib.ptr(u.all_values), _lib.ptr(v.all_values),
                                       _lib.ptr(w.all_values), _lib.ptr(nuh.all_values), float(timestep),
                                       float(molecular))
        _lib.check(rc, "ensemble_bind")
        self._keep[d] = (tracers, u, v, w, nuh)
        self.tracers[d] = tracers

    def set_tuning(self, tile_y: int = 0, tracers_per_block_h: int = 0, tracers_per_block_w: int = 16,
                   tracers_per_block_d: int = 16) -> None:
        _lib.check(_lib.load().ensemble_set_tuning(self.p, int(tile_y), int(tracers_per_block_h),
                                                   int(tracers_per_block_w), int(tracers_per_block_d)),
                   "ensemble_set_tuning")
        self.tuning = (tile_y, tracers_per_block_h, tracers_per_block_w, tracers_per_block_d)

    def set_tile_x(self, tile_x: int = 64) -> None:
        """Width of the horizontal-stage tiles: 32, 64 or 128 columns (32 with 8 or 16
        rows, 64 with 4, 8 or 16, 128 with 4 or 8); takes effect at commit and does
        not change results."""
        _lib.check(_lib.load().ensemble_set_tile_x(self.p, int(tile_x)), "ensemble_set_tile_x")
        self.tile_x = int(tile_x)

    def set_solver(self, pcr_max_nz: int = 0) -> None:
        """Column solver per domain by depth (TracerCollection.set_solver); takes effect at commit."""
        _lib.check(_lib.load().ensemble_set_solver(self.p, int(pcr_max_nz)), "ensemble_set_solver")
        self.pcr_max_nz = int(pcr_max_nz)

    def set_rounding(self, mode: str = "stochastic", counters: Optional[Sequence[Optional[torch.Tensor]]] = None) -> None:
        """Rounding of the FP32 tracer updates (TracerCollection.set_rounding), per
        domain with its own step counter: ``counters[d]`` is a one-element int64
        device tensor (distinct per domain) or None for the ensemble's own counter
        (zero at creation).  Takes effect at commit."""
        code = _ROUNDING[mode]
        n = len(self.grids)
        if counters is not None:
            if len(counters) != n:
                raise ValueError("one counter (or None) per domain is needed")
            for c in counters:
                if c is not None and (c.dtype != torch.int64 or c.numel() != 1 or not c.is_cuda):
                    raise ValueError("step counters are one-element int64 device tensors")
        used = [c if c is not None else own
                for c, own in zip(counters if counters is not None else [None] * n, self._own_counters)]
        arr = (ctypes.c_void_p * n)(*[c.data_ptr() for c in used])
        _lib.check(_lib.load().ensemble_set_rounding(self.p, code, arr), "ensemble_set_rounding")
        self._counters = used
        self.rounding = mode

    def snapshot(self):
        """Device copies of the state a step changes: every bound tracer tensor and
        every domain's rounding step counter."""
        return ([t.clone() if t is not None else None for t in self.tracers],
                [c.clone() for c in self._counters])

    def restore(self, snap) -> None:
        """Copy a snapshot back into the bound tensors (in place: the library holds
        their addresses)."""
        tracers, counters = snap
        for t, saved in zip(self.tracers, tracers):
            if t is not None:
                t.copy_(saved)
        for c, saved in zip(self._counters, counters):
            c.copy_(saved)

    def commit(self) -> None:
        """Allocate scratch and upload the work tables (synchronous)."""
        _lib.check(_lib.load().ensemble_commit(self.p, _lib.stream_handle()), "ensemble_commit")

    def step(self) -> None:
        _lib.check(_lib.load().ensemble_step(self.p, _lib.stream_handle()), "ensemble_step")

    def graph_chunk(self, nsteps: int) -> "torch.cuda.CUDAGraph":
        """Capture nsteps steps as one hipGraph; replay with ``g.replay()``.  The
        warm-up step that sizes every lazily allocated buffer before the capture
        is undone (tracers and step counters restored), and capturing runs
        nothing, so the state is the one before the call."""
        snap = self.snapshot()
        self.step()  # every lazily sized buffer exists before capture
        self.restore(snap)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            for _ in range(nsteps):
                self.step()
        return g

    @property
    def cells(self) -> int:
        """Wet tracer cells updated per step (wet columns x nz x tracers, summed)."""
        tot = 0
        for g, n in zip(self.grids, self.ntracers):
            tot += int((g.mask.all_values == 1).sum().item()) * g.nz_ * n
        return tot
