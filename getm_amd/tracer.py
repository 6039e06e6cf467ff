"""Tracer transport driver mirroring pygetm.tracer.TracerCollection
(python/pygetm/tracer.py:125-276): split advection of every tracer followed by
implicit vertical diffusion of every tracer.

Tracers live in one stacked device tensor ``(T, nz, ny_, nx_)``.  With
``batched=True`` (default) a step is one ``transport_step`` call of the
C-ABI, which streams all tracers through shared stage coefficients; with
``batched=False`` it is the reference's call sequence of per-tracer operators
(FP64 only).  ``precision="f32"`` stores tracers in FP32 with the thickness
update kept in FP64.

FP32 anomalies (``set_values(..., offsets=...)``): every FP32 kernel updates a
tracer by increments formed from differences of its values (limited face values
shift with the field, the implicit column solve acts on increments), so a step
of c - o equals the step of c minus o up to rounding.  Storing each tracer as
its anomaly about an offset o_t near its smallest value makes the stored
magnitudes, and with them the ulp that every rounding of the stored values
scales with, several times smaller (salinity 30-35 stored about 30: 8x).  The
stored array then holds c - o_t; ``absolute`` adds the offsets back in FP64 and
``inventory_device(absolute=True)`` adds o_t times the water volume.
"""
from __future__ import annotations

import ctypes
import json
from typing import List, Optional

import numpy as np
import torch

from . import _lib, operators
from .core import CENTERS, INTERFACES, Array, Grid

_DTYPES = {"f64": (torch.float64, 1), "f32": (torch.float32, 2)}
_ROUNDING = {"nearest": 0, "stochastic": 1}


class TracerCollection:
    def __init__(self, grid: Grid, advection_scheme=operators.AdvectionScheme.DEFAULT, cnpar: float = 1.0,
                 precision: str = "f64", batched: bool = True, ufirst: bool = False, conserve: bool = False):
        if precision not in _DTYPES:
            raise ValueError(f"precision must be one of {list(_DTYPES)}")
        if precision == "f32" and not batched:
            raise ValueError("the FP32 form is only available as the batched step")
        self.grid = grid
        self.precision = precision
        self.batched = batched
        self.cnpar = float(cnpar)
        self.dtype, code = _DTYPES[precision]
        self._molecular: List[float] = []
        self._names: List[str] = []
        self._extra: List[tuple] = []  # (source, surface_flux, source_scale, vertical_velocity) per tracer
        self.values = torch.empty((0, grid.nz_, grid.ny_, grid.nx_), dtype=self.dtype, device=grid.device)
        self.offsets = np.zeros(0)  # FP32 anomalies: stored value = c - offsets[t]
        self._ones: Optional[torch.Tensor] = None
        self.h = torch.empty((grid.nz_, grid.ny_, grid.nx_), dtype=torch.float64, device=grid.device)
        self.Ah_u: Optional[Array] = None
        self.Ah_v: Optional[Array] = None
        # FP32 only: restore every tracer's inventory to its value before the step
        self.conserve = bool(conserve) and precision == "f32"
        self._target: Optional[torch.Tensor] = None
        self.rounding = "stochastic"
        self._counter: Optional[torch.Tensor] = None
        lib = _lib.load()
        if batched:
            self.p = lib.transport_create(grid.p, grid.ugrid.p, grid.vgrid.p, int(advection_scheme), self.cnpar, code,
                                          int(ufirst))
            if not self.p:
                raise ValueError(f"transport_create failed (scheme {advection_scheme!r})")
            # the collection's own step counter of the stochastic rounding, visible to
            # state_dict (a restart continues the same stream of random bits)
            self._own_counter = torch.zeros(1, dtype=torch.int64, device=grid.device)
            if precision == "f32":
                _lib.check(lib.transport_set_rounding(self.p, _ROUNDING["stochastic"], _lib.ptr(self._own_counter)),
                           "transport_set_rounding")
        else:
            self.p = None
        self._scheme, self._ufirst = advection_scheme, ufirst
        self._advection = None
        if not batched:
            self._sequence_operators()

    def _sequence_operators(self) -> None:
        """Operators of the reference's per-tracer call sequence (made on first use
        by an FP64 batched collection whose tracers have sources, surface fluxes
        or vertical velocities)."""
        if self._advection is not None:
            return
        grid = self.grid
        self._advection = operators.Advection(grid, scheme=self._scheme)
        self._advection.ufirst = self._ufirst
        self._vertical_diffusion = operators.VerticalDiffusion(grid, cnpar=self.cnpar)
        self._wcomb = grid.array(z=INTERFACES, fill=0.0)  # tracer.py:143 _w
        self._source = grid.array(z=CENTERS, fill=0.0)  # tracer.py:140 _source

    def __del__(self):
        p = getattr(self, "p", None)
        if p and _lib._lib is not None:
            _lib._lib.transport_finalize(p)
            self.p = None

    def __len__(self) -> int:
        return self.values.shape[0]

    def __getitem__(self, i: int) -> Array:
        return Array(self.grid, self.values[i], CENTERS, self._names[i])

    def add(self, values=None, molecular_diffusivity: float = 0.0, name: str = "", source: Optional[Array] = None,
            surface_flux: Optional[Array] = None, source_scale: float = 1.0,
            vertical_velocity: Optional[Array] = None) -> Array:
        """Append a tracer (device tensor or array of shape (nz, ny_, nx_)); returns its view.

        ``source`` (layers), ``surface_flux`` (2-D), ``source_scale`` and
        ``vertical_velocity`` (layers; movement through the water, e.g. sinking)
        have the meaning of pygetm.tracer.Tracer's (tracer.py:35-95) and are
        applied by :meth:`advance` as TracerCollection.advance does: in FP64 by the
        per-tracer operator sequence, in FP32 (sources and surface fluxes, cnpar =
        1) by the batched step's column solve."""
        g = self.grid
        if source is not None and (source.grid is not g or source.z != CENTERS):
            raise ValueError("source must be a layer field of the tracer's grid")
        if surface_flux is not None and (surface_flux.grid is not g or surface_flux.z):
            raise ValueError("surface_flux must be a 2-D field of the tracer's grid")
        if vertical_velocity is not None and (vertical_velocity.grid is not g or vertical_velocity.z != CENTERS):
            raise ValueError("vertical_velocity must be a layer field of the tracer's grid")
        new = torch.empty((len(self) + 1, g.nz_, g.ny_, g.nx_), dtype=self.dtype, device=g.device)
        new[:-1].copy_(self.values)
        if values is not None:
            new[-1].copy_(torch.as_tensor(values, device=g.device))
        self.values = new
        self.offsets = np.append(self.offsets, 0.0)
        self._molecular.append(float(molecular_diffusivity))
        self._names.append(name)
        self._extra.append((source, surface_flux, float(source_scale), vertical_velocity))
        return self[len(self) - 1]

    def set_values(self, values: torch.Tensor, molecular_diffusivity: float = 0.0, offsets=None) -> None:
        """Replace all tracers by a stacked (T, nz, ny_, nx_) tensor (copied).

        ``offsets`` (FP32 only) stores each tracer as its anomaly c - o_t: None
        stores c itself; "auto" takes o_t = floor of the tracer's smallest value
        over wet cells (integer offsets keep c - o_t exact for values above o_t);
        a sequence of T numbers gives them."""
        g = self.grid
        if tuple(values.shape[1:]) != (g.nz_, g.ny_, g.nx_):
            raise ValueError(f"expected (T, {g.nz_}, {g.ny_}, {g.nx_}), got {tuple(values.shape)}")
        ntr = values.shape[0]
        if offsets is None:
            off = np.zeros(ntr)
        else:
            if self.precision != "f32":
                raise ValueError("offsets apply to the FP32 form only (the FP64 form is the reference's method)")
            if isinstance(offsets, str):
                if offsets != "auto":
                    raise ValueError('offsets: None, "auto" or one number per tracer')
                wet = (g.mask.all_values == 1).to(g.device)
                off = np.array([float(np.floor(values[t].to(g.device)[..., wet].min().item())) if bool(wet.any())
                                else 0.0 for t in range(ntr)])
            else:
                off = np.asarray(offsets, dtype=np.float64).reshape(-1)
                if off.size != ntr:
                    raise ValueError(f"expected {ntr} offsets, got {off.size}")
        src = values.to(device=g.device)
        if off.any():
            o = torch.as_tensor(off, dtype=torch.float64, device=g.device).view(-1, 1, 1, 1)
            src = src.double() - o
        self.values = src.to(dtype=self.dtype).contiguous().clone()
        self.offsets = off
        self._molecular = [float(molecular_diffusivity)] * values.shape[0]
        self._names = [""] * values.shape[0]
        self._extra = [(None, None, 1.0, None)] * values.shape[0]

    def advance(self, timestep: float, u: Array, v: Array, w: Array, diffusivity: Array) -> None:
        """tracer.py:193-246: advection of all tracers with velocities at t + dt/2,
        then vertical diffusion with dzo = dzn = hn (tracer.py:248-276).  An FP64
        collection with tracer sources, surface fluxes or vertical velocities runs
        the per-tracer sequence (bit-identical to the batched step for the other
        tracers); the FP32 form takes sources and surface fluxes in its column
        solve (cnpar = 1) and rejects per-tracer vertical velocities."""
        vv_extra = any(e[3] is not None for e in self._extra)
        src_extra = any(e[0] is not None or e[1] is not None for e in self._extra)
        if self.precision == "f32":
            if vv_extra:
                raise ValueError("per-tracer vertical velocities need the FP64 form")
            if src_extra and self.cnpar != 1.0:
                raise ValueError("tracer sources in the FP32 form need cnpar = 1")
            if src_extra and self.conserve:
                raise ValueError("the inventory fixer restores pre-step inventories; with sources call "
                                 "transport_conserve with targets that include them")
            sequence = False
        else:
            sequence = vv_extra or src_extra
        if not self.batched or sequence:
            self._sequence_operators()
            tracers = [self[i] for i in range(len(self))]
            extra = dict(zip(map(id, tracers), self._extra))

            def combined_w(tracer: Array) -> Array:
                # tracer.py:223-231: w + the tracer's own vertical velocity at interfaces
                # (interpolation.F90:41-54: 0.5 (c_k + c_{k+1}); end interfaces kept)
                vv = extra[id(tracer)][3]
                if vv is None or not bool(vv.all_values.any()):
                    return w
                wc = self._wcomb.all_values
                wc[1:-1] = 0.5 * (vv.all_values[:-1] + vv.all_values[1:])
                wc += w.all_values
                return self._wcomb

            self._advection.apply_3d_batch(u, v, w, timestep, tracers, Ah_u=self.Ah_u, Ah_v=self.Ah_v,
                                           get_w=combined_w)
            self.h.copy_(self._advection.h)
            avmol = -1.0
            for tracer, mol, (source, flux, scale, _) in zip(tracers, self._molecular, self._extra):
                ea4 = None
                if source is not None or flux is not None:
                    # tracer.py:263-272
                    ea4 = self._source.all_values
                    if source is not None:
                        ea4.copy_(source.all_values)
                    else:
                        ea4.zero_()
                    if flux is not None:
                        ea4[-1] += flux.all_values
                    ea4 *= timestep * scale
                if mol != avmol:
                    avmol = mol
                    self._vertical_diffusion.prepare(diffusivity, timestep, avmol)
                self._vertical_diffusion.apply(tracer, ea4=self._source if ea4 is not None else None)
            return
        mols = set(self._molecular)
        if len(mols) > 1:
            raise ValueError("the batched step needs one molecular diffusivity for all tracers")
        mol = mols.pop() if mols else 0.0
        if self.conserve and len(self):
            # closed domain without sources: sum(area ho c) before = sum(area hn c) after
            self._target = self.inventory_device(self.grid.ho.all_values)
        ea4 = self._fp32_sources(timestep) if src_extra else None
        if ea4 is not None:
            _lib.check(_lib.load().transport_set_sources(self.p, _lib.ptr(ea4)), "transport_set_sources")
        try:
            rc = _lib.load().transport_step(
                self.p, len(self), _lib.ptr(self.values), _lib.ptr(u.all_values), _lib.ptr(v.all_values),
                _lib.ptr(w.all_values), _lib.ptr(diffusivity.all_values),
                _lib.ptr(self.Ah_u.all_values) if self.Ah_u is not None else None,
                _lib.ptr(self.Ah_v.all_values) if self.Ah_v is not None else None, float(timestep), mol,
                _lib.ptr(self.h), _lib.stream_handle())
            _lib.check(rc, "transport_step")
        finally:
            if ea4 is not None:
                _lib.load().transport_set_sources(self.p, None)
        if self.conserve and len(self):
            _lib.check(_lib.load().transport_conserve(self.p, len(self), _lib.ptr(self.values), _lib.ptr(self._target),
                                                      _lib.stream_handle()), "transport_conserve")

    def _fp32_sources(self, timestep: float) -> torch.Tensor:
        """The ea4 terms of every tracer stacked (FP64; zero for tracers without
        sources), formed as the reference's _diffuse does (tracer.py:263-272)."""
        g = self.grid
        shape = (len(self), g.nz_, g.ny_, g.nx_)
        if getattr(self, "_ea4", None) is None or tuple(self._ea4.shape) != shape:
            self._ea4 = torch.zeros(shape, dtype=torch.float64, device=g.device)
        for t, (source, flux, scale, _) in enumerate(self._extra):
            e = self._ea4[t]
            if source is None and flux is None:
                e.zero_()
                continue
            if source is not None:
                e.copy_(source.all_values)
            else:
                e.zero_()
            if flux is not None:
                e[-1] += flux.all_values
            e *= timestep * scale
        return self._ea4

    def inventory_device(self, h: Optional[torch.Tensor] = None, absolute: bool = False) -> torch.Tensor:
        """Per-tracer sum of area * h * c over wet cells (FP64, compensated,
        deterministic) as a device tensor; h defaults to the T grid's hn.  Of the
        stored values; with ``absolute`` and anomalies stored, of c itself (the
        stored inventory plus o_t times the water volume sum(area * h))."""
        if not self.batched:
            raise ValueError("inventories are computed by the batched transport object")
        out = torch.empty(len(self), dtype=torch.float64, device=self.grid.device)
        _lib.check(_lib.load().transport_inventory(self.p, len(self), _lib.ptr(self.values),
                                                   _lib.ptr(h) if h is not None else None, _lib.ptr(out),
                                                   _lib.stream_handle()), "transport_inventory")
        if absolute and self.offsets.any():
            g = self.grid
            if self._ones is None:
                self._ones = torch.ones((1, g.nz_, g.ny_, g.nx_), dtype=self.dtype, device=g.device)
            vol = torch.empty(1, dtype=torch.float64, device=g.device)
            _lib.check(_lib.load().transport_inventory(self.p, 1, _lib.ptr(self._ones),
                                                       _lib.ptr(h) if h is not None else None, _lib.ptr(vol),
                                                       _lib.stream_handle()), "transport_inventory")
            out = out + torch.as_tensor(self.offsets, dtype=torch.float64, device=g.device) * vol
        return out

    def absolute(self, i: Optional[int] = None) -> torch.Tensor:
        """Tracer values c in FP64 (stored value plus offset): tracer i, or all
        stacked."""
        if i is None:
            o = torch.as_tensor(self.offsets, dtype=torch.float64, device=self.grid.device).view(-1, 1, 1, 1)
            return self.values.double() + o
        return self.values[i].double() + float(self.offsets[i])

    def inventory(self, h: Optional[torch.Tensor] = None):
        """Host copy of inventory_device (simulation.py:1408-1466 report these)."""
        return self.inventory_device(h).cpu().numpy()

    def set_tuning(self, tracers_per_block_h: int = 0, tracers_per_block_w: int = 16,
                   tracers_per_block_d: int = 16, tile_y: int = 0) -> None:
        """Launch shape of the batched kernels: tracers per block of the three kernel
        kinds (0 = all) and the height of the horizontal-stage tiles (4, 8, 16; 0 = the
        default, 8 rows).  Does not change results."""
        if not self.batched:
            return
        _lib.check(_lib.load().transport_set_tuning(self.p, int(tile_y), int(tracers_per_block_h),
                                                    int(tracers_per_block_w), int(tracers_per_block_d)),
                   "transport_set_tuning")

    def set_solver(self, pcr_max_nz: int = 0) -> None:
        """FP32 column solve: wave-level parallel cyclic reduction for columns of
        2 <= nz <= pcr_max_nz (at most 32), per-lane Thomas elimination otherwise
        (0: Thomas only).  The two agree to FP32 rounding."""
        if not self.batched:
            return
        _lib.check(_lib.load().transport_set_solver(self.p, int(pcr_max_nz)), "transport_set_solver")
        self.pcr_max_nz = int(pcr_max_nz)

    def set_rounding(self, mode: str = "stochastic", counter: Optional[torch.Tensor] = None) -> None:
        """Rounding of the FP32 tracer updates: "stochastic" (the default) rounds
        each stored c + dc up or down with probability given by the exact
        residual, so rounding errors do not accumulate in a steady flow;
        "nearest" is plain FP32 arithmetic.  The random bits are a hash of (step
        count, stage, tracer, cell) -- deterministic, independent of launch shape
        and batching.  ``counter`` (a one-element int64 device tensor) holds the
        step count and is advanced by every step; None uses the collection's own
        counter (zero at creation, saved by state_dict).  No effect on FP64."""
        if not self.batched:
            return
        if counter is not None and (counter.dtype != torch.int64 or counter.numel() != 1 or not counter.is_cuda):
            raise ValueError("the step counter is a one-element int64 device tensor")
        ctr = counter if counter is not None else self._own_counter
        rc = _lib.load().transport_set_rounding(self.p, _ROUNDING[mode], _lib.ptr(ctr))
        _lib.check(rc, "transport_set_rounding")
        self._counter = counter
        self.rounding = mode

    # ---- restart: snapshot and restore of the state the step depends on ----
    def state_dict(self) -> dict:
        """Host copy of the collection's state: tracer values (in their storage
        precision), per-tracer names and molecular diffusivities, and for the FP32
        form the rounding mode and step count.  A collection restored from it
        continues bit for bit as the original would (the grid fields and the
        velocities are inputs of every step, not state of the collection)."""
        torch.cuda.synchronize(self.grid.device)
        ctr = self._counter if self._counter is not None else getattr(self, "_own_counter", None)
        return {"values": self.values.detach().cpu().numpy().copy(), "precision": self.precision,
                "names": list(self._names), "molecular": list(self._molecular),
                "rounding": self.rounding, "step": int(ctr.item()) if ctr is not None else 0,
                "offsets": [float(o) for o in self.offsets]}

    def load_state_dict(self, state: dict) -> None:
        """Restore a state_dict; the collection must have the same grid shape and
        precision.  Tracer sources and vertical velocities are not part of it
        (re-add them with ``add`` semantics if the run uses them)."""
        if state["precision"] != self.precision:
            raise ValueError(f"state is {state['precision']}, collection is {self.precision}")
        self.set_values(torch.from_numpy(np.ascontiguousarray(state["values"])))
        self.offsets = np.asarray(state.get("offsets", np.zeros(len(self))), dtype=np.float64)
        self._names = list(state["names"])
        self._molecular = [float(m) for m in state["molecular"]]
        if self.batched and self.precision == "f32":
            self.set_rounding(state["rounding"], self._counter)
            ctr = self._counter if self._counter is not None else self._own_counter
            ctr.fill_(int(state["step"]))

    def save(self, path: str) -> None:
        """Write state_dict to an .npz file (arrays and a JSON header; no pickles)."""
        st = self.state_dict()
        meta = {k: v for k, v in st.items() if k != "values"}
        np.savez(path, values=st["values"], meta=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8))

    def load(self, path: str) -> None:
        """Restore a file written by ``save``."""
        with np.load(path, allow_pickle=False) as f:
            st = json.loads(bytes(f["meta"]).decode())
            st["values"] = f["values"]
        self.load_state_dict(st)

    def stats(self, reset: bool = True):
        """(largest z sub-iteration count, columns capped at 200) since the last reset."""
        if not self.batched:
            return self._advection.stats(reset)
        mi, cap = ctypes.c_int(0), ctypes.c_int(0)
        _lib.check(_lib.load().transport_get_stats(self.p, ctypes.byref(mi), ctypes.byref(cap), int(reset),
                                                   _lib.stream_handle()), "transport_get_stats")
        return mi.value, cap.value
