"""Tracer transport driver mirroring pygetm.tracer (python/pygetm/tracer.py:22-276):
``Tracer`` fields with their sources, surface fluxes, vertical velocities and
molecular diffusivities, and ``TracerCollection``, whose ``advance`` is split
advection of every tracer followed by implicit vertical diffusion of every
tracer.

Tracers live in one stacked device tensor ``values`` of shape
``(T, nz, ny_, nx_)``.  With ``batched=True`` (default) a step is one call of
the C-ABI's batched step, which streams all tracers through shared stage
coefficients; with ``batched=False`` it is the reference's call sequence of
per-tracer operators (FP64 only).

Precisions:

- ``"f64"``: the reference's method, bit-identical to it.
- ``"f32"``: the FP32 form with compensated storage.  Each tracer value is an
  FP32 number (``values``) plus an int16 residual (``residuals``) that extends
  it by 15 bits.  The limiter and the tridiagonal elimination run in FP32 on
  differences of the decoded values; fluxes, cell updates and the thickness
  run in FP64, and every launch rounds its results once, to 2^-16 of an FP32
  ulp.  Neither rounding errors nor inventory defects accumulate.
  ``absolute()`` returns the decoded values.
- ``"f32-sr"``: plain FP32 storage, every store rounded stochastically.  This
  is the form of the ensembles and the serving loop.

FP32 anomalies (``set_values(..., offsets=...)``): every FP32 update is formed
from differences of a tracer's values, so a step of c - o equals the step of
c minus o up to rounding.  Storing a tracer as its anomaly about an offset o_t
near its smallest value makes the stored magnitudes smaller, and with them
the ulp that every rounding scales with.  ``absolute`` adds the offsets back
in FP64, and ``inventory_device(absolute=True)`` adds o_t times the water
volume.
"""
from __future__ import annotations

import ctypes
import json
import logging
from typing import List, Optional

import numpy as np
import torch

from . import _lib, operators
from .core import CENTERS, INTERFACES, Array, Grid

_PRECISIONS = {"f64": (torch.float64, _lib.GETM_F64), "f32": (torch.float32, _lib.GETM_F32C),
               "f32-sr": (torch.float32, _lib.GETM_F32)}
_ROUNDING = {"nearest": 0, "stochastic": 1}


def residual_unit(hi: torch.Tensor) -> torch.Tensor:
    """2^(e - 38) for the exponent e of each FP32 value (e = -126 for zero and
    subnormals): the unit of the int16 residuals of compensated storage."""
    eb = ((hi.view(torch.int32) >> 23) & 0xFF).clamp_(min=1)
    return torch.ldexp(torch.ones(hi.shape, dtype=torch.float64, device=hi.device), eb - 165)


def decode(hi: torch.Tensor, lo: torch.Tensor) -> torch.Tensor:
    """FP64 values of compensated FP32 storage: hi + lo 2^(e - 38), exact."""
    return hi.double() + lo.double() * residual_unit(hi)


def encode(x: torch.Tensor, hi: torch.Tensor, lo: torch.Tensor) -> None:
    """Store FP64 values as hi = (float)x (rounded to nearest) and
    lo = rint((x - hi) / 2^(e - 38)), the kernels' encoding."""
    h = x.to(torch.float32)
    r = torch.nan_to_num(x - h.double(), nan=0.0, posinf=0.0, neginf=0.0)
    hi.copy_(h)
    lo.copy_(torch.round(r / residual_unit(h)).clamp_(-32768, 32767).to(torch.int16))


def encode_host(x: np.ndarray):
    """The same encoding on the host (numpy): returns (hi float32, lo int16).
    FP32 input is exact in hi, so its residuals are zero."""
    x = np.asarray(x)
    if x.dtype == np.float32:
        return x, np.zeros(x.shape, dtype=np.int16)
    x = x.astype(np.float64, copy=False)
    hi = x.astype(np.float32)
    r = np.nan_to_num(x - hi.astype(np.float64), nan=0.0, posinf=0.0, neginf=0.0)
    eb = np.maximum((hi.view(np.int32) >> 23) & 0xFF, 1)
    lo = np.clip(np.round(r / np.ldexp(1.0, eb - 165)), -32768, 32767).astype(np.int16)
    return hi, lo


class _Storage:
    """The collection's tracer arrays, shared with its Tracer objects (which hold
    this, not the collection, so that no reference cycle keeps a collection and
    its device buffers alive after its last user lets go)."""

    __slots__ = ("values", "residuals")

    def __init__(self):
        self.values = None
        self.residuals = None

    def put(self, index, x) -> None:
        """Write values (any float dtype; a device tensor, or host data as a CPU
        tensor or array) into tracer slot ``index`` (an int) or every slot (a
        slice, x stacked).  Compensated storage keeps what FP32 cannot hold in
        the residuals, encoding one tracer at a time so that the FP64
        temporaries stay the size of one tracer.  Host data are converted and
        encoded on the host and then copied, so no device kernel runs."""
        comp = self.residuals is not None
        one = isinstance(index, int)
        slots = [index] if one else range(self.values.shape[0])
        if isinstance(x, np.ndarray) or not isinstance(x, torch.Tensor) or x.device.type == "cpu":
            x = x.numpy() if isinstance(x, torch.Tensor) else np.asarray(x)
            for t in slots:
                xt = x if one else x[t]
                if comp:
                    hi, lo = encode_host(xt)
                    self.values[t].copy_(torch.from_numpy(np.ascontiguousarray(hi)))
                    self.residuals[t].copy_(torch.from_numpy(np.ascontiguousarray(lo)))
                else:
                    dt = np.float64 if self.values.dtype == torch.float64 else np.float32
                    self.values[t].copy_(torch.from_numpy(np.ascontiguousarray(xt, dtype=dt)))
        elif not comp:
            self.values[index].copy_(x)
        else:
            dev = self.values.device
            for t in slots:
                xt = x if one else x[t]
                encode(xt.to(device=dev, dtype=torch.float64), self.values[t], self.residuals[t])


class Tracer(Array):
    """A transported tracer, pygetm.tracer.Tracer (tracer.py:22-123): its field
    (the collection's ``values[index]``) and the attributes that
    ``TracerCollection.advance`` reads at every step -- ``source`` and
    ``surface_flux`` (multiplied by ``source_scale``, in tracer units per time
    times layer thickness), ``vertical_velocity`` (movement through the water,
    positive upwards) and ``molecular_diffusivity``.  Rivers, open boundaries
    and precipitation are outside the transport step and not modelled."""

    __slots__ = ("source", "surface_flux", "source_scale", "vertical_velocity", "molecular_diffusivity",
                 "_store", "_index", "attrs")

    def __init__(self, collection: "TracerCollection", index: int, name: str = "", source: Optional[Array] = None,
                 surface_flux: Optional[Array] = None, source_scale: float = 1.0,
                 vertical_velocity: Optional[Array] = None, molecular_diffusivity: float = 0.0, **kwargs):
        g = collection.grid
        if source is not None and (source.grid is not g or source.z != CENTERS):
            raise ValueError("source must be a layer field of the tracer's grid")
        if surface_flux is not None and (surface_flux.grid is not g or surface_flux.z):
            raise ValueError("surface_flux must be a 2-D field of the tracer's grid")
        if vertical_velocity is not None and (vertical_velocity.grid is not g or vertical_velocity.z != CENTERS):
            raise ValueError("vertical_velocity must be a layer field of the tracer's grid")
        self.grid = g
        self.z = CENTERS
        self.name = name
        self._store = collection._storage
        self._index = index
        self.source = source
        self.surface_flux = surface_flux
        self.source_scale = float(source_scale)
        self.vertical_velocity = vertical_velocity
        self.molecular_diffusivity = float(molecular_diffusivity)
        # units, long_name, fill_value and the river/precipitation switches of the
        # reference's Tracer are kept as metadata only
        self.attrs = dict(kwargs)

    @property
    def all_values(self) -> torch.Tensor:
        return self._store.values[self._index]

    @all_values.setter
    def all_values(self, value) -> None:
        # compensated storage: the residuals are encoded with the values
        self._store.put(self._index, value)


class TracerCollection:
    """pygetm.tracer.TracerCollection (tracer.py:125-276) on the GPU.

    ``logger`` is optional; ``precision``, ``batched``, ``ufirst`` and
    ``conserve`` select the implementation (see the module docstring).
    ``conserve=True`` (FP32 forms) restores every tracer's inventory after each
    step to its value before the step (sum of area * ho * c before, area * hn * c
    after), so the collection must be closed and without sources then."""

    def __init__(self, grid: Grid, logger: Optional[logging.Logger] = None,
                 advection_scheme=operators.AdvectionScheme.DEFAULT, cnpar: float = 1.0, precision: str = "f64",
                 batched: bool = True, ufirst: bool = False, conserve: bool = False):
        if precision not in _PRECISIONS:
            raise ValueError(f"precision must be one of {list(_PRECISIONS)}")
        if precision != "f64" and not batched:
            raise ValueError("the FP32 forms are only available as the batched step")
        self.grid = grid
        self.logger = logger or logging.getLogger("getm_amd.tracer")
        self.precision = precision
        self.compensated = precision == "f32"
        self.batched = batched
        self.cnpar = float(cnpar)
        self.dtype, code = _PRECISIONS[precision]
        self._storage = _Storage()
        self._tracers: List[Tracer] = []
        self._alloc(0)
        self.offsets = np.zeros(0)  # FP32 anomalies: stored value = c - offsets[t]
        self._ones: Optional[torch.Tensor] = None
        self.h = torch.empty((grid.nz_, grid.ny_, grid.nx_), dtype=torch.float64, device=grid.device)
        self.Ah_u: Optional[Array] = None
        self.Ah_v: Optional[Array] = None
        self.conserve = bool(conserve) and precision != "f64"
        self._target: Optional[torch.Tensor] = None
        self.rounding = "stochastic" if precision == "f32-sr" else "nearest"
        self._counter: Optional[torch.Tensor] = None
        self._wcombs: List[Array] = []
        lib = _lib.load()
        if batched:
            self.p = lib.transport_create(grid.p, grid.ugrid.p, grid.vgrid.p, int(advection_scheme), self.cnpar, code,
                                          int(ufirst))
            if not self.p:
                raise ValueError(f"transport_create failed (scheme {advection_scheme!r})")
            # the collection's own step counter of the stochastic rounding, visible to
            # state_dict (a restart continues the same stream of random bits)
            self._own_counter = torch.zeros(1, dtype=torch.int64, device=grid.device)
            if precision == "f32-sr":
                _lib.check(lib.transport_set_rounding(self.p, _ROUNDING["stochastic"], _lib.ptr(self._own_counter)),
                           "transport_set_rounding")
        else:
            self.p = None
        self._scheme, self._ufirst = advection_scheme, ufirst
        self._advection = None
        if not batched:
            self._sequence_operators()

    # ---- storage ----------------------------------------------------------
    def _alloc(self, ntr: int) -> None:
        """Uninitialised storage for ntr tracers: ``values`` and, compensated,
        ``residuals`` as views of one buffer (the FP32 planes followed by the
        int16 planes, the layout the C-ABI takes)."""
        g = self.grid
        shape = (ntr, g.nz_, g.ny_, g.nx_)
        if self.compensated:
            n = ntr * g.nz_ * g.ny_ * g.nx_
            self._buf = torch.empty(3 * n, dtype=torch.int16, device=g.device)
            self.values = self._buf[:2 * n].view(torch.float32).view(shape)
            self.residuals = self._buf[2 * n:].view(shape)
        else:
            self.values = torch.empty(shape, dtype=self.dtype, device=g.device)
            self._buf = self.values
            self.residuals = None
        self._storage.values, self._storage.residuals = self.values, self.residuals

    def _store(self, index, x) -> None:
        """Write values into tracer slot ``index`` (an int) or every slot (a
        slice, x stacked); see _Storage.put."""
        self._storage.put(index, x)

    def _sequence_operators(self) -> None:
        """Operators of the reference's per-tracer call sequence (made on first use
        by an FP64 batched collection whose tracers have sources or surface fluxes)."""
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
        return len(self._tracers)

    def __getitem__(self, i: int) -> Tracer:
        return self._tracers[i]

    def __iter__(self):
        return iter(self._tracers)

    @property
    def names(self) -> List[str]:
        return [t.name for t in self._tracers]

    def add(self, name: str = "", data=None, **kwargs) -> Tracer:
        """Add a tracer (tracer.py:179-190): ``data`` (values of shape
        (nz, ny_, nx_), device tensor or array; NaN if not given, as the
        reference) and the keyword arguments of :class:`Tracer` (``source``,
        ``surface_flux``, ``source_scale``, ``vertical_velocity``,
        ``molecular_diffusivity``, and metadata).  Returns the tracer."""
        g = self.grid
        old_v, old_r, n = self.values, self.residuals, len(self)
        tracer = Tracer(self, n, name=name, **kwargs)
        self._alloc(n + 1)
        self.values[:n].copy_(old_v)
        if self.compensated:
            self.residuals[:n].copy_(old_r)
        x = torch.full((g.nz_, g.ny_, g.nx_), float("nan"), dtype=torch.float64, device=g.device) if data is None \
            else data
        self._store(n, x)
        self.offsets = np.append(self.offsets, 0.0)
        self._tracers.append(tracer)
        return tracer

    def set_values(self, values: torch.Tensor, molecular_diffusivity: Optional[float] = None, offsets=None) -> None:
        """Replace the values of all tracers by a stacked (T, nz, ny_, nx_) tensor
        (copied; FP64 input keeps its precision in the compensated form).  With
        as many tracers as the collection holds, the tracers keep their names,
        sources, vertical velocities and molecular diffusivities; otherwise the
        collection is rebuilt with plain tracers.  ``molecular_diffusivity``
        given sets it for every tracer.

        ``offsets`` (FP32 forms only) stores each tracer as its anomaly c - o_t:
        None stores c itself; "auto" takes o_t = floor of the tracer's smallest
        value over wet cells (integer offsets keep c - o_t exact for values above
        o_t); a sequence of T numbers gives them."""
        g = self.grid
        if tuple(values.shape[1:]) != (g.nz_, g.ny_, g.nx_):
            raise ValueError(f"expected (T, {g.nz_}, {g.ny_}, {g.nx_}), got {tuple(values.shape)}")
        ntr = values.shape[0]
        if offsets is None:
            off = np.zeros(ntr)
        else:
            if self.precision == "f64":
                raise ValueError("offsets apply to the FP32 forms only (the FP64 form is the reference's method)")
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
        # host data stay on the host until stored (converted and encoded there)
        src = values if isinstance(values, torch.Tensor) else torch.as_tensor(np.asarray(values))
        if off.any():
            o = torch.as_tensor(off, dtype=torch.float64, device=src.device).view(-1, 1, 1, 1)
            src = src.double() - o
        if ntr != len(self):
            self._tracers = [Tracer(self, t, name=f"tracer{t}") for t in range(ntr)]
            self._alloc(ntr)
        self._store(slice(None), src)
        self.offsets = off
        if molecular_diffusivity is not None:
            for t in self._tracers:
                t.molecular_diffusivity = float(molecular_diffusivity)

    def resize(self, ntr: int) -> None:
        """Hold ntr tracers: the first min(ntr, len) keep their values, names and
        attributes, new ones are plain tracers whose values are undefined until
        set (``set_tracer``).  Lets a caller fill a large collection one tracer
        at a time, e.g. from host data, without a stacked copy of all of them."""
        n = min(ntr, len(self))
        old_v, old_r = self.values, self.residuals
        self._alloc(ntr)
        if n:
            self.values[:n].copy_(old_v[:n])
            if self.compensated:
                self.residuals[:n].copy_(old_r[:n])
        self._tracers = self._tracers[:n] + [Tracer(self, t, name=f"tracer{t}") for t in range(n, ntr)]
        self.offsets = np.concatenate([self.offsets[:n], np.zeros(ntr - n)])

    def set_tracer(self, index: int, data) -> None:
        """Store the values (nz, ny_, nx_) of tracer ``index``: a device tensor,
        or host data (CPU tensor or array), which are converted and encoded on
        the host and copied in."""
        if not 0 <= index < len(self):
            raise IndexError(f"tracer {index} of {len(self)}")
        g = self.grid
        if tuple(data.shape) != (g.nz_, g.ny_, g.nx_):
            raise ValueError(f"expected ({g.nz_}, {g.ny_}, {g.nx_}), got {tuple(data.shape)}")
        self._store(int(index), data)

    # ---- the step ---------------------------------------------------------
    def _combined_w(self, w: Array) -> List[Optional[torch.Tensor]]:
        """Per tracer, w + its vertical velocity at interfaces (tracer.py:223-231,
        interpolation.F90:41-54: 0.5 (c_k + c_{k+1}), end interfaces of w kept),
        or None where a tracer has none or it is zero everywhere."""
        out: List[Optional[torch.Tensor]] = []
        k = 0
        for tr in self._tracers:
            vv = tr.vertical_velocity
            if vv is None or not bool(vv.all_values.any()):
                out.append(None)
                continue
            if k == len(self._wcombs):
                self._wcombs.append(self.grid.array(z=INTERFACES, fill=0.0))
            wc = self._wcombs[k].all_values
            k += 1
            wc.zero_()
            wc[1:-1] = 0.5 * (vv.all_values[:-1] + vv.all_values[1:])
            wc += w.all_values
            out.append(wc)
        return out

    def advance(self, timestep: float, u: Array, v: Array, w: Array, diffusivity: Array) -> None:
        """tracer.py:193-246: advection of all tracers with velocities at t + dt/2,
        then vertical diffusion with dzo = dzn = hn (tracer.py:248-276), each
        tracer with its own vertical velocity and molecular diffusivity.  The
        FP64 batched step runs the per-tracer sequence when a tracer has sources
        or a surface flux (bit-identical to the batched step for the others); the
        FP32 forms take them in their column solve."""
        src_extra = any(t.source is not None or t.surface_flux is not None for t in self._tracers)
        if self.precision != "f64":
            if src_extra and self.conserve:
                raise ValueError("the inventory fixer restores the inventory the tracers had; with sources call "
                                 "transport_conserve with targets that include them")
            sequence = False
        else:
            sequence = src_extra
        if not self.batched or sequence:
            self._advance_sequence(timestep, u, v, w, diffusivity)
            return
        if not len(self):
            return
        if self.conserve:
            # closed domain without sources: sum(area ho c) before = sum(area hn c) after
            self._target = self.inventory_device(self.grid.ho.all_values)
        ea4 = self._fp32_sources(timestep) if src_extra else None
        lib = _lib.load()
        if ea4 is not None:
            _lib.check(lib.transport_set_sources(self.p, _lib.ptr(ea4)), "transport_set_sources")
        wv = self._combined_w(w)
        ntr = len(self)
        wv_arr = (ctypes.c_void_p * ntr)(*[t.data_ptr() if t is not None else None for t in wv])
        mol_arr = (ctypes.c_double * ntr)(*[t.molecular_diffusivity for t in self._tracers])
        try:
            rc = lib.transport_step_mixed(
                self.p, ntr, _lib.ptr(self._buf), _lib.ptr(u.all_values), _lib.ptr(v.all_values),
                _lib.ptr(w.all_values), ctypes.cast(wv_arr, ctypes.POINTER(ctypes.c_void_p)),
                _lib.ptr(diffusivity.all_values),
                _lib.ptr(self.Ah_u.all_values) if self.Ah_u is not None else None,
                _lib.ptr(self.Ah_v.all_values) if self.Ah_v is not None else None, float(timestep),
                ctypes.cast(mol_arr, ctypes.POINTER(ctypes.c_double)), _lib.ptr(self.h), _lib.stream_handle())
            _lib.check(rc, "transport_step_mixed")
        finally:
            if ea4 is not None:
                lib.transport_set_sources(self.p, None)
        if self.conserve:
            _lib.check(lib.transport_conserve(self.p, ntr, _lib.ptr(self._buf), _lib.ptr(self._target),
                                              _lib.stream_handle()), "transport_conserve")

    def _advance_sequence(self, timestep, u, v, w, diffusivity) -> None:
        """The reference's per-tracer operator sequence (FP64): apply_3d_batch with
        each tracer's combined w, then prepare/apply per tracer with its sources."""
        self._sequence_operators()
        tracers = list(self._tracers)
        wv = {id(t): c for t, c in zip(tracers, self._combined_w(w))}

        def combined_w(tracer: Array) -> Array:
            c = wv[id(tracer)]
            return w if c is None else Array(self.grid, c, INTERFACES)

        self._advection.apply_3d_batch(u, v, w, timestep, tracers, Ah_u=self.Ah_u, Ah_v=self.Ah_v,
                                       get_w=combined_w)
        self.h.copy_(self._advection.h)
        avmol = -1.0
        for tracer in tracers:
            ea4 = None
            if tracer.source is not None or tracer.surface_flux is not None:
                # tracer.py:263-272
                ea4 = self._source.all_values
                if tracer.source is not None:
                    ea4.copy_(tracer.source.all_values)
                else:
                    ea4.zero_()
                if tracer.surface_flux is not None:
                    ea4[-1] += tracer.surface_flux.all_values
                ea4 *= timestep * tracer.source_scale
            if tracer.molecular_diffusivity != avmol:
                avmol = tracer.molecular_diffusivity
                self._vertical_diffusion.prepare(diffusivity, timestep, avmol)
            self._vertical_diffusion.apply(tracer, ea4=self._source if ea4 is not None else None)

    def _fp32_sources(self, timestep: float) -> torch.Tensor:
        """The ea4 terms of every tracer stacked (FP64; zero for tracers without
        sources), formed as the reference's _diffuse does (tracer.py:263-272)."""
        g = self.grid
        shape = (len(self), g.nz_, g.ny_, g.nx_)
        if getattr(self, "_ea4", None) is None or tuple(self._ea4.shape) != shape:
            self._ea4 = torch.zeros(shape, dtype=torch.float64, device=g.device)
        for t, tr in enumerate(self._tracers):
            e = self._ea4[t]
            if tr.source is None and tr.surface_flux is None:
                e.zero_()
                continue
            if tr.source is not None:
                e.copy_(tr.source.all_values)
            else:
                e.zero_()
            if tr.surface_flux is not None:
                e[-1] += tr.surface_flux.all_values
            e *= timestep * tr.source_scale
        return self._ea4

    # ---- inventories and values -------------------------------------------
    def inventory_device(self, h: Optional[torch.Tensor] = None, absolute: bool = False) -> torch.Tensor:
        """Per-tracer sum of area * h * c over wet cells (FP64, compensated,
        deterministic) as a device tensor; h defaults to the T grid's hn.  Of the
        stored values; with ``absolute`` and anomalies stored, of c itself (the
        stored inventory plus o_t times the water volume sum(area * h))."""
        if not self.batched:
            raise ValueError("inventories are computed by the batched transport object")
        out = torch.empty(len(self), dtype=torch.float64, device=self.grid.device)
        if not len(self):
            return out
        _lib.check(_lib.load().transport_inventory(self.p, len(self), _lib.ptr(self._buf),
                                                   _lib.ptr(h) if h is not None else None, _lib.ptr(out),
                                                   _lib.stream_handle()), "transport_inventory")
        if absolute and self.offsets.any():
            g = self.grid
            vol = self.volume(h)
            out = out + torch.as_tensor(self.offsets, dtype=torch.float64, device=g.device) * vol
        return out

    def volume(self, h: Optional[torch.Tensor] = None) -> torch.Tensor:
        """Water volume sum(area * h) over wet cells (one-element FP64 tensor)."""
        g = self.grid
        hh = g.hn.all_values if h is None else h
        wet = (g.mask.all_values == 1).to(hh.device)
        return (g.area.all_values * hh * wet).sum().reshape(1)

    def absolute(self, i: Optional[int] = None) -> torch.Tensor:
        """Tracer values c in FP64 (decoded, plus the offset): tracer i, or all
        stacked."""
        if i is None:
            o = torch.as_tensor(self.offsets, dtype=torch.float64, device=self.grid.device).view(-1, 1, 1, 1)
            x = decode(self.values, self.residuals) if self.compensated else self.values.double()
            return x + o
        x = decode(self.values[i], self.residuals[i]) if self.compensated else self.values[i].double()
        return x + float(self.offsets[i])

    def inventory(self, h: Optional[torch.Tensor] = None):
        """Host copy of inventory_device (simulation.py:1408-1466 report these)."""
        return self.inventory_device(h).cpu().numpy()

    def non_finite(self) -> np.ndarray:
        """Per tracer (rows): the number of non-finite values in interior cells
        whose mask is 1 or 2, and the (k, j, i) of the first in memory order
        (-1s if none); counted on the device (``transport_check_finite``)."""
        n = len(self)
        out = torch.empty(2 * max(n, 1), dtype=torch.int64, device=self.grid.device)
        if n:
            _lib.check(_lib.load().transport_check_finite(self.p, n, _lib.ptr(self._buf), _lib.ptr(out),
                                                          _lib.stream_handle()), "transport_check_finite")
        res = out[:2 * n].view(n, 2).cpu().numpy()
        g = self.grid
        rows = np.full((n, 4), -1, dtype=np.int64)
        for t, (count, first) in enumerate(res):
            rows[t, 0] = count
            if count:
                rows[t, 1:] = (first // (g.ny_ * g.nx_), (first // g.nx_) % g.ny_, first % g.nx_)
        return rows

    def check_finite(self) -> bool:
        """The tracers' part of Simulation.check_finite (simulation.py:348-410):
        logs every tracer with non-finite values in unmasked interior cells,
        with their count and the first one's (k, j, i), and returns whether all
        values were finite."""
        if not self.batched:
            raise ValueError("the finiteness check runs on the batched transport object")
        ok = True
        g = self.grid
        unmasked = int(((g.mask.values == 1) | (g.mask.values == 2)).sum()) * g.nz_
        for tr, (count, k, j, i) in zip(self._tracers, self.non_finite()):
            if count:
                ok = False
                self.logger.error(f"Field {tr.name} has {count} non-finite values (out of {unmasked} unmasked "
                                  f"values); the first at (k, j, i) = ({k}, {j}, {i}).")
        return ok

    # ---- launch shape and rounding ------------------------------------------
    def set_tuning(self, tracers_per_block_h: int = 0, tracers_per_block_w: int = 16,
                   tracers_per_block_d: int = 16, tile_y: int = 0) -> None:
        """Launch shape of the batched kernels: tracers per block of the three kernel
        kinds (0 = all) and the height of the horizontal-stage tiles (4, 8, 16, and 12
        for the compensated form; 0 = the default: 16 rows with two cells per lane for
        the FP32 forms with 16 or more tracers, more than 8 per block, the compensated
        form's y-first pair on 12 rows; 8 rows otherwise).  Does not change results."""
        if not self.batched:
            return
        _lib.check(_lib.load().transport_set_tuning(self.p, int(tile_y), int(tracers_per_block_h),
                                                    int(tracers_per_block_w), int(tracers_per_block_d)),
                   "transport_set_tuning")

    def set_solver(self, pcr_max_nz: int = 0) -> None:
        """FP32 column solve of the "f32-sr" form: wave-level parallel cyclic
        reduction for columns of 2 <= nz <= pcr_max_nz (at most 32), per-lane
        Thomas elimination otherwise (0: Thomas only).  The two agree to FP32
        rounding.  The compensated form always eliminates per lane."""
        if not self.batched:
            return
        _lib.check(_lib.load().transport_set_solver(self.p, int(pcr_max_nz)), "transport_set_solver")
        self.pcr_max_nz = int(pcr_max_nz)

    def set_rounding(self, mode: str = "stochastic", counter: Optional[torch.Tensor] = None) -> None:
        """Rounding of the "f32-sr" form's stores: "stochastic" (its default)
        rounds each stored c + dc up or down with probability given by the exact
        residual, so rounding errors do not accumulate in a steady flow;
        "nearest" is plain FP32 arithmetic.  The random bits are a hash of (step
        count, stage, tracer, cell) -- deterministic, independent of launch shape
        and batching.  ``counter`` (a one-element int64 device tensor) holds the
        step count and is advanced by every step; None uses the collection's own
        counter (zero at creation, saved by state_dict).  The other forms have
        nothing to choose: FP64 is exact, the compensated form rounds at 2^-16 of
        an FP32 ulp."""
        if not self.batched or self.precision != "f32-sr":
            return
        if counter is not None and (counter.dtype != torch.int64 or counter.numel() != 1 or not counter.is_cuda):
            raise ValueError("the step counter is a one-element int64 device tensor")
        ctr = counter if counter is not None else self._own_counter
        rc = _lib.load().transport_set_rounding(self.p, _ROUNDING[mode], _lib.ptr(ctr))
        _lib.check(rc, "transport_set_rounding")
        self._counter = counter
        self.rounding = mode

    def snapshot(self):
        """Device copies of the state a step changes (the tracer buffer, values and
        residuals, and the rounding step counter), for ``restore``."""
        ctr = self._counter if self._counter is not None else getattr(self, "_own_counter", None)
        return self._buf.clone(), (ctr.clone() if ctr is not None else None)

    def restore(self, snap) -> None:
        """Copy a snapshot back in place (same tracer count as when it was taken)."""
        buf, ctr = snap
        if buf.shape != self._buf.shape:
            raise ValueError("the snapshot holds another number of tracers")
        self._buf.copy_(buf)
        mine = self._counter if self._counter is not None else getattr(self, "_own_counter", None)
        if mine is not None and ctr is not None:
            mine.copy_(ctr)

    # ---- restart: snapshot and restore of the state the step depends on ----
    def state_dict(self) -> dict:
        """Host copy of the collection's state: tracer values in their storage
        precision (and the residuals of the compensated form), per-tracer names and
        molecular diffusivities, offsets, for the "f32-sr" form the rounding mode
        and step count, and the grid's layer thicknesses ho and hn (the reference
        keeps T.ho in its restarts as part of the model state, simulation.py:620-622;
        hn is the thickness the tracers were last stepped to).  A collection
        restored from it continues bit for bit as the original would (velocities
        and diffusivity are inputs of every step; tracer sources and vertical
        velocities are references to the caller's fields, which a restore keeps)."""
        torch.cuda.synchronize(self.grid.device)
        ctr = self._counter if self._counter is not None else getattr(self, "_own_counter", None)
        st = {"values": self.values.detach().cpu().numpy().copy(), "precision": self.precision,
              "names": self.names, "molecular": [t.molecular_diffusivity for t in self._tracers],
              "rounding": self.rounding, "step": int(ctr.item()) if ctr is not None else 0,
              "offsets": [float(o) for o in self.offsets],
              "ho": self.grid.ho.all_values.detach().cpu().numpy().copy(),
              "hn": self.grid.hn.all_values.detach().cpu().numpy().copy()}
        if self.compensated:
            st["residuals"] = self.residuals.detach().cpu().numpy().copy()
        return st

    def load_state_dict(self, state: dict) -> None:
        """Restore a state_dict; the collection must have the same grid shape and
        precision.  With as many tracers as the state holds, the collection's
        tracers keep their sources, surface fluxes and vertical velocities;
        otherwise the collection is rebuilt with plain tracers."""
        if state["precision"] != self.precision:
            raise ValueError(f"state is {state['precision']}, collection is {self.precision}")
        missing = [k for k in ("values", "ho", "hn") if k not in state]
        if missing:
            # as the reference's restart (simulation.py:185-197)
            raise ValueError(f"the state lacks field(s) of the model state: {', '.join(missing)}")
        for name in ("ho", "hn"):
            dst = getattr(self.grid, name).all_values
            src = np.asarray(state[name])
            if tuple(src.shape) != tuple(dst.shape):
                raise ValueError(f"{name}: state has shape {src.shape}, grid has {tuple(dst.shape)}")
            dst.copy_(torch.from_numpy(np.ascontiguousarray(src)).to(dst.device))
        vals = torch.from_numpy(np.ascontiguousarray(state["values"]))
        ntr = vals.shape[0]
        if ntr != len(self):
            self._tracers = [Tracer(self, t) for t in range(ntr)]
            self._alloc(ntr)
        self.values.copy_(vals.to(self.grid.device))
        if self.compensated:
            self.residuals.copy_(torch.from_numpy(np.ascontiguousarray(state["residuals"])).to(self.grid.device))
        self.offsets = np.asarray(state.get("offsets", np.zeros(ntr)), dtype=np.float64)
        for t, name, mol in zip(self._tracers, state["names"], state["molecular"]):
            t.name = name
            t.molecular_diffusivity = float(mol)
        if self.batched and self.precision == "f32-sr":
            self.set_rounding(state["rounding"], self._counter)
            ctr = self._counter if self._counter is not None else self._own_counter
            ctr.fill_(int(state["step"]))

    def save(self, path: str) -> None:
        """Write state_dict to an .npz file (arrays and a JSON header; no pickles)."""
        st = self.state_dict()
        arrays = {k: st.pop(k) for k in ("values", "residuals", "ho", "hn") if k in st}
        np.savez(path, **arrays, meta=np.frombuffer(json.dumps(st).encode(), dtype=np.uint8))

    def load(self, path: str) -> None:
        """Restore a file written by ``save``."""
        with np.load(path, allow_pickle=False) as f:
            st = json.loads(bytes(f["meta"]).decode())
            for k in ("values", "residuals", "ho", "hn"):
                if k in f:
                    st[k] = f[k]
        self.load_state_dict(st)

    def stats(self, reset: bool = True):
        """(largest z sub-iteration count, columns capped at 200) since the last reset."""
        if not self.batched:
            return self._advection.stats(reset)
        mi, cap = ctypes.c_int(0), ctypes.c_int(0)
        _lib.check(_lib.load().transport_get_stats(self.p, ctypes.byref(mi), ctypes.byref(cap), int(reset),
                                                   _lib.stream_handle()), "transport_get_stats")
        return mi.value, cap.value
