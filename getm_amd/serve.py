"""Serving loop for transport scenarios with continuous batching.

Requests (dye-release scenarios: a domain, 1..T point-released FP32 tracers,
a number of steps and a limiter) arrive over time.  A request is admitted as
soon as a slot and HBM budget are free; its state then stays on the device.
Every tick advances all in-flight requests -- by a share of work each (fair
policy: 1 to 64 steps, by request size) or by one step each (lockstep) -- with
one ensemble launch per kernel kind and (limiter, steps per tick) group
(getm_amd.ensemble), so requests that arrive while others run share the same
launches, and finished requests leave without waiting for the others.
Per-request arithmetic does not depend on what else is in flight: a result is
the same alone and in the mix, bit for bit.

With ``conserve=True`` each request's tracer inventories are restored to their
values at admission (FP64 inventories, FP32 ulp fixer), which holds because the
scenario fields keep ho == hn (closed, steady flow): once, when the request
completes (``fix="end"``, the default: 5 launches per request in all), or after
every tick (``fix="tick"``: 5 launches per request and tick).

Requests may carry a latency budget (``deadline``, from arrival) and a client
cancellation time (``cancel_after``, from arrival); ``Server.cancel(rid)``
cancels one from outside (e.g. an ``on_tick`` callback).  A cancelled or expired
request leaves at the next tick boundary, frees its slot and budget, and is
reported with its status.

Latency is measured from a request's arrival to its result being complete on
the device.  ``verify`` re-runs completed requests with the FP64 plain method
(the batched FP64 step, equal bit for bit to the reference algorithm) and
reports each FP32 result's error against it.

    python -m getm_amd.serve --requests 48 --steps 50,400 --tracers 1,64
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import math
import sys
import time
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from . import _lib, inputs
from .core import CENTERS, INTERFACES, Grid
from .ensemble import Ensemble
from .tracer import TracerCollection


@dataclasses.dataclass
class Request:
    rid: int
    nx: int
    ny: int
    nz: int
    ntracers: int
    nsteps: int
    scheme: int
    seed: int
    arrival: float = 0.0  # seconds after the start of the run
    deadline: Optional[float] = None  # latency budget from arrival; expired requests are dropped
    cancel_after: Optional[float] = None  # the client cancels this long after arrival

    @property
    def cells(self) -> int:
        return (self.nx + 4) * (self.ny + 4) * self.nz

    def domain(self) -> inputs.Domain:
        return inputs.steady_release(nx=self.nx, ny=self.ny, nz=self.nz, scheme=self.scheme, seed=self.seed,
                                     ntracers=self.ntracers, dense=False)


@dataclasses.dataclass
class Result:
    rid: int
    arrival: float
    admitted: float
    finished: float
    steps: int
    wet_updates: int  # wet cells x nz x tracers x steps
    inventory_drift: float  # max over tracers of |I - I0| / I0 at the end
    tracers: Optional[torch.Tensor] = None  # (T, nz, ny_, nx_) float32 on the device
    error: str = ""
    status: str = "done"  # done | rejected | cancelled | expired

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def initial_tracers(dom: inputs.Domain, device, dtype=torch.float32) -> torch.Tensor:
    """(T, nz, ny_, nx_) device tensor of a domain's initial tracers (dense or from its releases)."""
    if dom.releases is None:
        return torch.from_numpy(np.ascontiguousarray(dom.tracers)).to(device=device, dtype=dtype).contiguous()
    out = torch.zeros((dom.ntracers, dom.nz, dom.ny_, dom.nx_), dtype=dtype, device=device)
    rel = torch.from_numpy(dom.releases).to(device)
    idx = rel[:, :3].long()
    out[torch.arange(dom.ntracers, device=device), idx[:, 0], idx[:, 1], idx[:, 2]] = rel[:, 3].to(dtype)
    return out


def request_bytes(req: Request) -> int:
    """Device bytes a request holds while in flight: FP64 fields (ho, hn, hu, hv, u, v,
    w, nuh), FP32 tracers and their stage scratch, FP64 stage thicknesses, masks/metrics."""
    n2 = (req.nx + 4) * (req.ny + 4)
    n3 = n2 * req.nz
    return n3 * (6 * 8 + 2 * 8 + 2 * 4 * req.ntracers) + 2 * n2 * (req.nz + 1) * 8 + n2 * 64


def draw_requests(n: int, seed: int = 0, size_range=(32, 192), nz_range=(10, 60), tracer_range=(1, 32),
                  step_range=(20, 200)) -> List[Request]:
    """Scenarios drawn like the uneven ensemble: nx, ny log-uniform, nz uniform,
    tracers and steps uniform, SUPERBEE or HSIMT."""
    rng = np.random.default_rng(seed)
    lo, hi = math.log(size_range[0]), math.log(size_range[1])
    out = []
    for r in range(n):
        out.append(Request(rid=r, nx=int(round(math.exp(rng.uniform(lo, hi)))),
                           ny=int(round(math.exp(rng.uniform(lo, hi)))),
                           nz=int(rng.integers(nz_range[0], nz_range[1] + 1)),
                           ntracers=int(rng.integers(tracer_range[0], tracer_range[1] + 1)),
                           nsteps=int(rng.integers(step_range[0], step_range[1] + 1)),
                           scheme=int(rng.choice([5, 1])), seed=10_000 + r))
    return out


def relative_drift(now: torch.Tensor, target: torch.Tensor) -> float:
    """max over tracers of |I - I0| / |I0|; a tracer whose inventory at admission
    is zero (no mass, or no water) contributes its absolute change instead."""
    scale = torch.where(target != 0, target.abs(), torch.ones_like(target))
    return float(((now - target).abs() / scale).max().item())


class _Slot:
    _serial = 0

    def __init__(self, req: Request, dom: inputs.Domain, device, conserve: bool, quantum: Optional[float],
                 horizon: Optional[int] = None):
        _Slot._serial += 1
        self.serial = _Slot._serial
        self.req = req
        # steps per loop tick, a property of the request alone: 1 (lockstep); with a
        # work quantum Q the power of two nearest Q / (updates per step), 1 .. 64
        # (fair); with a horizon H, ceil(steps / H), so that every request completes
        # in at most H ticks, and in about H, whatever its size (uniform)
        self.cost = dom.n_wet_columns * dom.nz * req.ntracers
        self.k = 1
        if quantum:
            self.k = int(min(64, max(1, 2 ** round(math.log2(max(quantum / max(self.cost, 1), 1e-9))))))
        if horizon:
            self.k = max(1, math.ceil(req.nsteps / horizon))
        self.exact = bool(horizon)
        self.grid = Grid.from_domain(dom, device=device)
        g = self.grid
        dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)
        self.u = g.ugrid.wrap(dev(dom.u), CENTERS)
        self.v = g.vgrid.wrap(dev(dom.v), CENTERS)
        self.w = g.wrap(dev(dom.w), INTERFACES)
        self.nuh = g.wrap(dev(dom.nuh), INTERFACES)
        self.tracers = initial_tracers(dom, device)
        self.dt = dom.dt
        self.steps = 0
        self.wet_cols = dom.n_wet_columns
        # a transport object of this grid provides inventories and the FP32 fixer
        self.tc = TracerCollection(g, advection_scheme=req.scheme, precision="f32-sr")
        self.target = self._inventory()
        self.conserve = conserve
        # step count of the request's stochastic rounding, carried by the request
        # from ensemble to ensemble, so its results never depend on the mix
        self.counter = torch.zeros(1, dtype=torch.int64, device=device)

    def _inventory(self) -> torch.Tensor:
        out = torch.empty(self.req.ntracers, dtype=torch.float64, device=self.grid.device)
        _lib.check(_lib.load().transport_inventory(self.tc.p, self.req.ntracers, _lib.ptr(self.tracers), None,
                                                   _lib.ptr(out), _lib.stream_handle()), "transport_inventory")
        return out

    def fix(self) -> None:
        _lib.check(_lib.load().transport_conserve(self.tc.p, self.req.ntracers, _lib.ptr(self.tracers),
                                                  _lib.ptr(self.target), _lib.stream_handle()), "transport_conserve")

    def drift(self) -> float:
        return relative_drift(self._inventory(), self.target)

    def steps_this_tick(self) -> int:
        """k, or (uniform) the steps left, or the largest power of two not above them."""
        left = self.req.nsteps - self.steps
        if self.exact:
            return max(1, min(self.k, left))
        kk = self.k
        while kk > left:
            kk //= 2
        return max(kk, 1)


class Server:
    """policy "lockstep": every request advances one step per tick and is fixed
    after every step.  policy "fair": every tick each request advances about
    ``quantum`` tracer cell-updates (k = 1 .. 64 steps, by its size), so a small
    request is not held to the pace of the largest one in the mix.  policy
    "uniform": every request advances k = ceil(steps / ``horizon``) steps per tick
    (its last tick the steps left), so each completes in at most ``horizon`` ticks,
    and in about that many, whatever its size: latency no longer follows request
    size (the tail narrows), at the cost of small requests' latency.  Under "fair" and "uniform" a request's inventory
    is restored once per tick.  Under every policy a request's result depends only
    on the request itself, never on the mix."""

    def __init__(self, hbm_budget: float = 250e9, max_slots: int = 64, conserve: bool = True,
                 keep_results: bool = False, tuning: Optional[Sequence[int]] = None, device=None,
                 policy: str = "lockstep", quantum: float = 2.0 ** 25, horizon: int = 8, fix: str = "end"):
        if policy not in ("fair", "lockstep", "uniform"):
            raise ValueError("policy must be 'fair', 'lockstep' or 'uniform'")
        if fix not in ("end", "tick"):
            raise ValueError("fix must be 'end' (once per request) or 'tick' (after every tick)")
        self.fix = fix
        self._cancel: set = set()
        if policy == "uniform" and horizon < 1:
            raise ValueError("horizon must be at least one tick")
        self.policy = policy
        self.quantum = float(quantum) if policy == "fair" else None
        self.horizon = int(horizon) if policy == "uniform" else None
        self.hbm_budget = float(hbm_budget)
        self.max_slots = int(max_slots)
        self.conserve = conserve
        self.keep_results = keep_results
        self.tuning = tuning
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        self._groups: Dict[int, Ensemble] = {}
        self._members: Dict[int, tuple] = {}

    def _ensemble(self, group: tuple, slots: List[_Slot]) -> Ensemble:
        """The ensemble of one (limiter, steps per tick) group, rebuilt when its members change."""
        key = tuple(s.serial for s in slots)
        if self._members.get(group) != key:
            torch.cuda.synchronize()  # the previous ensemble of this group may still be running
            ens = Ensemble([s.grid for s in slots], [s.req.ntracers for s in slots], scheme=group[0])
            for i, s in enumerate(slots):
                ens.bind(i, s.tracers, s.u, s.v, s.w, s.nuh, s.dt)
            ens.set_rounding("stochastic", [s.counter for s in slots])
            if self.tuning:
                ens.set_tuning(*self.tuning[:4])
                if len(self.tuning) > 4:
                    ens.set_solver(self.tuning[4])
            ens.commit()
            self._groups[group] = ens
            self._members[group] = key
        return self._groups[group]

    def cancel(self, rid: int) -> None:
        """Cancel request ``rid``: it leaves (or is never admitted) at the next tick
        boundary and is reported with status "cancelled"."""
        self._cancel.add(rid)

    def run(self, requests: Sequence[Request], domains: Optional[Dict[int, inputs.Domain]] = None,
            log=None, clock: str = "wall", on_tick=None) -> List[Result]:
        """Serve ``requests`` (arrival times relative to the call) until all are done,
        cancelled or expired.  ``domains`` may hold pre-generated inputs by request
        id (generated up front otherwise).  With ``clock="ticks"`` arrivals,
        deadlines and latencies count steps of the loop instead of seconds, which
        makes admission order reproducible.  ``on_tick(server, tick)`` is called
        before every tick."""
        if clock not in ("wall", "ticks"):
            raise ValueError("clock must be 'wall' or 'ticks'")
        self._groups.clear()
        self._members.clear()
        self.peak_active, self.peak_bytes = 0, 0
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        domains = {} if domains is None else domains
        for r in pending:
            if r.rid not in domains:
                domains[r.rid] = r.domain()
        active: List[_Slot] = []
        results: List[Result] = []
        used = 0
        ticks = 0
        t0 = time.perf_counter()
        now = (lambda: float(ticks)) if clock == "ticks" else (lambda: time.perf_counter() - t0)
        admitted_at: Dict[int, float] = {}
        def leaving(req: Request, t: float) -> str:
            """"cancelled" / "expired" if the request must leave at time t, else ""."""
            if req.rid in self._cancel or (req.cancel_after is not None and t >= req.arrival + req.cancel_after):
                return "cancelled"
            if req.deadline is not None and t > req.arrival + req.deadline:
                return "expired"
            return ""

        while pending or active:
            if on_tick:
                on_tick(self, ticks)
            # cancellations and expiry at the tick boundary, of arrived requests
            t_now = now()
            keep = []
            for req in pending:
                why = leaving(req, t_now) if req.arrival <= t_now else ""
                if why:
                    results.append(Result(req.rid, req.arrival, float("nan"), t_now, 0, 0, float("nan"),
                                          error=why, status=why))
                else:
                    keep.append(req)
            pending = keep
            for s in list(active):
                why = leaving(s.req, t_now)
                if why:
                    results.append(Result(s.req.rid, s.req.arrival, admitted_at[s.req.rid], t_now, s.steps,
                                          s.wet_cols * s.req.nz * s.req.ntracers * s.steps, float("nan"),
                                          error=why, status=why))
                    used -= request_bytes(s.req)
                    active.remove(s)
                    if log:
                        log(f"[serve] request {s.req.rid} {why} after {s.steps} steps")
            # admission: arrived requests, in arrival order, while slots and budget allow
            while pending and pending[0].arrival <= now() and len(active) < self.max_slots:
                req = pending[0]
                need = request_bytes(req)
                if need > self.hbm_budget:
                    pending.pop(0)
                    results.append(Result(req.rid, req.arrival, now(), now(), 0, 0, float("nan"),
                                          error="request exceeds the HBM budget", status="rejected"))
                    continue
                if used + need > self.hbm_budget:
                    break
                pending.pop(0)
                active.append(_Slot(req, domains[req.rid], self.device, self.conserve, self.quantum, self.horizon))
                used += need
                admitted_at[req.rid] = now()
            self.peak_active = max(self.peak_active, len(active))
            self.peak_bytes = max(self.peak_bytes, used)
            if not active:
                if pending:
                    if clock == "ticks":
                        ticks = max(ticks, math.ceil(pending[0].arrival))
                    else:
                        time.sleep(min(max(pending[0].arrival - now(), 0.0), 0.05))
                continue
            # one tick: every in-flight request advances its steps for this tick, the
            # requests of one (limiter, steps) group in one ensemble launch per kernel kind
            groups: Dict[tuple, List[_Slot]] = {}
            for s in active:
                groups.setdefault((s.req.scheme, s.steps_this_tick()), []).append(s)
            for group in sorted(groups):
                ens = self._ensemble(group, groups[group])
                for _ in range(group[1]):
                    ens.step()
            for group in sorted(groups):
                for s in groups[group]:
                    if self.conserve and self.fix == "tick":
                        s.fix()
                    s.steps += group[1]
            ticks += 1
            done = [s for s in active if s.steps >= s.req.nsteps]
            if done:
                if self.conserve and self.fix == "end":
                    for s in done:
                        s.fix()
                torch.cuda.synchronize()
                for s in done:
                    res = Result(s.req.rid, s.req.arrival, admitted_at[s.req.rid], now(), s.steps,
                                 s.wet_cols * s.req.nz * s.req.ntracers * s.steps, s.drift())
                    if self.keep_results:
                        res.tracers = s.tracers  # device tensor, no longer stepped
                    results.append(res)
                    used -= request_bytes(s.req)
                    if log:
                        log(f"[serve] request {s.req.rid} done: {s.steps} steps, latency {res.latency:.3f}")
                active = [s for s in active if s.steps < s.req.nsteps]
            live = {(s.req.scheme, s.steps_this_tick()) for s in active}
            for group in list(self._groups):
                if group not in live:
                    del self._groups[group]  # releases the group's scratch and bound tensors
                    del self._members[group]
        self._groups.clear()
        self._members.clear()
        self._cancel.clear()
        return results


def reference_fp64(req: Request, dom: inputs.Domain, device=None) -> torch.Tensor:
    """The request's tracers after ``req.nsteps`` steps of the FP64 plain method
    (no inventory fixer: FP64 conserves these closed, steady scenarios by itself)."""
    g = Grid.from_domain(dom, device=device)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(g.device)
    u, v = g.ugrid.wrap(dev(dom.u), CENTERS), g.vgrid.wrap(dev(dom.v), CENTERS)
    w, nuh = g.wrap(dev(dom.w), INTERFACES), g.wrap(dev(dom.nuh), INTERFACES)
    tc = TracerCollection(g, advection_scheme=req.scheme, precision="f64")
    tc.set_values(initial_tracers(dom, g.device, torch.float64))
    for _ in range(req.nsteps):
        tc.advance(dom.dt, u, v, w, nuh)
    return tc.values


def verify(results: Sequence[Result], requests: Sequence[Request], domains: Dict[int, inputs.Domain],
           max_updates: Optional[float] = None) -> dict:
    """Error of every completed result (kept with ``keep_results``) against the FP64
    plain method: max over tracers of max|c32 - c64| / max|c64| on wet cells.
    ``max_updates`` limits the check to results of at most that many tracer
    cell-updates (the FP64 rerun of the largest requests takes minutes)."""
    by_id = {r.rid: r for r in requests}
    errs = {}
    for res in results:
        if res.error or res.tracers is None:
            continue
        if max_updates is not None and res.wet_updates > max_updates:
            continue
        req = by_id[res.rid]
        ref = reference_fp64(req, domains[req.rid], device=res.tracers.device)
        wet = torch.from_numpy(domains[req.rid].tmask == 1).to(ref.device)
        e = 0.0
        for t in range(ref.shape[0]):
            r64 = ref[t][:, wet]
            scale = float(r64.abs().max().item())
            diff = float((res.tracers[t][:, wet].double() - r64).abs().max().item())
            e = max(e, diff / scale if scale > 0 else diff)
        errs[res.rid] = e
        del ref
    return errs


def percentile(xs: Sequence[float], q: float) -> float:
    return float(np.percentile(np.asarray(xs), q)) if len(xs) else float("nan")


def summary(results: Sequence[Result], wall: float, rate: Optional[float] = None) -> dict:
    """Throughput and latency percentiles; with ``rate`` (tracer cell-updates/s at
    saturation) also the slowdown, latency / (work / rate), a size-free latency."""
    ok = [r for r in results if not r.error]
    lat = [r.latency for r in ok]
    upd = sum(r.wet_updates for r in ok)
    out = {"completed": len(ok), "rejected": sum(r.status == "rejected" for r in results),
           "cancelled": sum(r.status == "cancelled" for r in results),
           "expired": sum(r.status == "expired" for r in results), "wall_s": wall,
           "tracer_cell_updates_per_s": upd / wall if wall > 0 else float("nan"),
           "latency_p50_s": percentile(lat, 50), "latency_p99_s": percentile(lat, 99),
           "latency_p999_s": percentile(lat, 99.9),
           "max_inventory_drift": max((r.inventory_drift for r in ok), default=float("nan"))}
    if rate:
        slow = [r.latency * rate / r.wet_updates for r in ok]
        out.update(slowdown_p50=percentile(slow, 50), slowdown_p99=percentile(slow, 99))
    return out


def sweep(reqs: Sequence[Request], domains: Dict[int, inputs.Domain], fractions: Sequence[float],
          server: Server, verify_results: bool = True, seed: int = 0, log=None,
          verify_max_updates: Optional[float] = None) -> dict:
    """Saturation run (every request arrives at t = 0), optionally checked against
    the FP64 method, then one Poisson-arrival run per fraction of the measured
    saturation throughput.  Returns the report as a dict."""
    log = log or (lambda m: None)
    # warm-up (kernel loading)
    server.run([dataclasses.replace(reqs[0], rid=-1, nsteps=2)], {-1: domains[reqs[0].rid]})
    torch.cuda.synchronize()
    keep = server.keep_results
    server.keep_results = verify_results
    t = time.perf_counter()
    sat = server.run(reqs, domains, log=log)
    wall = time.perf_counter() - t
    server.keep_results = keep
    s0 = summary(sat, wall)
    sat_rate = s0["tracer_cell_updates_per_s"]
    s0.update(summary(sat, wall, sat_rate))
    work = [r.wet_updates for r in sat if not r.error]
    report = {"requests": {"n": len(reqs), "mean_work_updates": float(np.mean(work)),
                           "budget_gb": server.hbm_budget / 1e9, "max_slots": server.max_slots},
              "saturation": s0, "sweep": []}
    log(f"[serve] saturation {sat_rate:.3e} tracer cell-updates/s over {wall:.1f} s, "
        f"p50 {s0['latency_p50_s']:.3f} s, drift {s0['max_inventory_drift']:.1e}")
    errs = {}
    if verify_results:
        errs = verify(sat, reqs, domains, verify_max_updates)
        s0["max_error_vs_fp64"] = max(errs.values()) if errs else float("nan")
        s0["verified"] = len(errs)
        log(f"[serve] {len(errs)} results checked against FP64: max error {s0['max_error_vs_fp64']:.2e}")
    by_id = {r.rid: r for r in reqs}
    s0["per_request"] = [
        {"rid": r.rid, "nx": by_id[r.rid].nx, "ny": by_id[r.rid].ny, "nz": by_id[r.rid].nz,
         "tracers": by_id[r.rid].ntracers, "steps": r.steps, "scheme": by_id[r.rid].scheme, "latency_s": r.latency,
         "inventory_drift": r.inventory_drift, "error_vs_fp64": errs.get(r.rid)}
        for r in sorted(sat, key=lambda r: r.rid) if not r.error]
    del sat
    for frac in fractions:
        lam = frac * sat_rate / float(np.mean(work))  # requests per second
        rng = np.random.default_rng(seed + 1)
        arrivals = np.cumsum(rng.exponential(1.0 / lam, size=len(reqs)))
        batch = [dataclasses.replace(r, arrival=float(a)) for r, a in zip(reqs, arrivals)]
        t = time.perf_counter()
        res = server.run(batch, domains, log=log)
        wall = time.perf_counter() - t
        s = summary(res, wall, sat_rate)
        s["fraction_of_saturation"] = frac
        s["arrival_rate_per_s"] = lam
        report["sweep"].append(s)
        log(f"[serve] load {frac:.2f}: p50 {s['latency_p50_s']:.3f} s, p99 {s['latency_p99_s']:.3f} s, "
            f"p99.9 {s['latency_p999_s']:.3f} s, slowdown p50 {s['slowdown_p50']:.2f} p99 {s['slowdown_p99']:.2f}, "
            f"{s['tracer_cell_updates_per_s']:.3e} updates/s, max drift {s['max_inventory_drift']:.1e}")
    return report


def main(argv=None):
    ap = argparse.ArgumentParser(description="Serving-loop sweep: saturation throughput, then Poisson arrivals "
                                             "at fractions of it, with latency percentiles.")
    ap.add_argument("--requests", type=int, default=48)
    ap.add_argument("--fractions", default="0.1,0.5,0.8,1.0,1.2", help="arrival rates as fractions of saturation")
    ap.add_argument("--budget-gb", type=float, default=250.0)
    ap.add_argument("--max-slots", type=int, default=64)
    ap.add_argument("--size", default="64,768", help="nx, ny range (log-uniform)")
    ap.add_argument("--layers", default="10,60")
    ap.add_argument("--tracers", default="1,200")
    ap.add_argument("--steps", default="50,2000")
    ap.add_argument("--no-verify", action="store_true", help="skip the FP64 check of the saturation run")
    ap.add_argument("--no-conserve", action="store_true", help="no inventory fixer (raw FP32 steps)")
    ap.add_argument("--fix", choices=["end", "tick"], default="end",
                    help="inventory fixer once per request at completion, or after every tick")
    ap.add_argument("--verify-max-updates", type=float, default=0.0,
                    help="check against FP64 only results of at most this many tracer cell-updates (0: all)")
    ap.add_argument("--policy", choices=["fair", "lockstep", "uniform"], default="lockstep")
    ap.add_argument("--horizon", type=int, default=8, help="ticks per request under the uniform policy")
    ap.add_argument("--quantum", type=float, default=2.0 ** 25,
                    help="tracer cell-updates per request per tick under the fair policy")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default="")
    args = ap.parse_args(argv)
    rng_range = lambda s: tuple(int(x) for x in s.split(","))
    reqs = draw_requests(args.requests, seed=args.seed, size_range=rng_range(args.size),
                         nz_range=rng_range(args.layers), tracer_range=rng_range(args.tracers),
                         step_range=rng_range(args.steps))
    log = lambda m: print(m, file=sys.stderr, flush=True)
    log(f"[serve] generating {len(reqs)} scenarios")
    domains = {r.rid: r.domain() for r in reqs}
    server = Server(hbm_budget=args.budget_gb * 1e9, max_slots=args.max_slots, conserve=not args.no_conserve,
                    policy=args.policy, quantum=args.quantum, horizon=args.horizon, fix=args.fix)
    report = sweep(reqs, domains, [float(x) for x in args.fractions.split(",") if x], server,
                   verify_results=not args.no_verify, seed=args.seed, log=log,
                   verify_max_updates=args.verify_max_updates or None)
    report["requests"].update(size=args.size, layers=args.layers, tracers=args.tracers, steps=args.steps,
                              conserve=not args.no_conserve, fix=args.fix, policy=args.policy, quantum=args.quantum,
                              horizon=args.horizon if args.policy == "uniform" else None)
    print(json.dumps(report))
    if args.out:
        with open(args.out, "w") as fh:
            json.dump(report, fh, indent=1)


if __name__ == "__main__":
    main()
