"""Benchmark of the tracer-transport step (advection + vertical diffusion).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload multi|coastal|ensemble|producers]
                    [--precision f32|f32-sr|f64] [--tracers T]

A step is one TracerCollection.advance (python/pygetm/tracer.py:193-276).  The
default workload is the multi-tracer batch: 256 passive tracers, each a sum of
4 Gaussian blobs with amplitudes log-uniform in [1e-3, 10], on the synthetic
coastal grid (1024 x 768 x 40 cells, 2 km, 30% land, 40 sigma layers, dt 30 s),
P2_PDM, in the compensated FP32 form (FP32 values plus int16 residuals, FP64
thickness) with the inventory fixer on: every timed step includes the pre-step
inventory and the fixer.  Inputs are resident in HBM before the timed region.
"--workload coastal" is the same grid with temperature and salinity, SUPERBEE.
Prints one JSON line.

metric: tracer cell-updates/s = wet cells (mask == 1) x nz x tracers x steps / time.
roofline: the dominant kernel's algorithmic bytes per launch / its mean launch
time, timed with HIP events on its launch stream during the timed steps.
traffic: HBM bytes per launch of that kernel from the rocprofv3 FETCH_SIZE and
WRITE_SIZE counters, calibrated on a plain copy (getm_amd.hbm_copy), as recorded
in profiles/hbm_traffic_*.md for this workload (null if none is recorded).
serving: a short run of the serving loop (getm_amd.serve) on a bounded sample
(8 requests) of the dye-release request distribution (nx, ny 64-768, 1-200
tracers, 50-2000 steps): saturation throughput, then Poisson arrivals at 80% of
it, with latency percentiles, inventory drift and the error against the FP64
method of the results of at most 3e10 tracer cell-updates.
accuracy (FP32 runs): the first (up to 16) tracers, which evolve independently
of the others, rerun from the same (FP32) initial values for --accuracy-steps
steps in FP64 (the plain method), in the benchmarked FP32 form without and with
the inventory fixer, and for the stochastic-rounding form also storing anomalies
about per-tracer offsets (the same kernels; offsets "auto"):
max |c32 - c64| / max |c64| over wet cells, and the largest per-step
inventory defect |I_after - I_before| / |I_before| (FP64 compensated sums, of the
absolute values).
fp64: the plain method (the batched FP64 step, bit-identical to the reference's
algorithm) on the coastal grid with temperature and salinity, SUPERBEE, timed
over --fp64-steps steps after the main run.
cpu_baseline: the reference's Fortran sweeps (oracle/_ref) with the C restatement
of the diffusion, FP64, on bounded slabs of the same grid with the workload's
limiter, one tracer, one process per host core (the C restatement alone where
oracle/_ref is not built); the host CPU model is reported with it.

--workload producers times the kernels on either side of the step at the same
coastal size (one launch each, HIP events on the launch stream, median of
--steps launches): w from continuity, velocities from transports, sinking into
sources, GVC layer thicknesses, layer coordinates, a grid interpolation and the
transport change of the velocity self-advection.  Each is reported against its algorithmic bytes
and the HBM peak, beside the C restatement timed on one host core on a slab.
"""
from __future__ import annotations

import argparse
import json
import math
import multiprocessing as mp
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
FORM_NAMES = {"f32": "compensated FP32 tracers (FP32 values + int16 residuals) with FP64 thickness",
              "f32-sr": "FP32 tracers with stochastic rounding and FP64 thickness",
              "f64": "FP64 bit-exact"}
NX, NY, NZ = 1024, 768, 40
METRIC = "tracer cell-updates/s (adv+diff step), % of HBM roofline, p50/p99 solve latency"


def recorded_traffic(kernel: str, workload: str):
    """Bytes per launch of ``kernel`` recorded for ``workload`` in profiles/hbm_traffic_*.md
    (lines 'traffic <workload> <kernel> <bytes>'), or None."""
    import glob
    import re

    for fn in sorted(glob.glob(os.path.join(ROOT, "profiles", "hbm_traffic_*.md"))):
        for line in open(fn):
            m = re.match(r"^traffic\s+(\S+)\s+(\S+)\s+([0-9.eE+]+)\s*$", line.strip())
            if m and m.group(1) == workload and m.group(2) == kernel:
                return float(m.group(3)), os.path.relpath(fn, ROOT)
    return None


def accuracy_check(d, grid, fields, init, scheme, steps, form, rounding, log):
    """FP32 error against the FP64 method and inventory defects (module docstring)."""
    import numpy as np
    import torch

    from getm_amd import precision

    wet = torch.from_numpy(np.broadcast_to(d.tmask == 1, (grid.nz_, grid.ny_, grid.nx_)).copy()).to(grid.device)
    ref, d64, _ = precision.run(d, grid, fields, init, scheme, steps, "f64", False)
    out = {"steps": steps, "tracers_checked": int(init.shape[0]), "fp64_max_step_inventory_defect": d64}

    def err(tc):
        e = 0.0
        for t in range(init.shape[0]):
            r = ref.values[t][wet]
            e = max(e, float((tc.absolute(t)[wet] - r).abs().max() / r.abs().max()))
        return e

    runs = [("plain", False, None), ("with_fixer", True, None)]
    if form == "f32-sr":
        runs.append(("anomalies", False, "auto"))
    for name, conserve, offsets in runs:
        tc, dd, _ = precision.run(d, grid, fields, init, scheme, steps, form, conserve, rounding, offsets)
        out[name] = {"max_error_vs_fp64": err(tc), "max_step_inventory_defect": dd}
        if offsets is not None:
            out[name]["offsets"] = [float(o) for o in tc.offsets]
        del tc
        torch.cuda.empty_cache()
    log(f"[bench] accuracy over {steps} steps: {out}")
    return out


def serving_run(nreq: int, log, policy: str = "uniform", seed: int = 3):
    """Serving loop on a bounded sample of the dye-release request distribution
    (nx, ny log-uniform in [64, 768], 10-60 layers, 1-200 tracers, 50-2000 steps,
    SUPERBEE or HSIMT; ``nreq`` requests drawn with ``seed``): saturation, FP64
    check of the results of at most 3e10 tracer cell-updates (the FP64 rerun of
    the largest takes minutes), then Poisson arrivals at 80% of saturation.
    Policy "uniform" (every request completes in at most 8 ticks, about 8,
    whatever its size); the inventory fixer runs once per request, at completion."""
    from getm_amd import serve

    reqs = serve.draw_requests(nreq, seed=seed, size_range=(64, 768), nz_range=(10, 60), tracer_range=(1, 200),
                               step_range=(50, 2000))
    domains = {r.rid: r.domain() for r in reqs}
    server = serve.Server(hbm_budget=250e9, max_slots=64, policy=policy)
    rep = serve.sweep(reqs, domains, [0.8], server, verify_results=True, log=log, verify_max_updates=3e10)
    sat, run = rep["saturation"], rep["sweep"][0]
    return {
        "requests": f"{nreq} dye releases drawn with seed {seed}: nx, ny logU[64,768], nz U[10,60], 1-200 tracers, "
                    "50-2000 steps, SUPERBEE/HSIMT, FP32 (stochastic rounding) with FP64 h, inventory fixer per "
                    "request at completion",
        "policy": policy,
        "saturation_updates_per_s": sat["tracer_cell_updates_per_s"],
        "load": run["fraction_of_saturation"],
        "arrival_rate_per_s": run["arrival_rate_per_s"],
        "updates_per_s": run["tracer_cell_updates_per_s"],
        "latency_p50_s": run["latency_p50_s"],
        "latency_p99_s": run["latency_p99_s"],
        "latency_p999_s": run["latency_p999_s"],
        "slowdown_p50": run["slowdown_p50"],
        "slowdown_p99": run["slowdown_p99"],
        "max_inventory_drift": max(sat["max_inventory_drift"], run["max_inventory_drift"]),
        "max_error_vs_fp64": sat.get("max_error_vs_fp64"),
        "verified": sat.get("verified"),
    }


def algorithmic_bytes(kernel: str, n_wet_cols: int, nz: int, s: int, tracers: int) -> float:
    """Compulsory HBM bytes of one launch, counting wet cells/columns only
    (s = bytes per tracer value, tracers = tracers per launch)."""
    cells = n_wet_cols * nz
    tr = tracers * 2 * s  # every tracer read once and written once
    if kernel in ("hpair_vu", "hpair_uv"):  # + h in/out, u, v, hu, hv; 3 masks + 5 metrics per column
        return cells * (tr + 16 + 32) + n_wet_cols * 52
    if kernel == "wcol":  # + h in/out, w at interfaces; mask
        return cells * (tr + 16) + n_wet_cols * (8 * (nz + 1) + 4)
    if kernel == "vdiff_f32":  # + hn, nuh at the nz-1 inner interfaces; mask
        return cells * (tr + 8) + n_wet_cols * (8 * (nz - 1) + 4)
    if kernel == "inventory":  # tracers and h read; area and mask per column
        return cells * (tracers * s + 8) + n_wet_cols * 12
    if kernel == "fix":  # tracers read and written; mask per column
        return cells * tracers * 2 * s + n_wet_cols * 4
    if kernel == "adv_uv_f64":   # f, h read+write; velocity, face thickness; masks + 3 metrics per column
        return cells * (16 + 16 + 16) + n_wet_cols * (8 + 24)
    if kernel == "adv_w_f64":    # f, h read+write; w at interfaces; mask
        return cells * 32 + n_wet_cols * (8 * (nz + 1) + 4)
    if kernel == "vdiff_apply_f64":  # var read+write; hn, auxo, auxn read; mask
        return cells * 40 + n_wet_cols * 4
    if kernel == "vdiff_prepare_f64":  # nuh, hn read; auxo, auxn write
        return cells * 32
    raise KeyError(kernel)


def step_bytes(n_wet_cols: int, nz: int, s: int, tracers: int) -> float:
    """B_step: every tracer read and written once, shared FP64 u, v, hu, hv, ho, hn
    once, w and nuh at interfaces once, 3 masks + 5 metrics per column."""
    cells = n_wet_cols * nz
    return cells * (tracers * 2 * s + 48 + 16 * (nz + 1) / nz) + n_wet_cols * 52


# ------------------------------------------------------------------ CPU baseline
def _cpu_worker(args):
    """One process: a 256x192x40 slab of the coastal workload, 1 tracer, stepped for
    ``seconds``.  impl "ref": the reference's own Fortran sweeps (oracle/_ref, built
    from the reference sources) in the reference call order (v/2, u/2, w, u/2, v/2),
    then the C restatement of the diffusion; impl "c": the C restatement throughout.
    Both give bit-identical tracers."""
    seed, seconds, impl, scheme = args
    import numpy as np

    import oracle
    from getm_amd import inputs

    d = inputs.coastal(nx=NX // 4, ny=NY // 4, nz=NZ, seed=100 + seed, ntracers=1, scheme=scheme)
    half = 0.5 * d.dt
    ua = (d.umask, d.idxu, d.dyu, d.hu, d.u, d.tmask, d.iarea, None, half)
    va = (d.vmask, d.dxv, d.idyv, d.hv, d.v, d.tmask, d.iarea, None, half)
    f = np.ascontiguousarray(d.tracers[0].copy())
    tr = d.tracers.copy()

    def step():
        if impl == "ref":
            h = d.ho.copy()
            oracle.advection_uv(2, d.scheme, *va, h, f, impl="ref")
            oracle.advection_uv(1, d.scheme, *ua, h, f, impl="ref")
            oracle.advection_w(d.scheme, d.w, d.w, d.tmask, d.dt, h, f, impl="ref")
            oracle.advection_uv(1, d.scheme, *ua, h, f, impl="ref")
            oracle.advection_uv(2, d.scheme, *va, h, f, impl="ref")
            oracle.vertical_diffusion(d.nuh, d.dt, f, d.tmask, d.hn, cnpar=1.0)
        else:
            oracle.transport_step(d, tr)

    step()  # warm caches
    n, t0 = 0, time.perf_counter()
    while True:
        step()
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    return d.n_wet_cells * n / el


def cpu_model() -> str:
    """The host CPU's model name (/proc/cpuinfo), or the platform's processor string."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.lower().startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform

    return platform.processor() or "unknown"


def cpu_baseline(seconds: float, cores: int, scheme: int):
    """Summed rate of ``cores`` processes; the reference Fortran sweeps when oracle/_ref
    is built (shipped with the tree), else the C restatement."""
    import oracle

    oracle.lib()
    impl = "ref" if oracle.ref_available() else "c"
    ctx = mp.get_context("spawn")
    with ctx.Pool(cores) as pool:
        rates = pool.map(_cpu_worker, [(i, seconds, impl, scheme) for i in range(cores)])
    return sum(rates), cores, impl


# ------------------------------------------------------------------ GPU
def blob_tracers(ntr, grid, seed, dtype):
    """Tracers generated on the device: each a sum of 4 Gaussian blobs with
    amplitude 10**U(-3, 1) (as inputs.make_tracers(kind="blobs"))."""
    import numpy as np
    import torch

    rng = np.random.default_rng(seed)
    nz, ny_, nx_ = grid.nz_, grid.ny_, grid.nx_
    dev = grid.device
    xs = torch.arange(nx_, device=dev, dtype=torch.float64)
    ys = torch.arange(ny_, device=dev, dtype=torch.float64)
    zs = torch.arange(nz, device=dev, dtype=torch.float64)
    out = torch.empty((ntr, nz, ny_, nx_), dtype=dtype, device=dev)
    for t in range(ntr):
        acc = torch.zeros((nz, ny_, nx_), dtype=torch.float64, device=dev)
        for _ in range(4):
            amp = 10.0 ** rng.uniform(-3.0, 1.0)
            x0, y0, z0 = rng.uniform(0, nx_), rng.uniform(0, ny_), rng.uniform(0, nz)
            sx, sy, sz = rng.uniform(0.03, 0.15) * nx_, rng.uniform(0.03, 0.15) * ny_, rng.uniform(0.1, 0.5) * nz
            gx = torch.exp(-0.5 * ((xs - x0) / sx) ** 2)
            gy = torch.exp(-0.5 * ((ys - y0) / sy) ** 2)
            gz = torch.exp(-0.5 * ((zs - z0) / sz) ** 2)
            acc += amp * gz[:, None, None] * gy[None, :, None] * gx[None, None, :]
        out[t] = acc.to(dtype)
    return out


def blob_tracers_host(ntr, grid, seed):
    """The same tracers generated on the host with numpy (FP64, rounded to FP32),
    yielded one at a time as (index, array) in order; several tracers are
    computed at once in threads.  The parameters are drawn in the order of
    blob_tracers."""
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np

    rng = np.random.default_rng(seed)
    nz, ny_, nx_ = grid.nz_, grid.ny_, grid.nx_
    xs, ys, zs = np.arange(nx_, dtype=np.float64), np.arange(ny_, dtype=np.float64), np.arange(nz, dtype=np.float64)
    params = []
    for t in range(ntr):
        blobs = []
        for _ in range(4):
            amp = 10.0 ** rng.uniform(-3.0, 1.0)
            x0, y0, z0 = rng.uniform(0, nx_), rng.uniform(0, ny_), rng.uniform(0, nz)
            sx, sy, sz = rng.uniform(0.03, 0.15) * nx_, rng.uniform(0.03, 0.15) * ny_, rng.uniform(0.1, 0.5) * nz
            blobs.append((amp, x0, y0, z0, sx, sy, sz))
        params.append(blobs)

    def one(t):
        acc = np.zeros((nz, ny_, nx_), dtype=np.float64)
        for amp, x0, y0, z0, sx, sy, sz in params[t]:
            gx = np.exp(-0.5 * ((xs - x0) / sx) ** 2)
            gy = np.exp(-0.5 * ((ys - y0) / sy) ** 2)
            gz = np.exp(-0.5 * ((zs - z0) / sz) ** 2)
            acc += amp * (gz[:, None, None] * gy[None, :, None]) * gx[None, None, :]
        return acc.astype(np.float32)

    workers = min(16, os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=workers) as ex:
        for t0 in range(0, ntr, workers):
            for t, x in zip(range(t0, min(ntr, t0 + workers)), ex.map(one, range(t0, min(ntr, t0 + workers)))):
                yield t, x


def fp64_run(steps: int, warmup: int, log):
    """The plain FP64 batched step on the coastal grid with temperature and
    salinity, SUPERBEE: tracer cell-updates/s and ms per step."""
    import numpy as np
    import torch

    from getm_amd import inputs
    from getm_amd.core import CENTERS, INTERFACES, Grid
    from getm_amd.tracer import TracerCollection

    d = inputs.coastal(nx=NX, ny=NY, nz=NZ, scheme=5, ntracers=2, tracer_kind="TS")
    grid = Grid.from_domain(d, device="cuda:0")
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    u, v = grid.ugrid.wrap(dev(d.u), CENTERS), grid.vgrid.wrap(dev(d.v), CENTERS)
    w, nuh = grid.wrap(dev(d.w), INTERFACES), grid.wrap(dev(d.nuh), INTERFACES)
    tc = TracerCollection(grid, advection_scheme=5, cnpar=1.0, precision="f64")
    tc.set_values(dev(d.tracers), molecular_diffusivity=0.0)
    for _ in range(warmup):
        tc.advance(d.dt, u, v, w, nuh)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        tc.advance(d.dt, u, v, w, nuh)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    out = {"workload": f"coastal {NX}x{NY}x{NZ}, temperature and salinity, SUPERBEE, dt 30 s, FP64 batched step "
                       "(bit-identical to the reference's algorithm)",
           "value": d.n_wet_columns * NZ * 2 * steps / el, "unit": "tracer cell-updates/s",
           "ms_per_step": el * 1e3 / steps, "steps": steps}
    log(f"[bench] fp64: {out}")
    return out


def run_ensemble(args, log):
    """96 uneven domains (nx, ny logU[64, 768], nz U[10, 60], 1-32 blob tracers, fields as
    the coastal workload from per-domain seeds), SUPERBEE, FP32 with FP64 thickness,
    stepped by one launch per kernel kind for all domains; the timed region replays one
    hipGraph of ``steps`` steps (captured once, after the warm-up steps)."""
    import numpy as np
    import torch

    from getm_amd import inputs
    from getm_amd.autotune import Autotuner
    from getm_amd.core import CENTERS, INTERFACES, Grid
    from getm_amd.ensemble import Ensemble

    torch.cuda.set_device(0)
    specs = inputs.ensemble_specs(ndomains=args.domains, seed=3)
    grids, tracers, keep = [], [], []
    t_gen = time.perf_counter()
    for nx, ny, nz, nt, seed in specs:
        d = inputs.coastal(nx=nx, ny=ny, nz=nz, dx=2000.0, dt=30.0, scheme=5, seed=seed, ntracers=0)
        g = Grid.from_domain(d, device="cuda:0")
        dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
        keep.append((d.dt, g.ugrid.wrap(dev(d.u), CENTERS), g.vgrid.wrap(dev(d.v), CENTERS),
                     g.wrap(dev(d.w), INTERFACES), g.wrap(dev(d.nuh), INTERFACES)))
        grids.append(g)
        tracers.append(blob_tracers(nt, g, seed, torch.float32))
    log(f"[bench] {len(specs)} domains generated in {time.perf_counter() - t_gen:.1f} s")
    ens = Ensemble(grids, [s[3] for s in specs], scheme=5)
    for i, (dt, u, v, w, nuh) in enumerate(keep):
        ens.bind(i, tracers[i], u, v, w, nuh, dt)
    ens.set_rounding(args.rounding)
    ens.commit()
    shape = None
    if args.autotune:
        shape = Autotuner(None).tune_ensemble(ens, reps=2)
        log(f"[bench] autotuner picked {shape}")
    tuning, tile_x = ens.tuning, ens.tile_x
    for _ in range(args.warmup):
        ens.step()
    torch.cuda.synchronize()
    graph = ens.graph_chunk(args.steps)
    graph.replay()  # first replay uploads the graph
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    graph.replay()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    cells = ens.cells
    value = cells * args.steps / elapsed
    # roofline: one more (eager) step with every kernel timed by HIP events on its
    # launch stream; algorithmic bytes summed over the domains
    from getm_amd import _lib

    _lib.profile_enable(True)
    for k in _lib.KERNELS:
        _lib.profile_read(k, reset=True)
    ens.step()
    torch.cuda.synchronize()
    _lib.profile_enable(False)
    ktimes = {k: v for k, v in ((k, _lib.profile_read(k)) for k in _lib.KERNELS) if v[1] > 0}
    dom = max(ktimes, key=lambda k: ktimes[k][0])
    wet_cols = [int((g.mask.all_values == 1).sum().item()) for g in grids]
    alg = sum(algorithmic_bytes(dom, wc, g.nz_, 4, nt) for wc, g, nt in zip(wet_cols, grids, [s[3] for s in specs]))
    achieved = alg / (ktimes[dom][0] / 1e3) / 1e9  # every kernel kind is one launch per step
    log("[bench] ensemble kernel times of one eager step (ms, launches): " +
        ", ".join(f"{k}={v[0]:.3f}/{v[1]}" for k, v in ktimes.items()))
    del ens, graph, tracers, keep
    torch.cuda.empty_cache()
    cpu = None
    if not args.no_cpu:
        cores = args.cpu_cores or min(16, os.cpu_count() or 1)
        rate, cores, impl = cpu_baseline(args.cpu_seconds, cores, 5)
        model = cpu_model()
        cpu = {"value": rate, "unit": "tracer cell-updates/s", "cores": cores,
               "kind": "reference" if impl == "ref" else "port", "cpu_model": model,
               "sample": f"{'the reference Fortran sweeps (oracle/_ref) + the C restatement of diffusion.F90' if impl == 'ref' else 'FP64 oracle'}; "
                         f"{cores} processes on {model}, each stepping a {NX // 4}x{NY // 4}x{NZ} coastal slab, 1 tracer, "
                         f"SUPERBEE, FP64, ~{args.cpu_seconds:.0f} s each (the reference steps one domain and one "
                         "tracer at a time, so its rate per tracer cell does not depend on the ensemble)"}
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "tracer cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded coastal-like domains, device-generated blob tracers)",
        "config": {
            "workload": f"uneven ensemble of {len(specs)} domains (nx, ny logU[64,768], nz U[10,60], 1-32 "
                        f"tracers), SUPERBEE, FP32 tracers with FP64 thickness, {args.steps}-step hipGraph chunk",
            "domains": len(specs),
            "tracer_cells_per_step": cells,
            "tuning": list(shape) if shape else list(tuning),
            "tile_x": tile_x,
            "autotuned": bool(args.autotune),
            "rounding": args.rounding,
            "kernel_ms_per_step": {k: v[0] for k, v in ktimes.items()},
        },
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "traffic_source": None,
                     "algorithmic_bytes": alg},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def _producer_cases(d, grid, torch, np):
    """(name, callable launching the kernel once, algorithmic bytes per launch,
    points per launch, oracle callable on host copies of a slab) for the kernels
    on either side of the step."""
    from getm_amd import depth, momentum, operators
    from getm_amd.core import CENTERS, INTERFACES
    from getm_amd.vertical_coordinates import GVC

    nz, ny_, nx_ = d.hn.shape
    n2, n3 = ny_ * nx_, nz * ny_ * nx_
    wet2 = int((d.tmask != 0).sum())
    wetc = int((d.tmask == 1).sum())
    nxi, nyi = nx_ - 2 * d.halo, ny_ - 2 * d.halo
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    rng = np.random.default_rng(5)
    pk = grid.ugrid.wrap(dev(d.u * d.hu), CENTERS)
    qk = grid.vgrid.wrap(dev(d.v * d.hv), CENTERS)
    w = grid.wrap(dev(np.zeros((nz + 1, ny_, nx_))), INTERFACES)
    mask3 = dev(np.broadcast_to(d.tmask, (nz, ny_, nx_)).astype(np.int32))
    conc = dev(rng.uniform(0.0, 5.0, (nz, ny_, nx_)))
    wsink = dev(np.full((nz, ny_, nx_), -1e-4))
    hn = dev(d.hn)
    src = dev(np.zeros((nz, ny_, nx_)))
    gvc = GVC(nz, ddl=0.5, ddu=1.5, gamma_surf=True, Dgamma=20.0)
    D = grid.D.all_values
    H = grid.wrap(dev(d.H))
    zc = grid.array(CENTERS)
    zf = grid.array(INTERFACES)
    hgvc = torch.empty((nz, ny_, nx_), dtype=torch.float64, device=D.device)
    f3 = dev(rng.standard_normal((nz, ny_, nx_)))
    g3 = torch.empty_like(f3)
    un = grid.ugrid.wrap(dev(d.u), CENTERS)
    Uo = grid.ugrid.wrap(dev(d.u * d.hu), CENTERS)
    hu = dev(d.hu)
    uk, vk = grid.ugrid.array(CENTERS), grid.vgrid.array(CENTERS)
    interp = lambda: grid.wrap(f3, CENTERS).interp(grid.ugrid.wrap(g3, CENTERS))
    return [
        ("w_momentum_3d", lambda: momentum.w_momentum_3d(pk, qk, d.dt, w),
         wetc * (28 + 40 * (nz - 1)), wetc * nz),
        # activity pass: w of the top layer of each interior column (it stops at the
        # first non-zero value); column pass: mask and s (read and written) of every
        # interior cell, and w, the upstream c and h of wet cells
        ("vertical_advection_to_sources",
         lambda: operators.vertical_advection_to_sources(d.halo, d.halo, mask3, conc, wsink, hn, src),
         nxi * nyi * (8 + nz * 20) + wetc * nz * 24, nxi * nyi * nz),
        ("update_gvc", lambda: gvc(D, hgvc, grid.mask.all_values), n2 * 12 + wet2 * nz * 8, n2 * nz),
        ("thickness2vertical_coordinates",
         lambda: depth.thickness2vertical_coordinates(grid.mask, H, grid.hn, zc, zf),
         n2 * 4 + wet2 * (8 + 8 * nz + 8 * nz + 8 * (nz + 1)), n2 * nz),
        ("grid_interp_x", interp, 16 * n3, n3),
        ("reconstruct_transport_change", lambda: momentum.reconstruct_transport_change(un, hu, Uo, d.dt),
         32 * nx_ * nyi * nz, nx_ * nyi * nz),
        ("velocities_from_transports", lambda: momentum.velocities_from_transports(pk, qk, uk, vk),
         2 * 24 * n3, 2 * n3),
    ]


def producers_cpu_baseline(seconds_each: float = 1.5):
    """The C restatement of each producer kernel on one host core, on a 256 x 192 x 40
    slab of the coastal workload: points per second."""
    import numpy as np

    import oracle
    from getm_amd import inputs
    from getm_amd.vertical_coordinates import calculate_sigma

    d = inputs.coastal(nx=256, ny=192, nz=NZ, ntracers=0)
    nz, ny_, nx_ = d.hn.shape
    rng = np.random.default_rng(5)
    n2, n3 = ny_ * nx_, nz * ny_ * nx_
    mask3 = np.ascontiguousarray(np.broadcast_to(d.tmask, (nz, ny_, nx_)).astype(np.int32))
    conc = rng.uniform(0.0, 5.0, (nz, ny_, nx_))
    wsink = np.full((nz, ny_, nx_), -1e-4)
    s = np.zeros((nz, ny_, nx_))
    dbeta = np.ascontiguousarray(calculate_sigma(nz, 0.5, 1.5))
    D = np.ascontiguousarray(d.hn.sum(axis=0))
    h = np.empty((nz, ny_, nx_))
    zc, zf = np.empty((nz, ny_, nx_)), np.empty((nz + 1, ny_, nx_))
    f3 = rng.standard_normal((nz, ny_, nx_))
    g3 = np.empty_like(f3)
    w = np.zeros((nz + 1, ny_, nx_))
    pk, qk = np.ascontiguousarray(d.u * d.hu), np.ascontiguousarray(d.v * d.hv)
    un = np.ascontiguousarray(d.u.copy())
    wetc = int((d.tmask == 1).sum())
    nxi, nyi = nx_ - 2 * d.halo, ny_ - 2 * d.halo
    cases = {
        "w_momentum_3d": (lambda: oracle.w_momentum_3d(d.tmask, d.dyu, d.dxv, d.iarea, d.ho, d.hn, pk, qk, d.dt, w),
                          wetc * nz),
        "vertical_advection_to_sources": (
            lambda: oracle.vertical_advection_to_sources(d.halo, d.halo, mask3, conc, wsink, d.hn, s), nxi * nyi * nz),
        "update_gvc": (lambda: oracle.update_gvc(1.0 / nz, dbeta, 20.0, -1, D, d.tmask, h), n2 * nz),
        "thickness2vertical_coordinates": (
            lambda: oracle.thickness2vertical_coordinates(d.tmask, d.H, d.hn, zc, zf), n2 * nz),
        "grid_interp_x": (lambda: oracle.interp_x(f3, g3, 0), n3),
        "reconstruct_transport_change": (
            lambda: oracle.reconstruct_transport_change(un, d.hu, pk, d.dt, d.halo + 1, d.halo + nyi), nx_ * nyi * nz),
        # numpy's divide, as the reference itself computes it
        "velocities_from_transports": (lambda: (np.divide(pk, d.hu, out=g3), np.divide(qk, d.hv, out=g3)), 2 * n3),
    }
    out = {}
    for name, (fn, pts) in cases.items():
        fn()
        n, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < seconds_each:
            fn()
            n += 1
        out[name] = pts * n / (time.perf_counter() - t0)
    return out


def run_producers(args, log):
    import numpy as np
    import torch

    from getm_amd import inputs
    from getm_amd.core import Grid

    torch.cuda.set_device(0)
    d = inputs.coastal(nx=NX, ny=NY, nz=NZ, ntracers=0)
    grid = Grid.from_domain(d, device="cuda:0")
    grid.set_field("H", d.H)
    cases = _producer_cases(d, grid, torch, np)
    rows = {}
    total_bytes = total_s = 0.0
    for name, fn, nbytes, pts in cases:
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        times = []
        for _ in range(args.steps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            e1.synchronize()
            times.append(e0.elapsed_time(e1) * 1e-3)
        t = float(np.median(times))
        rows[name] = {"ms": t * 1e3, "algorithmic_bytes": nbytes, "achieved_GBs": nbytes / t / 1e9,
                      "frac": nbytes / t / 1e9 / HBM_PEAK_GBS, "points_per_s": pts / t}
        total_bytes += nbytes
        total_s += t
        log(f"[bench] {name}: {t * 1e3:.3f} ms, {nbytes / t / 1e9:.0f} GB/s of algorithmic bytes")
    cpu = None
    if not args.no_cpu:
        rates = producers_cpu_baseline()
        for name, r in rates.items():
            rows[name]["cpu_points_per_s"] = r
            rows[name]["gpu_over_cpu"] = rows[name]["points_per_s"] / r
        cpu = {"value": None, "unit": "points/s per kernel (config.kernels.*.cpu_points_per_s)", "cores": 1,
               "kind": "port", "sample": "the C restatement of each kernel on a 256x192x40 slab of the coastal "
                                         "workload, ~1.5 s each, one host core"}
    out = {
        "metric": "producer kernels: algorithmic bytes / launch time",
        "value": total_bytes / total_s / 1e9,
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded coastal generator)",
        "config": {"workload": "kernels either side of the transport step, coastal 1024x768x40, one launch each "
                               "(median of the timed launches)", "kernels": rows},
        "roofline": None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0,
                    help="timed steps (default 20; the ensemble workload: one 100-step hipGraph chunk)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=["multi", "coastal", "ensemble", "producers"], default="multi",
                    help="multi: 1024x768x40, 256 blob tracers, P2_PDM; coastal: same grid with T and S, SUPERBEE; "
                         "ensemble: 96 uneven domains in one launch per kernel kind, hipGraph chunk")
    ap.add_argument("--domains", type=int, default=96, help="ensemble workload: number of domains")
    ap.add_argument("--autotune", action="store_true", help="ensemble workload: autotune the launch shape first")
    ap.add_argument("--precision", choices=["f32", "f32-sr", "f64"], default="f32")
    ap.add_argument("--tracers", type=int, default=0, help="override the workload's tracer count")
    ap.add_argument("--scheme", type=int, default=0, help="override the workload's scheme")
    ap.add_argument("--per-tracer", action="store_true", help="FP64 per-tracer operator sequence (reference order)")
    ap.add_argument("--tuning", default="", help="tracers per block h,w,d of the batched kernels (0 = all)[,tile height]")
    ap.add_argument("--solver", type=int, default=0,
                    help="wave-PCR column solve for columns of up to this many layers (0: Thomas only)")
    ap.add_argument("--rounding", choices=["stochastic", "nearest"], default="stochastic",
                    help="f32-sr tracer stores: stochastic (default) or round-to-nearest")
    ap.add_argument("--no-conserve", action="store_true",
                    help="FP32 forms: leave the inventory fixer out of the step")
    ap.add_argument("--accuracy-steps", type=int, default=200,
                    help="FP32 runs: steps of the error check against the FP64 method")
    ap.add_argument("--fp64-steps", type=int, default=10, help="steps of the FP64 block (0: skip it)")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--host-inputs", action="store_true",
                    help="generate and encode the tracers on the host and copy them in, so that no torch "
                         "compute kernel runs before the timed steps (hardware-counter runs)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-cores", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-loop run")
    ap.add_argument("--serve-requests", type=int, default=8)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the step runs on a single undivided domain; "
              "sharding over several GPUs is not implemented.", file=sys.stderr)
        sys.exit(2)

    if not args.steps:
        args.steps = 100 if args.workload == "ensemble" else 20
    if args.workload == "ensemble":
        return run_ensemble(args, lambda m: print(m, file=sys.stderr, flush=True))
    if args.workload == "producers":
        return run_producers(args, lambda m: print(m, file=sys.stderr, flush=True))

    import numpy as np
    import torch

    from getm_amd import _lib, inputs
    from getm_amd.core import CENTERS, INTERFACES, Grid
    from getm_amd.tracer import TracerCollection

    if args.workload == "coastal":
        ntr, scheme, kind = 2, 5, "TS"
    else:
        ntr, scheme, kind = 256, 3, "blobs"
    ntr = args.tracers or ntr
    scheme = args.scheme or scheme
    args.tracers, args.scheme = ntr, scheme
    t_gen = time.perf_counter()
    d = inputs.coastal(nx=NX, ny=NY, nz=NZ, scheme=scheme, ntracers=0 if kind == "blobs" else ntr,
                       tracer_kind="TS")
    t_gen = time.perf_counter() - t_gen
    print(f"[bench] inputs generated in {t_gen:.1f} s: {NX}x{NY}x{NZ}, {d.n_wet_columns} wet columns",
          file=sys.stderr, flush=True)
    torch.cuda.set_device(0)
    grid = Grid.from_domain(d, device="cuda:0")
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    u = grid.ugrid.wrap(dev(d.u), CENTERS)
    v = grid.vgrid.wrap(dev(d.v), CENTERS)
    w = grid.wrap(dev(d.w), INTERFACES)
    nuh = grid.wrap(dev(d.nuh), INTERFACES)
    conserve = args.precision != "f64" and not args.no_conserve
    tc = TracerCollection(grid, advection_scheme=scheme, cnpar=1.0, precision=args.precision,
                          batched=not args.per_tracer, conserve=conserve)
    if args.tuning:
        tc.set_tuning(*[int(x) for x in args.tuning.split(",")])
    if args.solver:
        tc.set_solver(args.solver)
    tc.set_rounding(args.rounding)
    if kind == "blobs" and args.host_inputs:
        tc.resize(ntr)
        for t, x in blob_tracers_host(ntr, grid, seed=7):
            tc.set_tracer(t, x)
    elif kind == "blobs":
        tc.set_values(blob_tracers(ntr, grid, seed=7, dtype=tc.dtype))
    elif args.host_inputs:
        tc.set_values(torch.from_numpy(np.ascontiguousarray(d.tracers)))
    else:
        tc.set_values(dev(d.tracers))
    del d.u, d.v, d.w, d.nuh, d.tracers
    s_bytes = {"f32": 6, "f32-sr": 4, "f64": 8}[args.precision]  # stored bytes per tracer value
    # initial values of the tracers whose FP32 error is checked (tracers evolve independently)
    n_check = min(ntr, 16)
    init_check = tc.values[:n_check].clone() if args.precision != "f64" and not args.no_accuracy else None

    def step():
        tc.advance(d.dt, u, v, w, nuh)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    _lib.profile_enable(True)
    for k in _lib.KERNELS:
        _lib.profile_read(k, reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    _lib.profile_enable(False)
    ktimes = {k: _lib.profile_read(k) for k in _lib.KERNELS}
    # profiling adds two event records per launch; report the clock of the timed region itself
    ms_per_step = elapsed * 1e3 / args.steps
    n_wet_cols = d.n_wet_columns
    updates = n_wet_cols * NZ * args.tracers * args.steps
    value = updates / elapsed
    ktimes = {k: v for k, v in ktimes.items() if v[1] > 0}
    dom = max(ktimes, key=lambda k: ktimes[k][0])
    dom_ms, dom_n = ktimes[dom]
    avg_s = dom_ms / 1e3 / max(dom_n, 1)
    per_launch_tracers = 1 if dom.endswith("_f64") and dom != "vdiff_prepare_f64" else args.tracers
    alg = algorithmic_bytes(dom, n_wet_cols, NZ, s_bytes, per_launch_tracers)
    achieved = alg / avg_s / 1e9
    b_step = step_bytes(n_wet_cols, NZ, s_bytes, args.tracers)
    print("[bench] kernel times (ms total, launches): " +
          ", ".join(f"{k}={v[0]:.2f}/{v[1]}" for k, v in ktimes.items()), file=sys.stderr, flush=True)

    log = lambda m: print(m, file=sys.stderr, flush=True)
    accuracy = None
    if init_check is not None:
        accuracy = accuracy_check(d, grid, (u, v, w, nuh), init_check, scheme, args.accuracy_steps, args.precision,
                                  args.rounding, log)
        del init_check
    workload_key = f"{args.workload}-{args.precision}-T{args.tracers}-s{args.scheme}" + (
        "-pertracer" if args.per_tracer else "")
    rec = recorded_traffic(dom, workload_key)
    del tc, u, v, w, nuh, grid
    torch.cuda.empty_cache()
    fp64 = fp64_run(args.fp64_steps, 2, log) if args.fp64_steps > 0 else None
    torch.cuda.empty_cache()
    serving = None
    if not args.no_serve:
        serving = serving_run(args.serve_requests, log)

    cpu = None
    if not args.no_cpu:
        cores = args.cpu_cores or min(16, os.cpu_count() or 1)
        rate, cores, impl = cpu_baseline(args.cpu_seconds, cores, scheme)
        what = ("the reference's Fortran u2d/v2d/w3d (oracle/_ref, built from the reference sources) in the "
                "reference call order + the C restatement of diffusion.F90 (not buildable here)"
                if impl == "ref" else "FP64 oracle (C restatement of the reference Fortran)")
        model = cpu_model()
        cpu = {"value": rate, "unit": "tracer cell-updates/s", "cores": cores,
               "kind": "reference" if impl == "ref" else "port", "cpu_model": model,
               "sample": f"{what}; {cores} processes on {model}, each stepping a {NX // 4}x{NY // 4}x{NZ} "
                         f"coastal slab, 1 tracer, scheme {scheme}, FP64, ~{args.cpu_seconds:.0f} s each (the "
                         "reference steps tracers one at a time, so its rate per tracer does not depend on their "
                         "number)"}

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "tracer cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f32": "f32+f64", "f32-sr": "f32", "f64": "f64"}[args.precision],
        "data": "synthetic (seeded coastal generator, continuity-consistent)",
        "config": {
            "workload": (("multi-tracer batch: " if kind == "blobs" else "single large domain: ") +
                         f"coastal {NX}x{NY}x{NZ}, 2 km grid, 30% land, scheme {scheme}, dt 30 s, {ntr} " +
                         ("blob tracers" if kind == "blobs" else "tracer(s) (temperature, salinity)") + ", " +
                         FORM_NAMES[args.precision] + (" with the inventory fixer in every step" if conserve else "") +
                         (", per-tracer operators" if args.per_tracer else ", batched step")),
            "grid": [NX, NY, NZ],
            "wet_columns": n_wet_cols,
            "tracers": args.tracers,
            "scheme": args.scheme,
            "precision": args.precision,
            "conserve": conserve,
            "bytes_per_tracer_value": s_bytes,
            "rounding": args.rounding if args.precision == "f32-sr" else None,
            "step_hbm_fraction": b_step / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBS,
            "kernel_ms_per_step": {k: v[0] / args.steps for k, v in ktimes.items()},
            "tuning": args.tuning or "default",
            "pcr_max_nz": args.solver,
            "key": workload_key,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": rec[0] if rec else None,
            "traffic_source": rec[1] if rec else None,
            "algorithmic_bytes": alg,
        },
        "accuracy": accuracy,
        "fp64": fp64,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
