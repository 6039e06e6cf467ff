"""FP32 form against the FP64 method on the large workloads.

    python -m getm_amd.precision --workload coastal --steps 500 --schemes 1,2,3,4,5,6
    python -m getm_amd.precision --workload multi --steps 200

Runs the batched step from identical inputs on the GPU: FP64 (the plain method,
bit-identical to the reference algorithm -- see the GPU parity tests) against
the FP32 forms, all with FP64 thickness: the compensated form (FP32 values plus
int16 residuals, "f32"), with the inventory fixer ("f32_fixed"); FP32 stores
rounded stochastically ("f32-sr"), with the fixer ("f32-sr_fixed") and as
anomalies about per-tracer offsets (``offsets="auto"``, "f32-sr_offsets"); and
FP32 stores rounded to nearest ("f32-sr_nearest"), compared on their absolute
values and absolute inventories.  Reports,
per limiter, the error max |c32 - c64| / max |c64| over wet cells (max and
median over tracers) after the run, and the largest per-step conservation
defect |sum(area hn c_after) - sum(area ho c_before)| / |sum(area ho c_before)|
(FP64 compensated sums) over tracers and steps.

coastal: 1024 x 768 x 40, T and S, dt 30 s.  multi: the same grid, 256 blob
tracers, P2_PDM.  The inputs are continuity-consistent for one step ho -> hn,
which every step repeats.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np
import torch

from . import inputs
from .core import CENTERS, INTERFACES, Grid
from .tracer import TracerCollection


def _blob_tracers(ntr, grid, seed, dtype):
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
            acc += amp * (torch.exp(-0.5 * ((zs - z0) / sz) ** 2)[:, None, None] *
                          torch.exp(-0.5 * ((ys - y0) / sy) ** 2)[None, :, None] *
                          torch.exp(-0.5 * ((xs - x0) / sx) ** 2)[None, None, :])
        out[t] = acc.to(dtype)
    return out


def run(d, grid, fields, init, scheme, steps, precision, conserve, rounding="stochastic", offsets=None):
    """Advance ``steps`` steps; return the collection, the largest per-step
    conservation defect |sum(area hn c_after) - sum(area ho c_before)| / |..|
    over tracers and steps (of the absolute values), and the wall time."""
    u, v, w, nuh = fields
    tc = TracerCollection(grid, advection_scheme=scheme, cnpar=1.0, precision=precision, conserve=conserve)
    tc.set_rounding(rounding)
    tc.set_values(init, offsets=offsets)
    ho = grid.ho.all_values
    worst = torch.zeros((), dtype=torch.float64, device=grid.device)
    t0 = time.perf_counter()
    for _ in range(steps):
        before = tc.inventory_device(ho, absolute=True)
        tc.advance(d.dt, u, v, w, nuh)
        after = tc.inventory_device(absolute=True)
        worst = torch.maximum(worst, ((after - before).abs() / before.abs()).max())
    torch.cuda.synchronize()
    return tc, float(worst.item()), time.perf_counter() - t0


# (report name, precision, inventory fixer, rounding of the FP32 stores, offsets)
FORMS = (("f32", "f32", False, "nearest", None),
         ("f32_fixed", "f32", True, "nearest", None),
         ("f32-sr", "f32-sr", False, "stochastic", None),
         ("f32-sr_fixed", "f32-sr", True, "stochastic", None),
         ("f32-sr_offsets", "f32-sr", False, "stochastic", "auto"),
         ("f32-sr_nearest", "f32-sr", False, "nearest", None))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=["coastal", "multi"], default="coastal")
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--schemes", default="")
    ap.add_argument("--tracers", type=int, default=0)
    ap.add_argument("--forms", default="", help="comma-separated subset of the FP32 forms (default: all)")
    ap.add_argument("--out", default="")
    args = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    kind = "TS" if args.workload == "coastal" else "blobs"
    schemes = [int(x) for x in args.schemes.split(",") if x] or ([5] if kind == "TS" else [3])
    ntr = args.tracers or (2 if kind == "TS" else 256)
    d = inputs.coastal(nx=1024, ny=768, nz=40, ntracers=ntr if kind == "TS" else 0, tracer_kind="TS")
    grid = Grid.from_domain(d)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(grid.device)
    fields = (grid.ugrid.wrap(dev(d.u), CENTERS), grid.vgrid.wrap(dev(d.v), CENTERS),
              grid.wrap(dev(d.w), INTERFACES), grid.wrap(dev(d.nuh), INTERFACES))
    # FP32-representable initial values, so both precisions start from the same state
    init = dev(d.tracers).float() if kind == "TS" else _blob_tracers(ntr, grid, 7, torch.float32)
    del d.tracers
    wet = torch.from_numpy(np.broadcast_to(d.tmask == 1, d.ho.shape).copy()).to(grid.device)
    report = {"workload": f"{args.workload} 1024x768x40, {ntr} tracers, dt {d.dt} s", "steps": args.steps,
              "schemes": {}}
    for scheme in schemes:
        ref, d64, el64 = run(d, grid, fields, init, scheme, args.steps, "f64", False)
        log(f"[precision] scheme {scheme} f64: {args.steps} steps in {el64:.1f} s, per-step defect {d64:.1e}")
        rows = {"f64": {"max_step_conservation_defect": d64, "seconds": el64}}
        for name, form, conserve, rounding, offsets in FORMS:
            if args.forms and name not in args.forms.split(","):
                continue
            tc, dd, el = run(d, grid, fields, init, scheme, args.steps, form, conserve, rounding, offsets)
            err = []
            for t in range(ntr):
                r = ref.values[t][wet]
                err.append(float((tc.absolute(t)[wet] - r).abs().max() / r.abs().max()))
            rows[name] = {"max_error": max(err), "median_error": float(np.median(err)),
                          "max_step_conservation_defect": dd, "seconds": el}
            log(f"[precision] scheme {scheme} {name}: {args.steps} steps in {el:.1f} s, max error {max(err):.2e}, "
                f"per-step defect {dd:.1e}")
            del tc
            torch.cuda.empty_cache()
        report["schemes"][scheme] = rows
        del ref
        torch.cuda.empty_cache()
    print(json.dumps(report))
    if args.out:
        with open(args.out, "w") as fh:
            json.dump(report, fh, indent=1)


if __name__ == "__main__":
    main()
