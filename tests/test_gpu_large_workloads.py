"""The FP32 form at full size against the FP64 batched step (bit-identical to the
oracle, tests/test_gpu_batched.py), on the 1024 x 768 x 40 coastal grid:

- 256 passive blob tracers, P2_PDM, 200 steps, the inventory fixer on: every
  tracer within 1e-5 of FP64 (max |error| / max |value|), every per-step
  inventory defect within 1e-12;
- temperature and salinity, each of the six limiters, 500 steps.

Both FP32 runs start from the same FP32-representable values as the FP64 run.
The 256 tracers run in four batches of 64 (tracers are independent problems;
that bounds the memory of the FP64 reference run).
"""
import numpy as np
import pytest
import torch

from getm_amd import inputs, precision
from getm_amd.core import CENTERS, INTERFACES, Grid

pytestmark = pytest.mark.gpu


def _coastal(ntracers, kind):
    d = inputs.coastal(nx=1024, ny=768, nz=40, ntracers=ntracers, tracer_kind=kind, scheme=3)
    grid = Grid.from_domain(d)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(grid.device)
    fields = (grid.ugrid.wrap(dev(d.u), CENTERS), grid.vgrid.wrap(dev(d.v), CENTERS),
              grid.wrap(dev(d.w), INTERFACES), grid.wrap(dev(d.nuh), INTERFACES))
    wet = torch.from_numpy(np.broadcast_to(d.tmask == 1, d.ho.shape).copy()).to(grid.device)
    return d, grid, fields, wet


def _errors(tc32, ref, wet):
    out = []
    for t in range(len(ref)):
        r = ref.absolute(t)[wet]
        out.append(float((tc32.absolute(t)[wet] - r).abs().max() / r.abs().max()))
    return out


@pytest.fixture(scope="module")
def blobs_domain():
    return _coastal(0, "blobs")


@pytest.fixture(scope="module")
def ts_domain():
    return _coastal(2, "TS")


def test_256_blob_tracers_200_steps_within_1e5_and_conserved(blobs_domain):
    d, grid, fields, wet = blobs_domain
    errs, defects = [], []
    for batch in range(4):
        init = precision._blob_tracers(64, grid, 7 + batch, torch.float32)
        ref, d64, _ = precision.run(d, grid, fields, init, 3, 200, "f64", False)
        tc, d32, el = precision.run(d, grid, fields, init, 3, 200, "f32", True)
        e = _errors(tc, ref, wet)
        print(f"batch {batch}: max error {max(e):.2e}, median {np.median(e):.2e}, per-step inventory defect "
              f"{d32:.1e} (FP64 {d64:.1e}), {el:.1f} s for 200 FP32 steps with the fixer and two inventories each",
              flush=True)
        errs += e
        defects.append(d32)
        del ref, tc, init
        torch.cuda.empty_cache()
    print(f"256 tracers: max error {max(errs):.2e}, median {np.median(errs):.2e}, "
          f"largest per-step defect {max(defects):.1e}")
    assert max(errs) <= 1e-5, max(errs)
    assert max(defects) <= 1e-12, max(defects)


@pytest.mark.parametrize("scheme", range(1, 7))
def test_temperature_salinity_500_steps_within_1e5(scheme, ts_domain):
    d, grid, fields, wet = ts_domain
    init = torch.from_numpy(np.ascontiguousarray(d.tracers)).to(grid.device).float()
    ref, _, _ = precision.run(d, grid, fields, init, scheme, 500, "f64", False)
    tc, d32, _ = precision.run(d, grid, fields, init, scheme, 500, "f32", False)
    e = _errors(tc, ref, wet)
    print(f"scheme {scheme}: T {e[0]:.2e}, S {e[1]:.2e}, per-step inventory defect without the fixer {d32:.1e}")
    assert max(e) <= 1e-5, e
    assert d32 <= 1e-12, d32
