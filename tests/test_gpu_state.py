"""Model-state checks of the tracer collection on the GPU:

- the device-side finiteness check (the tracers' part of the reference's
  Simulation.check_finite, python/pygetm/simulation.py:348-410): counts and
  first cell of non-finite values in unmasked interior cells, land and halo
  cells ignored, in every storage form;
- restart of the whole state the step depends on, including the layer
  thicknesses ho and hn (the reference keeps T.ho in its restarts,
  simulation.py:620-622), while the thicknesses change from step to step.
"""
import logging

import numpy as np
import pytest
import torch

from getm_amd import inputs
from getm_amd.core import CENTERS, INTERFACES, Grid

from helpers import assert_bitwise

pytestmark = pytest.mark.gpu


def _setup(nx, ny, nz, seed, ntracers):
    d = inputs.coastal(nx=nx, ny=ny, nz=nz, seed=seed, ntracers=ntracers, tracer_kind="TS")
    grid = Grid.from_domain(d)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(grid.device)
    fields = (grid.ugrid.wrap(dev(d.u), CENTERS), grid.vgrid.wrap(dev(d.v), CENTERS),
              grid.wrap(dev(d.w), INTERFACES), grid.wrap(dev(d.nuh), INTERFACES))
    return d, grid, dev, fields


@pytest.mark.parametrize("precision", ["f64", "f32", "f32-sr"])
def test_check_finite_counts_unmasked_interior_cells(precision, caplog):
    from getm_amd.tracer import TracerCollection

    d, grid, dev, (u, v, w, nuh) = _setup(40, 32, 6, 5, 3)
    tc = TracerCollection(grid, advection_scheme=3, precision=precision)
    tc.set_values(dev(d.tracers))
    tc.advance(d.dt, u, v, w, nuh)
    assert tc.check_finite()
    assert (tc.non_finite()[:, 0] == 0).all()

    mask = grid.mask.all_values.cpu().numpy()
    hx, hy = grid.halox, grid.haloy
    interior = np.zeros_like(mask, dtype=bool)
    interior[hy:hy + grid.ny, hx:hx + grid.nx] = True
    wet = np.argwhere(interior & (mask == 1))
    land = np.argwhere(interior & (mask == 0))
    assert len(wet) > 10 and len(land) > 0
    (j0, i0), (j1, i1), (j2, i2) = wet[len(wet) // 2], wet[3], wet[-2]
    jl, il = land[0]
    vals = tc.values  # FP64 values, or the FP32 values of the FP32 forms
    with torch.no_grad():
        vals[0, 2, jl, il] = float("nan")         # land: ignored
        vals[0, 1, 0, 0] = float("inf")           # halo corner: ignored
        vals[1, 3, j0, i0] = float("nan")
        vals[2, 4, j2, i2] = float("-inf")
        vals[2, 1, j1, i1] = float("nan")
    caplog.set_level(logging.ERROR)
    assert not tc.check_finite()
    res = tc.non_finite()
    assert res[0].tolist() == [0, -1, -1, -1]
    assert res[1].tolist() == [1, 3, j0, i0]
    first = min((4, j2, i2), (1, j1, i1))   # memory order: k, then j, then i
    assert res[2].tolist() == [2, *first]
    msgs = [r.getMessage() for r in caplog.records]
    assert len(msgs) == 2 and "2 non-finite values" in msgs[1], msgs


@pytest.mark.parametrize("precision", ["f64", "f32", "f32-sr"])
def test_restart_with_changing_thickness_continues_bit_for_bit(tmp_path, precision):
    """The thicknesses change every step (ho <- hn, hn scaled); a collection on a
    fresh grid whose ho and hn still hold the initial thicknesses continues bit
    for bit after loading a state saved at step 3, because the state carries
    the grid's ho and hn."""
    from getm_amd.tracer import TracerCollection

    def run(grid, tc, u, v, w, nuh, dt, steps, start):
        for n in range(start, start + steps):
            grid.ho.all_values.copy_(grid.hn.all_values)
            grid.hn.all_values.mul_(1.0 + 1e-3 * ((n % 3) - 1))
            tc.advance(dt, u, v, w, nuh)

    d, grid, dev, fields = _setup(52, 40, 8, 19, 2)
    ref = TracerCollection(grid, advection_scheme=2, precision=precision)
    ref.set_values(dev(d.tracers), molecular_diffusivity=1e-6)
    run(grid, ref, *fields, d.dt, 7, 0)
    expected = ref.absolute().cpu().numpy()

    d, grid, dev, fields = _setup(52, 40, 8, 19, 2)
    first = TracerCollection(grid, advection_scheme=2, precision=precision)
    first.set_values(dev(d.tracers), molecular_diffusivity=1e-6)
    run(grid, first, *fields, d.dt, 3, 0)
    path = str(tmp_path / "state.npz")
    first.save(path)

    d, grid2, dev, fields2 = _setup(52, 40, 8, 19, 2)
    assert not torch.equal(grid2.hn.all_values, grid.hn.all_values)
    second = TracerCollection(grid2, advection_scheme=2, precision=precision)
    second.load(path)
    assert torch.equal(grid2.ho.all_values, grid.ho.all_values)
    assert torch.equal(grid2.hn.all_values, grid.hn.all_values)
    run(grid2, second, *fields2, d.dt, 4, 3)
    assert_bitwise(second.absolute().cpu().numpy(), expected, "restarted run")


def test_state_without_thickness_is_rejected():
    from getm_amd.tracer import TracerCollection

    d, grid, dev, _ = _setup(20, 16, 4, 2, 1)
    tc = TracerCollection(grid, precision="f64")
    tc.set_values(dev(d.tracers))
    st = tc.state_dict()
    del st["ho"]
    with pytest.raises(ValueError, match="ho"):
        tc.load_state_dict(st)


def _same_storage(a, b):
    assert_bitwise(a.values.cpu().numpy(), b.values.cpu().numpy(), "values")
    if a.residuals is not None:
        assert np.array_equal(a.residuals.cpu().numpy(), b.residuals.cpu().numpy()), "residuals"


@pytest.mark.parametrize("precision", ["f64", "f32", "f32-sr"])
def test_host_inputs_store_the_bits_of_device_inputs(precision):
    """Host data (encoded on the host, then copied) and device data give the same
    storage bit for bit, through set_values, resize + set_tracer, add and the
    Tracer.all_values setter; the compensated form's residuals included."""
    from getm_amd.tracer import TracerCollection

    d, grid, dev, (u, v, w, nuh) = _setup(40, 32, 6, 7, 3)
    x = d.tracers * (1.0 + 1e-9 * np.arange(d.tracers.size).reshape(d.tracers.shape) % 7)  # not FP32-exact
    a = TracerCollection(grid, advection_scheme=3, precision=precision)
    a.set_values(dev(x))
    b = TracerCollection(grid, advection_scheme=3, precision=precision)
    b.set_values(torch.from_numpy(x))
    _same_storage(b, a)
    c = TracerCollection(grid, advection_scheme=3, precision=precision)
    c.resize(3)
    for t in range(3):
        c.set_tracer(t, x[t])
    _same_storage(c, a)
    e = TracerCollection(grid, advection_scheme=3, precision=precision)
    for t in range(3):
        e.add(f"t{t}", data=dev(x[t]) if t % 2 else x[t])
    _same_storage(e, a)
    e[1].all_values = x[0]
    e[0].all_values = dev(x[1])
    a.set_values(dev(x[[1, 0, 2]]))
    _same_storage(e, a)
    # resize keeps the first tracers and their attributes
    e[0].molecular_diffusivity = 1e-7
    e.resize(2)
    assert len(e) == 2 and e[0].molecular_diffusivity == 1e-7
    assert_bitwise(e.absolute()[:2].cpu().numpy(), a.absolute()[:2].cpu().numpy())
