"""Input producer w_momentum_3d (c_w_momentum_3d) against the oracle, bit for bit."""
import numpy as np
import pytest

from getm_amd import inputs
from helpers import assert_bitwise
from test_gpu_parity import dev, host, make_grid
from test_oracle import _momentum_fields

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed", [1, 6])
def test_w_momentum_3d_bitwise(oracle_mod, seed):
    from getm_amd.core import CENTERS, INTERFACES
    from getm_amd.momentum import velocities_from_transports, w_momentum_3d

    d = inputs.coastal(nx=70, ny=52, nz=9, seed=seed, ntracers=1)
    pk, qk = _momentum_fields(d)
    rng = np.random.default_rng(seed)
    w0 = rng.standard_normal((d.nz + 1,) + d.ho.shape[1:]) * 1e-3  # nonzero start: untouched entries stay
    ref = w0.copy()
    oracle_mod.w_momentum_3d(d.tmask, d.dyu, d.dxv, d.iarea, d.ho, d.hn, pk, qk, d.dt, ref)
    grid = make_grid(d)
    w = grid.wrap(dev(w0), INTERFACES)
    w_momentum_3d(grid.ugrid.wrap(dev(pk), CENTERS), grid.vgrid.wrap(dev(qk), CENTERS), d.dt, w)
    assert_bitwise(host(w.all_values), ref, "w")
    # velocities from transports
    uk = grid.ugrid.wrap(dev(np.zeros_like(pk)), CENTERS)
    vk = grid.vgrid.wrap(dev(np.zeros_like(qk)), CENTERS)
    velocities_from_transports(grid.ugrid.wrap(dev(pk), CENTERS), grid.vgrid.wrap(dev(qk), CENTERS), uk, vk)
    with np.errstate(divide="ignore", invalid="ignore"):
        assert_bitwise(host(uk.all_values), pk / d.hu, "uk")
        assert_bitwise(host(vk.all_values), qk / d.hv, "vk")
