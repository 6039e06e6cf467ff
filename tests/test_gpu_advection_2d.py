"""2D split advection, Advection.__call__ (python/pygetm/operators.py:52-93), on
the T grid against the oracle's composition of single sweeps, bit for bit, for
the three split modes and both sweep orders:
    FULL:        a(dt), b(dt)
    HALF:        a(dt/2), b(dt), a(dt/2)
    HALF_ALWAYS: a(dt/2), b(dt/2), b(dt/2), a(dt/2)
with (a, b) = (v, u), or (u, v) when ufirst is set, and the depth D as the
thickness that the sweeps carry from one to the next."""
import numpy as np
import pytest

from getm_amd import inputs
from helpers import assert_bitwise
from test_gpu_parity import dev, host, make_grid

pytestmark = pytest.mark.gpu

SEQUENCES = {
    "FULL": [(0, 1.0), (1, 1.0)],
    "HALF": [(0, 0.5), (1, 1.0), (0, 0.5)],
    "HALF_ALWAYS": [(0, 0.5), (1, 0.5), (1, 0.5), (0, 0.5)],
}


@pytest.mark.parametrize("ufirst", [False, True])
@pytest.mark.parametrize("mode", list(SEQUENCES))
@pytest.mark.parametrize("scheme", [5, 3, 1])
def test_split_2d_bitwise(oracle_mod, scheme, mode, ufirst):
    from getm_amd.operators import Advection, AdvectionSplit

    d = inputs.tiny(nx=37, ny=29, nz=1, seed=40 + scheme, scheme=scheme, ntracers=1, cfl=0.4)
    grid = make_grid(d)
    adv = Advection(grid, scheme, split_2d=getattr(AdvectionSplit, mode))
    adv.ufirst = ufirst
    u2, v2, f2 = d.u[0], d.v[0], d.tracers[0][0]
    u = grid.ugrid.wrap(dev(u2))
    v = grid.vgrid.wrap(dev(v2))
    var = grid.wrap(dev(f2))
    adv(u, v, d.dt, var)

    D = d.hn.sum(axis=0)[None].copy()  # Grid.from_domain's D of each grid: the column sums
    f = f2[None].copy()
    Du, Dv = d.hu.sum(axis=0)[None], d.hv.sum(axis=0)[None]
    sweeps = [(2, (d.vmask, d.dxv, d.idyv, Dv, v2[None])), (1, (d.umask, d.idxu, d.dyu, Du, u2[None]))]
    if ufirst:
        sweeps.reverse()
    for which, frac in SEQUENCES[mode]:
        direction, fields = sweeps[which]
        oracle_mod.advection_uv(direction, scheme, *fields, d.tmask, d.iarea, None, frac * d.dt, D, f)
    assert_bitwise(host(var.all_values), f[0], "tracer")
    assert_bitwise(host(adv.D), D[0], "depth carried through the sweeps")
    assert not np.array_equal(f[0], f2), "the flow moves the tracer"
