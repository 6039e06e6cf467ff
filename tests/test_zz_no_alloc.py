"""transport_conserve_noalloc, the inventory fixer with its working memory in a
fixed array inside the library, against transport_conserve through the C-ABI:

- bit-identical tracers on two calls in a row, the second with fewer tracers
  on a smaller grid (the array keeps the first call's partial sums), for both
  FP32 forms;
- the size limit: 4 Mi doubles of working memory, i.e. with B blocks of 64 x 4
  columns per layer, floor(4 Mi / (4 B + 1)) compensated tracers or
  floor(4 Mi / (8 B + 3)) stochastic-rounding tracers; one tracer more is
  rejected with GETM_ERR_CAPACITY before anything runs.
"""
import numpy as np
import pytest
import torch

from getm_amd import _lib, inputs
from getm_amd.core import CENTERS, INTERFACES, Grid

pytestmark = pytest.mark.gpu

GETM_ERR_CAPACITY = 5  # include/getm_amd.h
LIMIT = 4 << 20
PER_TRACER = {"f32": (4, 1), "f32-sr": (8, 3)}  # doubles per tracer: a B + b


def _collection(nx, ny, nz, seed, ntr, precision):
    from getm_amd.tracer import TracerCollection

    d = inputs.coastal(nx=nx, ny=ny, nz=nz, seed=seed, ntracers=2, tracer_kind="TS")
    grid = Grid.from_domain(d)
    tc = TracerCollection(grid, advection_scheme=3, precision=precision)
    gen = torch.Generator(device=grid.device).manual_seed(seed)
    vals = 1.0 + 9.0 * torch.rand((ntr, grid.nz_, grid.ny_, grid.nx_), generator=gen, dtype=torch.float64,
                                  device=grid.device)
    tc.set_values(vals)
    return d, grid, tc


def _drifted(nx, ny, nz, seed, ntr, precision):
    """A collection one step on, and the inventories it had before the step."""
    d, grid, tc = _collection(nx, ny, nz, seed, ntr, precision)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(grid.device)
    u, v = grid.ugrid.wrap(dev(d.u), CENTERS), grid.vgrid.wrap(dev(d.v), CENTERS)
    w, nuh = grid.wrap(dev(d.w), INTERFACES), grid.wrap(dev(d.nuh), INTERFACES)
    target = tc.inventory_device(grid.ho.all_values)
    tc.advance(d.dt, u, v, w, nuh)
    return tc, target


def _same_bits(a, b, what):
    a, b = a.view(torch.int16), b.view(torch.int16)
    ndiff = int((a != b).sum())
    assert ndiff == 0, f"{what}: {ndiff} of {a.numel()} 16-bit words differ"


def _call(name, tc, ntr, target):
    return getattr(_lib.load(), name)(tc.p, ntr, _lib.ptr(tc._buf), _lib.ptr(target), _lib.stream_handle())


@pytest.mark.parametrize("precision", ["f32", "f32-sr"])
def test_noalloc_fixer_equals_fixer_on_two_calls(precision):
    cases = [_drifted(200, 150, 20, 3, 6, precision), _drifted(90, 60, 10, 4, 2, precision)]
    start = [tc._buf.clone() for tc, _ in cases]
    expected = []
    for (tc, target), s in zip(cases, start):
        _lib.check(_call("transport_conserve", tc, len(tc), target), "transport_conserve")
        expected.append(tc._buf.clone())
        tc._buf.copy_(s)
    torch.cuda.synchronize()
    for tc, target in cases:  # two calls in a row, the second smaller
        _lib.check(_call("transport_conserve_noalloc", tc, len(tc), target), "transport_conserve_noalloc")
    torch.cuda.synchronize()
    for n, ((tc, target), e) in enumerate(zip(cases, expected)):
        _same_bits(tc._buf, e, f"call {n}")
        after = tc.inventory_device()
        assert float(((after - target).abs() / target.abs()).max()) <= 1e-12, n


@pytest.mark.parametrize("precision", ["f32", "f32-sr"])
def test_noalloc_fixer_size_limit(precision):
    nx, ny = 1024, 768
    blocks = -(-(nx + 4) // 64) * -(-(ny + 4) // 4)
    a, b = PER_TRACER[precision]
    nmax = LIMIT // (a * blocks + b)
    _, _, big = _collection(nx, ny, 1, 5, nmax + 1, precision)
    before = big._buf.clone()
    target = big.inventory_device() * (1.0 + 1e-9)
    assert _call("transport_conserve_noalloc", big, nmax + 1, target) == GETM_ERR_CAPACITY
    _same_bits(big._buf, before, "rejected call")
    del big, before, target
    _, _, tc = _collection(nx, ny, 1, 6, nmax, precision)
    target = tc.inventory_device() * (1.0 + 1e-9)
    start = tc._buf.clone()
    _lib.check(_call("transport_conserve", tc, nmax, target), "transport_conserve")
    expected = tc._buf.clone()
    tc._buf.copy_(start)
    _lib.check(_call("transport_conserve_noalloc", tc, nmax, target), "transport_conserve_noalloc")
    _same_bits(tc._buf, expected, f"{nmax} tracers")
    assert not torch.equal(tc._buf, start)  # the fixer did change cells
