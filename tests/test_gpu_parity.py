"""GPU parity: every C-ABI entry point against the FP64 oracle, bit for bit.

Inputs are seeded and identical for both sides; the FP64 HIP kernels must
reproduce the oracle's (and thereby the reference Fortran's) bits, including
NaN positions, masks and limiter branch selection.
"""
import numpy as np
import pytest

from getm_amd import inputs
from helpers import GOLDEN_CASES, assert_bitwise, golden_domain

pytestmark = pytest.mark.gpu
SCHEMES = range(1, 7)


def _torch():
    import torch

    return torch


def dev(a):
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    _torch().cuda.synchronize()
    return t.detach().cpu().numpy()


def make_grid(d):
    from getm_amd.core import Grid

    class D:
        pass

    full = D()
    full.__dict__.update(d.__dict__)
    if not hasattr(d, "hn"):
        full.hn = d.ho
    if not hasattr(d, "area"):
        full.area = 1.0 / d.iarea
    return Grid.from_domain(full)


def arrays(grid, d, tracer):
    from getm_amd.core import CENTERS, INTERFACES

    u = grid.ugrid.wrap(dev(d.u), CENTERS)
    v = grid.vgrid.wrap(dev(d.v), CENTERS)
    w = grid.wrap(dev(d.w), INTERFACES)
    var = grid.wrap(dev(tracer), CENTERS)
    return u, v, w, var


def _domains():
    out = []
    for seed, kw in ((31, dict(cfl=0.1, wscale=1.0)), (32, dict(cfl=0.9, wscale=0.0)),
                     (33, dict(cfl=0.3, wscale=2.0, mask_codes=True))):
        out.append((seed, kw))
    return out


@pytest.mark.parametrize("seed,kw", _domains())
@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("direction", [1, 2])
@pytest.mark.parametrize("with_ah", [False, True])
def test_uv_sweep_bitwise(oracle_mod, seed, kw, scheme, direction, with_ah):
    from getm_amd.operators import Advection

    d = inputs.tiny(nx=37, ny=29, nz=7, seed=seed, ntracers=1, **kw)
    grid = make_grid(d)
    u, v, w, var = arrays(grid, d, d.tracers[0])
    adv = Advection(grid, scheme)
    adv.h.copy_(dev(d.ho))
    ah_np = np.random.default_rng(seed).uniform(0, 50, size=d.idxu.shape) if with_ah else None
    if direction == 1:
        ah = grid.ugrid.wrap(dev(ah_np)) if with_ah else None
        adv.u_3d(u, 0.5 * d.dt, var, ah)
    else:
        ah = grid.vgrid.wrap(dev(ah_np)) if with_ah else None
        adv.v_3d(v, 0.5 * d.dt, var, ah)
    h_ref, f_ref = d.ho.copy(), d.tracers[0].copy()
    if direction == 1:
        oracle_mod.advection_uv(1, scheme, d.umask, d.idxu, d.dyu, d.hu, d.u, d.tmask, d.iarea, ah_np,
                                0.5 * d.dt, h_ref, f_ref)
    else:
        oracle_mod.advection_uv(2, scheme, d.vmask, d.dxv, d.idyv, d.hv, d.v, d.tmask, d.iarea, ah_np,
                                0.5 * d.dt, h_ref, f_ref)
    assert_bitwise(host(var.all_values), f_ref, "tracer")
    assert_bitwise(host(adv.h), h_ref, "h")


@pytest.mark.parametrize("seed,kw", _domains())
@pytest.mark.parametrize("scheme", SCHEMES)
def test_w_sweep_bitwise(oracle_mod, seed, kw, scheme):
    from getm_amd.operators import Advection

    d = inputs.tiny(nx=37, ny=29, nz=11, seed=seed, ntracers=1, **kw)
    grid = make_grid(d)
    u, v, w, var = arrays(grid, d, d.tracers[0])
    adv = Advection(grid, scheme)
    adv.stats(reset=True)
    adv.h.copy_(dev(d.ho))
    adv.w_3d(w, w, d.dt, var)
    h_ref, f_ref = d.ho.copy(), d.tracers[0].copy()
    mi = oracle_mod.advection_w(scheme, d.w, d.w, d.tmask, d.dt, h_ref, f_ref)
    assert_bitwise(host(var.all_values), f_ref, "tracer")
    assert_bitwise(host(adv.h), h_ref, "h")
    assert adv.stats()[0] == mi


@pytest.mark.parametrize("case", GOLDEN_CASES)
@pytest.mark.parametrize("scheme", SCHEMES)
def test_golden_vectors_on_gpu(golden, case, scheme):
    """Reference-Fortran outputs (tests/golden) reproduced by apply_3d on the GPU."""
    from getm_amd.core import CENTERS
    from getm_amd.operators import Advection

    d = golden_domain(golden, case)
    grid = make_grid(d)
    u, v, w, var = arrays(grid, d, d.tracer)
    adv = Advection(grid, scheme)
    ah_u = grid.ugrid.wrap(dev(d.ah)) if d.ah is not None else None
    ah_v = grid.vgrid.wrap(dev(d.ah)) if d.ah is not None else None
    adv.apply_3d(u, v, w, d.dt, var, ah_u, ah_v)
    assert_bitwise(host(var.all_values), golden[f"{case}/s{scheme}/f"], "tracer")
    assert_bitwise(host(adv.h), golden[f"{case}/s{scheme}/h"], "h")
    var2 = grid.wrap(dev(d.tracer), CENTERS)
    adv.h.copy_(dev(d.ho))
    adv.w_3d(w, w, d.dt, var2)
    assert_bitwise(host(var2.all_values), golden[f"{case}/s{scheme}/f_w"], "tracer after w")


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("ufirst", [False, True])
def test_apply_3d_matches_single_sweeps(oracle_mod, scheme, ufirst):
    """The fused stage sequence equals the reference's call order of single sweeps."""
    from getm_amd.operators import Advection

    d = inputs.tiny(nx=41, ny=33, nz=9, seed=40 + scheme, ntracers=1, cfl=0.4, wscale=1.5)
    grid = make_grid(d)
    u, v, w, var = arrays(grid, d, d.tracers[0])
    adv = Advection(grid, scheme)
    adv.ufirst = ufirst
    adv.apply_3d(u, v, w, d.dt, var)
    h, f = d.ho.copy(), d.tracers[0].copy()
    half = 0.5 * d.dt
    uargs = (d.umask, d.idxu, d.dyu, d.hu, d.u, d.tmask, d.iarea, None, half)
    vargs = (d.vmask, d.dxv, d.idyv, d.hv, d.v, d.tmask, d.iarea, None, half)
    first, second = ((1, uargs), (2, vargs)) if ufirst else ((2, vargs), (1, uargs))
    oracle_mod.advection_uv(first[0], scheme, *first[1], h, f)
    oracle_mod.advection_uv(second[0], scheme, *second[1], h, f)
    oracle_mod.advection_w(scheme, d.w, d.w, d.tmask, d.dt, h, f)
    oracle_mod.advection_uv(second[0], scheme, *second[1], h, f)
    oracle_mod.advection_uv(first[0], scheme, *first[1], h, f)
    assert_bitwise(host(var.all_values), f, "tracer")
    assert_bitwise(host(adv.h), h, "h")


@pytest.mark.parametrize("cnpar", [0.5, 0.75, 1.0])
@pytest.mark.parametrize("sources", ["none", "ea2", "ea4", "both"])
@pytest.mark.parametrize("nz", [1, 2, 25])
def test_vertical_diffusion_bitwise(oracle_mod, cnpar, sources, nz):
    from getm_amd.core import CENTERS, INTERFACES
    from getm_amd.operators import VerticalDiffusion

    rng = np.random.default_rng(nz * 100 + int(cnpar * 10))
    d = inputs.tiny(nx=23, ny=19, nz=nz, seed=nz, ntracers=1)
    hn = d.hn * rng.uniform(0.8, 1.2, size=d.hn.shape)
    d.hn = hn
    grid = make_grid(d)
    nuh_np = 10.0 ** rng.uniform(-6, -1, size=(nz + 1,) + hn.shape[1:])
    nuh = grid.wrap(dev(nuh_np), INTERFACES)
    var = grid.wrap(dev(d.tracers[0]), CENTERS)
    ea2_np = rng.uniform(0, 0.05, size=hn.shape) * hn if sources in ("ea2", "both") else None
    ea4_np = rng.uniform(-1, 1, size=hn.shape) if sources in ("ea4", "both") else None
    ea2 = grid.wrap(dev(ea2_np), CENTERS) if ea2_np is not None else None
    ea4 = grid.wrap(dev(ea4_np), CENTERS) if ea4_np is not None else None
    vd = VerticalDiffusion(grid, cnpar=cnpar)
    vd(nuh, 600.0, var, molecular=1e-6, ea2=ea2, ea4=ea4)
    ref = d.tracers[0].copy()
    oracle_mod.vertical_diffusion(nuh_np, 600.0, ref, d.tmask, hn, cnpar=cnpar, molecular=1e-6, ea2=ea2_np,
                                  ea4=ea4_np)
    assert_bitwise(host(var.all_values), ref, "tracer")


def test_vertical_diffusion_use_ho(oracle_mod):
    from getm_amd.core import CENTERS, INTERFACES
    from getm_amd.operators import VerticalDiffusion

    d = inputs.tiny(nx=17, ny=13, nz=6, seed=9, ntracers=1)
    grid = make_grid(d)
    rng = np.random.default_rng(9)
    nuh_np = 10.0 ** rng.uniform(-6, -1, size=(7,) + d.hn.shape[1:])
    var = grid.wrap(dev(d.tracers[0]), CENTERS)
    vd = VerticalDiffusion(grid, cnpar=0.6)
    vd(grid.wrap(dev(nuh_np), INTERFACES), 300.0, var, use_ho=True)
    ref = d.tracers[0].copy()
    oracle_mod.vertical_diffusion(nuh_np, 300.0, ref, d.tmask, d.hn, ho=d.ho, cnpar=0.6)
    assert_bitwise(host(var.all_values), ref, "tracer")


def test_apply_before_prepare_is_an_error():
    from getm_amd import _lib
    from getm_amd.core import CENTERS
    from getm_amd.operators import VerticalDiffusion

    d = inputs.tiny(nx=8, ny=7, nz=4, seed=1, ntracers=1)
    grid = make_grid(d)
    vd = VerticalDiffusion(grid)
    with pytest.raises(_lib.GetmError):
        vd.apply(grid.wrap(dev(d.tracers[0]), CENTERS))


def test_shape_mismatch_is_an_error():
    from getm_amd.core import Grid

    d = inputs.tiny(nx=8, ny=7, nz=4, seed=1, ntracers=1)
    grid = make_grid(d)
    with pytest.raises(ValueError):
        grid.wrap(dev(np.zeros((4, 5, 5))))
    with pytest.raises(ValueError):
        Grid.from_domain(d).set_field("iarea", np.zeros((3, 3)))


def test_seamount_100_steps_bitwise(oracle_mod):
    """The 100x100x30 seamount case for 100 steps (SUPERBEE, cnpar 1, dt 60 s):
    apply_3d_batch + vertical diffusion, bit-identical to the oracle."""
    from getm_amd.core import CENTERS, INTERFACES
    from getm_amd.operators import Advection, VerticalDiffusion

    d = inputs.seamount()
    grid = make_grid(d)
    u, v, w, var = arrays(grid, d, d.tracers[0])
    nuh = grid.wrap(dev(d.nuh), INTERFACES)
    adv = Advection(grid, d.scheme)
    vd = VerticalDiffusion(grid, cnpar=d.cnpar)
    ref = d.tracers.copy()
    for _ in range(100):
        adv.apply_3d_batch(u, v, w, d.dt, [var])
        vd(nuh, d.dt, var)
        oracle_mod.transport_step(d, ref)
    assert_bitwise(host(var.all_values), ref[0], "tracer after 100 steps")
    # thickness work array ends at hn (test_3d_barotropic.py:121-128)
    wet = np.broadcast_to(d.tmask == 1, d.ho.shape)
    np.testing.assert_allclose(host(adv.h)[wet], d.hn[wet], rtol=5e-14, atol=0)


def _stage_outer_batch(adv, u, v, w, dt, vars, w_vars, Ah_u, Ah_v, ufirst):
    """operators.py:135-193 as written: stage-outer, tracer-inner, the thickness
    work array reset to the stage's starting thickness before every tracer."""
    adv1 = lambda var: adv.u_3d(u, 0.5 * dt, var, Ah_u)
    adv2 = lambda var: adv.v_3d(v, 0.5 * dt, var, Ah_v)
    if not ufirst:
        adv1, adv2 = adv2, adv1
    current_h = adv.grid.ho.all_values.clone()
    for stage in range(5):
        for var, w_var in zip(vars, w_vars):
            adv.h.copy_(current_h)
            if stage in (0, 4):
                adv1(var)
            elif stage in (1, 3):
                adv2(var)
            else:
                adv.w_3d(w, w_var, dt, var)
        current_h.copy_(adv.h)


@pytest.mark.parametrize("w_vars", ["shared", "own"])
@pytest.mark.parametrize("ah", ["both", "u-only", "none"])
@pytest.mark.parametrize("ufirst", [False, True])
def test_apply_3d_batch_equals_the_stage_outer_loop(ah, ufirst, w_vars):
    """apply_3d_batch against the reference's stage-outer loop of single sweeps,
    three tracers, on a domain whose w stage sub-steps: with one vertical
    velocity for all tracers (each tracer's five stages in one
    advection_apply_3d call) and with their own (sinking and rising tracers,
    whose sub-step counts differ, so that the last tracer's thickness carries
    into stages 4-5): the same bits and the same final thickness."""
    from getm_amd.core import CENTERS, INTERFACES
    from getm_amd.operators import Advection

    d = inputs.tiny(nx=40, ny=34, nz=9, seed=23, scheme=4, ntracers=3, cfl=0.9, wscale=2.0)
    results = []
    for batched in (True, False):
        grid = make_grid(d)
        u, v, w, _ = arrays(grid, d, d.tracers[0])
        vars = [grid.wrap(dev(d.tracers[t]), CENTERS) for t in range(3)]
        if w_vars == "shared":
            wv = [w, w, w]
        else:
            wv = [w, grid.wrap(dev(1.6 * d.w), INTERFACES), grid.wrap(dev(0.4 * d.w), INTERFACES)]
        Ah_u = grid.ugrid.wrap(dev(np.full(d.hu.shape[1:], 30.0))) if ah != "none" else None
        Ah_v = grid.vgrid.wrap(dev(np.full(d.hv.shape[1:], 20.0))) if ah == "both" else None
        adv = Advection(grid, d.scheme)
        adv.ufirst = ufirst
        adv.stats()
        for _ in range(3):
            if batched:
                by = {id(var): x for var, x in zip(vars, wv)}
                adv.apply_3d_batch(u, v, w, d.dt, vars, Ah_u=Ah_u, Ah_v=Ah_v, get_w=lambda var: by[id(var)])
            else:
                _stage_outer_batch(adv, u, v, w, d.dt, vars, wv, Ah_u, Ah_v, ufirst)
        assert adv.stats()[0] > 1  # the w stage sub-stepped
        results.append(([host(var.all_values) for var in vars], host(adv.h)))
    for t in range(3):
        assert_bitwise(results[0][0][t], results[1][0][t], f"tracer {t}")
    assert_bitwise(results[0][1], results[1][1], "thickness")
