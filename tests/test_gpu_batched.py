"""Batched T-tracer step (transport_step) against the oracle.

FP64: bit-identical to the oracle's TracerCollection.advance restatement for
every scheme, with sub-stepping, restarts, the 200 cap, mask codes and Ah.
FP32 forms: within 1e-5 of the FP64 oracle (max |error| / max |value| per
tracer); the compensated form ("f32") within 1e-8.
"""
import numpy as np
import pytest
import torch

from getm_amd import inputs
from helpers import assert_bitwise
from test_gpu_parity import dev, host, make_grid

pytestmark = pytest.mark.gpu
SCHEMES = range(1, 7)
F32_TOL = 1e-5
# compensated FP32 storage: FP32 limiter and elimination on differences, FP64 fluxes
# and updates, one rounding per launch at 2^-16 of an FP32 ulp.  On smooth fields its
# error is the FP32 limiter's and solve's (~1e-8); on the random-noise fields of
# inputs.tiny the FP32 elimination of increments as large as the values themselves
# (strong diffusion of white noise) reaches ~1e-6 in single cells.
C32_TOL = 1e-7
C32_NOISE_TOL = 3e-6
F32_FORMS = ["f32", "f32-sr"]


def tol_of(precision, noise=False):
    if precision != "f32":
        return F32_TOL
    return C32_NOISE_TOL if noise else C32_TOL


def run_batched(d, precision, nsteps=1, ah=None, ufirst=False, solver=0, tuning=None):
    from getm_amd.core import CENTERS, INTERFACES
    from getm_amd.tracer import TracerCollection

    grid = make_grid(d)
    u = grid.ugrid.wrap(dev(d.u), CENTERS)
    v = grid.vgrid.wrap(dev(d.v), CENTERS)
    w = grid.wrap(dev(d.w), INTERFACES)
    nuh = grid.wrap(dev(d.nuh), INTERFACES)
    tc = TracerCollection(grid, advection_scheme=d.scheme, cnpar=d.cnpar, precision=precision, ufirst=ufirst)
    tc.set_values(dev(d.tracers))
    if solver:
        tc.set_solver(solver)
    if tuning:
        tc.set_tuning(*tuning)
    if ah is not None:
        tc.Ah_u = grid.ugrid.wrap(dev(ah))
        tc.Ah_v = grid.vgrid.wrap(dev(ah))
    for _ in range(nsteps):
        tc.advance(d.dt, u, v, w, nuh)
    return host(tc.absolute()), host(tc.h), tc


def run_oracle(oracle_mod, d, nsteps=1, ah=None):
    tr = d.tracers.copy()
    h = None
    for _ in range(nsteps):
        h = oracle_mod.transport_step(d, tr, ah_u=ah, ah_v=ah, return_h=True)
    return tr, h


CASES = {
    "gentle": dict(nx=45, ny=38, nz=9, cfl=0.2, wscale=1.0),
    "violent": dict(nx=33, ny=27, nz=7, cfl=0.9, wscale=0.0),
    "codes": dict(nx=41, ny=35, nz=8, cfl=0.4, wscale=2.0, mask_codes=True),
}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("scheme", SCHEMES)
def test_batched_f64_bitwise(oracle_mod, case, scheme):
    d = inputs.tiny(seed=50 + scheme, scheme=scheme, ntracers=3, **CASES[case])
    ah = np.random.default_rng(scheme).uniform(0, 30, size=d.idxu.shape) if case == "codes" else None
    got, h, tc = run_batched(d, "f64", ah=ah)
    ref, href = run_oracle(oracle_mod, d, ah=ah)
    assert_bitwise(got, ref, "tracers")
    assert_bitwise(h, href, "advected thickness")
    if case == "violent":
        mi, capped = tc.stats()
        assert mi > 1


# degenerate and ragged domains: no water at all, no land at all, one row, one
# column, one cell of one layer (diffusion.F90:224-231 then divides the whole
# array, land included, so land cells turn to NaN), and a row wider than one
# 64-cell tile with a one-cell tail
EDGE_CASES = {
    "all_land": dict(nx=20, ny=17, nz=5, land_fraction=1.0),
    "all_water": dict(nx=19, ny=16, nz=6, land_fraction=-1.0, cfl=0.3),
    "one_row": dict(nx=40, ny=1, nz=3, cfl=0.5),
    "one_column": dict(nx=1, ny=37, nz=4, cfl=0.5),
    "one_cell": dict(nx=1, ny=1, nz=1, land_fraction=-1.0),
    "ragged": dict(nx=65, ny=3, nz=2, cfl=0.7, wscale=2.0),
}


@pytest.mark.parametrize("case", list(EDGE_CASES))
@pytest.mark.parametrize("scheme", SCHEMES)
def test_batched_f64_edge_domains_bitwise(oracle_mod, case, scheme):
    d = inputs.tiny(seed=60 + scheme, scheme=scheme, ntracers=3, **EDGE_CASES[case])
    got, h, _ = run_batched(d, "f64", nsteps=2)
    ref, href = run_oracle(oracle_mod, d, nsteps=2)
    assert_bitwise(got, ref, "tracers")
    assert_bitwise(h, href, "advected thickness")
    if case == "all_land":
        assert_bitwise(got, d.tracers, "tracers of a domain without water")


@pytest.mark.parametrize("case", list(EDGE_CASES))
@pytest.mark.parametrize("scheme", [1, 3, 5])
def test_per_tracer_operators_edge_domains_bitwise(oracle_mod, case, scheme):
    """The reference's per-tracer operator sequence (Advection.apply_3d_batch and
    VerticalDiffusion on the single-sweep kernels) on the same domains."""
    from getm_amd.core import CENTERS, INTERFACES
    from getm_amd.tracer import TracerCollection

    d = inputs.tiny(seed=60 + scheme, scheme=scheme, ntracers=2, **EDGE_CASES[case])
    grid = make_grid(d)
    tc = TracerCollection(grid, advection_scheme=d.scheme, cnpar=d.cnpar, precision="f64", batched=False)
    tc.set_values(dev(d.tracers))
    args = (grid.ugrid.wrap(dev(d.u), CENTERS), grid.vgrid.wrap(dev(d.v), CENTERS), grid.wrap(dev(d.w), INTERFACES),
            grid.wrap(dev(d.nuh), INTERFACES))
    for _ in range(2):
        tc.advance(d.dt, *args)
    ref, href = run_oracle(oracle_mod, d, nsteps=2)
    assert_bitwise(host(tc.values), ref, "tracers")
    assert_bitwise(host(tc.h), href, "advected thickness")


@pytest.mark.parametrize("case", list(EDGE_CASES))
@pytest.mark.parametrize("precision,solver", [("f32", 0), ("f32-sr", 0), ("f32-sr", 16)])
def test_batched_f32_edge_domains(oracle_mod, case, precision, solver):
    kw = dict(EDGE_CASES[case], cfl=0.05, wscale=0.2)
    d = inputs.tiny(seed=70, scheme=5, ntracers=3, **kw)
    got, h, _ = run_batched(d, precision, nsteps=2, solver=solver)
    ref, href = run_oracle(oracle_mod, d, nsteps=2)
    assert_bitwise(h, href, "FP64 thickness in the FP32 form")
    fin = np.isfinite(ref)
    np.testing.assert_array_equal(np.isfinite(got), fin)
    if fin.any():
        assert np.abs(got[fin] - ref[fin]).max() / np.abs(ref[fin]).max() <= tol_of(precision, noise=True)


def test_batched_f64_matches_per_tracer_operators():
    """The batched step equals the reference's per-tracer call sequence (batched=False)."""
    d = inputs.seamount(nx=48, ny=40, nz=12, ntracers=2)
    a, ha, _ = run_batched(d, "f64")
    from getm_amd.core import CENTERS, INTERFACES
    from getm_amd.tracer import TracerCollection

    grid = make_grid(d)
    tc = TracerCollection(grid, advection_scheme=d.scheme, cnpar=d.cnpar, precision="f64", batched=False)
    tc.set_values(dev(d.tracers))
    tc.advance(d.dt, grid.ugrid.wrap(dev(d.u), CENTERS), grid.vgrid.wrap(dev(d.v), CENTERS),
               grid.wrap(dev(d.w), INTERFACES), grid.wrap(dev(d.nuh), INTERFACES))
    assert_bitwise(a, host(tc.values), "tracers")
    assert_bitwise(ha, host(tc.h), "thickness")


@pytest.mark.parametrize("precision", F32_FORMS)
@pytest.mark.parametrize("scheme", SCHEMES)
def test_batched_f32_error(oracle_mod, scheme, precision):
    d = inputs.coastal(nx=96, ny=72, nz=12, scheme=scheme, ntracers=3, tracer_kind="blobs", seed=scheme)
    got, h, _ = run_batched(d, precision, nsteps=10)
    ref, href = run_oracle(oracle_mod, d, nsteps=10)
    assert_bitwise(h, href, "FP64 thickness in the FP32 form")
    wet = np.broadcast_to(d.tmask == 1, d.ho.shape)
    for t in range(d.ntracers):
        err = np.abs(got[t][wet] - ref[t][wet]).max() / np.abs(ref[t][wet]).max()
        print(f"{precision} scheme {scheme} tracer {t}: {err:.2e}")
        assert err <= tol_of(precision), (t, err)


# vertical Courant numbers of 1-2 that the iteration control resolves in 2-14
# sub-steps (with restarts) without reaching the cap of 200, with and without mask
# codes 2/3/4 and horizontal diffusion (seeds chosen so that no column is capped)
SUBSTEP_CASES = {
    "substep": (100, dict(nx=33, ny=27, nz=9, cfl=0.1, wscale=1.3)),
    "substep_codes": (110, dict(nx=33, ny=27, nz=9, cfl=0.1, wscale=1.3, mask_codes=True)),
}


@pytest.mark.parametrize("precision", F32_FORMS)
@pytest.mark.parametrize("case", list(SUBSTEP_CASES))
@pytest.mark.parametrize("scheme", SCHEMES)
def test_batched_f32_substepping(oracle_mod, case, scheme, precision):
    """The FP32 forms against the oracle where the w stage sub-steps (vertical
    Courant numbers above 1, the iteration control of template:286-322 with its
    restarts), with mask codes and horizontal diffusion in the second case; the
    thickness stays bit-identical.  Columns capped at 200 sub-steps (the
    "violent" and "codes" cases of the FP64 tests) are left out: there the
    reference's scheme is unstable and amplifies any difference in the last bit,
    which only the bit-exact FP64 form can follow."""
    seed, kw = SUBSTEP_CASES[case]
    d = inputs.tiny(seed=seed + scheme, scheme=scheme, ntracers=3, **kw)
    ah = np.random.default_rng(scheme).uniform(0, 30, size=d.idxu.shape) if kw.get("mask_codes") else None
    got, h, tc = run_batched(d, precision, ah=ah)
    ref, href = run_oracle(oracle_mod, d, ah=ah)
    assert_bitwise(h, href, "FP64 thickness in the FP32 form")
    mi, capped = tc.stats()
    assert mi > 1 and capped == 0, (mi, capped)
    fin = np.isfinite(ref)
    np.testing.assert_array_equal(np.isfinite(got), fin)
    err = np.abs(got[fin] - ref[fin]).max() / np.abs(ref[fin]).max()
    print(f"{precision} {case} scheme {scheme}: largest sub-step count {mi}, error {err:.2e}")
    assert err <= tol_of(precision, noise=True), err


@pytest.mark.parametrize("precision", F32_FORMS)
@pytest.mark.parametrize("nz", [1, 2, 16, 17, 33, 48, 64])
def test_batched_f32_depths(oracle_mod, nz, precision):
    # gentle flow: this test covers the register-column depth buckets, not ill-conditioned sub-stepping
    d = inputs.tiny(nx=30, ny=26, nz=nz, seed=nz, ntracers=2, cfl=0.05, wscale=0.2)
    d.cnpar = 0.7
    got, h, _ = run_batched(d, precision)
    ref, href = run_oracle(oracle_mod, d)
    assert_bitwise(h, href, "thickness")
    fin = np.isfinite(ref)
    np.testing.assert_array_equal(np.isfinite(got), fin)
    scale = np.abs(ref[fin]).max()
    assert np.abs(got[fin] - ref[fin]).max() / scale <= tol_of(precision, noise=True)


@pytest.mark.parametrize("precision", F32_FORMS)
def test_batched_f32_with_ah_and_ufirst(oracle_mod, precision):
    d = inputs.tiny(nx=40, ny=31, nz=10, seed=77, ntracers=2, cfl=0.1, wscale=0.3, mask_codes=True)
    ah = np.random.default_rng(1).uniform(0, 30, size=d.idxu.shape)
    got, _, _ = run_batched(d, precision, ah=ah)
    ref, _ = run_oracle(oracle_mod, d, ah=ah)
    fin = np.isfinite(ref)
    assert np.abs(got[fin] - ref[fin]).max() / np.abs(ref[fin]).max() <= tol_of(precision, noise=True)


def test_batched_f64_ufirst(oracle_mod):
    """ufirst = True runs u-v-w-v-u; compare with single sweeps in that order."""
    d = inputs.tiny(nx=29, ny=23, nz=6, seed=5, ntracers=1, cfl=0.4)
    got, h, _ = run_batched(d, "f64", ufirst=True)
    hh, f = d.ho.copy(), d.tracers[0].copy()
    half = 0.5 * d.dt
    ua = (d.umask, d.idxu, d.dyu, d.hu, d.u, d.tmask, d.iarea, None, half)
    va = (d.vmask, d.dxv, d.idyv, d.hv, d.v, d.tmask, d.iarea, None, half)
    oracle_mod.advection_uv(1, d.scheme, *ua, hh, f)
    oracle_mod.advection_uv(2, d.scheme, *va, hh, f)
    oracle_mod.advection_w(d.scheme, d.w, d.w, d.tmask, d.dt, hh, f)
    oracle_mod.advection_uv(2, d.scheme, *va, hh, f)
    oracle_mod.advection_uv(1, d.scheme, *ua, hh, f)
    oracle_mod.vertical_diffusion(d.nuh, d.dt, f, d.tmask, d.hn, cnpar=d.cnpar)
    assert_bitwise(got[0], f, "tracer")
    assert_bitwise(h, hh, "thickness")


@pytest.mark.parametrize("ntracers", [7, 10])
@pytest.mark.parametrize("precision", ["f64", "f32", "f32-sr"])
def test_batched_launch_shape_does_not_change_results(precision, ntracers):
    """Tracer batches spread over the grid (transport_set_tuning) give identical
    bits, and so do the tile shapes they select (FP32 stages 4-5: 32-wide tiles
    for up to 8 tracers per block, 64-wide beyond; tile heights 4, 8, 16)."""
    d = inputs.tiny(nx=37, ny=30, nz=9, seed=91, ntracers=ntracers, cfl=0.5, wscale=2.0)
    outs = []
    for tuning in [(0, 0, 0, 16), (1, 1, 1, 8), (2, 3, 5, 4), (4, 16, 16, 16), (0, 0, 0, 8), (10, 4, 4, 4)]:
        from getm_amd.core import CENTERS, INTERFACES
        from getm_amd.tracer import TracerCollection

        grid = make_grid(d)
        tc = TracerCollection(grid, advection_scheme=d.scheme, cnpar=d.cnpar, precision=precision)
        tc.set_tuning(*tuning)
        tc.set_values(dev(d.tracers))
        tc.advance(d.dt, grid.ugrid.wrap(dev(d.u), CENTERS), grid.vgrid.wrap(dev(d.v), CENTERS),
                   grid.wrap(dev(d.w), INTERFACES), grid.wrap(dev(d.nuh), INTERFACES))
        outs.append((host(tc.absolute()), host(tc.h)))
    for o in outs[1:]:
        assert_bitwise(o[0], outs[0][0], "tracers")
        assert_bitwise(o[1], outs[0][1], "thickness")


@pytest.mark.parametrize("precision", ["f32", "f32-sr"])
def test_many_tracer_tile_heights_do_not_change_results(precision):
    """With 16 or more tracers per block the FP32 forms run 16-row tiles with two
    cells per lane (the compensated form's y-first pair: 12-row tiles, 768-lane
    blocks); 12-row tiles for both pairs (compensated form only), 16-row tiles of
    layer groups (4 tracers per block), 8- and 4-row tiles give the same bits.  The
    other forms reject 12-row tiles."""
    from getm_amd._lib import GetmError
    from getm_amd.core import CENTERS, INTERFACES
    from getm_amd.tracer import TracerCollection

    d = inputs.tiny(nx=75, ny=53, nz=7, seed=93, ntracers=20, cfl=0.5, wscale=2.0)
    tunings = [None, (0, 16, 16, 16), (4, 16, 16, 16), (0, 16, 16, 8), (3, 16, 16, 4)]
    if precision == "f32":
        tunings.append((0, 16, 16, 12))
    outs = []
    for tuning in tunings:
        grid = make_grid(d)
        tc = TracerCollection(grid, advection_scheme=d.scheme, cnpar=d.cnpar, precision=precision)
        if tuning:
            tc.set_tuning(*tuning)
        tc.set_values(dev(d.tracers))
        for _ in range(2):
            tc.advance(d.dt, grid.ugrid.wrap(dev(d.u), CENTERS), grid.vgrid.wrap(dev(d.v), CENTERS),
                       grid.wrap(dev(d.w), INTERFACES), grid.wrap(dev(d.nuh), INTERFACES))
        res = host(tc.residuals) if tc.residuals is not None else None
        outs.append((host(tc.values), res, host(tc.h)))
    for o in outs[1:]:
        assert_bitwise(o[0], outs[0][0], "values")
        if o[1] is not None:
            assert np.array_equal(o[1], outs[0][1]), "residuals"
        assert_bitwise(o[2], outs[0][2], "thickness")
    if precision == "f32-sr":
        for other in ("f64", "f32-sr"):
            tc = TracerCollection(make_grid(d), advection_scheme=d.scheme, precision=other)
            with pytest.raises(GetmError):
                tc.set_tuning(0, 16, 16, 12)


@pytest.mark.parametrize("nx", [70, 96])
@pytest.mark.parametrize("precision", ["f64", "f32", "f32-sr"])
def test_inventory_matches_exact_sum(precision, nx):
    """Device inventories equal the exactly rounded host sum (math.fsum) to 1e-15,
    on rows of 74 and of 100 cells (a row length that is and one that is not a
    multiple of 4)."""
    from getm_amd.tracer import TracerCollection

    d = inputs.coastal(nx=nx, ny=52, nz=9, ntracers=3, tracer_kind="blobs", seed=4)
    grid = make_grid(d)
    tc = TracerCollection(grid, advection_scheme=d.scheme, precision=precision)
    tc.set_values(dev(d.tracers))
    got = tc.inventory()
    ref = d.inventory(host(tc.absolute()), h=d.hn)
    np.testing.assert_allclose(got, ref, rtol=1e-15)
    got_o = tc.inventory(dev(d.ho))
    np.testing.assert_allclose(got_o, d.inventory(host(tc.absolute()), h=d.ho), rtol=1e-15)


@pytest.mark.parametrize("precision", F32_FORMS)
def test_f32_inventory_conserved_to_1e12(precision):
    """FP32 tracers with the inventory fixer: every step conserves sum(area h c) to
    1e-12 relative (closed, continuity-consistent domain), and the fixed tracers stay
    within 1e-5 of the unfixed FP32 run.  The compensated form conserves to 1e-12
    without the fixer too (FP64 fluxes, column content corrected in the solve)."""
    from getm_amd.core import CENTERS, INTERFACES
    from getm_amd.tracer import TracerCollection

    d = inputs.coastal(nx=96, ny=80, nz=12, ntracers=4, tracer_kind="blobs", seed=8, scheme=3)
    runs = {}
    for conserve in (False, True):
        grid = make_grid(d)
        u = grid.ugrid.wrap(dev(d.u), CENTERS)
        v = grid.vgrid.wrap(dev(d.v), CENTERS)
        w = grid.wrap(dev(d.w), INTERFACES)
        nuh = grid.wrap(dev(d.nuh), INTERFACES)
        tc = TracerCollection(grid, advection_scheme=d.scheme, precision=precision, conserve=conserve)
        tc.set_values(dev(d.tracers))
        drift = []
        for _ in range(20):
            before = tc.inventory(dev(d.ho))
            tc.advance(d.dt, u, v, w, nuh)
            after = tc.inventory()
            drift.append(np.abs(after - before) / np.abs(before))
        runs[conserve] = (host(tc.absolute()), np.max(drift))
    print(f"{precision} inventory drift per step: {runs[False][1]:.2e}, with the fixer {runs[True][1]:.2e}")
    assert runs[True][1] <= 1e-12, runs[True][1]
    if precision == "f32":
        assert runs[False][1] <= 1e-12, runs[False][1]
    else:
        assert runs[False][1] > runs[True][1]
    wet = np.broadcast_to(d.tmask == 1, d.ho.shape)
    a, b = runs[True][0], runs[False][0]
    for t in range(d.ntracers):
        assert np.abs(a[t][wet] - b[t][wet]).max() / np.abs(b[t][wet]).max() <= 1e-5


@pytest.mark.parametrize("scheme", [5, 1])
def test_f32_error_does_not_grow_in_steady_flow(scheme):
    """Steady flow and diffusivity: the FP32 update factors are the same every
    step, so their rounding must not multiply the tracer itself (it would add up
    to ~2^-25 of the value per step, 1e-5 after a few hundred steps).  A point
    release after 200 steps stays within 5e-6 of the FP64 method (the batched
    FP64 step, bit-identical to the oracle in the tests above)."""
    from getm_amd.serve import initial_tracers

    d = inputs.steady_release(nx=96, ny=80, nz=20, scheme=scheme, seed=3, ntracers=2, dense=False)
    outs = {}
    for precision in ("f64", "f32", "f32-sr"):
        grid = make_grid(d)
        from getm_amd.core import CENTERS, INTERFACES
        from getm_amd.tracer import TracerCollection

        tc = TracerCollection(grid, advection_scheme=scheme, precision=precision)
        tc.set_values(initial_tracers(d, grid.device, torch.float64))
        u, v = grid.ugrid.wrap(dev(d.u), CENTERS), grid.vgrid.wrap(dev(d.v), CENTERS)
        w, nuh = grid.wrap(dev(d.w), INTERFACES), grid.wrap(dev(d.nuh), INTERFACES)
        for _ in range(200):
            tc.advance(d.dt, u, v, w, nuh)
        outs[precision] = host(tc.absolute())
    wet = np.broadcast_to(d.tmask == 1, d.ho.shape)
    for t in range(2):
        ref = outs["f64"][t][wet]
        for precision, bound in (("f32-sr", 5e-6), ("f32", C32_TOL)):
            err = np.abs(outs[precision][t][wet] - ref).max() / np.abs(ref).max()
            print(f"{precision} tracer {t} after 200 steps: {err:.2e}")
            assert err <= bound, (precision, t, err)


@pytest.mark.parametrize("cnpar", [1.0, 0.7])
@pytest.mark.parametrize("nz", [2, 5, 8, 9, 16, 17, 31, 32])
def test_batched_f32_pcr_solver(oracle_mod, nz, cnpar):
    """Wave-PCR column solve (every padding class 8/16/32 and its edges): within
    1e-5 of the FP64 oracle, like the per-lane Thomas solve it replaces."""
    d = inputs.tiny(nx=30, ny=26, nz=nz, seed=200 + nz, ntracers=3, cfl=0.05, wscale=0.2)
    d.cnpar = cnpar
    got, h, _ = run_batched(d, "f32-sr", solver=32)
    thomas, _, _ = run_batched(d, "f32-sr", solver=0)
    ref, href = run_oracle(oracle_mod, d)
    assert_bitwise(h, href, "thickness")
    fin = np.isfinite(ref)
    np.testing.assert_array_equal(np.isfinite(got), fin)
    scale = np.abs(ref[fin]).max()
    assert np.abs(got[fin] - ref[fin]).max() / scale <= F32_TOL
    assert np.abs(got[fin] - thomas[fin]).max() / scale <= F32_TOL


def test_pcr_solver_launch_shape_independent():
    """Tracer batches of the PCR solve give identical bits; a solver bound below nz
    selects the Thomas solve (identical to solver 0)."""
    d = inputs.tiny(nx=37, ny=30, nz=12, seed=91, ntracers=7, cfl=0.3, wscale=1.0)
    outs = [run_batched(d, "f32-sr", solver=16, tuning=t)[0] for t in [(0, 0, 0), (1, 16, 2), (4, 4, 3)]]
    for o in outs[1:]:
        assert_bitwise(o, outs[0], "PCR tracers")
    below, _, _ = run_batched(d, "f32-sr", solver=11)
    thomas, _, _ = run_batched(d, "f32-sr", solver=0)
    assert_bitwise(below, thomas, "Thomas when nz > pcr_max_nz")


def test_pcr_solver_bound_is_checked():
    from getm_amd import _lib
    from getm_amd.tracer import TracerCollection

    d = inputs.tiny(nx=12, ny=10, nz=4, seed=3, ntracers=1)
    tc = TracerCollection(make_grid(d), precision="f32")
    with pytest.raises(_lib.GetmError):
        tc.set_solver(33)
    tc.set_solver(32)


def _ts_run(d, precision, steps, rounding="stochastic", counter=None, conserve=False, offsets=None,
            collection=False):
    from getm_amd.core import CENTERS, INTERFACES
    from getm_amd.tracer import TracerCollection

    grid = make_grid(d)
    tc = TracerCollection(grid, advection_scheme=5, precision=precision, conserve=conserve)
    tc.set_values(dev(d.tracers).float(), offsets=offsets)  # both precisions start from the same FP32 values
    if precision == "f32-sr":
        tc.set_rounding(rounding, counter)
    u, v = grid.ugrid.wrap(dev(d.u), CENTERS), grid.vgrid.wrap(dev(d.v), CENTERS)
    w, nuh = grid.wrap(dev(d.w), INTERFACES), grid.wrap(dev(d.nuh), INTERFACES)
    for _ in range(steps):
        tc.advance(d.dt, u, v, w, nuh)
    if collection:
        return tc, (u, v, w, nuh)
    return host(tc.absolute())


def test_stochastic_rounding_keeps_error_from_accumulating():
    """Temperature and salinity (means 12.5 and 32.5) over 300 steps: rounding to
    nearest loses the sub-half-ulp part of slow increments the same way every step,
    so its error grows linearly; stochastic rounding makes it a random walk.  The
    stochastic error stays below 1e-5 and well below the nearest-rounding error.
    With the inventory fixer the error is
    that of the plain FP32 step: the fixer's one-ulp changes are spread over
    pseudo-randomly chosen cells instead of piling up in the same few cells."""
    d = inputs.coastal(nx=128, ny=96, nz=30, seed=11, ntracers=2, tracer_kind="TS")
    wet = np.broadcast_to(d.tmask == 1, d.ho.shape)
    ref = _ts_run(d, "f64", 300)
    got = {r: _ts_run(d, "f32-sr", 300, r) for r in ("nearest", "stochastic")}
    got["fixed"] = _ts_run(d, "f32-sr", 300, "stochastic", conserve=True)
    for t in range(2):
        r = ref[t][wet]
        scale = np.abs(r).max()
        e = {k: np.abs(v[t][wet] - r).max() / scale for k, v in got.items()}
        assert e["stochastic"] <= 1e-5, (t, e)
        assert e["stochastic"] <= 0.6 * e["nearest"], (t, e)
        assert e["fixed"] <= 1.25 * e["stochastic"], (t, e)


def test_stochastic_rounding_follows_the_step_counter():
    """The random bits are a function of the step count: a caller's counter is
    advanced once per step; two runs from the same count agree bit for bit; a run
    from another count differs in its last bits only."""
    d = inputs.coastal(nx=64, ny=48, nz=12, seed=12, ntracers=2, tracer_kind="TS")
    c1 = torch.zeros(1, dtype=torch.int64, device="cuda")
    c2 = torch.zeros(1, dtype=torch.int64, device="cuda")
    c3 = torch.full((1,), 1000, dtype=torch.int64, device="cuda")
    a = _ts_run(d, "f32-sr", 5, counter=c1)
    b = _ts_run(d, "f32-sr", 5, counter=c2)
    c = _ts_run(d, "f32-sr", 5, counter=c3)
    own = _ts_run(d, "f32-sr", 5)
    assert int(c1.item()) == 5 and int(c3.item()) == 1005
    assert_bitwise(a, b, "same counter start")
    assert_bitwise(a, own, "own counter starts at zero")
    assert not np.array_equal(a, c)
    assert np.abs(a - c).max() <= 1e-5 * np.abs(a).max()


def _interior_wet(d):
    wet2 = d.tmask == 1
    inner = np.zeros_like(wet2)
    h = 2
    inner[h:-h, h:-h] = True
    return np.broadcast_to(wet2 & inner, d.ho.shape)


@pytest.mark.parametrize("nz,solver,precision", [(1, 0, "f32-sr"), (12, 0, "f32-sr"), (12, 16, "f32-sr"),
                                                 (1, 0, "f32"), (12, 0, "f32")])
def test_f32_cells_not_updated_keep_their_values(nz, solver, precision):
    """The FP32 stages store only the cells they update: land and halo cells keep
    their values bit for bit (NaN included), and nothing read from them (the
    scratch buffer's copies are never read) reaches a wet cell: wet results are
    the same whether the land holds numbers or NaN.  With nz = 1 the diffusion is
    the reference's (dz c)/dz over the whole array (diffusion.F90:224-231), which
    every step applies to the cells that are not updated as well (0/0 = NaN on
    dry land)."""
    d = inputs.coastal(nx=80, ny=64, nz=nz, seed=21, ntracers=2, tracer_kind="TS")
    upd = _interior_wet(d)
    outs = {}
    for fill in ("finite", "nan"):
        dd = d.copy()
        tr = dd.tracers.astype(np.float32).astype(np.float64)
        if fill == "nan":
            tr = np.where(upd[None], tr, np.nan)
        dd.tracers = tr
        got, _, _ = run_batched(dd, precision, nsteps=4, solver=solver)
        outs[fill] = (got, tr)
    for fill, (got, tr) in outs.items():
        # cells that are not updated keep their initial (FP32) values, NaN positions included
        want = tr.astype(np.float32)
        if nz == 1:
            dz = d.hn.astype(np.float64)
            with np.errstate(invalid="ignore", divide="ignore"):
                for _ in range(4):
                    want = ((dz[None] * want.astype(np.float64)) / dz[None]).astype(np.float32)
        assert_bitwise(got[:, ~upd].astype(np.float32), want[:, ~upd], f"{fill}: not updated")
        assert np.isfinite(got[:, upd]).all(), fill
    assert_bitwise(outs["finite"][0][:, upd], outs["nan"][0][:, upd], "wet cells with finite vs NaN land")


def test_ensemble_cells_not_updated_keep_their_values():
    """As above for an ensemble of two domains: NaN land stays NaN, wet cells are finite
    and equal to the run with finite land."""
    from getm_amd.core import CENTERS, INTERFACES
    from getm_amd.ensemble import Ensemble

    doms = [inputs.coastal(nx=70, ny=50, nz=10, seed=31, ntracers=2, tracer_kind="TS"),
            inputs.coastal(nx=90, ny=40, nz=14, seed=32, ntracers=3, tracer_kind="TS")]
    res = {}
    for fill in ("finite", "nan"):
        grids, keep, tracers = [], [], []
        for d in doms:
            g = make_grid(d)
            upd = _interior_wet(d)
            tr = d.tracers.astype(np.float32)
            if fill == "nan":
                tr = np.where(upd[None], tr, np.float32(np.nan))
            grids.append(g)
            tracers.append(dev(np.ascontiguousarray(tr)))
            keep.append((d.dt, g.ugrid.wrap(dev(d.u), CENTERS), g.vgrid.wrap(dev(d.v), CENTERS),
                         g.wrap(dev(d.w), INTERFACES), g.wrap(dev(d.nuh), INTERFACES)))
        ens = Ensemble(grids, [d.ntracers for d in doms], scheme=5)
        for i, (dt, u, v, w, nuh) in enumerate(keep):
            ens.bind(i, tracers[i], u, v, w, nuh, dt)
        ens.commit()
        for _ in range(4):
            ens.step()
        res[fill] = [host(t) for t in tracers]
    for i, d in enumerate(doms):
        upd = _interior_wet(d)
        init = d.tracers.astype(np.float32)
        assert_bitwise(res["finite"][i][:, ~upd], init[:, ~upd], f"domain {i}: not updated")
        assert np.isnan(res["nan"][i][:, ~upd]).all(), i
        assert np.isfinite(res["nan"][i][:, upd]).all(), i
        assert_bitwise(res["finite"][i][:, upd], res["nan"][i][:, upd], f"domain {i}: wet cells")


@pytest.mark.parametrize("precision", ["f32", "f32-sr", "f64"])
def test_restart_continues_bit_for_bit(tmp_path, precision):
    """Snapshot after 3 steps, restore into a fresh collection, continue 4 steps:
    the result equals 7 uninterrupted steps bit for bit (the compensated form
    restores the residuals, the stochastic form its rounding step count)."""
    from getm_amd.core import CENTERS, INTERFACES
    from getm_amd.tracer import TracerCollection

    d = inputs.coastal(nx=60, ny=44, nz=9, seed=41, ntracers=3, tracer_kind="TS")
    grid = make_grid(d)
    u, v = grid.ugrid.wrap(dev(d.u), CENTERS), grid.vgrid.wrap(dev(d.v), CENTERS)
    w, nuh = grid.wrap(dev(d.w), INTERFACES), grid.wrap(dev(d.nuh), INTERFACES)

    def fresh():
        tc = TracerCollection(grid, advection_scheme=5, precision=precision)
        tc.set_values(dev(d.tracers), molecular_diffusivity=1e-6)
        return tc

    ref = fresh()
    for _ in range(7):
        ref.advance(d.dt, u, v, w, nuh)
    first = fresh()
    for _ in range(3):
        first.advance(d.dt, u, v, w, nuh)
    path = str(tmp_path / "state.npz")
    first.save(path)
    second = TracerCollection(grid, advection_scheme=5, precision=precision)
    second.load(path)
    assert [t.molecular_diffusivity for t in second] == [t.molecular_diffusivity for t in first]
    for _ in range(4):
        second.advance(d.dt, u, v, w, nuh)
    assert_bitwise(host(second.absolute()), host(ref.absolute()), "restarted run")
    if precision == "f32-sr":
        assert second.state_dict()["step"] == 7


def test_f32_anomalies_start_exact_and_track_fp64_more_closely():
    """Tracers stored as anomalies about integer offsets below their minima start
    from exactly the given values (c - o is exact in FP32 there) and, over 300
    steps, stay closer to the FP64 method than the plain FP32 form: the stored
    magnitudes, and with them the size of every stochastic rounding, are smaller
    (salinity 30-35 about 30: 8x; temperature 5-20 about 5: up to 2x)."""
    d = inputs.coastal(nx=128, ny=96, nz=30, seed=11, ntracers=2, tracer_kind="TS")
    wet = np.broadcast_to(d.tmask == 1, d.ho.shape)
    tc0, _ = _ts_run(d, "f32-sr", 0, offsets="auto", collection=True)
    assert tc0.offsets[1] >= 29.0 and tc0.offsets[0] >= 4.0, tc0.offsets
    assert (host(tc0.values)[:, wet] >= 0).all()
    start = d.tracers.astype(np.float32).astype(np.float64)
    assert_bitwise(host(tc0.absolute())[:, wet], start[:, wet], "anomalies at step 0")
    ref = _ts_run(d, "f64", 300)
    plain = _ts_run(d, "f32-sr", 300)
    anom = _ts_run(d, "f32-sr", 300, offsets="auto")
    for t in range(2):
        r = ref[t][wet]
        scale = np.abs(r).max()
        e_plain = np.abs(plain[t][wet] - r).max() / scale
        e_anom = np.abs(anom[t][wet] - r).max() / scale
        assert e_anom <= 1e-5, (t, e_anom, e_plain)
        print(f"tracer {t}: error plain {e_plain:.2e}, anomalies {e_anom:.2e}")
        if t == 1:  # salinity: 8x smaller ulp
            assert e_anom <= 0.5 * e_plain, (t, e_anom, e_plain)


def test_f32_anomalies_with_fixer_conserve_absolute_inventories():
    """With the fixer, the absolute inventories (stored anomalies plus offset
    times water volume) are kept to 1e-12 per step; a restart keeps the offsets."""
    d = inputs.coastal(nx=96, ny=80, nz=20, seed=13, ntracers=2, tracer_kind="TS")
    tc, (u, v, w, nuh) = _ts_run(d, "f32-sr", 0, conserve=True, offsets="auto", collection=True)
    ho = tc.grid.ho.all_values
    for _ in range(20):
        before = tc.inventory_device(ho, absolute=True)
        tc.advance(d.dt, u, v, w, nuh)
        after = tc.inventory_device(absolute=True)
        assert float(((after - before).abs() / before.abs()).max()) <= 1e-12
    stored = tc.inventory_device()
    assert float((stored - tc.inventory_device(absolute=True)).abs().min()) > 0  # offsets do count
    st = tc.state_dict()
    tc2, _ = _ts_run(d, "f32-sr", 0, conserve=True, collection=True)
    tc2.load_state_dict(st)
    assert list(tc2.offsets) == list(tc.offsets)
    assert_bitwise(host(tc2.absolute()), host(tc.absolute()), "restored anomalies")


def test_offsets_are_for_the_f32_form_only():
    from getm_amd.tracer import TracerCollection

    d = inputs.tiny(nx=12, ny=10, nz=4, seed=3, ntracers=1)
    tc = TracerCollection(make_grid(d), precision="f64")
    with pytest.raises(ValueError):
        tc.set_values(dev(d.tracers), offsets="auto")
    tc32 = TracerCollection(make_grid(d), precision="f32")
    with pytest.raises(ValueError):
        tc32.set_values(dev(d.tracers), offsets=[1.0, 2.0])


@pytest.mark.parametrize("precision", F32_FORMS)
@pytest.mark.parametrize("tpb", [0, 1, 16])
def test_f32_step_without_thickness_output_equals_with_it(tpb, precision):
    """With h_out NULL the FP32 step keeps the final thickness in its own buffer,
    where the cells that no stage updates already hold theirs (they are neither
    copied by the w stage nor stored again); the tracers equal those of a step
    that writes the thickness out, bit for bit, over several steps."""
    from getm_amd import _lib
    from getm_amd.core import CENTERS, INTERFACES
    from getm_amd.tracer import TracerCollection

    d = inputs.tiny(nx=45, ny=38, nz=12, seed=71, ntracers=5, cfl=0.6, wscale=2.0)
    outs = []
    for with_h in (True, False):
        grid = make_grid(d)
        tc = TracerCollection(grid, advection_scheme=d.scheme, cnpar=d.cnpar, precision=precision)
        tc.set_tuning(tpb, tpb, tpb, 0)
        tc.set_values(dev(d.tracers))
        u, v = grid.ugrid.wrap(dev(d.u), CENTERS), grid.vgrid.wrap(dev(d.v), CENTERS)
        w, nuh = grid.wrap(dev(d.w), INTERFACES), grid.wrap(dev(d.nuh), INTERFACES)
        for _ in range(3):
            rc = _lib.load().transport_step(tc.p, len(tc), _lib.ptr(tc._buf), _lib.ptr(u.all_values),
                                            _lib.ptr(v.all_values), _lib.ptr(w.all_values),
                                            _lib.ptr(nuh.all_values), None, None, float(d.dt), 0.0,
                                            _lib.ptr(tc.h) if with_h else None, _lib.stream_handle())
            _lib.check(rc, "transport_step")
        outs.append(host(tc.absolute()))
    assert_bitwise(outs[1], outs[0], "tracers without h_out")


@pytest.mark.parametrize("precision", ["f32", "f32-sr"])
def test_fixer_of_a_mean_zero_tracer_moves_cells_by_the_absolute_mass_ratio(precision):
    """transport_conserve on a tracer of mixed signs whose inventory is ~0 next to
    a positive one: the defect D is shared in proportion to each cell's absolute
    mass (ratio D / I_abs, I_abs = sum area h |c|), so both reach their targets
    and no cell of the mean-zero tracer moves by more than |D| / I_abs of its
    value plus a storage unit.  Weights of signed mass (ratio D / I with I ~ 0)
    would scale its cells by orders of magnitude."""
    from getm_amd import _lib
    from getm_amd.core import Grid
    from getm_amd.tracer import TracerCollection

    d = inputs.coastal(nx=120, ny=90, nz=12, seed=5, ntracers=2, tracer_kind="TS")
    grid = Grid.from_domain(d)
    shape = (grid.nz_, grid.ny_, grid.nx_)
    gen = torch.Generator(device=grid.device).manual_seed(11)
    pos = 1.0 + torch.rand(shape, generator=gen, dtype=torch.float64, device=grid.device)
    anom = 2.0 * torch.rand(shape, generator=gen, dtype=torch.float64, device=grid.device) - 1.0

    def collection(vals):
        tc = TracerCollection(grid, advection_scheme=3, precision=precision)
        tc.set_values(torch.stack(vals))
        return tc

    inv = collection([pos, anom, torch.ones_like(pos)]).inventory_device()
    anom = anom - float(inv[1] / inv[2])  # remove the volume mean: inventory ~0
    tc = collection([pos, anom])
    before = tc.absolute().clone()
    absmass = collection([before[0].abs(), before[1].abs()]).inventory_device()
    inv = tc.inventory_device()
    assert float(inv[1].abs() / absmass[1]) < 1e-6, "the anomaly's inventory should be ~0"
    dfix = 1e-9 * absmass
    target = inv + dfix
    rc = _lib.load().transport_conserve(tc.p, len(tc), _lib.ptr(tc._buf), _lib.ptr(target), _lib.stream_handle())
    _lib.check(rc, "transport_conserve")
    after_inv = tc.inventory_device()
    err = ((after_inv - target).abs() / absmass).cpu().numpy()
    assert (err <= 1e-12).all(), err
    after = tc.absolute()
    wet = before[1] != after[1]
    assert int(wet.sum()) > 0, "the fixer changed nothing"
    rel = ((after[1] - before[1]).abs() / before[1].abs().clamp_min(1e-30))
    if precision == "f32":
        # x + r |x|, r = D / I_abs = 1e-9, encoded again (residual unit 2^-39 |x|)
        assert rel.max().item() <= 1e-9 + 2.0 ** -38, rel.max().item()
    else:
        # dithering moves cells by at most one ulp (2^-23 relative), the greedy
        # pass that takes the remainder a few cells by at most 4096 ulps
        assert torch.quantile(rel.flatten()[::7].float(), 0.999).item() <= 2.0 ** -22
        assert rel.max().item() <= 4096 * 2.0 ** -23, rel.max().item()
