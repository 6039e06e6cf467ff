"""TracerCollection with tracer sources, surface fluxes, vertical velocities and
per-tracer molecular diffusivities (tracer.py:22-276) against the oracle: bit for
bit in FP64 (the per-tracer operator sequence, and the batched step for
vertical velocities and molecular diffusivities), within tolerance in FP32."""
import numpy as np
import pytest

from getm_amd import inputs
from helpers import assert_bitwise
from test_gpu_parity import dev, host, make_grid

pytestmark = pytest.mark.gpu


def oracle_advance(oracle_mod, d, tracers, mols, extras):
    """TracerCollection.advance restated with oracle operators (operators.py:135-193,
    ufirst = False: v, u, w, u, v; tracer.py:223-231 combined w; tracer.py:248-276)."""
    half = 0.5 * d.dt
    ua = (d.umask, d.idxu, d.dyu, d.hu, d.u, d.tmask, d.iarea, None, half)
    va = (d.vmask, d.dxv, d.idyv, d.hv, d.v, d.tmask, d.iarea, None, half)
    current = d.ho.copy()
    for stage in range(5):
        hh = None
        for f, (src, flux, scale, vv) in zip(tracers, extras):
            hh = current.copy()
            if stage in (0, 4):
                oracle_mod.advection_uv(2, d.scheme, *va, hh, f)
            elif stage in (1, 3):
                oracle_mod.advection_uv(1, d.scheme, *ua, hh, f)
            else:
                w_var = d.w
                if vv is not None:
                    w_var = np.zeros_like(d.w)
                    w_var[1:-1] = 0.5 * (vv[:-1] + vv[1:])
                    w_var += d.w
                oracle_mod.advection_w(d.scheme, d.w, w_var, d.tmask, d.dt, hh, f)
        current = hh
    for f, mol, (src, flux, scale, vv) in zip(tracers, mols, extras):
        ea4 = None
        if src is not None or flux is not None:
            ea4 = src.copy() if src is not None else np.zeros_like(f)
            if flux is not None:
                ea4[-1] += flux
            ea4 *= d.dt * scale
        oracle_mod.vertical_diffusion(d.nuh, d.dt, f, d.tmask, d.hn, cnpar=d.cnpar, molecular=mol, ea4=ea4)
    return current


@pytest.mark.parametrize("batched", [False, True])
@pytest.mark.parametrize("scheme", [5, 1])
def test_sources_fluxes_and_sinking_bitwise(oracle_mod, scheme, batched):
    """Per-tracer sources, surface fluxes and vertical velocities against the
    oracle; an FP64 collection created batched runs the per-tracer sequence for
    them, with the same bits."""
    from getm_amd.core import CENTERS, INTERFACES
    from getm_amd.tracer import TracerCollection

    d = inputs.tiny(nx=31, ny=27, nz=8, seed=17 + scheme, scheme=scheme, ntracers=3, cfl=0.3, wscale=0.5)
    rng = np.random.default_rng(scheme)
    src = rng.standard_normal(d.ho.shape) * 1e-4
    flux = rng.standard_normal(d.ho.shape[1:]) * 1e-3
    vv = -np.abs(rng.standard_normal(d.ho.shape)) * 2e-4
    extras = [(src, flux, 0.5, None), (None, None, 1.0, vv), (None, None, 1.0, None)]
    mols = [0.0, 1e-6, 1e-6]

    grid = make_grid(d)
    tc = TracerCollection(grid, advection_scheme=scheme, cnpar=d.cnpar, precision="f64", batched=batched)
    tc.add("heat", data=dev(d.tracers[0]), molecular_diffusivity=mols[0], source=grid.wrap(dev(src), CENTERS),
           surface_flux=grid.wrap(dev(flux)), source_scale=0.5)
    tc.add("sinking", data=dev(d.tracers[1]), molecular_diffusivity=mols[1],
           vertical_velocity=grid.wrap(dev(vv), CENTERS))
    tc.add("dye", data=dev(d.tracers[2]), molecular_diffusivity=mols[2])
    tc.advance(d.dt, grid.ugrid.wrap(dev(d.u), CENTERS), grid.vgrid.wrap(dev(d.v), CENTERS),
               grid.wrap(dev(d.w), INTERFACES), grid.wrap(dev(d.nuh), INTERFACES))

    ref = [np.ascontiguousarray(t.copy()) for t in d.tracers]
    h = oracle_advance(oracle_mod, d, ref, mols, extras)
    for i in range(3):
        assert_bitwise(host(tc.values[i]), ref[i], f"tracer {i}")
    assert_bitwise(host(tc.h), h, "thickness")


def _sinking_and_ts(scheme, nz=8):
    """Three tracers: temperature- and salinity-like molecular diffusivities
    (1.4e-7 and 1.1e-9, the reference's defaults, simulation.py:847, 857), a
    sinking tracer with a third, and a floating one sharing it."""
    # a flow whose w stage sub-steps 2-3 times without reaching the cap of 200 (capped
    # columns make the reference's scheme unstable, which only FP64 can follow)
    d = inputs.tiny(nx=31, ny=27, nz=nz, seed=27 + scheme, scheme=scheme, ntracers=4, cfl=0.08, wscale=0.5)
    rng = np.random.default_rng(7 + scheme)
    vv_sink = -np.abs(rng.standard_normal(d.ho.shape)) * 2e-4
    vv_float = np.abs(rng.standard_normal(d.ho.shape)) * 1e-4
    extras = [(None, None, 1.0, None), (None, None, 1.0, None), (None, None, 1.0, vv_sink),
              (None, None, 1.0, vv_float)]
    mols = [1.4e-7, 1.1e-9, 1e-6, 1e-6]
    return d, extras, mols


def _collection(d, precision, extras, mols):
    from getm_amd.core import CENTERS
    from getm_amd.tracer import TracerCollection

    grid = make_grid(d)
    tc = TracerCollection(grid, advection_scheme=d.scheme, cnpar=d.cnpar, precision=precision)
    for i, ((src, flux, scale, vv), mol) in enumerate(zip(extras, mols)):
        tc.add(f"tracer{i}", data=dev(d.tracers[i]), molecular_diffusivity=mol, source_scale=scale,
               source=grid.wrap(dev(src), CENTERS) if src is not None else None,
               surface_flux=grid.wrap(dev(flux)) if flux is not None else None,
               vertical_velocity=grid.wrap(dev(vv), CENTERS) if vv is not None else None)
    return tc, grid


def _fields(d, grid):
    from getm_amd.core import CENTERS, INTERFACES

    return (grid.ugrid.wrap(dev(d.u), CENTERS), grid.vgrid.wrap(dev(d.v), CENTERS), grid.wrap(dev(d.w), INTERFACES),
            grid.wrap(dev(d.nuh), INTERFACES))


@pytest.mark.parametrize("scheme", [5, 3, 1])
def test_batched_f64_vertical_velocities_and_molecular_bitwise(oracle_mod, scheme):
    """The batched FP64 step takes per-tracer vertical velocities (its w stage
    runs once per tracer group with the group's w_var, the thickness following
    the last group as the reference's apply_3d_batch leaves it) and per-tracer
    molecular diffusivities (one diffusion per group, tracer.py:273-275), bit for
    bit against the oracle's per-tracer sequence, over two steps."""
    d, extras, mols = _sinking_and_ts(scheme)
    tc, grid = _collection(d, "f64", extras, mols)
    assert tc._advection is None  # the batched step, not the per-tracer sequence
    args = _fields(d, grid)
    ref = [np.ascontiguousarray(t.copy()) for t in d.tracers]
    for _ in range(2):
        tc.advance(d.dt, *args)
        h = oracle_advance(oracle_mod, d, ref, mols, extras)
    assert tc._advection is None
    for i in range(len(ref)):
        assert_bitwise(host(tc.values[i]), ref[i], f"tracer {i}")
    assert_bitwise(host(tc.h), h, "thickness")


@pytest.mark.parametrize("precision", ["f32", "f32-sr"])
@pytest.mark.parametrize("scheme", [5, 3])
def test_f32_vertical_velocities_and_molecular_follow_fp64(oracle_mod, scheme, precision):
    """The FP32 forms with the same tracers: within 1e-5 (stochastic rounding) or
    3e-6 (compensated storage, on these random-noise fields) of the oracle,
    thickness bit for bit."""
    d, extras, mols = _sinking_and_ts(scheme)
    tc, grid = _collection(d, precision, extras, mols)
    args = _fields(d, grid)
    ref = [np.ascontiguousarray(t.copy()) for t in d.tracers]
    for _ in range(3):
        tc.advance(d.dt, *args)
        h = oracle_advance(oracle_mod, d, ref, mols, extras)
    assert_bitwise(host(tc.h), h, "thickness")
    got = host(tc.absolute())
    tol = 3e-6 if precision == "f32" else 1e-5  # random-noise fields (inputs.tiny)
    for i in range(len(ref)):
        fin = np.isfinite(ref[i])
        err = np.abs(got[i][fin] - ref[i][fin]).max() / np.abs(ref[i][fin]).max()
        print(f"{precision} scheme {scheme} tracer {i}: {err:.2e}")
        assert err <= tol, (i, err)


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_restart_keeps_sources_and_vertical_velocities(tmp_path, precision):
    """A collection restored from a snapshot keeps its tracers' sources, surface
    fluxes and vertical velocities: 2 steps, save, load into the same collection,
    3 more steps equal 5 uninterrupted steps bit for bit."""
    d = inputs.tiny(nx=25, ny=21, nz=6, seed=5, ntracers=3, cfl=0.08, wscale=0.5)
    rng = np.random.default_rng(3)
    src = rng.standard_normal(d.ho.shape) * 1e-5
    flux = rng.standard_normal(d.ho.shape[1:]) * 1e-4
    vv = -np.abs(rng.standard_normal(d.ho.shape)) * 1e-4
    extras = [(src, flux, 0.5, None), (None, None, 1.0, vv), (None, None, 1.0, None)]
    d.cnpar = 1.0
    mols = [1e-6, 1e-6, 2e-6]
    outs = []
    for split in (False, True):
        tc, grid = _collection(d, precision, extras, mols)
        args = _fields(d, grid)
        for _ in range(2):
            tc.advance(d.dt, *args)
        if split:
            path = str(tmp_path / "state.npz")
            tc.save(path)
            tc.advance(d.dt, *args)  # a step the restore must undo
            tc.load(path)
            assert tc[0].source is not None and tc[1].vertical_velocity is not None
        for _ in range(3):
            tc.advance(d.dt, *args)
        outs.append(host(tc.absolute()))
    assert_bitwise(outs[1], outs[0], "restored run")


@pytest.mark.parametrize("cnpar", [1.0, 0.5])
@pytest.mark.parametrize("nz", [1, 12, 30])
def test_f32_sources_and_fluxes_follow_fp64(oracle_mod, nz, cnpar):
    """The FP32 forms' sources and surface fluxes (the implicit part of the
    column solve acting on c + ea4/dz, the explicit part of cnpar < 1 on c) against
    the FP64 per-tracer sequence on the smooth coastal fields (T- and S-like
    values): within 1e-5 of the tracer's range over 5 steps, for a tracer with
    sources and a surface flux, one with a flux only and one with neither."""
    from getm_amd.core import CENTERS, INTERFACES
    from getm_amd.tracer import TracerCollection

    d = inputs.coastal(nx=96, ny=80, nz=nz, seed=13, ntracers=3, tracer_kind="TS")
    rng = np.random.default_rng(nz)
    # with cnpar < 1 sources large enough that acting on c + ea4/dz in the explicit
    # part instead of on c would move the result by more than the tolerance
    src = rng.standard_normal(d.ho.shape) * (1e-5 if cnpar == 1.0 else 1e-3)
    flux = rng.standard_normal(d.ho.shape[1:]) * 1e-4
    wet = np.broadcast_to(d.tmask == 1, d.ho.shape)
    out = {}
    for precision in ("f64", "f32", "f32-sr"):
        grid = make_grid(d)
        tc = TracerCollection(grid, advection_scheme=5, cnpar=cnpar, precision=precision)
        tc.add("heat", data=dev(d.tracers[0]), molecular_diffusivity=1e-6, source=grid.wrap(dev(src), CENTERS),
               surface_flux=grid.wrap(dev(flux)), source_scale=0.5)
        tc.add("flux", data=dev(d.tracers[1]), molecular_diffusivity=1e-6, surface_flux=grid.wrap(dev(flux)))
        tc.add("plain", data=dev(d.tracers[2]), molecular_diffusivity=1e-6)
        args = (grid.ugrid.wrap(dev(d.u), CENTERS), grid.vgrid.wrap(dev(d.v), CENTERS),
                grid.wrap(dev(d.w), INTERFACES), grid.wrap(dev(d.nuh), INTERFACES))
        for _ in range(5):
            tc.advance(d.dt, *args)
        out[precision] = host(tc.absolute())
    for t in range(3):
        r = out["f64"][t][wet]
        for precision, tol in (("f32-sr", 1e-5), ("f32", 1e-6)):
            err = np.abs(out[precision][t][wet] - r).max() / np.abs(r).max()
            print(f"{precision} nz {nz} cnpar {cnpar} tracer {t}: {err:.2e}")
            assert err <= tol, (precision, t, err)
    # the sources did change the first two tracers
    plain = single_f64_without_sources(d, cnpar)
    assert np.abs(out["f64"][0][wet] - plain[0][wet]).max() > 1e-6 * np.abs(plain[0][wet]).max()


def single_f64_without_sources(d, cnpar=1.0):
    from getm_amd.core import CENTERS, INTERFACES
    from getm_amd.tracer import TracerCollection

    grid = make_grid(d)
    tc = TracerCollection(grid, advection_scheme=5, cnpar=cnpar, precision="f64")
    tc.set_values(dev(d.tracers), molecular_diffusivity=1e-6)
    args = (grid.ugrid.wrap(dev(d.u), CENTERS), grid.vgrid.wrap(dev(d.v), CENTERS),
            grid.wrap(dev(d.w), INTERFACES), grid.wrap(dev(d.nuh), INTERFACES))
    for _ in range(5):
        tc.advance(d.dt, *args)
    return host(tc.values)


def test_f32_sources_with_the_fixer_are_rejected():
    """The fixer restores the inventory the tracers had before the step, which
    sources change: a collection with both raises."""
    from getm_amd.core import CENTERS
    from getm_amd.tracer import TracerCollection

    d = inputs.tiny(nx=12, ny=10, nz=4, seed=3, ntracers=1)
    grid = make_grid(d)
    for kw in (dict(cnpar=0.5, conserve=True), dict(cnpar=1.0, conserve=True)):
        for precision in ("f32", "f32-sr"):
            tc = TracerCollection(grid, precision=precision, **kw)
            tc.add("s", data=dev(d.tracers[0]), source=grid.wrap(dev(np.zeros_like(d.ho)), CENTERS))
            with pytest.raises(ValueError):
                tc.advance(d.dt, None, None, None, None)
