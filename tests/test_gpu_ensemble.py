"""Ensembles of uneven domains (one launch per kernel kind for all domains).

Each domain's result equals the single-domain FP32 batched step bit for bit,
alone or in the mix, for any launch shape, and hipGraph-captured chunks
replayed several times (with eager work in between) equal eager stepping."""
import numpy as np
import pytest

from getm_amd import inputs
from helpers import assert_bitwise
from test_gpu_parity import dev, host, make_grid

pytestmark = pytest.mark.gpu


def small_ensemble(scheme, seed=5):
    return inputs.ensemble_domains(ndomains=5, seed=seed, size_range=(24, 70), nz_range=(3, 40), tracer_range=(1, 5),
                                   scheme=scheme)


def make_ensemble(doms, tuning=None, tile_x=None):
    from getm_amd.core import CENTERS, INTERFACES
    from getm_amd.ensemble import Ensemble

    grids = [make_grid(d) for d in doms]
    ens = Ensemble(grids, [d.ntracers for d in doms], scheme=doms[0].scheme)
    for i, (d, g) in enumerate(zip(doms, grids)):
        ens.bind(i, dev(d.tracers).float().contiguous(), g.ugrid.wrap(dev(d.u), CENTERS),
                 g.vgrid.wrap(dev(d.v), CENTERS), g.wrap(dev(d.w), INTERFACES), g.wrap(dev(d.nuh), INTERFACES), d.dt)
    if tuning:
        ens.set_tuning(*tuning)
    if tile_x:
        ens.set_tile_x(tile_x)
    ens.commit()
    return ens


def single_domain(d, nsteps, solver=0):
    from getm_amd.core import CENTERS, INTERFACES
    from getm_amd.tracer import TracerCollection

    g = make_grid(d)
    tc = TracerCollection(g, advection_scheme=d.scheme, precision="f32-sr")
    tc.set_values(dev(d.tracers))
    if solver:
        tc.set_solver(solver)
    args = (g.ugrid.wrap(dev(d.u), CENTERS), g.vgrid.wrap(dev(d.v), CENTERS), g.wrap(dev(d.w), INTERFACES),
            g.wrap(dev(d.nuh), INTERFACES))
    for _ in range(nsteps):
        tc.advance(d.dt, *args)
    return host(tc.values)


@pytest.mark.parametrize("scheme", [5, 1, 3])
def test_ensemble_equals_single_domains(scheme):
    doms = small_ensemble(scheme)
    ens = make_ensemble(doms)
    for _ in range(3):
        ens.step()
    for i, d in enumerate(doms):
        assert_bitwise(host(ens.tracers[i]), single_domain(d, 3), f"domain {i}")


def test_ensemble_with_degenerate_domains_equals_single_domains():
    """Domains without water, without land, of one row, one column or one cell of
    one layer, mixed with an ordinary one in a single ensemble."""
    shapes = [dict(nx=20, ny=17, nz=5, land_fraction=1.0), dict(nx=19, ny=16, nz=6, land_fraction=-1.0),
              dict(nx=40, ny=1, nz=3), dict(nx=1, ny=37, nz=4), dict(nx=1, ny=1, nz=1, land_fraction=-1.0),
              dict(nx=65, ny=3, nz=2), dict(nx=48, ny=40, nz=12)]
    doms = [inputs.tiny(seed=80 + i, scheme=5, ntracers=1 + i % 3, cfl=0.3, **kw) for i, kw in enumerate(shapes)]
    ens = make_ensemble(doms)
    for _ in range(2):
        ens.step()
    for i, d in enumerate(doms):
        assert_bitwise(host(ens.tracers[i]), single_domain(d, 2), f"domain {i}")


def test_ensemble_launch_shape_independent():
    doms = small_ensemble(5, seed=9)
    outs = []
    for tuning in [(16, 0, 16, 16), (4, 1, 2, 1), (8, 2, 64, 3)]:
        ens = make_ensemble(doms, tuning)
        ens.step()
        ens.step()
        outs.append([host(t) for t in ens.tracers])
    for o in outs[1:]:
        for i in range(len(doms)):
            assert_bitwise(o[i], outs[0][i], f"domain {i}")


def test_ensemble_tile_widths_do_not_change_results():
    """Every tile shape of the horizontal stages (32 x {8, 16}, 64 x {4, 8, 16},
    128 x {4, 8}) gives the single-domain step's bits, also on domains narrower
    than a tile and with a ragged last tile."""
    from getm_amd.autotune import ENSEMBLE_TILES

    doms = small_ensemble(3, seed=17) + [inputs.tiny(seed=90, scheme=3, ntracers=2, nx=130, ny=9, nz=4)]
    ref = [single_domain(d, 2) for d in doms]
    for tx, ty in ENSEMBLE_TILES:
        ens = make_ensemble(doms, (ty, 0, 16, 16), tile_x=tx)
        ens.step()
        ens.step()
        for i in range(len(doms)):
            assert_bitwise(host(ens.tracers[i]), ref[i], f"tile {tx}x{ty}, domain {i}")


def test_ensemble_rejects_unsupported_tile_shapes():
    from getm_amd._lib import GetmError

    doms = small_ensemble(5, seed=19)
    ens = make_ensemble(doms)
    with pytest.raises(GetmError):
        ens.set_tile_x(48)
    ens.set_tile_x(128)
    ens.set_tuning(16, 0, 16, 16)  # 128 x 16 would be 2048 lanes
    with pytest.raises(GetmError):
        ens.commit()


def test_ensemble_graph_chunks_replay_equal_eager():
    import torch

    doms = small_ensemble(5, seed=11)
    eager = make_ensemble(doms)
    graphed = make_ensemble(doms)
    g = graphed.graph_chunk(2)  # its warm-up step is undone: the state is the initial one
    for i in range(len(doms)):
        assert_bitwise(host(graphed.tracers[i]), host(eager.tracers[i]), f"after capture, domain {i}")
    for rep in range(3):
        g.replay()
        torch.cuda.synchronize()
        _ = torch.ones(1000, device="cuda").sum().item()  # eager work between replays
        eager.step()
        eager.step()
        for i in range(len(doms)):
            assert_bitwise(host(graphed.tracers[i]), host(eager.tracers[i]), f"replay {rep} domain {i}")


def test_autotuner_keeps_results(tmp_path):
    """Tuning leaves the tracers and rounding step counters as they were, whether it
    times candidates (PCR and Thomas solves round differently) or applies a cached
    shape: the tuned ensemble then steps exactly as an untuned one."""
    from getm_amd.autotune import Autotuner

    doms = small_ensemble(5, seed=13)
    cands = [(16, 0, 16, 16, 0, 32), (8, 2, 4, 4, 32, 128), (4, 0, 16, 16, 0, 64)]
    cache = str(tmp_path / "tune.json")
    for run in ("timed", "cached"):
        ens = make_ensemble(doms)
        ref = make_ensemble(doms)
        tuner = Autotuner(cache)
        shape = tuner.tune_ensemble(ens, candidates=cands, reps=1)
        assert shape in cands
        for i in range(len(doms)):
            assert_bitwise(host(ens.tracers[i]), host(ref.tracers[i]), f"{run}: after tuning, domain {i}")
        ref.set_tuning(*shape[:4])
        ref.set_solver(shape[4])
        ref.set_tile_x(shape[5])
        ref.commit()
        for _ in range(2):
            ens.step()
            ref.step()
        for i in range(len(doms)):
            assert_bitwise(host(ens.tracers[i]), host(ref.tracers[i]), f"{run}: two steps later, domain {i}")


def test_ensemble_pcr_by_depth_equals_single_domains():
    """Domains of up to 24 layers take the wave-PCR solve, deeper ones Thomas, in one
    ensemble step; each equals its single-domain step with the same solver choice."""
    doms = small_ensemble(5, seed=11)
    assert min(d.nz for d in doms) <= 24 < max(d.nz for d in doms)
    ens = make_ensemble(doms)
    ens.set_solver(24)
    ens.commit()
    for _ in range(3):
        ens.step()
    for i, d in enumerate(doms):
        assert_bitwise(host(ens.tracers[i]), single_domain(d, 3, solver=24), f"domain {i} (nz {d.nz})")


@pytest.mark.parametrize("scheme", [5, 1, 3])
def test_ensemble_domains_follow_the_oracle(oracle_mod, scheme):
    """Every domain of an uneven ensemble against the FP64 oracle (its restatement of
    TracerCollection.advance): within 1e-5 of each tracer's largest value after 3 steps."""
    doms = small_ensemble(scheme, seed=21)
    ens = make_ensemble(doms)
    for _ in range(3):
        ens.step()
    for i, d in enumerate(doms):
        ref = d.tracers.astype(np.float32).astype(np.float64)  # the ensemble starts from FP32 values
        for _ in range(3):
            oracle_mod.transport_step(d, ref)
        got = host(ens.tracers[i]).astype(np.float64)
        wet = np.broadcast_to(d.tmask == 1, d.ho.shape)
        for t in range(d.ntracers):
            err = np.abs(got[t][wet] - ref[t][wet]).max() / np.abs(ref[t][wet]).max()
            assert err <= 1e-5, (i, t, err)


def test_96_domain_ensemble_graph_replay_equals_eager():
    """The full uneven ensemble: 96 domains with nx, ny log-uniform in [64, 768], nz
    uniform in [10, 60] and 1-32 blob tracers each.  A 100-step hipGraph replay equals
    100 eager steps bit for bit, and sampled domains equal their single-domain steps."""
    import torch

    from getm_amd.core import CENTERS, INTERFACES, Grid
    from getm_amd.ensemble import Ensemble
    from getm_amd.precision import _blob_tracers

    specs = inputs.ensemble_specs(96, seed=3)
    grids, keep, fields = [], [], []
    for nx, ny, nz, nt, seed in specs:
        d = inputs.coastal(nx=nx, ny=ny, nz=nz, dx=2000.0, dt=30.0, scheme=5, seed=seed, ntracers=0)
        g = Grid.from_domain(d)
        f = (g.ugrid.wrap(dev(d.u), CENTERS), g.vgrid.wrap(dev(d.v), CENTERS), g.wrap(dev(d.w), INTERFACES),
             g.wrap(dev(d.nuh), INTERFACES))
        grids.append(g)
        fields.append((d.dt, f))
        keep.append(_blob_tracers(nt, g, seed, torch.float32))
    print(f"96 domains: {sum(s[3] for s in specs)} tracers, nz {min(s[2] for s in specs)}-{max(s[2] for s in specs)}",
          flush=True)
    ens = Ensemble(grids, [s[3] for s in specs], scheme=5)
    for i, (dt, (u, v, w, nuh)) in enumerate(fields):
        ens.bind(i, keep[i], u, v, w, nuh, dt)
    ens.commit()
    start = ens.snapshot()
    for _ in range(100):
        ens.step()
    eager = [host(t) for t in ens.tracers]
    ens.restore(start)
    graph = ens.graph_chunk(100)
    graph.replay()
    torch.cuda.synchronize()
    for i in range(96):
        assert_bitwise(host(ens.tracers[i]), eager[i], f"domain {i}")
    print("100-step replay equals eager stepping", flush=True)
    # sampled domains (the smallest, the largest, the deepest) against their single-domain step
    ncells = [s[0] * s[1] * s[2] for s in specs]
    sample = {int(np.argmin(ncells)), int(np.argmax(ncells)), int(np.argmax([s[2] for s in specs]))}
    from getm_amd.tracer import TracerCollection

    ens.restore(start)
    for _ in range(3):
        ens.step()
    for i in sorted(sample):
        tc = TracerCollection(grids[i], advection_scheme=5, precision="f32-sr")
        tc.set_values(start[0][i])
        dt, args = fields[i]
        for _ in range(3):
            tc.advance(dt, *args)
        assert_bitwise(host(ens.tracers[i]), host(tc.values), f"domain {i} {specs[i][:4]}")
        mi, capped = tc.stats()
        assert capped == 0, (i, mi, capped)
