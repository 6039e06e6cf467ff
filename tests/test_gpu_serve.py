"""Serving loop: continuous batching gives every request the result it gets alone
(bit for bit), results match the FP64 oracle to 1e-5 and keep their inventory to
1e-12, and admission respects the slot count and the HBM budget."""
import dataclasses

import numpy as np
import pytest
import torch

from getm_amd import inputs
from getm_amd.serve import Request, Server, request_bytes

pytestmark = pytest.mark.gpu

# small scenarios of both limiters, arriving while others run (arrivals in loop ticks)
REQUESTS = [
    Request(rid=0, nx=40, ny=32, nz=10, ntracers=3, nsteps=12, scheme=5, seed=2, arrival=0),
    Request(rid=1, nx=29, ny=45, nz=7, ntracers=1, nsteps=9, scheme=1, seed=3, arrival=0),
    Request(rid=2, nx=52, ny=21, nz=13, ntracers=4, nsteps=10, scheme=5, seed=4, arrival=3),
    Request(rid=3, nx=33, ny=33, nz=5, ntracers=2, nsteps=6, scheme=1, seed=5, arrival=5),
    Request(rid=4, nx=24, ny=40, nz=16, ntracers=5, nsteps=8, scheme=5, seed=6, arrival=11),
]


@pytest.fixture(scope="module")
def domains():
    return {r.rid: r.domain() for r in REQUESTS}


def run_alone(domains, **kw):
    server = Server(keep_results=True, **kw)
    out = {}
    for r in REQUESTS:
        (res,) = server.run([dataclasses.replace(r, arrival=0)], domains, clock="ticks")
        assert not res.error
        out[r.rid] = res
    return out


@pytest.fixture(scope="module")
def alone(domains):
    return run_alone(domains, policy="lockstep")


def run_mix(domains, **kw):
    kw.setdefault("policy", "lockstep")
    server = Server(keep_results=True, **kw)
    results = server.run(REQUESTS, domains, clock="ticks")
    return server, {r.rid: r for r in results}


def test_mix_equals_alone_bitwise(domains, alone):
    server, mix = run_mix(domains)
    assert server.peak_active >= 3, "requests must overlap on the device"
    for rid, res in mix.items():
        assert not res.error
        assert res.steps == alone[rid].steps
        assert torch.equal(res.tracers, alone[rid].tracers), f"request {rid} differs in the mix"
    # request 2 arrived at tick 3 while 0 and 1 were running; it was admitted at once
    assert mix[2].admitted == 3 and mix[2].finished == 3 + REQUESTS[2].nsteps


def test_budget_and_slot_limits_keep_results(domains, alone):
    sizes = sorted(request_bytes(r) for r in REQUESTS)
    budget = sizes[-1] + sizes[-2]  # at most two of the largest in flight
    server, mix = run_mix(domains, hbm_budget=budget)
    assert server.peak_bytes <= budget
    for rid, res in mix.items():
        assert torch.equal(res.tracers, alone[rid].tracers)
    server, mix = run_mix(domains, max_slots=1)
    assert server.peak_active == 1
    for rid, res in mix.items():
        assert torch.equal(res.tracers, alone[rid].tracers)


def test_oversized_request_is_rejected(domains):
    server = Server(hbm_budget=request_bytes(REQUESTS[0]) - 1, keep_results=True)
    results = server.run(REQUESTS[:2], domains, clock="ticks")
    by_id = {r.rid: r for r in results}
    assert "budget" in by_id[0].error
    assert not by_id[1].error  # smaller request still served


def test_results_match_fp64_oracle_and_conserve(oracle_mod, alone):
    for r in REQUESTS:
        d = inputs.steady_release(nx=r.nx, ny=r.ny, nz=r.nz, scheme=r.scheme, seed=r.seed, ntracers=r.ntracers)
        tr = np.ascontiguousarray(d.tracers)
        inv0 = d.inventory(tr)
        for _ in range(r.nsteps):
            oracle_mod.transport_step(d, tr)
        got = alone[r.rid].tracers.double().cpu().numpy()
        wet = np.broadcast_to(d.tmask == 1, tr.shape[1:])
        for t in range(r.ntracers):
            scale = np.abs(tr[t][wet]).max()
            err = np.abs(got[t][wet] - tr[t][wet]).max() / scale
            assert err <= 1e-5, (r.rid, t, err)
        # the FP64 oracle conserves these closed, steady scenarios by itself
        assert np.max(np.abs(d.inventory(tr) - inv0) / inv0) <= 1e-12
        assert alone[r.rid].inventory_drift <= 1e-12, (r.rid, alone[r.rid].inventory_drift)


def test_large_point_release_conserved_and_close_to_fp64():
    """A release on a grid far larger than the fixer's neighbourhood walk: every
    result keeps its inventory to 1e-12 and stays within 1e-5 of the FP64 method."""
    from getm_amd.serve import verify

    req = Request(rid=7, nx=300, ny=260, nz=40, ntracers=6, nsteps=8, scheme=5, seed=21)
    dom = {7: req.domain()}
    (res,) = Server(keep_results=True).run([req], dom, clock="ticks")
    assert res.inventory_drift <= 1e-12, res.inventory_drift
    errs = verify([res], [req], dom)
    assert errs[7] <= 1e-5, errs


def test_fair_policy_mix_equals_alone_bitwise(domains):
    """Fair sharing: small requests take several steps per tick (their share of the
    work quantum), so they finish in fewer ticks than steps; results still do not
    depend on the mix."""
    kw = dict(policy="fair", quantum=6e4)
    solo = run_alone(domains, **kw)
    server, mix = run_mix(domains, **kw)
    assert server.peak_active >= 2
    for rid, res in mix.items():
        assert torch.equal(res.tracers, solo[rid].tracers), f"request {rid} differs in the mix"
        assert res.inventory_drift <= 1e-12
    # at least one small request advanced more than one step per tick
    assert any(r.finished - r.admitted < r.steps for r in mix.values())


def test_uniform_policy_mix_equals_alone_bitwise(domains):
    """Uniform completion time: every request takes ceil(steps/horizon) steps per
    tick, so each completes within horizon ticks of its admission whatever its
    step count; results still do not depend on the mix."""
    horizon = 4
    kw = dict(policy="uniform", horizon=horizon)
    solo = run_alone(domains, **kw)
    server, mix = run_mix(domains, **kw)
    assert server.peak_active >= 2
    for rid, res in mix.items():
        assert not res.error
        assert torch.equal(res.tracers, solo[rid].tracers), f"request {rid} differs in the mix"
        assert res.inventory_drift <= 1e-12
        assert res.finished - res.admitted <= horizon, (rid, res.admitted, res.finished)


def test_fix_per_tick_mix_equals_alone_and_conserves(domains):
    """The fixer after every tick instead of once at the end: results still do not
    depend on the mix, and inventories hold to 1e-12."""
    solo = run_alone(domains, fix="tick")
    server, mix = run_mix(domains, fix="tick")
    for rid, res in mix.items():
        assert torch.equal(res.tracers, solo[rid].tracers), f"request {rid} differs in the mix"
        assert res.inventory_drift <= 1e-12


def test_cancelled_and_expired_requests_leave_and_free_their_slot(domains, alone):
    """A request cancelled by its client, one cancelled through Server.cancel and
    one past its deadline leave at a tick boundary with their status; the others'
    results are unchanged, and the freed slot admits a waiting request."""
    reqs = [dataclasses.replace(REQUESTS[0], cancel_after=4),  # client cancels at tick 4
            dataclasses.replace(REQUESTS[1], deadline=5),  # 9 steps cannot finish within 5 ticks
            REQUESTS[2], REQUESTS[3], REQUESTS[4]]

    def on_tick(server, tick):
        if tick == 7:
            server.cancel(3)

    server = Server(keep_results=True, max_slots=2)
    results = {r.rid: r for r in server.run(reqs, domains, clock="ticks", on_tick=on_tick)}
    assert results[0].status == "cancelled" and results[0].steps == 4
    assert results[1].status == "expired" and results[1].steps < REQUESTS[1].nsteps
    assert results[3].status == "cancelled"
    for rid in (2, 4):
        assert results[rid].status == "done" and not results[rid].error
        assert torch.equal(results[rid].tracers, alone[rid].tracers)
    # with two slots, request 2 (arrival 3) could only start once 0 or 1 had left
    assert results[2].admitted >= 4
