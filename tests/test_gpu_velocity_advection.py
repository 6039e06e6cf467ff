"""Momentum self-advection on the velocity grids (momentum.py:928-961,
1081-1097) and its pieces -- the grid interpolations (interpolation.F90) and
reconstruct_transport_change (momentum.F90:323-340) -- on the GPU, bit for bit
against the oracle and the reference's golden interpolation vectors."""
import os

import numpy as np
import pytest
import torch

from helpers import assert_bitwise

pytestmark = pytest.mark.gpu
HALO = 2


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def _lib():
    from getm_amd import _lib

    return _lib


def _interp_gpu(kind, src, dst, offs):
    L = _lib()
    lib, s, d, st = L.load(), dev(src), dev(dst), L.stream_handle()
    if kind == "x":
        rc = lib.grid_interp_x(src.shape[2], src.shape[1], src.shape[0], L.ptr(s), L.ptr(d), offs[0], st)
    elif kind == "y":
        rc = lib.grid_interp_y(src.shape[2], src.shape[1], src.shape[0], L.ptr(s), L.ptr(d), offs[0], st)
    elif kind == "z":
        rc = lib.grid_interp_z(src.shape[2], src.shape[1], src.shape[0], dst.shape[0], L.ptr(s), L.ptr(d), offs[0],
                               st)
    else:
        rc = lib.grid_interp_xy(src.shape[2], src.shape[1], dst.shape[2], dst.shape[1], src.shape[0], L.ptr(s),
                                L.ptr(d), offs[0], offs[1], st)
    L.check(rc, f"interp {kind}")
    return host(d)


def test_interp_golden_vectors_on_gpu():
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "interp_golden.npz"))
    names = sorted({k.split("/")[0] for k in g.files})
    for name in names:
        got = _interp_gpu(str(g[f"{name}/kind"]), g[f"{name}/source"], g[f"{name}/before"].copy(),
                          tuple(int(o) for o in g[f"{name}/offsets"]))
        assert_bitwise(got, g[f"{name}/after"], name)


@pytest.mark.parametrize("kind,sshape,tshape,offs", [
    ("x", (40, 31, 300), (40, 31, 300), (0,)), ("x", (1, 5, 513), (1, 5, 513), (1,)),
    ("y", (7, 260, 33), (7, 260, 33), (1,)), ("y", (3, 2, 9), (3, 2, 9), (0,)),
    ("z", (41, 20, 70), (40, 20, 70), (0,)), ("z", (40, 20, 70), (41, 20, 70), (1,)),
    ("xy", (9, 35, 257), (9, 35, 257), (1, 1)), ("xy", (9, 35, 257), (9, 35, 257), (0, 1)),
    ("xy", (2, 30, 40), (2, 31, 41), (1, 1)), ("xy", (1, 1, 8), (1, 1, 8), (0, 0))])
def test_interp_bitwise(oracle_mod, kind, sshape, tshape, offs):
    rng = np.random.default_rng(sum(sshape) + len(offs))
    src = rng.standard_normal(sshape) * 10.0 ** rng.integers(-4, 5, sshape)
    want = np.full(tshape, -3.25)
    {"x": lambda: oracle_mod.interp_x(src, want, offs[0]), "y": lambda: oracle_mod.interp_y(src, want, offs[0]),
     "z": lambda: oracle_mod.interp_z(src, want, offs[0]),
     "xy": lambda: oracle_mod.interp_xy(src, want, offs[0], offs[1])}[kind]()
    got = _interp_gpu(kind, src, np.full(tshape, -3.25), offs)
    assert_bitwise(got, want, f"{kind} {sshape} {offs}")


def test_array_interp_picks_the_reference_routine_for_every_grid_pair(oracle_mod):
    """Array.interp between T, U, V, X, UU and VV points and between layer
    centres and interfaces, against the routine core.py:407-462 selects."""
    from getm_amd.core import CENTERS, INTERFACES, Grid

    nx, ny, nz = 21, 13, 4
    grids = {n: Grid(nx, ny, nz, HALO, HALO, torch.device("cuda", 0), n) for n in ("T", "U", "V", "X", "UU", "VV")}
    rng = np.random.default_rng(5)
    shape = (nz, ny + 2 * HALO, nx + 2 * HALO)
    src = rng.standard_normal(shape)

    def check(a, b, oracle_call):
        sa = grids[a].wrap(dev(src), CENTERS)
        ta = grids[b].wrap(dev(np.full(shape, 9.5)), CENTERS)
        sa.interp(ta)
        want = np.full(shape, 9.5)
        oracle_call(want)
        assert_bitwise(host(ta.all_values), want, f"{a} -> {b}")

    check("U", "T", lambda w: oracle_mod.interp_x(src, w, 1))
    check("T", "U", lambda w: oracle_mod.interp_x(src, w, 0))
    check("V", "T", lambda w: oracle_mod.interp_y(src, w, 1))
    check("T", "V", lambda w: oracle_mod.interp_y(src, w, 0))
    check("X", "T", lambda w: oracle_mod.interp_xy(src, w, 0, 0))
    check("T", "X", lambda w: oracle_mod.interp_xy(src, w, 1, 1))
    check("V", "U", lambda w: oracle_mod.interp_xy(src, w, 0, 1))
    check("U", "V", lambda w: oracle_mod.interp_xy(src, w, 1, 0))
    check("U", "UU", lambda w: oracle_mod.interp_x(src, w, 0))

    def copy_x(w):
        w[..., :-1] = src[..., 1:]

    def copy_y(w):
        w[..., :-1, :] = src[..., 1:, :]

    check("T", "UU", copy_x)
    check("T", "VV", copy_y)
    # 2D fields go through the same routines with one level
    s2 = grids["T"].wrap(dev(src[0]))
    t2 = grids["U"].wrap(dev(np.full(shape[1:], 9.5)))
    s2.interp(t2)
    want = np.full((1,) + shape[1:], 9.5)
    oracle_mod.interp_x(np.ascontiguousarray(src[:1]), want, 0)
    assert_bitwise(host(t2.all_values), want[0], "2D T -> U")
    # layer centres <-> interfaces of one grid
    T = grids["T"]
    c = T.wrap(dev(src), CENTERS)
    f = T.wrap(dev(np.full((nz + 1,) + shape[1:], 9.5)), INTERFACES)
    c.interp(f)
    want = np.full((nz + 1,) + shape[1:], 9.5)
    oracle_mod.interp_z(src, want, 1)
    assert_bitwise(host(f.all_values), want, "centres -> interfaces")
    back = T.wrap(dev(np.full(shape, 9.5)), CENTERS)
    f.interp(back)
    want2 = np.full(shape, 9.5)
    oracle_mod.interp_z(np.ascontiguousarray(host(f.all_values)), want2, 0)
    assert_bitwise(host(back.all_values), want2, "interfaces -> centres")


def test_reconstruct_transport_change_bitwise(oracle_mod):
    from getm_amd.core import CENTERS, Grid
    from getm_amd.momentum import reconstruct_transport_change

    g = Grid(33, 17, 6, HALO, HALO, torch.device("cuda", 0), "U")
    rng = np.random.default_rng(12)
    shape = (6, 17 + 2 * HALO, 33 + 2 * HALO)
    un, hn, Uo = rng.standard_normal(shape), rng.uniform(0.1, 5.0, shape), rng.standard_normal(shape) * 3
    for dt in (30.0, 7.3):
        want = un.copy()
        oracle_mod.reconstruct_transport_change(want, hn, Uo, dt, HALO + 1, HALO + 17)
        a = g.wrap(dev(un), CENTERS)
        reconstruct_transport_change(a, dev(hn), g.wrap(dev(Uo), CENTERS), dt)
        assert_bitwise(host(a.all_values), want, f"dt {dt}")
        assert np.array_equal(host(a.all_values)[:, :HALO], un[:, :HALO])
    # 2D transports: one level
    a2 = g.wrap(dev(un[0]))
    reconstruct_transport_change(a2, dev(hn[0]), g.wrap(dev(Uo[0])), 30.0)
    want = un[:1].copy()
    oracle_mod.reconstruct_transport_change(want, np.ascontiguousarray(hn[:1]), np.ascontiguousarray(Uo[:1]), 30.0,
                                            HALO + 1, HALO + 17)
    assert_bitwise(host(a2.all_values), want[0], "2D")


def _ring(a, h=HALO):
    a[..., :h, :] = 0
    a[..., -h:, :] = 0
    a[..., :, :h] = 0
    a[..., :, -h:] = 0
    return a


class _VelocityCase:
    """U and V grids with their advection grids (UU, UV and VU, VV), random masks
    that are 0 on the outer halo ring (as a domain's are), metrics, thicknesses
    and depths, plus velocities, transports and ww on the T grid."""

    def __init__(self, seed, nx=37, ny=29, nz=7, cfl=0.3):
        from getm_amd.core import CENTERS, INTERFACES, Grid

        rng = np.random.default_rng(seed)
        self.nz, ny_, nx_ = nz, ny + 2 * HALO, nx + 2 * HALO
        self.shape3, self.shape2 = (nz, ny_, nx_), (ny_, nx_)
        d0 = torch.device("cuda", 0)
        mk = lambda name: Grid(nx, ny, nz, HALO, HALO, d0, name)
        self.T, self.U, self.V = mk("T"), mk("U"), mk("V")
        self.UU, self.UV, self.VU, self.VV = mk("UU"), mk("UV"), mk("VU"), mk("VV")
        self.U.ugrid, self.U.vgrid, self.V.ugrid, self.V.vgrid = self.UU, self.UV, self.VU, self.VV
        self.np = {}
        dx = 500.0
        dt = 20.0
        for name, g in (("T", self.T), ("U", self.U), ("V", self.V), ("UU", self.UU), ("UV", self.UV),
                        ("VU", self.VU), ("VV", self.VV)):
            f = {"mask": _ring((rng.random(self.shape2) > 0.15).astype(np.int32))}
            f["iarea"] = rng.uniform(0.8, 1.2, self.shape2) / dx ** 2
            f["idx"] = rng.uniform(0.8, 1.2, self.shape2) / dx
            f["idy"] = rng.uniform(0.8, 1.2, self.shape2) / dx
            f["dx"] = rng.uniform(0.8, 1.2, self.shape2) * dx
            f["dy"] = rng.uniform(0.8, 1.2, self.shape2) * dx
            f["hn"] = rng.uniform(0.5, 4.0, self.shape3)
            f["ho"] = f["hn"] * rng.uniform(0.95, 1.05, self.shape3)
            f["D"] = f["hn"].sum(axis=0)
            for k, v in f.items():
                g.set_field(k, v)
            self.np[name] = f
        umax = cfl * dx / dt
        self.dt = dt
        self.uk = rng.uniform(-umax, umax, self.shape3)
        self.vk = rng.uniform(-umax, umax, self.shape3)
        self.pk = self.uk * self.np["U"]["hn"] * rng.uniform(0.9, 1.1, self.shape3)
        self.qk = self.vk * self.np["V"]["hn"] * rng.uniform(0.9, 1.1, self.shape3)
        self.ww = rng.uniform(-2e-3, 2e-3, (nz + 1,) + self.shape2)
        self.ww[0] = 0.0
        self.ww[-1] = 0.0
        self.CENTERS, self.INTERFACES = CENTERS, INTERFACES


def _oracle_apply_3d(oracle_mod, case, g, ug, vg, u_face, v_face, w, var, scheme):
    """Advection.apply_3d(new_h=True) as the reference's call order of single
    sweeps: v dt/2, u dt/2, w dt, u dt/2, v dt/2 (ufirst False)."""
    G, GU, GV = case.np[g], case.np[ug], case.np[vg]
    h = G["hn"].copy()
    half = 0.5 * case.dt
    uargs = (GU["mask"], GU["idx"], GU["dy"], GU["hn"], u_face, G["mask"], G["iarea"], None, half)
    vargs = (GV["mask"], GV["dx"], GV["idy"], GV["hn"], v_face, G["mask"], G["iarea"], None, half)
    oracle_mod.advection_uv(2, scheme, *vargs, h, var)
    oracle_mod.advection_uv(1, scheme, *uargs, h, var)
    oracle_mod.advection_w(scheme, w, w, G["mask"], case.dt, h, var)
    oracle_mod.advection_uv(1, scheme, *uargs, h, var)
    oracle_mod.advection_uv(2, scheme, *vargs, h, var)
    return h


@pytest.mark.parametrize("scheme", [1, 5, 6])
@pytest.mark.parametrize("seed", [3, 4])
def test_velocity_advection_3d_bitwise(oracle_mod, seed, scheme):
    from getm_amd.momentum import VelocityAdvection

    c = _VelocityCase(seed)
    C, I = c.CENTERS, c.INTERFACES
    va = VelocityAdvection(c.U, c.V, scheme)
    uk, vk = c.U.wrap(dev(c.uk), C), c.V.wrap(dev(c.vk), C)
    pk, qk = c.U.wrap(dev(c.pk), C), c.V.wrap(dev(c.qk), C)
    ww = c.T.wrap(dev(c.ww), I)
    advpk, advqk = c.U.array(C), c.V.array(C)
    va.advect_3d(uk, vk, ww, pk, qk, c.dt, advpk, advqk)

    nan3 = lambda: np.full(c.shape3, np.nan)
    # U velocity on its own advection grids: U -> UU and V -> UV are x averages, T -> U for ww
    uua, uva, wwu = nan3(), nan3(), np.zeros(c.ww.shape)
    oracle_mod.interp_x(c.uk, uua, 0)
    oracle_mod.interp_x(c.vk, uva, 0)
    oracle_mod.interp_x(c.ww, wwu, 0)
    fu = c.uk.copy()
    hu = _oracle_apply_3d(oracle_mod, c, "U", "UU", "UV", uua, uva, wwu, fu, scheme)
    oracle_mod.reconstruct_transport_change(fu, hu, c.pk, c.dt, HALO + 1, HALO + c.shape3[1] - 2 * HALO)
    # V velocity: U -> VU and V -> VV are y averages, T -> V for ww
    vua, vva, wwv = nan3(), nan3(), np.zeros(c.ww.shape)
    oracle_mod.interp_y(c.uk, vua, 0)
    oracle_mod.interp_y(c.vk, vva, 0)
    oracle_mod.interp_y(c.ww, wwv, 0)
    fv = c.vk.copy()
    hv = _oracle_apply_3d(oracle_mod, c, "V", "VU", "VV", vua, vva, wwv, fv, scheme)
    oracle_mod.reconstruct_transport_change(fv, hv, c.qk, c.dt, HALO + 1, HALO + c.shape3[1] - 2 * HALO)

    assert_bitwise(host(va.uadv.h), hu, "U thickness")
    assert_bitwise(host(va.vadv.h), hv, "V thickness")
    assert_bitwise(host(advpk.all_values), fu, "advpk")
    assert_bitwise(host(advqk.all_values), fv, "advqk")
    assert np.isfinite(fu[:, HALO:-HALO, HALO:-HALO]).all()
    assert np.abs(fu[:, HALO:-HALO, HALO:-HALO]).max() > 0


@pytest.mark.parametrize("scheme", [5, 2])
def test_velocity_advection_2d_bitwise(oracle_mod, scheme):
    """advect_2d: transports interpolated and divided by the face-grid depth, the
    2D split advection (HALF: v dt/2, u dt, v dt/2), then the transport change."""
    from getm_amd.momentum import VelocityAdvection

    c = _VelocityCase(8, nz=1)
    va = VelocityAdvection(c.U, c.V, scheme)
    U2, V2 = c.pk[0] * 3.0, c.qk[0] * 3.0
    u1, v1 = c.uk[0].copy(), c.vk[0].copy()
    Ua, Va = c.U.wrap(dev(U2)), c.V.wrap(dev(V2))
    advU, advV = c.U.array(), c.V.array()
    va.advect_2d(Ua, Va, c.U.wrap(dev(u1)), c.V.wrap(dev(v1)), c.dt, advU, advV)

    def face(src, kind, g):
        t = np.full((1,) + c.shape2, np.nan)
        (oracle_mod.interp_x if kind == "x" else oracle_mod.interp_y)(np.ascontiguousarray(src[None]), t, 0)
        return t / c.np[g]["D"][None]

    def run(g, ug, vg, u_face, v_face, var0, Uo):
        G, GU, GV = c.np[g], c.np[ug], c.np[vg]
        D = G["D"][None].copy()
        var = var0[None].copy()
        step = lambda direction, dt: oracle_mod.advection_uv(
            direction, scheme, *((GU["mask"], GU["idx"], GU["dy"], GU["D"][None], u_face) if direction == 1 else
                                 (GV["mask"], GV["dx"], GV["idy"], GV["D"][None], v_face)),
            G["mask"], G["iarea"], None, dt, D, var)
        step(2, 0.5 * c.dt)
        step(1, c.dt)
        step(2, 0.5 * c.dt)
        oracle_mod.reconstruct_transport_change(var, D, np.ascontiguousarray(Uo[None]), c.dt, HALO + 1,
                                                HALO + c.shape2[0] - 2 * HALO)
        return var[0]

    fu = run("U", "UU", "UV", face(U2, "x", "UU"), face(V2, "x", "UV"), u1, U2)
    fv = run("V", "VU", "VV", face(U2, "y", "VU"), face(V2, "y", "VV"), v1, V2)
    assert_bitwise(host(advU.all_values), fu, "advU")
    assert_bitwise(host(advV.all_values), fv, "advV")
