"""Seeded synthetic inputs for the tracer-transport step.

Each generator returns a :class:`Domain` of numpy FP64/int32 fields in the
pygetm layout (C-order ``(nz, ny_, nx_)``, i fastest, halo 2) that is
continuity-consistent, so the advected layer thickness ends at ``hn`` and the
tracer inventory is conserved:

1. ``ho_k = sigma_k (H + z_old)``; U/V face thicknesses ``hu_k = sigma_k Du``
   with ``Du`` the mean of the two adjacent T-column depths.
2. layer transports ``p_k = u hu dyu`` and ``q_k = v hv dxv`` on open faces
   (mask 1 or 2), zero elsewhere; ``div_k`` their discrete divergence.
3. ``z_new = z_old - dt sum_k div_k iA``; ``hn_k = sigma_k (H + z_new)``.
4. ``w_0 = 0``, ``w_k = w_{k-1} - (hn_k - ho_k)/dt - div_k iA`` for
   k = 1..nz-1 and ``w_nz = 0`` (the recipe of python/src/momentum.F90:191-216).

The tracer-transport step reads: masks (T: 0 land / 1 water; U, V: 1 where
both T neighbours are water, domain.py:918-940), metrics ``idxu, dyu, dxv,
idyv, iarea`` (core.py:324-378), thicknesses ``ho, hn, hu, hv``, velocities
``u, v, w``, diffusivity ``nuh`` and the tracers.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Optional

import numpy as np

from .vertical_coordinates import calculate_sigma

HALO = 2
M2_PERIOD = 44712.0


@dataclasses.dataclass
class Domain:
    nx: int
    ny: int
    nz: int
    halo: int
    tmask: np.ndarray
    umask: np.ndarray
    vmask: np.ndarray
    idxu: np.ndarray
    dyu: np.ndarray
    dxv: np.ndarray
    idyv: np.ndarray
    iarea: np.ndarray
    area: np.ndarray
    H: np.ndarray
    z: np.ndarray
    sigma: np.ndarray
    ho: np.ndarray
    hn: np.ndarray
    hu: np.ndarray
    hv: np.ndarray
    u: np.ndarray
    v: np.ndarray
    w: np.ndarray
    nuh: np.ndarray
    tracers: np.ndarray  # (T, nz, ny_, nx_)
    dt: float
    scheme: int
    cnpar: float = 1.0
    name: str = ""
    # sparse initial tracers (steady_release(dense=False)): rows (k, j, i, value), one per tracer,
    # every other cell zero; ``tracers`` is then empty
    releases: Optional[np.ndarray] = None

    @property
    def nx_(self) -> int:
        return self.nx + 2 * self.halo

    @property
    def ny_(self) -> int:
        return self.ny + 2 * self.halo

    @property
    def ntracers(self) -> int:
        return self.tracers.shape[0] if self.releases is None else self.releases.shape[0]

    @property
    def n_wet_columns(self) -> int:
        return int((self.tmask == 1).sum())

    @property
    def n_wet_cells(self) -> int:
        return self.n_wet_columns * self.nz

    def inventory(self, tracers: Optional[np.ndarray] = None, h: Optional[np.ndarray] = None) -> np.ndarray:
        """Per-tracer sum of area * h * c over wet cells (simulation.py:1408-1466), in FP64
        with exact (math.fsum) accumulation."""
        tr = self.tracers if tracers is None else tracers
        h = self.hn if h is None else h
        wet = np.broadcast_to(self.tmask == 1, h.shape)
        w = (h * self.area)[wet]
        return np.array([math.fsum((w * t[wet]).tolist()) for t in tr])

    def copy(self) -> "Domain":
        return dataclasses.replace(self, **{f.name: getattr(self, f.name).copy()
                                            for f in dataclasses.fields(self)
                                            if isinstance(getattr(self, f.name), np.ndarray)})


def uv_masks(tmask: np.ndarray):
    """U point i sits between T points i and i+1; V point j between j and j+1."""
    wet = tmask == 1
    umask = np.zeros_like(tmask)
    vmask = np.zeros_like(tmask)
    umask[:, :-1] = (wet[:, :-1] & wet[:, 1:]).astype(np.int32)
    vmask[:-1, :] = (wet[:-1, :] & wet[1:, :]).astype(np.int32)
    return umask, vmask


def _continuity(tmask, umask, vmask, dyu, dxv, iarea, H, z_old, sigma, u, v, dt):
    """Steps 1-4 of the module docstring.  Returns ho, hn, hu, hv, w, z_new."""
    nz = sigma.size
    s = sigma[:, None, None]
    D = np.where(tmask == 1, H + z_old, 0.0)
    Du = np.zeros_like(D)
    Du[:, :-1] = 0.5 * (D[:, :-1] + D[:, 1:])
    Dv = np.zeros_like(D)
    Dv[:-1, :] = 0.5 * (D[:-1, :] + D[1:, :])
    ho = s * D
    hu = s * Du
    hv = s * Dv
    open_u = (umask == 1) | (umask == 2)
    open_v = (vmask == 1) | (vmask == 2)
    p = np.where(open_u, (u * hu) * dyu, 0.0)
    q = np.where(open_v, (v * hv) * dxv, 0.0)
    div = np.zeros_like(p)
    div[:, :, 1:] += p[:, :, 1:] - p[:, :, :-1]
    div[:, 1:, :] += q[:, 1:, :] - q[:, :-1, :]
    div *= iarea
    wet = tmask == 1
    div = np.where(wet, div, 0.0)
    z_new = np.where(wet, z_old - dt * div.sum(axis=0), z_old)
    Dn = np.where(wet, H + z_new, 0.0)
    hn = s * Dn
    w = np.zeros((nz + 1,) + D.shape)
    dh = (hn - ho) / dt + div
    w[1:nz] = -np.cumsum(dh[:-1], axis=0)
    w[:, ~wet] = 0.0
    return ho, hn, hu, hv, w, z_new


def _smooth_noise(rng, shape, sigma_cells):
    from scipy.ndimage import gaussian_filter

    a = gaussian_filter(rng.standard_normal(shape), sigma_cells, mode="wrap")
    return (a - a.min()) / max(a.max() - a.min(), 1e-300)


def _pad(a2d: np.ndarray, halo: int, value=0):
    return np.pad(a2d, halo, constant_values=value)


def _metrics(ny_, nx_, dx, dy):
    full = lambda x: np.full((ny_, nx_), x, dtype=np.float64)
    return dict(idxu=full(1.0 / dx), dyu=full(dy), dxv=full(dx), idyv=full(1.0 / dy),
                iarea=full(1.0 / (dx * dy)), area=full(dx * dy))


def seamount(nx=100, ny=100, nz=30, dx=1000.0, dt=60.0, scheme=5, seed=0, cfl=0.4,
             ntracers=1) -> Domain:
    """Seamount / solid-body-rotation case (100x100x30 at full size).

    Bathymetry H = 4500 (1 - 0.9 exp(-r^2/(25 km)^2)) centred; outer ring of the
    interior is land; 30 equal sigma layers; u = -omega (y - yc), v = omega (x - xc)
    with omega set so the largest horizontal CFL at the full dt is ``cfl``; tracer =
    Gaussian (sigma 10 km, amplitude 1, centred 25 km east of the centre) + U(0, 0.01)
    noise; nuh = 10**U(-6, -2) per interface; SUPERBEE by default; cnpar 1; dt 60 s."""
    rng = np.random.default_rng(seed)
    h = HALO
    ny_, nx_ = ny + 2 * h, nx + 2 * h
    dy = dx
    # interior T centres at (i + 0.5) dx; U points at (i + 1) dx; V points at (j + 1) dy
    xi = (np.arange(nx_) - h + 0.5) * dx
    yj = (np.arange(ny_) - h + 0.5) * dy
    xc, yc = 0.5 * nx * dx, 0.5 * ny * dy
    X, Y = np.meshgrid(xi, yj)
    r2 = (X - xc) ** 2 + (Y - yc) ** 2
    Hfull = 4500.0 * (1.0 - 0.9 * np.exp(-r2 / (25e3) ** 2))
    interior = np.zeros((ny, nx), dtype=np.int32)
    interior[1:-1, 1:-1] = 1
    tmask = _pad(interior, h).astype(np.int32)
    H = np.where(tmask == 1, Hfull, 0.0)
    umask, vmask = uv_masks(tmask)
    m = _metrics(ny_, nx_, dx, dy)
    sigma = calculate_sigma(nz)
    Yu = Y
    Xv = X
    u2 = -(Yu - yc)
    v2 = (Xv - xc)
    umax = max(np.abs(u2[umask == 1]).max(), np.abs(v2[vmask == 1]).max())
    omega = cfl * dx / (dt * umax)
    u = np.broadcast_to(omega * u2, (nz, ny_, nx_)).copy()
    v = np.broadcast_to(omega * v2, (nz, ny_, nx_)).copy()
    z = np.zeros((ny_, nx_))
    ho, hn, hu, hv, w, _ = _continuity(tmask, umask, vmask, m["dyu"], m["dxv"], m["iarea"], H, z, sigma,
                                       u, v, dt)
    nuh = 10.0 ** rng.uniform(-6.0, -2.0, size=(nz + 1, ny_, nx_))
    blob = np.exp(-((X - xc - 25e3) ** 2 + (Y - yc) ** 2) / (2 * (10e3) ** 2))
    tr = np.empty((ntracers, nz, ny_, nx_))
    for t in range(ntracers):
        tr[t] = blob[None] + rng.uniform(0.0, 0.01, size=(nz, ny_, nx_))
    return Domain(nx, ny, nz, h, tmask, umask, vmask, m["idxu"], m["dyu"], m["dxv"], m["idyv"], m["iarea"],
                  m["area"], H, z, sigma, ho, hn, hu, hv, u, v, w, nuh, tr, dt, scheme, 1.0, "seamount")


def coastal(nx=1024, ny=768, nz=40, dx=2000.0, dt=30.0, scheme=5, seed=1, land_fraction=0.3,
            ntracers=2, time=None, nuh_range=(-6.0, -1.0), tracer_kind="TS") -> Domain:
    """Large coastal case (1024x768x40, 2 km grid at full size).

    Land = Gaussian-smoothed (sigma 8 cells) seeded noise below its ``land_fraction``
    quantile; depth 10-200 m log-uniform from noise smoothed the same way; 40 sigma
    layers with zooming factors ddl = ddu = 1; horizontal velocity = discrete curl of
    a smoothed streamfunction scaled to 1.0 m/s maximum, times a linear vertical shear
    profile (0.5 at the bottom to 1.5 at the surface), plus a uniform 0.5 m/s x-current
    oscillating with the M2 period (sin(2 pi t / 44712 s), t = ``time``, default T/4);
    w from continuity.  Tracers: "TS" gives T ~ U(5, 20) and S ~ U(30, 35) smoothed,
    "blobs" gives ``ntracers`` sums of 4 Gaussian blobs with amplitude 10**U(-3, 1).
    nuh = 10**U(nuh_range) per interface."""
    rng = np.random.default_rng(seed)
    h = HALO
    ny_, nx_ = ny + 2 * h, nx + 2 * h
    dy = dx
    sm = _smooth_noise(rng, (ny, nx), 8.0)
    thr = np.quantile(sm, land_fraction)
    interior = (sm > thr).astype(np.int32)
    tmask = _pad(interior, h).astype(np.int32)
    dep = _smooth_noise(rng, (ny, nx), 8.0)
    Hi = 10.0 ** (1.0 + dep * (math.log10(200.0) - 1.0))
    H = np.where(tmask == 1, _pad(Hi, h), 0.0)
    umask, vmask = uv_masks(tmask)
    m = _metrics(ny_, nx_, dx, dy)
    sigma = calculate_sigma(nz, 1.0, 1.0)
    psi = _pad(_smooth_noise(rng, (ny, nx), 8.0), h)
    # u = -dpsi/dy at U points, v = dpsi/dx at V points (psi at T points, averaged)
    psi_c = np.zeros((ny_ + 1, nx_ + 1))
    psi_c[1:, 1:] = psi
    u2 = np.zeros((ny_, nx_))
    v2 = np.zeros((ny_, nx_))
    u2[1:-1, :-1] = -((psi[2:, :-1] + psi[2:, 1:]) - (psi[:-2, :-1] + psi[:-2, 1:])) / (4.0 * dy)
    v2[:-1, 1:-1] = ((psi[:-1, 2:] + psi[1:, 2:]) - (psi[:-1, :-2] + psi[1:, :-2])) / (4.0 * dx)
    sc = max(np.abs(u2).max(), np.abs(v2).max(), 1e-300)
    u2 *= 1.0 / sc
    v2 *= 1.0 / sc
    t = M2_PERIOD / 4 if time is None else time
    u2 = u2 + 0.5 * math.sin(2 * math.pi * t / M2_PERIOD)
    prof = (0.5 + (np.arange(nz) + 0.5) / nz)[:, None, None]
    u = prof * u2[None]
    v = prof * v2[None]
    z = np.zeros((ny_, nx_))
    ho, hn, hu, hv, w, _ = _continuity(tmask, umask, vmask, m["dyu"], m["dxv"], m["iarea"], H, z, sigma,
                                       u, v, dt)
    nuh = 10.0 ** rng.uniform(nuh_range[0], nuh_range[1], size=(nz + 1, ny_, nx_))
    tr = make_tracers(rng, tracer_kind, ntracers, nx, ny, nz, h)
    return Domain(nx, ny, nz, h, tmask, umask, vmask, m["idxu"], m["dyu"], m["dxv"], m["idyv"], m["iarea"],
                  m["area"], H, z, sigma, ho, hn, hu, hv, u, v, w, nuh, tr, dt, scheme, 1.0, "coastal")


def make_tracers(rng, kind, ntracers, nx, ny, nz, halo=HALO) -> np.ndarray:
    ny_, nx_ = ny + 2 * halo, nx + 2 * halo
    tr = np.empty((ntracers, nz, ny_, nx_))
    if kind == "TS":
        ranges = [(5.0, 20.0), (30.0, 35.0)]
        for t in range(ntracers):
            lo, hi = ranges[t % 2]
            base = np.stack([_pad(_smooth_noise(rng, (ny, nx), 8.0), halo) for _ in range(1)])[0]
            tr[t] = lo + (hi - lo) * (0.7 * base[None] + 0.3 * rng.uniform(size=(nz, ny_, nx_)))
    elif kind == "blobs":
        xs = np.arange(nx_, dtype=np.float64)
        ys = np.arange(ny_, dtype=np.float64)
        zs = np.arange(nz, dtype=np.float64)
        for t in range(ntracers):
            acc = np.zeros((nz, ny_, nx_))
            for _ in range(4):
                amp = 10.0 ** rng.uniform(-3.0, 1.0)
                x0, y0, z0 = rng.uniform(0, nx_), rng.uniform(0, ny_), rng.uniform(0, nz)
                sx, sy, sz = rng.uniform(0.03, 0.15) * nx, rng.uniform(0.03, 0.15) * ny, rng.uniform(0.1, 0.5) * nz
                gx = np.exp(-0.5 * ((xs - x0) / sx) ** 2)
                gy = np.exp(-0.5 * ((ys - y0) / sy) ** 2)
                gz = np.exp(-0.5 * ((zs - z0) / sz) ** 2)
                acc += amp * gz[:, None, None] * gy[None, :, None] * gx[None, None, :]
            tr[t] = acc
    else:
        raise ValueError(kind)
    return tr


def tiny(nx=24, ny=20, nz=8, seed=0, scheme=5, dt=None, land_fraction=0.25, cfl=0.1, wscale=1.0,
         ntracers=2, mask_codes=False) -> Domain:
    """Small random case for bitwise tests: random land mask, random depths, random
    signed velocities (both signs of u, v, w), random tracers with sharp features so
    every limiter branch is visited.  ``wscale`` > 1 drives vertical CFL above 1 so
    the w sweep sub-steps and restarts.  ``mask_codes`` additionally marks some open
    faces with mask 2 and some T cells with 2 (open-boundary codes) to exercise the
    mask semantics."""
    rng = np.random.default_rng(seed)
    h = HALO
    dx = 500.0
    dt = 20.0 if dt is None else dt
    ny_, nx_ = ny + 2 * h, nx + 2 * h
    interior = (rng.uniform(size=(ny, nx)) > land_fraction).astype(np.int32)
    tmask = _pad(interior, h).astype(np.int32)
    H = np.where(tmask == 1, rng.uniform(5.0, 50.0, size=(ny_, nx_)), 0.0)
    umask, vmask = uv_masks(tmask)
    if mask_codes:
        umask[(umask == 1) & (rng.uniform(size=umask.shape) < 0.2)] = 2
        vmask[(vmask == 1) & (rng.uniform(size=vmask.shape) < 0.2)] = 2
        umask[(umask == 0) & (rng.uniform(size=umask.shape) < 0.1)] = 3
        vmask[(vmask == 0) & (rng.uniform(size=vmask.shape) < 0.1)] = 4
    m = _metrics(ny_, nx_, dx, dx)
    m["idxu"] = m["idxu"] * rng.uniform(0.8, 1.2, size=(ny_, nx_))
    m["idyv"] = m["idyv"] * rng.uniform(0.8, 1.2, size=(ny_, nx_))
    sigma = rng.uniform(0.5, 1.5, size=nz)
    sigma /= sigma.sum()
    umax = cfl * dx / dt
    u = rng.uniform(-umax, umax, size=(nz, ny_, nx_))
    v = rng.uniform(-umax, umax, size=(nz, ny_, nx_))
    z = np.zeros((ny_, nx_))
    ho, hn, hu, hv, w, _ = _continuity(tmask, umask, vmask, m["dyu"], m["dxv"], m["iarea"], H, z, sigma,
                                       u, v, dt)
    # extra vertical velocity, smooth in k so layers stay positive: vertical CFL up to ~wscale
    kk = np.arange(1, nz)
    hmin = np.where(tmask == 1, ho.min(axis=0), 0.0)
    r = rng.uniform(-1.0, 1.0, size=(ny_, nx_))
    w[1:-1] += wscale * r[None] * np.sin(np.pi * kk / nz)[:, None, None] * hmin[None] / dt
    w[:, tmask != 1] = 0.0
    nuh = 10.0 ** rng.uniform(-6.0, -1.0, size=(nz + 1, ny_, nx_))
    tr = rng.uniform(0.0, 1.0, size=(ntracers, nz, ny_, nx_))
    tr[:, :, ::3, :] += 2.0  # stripes so slope ratios take every sign and size
    tr *= np.where(rng.uniform(size=tr.shape) < 0.1, 5.0, 1.0)
    if mask_codes:
        tm = tmask.copy()
        tm[(tm == 1) & (rng.uniform(size=tm.shape) < 0.05)] = 2
        tmask = tm
    return Domain(nx, ny, nz, h, tmask, umask, vmask, m["idxu"], m["dyu"], m["dxv"], m["idyv"], m["iarea"],
                  m["area"], H, z, sigma, ho, hn, hu, hv, u, v, w, nuh, tr, dt, scheme, 1.0, "tiny")


def ensemble_specs(ndomains=96, seed=3, size_range=(64, 768), nz_range=(10, 60), tracer_range=(1, 32)):
    """(nx, ny, nz, tracers, field seed) per domain of an uneven ensemble: nx, ny
    log-uniform in ``size_range``, nz uniform in ``nz_range``, tracers uniform in
    ``tracer_range``."""
    rng = np.random.default_rng(seed)
    out = []
    lo, hi = math.log(size_range[0]), math.log(size_range[1])
    for d in range(ndomains):
        nx = int(round(math.exp(rng.uniform(lo, hi))))
        ny = int(round(math.exp(rng.uniform(lo, hi))))
        nz = int(rng.integers(nz_range[0], nz_range[1] + 1))
        nt = int(rng.integers(tracer_range[0], tracer_range[1] + 1))
        out.append((nx, ny, nz, nt, 1000 + d))
    return out


def ensemble_domains(ndomains=96, seed=3, size_range=(64, 768), nz_range=(10, 60), tracer_range=(1, 32),
                     scheme=5, dt=30.0):
    """Uneven ensemble (:func:`ensemble_specs`), fields and blob tracers generated as
    by :func:`coastal` from a per-domain seed."""
    return [coastal(nx=nx, ny=ny, nz=nz, dx=2000.0, dt=dt, scheme=scheme, seed=s, ntracers=nt, tracer_kind="blobs")
            for nx, ny, nz, nt, s in ensemble_specs(ndomains, seed, size_range, nz_range, tracer_range)]


def steady_release(nx=96, ny=80, nz=20, dx=2000.0, dt=30.0, scheme=5, seed=0, ntracers=1, amount=1e3,
                   land_fraction=0.3, umax=0.5, baroclinic=0.3, dense=True) -> Domain:
    """Dye-release scenario with fields that are valid for every step.

    Land and depth as in :func:`coastal`.  Layer transports are a discretely
    divergence-free barotropic part (from a corner streamfunction) plus a
    baroclinic part a_k * P with sum_k a_k = 0, so every layer diverges but no
    column does: the surface stays put, ``ho == hn`` and w follows from
    continuity -- repeated steps conserve each tracer's inventory exactly.
    Each tracer starts as ``amount`` units released in one random wet cell;
    with ``dense=False`` these come as the ``releases`` table instead of a dense
    ``tracers`` array."""
    rng = np.random.default_rng(seed)
    h = HALO
    ny_, nx_ = ny + 2 * h, nx + 2 * h
    sm = _smooth_noise(rng, (ny, nx), 8.0)
    tmask = _pad((sm > np.quantile(sm, land_fraction)).astype(np.int32), h).astype(np.int32)
    dep = _smooth_noise(rng, (ny, nx), 8.0)
    H = np.where(tmask == 1, _pad(10.0 ** (1.0 + dep * (math.log10(200.0) - 1.0)), h), 0.0)
    umask, vmask = uv_masks(tmask)
    m = _metrics(ny_, nx_, dx, dx)
    sigma = calculate_sigma(nz, 1.0, 1.0)
    s = sigma[:, None, None]
    D = np.where(tmask == 1, H, 0.0)
    Du = np.zeros_like(D)
    Du[:, :-1] = 0.5 * (D[:, :-1] + D[:, 1:])
    Dv = np.zeros_like(D)
    Dv[:-1, :] = 0.5 * (D[:-1, :] + D[1:, :])
    hu, hv, ho = s * Du, s * Dv, s * D
    # barotropic transports per unit depth from a streamfunction on cell corners (i+1/2, j+1/2)
    psi = _smooth_noise(rng, (ny_, nx_), 6.0) - 0.5
    # psi[j, i] sits on corner (i+1/2, j+1/2); zero on corners touching land so that
    # closed faces carry no transport and every column stays divergence-free
    land = tmask != 1
    touch = land.copy()
    touch[:, :-1] |= land[:, 1:]
    touch[:-1, :] |= land[1:, :]
    touch[:-1, :-1] |= land[1:, 1:]
    psi = np.where(touch, 0.0, psi)
    pu = np.zeros((ny_, nx_))
    qv = np.zeros((ny_, nx_))
    pu[1:, :] = -(psi[1:, :] - psi[:-1, :])   # U point (i+1/2, j): corners (i+1/2, j +- 1/2)
    qv[:, 1:] = psi[:, 1:] - psi[:, :-1]      # V point (i, j+1/2): corners (i +- 1/2, j+1/2)
    # baroclinic, zero-sum over layers: a_k = cos(pi (k + 1/2) / nz) sigma-weighted to sum 0
    a = np.cos(np.pi * (np.arange(nz) + 0.5) / nz)
    a = a - (a * sigma).sum() / sigma.sum()
    P = _smooth_noise(rng, (ny_, nx_), 6.0) - 0.5
    open_u, open_v = (umask == 1), (vmask == 1)
    p = np.zeros((nz, ny_, nx_))
    q = np.zeros((nz, ny_, nx_))
    for k in range(nz):
        p[k] = np.where(open_u, sigma[k] * pu + baroclinic * sigma[k] * a[k] * P, 0.0)
        q[k] = np.where(open_v, sigma[k] * qv + baroclinic * sigma[k] * a[k] * np.roll(P, 1, axis=1), 0.0)
    # velocities reproducing these transports: u = p / (hu dyu)
    with np.errstate(divide="ignore", invalid="ignore"):
        u = np.where(open_u & (hu > 0), p / (hu * m["dyu"]), 0.0)
        v = np.where(open_v & (hv > 0), q / (hv * m["dxv"]), 0.0)
    sc = max(np.abs(u).max(), np.abs(v).max(), 1e-300)
    u *= umax / sc
    v *= umax / sc
    p = np.where(open_u, (u * hu) * m["dyu"], 0.0)
    q = np.where(open_v, (v * hv) * m["dxv"], 0.0)
    div = np.zeros_like(p)
    div[:, :, 1:] += p[:, :, 1:] - p[:, :, :-1]
    div[:, 1:, :] += q[:, 1:, :] - q[:, :-1, :]
    div *= m["iarea"]
    wet = tmask == 1
    div = np.where(wet, div, 0.0)
    w = np.zeros((nz + 1, ny_, nx_))
    w[1:nz] = -np.cumsum(div[:-1], axis=0)
    w[:, ~wet] = 0.0
    nuh = 10.0 ** rng.uniform(-6.0, -2.0, size=(nz + 1, ny_, nx_))
    jj, ii = np.nonzero(wet)
    rel = np.zeros((ntracers, 4))
    for t in range(ntracers):
        c = rng.integers(len(ii))
        k = rng.integers(nz)
        rel[t] = (k, jj[c], ii[c], amount / (m["area"][jj[c], ii[c]] * ho[k, jj[c], ii[c]]))
    if dense:
        tr = np.zeros((ntracers, nz, ny_, nx_))
        for t, (k, j, i, val) in enumerate(rel):
            tr[t, int(k), int(j), int(i)] = val
    else:
        tr = np.zeros((0, nz, ny_, nx_))
    z = np.zeros((ny_, nx_))
    return Domain(nx, ny, nz, h, tmask, umask, vmask, m["idxu"], m["dyu"], m["dxv"], m["idyv"], m["iarea"],
                  m["area"], H, z, sigma, ho, ho.copy(), hu, hv, u, v, w, nuh, tr, dt, scheme, 1.0, "release",
                  None if dense else rel)
