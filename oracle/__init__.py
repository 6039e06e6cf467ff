"""FP64 CPU oracle for pygetm's 3D tracer-transport step.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker (or the timed CPU baseline); the product path (``getm_amd``) never
imports, links or calls it.

Two implementations live here:

* ``libgetm_oracle.so`` -- a C restatement (getm_oracle.c) of the reference
  Fortran in its FP64 operation order (file:line citations in that file).
* ``_ref/libgetm_ref.so`` -- the reference's own advection Fortran
  (src/operators/advection_schemes/) compiled from /root/reference with
  amdflang by ``oracle/Makefile`` plus the bind(c) driver ``ref_shim.F90``.
  It pins the C restatement of the x/y/z sweeps and the six limiters bit for
  bit; python/src/interpolation.F90 (self-contained bind(c) routines) is
  compiled into the same library and pins the grid interpolations.  The diffusion Fortran (src/operators/diffusion.F90) cannot be built
  here without stand-ins for modules the image lacks (it needs getm_domain ->
  flexout), so its restatement is pinned by closed forms and invariants in
  tests/test_oracle.py instead.

All arrays are numpy, C-order, padded with the halo: 3D (nz, ny_, nx_),
interface fields (nz+1, ny_, nx_), 2D (ny_, nx_); float64 and int32.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgetm_oracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libgetm_ref.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lib = None
_ref = None


def build(ref: bool = False) -> None:
    """Compile the C restatement (and, if asked and /root/reference exists, _ref)."""
    import subprocess

    subprocess.run(["make", "-s", "-C", HERE, "c"], check=True)
    if ref and os.path.isdir("/root/reference"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.oracle_limiter.restype = ctypes.c_double
        _lib.oracle_limiter.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref_lib():
    global _ref
    if _ref is None:
        _ref = ctypes.CDLL(REF_PATH)
    return _ref


def _d(a: Optional[np.ndarray]):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(_dp)


def _i(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_ip)


def _check(rc: int, what: str):
    if rc != 0:
        raise ValueError(f"{what} failed with status {rc}")


def _dims(var: np.ndarray, halo: int):
    nz, ny_, nx_ = var.shape
    return nx_ - 2 * halo, ny_ - 2 * halo, nz


def limiter(scheme: int, cfl: float, ratio: float) -> float:
    return lib().oracle_limiter(int(scheme), float(cfl), float(ratio))


def advection_uv(direction, scheme, uvmask, metric1, metric2, huv, uv, tmask, iarea, ah, dt, h, var,
                 halo=2, impl="c"):
    """In-place x (direction 1) or y (direction 2) half-sweep over all layers.

    direction 1: metric1 = idxu, metric2 = dyu; direction 2: metric1 = dxv, metric2 = idyv.
    impl="ref" runs the reference Fortran (oracle/_ref) instead of the C restatement."""
    nx, ny, nk = _dims(var, halo)
    if impl == "ref":
        dummy = ah if ah is not None else metric1
        rc = ref_lib().ref_advection_uv(
            int(direction), int(scheme), nx, ny, nk, halo, halo, _i(uvmask), _d(metric1), _d(metric2),
            _d(huv), _d(uv), _i(tmask), _d(iarea), int(ah is not None), _d(dummy), ctypes.c_double(dt),
            _d(h), _d(var))
    else:
        rc = lib().oracle_advection_uv(
            int(direction), int(scheme), nx, ny, nk, halo, halo, _i(uvmask), _d(metric1), _d(metric2),
            _d(huv), _d(uv), _i(tmask), _d(iarea), _d(ah), ctypes.c_double(dt), _d(h), _d(var))
    _check(rc, "advection_uv")


def advection_w(scheme, w, w_var, tmask, dt, h, var, halo=2, impl="c") -> int:
    """In-place vertical sweep; returns the largest sub-iteration count (C impl only)."""
    nx, ny, nz = _dims(var, halo)
    if impl == "ref":
        rc = ref_lib().ref_advection_w(int(scheme), nx, ny, nz, halo, halo, _d(w), _d(w_var), _i(tmask),
                                       ctypes.c_double(dt), _d(h), _d(var))
        _check(rc, "ref_advection_w")
        return -1
    mi = ctypes.c_int(0)
    rc = lib().oracle_advection_w(int(scheme), nx, ny, nz, halo, halo, _d(w), _d(w_var), _i(tmask),
                                  ctypes.c_double(dt), _d(h), _d(var), ctypes.byref(mi))
    _check(rc, "advection_w")
    return mi.value


def vdiff_prepare(nuh, dt, cnpar, dzo, dzn, molecular=0.0, halo=2):
    nx, ny, nz = _dims(dzn, halo)
    auxo = np.zeros_like(dzn)
    auxn = np.zeros_like(dzn)
    rc = lib().oracle_vdiff_prepare(nx, ny, nz, halo, halo, ctypes.c_double(molecular), _d(nuh),
                                    ctypes.c_double(dt), ctypes.c_double(cnpar), _d(dzo), _d(dzn),
                                    _d(auxo), _d(auxn))
    _check(rc, "vdiff_prepare")
    return auxo, auxn


def vdiff_apply(mask, dzo, dzn, auxo, auxn, var, ea2=None, ea4=None, halo=2):
    nx, ny, nz = _dims(var, halo)
    rc = lib().oracle_vdiff_apply(nx, ny, nz, halo, halo, _i(mask), _d(dzo), _d(dzn), _d(auxo),
                                  _d(auxn), _d(var), _d(ea2), _d(ea4))
    _check(rc, "vdiff_apply")


def vertical_diffusion(nuh, dt, var, mask, hn, ho=None, cnpar=1.0, molecular=0.0, ea2=None, ea4=None,
                       halo=2):
    """VerticalDiffusion.__call__ (operators.py:218-230): prepare + apply, in place."""
    dzo = hn if ho is None else ho
    auxo, auxn = vdiff_prepare(nuh, dt, cnpar, dzo, hn, molecular, halo)
    vdiff_apply(mask, dzo, hn, auxo, auxn, var, ea2, ea4, halo)


def w_momentum_3d(mask, dyu, dxv, iarea, ho, hn, pk, qk, dt, w, halo=2, extra=1):
    """In-place w (nz+1, ny_, nx_) from layer transports (momentum.F90:191-216).
    Points: the interior plus ``extra`` columns/rows on the high side, as
    _pygetm.pyx:454 passes (i in [halo, halo + nx], j likewise)."""
    nz, ny_, nx_ = hn.shape
    nx, ny = nx_ - 2 * halo, ny_ - 2 * halo
    rc = lib().oracle_w_momentum_3d(nx_, ny_, nz, halo, halo + nx - 1 + extra, halo, halo + ny - 1 + extra,
                                    _i(mask), _d(dyu), _d(dxv), _d(iarea), _d(ho), _d(hn), _d(pk), _d(qk),
                                    ctypes.c_double(dt), _d(w))
    _check(rc, "w_momentum_3d")


def vertical_advection_to_sources(halox, haloy, mask, c, w, h, s):
    """In-place s (nz, ny_, nx_): upstream sinking/floating into sources
    (wrapper.F90:447-486), with the reference's argument order (_pygetm.pyx:574-575)."""
    nz, ny_, nx_ = c.shape
    for a in (mask, c, w, h, s):
        assert a.shape == (nz, ny_, nx_) and a.flags["C_CONTIGUOUS"]
    rc = lib().oracle_vertical_advection_to_sources(nx_, ny_, nz, int(halox), int(haloy), _i(mask), _d(c), _d(w),
                                                    _d(h), _d(s))
    _check(rc, "vertical_advection_to_sources")


def elevation2depth(z, H, Dmin, mask, D):
    """In-place D = max(H + z, Dmin) where mask != 0 (wrapper.F90 c_elevation2depth)."""
    n = z.size
    for a in (z, H, mask, D):
        assert a.size == n and a.flags["C_CONTIGUOUS"]
    _check(lib().oracle_elevation2depth(n, _d(z), _d(H), ctypes.c_double(Dmin), _i(mask), _d(D)), "elevation2depth")


def thickness2vertical_coordinates(mask, H, h, zc, zf):
    """In-place zc (nz, ny_, nx_) and zf (nz+1, ny_, nx_) (wrapper.F90 c_thickness2vertical_coordinates)."""
    nz, ny_, nx_ = h.shape
    assert zc.shape == h.shape and zf.shape == (nz + 1, ny_, nx_) and mask.shape == H.shape == (ny_, nx_)
    _check(lib().oracle_thickness2vertical_coordinates(nx_, ny_, nz, _i(mask), _d(H), _d(h), _d(zc), _d(zf)),
           "thickness2vertical_coordinates")


def alpha(D, Dmin, Dcrit, mask, out):
    """In-place out = max(0, min(1, (D - Dmin) / (Dcrit - Dmin))) where mask != 0 (wrapper.F90 c_alpha)."""
    n = D.size
    for a in (D, mask, out):
        assert a.size == n and a.flags["C_CONTIGUOUS"]
    _check(lib().oracle_alpha(n, _d(D), ctypes.c_double(Dmin), ctypes.c_double(Dcrit), _i(mask), _d(out)), "alpha")


def update_gvc(dsigma, dbeta, Dgamma, kk, D, mask, h):
    """In-place h (nz, ny, nx) of generalized vertical coordinates (wrapper.F90 c_update_gvc);
    kk is the reference layer as _pygetm.update_gvc takes it (negative counts from the top)."""
    nz = dbeta.size
    if kk < 0:
        kk = nz + kk
    ny, nx = D.shape
    assert h.shape == (nz, ny, nx) and mask.shape == D.shape
    _check(lib().oracle_update_gvc(nx, ny, nz, ctypes.c_double(dsigma), _d(np.ascontiguousarray(dbeta)),
                                   ctypes.c_double(Dgamma), 1 + kk, _d(D), _i(mask), _d(h)), "update_gvc")


def update_sigma(dga, D, mask, h):
    nz = dga.size
    ny, nx = D.shape
    assert h.shape == (nz, ny, nx) and mask.shape == D.shape
    _check(lib().oracle_update_sigma(nx, ny, nz, _d(np.ascontiguousarray(dga)), _d(D), _i(mask), _d(h)),
           "update_sigma")


def _interp_args(source, target):
    assert source.dtype == target.dtype == np.float64 and source.flags["C_CONTIGUOUS"] and target.flags["C_CONTIGUOUS"]
    assert source.ndim == target.ndim == 3


def interp_x(source, target, offset, impl="c"):
    """In-place target[..., i + offset] = 0.5 (source[..., i] + source[..., i + 1])
    (interpolation.F90 grid_interp_x); impl="ref" runs the reference Fortran."""
    _interp_args(source, target)
    assert source.shape == target.shape
    nz, ny, nx = source.shape
    if impl == "ref":
        ref_lib().grid_interp_x(nx, ny, nz, _d(source), _d(target), int(offset))
        return
    _check(lib().oracle_interp_x(nx, ny, nz, _d(source), _d(target), int(offset)), "interp_x")


def interp_y(source, target, offset, impl="c"):
    _interp_args(source, target)
    assert source.shape == target.shape
    nz, ny, nx = source.shape
    if impl == "ref":
        ref_lib().grid_interp_y(nx, ny, nz, _d(source), _d(target), int(offset))
        return
    _check(lib().oracle_interp_y(nx, ny, nz, _d(source), _d(target), int(offset)), "interp_y")


def interp_z(source, target, offset, impl="c"):
    _interp_args(source, target)
    nz1, ny, nx = source.shape
    assert target.shape[1:] == (ny, nx)
    if impl == "ref":
        ref_lib().grid_interp_z(nx, ny, nz1, target.shape[0], _d(source), _d(target), int(offset))
        return
    _check(lib().oracle_interp_z(nx, ny, nz1, target.shape[0], _d(source), _d(target), int(offset)), "interp_z")


def interp_xy(source, target, ioffset, joffset, impl="c"):
    _interp_args(source, target)
    nz, ny1, nx1 = source.shape
    assert target.shape[0] == nz
    ny2, nx2 = target.shape[1:]
    if impl == "ref":
        ref_lib().grid_interp_xy(nx1, ny1, nx2, ny2, nz, _d(source), _d(target), int(ioffset), int(joffset))
        return
    _check(lib().oracle_interp_xy(nx1, ny1, nx2, ny2, nz, _d(source), _d(target), int(ioffset), int(joffset)),
           "interp_xy")


def reconstruct_transport_change(un, hn, Uo, timestep, jmin, jmax):
    """In-place un = (un hn - Uo) / timestep over Fortran rows jmin..jmax (momentum.F90:323-340)."""
    nz, ny, nx = un.shape
    assert hn.shape == Uo.shape == un.shape
    _check(lib().oracle_reconstruct_transport_change(nx, ny, nz, int(jmin), int(jmax), _d(un), _d(hn), _d(Uo),
                                                     ctypes.c_double(timestep)), "reconstruct_transport_change")


def transport_step(d, tracers: np.ndarray, scheme=None, dt=None, cnpar=None, molecular=0.0,
                   ah_u=None, ah_v=None, return_h=False):
    """One adv+diff step on stacked FP64 tracers (T, nz, ny_, nx_), in place.

    ``d`` is a getm_amd.inputs.Domain (numpy fields)."""
    scheme = d.scheme if scheme is None else scheme
    dt = d.dt if dt is None else dt
    cnpar = d.cnpar if cnpar is None else cnpar
    assert tracers.dtype == np.float64 and tracers.flags["C_CONTIGUOUS"]
    assert tracers.shape[1:] == d.ho.shape
    h_out = np.empty_like(d.ho) if return_h else None
    rc = lib().oracle_transport_step(
        int(scheme), d.nx, d.ny, d.nz, d.halo, d.halo, _i(d.tmask), _i(d.umask), _i(d.vmask),
        _d(d.idxu), _d(d.dyu), _d(d.dxv), _d(d.idyv), _d(d.iarea), _d(d.ho), _d(d.hn), _d(d.hu),
        _d(d.hv), _d(d.u), _d(d.v), _d(d.w), _d(d.nuh), _d(ah_u), _d(ah_v), ctypes.c_double(dt),
        ctypes.c_double(cnpar), ctypes.c_double(molecular), int(tracers.shape[0]), _d(tracers), _d(h_out))
    _check(rc, "transport_step")
    return h_out
