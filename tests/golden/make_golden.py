"""Write golden vectors of the reference's own advection Fortran.

Runs the x/y/z sweeps of src/operators/advection_schemes/ (built from
/root/reference into oracle/_ref/libgetm_ref.so by oracle/Makefile) on small
seeded inputs and stores inputs and outputs in ``advection_golden.npz``.  The
fixture lets the C restatement and the GPU path be checked against reference
outputs wherever the reference build is absent.

Cases (all 6 schemes each):
  gentle  12x10x5, horizontal CFL <= 0.1, vertical CFL up to ~1 (1-2 sub-steps)
  violent 10x8x5,  horizontal CFL <= 0.9, vertical CFL up to ~15 (restarts, some at the cap)
  codes   12x10x5, open-boundary mask codes (U/V = 2, 3, 4; T = 2) and Ah horizontal diffusion
Per case and scheme the stored outputs are h and the tracer after
v dt/2 -> u dt/2 -> w dt -> u dt/2 -> v dt/2 (operators.py:124-133 with ufirst=False),
and after the w sweep alone.

Usage: python tests/golden/make_golden.py   (needs oracle/_ref built: make -C oracle ref)
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from getm_amd import inputs  # noqa: E402

CASES = {
    "gentle": dict(nx=12, ny=10, nz=5, seed=11, cfl=0.1, wscale=1.0),
    "violent": dict(nx=10, ny=8, nz=5, seed=12, cfl=0.9, wscale=0.0),
    "codes": dict(nx=12, ny=10, nz=5, seed=13, cfl=0.3, wscale=1.5, mask_codes=True),
}
INPUT_FIELDS = ("tmask", "umask", "vmask", "idxu", "dyu", "dxv", "idyv", "iarea", "ho", "hu", "hv", "u", "v", "w")


def run_sequence(d, scheme, ah, impl):
    h = d.ho.copy()
    f = d.tracers[0].copy()
    half = 0.5 * d.dt
    out = {}
    oracle.advection_uv(2, scheme, d.vmask, d.dxv, d.idyv, d.hv, d.v, d.tmask, d.iarea, ah, half, h, f, impl=impl)
    oracle.advection_uv(1, scheme, d.umask, d.idxu, d.dyu, d.hu, d.u, d.tmask, d.iarea, ah, half, h, f, impl=impl)
    oracle.advection_w(scheme, d.w, d.w, d.tmask, d.dt, h, f, impl=impl)
    oracle.advection_uv(1, scheme, d.umask, d.idxu, d.dyu, d.hu, d.u, d.tmask, d.iarea, ah, half, h, f, impl=impl)
    oracle.advection_uv(2, scheme, d.vmask, d.dxv, d.idyv, d.hv, d.v, d.tmask, d.iarea, ah, half, h, f, impl=impl)
    out["f"], out["h"] = f, h
    hw, fw = d.ho.copy(), d.tracers[0].copy()
    oracle.advection_w(scheme, d.w, d.w, d.tmask, d.dt, hw, fw, impl=impl)
    out["f_w"], out["h_w"] = fw, hw
    return out


def case_domain(name):
    d = inputs.tiny(ntracers=1, **CASES[name])
    ah = np.full_like(d.idxu, 40.0) if name == "codes" else None
    return d, ah


def main():
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref/libgetm_ref.so missing: run `make -C oracle ref` first")
    data = {}
    for name in CASES:
        d, ah = case_domain(name)
        for f in INPUT_FIELDS:
            data[f"{name}/in/{f}"] = getattr(d, f)
        data[f"{name}/in/tracer"] = d.tracers[0]
        data[f"{name}/in/dt"] = np.float64(d.dt)
        if ah is not None:
            data[f"{name}/in/ah"] = ah
        for scheme in range(1, 7):
            for k, v in run_sequence(d, scheme, ah, "ref").items():
                data[f"{name}/s{scheme}/{k}"] = v
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "advection_golden.npz")
    np.savez_compressed(path, **data)
    print(f"wrote {path} ({os.path.getsize(path)} bytes, {len(data)} arrays)")


if __name__ == "__main__":
    main()
