"""Write golden vectors of the reference's own grid interpolations.

Runs grid_interp_x/_y/_z/_xy of python/src/interpolation.F90 (compiled from
/root/reference into oracle/_ref/libgetm_ref.so by oracle/Makefile) on small
seeded fields and stores inputs and outputs in ``interp_golden.npz``: for each
routine and offset the source, the target before (a sentinel everywhere, so the
points the routine leaves alone are checked too) and the target after.

Usage: python tests/golden/make_interp_golden.py   (needs oracle/_ref built: make -C oracle ref)
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

SENTINEL = -7.5e300

# name -> (routine, source shape, target shape, offsets)
CASES = {
    "x0": ("x", (4, 6, 9), (4, 6, 9), (0,)),
    "x1": ("x", (4, 6, 9), (4, 6, 9), (1,)),
    "y0": ("y", (3, 7, 5), (3, 7, 5), (0,)),
    "y1": ("y", (3, 7, 5), (3, 7, 5), (1,)),
    "z0": ("z", (6, 5, 4), (5, 5, 4), (0,)),  # interfaces -> centres
    "z1": ("z", (5, 5, 4), (6, 5, 4), (1,)),  # centres -> interfaces
    "xy00": ("xy", (3, 6, 8), (3, 6, 8), (0, 0)),
    "xy11": ("xy", (3, 6, 8), (3, 6, 8), (1, 1)),
    "xy01": ("xy", (3, 6, 8), (3, 6, 8), (0, 1)),
    "xy10": ("xy", (3, 6, 8), (3, 6, 8), (1, 0)),
    "xy_small_target": ("xy", (2, 6, 8), (2, 5, 7), (0, 0)),
}


def run(kind, src, dst, offs, impl):
    if kind == "x":
        oracle.interp_x(src, dst, offs[0], impl=impl)
    elif kind == "y":
        oracle.interp_y(src, dst, offs[0], impl=impl)
    elif kind == "z":
        oracle.interp_z(src, dst, offs[0], impl=impl)
    else:
        oracle.interp_xy(src, dst, offs[0], offs[1], impl=impl)


def main():
    if not oracle.ref_available():
        sys.exit("oracle/_ref/libgetm_ref.so is not built (make -C oracle ref)")
    rng = np.random.default_rng(21)
    out = {}
    for name, (kind, sshape, tshape, offs) in CASES.items():
        src = rng.uniform(-3.0, 3.0, sshape) * 10.0 ** rng.integers(-3, 4, sshape)
        before = np.full(tshape, SENTINEL)
        after = before.copy()
        run(kind, src, after, offs, "ref")
        out[f"{name}/source"] = src
        out[f"{name}/before"] = before
        out[f"{name}/after"] = after
        out[f"{name}/offsets"] = np.array(offs, dtype=np.int32)
        out[f"{name}/kind"] = np.array(kind)
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "interp_golden.npz"), **out)
    print(f"wrote {len(CASES)} cases")


if __name__ == "__main__":
    main()
