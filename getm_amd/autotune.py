"""Launch-shape and solver autotuner for the batched step and ensembles.

Searches the height of the horizontal-stage tiles (4, 8 or 16 rows; for
ensembles also their width, 32, 64 or 128 columns), the tracers per block of
the three kernel kinds, and the column solver by depth: per-lane Thomas elimination (the reference's algorithm,
diffusion.F90:276-313) or wave-level parallel cyclic reduction for columns of
up to ``pcr_max_nz`` levels.  Each candidate is timed with HIP events on the
current stream, and the best is cached per workload signature (grid size
bucket, nz, tracers) in a JSON file.  Launch shapes never change results; the
two solvers agree to FP32 rounding (see tests).  The timing steps run on the
caller's tracers, which are restored afterwards with their rounding step
counters: a tuned collection or ensemble continues from the state it had
before tuning, whether the shape came from the cache or from timing.
"""
from __future__ import annotations

import itertools
import json
import math
import os
from typing import Callable, Dict, Iterable, Optional, Tuple

import torch

Shape = Tuple[int, ...]  # (tile_y, tracers/block h, w, d[, pcr_max_nz[, tile_x]])

DEFAULT_CANDIDATES = [
    (ty, h, w, d, pcr)
    for ty, h, (w, d), pcr in itertools.product((16, 8, 4), (0, 32), ((16, 16), (64, 64)), (0, 16, 32))
]

# ensembles: tile width x height (the shapes ensemble_set_tile_x accepts), the
# column solver by depth, and two tracer batchings
ENSEMBLE_TILES = ((32, 8), (32, 16), (64, 4), (64, 8), (64, 16), (128, 4), (128, 8))
ENSEMBLE_CANDIDATES = [
    (ty, h, w, d, pcr, tx)
    for (tx, ty), (h, w, d), pcr in itertools.product(ENSEMBLE_TILES, ((0, 16, 16), (32, 64, 64)), (0, 32))
]


def signature(ncells: int, nz: int, ntracers: int) -> str:
    """Bucketed workload key: log2 of the column count, nz rounded up to 8, log2 of T."""
    cols = max(1, ncells // max(nz, 1))
    return f"c{int(math.log2(cols))}-z{8 * math.ceil(nz / 8)}-t{int(math.log2(max(ntracers, 1)))}"


def time_step(step: Callable[[], None], reps: int = 3) -> float:
    """Mean milliseconds of ``step`` over ``reps`` calls (after one warm-up call)."""
    step()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(reps):
        step()
    stop.record()
    stop.synchronize()
    return start.elapsed_time(stop) / reps


class Autotuner:
    def __init__(self, cache_path: Optional[str] = None):
        self.cache_path = cache_path
        self.cache: Dict[str, Shape] = {}
        if cache_path and os.path.exists(cache_path):
            with open(cache_path) as fh:
                self.cache = {k: tuple(v) for k, v in json.load(fh).items()}

    def _save(self):
        if self.cache_path:
            tmp = self.cache_path + ".tmp"
            with open(tmp, "w") as fh:
                json.dump(self.cache, fh, indent=1)
            os.replace(tmp, self.cache_path)

    def tune(self, key: str, apply: Callable[[Shape], None], step: Callable[[], None],
             candidates: Iterable[Shape] = DEFAULT_CANDIDATES, reps: int = 3, force: bool = False) -> Shape:
        """Apply the cached shape for ``key`` or time every candidate and keep the fastest.
        ``apply(shape)`` must make the next ``step()`` use the shape."""
        if key in self.cache and not force:
            apply(self.cache[key])
            return self.cache[key]
        best, best_ms = None, float("inf")
        self.timings = {}
        for cand in candidates:
            apply(cand)
            ms = time_step(step, reps)
            self.timings[cand] = ms
            if ms < best_ms:
                best, best_ms = cand, ms
        apply(best)
        self.cache[key] = best
        self._save()
        return best

    def tune_collection(self, tc, advance: Callable[[], None], **kw) -> Shape:
        """Tune a batched getm_amd.tracer.TracerCollection; ``advance`` runs one step.
        The collection's state is restored afterwards."""
        g = tc.grid
        key = "step-" + tc.precision + "-" + signature(g.nx_ * g.ny_ * g.nz_, g.nz_, len(tc))
        def apply(s):
            tc.set_tuning(s[1], s[2], s[3], tile_y=s[0])
            tc.set_solver(s[4] if len(s) > 4 else 0)

        snap = tc.snapshot()
        try:
            return self.tune(key, apply, advance, **kw)
        finally:
            tc.restore(snap)

    def tune_ensemble(self, ens, **kw) -> Shape:
        """Tune a getm_amd.ensemble.Ensemble (rebuilds its work tables per candidate).
        The ensemble's state is restored afterwards."""
        ncells = sum(g.nx_ * g.ny_ * g.nz_ for g in ens.grids)
        key = "ens-" + signature(ncells, max(g.nz_ for g in ens.grids), sum(ens.ntracers)) + f"-d{len(ens)}"

        def apply(s):
            ens.set_tuning(*s[:4])
            ens.set_solver(s[4] if len(s) > 4 else 0)
            ens.set_tile_x(s[5] if len(s) > 5 else 64)
            ens.commit()

        kw.setdefault("candidates", ENSEMBLE_CANDIDATES)
        snap = ens.snapshot()
        try:
            return self.tune(key, apply, ens.step, **kw)
        finally:
            ens.restore(snap)
