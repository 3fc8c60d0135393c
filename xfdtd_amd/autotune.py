"""Tile-shape autotuner for the curl kernels.

For each grid-size class it times every candidate launch geometry on a
representative scene of that class (a vacuum cube with a dielectric sphere,
TFSF and CPML, the class's middle edge length) with the engine's own HIP-event
timer, and keeps the fastest.  All candidates use 256-thread workgroups, so
instances tuned into different classes can share one batched launch.

The result maps a class (lo, hi) to (tz, ty, tx, flags) for
`Simulation.set_tiling`; `tiling_for(table, n)` picks an instance's entry.
"""
from __future__ import annotations

import json
import os

from . import scene as S

SIZE_CLASSES = ((0, 80), (81, 120), (121, 100000))
CANDIDATES = ((64, 4, 8, 0), (64, 4, 4, 0), (64, 4, 16, 0), (128, 2, 8, 0), (256, 0, 8, 0),
              (256, 0, 4, 0), (64, 4, 8, 16))


def _representative(n, dl=5e-3, pml=8):
    objs = [S.Object("s", S.Sphere((0.0, 0.0, 0.0), 0.2 * n * dl), S.Material("d", 3.0, 1.0, 0.0, 0.0))]
    return S.cube_domain_scene(n, dl, pml, cfl=0.9, objects=objs, tfsf_gap=pml + 4)


def tune(precision=32, classes=SIZE_CLASSES, candidates=CANDIDATES, steps=12, warmup=3, log=None):
    """Measure every candidate on each class; return {(lo, hi): tiling}."""
    from .sim import Simulation
    table = {}
    for lo, hi in classes:
        n = (max(lo, 48) + min(hi, 160)) // 2
        best = None
        for cand in candidates:
            sim = Simulation(_representative(n), warmup + steps, precision=precision)
            sim.set_tiling(*cand)
            sim.step(warmup)
            sim.set_timing(True)
            sim.step(steps)
            ms_e, ms_h, _ = sim.kernel_times()
            sim.close()
            t = ms_e + ms_h
            if log:
                log(f"autotune n={n} tiling={cand}: {t * 1e3:.1f} us per step (E+H)")
            if best is None or t < best[0]:
                best = (t, cand)
        table[(lo, hi)] = best[1]
    return table


BATCH_CANDIDATES = ((64, 4, 8, 0), (256, 0, 8, 0), (64, 4, 16, 0), (256, 0, 16, 0), (64, 2, 16, 0))


def tune_batch(sims, candidates=BATCH_CANDIDATES, steps=40, warmup=5, log=None):
    """Time each candidate tiling on the batched launch itself: every member
    of `sims` (set-up Simulations of one precision, used up by the timing
    runs: they advance by warm-up + steps per candidate) takes the
    candidate, one Batch steps them all.  A batch fills the GPU with any
    tile size, so its best tiling can differ from the best plain-run tiling
    of each member's size.  Returns the fastest (tz, ty, tx, flags)."""
    import time

    from .sim import Batch
    best = None
    for cand in candidates:
        for s in sims:
            s.set_tiling(*cand)
        b = Batch(sims)
        b.step(warmup)
        b.sync()
        t0 = time.perf_counter()
        b.step(steps)
        b.sync()
        t = (time.perf_counter() - t0) / steps
        b.close()
        if log:
            log(f"autotune batch of {len(sims)} tiling={cand}: {t * 1e3:.3f} ms per step")
        if best is None or t < best[0]:
            best = (t, cand)
    return best[1]


def tiling_for(table, n):
    for (lo, hi), til in table.items():
        if lo <= n <= hi:
            return til
    raise KeyError(f"no size class holds n = {n}")


def save(table, path):
    with open(path, "w") as f:
        json.dump([{"lo": lo, "hi": hi, "tiling": list(t)} for (lo, hi), t in table.items()], f, indent=1)


def load(path):
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return {(e["lo"], e["hi"]): tuple(e["tiling"]) for e in json.load(f)}
