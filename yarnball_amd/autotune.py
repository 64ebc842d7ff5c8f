"""Launch-shape and tile autotuner.

Picks, by timing each candidate with HIP events on the stepper's stream
(Sim.profile) on a scratch handle loaded with the caller's workload:
  - the threads per block of the Newton sweep (k_sweep) and of the contact
    query (k_query);
  - the refit tile: sorted leaves per refit workgroup (256, 512, 1024);
  - the query's LDS shape: traversal-stack entries kept in LDS and candidate
    leaves buffered per thread, which together set the waves a CU can hold.
The caller's simulation is never touched: apply the result with
Sim.setLaunchConfig and Sim.setTileConfig.  The candidates are timed with
the contact skin off, so that every substep walks the tree.  Every kernel's output per vertex
or segment is a function of its inputs only, refit boxes are min/max unions
and contact lists keep their order, so the choice changes the speed and never
the results (checked by tests/test_gpu_launch_modes.py).
"""
import numpy as np

from . import scenes
from .sim import Sim

SWEEP_THREADS = (64, 128, 256, 512)
QUERY_THREADS = (64, 128, 256)
REFIT_LEAVES = (256, 512, 1024)
QUERY_SHAPES = ((4, 4), (8, 6), (8, 8), (16, 8))


def _stage_us(prof, stage):
    ms, n = prof[stage]
    return ms / n * 1e3 if n else float("inf")


def tune(scene, offsets=None, residual=0.0, max_sweeps=6000, substeps=2, rounds=2, log=None):
    """Time every candidate block size on `scene` (a configured HostScene, or
    a batch with its scene offsets) and return the fastest:
    {"sweep_threads", "query_threads", "sweep_us": {...}, "query_us": {...}}."""
    sim = Sim(scene.n)
    sim.printErrors = False
    scenes.load_into(sim, scene)
    # the tree is refit and walked in every timed substep (the contact skin,
    # which reuses the candidates of a walk, would skip both in most of them)
    sim.setContactSkin(0.0)
    if offsets is not None:
        sim.setScenes(offsets)
    if residual > 0:
        sim.setSolver(residual, maxSweeps=max_sweeps, chunk=16, checkEvery=4)
    sim.advance(np.float32(2 * scene.maxH))  # first tree build, warm caches
    sweep = {t: float("inf") for t in SWEEP_THREADS}
    query = {t: float("inf") for t in QUERY_THREADS}
    for _ in range(rounds):
        for t in SWEEP_THREADS:
            sim.setLaunchConfig(t, 128)
            sweep[t] = min(sweep[t], _stage_us(sim.profile(substeps), "sweep"))
    best_sweep = min(sweep, key=sweep.get)
    for _ in range(rounds):
        for t in QUERY_THREADS:
            sim.setLaunchConfig(best_sweep, t)
            query[t] = min(query[t], _stage_us(sim.profile(substeps), "query"))
    best_query = min(query, key=query.get)
    sim.setLaunchConfig(best_sweep, best_query)
    # refit tile: the first substep after a change rebuilds the tree, so each
    # candidate is timed over substeps + 1 and the first one dropped
    refit = {r: float("inf") for r in REFIT_LEAVES}
    for _ in range(rounds):
        for r in REFIT_LEAVES:
            sim.setTileConfig(r, 8, 6)
            sim.profile(1)
            refit[r] = min(refit[r], _stage_us(sim.profile(substeps), "refit"))
    best_refit = min(refit, key=refit.get)
    sim.setTileConfig(best_refit, 8, 6)
    sim.profile(1)
    shape = {c: float("inf") for c in QUERY_SHAPES}
    for _ in range(rounds):
        for c in QUERY_SHAPES:
            sim.setTileConfig(best_refit, *c)
            shape[c] = min(shape[c], _stage_us(sim.profile(substeps), "query"))
    best_shape = min(shape, key=shape.get)
    sim.close()
    out = {"sweep_threads": best_sweep, "query_threads": best_query,
           "refit_leaves": best_refit, "query_shape": list(best_shape),
           "sweep_us": {str(k): round(v, 2) for k, v in sweep.items()},
           "query_us": {str(k): round(v, 2) for k, v in query.items()},
           "refit_us": {str(k): round(v, 2) for k, v in refit.items()},
           "query_shape_us": {f"{a}x{b}": round(v, 2) for (a, b), v in shape.items()}}
    if log:
        log(f"[autotune] sweep {out['sweep_us']} -> {best_sweep}; query {out['query_us']} -> "
            f"{best_query}; refit tile {out['refit_us']} -> {best_refit}; query shape "
            f"{out['query_shape_us']} -> {best_shape}")
    return out
