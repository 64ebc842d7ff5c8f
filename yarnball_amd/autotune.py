"""Launch-shape autotuner.

Picks the threads per block of the Newton sweep (k_sweep) and of the contact
query (k_query) by timing each candidate, with HIP events on the stepper's
stream (Sim.profile), on a scratch handle loaded with the caller's workload.
The caller's simulation is never touched: apply the result with
Sim.setLaunchConfig.  Every kernel's output per vertex or segment is a
function of its inputs only, so the choice changes the speed and never the
results (checked by tests/test_gpu_launch_modes.py).
"""
import numpy as np

from . import scenes
from .sim import Sim

SWEEP_THREADS = (64, 128, 256, 512)
QUERY_THREADS = (64, 128, 256)


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
    sim.close()
    out = {"sweep_threads": best_sweep, "query_threads": best_query,
           "sweep_us": {str(k): round(v, 2) for k, v in sweep.items()},
           "query_us": {str(k): round(v, 2) for k, v in query.items()}}
    if log:
        log(f"[autotune] sweep {out['sweep_us']} -> {best_sweep}; query {out['query_us']} -> "
            f"{best_query}")
    return out
