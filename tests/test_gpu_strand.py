"""Strand solve (a warm-start option with no counterpart in the reference):
the step solves the strand system it assembles -- cyclic reduction inside each
62-vertex tile plus the boundary-pair iterations against an exact banded
solve of the same rows -- and a schedule of strand steps reaches the plain
solver's residual in fewer sweeps with no penetration."""
import ctypes

import numpy as np
import pytest
from scipy.linalg import solve_banded

from yarnball_amd import Sim, learned, scenes
from yarnball_amd._lib import check, ptr

FLAG_HAS_PREV, FLAG_HAS_NEXT = 1, 2


def exact_strand_solve(rows):
    """FP64 solution of a_i x_i - c_(i-1) x_(i-1) - c_i x_(i+1) = r_i over the
    whole index range (c_i = 0 where a strand ends or a vertex is pinned)."""
    a = rows[:, 0].astype(np.float64)
    c = rows[:, 1].astype(np.float64)
    n = len(a)
    ab = np.zeros((3, n))
    ab[0, 1:] = -c[:-1]
    ab[1] = a
    ab[2, :-1] = -c[:-1]
    return solve_banded((1, 1), ab, rows[:, 2:5].astype(np.float64))


def step_once(sim, iters):
    """dx before and after one strand step on the prologue's warm start, and
    the system rows the step assembled."""
    n = sim.numVerts
    q = sim.meta.copy()
    check(sim._lib.yb_debug_prologue(sim.handle, ctypes.byref(q), np.float32(sim.maxH), 1, 1))
    learned.strand_schedule(sim, every=0, pair_iters=iters)
    before = np.zeros((n, 4), np.float32)
    check(sim._lib.yb_debug_read(sim.handle, ptr(before), None, None, None, None, None))
    rows = np.zeros((n, 5), np.float32)
    check(sim._lib.yb_debug_strand(sim.handle, ptr(rows)))
    after = np.zeros((n, 4), np.float32)
    check(sim._lib.yb_debug_read(sim.handle, ptr(after), None, None, None, None, None))
    return before, after, rows


@pytest.mark.gpu
def test_strand_step_solves_its_system():
    sc = scenes.knit_patch(rows=16, per=1024)
    sim = Sim(sc.n)
    scenes.load_into(sim, sc)
    sim.advance(np.float32(3 * sc.maxH))  # motion and contacts
    sim.download()
    fl = sim.verts["flags"].astype(np.int64)
    free = sim.verts["invMass"] != 0
    errs = {}
    for iters in (1, 2, 8):
        before, after, rows = step_once(sim, iters)
        got = (after - before)[:, :3].astype(np.float64)
        want = exact_strand_solve(rows)
        # rows the step limit clamped are left out (a sweep clamps the same way)
        w = before[:, 3].astype(np.float64)
        lim = np.where((fl & FLAG_HAS_NEXT) != 0, w, np.inf)
        lim = np.where((fl & FLAG_HAS_PREV) != 0, np.minimum(lim, np.roll(w, 1)), lim)
        unclamped = np.linalg.norm(after[:, :3], axis=1) < lim * (1 - 1e-5)
        m = free & unclamped
        assert m.sum() > 0.9 * free.sum()
        errs[iters] = np.max(np.abs(got[m] - want[m])) / np.max(np.abs(want[m]))
        assert np.array_equal(after[~free], before[~free]), "pinned vertices must keep dx"
        # the pinned rows are identity rows with no coupling
        assert np.all(rows[~free, 0] == 1) and np.all(rows[~free, 2:] == 0)
        # The rows are diagonally dominant, which the unpivoted cyclic
        # reduction relies on: the diagonal (the local Hessian's curvature
        # along the force) holds the stretch stiffness of both segments.
        c = rows[:, 1].astype(np.float64)
        off = c + np.concatenate([[0.0], c[:-1]])
        a = rows[:, 0].astype(np.float64)
        margin = (a[m] - off[m]) / a[m]
        print(f"pair iterations {iters}: min diagonal-dominance margin {margin.min():.3e}")
        assert margin.min() > -1e-6
    print("strand step vs exact banded solve, relative max error by pair iterations:", errs)
    # two pair iterations already reach the FP32 accuracy of the tile solves
    assert errs[2] < 1e-5 and errs[8] < 1e-5
    assert errs[1] < 1e-3
    sim.close()


def _run(sc, frames, strand_every=None):
    sim = Sim(sc.n)
    scenes.load_into(sim, sc)
    sim.setSolver(1e-6, maxSweeps=6000, chunk=16, checkEvery=4)
    if strand_every is not None:
        learned.strand(sim, every=strand_every)
    sim.minSeparation(reset=True)
    sim.solverStats(reset=True)
    sim.advance(np.float32(frames * sc.frame_dt))
    sim.download()
    out = dict(sim.solverStats())
    out["minSep"] = sim.minSeparation()
    out["pos"] = sim.verts["pos"].astype(np.float64).copy()
    sim.close()
    return out


@pytest.mark.gpu
def test_strand_schedule_reaches_the_residual_in_fewer_sweeps():
    sc = scenes.knit_patch(rows=32, per=1024, seed=2)
    plain = _run(sc, 2)
    strand = _run(sc, 2, strand_every=8)
    for name, r in (("plain", plain), ("strand", strand)):
        print(name, {k: r[k] for k in ("sweeps", "substeps", "notConverged", "maxResidual",
                                       "maxSweepsUsed", "minSep")})
    for r in (plain, strand):
        assert r["notConverged"] == 0 and r["maxResidual"] <= 1e-6
        assert r["minSep"] > 1.0  # no penetration
    sp, ss = plain["sweeps"] / plain["substeps"], strand["sweeps"] / strand["substeps"]
    seg = float(np.median(sc.verts["lRest"][sc.verts["lRest"] > 0]))
    d = np.linalg.norm(plain["pos"] - strand["pos"], axis=1) / seg
    rms = float(np.sqrt(np.mean(d ** 2)))
    print(f"sweeps per substep: plain {sp:.1f}, strand every 8 {ss:.1f}; position difference "
          f"after {plain['substeps']} substeps: rms {rms:.2e}, max {d.max():.2e} segment lengths")
    assert ss * 3 < sp
    # Both stop at the same per-sweep update size, but a block-Jacobi sweep
    # barely moves the smooth modes along a stiff strand, so the plain solver
    # stops with more of them in place; over many substeps the two motions
    # drift apart by a fraction of a segment.
    assert rms < 0.05 and d.max() < 0.3
