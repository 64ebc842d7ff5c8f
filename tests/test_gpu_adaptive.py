"""Residual-targeted mode: the device sweeps until the relative Newton
residual RMS |dx_k+1 - dx_k| / rest length <= tol; the oracle does the same in FP64
from the same state.  Both converged iterates must agree (positions
normwise <= 1e-5, frames <= 1e-5) and the sweep counts must be close."""
import numpy as np
import pytest

from oracle import Oracle
from yarnball_amd import scenes

from helpers import gpu_sim, gpu_state, rel_err


@pytest.mark.gpu
@pytest.mark.parametrize("name,make,tol", [
    ("readme", scenes.readme_scene, 1e-6),
    ("knit_8x256", lambda: scenes.knit_patch(rows=8, per=256), 1e-5),
])
def test_adaptive_parity(name, make, tol):
    sc = make()
    sim = gpu_sim(sc)
    sim.setSolver(tol, maxSweeps=4000, chunk=8)
    orc = Oracle(n=sc.n)
    orc.set_solver(tol, 4000)
    worst = 0.0
    for k in range(6):
        verts, vels, qs, lv = gpu_state(sim)
        counter = sim.stepCount()
        sim.advance(sim.maxH)
        st = sim.solverStats(reset=True)
        v1, _, q1, _ = gpu_state(sim)
        orc.load_state(verts, vels, qs, lv, sc.qRests if k == 0 else None)
        orc.set_step_counter(counter)
        p = sim.meta.copy()
        orc.substep(p, p.detectionPeriod > 0 and counter % p.detectionPeriod == 0)
        pos, _, _, q = orc.state()
        ep = rel_err(v1["pos"], pos)
        eq = float(np.max(np.abs(q1 - q)))
        worst = max(worst, ep)
        print(f"{name} substep {k}: gpu sweeps {st['sweeps']} resid {st['lastResidual']:.2e} "
              f"| oracle sweeps {orc.last_sweeps()} resid {orc.residual():.2e} | pos {ep:.2e} q {eq:.2e}")
        assert st["substeps"] == 1 and st["notConverged"] == 0
        assert st["lastResidual"] <= tol
        assert abs(st["sweeps"] - orc.last_sweeps()) <= max(3, 0.1 * orc.last_sweeps())
        assert ep <= 1e-5 and eq <= 1e-5
    sim.close()


@pytest.mark.gpu
def test_adaptive_matches_fixed_when_capped():
    """maxSweeps = numItr with an unreachable tolerance reproduces the fixed
    schedule bit for bit (same kernels, same buffers)."""
    sc = scenes.knit_patch(rows=4, per=200)
    a = gpu_sim(sc)
    b = gpu_sim(sc)
    b.setSolver(1e-30, maxSweeps=int(sc.params.numItr), chunk=3)
    for _ in range(5):
        a.advance(a.maxH)
        b.advance(b.maxH)
    va, _, qa, _ = gpu_state(a)
    vb, _, qb, _ = gpu_state(b)
    assert np.array_equal(va["pos"], vb["pos"]) and np.array_equal(qa, qb)
    st = b.solverStats()
    assert st["notConverged"] == 5 and st["sweeps"] == 5 * sc.params.numItr
    a.close()
    b.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name,make,tol", [
    ("readme", scenes.readme_scene, 3e-2),
    ("knit_8x256", lambda: scenes.knit_patch(rows=8, per=256), 3e-2),
])
def test_newton_residual_stop_matches_the_oracle(name, make, tol):
    """Stopping on the Newton residual |f(dx_k)| / |f(dx_0)| (tolerance well
    above the FP32 floor): the device and the FP64 oracle, from the same
    state, stop after nearly the same number of sweeps and report nearly the
    same residual; both residuals are reported whichever one stops."""
    sc = make()
    sim = gpu_sim(sc)
    sim.setSolver(tol, maxSweeps=4000, chunk=8, kind="newton")
    orc = Oracle(n=sc.n)
    orc.set_solver(tol, 4000)
    orc.set_residual_kind(1)
    for k in range(6):
        verts, vels, qs, lv = gpu_state(sim)
        counter = sim.stepCount()
        sim.advance(sim.maxH)
        st = sim.solverStats(reset=True)
        orc.load_state(verts, vels, qs, lv, sc.qRests if k == 0 else None)
        orc.set_step_counter(counter)
        p = sim.meta.copy()
        orc.substep(p, p.detectionPeriod > 0 and counter % p.detectionPeriod == 0)
        print(f"{name} substep {k}: gpu sweeps {st['sweeps']} Newton {st['lastNewtonResidual']:.3e} "
              f"update {st['lastResidual']:.2e} | oracle sweeps {orc.last_sweeps()} "
              f"Newton {orc.newton_residual():.3e} update {orc.update_residual():.2e}")
        assert st["notConverged"] == 0 and st["lastNewtonResidual"] <= tol
        assert st["lastResidual"] == st["lastNewtonResidual"]  # the stopping one
        assert abs(st["sweeps"] - orc.last_sweeps()) <= max(3, 0.05 * orc.last_sweeps())
        if st["sweeps"] == orc.last_sweeps():
            assert st["lastNewtonResidual"] == pytest.approx(orc.newton_residual(), rel=2e-2)
    sim.close()


@pytest.mark.gpu
def test_substep_timeline_records_each_substep():
    """The residual-mode timeline (yb_debug_timeline): one row per substep,
    non-negative durations, the converged chunk consistent with the sweep
    count, and the recorded sweeps equal to the solver's own count."""
    sc = scenes.knit_patch(rows=8, per=256)
    sim = gpu_sim(sc)
    chunk = 8
    sim.setSolver(1e-5, maxSweeps=4000, chunk=chunk)
    sim.advance(sim.maxH)  # graphs built, tree built
    sim.solverStats(reset=True)
    sim.startTimeline(4)
    for _ in range(4):
        sim.advance(sim.maxH)
    tl = sim.timeline()
    st = sim.solverStats()
    print(tl)
    assert tl.shape == (4, 8)
    assert np.all(tl[:, :5] >= 0.0) and np.all(tl[:, 1] > 0.0) and np.all(tl[:, 4] > 0.0)
    sweeps = tl[:, 5].astype(int)
    assert np.all(sweeps >= 1) and int(sweeps.sum()) == st["sweeps"]
    assert np.array_equal(tl[:, 7].astype(int), (sweeps - 1) // chunk)
    assert np.all(tl[:, 6] >= tl[:, 7] + 1)
    # recording stopped: nothing more is kept
    sim.advance(sim.maxH)
    assert sim.timeline().shape == (0, 8)
    sim.close()
