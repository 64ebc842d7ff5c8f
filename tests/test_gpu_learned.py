"""The learned warm start inside the stepper: the correction the solver
applies equals the standalone forward kernel on the same features (after the
[1 2 1]/4 strand filter), a model trained on one knit instance saves sweeps on
another at the same residual, and the substep-level fallback bounds the cost
of a useless model."""
import ctypes
import os
import time

import numpy as np
import pytest

from oracle import Oracle
from yarnball_amd import Sim, learned, scenes, train_warmstart
from yarnball_amd._lib import check, ptr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

FLAG_HAS_PREV, FLAG_HAS_NEXT = 1, 2


def prologue_dx(sim):
    """Warm start (after any correction), the plain warm start and the raw
    features of one prologue from the current state."""
    n = sim.numVerts
    q = sim.meta.copy()
    check(sim._lib.yb_debug_prologue(sim.handle, ctypes.byref(q), np.float32(sim.maxH), 1, 1))
    dx = np.zeros((n, 4), np.float32)
    check(sim._lib.yb_debug_read(sim.handle, ptr(dx), None, None, None, None, None))
    plain = np.zeros((n, 4), np.float32)
    X = np.zeros((n, 48), np.float32)
    check(sim._lib.yb_debug_read_mlp(sim.handle, ptr(plain), ptr(X)))
    return dx, plain, X


@pytest.mark.gpu
def test_applied_correction_equals_forward_kernel():
    sc = scenes.knit_patch(rows=8, per=256)
    rng = np.random.default_rng(0)
    a = Sim(sc.n)
    scenes.load_into(a, sc)
    a.advance(np.float32(5 * sc.maxH))  # some history (velocities, contacts)
    a.download()
    lv = np.zeros((sc.n, 3), np.float32)
    check(a._lib.yb_debug_read_last_vels(a.handle, ptr(lv)))
    state = sc.copy()
    state.verts[:] = a.verts
    state.vels[:] = a.vels
    state.qs[:] = a.qs
    state.params = a.meta.copy()
    a.close()

    def fresh(mode, p=None):
        s = Sim(sc.n)
        scenes.load_into(s, state)
        check(s._lib.yb_debug_write_last_vels(s.handle, ptr(lv)))
        learned.record(s, sc.n)
        if mode == "learned":
            learned.use(s, p)
        return s

    plain = fresh("plain")
    dx_plain, _, X = prologue_dx(plain)
    # a random (untrained) model with the data's normalisation
    T = rng.normal(0, 1e-3, (sc.n, 3)).astype(np.float32)
    p = learned.init_params(X, T, seed=1)
    model = fresh("learned", p)
    dx_l, dx_plain2, X2 = prologue_dx(model)
    assert np.array_equal(X, X2), "features must not depend on the warm-start mode"
    assert np.array_equal(dx_plain, dx_plain2), "plain warm start must not depend on the mode"
    pred = learned.forward(X, p) * p[learned.OSCALE:] + X[:, learned.CPREV]
    fl = state.verts["flags"]
    # strand filter: a missing neighbour counts as the vertex itself
    prev = np.where(((fl & FLAG_HAS_PREV) != 0)[:, None], np.roll(pred, 1, axis=0), pred)
    nxt = np.where(((fl & FLAG_HAS_NEXT) != 0)[:, None], np.roll(pred, -1, axis=0), pred)
    pred = np.float32(0.25) * (prev + np.float32(2.0) * pred + nxt)
    L = state.verts["lRest"].astype(np.float64).copy()
    ends = ((fl & 2) == 0) & ((fl & 1) != 0)
    L[ends] = np.roll(state.verts["lRest"], 1)[ends]
    free = state.verts["invMass"] != 0
    want = dx_plain[:, :3] + pred * L[:, None]
    lim = dx_plain[:, 3]
    ln = np.linalg.norm(want, axis=1)
    scale = np.where((ln > lim) & (ln > 0), lim / np.maximum(ln, 1e-30), 1.0)
    want = want * scale[:, None]
    got = dx_l[:, :3]
    err = np.max(np.abs(got[free] - want[free])) / np.max(np.abs(want[free]))
    print(f"applied vs forward: rel err {err:.2e}")
    assert err < 1e-5
    assert np.array_equal(got[~free], dx_plain[~free, :3])
    plain.close()
    model.close()


def _run(sc, frames, model=None, fallback=True, tol=1e-6):
    sim = Sim(sc.n)
    scenes.load_into(sim, sc)
    sim.setSolver(tol, maxSweeps=6000, chunk=16, checkEvery=4)
    if model is not None:
        learned.use(sim, model, fallback=fallback)
    sim.minSeparation(reset=True)
    sim.solverStats(reset=True)
    t = time.perf_counter()
    sim.advance(np.float32(frames * sc.frame_dt))
    sim.download()
    st = sim.solverStats()
    out = {"sweeps": st["sweeps"], "substeps": st["substeps"], "notConverged": st["notConverged"],
           "maxResidual": st["maxResidual"], "seconds": time.perf_counter() - t,
           "minSep": sim.minSeparation(), "pos": sim.verts["pos"].astype(np.float64).copy()}
    if model is not None:
        out["warm"] = learned.warm_substeps(sim)
        out["counters"] = learned.counters(sim)
    sim.close()
    return out


@pytest.mark.gpu
def test_trained_model_saves_sweeps_on_unseen_instance():
    """Train on the plain solver's converged corrections of one knit instance
    (jitter seed 1) on the device, then run another instance (seed 2) to the
    same residual: fewer sweeps and no interpenetration; and from one shared
    state, one learned and one plain substep agree to the solver tolerance."""
    X, C = train_warmstart.record_scene(scenes.knit_patch(32, 1024, seed=1), 2, 2, 1 << 30,
                                        log=lambda m: None)
    p, hist = train_warmstart.fit(X, C, epochs=40, log=lambda m: None)
    assert hist[-1] < 0.25 * hist[0]
    sc = scenes.knit_patch(32, 1024, seed=2)
    plain = _run(sc, 2)
    net = _run(sc, 2, p, fallback=False)
    auto = _run(sc, 2, p, fallback=True)
    for r in (plain, net, auto):
        assert r["notConverged"] == 0 and r["maxResidual"] <= 1e-6
        assert r["minSep"] >= 1.0
    per = lambda r: r["sweeps"] / r["substeps"]
    print(f"sweeps/substep plain {per(plain):.1f}, learned {per(net):.1f}, with fallback "
          f"{per(auto):.1f} (substeps {auto['warm']}); vertex fallbacks "
          f"{net['counters']['fallback']}")
    assert net["sweeps"] < 0.8 * plain["sweeps"]
    assert auto["sweeps"] < 0.9 * plain["sweeps"]

    # one substep from a shared state (the learned sim keeps its own history
    # of converged corrections, which only its warm start reads)
    a, b = Sim(sc.n), Sim(sc.n)
    for s_ in (a, b):
        scenes.load_into(s_, sc)
        s_.setSolver(1e-6, maxSweeps=6000, chunk=16, checkEvery=4)
    learned.use(b, p, fallback=False)
    a.advance(np.float32(sc.frame_dt))
    b.advance(np.float32(sc.frame_dt))
    a.download()
    lv = np.zeros((sc.n, 3), np.float32)
    check(a._lib.yb_debug_read_last_vels(a.handle, ptr(lv)))
    b.verts[:] = a.verts
    b.vels[:] = a.vels
    b.qs[:] = a.qs
    ctypes.memmove(ctypes.byref(b.meta), ctypes.byref(a.meta), ctypes.sizeof(a.meta))
    b.upload()
    check(b._lib.yb_debug_write_last_vels(b.handle, ptr(lv)))
    a.advance(np.float32(sc.maxH))
    b.advance(np.float32(sc.maxH))
    a.download()
    b.download()
    L = float(np.median(sc.verts["lRest"]))
    d = np.linalg.norm(a.verts["pos"].astype(np.float64) - b.verts["pos"], axis=1) / L
    print(f"one substep from a shared state: |x_learned - x_plain| rms {np.sqrt(np.mean(d**2)):.2e}, "
          f"max {d.max():.2e} rest lengths")
    assert np.sqrt(np.mean(d ** 2)) < 2e-3 and d.max() < 2e-2
    a.close()
    b.close()


@pytest.mark.gpu
def test_fallback_bounds_cost_of_useless_model():
    """An untrained network is a poor warm start; with the substep-level
    fallback the run costs at most a little more than the plain solver."""
    sc = scenes.knit_patch(16, 512, seed=3)
    probe = Sim(sc.n)
    scenes.load_into(probe, sc)
    learned.record(probe, sc.n)
    probe.advance(np.float32(sc.maxH))
    X, _ = learned.fetch(probe)
    probe.close()
    rng = np.random.default_rng(0)
    p = learned.init_params(X, rng.normal(0, 0.05, (len(X), 3)).astype(np.float32), seed=4)
    plain = _run(sc, 2)
    net = _run(sc, 2, p, fallback=False)
    auto = _run(sc, 2, p, fallback=True)
    print(f"sweeps plain {plain['sweeps']}, untrained {net['sweeps']}, with fallback "
          f"{auto['sweeps']} (substeps {auto['warm']})")
    assert net["sweeps"] > 1.2 * plain["sweeps"], "the untrained model should hurt"
    assert auto["sweeps"] < 1.25 * plain["sweeps"]
    assert auto["warm"]["plain"] > auto["warm"]["learned"]


@pytest.mark.gpu
def test_learned_apply_refreshes_the_frames():
    """The learned start's apply also refreshes the frames, in the same pass:
    after a prologue the frames equal one frame update (cosserat.cu:209-253,
    the FP64 oracle's) of the previous frames at the learned warm start; with
    the refresh off they are the previous frames, bit for bit."""
    sc = scenes.knit_patch(rows=8, per=256)
    a = Sim(sc.n)
    scenes.load_into(a, sc)
    a.advance(np.float32(5 * sc.maxH))
    a.download()
    lv = np.zeros((sc.n, 3), np.float32)
    check(a._lib.yb_debug_read_last_vels(a.handle, ptr(lv)))
    state = sc.copy()
    state.verts[:] = a.verts
    state.vels[:] = a.vels
    state.qs[:] = a.qs
    state.params = a.meta.copy()
    a.close()
    model = np.load(os.path.join(ROOT, "yarnball_amd", "models", "warmstart.npy"))
    out = {}
    for refresh in (True, False):
        s = Sim(sc.n)
        scenes.load_into(s, state)
        check(s._lib.yb_debug_write_last_vels(s.handle, ptr(lv)))
        learned.record(s, sc.n)  # (keeps the plain start readable)
        learned.use(s, model, frame_refresh=refresh)
        dx, plain, _ = prologue_dx(s)
        s.download()
        out[refresh] = (dx, plain, s.qs.astype(np.float64).copy())
        s.close()
    dx, plain, q = out[True]
    assert np.array_equal(dx, out[False][0]), "the refresh must not change the warm start"
    assert np.array_equal(out[False][2], state.qs.astype(np.float64)), "frames left alone"
    assert not np.array_equal(dx, plain), "the learned start moved the iterate"
    o = Oracle(n=sc.n)
    o.load_state(state.verts, state.vels, state.qs, lv, state.qRests)
    want = o.frame_update(dx[:, :3])
    err = np.abs(q - want).max()
    moved = np.abs(want - state.qs).max()
    print(f"refreshed frames vs the oracle's frame update: max err {err:.2e} "
          f"(frames moved by up to {moved:.2e})")
    assert err <= 1e-5 and moved > 10 * err
    o.close()


def _headline_run(sc, frames, newton):
    sim = Sim(sc.n)
    scenes.load_into(sim, sc)
    sim.setSolver(1e-6, maxSweeps=6000, chunk=16, checkEvery=4)
    model = np.load(os.path.join(ROOT, "yarnball_amd", "models", "warmstart_strand.npy"))
    learned.use(sim, model, strand=True)
    learned.strand_schedule(sim, every=4, first=2)
    # out-of-range settings are refused with the handle left as it was
    assert sim._lib.yb_set_newton_residual(sim.handle, 2) != 0
    assert sim._lib.yb_set_learned_start_cost(sim.handle, -1.0) != 0
    assert sim._lib.yb_set_learned_start_cost(sim.handle, float("nan")) != 0
    sim.setNewtonResidual(newton)
    sim.solverStats(reset=True)
    learned.warm_substeps(sim, reset=True)
    sim.advance(np.float32(frames * sc.frame_dt))
    sim.download()
    out = {"stats": sim.solverStats(), "warm": learned.warm_substeps(sim),
           "pos": sim.verts["pos"].copy(), "qs": sim.qs.copy(), "vels": sim.vels.copy()}
    sim.close()
    return out


@pytest.mark.gpu
def test_newton_residual_at_plain_starts_only():
    """Recording the Newton residual only in substeps that start from the plain
    warm start skips the learned substeps' extra force pass and changes no
    iterate: the learned start + strand solve ends in the same state, bit for
    bit, with the same sweeps, and records the residual for exactly the
    plain-start substeps."""
    sc = scenes.knit_patch(16, 512, seed=3)
    a = _headline_run(sc, 2, "all")
    b = _headline_run(sc, 2, "plain")
    sa, sb = a["stats"], b["stats"]
    print(f"substeps {sa['substeps']}, learned starts {b['warm']['learned']}, plain starts "
          f"{b['warm']['plain']}; Newton residual recorded in {sa['newtonSubsteps']} / "
          f"{sb['newtonSubsteps']} substeps, mean {sa['meanNewtonResidual']:.3e} / "
          f"{sb['meanNewtonResidual']:.3e}")
    for k in ("pos", "qs", "vels"):
        assert np.array_equal(a[k], b[k]), k
    assert sa["sweeps"] == sb["sweeps"] and a["warm"] == b["warm"]
    assert b["warm"]["learned"] > 0 and b["warm"]["plain"] > 0
    assert sa["newtonSubsteps"] == sa["substeps"]
    assert sb["newtonSubsteps"] == b["warm"]["plain"]
    assert sb["maxNewtonResidual"] <= sa["maxNewtonResidual"]
    assert 0.0 < sb["sumNewtonResidual"] <= sa["sumNewtonResidual"]
