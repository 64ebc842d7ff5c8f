"""Device stepper vs FP64 CPU oracle, substep by substep.

Each substep starts both sides from the device's FP32 state (re-synced, so
chaotic divergence cannot accumulate), runs one full substep (warm start,
detection, step limit, numItr sweeps, end) and compares:
  * contact lists: identical sets (bit-exact discrete stages);
  * positions: normwise relative error <= 1e-5;
  * frames (unit quaternions): max abs error <= 1e-5.
"""
import numpy as np
import pytest

from oracle import Oracle
from yarnball_amd import scenes

from helpers import contact_sets, gpu_contacts, gpu_sim, gpu_state, rel_err

POS_TOL = 1e-5
Q_TOL = 1e-5


def run_parity(sc, substeps, check_contacts=True, label=""):
    sim = gpu_sim(sc)
    orc = Oracle(n=sc.n)
    worst_p = worst_q = 0.0
    total_contacts = 0
    for k in range(substeps):
        verts, vels, qs, lv = gpu_state(sim)
        counter = sim.stepCount()
        sim.advance(sim.maxH)
        v1, u1, q1, _ = gpu_state(sim)
        orc.load_state(verts, vels, qs, lv, sc.qRests if k == 0 else None)
        orc.set_step_counter(counter)
        p = sim.meta.copy()
        det = p.detectionPeriod > 0 and counter % p.detectionPeriod == 0
        orc.substep(p, det)
        pos, vel, _, q = orc.state()
        ep = rel_err(v1["pos"], pos)
        eq = float(np.max(np.abs(q1.astype(np.float64) - q)))
        worst_p = max(worst_p, ep)
        worst_q = max(worst_q, eq)
        if check_contacts and det:
            gnc, gcols = gpu_contacts(sim)
            onc, ocols = orc.contacts()
            assert contact_sets(gnc, gcols) == contact_sets(onc, ocols), f"{label} substep {k}"
            total_contacts += int(gnc.sum())
        assert ep <= POS_TOL, f"{label} substep {k}: position rel err {ep:.3e}"
        assert eq <= Q_TOL, f"{label} substep {k}: frame err {eq:.3e}"
    sim.close()
    return worst_p, worst_q, total_contacts


@pytest.mark.gpu
def test_readme_scene_parity():
    sc = scenes.readme_scene()
    ep, eq, _ = run_parity(sc, 60, label="readme")
    print(f"readme: worst pos rel {ep:.2e}, worst q {eq:.2e}")


@pytest.mark.gpu
def test_readme_scene_parity_late():
    """Later in the README run, when the strands swing into contact."""
    sc = scenes.readme_scene()
    sim = gpu_sim(sc)
    sim.advance(0.6)
    sim.download()
    sc2 = sc.copy()
    sc2.verts[:] = sim.verts
    sc2.vels[:] = sim.vels
    sc2.qs[:] = sim.qs
    sc2.params = sim.meta.copy()
    sim.close()
    ep, eq, ncont = run_parity(sc2, 30, label="readme-late")
    print(f"readme late: worst pos rel {ep:.2e}, worst q {eq:.2e}, contacts {ncont}")


@pytest.mark.gpu
def test_knit_patch_small_parity():
    """Dense contact from t=0: 8 rows x 256 vertices of the knit patch."""
    sc = scenes.knit_patch(rows=8, per=256)
    ep, eq, ncont = run_parity(sc, 12, label="knit-small")
    assert ncont > 0
    print(f"knit small: worst pos rel {ep:.2e}, worst q {eq:.2e}, contacts {ncont}")


@pytest.mark.gpu
def test_twisted_cable_small_parity():
    sc = scenes.twisted_cable(strands=7, per=256)
    ep, eq, ncont = run_parity(sc, 12, label="cable-small")
    print(f"cable small: worst pos rel {ep:.2e}, worst q {eq:.2e}, contacts {ncont}")


@pytest.mark.gpu
def test_frames_negated_give_the_same_substeps():
    """q and -q are the same rotation: uploading the knit patch with a random
    half of its frames negated gives bit-identical positions after several
    substeps, and every frame equal up to its sign (the device restates the
    sign choices of cosserat.cu:228-245; rotate_e1 is quadratic in q)."""
    sc = scenes.knit_patch(rows=8, per=256)
    flip = np.random.default_rng(11).random(sc.n) < 0.5
    sc2 = sc.copy()
    sc2.qs = np.where(flip[:, None], -sc.qs, sc.qs).astype(np.float32)
    a, b = gpu_sim(sc), gpu_sim(sc2)
    for _ in range(6):
        a.advance(a.maxH)
        b.advance(b.maxH)
    va, _, qa, _ = gpu_state(a)
    vb, _, qb, _ = gpu_state(b)
    assert np.array_equal(va["pos"], vb["pos"])
    same = np.all(qa == qb, axis=1) | np.all(qa == -qb, axis=1)
    assert same.all(), f"{int((~same).sum())} frames differ beyond their sign"
    a.close()
    b.close()


@pytest.mark.gpu
def test_advance_splits_h_into_equal_substeps():
    """advance(h) runs ceil(h / maxH) substeps of h / steps each (step.cpp:36-55),
    and each substep's warm start uses the previous substep's length (lastH):
    the device against the oracle's restatement of advance, for h = 2.5 maxH,
    then h = maxH / 3, then h = 4 maxH exactly."""
    sc = scenes.readme_scene()
    sim = gpu_sim(sc)
    orc = Oracle(sc)
    p = sc.params.copy()
    steps = 0
    for ratio, n in ((2.5, 3), (1.0 / 3.0, 1), (4.0, 4)):
        h = float(np.float32(ratio * sc.maxH))
        sim.advance(h)
        orc.advance(p, h, sc.maxH)
        steps += n
        assert sim.stepCount() == steps
        v, _, q, _ = gpu_state(sim)
        pos, _, _, oq = orc.state()
        assert rel_err(v["pos"], pos) <= 1e-5, ratio
        assert float(np.max(np.abs(q - oq))) <= 1e-5, ratio
    sim.close()
