"""Device paths the plain parity scenes do not reach, each against the oracle
or an independent brute force:

* glue: the glue spring (cosserat.cu:47-53) and the glued-pair exemption of
  the contact filter (collision.cu:51-52) on closed rings whose end vertices
  are glued -- substep parity with the FP64 oracle, contact sets bit-exact
  with the FP32 twin, and the glued pairs absent only because of the glue;
* the contact-list overflow through yb_advance (collision.cu:80-89,
  YarnBall.cu:283-286): status 1, the Python and C++ facades raise with the
  reference's message, the lists saturate at 128;
* the traversal-stack overflow, forced by a host-set stack limit: its own
  status 3;
* the penetration warning (YarnBall.h:121, YarnBall.cu:292-294): a crossing
  pair closer than 2 radius raises warning 2, and the reported minimum
  separation equals the exact segment distance / (2 radius);
* the step limit (collision.cu:121-159) value for value against the oracle.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from oracle import Oracle, orc32
from yarnball_amd import Sim, scenes
from yarnball_amd._lib import YarnBallError, check, ptr

from helpers import contact_sets, gpu_contacts, gpu_sim, gpu_state, rel_err
from test_cpu_oracle_pins import seg_dist

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def glued_pairs(sc):
    """Segment pairs (x, y) the glue exempts: a vertex of segment x is glued to
    a vertex of segment y (collision.cu:51-52, with x < y)."""
    cid = sc.verts["connectionIndex"]
    seg = (sc.verts["flags"] & 2) != 0
    out = set()
    for x in np.flatnonzero(seg):
        for c in (cid[x], cid[x + 1]):
            if c >= 0:
                for y in (c, c - 1):
                    if 0 <= y < sc.n and seg[y] and abs(x - y) > 2:
                        out.add((int(min(x, y)), int(max(x, y))))
    return out


@pytest.mark.gpu
def test_glued_rings_parity_and_exemption():
    sc = scenes.glued_rings()
    assert np.count_nonzero(sc.verts["connectionIndex"] >= 0) == 8
    sim = gpu_sim(sc)
    orc = Oracle(n=sc.n)
    exempt = glued_pairs(sc)
    assert exempt  # the ring ends' segments touch
    worst = 0.0
    for k in range(20):
        verts, vels, qs, lv = gpu_state(sim)
        counter = sim.stepCount()
        sim.advance(sim.maxH)
        v1, _, q1, _ = gpu_state(sim)
        orc.load_state(verts, vels, qs, lv, sc.qRests if k == 0 else None)
        orc.set_step_counter(counter)
        p = sim.meta.copy()
        orc.substep(p, True)
        pos, _, _, q = orc.state()
        gnc, gcols = gpu_contacts(sim)
        onc, ocols = orc.contacts()
        gsets = contact_sets(gnc, gcols)
        assert gsets == contact_sets(onc, ocols), f"substep {k}"
        for x, y in exempt:
            assert y not in gsets[x] and x not in gsets[y]
        ep = rel_err(v1["pos"], pos)
        eq = float(np.max(np.abs(q1.astype(np.float64) - q)))
        worst = max(worst, ep)
        assert ep <= 1e-5 and eq <= 1e-5, f"substep {k}: pos {ep:.2e} q {eq:.2e}"
    sim.download()
    per = 64
    gaps = [np.linalg.norm(sim.verts["pos"][r * per] - sim.verts["pos"][r * per + per - 1])
            for r in range(4)]
    print(f"glued rings: worst pos rel {worst:.2e}, contacts {int(gnc.sum())}, "
          f"end gaps {max(gaps):.2e} m")
    sim.close()

    # without the glue the same pairs are contacts (coincident ring ends)
    free = sc.copy()
    free.verts["connectionIndex"][:] = -1
    dxw = np.zeros((sc.n, 3), np.float32)
    ab = orc32.aabb(free.params, free.verts["pos"], free.verts["flags"], dxw)
    nc, cols, _ = orc32.contacts(free.params, free.verts["pos"], free.verts["flags"],
                                 free.verts["connectionIndex"], dxw, ab)
    sets = contact_sets(nc, cols)
    assert all(y in sets[x] for x, y in exempt)


def overflow_scene():
    """One long segment crossed by 200 short segments (more than 128 contacts)."""
    m = 200
    n = 2 + 2 * m
    pos = np.zeros((n, 3), np.float32)
    pos[0] = (-1e-2, 0, 0)
    pos[1] = (1e-2, 0, 0)
    for k in range(m):
        x = -9e-3 + k * 9e-5
        pos[2 + 2 * k] = (x, -5e-4, 3e-4)
        pos[3 + 2 * k] = (x, 5e-4, 3e-4)
    breaks = [1] + [3 + 2 * k for k in range(m)]
    return scenes._configure("overflow", pos, breaks_after=breaks, pinned=[0, 1],
                             kCollision=1e-5, kBend=1e-8, kStretch=2e-2, maxH=1e-3,
                             gravity=(0.0, 0.0, 0.0))


@pytest.mark.gpu
def test_contact_overflow_through_advance():
    sc = overflow_scene()
    sim = gpu_sim(sc)
    with pytest.raises(YarnBallError) as ei:
        sim.advance(sim.maxH)
    assert ei.value.code == 1 and "MAX_COLLISIONS_PER_SEGMENT" in str(ei.value)
    assert sim.lastErrorCode == 1
    nc, _ = gpu_contacts(sim)
    assert nc[0] == 128 and nc.max() == 128
    # The error word was cleared after the report: the barrier has pushed the
    # crossing segments apart, the next substep lists fewer contacts and
    # advances without an error.
    sim.lastErrorCode = 0
    sim.advance(sim.maxH)
    nc, _ = gpu_contacts(sim)
    assert sim.lastErrorCode == 0 and nc[0] <= 128
    sim.close()


@pytest.mark.gpu
def test_contact_overflow_through_the_cpp_facade():
    exe = os.path.join(ROOT, "tests", "cpp", "overflow_scene")
    assert os.path.exists(exe), "build with make -C tests/cpp"
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0
    assert "caught: MAX_COLLISIONS_PER_SEGMENT exceeded" in out.stdout
    assert "lastErrorCode 1" in out.stdout
    assert "ERROR: MAX_COLLISIONS_PER_SEGMENT exceeded" in out.stderr


@pytest.mark.gpu
def test_traversal_stack_overflow_has_its_own_status():
    sc = scenes.knit_patch(rows=8, per=256)
    sim = gpu_sim(sc)
    check(sim._lib.yb_debug_set_query_stack(sim.handle, 1))
    with pytest.raises(YarnBallError) as ei:
        sim.advance(sim.maxH)
    assert ei.value.code == 3 and "traversal stack" in str(ei.value)
    assert sim.lastErrorCode == 3
    check(sim._lib.yb_debug_set_query_stack(sim.handle, 64))
    sim.lastErrorCode = 0
    sim.advance(sim.maxH)  # the default depth is enough
    assert sim.lastErrorCode == 0
    assert sim._lib.yb_debug_set_query_stack(sim.handle, 65) != 0
    sim.close()


@pytest.mark.gpu
@pytest.mark.parametrize("gap_over_2r", [0.75, 1.3])
def test_penetration_warning_and_minimum_separation(gap_over_2r):
    from test_cpu_oracle_pins import crossing_scene
    r = 1e-4
    sc = crossing_scene(gap_over_2r * 2 * r)
    sim = gpu_sim(sc)
    sim.minSeparation(reset=True)
    sim.advance(sim.maxH)
    sep = sim.minSeparation()
    x = sc.verts["pos"].astype(np.float64)
    brute = min(seg_dist(x[i], x[i + 1], x[j], x[j + 1])
                for i in (0, 1) for j in (3, 4)) / (2 * float(np.float32(r)))
    print(f"gap {gap_over_2r}: min separation {sep:.6f}, brute force {brute:.6f}, "
          f"warning {sim.lastWarningCode}")
    assert sep == pytest.approx(brute, rel=1e-5)
    assert sim.lastWarningCode == (2 if gap_over_2r < 1 else 0)
    sim.close()


@pytest.mark.gpu
@pytest.mark.parametrize("make", [scenes.readme_scene, lambda: scenes.knit_patch(8, 256),
                                  scenes.glued_rings], ids=["readme", "knit8x256", "rings"])
def test_step_limit_matches_the_oracle(make):
    sc = make()
    sim = gpu_sim(sc)
    for _ in range(3):
        sim.advance(sim.maxH)
    verts, vels, qs, lv = gpu_state(sim)
    p = sim.meta.copy()
    check(sim._lib.yb_debug_prologue(sim.handle, ctypes.byref(p), np.float32(sc.maxH), 1, 0))
    dx = np.zeros((sc.n, 4), np.float32)
    check(sim._lib.yb_debug_read(sim.handle, ptr(dx), None, None, None, None, None))
    orc = Oracle(n=sc.n)
    orc.load_state(verts, vels, qs, lv, sc.qRests)
    orc.set_step_counter(0)
    q = sim.meta.copy()
    q.numItr = 0
    orc.substep(q, True)
    lim, _ = orc.prologue()
    seg = (sc.verts["flags"] & 2) != 0
    assert np.all(np.isinf(dx[~seg, 3])) and np.all(np.isinf(lim[~seg]))
    err = np.max(np.abs(dx[seg, 3] - lim[seg]) / lim[seg])
    bound = lim[seg] < lim[seg].max()  # limited by a contact, not the padding
    print(f"{sc.name}: step limit rel err {err:.2e}, contact-bound segments {int(bound.sum())}")
    assert err <= 1e-5
    sim.close()
