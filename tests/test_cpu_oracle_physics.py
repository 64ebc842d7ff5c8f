"""Closed-form behaviour of the FP64 oracle stepper."""
import numpy as np
import pytest

from oracle import Oracle
from yarnball_amd import scenes


def free_fall_scene(n=6, drag=0.2):
    pos = np.zeros((n, 3), np.float32)
    pos[:, 0] = np.arange(n) * 0.002
    sc = scenes._configure("free_fall", pos, breaks_after=[], pinned=[], kCollision=1e-7,
                           kBend=1e-12, kStretch=0.0, maxH=1e-3, detectionPeriod=-1,
                           useStepSizeLimit=0)
    sc.params.drag = drag
    sc.vels[:] = np.array([0.1, 0.2, -0.3], np.float32)
    return sc


def free_fall_reference(sc, steps, h):
    g = np.array(sc.params.gravity, np.float64)
    x = sc.verts["pos"].astype(np.float64)
    v = sc.vels.astype(np.float64)
    drag = float(sc.params.drag)
    for _ in range(steps):
        dx = h * v + h * h * g  # y - x: exact Newton step with no other force
        v = dx / h * (1 - drag * h)
        x = x + dx
    return x, v


def test_free_fall_closed_form():
    sc = free_fall_scene()
    o = Oracle(sc)
    p = sc.params.copy()
    for _ in range(40):
        o.advance(p, sc.maxH, sc.maxH)
    pos, vel, _, _ = o.state()
    h = float(np.float32(sc.maxH))
    xr, vr = free_fall_reference(sc, 40, h)
    assert np.allclose(pos, xr, rtol=0, atol=1e-12)
    assert np.allclose(vel, vr, rtol=1e-10, atol=1e-12)


def test_pinned_vertices_move_with_their_velocity():
    # iteration.cu:23-35, cosserat.cu:195: pinned vertices advect by h*v exactly
    sc = scenes.readme_scene()
    sc.vels[0] = np.array([0.01, 0.0, 0.0], np.float32)
    o = Oracle(sc)
    p = sc.params.copy()
    x0 = sc.verts["pos"][0].astype(np.float64)
    for _ in range(10):
        o.advance(p, sc.maxH, sc.maxH)
    pos, vel, _, _ = o.state()
    h = float(np.float32(sc.maxH))
    assert np.allclose(pos[0], x0 + 10 * h * np.array([np.float32(0.01), 0, 0]), atol=1e-15)
    assert np.allclose(pos[32], sc.verts["pos"][32], atol=0)


def test_rest_configuration_is_a_fixed_point():
    """A straight, unloaded rod at rest with no gravity stays put: stretch
    constraint satisfied (c = 0) and frames equal their rest rotation."""
    pos = np.zeros((10, 3), np.float32)
    pos[:, 0] = np.arange(10) * 0.002
    sc = scenes._configure("rest", pos, breaks_after=[], pinned=[0], kCollision=1e-7,
                           kBend=1e-8, kStretch=2e-2, maxH=1e-3, gravity=(0.0, 0.0, 0.0))
    o = Oracle(sc)
    p = sc.params.copy()
    for _ in range(20):
        o.advance(p, sc.maxH, sc.maxH)
    pos2, vel, _, q = o.state()
    assert np.max(np.abs(pos2 - sc.verts["pos"])) < 1e-12
    assert np.max(np.abs(q - sc.qs)) < 1e-6


def test_frame_update_is_rotation_equivariant():
    """Rotating the whole scene (positions, frames, gravity) rotates the result."""
    sc = scenes.readme_scene()
    ang = 0.7
    R = np.array([[np.cos(ang), 0, np.sin(ang)], [0, 1, 0], [-np.sin(ang), 0, np.cos(ang)]])
    qR = np.array([0, np.sin(ang / 2), 0, np.cos(ang / 2)])

    def qmul(a, b):
        ax, ay, az, aw = a
        bx, by, bz, bw = b
        return np.array([aw * bx + bw * ax + ay * bz - az * by,
                         aw * by + bw * ay + az * bx - ax * bz,
                         aw * bz + bw * az + ax * by - ay * bx,
                         aw * bw - ax * bx - ay * by - az * bz])

    o1 = Oracle(sc)
    p1 = sc.params.copy()
    p1.gravity[0], p1.gravity[1], p1.gravity[2] = 0.0, -9.8, 0.0
    sc2 = sc.copy()
    o2 = Oracle(n=sc.n)
    o2.load_scene(sc)
    pos, vel, lv, q = o1.state()
    o2.lib.orc_load64(o2.o, *(np.ascontiguousarray(a).ctypes.data_as(__import__("ctypes").c_void_p)
                              for a in ((pos @ R.T), (vel @ R.T), (lv @ R.T),
                                        np.array([qmul(qR, qq) for qq in q]))))
    p2 = p1.copy()
    g = R @ np.array([0.0, -9.8, 0.0])
    p2.gravity[0], p2.gravity[1], p2.gravity[2] = g
    p1.detectionPeriod = p2.detectionPeriod = -1
    for _ in range(5):
        o1.advance(p1, sc.maxH, sc.maxH)
        o2.advance(p2, sc.maxH, sc.maxH)
    a, _, _, qa = o1.state()
    b, _, _, qb = o2.state()
    assert np.allclose(a @ R.T, b, atol=1e-9)
    qa_rot = np.array([qmul(qR, qq) for qq in qa])
    assert np.allclose(qa_rot, qb, atol=1e-6)


def test_frame_update_is_sign_flip_robust():
    """q and -q are the same rotation: negating any frames of the state leaves
    the positions of the following substeps unchanged and every frame equal up
    to its sign (the update picks each neighbour term's sign from a dot
    product, cosserat.cu:228-245)."""
    sc = scenes.readme_scene()
    p = sc.params.copy()
    a = Oracle(sc)
    b = Oracle(sc)
    pos, vel, lv, q = a.state()
    flip = np.zeros(sc.n, bool)
    flip[np.random.default_rng(5).random(sc.n) < 0.5] = True
    q2 = np.where(flip[:, None], -q, q)
    b.load64(pos=pos, vels=vel, lastVels=lv, qs=q2)
    for _ in range(4):
        a.advance(p, sc.maxH, sc.maxH)
        b.advance(p, sc.maxH, sc.maxH)
    pa, _, _, qa = a.state()
    pb, _, _, qb = b.state()
    assert np.max(np.abs(pa - pb)) <= 1e-12 * np.max(np.abs(pa))
    same = np.minimum(np.abs(qa - qb).max(axis=1), np.abs(qa + qb).max(axis=1))
    assert same.max() <= 1e-12


@pytest.mark.parametrize("sweeps", [1, 3])
def test_frame_update_export_is_the_sweeps_frame_update(sweeps):
    """orc_frame_update, on the state before a substep and at the iterate that
    substep ends with (its position change), gives the frames the substep
    leaves: the frame update of its last sweep, whose dx is that iterate.  With
    one sweep the frames before it are the substep's own; with three, the
    export is applied to the frames after the second sweep, read from a
    two-sweep run of the same state."""
    sc = scenes.readme_scene()
    sc.vels[:, 2] = np.float32(0.05)  # some motion besides gravity
    p = sc.params.copy()
    p.h = np.float32(sc.maxH)
    p.detectionPeriod = -1
    start = Oracle(sc)
    pos0, vel0, lv0, q0 = start.state()
    qs_before = q0
    if sweeps > 1:
        pre = Oracle(sc)
        pp = p.copy()
        pp.numItr = sweeps - 1
        pre.substep(pp, False)
        # the frames after sweep (sweeps - 1) of this substep
        qs_before = pre.state()[3]
    p.numItr = sweeps
    run = Oracle(sc)
    run.substep(p, False)
    pos1, _, _, q1 = run.state()
    dx = pos1 - pos0
    probe = Oracle(sc)
    probe.load64(pos=pos0, vels=vel0, lastVels=lv0, qs=qs_before)
    got = probe.frame_update(dx)
    assert np.abs(got - q1).max() <= 1e-10
    assert np.abs(q1 - q0).max() > 1e-6  # the frames did turn
    # the rest configuration: a straight rod with rest frames stays put
    rod = np.zeros((10, 3), np.float32)
    rod[:, 0] = np.arange(10) * 0.002
    rs = scenes._configure("rest", rod, breaks_after=[], pinned=[0], kCollision=1e-7,
                           kBend=1e-8, kStretch=2e-2, maxH=1e-3, gravity=(0.0, 0.0, 0.0))
    ro = Oracle(rs)
    qr = ro.state()[3]
    assert np.abs(ro.frame_update(np.zeros((10, 3))) - qr).max() <= 1e-12


@pytest.mark.parametrize("ratio,s", [(1.0, 1.0), (0.5, 0.5), (-1.0, 0.0), (2.0, 1.0), (0.0, 0.0)])
def test_warm_start_acceleration_clamp(ratio, s):
    """The plain warm start (iteration.cu:7-41) in closed form:
    dx0 = h v + h^2 s g with s = clamp(a.g / |g|^2, 0, 1), a = (v - v_last) /
    lastH, for a free vertex; y - x = h v + h^2 g; a pinned vertex gets h v.
    Cases: a = ratio * g along g (s = ratio clamped), and a perpendicular to g
    (s = 0), with lastH != h."""
    from oracle import orc32
    from yarnball_amd._lib import Params

    p = Params()
    p.h = 1e-3
    p.lastH = 2e-3
    g = np.array([0.0, 0.0, -9.8])
    p.gravity[0], p.gravity[1], p.gravity[2] = g
    v = np.array([[0.3, -0.2, 0.1]] * 3)
    a_par = ratio * g
    a_perp = np.array([2.0, 1.0, 0.0])  # perpendicular to g
    lastV = np.stack([v[0] - a_par * p.lastH, v[1] - a_perp * p.lastH, v[2] - a_par * p.lastH])
    invMass = np.array([1.0, 1.0, 0.0])
    dx, y = orc32.warm_start(p, invMass, v.astype(np.float32), lastV.astype(np.float32))
    h = p.h
    expect = [h * v[0] + h * h * s * g, h * v[1], h * v[2]]
    for k in range(3):
        assert np.allclose(dx[k], expect[k], rtol=1e-5, atol=1e-9), (k, dx[k], expect[k])
    assert np.allclose(y[0], h * v[0] + h * h * g, rtol=1e-5, atol=1e-9)
