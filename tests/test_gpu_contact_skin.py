"""Contact skin: the broadphase keeps each segment's candidate leaves from a
walk with grown boxes and reuses them while every box stays inside its grown
box.  The contact lists must equal those of a walk in every substep, slot for
slot (the order fixes the sweep's summation order), so a run with the skin is
bit-identical to a run without it -- in both schedules, with the strand
solve, with glue, across the scenes of a batch (which walk per scene), when
candidate lists fill up, and when the margin is so small that segments leave
their grown boxes all the time.  The reference refits and walks in every
detecting substep (collision.cu:94-119); these runs hold the skin to that."""
import numpy as np
import pytest

from yarnball_amd import Sim, learned, scenes
from yarnball_amd._lib import DRIVE_TWIST, YarnBallError

from helpers import gpu_contacts


def _sim(sc, skin, residual=0.0, strand=False, offs=None):
    s = Sim(sc.n)
    s.printErrors = False
    scenes.load_into(s, sc)
    if offs is not None:
        s.setScenes(offs)
    if residual > 0:
        s.setSolver(residual, maxSweeps=4000, chunk=16, checkEvery=4)
    if strand:
        learned.strand(s, every=4)
    if skin is None:
        s.setContactSkin(0.0)
    else:
        s.setContactSkin(*skin)
    return s


def _run_pair(sc, substeps, skin=(0.05, 8.0), residual=0.0, strand=False, offs=None,
              lists_every=1):
    """Advance a handle with the skin and one without, one substep at a time,
    comparing the contact lists after every `lists_every` substeps and the
    states at the end; returns the skin handle's (walks, reuses)."""
    a = _sim(sc, skin, residual, strand, offs)
    b = _sim(sc, None, residual, strand, offs)
    for k in range(substeps):
        a.advance(np.float32(sc.maxH))
        b.advance(np.float32(sc.maxH))
        if (k + 1) % lists_every == 0 or k == substeps - 1:
            na, ca = gpu_contacts(a)
            nb, cb = gpu_contacts(b)
            assert np.array_equal(na, nb), f"substep {k}: contact counts"
            m = int(min(na.max(), 128)) if len(na) else 0
            for slot in range(m):
                live = na > slot
                assert np.array_equal(ca[slot][live], cb[slot][live]), f"substep {k} slot {slot}"
    a.download()
    b.download()
    assert np.array_equal(a.verts["pos"], b.verts["pos"])
    assert np.array_equal(a.vels, b.vels)
    assert np.array_equal(a.qs, b.qs)
    stats = a.contactSkinStats()
    assert b.contactSkinStats() == (0, 0)
    total = int(gpu_contacts(a)[0].sum())
    a.close()
    b.close()
    return stats, total


@pytest.mark.gpu
def test_skin_is_on_by_default_and_settable():
    sc = scenes.readme_scene()
    s = Sim(sc.n)
    scenes.load_into(s, sc)
    assert s.contactSkin() == pytest.approx((0.05, 8.0))
    s.setContactSkin(0.5, 3.0)
    assert s.contactSkin() == pytest.approx((0.5, 3.0))
    s.setContactSkin(0.0)
    assert s.contactSkin()[0] == 0.0
    with pytest.raises(Exception):
        s.setContactSkin(0.25, -1.0)
    s.close()


@pytest.mark.gpu
def test_batches_start_without_the_skin():
    """A handle given several scenes starts with the skin off (unless the
    caller chose a setting); back to one scene, the default returns."""
    scs = scenes.random_walks([300, 400], seed=2)
    big, offs = scenes.batch(scs)
    s = Sim(big.n)
    scenes.load_into(s, big)
    assert s.contactSkin()[0] == pytest.approx(0.05)
    s.setScenes(offs)
    assert s.contactSkin()[0] == 0.0
    s.setScenes([0, big.n])
    assert s.contactSkin()[0] == pytest.approx(0.05)
    s.setContactSkin(0.1, 4.0)
    s.setScenes(offs)
    assert s.contactSkin() == pytest.approx((0.1, 4.0))
    s.close()


@pytest.mark.gpu
@pytest.mark.parametrize("residual", [0.0, 1e-5])
@pytest.mark.parametrize("skin", [(0.05, 8.0), (0.5, 8.0)], ids=["default", "wide"])
def test_knit_lists_and_states_equal_a_walk_every_substep(residual, skin):
    """The pulled knit (one edge moving at 5 cm/s), settling in its first
    frames: every substep's lists equal the walk's.  With the default margin
    its lists rarely last in this settling motion; with a wide one some do,
    so both ways of detecting are compared."""
    sc = scenes.knit_patch(rows=8, per=512)
    (walks, reuses), total = _run_pair(sc, 60, skin=skin, residual=residual)
    print(f"knit 8x512, residual {residual}, skin {skin}: {walks} walks, {reuses} reuses, "
          f"{total} contact entries")
    assert total > 0
    assert walks >= 1 and walks + reuses == 60
    if skin[0] >= 0.5:
        assert reuses >= 3


@pytest.mark.gpu
@pytest.mark.parametrize("skin", [(0.05, 8.0), (0.5, 8.0)], ids=["default", "wide"])
def test_strand_solve_with_the_skin(skin):
    sc = scenes.knit_patch(rows=8, per=512)
    (walks, reuses), _ = _run_pair(sc, 40, skin=skin, residual=1e-6, strand=True)
    print(f"strand solve, skin {skin}: {walks} walks, {reuses} reuses")
    if skin[0] >= 0.5:
        assert reuses >= 1


@pytest.mark.gpu
def test_tiny_margin_walks_often_and_stays_exact():
    """Margin 1e-3 radii and no sweep: segments leave their grown boxes in
    most substeps, and every mix of walking and reusing gives the same lists."""
    sc = scenes.knit_patch(rows=8, per=512)
    (walks, reuses), _ = _run_pair(sc, 16, skin=(1e-3, 0.0))
    print(f"tiny margin: {walks} walks, {reuses} reuses")
    assert walks >= 8


@pytest.mark.gpu
def test_glued_rings_with_the_skin():
    sc = scenes.glued_rings()
    (walks, reuses), total = _run_pair(sc, 30)
    print(f"glued rings: {walks} walks, {reuses} reuses")
    assert total > 0


@pytest.mark.gpu
def test_cable_twist_with_the_skin():
    """The twisted cable with its top end turned by the device driver (1 rev/s
    about the y axis): the driven vertices rotate, and every substep's lists
    still equal the walk's."""
    sc = scenes.twisted_cable(strands=7, per=256)
    frame = np.float32(sc.frame_dt)
    a = _sim(sc, (0.5, 8.0))
    b = _sim(sc, None)
    d = Sim.defaultDriver(DRIVE_TWIST)
    d.speed, d.start, d.duration, d.release, d.side = 2 * np.pi, 0.0, 1e9, -1.0, 1
    d.center[0], d.center[1], d.center[2] = 0.0, float(np.mean(sc.verts["pos"][:, 1])), 0.0
    for s in (a, b):
        s.driverInit()
    t = 0.0
    for f in range(3):
        for s in (a, b):
            s.driveFrame(d, t, float(frame))
            s.advance(frame)
        t += float(frame)
        na, ca = gpu_contacts(a)
        nb, cb = gpu_contacts(b)
        assert np.array_equal(na, nb), f"frame {f}"
        for slot in range(int(min(na.max(), 128))):
            live = na > slot
            assert np.array_equal(ca[slot][live], cb[slot][live]), f"frame {f} slot {slot}"
    a.download()
    b.download()
    assert np.array_equal(a.verts["pos"], b.verts["pos"])
    assert np.array_equal(a.qs, b.qs)
    walks, reuses = a.contactSkinStats()
    print(f"cable, 3 frames: {walks} walks, {reuses} reuses")
    assert reuses >= 10
    a.close()
    b.close()


@pytest.mark.gpu
def test_batched_scenes_walk_per_scene():
    scs = scenes.random_walks([300, 1000, 257, 2500, 600, 1024], seed=3, turn=0.8)
    for sc in scs:
        sc.params.bvhRebuildPeriod = 0.01  # a tree rebuild every 10 substeps
    big, offs = scenes.batch(scs)
    (walks, reuses), total = _run_pair(big, 25, offs=offs)
    print(f"batch of {len(scs)}: {walks} substeps with a walking scene, {reuses} without")
    assert total > 0


@pytest.mark.gpu
def test_batched_residual_schedule_with_the_skin():
    scs = scenes.random_walks([300, 1000, 257, 2500], seed=5, turn=0.8)
    big, offs = scenes.batch(scs)
    _run_pair(big, 12, residual=1e-5, strand=True, offs=offs, lists_every=4)


@pytest.mark.gpu
def test_full_candidate_lists_walk_again():
    """200 segments cross one long segment: its candidate list cannot hold them
    (YB_SKIN_CANDIDATES), so it walks in every substep, with the same lists
    and the same contact-overflow error as without the skin."""
    from test_gpu_glue_and_errors import overflow_scene
    sc = overflow_scene()
    a = _sim(sc, (0.05, 8.0))
    b = _sim(sc, None)
    for k in range(4):
        codes = []
        for s in (a, b):
            try:
                s.advance(np.float32(sc.maxH))
                codes.append(0)
            except YarnBallError as e:
                codes.append(e.code)
        assert codes[0] == codes[1], (k, codes)
        if k == 0:
            assert codes[0] == 1
        na, ca = gpu_contacts(a)
        nb, cb = gpu_contacts(b)
        assert np.array_equal(na, nb), k
        for slot in range(int(min(na.max(), 128))):
            live = na > slot
            assert np.array_equal(ca[slot][live], cb[slot][live]), (k, slot)
    walks, reuses = a.contactSkinStats()
    assert walks == 4 and reuses == 0, (walks, reuses)
    a.close()
    b.close()


def _same_lists(a, b, tag):
    na, ca = gpu_contacts(a)
    nb, cb = gpu_contacts(b)
    assert np.array_equal(na, nb), tag
    for slot in range(int(min(na.max(), 128)) if len(na) else 0):
        live = na > slot
        assert np.array_equal(ca[slot][live], cb[slot][live]), (tag, slot)


@pytest.mark.gpu
def test_caller_changes_between_advances():
    """What a caller may change between advances, each applied to a handle
    with the skin and one without: moved vertices (download, shift a strand
    by a radius, upload), a larger detection radius, zeroed velocities, a
    strand cut in two (a vertex loses its next segment) and joined again.
    Kept lists must not survive any change that a walk would see; every
    substep's lists and the final states stay bit-identical."""
    sc = scenes.knit_patch(rows=8, per=512)
    a = _sim(sc, (0.5, 8.0))
    b = _sim(sc, None)

    def both(fn):
        for s in (a, b):
            fn(s)

    def step(k, tag):
        for _ in range(k):
            both(lambda s: s.advance(np.float32(sc.maxH)))
            _same_lists(a, b, tag)

    step(6, "start")

    def shift(s):
        s.download()
        s.verts["pos"][512:1024, 1] += np.float32(1e-4)
        s.upload()
    both(shift)
    step(4, "shifted strand")

    def wider(s):
        s.meta.detectionScaler = np.float32(1.6)
    both(wider)
    step(4, "wider detection")

    both(lambda s: s.zeroVelocities())
    step(4, "zeroed velocities")

    # cut segment c out of its strand, with the flags configure() would give
    # (YB_FLAG_HAS_PREV 1, HAS_NEXT 2, HAS_NEXT_ORIENTATION 4), then restore them
    c = 1024 + 100
    saved = {}

    def cut(s, on):
        s.download()
        fl = s.verts["flags"]
        if on:
            saved[id(s)] = fl[c - 1:c + 2].copy()
            fl[c] &= np.uint32(~(2 | 4) & 0xffffffff)
            fl[c + 1] &= np.uint32(~1 & 0xffffffff)
            fl[c - 1] &= np.uint32(~4 & 0xffffffff)
        else:
            fl[c - 1:c + 2] = saved[id(s)]
        s.upload()
    both(lambda s: cut(s, True))
    step(4, "cut strand")
    both(lambda s: cut(s, False))
    step(4, "joined strand")

    a.download()
    b.download()
    assert np.array_equal(a.verts["pos"], b.verts["pos"])
    assert np.array_equal(a.qs, b.qs)
    walks, reuses = a.contactSkinStats()
    print(f"caller changes: {walks} walks, {reuses} reuses")
    assert reuses >= 2
    a.close()
    b.close()
