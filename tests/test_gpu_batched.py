"""Batched scenes in one handle: every kernel covers all scenes in one launch,
contact detection stays inside each scene, and with the fixed sweep schedule
each scene evolves bit-identically to a standalone run of that scene."""
import numpy as np
import pytest

from yarnball_amd import Sim, scenes

from helpers import contact_sets, gpu_contacts


def _scenes():
    # compact walks (1 rad turns) so that the strands touch themselves
    scs = scenes.random_walks([300, 1000, 257, 2500, 600, 1024], seed=3, turn=0.8)
    for sc in scs:
        sc.params.bvhRebuildPeriod = 0.01  # a tree rebuild every 10 substeps
    return scs


@pytest.mark.gpu
def test_batched_scenes_equal_standalone_runs():
    scs = _scenes()
    big, offs = scenes.batch(scs)
    steps = 25
    b = Sim(big.n)
    b.printErrors = False  # compact random walks start with a few crossings
    scenes.load_into(b, big)
    b.setScenes(offs)
    b.advance(np.float32(steps * big.maxH))
    b.download()
    bnc, bcols = gpu_contacts(b)
    assert int(bnc.sum()) > 0, "the test scenes should have contacts"
    total = 0
    for k, sc in enumerate(scs):
        s = Sim(sc.n)
        s.printErrors = False
        scenes.load_into(s, sc)
        s.advance(np.float32(steps * sc.maxH))
        s.download()
        sl = slice(offs[k], offs[k + 1])
        assert np.array_equal(b.verts["pos"][sl], s.verts["pos"]), f"scene {k} positions"
        assert np.array_equal(b.vels[sl], s.vels), f"scene {k} velocities"
        assert np.array_equal(b.qs[sl], s.qs), f"scene {k} frames"
        nc, cols = gpu_contacts(s)
        mine = [tuple(j - offs[k] for j in c) for c in contact_sets(bnc[sl], bcols[:, sl])]
        assert mine == contact_sets(nc, cols), f"scene {k} contact sets"
        total += int(nc.sum())
        s.close()
    print(f"{len(scs)} scenes, {big.n} vertices, {total} contact entries: batched == standalone")
    b.close()


@pytest.mark.gpu
def test_set_scenes_rejects_bad_layouts():
    scs = scenes.random_walks([50, 60], seed=2)
    big, offs = scenes.batch(scs)
    s = Sim(big.n)
    scenes.load_into(s, big)
    with pytest.raises(Exception):
        s.setScenes([0, 50])            # does not end at the vertex count
    with pytest.raises(Exception):
        s.setScenes([0, 1, big.n])      # a scene of one vertex
    s.setScenes(offs)
    s.verts["flags"][offs[1] - 1] |= 2  # a segment from scene 0 into scene 1
    with pytest.raises(Exception):
        s.upload()
    s.setScenes([0, big.n])             # single layout: the array end is the only check
    s.upload()
    with pytest.raises(Exception):
        s.setScenes(offs)               # checked against the uploaded flags
    s.close()


def _walk_sizes(count, lo, hi, seed):
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(np.log(lo), np.log(hi), count)).astype(np.int64)


@pytest.mark.gpu
def test_random_walk_batch_against_the_oracle():
    """Eight uneven random-walk scenes (the batched workload's scene type) in
    one handle, substep by substep against the FP64 oracle run on each scene
    alone from the same state: positions and frames <= 1e-5, contact sets
    bit-exact, and no contact across scenes."""
    from oracle import Oracle

    from helpers import gpu_state, rel_err

    scs = scenes.random_walks(_walk_sizes(8, 256, 4096, 5), seed=5, turn=0.8)
    big, offs = scenes.batch(scs)
    b = Sim(big.n)
    b.printErrors = False
    scenes.load_into(b, big)
    b.setScenes(offs)
    orcs = [Oracle(n=sc.n) for sc in scs]
    worst_p = worst_q = 0.0
    contacts = 0
    for k in range(10):
        verts, vels, qs, lv = gpu_state(b)
        counter = b.stepCount()
        b.advance(b.maxH)
        v1, _, q1, _ = gpu_state(b)
        bnc, bcols = gpu_contacts(b)
        sets = contact_sets(bnc, bcols)
        p = b.meta.copy()
        for s, (sc, o) in enumerate(zip(scs, orcs)):
            sl = slice(offs[s], offs[s + 1])
            cid = verts["connectionIndex"][sl]
            assert np.all(cid < 0)
            o.load_state(verts[sl], vels[sl], qs[sl], lv[sl], sc.qRests if k == 0 else None)
            o.set_step_counter(counter)
            ps = p.copy()
            ps.numVerts = sc.n
            o.substep(ps, True)
            pos, _, _, q = o.state()
            worst_p = max(worst_p, rel_err(v1["pos"][sl], pos))
            worst_q = max(worst_q, float(np.max(np.abs(q1[sl].astype(np.float64) - q))))
            onc, ocols = o.contacts()
            mine = [tuple(j - offs[s] for j in c) for c in sets[offs[s]:offs[s + 1]]]
            assert all(0 <= j < sc.n for c in mine for j in c)  # within the scene
            assert mine == contact_sets(onc, ocols), f"substep {k} scene {s}"
            contacts += int(onc.sum())
    print(f"8 random walks ({big.n} vertices): worst pos rel {worst_p:.2e}, frames {worst_q:.2e}, "
          f"{contacts} contact entries")
    assert contacts > 0
    assert worst_p <= 1e-5 and worst_q <= 1e-5
    b.close()


@pytest.mark.gpu
def test_full_512_scene_batch_properties():
    """The batched workload at full size (512 scenes, sizes log-uniform in
    [256, 65536]) for one frame: every value finite, the pinned vertices
    unmoved, contacts only inside each scene, and three scenes (smallest,
    median, largest) bit-identical to standalone runs of themselves."""
    sizes = _walk_sizes(512, 256, 65536, 0)
    scs = scenes.random_walks(sizes, seed=0)
    big, offs = scenes.batch(scs)
    frame = np.float32(big.frame_dt)
    b = Sim(big.n)
    b.printErrors = False
    scenes.load_into(b, big)
    b.setScenes(offs)
    b.advance(frame)
    b.download()
    assert np.all(np.isfinite(b.verts["pos"])) and np.all(np.isfinite(b.qs))
    pinned = big.verts["invMass"] == 0
    assert np.array_equal(b.verts["pos"][pinned], big.verts["pos"][pinned])
    nc, cols = gpu_contacts(b)
    scene_of = np.repeat(np.arange(len(scs)), np.diff(offs))
    seg = np.flatnonzero(nc > 0)
    for i in seg[:: max(1, len(seg) // 20000)]:
        part = cols[: nc[i], i]
        assert np.all(scene_of[part] == scene_of[i])
    order = np.argsort(sizes)
    for s in (order[0], order[len(order) // 2], order[-1]):
        one = Sim(scs[s].n)
        one.printErrors = False
        scenes.load_into(one, scs[s])
        one.advance(frame)
        one.download()
        sl = slice(offs[s], offs[s + 1])
        assert np.array_equal(b.verts["pos"][sl], one.verts["pos"]), f"scene {s}"
        assert np.array_equal(b.qs[sl], one.qs), f"scene {s}"
        one.close()
    print(f"512 scenes, {big.n} vertices, one frame: {int(nc.sum())} contact entries, "
          f"all inside their scenes; scenes {order[0]}, {order[len(order) // 2]}, {order[-1]} "
          f"equal their standalone runs")
    b.close()


def _residual_runs(scs, tol, strand, steps):
    """The batch and a standalone handle per scene, each in the residual-
    targeted schedule (optionally with the strand solve), advanced by `steps`
    substeps; returns the downloaded batch, its stats and the standalone
    handles' states and stats."""
    from yarnball_amd import learned

    def setup(sim):
        sim.printErrors = False
        sim.setSolver(tol, maxSweeps=2000, chunk=8, checkEvery=4)
        if strand:
            learned.strand(sim, every=4, first=1)

    big, offs = scenes.batch(scs)
    b = Sim(big.n)
    scenes.load_into(b, big)
    b.setScenes(offs)
    setup(b)
    b.advance(np.float32(steps * big.maxH))
    b.download()
    bst = b.solverStats()
    alone = []
    for sc in scs:
        s = Sim(sc.n)
        scenes.load_into(s, sc)
        setup(s)
        s.advance(np.float32(steps * sc.maxH))
        s.download()
        alone.append((s.verts["pos"].copy(), s.vels.copy(), s.qs.copy(), s.solverStats()))
        s.close()
    return b, offs, bst, alone


@pytest.mark.gpu
def test_batched_residual_schedule_converges_per_scene():
    """Each scene of a batch stops at its own sweep: with the plain start the
    scene's state equals a standalone run's bit for bit (the per-scene
    residual sums the same per-vertex terms as the standalone residual, in
    tiles that may straddle two scenes), and the scenes' own sweep counts add
    up to the standalone runs' totals, below the batch's sweeps x scenes."""
    scs = _scenes()
    steps = 8
    b, offs, bst, alone = _residual_runs(scs, 1e-5, False, steps)
    total = 0
    for k, (pos, vel, q, st) in enumerate(alone):
        sl = slice(offs[k], offs[k + 1])
        assert st["notConverged"] == 0
        assert np.array_equal(b.verts["pos"][sl], pos), f"scene {k} positions"
        assert np.array_equal(b.vels[sl], vel), f"scene {k} velocities"
        assert np.array_equal(b.qs[sl], q), f"scene {k} frames"
        total += st["sweeps"]
        print(f"scene {k}: {st['sweeps'] / steps:.1f} sweeps per substep standalone")
    assert bst["notConverged"] == 0 and bst["substeps"] == steps
    assert bst["sceneSubsteps"] == steps * len(scs)
    assert bst["sceneSweeps"] == total
    assert bst["sceneSweeps"] < bst["sweeps"] * len(scs)  # scenes stopped apart
    print(f"batch: {bst['sweeps'] / steps:.1f} sweeps per substep, scenes "
          f"{bst['meanSceneSweeps']:.1f} on average")
    b.close()


@pytest.mark.gpu
def test_batched_strand_solve_on_tile_aligned_scenes():
    """Scenes whose sizes are multiples of 62 start on a wave-tile boundary,
    so the batch's strand tiles are the standalone runs' tiles: with the
    strand solve every scene's state equals its standalone run's bit for bit,
    each scene stopping at its own sweep."""
    scs = scenes.random_walks([310, 992, 248, 2480, 620, 1054], seed=3, turn=0.8)
    for sc in scs:
        sc.params.bvhRebuildPeriod = 0.01
    steps = 8
    b, offs, bst, alone = _residual_runs(scs, 1e-5, True, steps)
    total = 0
    for k, (pos, vel, q, st) in enumerate(alone):
        sl = slice(offs[k], offs[k + 1])
        assert st["notConverged"] == 0
        assert np.array_equal(b.verts["pos"][sl], pos), f"scene {k} positions"
        assert np.array_equal(b.qs[sl], q), f"scene {k} frames"
        total += st["sweeps"]
    assert bst["notConverged"] == 0 and bst["sceneSweeps"] == total
    assert bst["sceneSweeps"] < bst["sweeps"] * len(scs)
    print(f"aligned strand batch: {bst['sweeps'] / steps:.1f} sweeps per substep, scenes "
          f"{bst['meanSceneSweeps']:.1f}")
    b.close()


@pytest.mark.gpu
def test_batched_strand_solve_on_unaligned_scenes():
    """Scenes that start inside a wave tile share strand tiles with their
    neighbours, whose tile-boundary solves differ from a standalone run's: the
    states then agree to the solver's accuracy -- no further from the
    standalone strand runs than those are from the plain solver at the same
    tolerance -- and every scene converges at its own sweep."""
    from helpers import rel_err

    scs = _scenes()
    steps = 8
    b, offs, bst, alone = _residual_runs(scs, 1e-5, True, steps)
    _, _, _, plain = _residual_runs(scs, 1e-5, False, steps)
    worst = scale = 0.0
    for k, ((pos, _, _, st), (ppos, _, _, _)) in enumerate(zip(alone, plain)):
        sl = slice(offs[k], offs[k + 1])
        assert st["notConverged"] == 0
        assert np.isfinite(b.verts["pos"][sl]).all()
        worst = max(worst, rel_err(b.verts["pos"][sl], pos))
        scale = max(scale, rel_err(ppos, pos))
    print(f"unaligned strand batch: worst position error {worst:.2e} (standalone plain vs "
          f"strand {scale:.2e}), {bst['sweeps'] / steps:.1f} sweeps per substep, scenes "
          f"{bst['meanSceneSweeps']:.1f}")
    assert bst["notConverged"] == 0 and worst <= 2.0 * scale
    assert bst["sceneSweeps"] < bst["sweeps"] * len(scs)
    b.close()


@pytest.mark.gpu
def test_batched_residual_schedule_needs_scenes_of_64_vertices():
    scs = scenes.random_walks([50, 300], seed=2)
    big, offs = scenes.batch(scs)
    b = Sim(big.n)
    b.printErrors = False
    scenes.load_into(b, big)
    b.setScenes(offs)
    b.setSolver(1e-5, maxSweeps=100, chunk=8, checkEvery=4)
    with pytest.raises(Exception, match="64 vertices"):
        b.advance(np.float32(big.maxH))
    b.setSolver(0.0, maxSweeps=100, chunk=8, checkEvery=4)  # the fixed schedule runs
    b.advance(np.float32(big.maxH))
    b.close()
