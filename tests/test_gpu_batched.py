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
