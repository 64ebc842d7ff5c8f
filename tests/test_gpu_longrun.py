"""Long runs: no NaN, no interpenetration (every listed contact stays at least
2 * radius apart), the README scene for its full 5 s (5,000 substeps)."""
import numpy as np
import pytest

from yarnball_amd import scenes

from helpers import gpu_sim


@pytest.mark.gpu
def test_readme_full_five_seconds():
    sc = scenes.readme_scene()
    sim = gpu_sim(sc)
    sim.advance(5.0)
    sim.download()
    assert sim.stepCount() == 5000
    assert np.all(np.isfinite(sim.verts["pos"]))
    assert sim.lastWarningCode == 0
    sep = sim.minSeparation()
    print(f"readme 5 s: min separation {sep:.3f} x 2r, contacts now {int(sim.collisionCounts().sum())}")
    assert sep >= 1.0
    # pinned vertices did not move
    assert np.array_equal(sim.verts["pos"][[0, 32, 63]], sc.verts["pos"][[0, 32, 63]])
    sim.close()


@pytest.mark.gpu
def test_knit_patch_frames_no_penetration():
    sc = scenes.knit_patch(rows=32, per=1024)
    sim = gpu_sim(sc)
    for f in range(6):
        sim.advance(np.float32(1 / 30))
    sim.download()
    assert np.all(np.isfinite(sim.verts["pos"]))
    sep = sim.minSeparation()
    print(f"knit 32x1024 6 frames: min separation {sep:.3f} x 2r")
    assert sim.lastWarningCode == 0 and sep >= 1.0
    # the pulled edge moved by 5 cm/s * 0.2 s
    right = sc.drive_ids
    moved = sim.verts["pos"][right, 0] - sc.verts["pos"][right, 0]
    assert np.allclose(moved, 0.05 * 6 * np.float32(1 / 30), rtol=1e-3)
    sim.close()


@pytest.mark.gpu
def test_knit_patch_strand_solve_no_penetration():
    """The strand solve (before sweeps 1, 5, 9, ..., at residual 1e-6) over 12 frames: every
    substep converges, no NaN, no interpenetration, and the driven edge moves
    as prescribed."""
    from yarnball_amd import learned

    sc = scenes.knit_patch(rows=32, per=1024, seed=4)
    sim = gpu_sim(sc)
    sim.setSolver(1e-6, maxSweeps=6000, chunk=16, checkEvery=4)
    learned.strand(sim, every=4, first=1, frame_refresh=False)
    sim.solverStats(reset=True)
    for f in range(12):
        sim.advance(np.float32(1 / 30))
    sim.download()
    st = sim.solverStats()
    assert st["notConverged"] == 0 and st["maxResidual"] <= 1e-6
    assert np.all(np.isfinite(sim.verts["pos"]))
    sep = sim.minSeparation()
    print(f"knit 32x1024 12 frames with the strand solve: {st['sweeps'] / st['substeps']:.1f} "
          f"sweeps per substep, min separation {sep:.3f} x 2r")
    assert sim.lastWarningCode == 0 and sep >= 1.0
    right = sc.drive_ids
    moved = sim.verts["pos"][right, 0] - sc.verts["pos"][right, 0]
    assert np.allclose(moved, 0.05 * 12 * np.float32(1 / 30), rtol=1e-3)
    sim.close()
