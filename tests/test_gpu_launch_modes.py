"""Eager launches (every kernel enqueued directly) and captured HIP graphs run
the same kernels in the same order, and any block sizes, refit tiles and query
shapes the autotuner may pick compute the same values: bit-identical states,
in the fixed-sweep and in the residual-targeted schedule, with and without the
strand solve (its tiles are the sweep's 62-vertex wave tiles, whatever the
block size)."""
import numpy as np
import pytest

from yarnball_amd import Sim, learned, scenes


def _state(sc, graphs, residual, config=None, strand=False, tile=None):
    sim = Sim(sc.n)
    scenes.load_into(sim, sc)
    if residual > 0:
        sim.setSolver(residual, maxSweeps=4000, chunk=16, checkEvery=4)
    if strand:
        learned.strand(sim, every=6)
    sim.setGraphs(graphs)
    if config:
        sim.setLaunchConfig(*config)
    if tile:
        sim.setTileConfig(*tile)
        assert sim.tileConfig() == tuple(tile)
    sim.advance(np.float32(6 * sc.maxH))
    sim.download()
    st = sim.solverStats() if residual > 0 else None
    out = (sim.verts["pos"].copy(), sim.vels.copy(), sim.qs.copy(), st)
    sim.close()
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("residual", [0.0, 1e-5])
def test_eager_equals_graphs(residual):
    sc = scenes.knit_patch(rows=8, per=256)
    a = _state(sc, True, residual)
    b = _state(sc, False, residual)
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(x, y)
    if residual > 0:
        assert a[3]["sweeps"] == b[3]["sweeps"] and a[3]["substeps"] == b[3]["substeps"] == 6
        print(f"residual {residual}: {a[3]['sweeps']} sweeps in both launch modes")


@pytest.mark.gpu
@pytest.mark.parametrize("residual", [0.0, 1e-5])
def test_launch_shape_does_not_change_results(residual):
    """The autotuner's candidates: block sizes of the sweep and the query."""
    sc = scenes.knit_patch(rows=8, per=256)
    ref = _state(sc, True, residual)
    for cfg in [(64, 64), (128, 256), (512, 128)]:
        got = _state(sc, True, residual, cfg)
        for x, y in zip(ref[:3], got[:3]):
            assert np.array_equal(x, y), cfg
        if residual > 0:
            assert ref[3]["sweeps"] == got[3]["sweeps"], cfg


@pytest.mark.gpu
def test_strand_solve_is_independent_of_launch_mode_and_shape():
    sc = scenes.knit_patch(rows=8, per=256)
    ref = _state(sc, True, 1e-5, strand=True)
    for graphs, cfg in [(False, None), (True, (64, 64)), (True, (512, 128))]:
        got = _state(sc, graphs, 1e-5, cfg, strand=True)
        for x, y in zip(ref[:3], got[:3]):
            assert np.array_equal(x, y), (graphs, cfg)
        assert ref[3]["sweeps"] == got[3]["sweeps"], (graphs, cfg)
    print(f"strand solve: {ref[3]['sweeps']} sweeps in every launch mode and shape")


@pytest.mark.gpu
@pytest.mark.parametrize("residual", [0.0, 1e-5])
def test_tile_shape_does_not_change_results(residual):
    """The autotuner's tile candidates: refit leaves per workgroup (2, 4 or 8
    refit blocks on this patch) and the query's LDS shape."""
    sc = scenes.knit_patch(rows=8, per=256)
    ref = _state(sc, True, residual)
    for tile in [(256, 4, 4), (512, 8, 8), (1024, 16, 8), (256, 8, 6)]:
        got = _state(sc, True, residual, tile=tile)
        for x, y in zip(ref[:3], got[:3]):
            assert np.array_equal(x, y), tile
        if residual > 0:
            assert ref[3]["sweeps"] == got[3]["sweeps"], tile


@pytest.mark.gpu
def test_small_refit_tile_on_a_large_patch():
    """537,600 vertices with 256-leaf refit tiles: 2,100 refit blocks, so the
    block-union table needs a level past the one-workgroup part."""
    sc = scenes.knit_patch(rows=128, per=4200)
    ref = _state(sc, True, 0.0)
    got = _state(sc, True, 0.0, tile=(256, 8, 6))
    for x, y in zip(ref[:3], got[:3]):
        assert np.array_equal(x, y)
