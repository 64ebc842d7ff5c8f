"""The device stepper against the committed golden fixtures of the oracle:
the README scene after 1, 10 and 100 substeps (no re-synchronisation in
between) and one prologue of a 256-vertex knit patch (bit-exact)."""
import ctypes

import numpy as np
import pytest

from golden_util import load, scene
from helpers import contact_sets, gpu_sim, rel_err
from yarnball_amd._lib import check, ptr


@pytest.mark.gpu
def test_readme_substeps_match_golden():
    g = load("readme_substeps.npz")
    sc = scene(g)
    sim = gpu_sim(sc)
    done = 0
    for k, tol in zip(g["steps"], (1e-5, 1e-5, 1e-5)):
        while done < k:
            sim.advance(np.float32(sc.maxH))
            done += 1
        sim.download()
        ep = rel_err(sim.verts["pos"], g[f"pos_{k}"])
        eq = rel_err(sim.qs, g[f"q_{k}"])
        print(f"README after {k} substeps: pos rel {ep:.2e}, q rel {eq:.2e}")
        assert ep < tol and eq < tol
    sim.close()


@pytest.mark.gpu
def test_knit_prologue_matches_golden():
    g = load("knit256_discrete.npz")
    sc = scene(g)
    sim = gpu_sim(sc)
    n = sc.n
    check(sim._lib.yb_debug_write_last_vels(sim.handle, ptr(np.ascontiguousarray(g["lastVels"]))))
    p = g["params"].copy()
    check(sim._lib.yb_debug_prologue(sim.handle, ctypes.byref(p), np.float32(p.h), 1, 1))
    dx = np.zeros((n, 4), np.float32)
    aabb = np.zeros((n, 8), np.float32)
    morton = np.zeros(n, np.uint32)
    order = np.zeros(n, np.int32)
    nc = np.zeros(n, np.int32)
    cols = np.zeros(n * 128, np.int32)
    check(sim._lib.yb_debug_read(sim.handle, ptr(dx), ptr(aabb), ptr(morton), ptr(order), ptr(nc),
                                 ptr(cols)))
    cols = cols.reshape(128, n)
    assert np.array_equal(dx[:, :3].view(np.uint32), g["dx"].view(np.uint32))
    gab = np.concatenate([aabb[:, 0:3], aabb[:, 4:7]], axis=1)
    assert np.array_equal(gab.view(np.uint32), g["aabb"].view(np.uint32))
    assert np.array_equal(morton, g["morton"])
    assert np.array_equal(order, g["order"])
    gcols = np.zeros((128, n), np.int32)
    gcols[: len(g["cols"])] = g["cols"]
    assert contact_sets(nc, cols) == contact_sets(g["numCols"], gcols)
    print(f"knit256 prologue: {int(nc.sum())} contact entries bit-exact")
    sim.close()
