"""The device's solvers against converged FP64 substeps.

tests/golden/converged_substeps.npz (written by tests/golden/make_converged.py)
holds FP32 start states of the README scene, the knit patch 8 x 256 and the
twisted cable 7 x 256, and the FP64 oracle's solution of the next substep
solved until its Newton residual |f| / |f0| <= 1e-9 (cosserat.cu:20-253 run to
its fixed point).  Each device method starts from the same FP32 state (re-
synced for every substep) and stops at its own tolerance; its displacement dx
is compared with the converged one:

  dx error   max |dx - dx*| / max |dx*|
  pos error  max |x - x*| / max |x*|   (the normwise measure of the parity tests)

Measured on MI355X (update residual 1e-6 for both methods):

  scene         plain: sweeps, dx error   strand: sweeps, dx error
  README            40, 3.7e-3                16, 9.4e-3
  knit 8 x 256    2076, 4.8e-2              1736, 5.1e-2
  cable 7 x 256   1000, 5.3e-3               568, 2.5e-3

So the update residual 1e-6 stops both solvers 0.3-5% of a substep's
displacement short of convergence, the strand solve within a factor 2.6 of
the plain solver's error either way, in fewer sweeps.  FP32 state cannot come
closer than about 1e-4 in dx (an FP32 oracle run for 1.5e5 sweeps stays ~1e-4
away), so no method could be held to less.
"""
import os

import numpy as np
import pytest

from yarnball_amd import learned, scenes
from yarnball_amd._lib import Params, check, ptr

from helpers import gpu_sim

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                    "converged_substeps.npz")
BUILDERS = {
    "readme": scenes.readme_scene,
    "knit8x256": lambda: scenes.knit_patch(rows=8, per=256),
    "cable7x256": lambda: scenes.twisted_cable(strands=7, per=256),
}
CASES = [(name, k) for name in BUILDERS for k in (0, 1)]


def load_case(name, k):
    z = np.load(GOLD, allow_pickle=False)
    sc = BUILDERS[name]().copy()
    sc.params = Params.from_buffer_copy(z[f"{name}/params"].tobytes())
    base = f"{name}/{k}/"
    sc.verts["pos"] = z[base + "pos"]
    sc.vels[:] = z[base + "vels"]
    sc.qs[:] = z[base + "qs"]
    return sc, z[base + "lastVels"], z[base + "dx"], float(z[base + "newton"])


def run(sc, lastVels, method):
    sim = gpu_sim(sc)
    check(sim._lib.yb_debug_write_last_vels(sim.handle, ptr(np.ascontiguousarray(lastVels))))
    if method == "plain":
        sim.setSolver(1e-6, maxSweeps=20000, chunk=16, checkEvery=4)
    elif method == "strand":
        sim.setSolver(1e-6, maxSweeps=20000, chunk=16, checkEvery=4)
        learned.strand(sim, every=4, first=1)
    x0 = sc.verts["pos"].astype(np.float64)
    sim.advance(sim.maxH)
    sim.download()
    st = sim.solverStats()
    dx = sim.verts["pos"].astype(np.float64) - x0
    sim.close()
    return dx, st


@pytest.mark.gpu
@pytest.mark.parametrize("name,k", CASES, ids=[f"{n}-{k}" for n, k in CASES])
def test_solvers_against_the_converged_substep(name, k):
    sc, lv, dxref, newton_ref = load_case(name, k)
    assert newton_ref <= 1.01e-9
    xref = sc.verts["pos"].astype(np.float64) + dxref
    res = {}
    for method in ("plain", "strand"):
        dx, st = run(sc, lv, method)
        x = sc.verts["pos"].astype(np.float64) + dx
        res[method] = dict(
            dx=float(np.abs(dx - dxref).max() / np.abs(dxref).max()),
            pos=float(np.abs(x - xref).max() / np.abs(xref).max()),
            sweeps=int(st["sweeps"]), newton=float(st["lastNewtonResidual"]),
            conv=int(st["notConverged"]) == 0)
        print(f"{name} substep {k} {method:14s}: dx error {res[method]['dx']:.2e}, "
              f"pos error {res[method]['pos']:.2e}, sweeps {res[method]['sweeps']}, "
              f"Newton residual {res[method]['newton']:.2e}")
    for r in res.values():
        assert r["conv"] and np.isfinite(r["dx"])
    s, p = res["strand"], res["plain"]
    assert p["dx"] <= 0.1 and s["dx"] <= 0.1          # both within 10% of the displacement
    assert s["pos"] <= 1e-4 and p["pos"] <= 1e-4
    assert s["dx"] <= 3.0 * p["dx"] + 1e-3            # as close as plain, within a factor 3
    assert s["sweeps"] <= p["sweeps"]
