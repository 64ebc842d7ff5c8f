"""A scene described in JSON (curve file + simulation block), built and
stepped through the drop-in C++ header (buildFromJSON, advance, exportToOBJ)
and through the Python mirror: identical states, physically sane."""
import os
import subprocess

import numpy as np
import pytest

from yarnball_amd import Sim, scene_io, scenes

from test_cpu_scene_io import write_poly

HERE = os.path.dirname(os.path.abspath(__file__))
PROG = os.path.join(HERE, "cpp", "json_scene")


def make_scene(tmp_path):
    t = np.linspace(0, 6 * np.pi, 200)
    helix = np.stack([np.cos(t), 0.15 * t, np.sin(t)], axis=1) * 2.0   # file units (cm)
    a = np.linspace(0, 2 * np.pi, 90, endpoint=False)
    ring = np.stack([1.2 * np.cos(a), np.full_like(a, 1.4), 1.2 * np.sin(a)], axis=1) * 2.0
    write_poly(tmp_path / "coil.poly", [(helix, False), (ring, True)])
    (tmp_path / "coil.json").write_text("""{
      "curveFile": "coil.poly",
      "resampleLength": 1e-3,
      "curveRadius": 0.0004,
      "fixVertex": [[0.02, 0.0, 0.0], [0.02, 0.0565, 0.0]],
      "simulation": {
        "density": 1e-3, "frictionCoeff": 0.1, "kStretch": 1e-1, "kBend": 2e-7,
        "kCollision": 2e-4, "kFriction": 5, "damping": 1e-7,
        "gravity": [0, -9.8, 0], "drag": 0.2,
        "maxTimeStep": 4e-4, "numIterations": 4, "detectionPeriod": 4,
        "detectionScaler": 1.1, "stepLimit": true, "velStepLimit": true,
      },
    }""")
    return tmp_path / "coil.json"


def read_obj_positions(path):
    return np.array([[float(x) for x in l.split()[1:4]] for l in open(path)
                     if l.startswith("v ")], np.float32)


@pytest.mark.gpu
def test_json_scene_cpp_and_python_agree(tmp_path):
    if not os.path.exists(PROG):
        pytest.skip("tests/cpp/json_scene not built (make -C tests/cpp)")
    js = make_scene(tmp_path)
    seconds = "0.05"
    out = subprocess.run([PROG, str(js), seconds, str(tmp_path / "out.obj")], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    cpp = read_obj_positions(tmp_path / "out.obj")

    sc = scene_io.build_from_json(js)
    sim = Sim(sc.n)
    sim.printErrors = False
    scenes.load_into(sim, sc)
    sim.minSeparation(reset=True)
    sim.advance(np.float32(seconds))
    sim.download()
    py = sim.verts["pos"].copy()
    assert cpp.shape == py.shape
    assert np.array_equal(cpp, py), "C++ facade and Python mirror differ"
    assert np.isfinite(py).all()
    pinned = sc.verts["invMass"] == 0
    assert pinned.sum() == 2 and np.array_equal(py[pinned], sc.verts["pos"][pinned])
    moved = np.linalg.norm(py - sc.verts["pos"], axis=1)
    print(f"{sc.n} vertices, {sim.stepCount()} substeps, max displacement {moved.max():.3e} m, "
          f"min separation {sim.minSeparation():.2f} x 2r")
    assert moved.max() > 1e-4  # it fell under gravity
    sim.close()
