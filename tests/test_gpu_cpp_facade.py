"""The reference's README program, compiled against the drop-in C++ header,
runs on the device and matches the Python mirror bit for bit (both drive
the same C ABI)."""
import os
import subprocess

import numpy as np
import pytest

from yarnball_amd import Sim

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "readme_scene")


@pytest.mark.gpu
def test_readme_program_matches_python_mirror():
    assert os.path.exists(BIN), "build tests/cpp first (__graft_entry__.build())"
    out = subprocess.run([BIN, "0.25"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.splitlines()
    assert lines[0] == "steps 250", out.stdout[-500:]
    cpp = np.array([[float(x) for x in l.split()] for l in lines[1:65]], np.float32)
    assert "errors 2" in lines, out.stdout[-800:]

    n = 64
    seg = 0.002
    s = Sim(n)
    for i in range(32):
        x = (np.float32(0.00002) * np.float32(i)) * np.float32(i - 16) + np.float32(seg)
        s.verts["pos"][i] = (x, -np.float32(seg) * np.float32(i), 0)
        s.verts["pos"][i + 32] = (-x, -np.float32(seg) * np.float32(i), 0)
    s.verts["invMass"][[0, 32, 63]] = 0
    s.verts["flags"][31] = 0
    s.meta.kCollision = 1e-7
    s.configure()
    s.setKBend(1e-8)
    s.setKStretch(2e-2)
    s.maxH = 1e-3
    s.upload()
    s.advance(0.25)
    s.download()
    assert s.stepCount() == 250
    assert np.array_equal(s.verts["pos"], cpp)
    s.close()
