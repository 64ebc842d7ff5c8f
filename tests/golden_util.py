"""Loading the golden fixtures of tests/golden (written by make_golden.py)."""
import ctypes
import os

import numpy as np

from yarnball_amd._lib import Params, VERTEX_DTYPE

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        g = {k: z[k] for k in z.files}
    g["verts"] = g["verts"].view(VERTEX_DTYPE).copy()
    p = Params()
    raw = g["params"].tobytes()
    assert len(raw) == ctypes.sizeof(Params)
    ctypes.memmove(ctypes.byref(p), raw, len(raw))
    g["params"] = p
    return g


def scene(g):
    """A HostScene-like object for scenes.load_into / Oracle(scene)."""
    from yarnball_amd.scenes import HostScene

    return HostScene("golden", g["verts"].copy(), g["vels"].copy(), g["qs"].copy(),
                     g["qRests"].copy(), g["params"].copy(), float(g["maxH"]),
                     np.zeros(len(g["verts"]), np.float32))
