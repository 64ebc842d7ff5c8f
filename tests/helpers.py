"""Shared helpers of the parity tests: run the device stepper and the CPU
oracle side by side from identical state."""
import ctypes

import numpy as np

from yarnball_amd import Sim, scenes
from yarnball_amd._lib import check, ptr


def gpu_sim(sc):
    s = Sim(sc.n)
    scenes.load_into(s, sc)
    return s


def gpu_state(sim):
    sim.download()
    lv = np.zeros((sim.numVerts, 3), np.float32)
    check(sim._lib.yb_debug_read_last_vels(sim.handle, ptr(lv)))
    return sim.verts.copy(), sim.vels.copy(), sim.qs.copy(), lv


def gpu_contacts(sim):
    n = sim.numVerts
    nc = np.zeros(n, np.int32)
    cols = np.zeros(n * 128, np.int32)
    check(sim._lib.yb_debug_read(sim.handle, None, None, None, None, ptr(nc), ptr(cols)))
    return nc, cols.reshape(128, n)


def contact_sets(nc, cols):
    """Per-segment contact lists as sorted tuples (order-free comparison)."""
    return [tuple(sorted(cols[: nc[i], i].tolist())) for i in range(len(nc))]


def rel_err(a, b):
    """Normwise relative error max|a-b| / max|b|."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
