"""Learned warm start: parameter vector, data recording, device training.

The model (48 -> 128 -> 128 -> 3, ReLU) predicts each vertex's correction
(dx* - dx0)/L from 48 local features (listed in csrc/mlp.hip).  Training data
are recorded from the plain residual-targeted solver: the features at the
warm start of every substep and the converged correction at its end.
Training, forward and backward run on the device (csrc/mlp.hip); this module
only moves arrays and chooses hyper-parameters.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import MLP_HIDDEN as H, MLP_INPUTS as F, MLP_OUTPUTS as O, MLP_PARAMS, MLP_TRAINABLE
from ._lib import check, ptr

# offsets in the parameter vector
W1, B1 = 0, F * H
W2 = B1 + H
B2 = W2 + H * H
W3 = B2 + H
B3 = W3 + H * O
MEAN = B3 + O
ISTD = MEAN + F
OSCALE = ISTD + F
assert OSCALE + O == MLP_PARAMS and B3 + O == MLP_TRAINABLE


def unpack(p):
    """Split a parameter vector into (W1, b1, W2, b2, W3, b3, mean, istd, oscale)."""
    p = np.asarray(p, np.float32)
    return (p[W1:B1].reshape(F, H), p[B1:W2], p[W2:B2].reshape(H, H), p[B2:W3],
            p[W3:B3].reshape(H, O), p[B3:MEAN], p[MEAN:ISTD], p[ISTD:OSCALE], p[OSCALE:])


CPREV = slice(36, 39)  # feature columns of the previous converged correction


def residual_targets(X, C):
    """Network targets for recorded corrections C: the model adds its output to
    the previous substep's correction (features 36-38)."""
    return (np.asarray(C, np.float32) - np.asarray(X, np.float32)[:, CPREV]).astype(np.float32)


def init_params(X, T, seed=0):
    """He-initialised weights plus normalisation constants from the data."""
    rng = np.random.default_rng(seed)
    p = np.zeros(MLP_PARAMS, np.float32)
    p[W1:B1] = rng.normal(0, np.sqrt(2.0 / F), F * H)
    p[W2:B2] = rng.normal(0, np.sqrt(2.0 / H), H * H)
    p[W3:B3] = rng.normal(0, np.sqrt(1.0 / H), H * O)
    mean = X.mean(axis=0)
    std = X.std(axis=0)
    p[MEAN:ISTD] = mean
    p[ISTD:OSCALE] = np.where(std > 1e-12, 1.0 / np.maximum(std, 1e-12), 0.0)
    p[OSCALE:] = np.maximum(np.sqrt((T.astype(np.float64) ** 2).mean(axis=0)), 1e-12)
    return p


def record(sim, capacity):
    check(sim._lib.yb_record(sim.handle, int(capacity)))


def record_enable(sim, on):
    check(sim._lib.yb_record_enable(sim.handle, int(bool(on))))


def fetch(sim):
    n = ctypes.c_int64(0)
    check(sim._lib.yb_record_fetch(sim.handle, None, None, ctypes.byref(n)))
    X = np.zeros((n.value, F), np.float32)
    T = np.zeros((n.value, O), np.float32)
    check(sim._lib.yb_record_fetch(sim.handle, ptr(X), ptr(T), ctypes.byref(n)))
    return X, T


def train(X, T, params, epochs=8, batch=16384, lr=1e-3, seed=0):
    """Device training (Adam); returns (params, per-epoch mean loss)."""
    lib = _lib.load()
    p = np.ascontiguousarray(params, np.float32).copy()
    hist = np.zeros(epochs, np.float32)
    check(lib.yb_mlp_train(ptr(np.ascontiguousarray(X, np.float32)),
                           ptr(np.ascontiguousarray(T, np.float32)), len(X), ptr(p), int(epochs),
                           int(batch), float(lr), int(seed), ptr(hist)), "train")
    return p, hist


def forward(X, params):
    lib = _lib.load()
    Y = np.zeros((len(X), O), np.float32)
    check(lib.yb_mlp_forward(ptr(np.ascontiguousarray(X, np.float32)), len(X),
                             ptr(np.ascontiguousarray(params, np.float32)), ptr(Y)))
    return Y


def grad(X, T, params):
    lib = _lib.load()
    g = np.zeros(MLP_TRAINABLE, np.float32)
    loss = ctypes.c_float(0)
    check(lib.yb_mlp_grad(ptr(np.ascontiguousarray(X, np.float32)),
                          ptr(np.ascontiguousarray(T, np.float32)), len(X),
                          ptr(np.ascontiguousarray(params, np.float32)), ptr(g),
                          ctypes.byref(loss)))
    return g, loss.value


NO_FRAME_REFRESH = 0x100
NO_SMOOTHING = 0x200
NO_FALLBACK = 0x400
STRAND_SOLVE = 0x800


def use(sim, params, frame_refresh=True, smooth=True, fallback=True, strand=False):
    """Load a model into a simulation and switch its warm start to learned.

    fallback: in residual-targeted mode, start each substep from whichever of
    the learned and the plain warm start has needed fewer sweeps lately.
    strand: follow the warm start with the strand solve (see strand())."""
    p = np.ascontiguousarray(params, np.float32)
    check(sim._lib.yb_mlp_load(sim.handle, ptr(p), MLP_PARAMS))
    mode = (1 | (0 if frame_refresh else NO_FRAME_REFRESH) | (0 if smooth else NO_SMOOTHING)
            | (0 if fallback else NO_FALLBACK) | (STRAND_SOLVE if strand else 0))
    check(sim._lib.yb_set_warm_start(sim.handle, mode))


def strand_schedule(sim, every=0, pair_iters=2, first=0):
    """Strand-step schedule: every > 0 (residual mode) repeats the step before
    sweeps first, first + every, ... (first 0: first = every), at most 16 times
    per substep; pair_iters boundary-pair iterations per step."""
    check(sim._lib.yb_set_strand_schedule(sim.handle, int(every), int(pair_iters), int(first)))


def strand(sim, frame_refresh=True, every=0, pair_iters=2, first=0):
    """Plain warm start followed by one Newton step solved along whole strands
    (each local Hessian's curvature along the vertex's force on the
    diagonal, stretch coupling between neighbours): it removes the smooth
    error modes that block-Jacobi sweeps damp slowly on long stiff strands.
    No network.  every > 0 (residual mode): repeat the step before sweeps
    first, first + every, ... (the 1M knit: every=4, first=2)."""
    strand_schedule(sim, every, pair_iters, first)
    check(sim._lib.yb_set_warm_start(
        sim.handle, STRAND_SOLVE | (0 if frame_refresh else NO_FRAME_REFRESH)))


def start_cost(sim, sweeps):
    """The chooser's offset: run the learned start only while its sweep
    average plus `sweeps` (its prologue's extra time, in sweeps) is at most
    the plain start's."""
    check(sim._lib.yb_set_learned_start_cost(sim.handle, float(sweeps)))


def warm_substeps(sim, reset=False):
    """Substeps started from the session's warm start and from the plain fallback."""
    out = np.zeros(2, np.int64)
    check(sim._lib.yb_warm_start_substeps(sim.handle, ptr(out), int(bool(reset))))
    return {"learned": int(out[0]), "plain": int(out[1])}


def plain(sim):
    check(sim._lib.yb_set_warm_start(sim.handle, 0))


def temporal(sim, frame_refresh=True):
    """Warm start from the previous substep's converged correction (no network)."""
    check(sim._lib.yb_set_warm_start(sim.handle, 2 | (0 if frame_refresh else NO_FRAME_REFRESH)))


def counters(sim, reset=False):
    out = np.zeros(2, np.int64)
    check(sim._lib.yb_mlp_counters(sim.handle, ptr(out), int(bool(reset))))
    return {"applied": int(out[0]), "fallback": int(out[1])}
