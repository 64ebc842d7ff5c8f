"""yb_mlp_forward_device (the allocation-free forward pass: device buffers,
the caller's stream, a static per-GPU scratch) against yb_mlp_forward (host
buffers, its own allocations): bit-identical outputs on two calls in a row,
the second with fewer rows and different inputs; sizes beyond the scratch are
rejected with a clear error.  The device buffers and the stream come from the
HIP runtime the library itself links (libamdhip64.so.7, already loaded with
it), so the test never brings a second HIP runtime into the process."""
import ctypes
import os

import numpy as np
import pytest

from yarnball_amd import _lib
from yarnball_amd._lib import check, ptr
from yarnball_amd.learned import ISTD, MEAN, OSCALE

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def features(params, n, seed):
    """Raw feature rows whose normalised values are standard normal."""
    rng = np.random.default_rng(seed)
    mean = params[MEAN:ISTD].astype(np.float64)
    istd = params[ISTD:OSCALE].astype(np.float64)
    scale = np.where(istd > 0, 1.0 / np.where(istd > 0, istd, 1.0), 1.0)
    return (mean + rng.normal(size=(n, 48)) * scale).astype(np.float32)


class Hip:
    """hipMalloc / hipMemcpy / streams of the library's own HIP runtime."""

    H2D, D2H = 1, 2

    def __init__(self):
        self.rt = ctypes.CDLL("libamdhip64.so.7")
        self.bufs = []

    def ok(self, e, what):
        assert e == 0, f"{what}: hip error {e}"

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        self.ok(self.rt.hipMalloc(ctypes.byref(p), ctypes.c_size_t(a.nbytes)), "hipMalloc")
        self.bufs.append(p)
        self.ok(self.rt.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes),
                                  self.H2D), "hipMemcpy")
        return p

    def download(self, p, a):
        self.ok(self.rt.hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), p, ctypes.c_size_t(a.nbytes),
                                  self.D2H), "hipMemcpy")
        return a

    def close(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


@pytest.mark.gpu
def test_forward_device_equals_the_allocating_forward():
    lib = _lib.load()
    hip = Hip()
    params = np.ascontiguousarray(
        np.load(os.path.join(ROOT, "yarnball_amd", "models", "warmstart.npy"), allow_pickle=False),
        np.float32)
    dP = hip.upload(params)
    stream = ctypes.c_void_p()
    hip.ok(hip.rt.hipStreamCreate(ctypes.byref(stream)), "hipStreamCreate")
    for n, seed in ((40000, 1), (3001, 2)):  # the second call is smaller
        X = features(params, n, seed)
        want = np.zeros((n, 3), np.float32)
        check(lib.yb_mlp_forward(ptr(X), n, ptr(params), ptr(want)))
        dX = hip.upload(X)
        dY = hip.upload(np.full((n, 3), np.nan, np.float32))
        print(f"[no-alloc] n = {n}: yb_mlp_forward_device", flush=True)
        check(lib.yb_mlp_forward_device(dX, n, dP, dY, stream))
        hip.ok(hip.rt.hipStreamSynchronize(stream), "hipStreamSynchronize")
        got = hip.download(dY, np.zeros((n, 3), np.float32))
        assert np.all(np.isfinite(got))
        assert np.array_equal(got, want), f"n = {n}: max diff {np.abs(got - want).max()}"
    rc = lib.yb_mlp_forward_device(dX, 65536 + 1, dP, dY, stream)
    assert rc == _lib.YB_ERR_INVALID_ARGUMENT and "YB_MLP_NOALLOC_MAX_ROWS" in _lib.last_message()
    assert lib.yb_mlp_forward_device(None, 10, dP, dY, stream) == _lib.YB_ERR_INVALID_ARGUMENT
    hip.ok(hip.rt.hipStreamDestroy(stream), "hipStreamDestroy")
    hip.close()
