"""ctypes wrapper of liboracle.so (the CPU oracle; test infrastructure only).

Build with ``make -C oracle``.  See yb_oracle.h for what each function restates.
"""
import ctypes
import os

import numpy as np

from yarnball_amd._lib import Params, ptr

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_P = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float
_PP = ctypes.POINTER(Params)

SIGS = {
    "orc_set_threads": (None, [_I]),
    "orc_get_threads": (_I, []),
    "orc_new": (_P, [_I]),
    "orc_free": (None, [_P]),
    "orc_load": (None, [_P, _P, _P, _P, _P, _P]),
    "orc_load64": (None, [_P, _P, _P, _P, _P]),
    "orc_store64": (None, [_P, _P, _P, _P, _P]),
    "orc_set_step_counter": (None, [_P, ctypes.c_int64]),
    "orc_substep": (_I, [_P, _PP, _I]),
    "orc_advance": (_I, [_P, _PP, _F, _F]),
    "orc_contacts": (None, [_P, _P, _P]),
    "orc_last_prologue": (None, [_P, _P, _P]),
    "orc_last_residual": (ctypes.c_double, [_P]),
    "orc_set_solver": (None, [_P, ctypes.c_double, _I]),
    "orc_last_sweeps": (_I, [_P]),
    "orc_last_force_norm0": (ctypes.c_double, [_P]),
    "orc_last_newton_residual": (ctypes.c_double, [_P]),
    "orc_last_update_residual": (ctypes.c_double, [_P]),
    "orc_set_residual_kind": (None, [_P, _I]),
    "orc_set_contacts": (None, [_P, _P, _P]),
    "orc_set_iterate": (None, [_P, _P, _P, _P]),
    "orc_seg_terms": (None, [_P, _PP, _I, _P, _P]),
    "orc_frame_update": (None, [_P, _P, _P]),
    "orc_vertex_terms": (None, [_P, _PP, _I, _P, _P]),
    "orc_force_norm2": (ctypes.c_double, [_P, _PP, _P]),
    "orc32_warm_start": (None, [_I, _PP, _P, _P, _P, _P, _P]),
    "orc32_aabb": (None, [_I, _PP, _P, _P, _P, _P]),
    "orc32_scene_box": (None, [_I, _P, _P]),
    "orc32_morton": (None, [_I, _P, _P, _P, _P]),
    "orc32_sort": (None, [_I, _P, _P]),
    "orc32_contacts": (_I, [_I, _PP, _P, _P, _P, _P, _P, _P, _P]),
    "orc_closest": (None, [_P, _P, _P, _P, _P]),
    "orc_barrier": (None, [ctypes.c_double, _P]),
    "orc_solve3": (None, [_P, _P, _P]),
    "orc_configure": (_I, [_I, _P, _P, _P, ctypes.c_double, _P, _P]),
    "orc_rotor_mul": (None, [_P, _P, _P]),
    "orc_rotor_rotate": (None, [_P, _P, _P]),
    "orc_from_to": (None, [_P, _P, _P]),
}

# liboracle32.so: the same stepper built with float arithmetic (-DORC_FP32), a
# CPU timing baseline only; it has no test helpers.
LIB32_PATH = os.path.join(_HERE, "liboracle32.so")
_TEST_ONLY = {"orc_closest", "orc_barrier", "orc_solve3", "orc_configure", "orc_rotor_mul",
              "orc_rotor_rotate", "orc_from_to"}
_libs = {}


def load(fp32=False):
    path = LIB32_PATH if fp32 else LIB_PATH
    if path not in _libs:
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run `make -C oracle`")
        lib = ctypes.CDLL(path)
        for k, (r, a) in SIGS.items():
            if fp32 and k in _TEST_ONLY:
                continue
            f = getattr(lib, k)
            f.restype = r
            f.argtypes = a
        _libs[path] = lib
    return _libs[path]


def _f64(a, shape):
    return np.ascontiguousarray(np.asarray(a, np.float64).reshape(shape))


def _sym(a6):
    """Packed symmetric {xx, yy, zz, xy, xz, yz} -> 3x3."""
    xx, yy, zz, xy, xz, yz = a6
    return np.array([[xx, xy, xz], [xy, yy, yz], [xz, yz, zz]])


class Oracle:
    """FP64 oracle holding one scene's state (fp32=True: the float build)."""

    def __init__(self, scene=None, n=None, fp32=False):
        self.lib = load(fp32)
        self.n = n if n is not None else scene.n
        self.o = self.lib.orc_new(self.n)
        if scene is not None:
            self.load_scene(scene)

    def load_scene(self, sc, lastVels=None):
        lv = np.zeros((self.n, 3), np.float32) if lastVels is None else np.ascontiguousarray(
            lastVels, np.float32)
        self.lib.orc_load(self.o, ptr(np.ascontiguousarray(sc.verts)),
                          ptr(np.ascontiguousarray(sc.vels, np.float32)), ptr(lv),
                          ptr(np.ascontiguousarray(sc.qs, np.float32)),
                          ptr(np.ascontiguousarray(sc.qRests, np.float32)))

    def load_state(self, verts, vels, qs, lastVels, qRests=None):
        """Load FP32 device-format state (for per-substep resync)."""
        self.lib.orc_load(self.o, ptr(np.ascontiguousarray(verts)),
                          ptr(np.ascontiguousarray(vels, np.float32)),
                          ptr(np.ascontiguousarray(lastVels, np.float32)),
                          ptr(np.ascontiguousarray(qs, np.float32)),
                          None if qRests is None else ptr(np.ascontiguousarray(qRests, np.float32)))

    def state(self):
        pos = np.zeros((self.n, 3))
        vel = np.zeros((self.n, 3))
        lv = np.zeros((self.n, 3))
        q = np.zeros((self.n, 4))
        self.lib.orc_store64(self.o, ptr(pos), ptr(vel), ptr(lv), ptr(q))
        return pos, vel, lv, q

    def load64(self, pos=None, vels=None, lastVels=None, qs=None):
        """Overwrite parts of the FP64 state (None keeps an array)."""
        arrs = [None if a is None else _f64(a, (self.n, k))
                for a, k in ((pos, 3), (vels, 3), (lastVels, 3), (qs, 4))]
        self.lib.orc_load64(self.o, *[None if a is None else ptr(a) for a in arrs])

    def set_step_counter(self, c):
        self.lib.orc_set_step_counter(self.o, int(c))

    def substep(self, params, detect):
        return self.lib.orc_substep(self.o, ctypes.byref(params), int(bool(detect)))

    def advance(self, params, h, maxH):
        return self.lib.orc_advance(self.o, ctypes.byref(params), float(h), float(maxH))

    def contacts(self):
        nc = np.zeros(self.n, np.int32)
        cols = np.zeros(self.n * 128, np.int32)
        self.lib.orc_contacts(self.o, ptr(nc), ptr(cols))
        return nc, cols.reshape(128, self.n)

    def prologue(self):
        lim = np.zeros(self.n)
        dxw = np.zeros((self.n, 3))
        self.lib.orc_last_prologue(self.o, ptr(lim), ptr(dxw))
        return lim, dxw

    def set_solver(self, tol, max_sweeps):
        self.lib.orc_set_solver(self.o, float(tol), int(max_sweeps))

    def last_sweeps(self):
        return int(self.lib.orc_last_sweeps(self.o))

    def residual(self):
        """Stopping residual of the last sweep of the last substep."""
        return float(self.lib.orc_last_residual(self.o))

    def newton_residual(self):
        """|f(dx_k)| / |f(dx_0)| of the last sweep's input iterate."""
        return float(self.lib.orc_last_newton_residual(self.o))

    def update_residual(self):
        return float(self.lib.orc_last_update_residual(self.o))

    def set_residual_kind(self, kind):
        """0: stop on the update residual, 1: on the Newton residual."""
        self.lib.orc_set_residual_kind(self.o, int(kind))

    # -- the sweep's own functions (closed-form and finite-difference tests) --
    def set_contacts(self, numCols, cols):
        """numCols (n,), cols (128, n) slot-major."""
        self.lib.orc_set_contacts(self.o, ptr(np.ascontiguousarray(numCols, np.int32)),
                                  ptr(np.ascontiguousarray(cols, np.int32)))

    def set_iterate(self, dx=None, y=None, step_limit=None):
        def f(a, shape):
            return None if a is None else ptr(_f64(a, shape))
        self.lib.orc_set_iterate(self.o, f(dx, (self.n, 3)), f(y, (self.n, 3)),
                                 f(step_limit, (self.n,)))

    def frame_update(self, dx):
        """One sweep's frame update (cosserat.cu:209-253) of the loaded frames
        at iterate dx (n x 3), without changing the state: (n x 4)."""
        q = np.zeros((self.n, 4))
        self.lib.orc_frame_update(self.o, ptr(_f64(dx, (self.n, 3))), ptr(q))
        return q

    def seg_terms(self, params, s, dx):
        out = np.zeros(18)
        self.lib.orc_seg_terms(self.o, ctypes.byref(params), int(s), ptr(_f64(dx, (self.n, 3))),
                               ptr(out))
        return {"f0": out[0:3], "f1": out[3:6], "H0": _sym(out[6:12]), "H1": _sym(out[12:18])}

    def vertex_terms(self, params, v, dx):
        out = np.zeros(16)
        self.lib.orc_vertex_terms(self.o, ctypes.byref(params), int(v),
                                  ptr(_f64(dx, (self.n, 3))), ptr(out))
        return {"f": out[0:3], "H": _sym(out[3:9]), "limit": out[9], "next": out[10:13],
                "resid": out[13:16]}

    def force_norm2(self, params, dx):
        return float(self.lib.orc_force_norm2(self.o, ctypes.byref(params),
                                              ptr(_f64(dx, (self.n, 3)))))

    def close(self):
        if self.o:
            self.lib.orc_free(self.o)
            self.o = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class orc32:
    """FP32 twin of the discrete stages."""

    @staticmethod
    def warm_start(params, invMass, vel3, lastVel3):
        lib = load()
        n = len(invMass)
        dx = np.zeros((n, 3), np.float32)
        y = np.zeros((n, 3), np.float32)
        lib.orc32_warm_start(n, ctypes.byref(params), ptr(np.ascontiguousarray(invMass, np.float32)),
                             ptr(np.ascontiguousarray(vel3, np.float32)),
                             ptr(np.ascontiguousarray(lastVel3, np.float32)), ptr(dx), ptr(y))
        return dx, y

    @staticmethod
    def aabb(params, pos3, flags, dx3):
        lib = load()
        n = len(flags)
        out = np.zeros((n, 6), np.float32)
        lib.orc32_aabb(n, ctypes.byref(params), ptr(np.ascontiguousarray(pos3, np.float32)),
                       ptr(np.ascontiguousarray(flags, np.uint32)),
                       ptr(np.ascontiguousarray(dx3, np.float32)), ptr(out))
        return out

    @staticmethod
    def scene_box(aabb6):
        lib = load()
        out = np.zeros(6, np.float32)
        lib.orc32_scene_box(len(aabb6), ptr(np.ascontiguousarray(aabb6, np.float32)), ptr(out))
        return out

    @staticmethod
    def morton(flags, aabb6, box6):
        lib = load()
        n = len(flags)
        out = np.zeros(n, np.uint32)
        lib.orc32_morton(n, ptr(np.ascontiguousarray(flags, np.uint32)),
                         ptr(np.ascontiguousarray(aabb6, np.float32)),
                         ptr(np.ascontiguousarray(box6, np.float32)), ptr(out))
        return out

    @staticmethod
    def sort(codes):
        lib = load()
        n = len(codes)
        out = np.zeros(n, np.int32)
        lib.orc32_sort(n, ptr(np.ascontiguousarray(codes, np.uint32)), ptr(out))
        return out

    @staticmethod
    def contacts(params, pos3, flags, cid, dx3, aabb6):
        lib = load()
        n = len(flags)
        nc = np.zeros(n, np.int32)
        cols = np.zeros(n * 128, np.int32)
        err = lib.orc32_contacts(n, ctypes.byref(params), ptr(np.ascontiguousarray(pos3, np.float32)),
                                 ptr(np.ascontiguousarray(flags, np.uint32)),
                                 ptr(np.ascontiguousarray(cid, np.int32)),
                                 ptr(np.ascontiguousarray(dx3, np.float32)),
                                 ptr(np.ascontiguousarray(aabb6, np.float32)), ptr(nc), ptr(cols))
        return nc, cols.reshape(128, n), err
