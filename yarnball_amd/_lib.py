"""ctypes binding of libyarnball_hip.so (the C ABI declared in include/yarnball_hip.h).

The library is built in-tree by ``python -m yarnball_amd.build`` (or
``__graft_entry__.build()``).  Loading fails loudly when it is missing: there is
no CPU fallback for the stepper.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libyarnball_hip.so")

MAX_COLLISIONS_PER_SEGMENT = 128
MAX_VERTICES = 1 << 28

YB_OK = 0
YB_ERROR_MAX_COLLISIONS_PER_SEGMENT_EXCEEDED = 1
YB_ERROR_TRAVERSAL_STACK_EXCEEDED = 3
YB_WARNING_SEGMENT_INTERPENETRATION = 2
YB_ERR_INVALID_ARGUMENT = -1
YB_ERR_NOT_CONFIGURED = -2
YB_ERR_HIP = -3
YB_ERR_VALIDATION = -4
YB_ERR_NO_DEVICE = -5

FLAG_HAS_PREV = 1
FLAG_HAS_NEXT = 2
FLAG_HAS_NEXT_ORIENTATION = 4
FLAG_FIX_ORIENTATION = 8
FLAG_COLLIDING = 16

# 32-byte vertex record, YarnBall.h:52-61.
VERTEX_DTYPE = np.dtype(
    [
        ("pos", "<f4", (3,)),
        ("invMass", "<f4"),
        ("lRest", "<f4"),
        ("kStretch", "<f4"),
        ("connectionIndex", "<i4"),
        ("flags", "<u4"),
    ]
)
assert VERTEX_DTYPE.itemsize == 32


class SolverStats(ctypes.Structure):
    _fields_ = [
        ("substeps", ctypes.c_int64),
        ("sweeps", ctypes.c_int64),
        ("maxSweepsUsed", ctypes.c_int32),
        ("notConverged", ctypes.c_int32),
        ("lastResidual", ctypes.c_float),
        ("maxResidual", ctypes.c_float),
        ("lastNewtonResidual", ctypes.c_float),
        ("maxNewtonResidual", ctypes.c_float),
        ("sumNewtonResidual", ctypes.c_double),
        ("sceneSubsteps", ctypes.c_int64),
        ("sceneSweeps", ctypes.c_int64),
        ("newtonSubsteps", ctypes.c_int64),
    ]


class Driver(ctypes.Structure):
    """yb_driver: scenario animation of the pinned vertices (main.cpp:54-121)."""
    _fields_ = [("kind", ctypes.c_int32), ("center", ctypes.c_float * 3), ("speed", ctypes.c_float),
                ("start", ctypes.c_float), ("duration", ctypes.c_float),
                ("release", ctypes.c_float), ("side", ctypes.c_int32)]


DRIVE_TWIST, DRIVE_PULL = 1, 2


class Params(ctypes.Structure):
    """yb_params: the scalar fields of MetaData (YarnBall.h:63-114)."""

    _fields_ = [
        ("gravity", ctypes.c_float * 3),
        ("numItr", ctypes.c_int32),
        ("worldFloorNormal", ctypes.c_float * 3),
        ("worldFloorPos", ctypes.c_float),
        ("h", ctypes.c_float),
        ("lastH", ctypes.c_float),
        ("time", ctypes.c_float),
        ("numVerts", ctypes.c_int32),
        ("damping", ctypes.c_float),
        ("drag", ctypes.c_float),
        ("frictionCoeff", ctypes.c_float),
        ("kCollision", ctypes.c_float),
        ("kFriction", ctypes.c_float),
        ("detectionRadius", ctypes.c_float),
        ("scaledDetectionRadius", ctypes.c_float),
        ("radius", ctypes.c_float),
        ("accelerationRatio", ctypes.c_float),
        ("barrierThickness", ctypes.c_float),
        ("detectionScaler", ctypes.c_float),
        ("bvhRebuildPeriod", ctypes.c_float),
        ("detectionPeriod", ctypes.c_int32),
        ("maxSegLen", ctypes.c_float),
        ("minSegLen", ctypes.c_float),
        ("useStepSizeLimit", ctypes.c_int32),
        ("useVelocityRadius", ctypes.c_int32),
    ]

    def copy(self):
        c = Params()
        ctypes.memmove(ctypes.byref(c), ctypes.byref(self), ctypes.sizeof(Params))
        return c


_P = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float

# name -> (restype, argtypes); every symbol the C header declares.
SIGNATURES = {
    "yb_default_params": (_I, [ctypes.POINTER(Params), _I, _F]),
    "yb_default_vertices": (_I, [_I, _P, _P, _P, _P]),
    "yb_configure_host": (_I, [_I, _P, _P, _P, _F, _P, ctypes.POINTER(Params)]),
    "yb_set_kstretch_host": (_I, [_I, _P, _F]),
    "yb_set_kbend_host": (_I, [_I, _P, _P, _F]),
    "yb_validate_params": (_I, [ctypes.POINTER(Params)]),
    "yb_create": (_I, [_I, ctypes.POINTER(_P)]),
    "yb_destroy": (None, [_P]),
    "yb_upload": (_I, [_P, _P, _P, _P, _P]),
    "yb_download": (_I, [_P, _P, _P, _P]),
    "yb_zero_velocities": (_I, [_P]),
    "yb_advance": (_I, [_P, ctypes.POINTER(Params), _F, _F, ctypes.POINTER(_I)]),
    "yb_step_count": (ctypes.c_int64, [_P]),
    "yb_bounds": (_I, [_P, _P]),
    "yb_collision_counts": (_I, [_P, _P]),
    "yb_last_message": (ctypes.c_char_p, []),
    "yb_debug_prologue": (_I, [_P, ctypes.POINTER(Params), _F, _I, _I]),
    "yb_debug_read": (_I, [_P, _P, _P, _P, _P, _P, _P]),
    "yb_debug_tree": (_I, [_P, _P, _P, _P]),
    "yb_debug_write_dx": (_I, [_P, _P]),
    "yb_debug_strand": (_I, [_P, _P]),
    "yb_debug_read_last_vels": (_I, [_P, _P]),
    "yb_debug_write_last_vels": (_I, [_P, _P]),
    "yb_debug_read_mlp": (_I, [_P, _P, _P]),
    "yb_profile": (_I, [_P, ctypes.POINTER(Params), _F, _I, _P, _P]),
    "yb_set_solver": (_I, [_P, _F, _I, _I, _I]),
    "yb_set_residual_kind": (_I, [_P, _I]),
    "yb_set_newton_residual": (_I, [_P, _I]),
    "yb_debug_set_query_stack": (_I, [_P, _I]),
    "yb_debug_query_counts": (_I, [_P, ctypes.POINTER(Params), _F, _P]),
    "yb_debug_timeline": (_I, [_P, _I]),
    "yb_debug_timeline_read": (_I, [_P, _P, _I, _P]),
    "yb_mlp_forward_device": (_I, [_P, ctypes.c_int64, _P, _P, _P]),
    "yb_min_separation": (_I, [_P, _P, _I]),
    "yb_mlp_load": (_I, [_P, _P, _I]),
    "yb_set_warm_start": (_I, [_P, _I]),
    "yb_set_learned_start_cost": (_I, [_P, _F]),
    "yb_set_strand_schedule": (_I, [_P, _I, _I, _I]),
    "yb_mlp_counters": (_I, [_P, _P, _I]),
    "yb_warm_start_substeps": (_I, [_P, _P, _I]),
    "yb_set_graphs": (_I, [_P, _I]),
    "yb_set_scenes": (_I, [_P, _I, _P]),
    "yb_set_launch_config": (_I, [_P, _I, _I]),
    "yb_driver_default": (_I, [_I, _P]),
    "yb_driver_init": (_I, [_P, _P, _P]),
    "yb_driver_frame": (_I, [_P, ctypes.POINTER(Params), _P, _F, _F]),
    "yb_build_from_json": (_I, [ctypes.c_char_p, _P]),
    "yb_read_curves": (_I, [ctypes.c_char_p, _F, _P, _I, _I, _P]),
    "yb_scene_size": (_I, [_P, _P]),
    "yb_scene_get": (_I, [_P, ctypes.POINTER(Params), _P, _P, _P, _P, _P, _P]),
    "yb_scene_free": (None, [_P]),
    "yb_glue_endpoints": (_I, [_I, _P, _F]),
    "yb_export_bcc": (_I, [ctypes.c_char_p, _I, _P, _I]),
    "yb_export_obj": (_I, [ctypes.c_char_p, _I, _P]),
    "yb_get_launch_config": (_I, [_P, _P, _P]),
    "yb_set_tile_config": (_I, [_P, _I, _I, _I]),
    "yb_get_tile_config": (_I, [_P, _P, _P, _P]),
    "yb_set_contact_skin": (_I, [_P, _F, _F]),
    "yb_get_contact_skin": (_I, [_P, _P, _P]),
    "yb_contact_skin_stats": (_I, [_P, _P, _I]),
    "yb_debug_read_skin": (_I, [_P, _P, _P, _P]),
    "yb_record": (_I, [_P, ctypes.c_int64]),
    "yb_record_enable": (_I, [_P, _I]),
    "yb_record_fetch": (_I, [_P, _P, _P, ctypes.POINTER(ctypes.c_int64)]),
    "yb_mlp_train": (_I, [_P, _P, ctypes.c_int64, _P, _I, _I, _F, ctypes.c_uint32, _P]),
    "yb_mlp_grad": (_I, [_P, _P, ctypes.c_int64, _P, _P, _P]),
    "yb_mlp_forward": (_I, [_P, ctypes.c_int64, _P, _P]),
    "yb_solver_stats": (_I, [_P, ctypes.POINTER(SolverStats), _I]),
}

STAGES = ["init", "aabb", "rebuild", "refit", "query", "step_limit", "sweep", "mlp", "end",
          "residual"]

MLP_INPUTS = 48
MLP_HIDDEN = 128
MLP_OUTPUTS = 3
MLP_TRAINABLE = 23171
MLP_PARAMS = 23270

_lib = None


def load():
    """Load libyarnball_hip.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `make -C yarnball_amd/csrc` "
            "(the stepper has no CPU fallback)"
        )
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def ptr(a):
    """Raw data pointer of a C-contiguous numpy array (or None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays passed to the C ABI must be contiguous"
    return a.ctypes.data_as(ctypes.c_void_p)


def last_message():
    return load().yb_last_message().decode(errors="replace")


class YarnBallError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def check(code, where=""):
    if code != YB_OK:
        msg = last_message() or where
        raise YarnBallError(code, msg)
    return code
