"""Scene files through the library's host-side entry points: the JSON scene
description (buildFromJSON), the curve readers (.poly, .obj, .bcc) with
Catmull-Rom resampling, endpoint gluing, and the BCC / OBJ writers
(KittenEngine/YarnBall/io/jsonBuilder.cpp, reader.cpp, resample.h,
fileExport.cpp)."""
import ctypes
import os

import numpy as np

from . import _lib
from ._lib import Params, VERTEX_DTYPE, check, ptr
from .scenes import HostScene


def _take(handle, name):
    lib = _lib.load()
    try:
        n = ctypes.c_int(0)
        check(lib.yb_scene_size(handle, ctypes.byref(n)))
        n = n.value
        p = Params()
        maxH = ctypes.c_float(0)
        verts = np.zeros(n, VERTEX_DTYPE)
        vels = np.zeros((n, 3), np.float32)
        qs = np.zeros((n, 4), np.float32)
        qRests = np.zeros((n, 4), np.float32)
        inv0 = np.zeros(n, np.float32)
        check(lib.yb_scene_get(handle, ctypes.byref(p), ctypes.byref(maxH), ptr(verts), ptr(vels),
                               ptr(qs), ptr(qRests), ptr(inv0)))
    finally:
        lib.yb_scene_free(handle)
    return HostScene(name, verts, vels, qs, qRests, p, float(maxH.value), inv0,
                     init_pos=verts["pos"].astype(np.float64).copy())


def build_from_json(path):
    """A configured scene from a JSON description (buildFromJSON)."""
    h = ctypes.c_void_p()
    check(_lib.load().yb_build_from_json(os.fsencode(path), ctypes.byref(h)), "buildFromJSON")
    return _take(h, os.path.basename(path))


def read_curves(path, targetSegLen, transform=None, breakUpClosedCurves=False,
                allowResample=True):
    """An unconfigured scene from a curve file (readFromPoly / OBJ / BCC).
    transform: 3x4 affine map (rows), default the reference's 0.01 scale."""
    tf = None if transform is None else np.ascontiguousarray(transform, np.float32).reshape(12)
    h = ctypes.c_void_p()
    check(_lib.load().yb_read_curves(os.fsencode(path), float(targetSegLen),
                                     None if tf is None else ptr(tf), int(bool(breakUpClosedCurves)),
                                     int(bool(allowResample)), ctypes.byref(h)), "readCurves")
    return _take(h, os.path.basename(path))


def glue_endpoints(verts, searchRadius):
    """Sim::glueEndpoints on a host vertex array (modified in place)."""
    check(_lib.load().yb_glue_endpoints(len(verts), ptr(verts), float(searchRadius)))


def export_bcc(path, verts, asPolyline=False):
    check(_lib.load().yb_export_bcc(os.fsencode(path), len(verts), ptr(verts),
                                    int(bool(asPolyline))), "exportToBCC")


def export_obj(path, verts):
    check(_lib.load().yb_export_obj(os.fsencode(path), len(verts), ptr(verts)), "exportToOBJ")
