"""The C ABI library loads without a GPU, exports every symbol the header
declares, and its host-side set-up math behaves like the reference's."""
import ctypes
import os
import re

import numpy as np
import pytest

from yarnball_amd import Sim, _lib, scenes
from yarnball_amd._lib import Params, VERTEX_DTYPE, check, ptr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "yarnball_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(yb_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s
        assert s in _lib.SIGNATURES, f"{s} missing from the ctypes binding"


def test_vertex_record_is_32_bytes():
    assert VERTEX_DTYPE.itemsize == 32
    assert ctypes.sizeof(Params) == 29 * 4  # 29 four-byte fields, no padding


def test_constructor_defaults():
    # YarnBall.cu:14-50
    s = Sim(10)
    m = s.meta
    assert tuple(m.gravity) == pytest.approx((0, -9.8, 0))
    assert m.drag == pytest.approx(0.2) and m.damping == pytest.approx(1e-6)
    assert m.radius == pytest.approx(1e-4) and m.barrierThickness == pytest.approx(8e-4)
    assert m.kCollision == pytest.approx(1e-5) and m.detectionScaler == pytest.approx(1.2)
    assert m.frictionCoeff == pytest.approx(0.1) and m.kFriction == pytest.approx(5)
    assert m.detectionPeriod == 1 and m.numItr == 8
    assert m.bvhRebuildPeriod == pytest.approx(0.1)
    assert np.all(s.verts["invMass"] == 1) and np.all(s.verts["kStretch"] == 100)
    assert np.all(s.verts["connectionIndex"] == -1)
    assert np.all(s.verts["flags"][:-1] == 2) and s.verts["flags"][-1] == 0
    assert np.allclose(s.qRests, [0, 0, 0, 5])


def test_errors_like_the_reference():
    with pytest.raises(RuntimeError, match="Too little vertices"):
        Sim(2)
    s = Sim(5)
    with pytest.raises(RuntimeError, match="Must call configure"):
        s.setKBend(1e-8)
    # vertex 2 isolated: dangling segment (YarnBall.cu:115-116)
    s.verts["pos"] = np.arange(15, dtype=np.float32).reshape(5, 3)
    s.verts["flags"][1] = 0
    s.verts["flags"][2] = 0
    with pytest.raises(RuntimeError, match="Dangling segment"):
        s.configure()
    s = Sim(4)
    s.verts["pos"] = 0.0
    with pytest.raises(RuntimeError, match="0 length segment"):
        s.configure()


def test_create_rejects_sizes_outside_the_limits():
    # Both checks come before any device call, so they hold without a GPU.
    lib = _lib.load()
    h = ctypes.c_void_p()
    for n, msg in ((2, "Too little vertices"), (_lib.MAX_VERTICES + 1, "YB_MAX_VERTICES")):
        assert lib.yb_create(n, ctypes.byref(h)) == _lib.YB_ERR_INVALID_ARGUMENT
        assert not h.value
        assert msg in lib.yb_last_message().decode()


def configured(pos, breaks=(), pinned=(), density=1e-3):
    lib = _lib.load()
    n = len(pos)
    p = Params()
    lib.yb_default_params(ctypes.byref(p), n, 1e-3)
    v = np.zeros(n, VERTEX_DTYPE)
    vel = np.zeros((n, 3), np.float32)
    q = np.zeros((n, 4), np.float32)
    qr = np.zeros((n, 4), np.float32)
    lib.yb_default_vertices(n, ptr(v), ptr(vel), ptr(q), ptr(qr))
    v["pos"] = pos
    v["flags"][list(breaks)] = 0
    v["invMass"][list(pinned)] = 0
    inv = np.zeros(n, np.float32)
    check(lib.yb_configure_host(n, ptr(v), ptr(q), ptr(qr), density, ptr(inv), ctypes.byref(p)))
    return v, q, qr, inv, p


def qmat(q):
    x, y, z, w = q / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def test_configure_straight_rod_closed_form():
    d = np.array([1.0, 2.0, -0.5])
    d /= np.linalg.norm(d)
    L = 0.003
    pos = (np.arange(12)[:, None] * L * d[None, :]).astype(np.float32)
    v, q, qr, inv, p = configured(pos, density=2e-3)
    # every frame maps e1 onto the rod direction and they are all equal
    for i in range(11):
        assert np.allclose(qmat(q[i].astype(np.float64)) @ [1, 0, 0], d, atol=1e-6)
        assert np.allclose(q[i], q[0], atol=1e-6)
    # rest Darboux rotor is the identity scaled by |qRest| = 5 (YarnBall.cu:157)
    assert np.allclose(qr[:11], [0, 0, 0, 5], atol=1e-5)
    assert np.allclose(v["lRest"][:11], L, rtol=1e-5)
    # mass = density * (l_prev + l_next) / 2; ends get half
    m = 1.0 / inv
    assert m[0] == pytest.approx(2e-3 * L / 2, rel=1e-5)
    assert m[5] == pytest.approx(2e-3 * L, rel=1e-5)
    assert m.sum() == pytest.approx(2e-3 * 11 * L, rel=1e-5)
    # flags: interior = prev|next|nextOrientation
    fl = v["flags"]
    assert fl[0] == 2 | 4 and fl[5] == 1 | 2 | 4 and fl[10] == 1 | 2 and fl[11] == 1
    assert p.minSegLen == pytest.approx(L, rel=1e-5) and p.maxSegLen == pytest.approx(L, rel=1e-5)


def test_configure_breaks_and_pins():
    pos = np.zeros((8, 3), np.float32)
    pos[:, 0] = np.arange(8) * 0.01
    v, q, qr, inv, p = configured(pos, breaks=[3], pinned=[0, 7])
    fl = v["flags"]
    assert fl[3] & 2 == 0 and fl[3] & 8 and fl[4] & 1 == 0
    assert fl[2] & 4 == 0  # segment 2 has no next segment across the break
    assert inv[0] == 0 and inv[7] == 0 and inv[1] > 0


def test_configure_matches_fp64_oracle():
    from oracle import load as oload

    sc = scenes.twisted_cable(strands=7, per=64)
    lib = _lib.load()
    # redo configure from raw positions on both sides
    n = sc.n
    v32 = np.zeros(n, VERTEX_DTYPE)
    v32["pos"] = sc.verts["pos"]
    v32["invMass"] = (sc.verts["invMass"] != 0).astype(np.float32)
    v32["flags"] = np.where(np.arange(n) % 64 == 63, 0, 2)
    v32["connectionIndex"] = -1
    v64 = v32.copy()
    q = np.zeros((n, 4), np.float32)
    qr = np.tile(np.array([0, 0, 0, 5], np.float32), (n, 1))
    qr64 = qr.astype(np.float64)
    inv = np.zeros(n, np.float32)
    p = Params()
    lib.yb_default_params(ctypes.byref(p), n, 1e-3)
    check(lib.yb_configure_host(n, ptr(v32), ptr(q), ptr(qr), 1e-3, ptr(inv), ctypes.byref(p)))
    q64 = np.zeros((n, 4))
    l64 = np.zeros(n)
    im64 = np.zeros(n)
    assert oload().orc_configure(n, ptr(v64), ptr(q64), ptr(qr64), 1e-3, ptr(l64), ptr(im64)) == 0
    assert np.array_equal(v32["flags"], v64["flags"])
    assert np.allclose(v32["lRest"], l64, rtol=1e-6)
    assert np.allclose(inv, im64, rtol=1e-5)
    # frames are transported along 64-vertex strands: FP32 error grows slowly
    assert np.max(np.abs(q - q64)) < 2e-5
    assert np.max(np.abs(qr - qr64)) < 1e-4


def test_set_kbend_and_kstretch():
    pos = np.zeros((6, 3), np.float32)
    pos[:, 0] = np.arange(6) * 0.002
    v, q, qr, inv, p = configured(pos)
    lib = _lib.load()
    lib.yb_set_kbend_host(6, ptr(v), ptr(qr), ctypes.c_float(1e-8))
    lib.yb_set_kstretch_host(6, ptr(v), ctypes.c_float(2e-2))
    # |qRest| = 4 kBend / lRest (YarnBall.cu:216-223); kStretch = k * lRest (:208-209)
    assert np.allclose(np.linalg.norm(qr[:5], axis=1), 4e-8 / 0.002, rtol=1e-5)
    assert np.allclose(v["kStretch"][:5], 2e-2 * 0.002, rtol=1e-5)


def test_validate_params():
    lib = _lib.load()
    p = Params()
    lib.yb_default_params(ctypes.byref(p), 10, 1e-3)
    p.minSegLen = 1e-3
    assert lib.yb_validate_params(ctypes.byref(p)) == 0
    assert p.detectionRadius == pytest.approx(1e-4 + 4e-4)
    assert p.scaledDetectionRadius == pytest.approx(5e-4 * 1.2)
    p.detectionScaler = 1.0
    assert lib.yb_validate_params(ctypes.byref(p)) != 0
    assert "Detection scaler" in _lib.last_message()
    p.detectionScaler = 1.2
    p.minSegLen = 1e-4
    assert lib.yb_validate_params(ctypes.byref(p)) != 0
    assert "thinner yarn" in _lib.last_message()
