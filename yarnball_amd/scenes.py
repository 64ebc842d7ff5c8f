"""Seeded synthetic scenes for the stepper (the workloads the tests and the
benchmark run).  Each builder returns a ``HostScene``: the fully configured host
arrays (after the library's host-side configure/setK* math) plus parameters, so
the device and the CPU oracle start from byte-identical state.

* ``readme_scene``   two stiff curved strands, 64 vertices (README.md:104-126)
* ``twisted_cable``  helical strands on a hexagonal bundle, both ends pinned,
                     top end rotated (the twist driver pattern of main.cpp:54-87)
* ``knit_patch``     interlocked sinusoidal rows, one edge pinned, the opposite
                     edge pulled (the pull driver pattern of main.cpp:90-121)
* ``glued_rings``    a chain of interlocked closed rings, ends glued (glue paths)
* ``random_walk``    one strand as a seeded random walk (batched scenes)
"""
import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import FLAG_HAS_NEXT, Params, VERTEX_DTYPE, check, ptr


@dataclass
class HostScene:
    name: str
    verts: np.ndarray
    vels: np.ndarray
    qs: np.ndarray
    qRests: np.ndarray
    params: Params
    maxH: float
    initialInvMasses: np.ndarray
    frame_dt: float = 1.0 / 30.0
    frames: int = 1
    # kinematic driver: pinned vertex ids and their motion
    drive_ids: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    drive_kind: str = ""  # "", "twist", "pull"
    drive_axis_point: np.ndarray = field(default_factory=lambda: np.zeros(3))
    drive_rate: float = 0.0  # rad/s (twist) or m/s along +x (pull)
    init_pos: np.ndarray = None

    @property
    def n(self):
        return len(self.verts)

    def copy(self):
        return HostScene(self.name, self.verts.copy(), self.vels.copy(), self.qs.copy(),
                         self.qRests.copy(), self.params.copy(), self.maxH,
                         self.initialInvMasses.copy(), self.frame_dt, self.frames,
                         self.drive_ids.copy(), self.drive_kind, self.drive_axis_point.copy(),
                         self.drive_rate, None if self.init_pos is None else self.init_pos.copy())


def _configure(name, pos, breaks_after, pinned, kCollision, kBend, kStretch, maxH,
               density=1e-3, radius=None, barrier=None, gravity=None, numItr=None, **kw):
    """Sim(n) + user setup + configure() + setKBend/setKStretch on the host.

    breaks_after: vertex ids whose hasNext flag is cleared (strand ends)."""
    lib = _lib.load()
    n = len(pos)
    p = Params()
    check(lib.yb_default_params(ctypes.byref(p), n, 1e-3))
    verts = np.zeros(n, VERTEX_DTYPE)
    vels = np.zeros((n, 3), np.float32)
    qs = np.zeros((n, 4), np.float32)
    qRests = np.zeros((n, 4), np.float32)
    check(lib.yb_default_vertices(n, ptr(verts), ptr(vels), ptr(qs), ptr(qRests)))
    verts["pos"] = np.asarray(pos, np.float32)
    verts["invMass"][np.asarray(pinned, np.int64)] = 0.0
    fl = verts["flags"]
    fl[np.asarray(breaks_after, np.int64)] = 0
    p.kCollision = kCollision
    if radius is not None:
        p.radius = radius
    if barrier is not None:
        p.barrierThickness = barrier
    if gravity is not None:
        p.gravity[0], p.gravity[1], p.gravity[2] = gravity
    if numItr is not None:
        p.numItr = numItr
    for k, v in kw.items():
        setattr(p, k, v)
    inv0 = np.zeros(n, np.float32)
    check(lib.yb_configure_host(n, ptr(verts), ptr(qs), ptr(qRests), float(density), ptr(inv0),
                                ctypes.byref(p)), "configure")
    check(lib.yb_validate_params(ctypes.byref(p)), "uploadMeta")
    check(lib.yb_set_kbend_host(n, ptr(verts), ptr(qRests), float(kBend)))
    check(lib.yb_set_kstretch_host(n, ptr(verts), float(kStretch)))
    return HostScene(name, verts, vels, qs, qRests, p, maxH, inv0,
                     init_pos=verts["pos"].astype(np.float64).copy())


def readme_scene():
    """The README example (README.md:104-126): two strands of 32 vertices."""
    n = 64
    seg = 0.002
    pos = np.zeros((n, 3), np.float32)
    for i in range(32):
        x = (np.float32(0.00002) * np.float32(i)) * np.float32(i - 16) + np.float32(seg)
        pos[i] = (x, -np.float32(seg) * np.float32(i), 0.0)
        pos[i + 32] = (-x, -np.float32(seg) * np.float32(i), 0.0)
    sc = _configure("readme_64", pos, breaks_after=[31], pinned=[0, 32, 63], kCollision=1e-7,
                    kBend=1e-8, kStretch=2e-2, maxH=1e-3)
    sc.frame_dt = 5.0
    sc.frames = 1
    return sc


# curveRadius convention of the JSON builder (jsonBuilder.cpp:44-49)
def _curve_radius(r):
    return 0.05 * r, 2.0 * (1.0 - 0.05) * r


def _hex_points(count, spacing):
    pts = []
    R = int(np.ceil(np.sqrt(count))) + 2
    for a in range(-R, R + 1):
        for b in range(-R, R + 1):
            x = spacing * (a + 0.5 * b)
            y = spacing * (np.sqrt(3.0) / 2.0) * b
            pts.append((x * x + y * y, x, y))
    pts.sort()
    return np.array([(x, y) for _, x, y in pts[:count]])


def twisted_cable(strands=64, per=2048, yarn_r=1e-4, bundle_r=2e-3, seg_len=5e-4, pitch=1e-2,
                  seed=0, numItr=8):
    """Helical strands on a hexagonal bundle along -y; ends pinned; top rotated."""
    rng = np.random.default_rng(seed)
    cross = _hex_points(strands, 1.0)
    cross *= (bundle_r - yarn_r) / max(np.max(np.linalg.norm(cross, axis=1)), 1e-12)
    rho_max = np.max(np.linalg.norm(cross, axis=1))
    # common axial step so the outermost strand has segments of seg_len
    dz = seg_len / np.sqrt(1.0 + (2 * np.pi * rho_max / pitch) ** 2)
    n = strands * per
    pos = np.zeros((n, 3))
    for s in range(strands):
        rho = np.hypot(*cross[s])
        th0 = np.arctan2(cross[s][1], cross[s][0])
        k = np.arange(per)
        z = k * dz
        th = th0 + 2 * np.pi * z / pitch
        pos[s * per:(s + 1) * per, 0] = rho * np.cos(th)
        pos[s * per:(s + 1) * per, 2] = rho * np.sin(th)
        pos[s * per:(s + 1) * per, 1] = -z
    pos += rng.normal(0.0, 0.02 * yarn_r, pos.shape)
    breaks = [s * per + per - 1 for s in range(strands)]
    top = np.concatenate([[s * per, s * per + 1] for s in range(strands)])
    bot = np.concatenate([[s * per + per - 2, s * per + per - 1] for s in range(strands)])
    radius, barrier = _curve_radius(yarn_r)
    sc = _configure(f"twisted_cable_{n}", pos.astype(np.float32), breaks,
                    np.concatenate([top, bot]), kCollision=1e-7, kBend=1e-8, kStretch=2e-2,
                    maxH=1e-3, radius=radius, barrier=barrier, numItr=numItr)
    sc.frame_dt = 1.0 / 30.0
    sc.frames = 300
    sc.drive_ids = top.astype(np.int64)
    sc.drive_kind = "twist"
    sc.drive_axis_point = np.array([0.0, 0.0, 0.0])
    sc.drive_rate = 2 * np.pi  # 1 rev/s about +y
    return sc


def knit_patch(rows=256, per=4096, yarn_r=1e-4, seg_len=2e-4, seed=0, numItr=8):
    """Interlocked sinusoidal rows in the x-y plane (out-of-plane waviness),
    gravity along -z, x-min ends pinned, x-max ends pulled at 5 cm/s."""
    rng = np.random.default_rng(seed)
    n = rows * per
    k = np.arange(per)
    pos = np.zeros((n, 3))
    for r in range(rows):
        sl = slice(r * per, (r + 1) * per)
        pos[sl, 0] = k * seg_len
        pos[sl, 1] = r * 1.8 * yarn_r
        pos[sl, 2] = 1.1 * yarn_r * np.sin(2 * np.pi * k / 16.0 + np.pi * r)
    pos += rng.normal(0.0, 0.05 * yarn_r, pos.shape)
    breaks = [r * per + per - 1 for r in range(rows)]
    left = np.concatenate([[r * per, r * per + 1] for r in range(rows)])
    right = np.concatenate([[r * per + per - 2, r * per + per - 1] for r in range(rows)])
    radius, barrier = _curve_radius(yarn_r)
    sc = _configure(f"knit_patch_{n}", pos.astype(np.float32), breaks,
                    np.concatenate([left, right]), kCollision=1e-7, kBend=1e-8, kStretch=2e-2,
                    maxH=1e-3, radius=radius, barrier=barrier, gravity=(0.0, 0.0, -9.8),
                    numItr=numItr)
    sc.frame_dt = 1.0 / 30.0
    sc.frames = 150
    sc.drive_ids = right.astype(np.int64)
    sc.drive_kind = "pull"
    sc.drive_rate = 0.05
    # The pull is a constant velocity: set it once.
    sc.vels[right, 0] = np.float32(0.05)
    return sc


def glued_rings(rings=4, per=64, yarn_r=1e-4, ring_r=2e-3, seed=0, numItr=8):
    """A chain of interlocked closed rings: ring k is centred on the x axis at
    k (2 R - 1.8 r) and lies in the x-y plane (even k) or the x-z plane (odd
    k), so each ring passes through its neighbours' holes and touches each of
    them at two points where their segments cross, 1.8 yarn radii apart.  Each
    ring is one strand whose last vertex sits on its first and is glued to it
    (connectionIndex both ways, as the reference connects closed curves,
    reader.cpp:38-42), with its ends a quarter turn from the contacts; the
    vertex opposite the ends is pinned and gravity pulls along -y, so the glue
    spring carries load."""
    rng = np.random.default_rng(seed)
    n = rings * per
    pos = np.zeros((n, 3))
    spacing = 2.0 * ring_r - 1.8 * yarn_r
    ang = np.pi / 2 + 2 * np.pi * np.arange(per) / (per - 1)
    for r in range(rings):
        sl = slice(r * per, (r + 1) * per)
        pos[sl, 0] = r * spacing + ring_r * np.cos(ang)
        pos[sl, 1 if r % 2 == 0 else 2] = ring_r * np.sin(ang)
        pos[r * per + 1: (r + 1) * per - 1] += rng.normal(0.0, 0.02 * yarn_r, (per - 2, 3))
        pos[(r + 1) * per - 1] = pos[r * per]
    breaks = [r * per + per - 1 for r in range(rings)]
    pinned = [r * per + per // 2 for r in range(rings)]
    radius, barrier = _curve_radius(yarn_r)
    sc = _configure(f"glued_rings_{n}", pos.astype(np.float32), breaks, pinned, kCollision=1e-7,
                    kBend=1e-8, kStretch=2e-2, maxH=1e-3, radius=radius, barrier=barrier,
                    numItr=numItr)
    cid = sc.verts["connectionIndex"]
    for r in range(rings):
        cid[r * per] = r * per + per - 1
        cid[r * per + per - 1] = r * per
    return sc


def random_walk(nverts, seed=0, yarn_r=1e-4, seg_len=2e-4, kBend=1e-8, numItr=8, turn=0.3):
    """One strand whose tangent turns by N(0, turn) rad per segment (curvature
    ~ N(0, turn/segLen)); first two vertices pinned."""
    rng = np.random.default_rng(seed)
    t = np.array([1.0, 0.0, 0.0])
    pos = np.zeros((nverts, 3))
    for i in range(1, nverts):
        axis = rng.normal(size=3)
        axis -= axis.dot(t) * t
        axis /= max(np.linalg.norm(axis), 1e-12)
        ang = rng.normal(0.0, turn)
        t = t * np.cos(ang) + np.cross(axis, t) * np.sin(ang)
        t /= np.linalg.norm(t)
        pos[i] = pos[i - 1] + seg_len * t
    return _walk_scene(pos, seed, yarn_r, kBend, numItr)


def _walk_scene(pos, seed, yarn_r, kBend, numItr):
    nverts = len(pos)
    radius, barrier = _curve_radius(yarn_r)
    return _configure(f"random_walk_{nverts}_{seed}", pos.astype(np.float32), [nverts - 1],
                      [0, 1], kCollision=1e-7, kBend=kBend, kStretch=2e-2, maxH=1e-3,
                      radius=radius, barrier=barrier, numItr=numItr)


def random_walks(sizes, seed=0, yarn_r=1e-4, seg_len=2e-4, kbend=(1e-9, 1e-7), numItr=8,
                 turn=0.3):
    """Many random-walk strands (as random_walk), generated together: one
    seeded stream, the tangent update vectorised across the strands, kBend
    log-uniform in kbend.  Returns a list of configured scenes."""
    rng = np.random.default_rng(seed)
    sizes = np.asarray(sizes, np.int64)
    m = len(sizes)
    order = np.argsort(-sizes, kind="stable")  # active strands form a prefix
    ssz = sizes[order]
    # step-major storage: step i holds the positions of the strands longer than i
    active = np.searchsorted(-ssz, -np.arange(int(ssz[0])), side="left")
    base = np.concatenate([[0], np.cumsum(active)])
    flat = np.zeros((int(base[-1]), 3))
    t = np.tile(np.array([1.0, 0.0, 0.0]), (m, 1))
    cur = np.zeros((m, 3))
    for i in range(1, int(ssz[0])):
        a = int(active[i])
        tt = t[:a]
        axis = rng.normal(size=(a, 3))
        axis -= np.sum(axis * tt, axis=1, keepdims=True) * tt
        axis /= np.maximum(np.linalg.norm(axis, axis=1, keepdims=True), 1e-12)
        ang = rng.normal(0.0, turn, size=(a, 1))
        tt = tt * np.cos(ang) + np.cross(axis, tt) * np.sin(ang)
        tt /= np.linalg.norm(tt, axis=1, keepdims=True)
        t[:a] = tt
        cur[:a] += seg_len * tt
        flat[base[i]:base[i] + a] = cur[:a]
    walks = [flat[base[:int(ssz[r])] + r] for r in range(m)]
    kb = np.exp(rng.uniform(np.log(kbend[0]), np.log(kbend[1]), size=m))
    out = [None] * m
    for r, k in enumerate(order):
        out[k] = _walk_scene(walks[r], seed * 100000 + int(k), yarn_r, float(kb[k]), numItr)
    return out


def batch(scenes_, name="batch"):
    """Concatenate configured scenes into one HostScene plus scene offsets
    (for Sim.setScenes).  The scenes must share the global parameters
    (gravity, radius, barrier, kCollision, maxH, schedule); the batch takes
    the first scene's, with the smallest minSegLen and largest maxSegLen.
    Glue indices (connectionIndex >= 0) are shifted to the batch's indexing."""
    offs = np.zeros(len(scenes_) + 1, np.int32)
    offs[1:] = np.cumsum([sc.n for sc in scenes_])
    verts = np.concatenate([sc.verts for sc in scenes_])
    cid = verts["connectionIndex"]
    for k, sc in enumerate(scenes_):
        sl = slice(offs[k], offs[k + 1])
        cid[sl] = np.where(cid[sl] >= 0, cid[sl] + offs[k], cid[sl])
    p = scenes_[0].params.copy()
    p.numVerts = int(offs[-1])
    p.minSegLen = min(float(sc.params.minSegLen) for sc in scenes_)
    p.maxSegLen = max(float(sc.params.maxSegLen) for sc in scenes_)
    out = HostScene(name, verts, np.concatenate([sc.vels for sc in scenes_]),
                    np.concatenate([sc.qs for sc in scenes_]),
                    np.concatenate([sc.qRests for sc in scenes_]), p, scenes_[0].maxH,
                    np.concatenate([sc.initialInvMasses for sc in scenes_]),
                    frame_dt=scenes_[0].frame_dt, frames=scenes_[0].frames,
                    init_pos=np.concatenate([sc.init_pos for sc in scenes_]))
    return out, offs


def drive_frame(sc, pos, vels, t0, dt):
    """Per-frame kinematic update of the pinned driver vertices (host side),
    as main.cpp:54-121 does between download() and upload()."""
    ids = sc.drive_ids
    if sc.drive_kind == "twist" and len(ids):
        p0 = sc.init_pos[ids]
        def rot(a):
            c, s = np.cos(a), np.sin(a)
            x, y, z = p0[:, 0], p0[:, 1], p0[:, 2]
            return np.stack([c * x - s * z, y, s * x + c * z], axis=1)
        a0 = sc.drive_rate * t0
        a1 = sc.drive_rate * (t0 + dt)
        q0, q1 = rot(a0), rot(a1)
        pos[ids] = q0.astype(np.float32)
        vels[ids] = ((q1 - q0) / dt).astype(np.float32)
    # "pull" uses a constant velocity set at construction; nothing per frame.


def load_into(sim, sc):
    """Copy a HostScene into a yarnball_amd.Sim (configure already applied)."""
    import ctypes as _ct
    sim.verts[:] = sc.verts
    sim.vels[:] = sc.vels
    sim.qs[:] = sc.qs
    sim.qRests[:] = sc.qRests
    _ct.memmove(_ct.byref(sim.meta), _ct.byref(sc.params), _ct.sizeof(Params))
    sim.maxH = sc.maxH
    sim.initialInvMasses[:] = sc.initialInvMasses
    if sim._h is None:
        h = _ct.c_void_p()
        check(sim._lib.yb_create(sc.n, _ct.byref(h)), "create")
        sim._h = h
    sim.upload()
    return sim
