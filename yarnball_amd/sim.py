"""Python mirror of YarnBall::Sim (KittenEngine/YarnBall/YarnBall.h:116-206).

Same names, argument meaning and error behaviour as the reference class; every
call goes through the C ABI of libyarnball_hip.so.  Host arrays are numpy:

* ``verts``   structured array of VERTEX_DTYPE (pos, invMass, lRest, kStretch,
              connectionIndex, flags) -- the 32-byte Vertex record;
* ``vels``    (n, 3) float32;
* ``qs``      (n, 4) float32 rotors (x, y, z, w);
* ``qRests``  (n, 4) float32 rest Darboux rotors scaled by the bend stiffness;
* ``meta``    a ``Params`` ctypes struct with MetaData's field names.

As in the reference, the device copy is authoritative between ``upload()`` and
``download()`` (README.md:100-102).
"""
import ctypes
import sys

import numpy as np

from . import _lib
from ._lib import Params, VERTEX_DTYPE, YarnBallError, check, ptr


class Sim:
    ERROR_NONE = 0
    ERROR_MAX_COLLISIONS_PER_SEGMENT_EXCEEDED = 1
    WARNING_SEGMENT_INTERPENETRATION = 2

    def __init__(self, numVerts):
        # YarnBall.cu:11-51
        if numVerts < 3:
            raise RuntimeError("Too little vertices")
        lib = _lib.load()
        self._lib = lib
        n = int(numVerts)
        self.maxH = 1e-3
        self.meta = Params()
        lib.yb_default_params(ctypes.byref(self.meta), n, self.maxH)
        self.verts = np.zeros(n, dtype=VERTEX_DTYPE)
        self.vels = np.zeros((n, 3), dtype=np.float32)
        self.qs = np.zeros((n, 4), dtype=np.float32)
        self.qRests = np.zeros((n, 4), dtype=np.float32)
        check(lib.yb_default_vertices(n, ptr(self.verts), ptr(self.vels), ptr(self.qs),
                                      ptr(self.qRests)))
        self.initialInvMasses = np.zeros(n, dtype=np.float32)
        inf = np.float32(np.inf)
        self.currentBounds = (np.array([inf] * 3, np.float32), np.array([-inf] * 3, np.float32))
        self.lastErrorCode = self.ERROR_NONE
        self.lastWarningCode = self.ERROR_NONE
        self.printErrors = True
        self._h = None

    # -- lifecycle -------------------------------------------------------------
    @property
    def numVerts(self):
        return int(self.meta.numVerts)

    def configure(self, density=1e-3):
        """Sim::configure (YarnBall.cu:91-200)."""
        n = self.numVerts
        lib = self._lib
        check(lib.yb_configure_host(n, ptr(self.verts), ptr(self.qs), ptr(self.qRests),
                                    float(density), ptr(self.initialInvMasses),
                                    ctypes.byref(self.meta)), "configure")
        if self._h is None:
            h = ctypes.c_void_p()
            check(lib.yb_create(n, ctypes.byref(h)), "configure")
            self._h = h
        check(lib.yb_validate_params(ctypes.byref(self.meta)), "uploadMeta")
        self.upload()

    def _need_handle(self):
        if self._h is None:
            raise RuntimeError("No rest length. Must call configure()")

    def setKStretch(self, k=1e-2):
        """Sim::setKStretch (YarnBall.cu:202-210)."""
        self._need_handle()
        check(self._lib.yb_set_kstretch_host(self.numVerts, ptr(self.verts), float(k)))

    def setKBend(self, k=3e-9):
        """Sim::setKBend (YarnBall.cu:212-224)."""
        self._need_handle()
        check(self._lib.yb_set_kbend_host(self.numVerts, ptr(self.verts), ptr(self.qRests),
                                          float(k)))

    def upload(self):
        """Sim::upload (YarnBall.cu:247-254)."""
        self._need_handle()
        check(self._lib.yb_upload(self._h, ptr(self.verts), ptr(self.vels), ptr(self.qs),
                                  ptr(self.qRests)), "upload")

    def download(self):
        """Sim::download (YarnBall.cu:256-261)."""
        if self._h is None:
            return
        check(self._lib.yb_download(self._h, ptr(self.verts), ptr(self.vels), ptr(self.qs)),
              "download")

    def zeroVelocities(self):
        """Sim::zeroVelocities (YarnBall.cu:271-274)."""
        if self._h is None:
            return
        check(self._lib.yb_zero_velocities(self._h))

    # -- stepping --------------------------------------------------------------
    def advance(self, dt):
        """Sim::advance (step.cpp:36-55) + checkErrors (YarnBall.cu:276-304).

        Returns the simulated time advanced."""
        if dt <= 0:
            return 0.0
        self._need_handle()
        warn = ctypes.c_int(0)
        rc = self._lib.yb_advance(self._h, ctypes.byref(self.meta), float(dt), float(self.maxH),
                                  ctypes.byref(warn))
        if warn.value == self.WARNING_SEGMENT_INTERPENETRATION:
            self.lastWarningCode = warn.value
            if self.printErrors:
                print("WARNING: Interpenetration detection. This can be due to unstable contacts",
                      file=sys.stderr)
        if rc == self.ERROR_MAX_COLLISIONS_PER_SEGMENT_EXCEEDED:
            self.lastErrorCode = rc
            print("ERROR: MAX_COLLISIONS_PER_SEGMENT exceeded. Current simulation state may be "
                  "corrupted!", file=sys.stderr)
            raise YarnBallError(rc, "MAX_COLLISIONS_PER_SEGMENT exceeded")
        if rc == _lib.YB_ERROR_TRAVERSAL_STACK_EXCEEDED:
            self.lastErrorCode = rc
            raise YarnBallError(rc, "broadphase traversal stack exceeded")
        check(rc, "advance")
        return float(dt)

    def step(self, dt=None):
        """Sim::step (step.cpp:57-59): one maxH step."""
        self.advance(self.maxH)

    def stepCount(self):
        return int(self._lib.yb_step_count(self._h)) if self._h is not None else 0

    def bounds(self):
        """Root box of the broadphase tree (Sim::bounds, YarnBall.cu:87-89)."""
        out = np.zeros(6, np.float32)
        if self._h is not None:
            check(self._lib.yb_bounds(self._h, ptr(out)))
        return out[:3].copy(), out[3:].copy()

    def collisionCounts(self):
        out = np.zeros(self.numVerts, np.int32)
        self._need_handle()
        check(self._lib.yb_collision_counts(self._h, ptr(out)))
        return out

    def printCollisionStats(self):
        """Sim::printCollisionStats (statistics.cpp:5-24)."""
        c = self.collisionCounts()
        print(f"Total collisions: {int(c[c > 0].sum()) // 2}")
        print(f"Segments with collisions: {int((c > 0).sum())}/{self.numVerts}")

    # -- residual-targeted solving (no counterpart in the reference) -----------
    def setSolver(self, residualTol, maxSweeps=512, chunk=8, checkEvery=1, kind=None):
        """Sweep every substep until the stopping residual <= residualTol (or
        maxSweeps); residualTol <= 0 restores the fixed meta.numItr schedule.
        kind: "update" -- RMS |dx_k+1 - dx_k| / rest length (the default);
        "newton" -- |f(dx_k)| / |f(dx_0)| over free vertices, dx_0 the plain warm
        start.  Both are measured and reported by solverStats()."""
        self._need_handle()
        check(self._lib.yb_set_solver(self._h, float(residualTol), int(maxSweeps), int(chunk),
                                      int(checkEvery)))
        if kind is not None:
            k = {"update": 0, "newton": 1}[kind]
            check(self._lib.yb_set_residual_kind(self._h, k))

    def setNewtonResidual(self, starts="all"):
        """Which substeps record the Newton residual |f| / |f0| while the
        stopping test is the update residual: "all" (the default) or "plain"
        -- only substeps starting from the plain warm start, whose prologue
        evaluates f0 anyway; a learned or temporal start then skips that extra
        force pass.  The iterates do not change."""
        self._need_handle()
        check(self._lib.yb_set_newton_residual(self._h, {"all": 1, "plain": 0}[starts]))

    def solverStats(self, reset=False):
        """Counters of the residual-targeted mode since the last reset;
        meanNewtonResidual is the mean over the substeps that recorded one
        (newtonSubsteps: every substep, or scene-substep in a batch, unless
        setNewtonResidual("plain")) of the final |f| / |f0|; meanSceneSweeps
        the mean sweeps a scene ran per substep (= sweeps per substep without
        batching)."""
        st = _lib.SolverStats()
        self._need_handle()
        check(self._lib.yb_solver_stats(self._h, ctypes.byref(st), int(bool(reset))))
        out = {f: getattr(st, f) for f, _ in st._fields_}
        out["meanNewtonResidual"] = (out["sumNewtonResidual"] / out["newtonSubsteps"]
                                     if out["newtonSubsteps"] > 0 else float("nan"))
        out["meanSceneSweeps"] = out["sceneSweeps"] / max(out["sceneSubsteps"], 1)
        return out

    def setScenes(self, offsets):
        """Batched scenes: scene k owns vertices [offsets[k], offsets[k+1])."""
        self._need_handle()
        off = np.ascontiguousarray(offsets, np.int32)
        check(self._lib.yb_set_scenes(self._h, len(off) - 1, ptr(off)), "setScenes")

    def setLaunchConfig(self, sweepThreads=256, queryThreads=128):
        """Threads per block of the sweep and the contact query (speed only)."""
        self._need_handle()
        check(self._lib.yb_set_launch_config(self._h, int(sweepThreads), int(queryThreads)))

    def launchConfig(self):
        self._need_handle()
        a, b = ctypes.c_int(0), ctypes.c_int(0)
        check(self._lib.yb_get_launch_config(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def setTileConfig(self, refitLeaves=1024, queryStackLds=8, queryCandidates=6):
        """Refit tile (leaves per workgroup) and the contact query's LDS shape
        (speed only; a new refit tile rebuilds the tree at the next detection)."""
        self._need_handle()
        check(self._lib.yb_set_tile_config(self._h, int(refitLeaves), int(queryStackLds),
                                           int(queryCandidates)), "setTileConfig")

    def tileConfig(self):
        self._need_handle()
        a, b, c = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
        check(self._lib.yb_get_tile_config(self._h, ctypes.byref(a), ctypes.byref(b),
                                           ctypes.byref(c)))
        return a.value, b.value, c.value

    def setContactSkin(self, margin=0.05, sweep=8.0):
        """Contact skin: reuse each segment's broadphase candidates while its
        box stays inside its box of the last walk grown by `sweep` substeps of
        warm-start motion and by margin x detectionRadius (margin <= 0: walk the tree in
        every detecting substep).  The contact lists are the same either way."""
        self._need_handle()
        check(self._lib.yb_set_contact_skin(self._h, float(margin), float(sweep)),
              "setContactSkin")

    def contactSkin(self):
        self._need_handle()
        a, b = ctypes.c_float(0), ctypes.c_float(0)
        check(self._lib.yb_get_contact_skin(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def contactSkinStats(self, reset=False):
        """(detecting substeps that walked the tree, that reused the kept
        candidates) since the last reset."""
        self._need_handle()
        out = np.zeros(2, np.int64)
        check(self._lib.yb_contact_skin_stats(self._h, ptr(out), int(bool(reset))))
        return int(out[0]), int(out[1])

    def profile(self, substeps=2):
        """Run `substeps` substeps of maxH with HIP events around every stage
        (yb_profile); returns {stage: (milliseconds, launches)}."""
        self._need_handle()
        ms = np.zeros(len(_lib.STAGES), np.float64)
        nl = np.zeros(len(_lib.STAGES), np.int32)
        check(self._lib.yb_profile(self._h, ctypes.byref(self.meta), float(np.float32(self.maxH)),
                                   int(substeps), ptr(ms), ptr(nl)), "profile")
        return {st: (float(ms[i]), int(nl[i])) for i, st in enumerate(_lib.STAGES)}

    TIMELINE_FIELDS = ("idle_us", "prologue_us", "converge_us", "tail_us", "end_us",
                       "sweeps", "chunks", "converged_chunk")

    def startTimeline(self, substeps):
        """Record a GPU timeline of the next `substeps` residual-mode substeps
        (yb_debug_timeline); read it with timeline()."""
        self._need_handle()
        check(self._lib.yb_debug_timeline(self._h, int(substeps)), "timeline")

    def timeline(self, stop=True):
        """Recorded substeps as an (rows, 8) array with columns TIMELINE_FIELDS."""
        self._need_handle()
        cap = 1 << 16
        out = np.zeros((cap, 8), np.float32)
        rows = ctypes.c_int(0)
        check(self._lib.yb_debug_timeline_read(self._h, ptr(out), cap, ctypes.byref(rows)),
              "timeline")
        if stop:
            check(self._lib.yb_debug_timeline(self._h, 0), "timeline")
        return out[:rows.value].copy()

    def driverInit(self, initialInvMasses=None):
        """Record the current device positions (and initial inverse masses) for
        the scenario drivers; returns the centre of their bounding box."""
        self._need_handle()
        c = (ctypes.c_float * 3)()
        inv = None if initialInvMasses is None else np.ascontiguousarray(initialInvMasses, np.float32)
        check(self._lib.yb_driver_init(self._h, None if inv is None else ptr(inv), c))
        return np.array(list(c), np.float32)

    def driveFrame(self, driver, t, dt):
        """One frame of a scenario driver on the device (before advance(dt))."""
        self._need_handle()
        check(self._lib.yb_driver_frame(self._h, ctypes.byref(self.meta), ctypes.byref(driver),
                                        float(t), float(dt)))

    @staticmethod
    def defaultDriver(kind):
        d = _lib.Driver()
        check(_lib.load().yb_driver_default(int(kind), ctypes.byref(d)))
        return d

    def setGraphs(self, on=True):
        """Run substeps as captured HIP graphs (default) or enqueue every
        kernel directly (same results; for tools that cannot trace graphs)."""
        self._need_handle()
        check(self._lib.yb_set_graphs(self._h, int(bool(on))))

    def minSeparation(self, reset=False):
        """Smallest contact distance / (2 radius) since the last reset (below 1:
        interpenetration, reported as WARNING_SEGMENT_INTERPENETRATION)."""
        out = np.zeros(1, np.float32)
        self._need_handle()
        check(self._lib.yb_min_separation(self._h, ptr(out), int(bool(reset))))
        return float(out[0])

    # -- stage-level access (parity tests) --------------------------------------
    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h is not None:
            self._lib.yb_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
