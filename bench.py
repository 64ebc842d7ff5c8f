#!/usr/bin/env python3
"""Throughput benchmark of the YarnBall stepper on one MI355X.

Workload: the knit patch, 256 rows x 4096 vertices = 1,048,576 vertices
(interlocked sinusoidal rows, dense contact from t = 0, one edge pinned, the
opposite edge pulled at 5 cm/s, gravity -z), 1/30 s frames, maxH = 1e-3
(34 substeps per frame), detection every substep, numItr sweeps per substep.

A step is one frame: Sim.advance(1/30) over the whole patch, state resident
in HBM, every substep solved until the update residual (RMS change of the
iterate per sweep in rest lengths) is <= 1e-6; the Newton residual |f|/|f0|
each method ends at is reported beside it (by default recorded in the timed
frames for the substeps that start from the plain warm start, whose prologue
evaluates |f0| anyway, and for every substep of one more, untimed frame of
the learned methods).  Four methods run the same frames
from the same initial state: the plain warm start, the learned one (the
committed model, with the substep-level plain fallback), and each of them
followed by the strand solve (a tridiagonal Newton step along the strands
after the warm start and before sweeps 2, 6, 10, ...).  value = vertices x
substeps / second over the timed frames of the learned warm start (the
faster of its two forms); every method's rate, sweeps per substep, residuals
and ratio to the plain one are beside it.

Also reported:
  roofline     the dominant kernel's algorithmic bytes per launch / its average
               launch time, measured live with HIP events on the stepper's
               stream (yb_profile), against MI355X's 8 TB/s HBM peak;
  broadphase   the same for the contact query (p candidates and k contacts per
               segment counted on the device) and the tree refit;
  cpu_baseline the CPU oracle (FP64, OpenMP) on a bounded sample of the same
               workload, on this host's cores: one substep with its sweeps
               capped at the plain GPU run's mean per substep;
  batched      the batched workload: 512 uneven random-walk scenes in one
               handle (fixed 8-sweep schedule), against one handle per scene.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per vertex of each stage (compulsory FP32 SoA traffic;
# neighbour and contact gathers excluded), c = mean contacts per segment.
def stage_bytes(stage, c):
    return {
        "init": 76.0,
        "step_limit": 36.0 + 4.0 * c,
        # fused sweep: pos+invMass 16, lRest/kStretch/cid/flags 16, y 12,
        # dx+stepLimit 16, q 16, qRest 16, numCols 4, ids 4c; writes dx 12, q 16
        "sweep": 124.0 + 4.0 * c,
        "end": 52.0,
        "aabb": 52.0,
    }.get(stage)


# HBM bytes per vertex of one launch, from rocprofv3 FETCH_SIZE and WRITE_SIZE
# scaled by a copy kernel's counts: k_sweep and k_query (with the step limit it
# computes; reusing its kept candidates, and walking) on the knit patch
# (256 x 4096, residual 1e-6, learned + strand, frames 0-6;
# profiles/substep_knit1M_hbm_traffic.md) and k_sweep on the 512-scene batch
# (8 sweeps, c = 0.24; profiles/sweep_batched512_hbm_traffic.md).
MEASURED_TRAFFIC = {("knit_patch", "sweep"): 160.5, ("batched_walks", "sweep"): 131.3,
                    ("knit_patch", "query_reuse"): 285.6, ("knit_patch", "query_walk"): 526.9}


def build_workload(name, numItr):
    from yarnball_amd import scenes

    if name == "knit_patch":
        return scenes.knit_patch(rows=256, per=4096, numItr=numItr)
    if name == "twisted_cable":
        return scenes.twisted_cable(strands=64, per=2048, numItr=numItr)
    if name == "knit_patch_small":
        return scenes.knit_patch(rows=32, per=1024, numItr=numItr)
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(sc, substeps, threads, residual, max_sweeps, fp32=False):
    """The CPU oracle (FP64; fp32=True: its float build) on `substeps` substeps
    of the scene: (vertex-substeps/s, threads, seconds, sweeps)."""
    from oracle import Oracle, load

    lib = load(fp32)
    if threads > 0:
        lib.orc_set_threads(threads)
    used = lib.orc_get_threads()
    orc = Oracle(sc, fp32=fp32)
    if residual > 0:
        orc.set_solver(residual, max_sweeps)
    p = sc.params.copy()
    # one untimed substep at 2 sweeps (first-touch of the oracle's arrays)
    if residual > 0:
        orc.set_solver(residual, 2)
    orc.advance(p, sc.maxH, sc.maxH)
    if residual > 0:
        orc.set_solver(residual, max_sweeps)
    t0 = time.perf_counter()
    sweeps = 0
    for _ in range(substeps):
        orc.advance(p, sc.maxH, sc.maxH)
        sweeps += lib.orc_last_sweeps(orc.o) if residual > 0 else int(p.numItr)
    dt = time.perf_counter() - t0
    orc.close()
    return sc.n * substeps / dt, used, dt, sweeps


def bench_batched(args):
    print(json.dumps(measure_batched(args, args.steps, args.warmup, args.steps,
                                     not args.no_cpu)))


def measure_batched(args, steps, warmup, per_scene_steps, with_cpu):
    """Batched uneven scenes: many random-walk strands (sizes log-uniform in
    [256, 65536], kBend log-uniform in [1e-9, 1e-7]) in one handle, every
    kernel launched once over all scenes, against the same scenes advanced
    one handle at a time (per_scene_steps frames).  Fixed sweep schedule
    (numItr), in which batched and per-scene runs are bit-identical.  Returns
    the JSON record."""
    from yarnball_amd import Sim, scenes
    from yarnball_amd._lib import STAGES, check, ptr

    rng = np.random.default_rng(0)
    sizes = np.exp(rng.uniform(np.log(256), np.log(65536), args.scenes)).astype(np.int64)
    t0 = time.perf_counter()
    scs = scenes.random_walks(sizes, seed=0, numItr=args.numItr)
    big, offs = scenes.batch(scs, name=f"random_walks_{args.scenes}")
    print(f"[bench] {args.scenes} scenes, {big.n} vertices, built in "
          f"{time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    frame = np.float32(big.frame_dt)
    spf = int(np.ceil(frame / np.float32(big.maxH)))
    if args.residual > 0:
        print("[bench] batched scenes run the fixed sweep schedule (numItr)", file=sys.stderr)

    tuned = None
    launch_config = args.launch_config
    tile_config = args.tile_config
    if not args.no_autotune:
        from yarnball_amd import autotune
        tuned = autotune.tune(big, offsets=offs, substeps=2,
                              log=lambda m: print(m, file=sys.stderr, flush=True))
        launch_config = (tuned["sweep_threads"], tuned["query_threads"])
        tile_config = (tuned["refit_leaves"], *tuned["query_shape"])

    b = Sim(big.n)
    # the random walks cross themselves here and there at t = 0: reported as
    # min_separation below instead of a warning per advance
    b.printErrors = False
    scenes.load_into(b, big)
    b.setScenes(offs)
    b.minSeparation(reset=True)
    if args.no_graphs:
        b.setGraphs(False)
    if launch_config:
        b.setLaunchConfig(*launch_config)
    if tile_config:
        b.setTileConfig(*tile_config)
    if args.skin:
        b.setContactSkin(*args.skin)
    skin_m = b.contactSkin()[0]  # a batch runs without the skin unless asked
    for w in range(warmup):
        b.advance(frame)
    t0 = time.perf_counter()
    for k in range(steps):
        b.advance(frame)
        print(f"[bench] batched: frame {k} done", file=sys.stderr, flush=True)
    el_b = time.perf_counter() - t0
    value = big.n * spf * steps / el_b
    min_sep = b.minSeparation()

    # live per-stage profile of the batched handle
    ms = np.zeros(len(STAGES), np.float64)
    nl = np.zeros(len(STAGES), np.int32)
    p = b.meta
    check(b._lib.yb_profile(b.handle, ctypes.byref(p), np.float32(frame / spf),
                            args.profile_substeps, ptr(ms), ptr(nl)))
    counts = b.collisionCounts()
    segs = (b.verts["flags"] & 2) != 0
    c = float(counts[segs].mean()) if segs.any() else 0.0
    i_sw = STAGES.index("sweep")
    sweep_us = ms[i_sw] / max(nl[i_sw], 1) * 1e3
    for s_ in STAGES:
        i = STAGES.index(s_)
        if nl[i]:
            print(f"[bench] stage {s_:11s} launches {nl[i]:5d} avg {ms[i] / nl[i] * 1e3:9.1f} us",
                  file=sys.stderr)
    b.close()

    # The same batch in the residual-targeted schedule with the strand solve:
    # every scene stops at its own sweep (per-scene convergence), timed over a
    # few substeps after one untimed substep.
    res = None
    if args.batched_residual > 0 and args.batched_residual_substeps > 0:
        from yarnball_amd import learned
        r = Sim(big.n)
        r.printErrors = False
        scenes.load_into(r, big)
        r.setScenes(offs)
        r.setSolver(args.batched_residual, maxSweeps=args.max_sweeps, chunk=args.chunk,
                    checkEvery=args.check_every)
        learned.strand(r, every=args.strand_every, first=args.strand_first)
        if args.no_graphs:
            r.setGraphs(False)
        if launch_config:
            r.setLaunchConfig(*launch_config)
        if tile_config:
            r.setTileConfig(*tile_config)
        h = np.float32(big.maxH)
        r.advance(h)
        r.solverStats(reset=True)
        nsub = args.batched_residual_substeps
        t0 = time.perf_counter()
        for k in range(nsub):
            r.advance(h)
            print(f"[bench] batched residual: substep {k} done", file=sys.stderr, flush=True)
        el_r = time.perf_counter() - t0
        st = r.solverStats()
        r.close()
        res = {"residual_tol": args.batched_residual,
               "residual_kind": "update (RMS sweep change / rest length), per scene",
               "method": f"plain warm start + strand solve (after the warm start and before "
                         f"sweeps {args.strand_first}, {args.strand_first}+{args.strand_every}k)",
               "substeps": nsub, "value": round(big.n * nsub / el_r, 1),
               "ms_per_substep": round(el_r / nsub * 1e3, 3),
               "batch_sweeps_per_substep": round(st["sweeps"] / max(st["substeps"], 1), 1),
               "scene_sweeps_per_substep": round(st["meanSceneSweeps"], 1),
               "not_converged_scene_substeps": int(st["notConverged"]),
               "newton_residual_mean": float(st["meanNewtonResidual"])}
        print(f"[bench] batched residual {args.batched_residual:g}: {res['value']:.4e} "
              f"vertex-substeps/s, batch {res['batch_sweeps_per_substep']} sweeps per substep, "
              f"scenes {res['scene_sweeps_per_substep']}, not converged "
              f"{res['not_converged_scene_substeps']}", file=sys.stderr, flush=True)

    per = None
    if not args.no_per_scene and per_scene_steps > 0:
        sims = []
        for sc in scs:
            s_ = Sim(sc.n)
            s_.printErrors = False
            scenes.load_into(s_, sc)
            if launch_config:
                s_.setLaunchConfig(*launch_config)
            if tile_config:
                s_.setTileConfig(*tile_config)
            sims.append(s_)
        for w in range(min(warmup, 1)):
            for s_ in sims:
                s_.advance(frame)
        t0 = time.perf_counter()
        for k in range(per_scene_steps):
            for s_ in sims:
                s_.advance(frame)
            print(f"[bench] per-scene: frame {k} done", file=sys.stderr, flush=True)
        el_p = time.perf_counter() - t0
        for s_ in sims:
            s_.close()
        per = {"value": round(big.n * spf * per_scene_steps / el_p, 1),
               "ms_per_step": round(el_p / per_scene_steps * 1e3, 3), "steps": per_scene_steps}
        print(f"[bench] per-scene launches: {per['value']:.4e} vertex-substeps/s", file=sys.stderr)

    cpu = None
    if with_cpu:
        from oracle import Oracle, load
        sub = scs[:32]
        nv = sum(sc.n for sc in sub)

        def per_scene_oracle(fp32, seconds):
            lib = load(fp32)
            if args.cpu_threads > 0:
                lib.orc_set_threads(args.cpu_threads)
            orcs = [(Oracle(sc, fp32=fp32), sc.params.copy(), sc) for sc in sub]
            rounds = 0
            t0 = time.perf_counter()
            while rounds == 0 or time.perf_counter() - t0 < seconds:
                for o, pp, sc in orcs:
                    o.advance(pp, sc.maxH, sc.maxH)
                rounds += 1
            dt = time.perf_counter() - t0
            for o, _, _ in orcs:
                o.close()
            return nv * rounds / dt, rounds, dt, int(lib.orc_get_threads())

        v, rounds, dt, used = per_scene_oracle(False, 10.0)
        cpu = {"value": round(v, 1), "unit": "vertex-substeps/s", "cores": used, "kind": "port",
               "sample": f"FP64 OpenMP oracle, the first 32 scenes ({nv} vertices) one substep "
                         f"({args.numItr} sweeps) each in turn, {rounds} rounds in {dt:.1f}s"}
        v32, rounds32, dt32, _ = per_scene_oracle(True, 5.0)
        cpu["fp32"] = {"value": round(v32, 1),
                       "sample": f"the same stepper built in float (liboracle32.so), same scenes: "
                                 f"{rounds32} rounds in {dt32:.1f}s"}

    alg = (124 + 4 * c) * big.n
    out = {
        "metric": "vertex-substeps/s (batched uneven scenes)",
        "value": round(value, 1),
        "unit": "vertex-substeps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(el_b / steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": big.name, "scenes": int(args.scenes), "vertices": int(big.n),
                   "scene_vertices": [int(sizes.min()), int(sizes.max())],
                   "substeps_per_step": spf, "numItr": int(args.numItr),
                   "launch": "eager" if args.no_graphs else "hip graphs",
                   "contact_skin_margin": round(float(skin_m), 4),
                   "min_separation_x2r": round(min_sep, 3)},
        "per_scene_launches": per,
        "launch_config": {"sweep_threads": (launch_config or (256, 128))[0],
                          "query_threads": (launch_config or (256, 128))[1],
                          "refit_leaves": (tile_config or (1024, 8, 6))[0],
                          "query_shape": list((tile_config or (1024, 8, 6))[1:]),
                          "autotune": tuned},
        "batched_vs_per_scene": round(value / per["value"], 3) if per else None,
        "residual_schedule": res,
        "roofline": {"kernel": "sweep", "bound": "hbm",
                     "achieved": round(alg / (sweep_us * 1e-6) / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(alg / (sweep_us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": round(MEASURED_TRAFFIC[("batched_walks", "sweep")] * big.n),
                     "alg_bytes_per_launch": int(alg),
                     "avg_launch_us": round(sweep_us, 2), "mean_contacts_per_segment": round(c, 3)},
        "cpu_baseline": cpu,
    }
    return out


def live_profile(sim, sc, args, substeps_per_frame):
    """yb_profile over args.profile_substeps substeps: per-stage milliseconds
    and launch counts (stage order of _lib.STAGES), and the mean contacts per
    segment of the state."""
    from yarnball_amd._lib import STAGES, check, ptr

    ms = np.zeros(len(STAGES), np.float64)
    nl = np.zeros(len(STAGES), np.int32)
    p = sim.meta
    frame = np.float32(sc.frame_dt)
    check(sim._lib.yb_profile(sim.handle, ctypes.byref(p), np.float32(frame / substeps_per_frame),
                              args.profile_substeps, ptr(ms), ptr(nl)))
    counts = sim.collisionCounts()
    segs = (sim.verts["flags"] & 2) != 0
    c = float(counts[segs].mean()) if segs.any() else 0.0
    return ms, nl, c


def broadphase_modes(sim, sc, reps=4):
    """The contact query and the refit in each way a detecting substep runs
    on the handle's current state, timed live (yb_profile, one substep each,
    the fastest of `reps`) and counted on the device (yb_debug_query_counts):
      reuse  the kept candidates of the last walk are tested (the skin's common
             case; no refit);
      walk   every segment walks with its grown box and keeps its candidates
             (the kept lists are dropped first, so this substep walks);
      plain  the skin off: the tree is refit on the segment boxes and walked
             with them (the reference's way).
    Each mode's counts come from one more substep run the same way."""
    from yarnball_amd._lib import check, ptr

    margin, sweep = sim.contactSkin()
    h = np.float32(sc.maxH)

    def counts(reuse):
        # a prologue moves nothing: when one walks, the next reuses its lists
        for _ in range(3):
            qc = np.zeros(6, np.int64)
            qp = sim.meta.copy()
            sim.contactSkinStats(reset=True)
            check(sim._lib.yb_debug_query_counts(sim.handle, ctypes.byref(qp), h, ptr(qc)))
            if not reuse or sim.contactSkinStats(reset=True)[1] == 1:
                return qc
        return None

    def drop_lists():
        sim.setContactSkin(0.0)
        sim.setContactSkin(margin, sweep)

    out = {}
    modes = (("reuse", None), ("walk", drop_lists), ("plain", None)) if margin > 0 else \
        (("plain", None),)
    for mode, before in modes:
        if mode == "plain" and margin > 0:
            sim.setContactSkin(0.0)
        q, r, got = [], [], 0
        for _ in range(reps * 3):
            if before:
                before()
            sim.contactSkinStats(reset=True)
            pr = sim.profile(1)
            w_, u_ = sim.contactSkinStats(reset=True)
            if mode == "reuse" and u_ != 1:
                continue  # this substep walked: not a reuse sample
            q.append(pr["query"][0] * 1e3)
            r.append(pr["refit"][0] * 1e3)
            got += 1
            if got == reps:
                break
        if before:
            before()
        qc = counts(mode == "reuse") if q else None
        out[mode] = {"query_us": min(q) if q else None, "refit_us": min(r) if r else None,
                     "counts": qc}
    if margin > 0:
        sim.setContactSkin(margin, sweep)
    return out


def broadphase_records(bp, workload, n, skin_stats):
    """Roofline records of the contact query per mode and of the refit.
    Algorithmic bytes per segment, in the survey's convention (p leaves
    tested by the narrowphase, k contacts and pk kept candidates scanned per
    segment, as counted on the device): a walk is traversal 24 + 8p, pair
    filter 8p + 8k + 4 and step limit 20 (the warm start read, the limit
    written): 48 + 16p + 8k; a reuse reads its list length and entries
    (4 + 4pk) in place of the traversal: 28 + 4pk + 8p + 8k.  The refit reads
    a box and writes two child boxes per segment (52); a walk with the skin
    also reads the warm start at both ends (+ 32)."""
    recs = {}
    for mode, r in bp.items():
        qc = r["counts"]
        if r["query_us"] is None or qc is None:
            continue
        segs = max(int(qc[0]), 1)
        p, k = qc[1] / segs, qc[2] / segs
        scanned = qc[3] / segs
        b_seg = (28.0 + 4.0 * scanned + 8.0 * p + 8.0 * k) if mode == "reuse" else \
            (48.0 + 16.0 * p + 8.0 * k)
        us = r["query_us"]
        ach = b_seg * segs / (us * 1e-6) / 1e9
        rec = {"kernel": "query", "mode": mode, "bound": "hbm", "achieved": round(ach, 1),
               "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
               "traffic": (round(MEASURED_TRAFFIC[(workload, "query_" + mode)] * n)
                           if (workload, "query_" + mode) in MEASURED_TRAFFIC else None),
               "alg_bytes_per_launch": int(b_seg * segs), "avg_launch_us": round(us, 2),
               "tested_per_segment": round(float(p), 3), "accepted_per_segment": round(float(k), 3)}
        if mode == "reuse":
            rec["kept_scanned_per_segment"] = round(float(scanned), 2)
        else:
            rec["node_visits_per_segment"] = round(float(scanned), 2)
            rec["stack_past_lds_fraction"] = round(float(qc[4] / segs), 4)
            rec["max_stack_depth"] = int(qc[5])
        recs["query_" + mode] = rec
        if mode != "reuse" and r["refit_us"]:
            rb = 52.0 + (32.0 if mode == "walk" else 0.0)
            ach_r = rb * segs / (r["refit_us"] * 1e-6) / 1e9
            recs["refit_" + mode] = {"kernel": "refit", "mode": mode, "bound": "hbm",
                                     "achieved": round(ach_r, 1), "peak": HBM_PEAK_GBS,
                                     "unit": "GB/s", "frac": round(ach_r / HBM_PEAK_GBS, 4),
                                     "traffic": None, "alg_bytes_per_launch": int(rb * segs),
                                     "avg_launch_us": round(r["refit_us"], 2)}
    walks, reuses = skin_stats
    recs["skin"] = {"walking_substeps": int(walks), "reusing_substeps": int(reuses),
                    "refit_skipped_us": round(bp["reuse"]["refit_us"], 2)
                    if "reuse" in bp and bp["reuse"]["refit_us"] is not None else None}
    return recs


def substep_timeline(sim, frame, substeps_per_frame):
    """One more frame (untimed) with the residual-mode substep timeline on:
    mean microseconds per substep of GPU idle before it, prologue, sweeps up
    to the chunk that converged, chunks queued after that one, and the end
    kernels, from HIP events between the graph launches; wall = the frame's
    host clock per substep."""
    sim.startTimeline(substeps_per_frame)
    t0 = time.perf_counter()
    sim.advance(frame)
    wall = (time.perf_counter() - t0) / substeps_per_frame * 1e6
    tl = sim.timeline()
    names = ("idle_us", "prologue_us", "converge_us", "tail_us", "end_us")
    out = {k: round(float(tl[1:, i].mean()) if len(tl) > 1 else 0.0, 1)
           for i, k in enumerate(names)}
    out["wall_us"] = round(wall, 1)
    out["substeps"] = int(len(tl))
    out["mean_chunks"] = round(float(tl[:, 6].mean()), 2) if len(tl) else 0.0
    return out


def run_method(sc, args, method, model, strand_model=None):
    """Warm-up frames, then the timed frames, of one warm-start method on a
    fresh simulation; returns the open simulation and the measurements."""
    from yarnball_amd import Sim, learned, scenes

    sim = Sim(sc.n)
    scenes.load_into(sim, sc)
    if args.residual > 0:
        sim.setSolver(args.residual, maxSweeps=args.max_sweeps, chunk=args.chunk, checkEvery=args.check_every)
    if method == "learned":
        learned.use(sim, model)
    # the strand methods refresh the frames from the warm start before the
    # strand step (held frames would undo the rotation the warm start predicts)
    elif method == "strand":
        learned.strand(sim, every=args.strand_every, first=args.strand_first)
    elif method == "learned+strand":
        learned.use(sim, strand_model, strand=True)
        learned.strand_schedule(sim, every=args.strand_every, first=args.strand_first)
    if method.startswith("learned") and args.learned_cost > 0:
        learned.start_cost(sim, args.learned_cost)
    if args.no_graphs:
        sim.setGraphs(False)
    if args.launch_config:
        sim.setLaunchConfig(*args.launch_config)
    if args.tile_config:
        sim.setTileConfig(*args.tile_config)
    if args.skin:
        sim.setContactSkin(*args.skin)
    if args.residual > 0:
        sim.setNewtonResidual(args.newton_residual)
    frame = np.float32(sc.frame_dt)
    substeps_per_frame = int(np.ceil(frame / np.float32(sc.maxH)))
    for w in range(args.warmup):
        sim.advance(frame)
        print(f"[bench] {method}: warmup frame {w} done", file=sys.stderr, flush=True)
    if args.residual > 0:
        sim.solverStats(reset=True)
    if method.startswith("learned"):
        learned.warm_substeps(sim, reset=True)
        learned.counters(sim, reset=True)
    sim.minSeparation(reset=True)
    sim.contactSkinStats(reset=True)
    t0 = time.perf_counter()
    for k in range(args.steps):
        sim.advance(frame)  # synchronous: returns after the stream has drained
        print(f"[bench] {method}: frame {k} done", file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t0
    r = {"sim": sim, "elapsed": elapsed, "substeps_per_frame": substeps_per_frame,
         "value": sc.n * substeps_per_frame * args.steps / elapsed,
         "min_separation": sim.minSeparation(), "skin": sim.contactSkinStats(reset=True),
         "sim_skin": sim.contactSkin()}
    print(f"[bench] {method}: contact skin: {r['skin'][0]} walking, {r['skin'][1]} reusing "
          f"substeps", file=sys.stderr, flush=True)
    if args.residual > 0:
        st = sim.solverStats()
        r["stats"] = st
        r["sweeps_per_substep"] = st["sweeps"] / max(st["substeps"], 1)
        print(f"[bench] {method}: {r['sweeps_per_substep']:.1f} sweeps/substep, max "
              f"{st['maxSweepsUsed']}, not converged {st['notConverged']}, max final residual "
              f"{st['maxResidual']:.2e}, final Newton residual mean "
              f"{st['meanNewtonResidual']:.2e} max {st['maxNewtonResidual']:.2e}",
              file=sys.stderr, flush=True)
    if method.startswith("learned"):
        r["warm_substeps"] = learned.warm_substeps(sim)
        r["vertex_counters"] = learned.counters(sim)
    if args.residual > 0:
        r["timeline"] = substep_timeline(sim, frame, substeps_per_frame)
        t = r["timeline"]
        print(f"[bench] {method}: substep timeline (us): idle {t['idle_us']}, prologue "
              f"{t['prologue_us']}, sweeps to the converged chunk {t['converge_us']}, chunks "
              f"after it {t['tail_us']}, end {t['end_us']}, wall {t['wall_us']}",
              file=sys.stderr, flush=True)
    # live per-stage timing right after the timed frames, while the GPU runs at
    # its working clocks (a handle left idle first starts with slower kernels)
    sim.contactSkinStats(reset=True)
    r["profile"] = live_profile(sim, sc, args, substeps_per_frame)
    r["profile_skin"] = sim.contactSkinStats(reset=True)
    if args.residual > 0 and method.startswith("learned") and args.newton_residual == "plain":
        # the timed frames recorded the Newton residual of the plain-start
        # substeps only; one more frame (untimed, after every timing) records
        # it for every start
        sim.setNewtonResidual("all")
        sim.solverStats(reset=True)
        sim.advance(frame)
        sa = sim.solverStats()
        r["newton_all_starts"] = {"substeps": int(sa["newtonSubsteps"]),
                                  "mean": float(sa["meanNewtonResidual"]),
                                  "max": float(sa["maxNewtonResidual"])}
        sim.setNewtonResidual(args.newton_residual)
        print(f"[bench] {method}: one more frame with the Newton residual of every start: mean "
              f"{sa['meanNewtonResidual']:.2e} max {sa['maxNewtonResidual']:.2e} over "
              f"{sa['newtonSubsteps']} substeps", file=sys.stderr, flush=True)
    print(f"[bench] {method}: timed {args.steps} frames in {elapsed:.3f}s "
          f"({r['value']:.4e} vertex-substeps/s), min separation {r['min_separation']:.3f} x 2r",
          file=sys.stderr, flush=True)
    return r


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="knit_patch")
    ap.add_argument("--numItr", type=int, default=8)
    ap.add_argument("--cpu-substeps", type=int, default=1)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-substeps", type=int, default=2)
    ap.add_argument("--residual", type=float, default=1e-6,
                    help="relative Newton residual target; <= 0 runs numItr sweeps")
    ap.add_argument("--max-sweeps", type=int, default=6000)
    ap.add_argument("--chunk", type=int, default=16,
                    help="sweeps per chunk graph in residual mode")
    ap.add_argument("--learned-cost", type=float, default=0.0,
                    help="the substep chooser's offset: a learned start's extra prologue time "
                         "in sweeps")
    ap.add_argument("--newton-residual", choices=["all", "plain"], default="plain",
                    help="residual mode: which substeps record the Newton residual (a "
                         "diagnostic; the stopping test is the update residual): every one, or "
                         "those starting from the plain warm start, whose prologue evaluates "
                         "its denominator anyway")
    ap.add_argument("--check-every", type=int, default=4,
                    help="residual mode: sweeps between residual evaluations")
    ap.add_argument("--method", choices=["plain", "learned", "both", "strand", "learned+strand",
                                         "all"],
                    default="all",
                    help="warm start(s) to run; value is the fastest non-plain method's rate")
    ap.add_argument("--no-graphs", action="store_true",
                    help="enqueue kernels directly instead of HIP graphs (profiler runs)")
    ap.add_argument("--model", default=os.path.join(ROOT, "yarnball_amd", "models",
                                                    "warmstart.npy"))
    ap.add_argument("--strand-model", default=os.path.join(ROOT, "yarnball_amd", "models",
                                                           "warmstart_strand.npy"),
                    help="model of the learned + strand method (trained on that solver's "
                         "converged corrections)")
    ap.add_argument("--strand-every", type=int, default=4,
                    help="strand methods: also run the strand step before every k-th sweep "
                         "(0: only after the warm start)")
    ap.add_argument("--strand-first", type=int, default=2,
                    help="strand methods: the sweep before which the first such step runs "
                         "(0: the strand-every-th)")
    ap.add_argument("--scenes", type=int, default=512, help="batched_walks: number of scenes")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default block sizes instead of timing the candidates")
    ap.add_argument("--no-per-scene", action="store_true",
                    help="batched_walks: skip the one-handle-per-scene comparison")
    ap.add_argument("--no-batched", action="store_true",
                    help="default workload: skip the batched-scenes record")
    ap.add_argument("--batched-residual", type=float, default=1e-5,
                    help="batched scenes: residual target of the per-scene convergence record "
                         "(<= 0 skips it)")
    ap.add_argument("--batched-residual-substeps", type=int, default=4,
                    help="batched scenes: timed substeps of the per-scene convergence record")
    ap.add_argument("--launch-config", type=int, nargs=2, default=None,
                    metavar=("SWEEP_THREADS", "QUERY_THREADS"),
                    help="fixed block sizes (skips the autotuner)")
    ap.add_argument("--skin", type=float, nargs=2, default=None, metavar=("MARGIN", "SWEEP"),
                    help="contact skin: margin in detection radii and sweep in substeps (default: the "
                         "library's; margin 0 walks the tree in every substep)")
    ap.add_argument("--tile-config", type=int, nargs=3, default=None,
                    metavar=("REFIT_LEAVES", "QUERY_STACK_LDS", "QUERY_CANDIDATES"),
                    help="fixed refit tile and query shape (default when the autotuner is "
                         "skipped: 1024 8 6)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the stepper runs on a single MI355X "
              "(multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)
    if args.launch_config:
        args.launch_config = tuple(args.launch_config)
        args.no_autotune = True
    if args.tile_config:
        args.tile_config = tuple(args.tile_config)
    if args.workload == "batched_walks":
        return bench_batched(args)

    from yarnball_amd._lib import STAGES, check, ptr

    t_build = time.perf_counter()
    sc = build_workload(args.workload, args.numItr)
    print(f"[bench] workload {sc.name}: {sc.n} vertices, built in "
          f"{time.perf_counter() - t_build:.1f}s", file=sys.stderr, flush=True)
    methods = {"both": ["plain", "learned"],
               "all": ["plain", "learned", "strand", "learned+strand"]}.get(args.method,
                                                                             [args.method])
    tuned = None
    if not args.no_autotune:
        from yarnball_amd import autotune
        tuned = autotune.tune(sc, residual=args.residual, max_sweeps=args.max_sweeps, substeps=1,
                              log=lambda m: print(m, file=sys.stderr, flush=True))
        args.launch_config = (tuned["sweep_threads"], tuned["query_threads"])
        args.tile_config = (tuned["refit_leaves"], *tuned["query_shape"])
    model = strand_model = None
    if "learned" in methods:
        model = np.load(args.model, allow_pickle=False)
    if "learned+strand" in methods:
        strand_model = np.load(args.strand_model, allow_pickle=False)
    runs = {}
    for m in methods:
        runs[m] = run_method(sc, args, m, model, strand_model)
    # headline: the learned warm start the metric names -- the faster of its
    # two forms (alone, or refined by the strand step) -- at the common
    # residual; another method only when no learned one ran
    learned_ms = [m for m in methods if m.startswith("learned")]
    main_m = max(learned_ms or [m for m in methods if m != "plain"] or methods,
                 key=lambda m: runs[m]["value"])
    ref = runs["plain" if "plain" in runs else main_m]
    # broadphase on the headline method's final state, in each way a
    # detecting substep can run (counts for the byte models, and times)
    qsim = runs[main_m]["sim"]
    bp = broadphase_modes(qsim, sc)
    for m in methods:
        if runs[m] is not ref:
            runs[m]["sim"].close()
    sim = ref["sim"]
    frame = np.float32(sc.frame_dt)
    substeps_per_frame = ref["substeps_per_frame"]
    value = runs[main_m]["value"]
    elapsed = runs[main_m]["elapsed"]
    stats = runs[main_m].get("stats")

    # Live per-stage timing on the stepper's stream, of the main method's
    # handle right after its timed frames.
    ms, nl, c = runs[main_m]["profile"]
    per_launch = {s: (ms[i] / nl[i] if nl[i] else 0.0) for i, s in enumerate(STAGES)}
    total = {s: ms[i] for i, s in enumerate(STAGES)}
    dominant = max(total, key=total.get)
    if stage_bytes(dominant, 0.0) is None:
        # roofline is stated for the dominant kernel with an algorithmic byte model
        dominant = max((s for s in STAGES if stage_bytes(s, 0.0) is not None and nl[STAGES.index(s)]),
                       key=total.get)
    for s in STAGES:
        i = STAGES.index(s)
        if nl[i]:
            b = stage_bytes(s, c)
            gbs = (b * sc.n / (per_launch[s] * 1e-3) / 1e9) if b else float("nan")
            print(f"[bench] stage {s:11s} launches {nl[i]:5d} avg {per_launch[s]*1e3:9.1f} us "
                  f"total {ms[i]:8.2f} ms  alg GB/s {gbs:8.1f}", file=sys.stderr)
    b = stage_bytes(dominant, c) or 0.0
    alg_bytes = b * sc.n
    achieved = alg_bytes / (per_launch[dominant] * 1e-3) / 1e9 if per_launch[dominant] else 0.0
    roofline = {
        "kernel": dominant,
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": (round(MEASURED_TRAFFIC[(args.workload, dominant)] * sc.n)
                    if (args.workload, dominant) in MEASURED_TRAFFIC and args.residual > 0 else None),
        "alg_bytes_per_launch": int(alg_bytes),
        "avg_launch_us": round(per_launch[dominant] * 1e3, 2),
        "mean_contacts_per_segment": round(c, 3),
    }
    broad = broadphase_records(bp, args.workload, sc.n, runs[main_m]["skin"])
    print("[bench] broadphase: " + "; ".join(
        f"{k_} {v['avg_launch_us']:.1f} us {v['frac']:.3f} of peak" for k_, v in broad.items()
        if isinstance(v, dict) and "avg_launch_us" in v), file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        cap = args.max_sweeps
        if args.residual > 0:
            cap = max(1, int(round(ref["sweeps_per_substep"])))
        v, used, dt, sw = cpu_baseline(sc, args.cpu_substeps, args.cpu_threads, args.residual, cap)
        cpu = {"value": round(v, 1), "unit": "vertex-substeps/s", "cores": int(used),
               "kind": "port",
               "sample": f"FP64 OpenMP oracle, {args.cpu_substeps} substep(s) of {sc.name} from "
                         f"its initial state, {sw} sweeps (residual target {args.residual:g}, "
                         f"sweeps capped at the plain GPU run's mean per substep), {dt:.1f}s"}
        print(f"[bench] cpu baseline {v:.3e} vertex-substeps/s on {used} threads",
              file=sys.stderr)
        v32, _, dt32, sw32 = cpu_baseline(sc, args.cpu_substeps, args.cpu_threads, args.residual,
                                          cap, fp32=True)
        cpu["fp32"] = {"value": round(v32, 1),
                       "sample": f"the same stepper built in float (liboracle32.so), same sample: "
                                 f"{sw32} sweeps, {dt32:.1f}s"}
        print(f"[bench] cpu baseline fp32 {v32:.3e} vertex-substeps/s", file=sys.stderr)

    method_rows = {}
    for m, r in runs.items():
        row = {"value": round(r["value"], 1), "ms_per_step": round(r["elapsed"] / args.steps * 1e3, 3),
               "min_separation_x2r": round(r["min_separation"], 3),
               "skin_walking_substeps": int(r["skin"][0]),
               "skin_reusing_substeps": int(r["skin"][1])}
        if "stats" in r:
            row["sweeps_per_substep"] = round(r["sweeps_per_substep"], 1)
            row["not_converged_substeps"] = int(r["stats"]["notConverged"])
            row["max_final_residual"] = float(r["stats"]["maxResidual"])
            nr = r["stats"]["meanNewtonResidual"]
            row["newton_residual_mean"] = float(nr) if np.isfinite(nr) else None
            row["newton_residual_max"] = float(r["stats"]["maxNewtonResidual"])
            row["newton_residual_substeps"] = int(r["stats"]["newtonSubsteps"])
            if "newton_all_starts" in r:
                row["newton_residual_all_starts_frame"] = r["newton_all_starts"]
            row["substep_timeline"] = r.get("timeline")
        if m.startswith("learned"):
            ws = r["warm_substeps"]
            row["substeps_learned_start"] = ws["learned"]
            row["substeps_plain_fallback"] = ws["plain"]
            row["fallback_rate"] = round(ws["plain"] / max(ws["learned"] + ws["plain"], 1), 3)
            row["vertex_fallbacks"] = r["vertex_counters"]["fallback"]
        method_rows[m] = row
    def vs_plain(m):
        return (round(runs[m]["value"] / runs["plain"]["value"], 3)
                if m in runs and "plain" in runs else None)
    method_names = {
        "plain": "plain warm start",
        "learned": "learned warm start with substep-level plain fallback",
        "strand": f"plain warm start + strand solve (after the warm start and before sweeps "
                  f"{args.strand_first}, {args.strand_first}+{args.strand_every}k)",
        "learned+strand": f"learned warm start with substep-level plain fallback + strand solve "
                          f"(after the warm start and before sweeps {args.strand_first}, "
                          f"{args.strand_first}+{args.strand_every}k)",
    }

    sim.close()
    # the batched uneven-scene workload (512 random-walk scenes) in the same run
    batched = None
    if not args.no_batched:
        batched = measure_batched(args, args.steps, min(args.warmup, 2), 2, False)

    out = {
        "metric": "vertex-substeps/s (plain vs learned warm-start, equal residual) + HBM GB/s",
        "value": round(value, 1),
        "unit": "vertex-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": sc.name,
            "vertices": sc.n,
            "substeps_per_step": substeps_per_frame,
            "numItr": int(args.numItr) if args.residual <= 0 else None,
            "residual_tol": args.residual if args.residual > 0 else None,
            "residual_kind": "update (RMS sweep change / rest length)" if args.residual > 0 else None,
            "mean_sweeps_per_substep": (round(stats["sweeps"] / max(stats["substeps"], 1), 1)
                                        if stats else int(args.numItr)),
            "not_converged_substeps": stats["notConverged"] if stats else None,
            "method": method_names[main_m],
            "detection_period": int(sc.params.detectionPeriod),
            "contact_skin": dict(zip(("margin_detection_radii", "sweep_substeps"),
                                     (round(float(x), 4) for x in runs[main_m]["sim_skin"]))),
            "launch": "eager" if args.no_graphs else "hip graphs",
        },
        "methods": method_rows,
        "learned_vs_plain": vs_plain("learned"),
        "strand_vs_plain": vs_plain("strand"),
        "learned_strand_vs_plain": vs_plain("learned+strand"),
        "launch_config": {"sweep_threads": (args.launch_config or (256, 128))[0],
                          "query_threads": (args.launch_config or (256, 128))[1],
                          "refit_leaves": (args.tile_config or (1024, 8, 6))[0],
                          "query_shape": list((args.tile_config or (1024, 8, 6))[1:]),
                          "autotune": tuned},
        "roofline": roofline,
        "roofline_broadphase": broad.get("query_reuse") or broad.get("query_plain"),
        "broadphase": broad,
        "cpu_baseline": cpu,
        "batched": batched,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
