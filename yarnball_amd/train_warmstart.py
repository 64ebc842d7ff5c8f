"""Generate training data for the learned warm start from the plain solver and
train the model on the device.

Data: (features at the warm start, converged correction (dx* - dx0)/L) pairs
recorded by the residual-targeted plain solver, seeded and deterministic:
  * the README scene (two strands of 32 vertices, README.md:104-126), every
    substep of its first 0.5 s;
  * the twisted cable (64 helical strands x 2048 vertices, top end rotated at
    1 rev/s by the per-frame driver pattern of main.cpp:54-87), every 4th
    substep of 8 frames, a seeded vertex subsample of each;
  * a knit patch of the evaluation size (256 rows x 4096 vertices) with a
    different jitter seed (7), every substep of frames 1-4 (the first frame,
    the settling transient, is run but not recorded), a seeded vertex
    subsample of each.  The cable has no contacts, so without a knit the model
    never sees a contact state; a smaller knit does not do, because the
    dominant part of the correction is the patch-wide sag/tension mode, which
    depends on the span.  The transient is left out because its corrections
    are far larger than the steady state's: in a squared loss they crowd out
    the small steady-state corrections, and the model then starts the steady
    state worse than the plain warm start (1.01x plain over frames 1-2 with the
    transient recorded, 1.3x without it).
The evaluation patch (256 x 4096, jitter seed 0) is never recorded.

    python -m yarnball_amd.train_warmstart --out yarnball_amd/models/warmstart.npy

The model for the learned warm start followed by the strand solve is trained
the same way on the corrections that solver converges to, from more of the
knit (frames 1-12: 10M samples), for 800 epochs:

    python -m yarnball_amd.train_warmstart --strand-every 7 --knit-frames 13 \
        --epochs 800 --out yarnball_amd/models/warmstart_strand.npy

On the evaluation patch the learned + strand method then needs 13.2 sweeps per
substep with steps before sweeps 1, 5, 9, ...  More of the knit was the
lever: frames 1-4 gave 25.9-26.6 and a fallback in about half of the
substeps, frames 1-8 23.8 (steps every 6, an older diagonal).  Recording with
the bench's own schedule (--strand-every 4 --strand-first 1) gave a worse
model: 18.5 sweeps per substep.
"""
import argparse
import os
import sys
import time

import numpy as np

from . import learned, scenes
from .sim import Sim

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_MODEL = os.path.join(HERE, "models", "warmstart.npy")


def record_scene(sc, frames, stride, keep, tol=1e-6, max_sweeps=6000, seed=0, log=print,
                 first_frame=0, strand_every=None, strand_first=0):
    """Run `frames` frames of a scene with the plain solver, substep by substep,
    recording every `stride`-th substep from frame `first_frame` on; keep a
    seeded subsample of `keep` vertices of each recorded substep (all of them
    when keep >= n).  strand_every (not None): the solver runs with the strand
    solve after the plain warm start and before sweeps strand_first,
    strand_first + strand_every, ... (strand_first 0: strand_every), so the
    targets are the corrections that solver converges to."""
    n = sc.n
    frame = np.float32(sc.frame_dt)
    sub = int(np.ceil(frame / np.float32(sc.maxH)))
    h = np.float32(frame / np.float32(sub))
    total = frames * sub
    rec = lambda k: k >= first_frame * sub and k % stride == stride - 1
    nrec = sum(1 for k in range(total) if rec(k))
    sim = Sim(n)
    scenes.load_into(sim, sc)
    sim.setSolver(tol, maxSweeps=max_sweeps, chunk=16, checkEvery=4)
    if strand_every is not None:
        learned.strand(sim, every=strand_every, first=strand_first)
    learned.record(sim, nrec * n)
    t = 0.0
    for f in range(frames):
        if sc.drive_kind == "twist":
            sim.download()
            scenes.drive_frame(sc, sim.verts["pos"], sim.vels, t, float(frame))
            sim.upload()
        for k in range(sub):
            learned.record_enable(sim, rec(f * sub + k))
            sim.advance(h)
        t += float(frame)
    learned.record_enable(sim, False)
    X, C = learned.fetch(sim)
    sim.close()
    if keep < n:
        rng = np.random.default_rng(seed)
        idx = np.concatenate([r * n + np.sort(rng.choice(n, keep, replace=False))
                              for r in range(len(X) // n)])
        X, C = X[idx], C[idx]
    log(f"[train] {sc.name}: {nrec} substeps recorded, {len(X)} samples")
    return X, C


def generate(tol=1e-6, max_sweeps=6000, log=print, strand_every=None, knit_frames=5,
             strand_first=0, knit_stride=1):
    """knit_frames: frames of the knit run (the first, the transient, is not
    recorded)."""
    t0 = time.perf_counter()
    st = dict(strand_every=strand_every, strand_first=strand_first)
    parts = [
        _readme(tol, max_sweeps, log, **st),
        record_scene(scenes.twisted_cable(64, 2048), 8, 4, 8192, tol, max_sweeps, seed=1,
                     log=log, **st),
        record_scene(scenes.knit_patch(256, 4096, seed=7), knit_frames, knit_stride, 24576, tol,
                     max_sweeps, seed=2, log=log, first_frame=1, **st),
    ]
    X = np.concatenate([p[0] for p in parts])
    C = np.concatenate([p[1] for p in parts])
    log(f"[train] {len(X)} samples in {time.perf_counter() - t0:.1f}s")
    return X, C


def _readme(tol, max_sweeps, log, seconds=0.5, strand_every=None, strand_first=0):
    sc = scenes.readme_scene()
    sc.frame_dt = seconds
    return record_scene(sc, 1, 1, sc.n, tol, max_sweeps, log=log, strand_every=strand_every,
                        strand_first=strand_first)


def fit(X, C, epochs=400, batch=16384, lr=1e-3, seed=0, log=print):
    """Train on recorded (features X, converged corrections C)."""
    T = learned.residual_targets(X, C)
    p0 = learned.init_params(X, T, seed=seed)
    t0 = time.perf_counter()
    p, hist = learned.train(X, T, p0, epochs=epochs, batch=batch, lr=lr, seed=seed)
    # loss is per sample in normalised output units; the zero predictor scores ~3
    log(f"[train] {epochs} epochs in {time.perf_counter() - t0:.1f}s, loss {hist[0]:.4f} -> "
        f"{hist[-1]:.4f}")
    return p, hist


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_MODEL)
    ap.add_argument("--epochs", type=int, default=400)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--strand-every", type=int, default=None,
                    help="record with the strand solve (after the warm start and before every "
                         "k-th sweep): the model for learned + strand")
    ap.add_argument("--strand-first", type=int, default=0,
                    help="with --strand-every: the sweep before which the first such step "
                         "runs (0: the strand-every-th)")
    ap.add_argument("--knit-frames", type=int, default=5,
                    help="frames of the training knit (frame 0 is run, not recorded)")
    ap.add_argument("--knit-stride", type=int, default=1,
                    help="record every k-th substep of the training knit (device memory: "
                         "each recorded substep holds all 1M vertices' rows)")
    args = ap.parse_args()
    log = lambda m: print(m, file=sys.stderr, flush=True)
    X, T = generate(args.tol, log=log, strand_every=args.strand_every,
                    knit_frames=args.knit_frames, strand_first=args.strand_first,
                    knit_stride=args.knit_stride)
    p, hist = fit(X, T, epochs=args.epochs, log=log)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    np.save(args.out, p)
    log(f"[train] saved {args.out}")


if __name__ == "__main__":
    main()
