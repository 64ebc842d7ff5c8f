// kernels_sweep.hip — per-substep solver kernels of the stepper on gfx950.
//
//   k_step_limit   per-segment no-tunnelling step bound   (collision.cu:121-159)
//   k_sweep        one block-Jacobi sweep: Newton step on dx (cosserat.cu:20-207)
//                  fused with the closed-form frame update (cosserat.cu:209-253)
//   k_end          velocity update + advection             (iteration.cu:48-61)
//   k_upload / k_download   AoS <-> SoA transposes        (YarnBall.cu:239-261)
//
// Schedule: strict Jacobi.  Sweep k reads dx[k&1], q[k&1] and writes dx[~k&1];
// the frame update uses the new dx and the old frames and writes q[~k&1].  The
// reference updates both arrays in place (asynchronous across blocks); the
// double buffers make every iterate reproducible.
//
// This translation unit holds only continuous (tolerance-compared) arithmetic
// and is built with fast single-precision divide / sqrt.
#include "yb_internal.h"
#include "ybmath.h"

namespace yb {

__device__ __forceinline__ v3 xyz(float4 a) { return v3{a.x, a.y, a.z}; }
__device__ __forceinline__ rot asrot(float4 a) { return rot{a.x, a.y, a.z, a.w}; }
__device__ __forceinline__ float4 f4(v3 a, float w) { return make_float4(a.x, a.y, a.z, w); }
__device__ __forceinline__ float4 f4r(rot a) { return make_float4(a.x, a.y, a.z, a.w); }

// Hardware reciprocal, reciprocal square root and logarithm (1 ulp).  The
// compiler's fast divide adds a range-scaling step for denominators above
// 2^96 and its logf a compensated log2 -> ln product; neither is needed here:
// the denominators are lengths, determinants and masses far below 2^96, and
// ln is taken only of the barrier distance in [1e-3, 1].  A zero denominator
// still gives inf, as the division did.
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fln(float x) { return __builtin_amdgcn_logf(x) * 0.693147181f; }
__device__ __forceinline__ rot rnormalize_fast(rot a) {
    return rscale(a, __builtin_amdgcn_rsqf(rdot(a, a)));
}
// rotate(r, e1) (Rotor.h:96-105 with v = (1,0,0)) written out, so that no
// product with a constant 0 is left for the compiler (it may not drop x * 0).
__device__ __forceinline__ v3 rotate_e1(rot r) {
    return v3{(r.w * r.w + r.x * r.x) - (r.y * r.y + r.z * r.z), 2.0f * (r.x * r.y + r.z * r.w),
              2.0f * (r.x * r.z - r.y * r.w)};
}
// (V b) e1 for the pure rotor V = (v, 0) and e1 = Rotor(1) = (1,0,0 | 0): the two
// rotor products (Rotor.h:185-188) written out.  With m = V b,
//   m = (b.w v + v x b.xyz | -v . b.xyz),   m e1 = (m.w, m.z, -m.y | -m.x).
__device__ __forceinline__ rot vb_e1(v3 v, rot b) {
    const v3 bq = rq(b);
    const v3 mq = b.w * v + cross(v, bq);
    return rot{-dot(v, bq), mq.z, -mq.y, -mq.x};
}

// Closed-form frame update of segment i (cosserat.cu:209-253): q0 its frame,
// e its edge in the configuration x + dx, kl = -2 kStretch_i / l_i, and the
// neighbouring frames with the rest rotors between them (qr: of segments i and
// i + 1, qrPrev: of segments i - 1 and i).
__device__ __forceinline__ rot frame_update(rot q0, v3 e, float kl, uint32_t flags, rot qPrev,
                                            rot qrPrev, rot qNext, rot qr) {
    const v3 vv = e * kl;
    rot b = rot{0.f, 0.f, 0.f, 0.f};
    if (flags & YB_FLAG_HAS_PREV) {
        const float sg = rdot(rmul(rinv(qPrev), q0), qrPrev) > 0.0f ? 1.0f : -1.0f;
        b = rscale(rmul(qPrev, qrPrev), sg);
    }
    if (flags & YB_FLAG_HAS_NEXT_ORIENTATION) {
        const float sg = rdot(rmul(rinv(q0), qNext), qr) > 0.0f ? 1.0f : -1.0f;
        b = radd(b, rscale(rmul(qNext, rinv(qr)), sg));
    }
    const float lambda = len(vv) + rlen(b);
    return rnormalize_fast(radd(vb_e1(vv, b), rscale(b, lambda)));
}

// H^{-1} f through the adjugate (sym_solve with the hardware reciprocal).
__device__ __forceinline__ v3 sym_solve_fast(const sym3& m, v3 f) {
    const float a = m.a[0], b = m.a[1], c = m.a[2], d = m.a[3], e = m.a[4], g = m.a[5];
    const float A00 = b * c - g * g;
    const float A01 = e * g - d * c;
    const float A02 = d * g - b * e;
    const float A11 = a * c - e * e;
    const float A12 = d * e - a * g;
    const float A22 = a * b - d * d;
    const float inv = frcp((a * A00 + d * A01) + e * A02);
    return v3{((A00 * f.x + A01 * f.y) + A02 * f.z) * inv,
              ((A01 * f.x + A11 * f.y) + A12 * f.z) * inv,
              ((A02 * f.x + A12 * f.y) + A22 * f.z) * inv};
}

// ---------------------------------------------------------------------------
// Step limit (collision.cu:121-159): the largest move for which the padded
// broadphase query still sees every possible contact, and at most 45% of the
// distance to every listed contact.  The same contact distances also feed the
// interpenetration check the reference declares but never sets (YarnBall.h:121):
// a listed pair closer than 2 * radius raises WARNING_SEGMENT_INTERPENETRATION,
// and the smallest distance / (2 * radius) seen is kept in *minSep.  A substep
// that detects contacts gets all of this from the query, which has just
// measured the same distances; this kernel serves the substeps that reuse
// their lists.

template <bool WRITE_LIMIT>
__global__ __launch_bounds__(256) void k_step_limit(const DevParams* __restrict__ P,
                                                    const float4* __restrict__ pos,
                                                    const float4* __restrict__ prm,
                                                    float4* __restrict__ dx,
                                                    const int32_t* __restrict__ numCols,
                                                    const int32_t* __restrict__ cols,
                                                    int32_t* __restrict__ err,
                                                    float* __restrict__ minSep) {
    const int n = P->n;
    const int v = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    float minDist = INFINITY;
    float minContact = INFINITY;
    if (v < n) {
        const uint32_t flags = __float_as_uint(prm[v].w);
        if (flags & YB_FLAG_HAS_NEXT) {
            const v3 a0 = xyz(pos[v]);
            const v3 a1 = xyz(pos[v + 1]) - a0;
            const int nc = numCols[v];
            for (int i = 0; i < nc; i++) {
                const int o = cols[v + (size_t)n * i];
                const v3 b0 = xyz(gather4(pos, o)) - a0;
                const v3 b1 = xyz(gather4<true>(pos, o)) - a0;
                float u, w;
                closest_uv(v3{0.f, 0.f, 0.f}, a1, b0, b1, u, w);
                minContact = fminf(minContact, len(u * a1 - mixv(b0, b1, w)));
            }
            if (WRITE_LIMIT && P->useStepSizeLimit) {
                minDist = P->detectionRadius * (P->detectionScaler - 1.0f);
                if (P->useVelocityRadius) minDist += len(xyz(dx[v]));
                minDist = fminf(minDist, 0.45f * minContact);
            }
        }
        if (WRITE_LIMIT) dx[v].w = minDist;
        if (minContact < 2.0f * P->radius) err[1] = YB_WARNING_SEGMENT_INTERPENETRATION;
    }
    reduce_min_sep(minContact / (2.0f * P->radius), minSep);
}

__device__ __forceinline__ void wave_residual(float numU, float numF, double2* partial, int wave,
                                              int n);

// Closest points of segment e = (0 -> e) and (B0 -> B0 + A1), specialised from
// closest_uv (Common.h:318-334) with the segment's |e|^2 and 1/|e|^2 hoisted
// out of the contact loop.  The two reprojections are the same as in
// closest_uv, with their dot products expanded over the ones the line solve
// already has:  (B0 + A1 v).e = e.B0 + v (e.A1),  (e u - B0).A1 = u (e.A1) - A1.B0.
__device__ __forceinline__ void closest_e(v3 e, float a, float inva, v3 B0, v3 A1, float& u,
                                         float& w) {
    const float b = dot(e, A1);
    const float c = len2(A1);
    const float eB = dot(e, B0);
    const float AB = dot(A1, B0);
    float vy = (b * eB - a * AB) * frcp(a * c - b * b);
    if (!isfinite(vy)) vy = 0.5f;
    vy = clampf(vy, 0.f, 1.f);
    const float ux = clampf((eB + vy * b) * inva, 0.f, 1.f);
    vy = clampf((ux * b - AB) * frcp(c), 0.f, 1.f);
    u = ux;
    w = vy;
}

// A strand step between sweeps does nothing once the substep has converged
// (before `sweep`).  Every kernel of one step reads the same control word, so
// they agree.
__device__ __forceinline__ bool strand_skip(const SolverCtl* ctl, int sweep) {
    return ctl && ctl->doneAt < sweep;
}

// Outputs of the strand solve's tile kernel (line mode of the sweep body).
struct LineOut {
    float4* g;   // per vertex: particular solution xyz, left-boundary factor p
    float2* ql;  // per vertex: right-boundary factor q, step limit (-1: not a row)
    float* sys;  // optional (tests): per vertex diagonal, coupling to the next, rhs xyz
};

// Parallel cyclic reduction over the 64 rows of a wave, one row per lane:
//   A x_l - Lo x_(l-1) - Up x_(l+1) = B_l      (five right-hand sides)
// Rows outside the wave count as identity rows.  After strides 1..32 every
// row is decoupled and x_l = B_l / A.  Each step substitutes the equations of
// rows l - s and l + s into row l; the system is diagonally dominant, so no
// pivoting is needed.
__device__ __forceinline__ void pcr64(float& A, float& Lo, float& Up, float B[5]) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) {
        const bool hl = lane >= s, hr = lane + s < 64;
        const float Am = __shfl_up(A, s, 64), Lm = __shfl_up(Lo, s, 64), Um = __shfl_up(Up, s, 64);
        const float Ap = __shfl_down(A, s, 64), Lp = __shfl_down(Lo, s, 64),
                    Upp = __shfl_down(Up, s, 64);
        const float al = hl ? Lo / Am : 0.0f;
        const float be = hr ? Up / Ap : 0.0f;
        A = A - al * (hl ? Um : 0.0f) - be * (hr ? Lp : 0.0f);
        Lo = hl ? al * Lm : 0.0f;
        Up = hr ? be * Upp : 0.0f;
#pragma unroll
        for (int k = 0; k < 5; k++) {
            const float bm = __shfl_up(B[k], s, 64), bp = __shfl_down(B[k], s, 64);
            B[k] += al * bm + be * bp;
        }
    }
}

// ---------------------------------------------------------------------------
// One block-Jacobi sweep: Newton step on dx (cosserat.cu:20-207) followed by
// the closed-form frame update (cosserat.cu:209-253), fused.
//
// Vertex i:  H = m/h^2 I + glue + stretch(seg i-1, seg i) + barrier/friction
//            f = m/h^2 ((y-x) - dx) + ...
//            dx' = dx + accelerationRatio H^{-1} f, clamped to the step limit.
// Segment i: v = -2 kStretch e'/l with e' from dx', b = neighbour frames
//            rotated by the rest Darboux rotors (sign-corrected), then
//            q' = normalize((v b e1) + (|v| + |b|) b).
//
// Wave tile: lane l works on vertex 62 w + l - 1.  Lanes 1..62 own their
// vertex (write dx' and q').  Lane 0 is a halo that only hands its segment's
// endpoint-1 terms and its frame to lane 1.  Lane 63 computes dx' of the next
// tile's first vertex without writing it, so that lane 62's frame update has
// dx'(i+1); it also supplies q(i+1).  All exchange is cross-lane shifts --
// no LDS, no barrier (the reference uses a 2-sector LDS reduction).
//
// Modes: kSweep -- the sweep.  kLine -- the strand solve's tile part: the
// same local systems are assembled, then solved along the tile's strands
// instead of vertex by vertex (the line tail below); nothing of the sweep's own
// output is written.  kEval -- only the residual partials of the input iterate.
// kSweep and kEval write, when `partial` is set, the per-wave sums of the
// squared update |dx' - dx|^2 (kSweep) and of the squared residual force |r|^2
// at the input iterate (r: see the Newton step below).  kLineStart -- the
// first strand step of a substep, from the plain warm start: each lane first
// refreshes its segment's frame from the input dx (the sweep's frame update at
// dx' = dx) and writes it to qOut, then assembles and solves the strand system
// as kLine does, at the refreshed frames, and writes kEval's residual partials
// of the input iterate at the input frames; with dxOut set, it also copies the
// input dx there (the plain warm start a learned session measures its
// correction from).
constexpr int kSweep = 0, kLine = 1, kEval = 2, kLineStart = 3;

// Batched scenes in the residual-targeted schedule (kBatched): a wave tile
// holds at most two scenes (A: the scene of its first owned vertex, B: of its
// last); a lane of a scene that has converged computes but writes nothing,
// and a tile whose scenes have both converged exits at once (a strand tile
// then writes empty rows, so that its neighbours read finite values).  The
// residual partials are kept per scene side: sums over the lanes of scene A
// at partial[wave], of scene B at partial[tiles + wave].

__device__ __forceinline__ void wave_residual_scenes(float numU, float numF, bool sideA,
                                                     double2* partial, int wave, int n) {
    float uA = sideA ? numU : 0.0f, fA = sideA ? numF : 0.0f;
    float uB = sideA ? 0.0f : numU, fB = sideA ? 0.0f : numF;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        uA += __shfl_xor(uA, off, 64);
        fA += __shfl_xor(fA, off, 64);
        uB += __shfl_xor(uB, off, 64);
        fB += __shfl_xor(fB, off, 64);
    }
    const int tiles = (n + 61) / 62;
    if ((threadIdx.x & 63) == 0 && wave < tiles) {
        partial[wave] = make_double2((double)uA, (double)fA);
        partial[tiles + wave] = make_double2((double)uB, (double)fB);
    }
}

// Sum of nb partials by one wave in a fixed order of additions.  Eight running
// sums per lane keep eight loads in flight (the reduction is latency-bound: one
// wave reads 17k partials on the 1M-vertex knit).
__device__ __forceinline__ double2 reduce_partials(const double2* __restrict__ partial, int nb) {
    const int lane = threadIdx.x & 63;
    double2 acc[8];
#pragma unroll
    for (int j = 0; j < 8; j++) acc[j] = make_double2(0.0, 0.0);
    int k = lane;
    for (; k + 7 * 64 < nb; k += 8 * 64) {
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const double2 t = partial[k + 64 * j];
            acc[j].x += t.x;
            acc[j].y += t.y;
        }
    }
    for (; k < nb; k += 64) {
        const double2 t = partial[k];
        acc[0].x += t.x;
        acc[0].y += t.y;
    }
    double a = ((acc[0].x + acc[1].x) + (acc[2].x + acc[3].x)) +
               ((acc[4].x + acc[5].x) + (acc[6].x + acc[7].x));
    double b = ((acc[0].y + acc[1].y) + (acc[2].y + acc[3].y)) +
               ((acc[4].y + acc[5].y) + (acc[6].y + acc[7].y));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
    }
    return make_double2(a, b);
}

// Stopping decision of check sweep `sweep` from its sums a = (sum |dx' - dx|^2,
// sum |r|^2) and the Newton-residual denominator denF (one thread).  denF ==
// kNoNewton: the substep records no Newton residual (its stopping test is then
// the update residual).
__device__ __forceinline__ void decide(SolverCtl* __restrict__ ctl, double2 a, double denF,
                                       int sweep, float tol, int maxSweeps) {
    ctl->denomF = denF;
    const bool hasF = denF != kNoNewton;
    const double denU = ctl->denom;
    const float rU = denU > 0.0 ? (float)sqrt(a.x / denU) : 0.0f;
    const float rF = denF > 0.0 ? (float)sqrt(a.y / denF) : 0.0f;
    const float r = ctl->kind ? rF : rU;
    ctl->resid = r;
    ctl->fres = rF;
    if (r <= tol || sweep == maxSweeps - 1) {
        ctl->doneAt = sweep;
        ctl->fin = (sweep & 1) ^ 1;
        ctl->sweeps += (unsigned long long)(sweep + 1);
        ctl->substeps += 1ull;
        if (sweep + 1 > ctl->maxUsed) ctl->maxUsed = sweep + 1;
        if (r > tol) ctl->notConverged += 1;
        if (r > ctl->maxResid) ctl->maxResid = r;
        if (hasF) {
            if (rF > ctl->maxFres) ctl->maxFres = rF;
            ctl->sumFres += (double)rF;
            ctl->newtonSubsteps += 1ull;
        }
    }
}

// The in-sweep check (plain and learned sessions, which run hundreds of sweeps
// per substep): the first wave of the sweep after a check sweep reduces that
// sweep's partials in a fixed order and decides, while the rest of the launch
// runs.  It costs no launch of its own; a converged substep is recognised one
// sweep late (that sweep's output is not used).
struct CheckPrev {
    const double2* partial;    // the previous (check) sweep's partials, or null
    const double2* denomPart;  // the prologue's denominator partials, or null
    float tol;
    int maxSweeps;
};
__device__ __forceinline__ void residual_wave(SolverCtl* __restrict__ ctl, CheckPrev cp, int nb,
                                              int sweep) {
    const double2 a = reduce_partials(cp.partial, nb);
    double denF = ctl->denomF;  // uniform across the wave
    if (denF == -1.0) denF = cp.denomPart ? reduce_partials(cp.denomPart, nb).y : a.y;
    if ((threadIdx.x & 63) == 0) decide(ctl, a, denF, sweep, cp.tol, cp.maxSweeps);
}

// Tile-list kernels of a residual-targeted batch (after its first check): a
// fixed grid of waves loops over the listed tiles -- those of the scenes
// still running -- so that sweeps late in a substep, when few scenes remain,
// neither dispatch nor retire a wave per tile of the whole batch.  tileCount
// < 0 lists every tile.
__device__ __forceinline__ int listed_tiles(const int32_t* __restrict__ tileCount, int tiles) {
    const int c = tileCount[0];
    return c < 0 ? tiles : c;
}
__device__ __forceinline__ int listed_tile(const int32_t* __restrict__ tileList,
                                           const int32_t* __restrict__ tileCount, int idx) {
    return tileCount[0] < 0 ? idx : tileList[idx];
}

template <int kMode, bool kBatched = false>
__device__ __forceinline__ void sweep_body(const DevParams* __restrict__ P,
                                           const float4* __restrict__ pos,
                                           const float4* __restrict__ prm,
                                           const float4* __restrict__ yv,
                                           const float4* __restrict__ qIn,
                                           float4* __restrict__ qOut,
                                           const float4* __restrict__ qRest,
                                           const float4* __restrict__ dxIn,
                                           float4* __restrict__ dxOut,
                                           const int32_t* __restrict__ numCols,
                                           const int32_t* __restrict__ cols,
                                           SolverCtl* __restrict__ ctl, int sweep,
                                           double2* __restrict__ partial, LineOut lo,
                                           SceneArgs sa = SceneArgs{nullptr, nullptr},
                                           int tileIn = -1,
                                           CheckPrev cp = CheckPrev{nullptr, nullptr, 0.0f, 0}) {
    // Residual-targeted mode: sweeps after convergence do nothing.
    if (ctl && ctl->doneAt < sweep) return;
    const int n = P->n;
    // In-sweep check of the previous (check) sweep, by the launch's first wave.
    if (!kBatched && cp.partial && blockIdx.x == 0 && threadIdx.x < 64)
        residual_wave(ctl, cp, (n + 61) / 62, sweep - 1);
    const int lane = threadIdx.x & 63;
    // the wave tile: given by a tile-list kernel, else from the block index
    const int wave = tileIn >= 0 ? tileIn
                                 : (int)((xcd_tile(blockIdx.x, gridDim.x) * blockDim.x +
                                          threadIdx.x) >> 6);
    const int v = wave * 62 + lane - 1;
    const bool in = (v >= 0) && (v < n);

    // batched: the tile's scenes, whether each still runs, this lane's side
    bool live = true, sideA = true;
    if constexpr (kBatched) {
        const int scA = sa.sceneOf[min(wave * 62, n - 1)];
        const int scB = sa.sceneOf[min(wave * 62 + 61, n - 1)];
        const bool liveA = sa.ctl[scA].doneAt >= sweep;
        const bool liveB = sa.ctl[scB].doneAt >= sweep;
        if (!liveA && !liveB) {
            if constexpr (kMode == kLine || kMode == kLineStart) {
                if (in && lane >= 1 && lane <= 62) {
                    lo.g[v] = make_float4(0.f, 0.f, 0.f, 0.f);
                    lo.ql[v] = make_float2(0.0f, -1.0f);
                }
            }
            return;
        }
        sideA = sa.sceneOf[min(max(v, 0), n - 1)] == scA;
        live = sideA ? liveA : liveB;
    }

    const float h = P->h;
    const float damping = P->dampingH;

    v3 f2 = v3{0.f, 0.f, 0.f};
    sym3 H2 = sym_zero();
    float segLimit = INFINITY;
    float4 qr4 = make_float4(0.f, 0.f, 0.f, 0.f);
    v3 p1 = v3{0.f, 0.f, 0.f};
    float invl = 0.0f;  // 1 / rest length of segment i (set when it exists)

    // Lanes outside [0, n) (the halo before vertex 0, the tail) read a clamped
    // vertex with its flags cleared: they add no segment terms to their
    // neighbours and write nothing, and no code branches on the range.
    const int vc = min(max(v, 0), n - 1);
    const float4 p4 = pos[vc];
    const float4 r4 = prm[vc];
    const float4 d4 = dxIn[vc];
    const float4 q4 = qIn[vc];
    const uint32_t flags = in ? __float_as_uint(r4.w) : 0u;
    const int cid = in ? __float_as_int(r4.z) : -1;
    const v3 x0 = xyz(p4);
    const v3 dx = xyz(d4);

    // kLineStart: the segment's frame refreshed from the input dx, stored at
    // once; lane 0 reads the frame and rest rotor before it from memory (that
    // vertex is not in the wave).  dl = k (e1(qs) - e1(q)) is by how much the
    // stretch force at the input frame exceeds the one at qs.
    rot qs = asrot(q4);
    v3 dl = v3{0.f, 0.f, 0.f};
    if constexpr (kMode == kLineStart) {
        if (in) qr4 = qRest[v];
        rot qP = rot{__shfl_up(q4.x, 1, 64), __shfl_up(q4.y, 1, 64), __shfl_up(q4.z, 1, 64),
                     __shfl_up(q4.w, 1, 64)};
        rot qrP = rot{__shfl_up(qr4.x, 1, 64), __shfl_up(qr4.y, 1, 64), __shfl_up(qr4.z, 1, 64),
                      __shfl_up(qr4.w, 1, 64)};
        const rot qN = rot{__shfl_down(q4.x, 1, 64), __shfl_down(q4.y, 1, 64),
                           __shfl_down(q4.z, 1, 64), __shfl_down(q4.w, 1, 64)};
        if (lane == 0 && (flags & YB_FLAG_HAS_PREV)) {
            qP = asrot(qIn[v - 1]);
            qrP = asrot(qRest[v - 1]);
        }
        if (!(flags & YB_FLAG_FIX_ORIENTATION) && (flags & YB_FLAG_HAS_NEXT)) {
            const v3 e = (xyz(pos[vc + 1]) - x0) + (xyz(dxIn[vc + 1]) - dx);
            qs = frame_update(qs, e, -2.0f * r4.y * frcp(r4.x), flags, qP, qrP, qN, asrot(qr4));
        }
        if (in && lane >= 1 && lane <= 62 && live) {
            qOut[v] = f4r(qs);
            if (dxOut) dxOut[v] = d4;
        }
    }

    // Inertia (cosserat.cu:40-44); yv holds y - x.
    const float mh2 = frcp(p4.w * h * h);
    sym3 H = sym_diag(mh2);
    v3 f = mh2 * (xyz(yv[vc]) - dx);

    // Glue spring to a connected vertex (cosserat.cu:47-53).
    if (cid >= 0) {
        const float stiffness = 40.0f;
        const v3 pc = xyz(pos[cid]);
        const v3 dc = xyz(dxIn[cid]);
        f -= stiffness * (((x0 - pc) + (dx - dc)) + damping * dx);
        sym_add_diag(H, stiffness * (1.0f + damping));
    }

    if (flags & YB_FLAG_HAS_NEXT) {
        p1 = xyz(pos[vc + 1]);
        const v3 p1dx = xyz(dxIn[vc + 1]);
        segLimit = d4.w;

        // Cosserat stretch-shear (cosserat.cu:67-81).
        invl = frcp(r4.x);
        const v3 e = (p1 - x0) + (p1dx - dx);
        const v3 c = e * invl - rotate_e1(qs);
        const float k = r4.y * invl;
        if constexpr (kMode == kLineStart) dl = k * (rotate_e1(qs) - rotate_e1(asrot(q4)));
        float d = k * invl;
        f += k * c - (damping * d) * dx;
        f2 = -k * c - (damping * d) * p1dx;
        d *= 1.0f + damping;
        sym_add_diag(H, d);
        sym_add_diag(H2, d);

        // IPC barrier + lagged friction over this segment's contacts
        // (cosserat.cu:83-148).
        const int nc = numCols[vc];
        if (nc > 0) {
            const float fricK = P->kFriction;
            const float invb = 1.0f / P->barrierThickness;
            const float radius2 = 2.0f * P->radius;
            const float fricMu = P->frictionCoeff;
            const float kCol = P->kCollision * invb;
            const float ea = len2(e);
            const float inva = frcp(ea);
            // friction's isotropic Hessian part f1 w^2 I, summed as scalars
            float diag0 = 0.0f, diag1 = 0.0f;
            const v3 dseg = p1dx - dx;  // this segment's relative displacement
            int o = cols[vc];
            for (int i = 0; i < nc; i++) {
                // next partner id fetched one iteration ahead; slot-major
                // lists span n * 128 entries, past the int range
                const int oNext = (i + 1 < nc) ? cols[vc + (size_t)n * (i + 1)] : o;
                const v3 b0 = xyz(gather4(pos, o));
                const v3 b1 = xyz(gather4<true>(pos, o));
                const v3 db0 = xyz(gather4(dxIn, o));
                const v3 db1 = xyz(gather4<true>(dxIn, o));
                o = oNext;
                // partner segment relative to this vertex, in the
                // current configuration x + dx: start B0, direction A1
                const v3 ddb = db1 - db0;
                const v3 B0 = (b0 - x0) + (db0 - dx);
                const v3 A1 = (b1 - b0) + ddb;
                float u, w;
                closest_e(e, ea, inva, B0, A1, u, w);
                // closest-point difference (pos + dx of this segment minus
                // the partner's) and its length
                v3 nrm = e * u - (B0 + A1 * w);
                const float dist = len(nrm);
                float dd = (dist - radius2) * invb;
                if (dd > 1.0f) continue;  // outside the barrier
                dd = __builtin_amdgcn_fmed3f(dd, 1e-3f, 1.0f);  // max(dd, 1e-3), dd <= 1
                nrm *= frcp(dist);
                const float w1 = u;
                const float w0 = 1.0f - u;

                // b(d) = -(d-1)^2 ln d; ff = -kCol b'(d), dH = kCol/b b''(d)
                const float invd = frcp(dd);
                const float logd = fln(dd);
                float dH = (-3.0f + (2.0f + invd) * invd - 2.0f * logd) * kCol * invb;
                const float ff = -(1.0f - dd) * (dd - 1.0f + 2.0f * dd * logd) * invd * kCol;
                f += (ff * w0 - damping * dH * w0 * w0 * dot(nrm, dx)) * nrm;
                f2 += (ff * w1 - damping * dH * w1 * w1 * dot(nrm, p1dx)) * nrm;
                dH *= 1.0f + damping;

                // Lagged friction on the tangential relative displacement;
                // its Hessian f1 w^2 (I - n n^T) is merged with the
                // barrier's dH w^2 n n^T into one rank-1 update + diagonal.
                const v3 ddpos = (dx + dseg * u) - (db0 + ddb * w);
                const v3 uu = ddpos - dot(nrm, ddpos) * nrm;
                const float ul = len(uu);
                float f1 = 0.0f;
                if (ul > 0.0f) {
                    f1 = fminf(fricK, fricMu * ff * frcp(ul));
                    f -= (f1 * w0) * uu;
                    f2 -= (f1 * w1) * uu;
                }
                const sym3 op = sym_outer(nrm);
                const float s0 = w0 * w0, s1 = w1 * w1;
                sym_axpy(H, op, (dH - f1) * s0);
                sym_axpy(H2, op, (dH - f1) * s1);
                diag0 += f1 * s0;
                diag1 += f1 * s1;
            }
            sym_add_diag(H, diag0);
            sym_add_diag(H2, diag1);
        }
    }

    // The rest rotor is needed only from here on: loading it after the contact
    // loop keeps it out of the loop's register budget.
    if constexpr (kMode != kLineStart) {
        if (in) qr4 = qRest[v];
    }

    // Segment (i-1)'s endpoint-1 terms, frame and rest rotor (cosserat.cu:175-193).
    const float pf0 = __shfl_up(f2.x, 1, 64);
    const float pf1 = __shfl_up(f2.y, 1, 64);
    const float pf2 = __shfl_up(f2.z, 1, 64);
    float ph[6];
#pragma unroll
    for (int k = 0; k < 6; k++) ph[k] = __shfl_up(H2.a[k], 1, 64);
    const float pl = __shfl_up(segLimit, 1, 64);
    const rot qPrev = rot{__shfl_up(q4.x, 1, 64), __shfl_up(q4.y, 1, 64), __shfl_up(q4.z, 1, 64),
                          __shfl_up(q4.w, 1, 64)};
    const rot qrPrev = rot{__shfl_up(qr4.x, 1, 64), __shfl_up(qr4.y, 1, 64),
                           __shfl_up(qr4.z, 1, 64), __shfl_up(qr4.w, 1, 64)};
    v3 dlPrev = v3{0.f, 0.f, 0.f};
    if constexpr (kMode == kLineStart)
        dlPrev = v3{__shfl_up(dl.x, 1, 64), __shfl_up(dl.y, 1, 64), __shfl_up(dl.z, 1, 64)};

    // Newton step.
    v3 dxNew = xyz(d4);
    float num = 0.0f;   // |dx' - dx|^2
    float numF = 0.0f;  // |residual force|^2 of this vertex at the input iterate
    float limit = segLimit;
    const bool owner = in && lane >= 1 && lane <= 62;
    if (in && lane != 0) {
        if (flags & YB_FLAG_HAS_PREV) {
            limit = fminf(limit, pl);
            f += v3{pf0, pf1, pf2};
#pragma unroll
            for (int k = 0; k < 6; k++) H.a[k] += ph[k];
        }
        if (p4.w != 0.0f && (kMode == kSweep || partial)) {
            // Local 3x3 solve (cosserat.cu:195-205).  The residual force is f,
            // or, where the step limit shortens the step, H (dx' - dx) /
            // accelerationRatio: the part of f the limit does not absorb, which
            // vanishes at a clamped fixed point of the sweeps.  kLineStart
            // takes it at the input frames: only the stretch terms of the two
            // segments at the vertex depend on the frames (H does not).
            v3 fr = f;
            if constexpr (kMode == kLineStart) {
                fr += dl;
                if (flags & YB_FLAG_HAS_PREV) fr -= dlPrev;
            }
            dxNew += P->accelerationRatio * sym_solve_fast(H, fr);
            const float l = len(dxNew);
            v3 r = fr;
            if (l > limit && l > 0.0f) {
                dxNew *= limit / l;
                r = sym_apply(H, dxNew - xyz(d4)) * frcp(P->accelerationRatio);
            }
            num = len2(dxNew - xyz(d4));
            numF = len2(r);
        }
    }
    if constexpr (kMode == kEval) {
        if constexpr (kBatched)
            wave_residual_scenes(0.0f, owner ? numF : 0.0f, sideA, partial, wave, n);
        else
            wave_residual(0.0f, owner ? numF : 0.0f, partial, wave, n);
        if (!qOut) return;
        // frame refresh from the input iterate itself (the closed-form frame
        // update below, with dx' = dx)
        dxNew = dx;
    }

    if constexpr (kMode == kLineStart) {
        if constexpr (kBatched)
            wave_residual_scenes(0.0f, owner ? numF : 0.0f, sideA, partial, wave, n);
        else
            wave_residual(0.0f, owner ? numF : 0.0f, partial, wave, n);
    }

    if constexpr (kMode == kLine || kMode == kLineStart) {
        // Strand solve, tile part: the tile's owned vertices (lanes 1..62)
        // form a tridiagonal system along their strands -- on the diagonal a
        // scalar stiffness of each local Hessian, between neighbours the
        // stretch stiffness k/l^2 of each segment between its two free ends --
        // with the vertices just outside the tile (lanes 0 and 63) as unknown
        // boundary values.  Per vertex: x = g + p x_left + q x_right.
        // The scalar is the curvature along the vertex's force, f.H.f / f.f,
        // with which x = f / a is the exact minimiser of the local quadratic
        // along f.  The isotropic trace(H)/3 (used when f vanishes) is smaller
        // wherever the contact stiffness sits in the force direction: the
        // step then overshoots there and the sweeps that follow must take the
        // excess back, so steps could not come closer together than about 7
        // sweeps without losing ground.
        // A vertex the step limit holds (a sweep clamped it) keeps a nonzero
        // force at the sweeps' fixed point; as a row it would push its
        // neighbours every strand step and the sweeps would push them back.
        // It is held fixed instead, like a pinned vertex, so the converged
        // state of the sweeps is also a fixed point of the strand step.
        const bool free = in && p4.w != 0.0f && len2(dx) < limit * limit * (1.0f - 2e-5f);
        const bool row = free && lane >= 1 && lane <= 62;
        const float nextFree = __shfl_down(free ? 1.0f : 0.0f, 1, 64);
        const float cpl = (free && (flags & YB_FLAG_HAS_NEXT) && nextFree != 0.0f)
                              ? r4.y * invl * invl
                              : 0.0f;
        const float cPrev = __shfl_up(cpl, 1, 64);
        const float ff = len2(f);
        const float fHf = H.a[0] * f.x * f.x + H.a[1] * f.y * f.y + H.a[2] * f.z * f.z +
                          2.0f * (H.a[3] * f.x * f.y + H.a[4] * f.x * f.z + H.a[5] * f.y * f.z);
        const float diag = ff > 1e-30f ? fHf / ff : (H.a[0] + H.a[1] + H.a[2]) * (1.0f / 3.0f);
        float A = row ? diag : 1.0f;
        float Lo = (row && lane >= 2) ? cPrev : 0.0f;
        float Up = (row && lane <= 61) ? cpl : 0.0f;
        float B[5] = {row ? f.x : 0.0f, row ? f.y : 0.0f, row ? f.z : 0.0f,
                      (row && lane == 1) ? cPrev : 0.0f, (row && lane == 62) ? cpl : 0.0f};
        pcr64(A, Lo, Up, B);
        if (lo.sys && in && lane >= 1 && lane <= 62) {
            float* o = lo.sys + 5 * (size_t)v;
            o[0] = row ? diag : 1.0f;
            o[1] = cpl;
            o[2] = row ? f.x : 0.0f;
            o[3] = row ? f.y : 0.0f;
            o[4] = row ? f.z : 0.0f;
        }
        if (in && lane >= 1 && lane <= 62) {
            const float ia = 1.0f / A;
            lo.g[v] = row ? make_float4(B[0] * ia, B[1] * ia, B[2] * ia, B[3] * ia)
                          : make_float4(0.f, 0.f, 0.f, 0.f);
            lo.ql[v] = make_float2(row ? B[4] * ia : 0.0f, row ? limit : -1.0f);
        }
        return;
    }

    // Next vertex's new dx and the next segment's frame.
    const v3 dxNext = v3{__shfl_down(dxNew.x, 1, 64), __shfl_down(dxNew.y, 1, 64),
                         __shfl_down(dxNew.z, 1, 64)};
    const rot qNext = rot{__shfl_down(q4.x, 1, 64), __shfl_down(q4.y, 1, 64),
                          __shfl_down(q4.z, 1, 64), __shfl_down(q4.w, 1, 64)};

    if (owner && live) {
        if constexpr (kMode == kSweep) dxOut[v] = f4(dxNew, d4.w);
        rot q0 = asrot(q4);
        if (!(flags & YB_FLAG_FIX_ORIENTATION) && (flags & YB_FLAG_HAS_NEXT))
            q0 = frame_update(q0, (p1 - xyz(p4)) + (dxNext - dxNew), -2.0f * r4.y * invl, flags,
                              qPrev, qrPrev, qNext, asrot(qr4));
        qOut[v] = f4r(q0);
    }
    if constexpr (kMode == kSweep) {
        if (partial) {
            if constexpr (kBatched)
                wave_residual_scenes(owner ? num : 0.0f, owner ? numF : 0.0f, sideA, partial,
                                     wave, n);
            else
                wave_residual(owner ? num : 0.0f, owner ? numF : 0.0f, partial, wave, n);
        }
    }
}

template <bool kBatched>
__global__ __launch_bounds__(512) void k_sweep(const DevParams* __restrict__ P,
                                               const float4* __restrict__ pos,
                                               const float4* __restrict__ prm,
                                               const float4* __restrict__ yv,
                                               const float4* __restrict__ qIn,
                                               float4* __restrict__ qOut,
                                               const float4* __restrict__ qRest,
                                               const float4* __restrict__ dxIn,
                                               float4* __restrict__ dxOut,
                                               const int32_t* __restrict__ numCols,
                                               const int32_t* __restrict__ cols,
                                               SolverCtl* __restrict__ ctl, int sweep,
                                               double2* __restrict__ partial, SceneArgs sa,
                                               CheckPrev cp) {
    sweep_body<kSweep, kBatched>(P, pos, prm, yv, qIn, qOut, qRest, dxIn, dxOut, numCols, cols,
                                 ctl, sweep, partial, LineOut{nullptr, nullptr, nullptr}, sa, -1,
                                 cp);
}

// Strand solve, tile part (see the line mode of sweep_body).
template <bool kBatched>
__global__ __launch_bounds__(256) void k_line_tile(const DevParams* __restrict__ P,
                                                   const float4* __restrict__ pos,
                                                   const float4* __restrict__ prm,
                                                   const float4* __restrict__ yv,
                                                   const float4* __restrict__ qIn,
                                                   const float4* __restrict__ qRest,
                                                   const float4* __restrict__ dxIn,
                                                   const int32_t* __restrict__ numCols,
                                                   const int32_t* __restrict__ cols, LineOut lo,
                                                   const SolverCtl* __restrict__ ctl, int sweep,
                                                   SceneArgs sa) {
    if (strand_skip(ctl, sweep)) return;
    sweep_body<kLine, kBatched>(P, pos, prm, yv, qIn, nullptr, qRest, dxIn, nullptr, numCols,
                                cols, nullptr, sweep, nullptr, lo, sa);
}

// First strand step of a substep from the plain warm start, fused with the
// frame refresh and the Newton-residual denominator (the line-start mode of
// sweep_body): reads the frames qIn, writes the refreshed ones to qOut and,
// with dxPlain set, a copy of dxIn there.
template <bool kBatched>
__global__ __launch_bounds__(256) void k_line_start(const DevParams* __restrict__ P,
                                                    const float4* __restrict__ pos,
                                                    const float4* __restrict__ prm,
                                                    const float4* __restrict__ yv,
                                                    const float4* __restrict__ qIn,
                                                    float4* __restrict__ qOut,
                                                    const float4* __restrict__ qRest,
                                                    const float4* __restrict__ dxIn,
                                                    float4* __restrict__ dxPlain,
                                                    const int32_t* __restrict__ numCols,
                                                    const int32_t* __restrict__ cols, LineOut lo,
                                                    double2* __restrict__ partial, SceneArgs sa) {
    sweep_body<kLineStart, kBatched>(P, pos, prm, yv, qIn, qOut, qRest, dxIn, dxPlain, numCols,
                                     cols, nullptr, 0, partial, lo, sa);
}

__global__ __launch_bounds__(256) void k_sweep_list(
    const DevParams* __restrict__ P, const float4* __restrict__ pos,
    const float4* __restrict__ prm, const float4* __restrict__ yv,
    const float4* __restrict__ qIn, float4* __restrict__ qOut, const float4* __restrict__ qRest,
    const float4* __restrict__ dxIn, float4* __restrict__ dxOut,
    const int32_t* __restrict__ numCols, const int32_t* __restrict__ cols,
    SolverCtl* __restrict__ ctl, int sweep, double2* __restrict__ partial, SceneArgs sa,
    const int32_t* __restrict__ tileList, const int32_t* __restrict__ tileCount) {
    if (ctl->doneAt < sweep) return;
    const int tiles = (P->n + 61) / 62;
    const int count = listed_tiles(tileCount, tiles);
    const int waves = (int)(gridDim.x * blockDim.x >> 6);
    for (int k = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6); k < count; k += waves)
        sweep_body<kSweep, true>(P, pos, prm, yv, qIn, qOut, qRest, dxIn, dxOut, numCols, cols,
                                 ctl, sweep, partial, LineOut{nullptr, nullptr, nullptr}, sa,
                                 listed_tile(tileList, tileCount, k));
}

__global__ __launch_bounds__(256) void k_line_tile_list(
    const DevParams* __restrict__ P, const float4* __restrict__ pos,
    const float4* __restrict__ prm, const float4* __restrict__ yv,
    const float4* __restrict__ qIn, const float4* __restrict__ qRest,
    const float4* __restrict__ dxIn, const int32_t* __restrict__ numCols,
    const int32_t* __restrict__ cols, LineOut lo, const SolverCtl* __restrict__ ctl, int sweep,
    SceneArgs sa, const int32_t* __restrict__ tileList, const int32_t* __restrict__ tileCount) {
    if (strand_skip(ctl, sweep)) return;
    const int tiles = (P->n + 61) / 62;
    const int count = listed_tiles(tileCount, tiles);
    const int waves = (int)(gridDim.x * blockDim.x >> 6);
    for (int k = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6); k < count; k += waves)
        sweep_body<kLine, true>(P, pos, prm, yv, qIn, nullptr, qRest, dxIn, nullptr, numCols,
                                cols, nullptr, sweep, nullptr, lo, sa,
                                listed_tile(tileList, tileCount, k));
}

// The tile list after a check of a batch: the tiles of every scene that has
// not converged, in scene order.  A tile shared by two scenes is listed once,
// by the first of them that is still running.  k_tile_scan (one block)
// counts each scene's entries and sums them; k_tile_fill (a block per scene)
// writes them.
constexpr int kTileScanThreads = 1024;
__device__ __forceinline__ void scene_tiles(const SceneCtl* __restrict__ sc,
                                            const int32_t* __restrict__ sceneOff,
                                            const int32_t* __restrict__ sceneOf, int k,
                                            int& first, int& count) {
    const int t0 = sceneOff[k] / 62, t1 = (sceneOff[k + 1] - 1) / 62;
    first = t0;
    count = 0;
    if (sc[k].doneAt != 0x7fffffff) return;  // converged
    // the first tile starts in scene k - 1: that scene lists it if still running
    if (k > 0 && sceneOf[62 * t0] != k && sc[k - 1].doneAt == 0x7fffffff) first = t0 + 1;
    count = t1 - first + 1;
}

__global__ __launch_bounds__(kTileScanThreads) void k_tile_scan(const SceneCtl* __restrict__ sc,
                                                            const int32_t* __restrict__ sceneOff,
                                                            const int32_t* __restrict__ sceneOf,
                                                            int numScenes,
                                                            int32_t* __restrict__ tileCount) {
    __shared__ int part[kTileScanThreads];
    const int t = threadIdx.x;
    const int per = (numScenes + kTileScanThreads - 1) / kTileScanThreads;
    const int k0 = t * per, k1 = min(k0 + per, numScenes);
    int sum = 0;
    for (int k = k0; k < k1; k++) {
        int first, count;
        scene_tiles(sc, sceneOff, sceneOf, k, first, count);
        sum += count;
    }
    part[t] = sum;
    __syncthreads();
    // inclusive scan of the per-thread sums (Hillis-Steele)
    for (int off = 1; off < kTileScanThreads; off <<= 1) {
        const int add = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += add;
        __syncthreads();
    }
    int at = part[t] - sum;
    for (int k = k0; k < k1; k++) {
        int first, count;
        scene_tiles(sc, sceneOff, sceneOf, k, first, count);
        tileCount[1 + k] = at;
        at += count;
    }
    if (t == kTileScanThreads - 1) tileCount[0] = part[t];
}

__global__ __launch_bounds__(64) void k_tile_fill(const SceneCtl* __restrict__ sc,
                                                  const int32_t* __restrict__ sceneOff,
                                                  const int32_t* __restrict__ sceneOf,
                                                  const int32_t* __restrict__ tileCount,
                                                  int32_t* __restrict__ tileList) {
    const int k = blockIdx.x;
    int first, count;
    scene_tiles(sc, sceneOff, sceneOf, k, first, count);
    const int at = tileCount[1 + k];
    for (int j = threadIdx.x; j < count; j += 64) tileList[at + j] = first + j;
}

// Residual partials of an iterate without changing it: the denominator |f(dx0)|
// of the relative residual when the sweeps do not start from the plain warm
// start.  With qOut set, also the frames refreshed from that iterate (the
// sweep's closed-form frame update at dx' = dx), written to qOut.
template <bool kBatched>
__global__ __launch_bounds__(256) void k_force_eval(const DevParams* __restrict__ P,
                                                    const float4* __restrict__ pos,
                                                    const float4* __restrict__ prm,
                                                    const float4* __restrict__ yv,
                                                    const float4* __restrict__ qIn,
                                                    float4* __restrict__ qOut,
                                                    const float4* __restrict__ qRest,
                                                    const float4* __restrict__ dxIn,
                                                    const int32_t* __restrict__ numCols,
                                                    const int32_t* __restrict__ cols,
                                                    double2* __restrict__ partial,
                                                    SceneArgs sa) {
    sweep_body<kEval, kBatched>(P, pos, prm, yv, qIn, qOut, qRest, dxIn, nullptr, numCols, cols,
                                nullptr, 0, partial, LineOut{nullptr, nullptr, nullptr}, sa);
}

// Strand solve, boundary part: one iteration over the pairs of vertices that
// face each other across a tile boundary.  E_w is the last owned vertex of
// tile w, F_(w+1) the first of tile w + 1; their tile solutions
//   E_w = gE + pE E_(w-1) + qE F_(w+1),   F_(w+1) = gF + pF E_w + qF F_(w+2)
// are solved exactly for the pair, with E_(w-1) and F_(w+2) -- each a whole
// tile away, so weakly coupled -- from the previous iteration (zero in the
// first).
// The pair (E, F) of boundary b from its two rows and the outer neighbours'
// values eo = E_(b-1), fo = F_(b+1) of the previous iteration.
struct PairRows {
    float4 GE, GF;  // tile solutions (xyz) and left factors (w) of the two rows
    float qE, qF;   // their right factors
};
__device__ __forceinline__ PairRows pair_rows(const float4* __restrict__ g,
                                              const float2* __restrict__ ql, int b) {
    const int vE = 62 * b + 61, vF = 62 * (b + 1);  // both < n: tile b + 1 exists
    return PairRows{g[vE], g[vF], ql[vE].x, ql[vF].x};
}
__device__ __forceinline__ void pair_solve(const PairRows& r, v3 eo, v3 fo, v3& E, v3& F) {
    const v3 a = xyz(r.GE) + r.GE.w * eo;
    const v3 b = xyz(r.GF) + r.qF * fo;
    const float pF = r.GF.w;
    E = (a + r.qE * b) * (1.0f / (1.0f - r.qE * pF));
    F = b + pF * E;
}

__global__ void k_line_pairs(const float4* __restrict__ g, const float2* __restrict__ ql,
                             int tiles, const float4* __restrict__ Eold,
                             const float4* __restrict__ Fold, float4* __restrict__ Enew,
                             float4* __restrict__ Fnew, int first,
                             const SolverCtl* __restrict__ ctl, int sweep) {
    if (strand_skip(ctl, sweep)) return;
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= tiles - 1) return;
    v3 eo = v3{0.f, 0.f, 0.f}, fo = v3{0.f, 0.f, 0.f};
    if (!first && w >= 1) eo = xyz(Eold[w - 1]);
    if (!first && w + 2 < tiles) fo = xyz(Fold[w + 2]);
    v3 E, F;
    pair_solve(pair_rows(g, ql, w), eo, fo, E, F);
    Enew[w] = f4(E, 0.0f);
    Fnew[w + 1] = f4(F, 0.0f);
}

// One or two boundary-pair iterations and the apply, fused: one wave per
// 62-vertex tile (lane l: vertex 62 w + l - 1, as in the sweep).  Tile w needs
// E_(w-1) and F_w.  Lanes 0..3 solve the first iteration of boundaries
// w-2 .. w+1 (the outer values start at zero); with two iterations, lanes 1
// and 2 solve boundaries w-1 and w again from their neighbours' results.  The
// same arithmetic as k_line_pairs, so the result equals that path's.
__device__ __forceinline__ void pairs_apply_tile(int w, int n, int tiles, int iters,
                                                 const float4* __restrict__ g,
                                                 const float2* __restrict__ ql,
                                                 float4* __restrict__ dx, int sweep,
                                                 SceneArgs sa) {
    const int lane = threadIdx.x & 63;
    // The apply's own rows are loaded first, so that their latency overlaps
    // the pair solves (which depend only on the boundary rows).
    const int v = 62 * w + lane - 1;
    const bool own = lane >= 1 && lane <= 62 && v < n;
    float2 QL = make_float2(0.0f, -1.0f);
    float4 G = make_float4(0.f, 0.f, 0.f, 0.f), d4 = G;
    if (own) {
        QL = ql[v];
        G = g[v];
        d4 = dx[v];
    }
    // first iteration: lane l < 4 takes boundary w - 2 + l
    v3 E1 = v3{0.f, 0.f, 0.f}, F1 = v3{0.f, 0.f, 0.f};
    PairRows rows{};
    const int b = w - 2 + lane;
    const bool has = lane < 4 && b >= 0 && b <= tiles - 2;
    if (has) {
        rows = pair_rows(g, ql, b);
        pair_solve(rows, v3{0.f, 0.f, 0.f}, v3{0.f, 0.f, 0.f}, E1, F1);
    }
    v3 E = E1, F = F1;
    if (iters == 2) {
        // lane l needs E of lane l - 1 and F of lane l + 1
        const v3 eo = v3{__shfl_up(E1.x, 1, 64), __shfl_up(E1.y, 1, 64), __shfl_up(E1.z, 1, 64)};
        const v3 fo = v3{__shfl_down(F1.x, 1, 64), __shfl_down(F1.y, 1, 64),
                         __shfl_down(F1.z, 1, 64)};
        if (has && (lane == 1 || lane == 2)) pair_solve(rows, eo, fo, E, F);
    }
    // E_(w-1) is boundary w - 1's (lane 1), F_w boundary w's (lane 2); a missing
    // boundary left zeros there
    const v3 Ew = v3{__shfl(E.x, 1, 64), __shfl(E.y, 1, 64), __shfl(E.z, 1, 64)};
    const v3 Fw = v3{__shfl(F.x, 2, 64), __shfl(F.y, 2, 64), __shfl(F.z, 2, 64)};
    if (!own || QL.y < 0.0f) return;
    if (sa.ctl && sa.ctl[sa.sceneOf[v]].doneAt < sweep) return;  // its scene has converged
    v3 d = xyz(G);
    if (w >= 1) d += G.w * Ew;
    if (w + 1 < tiles) d += QL.x * Fw;
    if (!(isfinite(d.x) && isfinite(d.y) && isfinite(d.z))) return;
    v3 nd = xyz(d4) + d;
    const float l = len(nd);
    if (l > QL.y && l > 0.0f) nd *= QL.y / l;
    dx[v] = f4(nd, d4.w);
}

// qDst set (a prologue step whose tile pass read frames refreshed into qSrc):
// each lane also copies its vertex's frame qSrc -> qDst for the sweeps.
__global__ __launch_bounds__(256) void k_line_pairs_apply(int n, int tiles, int iters,
                                                          const float4* __restrict__ g,
                                                          const float2* __restrict__ ql,
                                                          float4* __restrict__ dx,
                                                          const SolverCtl* __restrict__ ctl,
                                                          int sweep, SceneArgs sa,
                                                          const float4* __restrict__ qSrc,
                                                          float4* __restrict__ qDst) {
    if (strand_skip(ctl, sweep)) return;
    const int w = (int)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 6);
    if (w >= tiles) return;
    if (qDst) {
        const int lane = threadIdx.x & 63;
        const int v = 62 * w + lane - 1;
        if (lane >= 1 && lane <= 62 && v < n) qDst[v] = qSrc[v];
    }
    pairs_apply_tile(w, n, tiles, iters, g, ql, dx, sweep, sa);
}

__global__ __launch_bounds__(256) void k_line_pairs_apply_list(
    int n, int tiles, int iters, const float4* __restrict__ g, const float2* __restrict__ ql,
    float4* __restrict__ dx, const SolverCtl* __restrict__ ctl, int sweep, SceneArgs sa,
    const int32_t* __restrict__ tileList, const int32_t* __restrict__ tileCount) {
    if (strand_skip(ctl, sweep)) return;
    const int count = listed_tiles(tileCount, tiles);
    const int waves = (int)(gridDim.x * blockDim.x >> 6);
    for (int k = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6); k < count; k += waves)
        pairs_apply_tile(listed_tile(tileList, tileCount, k), n, tiles, iters, g, ql, dx, sweep,
                         sa);
}

// Strand solve, apply: x = g + p E_(w-1) + q F_(w+1) for every vertex of tile
// w, added to dx and clamped to the step limit like a sweep's step.  Vertices
// that are not rows (pinned) and non-finite results keep dx.
__global__ __launch_bounds__(256) void k_line_apply(int n, int tiles,
                                                    const float4* __restrict__ g,
                                                    const float2* __restrict__ ql,
                                                    const float4* __restrict__ E,
                                                    const float4* __restrict__ F,
                                                    float4* __restrict__ dx,
                                                    const SolverCtl* __restrict__ ctl, int sweep,
                                                    SceneArgs sa) {
    if (strand_skip(ctl, sweep)) return;
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const float2 QL = ql[v];
    if (QL.y < 0.0f) return;
    if (sa.ctl && sa.ctl[sa.sceneOf[v]].doneAt < sweep) return;  // its scene has converged
    const int w = v / 62;
    const float4 G = g[v];
    v3 d = xyz(G);
    if (w >= 1) d += G.w * xyz(E[w - 1]);
    if (w + 1 < tiles) d += QL.x * xyz(F[w + 1]);
    if (!(isfinite(d.x) && isfinite(d.y) && isfinite(d.z))) return;
    const float4 d4 = dx[v];
    v3 nd = xyz(d4) + d;
    const float l = len(nd);
    if (l > QL.y && l > 0.0f) nd *= QL.y / l;
    dx[v] = f4(nd, d4.w);
}

// Residuals of a sweep, both from per-wave partial sums (one per wave tile of
// 62 vertices, so the partials and the residuals do not depend on the block
// size), reduced in a fixed order by one wave:
//   update residual  r_U = sqrt( sum_free |dx_k+1 - dx_k|^2 / sum_free l_i^2 ),
//     the RMS change of the iterate in units of the rest length (l_i: the rest
//     length of vertex i's own segment, or of the previous one at a strand end);
//   Newton residual  r_F = |f(dx_k)| / |f(dx_0)|, the residual force of the
//     sweep's input iterate relative to that of the plain warm start dx_0.
// The stopping test uses one of them (SolverCtl::kind); both are recorded.
__device__ __forceinline__ void wave_residual(float numU, float numF, double2* partial, int wave,
                                              int n) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        numU += __shfl_xor(numU, off, 64);
        numF += __shfl_xor(numF, off, 64);
    }
    if ((threadIdx.x & 63) == 0 && wave < (n + 61) / 62)
        partial[wave] = make_double2((double)numU, (double)numF);
}

// Fixed-order sum of partials [lo, hi) by one block: thread t sums
// lo + t, lo + t + blockDim.x, ... in order, then a tree over the threads with
// a fixed pairing.  Every thread returns the sum.
constexpr int kResThreads = 256;
__device__ __forceinline__ double2 block_sum(const double2* __restrict__ p, int lo, int hi) {
    __shared__ double2 red[kResThreads];
    const int t = threadIdx.x;
    double x = 0.0, y = 0.0;
    for (int k = lo + t; k < hi; k += kResThreads) {
        const double2 v = p[k];
        x += v.x;
        y += v.y;
    }
    __syncthreads();  // a previous call's values have been read
    red[t] = make_double2(x, y);
    __syncthreads();
    for (int half = kResThreads / 2; half > 0; half >>= 1) {
        if (t < half) red[t] = make_double2(red[t].x + red[t + half].x, red[t].y + red[t + half].y);
        __syncthreads();
    }
    return red[0];
}

// Residual check of sweep `sweep`, launched right after it: kResBlocks blocks
// each sum a fixed range of the per-tile partials (and of the prologue's
// denominator partials when the substep has no denominator yet); the last
// block to finish adds the block sums in block order and makes the stopping
// decision.  The next sweep, launched after this kernel, then exits at once
// when the substep has converged.  The sums do not depend on scheduling.
constexpr int kResBlocks = 32;
__global__ __launch_bounds__(kResThreads) void k_residual(SolverCtl* __restrict__ ctl,
                                                          const double2* __restrict__ partial,
                                                          const double2* __restrict__ denomPart,
                                                          int nb, int sweep, float tol,
                                                          int maxSweeps,
                                                          double2* __restrict__ blockSums,
                                                          unsigned int* __restrict__ counter) {
    if (ctl->doneAt < sweep) return;
    // No Newton-residual denominator yet (the substep's first check): it is the
    // sum of denomPart, the partials of the prologue's pass over the plain warm
    // start, or -- in a plain session, whose first check is sweep 0 and whose
    // sweep 0 input is the plain warm start itself -- this check's own sum.
    const bool needDen = ctl->denomF == -1.0 && denomPart;
    const int b = blockIdx.x, R = gridDim.x;
    const int lo = (int)((long long)nb * b / R), hi = (int)((long long)nb * (b + 1) / R);
    const double2 sum = block_sum(partial, lo, hi);
    const double2 den = needDen ? block_sum(denomPart, lo, hi) : make_double2(0.0, 0.0);
    __shared__ bool last;
    __shared__ double3 all[kResBlocks];
    if (threadIdx.x == 0) {
        blockSums[2 * b] = sum;
        blockSums[2 * b + 1] = den;
        __threadfence();
        last = atomicAdd(counter, 1u) == (unsigned)(R - 1);
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    // the last block: every block's sums, one per thread, read at device
    // scope (they were written by other CUs), then added in block order
    if (threadIdx.x < R) {
        const double* q = reinterpret_cast<const double*>(blockSums + 2 * threadIdx.x);
        all[threadIdx.x] =
            make_double3(__hip_atomic_load(q + 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                         __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                         __hip_atomic_load(q + 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double2 a = make_double2(0.0, 0.0);
    double d = 0.0;
    for (int k = 0; k < R; k++) {
        a.x += all[k].x;
        a.y += all[k].y;
        d += all[k].z;
    }
    *counter = 0u;  // for the next check (a later launch)
    const double denF = ctl->denomF != -1.0 ? ctl->denomF : (needDen ? d : a.y);
    decide(ctl, a, denF, sweep, tol, maxSweeps);
}

// noNewton: this substep records no Newton residual (single handles only).
__global__ void k_ctl_reset(SolverCtl* __restrict__ ctl, SceneCtl* __restrict__ sceneCtl,
                            int numScenes, int32_t* __restrict__ tileCount,
                            unsigned int* __restrict__ resCounter, int noNewton) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) {
        ctl->doneAt = 0x7fffffff;
        ctl->denomF = noNewton ? kNoNewton : -1.0;
        ctl->scenesDone = 0;
        if (tileCount) tileCount[0] = -1;  // every tile until the first check
        if (resCounter) *resCounter = 0u;
    }
    if (sceneCtl && k < numScenes) {
        sceneCtl[k].doneAt = 0x7fffffff;
        sceneCtl[k].denomF = -1.0;
    }
}

// Fixed-order sum of the partials [a, b) by one wave (as reduce_partials).
__device__ __forceinline__ double2 reduce_range(const double2* __restrict__ p, int a, int b) {
    const int lane = threadIdx.x & 63;
    double2 acc[4];
#pragma unroll
    for (int j = 0; j < 4; j++) acc[j] = make_double2(0.0, 0.0);
    int k = a + lane;
    for (; k + 3 * 64 < b; k += 4 * 64) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const double2 t = p[k + 64 * j];
            acc[j].x += t.x;
            acc[j].y += t.y;
        }
    }
    for (; k < b; k += 64) {
        const double2 t = p[k];
        acc[0].x += t.x;
        acc[0].y += t.y;
    }
    double x = (acc[0].x + acc[1].x) + (acc[2].x + acc[3].x);
    double y = (acc[0].y + acc[1].y) + (acc[2].y + acc[3].y);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        x += __shfl_xor(x, off, 64);
        y += __shfl_xor(y, off, 64);
    }
    return make_double2(x, y);
}

// Sum over scene sc of per-scene-side tile partials (wave_residual_scenes):
// its first tile contributes the side that holds it, every later tile its
// side A.  A block of kSceneThreads: each wave sums a fixed quarter of the
// tiles (reduce_range), and the quarters are added in order, so the result
// does not depend on scheduling.  Every thread returns the sum.
constexpr int kSceneThreads = 256;
__device__ __forceinline__ double2 scene_sum(const double2* __restrict__ part,
                                             const int32_t* __restrict__ sceneOff,
                                             const int32_t* __restrict__ sceneOf, int sc,
                                             int tiles) {
    __shared__ double2 waveSum[kSceneThreads / 64];
    const int v0 = sceneOff[sc], v1 = sceneOff[sc + 1] - 1;
    const int t0 = v0 / 62, t1 = v1 / 62;
    constexpr int kWaves = kSceneThreads / 64;
    const int w = threadIdx.x >> 6;
    const int len = t1 - t0;  // tiles t0 + 1 .. t1
    const int a = t0 + 1 + (int)((long long)len * w / kWaves);
    const int b = t0 + 1 + (int)((long long)len * (w + 1) / kWaves);
    const double2 q = reduce_range(part, a, b);
    __syncthreads();  // a previous call's sums have been read
    if ((threadIdx.x & 63) == 0) waveSum[w] = q;
    __syncthreads();
    const double2 first = sceneOf[62 * t0] == sc ? part[t0] : part[tiles + t0];
    double x = first.x, y = first.y;
#pragma unroll
    for (int k = 0; k < kWaves; k++) {
        x += waveSum[k].x;
        y += waveSum[k].y;
    }
    return make_double2(x, y);
}

// Residual check of sweep `sweep` for batched scenes: one wave per scene
// decides that scene's convergence (the test of residual_wave on the scene's
// own sums); the last scene to converge ends the substep for the batch.
__global__ __launch_bounds__(kSceneThreads) void k_residual_scenes(SolverCtl* __restrict__ ctl,
                                                        SceneCtl* __restrict__ sc,
                                                        const int32_t* __restrict__ sceneOff,
                                                        const int32_t* __restrict__ sceneOf,
                                                        const double2* __restrict__ part,
                                                        const double2* __restrict__ denomPart,
                                                        int tiles, int numScenes, int sweep,
                                                        float tol, int maxSweeps) {
    const int k = blockIdx.x;
    if (sc[k].doneAt < sweep) return;
    const double2 a = scene_sum(part, sceneOff, sceneOf, k, tiles);
    double denF = sc[k].denomF;
    if (denF < 0.0) denF = denomPart ? scene_sum(denomPart, sceneOff, sceneOf, k, tiles).y : a.y;
    if (threadIdx.x != 0) return;
    sc[k].denomF = denF;
    const double denU = sc[k].denom;
    const float rU = denU > 0.0 ? (float)sqrt(a.x / denU) : 0.0f;
    const float rF = denF > 0.0 ? (float)sqrt(a.y / denF) : 0.0f;
    const float r = ctl->kind ? rF : rU;
    if (!(r <= tol || sweep == maxSweeps - 1)) return;
    sc[k].doneAt = sweep;
    sc[k].fin = (sweep & 1) ^ 1;
    sc[k].sweeps += (unsigned long long)(sweep + 1);
    sc[k].sumFres += (double)rF;
    // batch statistics: counts and maxima of non-negative floats (as integers)
    // are independent of the order in which scenes finish
    if (r > tol) atomicAdd(&ctl->notConverged, 1);
    atomicMax(reinterpret_cast<int*>(&ctl->maxResid), __float_as_int(r));
    atomicMax(reinterpret_cast<int*>(&ctl->maxFres), __float_as_int(rF));
    __threadfence();
    if (atomicAdd(&ctl->scenesDone, 1) == numScenes - 1) {
        ctl->doneAt = sweep;
        ctl->resid = r;
        ctl->fres = rF;
        ctl->sweeps += (unsigned long long)(sweep + 1);
        ctl->substeps += 1ull;
        if (sweep + 1 > ctl->maxUsed) ctl->maxUsed = sweep + 1;
    }
}


// ---------------------------------------------------------------------------
// Frame refresh after a non-plain warm start: one closed-form frame update
// (cosserat.cu:209-253) from the warm-start dx, q[0] -> q[1], then copied
// back so that the first sweep sees frames consistent with its dx.
// The refreshed frame of segment v (flags and rest data r4 = prm[v]) whose
// edge in the configuration x + dx is e, from its frame q0, the neighbouring
// frames qPrev, qNext and the rest rotors qrPrev = qRest[v - 1], qr = qRest[v]
// (those of absent neighbours are not read).  The callers load every input
// first, from clamped indices, so that all loads are in flight at once.
__device__ __forceinline__ float4 refreshed_frame(float4 r4, uint32_t flags, v3 e, float4 q0f,
                                                  float4 qPrev, float4 qNext, float4 qrPrev,
                                                  float4 qrf) {
    const rot q0 = asrot(q0f);
    v3 vv = v3{e.x / r4.x, e.y / r4.x, e.z / r4.x};
    vv *= -2.0f * r4.y;
    rot b = rot{0.f, 0.f, 0.f, 0.f};
    if (flags & YB_FLAG_HAS_PREV) {
        const rot qr = asrot(qrPrev);
        const rot qq = asrot(qPrev);
        const float sg = rdot(rmul(rinv(qq), q0), qr) > 0.0f ? 1.0f : -1.0f;
        b = rscale(rmul(qq, qr), sg);
    }
    if (flags & YB_FLAG_HAS_NEXT_ORIENTATION) {
        const rot qr = asrot(qrf);
        const rot qq = asrot(qNext);
        const float sg = rdot(rmul(rinv(q0), qq), qr) > 0.0f ? 1.0f : -1.0f;
        b = radd(b, rscale(rmul(qq, rinv(qr)), sg));
    }
    const float lambda = len(vv) + rlen(b);
    return f4r(rnormalize(radd(vb_e1(vv, b), rscale(b, lambda))));
}

__global__ __launch_bounds__(256) void k_frames(const DevParams* __restrict__ P,
                                                const float4* __restrict__ pos,
                                                const float4* __restrict__ prm,
                                                const float4* __restrict__ dx,
                                                const float4* __restrict__ qIn,
                                                float4* __restrict__ qOut,
                                                const float4* __restrict__ qRest) {
    const int n = P->n;
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const int vm = v > 0 ? v - 1 : v, vp = v + 1 < n ? v + 1 : v;
    const float4 r4 = prm[v], q0 = qIn[v], qm = qIn[vm], qp = qIn[vp];
    const float4 qrm = qRest[vm], qr = qRest[v];
    const float4 x0 = pos[v], x1 = pos[vp], d0 = dx[v], d1 = dx[vp];
    const uint32_t flags = __float_as_uint(r4.w);
    if (!(flags & YB_FLAG_FIX_ORIENTATION) && (flags & YB_FLAG_HAS_NEXT)) {
        const v3 e = (xyz(x1) - xyz(x0)) + (xyz(d1) - xyz(d0));
        qOut[v] = refreshed_frame(r4, flags, e, q0, qm, qp, qrm, qr);
    } else {
        qOut[v] = q0;
    }
}

// ---------------------------------------------------------------------------
// The learned warm start's apply (no counterpart in the reference, whose warm
// start is iteration.cu:18-34): the predicted correction c of vertex v,
// low-passed along the strand with the binomial [1 2 1]/4 filter when passes
// > 0 (a missing neighbour counts as the vertex itself), in units of the rest
// length L, on top of the plain warm start: dx0 + L c, clamped to the step
// limit like a sweep step.  Block-Jacobi damps the alternating (odd-even) mode
// along a strand as slowly as the smoothest one, and a per-vertex regression
// error is close to white noise; the filter removes that mode from the
// prediction before the sweeps see it.  A pinned vertex keeps dx0 (kind 0),
// and so does one whose correction is not finite (kind 2, a fallback); kind 1:
// the correction was applied.
struct LearnedDx {
    float4 dx;
    int kind;
};
// Of a vertex with plain start d4, mass weight posW (0: pinned), rest data
// r4, the previous vertex's rest length prevLen, and the corrections cPrev,
// c0, cNext of it and its neighbours (those of absent neighbours not used).
__device__ __forceinline__ LearnedDx learned_dx(float4 d4, float posW, float4 r4, float prevLen,
                                                float4 cPrev, float4 c0, float4 cNext,
                                                int passes) {
    if (posW == 0.0f) return LearnedDx{d4, 0};
    const uint32_t fl = __float_as_uint(r4.w);
    v3 c = v3{c0.x, c0.y, c0.z};
    if (passes > 0) {
        const float4 cm = (fl & YB_FLAG_HAS_PREV) ? cPrev : c0;
        const float4 cq = (fl & YB_FLAG_HAS_NEXT) ? cNext : c0;
        c = v3{0.25f * (cm.x + 2.0f * c0.x + cq.x), 0.25f * (cm.y + 2.0f * c0.y + cq.y),
               0.25f * (cm.z + 2.0f * c0.z + cq.z)};
    }
    const float L = (!(fl & YB_FLAG_HAS_NEXT) && (fl & YB_FLAG_HAS_PREV)) ? prevLen : r4.x;
    const v3 cL = c * L;
    if (!(isfinite(cL.x) && isfinite(cL.y) && isfinite(cL.z))) return LearnedDx{d4, 2};
    v3 nd = xyz(d4) + cL;
    const float l = len(nd);
    if (l > d4.w && l > 0.0f) nd *= d4.w / l;
    return LearnedDx{make_float4(nd.x, nd.y, nd.z, d4.w), 1};
}

// dx = the learned warm start from the plain one in dx0 (the inference
// kernel's copy, so that no lane reads a dx another lane has replaced), and
// with kFrames the frames refreshed from it into qOut (as k_frames; vertex
// v + 1's warm start is recomputed for the edge).  Every input is loaded
// first, from clamped indices (values of absent neighbours are not used), so
// that a lane has all its loads in flight at once.  Per-wave counters of
// applied and fallback vertices: one pair per wave, so no two waves contend
// for an address; only this wave touches its pair and launches on the stream
// are ordered, so a plain read-modify-write is exact.
template <bool kFrames>
__global__ __launch_bounds__(256) void k_learned_apply(
    const DevParams* __restrict__ P, const float4* __restrict__ predC,
    const float4* __restrict__ pos, const float4* __restrict__ prm,
    const float4* __restrict__ dx0, float4* __restrict__ dx,
    unsigned long long* __restrict__ counters, int passes, const float4* __restrict__ qIn,
    float4* __restrict__ qOut, const float4* __restrict__ qRest) {
    const int n = P->n;
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    int kind = 0;
    if (v < n) {
        const int vm = v > 0 ? v - 1 : v, vp = v + 1 < n ? v + 1 : v;
        const int vpp = vp + 1 < n ? vp + 1 : vp;
        const float4 rm = prm[vm], r0 = prm[v], cm = predC[vm], c0 = predC[v], cp = predC[vp];
        const float4 x0 = pos[v], d0 = dx0[v];
        float4 r1, cpp, x1, d1, q0, qm, qp, qrm, qr;
        if (kFrames) {
            r1 = prm[vp];
            cpp = predC[vpp];
            x1 = pos[vp];
            d1 = dx0[vp];
            q0 = qIn[v];
            qm = qIn[vm];
            qp = qIn[vp];
            qrm = qRest[vm];
            qr = qRest[v];
        }
        const LearnedDx d = learned_dx(d0, x0.w, r0, rm.x, cm, c0, cp, passes);
        kind = d.kind;
        if (kind == 1) dx[v] = d.dx;
        if (kFrames) {
            const uint32_t flags = __float_as_uint(r0.w);
            if (!(flags & YB_FLAG_FIX_ORIENTATION) && (flags & YB_FLAG_HAS_NEXT)) {
                const float4 e1 = learned_dx(d1, x1.w, r1, r0.x, c0, cp, cpp, passes).dx;
                const v3 e = (xyz(x1) - xyz(x0)) + (xyz(e1) - xyz(d.dx));
                qOut[v] = refreshed_frame(r0, flags, e, q0, qm, qp, qrm, qr);
            } else {
                qOut[v] = q0;
            }
        }
    }
    unsigned applied = kind == 1 ? 1u : 0u, fallback = kind == 2 ? 1u : 0u;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        applied += __shfl_xor(applied, off, 64);
        fallback += __shfl_xor(fallback, off, 64);
    }
    if ((threadIdx.x & 63) == 0 && (applied | fallback)) {
        const int wave = v >> 6;
        counters[2 * wave] += applied;
        counters[2 * wave + 1] += fallback;
    }
}

__global__ void k_copy4(const float4* __restrict__ src, float4* __restrict__ dst, int n) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) dst[v] = src[v];
}

// ---------------------------------------------------------------------------
// End of substep (iteration.cu:48-61).
// When the sweep count is odd the final frames sit in the second buffer and are
// copied back so that every substep starts from q[0].
__global__ __launch_bounds__(256) void k_end(const DevParams* __restrict__ P,
                                             float4* __restrict__ pos,
                                             float4* __restrict__ vel,
                                             const float4* __restrict__ dxFin,
                                             const float4* __restrict__ qFin,
                                             float4* __restrict__ q0,
                                             const SolverCtl* __restrict__ ctl,
                                             const float4* __restrict__ dxAlt,
                                             const float4* __restrict__ qAlt, SceneArgs sa,
                                             const float4* __restrict__ prm,
                                             const float4* __restrict__ dxPlain,
                                             float4* __restrict__ corrPrev) {
    const int n = P->n;
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    float4 plain = make_float4(0.f, 0.f, 0.f, 0.f), r4 = plain, rm = plain;
    if (corrPrev) {  // issued before the final iterate's buffer is known
        plain = dxPlain[v];
        r4 = prm[v];
        rm = prm[v > 0 ? v - 1 : v];
    }
    // residual-targeted mode: final iterate in buffer 1 (per scene in a batch)
    const int fin = sa.ctl ? sa.ctl[sa.sceneOf[v]].fin : (ctl ? ctl->fin : 0);
    if (fin == 1) {
        dxFin = dxAlt;
        qFin = qAlt;
    }
    const float h = P->h;
    const float invH = 1.0f / h;
    const v3 dx = xyz(dxFin[v]);
    float4 p4 = pos[v];
    if (p4.w != 0.0f) {
        const float s = 1.0f - P->drag * h;
        vel[v] = f4((dx * invH) * s, 0.0f);
    }
    p4.x += dx.x;
    p4.y += dx.y;
    p4.z += dx.z;
    pos[v] = p4;
    if (qFin != q0) q0[v] = qFin[v];
    // A learned (or temporal) session: the converged correction (dx* - dx0) / L
    // that the next substep's warm start builds on, L the rest length of the
    // vertex's segment (of the previous one at a strand end).
    if (corrPrev) {
        const uint32_t fl = __float_as_uint(r4.w);
        const float L =
            (!(fl & YB_FLAG_HAS_NEXT) && (fl & YB_FLAG_HAS_PREV)) ? rm.x : r4.x;
        const float iL = 1.0f / L;
        corrPrev[v] = make_float4((dx.x - plain.x) * iL, (dx.y - plain.y) * iL,
                                  (dx.z - plain.z) * iL, 0.f);
    }
}

// ---------------------------------------------------------------------------
// Upload / download transposes.  Staging layout: n yb_vertex records, then
// 3n floats of velocity, 4n of frames, 4n of rest frames.
__global__ void k_upload(int n, const yb_vertex* __restrict__ sv, const float* __restrict__ svel,
                         const float* __restrict__ sq, const float* __restrict__ sqr,
                         float4* __restrict__ pos, float4* __restrict__ prm,
                         float4* __restrict__ vel, float4* __restrict__ q,
                         float4* __restrict__ qRest) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const yb_vertex a = sv[v];
    pos[v] = make_float4(a.pos[0], a.pos[1], a.pos[2], a.invMass);
    prm[v] = make_float4(a.lRest, a.kStretch, __int_as_float(a.connectionIndex),
                         __uint_as_float(a.flags));
    vel[v] = make_float4(svel[3 * v], svel[3 * v + 1], svel[3 * v + 2], 0.f);
    q[v] = make_float4(sq[4 * v], sq[4 * v + 1], sq[4 * v + 2], sq[4 * v + 3]);
    qRest[v] = make_float4(sqr[4 * v], sqr[4 * v + 1], sqr[4 * v + 2], sqr[4 * v + 3]);
}

__global__ void k_download(int n, yb_vertex* __restrict__ sv, float* __restrict__ svel,
                           float* __restrict__ sq, const float4* __restrict__ pos,
                           const float4* __restrict__ prm, const float4* __restrict__ vel,
                           const float4* __restrict__ q) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const float4 p = pos[v];
    const float4 r = prm[v];
    yb_vertex a;
    a.pos[0] = p.x;
    a.pos[1] = p.y;
    a.pos[2] = p.z;
    a.invMass = p.w;
    a.lRest = r.x;
    a.kStretch = r.y;
    a.connectionIndex = __float_as_int(r.z);
    a.flags = __float_as_uint(r.w);
    sv[v] = a;
    const float4 u = vel[v];
    svel[3 * v] = u.x;
    svel[3 * v + 1] = u.y;
    svel[3 * v + 2] = u.z;
    const float4 qq = q[v];
    sq[4 * v] = qq.x;
    sq[4 * v + 1] = qq.y;
    sq[4 * v + 2] = qq.z;
    sq[4 * v + 3] = qq.w;
}

__global__ void k_zero_vel(int n, float4* __restrict__ vel, float4* __restrict__ lastVel) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    vel[v] = make_float4(0.f, 0.f, 0.f, 0.f);
    lastVel[v] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// Scenario drivers (main.cpp:54-121) on the device state.
__global__ void k_drive(int n, yb_driver d, float t, float dt, float4* __restrict__ pos,
                        float4* __restrict__ vel, const float4* __restrict__ init) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    float4 p = pos[v];
    if (p.w != 0.0f) return;  // only pinned vertices are driven
    const float4 q = init[v];
    const float cx = d.center[0], cy = d.center[1], cz = d.center[2];
    const bool below = p.y < cy;
    if (d.kind == YB_DRIVE_TWIST) {
        if (d.side != 0 && (d.side < 0) != below) return;
        const float a0 = d.speed * fminf(fmaxf(t - d.start, 0.0f), d.duration);
        const float a1 = d.speed * fminf(fmaxf(t + dt - d.start, 0.0f), d.duration);
        const float rx = q.x - cx, rz = q.z - cz;
        const float c0 = cosf(a0), s0 = sinf(a0), c1 = cosf(a1), s1 = sinf(a1);
        const float x0 = cx + (c0 * rx - s0 * rz), z0 = cz + (s0 * rx + c0 * rz);
        const float x1 = cx + (c1 * rx - s1 * rz), z1 = cz + (s1 * rx + c1 * rz);
        const float y0 = cy + (q.y - cy);
        p.x = x0;
        p.y = y0;
        p.z = z0;
        const float4 u = vel[v];
        vel[v] = make_float4((x1 - x0) / dt, ((cy + (q.y - cy)) - y0) / dt, (z1 - z0) / dt, u.w);
        if (d.release >= 0.0f && t > d.release) p.w = q.w;
        pos[v] = p;
    } else if (d.kind == YB_DRIVE_PULL) {
        const float x0 = -d.speed * fminf(fmaxf(t - d.start, 0.0f), d.duration);
        const float x1 = -d.speed * fminf(fmaxf(t + dt - d.start, 0.0f), d.duration);
        const float y0 = q.y + (below ? x0 : -x0);
        const float y1 = q.y + (below ? x1 : -x1);
        p.y = y0;
        pos[v] = p;
        float4 u = vel[v];
        u.y = (y1 - y0) / dt;
        vel[v] = u;
    }
}

__global__ void k_fill_i32(int32_t* __restrict__ p, int32_t val, int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = val;
}

static inline int blocks(int n, int b) { return (n + b - 1) / b; }

// ---- launchers -------------------------------------------------------------
void launch_upload(yb_sim* s) {
    const int n = s->n;
    char* st = (char*)s->b.stage;
    const yb_vertex* sv = (const yb_vertex*)st;
    const float* svel = (const float*)(st + sizeof(yb_vertex) * n);
    const float* sq = svel + 3 * (size_t)n;
    const float* sqr = sq + 4 * (size_t)n;
    hipLaunchKernelGGL(k_upload, dim3(blocks(n, 256)), dim3(256), 0, s->stream, n, sv, svel, sq,
                       sqr, s->b.pos, s->b.prm, s->b.vel, s->b.q[0], s->b.qRest);
}

void launch_download(yb_sim* s) {
    const int n = s->n;
    char* st = (char*)s->b.stage;
    yb_vertex* sv = (yb_vertex*)st;
    float* svel = (float*)(st + sizeof(yb_vertex) * n);
    float* sq = svel + 3 * (size_t)n;
    hipLaunchKernelGGL(k_download, dim3(blocks(n, 256)), dim3(256), 0, s->stream, n, sv, svel, sq,
                       s->b.pos, s->b.prm, s->b.vel, s->b.q[0]);
}

void launch_drive(yb_sim* s, const yb_driver& d, float t, float dt) {
    hipLaunchKernelGGL(k_drive, dim3(blocks(s->n, 256)), dim3(256), 0, s->stream, s->n, d, t, dt,
                       s->b.pos, s->b.vel, s->b.drvInit);
}

void launch_zero_vel(yb_sim* s) {
    hipLaunchKernelGGL(k_zero_vel, dim3(blocks(s->n, 256)), dim3(256), 0, s->stream, s->n,
                       s->b.vel, s->b.lastVel);
}

void launch_fill_i32(yb_sim* s, int32_t* p, int32_t v, int count) {
    if (count <= 0) return;
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks(count, 256)), dim3(256), 0, s->stream, p, v, count);
}

void launch_step_limit(yb_sim* s) {
    hipLaunchKernelGGL(k_step_limit<true>, dim3(blocks(s->n, 256)), dim3(256), 0, s->stream,
                       s->b.params, s->b.pos, s->b.prm, s->b.dx[0], s->b.numCols, s->b.cols,
                       s->b.err, s->b.minSep);
}

void launch_separation(yb_sim* s) {
    hipLaunchKernelGGL(k_step_limit<false>, dim3(blocks(s->n, 256)), dim3(256), 0, s->stream,
                       s->b.params, s->b.pos, s->b.prm, s->b.dx[0], s->b.numCols, s->b.cols,
                       s->b.err, s->b.minSep);
}

// wave tiles of 62 vertices (a 64-lane wave with a one-vertex halo each side)
int sweep_waves(int n) { return blocks(n, 62); }

// Batched scenes in the residual-targeted schedule: per-scene convergence.
static bool per_scene(const yb_sim* s, bool adaptive) {
    return adaptive && s->numScenes > 1 && s->b.sceneCtl;
}
SceneArgs scene_args(const yb_sim* s, bool adaptive) {
    return per_scene(s, adaptive) ? SceneArgs{s->b.sceneOf, s->b.sceneCtl}
                                  : SceneArgs{nullptr, nullptr};
}
// Residual partials of sweep k: half k & 1 of b.partial (one double2 per
// tile), or of b.scenePart in a batch (two per tile: the tile's scene sides).
static double2* sweep_partials(yb_sim* s, int sweep, bool batched) {
    const size_t nbw = (size_t)sweep_waves(s->n);
    return batched ? s->b.scenePart + 2 * nbw * (sweep & 1) : s->b.partial + nbw * (sweep & 1);
}

// Grid of a tile-list kernel: enough 4-wave blocks to fill the device, at
// most one wave per tile.
static int list_blocks(const yb_sim* s) {
    const int waves = std::min(sweep_waves(s->n), s->numCUs * 24);
    return (waves + 3) / 4;
}
// In a batch, the sweeps and strand steps after the substep's first check run
// over the tile list.
static bool tile_list_mode(const yb_sim* s, bool adaptive, int sweep) {
    return per_scene(s, adaptive) && sweep > first_check(s);
}

void launch_sweep(yb_sim* s, int sweep, bool adaptive, bool partials, bool checkPrev) {
    const int in = sweep & 1;
    const int out = in ^ 1;
    const int threads = s->sweepThreads;
    const bool batched = per_scene(s, adaptive);
    // partials of sweep k go to half k & 1, so a sweep can reduce its
    // predecessor's while writing its own
    double2* part = sweep_partials(s, sweep, batched);
    const CheckPrev cp{checkPrev ? sweep_partials(s, sweep ^ 1, false) : nullptr,
                       denom_partials(s), s->tol, s->maxSweeps};
    const dim3 grid(blocks(sweep_waves(s->n) * 64, threads));
    if (tile_list_mode(s, adaptive, sweep))
        hipLaunchKernelGGL(k_sweep_list, dim3(list_blocks(s)), dim3(256), 0, s->stream,
                           s->b.params, s->b.pos, s->b.prm, s->b.yv, s->b.q[in], s->b.q[out],
                           s->b.qRest, s->b.dx[in], s->b.dx[out], s->b.numCols, s->b.cols,
                           s->b.ctl, sweep, partials ? part : nullptr, scene_args(s, adaptive),
                           s->b.tileList, s->b.tileCount);
    else if (batched)
        hipLaunchKernelGGL(k_sweep<true>, grid, dim3(threads), 0, s->stream, s->b.params,
                           s->b.pos, s->b.prm, s->b.yv, s->b.q[in], s->b.q[out], s->b.qRest,
                           s->b.dx[in], s->b.dx[out], s->b.numCols, s->b.cols, s->b.ctl, sweep,
                           partials ? part : nullptr, scene_args(s, adaptive),
                           CheckPrev{nullptr, nullptr, 0.0f, 0});
    else
        hipLaunchKernelGGL(k_sweep<false>, grid, dim3(threads), 0, s->stream, s->b.params,
                           s->b.pos, s->b.prm, s->b.yv, s->b.q[in], s->b.q[out], s->b.qRest,
                           s->b.dx[in], s->b.dx[out], s->b.numCols, s->b.cols,
                           adaptive ? s->b.ctl : nullptr, sweep, partials ? part : nullptr,
                           SceneArgs{nullptr, nullptr}, cp);
}

void launch_residual(yb_sim* s, int sweep) {
    const int nbw = sweep_waves(s->n);
    if (per_scene(s, true)) {
        hipLaunchKernelGGL(k_residual_scenes, dim3(s->numScenes), dim3(kSceneThreads), 0, s->stream,
                           s->b.ctl,
                           s->b.sceneCtl, s->b.sceneOff, s->b.sceneOf,
                           sweep_partials(s, sweep, true), denom_partials(s), nbw, s->numScenes,
                           sweep, s->tol, s->maxSweeps);
        // the tiles of the scenes still running, for the sweeps that follow
        hipLaunchKernelGGL(k_tile_scan, dim3(1), dim3(kTileScanThreads), 0, s->stream, s->b.sceneCtl,
                           s->b.sceneOff, s->b.sceneOf, s->numScenes, s->b.tileCount);
        hipLaunchKernelGGL(k_tile_fill, dim3(s->numScenes), dim3(64), 0, s->stream, s->b.sceneCtl,
                           s->b.sceneOff, s->b.sceneOf, s->b.tileCount, s->b.tileList);
        return;
    }
    hipLaunchKernelGGL(k_residual, dim3(std::min(kResBlocks, nbw)), dim3(kResThreads), 0,
                       s->stream, s->b.ctl, sweep_partials(s, sweep, false), denom_partials(s),
                       nbw, sweep, s->tol, s->maxSweeps, s->b.resBlocks, s->b.resCounter);
}

void launch_learned_apply(yb_sim* s, bool frames, bool copyFrames) {
    if (frames) {
        hipLaunchKernelGGL(k_learned_apply<true>, dim3(blocks(s->n, 256)), dim3(256), 0,
                           s->stream, s->b.params, s->b.predC, s->b.pos, s->b.prm, s->b.dxPlain,
                           s->b.dx[0], s->b.mlpCount, s->mlpSmooth, s->b.q[0], s->b.q[1],
                           s->b.qRest);
        if (copyFrames)
            hipLaunchKernelGGL(k_copy4, dim3(blocks(s->n, 256)), dim3(256), 0, s->stream,
                               s->b.q[1], s->b.q[0], s->n);
    } else {
        hipLaunchKernelGGL(k_learned_apply<false>, dim3(blocks(s->n, 256)), dim3(256), 0,
                           s->stream, s->b.params, s->b.predC, s->b.pos, s->b.prm, s->b.dxPlain,
                           s->b.dx[0], s->b.mlpCount, s->mlpSmooth, nullptr, nullptr, nullptr);
    }
}

void launch_frame_refresh(yb_sim* s) {
    hipLaunchKernelGGL(k_frames, dim3(blocks(s->n, 256)), dim3(256), 0, s->stream, s->b.params,
                       s->b.pos, s->b.prm, s->b.dx[0], s->b.q[0], s->b.q[1], s->b.qRest);
    hipLaunchKernelGGL(k_copy4, dim3(blocks(s->n, 256)), dim3(256), 0, s->stream, s->b.q[1],
                       s->b.q[0], s->n);
}

static void launch_strand_pairs(yb_sim* s, int buf, SolverCtl* ctl, int sweep, bool list,
                                bool copyFrames = false);

void launch_strand_solve(yb_sim* s, int buf, SolverCtl* ctl, int sweep, float* sys, int qbuf,
                         bool copyFrames) {
    const float4* qf = s->b.q[qbuf < 0 ? buf : qbuf];
    const int tiles = sweep_waves(s->n);
    const LineOut lo{s->b.lineG, s->b.lineQL, sys};
    // in a batch, converged scenes are left alone (before sweep `sweep`)
    const SceneArgs sa = scene_args(s, ctl != nullptr);
    const bool list = tile_list_mode(s, ctl != nullptr, sweep);
    if (list)
        hipLaunchKernelGGL(k_line_tile_list, dim3(list_blocks(s)), dim3(256), 0, s->stream,
                           s->b.params, s->b.pos, s->b.prm, s->b.yv, qf, s->b.qRest,
                           s->b.dx[buf], s->b.numCols, s->b.cols, lo, ctl, sweep, sa,
                           s->b.tileList, s->b.tileCount);
    else if (sa.ctl)
        hipLaunchKernelGGL(k_line_tile<true>, dim3(blocks(tiles * 64, 256)), dim3(256), 0,
                           s->stream, s->b.params, s->b.pos, s->b.prm, s->b.yv, qf,
                           s->b.qRest, s->b.dx[buf], s->b.numCols, s->b.cols, lo, ctl, sweep, sa);
    else
        hipLaunchKernelGGL(k_line_tile<false>, dim3(blocks(tiles * 64, 256)), dim3(256), 0,
                           s->stream, s->b.params, s->b.pos, s->b.prm, s->b.yv, qf,
                           s->b.qRest, s->b.dx[buf], s->b.numCols, s->b.cols, lo, ctl, sweep, sa);
    launch_strand_pairs(s, buf, ctl, sweep, list, copyFrames);
}

// The boundary part of a strand step on dx[buf], and the apply.  copyFrames:
// then copy the frames q[1] -> q[0] (refreshed by the prologue into q[1]), in
// the apply's pass where it is the one-launch kind.
static void launch_strand_pairs(yb_sim* s, int buf, SolverCtl* ctl, int sweep, bool list,
                                bool copyFrames) {
    const int tiles = sweep_waves(s->n);
    const SceneArgs sa = scene_args(s, ctl != nullptr);
    if (s->strandIters <= 2 && !list) {
        hipLaunchKernelGGL(k_line_pairs_apply, dim3(blocks(tiles * 64, 256)), dim3(256), 0,
                           s->stream, s->n, tiles, s->strandIters, s->b.lineG, s->b.lineQL,
                           s->b.dx[buf], ctl, sweep, sa, copyFrames ? s->b.q[1] : nullptr,
                           copyFrames ? s->b.q[0] : nullptr);
        return;
    }
    if (copyFrames)
        hipLaunchKernelGGL(k_copy4, dim3(blocks(s->n, 256)), dim3(256), 0, s->stream, s->b.q[1],
                           s->b.q[0], s->n);
    if (s->strandIters <= 2) {
        hipLaunchKernelGGL(k_line_pairs_apply_list, dim3(list_blocks(s)), dim3(256), 0,
                           s->stream, s->n, tiles, s->strandIters, s->b.lineG, s->b.lineQL,
                           s->b.dx[buf], ctl, sweep, sa, s->b.tileList, s->b.tileCount);
        return;
    }
    int cur = 0;
    for (int it = 0; it < s->strandIters; it++) {
        hipLaunchKernelGGL(k_line_pairs, dim3(blocks(std::max(tiles - 1, 1), 256)), dim3(256), 0,
                           s->stream, s->b.lineG, s->b.lineQL, tiles, s->b.lineE[cur],
                           s->b.lineF[cur], s->b.lineE[cur ^ 1], s->b.lineF[cur ^ 1], it == 0 ? 1 : 0,
                           ctl, sweep);
        cur ^= 1;
    }
    hipLaunchKernelGGL(k_line_apply, dim3(blocks(s->n, 256)), dim3(256), 0, s->stream, s->n, tiles,
                       s->b.lineG, s->b.lineQL, s->b.lineE[cur], s->b.lineF[cur], s->b.dx[buf], ctl,
                       sweep, sa);
}

CtlReset ctl_reset_args(const yb_sim* s, bool noNewton) {
    const bool batched = per_scene(s, true);
    CtlReset r;
    r.ctl = s->b.ctl;
    r.sceneCtl = batched ? s->b.sceneCtl : nullptr;
    r.numScenes = s->numScenes;
    r.tileCount = batched ? s->b.tileCount : nullptr;
    r.resCounter = s->b.resCounter;
    r.noNewton = (noNewton && !batched) ? 1 : 0;
    return r;
}

void launch_ctl_reset(yb_sim* s, bool noNewton) {
    const bool batched = per_scene(s, true);
    hipLaunchKernelGGL(k_ctl_reset, dim3(batched ? blocks(s->numScenes, 256) : 1),
                       dim3(batched ? 256 : 64), 0, s->stream, s->b.ctl,
                       batched ? s->b.sceneCtl : nullptr, s->numScenes,
                       batched ? s->b.tileCount : nullptr, s->b.resCounter,
                       (noNewton && !batched) ? 1 : 0);
}

// The Newton-residual denominator's partials: the third part of b.partial
// (of b.scenePart in a batch), written by the prologue's force evaluation in
// sessions whose sweeps do not start from the plain warm start; null in
// plain sessions.
const double2* denom_partials(const yb_sim* s) {
    if (!(s->warmStart != 0 || s->strandSolve)) return nullptr;
    const size_t nbw = (size_t)sweep_waves(s->n);
    return per_scene(s, true) ? s->b.scenePart + 4 * nbw : s->b.partial + 2 * nbw;
}

void launch_residual_denom(yb_sim* s, bool frames) {
    double2* part = const_cast<double2*>(denom_partials(s));
    const dim3 grid(blocks(sweep_waves(s->n) * 64, 256));
    if (per_scene(s, true))
        hipLaunchKernelGGL(k_force_eval<true>, grid, dim3(256), 0, s->stream, s->b.params,
                           s->b.pos, s->b.prm, s->b.yv, s->b.q[0], frames ? s->b.q[1] : nullptr,
                           s->b.qRest, s->b.dx[0], s->b.numCols, s->b.cols, part,
                           scene_args(s, true));
    else
        hipLaunchKernelGGL(k_force_eval<false>, grid, dim3(256), 0, s->stream, s->b.params,
                           s->b.pos, s->b.prm, s->b.yv, s->b.q[0], frames ? s->b.q[1] : nullptr,
                           s->b.qRest, s->b.dx[0], s->b.numCols, s->b.cols, part,
                           SceneArgs{nullptr, nullptr});
    if (frames)
        hipLaunchKernelGGL(k_copy4, dim3(blocks(s->n, 256)), dim3(256), 0, s->stream, s->b.q[1],
                           s->b.q[0], s->n);
}

void launch_line_start(yb_sim* s, bool keepPlain) {
    double2* part = const_cast<double2*>(denom_partials(s));
    float4* plain = keepPlain ? s->b.dxPlain : nullptr;
    const LineOut lo{s->b.lineG, s->b.lineQL, nullptr};
    const dim3 grid(blocks(sweep_waves(s->n) * 64, 256));
    if (per_scene(s, true))
        hipLaunchKernelGGL(k_line_start<true>, grid, dim3(256), 0, s->stream, s->b.params,
                           s->b.pos, s->b.prm, s->b.yv, s->b.q[0], s->b.q[1], s->b.qRest,
                           s->b.dx[0], plain, s->b.numCols, s->b.cols, lo, part,
                           scene_args(s, true));
    else
        hipLaunchKernelGGL(k_line_start<false>, grid, dim3(256), 0, s->stream, s->b.params,
                           s->b.pos, s->b.prm, s->b.yv, s->b.q[0], s->b.q[1], s->b.qRest,
                           s->b.dx[0], plain, s->b.numCols, s->b.cols, lo, part,
                           SceneArgs{nullptr, nullptr});
    launch_strand_pairs(s, 0, nullptr, 0, false, true);  // also the frames q[1] -> q[0]
}

void launch_end(yb_sim* s, int fin, bool adaptive, bool correction) {
    hipLaunchKernelGGL(k_end, dim3(blocks(s->n, 256)), dim3(256), 0, s->stream, s->b.params,
                       s->b.pos, s->b.vel, s->b.dx[fin], s->b.q[fin], s->b.q[0],
                       adaptive ? s->b.ctl : nullptr, s->b.dx[1], s->b.q[1],
                       scene_args(s, adaptive), s->b.prm, s->b.dxPlain,
                       correction ? s->b.corrPrev : nullptr);
}

}  // namespace yb
