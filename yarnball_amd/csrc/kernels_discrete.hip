// kernels_discrete.hip — substep prologue and the LBVH broadphase on gfx950.
//
// This translation unit is compiled with -ffp-contract=off: every kernel here
// feeds a discrete decision (contact accepted or not, Morton cell, sort order),
// and the CPU oracle's FP32 twin repeats the same IEEE operations in the same
// order, so codes, orders and contact sets match bit for bit.
//
//   k_init        inertial target + gravity-extrapolated warm start  (iteration.cu:7-41)
//   k_aabb        padded segment boxes                                 (collision.cu:10-37)
//   k_box_*       scene box reduction (the root box LBVH::bounds returns)
//   k_morton      30-bit Morton code of each box centre in the scene box
//   k_sort_*      stable LSD radix sort, 4 x 8-bit passes, LDS ranking
//   k_karras      Karras (2012) radix tree over the sorted codes
//   k_refit_*     fence-free refit: in-block LDS climbing + range unions
//   k_query       stack traversal + narrowphase filter (collision.cu:39-92)
//                 writing each segment's own contact list: no atomics, no
//                 host round trip for the pair count.
#include "yb_internal.h"
#include "ybmath.h"

namespace yb {

__device__ __forceinline__ v3 xyz(float4 a) { return v3{a.x, a.y, a.z}; }
__device__ __forceinline__ float4 f4(v3 a, float w) { return make_float4(a.x, a.y, a.z, w); }

// ---------------------------------------------------------------------------
// initItr (iteration.cu:7-41): yv = h v + h^2 g for free vertices,
// dx0 = h v + h^2 s g with s = clamp(a.g / |g|^2, 0, 1), a = (v - v_last)/lastH.
__global__ __launch_bounds__(256) void k_init(const DevParams* __restrict__ P,
                                              const float4* __restrict__ pos,
                                              const float4* __restrict__ vel,
                                              float4* __restrict__ lastVel,
                                              float4* __restrict__ yv,
                                              float4* __restrict__ dx0,
                                              int32_t* __restrict__ skinState,
                                              int4* __restrict__ skinPolicy, int skinScenes,
                                              CtlReset cr) {
    const int n = P->n;
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    // Residual-targeted schedule: the substep's solver state starts running
    // (nothing reads it before the substep's first residual check).
    if (cr.ctl) {
        if (v == 0) {
            cr.ctl->doneAt = 0x7fffffff;
            cr.ctl->denomF = cr.noNewton ? kNoNewton : -1.0;
            cr.ctl->scenesDone = 0;
            if (cr.tileCount) cr.tileCount[0] = -1;  // every tile until the first check
            if (cr.resCounter) *cr.resCounter = 0u;
        }
        if (cr.sceneCtl && v < cr.numScenes) {
            cr.sceneCtl[v].doneAt = 0x7fffffff;
            cr.sceneCtl[v].denomF = -1.0;
        }
    }
    // Contact skin, one thread per scene: the scene's walk flag of the last
    // substep feeds its policy (lists reused fewer than kSkinMinReuse times:
    // walk without the skin for a back-off of kSkinBackoff substeps, doubled
    // after each retry that fails too), then this substep's flags start
    // clear (k_aabb sets them).
    if (v < skinScenes) {
        const int f = skinScenes > 1 ? 1 + v : 0;
        int4 pol = skinPolicy[v];
        const int32_t walked = skinState[f];
        if (pol.x > 0) {
            pol.x--;
            pol.y = 0;
        } else if (walked == 2) {
            pol.y = 1;  // a new tree: the lists it made are fresh, not short-lived
        } else if (walked) {
            if (pol.y >= 1 && pol.y - 1 < kSkinMinReuse) {
                pol.z = pol.z > 0 ? min(2 * pol.z, kSkinBackoffMax) : kSkinBackoff;
                pol.x = pol.z;
                pol.y = 0;
            } else {
                if (pol.y - 1 >= kSkinMinReuse) pol.z = 0;  // those lists lasted
                pol.y = 1;
            }
        } else if (pol.y >= 1) {
            pol.y = min(pol.y + 1, 1 << 20);
        }
        skinPolicy[v] = pol;
        skinState[f] = 0;
        if (v == 0 && skinScenes > 1) skinState[0] = 0;
    }
    const float h = P->h;
    const v3 g = v3{P->gx, P->gy, P->gz};
    const v3 u = xyz(vel[v]);
    v3 dx = h * u;
    const v3 lv = xyz(lastVel[v]);
    lastVel[v] = f4(u, 0.f);
    if (pos[v].w != 0.0f) {
        yv[v] = f4(dx + (h * h) * g, 0.f);
        const float g2 = len2(g);
        if (g2 > 0.0f) {
            const float lh = P->lastH;
            const v3 a = v3{(u.x - lv.x) / lh, (u.y - lv.y) / lh, (u.z - lv.z) / lh};
            const float s = clampf(dot(a, g) / g2, 0.0f, 1.0f);
            dx += (h * h * s) * g;
        }
    }
    dx0[v] = f4(dx, 0.f);
}

// ---------------------------------------------------------------------------
// Segment boxes (collision.cu:10-37).  Vertices without a next vertex get the
// empty box (+inf, -inf), which overlaps nothing.  hi.w carries the segment's
// velocity pad sqrt(max(|dx_i|^2, |dx_i+1|^2)) for the narrowphase radius, with
// the sign bit set when either endpoint is glued (connectionIndex >= 0).
// Contact skin (sk.state set): a segment whose box has left the grown box of
// the last walk, or whose kept candidate list is incomplete, flags its scene
// (and sk.state[0]) to walk the tree in this substep; force flags every one.
// Every flagging lane stores the same value, so the stores need no atomics.
__global__ __launch_bounds__(256) void k_aabb(const DevParams* __restrict__ P,
                                              const float4* __restrict__ pos,
                                              const float4* __restrict__ prm,
                                              const float4* __restrict__ dx,
                                              float4* __restrict__ aabb, SkinArgs sk,
                                              const int32_t* __restrict__ sceneOf, int force) {
    const int n = P->n;
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const float4 r0 = prm[v];
    const uint32_t flags = __float_as_uint(r0.w);
    v3 lo = v3{INFINITY, INFINITY, INFINITY};
    v3 hi = v3{-INFINITY, -INFINITY, -INFINITY};
    float vpad = 0.0f;
    bool glued = false;
    if (flags & YB_FLAG_HAS_NEXT) {
        const v3 p0 = xyz(pos[v]);
        const v3 p1 = xyz(pos[v + 1]);
        lo = vmin(p0, p1);
        hi = vmax(p0, p1);
        vpad = sqrtf(fmaxf(len2(xyz(dx[v])), len2(xyz(dx[v + 1]))));
        float pad = P->scaledDetectionRadius;
        if (P->useVelocityRadius) pad += vpad;
        lo = lo - v3{pad, pad, pad};
        hi = hi + v3{pad, pad, pad};
        glued = (__float_as_int(r0.z) >= 0) || (__float_as_int(prm[v + 1].z) >= 0);
    }
    aabb[2 * v] = f4(lo, 0.f);
    aabb[2 * v + 1] = f4(hi, glued ? -vpad : vpad);
    if (sk.state) {
        // a scene walking without the skin walks in every detection
        bool walk = force != 0 || sk.policy[sceneOf ? sceneOf[v] : 0].x > 0;
        if (!walk && (flags & YB_FLAG_HAS_NEXT)) {
            const float4 s0 = sk.sbox[2 * v];
            const float4 s1 = sk.sbox[2 * v + 1];
            const bool inside = (s0.x <= lo.x) & (s0.y <= lo.y) & (s0.z <= lo.z) &
                                (hi.x <= s1.x) & (hi.y <= s1.y) & (hi.z <= s1.z);
            walk = !inside || sk.candN[v] < 0;
        }
        // 2: every segment walks (a new tree), 1: some segment's list ended
        if (walk) {
            const int32_t why = force ? 2 : 1;
            sk.state[0] = why;
            if (sceneOf) sk.state[1 + sceneOf[v]] = why;
        }
    }
}

// Slot c of segment seg's kept candidate list: the first kSkinHead slots of
// every segment are packed together (a segment's usual list is one 64-byte
// piece of a line shared with its neighbours), the rest lie in an overflow
// area.  Four consecutive slots from a multiple of 4 are one aligned int4.
constexpr int kSkinHead = 16;
__device__ __forceinline__ size_t skin_slot(int n, int seg, int c) {
    return c < kSkinHead ? (size_t)seg * kSkinHead + c
                         : (size_t)n * kSkinHead + (size_t)seg * (kSkinCand - kSkinHead) +
                               (c - kSkinHead);
}

// Grown box of a segment for the contact skin: its box swept by `sweep`
// substeps of its endpoints' warm-start displacement, at most `cap` along an
// axis (per axis, the box of both endpoints at t = 0 and at the end of the
// sweep lies inside it, and so does every box between), then grown by
// `margin` on every side.  hi.w (the velocity pad) is kept.  The empty box of
// a vertex without a next vertex stays empty.
__device__ __forceinline__ void skin_box(float4& lo, float4& hi, const float4* __restrict__ dx,
                                         int seg, int n, const DevParams* __restrict__ P) {
    if (!(lo.x <= hi.x) || seg + 1 >= n) return;
    const float sweep = P->skinSweep, margin = P->skinMargin, cap = P->skinCap;
    const float4 d0 = dx[seg];
    const float4 d1 = dx[seg + 1];
    lo.x += fmaxf(sweep * fminf(0.f, fminf(d0.x, d1.x)), -cap) - margin;
    lo.y += fmaxf(sweep * fminf(0.f, fminf(d0.y, d1.y)), -cap) - margin;
    lo.z += fmaxf(sweep * fminf(0.f, fminf(d0.z, d1.z)), -cap) - margin;
    hi.x += fminf(sweep * fmaxf(0.f, fmaxf(d0.x, d1.x)), cap) + margin;
    hi.y += fminf(sweep * fmaxf(0.f, fmaxf(d0.y, d1.y)), cap) + margin;
    hi.z += fminf(sweep * fmaxf(0.f, fmaxf(d0.z, d1.z)), cap) + margin;
}

// ---------------------------------------------------------------------------
// Scene box: min/max are exact, so the result is independent of the order.
constexpr int kBoxThreads = 256;
__global__ __launch_bounds__(kBoxThreads) void k_box_partial(int n, const float4* __restrict__ aabb,
                                                             float4* __restrict__ blockBox) {
    __shared__ float sm[6][kBoxThreads];
    v3 lo = v3{INFINITY, INFINITY, INFINITY};
    v3 hi = v3{-INFINITY, -INFINITY, -INFINITY};
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        lo = vmin(lo, xyz(aabb[2 * v]));
        hi = vmax(hi, xyz(aabb[2 * v + 1]));
    }
    const int t = threadIdx.x;
    sm[0][t] = lo.x; sm[1][t] = lo.y; sm[2][t] = lo.z;
    sm[3][t] = hi.x; sm[4][t] = hi.y; sm[5][t] = hi.z;
    __syncthreads();
    for (int s = kBoxThreads / 2; s > 0; s >>= 1) {
        if (t < s) {
            for (int k = 0; k < 3; k++) sm[k][t] = fminf(sm[k][t], sm[k][t + s]);
            for (int k = 3; k < 6; k++) sm[k][t] = fmaxf(sm[k][t], sm[k][t + s]);
        }
        __syncthreads();
    }
    if (t == 0) {
        blockBox[2 * blockIdx.x] = make_float4(sm[0][0], sm[1][0], sm[2][0], 0.f);
        blockBox[2 * blockIdx.x + 1] = make_float4(sm[3][0], sm[4][0], sm[5][0], 0.f);
    }
}

__global__ __launch_bounds__(kBoxThreads) void k_box_final(int nb, const float4* __restrict__ blockBox,
                                                           float4* __restrict__ sceneBox) {
    __shared__ float sm[6][kBoxThreads];
    v3 lo = v3{INFINITY, INFINITY, INFINITY};
    v3 hi = v3{-INFINITY, -INFINITY, -INFINITY};
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        lo = vmin(lo, xyz(blockBox[2 * b]));
        hi = vmax(hi, xyz(blockBox[2 * b + 1]));
    }
    const int t = threadIdx.x;
    sm[0][t] = lo.x; sm[1][t] = lo.y; sm[2][t] = lo.z;
    sm[3][t] = hi.x; sm[4][t] = hi.y; sm[5][t] = hi.z;
    __syncthreads();
    for (int s = kBoxThreads / 2; s > 0; s >>= 1) {
        if (t < s) {
            for (int k = 0; k < 3; k++) sm[k][t] = fminf(sm[k][t], sm[k][t + s]);
            for (int k = 3; k < 6; k++) sm[k][t] = fmaxf(sm[k][t], sm[k][t + s]);
        }
        __syncthreads();
    }
    if (t == 0) {
        sceneBox[0] = make_float4(sm[0][0], sm[1][0], sm[2][0], 0.f);
        sceneBox[1] = make_float4(sm[3][0], sm[4][0], sm[5][0], 0.f);
    }
}

// Batched scenes: one workgroup per scene reduces that scene's boxes
// [off[s], off[s+1]) into sceneBoxes[2s], [2s+1] (the same exact min/max as a
// standalone run of the scene).
__global__ __launch_bounds__(kBoxThreads) void k_box_scenes(const int32_t* __restrict__ off,
                                                            const float4* __restrict__ aabb,
                                                            float4* __restrict__ sceneBoxes) {
    __shared__ float sm[6][kBoxThreads];
    const int sc = blockIdx.x;
    v3 lo = v3{INFINITY, INFINITY, INFINITY};
    v3 hi = v3{-INFINITY, -INFINITY, -INFINITY};
    for (int v = off[sc] + threadIdx.x; v < off[sc + 1]; v += blockDim.x) {
        lo = vmin(lo, xyz(aabb[2 * v]));
        hi = vmax(hi, xyz(aabb[2 * v + 1]));
    }
    const int t = threadIdx.x;
    sm[0][t] = lo.x; sm[1][t] = lo.y; sm[2][t] = lo.z;
    sm[3][t] = hi.x; sm[4][t] = hi.y; sm[5][t] = hi.z;
    __syncthreads();
    for (int st = kBoxThreads / 2; st > 0; st >>= 1) {
        if (t < st) {
            for (int k = 0; k < 3; k++) sm[k][t] = fminf(sm[k][t], sm[k][t + st]);
            for (int k = 3; k < 6; k++) sm[k][t] = fmaxf(sm[k][t], sm[k][t + st]);
        }
        __syncthreads();
    }
    if (t == 0) {
        sceneBoxes[2 * sc] = make_float4(sm[0][0], sm[1][0], sm[2][0], 0.f);
        sceneBoxes[2 * sc + 1] = make_float4(sm[3][0], sm[4][0], sm[5][0], 0.f);
    }
}

// ---------------------------------------------------------------------------
// Morton codes: box centre quantised to 1024 cubic cells spanning the largest
// extent of the scene box (isotropic, so a long thin scene does not get
// sliver cells), bits interleaved x-major.  Non-segments get 0xFFFFFFFF and
// sort last.  Batched scenes (sceneOf != null) quantise in their own box.
__global__ __launch_bounds__(256) void k_morton(int n, const float4* __restrict__ prm,
                                                const float4* __restrict__ aabb,
                                                const float4* __restrict__ sceneBox,
                                                const int32_t* __restrict__ sceneOf,
                                                uint32_t* __restrict__ morton,
                                                uint32_t* __restrict__ keys,
                                                int32_t* __restrict__ vals) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    uint32_t code = 0xFFFFFFFFu;
    if (__float_as_uint(prm[v].w) & YB_FLAG_HAS_NEXT) {
        const int sc = sceneOf ? sceneOf[v] : 0;
        const v3 slo = xyz(sceneBox[2 * sc]);
        const v3 shi = xyz(sceneBox[2 * sc + 1]);
        const v3 e3 = shi - slo;
        const float ext = fmaxf(fmaxf(e3.x, e3.y), e3.z);
        const v3 lo = xyz(aabb[2 * v]);
        const v3 hi = xyz(aabb[2 * v + 1]);
        const v3 c = (lo + hi) * 0.5f;
        const uint32_t qx = quant10(c.x, slo.x, ext);
        const uint32_t qy = quant10(c.y, slo.y, ext);
        const uint32_t qz = quant10(c.z, slo.z, ext);
        code = (expand_bits10(qx) << 2) | (expand_bits10(qy) << 1) | expand_bits10(qz);
    }
    morton[v] = code;
    keys[v] = code;
    vals[v] = v;
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort of (key, value), 8 bits per pass.
// Pass = histogram per 2048-key tile -> one exclusive scan over the
// digit-major [256 x tiles] table -> scatter.  The scatter ranks keys of equal
// digit in index order: within a 64-lane wave by ballot matching of the 8
// digit bits, across the block's 4 waves through LDS counters, across strips
// by a running per-digit count.  With sceneOf != null the digit is taken from
// the scene of the value instead of the key: two such passes after the four
// key passes order batched scenes scene-major, Morton order within a scene.
__device__ __forceinline__ uint32_t sort_digit(uint32_t key, int32_t val, int shift,
                                               const int32_t* __restrict__ sceneOf) {
    const uint32_t x = sceneOf ? (uint32_t)sceneOf[val] : key;
    return (x >> shift) & (kSortBins - 1);
}

__global__ __launch_bounds__(kSortThreads) void k_sort_hist(int n, int shift, int tiles,
                                                            const uint32_t* __restrict__ keys,
                                                            const int32_t* __restrict__ vals,
                                                            const int32_t* __restrict__ sceneOf,
                                                            uint32_t* __restrict__ hist) {
    __shared__ uint32_t lh[kSortBins];
    const int t = threadIdx.x;
    lh[t] = 0;
    __syncthreads();
    const int base = blockIdx.x * kSortTile;
#pragma unroll
    for (int k = 0; k < kSortItems; k++) {
        const int idx = base + k * kSortThreads + t;
        if (idx < n)
            atomicAdd(&lh[sort_digit(keys[idx], sceneOf ? vals[idx] : 0, shift, sceneOf)], 1u);
    }
    __syncthreads();
    hist[t * tiles + blockIdx.x] = lh[t];
}

// Exclusive scan of the digit-major histogram (hist[digit * tiles + tile]), in
// three steps: per-workgroup sums of kScanBlock counters, a scan of those sums,
// and a scan within each workgroup's counters from its offset.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    return x;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(int total,
                                                              const uint32_t* __restrict__ hist,
                                                              uint32_t* __restrict__ part) {
    __shared__ uint32_t ws[kScanThreads / 64];
    const int t = threadIdx.x;
    const int base = blockIdx.x * kScanBlock;
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        const int i = base + k * kScanThreads + t;
        if (i < total) sum += hist[i];
    }
    sum = wave_incl_scan(sum);
    if ((t & 63) == 63) ws[t >> 6] = sum;
    __syncthreads();
    if (t == 0) {
        uint32_t tot = 0;
        for (int w = 0; w < kScanThreads / 64; w++) tot += ws[w];
        part[blockIdx.x] = tot;
    }
}

// One workgroup: exclusive scan of the per-workgroup sums, in place.
__global__ __launch_bounds__(1024) void k_scan_parts(int parts, uint32_t* __restrict__ part) {
    __shared__ uint32_t ws[16];
    const int t = threadIdx.x;
    const int per = (parts + 1023) / 1024;
    const int b = t * per, e = min(b + per, parts);
    uint32_t sum = 0;
    for (int i = b; i < e; i++) sum += part[i];
    const uint32_t inc = wave_incl_scan(sum);
    if ((t & 63) == 63) ws[t >> 6] = inc;
    __syncthreads();
    uint32_t run = inc - sum;
    for (int w = 0; w < (t >> 6); w++) run += ws[w];
    for (int i = b; i < e; i++) {
        const uint32_t c = part[i];
        part[i] = run;
        run += c;
    }
}

// Each thread scans kScanItems consecutive counters.
__global__ __launch_bounds__(kScanThreads) void k_scan_down(int total, uint32_t* __restrict__ hist,
                                                            const uint32_t* __restrict__ part) {
    __shared__ uint32_t ws[kScanThreads / 64];
    const int t = threadIdx.x;
    const int base = blockIdx.x * kScanBlock + t * kScanItems;
    uint32_t v[kScanItems];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        v[k] = base + k < total ? hist[base + k] : 0u;
        sum += v[k];
    }
    const uint32_t inc = wave_incl_scan(sum);
    if ((t & 63) == 63) ws[t >> 6] = inc;
    __syncthreads();
    uint32_t run = part[blockIdx.x] + inc - sum;
    for (int w = 0; w < (t >> 6); w++) run += ws[w];
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        if (base + k < total) hist[base + k] = run;
        run += v[k];
    }
}

__global__ __launch_bounds__(kSortThreads) void k_sort_scatter(int n, int shift, int tiles,
                                                               const uint32_t* __restrict__ kin,
                                                               const int32_t* __restrict__ vin,
                                                               uint32_t* __restrict__ kout,
                                                               int32_t* __restrict__ vout,
                                                               const int32_t* __restrict__ sceneOf,
                                                               const uint32_t* __restrict__ hist) {
    constexpr int kWaves = kSortThreads / 64;
    __shared__ uint32_t run[kSortBins];
    __shared__ uint32_t wcnt[kWaves][kSortBins];
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int w = t >> 6;
    run[t] = hist[t * tiles + blockIdx.x];
#pragma unroll
    for (int k = 0; k < kWaves; k++) wcnt[k][t] = 0;
    __syncthreads();
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int base = blockIdx.x * kSortTile;
    for (int k = 0; k < kSortItems; k++) {
        const int idx = base + k * kSortThreads + t;
        const bool valid = idx < n;
        uint32_t key = 0;
        int32_t val = 0;
        if (valid) {
            key = kin[idx];
            val = vin[idx];
        }
        const uint32_t d = valid ? sort_digit(key, val, shift, sceneOf) : 0u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < kSortBits; bit++) {
            const bool on = (d >> bit) & 1u;
            const uint64_t m = __ballot(on);
            peers &= on ? m : ~m;
        }
        const int rank = __popcll(peers & below);
        const bool leader = valid && ((peers & below) == 0ull);
        if (leader) wcnt[w][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t dst = run[d];
            for (int ww = 0; ww < w; ww++) dst += wcnt[ww][d];
            dst += (uint32_t)rank;
            kout[dst] = key;
            vout[dst] = val;
        }
        __syncthreads();
        uint32_t tot = 0;
#pragma unroll
        for (int ww = 0; ww < kWaves; ww++) {
            tot += wcnt[ww][t];
            wcnt[ww][t] = 0;
        }
        run[t] += tot;
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Karras (2012) radix tree.  Keys are made unique by the sorted position:
// delta(i,j) = clz(k_i ^ k_j), or 32 + clz(i ^ j) for equal codes.  Batched
// scenes (sceneOf != null) extend the key to (scene, code, position within
// the scene): delta = clz(s_i ^ s_j) across scenes, 32 + the single-scene
// delta within one, so every scene is one subtree shaped exactly as in a
// standalone run of that scene; k_karras records that subtree's root.
// Node i (internal, 0..n-2) stores its children's boxes inline:
// nodes[4i+0] = (lo_left, child code), nodes[4i+1] = (hi_left, leaf pad),
// nodes[4i+2] = (lo_right, child code), nodes[4i+3] = (hi_right, leaf pad),
// child code = internal index (>= 0) or ~segment for a leaf; "leaf pad" is the
// leaf's signed velocity pad (see k_aabb).  Also written per node: its sorted
// leaf range rng[i] = (first, last) and kids[i] with leaves as ~sorted position.
struct SceneKeys {
    const int32_t* sceneOf;   // scene of each vertex = of each sorted position (scene-major)
    const int32_t* sceneOff;  // first vertex of each scene, numScenes + 1 entries
    int32_t* sceneRoot;       // out: root node of each scene's subtree
};

__device__ __forceinline__ int kdelta(const uint32_t* __restrict__ keys, int n, int a, int b,
                                      const SceneKeys& sk) {
    if (b < 0 || b > n - 1) return -1;
    int la = a, lb = b;
    int base = 0;
    if (sk.sceneOf) {
        const int sa = sk.sceneOf[a], sb = sk.sceneOf[b];
        if (sa != sb) return __clz((uint32_t)(sa ^ sb));
        la -= sk.sceneOff[sa];
        lb -= sk.sceneOff[sa];
        base = 32;
    }
    const uint32_t ka = keys[a];
    const uint32_t kb = keys[b];
    if (ka == kb) return base + 32 + __clz((uint32_t)(la ^ lb));
    return base + __clz(ka ^ kb);
}

__global__ __launch_bounds__(256) void k_karras(int n, const uint32_t* __restrict__ keys,
                                                const int32_t* __restrict__ order,
                                                float4* __restrict__ nodes,
                                                int32_t* __restrict__ parent,
                                                int2* __restrict__ rng, int2* __restrict__ kids,
                                                int32_t* __restrict__ strad,
                                                int32_t* __restrict__ stradCount,
                                                int32_t* __restrict__ leafNeed, SceneKeys sk,
                                                int leafShift) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const int d = (kdelta(keys, n, i, i + 1, sk) - kdelta(keys, n, i, i - 1, sk)) >= 0 ? 1 : -1;
    const int dmin = kdelta(keys, n, i, i - d, sk);
    int lmax = 2;
    while (kdelta(keys, n, i, i + lmax * d, sk) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (kdelta(keys, n, i, i + (l + t) * d, sk) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = kdelta(keys, n, i, j, sk);
    int s = 0;
    int t = l;
    do {
        t = (t + 1) >> 1;
        if (kdelta(keys, n, i, i + (s + t) * d, sk) > dnode) s += t;
    } while (t > 1);
    const int gamma = i + s * d + min(d, 0);
    const int lo = min(i, j);
    const int hi = max(i, j);
    // the children's parent slots carry whether this node straddles refit blocks
    const bool straddles = (lo >> leafShift) != (hi >> leafShift);
    const int32_t sf = straddles ? kParentStrad : 0;
    int cl, cr, kl, kr;
    if (lo == gamma) {
        cl = ~order[gamma];
        kl = ~gamma;
        parent[(n - 1) + gamma] = (i << 1) | 0 | sf;
    } else {
        cl = kl = gamma;
        parent[gamma] = (i << 1) | 0 | sf;
    }
    if (hi == gamma + 1) {
        cr = ~order[gamma + 1];
        kr = ~(gamma + 1);
        parent[(n - 1) + gamma + 1] = (i << 1) | 1 | sf;
    } else {
        cr = kr = gamma + 1;
        parent[gamma + 1] = (i << 1) | 1 | sf;
    }
    nodes[4 * i + 0].w = __int_as_float(cl);
    nodes[4 * i + 2].w = __int_as_float(cr);
    rng[i] = make_int2(lo, hi);
    kids[i] = make_int2(kl, kr);
    if (i == 0) parent[0] = -1;
    if (sk.sceneOf) {
        const int sc = sk.sceneOf[lo];
        if (lo == sk.sceneOff[sc] && hi == sk.sceneOff[sc + 1] - 1) sk.sceneRoot[sc] = i;
    }
    // Nodes whose leaf range crosses a refit-block boundary are finished by
    // the range-union pass; their end leaves need suffix / prefix unions.
    if (straddles) {
        const int slot = atomicAdd(stradCount, 1);
        strad[slot] = i;
        atomicOr(&leafNeed[lo], 1);
        atomicOr(&leafNeed[hi], 2);
    }
}

// ---------------------------------------------------------------------------
// Refit without device-scope fences.  Sorted leaves are split into blocks of
// kLeaves (the refit tile: 256, 512 or 1024, chosen by timing).  (1) k_refit_local: one workgroup per block climbs from every
// leaf through the nodes whose range lies inside the block, with LDS arrival
// counters and LDS boxes (the second arrival at a node forms the union); it
// writes those nodes' child boxes, the box of every child whose parent
// straddles a block boundary, the block's union, and suffix / prefix unions
// at the leaves that end a straddling range.  (2) k_refit_table: a sparse
// table of block unions.  (3) k_refit_top: each straddling node's box is the
// union of a suffix, at most two overlapping sparse-table windows and a
// prefix -- computed directly, so no node waits on another.  Every step is a
// min/max, hence exact and independent of scheduling.
// One thread per leaf: each climb is a chain of dependent loads, and with a
// leaf per thread all of a block's climbs are in flight at once.  Every global
// input of a leaf is loaded up front, so a block waits on memory once.

__device__ __forceinline__ void box_union(float4& lo, float4& hi, float4 lo2, float4 hi2) {
    lo = make_float4(fminf(lo.x, lo2.x), fminf(lo.y, lo2.y), fminf(lo.z, lo2.z), 0.f);
    hi = make_float4(fmaxf(hi.x, hi2.x), fmaxf(hi.y, hi2.y), fmaxf(hi.z, hi2.z), 0.f);
}
__device__ __forceinline__ void store_xyz(float4* p, float4 v) {
    float* f = reinterpret_cast<float*>(p);
    f[0] = v.x;
    f[1] = v.y;
    f[2] = v.z;
}

template <int kLeaves>
__global__ __launch_bounds__(kLeaves) void k_refit_local(
    int n, const int32_t* __restrict__ order, const float4* __restrict__ aabb,
    float4* __restrict__ nodes, const int32_t* __restrict__ parent, const int2* __restrict__ rng,
    const int2* __restrict__ kids, const int32_t* __restrict__ leafNeed,
    float4* __restrict__ sufBox, float4* __restrict__ preBox, float4* __restrict__ blkBox,
    float4* __restrict__ sceneBox, const DevParams* __restrict__ P,
    const float4* __restrict__ dx, SkinArgs sk, const int32_t* __restrict__ sceneOf) {
    // contact skin: no segment walks in this substep, so nothing reads the tree
    if (sk.state && sk.state[0] == 0) return;
    __shared__ float4 lbox[kLeaves][2];
    __shared__ float4 nbox[kLeaves][2];
    __shared__ int2 lkids[kLeaves];          // kids[] of nodes base..end-1
    __shared__ int32_t lpar[kLeaves];        // parent[] of nodes base..end-1
    __shared__ uint32_t arrived[kLeaves / 32];  // first child done, per node
    const int t = threadIdx.x;
    const int base = blockIdx.x * kLeaves;
    const int end = min(base + kLeaves, n);
    const int cntLeaves = end - base;
    // A Karras node's range starts or ends at its own index, so every node
    // whose range lies in [base, end) has its index there too: its topology
    // is loaded here once, with coalesced reads, together with the leaf's
    // parent slot, the node's range and the leaf's suffix / prefix needs.
    if (t < kLeaves / 32) arrived[t] = 0u;
    int slot = -1, need = 0;
    int2 pr = make_int2(-1, -1);
    if (base + t < end) {
        const int seg = order[base + t];
        slot = parent[(n - 1) + base + t];
        need = leafNeed[base + t];
        if (base + t < n - 1) {
            lkids[t] = kids[base + t];
            lpar[t] = parent[base + t];
            pr = rng[base + t];
        }
        float4 lo = aabb[2 * seg];
        float4 hi = aabb[2 * seg + 1];
        const int sc = sceneOf ? sceneOf[seg] : 0;
        if (sk.state && sk.policy[sc].x == 0) {
            // a scene that keeps candidates has grown boxes in the tree; a
            // walking one keeps them for its next detections
            skin_box(lo, hi, dx, seg, n, P);
            if (!sceneOf || sk.state[1 + sc]) {
                sk.sbox[2 * seg] = lo;
                sk.sbox[2 * seg + 1] = hi;
            }
        }
        lbox[t][0] = lo;
        lbox[t][1] = hi;
    }
    __syncthreads();

    // Climb from each leaf through the nodes whose range lies in this block,
    // in LDS only; the second child to arrive at a node forms its box.
    if (base + t < end) {
        float4 lo = lbox[t][0];
        float4 hi = lbox[t][1];
        bool leaf = true;
        while (slot >= 0) {
            const int p = (slot & ~kParentStrad) >> 1;
            if (slot & kParentStrad) {
                // p straddles: publish this child's box in p's slot for traversal.
                const int side = slot & 1;
                store_xyz(&nodes[4 * p + 2 * side], lo);
                nodes[4 * p + 2 * side + 1] = make_float4(hi.x, hi.y, hi.z, leaf ? hi.w : 0.f);
                break;
            }
            const int q = p - base;
            const uint32_t bit = 1u << (q & 31);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            const uint32_t old = atomicOr(&arrived[q >> 5], bit);
            if (!(old & bit)) break;
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            const int2 kc = lkids[q];
            float4 lo1, hi1;
            if (kc.x < 0) { lo = lbox[~kc.x - base][0]; hi = lbox[~kc.x - base][1]; }
            else { lo = nbox[kc.x - base][0]; hi = nbox[kc.x - base][1]; }
            if (kc.y < 0) { lo1 = lbox[~kc.y - base][0]; hi1 = lbox[~kc.y - base][1]; }
            else { lo1 = nbox[kc.y - base][0]; hi1 = nbox[kc.y - base][1]; }
            box_union(lo, hi, lo1, hi1);
            nbox[q][0] = lo;
            nbox[q][1] = hi;
            leaf = false;
            slot = lpar[q];
        }
    }
    __syncthreads();

    // Child slots of every node completed in this block (coalesced pass).
    const int p = base + t;
    if (p < end && p < n - 1 && pr.x >= base && pr.y < end) {
        const int2 kc = lkids[t];
        float4 lo0, hi0, lo1, hi1;
        if (kc.x < 0) { lo0 = lbox[~kc.x - base][0]; hi0 = lbox[~kc.x - base][1]; }
        else { lo0 = nbox[kc.x - base][0]; hi0 = nbox[kc.x - base][1]; }
        if (kc.y < 0) { lo1 = lbox[~kc.y - base][0]; hi1 = lbox[~kc.y - base][1]; }
        else { lo1 = nbox[kc.y - base][0]; hi1 = nbox[kc.y - base][1]; }
        store_xyz(&nodes[4 * p + 0], lo0);
        nodes[4 * p + 1] = make_float4(hi0.x, hi0.y, hi0.z, kc.x < 0 ? hi0.w : 0.f);
        store_xyz(&nodes[4 * p + 2], lo1);
        nodes[4 * p + 3] = make_float4(hi1.x, hi1.y, hi1.z, kc.y < 0 ? hi1.w : 0.f);
        if (p == 0) {  // the whole tree fits in this block
            box_union(lo0, hi0, lo1, hi1);
            sceneBox[0] = lo0;
            sceneBox[1] = hi0;
        }
    }

    // Inclusive prefix and suffix unions over the block's leaves, for the ends
    // of straddling ranges: scans within each wave by shuffles, then across
    // the block's waves through LDS.  Leaves past the block's end are empty
    // boxes.  Min/max make every order of combination give the same box.
    static_assert(kLeaves % 64 == 0, "whole waves");
    constexpr int kWaves = kLeaves / 64;
    __shared__ float4 wpre[kWaves][2], wsuf[kWaves][2];
    const int lane = t & 63, w = t >> 6;
    float4 plo = make_float4(INFINITY, INFINITY, INFINITY, 0.f);
    float4 phi = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.f);
    if (t < cntLeaves) {
        plo = lbox[t][0];
        phi = lbox[t][1];
    }
    float4 slo = plo, shi = phi;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const float4 ulo = make_float4(__shfl_up(plo.x, off, 64), __shfl_up(plo.y, off, 64),
                                       __shfl_up(plo.z, off, 64), 0.f);
        const float4 uhi = make_float4(__shfl_up(phi.x, off, 64), __shfl_up(phi.y, off, 64),
                                       __shfl_up(phi.z, off, 64), 0.f);
        const float4 dlo = make_float4(__shfl_down(slo.x, off, 64), __shfl_down(slo.y, off, 64),
                                       __shfl_down(slo.z, off, 64), 0.f);
        const float4 dhi = make_float4(__shfl_down(shi.x, off, 64), __shfl_down(shi.y, off, 64),
                                       __shfl_down(shi.z, off, 64), 0.f);
        if (lane >= off) box_union(plo, phi, ulo, uhi);
        if (lane + off < 64) box_union(slo, shi, dlo, dhi);
    }
    if (lane == 63) {
        wpre[w][0] = plo;
        wpre[w][1] = phi;
    }
    if (lane == 0) {
        wsuf[w][0] = slo;
        wsuf[w][1] = shi;
    }
    __syncthreads();
    for (int k = 0; k < w; k++) box_union(plo, phi, wpre[k][0], wpre[k][1]);
    for (int k = w + 1; k < kWaves; k++) box_union(slo, shi, wsuf[k][0], wsuf[k][1]);
    {
        const int k = base + t;
        if (k < end) {
            if (need & 1) {
                sufBox[2 * k] = slo;
                sufBox[2 * k + 1] = shi;
            }
            if (need & 2) {
                preBox[2 * k] = plo;
                preBox[2 * k + 1] = phi;
            }
            if (t == cntLeaves - 1) {  // block union = prefix at the last leaf
                blkBox[2 * blockIdx.x] = plo;
                blkBox[2 * blockIdx.x + 1] = phi;
            }
        }
    }
}

// Sparse table over block unions: level L, entry b = union of blocks
// [b, b + 2^L) (clipped at nb).  Levels 1..kTabLocal: one workgroup per 1024
// entries, working in LDS on a 2048-entry window (entry b at level L needs
// level L-1 at b + 2^(L-1), at most 512 past the workgroup's last entry).
// Higher levels (more than 2048 blocks): one launch each over all entries.
constexpr int kTabLocal = 10;
__global__ __launch_bounds__(1024) void k_refit_table(int nb, int levels,
                                                      float4* __restrict__ tab,
                                                      const int32_t* __restrict__ skinState) {
    if (skinState && skinState[0] == 0) return;
    __shared__ float4 wlo[2048], whi[2048];
    const int t = threadIdx.x;
    const int b0 = blockIdx.x * 1024;
    for (int j = t; j < 2048; j += 1024) {
        const int b = b0 + j;
        wlo[j] = b < nb ? tab[2 * b] : make_float4(INFINITY, INFINITY, INFINITY, 0.f);
        whi[j] = b < nb ? tab[2 * b + 1] : make_float4(-INFINITY, -INFINITY, -INFINITY, 0.f);
    }
    const int top = min(levels - 1, kTabLocal);
    for (int L = 1; L <= top; L++) {
        const int half = 1 << (L - 1);
        __syncthreads();
        float4 lo[2], hi[2];
        for (int r = 0; r < 2; r++) {
            const int j = t + 1024 * r;
            lo[r] = wlo[j];
            hi[r] = whi[j];
            if (j + half < 2048) box_union(lo[r], hi[r], wlo[j + half], whi[j + half]);
        }
        __syncthreads();
        for (int r = 0; r < 2; r++) {
            wlo[t + 1024 * r] = lo[r];
            whi[t + 1024 * r] = hi[r];
        }
        const int b = b0 + t;
        if (b < nb) {
            float4* dst = tab + (size_t)2 * nb * L;
            dst[2 * b] = lo[0];
            dst[2 * b + 1] = hi[0];
        }
    }
}

__global__ __launch_bounds__(256) void k_refit_table_level(int nb, int L,
                                                           float4* __restrict__ tab,
                                                           const int32_t* __restrict__ skinState) {
    if (skinState && skinState[0] == 0) return;
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const int half = 1 << (L - 1);
    const float4* src = tab + (size_t)2 * nb * (L - 1);
    float4* dst = tab + (size_t)2 * nb * L;
    float4 lo = src[2 * b], hi = src[2 * b + 1];
    if (b + half < nb) box_union(lo, hi, src[2 * (b + half)], src[2 * (b + half) + 1]);
    dst[2 * b] = lo;
    dst[2 * b + 1] = hi;
}

__global__ __launch_bounds__(256) void k_refit_top(int n, int nb, int leafShift,
                                                   const int32_t* __restrict__ strad,
                                                   const int32_t* __restrict__ stradCount,
                                                   const int2* __restrict__ rng,
                                                   const int32_t* __restrict__ parent,
                                                   const float4* __restrict__ sufBox,
                                                   const float4* __restrict__ preBox,
                                                   const float4* __restrict__ tab,
                                                   float4* __restrict__ nodes,
                                                   float4* __restrict__ sceneBox,
                                                   const int32_t* __restrict__ skinState) {
    if (skinState && skinState[0] == 0) return;
    const int count = *stradCount;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < count; s += gridDim.x * blockDim.x) {
        const int i = strad[s];
        const int2 r = rng[i];
        float4 lo = sufBox[2 * r.x], hi = sufBox[2 * r.x + 1];
        box_union(lo, hi, preBox[2 * r.y], preBox[2 * r.y + 1]);
        const int b0 = (r.x >> leafShift) + 1;
        const int b1 = (r.y >> leafShift) - 1;
        if (b1 >= b0) {
            const int span = b1 - b0 + 1;
            const int L = 31 - __clz(span);
            const float4* T = tab + (size_t)2 * nb * L;
            box_union(lo, hi, T[2 * b0], T[2 * b0 + 1]);
            const int b2 = b1 - (1 << L) + 1;
            box_union(lo, hi, T[2 * b2], T[2 * b2 + 1]);
        }
        if (i == 0) {
            sceneBox[0] = lo;
            sceneBox[1] = hi;
        } else {
            const int slot = parent[i] & ~kParentStrad;
            const int p = slot >> 1;
            const int side = slot & 1;
            store_xyz(&nodes[4 * p + 2 * side], lo);
            nodes[4 * p + 2 * side + 1] = make_float4(hi.x, hi.y, hi.z, 0.f);
        }
    }
}

// ---------------------------------------------------------------------------
// Query + narrowphase (collision.cu:39-92).  One thread per segment.  Each overlapping
// leaf j is filtered as the unordered pair (x, y) = (min, max): glued
// endpoints and |x - y| <= 2 are exempt, and the pair is kept when the closest
// distance between the segments is below 2 * scaledDetectionRadius plus both
// segments' largest warm-start displacement.  Because the test is symmetric,
// segment i lists j exactly when j lists i -- the set the reference builds
// with two atomicAdds per pair.
// Traversal stack: pending right children.  The pending entries lie on one
// root-to-leaf path, so their number is below the tree depth: at most the 30
// Morton bits plus the index bits that split equal codes (28 at the largest
// handle) plus 16 scene bits of a batch -- 64 holds any tree whose levels do
// not all split in two.  A query that still runs out skips the subtree it
// cannot push and reports YB_ERROR_TRAVERSAL_STACK_EXCEEDED.
constexpr int kStack = YB_QUERY_STACK_DEPTH;

// All six comparisons are evaluated (no short circuit): six compares and five
// mask ANDs, with no branches.
__device__ __forceinline__ bool overlap(v3 qlo, v3 qhi, float4 lo, float4 hi) {
    return (qlo.x <= hi.x) & (lo.x <= qhi.x) & (qlo.y <= hi.y) & (lo.y <= qhi.y) &
           (qlo.z <= hi.z) & (lo.z <= qhi.z);
}

struct SegInfo {
    v3 p0, p1;
    float vpad;
    bool glued;
};

// d2: the squared distance of the two segments' closest points when the pair
// reaches the distance test.
__device__ __forceinline__ bool pair_accepted(const DevParams* __restrict__ P,
                                              const float4* __restrict__ prm, int i,
                                              const SegInfo& si, int j, const SegInfo& sj,
                                              float& d2) {
    const bool iFirst = i < j;
    const int x = iFirst ? i : j;
    const int y = iFirst ? j : i;
    if (y - x <= 2) return false;
    const SegInfo& sx = iFirst ? si : sj;
    const SegInfo& sy = iFirst ? sj : si;
    if (sx.glued) {
        const int c0 = __float_as_int(prm[x].z);
        const int c1 = __float_as_int(prm[x + 1].z);
        if (c0 == y || c1 == y || c0 == y + 1 || c1 == y + 1) return false;
    }
    const v3 a0 = sx.p0;
    const v3 a1 = sx.p1 - a0;
    const v3 b0 = sy.p0 - a0;
    const v3 b1 = sy.p1 - a0;
    float u, w;
    closest_uv(v3{0.f, 0.f, 0.f}, a1, b0, b1, u, w);
    const v3 nrm = u * a1 - mixv(b0, b1, w);
    d2 = len2(nrm);
    float r = 2.0f * P->scaledDetectionRadius;
    if (P->useVelocityRadius) {
        r += sx.vpad;
        r += sy.vpad;
    }
    return d2 < r * r;
}

// kStackL: stack entries kept in LDS (deeper ones go to a private array in
// scratch memory); kCand: candidate leaves buffered per thread before the
// narrowphase.  Both size the LDS footprint, which sets how many waves a CU
// holds: (kStackL + 2 kCand) x 4 B per thread.
// With dxLimit set, the query also does the step limit's work (collision.cu:
// 121-159) from the distances its pair tests have just measured: the limit of
// each vertex into dxLimit[i].w (dxLimit holds the warm start), the
// interpenetration warning into err[1] and the smallest separation into
// *minSep -- the contact lists are the pairs it accepts, so these are the
// distances the step limit kernel would measure again.
// Contact skin (sk.state set; the tree then holds grown boxes): a segment of a
// walking scene walks with its grown box and keeps every leaf it meets as a
// candidate; the kept candidates (fresh or from an earlier walk) are tested
// against the current box before the pair test.  Both visit the candidates
// in the tree's in-order leaf order, as the walk without the skin does, so
// every mode gives the same lists.  Two launches share the work, each with
// its own register budget: phase 0 makes every walk, phase 1 tests kept
// lists; a segment's lists, list length and step limit are written by the
// phase that finishes it, and a wave with nothing to do in a phase returns
// at once.
template <int kQueryThreads, int kStackL, int kCand, int kPhase>
__global__ __launch_bounds__(kQueryThreads) void k_query(const DevParams* __restrict__ P,
                                                         const int32_t* __restrict__ order,
                                                         const float4* __restrict__ aabb,
                                                         const float4* __restrict__ nodes,
                                                         const float4* __restrict__ pos,
                                                         const float4* __restrict__ prm,
                                                         int32_t* __restrict__ numCols,
                                                         int32_t* __restrict__ cols,
                                                         int32_t* __restrict__ err,
                                                         const int32_t* __restrict__ sceneOf,
                                                         const int32_t* __restrict__ sceneRoot,
                                                         int stackLimit,
                                                         int32_t* __restrict__ candCount,
                                                         float4* __restrict__ dxLimit,
                                                         float* __restrict__ minSep,
                                                         SkinArgs sk) {
    // One LDS array, kRows words per thread: the traversal stack (kStackL
    // rows), the narrowphase buffer (kCand rows of leaves, kCand of pads as
    // float bits); a reusing query lays its wave's work list over all of it.
    constexpr int kRows = kStackL + 2 * kCand;
    __shared__ int32_t lds[kRows][kQueryThreads];
    int32_t(*stk)[kQueryThreads] = lds;
    int32_t(*candJ)[kQueryThreads] = lds + kStackL;
    int32_t(*candPad)[kQueryThreads] = lds + kStackL + kCand;
    const int n = P->n;
    // Segments in index order: consecutive segments of a strand are spatial
    // neighbours (coherent traversals) and the slot-major list writes
    // cols[i + n * slot] coalesce across the wave.  Blocks walk the index
    // range XCD by XCD (xcd_tile), so the tree leaves and partner segments
    // that one stretch of strands visits stay in one L2.
    const int i = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const bool seg = i < n && (__float_as_uint(prm[min(i, n - 1)].w) & YB_FLAG_HAS_NEXT);
    // 0: walk (no skin), 1: walk with the grown box and keep the candidates,
    // 2: test the kept candidates, 3: walk without keeping (the scene's lists
    // did not last; the tree holds its current boxes)
    int mode = 0;
    if (sk.state) {
        const int sc = sceneOf ? sceneOf[min(i, n - 1)] : 0;
        mode = sk.policy[sc].x > 0 ? 3 : (sk.state[sceneOf ? 1 + sc : 0] ? 1 : 2);
        if (kPhase == 0 && sk.stats && blockIdx.x == 0 && threadIdx.x == 0)
            atomicAdd(&sk.stats[sk.state[0] ? 0 : 1], 1ull);
    }
    // fin: this phase writes the segment's lists and step limit; phase 0
    // finishes every mode but the reuse of a kept list, phase 1 that reuse
    // (of a list kept now, or earlier)
    bool fin;
    if constexpr (kPhase == 0)
        fin = i < n && mode != 2;
    else
        fin = i < n && (mode == 2 || (mode == 1 && seg && sk.candN[i] >= 0));
    if (!__any(fin)) return;  // the whole wave is another phase's
    float minD2 = INFINITY;  // closest listed contact, squared
    if (!seg) {
        if (fin) {
            numCols[i] = 0;
            if (mode == 1) sk.candN[i] = 0;
            if (candCount) {
                candCount[i] = 0;
                candCount[n + i] = 0;
                candCount[2 * n + i] = 0;
            }
        }
    } else {
        const float4 qlo4 = aabb[2 * i];
        const float4 qhi4 = aabb[2 * i + 1];
        const v3 qlo = xyz(qlo4);
        const v3 qhi = xyz(qhi4);
        SegInfo si;
        si.p0 = xyz(pos[i]);
        si.p1 = xyz(pos[i + 1]);
        si.vpad = fabsf(qhi4.w);
        si.glued = signbit(qhi4.w);
        const int t = threadIdx.x;
        int cnt = 0;
        bool overflowStack = false;
        // The traversal only collects overlapping leaves, in visit order; the
        // exact pair test runs over the buffered candidates when some thread of
        // the wave has nearly filled its buffer, so the wave evaluates pair tests
        // together instead of once per node visit at which any thread met a leaf.
        // Candidates keep their visit order, hence so do the contact lists.
        int nb = 0;
        // measurement only: leaves tested by the narrowphase, node visits (when
        // reusing: kept candidates scanned), deepest stack
        int cand = 0, visits = 0, maxSp = 0;
        // The first four partners stay in registers and are stored at the end,
        // slot by slot, by the whole wave at once (one coalesced 256-byte store
        // per slot); storing each partner when it is found left every slot's lines
        // partly written at different times, and the cache wrote them back several
        // times over.  Later partners are stored as they are found.
        int32_t r0 = 0, r1 = 0, r2 = 0, r3 = 0;
        // exact: the buffered leaves met the grown box only; their current
        // boxes are tested here first
        auto narrowphase = [&](bool exact) {
            cand += nb;
    #pragma unroll 1
            for (int k = 0; k < nb; k++) {
                const int j = candJ[k][t];
                float pw = __int_as_float(candPad[k][t]);
                bool hit = true;
                if (exact) {
                    const float4 blo = aabb[2 * j];
                    const float4 bhi = aabb[2 * j + 1];
                    hit = overlap(qlo, qhi, blo, bhi);
                    pw = bhi.w;
                }
                SegInfo sj;
                sj.p0 = xyz(gather4(pos, j));
                sj.p1 = xyz(gather4<true>(pos, j));
                sj.vpad = fabsf(pw);
                sj.glued = signbit(pw);
                float d2;
                if (pair_accepted(P, prm, i, si, j, sj, d2) && hit) {
                    minD2 = fminf(minD2, d2);
                    r0 = cnt == 0 ? j : r0;
                    r1 = cnt == 1 ? j : r1;
                    r2 = cnt == 2 ? j : r2;
                    r3 = cnt == 3 ? j : r3;
                    if (cnt >= 4 && cnt < kMaxCols) cols[i + (size_t)n * cnt] = j;
                    cnt++;
                }
            }
            nb = 0;
        };
        // Tree walk from `node` with the box (tlo, thi).  rec: keep every leaf
        // met (its grown box overlaps) in this segment's candidate list; nar:
        // test the leaves met (exact: against their current boxes first, when
        // the tree holds grown boxes).  Returns the leaves kept.
        // A recording walk keeps each lane's leaves in its LDS columns (the
        // narrowphase buffer, unused then: 2 kCand entries) and stores them
        // to its contiguous list in bursts, which the L2 merges into whole
        // lines; one store per leaf as it is met left most lines partly
        // written at different times (5x the list's bytes written).
        auto recSlot = [&](int k) -> int32_t& {
            return k < kCand ? candJ[k][t] : candPad[k - kCand][t];
        };
        auto walk = [&](v3 tlo, v3 thi, bool rec, bool nar, bool exact) {
            int nc = 0, nr = 0;
            auto flushRec = [&]() {
    #pragma unroll 1
                for (int k = 0; k < nr; k++)
                    if (nc + k < kSkinCand) sk.cand[skin_slot(n, i, nc + k)] = recSlot(k);
                nc += nr;
                nr = 0;
            };
            int sp = 0;
            volatile int32_t deep[kStack > kStackL ? kStack - kStackL : 1];  // entries past the LDS part
            // batched scenes: traverse only this segment's own scene subtree
            int node = sceneOf ? sceneRoot[sceneOf[i]] : 0;
            while (true) {
                if (candCount) {
                    visits++;
                    maxSp = max(maxSp, sp);
                }
                const float4 l0 = nodes[4 * node + 0];
                const float4 l1 = nodes[4 * node + 1];
                const float4 r0 = nodes[4 * node + 2];
                const float4 r1 = nodes[4 * node + 3];
                const int cl = __float_as_int(l0.w);
                const int cr = __float_as_int(r0.w);
                const bool hl = overlap(tlo, thi, l0, l1);
                const bool hr = overlap(tlo, thi, r0, r1);
                // leaves within two indices of i (itself and its strand neighbours,
                // which always overlap) are exempt from the pair test (collision.cu:54):
                // they are dropped here, before their positions are gathered.  Both
                // leaves are written without a branch: slot nb and the one after it
                // are free (nb <= kCand - 2 at every visit), and nb advances only for
                // an accepted leaf.
                const bool addL = hl && cl < 0 && abs(~cl - i) > 2;
                const bool addR = hr && cr < 0 && abs(~cr - i) > 2;
                if (rec) {
                    // slots nr and nr + 1 are free (nr <= 2 kCand - 2 here)
                    recSlot(nr) = ~cl;
                    nr += addL ? 1 : 0;
                    recSlot(nr) = ~cr;
                    nr += addR ? 1 : 0;
                    if (__any(nr > 2 * kCand - 2)) flushRec();
                }
                if (nar) {
                    candJ[nb][t] = ~cl;
                    candPad[nb][t] = __float_as_int(l1.w);
                    nb += addL ? 1 : 0;
                    candJ[nb][t] = ~cr;
                    candPad[nb][t] = __float_as_int(r1.w);
                    nb += addR ? 1 : 0;
                }
                // at most two leaves per visit: flush before a buffer could overflow
                if (__any(nb > kCand - 2)) narrowphase(exact);
                const bool tl = hl && cl >= 0;
                const bool tr = hr && cr >= 0;
                const bool descend = tl || tr;
                if (!descend && sp == 0) break;
                // Next node: descend left first (the right child is pushed when both
                // overlap), else pop.  stackLimit <= kStack (host-checked); tests lower
                // it to force the overflow path.  The LDS part of the stack is used
                // without branches: the slot above the top is free, so the write is
                // harmless when nothing is pushed, and the pop reads the top entry.
                const bool both = tl && tr;
                const bool push = both && sp < stackLimit;
                overflowStack = overflowStack || (both && !push);
                if (sp < kStackL) stk[sp][t] = cr;
                const int popped = stk[min(max(sp - 1, 0), kStackL - 1)][t];
                if (push && sp >= kStackL) deep[sp - kStackL] = cr;
                node = descend ? (tl ? cl : cr) : popped;
                sp = descend ? sp + (push ? 1 : 0) : sp - 1;
                if (!descend && sp >= kStackL) {
                    // keeps the compiler from merging the two pops into one generic
                    // (flat) load, which would make every pop a slow volatile one
                    asm volatile("" ::: "memory");
                    node = deep[sp - kStackL];
                }
            }
            narrowphase(exact);
            flushRec();
            return nc;
        };
        // Kept candidates, in the order the walk met them, kG per lane and
        // step.  Their current boxes are tested against this segment's; the
        // pair tests of the ones that overlap are then shared out over the
        // wave: each step lays the wave's overlapping candidates out in one
        // LDS list (by candidate slot, then lane), every lane tests one entry
        // per round (the owner segment's endpoints and pad read from LDS),
        // and each owner takes its results in slot order -- so the lists are
        // the same as testing them lane by lane, in about half the rounds
        // (a wave otherwise runs as many tests as its busiest lane needs).
        constexpr int kOwnRows = 7;                          // owner p0, p1, signed pad
        constexpr int kEntRows = (kRows - kOwnRows) / 3;     // rows per entry field
        constexpr int kG = kEntRows >= 4 ? 4 : 1;            // candidates per lane per step
        static_assert(kEntRows >= 1, "LDS shape too small for the reuse list");
        const int wb = threadIdx.x & ~63;  // the wave's first column
        const int ln = threadIdx.x & 63;
        const int iBase = i - ln;          // segment of the wave's lane 0
        auto wave_sync = []() {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        };
        auto ent = [&](int field, int e) -> int32_t& {
            return lds[kOwnRows + field * kEntRows + (e >> 6)][wb + (e & 63)];
        };
        auto reuse = [&](int cn, bool countScan) {
            lds[0][wb + ln] = __float_as_int(si.p0.x);
            lds[1][wb + ln] = __float_as_int(si.p0.y);
            lds[2][wb + ln] = __float_as_int(si.p0.z);
            lds[3][wb + ln] = __float_as_int(si.p1.x);
            lds[4][wb + ln] = __float_as_int(si.p1.y);
            lds[5][wb + ln] = __float_as_int(si.p1.z);
            lds[6][wb + ln] = __float_as_int(si.glued ? -si.vpad : si.vpad);
            for (int c = 0; __any(c < cn); c += kG) {
                int js[kG];
                if constexpr (kG == 4) {
                    const int4 q = c < cn ? *reinterpret_cast<const int4*>(sk.cand + skin_slot(n, i, c))
                                          : make_int4(i, i, i, i);
                    js[0] = c < cn ? q.x : i;
                    js[1] = c + 1 < cn ? q.y : i;
                    js[2] = c + 2 < cn ? q.z : i;
                    js[3] = c + 3 < cn ? q.w : i;
                } else {
                    js[0] = c < cn ? sk.cand[skin_slot(n, i, c)] : i;
                }
                float4 blo[kG], bhi[kG];
    #pragma unroll
                for (int k = 0; k < kG; k++) {
                    blo[k] = aabb[2 * js[k]];
                    bhi[k] = aabb[2 * js[k] + 1];
                }
                // list positions: slot k of every lane, then slot k + 1
                bool hitk[kG];
                int at[kG];
                int m = 0;
    #pragma unroll
                for (int k = 0; k < kG; k++) {
                    hitk[k] = c + k < cn && overlap(qlo, qhi, blo[k], bhi[k]);
                    const uint64_t b = __ballot(hitk[k]);
                    at[k] = m + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32),
                                                                __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0));
                    m += __popcll(b);
                    if (hitk[k]) {
                        ent(0, at[k]) = js[k];
                        ent(1, at[k]) = ln;
                        ent(2, at[k]) = __float_as_int(bhi[k].w);
                    }
                }
                wave_sync();
                // every lane tests entries rank, rank + A, ... of the m listed
                const uint64_t act = __ballot(1);
                const int A = __popcll(act);
                const int rank = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(act >> 32),
                                                                __builtin_amdgcn_mbcnt_lo((uint32_t)act, 0));
    #pragma unroll 1
                for (int e = rank; e < m; e += A) {
                    const int j = ent(0, e);
                    const int ow = ent(1, e);
                    const float pw = __int_as_float(ent(2, e));
                    SegInfo so;
                    so.p0 = v3{__int_as_float(lds[0][wb + ow]), __int_as_float(lds[1][wb + ow]),
                               __int_as_float(lds[2][wb + ow])};
                    so.p1 = v3{__int_as_float(lds[3][wb + ow]), __int_as_float(lds[4][wb + ow]),
                               __int_as_float(lds[5][wb + ow])};
                    const float vo = __int_as_float(lds[6][wb + ow]);
                    so.vpad = fabsf(vo);
                    so.glued = signbit(vo);
                    SegInfo sj;
                    sj.p0 = xyz(gather4(pos, j));
                    sj.p1 = xyz(gather4<true>(pos, j));
                    sj.vpad = fabsf(pw);
                    sj.glued = signbit(pw);
                    float d2;
                    const bool acc = pair_accepted(P, prm, iBase + ow, so, j, sj, d2);
                    ent(2, e) = __float_as_int(acc ? d2 : -1.0f);
                }
                wave_sync();
                // each owner takes its results in slot order
    #pragma unroll
                for (int k = 0; k < kG; k++) {
                    if (!hitk[k]) continue;
                    const float d2 = __int_as_float(ent(2, at[k]));
                    cand++;
                    if (d2 >= 0.0f) {
                        const int j = js[k];
                        minD2 = fminf(minD2, d2);
                        r0 = cnt == 0 ? j : r0;
                        r1 = cnt == 1 ? j : r1;
                        r2 = cnt == 2 ? j : r2;
                        r3 = cnt == 3 ? j : r3;
                        if (cnt >= 4 && cnt < kMaxCols) cols[i + (size_t)n * cnt] = j;
                        cnt++;
                    }
                }
                wave_sync();
                if (candCount && countScan) visits += min(kG, max(cn - c, 0));
            }
        };
        if constexpr (kPhase == 1) {
            if (fin) reuse(min(sk.candN[i], kSkinCand), mode == 2);
        } else if (mode == 1) {
            // Walk with the grown box (the refit kept it in sbox) only to keep
            // the candidates; phase 1 tests them as for a reusing substep.  A
            // list that is full or missed a subtree is not kept (the next
            // detection walks again) and is not used: its segment walks once
            // more with its current box, testing every leaf it meets.
            const int nc = walk(xyz(sk.sbox[2 * i]), xyz(sk.sbox[2 * i + 1]), true, false, false);
            const bool kept = nc <= kSkinCand && !overflowStack;
            sk.candN[i] = kept ? nc : -1;
            if (kept) {
                fin = false;
                if (candCount) {
                    candCount[n + i] = visits;
                    candCount[2 * n + i] = maxSp;
                }
            } else {
                // the stack overflow reported is the one of the walk that
                // makes the list, as without the skin
                overflowStack = false;
                walk(qlo, qhi, false, true, true);
            }
        } else if (mode != 2) {
            walk(qlo, qhi, false, true, false);
            if (mode == 3) sk.candN[i] = -1;  // an old list may follow an older tree
        }
      if (fin) {
        if (cnt > 0) cols[i] = r0;
        if (cnt > 1) cols[i + (size_t)n] = r1;
        if (cnt > 2) cols[i + 2 * (size_t)n] = r2;
        if (cnt > 3) cols[i + 3 * (size_t)n] = r3;
        // err[0] collects one bit per failure kind (kErrBit*), read by the host
        if (cnt > kMaxCols) {
            atomicOr(&err[0], kErrBitContacts);
            cnt = kMaxCols;
        }
        if (overflowStack) atomicOr(&err[0], kErrBitStack);
        numCols[i] = cnt;
        if (candCount) {
            candCount[i] = cand;
            if (kPhase == 0 || mode == 2) {
                candCount[n + i] = visits;
                candCount[2 * n + i] = maxSp;
            }
        }
      }
    }
    if (dxLimit) {
        // the step limit (collision.cu:121-159), every lane of the wave
        const float minContact = sqrtf(minD2);
        if (fin) {
            float minDist = INFINITY;
            if (seg && P->useStepSizeLimit) {
                minDist = P->detectionRadius * (P->detectionScaler - 1.0f);
                if (P->useVelocityRadius) minDist += len(xyz(dxLimit[i]));
                minDist = fminf(minDist, 0.45f * minContact);
            }
            dxLimit[i].w = minDist;
            if (minContact < 2.0f * P->radius) err[1] = YB_WARNING_SEGMENT_INTERPENETRATION;
        }
        reduce_min_sep(minContact / (2.0f * P->radius), minSep);
    }
}

// Flags of each scene's last vertex (which must not start a segment).
__global__ void k_scene_end_flags(int numScenes, const int32_t* __restrict__ off,
                                  const float4* __restrict__ prm, int32_t* __restrict__ out) {
    const int sc = blockIdx.x * blockDim.x + threadIdx.x;
    if (sc < numScenes) out[sc] = (int32_t)__float_as_uint(prm[off[sc + 1] - 1].w);
}

static inline int blocks(int n, int b) { return (n + b - 1) / b; }

// ---- launchers -------------------------------------------------------------
SkinArgs skin_args(const yb_sim* s) {
    SkinArgs a{nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    if (s->skinMargin > 0.f)
        a = SkinArgs{s->b.skinState, s->b.sbox,      s->b.cand,
                     s->b.candN,     s->b.skinStats, s->b.skinPolicy};
    return a;
}

void launch_init(yb_sim* s, const CtlReset& cr) {
    const int scenes = s->skinMargin > 0.f ? s->numScenes : 0;
    hipLaunchKernelGGL(k_init, dim3(blocks(s->n, 256)), dim3(256), 0, s->stream, s->b.params,
                       s->b.pos, s->b.vel, s->b.lastVel, s->b.yv, s->b.dx[0], s->b.skinState,
                       s->b.skinPolicy, scenes, cr);
}

void launch_aabb(yb_sim* s, bool force) {
    hipLaunchKernelGGL(k_aabb, dim3(blocks(s->n, 256)), dim3(256), 0, s->stream, s->b.params,
                       s->b.pos, s->b.prm, s->b.dx[0], s->b.aabb, skin_args(s),
                       s->numScenes > 1 ? s->b.sceneOf : nullptr, force ? 1 : 0);
}

void launch_rebuild(yb_sim* s) {
    const int n = s->n;
    const bool batched = s->numScenes > 1;
    const float4* boxes = s->b.sceneBox;
    if (batched) {
        hipLaunchKernelGGL(k_box_scenes, dim3(s->numScenes), dim3(kBoxThreads), 0, s->stream,
                           s->b.sceneOff, s->b.aabb, s->b.sceneBoxes);
        boxes = s->b.sceneBoxes;
    } else {
        hipLaunchKernelGGL(k_box_partial, dim3(s->boxBlocks), dim3(kBoxThreads), 0, s->stream, n,
                           s->b.aabb, s->b.blockBox);
        hipLaunchKernelGGL(k_box_final, dim3(1), dim3(kBoxThreads), 0, s->stream, s->boxBlocks,
                           s->b.blockBox, s->b.sceneBox);
    }
    hipLaunchKernelGGL(k_morton, dim3(blocks(n, 256)), dim3(256), 0, s->stream, n, s->b.prm,
                       s->b.aabb, boxes, batched ? s->b.sceneOf : nullptr, s->b.morton,
                       s->b.keys[0], s->b.vals[0]);
    const int tiles = s->sortTiles;
    // four 8-bit passes over the Morton code, then (batched) two over the scene
    // index: an even count, so the sorted pairs end in keys[0] / vals[0] == order
    const int passes = batched ? 6 : 4;
    for (int pass = 0; pass < passes; pass++) {
        const int in = pass & 1;
        const bool scenePass = pass >= 4;
        const int shift = (scenePass ? pass - 4 : pass) * kSortBits;
        const int32_t* sof = scenePass ? s->b.sceneOf : nullptr;
        hipLaunchKernelGGL(k_sort_hist, dim3(tiles), dim3(kSortThreads), 0, s->stream, n, shift,
                           tiles, s->b.keys[in], s->b.vals[in], sof, s->b.hist);
        const int total = tiles * kSortBins;
        const int parts = blocks(total, kScanBlock);
        hipLaunchKernelGGL(k_scan_reduce, dim3(parts), dim3(kScanThreads), 0, s->stream, total,
                           s->b.hist, s->b.scanPart);
        hipLaunchKernelGGL(k_scan_parts, dim3(1), dim3(1024), 0, s->stream, parts, s->b.scanPart);
        hipLaunchKernelGGL(k_scan_down, dim3(parts), dim3(kScanThreads), 0, s->stream, total,
                           s->b.hist, s->b.scanPart);
        hipLaunchKernelGGL(k_sort_scatter, dim3(tiles), dim3(kSortThreads), 0, s->stream, n, shift,
                           tiles, s->b.keys[in], s->b.vals[in], s->b.keys[in ^ 1],
                           s->b.vals[in ^ 1], sof, s->b.hist);
    }
    launch_fill_i32(s, s->b.stradCount, 0, 1);
    launch_fill_i32(s, s->b.leafNeed, 0, n);
    SceneKeys sk{batched ? s->b.sceneOf : nullptr, s->b.sceneOff, s->b.sceneRoot};
    hipLaunchKernelGGL(k_karras, dim3(blocks(n - 1, 256)), dim3(256), 0, s->stream, n,
                       s->b.keys[0], s->b.order, s->b.nodes, s->b.parent, s->b.rng, s->b.kids,
                       s->b.strad, s->b.stradCount, s->b.leafNeed, sk, s->refitShift);
}

void launch_scene_box(yb_sim* s) {
    hipLaunchKernelGGL(k_box_partial, dim3(s->boxBlocks), dim3(kBoxThreads), 0, s->stream, s->n,
                       s->b.aabb, s->b.blockBox);
    hipLaunchKernelGGL(k_box_final, dim3(1), dim3(kBoxThreads), 0, s->stream, s->boxBlocks,
                       s->b.blockBox, s->b.sceneBox);
}

void launch_scene_end_flags(yb_sim* s) {
    hipLaunchKernelGGL(k_scene_end_flags, dim3(blocks(s->numScenes, 256)), dim3(256), 0, s->stream,
                       s->numScenes, s->b.sceneOff, s->b.prm, s->b.sceneEndFlags);
}

template <int kLeaves>
static void launch_refit_local(yb_sim* s) {
    const SkinArgs sk = skin_args(s);
    // with the skin the tree holds grown boxes: its root box goes to a spare slot
    hipLaunchKernelGGL(k_refit_local<kLeaves>, dim3(s->refitBlocks), dim3(kLeaves), 0, s->stream,
                       s->n, s->b.order, s->b.aabb, s->b.nodes, s->b.parent, s->b.rng, s->b.kids,
                       s->b.leafNeed, s->b.sufBox, s->b.preBox, s->b.blkTab,
                       sk.state ? s->b.skinRoot : s->b.sceneBox, s->b.params, s->b.dx[0], sk,
                       s->numScenes > 1 ? s->b.sceneOf : nullptr);
}

void launch_refit(yb_sim* s) {
    const int n = s->n;
    if (s->refitLeaves == 256)
        launch_refit_local<256>(s);
    else if (s->refitLeaves == 512)
        launch_refit_local<512>(s);
    else
        launch_refit_local<1024>(s);
    if (s->refitBlocks > 1) {
        const int nb = s->refitBlocks;
        const int32_t* st = skin_args(s).state;
        hipLaunchKernelGGL(k_refit_table, dim3(blocks(nb, 1024)), dim3(1024), 0, s->stream, nb,
                           s->refitLevels, s->b.blkTab, st);
        for (int L = kTabLocal + 1; L < s->refitLevels; L++)
            hipLaunchKernelGGL(k_refit_table_level, dim3(blocks(nb, 256)), dim3(256), 0, s->stream,
                               nb, L, s->b.blkTab, st);
        hipLaunchKernelGGL(k_refit_top, dim3(s->refitTopBlocks), dim3(256), 0, s->stream, n,
                           s->refitBlocks, s->refitShift, s->b.strad, s->b.stradCount, s->b.rng, s->b.parent,
                           s->b.sufBox, s->b.preBox, s->b.blkTab, s->b.nodes,
                           st ? s->b.skinRoot : s->b.sceneBox, st);
    }
}

// The default shape, 8 stack entries in LDS and 6 buffered candidates: 80 B of
// LDS per thread, which lets a CU hold the full 8 waves per SIMD.  On the
// 1M-vertex knit and
// the 512-scene batch this ran the query 1.8x and 2.9x faster than a 64-entry
// LDS stack with 16 candidates (384 B per thread, 1.5 waves per SIMD): the
// traversal waits little on memory, but a SIMD needs several waves to
// interleave its VALU, scalar, branch and LDS instructions.
// Phase 0 (every walk) with the tuned LDS shape; with the contact skin,
// phase 1 (the tests of kept lists) follows, with the shape whose LDS holds
// its 256-entry work list.
template <int T, int SL, int C, int Ph>
static void launch_query_phase(yb_sim* s) {
    hipLaunchKernelGGL((k_query<T, SL, C, Ph>), dim3(blocks(s->n, T)), dim3(T), 0, s->stream,
                       s->b.params, s->b.order, s->b.aabb, s->b.nodes, s->b.pos, s->b.prm,
                       s->b.numCols, s->b.cols, s->b.err,
                       s->numScenes > 1 ? s->b.sceneOf : nullptr, s->b.sceneRoot, s->queryStack,
                       s->countCandidates ? s->b.candCount : nullptr, s->b.dx[0], s->b.minSep,
                       skin_args(s));
}

template <int T, int SL, int C>
static void launch_query_t(yb_sim* s) {
    launch_query_phase<T, SL, C, 0>(s);
    if (skin_args(s).state) launch_query_phase<T, 8, 6, 1>(s);
}

// Query shapes (LDS stack entries, buffered candidates) the autotuner picks
// from: 48, 80, 96 and 128 B of LDS per thread.
bool query_shape_supported(int stackLds, int candidates) {
    return (stackLds == 4 && candidates == 4) || (stackLds == 8 && candidates == 6) ||
           (stackLds == 8 && candidates == 8) || (stackLds == 16 && candidates == 8);
}

template <int SL, int C>
static void launch_query_shape(yb_sim* s) {
    if (s->queryThreads == 64)
        launch_query_t<64, SL, C>(s);
    else if (s->queryThreads == 256)
        launch_query_t<256, SL, C>(s);
    else
        launch_query_t<128, SL, C>(s);
}

void launch_query(yb_sim* s) {
    const int sl = s->queryStackLds, c = s->queryCandidates;
    if (sl == 4 && c == 4)
        launch_query_shape<4, 4>(s);
    else if (sl == 8 && c == 8)
        launch_query_shape<8, 8>(s);
    else if (sl == 16 && c == 8)
        launch_query_shape<16, 8>(s);
    else
        launch_query_shape<8, 6>(s);
}

}  // namespace yb
