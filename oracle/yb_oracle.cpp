// yb_oracle.cpp — CPU oracle of the YarnBall stepper.  TEST INFRASTRUCTURE
// ONLY: used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline;
// never linked into the product.  See yb_oracle.h for the reference map.
//
// Build: make -C oracle   (g++ -O2 -fopenmp -ffp-contract=off)
#include "yb_oracle.h"

#include <math.h>
#include <omp.h>
#include <string.h>

#include <algorithm>
#include <cfloat>
#include <cstdint>
#include <vector>

// Arithmetic type of the stepper: double for the oracle.  A second build with
// -DORC_FP32 (liboracle32.so) is an FP32 CPU timing baseline only; the test
// helpers at the end of this file exist in the double build alone.
#ifdef ORC_FP32
using real = float;
#else
using real = double;
#endif

namespace {

// ---------------------------------------------------------------------------
// Generic 3-vector / rotor arithmetic, instantiated for double (the FP64
// stepper) and float (the bit-exact twin of the discrete stages).  Operation
// order is written out explicitly; the build disables FMA contraction.
template <class T>
struct V3 {
    T x, y, z;
};
template <class T> inline V3<T> add(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> sub(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> mul(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> inline V3<T> mulr(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline T dot(V3<T> a, V3<T> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <class T> inline T len2(V3<T> a) { return dot(a, a); }
template <class T> inline T len(V3<T> a) { return std::sqrt(len2(a)); }
template <class T> inline V3<T> cross(V3<T> a, V3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// glm-style mix: a*(1-t) + b*t
template <class T> inline V3<T> mix(V3<T> a, V3<T> b, T t) {
    T s = T(1) - t;
    return {a.x * s + b.x * t, a.y * s + b.y * t, a.z * s + b.z * t};
}
template <class T> inline T clamp01(T x) { return std::fmin(std::fmax(x, T(0)), T(1)); }
template <class T> inline V3<T> vmin(V3<T> a, V3<T> b) {
    return {std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z)};
}
template <class T> inline V3<T> vmax(V3<T> a, V3<T> b) {
    return {std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)};
}

// segmentClosestPoints, Common.h:318-334.
template <class T>
inline void closest(V3<T> a0, V3<T> a1, V3<T> b0, V3<T> b1, T& u, T& v) {
    V3<T> A0 = sub(a1, a0);
    V3<T> A1 = sub(b1, b0);
    T a = len2(A0);
    T b = dot(A0, A1);
    T c = len2(A1);
    V3<T> diff = sub(b0, a0);
    T vy = b * dot(A0, diff) - a * dot(A1, diff);
    vy = vy / (a * c - b * b);
    if (!std::isfinite(vy)) vy = T(0.5);
    vy = clamp01(vy);
    T ux = clamp01(dot(sub(mix(b0, b1, vy), a0), A0) / a);
    vy = clamp01(dot(sub(mix(a0, a1, ux), b0), A1) / c);
    u = ux;
    v = vy;
}

template <class T>
struct R4 {
    T x, y, z, w;
};
template <class T> inline V3<T> rq(R4<T> r) { return {r.x, r.y, r.z}; }
template <class T> inline R4<T> mkr(V3<T> q, T w) { return {q.x, q.y, q.z, w}; }
template <class T> inline R4<T> rinv(R4<T> r) { return {-r.x, -r.y, -r.z, r.w}; }
// Rotor.h:185-188 (Hamilton product with the bivector part in x,y,z)
template <class T> inline R4<T> rmul(R4<T> a, R4<T> b) {
    V3<T> q = add(add(mul(a.w, rq(b)), mul(b.w, rq(a))), cross(rq(a), rq(b)));
    return mkr(q, a.w * b.w - dot(rq(a), rq(b)));
}
// Rotor.h:96-105
template <class T> inline V3<T> rrot(R4<T> r, V3<T> v) {
    V3<T> q = rq(r);
    V3<T> a = add(mul(r.w, v), cross(q, v));
    T c = dot(v, q);
    return add(add(mul(r.w, a), cross(q, a)), mul(c, q));
}
template <class T> inline T rdot(R4<T> a, R4<T> b) {
    return ((a.x * b.x + a.y * b.y) + a.z * b.z) + a.w * b.w;
}
template <class T> inline R4<T> rscale(R4<T> a, T s) { return {a.x * s, a.y * s, a.z * s, a.w * s}; }
template <class T> inline R4<T> radd(R4<T> a, R4<T> b) {
    return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w};
}
template <class T> inline T rlen(R4<T> a) { return std::sqrt(rdot(a, a)); }

// Rotor::fromTo (Rotor.h:57-64) with orthoBasisX (Common.h:235-250).
template <class T> inline R4<T> from_to(V3<T> from, V3<T> to) {
    V3<T> h = mulr(add(from, to), T(0.5));
    T l = len2(h);
    if (l > 0) {
        h = mulr(h, T(1) / std::sqrt(l));
    } else {
        V3<T> n = from;
        if (n.z >= n.y) {
            T a = T(1) / (T(1) + n.z);
            T b = -n.x * n.y * a;
            h = {b, T(1) - n.y * n.y * a, -n.y};
        } else {
            T a = T(1) / (T(1) + n.y);
            h = {T(1) - n.x * n.x * a, -n.x, -n.x * n.z * a};
        }
    }
    return mkr(cross(from, h), dot(from, h));
}

// Packed symmetric 3x3 {xx, yy, zz, xy, xz, yz} (SymMat.h:8-107).
struct S3 {
    real a[6];
};
inline void s3_diag(S3& m, real d) { m.a[0] += d; m.a[1] += d; m.a[2] += d; }
inline void s3_outer_axpy(S3& m, V3<real> v, real s) {
    m.a[0] += v.x * v.x * s;
    m.a[1] += v.y * v.y * s;
    m.a[2] += v.z * v.z * s;
    m.a[3] += v.x * v.y * s;
    m.a[4] += v.x * v.z * s;
    m.a[5] += v.y * v.z * s;
}
// 3x3 solve by Gaussian elimination with partial pivoting (FP64 textbook route;
// the device uses the adjugate).
inline V3<real> s3_solve(const S3& m, V3<real> f) {
    real A[3][4] = {{m.a[0], m.a[3], m.a[4], f.x},
                      {m.a[3], m.a[1], m.a[5], f.y},
                      {m.a[4], m.a[5], m.a[2], f.z}};
    for (int c = 0; c < 3; c++) {
        int p = c;
        for (int r = c + 1; r < 3; r++)
            if (std::fabs(A[r][c]) > std::fabs(A[p][c])) p = r;
        if (p != c)
            for (int k = 0; k < 4; k++) std::swap(A[c][k], A[p][k]);
        for (int r = c + 1; r < 3; r++) {
            real fct = A[r][c] / A[c][c];
            for (int k = c; k < 4; k++) A[r][k] -= fct * A[c][k];
        }
    }
    real x[3];
    for (int r = 2; r >= 0; r--) {
        real s = A[r][3];
        for (int k = r + 1; k < 3; k++) s -= A[r][k] * x[k];
        x[r] = s / A[r][r];
    }
    return {x[0], x[1], x[2]};
}

using D3 = V3<real>;
using F3 = V3<float>;
using DR = R4<real>;

inline bool has(uint32_t f, int b) { return (f & (uint32_t)b) != 0; }

// ---- FP32 twin: the operations of the device's discrete stages --------------
inline uint32_t expand10(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}
inline uint32_t quant(float c, float lo, float ext) {
    float t = (c - lo) / ext;
    float q = std::fmin(std::fmax(t * 1024.0f, 0.0f), 1023.0f);
    if (!(q == q)) q = 0.0f;
    return (uint32_t)q;
}

// Pair filter + narrowphase of buildCollisionList (collision.cu:45-77) for the
// unordered pair (x, y) = (min, max).
inline bool pair_ok(const yb_params* p, const float* pos3, const int32_t* cid, const float* dx3,
                    int i, int j) {
    const int x = std::min(i, j);
    const int y = std::max(i, j);
    const int c0 = cid[x];
    const int c1 = cid[x + 1];
    if (c0 == y || c1 == y || c0 == y + 1 || c1 == y + 1) return false;
    if (y - x <= 2) return false;
    auto P = [&](int k) { return F3{pos3[3 * k], pos3[3 * k + 1], pos3[3 * k + 2]}; };
    auto DX = [&](int k) { return F3{dx3[3 * k], dx3[3 * k + 1], dx3[3 * k + 2]}; };
    const F3 a0 = P(x);
    const F3 a1 = sub(P(x + 1), a0);
    const F3 b0 = sub(P(y), a0);
    const F3 b1 = sub(P(y + 1), a0);
    float u, w;
    closest<float>(F3{0.f, 0.f, 0.f}, a1, b0, b1, u, w);
    const F3 nrm = sub(mul(u, a1), mix(b0, b1, w));
    const float d2 = len2(nrm);
    float r = 2.0f * p->scaledDetectionRadius;
    if (p->useVelocityRadius) {
        r += std::sqrt(std::fmax(len2(DX(x)), len2(DX(x + 1))));
        r += std::sqrt(std::fmax(len2(DX(y)), len2(DX(y + 1))));
    }
    return d2 < r * r;
}

}  // namespace

// ============================================================================
// FP32 twin (C API)
extern "C" void orc32_warm_start(int n, const yb_params* p, const float* invMass,
                                 const float* vel3, const float* lastVel3, float* dx3,
                                 float* y3) {
    const float h = p->h;
    const F3 g = {p->gravity[0], p->gravity[1], p->gravity[2]};
#pragma omp parallel for schedule(static)
    for (int v = 0; v < n; v++) {
        const F3 u = {vel3[3 * v], vel3[3 * v + 1], vel3[3 * v + 2]};
        const F3 lv = {lastVel3[3 * v], lastVel3[3 * v + 1], lastVel3[3 * v + 2]};
        F3 dx = mul(h, u);
        F3 y = {0.f, 0.f, 0.f};
        if (invMass[v] != 0.0f) {
            y = add(dx, mul(h * h, g));
            const float g2 = len2(g);
            if (g2 > 0.0f) {
                const float lh = p->lastH;
                const F3 a = {(u.x - lv.x) / lh, (u.y - lv.y) / lh, (u.z - lv.z) / lh};
                const float s = clamp01(dot(a, g) / g2);
                dx = add(dx, mul(h * h * s, g));
            }
        }
        dx3[3 * v] = dx.x;
        dx3[3 * v + 1] = dx.y;
        dx3[3 * v + 2] = dx.z;
        if (y3) {
            y3[3 * v] = y.x;
            y3[3 * v + 1] = y.y;
            y3[3 * v + 2] = y.z;
        }
    }
}

extern "C" void orc32_aabb(int n, const yb_params* p, const float* pos3, const uint32_t* flags,
                           const float* dx3, float* aabb6) {
#pragma omp parallel for schedule(static)
    for (int v = 0; v < n; v++) {
        F3 lo = {INFINITY, INFINITY, INFINITY};
        F3 hi = {-INFINITY, -INFINITY, -INFINITY};
        if (has(flags[v], YB_FLAG_HAS_NEXT)) {
            const F3 p0 = {pos3[3 * v], pos3[3 * v + 1], pos3[3 * v + 2]};
            const F3 p1 = {pos3[3 * v + 3], pos3[3 * v + 4], pos3[3 * v + 5]};
            lo = vmin(p0, p1);
            hi = vmax(p0, p1);
            float pad = p->scaledDetectionRadius;
            if (p->useVelocityRadius) {
                const F3 d0 = {dx3[3 * v], dx3[3 * v + 1], dx3[3 * v + 2]};
                const F3 d1 = {dx3[3 * v + 3], dx3[3 * v + 4], dx3[3 * v + 5]};
                pad += std::sqrt(std::fmax(len2(d0), len2(d1)));
            }
            lo = sub(lo, F3{pad, pad, pad});
            hi = add(hi, F3{pad, pad, pad});
        }
        float* o = aabb6 + 6 * (size_t)v;
        o[0] = lo.x; o[1] = lo.y; o[2] = lo.z;
        o[3] = hi.x; o[4] = hi.y; o[5] = hi.z;
    }
}

extern "C" void orc32_scene_box(int n, const float* aabb6, float* box6) {
    F3 lo = {INFINITY, INFINITY, INFINITY};
    F3 hi = {-INFINITY, -INFINITY, -INFINITY};
    for (int v = 0; v < n; v++) {
        const float* a = aabb6 + 6 * (size_t)v;
        lo = vmin(lo, F3{a[0], a[1], a[2]});
        hi = vmax(hi, F3{a[3], a[4], a[5]});
    }
    box6[0] = lo.x; box6[1] = lo.y; box6[2] = lo.z;
    box6[3] = hi.x; box6[4] = hi.y; box6[5] = hi.z;
}

extern "C" void orc32_morton(int n, const uint32_t* flags, const float* aabb6, const float* box6,
                             uint32_t* codes) {
    const F3 slo = {box6[0], box6[1], box6[2]};
    const F3 shi = {box6[3], box6[4], box6[5]};
    const F3 e3 = sub(shi, slo);
    // cubic cells spanning the largest extent
    const float ext = std::fmax(std::fmax(e3.x, e3.y), e3.z);
#pragma omp parallel for schedule(static)
    for (int v = 0; v < n; v++) {
        uint32_t code = 0xFFFFFFFFu;
        if (has(flags[v], YB_FLAG_HAS_NEXT)) {
            const float* a = aabb6 + 6 * (size_t)v;
            const F3 c = mulr(add(F3{a[0], a[1], a[2]}, F3{a[3], a[4], a[5]}), 0.5f);
            code = (expand10(quant(c.x, slo.x, ext)) << 2) |
                   (expand10(quant(c.y, slo.y, ext)) << 1) | expand10(quant(c.z, slo.z, ext));
        }
        codes[v] = code;
    }
}

extern "C" void orc32_sort(int n, const uint32_t* codes, int32_t* order) {
    for (int i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order, order + n, [&](int a, int b) { return codes[a] < codes[b]; });
}

// Contact lists from an exact broadphase: every segment box is binned by its lo
// corner in a uniform grid whose cell is the largest box extent, so any box that
// overlaps box i has its lo corner in the cells spanning [lo_i - cell, hi_i].
// The candidates then go through the device's overlap predicate and pair filter.
extern "C" int orc32_contacts(int n, const yb_params* p, const float* pos3,
                              const uint32_t* flags, const int32_t* cid, const float* dx3,
                              const float* aabb6, int32_t* numCols, int32_t* cols) {
    std::vector<int> segs;
    for (int v = 0; v < n; v++)
        if (has(flags[v], YB_FLAG_HAS_NEXT)) segs.push_back(v);
    for (int v = 0; v < n; v++) numCols[v] = 0;
    if (segs.empty()) return 0;
    double gmin[3] = {INFINITY, INFINITY, INFINITY};
    double cell = 0.0;
    for (int v : segs) {
        const float* a = aabb6 + 6 * (size_t)v;
        for (int k = 0; k < 3; k++) {
            gmin[k] = std::min(gmin[k], (double)a[k]);
            cell = std::max(cell, (double)a[3 + k] - (double)a[k]);
        }
    }
    if (!(cell > 0)) cell = 1.0;
    auto cellOf = [&](double x, int k) { return (int64_t)std::floor((x - gmin[k]) / cell); };
    auto key = [](int64_t cx, int64_t cy, int64_t cz) {
        return ((uint64_t)(cx & 0x1FFFFF) << 42) | ((uint64_t)(cy & 0x1FFFFF) << 21) |
               (uint64_t)(cz & 0x1FFFFF);
    };
    std::vector<std::pair<uint64_t, int>> bins(segs.size());
#pragma omp parallel for schedule(static)
    for (size_t s = 0; s < segs.size(); s++) {
        const float* a = aabb6 + 6 * (size_t)segs[s];
        bins[s] = {key(cellOf(a[0], 0), cellOf(a[1], 1), cellOf(a[2], 2)), segs[s]};
    }
    std::sort(bins.begin(), bins.end());
    int err = 0;
    std::vector<std::vector<int>> lists(n);
#pragma omp parallel for schedule(dynamic, 256) reduction(| : err)
    for (size_t s = 0; s < segs.size(); s++) {
        const int i = segs[s];
        const float* qa = aabb6 + 6 * (size_t)i;
        int64_t c0[3], c1[3];
        for (int k = 0; k < 3; k++) {
            c0[k] = cellOf((double)qa[k] - cell, k) - 1;
            c1[k] = cellOf((double)qa[3 + k], k) + 1;
        }
        std::vector<int>& L = lists[i];
        for (int64_t cx = c0[0]; cx <= c1[0]; cx++)
            for (int64_t cy = c0[1]; cy <= c1[1]; cy++)
                for (int64_t cz = c0[2]; cz <= c1[2]; cz++) {
                    if (cx < 0 || cy < 0 || cz < 0) continue;
                    const uint64_t kk = key(cx, cy, cz);
                    auto it = std::lower_bound(bins.begin(), bins.end(),
                                               std::make_pair(kk, (int)-1));
                    for (; it != bins.end() && it->first == kk; ++it) {
                        const int j = it->second;
                        if (j == i) continue;
                        const float* b = aabb6 + 6 * (size_t)j;
                        const bool ov = (qa[0] <= b[3]) && (b[0] <= qa[3]) && (qa[1] <= b[4]) &&
                                        (b[1] <= qa[4]) && (qa[2] <= b[5]) && (b[2] <= qa[5]);
                        if (!ov) continue;
                        if (pair_ok(p, pos3, cid, dx3, i, j)) L.push_back(j);
                    }
                }
        std::sort(L.begin(), L.end());
        if ((int)L.size() > YB_MAX_COLLISIONS_PER_SEGMENT) err = 1;
    }
    for (int i = 0; i < n; i++) {
        const int c = std::min((int)lists[i].size(), YB_MAX_COLLISIONS_PER_SEGMENT);
        numCols[i] = c;
        if (cols)
            for (int k = 0; k < c; k++) cols[i + (size_t)n * k] = lists[i][k];
    }
    return err ? YB_ERROR_MAX_COLLISIONS_PER_SEGMENT_EXCEEDED : 0;
}

// ============================================================================
// FP64 stepper
struct orc_sim {
    int n = 0;
    std::vector<real> pos, vel, lastVel, q, qRest, invMass, lRest, kStretch;
    std::vector<int32_t> cid;
    std::vector<uint32_t> flags;
    std::vector<int32_t> numCols, cols;  // slot-major, n * 128
    std::vector<real> y, dx, dxNext, qNext, stepLimit, dxWarm;
    int64_t stepCounter = 0;
    double lastResidual = 0.0;
    double forceNorm0 = 0.0;  // |f|^2 at the plain warm start of the last substep
    double lastForceResidual = 0.0, lastUpdateResidual = 0.0;
    int kind = 0;  // stopping residual: 0 update, 1 Newton
    double tol = 0.0;     // > 0: sweep until the relative residual <= tol
    int maxSweeps = 0;
    int lastSweeps = 0;
    int64_t totalSweeps = 0;
};

extern "C" void orc_set_solver(orc_sim* o, double tol, int maxSweeps) {
    o->tol = tol;
    o->maxSweeps = maxSweeps;
}
extern "C" int orc_last_sweeps(const orc_sim* o) { return o->lastSweeps; }

extern "C" void orc_set_threads(int n) {
    if (n > 0) omp_set_num_threads(n);
}
extern "C" int orc_get_threads(void) { return omp_get_max_threads(); }

extern "C" orc_sim* orc_new(int n) {
    orc_sim* o = new orc_sim();
    o->n = n;
    const size_t N = (size_t)n;
    o->pos.assign(3 * N, 0.0);
    o->vel.assign(3 * N, 0.0);
    o->lastVel.assign(3 * N, 0.0);
    o->q.assign(4 * N, 0.0);
    o->qRest.assign(4 * N, 0.0);
    o->invMass.assign(N, 0.0);
    o->lRest.assign(N, 0.0);
    o->kStretch.assign(N, 0.0);
    o->cid.assign(N, -1);
    o->flags.assign(N, 0);
    o->numCols.assign(N, 0);
    o->cols.assign(N * YB_MAX_COLLISIONS_PER_SEGMENT, 0);
    o->y.assign(3 * N, 0.0);
    o->dx.assign(3 * N, 0.0);
    o->dxNext.assign(3 * N, 0.0);
    o->qNext.assign(4 * N, 0.0);
    o->stepLimit.assign(N, INFINITY);
    o->dxWarm.assign(3 * N, 0.0);
    return o;
}

extern "C" void orc_free(orc_sim* o) { delete o; }

extern "C" void orc_load(orc_sim* o, const yb_vertex* verts, const float* vels,
                         const float* lastVels, const float* qs, const float* qRests) {
    for (int i = 0; i < o->n; i++) {
        const yb_vertex& v = verts[i];
        for (int k = 0; k < 3; k++) {
            o->pos[3 * i + k] = v.pos[k];
            o->vel[3 * i + k] = vels[3 * i + k];
            o->lastVel[3 * i + k] = lastVels ? lastVels[3 * i + k] : 0.0;
        }
        for (int k = 0; k < 4; k++) {
            o->q[4 * i + k] = qs[4 * i + k];
            if (qRests) o->qRest[4 * i + k] = qRests[4 * i + k];
        }
        o->invMass[i] = v.invMass;
        o->lRest[i] = v.lRest;
        o->kStretch[i] = v.kStretch;
        o->cid[i] = v.connectionIndex;
        o->flags[i] = v.flags;
    }
}

extern "C" void orc_load64(orc_sim* o, const double* pos, const double* vels,
                           const double* lastVels, const double* qs) {
    const size_t N = (size_t)o->n;
    for (size_t k = 0; k < 3 * N; k++) {
        if (pos) o->pos[k] = (real)pos[k];
        if (vels) o->vel[k] = (real)vels[k];
        if (lastVels) o->lastVel[k] = (real)lastVels[k];
    }
    for (size_t k = 0; qs && k < 4 * N; k++) o->q[k] = (real)qs[k];
}

extern "C" void orc_store64(const orc_sim* o, double* pos, double* vels, double* lastVels,
                            double* qs) {
    const size_t N = (size_t)o->n;
    for (size_t k = 0; k < 3 * N; k++) {
        if (pos) pos[k] = o->pos[k];
        if (vels) vels[k] = o->vel[k];
        if (lastVels) lastVels[k] = o->lastVel[k];
    }
    for (size_t k = 0; qs && k < 4 * N; k++) qs[k] = o->q[k];
}

extern "C" void orc_set_step_counter(orc_sim* o, int64_t c) { o->stepCounter = c; }

namespace {

inline D3 g3(const std::vector<real>& a, int i) { return {a[3 * i], a[3 * i + 1], a[3 * i + 2]}; }
inline void s3(std::vector<real>& a, int i, D3 v) {
    a[3 * i] = v.x;
    a[3 * i + 1] = v.y;
    a[3 * i + 2] = v.z;
}
inline DR g4(const std::vector<real>& a, int i) {
    return {a[4 * i], a[4 * i + 1], a[4 * i + 2], a[4 * i + 3]};
}
inline void s4(std::vector<real>& a, int i, DR v) {
    a[4 * i] = v.x;
    a[4 * i + 1] = v.y;
    a[4 * i + 2] = v.z;
    a[4 * i + 3] = v.w;
}

// quaternionLambdaItr (cosserat.cu:209-253): every frame's closed-form update
// from the frames o->q and the iterate dx (3n), into qOut (4n).
static void frame_update(const orc_sim* o, const std::vector<real>& dx, std::vector<real>& qOut) {
    const int n = o->n;
#pragma omp parallel for schedule(static)
    for (int v = 0; v < n; v++) {
        DR q0 = g4(o->q, v);
        const uint32_t fl = o->flags[v];
        if (!has(fl, YB_FLAG_FIX_ORIENTATION) && has(fl, YB_FLAG_HAS_NEXT)) {
            const D3 e = add(sub(g3(o->pos, v + 1), g3(o->pos, v)),
                             sub(g3(dx, v + 1), g3(dx, v)));
            D3 vv = {e.x / o->lRest[v], e.y / o->lRest[v], e.z / o->lRest[v]};
            vv = mulr(vv, real(-2.0) * o->kStretch[v]);
            DR b = {0, 0, 0, 0};
            if (has(fl, YB_FLAG_HAS_PREV)) {
                const DR qr = g4(o->qRest, v - 1);
                const DR qq = g4(o->q, v - 1);
                const real sg = rdot(rmul(rinv(qq), q0), qr) > 0 ? 1.0 : -1.0;
                b = radd(b, rscale(rmul(qq, qr), sg));
            }
            if (has(fl, YB_FLAG_HAS_NEXT_ORIENTATION)) {
                const DR qr = g4(o->qRest, v);
                const DR qq = g4(o->q, v + 1);
                const real sg = rdot(rmul(rinv(q0), qq), qr) > 0 ? 1.0 : -1.0;
                b = radd(b, rscale(rmul(qq, rinv(qr)), sg));
            }
            const real lambda = len(vv) + rlen(b);
            const DR t = rmul(rmul(mkr(vv, real(0.0)), b), DR{1, 0, 0, 0});
            const DR r = radd(t, rscale(b, lambda));
            q0 = rscale(r, real(1.0) / rlen(r));
        }
        s4(qOut, v, q0);
    }
}

// Contact detection on the FP32 image of the current state.
int detect(orc_sim* o, const yb_params* p, const std::vector<float>& dxw32) {
    const int n = o->n;
    std::vector<float> pos32(3 * (size_t)n), aabb(6 * (size_t)n);
    for (size_t k = 0; k < pos32.size(); k++) pos32[k] = (float)o->pos[k];
    orc32_aabb(n, p, pos32.data(), o->flags.data(), dxw32.data(), aabb.data());
    return orc32_contacts(n, p, pos32.data(), o->flags.data(), o->cid.data(), dxw32.data(),
                          aabb.data(), o->numCols.data(), o->cols.data());
}

// Per-segment Newton terms: endpoint-0 and endpoint-1 force and Hessian of the
// stretch-shear energy and of every barrier/friction contact of segment s
// (cosserat.cu:63-148).
struct SegTerms {
    D3 f0, f1;
    S3 H0, H1;
};

SegTerms seg_terms(const orc_sim* o, const yb_params* p, real damping, int s,
                   const std::vector<real>& dx) {
    SegTerms t{};
    const int n = o->n;
    const D3 x0 = g3(o->pos, s);
    const D3 p1 = g3(o->pos, s + 1);
    const D3 d0 = g3(dx, s);
    const D3 d1 = g3(dx, s + 1);
    const real invl = 1.0 / o->lRest[s];
    const D3 e = add(sub(p1, x0), sub(d1, d0));
    const D3 c = sub(mulr(e, invl), rrot(g4(o->q, s), D3{1, 0, 0}));
    const real k = o->kStretch[s] * invl;
    real d = k * invl;
    t.f0 = sub(mul(k, c), mul(damping * d, d0));
    t.f1 = sub(mul(-k, c), mul(damping * d, d1));
    d *= 1.0 + damping;
    s3_diag(t.H0, d);
    s3_diag(t.H1, d);

    const real fricK = p->kFriction;
    const real invb = 1.0 / (real)p->barrierThickness;
    const real radius2 = 2.0 * (real)p->radius;
    const real mu = p->frictionCoeff;
    const real kCol = (real)p->kCollision * invb;
    const int nc = o->numCols[s];
    for (int i = 0; i < nc; i++) {
        const int oid = o->cols[s + (size_t)n * i];
        const D3 b0 = g3(o->pos, oid), b1 = g3(o->pos, oid + 1);
        const D3 db0 = g3(dx, oid), db1 = g3(dx, oid + 1);
        real u, w;
        closest<real>(D3{0, 0, 0}, e, add(sub(b0, x0), sub(db0, d0)),
                        add(sub(b1, x0), sub(db1, d0)), u, w);
        const D3 dpos = sub(mix(x0, p1, u), mix(b0, b1, w));
        const D3 ddpos = sub(mix(d0, d1, u), mix(db0, db1, w));
        D3 nrm = add(dpos, ddpos);
        const real dist = len(nrm);
        nrm = mulr(nrm, real(1.0) / dist);
        const real w1 = u, w0 = 1.0 - u;
        real dd = (dist - radius2) * invb;
        if (dd > 1.0) continue;
        dd = std::max(dd, real(1e-3));
        const real invd = 1.0 / dd;
        const real logd = std::log(dd);
        real dH = (-3.0 + (2.0 + invd) * invd - 2.0 * logd) * kCol * invb;
        const real ff = -(1.0 - dd) * (dd - 1.0 + 2.0 * dd * logd) * invd * kCol;
        t.f0 = add(t.f0, mul(ff * w0 - damping * dH * w0 * w0 * dot(nrm, d0), nrm));
        t.f1 = add(t.f1, mul(ff * w1 - damping * dH * w1 * w1 * dot(nrm, d1), nrm));
        dH *= 1.0 + damping;
        s3_outer_axpy(t.H0, nrm, dH * w0 * w0);
        s3_outer_axpy(t.H1, nrm, dH * w1 * w1);
        const D3 uu = sub(ddpos, mul(dot(nrm, ddpos), nrm));
        const real ul = len(uu);
        if (ul > 0) {
            const real f1 = std::min(fricK, mu * ff / ul);
            t.f0 = sub(t.f0, mul(f1 * w0, uu));
            t.f1 = sub(t.f1, mul(f1 * w1, uu));
            // H -= (nn^T - I) w^2 f1
            s3_outer_axpy(t.H0, nrm, -w0 * w0 * f1);
            s3_diag(t.H0, w0 * w0 * f1);
            s3_outer_axpy(t.H1, nrm, -w1 * w1 * f1);
            s3_diag(t.H1, w1 * w1 * f1);
        }
    }
    return t;
}

// Newton terms of vertex v at iterate dx: inertia (cosserat.cu:40-44), glue
// spring (cosserat.cu:47-53), endpoint-0 terms of segment v and endpoint-1
// terms of segment v-1 (cosserat.cu:63-193), and the step limit of the two
// segments (cosserat.cu:69,190).  seg holds seg_terms of every segment.
struct VertTerms {
    D3 f;
    S3 H;
    real limit;
};

VertTerms vertex_terms(const orc_sim* o, real h, real damping, int v,
                       const std::vector<SegTerms>& seg, const std::vector<real>& dx) {
    VertTerms t{};
    const D3 d = g3(dx, v);
    const real mh2 = 1.0 / (o->invMass[v] * h * h);
    s3_diag(t.H, mh2);
    t.f = mul(mh2, sub(g3(o->y, v), d));
    const int c = o->cid[v];
    if (c >= 0) {
        const real st = 40.0;
        t.f = sub(t.f, mul(st, add(add(sub(g3(o->pos, v), g3(o->pos, c)), sub(d, g3(dx, c))),
                                   mul(damping, d))));
        s3_diag(t.H, st * (1.0 + damping));
    }
    t.limit = INFINITY;
    if (has(o->flags[v], YB_FLAG_HAS_NEXT)) {
        t.f = add(t.f, seg[v].f0);
        for (int k = 0; k < 6; k++) t.H.a[k] += seg[v].H0.a[k];
        t.limit = o->stepLimit[v];
    }
    if (has(o->flags[v], YB_FLAG_HAS_PREV)) {
        t.f = add(t.f, seg[v - 1].f1);
        for (int k = 0; k < 6; k++) t.H.a[k] += seg[v - 1].H1.a[k];
        t.limit = std::min(t.limit, o->stepLimit[v - 1]);
    }
    return t;
}

inline D3 s3_apply(const S3& m, D3 x) {
    return {m.a[0] * x.x + m.a[3] * x.y + m.a[4] * x.z, m.a[3] * x.x + m.a[1] * x.y + m.a[5] * x.z,
            m.a[4] * x.x + m.a[5] * x.y + m.a[2] * x.z};
}

// Local Newton step of vertex v (cosserat.cu:195-205): the new iterate, and
// the vertex's residual force -- f itself, or, when the step limit shortens
// the step, H (dx' - dx) / accelerationRatio: the part of f the limit does
// not absorb, which vanishes at a clamped fixed point of the sweeps.
inline D3 local_step(const VertTerms& t, D3 d, real accel, D3& resid) {
    D3 nd = add(d, mul(accel, s3_solve(t.H, t.f)));
    const real l = len(nd);
    resid = t.f;
    if (l > t.limit && l > 0) {
        nd = mulr(nd, t.limit / l);
        resid = mulr(s3_apply(t.H, sub(nd, d)), real(1.0) / accel);
    }
    return nd;
}

// Sum over free vertices of |residual force|^2 at iterate dx (contacts, y and
// the step limits of the current substep).
double force_norm2(const orc_sim* o, const yb_params* p, const std::vector<real>& dx) {
    const int n = o->n;
    const real h = p->h;
    const real damping = (real)p->damping / h;
    std::vector<SegTerms> seg(n);
#pragma omp parallel for schedule(static)
    for (int s = 0; s < n; s++)
        if (has(o->flags[s], YB_FLAG_HAS_NEXT)) seg[s] = seg_terms(o, p, damping, s, dx);
    double acc = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : acc)
    for (int v = 0; v < n; v++) {
        if (o->invMass[v] == 0.0) continue;
        const VertTerms t = vertex_terms(o, h, damping, v, seg, dx);
        D3 r;
        local_step(t, g3(dx, v), (real)p->accelerationRatio, r);
        acc += (double)len2(r);
    }
    return acc;
}

}  // namespace

// One substep (step.cpp:46-51 + the captured graph of step.cpp:14-21).
extern "C" int orc_substep(orc_sim* o, const yb_params* p, int doDetect) {
    const int n = o->n;
    const real h = p->h;
    const real damping = (real)p->damping / h;
    const D3 g = {p->gravity[0], p->gravity[1], p->gravity[2]};
    int err = 0;

    // FP32 warm start for the discrete stages, from the pre-substep state.
    std::vector<float> dxw32;
    if (doDetect) {
        std::vector<float> vel32(3 * (size_t)n), lv32(3 * (size_t)n), im32(n);
        for (size_t k = 0; k < vel32.size(); k++) {
            vel32[k] = (float)o->vel[k];
            lv32[k] = (float)o->lastVel[k];
        }
        for (int v = 0; v < n; v++) im32[v] = (float)o->invMass[v];
        dxw32.resize(3 * (size_t)n);
        orc32_warm_start(n, p, im32.data(), vel32.data(), lv32.data(), dxw32.data(), nullptr);
    }

    // initItr (iteration.cu:7-41), FP64.
#pragma omp parallel for schedule(static)
    for (int v = 0; v < n; v++) {
        const D3 u = g3(o->vel, v);
        const D3 lv = g3(o->lastVel, v);
        D3 dx = mul(h, u);
        s3(o->lastVel, v, u);
        if (o->invMass[v] != 0.0) {
            s3(o->y, v, add(dx, mul(h * h, g)));
            const real g2 = len2(g);
            if (g2 > 0) {
                const real lh = p->lastH;
                const D3 a = {(u.x - lv.x) / lh, (u.y - lv.y) / lh, (u.z - lv.z) / lh};
                const real s = clamp01(dot(a, g) / g2);
                dx = add(dx, mul(h * h * s, g));
            }
        } else {
            s3(o->y, v, D3{0, 0, 0});
        }
        s3(o->dx, v, dx);
    }
    o->dxWarm = o->dx;

    // detectCollisions (collision.cu:94-119) on the FP32 image.
    if (doDetect) err = detect(o, p, dxw32);

    // recomputeStepLimit (collision.cu:121-159), FP64.
    const real detR = (real)p->radius + 0.5 * (real)p->barrierThickness;
#pragma omp parallel for schedule(static)
    for (int v = 0; v < n; v++) {
        real minDist = INFINITY;
        if (p->useStepSizeLimit && has(o->flags[v], YB_FLAG_HAS_NEXT)) {
            const D3 a0 = g3(o->pos, v);
            const D3 a1 = sub(g3(o->pos, v + 1), a0);
            minDist = detR * ((real)p->detectionScaler - 1.0);
            if (p->useVelocityRadius) minDist += len(g3(o->dx, v));
            for (int i = 0; i < o->numCols[v]; i++) {
                const int oid = o->cols[v + (size_t)n * i];
                const D3 b0 = sub(g3(o->pos, oid), a0);
                const D3 b1 = sub(g3(o->pos, oid + 1), a0);
                real u, w;
                closest<real>(D3{0, 0, 0}, a1, b0, b1, u, w);
                const D3 nrm = sub(mul(u, a1), mix(b0, b1, w));
                minDist = std::min(minDist, real(0.45) * len(nrm));
            }
        }
        o->stepLimit[v] = minDist;
    }

    // numItr strict-Jacobi sweeps: dx then frames (cosserat.cu:20-253).
    // Sweep k measures two residuals:
    //   update  r_U = sqrt( sum_free |dx_k+1 - dx_k|^2 / sum_free l_i^2 ),
    //   Newton  r_F = |f(dx_k)| / |f(dx_0)| (2-norms over free vertices, dx_0
    //           the plain warm start, f the residual force of local_step);
    // with a tolerance the substep stops after the first sweep whose stopping
    // residual (o->kind) is <= tol; that sweep's output is kept.
    std::vector<SegTerms> seg(n);
    double numU = 0.0, denU = 0.0, numF = 0.0;
    const bool adaptive = o->tol > 0;
    const int itMax = adaptive ? o->maxSweeps : p->numItr;
    const real accel = (real)p->accelerationRatio;
    int used = 0;
    o->lastResidual = 0.0;
    o->lastForceResidual = 0.0;
    for (int it = 0; it < itMax; it++) {
        used = it + 1;
#pragma omp parallel for schedule(static)
        for (int s = 0; s < n; s++)
            if (has(o->flags[s], YB_FLAG_HAS_NEXT)) seg[s] = seg_terms(o, p, damping, s, o->dx);
        numU = denU = numF = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : numU, denU, numF)
        for (int v = 0; v < n; v++) {
            const D3 d = g3(o->dx, v);
            if (o->invMass[v] == 0.0) {
                s3(o->dxNext, v, d);
                continue;
            }
            const VertTerms t = vertex_terms(o, h, damping, v, seg, o->dx);
            D3 r;
            const D3 nd = local_step(t, d, accel, r);
            s3(o->dxNext, v, nd);
            numU += (double)len2(sub(nd, d));
            numF += (double)len2(r);
            real li = o->lRest[v];
            if (!has(o->flags[v], YB_FLAG_HAS_NEXT) && has(o->flags[v], YB_FLAG_HAS_PREV))
                li = o->lRest[v - 1];
            denU += (double)(li * li);
        }
        if (it == 0) o->forceNorm0 = numF;
        std::swap(o->dx, o->dxNext);

        // quaternionLambdaItr with the new dx and the previous frames.
        frame_update(o, o->dx, o->qNext);
        std::swap(o->q, o->qNext);
        o->lastForceResidual = o->forceNorm0 > 0 ? std::sqrt(numF / o->forceNorm0) : 0.0;
        o->lastUpdateResidual = denU > 0 ? std::sqrt(numU / denU) : 0.0;
        o->lastResidual = o->kind ? o->lastForceResidual : o->lastUpdateResidual;
        if (adaptive && o->lastResidual <= o->tol) break;
    }
    o->lastSweeps = used;
    o->totalSweeps += used;

    // endItr (iteration.cu:48-61).
    const real s = 1.0 - (real)p->drag * h;
#pragma omp parallel for schedule(static)
    for (int v = 0; v < n; v++) {
        const D3 d = g3(o->dx, v);
        if (o->invMass[v] != 0.0) s3(o->vel, v, mulr(mulr(d, real(1.0) / h), s));
        s3(o->pos, v, add(g3(o->pos, v), d));
    }
    return err;
}

extern "C" int orc_advance(orc_sim* o, yb_params* p, float h, float maxH) {
    if (!(h > 0)) return 0;
    int steps = (int)ceilf(h / maxH);
    if (steps < 1) steps = 1;
    p->lastH = p->h;
    p->h = h / (float)steps;
    p->detectionRadius = p->radius + 0.5f * p->barrierThickness;
    p->scaledDetectionRadius = p->detectionRadius * p->detectionScaler;
    int err = 0;
    for (int st = 0; st < steps; st++, o->stepCounter++) {
        const int det = p->detectionPeriod > 0 && o->stepCounter % p->detectionPeriod == 0;
        err |= orc_substep(o, p, det);
    }
    p->time += h;
    return err;
}

extern "C" void orc_contacts(const orc_sim* o, int32_t* numCols, int32_t* cols) {
    const size_t N = (size_t)o->n;
    if (numCols) memcpy(numCols, o->numCols.data(), sizeof(int32_t) * N);
    if (cols) memcpy(cols, o->cols.data(), sizeof(int32_t) * N * YB_MAX_COLLISIONS_PER_SEGMENT);
}

extern "C" void orc_last_prologue(const orc_sim* o, double* stepLimit, double* dxWarm) {
    const size_t N = (size_t)o->n;
    for (size_t k = 0; stepLimit && k < N; k++) stepLimit[k] = o->stepLimit[k];
    for (size_t k = 0; dxWarm && k < 3 * N; k++) dxWarm[k] = o->dxWarm[k];
}

extern "C" double orc_last_residual(const orc_sim* o) { return o->lastResidual; }
extern "C" double orc_last_force_norm0(const orc_sim* o) { return o->forceNorm0; }
extern "C" double orc_last_newton_residual(const orc_sim* o) { return o->lastForceResidual; }
extern "C" double orc_last_update_residual(const orc_sim* o) { return o->lastUpdateResidual; }
extern "C" void orc_set_residual_kind(orc_sim* o, int kind) { o->kind = kind ? 1 : 0; }

// ---- access to the sweep's own functions, for the closed-form tests ----------
extern "C" void orc_set_contacts(orc_sim* o, const int32_t* numCols, const int32_t* cols) {
    const size_t N = (size_t)o->n;
    for (size_t i = 0; i < N; i++) {
        o->numCols[i] = std::min(std::max(numCols[i], 0), YB_MAX_COLLISIONS_PER_SEGMENT);
        for (int k = 0; k < o->numCols[i]; k++) o->cols[i + N * k] = cols[i + N * k];
    }
}

static std::vector<real> vec_of(const double* a, size_t count) {
    std::vector<real> v(count);
    for (size_t k = 0; k < count; k++) v[k] = (real)a[k];
    return v;
}

extern "C" void orc_set_iterate(orc_sim* o, const double* dx, const double* y,
                                const double* stepLimit) {
    const size_t N = (size_t)o->n;
    if (dx) o->dx = vec_of(dx, 3 * N);
    if (y) o->y = vec_of(y, 3 * N);
    if (stepLimit) o->stepLimit = vec_of(stepLimit, N);
}

extern "C" void orc_frame_update(const orc_sim* o, const double* dx, double* qOut) {
    const std::vector<real> d = vec_of(dx, 3 * (size_t)o->n);
    std::vector<real> q(4 * (size_t)o->n);
    frame_update(o, d, q);
    for (size_t i = 0; i < q.size(); i++) qOut[i] = (double)q[i];
}

// out: f0[3], f1[3], H0[6], H1[6] of segment s at iterate dx (3n).
extern "C" void orc_seg_terms(const orc_sim* o, const yb_params* p, int s, const double* dx,
                              double* out) {
    const std::vector<real> d = vec_of(dx, 3 * (size_t)o->n);
    const SegTerms t = seg_terms(o, p, (real)p->damping / (real)p->h, s, d);
    const D3 v[2] = {t.f0, t.f1};
    for (int k = 0; k < 2; k++) {
        out[3 * k] = v[k].x;
        out[3 * k + 1] = v[k].y;
        out[3 * k + 2] = v[k].z;
    }
    for (int k = 0; k < 6; k++) {
        out[6 + k] = t.H0.a[k];
        out[12 + k] = t.H1.a[k];
    }
}

// out: f[3], H[6], step limit, new iterate[3], residual force[3] of vertex v
// (free) at iterate dx (3n), with the contacts, y and step limits set.
extern "C" void orc_vertex_terms(const orc_sim* o, const yb_params* p, int v, const double* dx,
                                 double* out) {
    const int n = o->n;
    const std::vector<real> d = vec_of(dx, 3 * (size_t)n);
    const real damping = (real)p->damping / (real)p->h;
    std::vector<SegTerms> seg(n);
    for (int s = std::max(v - 1, 0); s <= v; s++)
        if (has(o->flags[s], YB_FLAG_HAS_NEXT)) seg[s] = seg_terms(o, p, damping, s, d);
    const VertTerms t = vertex_terms(o, (real)p->h, damping, v, seg, d);
    D3 r;
    const D3 nd = local_step(t, g3(d, v), (real)p->accelerationRatio, r);
    const D3 w[4] = {t.f, nd, r, nd};
    out[0] = t.f.x;
    out[1] = t.f.y;
    out[2] = t.f.z;
    for (int k = 0; k < 6; k++) out[3 + k] = t.H.a[k];
    out[9] = t.limit;
    for (int k = 1; k < 3; k++) {
        out[10 + 3 * (k - 1)] = w[k].x;
        out[11 + 3 * (k - 1)] = w[k].y;
        out[12 + 3 * (k - 1)] = w[k].z;
    }
}

// Sum over free vertices of |residual force|^2 at iterate dx (3n).
extern "C" double orc_force_norm2(const orc_sim* o, const yb_params* p, const double* dx) {
    return force_norm2(o, p, vec_of(dx, 3 * (size_t)o->n));
}

#ifndef ORC_FP32
// ============================================================================
// FP64 pieces for closed-form tests.
extern "C" void orc_closest(const double* a0, const double* a1, const double* b0,
                            const double* b1, double* uv) {
    double u, v;
    closest<double>(D3{a0[0], a0[1], a0[2]}, D3{a1[0], a1[1], a1[2]}, D3{b0[0], b0[1], b0[2]},
                    D3{b1[0], b1[1], b1[2]}, u, v);
    uv[0] = u;
    uv[1] = v;
}

// out = {b(d), ff/kCol, dH*b/kCol} with the sweep's clamp-free formulas:
// ff = -b'(d) * kCol, dH = b''(d) * kCol / b.
extern "C" void orc_barrier(double d, double* out3) {
    const double invd = 1.0 / d, logd = std::log(d);
    out3[0] = -(d - 1.0) * (d - 1.0) * logd;
    out3[1] = -(1.0 - d) * (d - 1.0 + 2.0 * d * logd) * invd;
    out3[2] = (-3.0 + (2.0 + invd) * invd - 2.0 * logd);
}

extern "C" void orc_solve3(const double* H6, const double* f, double* x) {
    S3 m;
    for (int k = 0; k < 6; k++) m.a[k] = H6[k];
    D3 r = s3_solve(m, D3{f[0], f[1], f[2]});
    x[0] = r.x;
    x[1] = r.y;
    x[2] = r.z;
}

extern "C" void orc_rotor_mul(const double* a, const double* b, double* out) {
    DR r = rmul(DR{a[0], a[1], a[2], a[3]}, DR{b[0], b[1], b[2], b[3]});
    out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}

extern "C" void orc_rotor_rotate(const double* r, const double* v, double* out) {
    D3 o = rrot(DR{r[0], r[1], r[2], r[3]}, D3{v[0], v[1], v[2]});
    out[0] = o.x; out[1] = o.y; out[2] = o.z;
}

extern "C" void orc_from_to(const double* from, const double* to, double* out) {
    DR r = from_to(D3{from[0], from[1], from[2]}, D3{to[0], to[1], to[2]});
    out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}

// configure (YarnBall.cu:91-157) in FP64: flags completed in verts, frames qs,
// rest Darboux qRests (magnitude kept), lRest and invMass (scaled by 1/mass).
extern "C" int orc_configure(int n, yb_vertex* verts, double* qs, double* qRests,
                             double density, double* lRest, double* invMass) {
    DR lastQ = {0, 0, 0, 1};
    verts[n - 1].flags &= ~(uint32_t)YB_FLAG_HAS_NEXT;
    for (int i = 0; i < n; i++) {
        yb_vertex& v = verts[i];
        if (i < n - 1) {
            const bool hn = has(v.flags, YB_FLAG_HAS_NEXT);
            uint32_t& nf = verts[i + 1].flags;
            nf = hn ? (nf | YB_FLAG_HAS_PREV) : (nf & ~(uint32_t)YB_FLAG_HAS_PREV);
            const bool hnn = hn && has(nf, YB_FLAG_HAS_NEXT);
            v.flags = hnn ? (v.flags | YB_FLAG_HAS_NEXT_ORIENTATION)
                          : (v.flags & ~(uint32_t)YB_FLAG_HAS_NEXT_ORIENTATION);
            if (!hn) v.flags |= YB_FLAG_FIX_ORIENTATION;
        }
        if (i == 0) v.flags &= ~(uint32_t)YB_FLAG_HAS_PREV;
        if (!has(v.flags, YB_FLAG_HAS_PREV) && !has(v.flags, YB_FLAG_HAS_NEXT)) return -4;
        lRest[i] = 1.0 / n;
        double mass = 0;
        if (has(v.flags, YB_FLAG_HAS_PREV)) mass += lRest[i - 1];
        if (has(v.flags, YB_FLAG_HAS_NEXT)) {
            const D3 seg = sub(D3{verts[i + 1].pos[0], verts[i + 1].pos[1], verts[i + 1].pos[2]},
                               D3{v.pos[0], v.pos[1], v.pos[2]});
            lRest[i] = len(seg);
            if (lRest[i] == 0 || !std::isfinite(lRest[i])) return -4;
            const DR qq = from_to(D3{1, 0, 0}, mulr(seg, 1.0 / lRest[i]));
            DR t = rmul(rinv(qq), lastQ);
            t = DR{t.x, 0, 0, t.w};
            t = rscale(t, 1.0 / rlen(t));
            lastQ = rmul(qq, t);
            mass += lRest[i];
        }
        qs[4 * i] = lastQ.x;
        qs[4 * i + 1] = lastQ.y;
        qs[4 * i + 2] = lastQ.z;
        qs[4 * i + 3] = lastQ.w;
        mass *= 0.5 * density;
        invMass[i] = mass != 0 ? (double)v.invMass / mass : 0.0;
    }
    for (int i = 0; i < n - 1; i++) {
        const DR qr = {qRests[4 * i], qRests[4 * i + 1], qRests[4 * i + 2], qRests[4 * i + 3]};
        const double m = rlen(qr);
        const DR a = {qs[4 * i], qs[4 * i + 1], qs[4 * i + 2], qs[4 * i + 3]};
        const DR b = {qs[4 * i + 4], qs[4 * i + 5], qs[4 * i + 6], qs[4 * i + 7]};
        const DR d = rscale(rmul(rinv(a), b), m);
        qRests[4 * i] = d.x;
        qRests[4 * i + 1] = d.y;
        qRests[4 * i + 2] = d.z;
        qRests[4 * i + 3] = d.w;
    }
    return 0;
}
#endif  // ORC_FP32
