// ybmath.h — small FP32 vector / rotor / symmetric-3x3 toolkit used by the
// stepper kernels and by the host-side set-up code.
//
// Restates the parts of the reference's device math that the stepper uses:
//   segmentClosestPoints     KittenEngine/includes/modules/Common.h:318-334
//   orthoBasisX              Common.h:235-250
//   Rotor (x,y,z | w)        KittenEngine/includes/modules/Rotor.h:57-105,181-188
//   hess3 (packed sym 3x3)   KittenEngine/includes/modules/SymMat.h:8-107
// Everything is plain IEEE float arithmetic in a fixed evaluation order.  The
// functions used by the discrete broadphase stages are written so that a CPU
// build with -ffp-contract=off produces bit-identical results.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define YB_HD __host__ __device__ __forceinline__
#define YB_UNROLL _Pragma("unroll")
#else
#define YB_HD inline
#define YB_UNROLL
#endif

namespace yb {

struct v3 {
    float x, y, z;
};

YB_HD v3 mk(float x, float y, float z) { return v3{x, y, z}; }
YB_HD v3 operator+(v3 a, v3 b) { return v3{a.x + b.x, a.y + b.y, a.z + b.z}; }
YB_HD v3 operator-(v3 a, v3 b) { return v3{a.x - b.x, a.y - b.y, a.z - b.z}; }
YB_HD v3 operator-(v3 a) { return v3{-a.x, -a.y, -a.z}; }
YB_HD v3 operator*(float s, v3 a) { return v3{s * a.x, s * a.y, s * a.z}; }
YB_HD v3 operator*(v3 a, float s) { return v3{a.x * s, a.y * s, a.z * s}; }
YB_HD v3& operator+=(v3& a, v3 b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
YB_HD v3& operator-=(v3& a, v3 b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; return a; }
YB_HD v3& operator*=(v3& a, float s) { a.x *= s; a.y *= s; a.z *= s; return a; }
YB_HD float dot(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
YB_HD float len2(v3 a) { return dot(a, a); }
YB_HD float len(v3 a) { return sqrtf(len2(a)); }
YB_HD v3 cross(v3 a, v3 b) {
    return v3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// a*(1-t) + b*t, the interpolation used by the reference's mix() calls.
YB_HD v3 mixv(v3 a, v3 b, float t) {
    float s = 1.0f - t;
    return v3{a.x * s + b.x * t, a.y * s + b.y * t, a.z * s + b.z * t};
}
YB_HD float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
YB_HD v3 vmin(v3 a, v3 b) { return v3{fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
YB_HD v3 vmax(v3 a, v3 b) { return v3{fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }

// Closest points between segments mix(a0,a1,u) and mix(b0,b1,v): the unrestricted
// line solution for v, clamped, then u re-projected and v re-projected
// (Common.h:318-334).  Non-finite v (parallel segments) falls back to v = 0.5.
YB_HD void closest_uv(v3 a0, v3 a1, v3 b0, v3 b1, float& u, float& v) {
    v3 A0 = a1 - a0;
    v3 A1 = b1 - b0;
    float a = len2(A0);
    float b = dot(A0, A1);
    float c = len2(A1);
    v3 diff = b0 - a0;
    float vy = b * dot(A0, diff) - a * dot(A1, diff);
    vy = vy / (a * c - b * b);
    if (!isfinite(vy)) vy = 0.5f;
    vy = clampf(vy, 0.f, 1.f);
    float ux = clampf(dot(mixv(b0, b1, vy) - a0, A0) / a, 0.f, 1.f);
    vy = clampf(dot(mixv(a0, a1, ux) - b0, A1) / c, 0.f, 1.f);
    u = ux;
    v = vy;
}

// ---- rotors: v = (x, y, z, w), bivector part {yz, zx, xy}, scalar w ----------
struct rot {
    float x, y, z, w;
};
YB_HD v3 rq(rot r) { return v3{r.x, r.y, r.z}; }
YB_HD rot mkrot(v3 q, float w) { return rot{q.x, q.y, q.z, w}; }
YB_HD rot rinv(rot r) { return rot{-r.x, -r.y, -r.z, r.w}; }
// Rotor.h:185-188
YB_HD rot rmul(rot a, rot b) {
    v3 q = (a.w * rq(b) + b.w * rq(a)) + cross(rq(a), rq(b));
    float w = a.w * b.w - dot(rq(a), rq(b));
    return mkrot(q, w);
}
// Rotor.h:96-105
YB_HD v3 rrotate(rot r, v3 v) {
    v3 q = rq(r);
    v3 a = r.w * v + cross(q, v);
    float c = dot(v, q);
    return (r.w * a + cross(q, a)) + c * q;
}
// rotate(r, e1) with e1 = (1,0,0), expanded.
YB_HD v3 rrotate_e1(rot r) { return rrotate(r, v3{1.f, 0.f, 0.f}); }
YB_HD float rdot(rot a, rot b) { return ((a.x * b.x + a.y * b.y) + a.z * b.z) + a.w * b.w; }
YB_HD float rlen(rot a) { return sqrtf(rdot(a, a)); }
YB_HD rot rscale(rot a, float s) { return rot{a.x * s, a.y * s, a.z * s, a.w * s}; }
YB_HD rot radd(rot a, rot b) { return rot{a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
YB_HD rot rnormalize(rot a) { return rscale(a, 1.0f / rlen(a)); }

// orthoBasisX(n)[1]  (Common.h:235-250)
YB_HD v3 ortho_basis_x1(v3 n) {
    if (n.z >= n.y) {
        float a = 1.0f / (1.0f + n.z);
        float b = -n.x * n.y * a;
        return v3{b, 1.0f - n.y * n.y * a, -n.y};
    } else {
        float a = 1.0f / (1.0f + n.y);
        return v3{1.0f - n.x * n.x * a, -n.x, -n.x * n.z * a};
    }
}
// Rotor::fromTo (Rotor.h:57-64)
YB_HD rot rfrom_to(v3 from, v3 to) {
    v3 h = (from + to) * 0.5f;
    float l = len2(h);
    if (l > 0)
        h *= 1.0f / sqrtf(l);
    else
        h = ortho_basis_x1(from);
    return mkrot(cross(from, h), dot(from, h));
}

// ---- packed symmetric 3x3: {xx, yy, zz, xy, xz, yz} (SymMat.h:8-107) ---------
struct sym3 {
    float a[6];
};
YB_HD sym3 sym_diag(float d) { return sym3{{d, d, d, 0.f, 0.f, 0.f}}; }
YB_HD sym3 sym_zero() { return sym3{{0.f, 0.f, 0.f, 0.f, 0.f, 0.f}}; }
YB_HD sym3 sym_outer(v3 v) {
    return sym3{{v.x * v.x, v.y * v.y, v.z * v.z, v.x * v.y, v.x * v.z, v.y * v.z}};
}
YB_HD void sym_add_diag(sym3& m, float d) { m.a[0] += d; m.a[1] += d; m.a[2] += d; }
YB_HD void sym_axpy(sym3& m, const sym3& o, float s) {
YB_UNROLL
    for (int i = 0; i < 6; i++) m.a[i] += o.a[i] * s;
}
YB_HD void sym_add(sym3& m, const sym3& o) {
YB_UNROLL
    for (int i = 0; i < 6; i++) m.a[i] += o.a[i];
}
YB_HD v3 sym_apply(const sym3& m, v3 x) {
    return v3{(m.a[0] * x.x + m.a[3] * x.y) + m.a[4] * x.z,
              (m.a[3] * x.x + m.a[1] * x.y) + m.a[5] * x.z,
              (m.a[4] * x.x + m.a[5] * x.y) + m.a[2] * x.z};
}
// H^{-1} f through the adjugate of the symmetric matrix.
YB_HD v3 sym_solve(const sym3& m, v3 f) {
    float a = m.a[0], b = m.a[1], c = m.a[2], d = m.a[3], e = m.a[4], g = m.a[5];
    float A00 = b * c - g * g;
    float A01 = e * g - d * c;
    float A02 = d * g - b * e;
    float A11 = a * c - e * e;
    float A12 = d * e - a * g;
    float A22 = a * b - d * d;
    float det = (a * A00 + d * A01) + e * A02;
    float inv = 1.0f / det;
    return v3{((A00 * f.x + A01 * f.y) + A02 * f.z) * inv,
              ((A01 * f.x + A11 * f.y) + A12 * f.z) * inv,
              ((A02 * f.x + A12 * f.y) + A22 * f.z) * inv};
}

// 30-bit Morton code of a point already quantised to [0,1023]^3.
YB_HD uint32_t expand_bits10(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}
YB_HD uint32_t quant10(float c, float lo, float ext) {
    float t = (c - lo) / ext;
    float q = clampf(t * 1024.0f, 0.0f, 1023.0f);
    if (!(q == q)) q = 0.0f;  // NaN (degenerate extent) maps to cell 0
    return (uint32_t)q;
}

}  // namespace yb
