// host_setup.cpp — host-side set-up math of the stepper (FP32, as the reference):
// constructor defaults, configure(), setKStretch/setKBend, uploadMeta checks.
#include <float.h>
#include <math.h>
#include <string.h>

#include <string>

#include "../../include/yarnball_hip.h"
#include "ybmath.h"

namespace yb {
void set_message(const std::string& m);
}

using namespace yb;

// Sim::Sim defaults (YarnBall.cu:14-35).  lastH is left equal to h so the first
// advance sees lastH = maxH, as the reference's first advance does.
extern "C" int yb_default_params(yb_params* p, int numVerts, float maxH) {
    if (!p) return YB_ERR_INVALID_ARGUMENT;
    memset(p, 0, sizeof(*p));
    p->numVerts = numVerts;
    p->gravity[0] = 0.f;
    p->gravity[1] = -9.8f;
    p->gravity[2] = 0.f;
    p->worldFloorNormal[1] = 1.f;
    p->h = maxH;
    p->lastH = maxH;
    p->drag = 0.2f;
    p->damping = 1e-6f;
    p->time = 0.f;
    p->radius = 1e-4f;
    p->barrierThickness = 8e-4f;
    p->accelerationRatio = 1.f;
    p->kCollision = 1e-5f;
    p->detectionScaler = 1.2f;
    p->frictionCoeff = 0.1f;
    p->kFriction = 5.f;
    p->detectionPeriod = 1;
    p->useStepSizeLimit = 1;
    p->useVelocityRadius = 1;
    p->bvhRebuildPeriod = 1 / 10.f;
    p->numItr = 8;
    p->maxSegLen = 0.f;
    p->minSegLen = FLT_MAX;
    return YB_OK;
}

// Per-vertex defaults (YarnBall.cu:42-50).
extern "C" int yb_default_vertices(int n, yb_vertex* verts, float* vels, float* qs,
                                   float* qRests) {
    if (n < 3) {
        set_message("Too little vertices");
        return YB_ERR_INVALID_ARGUMENT;
    }
    for (int i = 0; i < n; i++) {
        if (verts) {
            yb_vertex& v = verts[i];
            v.pos[0] = v.pos[1] = v.pos[2] = 0.f;
            v.invMass = v.lRest = 1.f;
            v.kStretch = 100.f;
            v.connectionIndex = -1;
            v.flags = YB_FLAG_HAS_NEXT;
        }
        if (vels) vels[3 * i] = vels[3 * i + 1] = vels[3 * i + 2] = 0.f;
        if (qs) {
            qs[4 * i] = qs[4 * i + 1] = qs[4 * i + 2] = 0.f;
            qs[4 * i + 3] = 1.f;
        }
        if (qRests) {
            qRests[4 * i] = qRests[4 * i + 1] = qRests[4 * i + 2] = 0.f;
            qRests[4 * i + 3] = 5.f;
        }
    }
    if (verts) verts[n - 1].flags = 0;
    return YB_OK;
}

static inline v3 vpos(const yb_vertex& v) { return v3{v.pos[0], v.pos[1], v.pos[2]}; }

// Sim::configure host part (YarnBall.cu:91-157).
extern "C" int yb_configure_host(int n, yb_vertex* verts, float* qs, float* qRests, float density,
                                 float* initialInvMasses, yb_params* p) {
    if (n < 3 || !verts || !qs || !qRests || !p) {
        set_message("yb_configure_host: invalid argument");
        return YB_ERR_INVALID_ARGUMENT;
    }
    float maxSeg = 0.f;
    float minSeg = FLT_MAX;
    rot lastQ = rot{0.f, 0.f, 0.f, 1.f};
    // The last vertex cannot start a segment (the reference never reads past it
    // because its constructor clears that vertex's flags).
    verts[n - 1].flags &= ~(uint32_t)YB_FLAG_HAS_NEXT;
    for (int i = 0; i < n; i++) {
        yb_vertex& v = verts[i];
        if (i < n - 1) {
            const bool hasNext = (v.flags & YB_FLAG_HAS_NEXT) != 0;
            uint32_t& nf = verts[i + 1].flags;
            nf = hasNext ? (nf | YB_FLAG_HAS_PREV) : (nf & ~(uint32_t)YB_FLAG_HAS_PREV);
            const bool hasNextNext = hasNext && (nf & YB_FLAG_HAS_NEXT);
            v.flags = hasNextNext ? (v.flags | YB_FLAG_HAS_NEXT_ORIENTATION)
                                  : (v.flags & ~(uint32_t)YB_FLAG_HAS_NEXT_ORIENTATION);
            if (!hasNext) v.flags |= YB_FLAG_FIX_ORIENTATION;
        }
        if (i == 0) v.flags &= ~(uint32_t)YB_FLAG_HAS_PREV;
        if (!(v.flags & YB_FLAG_HAS_PREV) && !(v.flags & YB_FLAG_HAS_NEXT)) {
            set_message("Dangling segment. Yarns must be atleast 2 segments long");
            return YB_ERR_VALIDATION;
        }
        v.lRest = 1.f / (float)n;
        float mass = 0.f;
        if (v.flags & YB_FLAG_HAS_PREV) mass += verts[i - 1].lRest;
        if (v.flags & YB_FLAG_HAS_NEXT) {
            const v3 seg0 = vpos(verts[i + 1]) - vpos(v);
            v.lRest = len(seg0);
            if (v.lRest == 0.f || !isfinite(v.lRest)) {
                set_message("0 length segment");
                return YB_ERR_VALIDATION;
            }
            // Minimal-twist frame transport (YarnBall.cu:131-137): the frame
            // that maps e1 onto the segment and is closest to the previous one.
            const v3 dir = seg0 * (1.0f / sqrtf(len2(seg0)));
            const rot qq = rfrom_to(v3{1.f, 0.f, 0.f}, dir);
            rot t = rmul(rinv(qq), lastQ);
            t = rot{t.x, 0.f, 0.f, t.w};
            t = rscale(t, 1.0f / rlen(t));
            lastQ = rmul(qq, t);
            mass += v.lRest;
            maxSeg = fmaxf(maxSeg, v.lRest);
            minSeg = fminf(minSeg, v.lRest);
        }
        qs[4 * i] = lastQ.x;
        qs[4 * i + 1] = lastQ.y;
        qs[4 * i + 2] = lastQ.z;
        qs[4 * i + 3] = lastQ.w;
        mass *= 0.5f * density;
        if (mass != 0.f)
            v.invMass *= 1.f / mass;
        else
            v.invMass = 0.f;
        if (initialInvMasses) initialInvMasses[i] = v.invMass;
    }
    // Rest Darboux rotors, keeping the stored magnitude (YarnBall.cu:156-157).
    for (int i = 0; i < n - 1; i++) {
        const rot qr = rot{qRests[4 * i], qRests[4 * i + 1], qRests[4 * i + 2], qRests[4 * i + 3]};
        const float m = rlen(qr);
        const rot a = rot{qs[4 * i], qs[4 * i + 1], qs[4 * i + 2], qs[4 * i + 3]};
        const rot b = rot{qs[4 * i + 4], qs[4 * i + 5], qs[4 * i + 6], qs[4 * i + 7]};
        const rot d = rscale(rmul(rinv(a), b), m);
        qRests[4 * i] = d.x;
        qRests[4 * i + 1] = d.y;
        qRests[4 * i + 2] = d.z;
        qRests[4 * i + 3] = d.w;
    }
    p->maxSegLen = maxSeg;
    p->minSegLen = minSeg;
    p->numVerts = n;
    return YB_OK;
}

// Sim::setKStretch (YarnBall.cu:202-210): per-segment energy density.
extern "C" int yb_set_kstretch_host(int n, yb_vertex* verts, float kStretch) {
    if (n < 1 || !verts) return YB_ERR_INVALID_ARGUMENT;
    for (int i = 0; i < n; i++) verts[i].kStretch = kStretch * verts[i].lRest;
    return YB_OK;
}

// Sim::setKBend (YarnBall.cu:212-224): |qRest| carries 4 kBend / lRest.
extern "C" int yb_set_kbend_host(int n, const yb_vertex* verts, float* qRests, float kBend) {
    if (n < 1 || !verts || !qRests) return YB_ERR_INVALID_ARGUMENT;
    kBend *= 4;
    for (int i = 0; i < n; i++) {
        rot r = rot{qRests[4 * i], qRests[4 * i + 1], qRests[4 * i + 2], qRests[4 * i + 3]};
        r = rscale(r, 1.0f / rlen(r));
        r = rscale(r, kBend / verts[i].lRest);
        qRests[4 * i] = r.x;
        qRests[4 * i + 1] = r.y;
        qRests[4 * i + 2] = r.z;
        qRests[4 * i + 3] = r.w;
    }
    return YB_OK;
}

// Sim::uploadMeta checks and derived radii (YarnBall.cu:226-235).
extern "C" int yb_validate_params(yb_params* p) {
    if (!p) return YB_ERR_INVALID_ARGUMENT;
    p->detectionRadius = p->radius + 0.5f * p->barrierThickness;
    p->scaledDetectionRadius = p->detectionRadius * p->detectionScaler;
    if (p->detectionScaler <= 1) {
        set_message("Detection scaler must be greater than 1.");
        return YB_ERR_VALIDATION;
    }
    if (2 * p->minSegLen < 2 * p->radius + p->barrierThickness) {
        set_message(
            "Use thinner yarn or use longer segments. (2 * Min seg length must be at least 2 * "
            "radius + barrierThickness");
        return YB_ERR_VALIDATION;
    }
    return YB_OK;
}
