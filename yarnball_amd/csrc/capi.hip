// capi.hip — extern "C" entry points of libyarnball_hip.so: device lifecycle,
// substep scheduling with captured HIP graphs, error reporting.
//
// Scheduling mirrors Sim::advance (step.cpp:36-55): a frame of length h is
// split into ceil(h / maxH) substeps; each substep is initItr, optional contact
// detection every detectionPeriod substeps (full tree rebuild when
// bvhRebuildPeriod of simulated time has passed, refit otherwise,
// collision.cu:101-108), the step limit, numItr sweeps and endItr.  Unlike the
// reference, detection needs no host round trip (the reference reads the
// root box and the pair count back at collision.cu:109,111), so every substep
// -- detection included -- is a single graph launch; the only host sync per
// advance is the error-flag read of checkErrors (YarnBall.cu:276-304).
#include <math.h>
#include <string.h>

#include <algorithm>
#include <functional>
#include <string>
#include <vector>

#include "yb_internal.h"

namespace yb {

static thread_local std::string g_msg;
void set_message(const std::string& m) { g_msg = m; }

}  // namespace yb

using namespace yb;

#define HIPCHECK(expr)                                                              \
    do {                                                                            \
        hipError_t e_ = (expr);                                                     \
        if (e_ != hipSuccess) {                                                     \
            set_message(std::string(#expr) + ": " + hipGetErrorString(e_));         \
            return YB_ERR_HIP;                                                      \
        }                                                                           \
    } while (0)

extern "C" const char* yb_last_message(void) { return g_msg.c_str(); }

// Substep timeline of the residual-targeted schedule (yb_debug_timeline): a
// HIP event before the prologue, after it, after every chunk and after the
// end kernels, all recorded between graph launches on the handle's stream.
struct TimelineRow {
    int first = 0;   // index of the row's first event in ev
    int chunks = 0;  // chunk events (launched chunks)
    int done = 0;    // sweeps to convergence (maxSweeps if not converged)
    int conv = -1;   // chunk holding the converged sweep, -1 if none
};
struct Timeline {
    int cap = 0;
    std::vector<hipEvent_t> ev;
    size_t used = 0;
    std::vector<TimelineRow> rows;
    bool open = false;  // a row is being recorded
};

static void timeline_free(yb_sim* s) {
    Timeline* t = (Timeline*)s->timeline;
    if (!t) return;
    for (hipEvent_t e : t->ev) (void)hipEventDestroy(e);
    delete t;
    s->timeline = nullptr;
}

// Records the next timeline event of the open row (no-op without one).
static int timeline_mark(yb_sim* s) {
    Timeline* t = (Timeline*)s->timeline;
    if (!t || !t->open) return YB_OK;
    if (t->used == t->ev.size()) {
        hipEvent_t e;
        HIPCHECK(hipEventCreate(&e));
        t->ev.push_back(e);
    }
    HIPCHECK(hipEventRecord(t->ev[t->used++], s->stream));
    return YB_OK;
}

static int reset_scene_ctl(yb_sim* s);
static int set_skin(yb_sim* s, float margin, float sweep);

static void free_all(yb_sim* s) {
    timeline_free(s);
    Buffers& b = s->b;
    void* ptrs[] = {b.pos,     b.prm,     b.vel,      b.lastVel, b.yv,      b.dx[0],
                    b.dx[1],   b.q[0],    b.q[1],     b.qRest,   b.numCols, b.cols,
                    b.aabb,    b.sceneBox, b.blockBox, b.morton, b.keys[0], b.keys[1],
                    b.vals[0], b.vals[1], b.hist, b.scanPart, b.nodes,   b.parent,  b.rng,
                    b.kids,    b.strad,   b.stradCount, b.leafNeed, b.sufBox, b.preBox, b.blkTab,
                    b.err,     b.params,  b.stage, b.ctl, b.partial, b.resBlocks, b.resCounter,
                    b.minSep,
                    b.mlpPrm,  b.mlpTiles, b.mlpBias, b.corrPrev, b.dxPlain, b.predC,
                    b.lineG,   b.lineQL,  b.lineE[0], b.lineE[1], b.lineF[0], b.lineF[1],
                    b.mlpCount, b.recX, b.recT, b.recCount, b.sceneOf, b.sceneOff,
                    b.sceneBoxes, b.sceneRoot, b.sceneEndFlags, b.drvInit, b.candCount,
                    b.sbox,    b.cand,    b.candN,    b.skinState, b.skinStats, b.skinRoot,
                    b.skinPolicy};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (b.hostFlag) (void)hipHostFree(b.hostFlag);
    b = Buffers();
    for (int i = 0; i < 3; i++) {
        if (s->graphs[i]) (void)hipGraphExecDestroy(s->graphs[i]);
        for (int w = 0; w < 2; w++) {
            if (s->proGraphs[w][i]) (void)hipGraphExecDestroy(s->proGraphs[w][i]);
            s->proGraphs[w][i] = nullptr;
        }
        s->graphs[i] = nullptr;
    }
    if (s->endGraph) (void)hipGraphExecDestroy(s->endGraph);
    s->endGraph = nullptr;
    if (s->chunkGraphs) {
        for (int c = 0; c < s->numChunks; c++)
            if (s->chunkGraphs[c]) (void)hipGraphExecDestroy(s->chunkGraphs[c]);
        delete[] s->chunkGraphs;
        s->chunkGraphs = nullptr;
    }
    for (int i = 0; i < 2; i++)
        if (s->chunkEv[i]) {
            (void)hipEventDestroy(s->chunkEv[i]);
            s->chunkEv[i] = nullptr;
        }
    if (s->stream) {
        (void)hipStreamDestroy(s->stream);
        s->stream = nullptr;
    }
}

template <typename T>
static hipError_t dalloc(T** p, size_t count) {
    return hipMalloc((void**)p, sizeof(T) * (count ? count : 1));
}

// Refit blocking for a tile of `leaves` sorted leaves per workgroup: block
// count, sparse-table levels and the grid of the range-union pass.
static void set_refit_geometry(yb_sim* s, int leaves) {
    const int n = s->n;
    s->refitLeaves = leaves;
    s->refitShift = 31 - __builtin_clz((unsigned)leaves);
    s->refitBlocks = (n + leaves - 1) / leaves;
    s->refitLevels = 1;
    while ((1 << s->refitLevels) <= s->refitBlocks) s->refitLevels++;
    s->refitTopBlocks = (n / leaves) * 64 / 256 + 1;
    if (s->refitTopBlocks > 2048) s->refitTopBlocks = 2048;
}

extern "C" int yb_create(int n, yb_sim** out) {
    if (!out) {
        set_message("yb_create: null output pointer");
        return YB_ERR_INVALID_ARGUMENT;
    }
    *out = nullptr;
    if (n < 3) {
        set_message("Too little vertices");  // YarnBall.cu:12
        return YB_ERR_INVALID_ARGUMENT;
    }
    if (n > YB_MAX_VERTICES) {
        set_message("yb_create: more than YB_MAX_VERTICES (268435456) vertices");
        return YB_ERR_INVALID_ARGUMENT;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        set_message("yb_create: no HIP device visible");
        return YB_ERR_NO_DEVICE;
    }
    yb_sim* s = new yb_sim();
    s->n = n;
    {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
            cus > 0)
            s->numCUs = cus;
    }
    s->sortTiles = (n + kSortTile - 1) / kSortTile;
    s->boxBlocks = (n + 255) / 256;
    if (s->boxBlocks > 1024) s->boxBlocks = 1024;
    set_refit_geometry(s, kRefitLeavesMin);  // sizes the table for the smallest tile
    const size_t tabEntries = 2 * (size_t)s->refitBlocks * s->refitLevels;
    set_refit_geometry(s, kRefitLeavesMax);
    Buffers& b = s->b;
    const size_t N = (size_t)n;
    hipError_t e = hipSuccess;
#define ALLOC(p, c) \
    if (e == hipSuccess) e = dalloc(&(p), (c))
    ALLOC(b.pos, N);
    ALLOC(b.prm, N);
    ALLOC(b.vel, N);
    ALLOC(b.lastVel, N);
    ALLOC(b.yv, N);
    ALLOC(b.dx[0], N);
    ALLOC(b.dx[1], N);
    ALLOC(b.q[0], N);
    ALLOC(b.q[1], N);
    ALLOC(b.qRest, N);
    ALLOC(b.numCols, N);
    ALLOC(b.cols, N * kMaxCols);
    ALLOC(b.aabb, 2 * N);
    ALLOC(b.sceneBox, 2);
    ALLOC(b.blockBox, 2 * (size_t)s->boxBlocks);
    ALLOC(b.morton, N);
    ALLOC(b.keys[0], N);
    ALLOC(b.keys[1], N);
    ALLOC(b.vals[0], N);
    ALLOC(b.vals[1], N);
    ALLOC(b.hist, (size_t)kSortBins * s->sortTiles);
    ALLOC(b.scanPart, ((size_t)kSortBins * s->sortTiles + kScanBlock - 1) / kScanBlock);
    ALLOC(b.nodes, 4 * (N - 1));
    ALLOC(b.parent, 2 * N - 1);
    ALLOC(b.rng, N - 1);
    ALLOC(b.kids, N - 1);
    ALLOC(b.strad, N - 1);
    ALLOC(b.stradCount, 1);
    ALLOC(b.leafNeed, N);
    ALLOC(b.sufBox, 2 * N);
    ALLOC(b.preBox, 2 * N);
    ALLOC(b.blkTab, tabEntries);
    ALLOC(b.err, 2);
    ALLOC(b.params, 1);
    ALLOC(b.ctl, 1);
    ALLOC(b.minSep, 1);
    ALLOC(b.mlpPrm, kMlpParams);
    ALLOC(b.mlpTiles, kMlpTileBufElems);
    ALLOC(b.mlpBias, 2 * kMlpHid + 16);
    ALLOC(b.corrPrev, N);
    ALLOC(b.dxPlain, N);
    ALLOC(b.predC, N);
    ALLOC(b.mlpCount, 2 * ((N + 63) / 64));
    ALLOC(b.recCount, 2);
    ALLOC(b.partial, 3 * (size_t)sweep_waves(n));
    ALLOC(b.resBlocks, 2 * 64);
    ALLOC(b.resCounter, 1);
    ALLOC(b.lineG, N);
    ALLOC(b.lineQL, N);
    for (int k = 0; k < 2; k++) {
        ALLOC(b.lineE[k], (size_t)sweep_waves(n));
        ALLOC(b.lineF[k], (size_t)sweep_waves(n));
    }
    ALLOC(b.sbox, 2 * N);
    ALLOC(b.cand, N * kSkinCand);
    ALLOC(b.candN, N);
    ALLOC(b.skinState, 1 + (size_t)kMaxScenes);
    ALLOC(b.skinStats, 2);
    ALLOC(b.skinRoot, 2);
    ALLOC(b.skinPolicy, (size_t)kMaxScenes);
#undef ALLOC
    if (e == hipSuccess) e = hipHostMalloc((void**)&b.hostFlag, 2 * sizeof(int32_t));
    if (e == hipSuccess) e = hipEventCreateWithFlags(&s->chunkEv[0], hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&s->chunkEv[1], hipEventDisableTiming);
    if (e == hipSuccess) {
        const float inf = INFINITY;
        e = hipMemcpy(b.minSep, &inf, sizeof(float), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) {
        SolverCtl z;
        memset(&z, 0, sizeof(z));
        z.doneAt = 0x7fffffff;
        e = hipMemcpy(b.ctl, &z, sizeof(z), hipMemcpyHostToDevice);
    }
    b.order = b.vals[0];
    b.stageBytes = N * (sizeof(yb_vertex) + 11 * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&b.stage, b.stageBytes);
    if (e == hipSuccess) e = hipStreamCreate(&s->stream);
    if (e == hipSuccess) e = hipMemset(b.lastVel, 0, sizeof(float4) * N);
    if (e == hipSuccess) e = hipMemset(b.vel, 0, sizeof(float4) * N);
    if (e == hipSuccess) e = hipMemset(b.dx[0], 0, sizeof(float4) * N);
    if (e == hipSuccess) e = hipMemset(b.dx[1], 0, sizeof(float4) * N);
    if (e == hipSuccess) e = hipMemset(b.yv, 0, sizeof(float4) * N);
    if (e == hipSuccess) e = hipMemset(b.numCols, 0, sizeof(int32_t) * N);
    if (e == hipSuccess) e = hipMemset(b.err, 0, sizeof(int32_t) * 2);
    if (e == hipSuccess) e = hipMemset(b.nodes, 0, sizeof(float4) * 4 * (N - 1));
    if (e == hipSuccess) e = hipMemset(b.corrPrev, 0, sizeof(float4) * N);
    if (e == hipSuccess)
        e = hipMemset(b.mlpCount, 0, sizeof(unsigned long long) * 2 * ((N + 63) / 64));
    if (e == hipSuccess) e = hipMemset(b.recCount, 0, 2 * sizeof(int64_t));
    // contact skin: no kept candidates yet (every segment walks first)
    if (e == hipSuccess) e = hipMemset(b.candN, 0xff, sizeof(int32_t) * N);
    if (e == hipSuccess) e = hipMemset(b.skinState, 0, sizeof(int32_t) * (1 + (size_t)kMaxScenes));
    if (e == hipSuccess) e = hipMemset(b.skinStats, 0, 2 * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemset(b.skinPolicy, 0, sizeof(int4) * (size_t)kMaxScenes);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        set_message(std::string("yb_create: ") + hipGetErrorString(e));
        free_all(s);
        delete s;
        return YB_ERR_HIP;
    }
    *out = s;
    return YB_OK;
}

extern "C" void yb_destroy(yb_sim* s) {
    if (!s) return;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    free_all(s);
    delete s;
}

extern "C" int yb_upload(yb_sim* s, const yb_vertex* verts, const float* vels, const float* qs,
                         const float* qRests) {
    if (!s || !verts || !vels || !qs || !qRests) {
        set_message("yb_upload: null argument");
        return YB_ERR_INVALID_ARGUMENT;
    }
    const size_t N = (size_t)s->n;
    // Every index the kernels follow must stay inside the vertex array.
    if (verts[N - 1].flags & YB_FLAG_HAS_NEXT) {
        set_message("yb_upload: the last vertex cannot have a next vertex");
        return YB_ERR_INVALID_ARGUMENT;
    }
    for (size_t i = 0; i < N; i++) {
        const int c = verts[i].connectionIndex;
        if (c < -1 || c >= (int)N) {
            set_message("yb_upload: connectionIndex out of range at vertex " + std::to_string(i));
            return YB_ERR_INVALID_ARGUMENT;
        }
    }
    for (int k = 0; k + 1 < (int)s->sceneOffHost.size(); k++) {
        if (verts[s->sceneOffHost[k + 1] - 1].flags & YB_FLAG_HAS_NEXT) {
            set_message("yb_upload: the last vertex of scene " + std::to_string(k) +
                        " has a next vertex (a segment would cross into the next scene)");
            return YB_ERR_INVALID_ARGUMENT;
        }
    }
    char* st = (char*)s->b.stage;
    HIPCHECK(hipMemcpyAsync(st, verts, sizeof(yb_vertex) * N, hipMemcpyHostToDevice, s->stream));
    st += sizeof(yb_vertex) * N;
    HIPCHECK(hipMemcpyAsync(st, vels, sizeof(float) * 3 * N, hipMemcpyHostToDevice, s->stream));
    st += sizeof(float) * 3 * N;
    HIPCHECK(hipMemcpyAsync(st, qs, sizeof(float) * 4 * N, hipMemcpyHostToDevice, s->stream));
    st += sizeof(float) * 4 * N;
    HIPCHECK(hipMemcpyAsync(st, qRests, sizeof(float) * 4 * N, hipMemcpyHostToDevice, s->stream));
    launch_upload(s);
    HIPCHECK(hipGetLastError());
    // Update-residual normaliser: sum over free vertices of l_i^2, l_i the rest
    // length of the vertex's own segment (or of the previous one at a strand end).
    double denom = 0.0;
    s->lsqHost.assign(N, 0.0);
    for (size_t i = 0; i < N; i++) {
        if (verts[i].invMass == 0.f) continue;
        double l = verts[i].lRest;
        if (!(verts[i].flags & YB_FLAG_HAS_NEXT) && (verts[i].flags & YB_FLAG_HAS_PREV) && i > 0)
            l = verts[i - 1].lRest;
        denom += l * l;
        s->lsqHost[i] = l * l;
    }
    s->denomHost = denom;
    HIPCHECK(hipMemcpyAsync(&s->b.ctl->denom, &s->denomHost, sizeof(double),
                            hipMemcpyHostToDevice, s->stream));
    HIPCHECK(hipStreamSynchronize(s->stream));
    s->uploaded = true;
    return reset_scene_ctl(s);
}

extern "C" int yb_driver_default(int kind, yb_driver* d) {
    if (!d || (kind != YB_DRIVE_TWIST && kind != YB_DRIVE_PULL)) return YB_ERR_INVALID_ARGUMENT;
    memset(d, 0, sizeof(*d));
    d->kind = kind;
    if (kind == YB_DRIVE_TWIST) {  // main.cpp:58-61, 73
        d->speed = 0.5f * 2.0f * (float)M_PI;
        d->start = 2.0f;
        d->duration = 12.0f;
        d->release = 12.0f + 3.0f;
        d->side = -1;
    } else {  // main.cpp:95-97
        d->speed = 0.25f;
        d->start = 2.0f;
        d->duration = 6.0f;
        d->release = -1.0f;
    }
    return YB_OK;
}

extern "C" int yb_driver_init(yb_sim* s, const float* initialInvMasses, float center[3]) {
    if (!s) return YB_ERR_INVALID_ARGUMENT;
    if (!s->uploaded) {
        set_message("yb_driver_init: upload the scene first");
        return YB_ERR_INVALID_ARGUMENT;
    }
    const size_t N = (size_t)s->n;
    if (!s->b.drvInit) HIPCHECK(hipMalloc(&s->b.drvInit, sizeof(float4) * N));
    std::vector<float4> p(N);
    HIPCHECK(hipMemcpyAsync(p.data(), s->b.pos, sizeof(float4) * N, hipMemcpyDeviceToHost,
                            s->stream));
    HIPCHECK(hipStreamSynchronize(s->stream));
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t i = 0; i < N; i++) {
        const float* q = &p[i].x;
        for (int k = 0; k < 3; k++) {
            lo[k] = fminf(lo[k], q[k]);
            hi[k] = fmaxf(hi[k], q[k]);
        }
        if (initialInvMasses) p[i].w = initialInvMasses[i];
    }
    HIPCHECK(hipMemcpy(s->b.drvInit, p.data(), sizeof(float4) * N, hipMemcpyHostToDevice));
    if (center)
        for (int k = 0; k < 3; k++) center[k] = 0.5f * (hi[k] + lo[k]);
    return YB_OK;
}

extern "C" int yb_driver_frame(yb_sim* s, yb_params* p, const yb_driver* d, float t, float dt) {
    if (!s || !p || !d || !(dt > 0.f)) return YB_ERR_INVALID_ARGUMENT;
    if (!s->b.drvInit) {
        set_message("yb_driver_frame: call yb_driver_init first");
        return YB_ERR_INVALID_ARGUMENT;
    }
    if (d->kind != YB_DRIVE_TWIST && d->kind != YB_DRIVE_PULL) return YB_ERR_INVALID_ARGUMENT;
    if (d->kind == YB_DRIVE_PULL) {  // main.cpp:93
        p->gravity[0] = 0.f;
        p->gravity[1] = 0.f;
        p->gravity[2] = -9.8f;
    }
    launch_drive(s, *d, t, dt);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(s->stream));
    return YB_OK;
}

// Per-scene solver state of a batch, reset: update-residual normalisers from
// the uploaded vertices, everything else cleared (yb_set_scenes, yb_upload,
// yb_solver_stats with reset).
static int reset_scene_ctl(yb_sim* s) {
    if (s->numScenes <= 1 || !s->b.sceneCtl) return YB_OK;
    std::vector<SceneCtl> c(s->numScenes);
    for (int k = 0; k < s->numScenes; k++) {
        memset(&c[k], 0, sizeof(SceneCtl));
        c[k].doneAt = 0x7fffffff;
        c[k].denomF = -1.0;
        double d = 0.0;
        if (!s->lsqHost.empty())
            for (int v = s->sceneOffHost[k]; v < s->sceneOffHost[k + 1]; v++) d += s->lsqHost[v];
        c[k].denom = d;
    }
    HIPCHECK(hipMemcpyAsync(s->b.sceneCtl, c.data(), sizeof(SceneCtl) * c.size(),
                            hipMemcpyHostToDevice, s->stream));
    HIPCHECK(hipStreamSynchronize(s->stream));
    return YB_OK;
}

// Back to the single-scene layout (frees the batched-scene buffers).
static int clear_scenes(yb_sim* s) {
    Buffers& b = s->b;
    for (void* p : {(void*)b.sceneOf, (void*)b.sceneOff, (void*)b.sceneBoxes, (void*)b.sceneRoot,
                    (void*)b.sceneEndFlags, (void*)b.sceneCtl, (void*)b.scenePart,
                    (void*)b.tileList, (void*)b.tileCount})
        if (p) HIPCHECK(hipFree(p));
    b.sceneOf = b.sceneOff = b.sceneRoot = b.sceneEndFlags = nullptr;
    b.sceneBoxes = nullptr;
    b.sceneCtl = nullptr;
    b.scenePart = nullptr;
    b.tileList = b.tileCount = nullptr;
    s->numScenes = 1;
    s->sceneOffHost.clear();
    s->treeBuilt = false;
    s->graphItr = -1;
    s->adaptiveReady = false;
    return YB_OK;
}

extern "C" int yb_set_scenes(yb_sim* s, int numScenes, const int32_t* offsets) {
    if (!s || numScenes < 1 || !offsets) {
        set_message("yb_set_scenes: invalid argument");
        return YB_ERR_INVALID_ARGUMENT;
    }
    if (numScenes > kMaxScenes) {
        set_message("yb_set_scenes: at most 65536 scenes");
        return YB_ERR_INVALID_ARGUMENT;
    }
    if (offsets[0] != 0 || offsets[numScenes] != s->n) {
        set_message("yb_set_scenes: offsets must start at 0 and end at the vertex count");
        return YB_ERR_INVALID_ARGUMENT;
    }
    for (int k = 0; k < numScenes; k++)
        if (offsets[k + 1] - offsets[k] < 2) {
            set_message("yb_set_scenes: scene " + std::to_string(k) + " has fewer than 2 vertices");
            return YB_ERR_INVALID_ARGUMENT;
        }
    HIPCHECK(hipStreamSynchronize(s->stream));
    int rc = clear_scenes(s);
    if (rc == YB_OK && numScenes == 1 && !s->skinSet)
        rc = set_skin(s, kSkinDefaultMargin, kSkinDefaultSweep);
    if (rc != YB_OK || numScenes == 1) return rc;
    Buffers& b = s->b;
    const size_t N = (size_t)s->n;
    HIPCHECK(hipMalloc(&b.sceneOf, sizeof(int32_t) * N));
    HIPCHECK(hipMalloc(&b.sceneOff, sizeof(int32_t) * (numScenes + 1)));
    HIPCHECK(hipMalloc(&b.sceneBoxes, sizeof(float4) * 2 * numScenes));
    HIPCHECK(hipMalloc(&b.sceneRoot, sizeof(int32_t) * numScenes));
    HIPCHECK(hipMalloc(&b.sceneEndFlags, sizeof(int32_t) * numScenes));
    HIPCHECK(hipMalloc(&b.sceneCtl, sizeof(SceneCtl) * numScenes));
    HIPCHECK(hipMalloc(&b.scenePart, sizeof(double2) * 6 * (size_t)sweep_waves(s->n)));
    HIPCHECK(hipMalloc(&b.tileList, sizeof(int32_t) * (size_t)sweep_waves(s->n)));
    HIPCHECK(hipMalloc(&b.tileCount, sizeof(int32_t) * (numScenes + 1)));
    std::vector<int32_t> of(N);
    for (int k = 0; k < numScenes; k++)
        for (int v = offsets[k]; v < offsets[k + 1]; v++) of[v] = k;
    HIPCHECK(hipMemcpy(b.sceneOf, of.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(b.sceneOff, offsets, sizeof(int32_t) * (numScenes + 1),
                       hipMemcpyHostToDevice));
    s->numScenes = numScenes;
    s->sceneOffHost.assign(offsets, offsets + numScenes + 1);
    rc = reset_scene_ctl(s);
    if (rc != YB_OK) return rc;
    if (s->uploaded) {
        launch_scene_end_flags(s);
        HIPCHECK(hipGetLastError());
        std::vector<int32_t> fl(numScenes);
        HIPCHECK(hipMemcpyAsync(fl.data(), b.sceneEndFlags, sizeof(int32_t) * numScenes,
                                hipMemcpyDeviceToHost, s->stream));
        HIPCHECK(hipStreamSynchronize(s->stream));
        for (int k = 0; k < numScenes; k++)
            if (fl[k] & YB_FLAG_HAS_NEXT) {
                rc = clear_scenes(s);
                set_message("yb_set_scenes: the last vertex of scene " + std::to_string(k) +
                            " has a next vertex (a segment would cross into the next scene)");
                return rc == YB_OK ? YB_ERR_INVALID_ARGUMENT : rc;
            }
    }
    // the contact skin's per-scene state starts over; a batch starts without
    // the skin unless the caller chose a setting (its scenes' motion varies,
    // and where most move too fast for kept lists the launches over the whole
    // batch cost more than the reuses save)
    HIPCHECK(hipMemsetAsync(b.skinPolicy, 0, sizeof(int4) * (size_t)kMaxScenes, s->stream));
    HIPCHECK(hipStreamSynchronize(s->stream));
    if (!s->skinSet) {
        rc = set_skin(s, 0.f, s->skinSweep);
        if (rc != YB_OK) return rc;
    }
    // the tree and the captured graphs depend on the scene layout
    s->treeBuilt = false;
    s->graphItr = -1;
    s->adaptiveReady = false;
    return YB_OK;
}

extern "C" int yb_download(yb_sim* s, yb_vertex* verts, float* vels, float* qs) {
    if (!s) {
        set_message("yb_download: null handle");
        return YB_ERR_INVALID_ARGUMENT;
    }
    const size_t N = (size_t)s->n;
    launch_download(s);
    HIPCHECK(hipGetLastError());
    const char* st = (const char*)s->b.stage;
    if (verts)
        HIPCHECK(hipMemcpyAsync(verts, st, sizeof(yb_vertex) * N, hipMemcpyDeviceToHost, s->stream));
    st += sizeof(yb_vertex) * N;
    if (vels)
        HIPCHECK(hipMemcpyAsync(vels, st, sizeof(float) * 3 * N, hipMemcpyDeviceToHost, s->stream));
    st += sizeof(float) * 3 * N;
    if (qs) HIPCHECK(hipMemcpyAsync(qs, st, sizeof(float) * 4 * N, hipMemcpyDeviceToHost, s->stream));
    HIPCHECK(hipStreamSynchronize(s->stream));
    return YB_OK;
}

extern "C" int yb_zero_velocities(yb_sim* s) {
    if (!s) return YB_ERR_INVALID_ARGUMENT;
    launch_zero_vel(s);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(s->stream));
    return YB_OK;
}

static DevParams make_dev_params(const yb_params* p, int n) {
    DevParams d;
    memset(&d, 0, sizeof(d));
    d.gx = p->gravity[0];
    d.gy = p->gravity[1];
    d.gz = p->gravity[2];
    d.h = p->h;
    d.lastH = p->lastH;
    d.dampingH = p->damping / p->h;
    d.drag = p->drag;
    d.frictionCoeff = p->frictionCoeff;
    d.kCollision = p->kCollision;
    d.kFriction = p->kFriction;
    d.detectionRadius = p->detectionRadius;
    d.scaledDetectionRadius = p->scaledDetectionRadius;
    d.radius = p->radius;
    d.accelerationRatio = p->accelerationRatio;
    d.barrierThickness = p->barrierThickness;
    d.detectionScaler = p->detectionScaler;
    d.useStepSizeLimit = p->useStepSizeLimit;
    d.useVelocityRadius = p->useVelocityRadius;
    d.n = n;
    return d;
}

static int upload_params(yb_sim* s, const yb_params* p) {
    DevParams d = make_dev_params(p, s->n);
    d.skinSweep = s->skinSweep;
    d.skinMargin = s->skinMargin * p->detectionRadius;
    d.skinCap = kSkinCapRadii * p->detectionRadius;
    HIPCHECK(hipMemcpyAsync(s->b.params, &d, sizeof(DevParams), hipMemcpyHostToDevice, s->stream));
    // The source is a stack object: wait for the (tiny) copy before returning.
    HIPCHECK(hipStreamSynchronize(s->stream));
    return YB_OK;
}

// Profiling hook: when s->prof is set (eager, uncaptured enqueue), an event
// is recorded after every kernel with the stage it belongs to.
struct ProfState {
    std::vector<hipEvent_t> ev;
    std::vector<int> stage;
    std::vector<int> count;  // launches between the previous event and this one
    std::vector<int> tag;    // first sweep index of a run of sweeps / residuals, else -1
    int used = 0;
    int curTag = -1;
    int doneAt = 0x7fffffff;  // runs reaching past the converged sweep are not harvested
    int runLen = 1;           // sweeps per tagged run
    int held = 0, heldStage = -1;
};

// Stage timing for yb_profile: an event after each stage.  A run of
// back-to-back launches of one stage (the sweeps between two residual checks)
// is timed by one interval: hold() counts a launch without an event, the next
// mark() of that stage closes the run.  No marker sits between the launches of
// a run, since each marker's release fence would write back the L2 between
// kernels that otherwise run back to back.  stage -1 restarts the clock after
// host-side waits and copies.
static void hold(yb_sim* s, int stage) {
    ProfState* ps = (ProfState*)s->prof;
    if (!ps) return;
    ps->heldStage = stage;
    ps->held++;
}

static void mark(yb_sim* s, int stage);
// Closes a run of held launches with an event of their own stage.  It is
// called before the launches of sweep curTag, so the run ends at sweep
// curTag - 1; mark() counts one launch of its own, which a flush does not add.
static void flush_held(yb_sim* s) {
    ProfState* ps = (ProfState*)s->prof;
    if (!ps || !ps->held) return;
    const int tag = ps->curTag;
    if (tag >= 0) ps->curTag = tag - 1;
    ps->held--;
    mark(s, ps->heldStage);
    ps->curTag = tag;
}

static void mark(yb_sim* s, int stage) {
    ProfState* ps = (ProfState*)s->prof;
    if (!ps) return;
    int count = 1;
    if (ps->held) {
        if (ps->heldStage == stage) count += ps->held;
        ps->held = 0;
    }
    if (ps->used == (int)ps->ev.size()) {
        hipEvent_t e;
        (void)hipEventCreate(&e);
        ps->ev.push_back(e);
        ps->stage.push_back(0);
        ps->count.push_back(0);
        ps->tag.push_back(-1);
    }
    ps->stage[ps->used] = stage;
    ps->count[ps->used] = count;
    ps->tag[ps->used] = ps->curTag;
    (void)hipEventRecord(ps->ev[ps->used++], s->stream);
}

// Substep prologue in detection mode m (0 none, 1 refit, 2 rebuild):
// initItr, detectCollisions, recomputeStepLimit (step.cpp:46-51, 15-16).
// warm: the warm start of this prologue (0 plain, 1 learned, 2 temporal); a
// plain prologue of a learned session still writes the plain start for the
// end-of-substep correction that the next learned substep reads.
static void enqueue_prologue(yb_sim* s, int mode, bool adaptive, int warm) {
    // Residual-targeted mode: the Newton residual's denominator is the force
    // at the plain warm start (now in dx[0], with the previous substep's
    // frames) for every warm start, so that all of them are held to the same
    // target.  A plain session gets it from sweep 0; any other session takes
    // it here, before the start is changed.
    // With a plain start this substep, the frame refresh below is made by the
    // same pass (both read the plain warm start and the previous frames).
    // In a strand session the first strand step joins that pass too, unless
    // features are recorded first; the plain start a learned session keeps
    // for its correction is then copied by the same pass.
    // With the Newton residual asked for at plain starts only (and the update
    // residual as the stopping test), a substep of a single handle that starts
    // elsewhere skips that pass and records no Newton residual.
    const bool refresh = (warm != 0 || s->strandSolve) && s->frameRefresh;
    const bool noNewton = warm != 0 && s->newtonMode == YB_NEWTON_PLAIN_STARTS &&
                          s->residualKind == YB_RESIDUAL_UPDATE && s->numScenes <= 1;
    // the residual schedule's solver state is reset by the first kernel
    launch_init(s, adaptive ? ctl_reset_args(s, noNewton) : CtlReset{});
    mark(s, YB_STAGE_INIT);
    if (mode >= 1) {
        launch_aabb(s, mode == 2);  // a new tree: every segment walks
        mark(s, YB_STAGE_AABB);
        if (mode == 2) {
            launch_rebuild(s);
            mark(s, YB_STAGE_REBUILD);
        }
        launch_refit(s);
        mark(s, YB_STAGE_REFIT);
        launch_query(s);  // also the step limit, from the distances it measures
        mark(s, YB_STAGE_QUERY);
    } else {
        launch_step_limit(s);
        mark(s, YB_STAGE_STEP_LIMIT);
    }
    const bool evalDenom = adaptive && (s->warmStart != 0 || s->strandSolve) && !noNewton;
    const bool fused = evalDenom && refresh && warm == 0;
    const bool learned = warm == 1, record = s->recCap > 0;
    const bool lineStart = fused && s->strandSolve && !record;
    // a learned start's apply refreshes the frames in the same pass (into
    // q[1]: the strand step that follows reads them there and copies them)
    const bool applyFrames = learned && refresh && !fused;
    const bool deferCopy = applyFrames && s->strandSolve;
    if (adaptive) {
        if (lineStart)
            launch_line_start(s, s->warmStart == 1);
        else if (evalDenom)
            launch_residual_denom(s, fused);
        mark(s, YB_STAGE_RESIDUAL);
    }
    if (learned || record || (s->warmStart == 1 && !lineStart)) {
        // the inference kernel computes its own features; the recording
        // kernel runs for training data and for the plain warm start alone
        if (record || !learned) launch_features(s, record);
        if (learned) launch_mlp_apply(s, applyFrames, !deferCopy);
        mark(s, YB_STAGE_MLP);
    } else if (warm == 2) {
        launch_temporal(s);
        mark(s, YB_STAGE_MLP);
    }
    // The frames are brought in line with the new warm start before a strand
    // step: the step solves for dx with the frames held, so frames left at
    // the previous substep would hold every segment to its old direction and
    // undo the rotation the warm start predicts (the sweeps restore it only
    // slowly where bending is soft).
    if (refresh && !fused && !applyFrames) {
        launch_frame_refresh(s);
        mark(s, YB_STAGE_MLP);
    }
    if (s->strandSolve && !lineStart) {
        launch_strand_solve(s, 0, nullptr, 0, nullptr, deferCopy ? 1 : -1, deferCopy);
        mark(s, YB_STAGE_MLP);
    }
}

// In residual-targeted mode the residual is evaluated after every
// s->checkEvery-th sweep, after the last allowed one, and -- in a plain
// session, whose prologue leaves the Newton-residual denominator to it --
// after sweep 0, whose input is the plain warm start.
static bool checks_residual(const yb_sim* s, int k) {
    const bool plainStart = s->warmStart == 0 && !s->strandSolve;
    return (k == 0 && plainStart) || (k + 1) % s->checkEvery == 0 || k == s->maxSweeps - 1;
}

int yb::first_check(const yb_sim* s) {
    for (int k = 0; k < s->maxSweeps; k++)
        if (checks_residual(s, k)) return k;
    return s->maxSweeps;
}

// The residual of a check sweep is reduced, and convergence decided, either by
// a launch of its own right after it (k_residual; per scene in a batch) or by
// the next sweep's first wave (separate_checks); the last sweep a substep may
// run is always checked by a launch of its own.
// last: the final sweep of a run that yb_profile times as one interval.
// A strand step between sweeps k - 1 and k (residual mode, every
// s->strandEvery sweeps) works on sweep k's input buffers.  Every kernel of
// the step does nothing once the solve has converged: a converged iterate is
// never changed.
// Steps at iterations first, first + every, ...; at most kStrandSteps per
// substep: a substep that has not converged by then finishes with plain sweeps.
constexpr int kStrandSteps = 16;
static void enqueue_sweep(yb_sim* s, int k, bool adaptive, bool last = false) {
    const bool sep = adaptive && separate_checks(s);
    const bool check = adaptive && checks_residual(s, k);
    bool checkPrev = adaptive && !sep && k > 0 && checks_residual(s, k - 1);
    const int first = s->strandFirst > 0 ? s->strandFirst : s->strandEvery;
    if (adaptive && s->strandSolve && s->strandEvery > 0 && k >= first &&
        (k - first) % s->strandEvery == 0 && (k - first) / s->strandEvery < kStrandSteps) {
        flush_held(s);
        if (checkPrev) {  // the step must not change a converged iterate
            launch_residual(s, k - 1);
            mark(s, YB_STAGE_RESIDUAL);
            checkPrev = false;
        }
        // no frame refresh: the sweep that follows updates the frames from the
        // new dx (a refresh here changed no sweep count, only the time)
        launch_strand_solve(s, k & 1, s->b.ctl, k);
        mark(s, YB_STAGE_MLP);
    }
    launch_sweep(s, k, adaptive, check, checkPrev);
    const bool standalone = check && (sep || k == s->maxSweeps - 1);
    if (last || standalone)
        mark(s, YB_STAGE_SWEEP);
    else
        hold(s, YB_STAGE_SWEEP);
    if (standalone) {
        launch_residual(s, k);
        mark(s, YB_STAGE_RESIDUAL);
    }
}

static void enqueue_end(yb_sim* s, int fin, bool adaptive) {
    // the converged correction: in the end pass, or (recording training data)
    // by the kernel that also writes the targets
    const bool record = s->recCap > 0;
    launch_end(s, fin, adaptive, s->warmStart != 0 && !record);
    if (record) launch_correction(s, fin, adaptive, true);
    mark(s, YB_STAGE_END);
}

static int capture(yb_sim* s, hipGraphExec_t* out, const std::function<void()>& body) {
    if (*out) {
        HIPCHECK(hipGraphExecDestroy(*out));
        *out = nullptr;
    }
    hipGraph_t g;
    HIPCHECK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    body();
    HIPCHECK(hipStreamEndCapture(s->stream, &g));
    hipError_t e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPCHECK(e);
    return YB_OK;
}

// Fixed-sweep mode: one graph per detection mode holds the whole substep
// (the reference's stepGraph plus initItr and detection, step.cpp:5-34).
static int rebuild_graphs(yb_sim* s, int numItr) {
    if (s->graphItr == numItr && s->graphs[0]) return YB_OK;
    s->graphItr = -1;
    HIPCHECK(hipStreamSynchronize(s->stream));
    for (int m = 0; m < 3; m++) {
        int rc = capture(s, &s->graphs[m], [&]() {
            enqueue_prologue(s, m, false, s->warmStart);
            for (int k = 0; k < numItr; k++) enqueue_sweep(s, k, false);
            enqueue_end(s, numItr & 1, false);
        });
        if (rc != YB_OK) return rc;
    }
    s->graphItr = numItr;
    return YB_OK;
}

// The residual-targeted schedule on a batch keeps each tile's residual
// partials for at most two scenes.
static int check_scene_sizes(const yb_sim* s) {
    for (int k = 0; k < s->numScenes && s->numScenes > 1; k++)
        if (s->sceneOffHost[k + 1] - s->sceneOffHost[k] < kMinSceneForResidual) {
            set_message("residual-targeted batches need scenes of at least 64 vertices (scene " +
                        std::to_string(k) + " has " +
                        std::to_string(s->sceneOffHost[k + 1] - s->sceneOffHost[k]) + ")");
            return YB_ERR_INVALID_ARGUMENT;
        }
    return YB_OK;
}

// probe gaps of the substep-level warm-start chooser (choose_start below)
constexpr int kProbeMin = 8, kProbeMax = 64;
// sweeps of margin on the convergence estimate before a look-ahead chunk is skipped
constexpr double kLookAheadMargin = 4.0;

// Residual-targeted mode: prologue graphs, chunk graphs of s->chunk sweeps
// with the residual reduction after each dx sweep, and the end graph.
static int rebuild_adaptive_graphs(yb_sim* s) {
    if (s->adaptiveReady) return YB_OK;
    HIPCHECK(hipStreamSynchronize(s->stream));
    const bool twoStarts = s->warmStart == 1 && s->fallback;
    for (int w = 0; w < 2; w++)
        for (int m = 0; m < 3; m++) {
            if (w == 1 && !twoStarts) {
                if (s->proGraphs[w][m]) (void)hipGraphExecDestroy(s->proGraphs[w][m]);
                s->proGraphs[w][m] = nullptr;
                continue;
            }
            int rc = capture(s, &s->proGraphs[w][m],
                             [&]() { enqueue_prologue(s, m, true, w == 0 ? s->warmStart : 0); });
            if (rc != YB_OK) return rc;
        }
    s->avgSweeps[0] = s->avgSweeps[1] = 0.0;
    s->sweepEstimate = 0.0;
    s->chooserStep = 0;
    s->probeGap = kProbeMin;
    s->nextProbe = kProbeMin;
    if (s->chunkGraphs) {
        for (int c = 0; c < s->numChunks; c++)
            if (s->chunkGraphs[c]) (void)hipGraphExecDestroy(s->chunkGraphs[c]);
        delete[] s->chunkGraphs;
    }
    s->numChunks = (s->maxSweeps + s->chunk - 1) / s->chunk;
    s->chunkGraphs = new hipGraphExec_t[s->numChunks]();
    for (int c = 0; c < s->numChunks; c++) {
        const int k0 = c * s->chunk;
        const int k1 = std::min(k0 + s->chunk, s->maxSweeps);
        int rc = capture(s, &s->chunkGraphs[c], [&]() {
            for (int k = k0; k < k1; k++) enqueue_sweep(s, k, true);
        });
        if (rc != YB_OK) return rc;
    }
    int rc = capture(s, &s->endGraph, [&]() { enqueue_end(s, 0, true); });
    if (rc != YB_OK) return rc;
    s->adaptiveReady = true;
    return YB_OK;
}

extern "C" int yb_set_solver(yb_sim* s, float residualTol, int maxSweeps, int chunk,
                             int checkEvery) {
    if (!s) return YB_ERR_INVALID_ARGUMENT;
    if (residualTol > 0 && (maxSweeps < 1 || chunk < 1 || checkEvery < 1)) {
        set_message("yb_set_solver: maxSweeps, chunk and checkEvery must be >= 1");
        return YB_ERR_INVALID_ARGUMENT;
    }
    if (residualTol != s->tol || maxSweeps != s->maxSweeps || chunk != s->chunk ||
        checkEvery != s->checkEvery)
        s->adaptiveReady = false;
    s->tol = residualTol;
    s->maxSweeps = maxSweeps;
    s->chunk = chunk;
    s->checkEvery = checkEvery;
    return YB_OK;
}

extern "C" int yb_debug_set_query_stack(yb_sim* s, int entries) {
    if (!s || entries < 1 || entries > YB_QUERY_STACK_DEPTH) {
        set_message("yb_debug_set_query_stack: 1 <= entries <= YB_QUERY_STACK_DEPTH");
        return YB_ERR_INVALID_ARGUMENT;
    }
    HIPCHECK(hipStreamSynchronize(s->stream));
    s->queryStack = entries;
    s->graphItr = -1;  // the captured queries carry the old limit
    s->adaptiveReady = false;
    return YB_OK;
}

extern "C" int yb_debug_timeline(yb_sim* s, int maxSubsteps) {
    if (!s || maxSubsteps < 0) {
        set_message("yb_debug_timeline: maxSubsteps >= 0");
        return YB_ERR_INVALID_ARGUMENT;
    }
    HIPCHECK(hipStreamSynchronize(s->stream));
    timeline_free(s);
    if (maxSubsteps > 0) {
        Timeline* t = new Timeline();
        t->cap = maxSubsteps;
        s->timeline = t;
    }
    return YB_OK;
}

extern "C" int yb_debug_timeline_read(yb_sim* s, float* out, int maxRows, int* rows) {
    if (!s || !out || !rows || maxRows < 0) {
        set_message("yb_debug_timeline_read: null argument");
        return YB_ERR_INVALID_ARGUMENT;
    }
    *rows = 0;
    Timeline* t = (Timeline*)s->timeline;
    if (!t) return YB_OK;
    HIPCHECK(hipStreamSynchronize(s->stream));
    auto dt = [&](int a, int b, float* us) -> int {
        float ms = 0.f;
        HIPCHECK(hipEventElapsedTime(&ms, t->ev[a], t->ev[b]));
        *us = 1e3f * ms;
        return YB_OK;
    };
    const int count = std::min((int)t->rows.size() - (t->open ? 1 : 0), maxRows);
    for (int r = 0; r < count; r++) {
        const TimelineRow& row = t->rows[r];
        const int e0 = row.first, e1 = e0 + 1, eLast = e1 + row.chunks, eEnd = eLast + 1;
        const int eConv = row.conv >= 0 ? e1 + 1 + row.conv : eLast;
        float* o = out + 8 * (size_t)r;
        o[0] = 0.f;
        int rc = r > 0 ? dt(t->rows[r - 1].first + t->rows[r - 1].chunks + 2, e0, &o[0]) : YB_OK;
        if (rc == YB_OK) rc = dt(e0, e1, &o[1]);
        if (rc == YB_OK) rc = dt(e1, eConv, &o[2]);
        if (rc == YB_OK) rc = dt(eConv, eLast, &o[3]);
        if (rc == YB_OK) rc = dt(eLast, eEnd, &o[4]);
        if (rc != YB_OK) return rc;
        o[5] = (float)row.done;
        o[6] = (float)row.chunks;
        o[7] = (float)row.conv;
    }
    *rows = count;
    return YB_OK;
}

extern "C" int yb_set_residual_kind(yb_sim* s, int kind) {
    if (!s || (kind != YB_RESIDUAL_UPDATE && kind != YB_RESIDUAL_NEWTON)) {
        set_message("yb_set_residual_kind: YB_RESIDUAL_UPDATE or YB_RESIDUAL_NEWTON");
        return YB_ERR_INVALID_ARGUMENT;
    }
    s->residualKind = kind;
    const int32_t k = kind;
    HIPCHECK(hipMemcpyAsync(&s->b.ctl->kind, &k, sizeof(k), hipMemcpyHostToDevice, s->stream));
    HIPCHECK(hipStreamSynchronize(s->stream));
    s->adaptiveReady = false;  // the prologue graphs depend on it (yb_set_newton_residual)
    return YB_OK;
}

extern "C" int yb_set_newton_residual(yb_sim* s, int mode) {
    if (!s || (mode != YB_NEWTON_ALL_STARTS && mode != YB_NEWTON_PLAIN_STARTS)) {
        set_message("yb_set_newton_residual: YB_NEWTON_ALL_STARTS or YB_NEWTON_PLAIN_STARTS");
        return YB_ERR_INVALID_ARGUMENT;
    }
    HIPCHECK(hipStreamSynchronize(s->stream));
    s->newtonMode = mode;
    s->adaptiveReady = false;
    return YB_OK;
}

extern "C" int yb_set_strand_schedule(yb_sim* s, int every, int pairIters, int first) {
    if (!s || every < 0 || pairIters < 1 || pairIters > 64 || first < 0) {
        set_message("yb_set_strand_schedule: every >= 0, 1 <= pairIters <= 64, first >= 0");
        return YB_ERR_INVALID_ARGUMENT;
    }
    if (every != s->strandEvery || pairIters != s->strandIters || first != s->strandFirst) {
        s->strandEvery = every;
        s->strandIters = pairIters;
        s->strandFirst = first;
        s->graphItr = -1;
        s->adaptiveReady = false;
    }
    return YB_OK;
}

extern "C" int yb_set_launch_config(yb_sim* s, int sweepThreads, int queryThreads) {
    if (!s) return YB_ERR_INVALID_ARGUMENT;
    const bool okS = sweepThreads == 64 || sweepThreads == 128 || sweepThreads == 256 ||
                     sweepThreads == 512;
    const bool okQ = queryThreads == 64 || queryThreads == 128 || queryThreads == 256;
    if (!okS || !okQ) {
        set_message("yb_set_launch_config: sweep threads 64/128/256/512, query threads 64/128/256");
        return YB_ERR_INVALID_ARGUMENT;
    }
    HIPCHECK(hipStreamSynchronize(s->stream));
    s->sweepThreads = sweepThreads;
    s->queryThreads = queryThreads;
    s->graphItr = -1;
    s->adaptiveReady = false;
    return YB_OK;
}

extern "C" int yb_get_launch_config(yb_sim* s, int* sweepThreads, int* queryThreads) {
    if (!s || !sweepThreads || !queryThreads) return YB_ERR_INVALID_ARGUMENT;
    *sweepThreads = s->sweepThreads;
    *queryThreads = s->queryThreads;
    return YB_OK;
}

extern "C" int yb_set_tile_config(yb_sim* s, int refitLeaves, int queryStackLds,
                                  int queryCandidates) {
    if (!s) return YB_ERR_INVALID_ARGUMENT;
    const bool okR = refitLeaves == 256 || refitLeaves == 512 || refitLeaves == 1024;
    const bool okQ = query_shape_supported(queryStackLds, queryCandidates);
    if (!okR || !okQ) {
        set_message("yb_set_tile_config: refit leaves 256/512/1024; query (LDS stack, "
                    "candidates) one of (4,4) (8,6) (8,8) (16,8)");
        return YB_ERR_INVALID_ARGUMENT;
    }
    HIPCHECK(hipStreamSynchronize(s->stream));
    if (refitLeaves != s->refitLeaves) {
        set_refit_geometry(s, refitLeaves);
        s->treeBuilt = false;  // the tree's straddle flags follow the refit tile
    }
    s->queryStackLds = queryStackLds;
    s->queryCandidates = queryCandidates;
    s->graphItr = -1;
    s->adaptiveReady = false;
    return YB_OK;
}

extern "C" int yb_get_tile_config(yb_sim* s, int* refitLeaves, int* queryStackLds,
                                  int* queryCandidates) {
    if (!s || !refitLeaves || !queryStackLds || !queryCandidates)
        return YB_ERR_INVALID_ARGUMENT;
    *refitLeaves = s->refitLeaves;
    *queryStackLds = s->queryStackLds;
    *queryCandidates = s->queryCandidates;
    return YB_OK;
}

extern "C" int yb_set_contact_skin(yb_sim* s, float margin, float sweep) {
    if (!s) return YB_ERR_INVALID_ARGUMENT;
    if (!(sweep >= 0.f && sweep <= 1e3f) || !(margin <= 1e3f)) {
        set_message("yb_set_contact_skin: sweep must lie in [0, 1000], margin at most 1000");
        return YB_ERR_INVALID_ARGUMENT;
    }
    s->skinSet = true;
    return set_skin(s, margin, sweep);
}

// Applies a contact-skin setting; a new on/off state drops the kept lists
// and the per-scene policy (and recaptures the graphs).
static int set_skin(yb_sim* s, float margin, float sweep) {
    HIPCHECK(hipStreamSynchronize(s->stream));
    const bool was = s->skinMargin > 0.f, on = margin > 0.f;
    s->skinMargin = on ? margin : 0.f;
    s->skinSweep = sweep;
    if (on != was) {
        // kept candidates from an earlier period may follow an older tree's
        // leaf order: every segment walks first
        HIPCHECK(hipMemsetAsync(s->b.candN, 0xff, sizeof(int32_t) * (size_t)s->n, s->stream));
        HIPCHECK(hipMemsetAsync(s->b.skinPolicy, 0, sizeof(int4) * (size_t)kMaxScenes, s->stream));
        HIPCHECK(hipStreamSynchronize(s->stream));
        s->graphItr = -1;
        s->adaptiveReady = false;
    }
    return YB_OK;
}

extern "C" int yb_get_contact_skin(yb_sim* s, float* margin, float* sweep) {
    if (!s || !margin || !sweep) return YB_ERR_INVALID_ARGUMENT;
    *margin = s->skinMargin;
    *sweep = s->skinSweep;
    return YB_OK;
}

extern "C" int yb_contact_skin_stats(yb_sim* s, int64_t out[2], int reset) {
    if (!s || !out) return YB_ERR_INVALID_ARGUMENT;
    unsigned long long c[2];
    HIPCHECK(hipMemcpyAsync(c, s->b.skinStats, sizeof(c), hipMemcpyDeviceToHost, s->stream));
    HIPCHECK(hipStreamSynchronize(s->stream));
    out[0] = (int64_t)c[0];
    out[1] = (int64_t)c[1];
    if (reset) {
        HIPCHECK(hipMemsetAsync(s->b.skinStats, 0, sizeof(c), s->stream));
        HIPCHECK(hipStreamSynchronize(s->stream));
    }
    return YB_OK;
}

extern "C" int yb_debug_read_skin(yb_sim* s, int32_t* candN, float* sbox, int32_t* policy) {
    if (!s) return YB_ERR_INVALID_ARGUMENT;
    const size_t N = (size_t)s->n;
    if (candN)
        HIPCHECK(hipMemcpyAsync(candN, s->b.candN, sizeof(int32_t) * N, hipMemcpyDeviceToHost,
                                s->stream));
    if (sbox)
        HIPCHECK(hipMemcpyAsync(sbox, s->b.sbox, sizeof(float4) * 2 * N, hipMemcpyDeviceToHost,
                                s->stream));
    if (policy)
        HIPCHECK(hipMemcpyAsync(policy, s->b.skinPolicy, sizeof(int4) * (size_t)s->numScenes,
                                hipMemcpyDeviceToHost, s->stream));
    HIPCHECK(hipStreamSynchronize(s->stream));
    return YB_OK;
}

extern "C" int yb_set_graphs(yb_sim* s, int on) {
    if (!s) return YB_ERR_INVALID_ARGUMENT;
    HIPCHECK(hipStreamSynchronize(s->stream));
    s->useGraphs = on ? 1 : 0;
    s->graphItr = -1;
    s->adaptiveReady = false;
    return YB_OK;
}

extern "C" int yb_solver_stats(yb_sim* s, yb_solver_stats_t* out, int reset) {
    if (!s || !out) return YB_ERR_INVALID_ARGUMENT;
    SolverCtl c;
    HIPCHECK(hipMemcpyAsync(&c, s->b.ctl, sizeof(c), hipMemcpyDeviceToHost, s->stream));
    HIPCHECK(hipStreamSynchronize(s->stream));
    out->substeps = (int64_t)c.substeps;
    out->sweeps = (int64_t)c.sweeps;
    out->maxSweepsUsed = c.maxUsed;
    out->notConverged = c.notConverged;
    out->lastResidual = c.resid;
    out->maxResidual = c.maxResid;
    out->lastNewtonResidual = c.fres;
    out->maxNewtonResidual = c.maxFres;
    out->sumNewtonResidual = c.sumFres;
    out->newtonSubsteps = (int64_t)c.newtonSubsteps;
    out->sceneSubsteps = out->substeps;
    out->sceneSweeps = out->sweeps;
    if (s->numScenes > 1 && s->b.sceneCtl) {
        // a batch: every scene converges once per substep, at its own sweep
        std::vector<SceneCtl> sc(s->numScenes);
        HIPCHECK(hipMemcpyAsync(sc.data(), s->b.sceneCtl, sizeof(SceneCtl) * sc.size(),
                                hipMemcpyDeviceToHost, s->stream));
        HIPCHECK(hipStreamSynchronize(s->stream));
        int64_t sw = 0;
        double fr = 0.0;
        for (const SceneCtl& k : sc) {
            sw += (int64_t)k.sweeps;
            fr += k.sumFres;
        }
        out->sceneSubsteps = out->substeps * s->numScenes;
        out->sceneSweeps = sw;
        out->sumNewtonResidual = fr;
        out->newtonSubsteps = out->sceneSubsteps;
    }
    if (reset) {
        SolverCtl z;
        memset(&z, 0, sizeof(z));
        z.doneAt = 0x7fffffff;
        z.denom = s->denomHost;
        z.denomF = -1.0;
        z.kind = s->residualKind;
        HIPCHECK(hipMemcpyAsync(s->b.ctl, &z, sizeof(z), hipMemcpyHostToDevice, s->stream));
        HIPCHECK(hipStreamSynchronize(s->stream));
        const int rc = reset_scene_ctl(s);
        if (rc != YB_OK) return rc;
    }
    return YB_OK;
}

// One residual-targeted substep: prologue, chunks until converged (the host
// reads the convergence flag after chunk c while chunk c+1 already runs; a
// chunk launched after convergence exits immediately), end.
// Warm start of the next substep of a learned session with the fallback on:
// 0 = learned, 1 = plain.  The start with the lower running sweep average is
// used; the other one is probed after s->probeGap substeps, a gap that starts
// at kProbeMin and doubles (up to kProbeMax) each time the probe loses.
static int best_start(const yb_sim* s) {
    return s->avgSweeps[0] + s->learnedCost <= s->avgSweeps[1] ? 0 : 1;
}
static int choose_start(const yb_sim* s) {
    if (!(s->warmStart == 1 && s->fallback)) return 0;
    if (s->avgSweeps[0] == 0.0) return 0;  // first a learned sample,
    if (s->avgSweeps[1] == 0.0) return 1;  // then a plain one
    const int best = best_start(s);
    return s->chooserStep >= s->nextProbe ? 1 - best : best;
}

static void update_chooser(yb_sim* s, int start, int sweeps) {
    if (!(s->warmStart == 1 && s->fallback)) return;
    const bool ready = s->avgSweeps[0] != 0.0 && s->avgSweeps[1] != 0.0;
    const bool probe = ready && start != best_start(s);
    // Sweep counts drift within a frame, so a probe's sample replaces the old
    // value of its start; the start in use keeps a running average (weight 1/2
    // on the newest sample) over its consecutive substeps.
    double& avg = s->avgSweeps[start];
    avg = (avg == 0.0 || probe) ? (double)sweeps : 0.5 * (avg + (double)sweeps);
    s->chooserStep++;
    if (probe || !ready) {
        const bool won = s->avgSweeps[1] != 0.0 && best_start(s) == start;
        s->probeGap = won ? kProbeMin : std::min(2 * s->probeGap, kProbeMax);
        s->nextProbe = s->chooserStep + s->probeGap;
    }
}

static int adaptive_substep(yb_sim* s, int mode) {
    const int start = choose_start(s);
    Timeline* tl = (Timeline*)s->timeline;
    if (tl && (int)tl->rows.size() < tl->cap) {
        tl->rows.emplace_back();
        tl->rows.back().first = (int)tl->used;
        tl->open = true;
    }
    int rc = timeline_mark(s);
    if (rc != YB_OK) return rc;
    if (s->useGraphs)
        HIPCHECK(hipGraphLaunch(s->proGraphs[start][mode], s->stream));
    else
        enqueue_prologue(s, mode, true, start == 0 ? s->warmStart : 0);
    rc = timeline_mark(s);
    if (rc != YB_OK) return rc;
    const int numChunks = (s->maxSweeps + s->chunk - 1) / s->chunk;
    int launched = 0;
    int done = s->maxSweeps;  // sweeps of this substep (all of them if not converged)
    bool converged = false;
    auto launch_chunk = [&](int c) -> int {
        if (s->useGraphs) {
            HIPCHECK(hipGraphLaunch(s->chunkGraphs[c], s->stream));
        } else {
            const int k1 = std::min((c + 1) * s->chunk, s->maxSweeps);
            for (int k = c * s->chunk; k < k1; k++) enqueue_sweep(s, k, true);
        }
        HIPCHECK(hipMemcpyAsync(&s->b.hostFlag[c & 1], &s->b.ctl->doneAt, sizeof(int32_t),
                                hipMemcpyDeviceToHost, s->stream));
        HIPCHECK(hipEventRecord(s->chunkEv[c & 1], s->stream));
        launched = c + 1;
        return timeline_mark(s);
    };
    // The next chunk is queued before the host waits for the current one, so
    // the device does not idle between chunks -- unless the recent substeps
    // say this one converges within the current chunk: a chunk queued after
    // convergence is a run of launches that exit at once (about 5 us each for
    // a sweep).  If the estimate was wrong, the next chunk is queued after the
    // wait.  This changes only when launches are queued, never the results.
    rc = launch_chunk(0);
    if (rc != YB_OK) return rc;
    for (int c = 0; c < numChunks; c++) {
        const bool expectDone =
            s->sweepEstimate > 0.0 && s->sweepEstimate + kLookAheadMargin <= (c + 1) * s->chunk;
        if (launched < numChunks && launched == c + 1 && !expectDone) {
            rc = launch_chunk(c + 1);
            if (rc != YB_OK) return rc;
        }
        HIPCHECK(hipEventSynchronize(s->chunkEv[c & 1]));
        if (s->b.hostFlag[c & 1] != 0x7fffffff) {
            done = s->b.hostFlag[c & 1] + 1;
            converged = true;
            break;
        }
        if (launched < numChunks && launched == c + 1) {
            rc = launch_chunk(c + 1);
            if (rc != YB_OK) return rc;
        }
    }
    // The last substep's count: consecutive substeps converge in similar
    // counts, and an estimate biased upwards (one that rose at once to each
    // new maximum) kept a look-ahead chunk of launches that exit at once
    // queued after most substeps (3-5% of the time on the 1M knit).
    s->sweepEstimate = (double)done;
    if (s->useGraphs)
        HIPCHECK(hipGraphLaunch(s->endGraph, s->stream));
    else
        enqueue_end(s, 0, true);
    HIPCHECK(hipGetLastError());
    rc = timeline_mark(s);
    if (rc != YB_OK) return rc;
    if (tl && tl->open) {
        TimelineRow& row = tl->rows.back();
        row.chunks = launched;
        row.done = done;
        row.conv = converged ? (done - 1) / s->chunk : -1;
        tl->open = false;
    }
    s->warmSubsteps[start]++;
    update_chooser(s, start, done);
    return YB_OK;
}

static int check_errors(yb_sim* s, int* warning) {
    int32_t err[2] = {0, 0};
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(err, s->b.err, sizeof(err), hipMemcpyDeviceToHost, s->stream));
    HIPCHECK(hipStreamSynchronize(s->stream));
    if (err[0] != 0 || err[1] != 0) {
        launch_fill_i32(s, s->b.err, 0, 2);
        HIPCHECK(hipStreamSynchronize(s->stream));
    }
    if (warning) *warning = err[1];
    // the reference's error (YarnBall.cu:283-286) takes precedence
    if (err[0] & kErrBitContacts) {
        set_message((err[0] & kErrBitStack)
                        ? "MAX_COLLISIONS_PER_SEGMENT exceeded (and a broadphase query ran out "
                          "of traversal stack)"
                        : "MAX_COLLISIONS_PER_SEGMENT exceeded");
        return YB_ERROR_MAX_COLLISIONS_PER_SEGMENT_EXCEEDED;
    }
    if (err[0] & kErrBitStack) {
        set_message("broadphase traversal stack exceeded: contacts may be missing");
        return YB_ERROR_TRAVERSAL_STACK_EXCEEDED;
    }
    if (err[0] != 0) {
        set_message("Indescript error");  // YarnBall.cu:289
        return YB_ERR_HIP;
    }
    return YB_OK;
}

// Host-side decision of the detection mode for the current substep
// (step.cpp:48-49, collision.cu:101-108).
static int next_mode(yb_sim* s, const yb_params* p) {
    if (!(p->detectionPeriod > 0 && s->stepCounter % p->detectionPeriod == 0)) return 0;
    if (!s->treeBuilt || s->lastBVHRebuild >= p->bvhRebuildPeriod) {
        s->lastBVHRebuild = 0.f;
        s->treeBuilt = true;
        return 2;
    }
    s->lastBVHRebuild += p->h * p->detectionPeriod;
    return 1;
}

extern "C" int yb_advance(yb_sim* s, yb_params* p, float h, float maxH, int* warning) {
    if (warning) *warning = 0;
    if (!s || !p) {
        set_message("yb_advance: null argument");
        return YB_ERR_INVALID_ARGUMENT;
    }
    if (!(h > 0)) return YB_OK;  // step.cpp:37
    if (!(maxH > 0)) {
        set_message("yb_advance: maxH must be positive");
        return YB_ERR_INVALID_ARGUMENT;
    }
    if (p->numItr < 0) {
        set_message("yb_advance: numItr must be >= 0");
        return YB_ERR_INVALID_ARGUMENT;
    }
    int steps = (int)ceilf(h / maxH);
    if (steps < 1) steps = 1;
    p->lastH = p->h;
    p->h = h / (float)steps;
    int rc = yb_validate_params(p);
    if (rc != YB_OK) return rc;
    if (s->tol <= 0 && s->useGraphs) {
        rc = rebuild_graphs(s, p->numItr);
        if (rc != YB_OK) return rc;
    }
    rc = upload_params(s, p);
    if (rc != YB_OK) return rc;
    const bool adaptive = s->tol > 0;
    if (adaptive) {
        rc = check_scene_sizes(s);
        if (rc != YB_OK) return rc;
    }
    if (adaptive && s->useGraphs) {
        rc = rebuild_adaptive_graphs(s);
        if (rc != YB_OK) return rc;
    }
    for (int st = 0; st < steps; st++, s->stepCounter++) {
        const int mode = next_mode(s, p);
        if (adaptive) {
            rc = adaptive_substep(s, mode);
            if (rc != YB_OK) return rc;
        } else if (s->useGraphs) {
            HIPCHECK(hipGraphLaunch(s->graphs[mode], s->stream));
        } else {
            enqueue_prologue(s, mode, false, s->warmStart);
            for (int k = 0; k < p->numItr; k++) enqueue_sweep(s, k, false);
            enqueue_end(s, p->numItr & 1, false);
            HIPCHECK(hipGetLastError());
        }
    }
    // Interpenetration check of the final state (the step limit checks every
    // substep's starting state).
    if (p->detectionPeriod > 0) {
        launch_separation(s);
        HIPCHECK(hipGetLastError());
    }
    p->time += h;
    return check_errors(s, warning);
}

extern "C" int yb_min_separation(yb_sim* s, float* out, int reset) {
    if (!s || !out) return YB_ERR_INVALID_ARGUMENT;
    HIPCHECK(hipMemcpyAsync(out, s->b.minSep, sizeof(float), hipMemcpyDeviceToHost, s->stream));
    HIPCHECK(hipStreamSynchronize(s->stream));
    if (reset) {
        const float inf = INFINITY;
        HIPCHECK(hipMemcpyAsync(s->b.minSep, &inf, sizeof(float), hipMemcpyHostToDevice, s->stream));
        HIPCHECK(hipStreamSynchronize(s->stream));
    }
    return YB_OK;
}

extern "C" int64_t yb_step_count(yb_sim* s) { return s ? s->stepCounter : 0; }

extern "C" int yb_bounds(yb_sim* s, float out[6]) {
    if (!s || !out) return YB_ERR_INVALID_ARGUMENT;
    float4 box[2];
    // with the contact skin the tree holds grown boxes and is not refit in
    // every substep: the union of the current boxes is formed here
    if (s->skinMargin > 0.f && s->treeBuilt) {
        launch_scene_box(s);
        HIPCHECK(hipGetLastError());
    }
    HIPCHECK(hipMemcpyAsync(box, s->b.sceneBox, sizeof(box), hipMemcpyDeviceToHost, s->stream));
    HIPCHECK(hipStreamSynchronize(s->stream));
    if (!s->treeBuilt) {
        box[0] = make_float4(INFINITY, INFINITY, INFINITY, 0.f);
        box[1] = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.f);
    }
    out[0] = box[0].x;
    out[1] = box[0].y;
    out[2] = box[0].z;
    out[3] = box[1].x;
    out[4] = box[1].y;
    out[5] = box[1].z;
    return YB_OK;
}

extern "C" int yb_collision_counts(yb_sim* s, int32_t* out) {
    if (!s || !out) return YB_ERR_INVALID_ARGUMENT;
    HIPCHECK(hipMemcpyAsync(out, s->b.numCols, sizeof(int32_t) * s->n, hipMemcpyDeviceToHost,
                            s->stream));
    HIPCHECK(hipStreamSynchronize(s->stream));
    return YB_OK;
}

// ---- stage-level access ------------------------------------------------------

extern "C" int yb_debug_prologue(yb_sim* s, yb_params* p, float h, int detect, int rebuild) {
    if (!s || !p) return YB_ERR_INVALID_ARGUMENT;
    p->h = h;
    int rc = yb_validate_params(p);
    if (rc != YB_OK) return rc;
    rc = upload_params(s, p);
    if (rc != YB_OK) return rc;
    int mode = 0;
    if (detect) {
        mode = (rebuild || !s->treeBuilt) ? 2 : 1;
        s->treeBuilt = true;
        if (mode == 2) s->lastBVHRebuild = 0.f;
    }
    enqueue_prologue(s, mode, false, s->warmStart);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(s->stream));
    return check_errors(s, nullptr);
}

extern "C" int yb_debug_query_counts(yb_sim* s, yb_params* p, float h, int64_t out[6]) {
    if (!s || !p || !out) return YB_ERR_INVALID_ARGUMENT;
    const size_t N = (size_t)s->n;
    if (!s->b.candCount) HIPCHECK(hipMalloc(&s->b.candCount, sizeof(int32_t) * 3 * N));
    s->countCandidates = true;
    int rc = yb_debug_prologue(s, p, h, 1, 0);
    s->countCandidates = false;
    if (rc != YB_OK) return rc;
    std::vector<int32_t> cand(3 * N), nc(N);
    std::vector<float4> prm(N);
    HIPCHECK(hipMemcpy(cand.data(), s->b.candCount, sizeof(int32_t) * 3 * N,
                       hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(nc.data(), s->b.numCols, sizeof(int32_t) * N, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(prm.data(), s->b.prm, sizeof(float4) * N, hipMemcpyDeviceToHost));
    int64_t segs = 0, c = 0, k = 0, visits = 0, deep = 0, maxDepth = 0;
    for (size_t i = 0; i < N; i++) {
        uint32_t fl;
        memcpy(&fl, &prm[i].w, sizeof(fl));
        if (fl & YB_FLAG_HAS_NEXT) segs++;
        c += cand[i];
        k += nc[i];
        visits += cand[N + i];
        deep += cand[2 * N + i] > 8 ? 1 : 0;
        maxDepth = std::max<int64_t>(maxDepth, cand[2 * N + i]);
    }
    out[0] = segs;
    out[1] = c;
    out[2] = k;
    out[3] = visits;
    out[4] = deep;
    out[5] = maxDepth;
    return YB_OK;
}

extern "C" int yb_debug_read(yb_sim* s, float* dx, float* aabb, uint32_t* morton, int32_t* order,
                             int32_t* numCols, int32_t* cols) {
    if (!s) return YB_ERR_INVALID_ARGUMENT;
    const size_t N = (size_t)s->n;
    const hipMemcpyKind k = hipMemcpyDeviceToHost;
    if (dx) HIPCHECK(hipMemcpyAsync(dx, s->b.dx[0], sizeof(float4) * N, k, s->stream));
    if (aabb) HIPCHECK(hipMemcpyAsync(aabb, s->b.aabb, sizeof(float4) * 2 * N, k, s->stream));
    if (morton) HIPCHECK(hipMemcpyAsync(morton, s->b.morton, sizeof(uint32_t) * N, k, s->stream));
    if (order) HIPCHECK(hipMemcpyAsync(order, s->b.order, sizeof(int32_t) * N, k, s->stream));
    if (numCols) HIPCHECK(hipMemcpyAsync(numCols, s->b.numCols, sizeof(int32_t) * N, k, s->stream));
    if (cols)
        HIPCHECK(hipMemcpyAsync(cols, s->b.cols, sizeof(int32_t) * N * kMaxCols, k, s->stream));
    HIPCHECK(hipStreamSynchronize(s->stream));
    return YB_OK;
}

extern "C" int yb_debug_tree(yb_sim* s, int32_t* child, int32_t* parent, float* box) {
    if (!s) return YB_ERR_INVALID_ARGUMENT;
    if (!s->treeBuilt) {
        set_message("yb_debug_tree: no tree built yet");
        return YB_ERR_INVALID_ARGUMENT;
    }
    const size_t N = (size_t)s->n;
    float4* nodes = new float4[4 * (N - 1)];
    hipError_t e = hipMemcpyAsync(nodes, s->b.nodes, sizeof(float4) * 4 * (N - 1),
                                  hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess && parent)
        e = hipMemcpyAsync(parent, s->b.parent, sizeof(int32_t) * (2 * N - 1),
                           hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) {
        delete[] nodes;
        set_message(std::string("yb_debug_tree: ") + hipGetErrorString(e));
        return YB_ERR_HIP;
    }
    if (parent)  // drop the refit-block flag: plain (node << 1) | side slots
        for (size_t i = 0; i < 2 * N - 1; i++)
            if (parent[i] >= 0) parent[i] &= ~kParentStrad;
    for (size_t i = 0; i + 1 < N; i++) {
        if (child) {
            int32_t cl, cr;
            memcpy(&cl, &nodes[4 * i].w, 4);
            memcpy(&cr, &nodes[4 * i + 2].w, 4);
            child[2 * i] = cl;
            child[2 * i + 1] = cr;
        }
        if (box) memcpy(&box[16 * i], &nodes[4 * i], sizeof(float4) * 4);
    }
    delete[] nodes;
    return YB_OK;
}

extern "C" int yb_debug_write_dx(yb_sim* s, const float* dx) {
    if (!s || !dx) return YB_ERR_INVALID_ARGUMENT;
    HIPCHECK(hipMemcpyAsync(s->b.dx[0], dx, sizeof(float4) * s->n, hipMemcpyHostToDevice, s->stream));
    HIPCHECK(hipStreamSynchronize(s->stream));
    return YB_OK;
}

extern "C" int yb_debug_read_last_vels(yb_sim* s, float* lastVels) {
    if (!s || !lastVels) return YB_ERR_INVALID_ARGUMENT;
    float4* tmp = new float4[s->n];
    hipError_t e = hipMemcpyAsync(tmp, s->b.lastVel, sizeof(float4) * s->n, hipMemcpyDeviceToHost,
                                  s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    for (int i = 0; i < s->n && e == hipSuccess; i++) {
        lastVels[3 * i] = tmp[i].x;
        lastVels[3 * i + 1] = tmp[i].y;
        lastVels[3 * i + 2] = tmp[i].z;
    }
    delete[] tmp;
    HIPCHECK(e);
    return YB_OK;
}

extern "C" int yb_debug_strand(yb_sim* s, float* system) {
    if (!s) return YB_ERR_INVALID_ARGUMENT;
    float* d = nullptr;
    if (system) HIPCHECK(hipMalloc(&d, sizeof(float) * 5 * (size_t)s->n));
    launch_strand_solve(s, 0, nullptr, 0, d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && system)
        e = hipMemcpyAsync(system, d, sizeof(float) * 5 * (size_t)s->n, hipMemcpyDeviceToHost,
                           s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (d) (void)hipFree(d);
    HIPCHECK(e);
    return YB_OK;
}

extern "C" int yb_debug_write_last_vels(yb_sim* s, const float* lastVels) {
    if (!s || !lastVels) return YB_ERR_INVALID_ARGUMENT;
    float4* tmp = new float4[s->n];
    for (int i = 0; i < s->n; i++)
        tmp[i] = make_float4(lastVels[3 * i], lastVels[3 * i + 1], lastVels[3 * i + 2], 0.f);
    hipError_t e = hipMemcpyAsync(s->b.lastVel, tmp, sizeof(float4) * s->n, hipMemcpyHostToDevice,
                                  s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    delete[] tmp;
    HIPCHECK(e);
    return YB_OK;
}

// ---- live per-stage timing ----------------------------------------------------
// Runs `substeps` substeps exactly as yb_advance schedules them, but enqueued
// eagerly with a HIP event between consecutive kernels on the handle's stream,
// and accumulates each stage's GPU time.  Stage ids are YB_STAGE_*.
extern "C" int yb_profile(yb_sim* s, yb_params* p, float h, int substeps, double* ms,
                          int32_t* launches) {
    if (!s || !p || !ms || !launches || substeps < 1) return YB_ERR_INVALID_ARGUMENT;
    for (int k = 0; k < YB_NUM_STAGES; k++) {
        ms[k] = 0.0;
        launches[k] = 0;
    }
    p->lastH = p->h;
    p->h = h;
    int rc = yb_validate_params(p);
    if (rc != YB_OK) return rc;
    rc = upload_params(s, p);
    if (rc != YB_OK) return rc;
    const bool adaptive = s->tol > 0;
    if (adaptive) {
        const int rcs = check_scene_sizes(s);
        if (rcs != YB_OK) return rcs;
    }
    ProfState ps;
    s->prof = &ps;
    // detach and free the markers on every exit path, early error returns included
    struct Detach {
        yb_sim* s;
        ProfState* ps;
        ~Detach() {
            s->prof = nullptr;
            for (auto& e : ps->ev) (void)hipEventDestroy(e);
        }
    } detach{s, &ps};
    auto harvest = [&]() -> int {
        for (int k = 1; k < ps.used; k++) {
            float t = 0.f;
            if (ps.stage[k] < 0) continue;  // gap marker: host latency, not a kernel
            // runs that reach past the converged sweep hold early-exit launches
            if (ps.tag[k] >= 0 && ps.tag[k] + ps.runLen - 1 > ps.doneAt) continue;
            HIPCHECK(hipEventElapsedTime(&t, ps.ev[k - 1], ps.ev[k]));
            ms[ps.stage[k]] += t;
            launches[ps.stage[k]] += ps.count[k];
        }
        return YB_OK;
    };
    for (int st = 0; st < substeps && rc == YB_OK; st++, s->stepCounter++) {
        const int mode = next_mode(s, p);
        ps.used = 0;
        mark(s, -1);
        enqueue_prologue(s, mode, adaptive, s->warmStart);
        int fin = p->numItr & 1;
        ps.doneAt = 0x7fffffff;
        if (adaptive) {
            // As in the stepper, sweeps run in chunks of s->chunk and the next
            // chunk is queued before the host waits for the current one's
            // flag, so the device does not idle between chunks.  Each event
            // is tagged with the sweep its interval ends at (one per chunk,
            // plus the strand steps'); the intervals that reach past the
            // converged sweep (they hold early-exit launches) are left out of
            // the harvest.
            const int ce = s->chunk;
            ps.runLen = 1;
            const int groups = (s->maxSweeps + ce - 1) / ce;
            auto enqueue_group = [&](int g) -> int {
                const int k1 = std::min((g + 1) * ce, s->maxSweeps);
                for (int k = g * ce; k < k1; k++) {
                    ps.curTag = k;
                    enqueue_sweep(s, k, true, k == k1 - 1);
                }
                ps.curTag = -1;
                HIPCHECK(hipMemcpyAsync(&s->b.hostFlag[g & 1], &s->b.ctl->doneAt, sizeof(int32_t),
                                        hipMemcpyDeviceToHost, s->stream));
                HIPCHECK(hipEventRecord(s->chunkEv[g & 1], s->stream));
                mark(s, -1);  // the readback belongs to no stage
                return YB_OK;
            };
            rc = enqueue_group(0);
            for (int g = 0; g < groups && rc == YB_OK; g++) {
                if (g + 1 < groups) rc = enqueue_group(g + 1);
                if (rc != YB_OK) break;
                HIPCHECK(hipEventSynchronize(s->chunkEv[g & 1]));
                if (s->b.hostFlag[g & 1] != 0x7fffffff) {
                    ps.doneAt = s->b.hostFlag[g & 1];
                    break;
                }
            }
            if (rc != YB_OK) break;
        } else {
            for (int k = 0; k < p->numItr; k++) enqueue_sweep(s, k, false, k == p->numItr - 1);
        }
        mark(s, -1);
        enqueue_end(s, fin, adaptive);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipStreamSynchronize(s->stream));
        rc = harvest();
    }
    if (rc != YB_OK) return rc;
    p->time += h * substeps;
    return check_errors(s, nullptr);
}
