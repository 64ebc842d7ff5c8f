// yb_internal.h — device-side state of one simulation handle and the kernel
// launchers shared between the translation units of libyarnball_hip.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/yarnball_hip.h"

namespace yb {

// Per-advance constants read by every kernel through one device pointer (the
// role MetaData* plays in the reference, YarnBall.h:63-114).  Uploaded once per
// advance outside the captured graphs, so graphs stay valid when the user edits
// kCollision, gravity, etc. between frames.
struct DevParams {
    float gx, gy, gz;
    float h;
    float lastH;
    float dampingH;  // meta.damping / h (cosserat.cu:29)
    float drag;
    float frictionCoeff;
    float kCollision;
    float kFriction;
    float detectionRadius;
    float scaledDetectionRadius;
    float radius;
    float accelerationRatio;
    float barrierThickness;
    float detectionScaler;
    int32_t useStepSizeLimit;
    int32_t useVelocityRadius;
    int32_t n;
    int32_t pad_;
    float skinSweep;   // contact skin: substeps of warm-start motion a grown box covers
    float skinMargin;  // contact skin: absolute margin on every side of a grown box
    float skinCap;     // contact skin: largest growth by motion along an axis
};

// Residual-targeted solver state (device side, one per handle).
struct SolverCtl {
    int32_t doneAt;  // sweep index at which the current substep converged (INT_MAX: running)
    int32_t fin;     // buffer (0/1) holding the final iterate
    float resid;     // relative residual of the last evaluated sweep
    int32_t maxUsed;
    unsigned long long sweeps;    // sweeps run, summed over substeps
    unsigned long long substeps;  // substeps run in this mode
    int32_t notConverged;         // substeps that hit maxSweeps above the tolerance
    float maxResid;               // largest final residual (of the stopping test)
    double denom;                 // update residual: sum over free vertices of l_i^2 (upload)
    double denomF;                // Newton residual: |f(dx_0)|^2 of the substep (< 0: unknown)
    double sumFres;               // final Newton residuals summed over substeps
    float fres;                   // Newton residual of the last evaluated sweep
    float maxFres;                // largest final Newton residual
    int32_t kind;                 // stopping test: 0 update residual, 1 Newton residual
    int32_t scenesDone;           // batched scenes: scenes converged in this substep
    unsigned long long newtonSubsteps;  // substeps with a Newton residual (sumFres, maxFres)
};
// SolverCtl::denomF of a substep that evaluates no Newton residual (a start
// other than the plain one, with the residual asked for at plain starts only);
// -1 marks a denominator still to be taken at the substep's first check.
constexpr double kNoNewton = -2.0;

// Per-scene solver state of a batch in the residual-targeted schedule: each
// scene stops at its own sweep and keeps its final iterate in its own buffer.
struct SceneCtl {
    int32_t doneAt;  // sweep at which the scene converged (INT_MAX: running)
    int32_t fin;     // buffer (0/1) holding the scene's final iterate
    double denom;    // update residual: sum over the scene's free vertices of l_i^2
    double denomF;   // Newton residual: |f(dx_0)|^2 of the scene (< 0: unknown)
    double sumFres;  // final Newton residuals summed over substeps
    unsigned long long sweeps;  // the scene's own sweeps summed over substeps
};
// Contact-skin arguments of the broadphase kernels (state null: skin off).
struct SkinArgs {
    int32_t* state;               // [0] some segment walks, [1 + scene] the scene walks
                                  // (1: a list ended, 2: a new tree)
    float4* sbox;                 // grown box of each segment's last walk
    int32_t* cand;                // kept candidates (kSkinCand per segment, skin_slot)
    int32_t* candN;               // kept candidates per segment (< 0: incomplete)
    unsigned long long* stats;    // [0] walking, [1] reusing detection substeps
    int4* policy;                 // per scene: (substeps left walking without the skin,
                                  //  0 or 1 + substeps its kept lists have lasted,
                                  //  length of the last back-off, unused)
};
constexpr int kSkinCand = YB_SKIN_CANDIDATES;
// A scene whose kept lists are reused fewer than kSkinMinReuse times before
// its next walk (a walk with grown boxes costs about two plain walks, a reuse
// about half of one) walks without keeping candidates for a while, then
// tries the skin again: kSkinBackoff substeps, doubled after every retry
// whose lists did not last either, up to kSkinBackoffMax; lists that last
// reset it.
constexpr int kSkinMinReuse = 2;
constexpr int kSkinBackoffMax = 1024;
constexpr float kSkinDefaultMargin = 0.05f;  // detection radii
constexpr float kSkinDefaultSweep = 8.0f;    // substeps
constexpr int kSkinBackoff = 16;
// Growth by motion of a grown box, per axis, in detection radii at most.
constexpr float kSkinCapRadii = 8.0f;
constexpr int kMaxScenes = 65536;  // scenes of one handle (yb_set_scenes)

// Per-scene arguments of the kernels that honour per-scene convergence (both
// null outside a residual-targeted batch).
struct SceneArgs {
    const int32_t* sceneOf;  // scene of each vertex
    SceneCtl* ctl;           // per-scene solver state
};

// Residual-targeted batches need scenes of at least this many vertices: a
// 62-vertex wave tile then holds at most two scenes, whose residual partials
// it keeps apart.
constexpr int kMinSceneForResidual = 64;

constexpr int kMaxCols = YB_MAX_COLLISIONS_PER_SEGMENT;
// Bits of the device error word err[0] (one per failure kind).
constexpr int32_t kErrBitContacts = 1;  // a segment had more than kMaxCols contacts
constexpr int32_t kErrBitStack = 2;     // a broadphase query ran out of traversal stack

// The 8 XCDs each have their own L2, and block b of a launch runs on XCD b % 8.
// Tile index of block b of nb such that every XCD works through one contiguous
// range of tiles: the strands a tile gathers from (the neighbouring rows of a
// patch lie a few thousand vertices away) are then fetched into one L2, not
// into all eight.  A bijection onto [0, nb) for any nb.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
    const int q = nb >> 3, r = nb & 7;
    const int x = b & 7, k = b >> 3;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + k;
}

// Element i (+ 1 when `next`) of a per-vertex float4 array through a 32-bit
// byte offset: yb_create caps n at YB_MAX_VERTICES = 2^28, so every offset fits,
// and a data-dependent gather becomes one SGPR-base + VGPR-offset load with no
// 64-bit address arithmetic.
template <bool next = false>
__device__ __forceinline__ float4 gather4(const float4* __restrict__ a, int i) {
    const uint32_t off = (uint32_t)i << 4;
    return *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(a) + off +
                                            (next ? 16 : 0));
}

// Rest length of vertex v's segment, or of the previous one at a strand end
// (the length unit of the learned warm start's features and corrections).
__device__ __forceinline__ float rest_len(const float4* prm, int v, uint32_t flags) {
    if (!(flags & YB_FLAG_HAS_NEXT) && (flags & YB_FLAG_HAS_PREV)) return prm[v - 1].x;
    return prm[v].x;
}

// The smallest contact distance / (2 * radius) over a wave, folded into the
// running minimum *minSep (the interpenetration monitor of the step limit).
// Positive floats order like their bit patterns.  Only a wave that beats the
// running minimum issues the atomic, so after the first substep almost none
// do.  Every lane of the wave must take part.
__device__ __forceinline__ void reduce_min_sep(float sep, float* minSep) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) sep = fminf(sep, __shfl_xor(sep, off, 64));
    if ((threadIdx.x & 63) == 0 && sep < INFINITY && sep < *(volatile float*)minSep)
        atomicMin(reinterpret_cast<int*>(minSep), __float_as_int(sep));
}

// Learned warm start: parameter vector layout (floats), see yarnball_hip.h.
constexpr int kMlpIn = YB_MLP_INPUTS;
constexpr int kMlpHid = YB_MLP_HIDDEN;
constexpr int kMlpOut = YB_MLP_OUTPUTS;
constexpr int kMlpW1 = 0;
constexpr int kMlpB1 = kMlpW1 + kMlpIn * kMlpHid;
constexpr int kMlpW2 = kMlpB1 + kMlpHid;
constexpr int kMlpB2 = kMlpW2 + kMlpHid * kMlpHid;
constexpr int kMlpW3 = kMlpB2 + kMlpHid;
constexpr int kMlpB3 = kMlpW3 + kMlpHid * kMlpOut;
constexpr int kMlpTrainable = kMlpB3 + kMlpOut;
constexpr int kMlpMean = kMlpTrainable;
constexpr int kMlpIStd = kMlpMean + kMlpIn;
constexpr int kMlpOScale = kMlpIStd + kMlpIn;
constexpr int kMlpParams = kMlpOScale + kMlpOut;
static_assert(kMlpParams == YB_MLP_PARAMS, "parameter layout");
// bf16 weight tiles: 168 training tiles of 256 elements (K = 16 operands),
// then 52 inference tiles of 512 elements (K = 32 operands).
constexpr int kMlpTileBufElems = 168 * 256 + 52 * 512;
constexpr int kSortBits = 8;
constexpr int kSortBins = 1 << kSortBits;
constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortThreads * kSortItems;
// sorted leaves per refit workgroup (the refit tile): 256, 512 or 1024
constexpr int kRefitLeavesMin = 256, kRefitLeavesMax = 1024;
// Flag in a parent slot ((node << 1) | side, node < 2^28): set when that parent
// node's leaf range crosses a boundary between refit workgroups.
constexpr int32_t kParentStrad = 1 << 30;
constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanBlock = kScanThreads * kScanItems;  // histogram counters per scan workgroup

// SoA device state.  Every per-vertex array is float4 (16 B, one dwordx4 per
// lane, fully coalesced):
//   pos   = (x, y, z, invMass)          prm = (lRest, kStretch, cid bits, flags bits)
//   vel   = (v, 0)   lastVel = (v_prev, 0)   yv = (y - x, 0)
//   dx[2] = ping-pong iterate (dx, stepLimit of segment i)
//   q[2]  = ping-pong segment frames,   qRest = rest Darboux * stiffness
struct Buffers {
    float4* pos = nullptr;
    float4* prm = nullptr;
    float4* vel = nullptr;
    float4* lastVel = nullptr;
    float4* yv = nullptr;
    float4* dx[2] = {nullptr, nullptr};
    float4* q[2] = {nullptr, nullptr};
    float4* qRest = nullptr;
    int32_t* numCols = nullptr;
    int32_t* cols = nullptr;  // slot-major: cols[seg + n*slot]
    // broadphase
    float4* aabb = nullptr;       // 2n: lo (xyz), hi (xyz) per segment
    float4* sceneBox = nullptr;   // 2 float4: union of all segment boxes
    // batched scenes (numScenes > 1): scene of each vertex, first vertex of each
    // scene (+ n), per-scene boxes (2 per scene), root node of each scene's subtree
    int32_t* sceneOf = nullptr;
    int32_t* sceneOff = nullptr;
    float4* sceneBoxes = nullptr;
    int32_t* sceneRoot = nullptr;
    int32_t* sceneEndFlags = nullptr;
    SceneCtl* sceneCtl = nullptr;  // per-scene solver state (residual-targeted batches)
    double2* scenePart = nullptr;  // 3 parts x 2 sides x tiles: residual partials per scene
    int32_t* tileList = nullptr;   // wave tiles of the scenes still running (after a check)
    int32_t* tileCount = nullptr;  // [0]: entries of tileList (< 0: every tile);
                                   // [1..numScenes]: each scene's first entry
    float4* drvInit = nullptr;  // scenario driver: initial position xyz, initial inverse mass
    float4* blockBox = nullptr;   // 2 per reduction block
    uint32_t* morton = nullptr;   // n, indexed by segment
    uint32_t* keys[2] = {nullptr, nullptr};
    int32_t* vals[2] = {nullptr, nullptr};
    uint32_t* hist = nullptr;     // kSortBins * numTiles
    uint32_t* scanPart = nullptr; // per-workgroup sums of the histogram scan
    int32_t* order = nullptr;     // sorted leaf -> segment
    float4* nodes = nullptr;      // 4 per internal node: lo0, hi0, lo1, hi1 (child ids in .w)
    int32_t* parent = nullptr;    // 2n-1: (parent << 1) | side
    int2* rng = nullptr;          // n-1: sorted leaf range of each internal node
    int2* kids = nullptr;         // n-1: children, leaves as ~sorted position
    int32_t* strad = nullptr;     // internal nodes whose range crosses a refit block
    int32_t* stradCount = nullptr;
    int32_t* leafNeed = nullptr;  // n: bit0 suffix union needed, bit1 prefix union
    float4* sufBox = nullptr;     // 2n
    float4* preBox = nullptr;     // 2n
    float4* blkTab = nullptr;     // sparse table of refit-block unions, 2*nb per level
    int32_t* candCount = nullptr; // n: broadphase candidates per segment (measurement, lazy)
    // contact skin: grown boxes of the last walk (2n), kept candidate leaves
    // (YB_SKIN_CANDIDATES x n, laid out by skin_slot), their counts (< 0: incomplete),
    // walk flags ([0] any, [1 + scene] per scene; set by k_aabb, cleared by
    // k_init), counts of walking / reusing substeps, and the refit's root box
    // (unused: the tree holds grown boxes)
    float4* sbox = nullptr;
    int32_t* cand = nullptr;
    int32_t* candN = nullptr;
    int32_t* skinState = nullptr;
    unsigned long long* skinStats = nullptr;
    int4* skinPolicy = nullptr;
    float4* skinRoot = nullptr;
    int32_t* err = nullptr;       // [0] error bits, [1] warning code
    float* minSep = nullptr;      // smallest contact distance / (2 radius) seen
    DevParams* params = nullptr;
    SolverCtl* ctl = nullptr;
    double2* partial = nullptr;  // residual partial sums (update, force) per sweep wave, three parts
    double2* resBlocks = nullptr;        // residual check: per-block sums (2 per block)
    unsigned int* resCounter = nullptr;  // residual check: blocks done
    int32_t* hostFlag = nullptr;  // pinned: doneAt read back after each chunk
    // learned warm start
    float* mlpPrm = nullptr;          // kMlpParams floats
    uint16_t* mlpTiles = nullptr;     // packed bf16 weight tiles
    float* mlpBias = nullptr;         // 2*hidden + 16
    float4* corrPrev = nullptr;       // n: last converged correction (dx* - dx0)/L
    float4* dxPlain = nullptr;        // n: plain warm start of the current substep
    float4* predC = nullptr;          // n: predicted correction before smoothing
    // strand solve (warm-start option)
    float4* lineG = nullptr;   // n: tile solution xyz + left-boundary factor
    float2* lineQL = nullptr;  // n: right-boundary factor + step limit
    float4* lineE[2] = {nullptr, nullptr};  // per tile: last owned vertex's value, two iterations
    float4* lineF[2] = {nullptr, nullptr};  // per tile: first owned vertex's value
    unsigned long long* mlpCount = nullptr;  // [applied, fallback] per 64-vertex wave
    float* recX = nullptr;            // recording: cap x 48 raw features
    float* recT = nullptr;            // recording: cap x 3 targets
    int64_t* recCount = nullptr;      // [rows stored, recording enabled]
    // upload / download staging
    void* stage = nullptr;
    size_t stageBytes = 0;
};

}  // namespace yb

struct yb_sim {
    int n = 0;
    hipStream_t stream = nullptr;
    yb::Buffers b;
    int sortTiles = 0;
    int boxBlocks = 0;
    int refitLeaves = yb::kRefitLeavesMax;  // refit tile (yb_set_tile_config)
    int refitShift = 10;                // log2(refitLeaves)
    int refitBlocks = 0;
    int refitLevels = 0;
    int refitTopBlocks = 0;
    int numCUs = 256;        // compute units of the device (persistent grids)
    int queryStackLds = 8;   // k_query: traversal-stack entries in LDS
    int queryCandidates = 6; // k_query: candidate leaves buffered per thread
    bool treeBuilt = false;
    bool uploaded = false;
    int numScenes = 1;
    std::vector<int32_t> sceneOffHost;  // first vertex of each scene, + n
    std::vector<double> lsqHost;  // per vertex: l_i^2 of a free vertex, 0 otherwise (upload)
    float lastBVHRebuild = INFINITY;  // sim time since the last full rebuild
    int64_t stepCounter = 0;
    // captured substep graphs: [detect mode 0=none,1=refit,2=rebuild]
    hipGraphExec_t graphs[3] = {nullptr, nullptr, nullptr};
    int graphItr = -1;
    int useGraphs = 1;  // 0: every kernel enqueued directly on the stream
    int sweepThreads = 256;  // block size of k_sweep (64, 128, 256 or 512)
    int queryThreads = 128;  // block size of k_query (64, 128 or 256)
    int queryStack = YB_QUERY_STACK_DEPTH;  // traversal stack entries a query may use
    // contact skin (yb_set_contact_skin): margin in detection radii (<= 0: off) and
    // the substeps of warm-start motion a grown box covers
    float skinMargin = yb::kSkinDefaultMargin;
    float skinSweep = yb::kSkinDefaultSweep;
    bool skinSet = false;  // yb_set_contact_skin was called (else batches start without it)
    bool countCandidates = false;           // k_query writes b.candCount (measurement)
    // residual-targeted mode (tol > 0): prologue graphs per detection mode,
    // one graph per chunk of sweeps, one end graph
    float tol = 0.f;
    int maxSweeps = 0;
    int chunk = 8;
    int checkEvery = 1;
    // [0]: prologue with the session's warm start; [1]: its plain-start variant
    // (learned sessions with the substep-level fallback)
    hipGraphExec_t proGraphs[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
    hipGraphExec_t endGraph = nullptr;
    hipGraphExec_t* chunkGraphs = nullptr;
    int numChunks = 0;
    bool adaptiveReady = false;
    hipEvent_t chunkEv[2] = {nullptr, nullptr};
    void* prof = nullptr;  // eager-profiling event log (capi.hip)
    void* timeline = nullptr;  // residual-mode substep timeline (capi.hip)
    double denomHost = 0.0;  // update-residual normaliser (sum of l_i^2 over free vertices)
    int residualKind = 0;    // stopping test: YB_RESIDUAL_UPDATE or YB_RESIDUAL_NEWTON
    int newtonMode = 1;      // YB_NEWTON_ALL_STARTS or YB_NEWTON_PLAIN_STARTS
    int warmStart = 0;     // 0 plain (iteration.cu:18-34), 1 learned, 2 temporal
    int frameRefresh = 1;  // refresh frames after a non-plain warm start
    int strandSolve = 0;   // one Newton step along the strands after the warm start
    int strandIters = 2;   // boundary-pair iterations of the strand solve
    double sweepEstimate = 0.0;  // residual mode: recent sweeps to convergence (look-ahead)
    int strandEvery = 0;   // residual mode: also before every strandEvery-th sweep (0: never)
    int strandFirst = 0;   // ... starting at this sweep (0: at strandEvery)
    int mlpSmooth = 1;     // [1 2 1]/4 passes on the predicted correction (0 or 1)
    int mlpGrid = 0;       // resident blocks of the inference kernel (set by yb_mlp_load)
    // Substep-level fallback of a learned session in residual-targeted mode:
    // the chooser runs the warm start whose sweep-count average is lower and
    // probes the other one from time to time.
    int fallback = 1;
    double avgSweeps[2] = {0.0, 0.0};  // [0] learned, [1] plain; 0 = no sample yet
    float learnedCost = 0.0f;  // extra cost of a learned start, in sweeps (the chooser's offset)
    int64_t chooserStep = 0, nextProbe = 8;
    int probeGap = 8;
    int64_t warmSubsteps[2] = {0, 0};  // substeps run with [0] learned, [1] plain start
    bool mlpLoaded = false;
    int64_t recCap = 0;
    int graphLimit = -1;
    int graphVel = -1;
    float hostBox[8];
};

namespace yb {

void set_message(const std::string& m);

// ---- launchers (all enqueue on s->stream) ---------------------------------
void launch_upload(yb_sim* s);     // staging AoS -> SoA
void launch_download(yb_sim* s);   // SoA -> staging AoS
void launch_zero_vel(yb_sim* s);
void launch_drive(yb_sim* s, const yb_driver& d, float t, float dt);
// The residual-targeted schedule's per-substep reset of its solver state
// (ctl null: none), made by the first prologue kernel.
struct CtlReset {
    SolverCtl* ctl = nullptr;
    SceneCtl* sceneCtl = nullptr;
    int numScenes = 0;
    int32_t* tileCount = nullptr;
    unsigned int* resCounter = nullptr;
    int noNewton = 0;
};
CtlReset ctl_reset_args(const yb_sim* s, bool noNewton);  // kernels_sweep.hip
void launch_init(yb_sim* s, const CtlReset& cr = CtlReset{});  // iteration.cu:7-41
void launch_step_limit(yb_sim* s);              // collision.cu:121-172
void launch_separation(yb_sim* s);              // contact distances of the current state
// Sweep k reads dx[k&1], q[k&1] and writes the other buffer.  adaptive: the
// kernel honours SolverCtl (early exit after convergence, residual partials).
// cosserat.cu:20-253; partials: write this sweep's residual partials;
// checkPrev: the launch's first wave also checks the previous sweep.
void launch_sweep(yb_sim* s, int sweep, bool adaptive, bool partials, bool checkPrev);
// The check of sweep `sweep` as a launch of its own (right after it):
// residuals and the stopping decision; in a batch, per scene, and the tile
// list of the scenes still running.
void launch_residual(yb_sim* s, int sweep);
// Whether every check is a launch of its own (the strand solve and batches,
// which converge within tens of sweeps, so that the sweep after the converged
// check exits at once), or is made by the next sweep's first wave (plain and
// learned sessions, hundreds of sweeps: no launch, one sweep late).
inline bool separate_checks(const yb_sim* s) { return s->strandSolve || s->numScenes > 1; }
void launch_scene_end_flags(yb_sim* s);  // flags of each scene's last vertex -> sceneEndFlags
// noNewton: the substep records no Newton residual (ignored in a batch)
void launch_ctl_reset(yb_sim* s, bool noNewton = false);
void launch_frame_refresh(yb_sim* s);  // closed-form frame update from dx[0]
// Tridiagonal Newton step along the strands on dx[buf] and q[buf]; with ctl
// (between sweeps), nothing once the substep converged before `sweep`.
// sys (tests): also write each vertex's system row (diagonal, coupling to the
// next vertex, right-hand side xyz).
// qbuf >= 0: the frames are read from q[qbuf]; copyFrames: the step's last
// pass also copies q[1] -> q[0] (frames refreshed into q[1] by the prologue).
void launch_strand_solve(yb_sim* s, int buf = 0, SolverCtl* ctl = nullptr, int sweep = 0,
                         float* sys = nullptr, int qbuf = -1, bool copyFrames = false);
// Newton-residual denominator of the substep: the partials of |f(dx_0)|^2 at
// the plain warm start in dx[0] (after launch_ctl_reset; before the warm start
// is changed), reduced by the substep's first residual check.  frames: also
// refresh the frames from dx[0] (q[0] updated in place through q[1]).
void launch_residual_denom(yb_sim* s, bool frames);
// Both of the above in one pass over the tiles, for a substep whose sweeps
// start from the plain warm start: refresh the frames from dx[0], take the
// denominator's partials at the frames before the refresh, and make the
// strand step at the refreshed ones.  keepPlain: also copy dx[0] to dxPlain
// (in place of launch_features in a learned session's plain substep).
void launch_line_start(yb_sim* s, bool keepPlain);
const double2* denom_partials(const yb_sim* s);
// iteration.cu:48-61; correction: also the learned start's converged
// correction (k_correction's, without its training record)
void launch_end(yb_sim* s, int fin, bool adaptive, bool correction = false);
SceneArgs scene_args(const yb_sim* s, bool adaptive);
// The first sweep of a residual-targeted substep whose residual is checked.
int first_check(const yb_sim* s);
int sweep_waves(int n);
// learned warm start (mlp.hip)
void launch_mlp_pack(yb_sim* s);
void launch_features(yb_sim* s, bool record);
// inference, then the apply into dx[0]; frames: the apply also refreshes the
// frames from the new dx[0] (q[0] updated in place through q[1])
// copyFrames (with frames): the refreshed frames q[1] -> q[0] in a pass of its
// own; false when the strand step that follows reads q[1] and copies them
void launch_mlp_apply(yb_sim* s, bool frames, bool copyFrames = true);
void launch_learned_apply(yb_sim* s, bool frames, bool copyFrames = true);  // kernels_sweep.hip
void launch_temporal(yb_sim* s);
void launch_correction(yb_sim* s, int fin, bool adaptive, bool record);
void launch_fill_i32(yb_sim* s, int32_t* p, int32_t v, int count);
// broadphase (collision.cu:10-119 + the LBVH the reference links)
void launch_aabb(yb_sim* s, bool force = false);  // force: every segment walks
SkinArgs skin_args(const yb_sim* s);
void launch_rebuild(yb_sim* s);    // scene box, Morton, sort, Karras
void launch_scene_box(yb_sim* s);  // union of the current segment boxes -> sceneBox
void launch_refit(yb_sim* s);
void launch_query(yb_sim* s);      // traversal + narrowphase -> contact lists, step limit
bool query_shape_supported(int stackLds, int candidates);

}  // namespace yb
