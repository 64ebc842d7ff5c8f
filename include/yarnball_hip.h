/*
 * yarnball_hip.h — C ABI of the MI355X (gfx950) YarnBall stepper.
 *
 * Every entry point is plain C: POD structs, raw pointers and sizes, int status
 * codes.  No C++ types or exceptions cross this boundary.  The C++ drop-in facade
 * (include/YarnBall/YarnBall.h, class YarnBall::Sim) and the Python mirror
 * (yarnball_amd/sim.py) are both thin layers over these functions.
 *
 * Reference interface being replaced: class YarnBall::Sim
 * (KittenEngine/YarnBall/YarnBall.h:116-206) and its implementation in
 * KittenEngine/YarnBall/YarnBall.cu and the files under KittenEngine/YarnBall/sim.
 *
 * Status codes: 0 = ok; 1 = YB_ERROR_MAX_COLLISIONS_PER_SEGMENT_EXCEEDED
 * (YarnBall.h:120); 2 = YB_WARNING_SEGMENT_INTERPENETRATION (YarnBall.h:121);
 * 3 = YB_ERROR_TRAVERSAL_STACK_EXCEEDED (a broadphase query ran out of stack;
 * the reference's LBVH is not part of its sources, so this has no counterpart);
 * negative values are API / HIP errors, described by yb_last_message().
 */
#ifndef YARNBALL_HIP_H
#define YARNBALL_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define YB_MAX_COLLISIONS_PER_SEGMENT 128 /* YarnBall.h:20 */
/* Pending subtrees one broadphase query may hold (its traversal stack). */
#define YB_QUERY_STACK_DEPTH 64
/* Largest vertex count of one handle (all scenes together): device byte
 * offsets of per-vertex float4 arrays stay below 2^32.  At this size the
 * device buffers take about 180 GiB, most of it the 128-slot contact lists. */
#define YB_MAX_VERTICES (1 << 28)

enum {
    YB_OK = 0,
    YB_ERROR_MAX_COLLISIONS_PER_SEGMENT_EXCEEDED = 1, /* YarnBall.h:120 */
    YB_WARNING_SEGMENT_INTERPENETRATION = 2,          /* YarnBall.h:121 */
    YB_ERROR_TRAVERSAL_STACK_EXCEEDED = 3,            /* no counterpart */
    YB_ERR_INVALID_ARGUMENT = -1,
    YB_ERR_NOT_CONFIGURED = -2,
    YB_ERR_HIP = -3,
    YB_ERR_VALIDATION = -4,
    YB_ERR_NO_DEVICE = -5,
};

/* Vertex flag bits, YarnBall.h:34-40. */
enum {
    YB_FLAG_HAS_PREV = 1,
    YB_FLAG_HAS_NEXT = 2,
    YB_FLAG_HAS_NEXT_ORIENTATION = 4,
    YB_FLAG_FIX_ORIENTATION = 8,
    YB_FLAG_COLLIDING = 16,
};

/* 32-byte AoS vertex record, byte-identical to YarnBall::Vertex (YarnBall.h:52-61). */
typedef struct yb_vertex {
    float pos[3];
    float invMass;
    float lRest;
    float kStretch;
    int32_t connectionIndex;
    uint32_t flags;
} yb_vertex;

/* Simulation parameters: the scalar part of YarnBall::MetaData (YarnBall.h:63-114)
 * without its device pointers.  h, lastH, time, detectionRadius and
 * scaledDetectionRadius are written by yb_advance(), as Sim::advance and
 * Sim::uploadMeta do (step.cpp:36-53, YarnBall.cu:226-228). */
typedef struct yb_params {
    float gravity[3];
    int32_t numItr;
    float worldFloorNormal[3];
    float worldFloorPos;
    float h;
    float lastH;
    float time;
    int32_t numVerts;
    float damping;
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
    float bvhRebuildPeriod;
    int32_t detectionPeriod;
    float maxSegLen;
    float minSegLen;
    int32_t useStepSizeLimit;
    int32_t useVelocityRadius;
} yb_params;

typedef struct yb_sim yb_sim; /* opaque device-side simulation handle */

/* ---- host-side set-up math (no device needed) ---------------------------- */

/* Constructor defaults of Sim::Sim (YarnBall.cu:11-51): fills *p and, when the
 * arrays are non-null, the per-vertex defaults.  vels: 3n floats, qRests: 4n. */
int yb_default_params(yb_params* p, int numVerts, float maxH);
int yb_default_vertices(int n, yb_vertex* verts, float* vels, float* qs, float* qRests);

/* Sim::configure host part (YarnBall.cu:91-157): completes flags, rest lengths,
 * minimal-twist frames qs (4n), masses, rest Darboux vectors qRests (4n),
 * initialInvMasses (n, may be null), and min/maxSegLen in *p.  Returns
 * YB_ERR_VALIDATION with a message for "Dangling segment" / "0 length segment". */
int yb_configure_host(int n, yb_vertex* verts, float* qs, float* qRests, float density,
                      float* initialInvMasses, yb_params* p);
/* Sim::setKStretch (YarnBall.cu:202-210). */
int yb_set_kstretch_host(int n, yb_vertex* verts, float kStretch);
/* Sim::setKBend (YarnBall.cu:212-224). */
int yb_set_kbend_host(int n, const yb_vertex* verts, float* qRests, float kBend);
/* Sim::uploadMeta validation and derived radii (YarnBall.cu:226-235). */
int yb_validate_params(yb_params* p);

/* ---- device lifecycle ---------------------------------------------------- */

/* Device allocation part of Sim::configure (YarnBall.cu:166-198): SoA buffers,
 * one HIP stream, contact lists, BVH storage.  3 <= n <= YB_MAX_VERTICES
 * (the lower bound is YarnBall.cu:12). */
int yb_create(int n, yb_sim** out);
/* Sim::~Sim (YarnBall.cu:53-85). */
void yb_destroy(yb_sim* s);

/* Sim::upload (YarnBall.cu:247-254): AoS host arrays -> device SoA.
 * verts: n records; vels: 3n; qs: 4n (x,y,z,w); qRests: 4n.  Synchronous. */
int yb_upload(yb_sim* s, const yb_vertex* verts, const float* vels, const float* qs,
              const float* qRests);
/* Sim::download (YarnBall.cu:256-261): device SoA -> AoS host.  Any pointer may
 * be null to skip that array.  Synchronous. */
int yb_download(yb_sim* s, yb_vertex* verts, float* vels, float* qs);
/* Sim::zeroVelocities (YarnBall.cu:263-274). */
int yb_zero_velocities(yb_sim* s);

/* Sim::advance (step.cpp:36-55) + Sim::checkErrors (YarnBall.cu:276-304).
 * Validates *p like uploadMeta, splits h into ceil(h/maxH) substeps, runs them on
 * the handle's stream and blocks until done.  Updates p->h, p->lastH, p->time and
 * the derived radii.  Returns 0, 1 (contact-list overflow; state may be corrupt),
 * or a negative error.  *warning receives 2 when a penetration was detected. */
int yb_advance(yb_sim* s, yb_params* p, float h, float maxH, int* warning);
/* Number of substeps run so far (Sim::stepCount, YarnBall.h:163). */
int64_t yb_step_count(yb_sim* s);

/* Root AABB of the broadphase tree, as Sim::bounds / LBVH::bounds
 * (YarnBall.cu:87-89; collision.cu:109): out[0..2] min, out[3..5] max. */
int yb_bounds(yb_sim* s, float out[6]);
/* Per-segment contact counts (the data printCollisionStats reads,
 * statistics.cpp:5-24).  out: n ints. */
int yb_collision_counts(yb_sim* s, int32_t* out);

/* Interpenetration monitor (WARNING_SEGMENT_INTERPENETRATION, declared at
 * YarnBall.h:121 but never raised by the reference): the smallest distance
 * between listed contact segments divided by 2 * radius, over every substep
 * start and every advance end since the last reset.  Below 1 means two yarn
 * cores overlap; advance then reports the warning.  INFINITY if no contact. */
int yb_min_separation(yb_sim* s, float* out, int reset);

/* Last error message (thread-local), for negative status codes. */
const char* yb_last_message(void);

/* ---- kinematic drivers on the device ---------------------------------------
 * The reference's scenario animations (main.cpp:54-121) move the pinned
 * vertices on the host between download() and upload() every frame; here one
 * kernel applies the same per-frame update to the device state.
 *   twist: every pinned vertex on the chosen side of center.y is placed at
 *          center + rotateY(initial - center, angle(t)) with velocity
 *          (position(t + dt) - position(t)) / dt, angle(t) = speed *
 *          clamp(t - start, 0, duration); after t > release it gets its
 *          initial inverse mass back (release < 0: never).
 *   pull:  gravity becomes (0, 0, -9.8) in *p; every pinned vertex gets
 *          y = initial.y + x(t) below center.y and initial.y - x(t) above, with
 *          x(t) = -speed * clamp(t - start, 0, duration), and the matching
 *          y velocity.
 * yb_driver_default fills main.cpp's constants (twist: pi rad/s from t = 2 s
 * for 12 s, released at 15 s, the vertices below the centre; pull: 0.25 m/s
 * from 2 s for 6 s).  yb_driver_init records the current device positions as
 * the initial ones, the initial inverse masses (null: the current ones), and
 * returns the centre of their bounding box (main.cpp:273-280). */
#define YB_DRIVE_TWIST 1
#define YB_DRIVE_PULL 2
typedef struct yb_driver {
    int32_t kind;
    float center[3];
    float speed;    /* rad/s (twist) or m/s (pull) */
    float start;    /* s */
    float duration; /* s */
    float release;  /* s, twist only; < 0: never */
    int32_t side;   /* twist: -1 below center.y, +1 above, 0 every pinned vertex */
} yb_driver;
int yb_driver_default(int kind, yb_driver* d);
int yb_driver_init(yb_sim* s, const float* initialInvMasses, float center[3]);
int yb_driver_frame(yb_sim* s, yb_params* p, const yb_driver* d, float t, float dt);

/* ---- scene files (host side, no device needed) ------------------------------
 * A yb_scene holds what Sim(numVerts) plus the caller's edits hold before
 * upload(): yb_params, maxH, the AoS vertices, velocities, frames, rest
 * frames and the initial inverse masses.  Read it out with yb_scene_get
 * (null pointers skip an array) and free it with yb_scene_free. */
typedef struct yb_scene yb_scene;
/* buildFromJSON (jsonBuilder.cpp:8-170): curve file + resampling, curveRadius,
 * the "simulation" block, configure(density), setKStretch/setKBend,
 * glueEndpoints, fixBorders, fixVertex.  Comments and trailing commas, as in
 * the shipped configs, are accepted.  A relative curveFile is looked up in the
 * working directory, then beside the JSON file and one directory above it. */
int yb_build_from_json(const char* path, yb_scene** out);
/* readFromPoly / readFromOBJ / readFromBCC (reader.cpp:50-295), chosen by the
 * file extension (.poly, .obj, else BCC): curves -> vertices with strand ends
 * and closed-curve connections (createFromCurves, reader.cpp:13-48), Catmull-
 * Rom resampling to targetSegLen (resample.h).  transform: 12 floats, a
 * row-major 3x4 affine map applied to every point, or null for the reference's
 * default 0.01 scale.  The scene is not configured. */
int yb_read_curves(const char* path, float targetSegLen, const float* transform,
                   int breakUpClosedCurves, int allowResample, yb_scene** out);
int yb_scene_size(const yb_scene* s, int* numVerts);
int yb_scene_get(const yb_scene* s, yb_params* p, float* maxH, yb_vertex* verts, float* vels,
                 float* qs, float* qRests, float* initialInvMasses);
void yb_scene_free(yb_scene* s);
/* Sim::glueEndpoints (jsonBuilder.cpp:174-212) on host vertices: each strand
 * endpoint, in index order, whose nearest vertex more than 2 indices away is
 * within searchRadius is connected to it (connectionIndex both ways) unless
 * either is connected already, and both move to their midpoint. */
int yb_glue_endpoints(int n, yb_vertex* verts, float searchRadius);
/* Sim::exportToBCC / exportToOBJ (fileExport.cpp:6-100) of host vertices
 * (call yb_download first, as the reference's methods do). */
int yb_export_bcc(const char* path, int n, const yb_vertex* verts, int asPolyline);
int yb_export_obj(const char* path, int n, const yb_vertex* verts);

/* ---- stage-level access used by the parity tests -------------------------
 * These expose intermediate device state so that each stage can be compared
 * with the CPU oracle in isolation.  They have no counterpart in the reference
 * (whose state is private), and are not needed by production callers. */

/* Run only the substep prologue: initItr (iteration.cu:7-41) and, when detect!=0,
 * the broadphase + narrowphase of detectCollisions (collision.cu:94-119) and
 * the step limit (collision.cu:121-172).  Uses p and h as yb_advance would. */
int yb_debug_prologue(yb_sim* s, yb_params* p, float h, int detect, int rebuild);
/* Broadphase counts for the query's byte model: runs yb_debug_prologue(s, p,
 * h, detect 1, refit) once more with counting on (so it changes the state as
 * that call does) and returns, summed over segments: out[0] segments, out[1]
 * candidate leaves whose boxes overlap a query box (leaves within two indices,
 * always exempt, not counted), out[2] accepted contacts
 * (the list lengths), out[3] tree nodes visited (with the contact skin, in a
 * substep that reuses the kept candidates: the candidates scanned); out[4]
 * the segments whose traversal stack grew past 8 entries, out[5] the deepest
 * stack. */
int yb_debug_query_counts(yb_sim* s, yb_params* p, float h, int64_t out[6]);
/* Substep timeline of the residual-targeted schedule (tol > 0), from HIP
 * events recorded on the handle's stream between its launches: one before
 * the prologue, one after it, one after each chunk of sweeps, one after the
 * end kernels.  yb_debug_timeline(s, m) starts recording the next m substeps
 * (0 stops and frees the events); yb_debug_timeline_read writes 8 floats per
 * recorded substep: [0] GPU idle before the substep in us (time from the
 * previous substep's end event to this one's first; 0 for the first row),
 * [1] prologue us, [2] us from the prologue to the end of the chunk that held
 * the converged sweep, [3] us of the chunks queued after it, [4] end kernels
 * us, [5] sweeps to convergence, [6] chunks launched, [7] the converged chunk
 * (-1: not converged).  Recording changes no results; each event adds a
 * little time of its own. */
int yb_debug_timeline(yb_sim* s, int maxSubsteps);
int yb_debug_timeline_read(yb_sim* s, float* out, int maxRows, int* rows);
/* Copy out intermediate arrays (any pointer may be null):
 * dx: 4n floats (xyz + step limit), aabb: 8n floats (lo xyz _, hi xyz _),
 * morton: n u32 indexed by segment, order: n i32 (sorted leaf -> segment),
 * numCols: n, cols: 128n slot-major (cols[seg + n*slot]). */
int yb_debug_read(yb_sim* s, float* dx, float* aabb, uint32_t* morton, int32_t* order,
                  int32_t* numCols, int32_t* cols);
/* Karras tree arrays after the last rebuild: child: 2(n-1) ints, (left, right)
 * per internal node, an internal child as its index and a leaf as ~segment;
 * parent: (2n-1) ints, (parent << 1 | side) for internal nodes [0, n-1) and then
 * for leaves in sorted order, -1 for the root; box: 16 floats per internal node
 * (left lo, left hi, right lo, right hi as xyzw; w of the lo entries holds the
 * child code bits) after the last refit.
 * Returns YB_ERR_INVALID_ARGUMENT if no tree was built. */
int yb_debug_tree(yb_sim* s, int32_t* child, int32_t* parent, float* box);
/* Plain warm start of the last prologue (4n floats) and the first n recorded
 * feature rows (48n floats) -- for the learned-warm-start tests. */
int yb_debug_read_mlp(yb_sim* s, float* dxPlain, float* features);
/* One strand step (see YB_WARM_STRAND_SOLVE) on the current iterate dx, with
 * the handle's pair-iteration count and no frame refresh; system (5n floats,
 * may be null): each vertex's row -- diagonal, coupling to the next vertex,
 * right-hand side xyz -- for checking the solve. */
int yb_debug_strand(yb_sim* s, float* system);
/* Traversal stack entries a broadphase query may use, 1..YB_QUERY_STACK_DEPTH
 * (the default).  Lowering it forces the YB_ERROR_TRAVERSAL_STACK_EXCEEDED
 * path on ordinary scenes (tests); a query that runs out skips the subtree it
 * cannot push. */
int yb_debug_set_query_stack(yb_sim* s, int entries);
/* Overwrite the device warm-start/iterate dx (4n floats) and lastVels (3n). */
int yb_debug_write_dx(yb_sim* s, const float* dx);
int yb_debug_read_last_vels(yb_sim* s, float* lastVels);
int yb_debug_write_last_vels(yb_sim* s, const float* lastVels);

/* ---- residual-targeted solving ---------------------------------------------
 * No counterpart in the reference, which always runs meta.numItr sweeps
 * (step.cpp:18-19).  With residualTol > 0 every substep runs sweeps until its
 * stopping residual is <= residualTol, or maxSweeps sweeps have run; numItr
 * is then ignored.  Two residuals are measured at every check:
 *   update residual (YB_RESIDUAL_UPDATE, the default stopping test)
 *     r_U = sqrt( sum_free |dx_{k+1} - dx_k|^2 / sum_free l_i^2 ),
 *     the RMS change of the iterate per sweep in units of the rest length l_i
 *     of the vertex's segment (the previous one at a strand end);
 *   Newton residual (YB_RESIDUAL_NEWTON)
 *     r_F = |f(dx_k)| / |f(dx_0)|   (2-norms over the free vertices),
 *     f the force of each vertex's local Newton step (inertia, glue,
 *     stretch-shear, barrier, lagged friction: the right-hand side of
 *     cosserat.cu:195-205), except at a vertex the step limit holds, where it
 *     is H (dx' - dx) / accelerationRatio, the part of f the limit does not
 *     absorb; dx_k the sweep's input iterate; dx_0 the plain warm start
 *     (iteration.cu:18-34) for every warm start mode, so all modes are held
 *     to the same target.
 * r_F of FP32 state levels off at a scene-dependent floor (frames and
 * positions carry 2^-24 relative rounding, which stiff contact forces
 * amplify; e.g. near 2e-3 on the two-strand README scene, above 1e-2 on a
 * twisted cable), so it is a usable stopping test only above that floor; the
 * final r_F of every substep is reported either way (yb_solver_stats).
 * Sweeps are launched in captured chunks of `chunk`; the residuals are
 * evaluated after sweep 0 and every checkEvery-th sweep.  residualTol <= 0
 * restores the fixed numItr schedule. */
#define YB_RESIDUAL_UPDATE 0
#define YB_RESIDUAL_NEWTON 1
typedef struct yb_solver_stats_t {
    int64_t substeps;     /* substeps run in residual-targeted mode */
    int64_t sweeps;       /* sweeps summed over those substeps */
    int32_t maxSweepsUsed;
    int32_t notConverged; /* substeps that stopped at maxSweeps above residualTol */
    float lastResidual;   /* stopping residual of the last check */
    float maxResidual;    /* largest final stopping residual of a substep */
    float lastNewtonResidual;  /* r_F of the last check */
    float maxNewtonResidual;   /* largest final r_F of a substep */
    double sumNewtonResidual;  /* final r_F summed over substeps (over scene-substeps
                                  in a batch) */
    int64_t sceneSubsteps;     /* batched scenes: substeps x scenes (else = substeps) */
    int64_t sceneSweeps;       /* batched scenes: each scene's own sweeps, summed
                                  (else = sweeps) */
    int64_t newtonSubsteps;    /* substeps (scene-substeps) that recorded r_F: the
                                  count behind sumNewtonResidual and maxNewtonResidual */
} yb_solver_stats_t;
int yb_set_solver(yb_sim* s, float residualTol, int maxSweeps, int chunk, int checkEvery);
/* Stopping test of the residual-targeted mode: YB_RESIDUAL_UPDATE (default) or
 * YB_RESIDUAL_NEWTON. */
int yb_set_residual_kind(yb_sim* s, int kind);
/* Which substeps record the Newton residual r_F while the stopping test is the
 * update residual (no counterpart in the reference, which measures no
 * residual).  r_F's denominator is the force at the plain warm start, which a
 * substep starting from the plain warm start evaluates in a pass it makes
 * anyway (its first check, or the pass that also refreshes the frames and
 * makes the first strand step).  A substep starting from the learned or
 * temporal warm start needs one more force pass over the plain start for it
 * (about one sweep).  YB_NEWTON_ALL_STARTS (default): every substep records
 * r_F.  YB_NEWTON_PLAIN_STARTS: only substeps that start from the plain warm
 * start; the others skip the extra pass.  The iterates are the same either
 * way.  Batches and the YB_RESIDUAL_NEWTON stopping test always record r_F. */
#define YB_NEWTON_ALL_STARTS 1
#define YB_NEWTON_PLAIN_STARTS 0
int yb_set_newton_residual(yb_sim* s, int mode);
int yb_solver_stats(yb_sim* s, yb_solver_stats_t* out, int reset);

/* ---- batched scenes ---------------------------------------------------------
 * No counterpart in the reference (one Sim per scene).  Many independent
 * scenes share one handle: their vertices are concatenated, scene k owning
 * [offsets[k], offsets[k+1]) with offsets[0] = 0 and offsets[numScenes] = n;
 * the last vertex of every scene must not have YB_FLAG_HAS_NEXT.  Every kernel
 * then covers all scenes in one launch.  Contact detection stays inside each
 * scene: per-scene bounding boxes and Morton codes, scene-major sorted order,
 * one subtree per scene whose shape equals a standalone run's, and queries
 * that start at their own scene's subtree root.  The scenes share one
 * yb_params (gravity, radii, stiffness scales, schedule); per-vertex
 * stiffness and rest shapes are per scene.  With the fixed sweep schedule each
 * scene's state evolves bit-identically to a standalone handle of that scene.
 * In the residual-targeted schedule every scene has its own residual and
 * stops at its own sweep (converged scenes are left alone while the others
 * run; the substep ends when the last scene has converged), with the strand
 * solve too; this needs scenes of at least 64 vertices (yb_advance fails
 * with YB_ERR_INVALID_ARGUMENT otherwise).  Sweep counts per scene are in
 * yb_solver_stats (sceneSweeps).
 * numScenes = 1 restores the single-scene layout.  Call after yb_create
 * (before or after yb_upload). */
int yb_set_scenes(yb_sim* s, int numScenes, const int32_t* offsets);

/* Launch shape (no counterpart in the reference): threads per block of the
 * Newton sweep (64, 128, 256 or 512; default 256) and of the contact query
 * (64, 128 or 256; default 128).  Results do not depend on it -- every
 * kernel's output per vertex or segment is a function of its inputs only, and
 * the residual partials are per 62-vertex wave tile -- only the speed does;
 * yarnball_amd/autotune.py picks it by timing. */
int yb_set_launch_config(yb_sim* s, int sweepThreads, int queryThreads);
int yb_get_launch_config(yb_sim* s, int* sweepThreads, int* queryThreads);

/* Tile shapes (no counterpart in the reference): the refit tile -- sorted
 * leaves per refit workgroup, 256, 512 or 1024 (default 1024) -- and the
 * contact query's LDS shape -- traversal-stack entries kept in LDS and
 * candidate leaves buffered per thread, one of (4,4) (8,6) (8,8) (16,8)
 * (default (8,6)).  Refit boxes are min/max unions and the query's contact
 * lists keep their order for every shape, so only the speed depends on them;
 * a new refit tile rebuilds the tree at the next detection.
 * yarnball_amd/autotune.py picks them by timing. */
int yb_set_tile_config(yb_sim* s, int refitLeaves, int queryStackLds, int queryCandidates);
int yb_get_tile_config(yb_sim* s, int* refitLeaves, int* queryStackLds, int* queryCandidates);

/* Contact skin (no counterpart in the reference, whose detectCollisions
 * refits the tree and walks it in every detecting substep, collision.cu:94-119;
 * the contact lists are the same either way).  With margin > 0, a detecting
 * substep that must walk the tree grows every segment's box by its warm-start
 * displacement times `sweep` (the motion of `sweep` substeps at that speed,
 * at most 8 detection radii along an axis) and by margin x
 * meta.detectionRadius on every side, walks the tree refit on those grown
 * boxes, and keeps each segment's candidate leaves (at most
 * YB_SKIN_CANDIDATES, in the walk's order).  Later detecting substeps skip
 * the refit and the walk while every segment's current box lies inside its
 * grown box, and test only the kept candidates: two segments whose current
 * boxes overlap have overlapping grown boxes, so the walk found the pair.
 * The leaf order of a walk is the tree's in-order leaf order, so the contact
 * lists equal those of a walk, slot for slot.  Every other case -- a box
 * outside its grown box (in a batch, per scene), a full candidate list, a
 * traversal-stack overflow, a tree rebuild, new scenes -- walks again.  A
 * scene whose lists are reused fewer than twice before it must walk again
 * walks the plain way (no grown boxes, nothing kept) for 16 substeps, then
 * tries the skin again (after each retry that fails too, for twice as long,
 * up to 1024 substeps; lists that last reset it).  margin <= 0 turns the skin off (the tree is refit
 * and walked in every detecting substep).  Default: margin 0.05, sweep 8 for
 * a handle of one scene; a handle given several scenes (yb_set_scenes)
 * starts with the skin off unless this was called. */
#define YB_SKIN_CANDIDATES 64
int yb_set_contact_skin(yb_sim* s, float margin, float sweep);
int yb_get_contact_skin(yb_sim* s, float* margin, float* sweep);
/* Detecting substeps that walked the tree (out[0]) and that reused the kept
 * candidates (out[1]) since the last reset; reset != 0 clears the counts. */
int yb_contact_skin_stats(yb_sim* s, int64_t out[2], int reset);
/* Contact skin state (tests): each segment's kept candidate count (n ints,
 * -1: incomplete, the segment walks at its next detection), grown box
 * (8n floats, lo xyzw, hi xyzw) and each scene's policy (4 ints per scene:
 * substeps left walking without the skin, 0 or 1 + substeps its kept lists
 * have lasted, length of the last back-off, 0); any pointer may be null. */
int yb_debug_read_skin(yb_sim* s, int32_t* candN, float* sbox, int32_t* policy);

/* Launch mode.  on = 1 (default): each substep's kernels run as captured HIP
 * graphs (step.cpp:5-34 likewise captures the sweep chain).  on = 0: the same
 * kernels are enqueued one by one on the handle's stream -- same results,
 * more host launch overhead; for tools that cannot trace graph launches. */
int yb_set_graphs(yb_sim* s, int on);

/* ---- learned warm start ------------------------------------------------------
 * No counterpart in the reference, whose warm start is the gravity
 * extrapolation of iteration.cu:18-34.  A 48 -> 128 -> 128 -> 3 ReLU MLP
 * (bf16 on the matrix cores, fp32 accumulation) predicts each free vertex's
 * correction (dx* - dx0)/L from 48 local features, as a residual on top of the
 * previous substep's converged correction (feature 36-38): the applied
 * correction is c_prev + MLP * output scale.  The corrected dx0, clamped to
 * the step limit, starts the sweeps.  Non-finite predictions keep dx0.
 * Parameter vector (YB_MLP_PARAMS floats): W1[48][128], b1[128], W2[128][128],
 * b2[128], W3[128][3], b3[3] (YB_MLP_TRAINABLE values), then the feature
 * means[48], inverse standard deviations[48] and output scales[3]. */
#define YB_MLP_INPUTS 48
#define YB_MLP_HIDDEN 128
#define YB_MLP_OUTPUTS 3
#define YB_MLP_TRAINABLE 23171
#define YB_MLP_PARAMS 23270
int yb_mlp_load(yb_sim* s, const float* params, int count);
/* mode 0: plain warm start, 1: learned (needs yb_mlp_load), 2: temporal --
 * dx0 plus the previous substep's converged correction, no network (the
 * ablation baseline for the learned mode).  After a non-plain warm start the
 * frames get one closed-form update from the new dx, so the first sweep
 * starts from consistent positions and frames; OR the mode with
 * YB_WARM_NO_FRAME_REFRESH to skip that (ablation). */
#define YB_WARM_NO_FRAME_REFRESH 0x100
/* The learned correction is low-passed along each strand with a [1 2 1]/4
 * filter before it is applied; YB_WARM_NO_SMOOTHING skips that (ablation). */
#define YB_WARM_NO_SMOOTHING 0x200
/* Substep-level fallback (residual-targeted mode, learned warm start): each
 * substep starts from the learned or the plain warm start, whichever has the
 * lower running average of sweeps to convergence; every 8th substep runs the
 * other one to keep its average current.  The plain solver still decides
 * convergence either way.  YB_WARM_NO_FALLBACK always uses the learned start. */
#define YB_WARM_NO_FALLBACK 0x400
/* Strand solve: after the warm start (any mode), one Newton step of the
 * linearised system restricted to the strands -- on the diagonal the
 * curvature of each vertex's local Hessian H along its force f, f.H.f / f.f
 * (trace(H)/3 where f vanishes), the stretch stiffness of each segment
 * between its free ends off the diagonal -- solved along whole strands: exactly
 * inside each 62-vertex sweep tile (cyclic reduction over the wave's lanes),
 * then by iterations over the vertex pairs that face each other across tile
 * boundaries (two by default, see yb_set_strand_schedule).  The correction is
 * clamped to the step limit like a sweep step, and the frames are refreshed
 * unless YB_WARM_NO_FRAME_REFRESH is set.  Block-Jacobi sweeps propagate a
 * correction one vertex per sweep and damp smooth modes along stiff strands
 * slowly; this step removes them.  It changes only the iterate the sweeps
 * start from: they still run to the residual tolerance.  No counterpart in
 * the reference. */
#define YB_WARM_STRAND_SOLVE 0x800
int yb_set_warm_start(yb_sim* s, int mode);
/* The substep-level chooser of a learned session with the fallback (residual-
 * targeted mode) runs the learned start while its running sweep average plus
 * `sweeps` is at most the plain start's: the extra time of a learned prologue
 * (inference, apply) expressed in sweeps.  Default 0 (fewest sweeps wins).
 * No counterpart in the reference. */
int yb_set_learned_start_cost(yb_sim* s, float sweeps);
/* Strand-solve schedule (with YB_WARM_STRAND_SOLVE): in residual-targeted
 * mode, also run the strand step before sweeps first, first + every,
 * first + 2 every, ... (every 0: only after the warm start; first 0: first =
 * every), on the current iterate, with no frame refresh (the next sweep
 * updates the frames); pairIters boundary-pair iterations per step (default 2).  The
 * residual of the sweep before such a step is reduced before it, and a
 * converged iterate is never changed.  At most 16 such steps per substep:
 * one that has not converged by then finishes with plain sweeps.  Vertices
 * held at their step limit take no part in a strand step (they stay where the
 * sweeps put them).  No counterpart in the reference. */
int yb_set_strand_schedule(yb_sim* s, int every, int pairIters, int first);
/* Vertex level: out[0] predictions applied, out[1] vertices that kept the
 * plain warm start because the prediction was not finite. */
int yb_mlp_counters(yb_sim* s, int64_t out[2], int reset);
/* Substep level (residual-targeted mode): out[0] substeps started from the
 * session's warm start, out[1] substeps started from the plain fallback. */
int yb_warm_start_substeps(yb_sim* s, int64_t out[2], int reset);
/* Training-data recording: each substep appends its n feature rows (raw,
 * unnormalised) and, after convergence, the n target corrections, until
 * `capacity` rows are stored.  capacity 0 stops recording. */
int yb_record(yb_sim* s, int64_t capacity);
/* Pause (0) / resume (1) recording without rebuilding the captured graphs. */
int yb_record_enable(yb_sim* s, int on);
int yb_record_fetch(yb_sim* s, float* X, float* T, int64_t* count);
/* Training on the device (own forward / backward kernels, Adam):
 * X n x 48 raw features, T n x 3 targets of the network itself (for the warm
 * start: the converged correction minus features 36-38); params in/out (normalisation
 * constants are read, the trainable part is updated).  lossHistory (epochs
 * floats, may be null) receives the mean squared error per sample. */
int yb_mlp_train(const float* X, const float* T, int64_t n, float* params, int epochs, int batch,
                 float lr, uint32_t seed, float* lossHistory);
/* Gradient of L = (1/n) sum_samples |y - t/scale|^2 over n samples (a multiple of
 * 16) at `params`: grad receives YB_MLP_TRAINABLE floats, *loss the summed
 * squared error.  Used to check the backward kernels. */
int yb_mlp_grad(const float* X, const float* T, int64_t n, const float* params, float* grad,
                float* loss);
/* Forward pass: Y (n x 3) in normalised output units. */
int yb_mlp_forward(const float* X, int64_t n, const float* params, float* Y);
/* The same forward pass without allocating: X (n x 48 floats), params
 * (YB_MLP_PARAMS floats) and Y (n x 3 floats) are device pointers; the work is
 * enqueued on `stream` (a hipStream_t; null: the default stream) and the call
 * returns without waiting.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops).  Its scratch -- packed weights and the normalised inputs, about
 * 6.4 MB -- is one static array per GPU, held until the process ends, so
 * n <= YB_MLP_NOALLOC_MAX_ROWS (larger n: YB_ERR_INVALID_ARGUMENT; split the
 * batch), and calls that overlap on one GPU race.  Y equals yb_mlp_forward's
 * bit for bit. */
#define YB_MLP_NOALLOC_MAX_ROWS 65536
int yb_mlp_forward_device(const float* X, int64_t n, const float* params, float* Y,
                          void* stream);

/* ---- live per-stage timing (measurement harness) --------------------------
 * Runs `substeps` substeps of length h with the yb_advance schedule, enqueued
 * without graphs and with a HIP event between consecutive kernels on the
 * handle's stream.  ms[k] receives the summed GPU time and launches[k] the
 * launch count of stage k (arrays of YB_NUM_STAGES).  No counterpart in the
 * reference (which has no timing instrument beyond a wall clock,
 * KittenEngine/main.cpp:123-149). */
enum {
    YB_STAGE_INIT = 0,        /* iteration.cu:7-41 */
    YB_STAGE_AABB = 1,        /* collision.cu:10-37 */
    YB_STAGE_REBUILD = 2,     /* LBVH compute: scene box, Morton, sort, Karras */
    YB_STAGE_REFIT = 3,       /* LBVH refit */
    YB_STAGE_QUERY = 4,       /* LBVH query + collision.cu:39-92 */
    YB_STAGE_STEP_LIMIT = 5,  /* collision.cu:121-159 */
    YB_STAGE_SWEEP = 6,       /* cosserat.cu:20-253 (dx step + frame update, fused) */
    YB_STAGE_MLP = 7,         /* warm starts: learned features + inference, temporal,
                                 frame refresh, strand steps */
    YB_STAGE_END = 8,         /* iteration.cu:48-61 */
    YB_STAGE_RESIDUAL = 9,    /* residual reduction / solver control */
    YB_NUM_STAGES = 10
};
int yb_profile(yb_sim* s, yb_params* p, float h, int substeps, double* ms, int32_t* launches);

#ifdef __cplusplus
}
#endif
#endif /* YARNBALL_HIP_H */
