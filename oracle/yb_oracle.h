/*
 * yb_oracle.h — CPU oracle of the YarnBall stepper.  TEST INFRASTRUCTURE ONLY.
 *
 * This library is the checker, never the product: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * there only to compare against or to time as the CPU baseline.  The product
 * path (libyarnball_hip.so) never links or calls it.
 *
 * It restates the reference's algorithm:
 *   continuous stepper in FP64 (OpenMP):
 *     initItr               KittenEngine/YarnBall/sim/iteration.cu:7-41
 *     recomputeStepLimit    KittenEngine/YarnBall/sim/collision.cu:121-159
 *     cosseratItr           KittenEngine/YarnBall/sim/cosserat.cu:20-207
 *     quaternionLambdaItr   KittenEngine/YarnBall/sim/cosserat.cu:209-253
 *     endItr                KittenEngine/YarnBall/sim/iteration.cu:48-61
 *     advance schedule      KittenEngine/YarnBall/sim/step.cpp:36-55
 *     configure/setK*       KittenEngine/YarnBall/YarnBall.cu:91-224
 *   discrete stages in FP32, same operation order as the device (bit-exact):
 *     warm start, padded boxes (collision.cu:10-37), Morton codes, stable sort,
 *     contact lists (collision.cu:39-92) from an exact grid broadphase.
 * Deliberate schedule choices shared with the device: strict Jacobi for dx and
 * q, contact lists as sets (sorted by partner here).
 */
#ifndef YB_ORACLE_H
#define YB_ORACLE_H
#include <stdint.h>

#include "../include/yarnball_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sim orc_sim;

void orc_set_threads(int n);
int orc_get_threads(void);

orc_sim* orc_new(int n);
void orc_free(orc_sim* o);
/* Load state from the device-format arrays (FP32 AoS). lastVels: 3n. */
void orc_load(orc_sim* o, const yb_vertex* verts, const float* vels, const float* lastVels,
              const float* qs, const float* qRests);
/* Load / store the FP64 state. Any pointer may be null. */
void orc_load64(orc_sim* o, const double* pos, const double* vels, const double* lastVels,
                const double* qs);
void orc_store64(const orc_sim* o, double* pos, double* vels, double* lastVels, double* qs);
void orc_set_step_counter(orc_sim* o, int64_t c);
/* One substep with p->h, p->lastH as given; detect != 0 refreshes contacts. */
int orc_substep(orc_sim* o, const yb_params* p, int detect);
/* Sim::advance schedule: ceil(h/maxH) substeps, detection every
 * detectionPeriod substeps; updates p->h, p->lastH, p->time, radii. */
int orc_advance(orc_sim* o, yb_params* p, float h, float maxH);
/* Contact lists of the last detection (slot-major, sorted by partner). */
void orc_contacts(const orc_sim* o, int32_t* numCols, int32_t* cols);
/* Step limits (per segment) and warm-start dx (3n) of the last substep. */
void orc_last_prologue(const orc_sim* o, double* stepLimit, double* dxWarm);
/* Residual-targeted mode (tol > 0): sweeps until the stopping residual of a
 * sweep is <= tol (that sweep's output is kept) or maxSweeps sweeps ran;
 * tol <= 0 runs p->numItr sweeps.  Residuals of sweep k:
 *   update (kind 0, default)  sqrt(sum_free |dx_k+1 - dx_k|^2 / sum_free l_i^2),
 *   Newton (kind 1)  |f(dx_k)| / |f(dx_0)|, 2-norms over free vertices, dx_0 the
 *     plain warm start, f the local Newton step's force -- at a vertex the step
 *     limit holds, H (dx' - dx) / accelerationRatio instead, the part of f the
 *     limit does not absorb. */
void orc_set_solver(orc_sim* o, double tol, int maxSweeps);
void orc_set_residual_kind(orc_sim* o, int kind);
int orc_last_sweeps(const orc_sim* o);
/* Residuals of the last sweep of the last substep (stopping one, update,
 * Newton), and |f(dx_0)|^2 of that substep. */
double orc_last_residual(const orc_sim* o);
double orc_last_update_residual(const orc_sim* o);
double orc_last_newton_residual(const orc_sim* o);
double orc_last_force_norm0(const orc_sim* o);

/* ---- the sweep's own functions, for closed-form and finite-difference tests ----
 * orc_set_contacts: contact lists (slot-major, numCols n, cols 128n).
 * orc_set_iterate: iterate dx (3n), inertial target y - x (3n), step limits (n);
 *   null skips an array.
 * orc_seg_terms: segment s's endpoint terms at dx (3n): f0[3] f1[3] H0[6] H1[6]
 *   (H packed xx yy zz xy xz yz).
 * orc_vertex_terms: free vertex v's assembled f[3], H[6], step limit, new
 *   iterate[3] and residual force[3] at dx.
 * orc_force_norm2: sum over free vertices of |residual force|^2 at dx. */
void orc_set_contacts(orc_sim* o, const int32_t* numCols, const int32_t* cols);
void orc_set_iterate(orc_sim* o, const double* dx, const double* y, const double* stepLimit);
void orc_seg_terms(const orc_sim* o, const yb_params* p, int s, const double* dx, double* out);
/* orc_frame_update: the frame update of one sweep (quaternionLambdaItr,
 * cosserat.cu:209-253) from the loaded frames and positions and the iterate
 * dx (3n), into qOut (4n); the state is not changed. */
void orc_frame_update(const orc_sim* o, const double* dx, double* qOut);
void orc_vertex_terms(const orc_sim* o, const yb_params* p, int v, const double* dx,
                      double* out);
double orc_force_norm2(const orc_sim* o, const yb_params* p, const double* dx);

/* ---- FP32 twin of the discrete stages (bit-exact with the device) ---- */
void orc32_warm_start(int n, const yb_params* p, const float* invMass, const float* vel3,
                      const float* lastVel3, float* dx3, float* y3);
void orc32_aabb(int n, const yb_params* p, const float* pos3, const uint32_t* flags,
                const float* dx3, float* aabb6);
void orc32_scene_box(int n, const float* aabb6, float* box6);
void orc32_morton(int n, const uint32_t* flags, const float* aabb6, const float* box6,
                  uint32_t* codes);
void orc32_sort(int n, const uint32_t* codes, int32_t* order);
int orc32_contacts(int n, const yb_params* p, const float* pos3, const uint32_t* flags,
                   const int32_t* cid, const float* dx3, const float* aabb6, int32_t* numCols,
                   int32_t* cols);

/* ---- FP64 pieces for closed-form tests ---- */
void orc_closest(const double* a0, const double* a1, const double* b0, const double* b1,
                 double* uv);
/* b(d) = -(d-1)^2 ln d and the force / Hessian factors the sweep uses. */
void orc_barrier(double d, double* out3);
void orc_solve3(const double* H6, const double* f, double* x);
int orc_configure(int n, yb_vertex* verts, double* qs, double* qRests, double density,
                  double* lRest, double* invMass);
void orc_rotor_mul(const double* a, const double* b, double* out);
void orc_rotor_rotate(const double* r, const double* v, double* out);
void orc_from_to(const double* from, const double* to, double* out);

#ifdef __cplusplus
}
#endif
#endif
