/*
 * picref.h -- CPU oracle for the PICurv per-step particle loop.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code; it is the checker used by tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py.
 *
 * It is a clean-room restatement, in plain C, of the reference's arithmetic
 * for the hot path (walking search, weights, centre->corner averaging,
 * trilinear interpolation, forward Euler, per-cell counts and the P average).
 * Each function cites the reference file:line it follows.  The parity build
 * is compiled with -O2 -ffp-contract=off so that every FP64 operation is a
 * single correctly rounded IEEE operation in the reference's order (the
 * reference is built -O0 without -march, Makefile:15-16, hence no FMA).
 *
 * Functions marked SPEC have no counterpart in the reference (face deposition,
 * RK2, the MLP initial guess); they state this project's own definitions in
 * FP64 and are pinned by closed forms and invariants in tests/, not by the
 * reference.
 */
#ifndef PICREF_H
#define PICREF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ParticleLocationStatus values (common.h:75-124) */
enum { REF_NEEDS_LOCATION = 0, REF_ACTIVE_AND_LOCATED = 1, REF_MIGRATING_OUT = 2,
       REF_LOST = 3, REF_UNINITIALIZED = 4 };
/* Face enum (common.h:129-137) */
enum { REF_LEFT = 0, REF_RIGHT = 1, REF_BOTTOM = 2, REF_TOP = 3, REF_FRONT = 4, REF_BACK = 5 };

/* error codes returned by the oracle */
enum { REF_OK = 0, REF_E_DEGENERATE = -3, REF_E_ARG = -1 };

typedef struct {
  int64_t IM, JM, KM;        /* cell counts; nodes are 0..IM etc. (grid.c:50-52) */
  const double *coor;        /* [(KM+1)(JM+1)(IM+1)][3], k slowest, i fastest */
} ref_grid_t;

/* ---- walking search primitives ---- */
void   ref_cell_vertices(const ref_grid_t *g, int64_t i, int64_t j, int64_t k, double V[8][3]);
double ref_cell_char_size(const double V[8][3]);
int    ref_signed_distance(const double v1[3], const double v2[3], const double v3[3],
                           const double v4[3], const double cc[3], const double p[3],
                           double thr, double *d);
int    ref_face_distances(const double V[8][3], const double p[3], double thr, double d[6]);
int    ref_point_position(const double d[6]);
int    ref_evaluate_position(const double V[8][3], const double p[3], double d[6], int *position);
void   ref_update_weights(double d[6], double w[3]);
void   ref_step_indices(const double d[6], int64_t *i, int64_t *j, int64_t *k);

/* Locate one particle (walkingsearch.c:1094-1242 on one rank).
 * cell[3] in: previous cell (or -1); out: located cell or -1.
 * w[3] out: weights (unchanged when LOST).
 * Returns the status (REF_ACTIVE_AND_LOCATED / REF_LOST) or a negative error.
 * *steps receives the number of walk iterations (cells evaluated or bounds-checked). */
int ref_locate_particle(const ref_grid_t *g, const double p[3], int64_t cell[3], double w[3],
                        int max_traversal, const int64_t *start_guess, int *steps);

/* Locate every particle whose status is not ACTIVE_AND_LOCATED
 * (ParticleMotion.c:1931-2135 on one rank).  guess may be NULL; when non-NULL,
 * guess[3p..3p+2] is the start cell used for particles without a valid cell
 * (the MLP slot, SPEC).  steps may be NULL. */
int ref_locate_all(const ref_grid_t *g, int64_t n, const double *pos, int64_t *cell, double *w,
                   int64_t *status, int max_traversal, const int64_t *guess, int32_t *steps);

/* ---- interpolation ---- */
/* centre->corner (interpolation.c:598-680); ucat and corner are node-slot
 * arrays [(KM+1)(JM+1)(IM+1)][3]; cell (i,j,k) lives at node slot (i,j,k). */
void ref_center_to_corner(int64_t IM, int64_t JM, int64_t KM, const double *ucat, double *corner);
/* trilinear (interpolation.c:839-864, 957-1042) */
void ref_trilinear(int64_t IM, int64_t JM, const double *corner, int64_t i, int64_t j, int64_t k,
                   double a1, double a2, double a3, double out[3]);
/* per-particle driver (interpolation.c:1341-1385): out-of-range cell -> 0 */
void ref_interp_all(int64_t IM, int64_t JM, int64_t KM, const double *corner, int64_t n,
                    const int64_t *cell, const double *w, double *vel);

/* ---- advection ---- */
void ref_advect_euler(int64_t n, double *pos, const double *vel, double dt); /* ParticleMotion.c:19-95 */
void ref_reset_statuses(int64_t n, int64_t *status);                         /* ParticleMotion.c:2149-2176 */

/* ---- removal ---- */
void ref_bbox(const ref_grid_t *g, double bmin[3], double bmax[3]);          /* grid.c:762-883 */
/* Removes LOST then out-of-bbox particles, compacting every non-NULL array in
 * place while keeping the survivors' relative order (the reference removes by
 * swapping the last point in; parity is compared per PID).  Returns new n. */
int64_t ref_remove(int64_t n, double *pos, double *vel, int64_t *cell, double *w, int64_t *status,
                   int64_t *pid, double *P, const double bmin[3], const double bmax[3],
                   int64_t *removed_lost, int64_t *removed_oob);

/* ---- deposition ---- */
/* ParticleCount (ParticleMotion.c:931-1067): count[k][j][i] += 1 at node slots */
void ref_count_cells(int64_t IM, int64_t JM, int64_t KM, int64_t n, const int64_t *cell, double *count);
/* P scatter: sum (interpolation.c:1838-1923) then normalise (1943-2036).
 * order[] (may be NULL) gives the particle order of the summation. */
void ref_scatter_P(int64_t IM, int64_t JM, int64_t KM, int64_t n, const int64_t *cell,
                   const double *P, const int64_t *order, const double *count, double *avg);
/* SPEC: face deposition.  Particle in cell (i,j,k) with weights (a1,a2,a3)
 * adds (1-a1) to i-face i and a1 to i-face i+1 (likewise j, k).  Outputs
 * count and momentum (3) per face.  A face is the sum of its lower cell's
 * contributions plus the sum of its upper cell's, each summed from zero in
 * ascending PID order (a missing cell contributes zero).
 * Layouts: iface [KM][JM][IM+1], jface [KM][JM+1][IM], kface [KM+1][JM][IM],
 * each entry 4 doubles (count, mx, my, mz). */
void ref_deposit_faces(int64_t IM, int64_t JM, int64_t KM, int64_t n, const int64_t *cell,
                       const double *w, const double *vel, const int64_t *pid,
                       double *iface, double *jface, double *kface);
/* SPEC: cell volume, six tetrahedra (V0, Va, Vb, V5) around the V0-V5
 * diagonal for the ring edges (1,2) (1,6) (3,2) (3,4) (7,6) (7,4), |det|
 * summed in that order, divided by 6. */
double ref_cell_volume(const double V[8][3]);
/* SPEC: face volume fraction of direction dir (0 i, 1 j, 2 k) from that
 * direction's face array (4 doubles per face, count first):
 * count * vp / V_dual, V_dual = 0.5 * (volume of the lower cell + volume of
 * the upper cell), lower first, a boundary face having one cell. */
void ref_face_volfrac(int64_t IM, int64_t JM, int64_t KM, const double *coor, int dir, const double *face,
                      double vp, double *out);

/* ---- full step drivers (AdvanceSimulation_TEST stage order, simulation.c:248-292) ---- */
typedef struct {
  int64_t IM, JM, KM;
  const double *coor;        /* node coordinates */
  const double *corner;      /* centre->corner result, node slots */
  int max_traversal;
  const double *bbox;        /* padded bounding box (min xyz, max xyz), or NULL: computed
                                from the nodes at every step (ref_bbox) */
} ref_ctx_t;

/* One reference step with forward Euler.  Arrays are compacted in place.
 * Returns the new particle count (or a negative error). */
int64_t ref_step_euler(const ref_ctx_t *c, int64_t n, double *pos, double *vel, int64_t *cell,
                       double *w, int64_t *status, int64_t *pid, double *P, double dt,
                       int64_t *removed);
/* SPEC: one Heun (RK2) step: x* = x + dt u; locate x* from cell_n; u*;
 * x1 = x + (u + u*)*(0.5 dt); locate x1 from cell*; remove; interpolate. */
int64_t ref_step_rk2(const ref_ctx_t *c, int64_t n, double *pos, double *vel, int64_t *cell,
                     double *w, int64_t *status, int64_t *pid, double *P, double dt,
                     int64_t *removed);

/* ---- Eulerian field preparation (eulerref.c) ----
 * Node-slot arrays [(KM+1)(JM+1)(IM+1)], vectors x3, i fastest. */
enum { REF_E_SINGULAR = -7, REF_E_GRID = -8 };
/* BCType (common.h:209-212) and BCHandlerType (common.h:215-222) values */
enum { REF_BCTYPE_WALL = 2, REF_BCTYPE_INLET = 4 };
enum { REF_BC_NOGRAD_COPY_GHOST = 1, REF_BC_WALL_NOSLIP = 2, REF_BC_INLET_PARABOLIC = 5,
       REF_BC_INLET_CONSTANT_VELOCITY = 6 };
typedef struct {
  int type[6];          /* BCType per face, faces -Xi +Xi -Eta +Eta -Zeta +Zeta */
  int handler[6];       /* BCHandlerType per face */
  double vel[6][3];     /* constant_velocity: vx vy vz */
  double vmax[6];       /* parabolic: vmax */
} ref_bc_t;
void ref_face_metrics(int64_t IM, int64_t JM, int64_t KM, const double *coor, double *csi, double *eta,
                      double *zet);                                               /* Metric.c:286-490 */
int ref_jacobian_inverse(int64_t IM, int64_t JM, int64_t KM, const double *coor, double *aj); /* Metric.c:491-622 */
int ref_grid_orientation(int64_t nnode, double *csi, double *eta, double *zet, double *aj,
                         int *orientation);                                       /* Metric.c:195-284 */
int ref_inlet_face(const ref_bc_t *bc);                                           /* io.c:2664-2672 */
void ref_initial_ucont(int64_t IM, int64_t JM, int64_t KM, const double *csi, const double *eta,
                       const double *zet, int finit, const double uin[3], int inlet_face, int orientation,
                       double *ucont);                                            /* AnalyticalSolution.c:986-1176 */
int ref_bc_initialize(int64_t IM, int64_t JM, int64_t KM, const ref_bc_t *bc, const double *csi,
                      const double *eta, const double *zet, double *ucat, double *ucont); /* Boundaries.c:467-565 */
int ref_bc_execute(int64_t IM, int64_t JM, int64_t KM, const ref_bc_t *bc, const double *csi,
                   const double *eta, const double *zet, double *ucat, double *ucont); /* Boundaries.c:566-646 */
int ref_contra2cart(int64_t IM, int64_t JM, int64_t KM, const double *ucont, const double *csi, const double *eta,
                    const double *zet, double *ucat);                             /* setup.c:1294-1420 */
int ref_set_eulerian_fields(int64_t IM, int64_t JM, int64_t KM, const ref_bc_t *bc, const double *csi,
                            const double *eta, const double *zet, int64_t step, int finit, const double uin[3],
                            int orientation, double *ucat, double *ucont);        /* simulation.c:38-107 */

/* ---- particle seeding (seedref.c) ---- */
uint64_t ref_srand48_state(int64_t seed);   /* POSIX srand48 */
double ref_drand48(uint64_t *state);        /* POSIX drand48 */
/* n particles from the stream srand48(seed): mode 1 volumetric
 * (ParticleSwarm.c:229-338), mode 0 on inlet face `face`
 * (Boundaries.c:224-390); positions by MetricLogicalToPhysical (Metric.c:68-91) */
void ref_seed_particles(int64_t IM, int64_t JM, int64_t KM, const double *coor, int64_t n, int mode, int face,
                        int64_t seed, double *pos);

/* ---- SPEC: MLP initial guess, 3 -> H -> H -> 3, GELU in its hard form z clamp(z/sqrt(2 pi) + 1/2, 0, 1) ---- */
/* params layout: W1[H][3], b1[H], W2[H][H], b2[H], W3[3][H], b3[3] (row = output) */
int64_t ref_mlp_param_count(int H);
/* forward; when round_bf16 != 0, weights and every matmul input are rounded to
 * bf16 (round-to-nearest-even) before the product, as the device does. */
void ref_mlp_forward(int H, const double *params, int64_t n, const double *x, double *y, int round_bf16);
/* mean-squared-error loss over n samples x 3 outputs and its gradient
 * w.r.t. params (same rounding rule).  Returns the loss. */
double ref_mlp_loss_grad(int H, const double *params, int64_t n, const double *x, const double *t,
                         double *grad, int round_bf16);

#ifdef __cplusplus
}
#endif
#endif
