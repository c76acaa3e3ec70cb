/*
 * picgpu.h -- C-ABI of the MI355X (gfx950) particle loop for PICurv.
 *
 * This is the drop-in boundary for the reference's per-step particle loop.
 * Plain C types only; the context owns device memory, the caller owns host
 * buffers.  Every call is synchronous at return unless stated otherwise.
 * One host thread per context.  Return value: PICGPU_OK (0) or a negative
 * PICGPU_E_* code; picgpu_last_error() gives a message.  Per-particle
 * outcomes are statuses with the reference's enum values
 * (common.h:75-124), never errors.
 *
 * Reference entry points each call replaces (one MPI rank):
 *   picgpu_reset_statuses     <- ResetAllParticleStatuses        ParticleMotion.c:2149-2176
 *   picgpu_advect             <- UpdateAllParticlePositions      ParticleMotion.c:38-95
 *   picgpu_locate             <- LocateAllParticlesInGrid_TEST   ParticleMotion.c:1931-2135
 *                                 (walk: LocateParticleOrFindMigrationTarget_TEST
 *                                  walkingsearch.c:1094-1242)
 *   picgpu_remove_lost        <- CheckAndRemoveLostParticles     ParticleMotion.c:532-625
 *   picgpu_remove_out_of_bounds <- CheckAndRemoveOutOfBoundsParticles ParticleMotion.c:283-383
 *   picgpu_interpolate        <- InterpolateAllFieldsToSwarm     interpolation.c:1660-1711
 *   picgpu_scatter            <- ScatterAllParticleFieldsToEulerFields interpolation.c:2223-2313
 *                                 (ParticleCount: CalculateParticleCountPerCell
 *                                  ParticleMotion.c:931-1067)
 *   picgpu_step               <- the particle stages of AdvanceSimulation_TEST simulation.c:248-292
 * Additions with no reference counterpart: picgpu_deposit_faces (face
 * deposition), the RK2 option of picgpu_step, the FP32 precision mode and
 * the MLP initial guess (picgpu_mlp_*).
 */
#ifndef PICGPU_H
#define PICGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- return codes ---- */
#define PICGPU_OK 0
#define PICGPU_E_ARG (-1)        /* bad argument (PETSC_ERR_ARG_*) */
#define PICGPU_E_STATE (-2)      /* call out of order (PETSC_ERR_ARG_WRONGSTATE) */
#define PICGPU_E_DEGENERATE (-3) /* |n| < 1e-12 on a face, walkingsearch.c:147-155 (PETSC_ERR_USER) */
#define PICGPU_E_HIP (-4)        /* HIP runtime error (PETSC_ERR_LIB) */
#define PICGPU_E_NOMEM (-5)      /* device allocation failed (PETSC_ERR_MEM) */
#define PICGPU_E_UNSUPPORTED (-6)/* (PETSC_ERR_SUP) */
#define PICGPU_E_SINGULAR (-7)   /* near-zero Jacobian (Metric.c:555-557) or metric determinant
                                    (setup.c:1366-1368) (PETSC_ERR_FLOP_COUNT) */
#define PICGPU_E_GRID (-8)       /* mixed Jacobian signs (Metric.c:217-220) (PETSC_ERR_USER) */

/* ---- particle status (common.h:75-124) ---- */
#define PICGPU_NEEDS_LOCATION 0
#define PICGPU_ACTIVE_AND_LOCATED 1
#define PICGPU_MIGRATING_OUT 2
#define PICGPU_LOST 3
#define PICGPU_UNINITIALIZED 4

/* ---- precision modes ---- */
#define PICGPU_FP64 0 /* plain: reference arithmetic, bit-exact cells/status/weights */
#define PICGPU_FP32 1 /* FP32 weights/interp/advect/deposit; FP64 positions and near-face tests */

/* ---- grid fields for picgpu_download_grid_field ---- */
#define PICGPU_FIELD_PARTICLE_COUNT 0 /* ParticleCount, node slots [(KM+1)(JM+1)(IM+1)] */
#define PICGPU_FIELD_P_AVERAGE 1      /* P scatter result, node slots */
#define PICGPU_FIELD_CORNER_VELOCITY 2/* centre->corner ucat, node slots x3 */
#define PICGPU_FIELD_IFACE 3          /* [KM][JM][IM+1] x4 (count, mx, my, mz) */
#define PICGPU_FIELD_JFACE 4          /* [KM][JM+1][IM] x4 */
#define PICGPU_FIELD_KFACE 5          /* [KM+1][JM][IM] x4 */
#define PICGPU_FIELD_IFACE_VOLFRAC 6  /* [KM][JM][IM+1]: count * V_p / V_dual */
#define PICGPU_FIELD_JFACE_VOLFRAC 7
#define PICGPU_FIELD_KFACE_VOLFRAC 8
/* Eulerian fields (picgpu_compute_metrics first), node slots [(KM+1)(JM+1)(IM+1)] */
#define PICGPU_FIELD_UCAT 9           /* Cartesian velocity Ucat x3 (ghost planes included) */
#define PICGPU_FIELD_UCONT 10         /* contravariant velocity Ucont x3 */
#define PICGPU_FIELD_CSI 11           /* i-face metric vectors x3 */
#define PICGPU_FIELD_ETA 12           /* j-face metric vectors x3 */
#define PICGPU_FIELD_ZET 13           /* k-face metric vectors x3 */
#define PICGPU_FIELD_AJ 14            /* inverse Jacobian */

/* ---- picgpu_step flags ---- */
#define PICGPU_STEP_RK2 0x1           /* Heun predictor-corrector instead of forward Euler */
#define PICGPU_STEP_DEPOSIT_CELLS 0x2 /* ParticleCount + P average */
#define PICGPU_STEP_DEPOSIT_FACES 0x4 /* face count + momentum */
#define PICGPU_STEP_FIELDS 0x8        /* each step first prepares Ucat as SetEulerianFields does for a
                                         step after the first (simulation.c:78-92: boundaries,
                                         Contra2Cart, boundaries); needs picgpu_boundary_create */
#define PICGPU_STEP_STAGED 0x100      /* run the stage kernels one by one (reference stage order)
                                         instead of the fused one-pass step; same results */
#define PICGPU_STEP_FACES_OF_START 0x20 /* with DEPOSIT_FACES, fused path, nsteps >= 2: the faces hold
                                         the deposition of the state at the start of the call's last
                                         step (each step deposits the state it starts from; the call's
                                         end state is then not deposited) */

typedef struct picgpu_ctx picgpu_ctx;

typedef struct {
  int precision;            /* PICGPU_FP64 or PICGPU_FP32 */
  int use_mlp_guess;        /* 1: MLP start cell for particles with no valid cell */
  int max_traversal;        /* walk cap, reference MAX_TRAVERSAL = 1000 (walkingsearch.c:10) */
  double distance_threshold;/* 1e-11 (walkingsearch.c:11) */
  double weight_clamp;      /* 1e-14 (ParticleSwarm.c:5) */
  double bbox_tolerance;    /* 1e-6 (grid.c:5) */
  int device;               /* HIP device ordinal */
  int collect_walk_stats;   /* 1: record per-particle walk iterations of the last locate */
} picgpu_options;

typedef struct {
  int64_t n_particles;      /* current swarm size */
  int64_t located;          /* ACTIVE_AND_LOCATED after the last locate */
  int64_t lost;             /* LOST after the last locate */
  int64_t walk_iterations;  /* sum of walk iterations in the last locate, or in the last
                               step of picgpu_step (both walks of an RK2 step) */
  int64_t walk_max;         /* max walk iterations of one particle in the last locate */
  int64_t fp64_fallbacks;   /* FP32 mode: faces re-evaluated in FP64 in the last locate
                               (or the last step of picgpu_step) */
  int64_t removed_lost;     /* by the last remove calls */
  int64_t removed_out_of_bounds;
  int64_t particle_steps;   /* particles processed, summed over picgpu_step steps */
  /* collect_walk_stats=1: walk iterations per particle in the last fused step
   * (both walks of an RK2 step) or the last locate; bin b counts particles
   * with b iterations, the last bin those with PICGPU_WALK_HIST_BINS-1 or more */
  int64_t walk_hist[64];
} picgpu_stats;
#define PICGPU_WALK_HIST_BINS 64

/* ---- lifetime ---- */
void picgpu_default_options(picgpu_options *opt);
int picgpu_create(picgpu_ctx **ctx, const picgpu_options *opt);
int picgpu_destroy(picgpu_ctx *ctx);
const char *picgpu_last_error(void);
int picgpu_synchronize(picgpu_ctx *ctx);

/* ---- grid and Eulerian input ---- */
/* coor: [(KM+1)(JM+1)(IM+1)][3] doubles, k slowest (the fda coordinate Vec, grid.c:386-395) */
int picgpu_upload_grid(picgpu_ctx *ctx, int64_t IM, int64_t JM, int64_t KM, const double *coor);
/* ucat: node-slot layout [(KM+1)(JM+1)(IM+1)][3]; cell (i,j,k), i<IM, j<JM, k<KM,
 * is stored at slot (i,j,k) (README.md:242-260); other slots are ignored. */
int picgpu_upload_ucat(picgpu_ctx *ctx, const double *ucat);

/* Same inputs already resident in device memory (same layouts; pointers are
 * HIP device pointers on the context's device).  For callers that build the
 * grid and field on the GPU. */
int picgpu_upload_grid_device(picgpu_ctx *ctx, int64_t IM, int64_t JM, int64_t KM, const double *d_coor);
int picgpu_upload_ucat_device(picgpu_ctx *ctx, const double *d_ucat);
/* Fresh particles at device positions d_pos [n][3]: cell -1, status
 * UNINITIALIZED, velocity and weights 0, P 101325, PID pid_base + p
 * (InitializeParticleBasicProperties, ParticleSwarm.c:448-454). */
int picgpu_init_particles_device(picgpu_ctx *ctx, int64_t n, const double *d_pos, int64_t pid_base);

/* ---- swarm I/O (DMSwarm field layout, ParticleSwarm.c:64-96) ---- */
/* pos, vel, weight: [n][3]; cell: [n][3] int64 (-1 = none); status: [n] int64;
 * pid: [n] int64; P: [n] (NULL -> 101325.0, ParticleSwarm.c:514-518).
 * vel, weight may be NULL (-> 0). */
int picgpu_upload_particles(picgpu_ctx *ctx, int64_t n, const double *pos, const double *vel,
                            const int64_t *cell, const double *weight, const int64_t *status,
                            const int64_t *pid, const double *P);
int picgpu_num_particles(picgpu_ctx *ctx, int64_t *n);
/* any output pointer may be NULL; arrays must hold the current n particles */
int picgpu_download_particles(picgpu_ctx *ctx, double *pos, double *vel, int64_t *cell,
                              double *weight, int64_t *status, int64_t *pid);
/* per-particle walk iterations of the last locate (collect_walk_stats=1) */
int picgpu_download_walk_steps(picgpu_ctx *ctx, int32_t *steps);

/* ---- stage-wise entry points (reference names in the header comment) ---- */
int picgpu_reset_statuses(picgpu_ctx *ctx);
int picgpu_advect(picgpu_ctx *ctx, double dt);
int picgpu_locate(picgpu_ctx *ctx);
int picgpu_remove_lost(picgpu_ctx *ctx, int64_t *removed);
int picgpu_remove_out_of_bounds(picgpu_ctx *ctx, int64_t *removed);
int picgpu_interpolate(picgpu_ctx *ctx);
int picgpu_scatter(picgpu_ctx *ctx);
int picgpu_deposit_faces(picgpu_ctx *ctx, double particle_volume);
/* Particle volume V_p of the face volume fractions (PICGPU_FIELD_*FACE_VOLFRAC
 * = count * V_p / V_dual); default 1.  picgpu_deposit_faces sets it too.
 * V_p <= 0 or not finite: PICGPU_E_ARG. */
int picgpu_set_particle_volume(picgpu_ctx *ctx, double particle_volume);

/* ---- a locate that allocates nothing ----
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): the walk of picgpu_locate over n particles the caller keeps in
 * device memory, in the library's own layouts:
 *   d_pos    double [n][3] positions (read)
 *   d_cell   int32 [n] packed cell i + IM*(j + JM*k), -1 none: the start cell
 *            in (-1: (0,0,0), walkingsearch.c:611-616), the located cell out
 *   d_weight [n][3] weights out, double in PICGPU_FP64 contexts, float in PICGPU_FP32
 *   d_status int32 [n] statuses in and out; ACTIVE_AND_LOCATED particles are
 *            skipped, the others come out ACTIVE_AND_LOCATED or LOST
 * The walk's counters and iteration histogram live in a fixed 40 KiB device
 * array of the library (one per device, held until the process ends); the
 * call makes no hipMalloc, hipFree or allocating library call.  Calls that
 * overlap on one device race on that array.  The grid's face-plane records
 * must already exist (any picgpu_locate or picgpu_step on the grid builds
 * them): else PICGPU_E_STATE.  n above PICGPU_NOALLOC_MAX_PARTICLES:
 * PICGPU_E_ARG.  st NULL: the call returns once the kernels are queued on the
 * context's stream (picgpu_synchronize waits); st given: it waits and fills
 * located, lost, walk_iterations, walk_max, fp64_fallbacks, n_particles = n
 * and walk_hist, and reports a degenerate face as PICGPU_E_DEGENERATE. */
#define PICGPU_NOALLOC_MAX_PARTICLES 2147483646LL
int picgpu_locate_device_noalloc(picgpu_ctx *ctx, int64_t n, const double *d_pos, int32_t *d_cell, void *d_weight,
                                 int32_t *d_status, picgpu_stats *st);

/* ---- whole steps: reset, advect(+locate) [Euler|RK2], remove, interpolate, deposit ---- */
int picgpu_step(picgpu_ctx *ctx, double dt, int nsteps, unsigned flags);

/* ---- grid outputs ---- */
int picgpu_grid_field_size(picgpu_ctx *ctx, int which, int64_t *count);
int picgpu_download_grid_field(picgpu_ctx *ctx, int which, double *out);
/* The diagnostics CalculateParticleCountPerCell logs (ParticleMotion.c:1021-1026:
 * VecSum and VecMax of ParticleCount), reduced on the device instead of read
 * from a downloaded field: *total = the particles counted, *max_count = the
 * largest count in a cell, *max_index = the first node slot holding it, in
 * the flat node-slot order of PICGPU_FIELD_PARTICLE_COUNT (the index VecMax
 * reports).  Needs the counts as that field does (PICGPU_E_STATE before). */
int picgpu_particle_count_summary(picgpu_ctx *ctx, double *total, double *max_count, int64_t *max_index);
/* LOG_INTERPOLATION_ERROR (logging.c:533-580), the reference's known-answer
 * check of the interpolation: *percent = 100 |a - v|_2 / |a|_2 over every
 * component of every particle in the swarm (any status), v the particles'
 * velocity as last interpolated, a the analytical velocity of
 * SetAnalyticalSolution (AnalyticalSolution.c:932-962): sin of each position
 * component for finit 1, (0, 0, 1) for finit 0 (the two the reference
 * defines; others: PICGPU_E_ARG).  Reduced on the device in a fixed order;
 * no particles give NaN (0/0), as the reference's norms do. */
int picgpu_interpolation_error(picgpu_ctx *ctx, int finit, double *percent);

/* ---- statistics and timing ---- */
int picgpu_get_stats(picgpu_ctx *ctx, picgpu_stats *st);
/* Turns the walk statistics of picgpu_options.collect_walk_stats on or off
 * (they cost a little time, so a benchmark enables them outside its timed steps). */
int picgpu_walk_stats_enable(picgpu_ctx *ctx, int on);
/* Kernel timing with HIP events on the library's stream: enable, then read the
 * summed milliseconds and launch count per kernel name. */
int picgpu_timing_enable(picgpu_ctx *ctx, int on);
int picgpu_timing_query(picgpu_ctx *ctx, const char *kernel, double *total_ms, int64_t *launches);
int picgpu_timing_reset(picgpu_ctx *ctx);
/* Times only the kernels whose name starts with kernel_prefix (null or "":
 * all of them).  Each timed launch records two events on the stream; timing
 * one kernel keeps the other launches free of them. */
int picgpu_timing_select(picgpu_ctx *ctx, const char *kernel_prefix);

/* ---- Eulerian field preparation: the stage before the particle loop every step ----
 * Reference entry points each call replaces (one MPI rank):
 *   picgpu_compute_metrics     <- ComputeFaceMetrics, ComputeCellCenteredJacobianInverse,
 *                                 CheckAndFixGridOrientation   Metric.c:195-622 (setup.c:279-281)
 *   picgpu_parse_boundary_file <- ParseAllBoundaryConditions   io.c:2515-2680 (+ handler parameters,
 *                                 BC_Handlers.c:238-256, 647-664)
 *   picgpu_boundary_create     <- BoundarySystem_Create         Boundaries.c:467-565
 *   picgpu_set_initial_ucont   <- SetInitialInteriorField("Ucont") AnalyticalSolution.c:986-1176
 *   picgpu_boundary_execute    <- BoundarySystem_ExecuteStep    Boundaries.c:566-646
 *   picgpu_contra2cart         <- Contra2Cart                   setup.c:1294-1420
 *   picgpu_set_eulerian_fields <- SetEulerianFields             simulation.c:38-107
 * All FP64 in both precision modes; the particle stages read the Ucat these
 * calls produce (or the one picgpu_upload_ucat* gave, whichever came last). */
#define PICGPU_FACE_NEG_X 0 /* BCFace, common.h:202-206 */
#define PICGPU_FACE_POS_X 1
#define PICGPU_FACE_NEG_Y 2
#define PICGPU_FACE_POS_Y 3
#define PICGPU_FACE_NEG_Z 4
#define PICGPU_FACE_POS_Z 5
#define PICGPU_BCTYPE_UNDEFINED 0 /* BCType, common.h:209-212 */
#define PICGPU_BCTYPE_NOGRAD 1
#define PICGPU_BCTYPE_WALL 2
#define PICGPU_BCTYPE_SYMMETRY 3
#define PICGPU_BCTYPE_INLET 4
#define PICGPU_BCTYPE_OUTLET 5
#define PICGPU_BC_UNDEFINED 0 /* BCHandlerType, common.h:215-222 */
#define PICGPU_BC_NOGRAD_COPY_GHOST 1
#define PICGPU_BC_WALL_NOSLIP 2
#define PICGPU_BC_WALL_MOVING 3
#define PICGPU_BC_SYMMETRY_PLANE 4
#define PICGPU_BC_INLET_PARABOLIC 5
#define PICGPU_BC_INLET_CONSTANT_VELOCITY 6
#define PICGPU_BC_OUTLET_CONSERVATION 9

typedef struct {
  int type[6];              /* PICGPU_BCTYPE_* per face (PICGPU_FACE_* order) */
  int handler[6];           /* PICGPU_BC_* per face */
  double velocity[6][3];    /* constant_velocity parameters vx vy vz (default 0) */
  double vmax[6];           /* parabolic parameter vmax (default 1) */
} picgpu_boundary_config;

/* every face WALL / noslip, parameters at their defaults (io.c:2531-2537) */
void picgpu_default_boundary_config(picgpu_boundary_config *bc);
/* "<face> <type> <handler> [key=value ...]" lines, '#' comments; faces -Xi +Xi
 * -Eta +Eta -Zeta +Zeta, types WALL SYMMETRY INLET OUTLET NOGRAD, handlers
 * noslip constant_velocity conservation allcopy parabolic (case-insensitive).
 * Unknown names, an invalid type/handler pair or an unreadable file: PICGPU_E_ARG. */
int picgpu_parse_boundary_file(const char *path, picgpu_boundary_config *bc);
/* Face metrics Csi/Eta/Zet, inverse Jacobian Aj, orientation check (a
 * left-handed grid is flipped; *orientation, may be NULL, receives the
 * reference's GridOrientation).  Sets Ucat and Ucont to zero (setup.c:250-268). */
int picgpu_compute_metrics(picgpu_ctx *ctx, int *orientation);
/* Stores the configuration and runs the inlet handlers' initial application.
 * Handlers other than allcopy, noslip, constant_velocity and parabolic:
 * PICGPU_E_UNSUPPORTED (Boundaries.c:402-466). */
int picgpu_boundary_create(picgpu_ctx *ctx, const picgpu_boundary_config *bc);
/* Ucont in the node-slot layout (e.g. read from a restart file) */
int picgpu_upload_ucont(picgpu_ctx *ctx, const double *ucont);
int picgpu_upload_ucont_device(picgpu_ctx *ctx, const double *d_ucont);
/* finit 0 zero, 1 constant, 2 parabolic profile (the -finit option); uin = -ucont_x/y/z */
int picgpu_set_initial_ucont(picgpu_ctx *ctx, int finit, const double uin[3]);
int picgpu_boundary_execute(picgpu_ctx *ctx);
int picgpu_contra2cart(picgpu_ctx *ctx);
/* step == start_step > 0 (restart): nothing (the caller uploaded Ucat and Ucont);
 * otherwise [step 0: initial Ucont], boundaries, Contra2Cart, boundaries. */
int picgpu_set_eulerian_fields(picgpu_ctx *ctx, int64_t step, int64_t start_step, int finit, const double uin[3]);

/* ---- on-disk formats: PETSc binary Vec files and the FDFGRID grid file ----
 * Host-side; the Vec files are PETSc 3.19's binary format (big-endian
 * VEC_FILE_CLASSID 1211214 and length as 64-bit PetscInt, then big-endian
 * doubles, plus "<file>.info"); 32-bit-PetscInt headers are read too.
 *   picgpu_read_grid_file   <- ReadGridFile io.c:101-176 + the file read of
 *                              AssignGridCoordinates grid.c:356-383
 *   picgpu_write_grid_file  <- none (writes the README.md:134-146 format)
 *   picgpu_field_file_name  <- the "results/<field>%05d_0.<ext>" names, io.c:680-682
 *   picgpu_petsc_vec_length <- the VecLoad size probe, ParticleMotion.c:1200-1228
 *   picgpu_read_petsc_vec   <- ReadFieldData io.c:253-436 (one rank)
 *   picgpu_write_petsc_vec  <- WriteFieldData io.c:657-745 (one rank)
 *   picgpu_write_output     <- WriteSwarmField(position, velocity) + WriteSimulationFields
 *                              simulation.c:304-306, io.c:887-936
 *   picgpu_read_restart     <- ReadSimulationFields io.c:437-488, ReadAllSwarmFields
 *                              io.c:1974-2057, PrepareLoadedSwarmForRelocation
 *                              ParticleSwarm.c:1073-1130
 * dir NULL or "" means "results".  Missing or mis-sized files: PICGPU_E_ARG. */
/* sizes of block 0 (and the block count); coor [(KM+1)(JM+1)(IM+1)][3] may be NULL */
int picgpu_read_grid_file(const char *path, int64_t *IM, int64_t *JM, int64_t *KM, int64_t *nblk, double *coor);
int picgpu_write_grid_file(const char *path, int64_t IM, int64_t JM, int64_t KM, const double *coor);
int picgpu_field_file_name(char *buf, size_t size, const char *dir, const char *field, int64_t step,
                           const char *ext);
int picgpu_petsc_vec_length(const char *path, int64_t *len);
int picgpu_read_petsc_vec(const char *path, int64_t len, double *out);
int picgpu_write_petsc_vec(const char *path, int64_t len, int block_size, const double *data);
int picgpu_write_output(picgpu_ctx *ctx, const char *dir, int64_t step);
int picgpu_read_restart(picgpu_ctx *ctx, const char *dir, int64_t step);
/* VTK XML output (io.c:1264-1815, postprocess.c:63-520): appended raw
 * little-endian blocks with 8-byte (PetscInt) byte counts, the VTKFile element
 * marked header_type="UInt64".  Structured grid: node coordinates and one
 * node field of 1 or 3 components; poly data: points, one point field and one
 * vertex per point.
 *   picgpu_write_vtk <- WriteEulerianVTK + WriteParticleVTK postprocess.c:451-520:
 *   <prefix>/Ucat_%05d.vts, P_, Ucont_, ParticleCount_ and velocity_%05d.vtp */
int picgpu_write_vts(const char *path, int64_t IM, int64_t JM, int64_t KM, const double *coor, const char *name,
                     int ncomp, const double *field);
int picgpu_write_vtp(const char *path, int64_t n, const double *pos, const char *name, int ncomp,
                     const double *field);
int picgpu_write_vtk(picgpu_ctx *ctx, const char *prefix, int64_t step);

/* ---- particle seeding, drawing the reference's random stream ----
 *   picgpu_seed_particles  <- InitializeParticleBasicProperties ParticleSwarm.c:340-499
 *                             (volume: DetermineVolumetricInitializationParameters :229-338;
 *                              inlet surface: GetRandomCellAndLogicOnInletFace Boundaries.c:224-390;
 *                              position: MetricLogicalToPhysical Metric.c:68-91)
 *   picgpu_reinit_on_inlet <- ReinitializeParticlesOnInletSurface ParticleMotion.c:1393-1517
 * The stream is the reference's PETSCRAND48 one: drand48 after srand48(seed);
 * the reference seeds it with 202403 + rank (ParticleSwarm.c:218).  Particle p
 * takes draws 6p..6p+5 (volume) or 4p..4p+3 (inlet face).  Seeded particles
 * are UNINITIALIZED with no cell, velocity 0, P 101325, PID pid_base + p;
 * the re-initialisation moves the current particles (in their array order)
 * onto the face and clears their cells, statuses unchanged. */
#define PICGPU_SEED_SURFACE 0 /* -pinit 0: on the inlet face */
#define PICGPU_SEED_VOLUME 1  /* -pinit 1: anywhere in the grid */
int picgpu_seed_particles(picgpu_ctx *ctx, int64_t n, int mode, int inlet_face, int64_t seed, int64_t pid_base);
int picgpu_reinit_on_inlet(picgpu_ctx *ctx, int inlet_face, int64_t seed);

/* ---- MLP initial guess (3 -> 64 -> 64 -> 3, hard GELU, BF16 weights, FP32 accumulation) ---- */
#define PICGPU_MLP_HIDDEN 64
/* number of FP32 parameters: W1[64][3] b1[64] W2[64][64] b2[64] W3[3][64] b3[3] */
int64_t picgpu_mlp_param_count(void);
int picgpu_mlp_set_params(picgpu_ctx *ctx, const float *params);
int picgpu_mlp_get_params(picgpu_ctx *ctx, float *params);
/* Train on (position, located cell + weights) pairs of the current swarm
 * (located particles only), Adam, minibatch `batch`, `iters` iterations. */
int picgpu_mlp_train(picgpu_ctx *ctx, int64_t max_pairs, int batch, int iters, float lr,
                     uint64_t seed, float *final_loss);
/* Device forward/backward for testing: x [n][3] normalised inputs, t [n][3]
 * targets; outputs y [n][3], loss and grad (FP32, param layout). */
int picgpu_mlp_forward(picgpu_ctx *ctx, int64_t n, const float *x, float *y);
int picgpu_mlp_loss_grad(picgpu_ctx *ctx, int64_t n, const float *x, const float *t, float *loss,
                         float *grad);
/* start cells the MLP proposes for positions pos [n][3]: cell [n][3] */
int picgpu_mlp_guess(picgpu_ctx *ctx, int64_t n, const double *pos, int64_t *cell);
/* use (1) or not (0) the MLP start cell for particles without a valid cell */
int picgpu_mlp_enable(picgpu_ctx *ctx, int on);

#ifdef __cplusplus
}
#endif
#endif /* PICGPU_H */
