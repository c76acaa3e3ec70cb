// picgpu_internal.hpp -- context and shared helpers of the gfx950 particle loop.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/picgpu.h"
#include "walk.hpp"

namespace picgpu {

void set_error(const char *fmt, ...);

#define PG_HIP(call)                                                                        \
  do {                                                                                      \
    hipError_t _e = (call);                                                                 \
    if (_e != hipSuccess) {                                                                 \
      ::picgpu::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(_e), __FILE__, __LINE__); \
      return PICGPU_E_HIP;                                                                  \
    }                                                                                       \
  } while (0)

#define PG_CHECK(expr)          \
  do {                          \
    int _r = (expr);            \
    if (_r != PICGPU_OK) return _r; \
  } while (0)

// A device buffer that grows on demand.
struct DevBuf {
  void *ptr = nullptr;
  size_t bytes = 0;
  int ensure(size_t want) {
    if (want <= bytes && ptr) return PICGPU_OK;
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
    if (want == 0) want = 256;
    if (hipMalloc(&ptr, want) != hipSuccess) {
      (void)hipGetLastError();
      set_error("hipMalloc(%zu) failed", want);
      return PICGPU_E_NOMEM;
    }
    bytes = want;
    return PICGPU_OK;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
  template <typename T>
  T *as() const {
    return static_cast<T *>(ptr);
  }
};

// Particle arrays (structure of arrays).  vel, w and P hold `real` elements:
// double in PICGPU_FP64 mode, float in PICGPU_FP32 mode.
struct SwarmBufs {
  DevBuf pos;     // double [n][3]
  DevBuf vel;     // real [n][3]
  DevBuf w;       // real [n][3]
  DevBuf cell;    // int32 [n], packed i + IM*(j + JM*k), -1 = none
  DevBuf status;  // int32 [n]
  DevBuf pid;     // int64 [n]
  DevBuf P;       // real [n]
  DevBuf pid32;   // int32 [n]: the PID's low 32 bits, while picgpu_ctx::pid32_valid
};

enum Counter : int {
  C_WALK_ITERS = 0,
  C_WALK_MAX = 1,
  C_FALLBACKS = 2,
  C_LOCATED = 3,
  C_LOST = 4,
  C_ERROR = 5,
  C_KEEP = 6,
  C_REMOVED_LOST = 7,
  C_REMOVED_OOB = 8,
  C_NUM = 16
};

// Device counters are striped over kCounterStripes copies (by workgroup
// number), so that the per-wavefront atomics of a large launch do not all
// meet on one address; the host sums the stripes.
constexpr int kCounterStripes = 64;
__device__ __forceinline__ unsigned long long *counter_stripe(unsigned long long *counters) {
  return counters + (blockIdx.x & (kCounterStripes - 1)) * C_NUM;
}

struct TimerRec {
  std::string name;
  hipEvent_t a, b;
};

}  // namespace picgpu

struct picgpu_ctx {
  picgpu_options opt;
  hipStream_t stream = nullptr;
  int device = 0;
  size_t rsize = 8;  // sizeof(real)

  // grid
  bool have_grid = false;
  int IM = 0, JM = 0, KM = 0;
  int64_t ncell = 0, nnode = 0;
  picgpu::DevBuf coor;  // double [nnode][3]
  double bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0};
  picgpu::DevBuf planes;  // [ncell] CellPlanes64 (FP64) or CellPlanes32 (FP32)
  bool have_planes = false;

  // Eulerian fields
  bool have_ucat = false;
  picgpu::DevBuf ucat;    // real [ncell][3] (compact, cell-linear)
  picgpu::DevBuf corner;  // real [nnode][3]

  // Eulerian field preparation (eulerian.hip); node-slot FP64 arrays
  struct {
    picgpu::DevBuf csi, eta, zet;  // double [nnode][3] face metric vectors
    picgpu::DevBuf aj;             // double [nnode] inverse Jacobian
    picgpu::DevBuf ucont, ucat;    // double [nnode][3] the reference's Ucont and Ucat
    picgpu::DevBuf flag;           // uint64 [2]: first singular node / scratch
    picgpu::DevBuf part;           // double [blocks][2] min/max partials
    bool have_metrics = false;     // metrics computed for the current grid
    bool have_bc = false;          // boundary system created
    int orientation = 1;           // GridOrientation
    bool flipped = false;          // a left-handed grid's metrics were negated
    picgpu_boundary_config bc{};
  } eul;

  // swarm (double-buffered for sorting / compaction)
  int64_t n = 0;
  picgpu::SwarmBufs sw, alt;
  picgpu::DevBuf rank;      // int32 [n] slot within the cell
  picgpu::DevBuf keep;      // int32 [n+1] keep flags / compaction offsets
  picgpu::DevBuf walk_steps;// int32 [n]
  picgpu::DevBuf walk_hist; // uint64 [kCounterStripes][PICGPU_WALK_HIST_BINS]
  picgpu::DevBuf guess;     // int32 [n] MLP start cell
  picgpu::DevBuf pos0, vel0;// RK2 saved state
  bool sorted = false;      // particles sorted by cell, cell_start valid
  int64_t n_live = 0;       // particles not removed (equals n outside picgpu_step)
  int steps_since_compact = 0;

  // fused step: index of the live particles by tile-major cell key
  // (double-buffered: the step reads one and builds the other)
  picgpu::DevBuf kcount;     // int32 [nkeys+2] histogram of the next index
  picgpu::DevBuf kstart[2];  // int32 [nkeys+2] bucket starts; bucket nkeys: live, no cell
  picgpu::DevBuf kidx[2];    // int32 [n] particle numbers in key order
  picgpu::DevBuf kslot;      // int32 [n] slot of each particle in its next bucket
  picgpu::DevBuf kbase;      // int32 [nkeys] a tile's base in each of its own buckets
  picgpu::DevBuf defer;      // int32 [n] particles deferred to the FP64 kernel, then their number
  picgpu::DevBuf step_counters;  // uint64 [steps of a batch][kCounterStripes][C_NUM]
  picgpu::DevBuf seams;      // real: halves of the faces between two tiles (FaceOut)
  int kcur = 0;              // index in use
  bool keyed = false;        // kidx[kcur] / kstart[kcur] describe the current swarm
  bool key_ordered = false;  // the swarm arrays are (nearly) in key order: compacted since the last change
  // sw.pid32 holds the low 32 bits of sw.pid, and pid_wide (int32 on the
  // device) is nonzero if some PID lies outside [0, 2^31): the PID order of
  // the buckets then compares the 64-bit PIDs
  bool pid32_valid = false;
  picgpu::DevBuf pid_wide;
  // a step call ended between compactions: kidx[kcur] lists the n_live live
  // particles and the arrays (length n) still hold removed ones; settle_swarm
  // compacts them before anything else reads or changes the arrays
  bool compact_pending = false;
  // a fused step met a degenerate face part-way: the arrays hold a mix of
  // that step's results and the state it started from; every entry point that
  // reads or changes the particles fails (PICGPU_E_STATE) until new particles
  // are uploaded, initialised or seeded
  bool swarm_invalid = false;
  int64_t kbegin_unloc = 0, kend_unloc = 0;  // range of bucket nkeys in kidx[kcur]

  // per-cell / per-face outputs
  picgpu::DevBuf cell_count;  // int32 [ncell+1]
  picgpu::DevBuf count_summary;  // k_count_summary's two u64, or k_interp_error_sums's block sums
  picgpu::DevBuf cell_start;  // int32 [ncell+1]
  picgpu::DevBuf pavg;        // real [ncell]
  picgpu::DevBuf face[3];     // real [nface_d][4]
  bool have_pavg = false, have_faces = false;
  double particle_volume = 1.0;

  // scratch for scans
  picgpu::DevBuf scan_tmp;

  // counters
  picgpu::DevBuf counters;  // uint64 [kCounterStripes][C_NUM]
  picgpu_stats stats{};

  // timing
  bool timing = false;
  std::string timing_only;  // when not empty: only kernels whose name starts with it are timed
  std::vector<picgpu::TimerRec> pending;
  std::vector<hipEvent_t> event_pool;
  std::map<std::string, std::pair<double, int64_t>> timing_acc;

  // MLP
  picgpu::DevBuf mlp_master;  // float [P]
  picgpu::DevBuf mlp_bf16;    // packed bf16 weights for kernels
  picgpu::DevBuf mlp_adam_m, mlp_adam_v, mlp_grad, mlp_partial;
  picgpu::DevBuf mlp_x, mlp_t, mlp_y;
  bool have_mlp = false;
  double mlp_lo[3] = {0, 0, 0}, mlp_scale[3] = {1, 1, 1};
};

namespace picgpu {

// kernel timing around a launch
hipEvent_t take_event(picgpu_ctx *c);
void timer_begin(picgpu_ctx *c, const char *name, TimerRec &rec);
void timer_end(picgpu_ctx *c, TimerRec &rec);
int flush_timers(picgpu_ctx *c);

// device exclusive scan of int32 (n entries -> out[0..n], out[n] = total)
int scan_exclusive_i32(picgpu_ctx *c, const int32_t *in, int32_t *out, int64_t n);

// Compacts a swarm that a step call left uncompacted (compact_pending); every
// entry point that reads or changes the particle arrays calls it first.
int settle_swarm(picgpu_ctx *c);
// particles of the swarm: n_live while a compaction is pending, else n
int64_t live_particles(const picgpu_ctx *c);

// Eulerian field preparation (eulerian.hip)
void eulerian_release(picgpu_ctx *c);
// device pointer and element count of an Eulerian node field (PICGPU_FIELD_UCAT..AJ)
int eulerian_field(picgpu_ctx *c, int which, const double **d_ptr, int64_t *count);
// SetEulerianFields of a step after the first, launched without waiting;
// eulerian_fields_begin checks the state and clears the error report,
// eulerian_fields_end waits and reports a singular cell system
int eulerian_fields_begin(picgpu_ctx *c);
int eulerian_fields_async(picgpu_ctx *c);
int eulerian_fields_end(picgpu_ctx *c);
// keeps the Eulerian Ucat (when metrics exist) equal to a freshly uploaded
// node-slot Ucat; kind: hipMemcpyHostToDevice or hipMemcpyDeviceToDevice
int eulerian_take_ucat(picgpu_ctx *c, const double *ucat, hipMemcpyKind kind);

// the device address of picgpu_locate_device_noalloc's scratch array on
// `device` (the current device), looked up once per device (noalloc.hip)
int noalloc_setup(int device);

// MLP
int mlp_guess_device(picgpu_ctx *c, int64_t n, const double *d_pos, const int32_t *d_cell_or_null,
                     int32_t *d_guess);

}  // namespace picgpu

#define PG_LAUNCH(ctx, name, grid, block, kernel, ...)                                    \
  do {                                                                                    \
    ::picgpu::TimerRec _rec;                                                              \
    ::picgpu::timer_begin((ctx), (name), _rec);                                           \
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, (ctx)->stream, __VA_ARGS__);   \
    hipError_t _le = hipGetLastError();                                                   \
    if (_le != hipSuccess) {                                                              \
      ::picgpu::set_error("launch %s failed: %s", (name), hipGetErrorString(_le));        \
      return PICGPU_E_HIP;                                                                \
    }                                                                                     \
    ::picgpu::timer_end((ctx), _rec);                                                     \
  } while (0)
