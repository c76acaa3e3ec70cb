// noalloc.hip -- a locate entry point that allocates no device memory.
//
// picgpu_locate_device_noalloc runs the walk of picgpu_locate (k_locate) over
// particles the caller keeps in device memory.  The only memory beyond the
// caller's inputs and outputs is the walk's counters and iteration
// histogram; they live in `scratch`, a fixed array in the library's device
// code object, so the call makes no hipMalloc/hipFree and no allocating
// library call.  Each device the library runs on holds its own copy of the
// array until the process ends; its device address is looked up once per
// device, when the first context on that device is created.
#include <string.h>

#include <mutex>

#include "kernels.hpp"

namespace picgpu {

// counters [kCounterStripes][C_NUM] then the walk-iteration histogram
// [kCounterStripes][PICGPU_WALK_HIST_BINS]: 64 x 16 + 64 x 64 words, 40 KiB
constexpr size_t kScratchCounterWords = (size_t)kCounterStripes * C_NUM;
constexpr size_t kScratchWords = kScratchCounterWords + (size_t)kCounterStripes * PICGPU_WALK_HIST_BINS;
static __device__ __align__(256) unsigned long long scratch[kScratchWords];

constexpr int kMaxDevices = 64;
static unsigned long long *g_scratch_dev[kMaxDevices];  // device address of `scratch`, per device
static std::mutex g_scratch_mu;

// With `device` current: the device address of `scratch` on it.
int noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) {
    set_error("device %d beyond the %d devices the scratch table holds", device, kMaxDevices);
    return PICGPU_E_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_dev[device]) return PICGPU_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    set_error("hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
    return PICGPU_E_HIP;
  }
  g_scratch_dev[device] = static_cast<unsigned long long *>(p);
  return PICGPU_OK;
}

static unsigned long long *scratch_on(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_dev[device];
}

// zero fill (a kernel, so that a caller may record the call in a graph)
__global__ __launch_bounds__(kBlock) void k_zero_u64(unsigned long long *__restrict__ p, int64_t n) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) p[q] = 0ull;
}

}  // namespace picgpu

using namespace picgpu;

extern "C" int picgpu_locate_device_noalloc(picgpu_ctx *c, int64_t n, const double *d_pos, int32_t *d_cell,
                                            void *d_weight, int32_t *d_status, picgpu_stats *st) {
  if (!c || n < 0 || (n > 0 && (!d_pos || !d_cell || !d_weight || !d_status))) {
    set_error("locate_device_noalloc: bad arguments");
    return PICGPU_E_ARG;
  }
  if (n > PICGPU_NOALLOC_MAX_PARTICLES) {
    set_error("locate_device_noalloc: %lld particles, more than the %lld the 32-bit indices allow",
              (long long)n, (long long)PICGPU_NOALLOC_MAX_PARTICLES);
    return PICGPU_E_ARG;
  }
  if (!c->have_grid) {
    set_error("locate_device_noalloc: no grid uploaded");
    return PICGPU_E_STATE;
  }
  if (!c->have_planes) {
    set_error("locate_device_noalloc: the grid's face-plane records are not built yet (a picgpu_locate or "
              "picgpu_step on this grid builds them)");
    return PICGPU_E_STATE;
  }
  PG_HIP(hipSetDevice(c->device));
  unsigned long long *ptr = scratch_on(c->device);
  if (!ptr) {
    set_error("locate_device_noalloc: no scratch address for device %d", c->device);
    return PICGPU_E_STATE;
  }
  hipLaunchKernelGGL(k_zero_u64, dim3((unsigned)((kScratchWords + kBlock - 1) / kBlock)), dim3(kBlock), 0, c->stream,
                     ptr, (int64_t)kScratchWords);
  PG_HIP(hipGetLastError());
  if (n > 0) {
    const int64_t chunk = 256;  // particles per wavefront, as picgpu_locate
    const int64_t waves = (n + chunk - 1) / chunk;
    const int64_t blocks = (waves * 64 + kBlock - 1) / kBlock;
    auto launch = [&](auto *tag) -> int {
      using R = std::remove_pointer_t<decltype(tag)>;
      LocateArgs<R> a;
      a.g = GridView{c->coor.as<double>(), c->IM, c->JM, c->KM, 1.0 / c->IM, 1.0 / ((double)c->IM * (double)c->JM)};
      a.planes = c->planes.ptr;
      a.n = n;
      a.chunk = chunk;
      a.pos = d_pos;
      a.cell = d_cell;
      a.w = static_cast<R *>(d_weight);
      a.status = d_status;
      a.guess = nullptr;
      a.order = nullptr;
      a.steps_out = nullptr;
      a.walk_hist = ptr + kScratchCounterWords;
      a.counters = ptr;
      a.max_traversal = c->opt.max_traversal;
      a.thresh_scale = c->opt.distance_threshold;
      a.clamp = (R)c->opt.weight_clamp;
      a.skip_lost = 0;
      constexpr bool EX = sizeof(R) == sizeof(double);
      hipLaunchKernelGGL((k_locate<R, EX>), dim3((unsigned)blocks), dim3(kBlock), 0, c->stream, a);
      PG_HIP(hipGetLastError());
      return PICGPU_OK;
    };
    PG_CHECK(c->opt.precision == PICGPU_FP32 ? launch((float *)nullptr) : launch((double *)nullptr));
  }
  if (!st) return PICGPU_OK;  // asynchronous: results after picgpu_synchronize
  unsigned long long h[kScratchWords];
  PG_HIP(hipMemcpyAsync(h, ptr, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  PG_HIP(hipStreamSynchronize(c->stream));
  unsigned long long v[C_NUM] = {0};
  for (int s = 0; s < kCounterStripes; ++s)
    for (int k = 0; k < C_NUM; ++k) {
      const unsigned long long x = h[(size_t)s * C_NUM + k];
      v[k] = k == C_WALK_MAX ? (x > v[k] ? x : v[k]) : v[k] + x;
    }
  memset(st, 0, sizeof(*st));
  st->n_particles = n;
  st->located = (int64_t)v[C_LOCATED];
  st->lost = (int64_t)v[C_LOST];
  st->walk_iterations = (int64_t)v[C_WALK_ITERS];
  st->walk_max = (int64_t)v[C_WALK_MAX];
  st->fp64_fallbacks = (int64_t)v[C_FALLBACKS];
  for (int b = 0; b < PICGPU_WALK_HIST_BINS; ++b) {
    int64_t t = 0;
    for (int s = 0; s < kCounterStripes; ++s) t += (int64_t)h[kScratchCounterWords + (size_t)s * PICGPU_WALK_HIST_BINS + b];
    st->walk_hist[b] = t;
  }
  if (v[C_ERROR]) {
    set_error("degenerate cell face (|n| < 1e-12) met during the walk");
    return PICGPU_E_DEGENERATE;
  }
  return PICGPU_OK;
}
