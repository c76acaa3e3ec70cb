// seed.hip -- particle seeding on gfx950, drawing the reference's random
// stream.  The reference places particles with PETSc's PETSCRAND48
// generators (ParticleSwarm.c:111-227): srand48 / drand48, one process-global
// POSIX 48-bit LCG (x' = 0x5DEECE66D x + 0xB mod 2^48) that its "i", "j" and
// "k" generators consume in call order, last seeded with 202403 + rank.
// Each thread jumps the LCG ahead to its particle's first draw (the draws per
// particle are fixed), so n particles are placed in parallel and every value
// equals the sequential stream's.
#include <math.h>

#include "picgpu_internal.hpp"

using namespace picgpu;

namespace {

constexpr uint64_t kLcgA = 0x5DEECE66DULL, kLcgC = 0xBULL, kMask48 = (1ULL << 48) - 1;
constexpr int kSBlock = 256;

// srand48(seed): the high 32 bits from the seed's low 32, low 16 bits 0x330E
inline uint64_t srand48_state(int64_t seed) { return ((uint64_t)(uint32_t)seed << 16 | 0x330EULL) & kMask48; }

// the LCG advanced by k steps: f^k(x), f(x) = a x + c, by squaring
__device__ __forceinline__ uint64_t lcg_jump(uint64_t x, uint64_t k) {
  uint64_t A = 1, C = 0, a = kLcgA, c = kLcgC;
  while (k) {
    if (k & 1) {
      A = (a * A) & kMask48;
      C = (a * C + c) & kMask48;
    }
    c = ((a + 1) * c) & kMask48;
    a = (a * a) & kMask48;
    k >>= 1;
  }
  return (A * x + C) & kMask48;
}

__device__ __forceinline__ double drand48_next(uint64_t &x) {
  x = (kLcgA * x + kLcgC) & kMask48;
  return (double)x * (1.0 / 281474976710656.0);  // x / 2^48, exact
}

// (PetscInt)(r * cells), clamped to [0, cells-1] (ParticleSwarm.c:262-270)
__device__ __forceinline__ int pick(double r, int n) {
  long long c = (long long)(r * (double)n);
  c = c < 0 ? 0 : c;
  return (int)(c > n - 1 ? n - 1 : c);
}

__device__ __forceinline__ double clamp_logic(double v) {  // Boundaries.c:364-374
  const double eps = 1.0e-7;
  v = v > 0.0 ? v : 0.0;
  return v < 1.0 - eps ? v : 1.0 - eps;
}

// MetricLogicalToPhysical: trilinear blend of the cell's vertices, corner c
// at (i + (c&1), j + (c>>1&1), k + (c>>2&1)) (Metric.c:24-91)
__device__ void blend(const double *X, int IM, int JM, int i, int j, int k, double xi, double eta, double zta,
                      double *out) {
  double x = 0.0, y = 0.0, z = 0.0;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int ii = i + (c & 1), jj = j + ((c >> 1) & 1), kk = k + ((c >> 2) & 1);
    const double *v = X + 3 * (((int64_t)kk * (JM + 1) + jj) * (IM + 1) + ii);
    const double N = ((c & 1) ? xi : 1.0 - xi) * ((c & 2) ? eta : 1.0 - eta) * ((c & 4) ? zta : 1.0 - zta);
    x += N * v[0];
    y += N * v[1];
    z += N * v[2];
  }
  out[0] = x;
  out[1] = y;
  out[2] = z;
}

// mode 1: volumetric, draws r_i r_j r_k xi eta zta (ParticleSwarm.c:229-338);
// mode 0: inlet face `face`, draws as GetRandomCellAndLogicOnInletFace
// (Boundaries.c:255-360): two cell draws and two logical draws, the face's
// own logical coordinate 1e-6 inside the face.
__global__ __launch_bounds__(kSBlock) void k_seed(int64_t n, int mode, int face, uint64_t x0, int IM, int JM, int KM,
                                                  const double *__restrict__ X, double *__restrict__ pos) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int draws = mode == 1 ? 6 : 4;
  uint64_t x = lcg_jump(x0, (uint64_t)p * draws);
  int ci, cj, ck;
  double xi, eta, zta;
  if (mode == 1) {
    ci = pick(drand48_next(x), IM);
    cj = pick(drand48_next(x), JM);
    ck = pick(drand48_next(x), KM);
    xi = drand48_next(x);
    eta = drand48_next(x);
    zta = drand48_next(x);
  } else {
    const int axis = face / 2, hi = face % 2;
    const double edge = hi ? 1.0 - 1.0e-6 : 1.0e-6;
    if (axis == 0) {
      ci = hi ? IM - 1 : 0;
      xi = edge;
      cj = pick(drand48_next(x), JM);
      ck = pick(drand48_next(x), KM);
      eta = clamp_logic(drand48_next(x));
      zta = clamp_logic(drand48_next(x));
    } else if (axis == 1) {
      cj = hi ? JM - 1 : 0;
      eta = edge;
      ci = pick(drand48_next(x), IM);
      ck = pick(drand48_next(x), KM);
      xi = clamp_logic(drand48_next(x));
      zta = clamp_logic(drand48_next(x));
    } else {
      ck = hi ? KM - 1 : 0;
      zta = edge;
      ci = pick(drand48_next(x), IM);
      cj = pick(drand48_next(x), JM);
      xi = clamp_logic(drand48_next(x));
      eta = clamp_logic(drand48_next(x));
    }
  }
  blend(X, IM, JM, ci, cj, ck, xi, eta, zta, pos + 3 * p);
}

// ReinitializeParticlesOnInletSurface's update of one particle: new position,
// no cell (ParticleMotion.c:1462-1476); the status is left as it is
__global__ __launch_bounds__(kSBlock) void k_replace_positions(int64_t n, const double *__restrict__ src,
                                                               double *__restrict__ pos, int32_t *__restrict__ cell) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  pos[3 * p] = src[3 * p];
  pos[3 * p + 1] = src[3 * p + 1];
  pos[3 * p + 2] = src[3 * p + 2];
  cell[p] = -1;
}

int seed_positions(picgpu_ctx *c, int64_t n, int mode, int face, int64_t seed, DevBuf &tmp) {
  if (mode != 0 && mode != 1) {
    set_error("seeding mode %d: 0 (inlet surface) or 1 (volume)", mode);
    return PICGPU_E_ARG;
  }
  if (mode == 0 && (face < 0 || face > 5)) {
    set_error("inlet face %d out of range 0..5", face);  // Boundaries.c:356-357
    return PICGPU_E_ARG;
  }
  PG_CHECK(tmp.ensure(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1)));
  if (n > 0)
    PG_LAUNCH(c, "seed", (n + kSBlock - 1) / kSBlock, kSBlock, k_seed, n, mode, face, srand48_state(seed), c->IM,
              c->JM, c->KM, c->coor.as<double>(), tmp.as<double>());
  return PICGPU_OK;
}

}  // namespace

extern "C" {

int picgpu_seed_particles(picgpu_ctx *c, int64_t n, int mode, int inlet_face, int64_t seed, int64_t pid_base) {
  if (!c || n < 0) return PICGPU_E_ARG;
  if (!c->have_grid) {
    set_error("no grid uploaded");
    return PICGPU_E_STATE;
  }
  PG_HIP(hipSetDevice(c->device));
  DevBuf tmp;
  int r = seed_positions(c, n, mode, inlet_face, seed, tmp);
  if (r == PICGPU_OK) r = picgpu_init_particles_device(c, n, tmp.as<double>(), pid_base);
  (void)hipStreamSynchronize(c->stream);
  tmp.release();
  return r;
}

int picgpu_reinit_on_inlet(picgpu_ctx *c, int inlet_face, int64_t seed) {
  if (!c) return PICGPU_E_ARG;
  if (!c->have_grid) {
    set_error("no grid uploaded");
    return PICGPU_E_STATE;
  }
  PG_HIP(hipSetDevice(c->device));
  PG_CHECK(settle_swarm(c));
  const int64_t n = c->n;
  DevBuf tmp;
  int r = seed_positions(c, n, 0, inlet_face, seed, tmp);
  if (r == PICGPU_OK && n > 0) {
    r = [&]() -> int {
      PG_LAUNCH(c, "replace_positions", (n + kSBlock - 1) / kSBlock, kSBlock, k_replace_positions, n,
                tmp.as<double>(), c->sw.pos.as<double>(), c->sw.cell.as<int32_t>());
      return PICGPU_OK;
    }();
  }
  (void)hipStreamSynchronize(c->stream);
  tmp.release();
  c->sorted = c->keyed = c->key_ordered = false;
  return r;
}

}  // extern "C"
