/*
 * seedref.c -- CPU oracle for particle seeding (test infrastructure only).
 *
 * The reference places particles with PETSc's PETSCRAND48 generators
 * (ParticleSwarm.c:111-227).  In PETSc 3.19.2 that type seeds with srand48
 * and draws with drand48, i.e. one process-global POSIX 48-bit LCG stream;
 * every PetscRandomSeed re-seeds it, so after InitializeLogicalSpaceRNGs
 * the stream is srand48(202403 + rank), and the "i", "j" and "k" generators
 * consume it in call order.  PETSc is not in this image: the stream is
 * restated from the POSIX definition (x' = 0x5DEECE66D x + 0xB mod 2^48,
 * srand48(s) sets x = s << 16 | 0x330E, drand48() = x / 2^48) and checked in
 * tests/ against the C library's own drand48.
 *
 * Placement: a random cell (int)(r * cells) clamped to the range, logical
 * coordinates, then the trilinear blend of the cell's 8 vertices in the
 * order of MetricGetCellVertices (Metric.c:24-45, 51-66); volumetric draws
 * r_i r_j r_k xi eta zta (ParticleSwarm.c:229-338), inlet-face draws per
 * face as GetRandomCellAndLogicOnInletFace (Boundaries.c:224-390).
 */
#include "picref.h"

#define LCG_A 0x5DEECE66DULL
#define LCG_C 0xBULL
#define MASK48 ((1ULL << 48) - 1)

uint64_t ref_srand48_state(int64_t seed) { return (((uint64_t)(uint32_t)seed) << 16 | 0x330EULL) & MASK48; }

double ref_drand48(uint64_t *x) {
  *x = (LCG_A * *x + LCG_C) & MASK48;
  return (double)*x * (1.0 / 281474976710656.0); /* 2^-48, exact */
}

/* TrilinearBlend (Metric.c:51-66): corner c = (i + (c&1), j + (c>>1&1), k + (c>>2&1)) */
static void blend(int64_t IM, int64_t JM, const double *coor, int64_t i, int64_t j, int64_t k, double xi,
                  double eta, double zta, double out[3]) {
  double x = 0, y = 0, z = 0;
  for (int c = 0; c < 8; ++c) {
    const int64_t ii = i + ((c & 1) ? 1 : 0), jj = j + ((c & 2) ? 1 : 0), kk = k + ((c & 4) ? 1 : 0);
    const double *v = coor + 3 * ((kk * (JM + 1) + jj) * (IM + 1) + ii);
    const double N = ((c & 1) ? xi : 1.0 - xi) * ((c & 2) ? eta : 1.0 - eta) * ((c & 4) ? zta : 1.0 - zta);
    x += N * v[0];
    y += N * v[1];
    z += N * v[2];
  }
  out[0] = x;
  out[1] = y;
  out[2] = z;
}

static int64_t pick(double r, int64_t n) {
  int64_t c = (int64_t)(r * (double)n);
  c = c < 0 ? 0 : c;
  return c > n - 1 ? n - 1 : c;
}

static double clamp_logic(double v) {
  const double eps = 1.0e-7; /* Boundaries.c:364 */
  v = v > 0.0 ? v : 0.0;
  return v < 1.0 - eps ? v : 1.0 - eps;
}

/* one particle on inlet face f (BCFace order -Xi +Xi -Eta +Eta -Zeta +Zeta) */
static void inlet_one(int64_t IM, int64_t JM, int64_t KM, const double *coor, int f, uint64_t *x, double out[3]) {
  int64_t ci = 0, cj = 0, ck = 0;
  double xi = 0.5, eta = 0.5, zta = 0.5;
  const int axis = f / 2, pos = f % 2;
  const double edge = pos ? 1.0 - 1.0e-6 : 1.0e-6;
  if (axis == 0) {
    ci = pos ? IM - 1 : 0;
    xi = edge;
    cj = pick(ref_drand48(x), JM);
    ck = pick(ref_drand48(x), KM);
    eta = ref_drand48(x);
    zta = ref_drand48(x);
    eta = clamp_logic(eta);
    zta = clamp_logic(zta);
  } else if (axis == 1) {
    cj = pos ? JM - 1 : 0;
    eta = edge;
    ci = pick(ref_drand48(x), IM);
    ck = pick(ref_drand48(x), KM);
    xi = ref_drand48(x);
    zta = ref_drand48(x);
    xi = clamp_logic(xi);
    zta = clamp_logic(zta);
  } else {
    ck = pos ? KM - 1 : 0;
    zta = edge;
    ci = pick(ref_drand48(x), IM);
    cj = pick(ref_drand48(x), JM);
    xi = ref_drand48(x);
    eta = ref_drand48(x);
    xi = clamp_logic(xi);
    eta = clamp_logic(eta);
  }
  blend(IM, JM, coor, ci, cj, ck, xi, eta, zta, out);
}

void ref_seed_particles(int64_t IM, int64_t JM, int64_t KM, const double *coor, int64_t n, int mode, int face,
                        int64_t seed, double *pos) {
  uint64_t x = ref_srand48_state(seed);
  for (int64_t p = 0; p < n; ++p) {
    if (mode == 1) {
      const int64_t ci = pick(ref_drand48(&x), IM), cj = pick(ref_drand48(&x), JM), ck = pick(ref_drand48(&x), KM);
      const double xi = ref_drand48(&x), eta = ref_drand48(&x), zta = ref_drand48(&x);
      blend(IM, JM, coor, ci, cj, ck, xi, eta, zta, pos + 3 * p);
    } else {
      inlet_one(IM, JM, KM, coor, face, &x, pos + 3 * p);
    }
  }
}
