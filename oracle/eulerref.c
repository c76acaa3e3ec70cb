/*
 * eulerref.c -- CPU oracle for the Eulerian field preparation that precedes
 * the particle loop every step (SetEulerianFields, simulation.c:38-107):
 * grid metrics, the grid-orientation check, the initial contravariant field,
 * the four boundary handlers and the contravariant -> Cartesian conversion.
 *
 * TEST INFRASTRUCTURE ONLY (see picref.h).  Clean-room restatement of the
 * reference's arithmetic on one rank, where the DMDA (DM_BOUNDARY_NONE,
 * grid.c:59-60) has no ghost layer outside the domain and every local array
 * is the global one.  All arrays are node-slot arrays [mz][my][mx] with
 * mx = IM+1, my = JM+1, mz = KM+1 (i fastest); vectors hold 3 doubles.
 * Parity build: gcc -O2 -ffp-contract=off, so every expression rounds in the
 * reference's order (left to right, no FMA).
 */
#include "picref.h"

#include <math.h>
#include <stddef.h>

typedef struct {
  int64_t mx, my, mz;
} dims_t;

static inline int64_t at(const dims_t *d, int64_t i, int64_t j, int64_t k) { return (k * d->my + j) * d->mx + i; }

/* 0.5 * (a + b - c - d) per component, summed left to right */
static inline void half_diff(const double *a, const double *b, const double *c, const double *e, double out[3]) {
  for (int q = 0; q < 3; ++q) out[q] = 0.5 * (a[q] + b[q] - c[q] - e[q]);
}

/* u x v with the component expressions of Metric.c:339-341 */
static inline void cross(const double u[3], const double v[3], double *out) {
  out[0] = u[1] * v[2] - u[2] * v[1];
  out[1] = u[2] * v[0] - u[0] * v[2];
  out[2] = u[0] * v[1] - u[1] * v[0];
}

/* Boundary copy of one vector field along one axis (Metric.c:408-476): the
 * first node plane takes the second, then the last takes the one before it. */
static void copy_planes(const dims_t *d, double *f, int ncomp, int axis) {
  const int64_t n[3] = {d->mx, d->my, d->mz};
  for (int side = 0; side < 2; ++side) {
    const int64_t dst = side == 0 ? 0 : n[axis] - 1;
    const int64_t src = side == 0 ? 1 : n[axis] - 2;
    if (src < 0 || src >= n[axis]) continue;
    for (int64_t k = 0; k < d->mz; ++k)
      for (int64_t j = 0; j < d->my; ++j)
        for (int64_t i = 0; i < d->mx; ++i) {
          const int64_t idx[3] = {i, j, k};
          if (idx[axis] != dst) continue;
          int64_t s[3] = {i, j, k};
          s[axis] = src;
          const int64_t to = at(d, i, j, k), from = at(d, s[0], s[1], s[2]);
          for (int q = 0; q < ncomp; ++q) f[ncomp * to + q] = f[ncomp * from + q];
        }
  }
}

/* ComputeFaceMetrics, Metric.c:286-490.  csi on i-faces (all i, j>=1, k>=1),
 * eta on j-faces (i>=1, all j, k>=1), zet on k-faces (i>=1, j>=1, all k), each
 * the cross product of two half-differences of the face's four nodes; then the
 * boundary planes copied in the order x (eta, zet), y (csi, zet), z (csi, eta).
 * Entries the loops never reach keep their input values (the reference zeroes
 * them first, setup.c:256-258). */
void ref_face_metrics(int64_t IM, int64_t JM, int64_t KM, const double *coor, double *csi, double *eta,
                      double *zet) {
  const dims_t d = {IM + 1, JM + 1, KM + 1};
#define N(i_, j_, k_) (coor + 3 * at(&d, (i_), (j_), (k_)))
#pragma omp parallel for schedule(static)
  for (int64_t k = 1; k < d.mz; ++k)
    for (int64_t j = 1; j < d.my; ++j)
      for (int64_t i = 0; i < d.mx; ++i) {
        double deta[3], dzeta[3];
        half_diff(N(i, j, k), N(i, j, k - 1), N(i, j - 1, k), N(i, j - 1, k - 1), deta);
        half_diff(N(i, j - 1, k), N(i, j, k), N(i, j - 1, k - 1), N(i, j, k - 1), dzeta);
        cross(deta, dzeta, csi + 3 * at(&d, i, j, k));
      }
#pragma omp parallel for schedule(static)
  for (int64_t k = 1; k < d.mz; ++k)
    for (int64_t j = 0; j < d.my; ++j)
      for (int64_t i = 1; i < d.mx; ++i) {
        double dxi[3], dzeta[3];
        half_diff(N(i, j, k), N(i, j, k - 1), N(i - 1, j, k), N(i - 1, j, k - 1), dxi);
        half_diff(N(i, j, k), N(i - 1, j, k), N(i, j, k - 1), N(i - 1, j, k - 1), dzeta);
        cross(dzeta, dxi, eta + 3 * at(&d, i, j, k));
      }
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < d.mz; ++k)
    for (int64_t j = 1; j < d.my; ++j)
      for (int64_t i = 1; i < d.mx; ++i) {
        double dxi[3], deta[3];
        half_diff(N(i, j, k), N(i, j - 1, k), N(i - 1, j, k), N(i - 1, j - 1, k), dxi);
        half_diff(N(i, j, k), N(i - 1, j, k), N(i, j - 1, k), N(i - 1, j - 1, k), deta);
        cross(dxi, deta, zet + 3 * at(&d, i, j, k));
      }
#undef N
  copy_planes(&d, eta, 3, 0);
  copy_planes(&d, zet, 3, 0);
  copy_planes(&d, csi, 3, 1);
  copy_planes(&d, zet, 3, 1);
  copy_planes(&d, csi, 3, 2);
  copy_planes(&d, eta, 3, 2);
}

/* 0.25 * ((a+b+c+e) - (f+g+h+l)) of one component (Metric.c:533-550) */
static inline double quarter_diff(const double *a, const double *b, const double *c, const double *e,
                                  const double *f, const double *g, const double *h, const double *l, int q) {
  return 0.25 * ((a[q] + b[q] + c[q] + e[q]) - (f[q] + g[q] + h[q] + l[q]));
}

/* ComputeCellCenteredJacobianInverse, Metric.c:491-622: for nodes i,j,k >= 1,
 * the determinant of the three quarter-differences over the 8 nodes of the
 * box below-left of the node, Aj = 1/det; |det| < 1e-12 is an error
 * (Metric.c:555-557).  Then boundary planes copied x, y, z. */
int ref_jacobian_inverse(int64_t IM, int64_t JM, int64_t KM, const double *coor, double *aj) {
  const dims_t d = {IM + 1, JM + 1, KM + 1};
  int err = REF_OK;
#define N(i_, j_, k_) (coor + 3 * at(&d, (i_), (j_), (k_)))
#pragma omp parallel for schedule(static) reduction(min : err)
  for (int64_t k = 1; k < d.mz; ++k)
    for (int64_t j = 1; j < d.my; ++j)
      for (int64_t i = 1; i < d.mx; ++i) {
        const double *p000 = N(i - 1, j - 1, k - 1), *p100 = N(i, j - 1, k - 1), *p010 = N(i - 1, j, k - 1),
                     *p110 = N(i, j, k - 1), *p001 = N(i - 1, j - 1, k), *p101 = N(i, j - 1, k),
                     *p011 = N(i - 1, j, k), *p111 = N(i, j, k);
        double dxi[3], deta[3], dzeta[3];
        for (int q = 0; q < 3; ++q) {
          dxi[q] = quarter_diff(p111, p101, p110, p100, p011, p001, p010, p000, q);
          deta[q] = quarter_diff(p111, p011, p110, p010, p101, p001, p100, p000, q);
          dzeta[q] = quarter_diff(p111, p101, p011, p001, p110, p100, p010, p000, q);
        }
        const double det = dxi[0] * (deta[1] * dzeta[2] - deta[2] * dzeta[1]) -
                           dxi[1] * (deta[0] * dzeta[2] - deta[2] * dzeta[0]) +
                           dxi[2] * (deta[0] * dzeta[1] - deta[1] * dzeta[0]);
        if (fabs(det) < 1.0e-12) err = REF_E_SINGULAR;
        aj[at(&d, i, j, k)] = 1.0 / det;
      }
#undef N
  if (err != REF_OK) return err;
  copy_planes(&d, aj, 1, 0);
  copy_planes(&d, aj, 1, 1);
  copy_planes(&d, aj, 1, 2);
  return REF_OK;
}

/* CheckAndFixGridOrientation, Metric.c:195-284: mixed Aj signs are an error;
 * an all-negative Aj flips csi, eta, zet and Aj.  The orientation it stores
 * ends as +1 in every case that does not fail (after a flip, Aj > 0 resets it
 * to +1, Metric.c:262-263). */
int ref_grid_orientation(int64_t nnode, double *csi, double *eta, double *zet, double *aj, int *orientation) {
  double lo = aj[0], hi = aj[0];
  for (int64_t n = 1; n < nnode; ++n) {
    lo = aj[n] < lo ? aj[n] : lo;
    hi = aj[n] > hi ? aj[n] : hi;
  }
  if (lo < 0.0 && hi > 0.0) return REF_E_GRID;
  int o = 1;
  if (hi < 0.0) {
    o = -1;
    for (int64_t n = 0; n < 3 * nnode; ++n) {
      csi[n] = csi[n] * -1.0;
      eta[n] = eta[n] * -1.0;
      zet[n] = zet[n] * -1.0;
    }
    for (int64_t n = 0; n < nnode; ++n) aj[n] = aj[n] * -1.0;
    lo = aj[0];
    hi = aj[0];
    for (int64_t n = 1; n < nnode; ++n) {
      lo = aj[n] < lo ? aj[n] : lo;
      hi = aj[n] > hi ? aj[n] : hi;
    }
    if (lo <= 0.0) return REF_E_GRID;
    if (lo != 0.0 && hi > 0.0) o = 1;
  }
  *orientation = o;
  return REF_OK;
}

/* First face whose mathematical type is INLET, else -Xi (io.c:2664-2672,
 * setup.c:136). */
int ref_inlet_face(const ref_bc_t *bc) {
  for (int f = 0; f < 6; ++f)
    if (bc->type[f] == REF_BCTYPE_INLET) return f;
  return 0;
}

/* SetInitialInteriorField("Ucont"), AnalyticalSolution.c:986-1176: every
 * interior node (0 < i < mx-1, likewise j, k) gets a contravariant vector
 * that is zero except along the inlet face's direction, where it is the
 * signed normal speed times the face area |csi|, |eta| or |zet|.  finit 0:
 * zero; 1: the inlet direction's component of uin; 2: that component times
 * the parabolic profile in (i, j), clipped at 0; other values: zero. */
void ref_initial_ucont(int64_t IM, int64_t JM, int64_t KM, const double *csi, const double *eta,
                       const double *zet, int finit, const double uin[3], int inlet_face, int orientation,
                       double *ucont) {
  const dims_t d = {IM + 1, JM + 1, KM + 1};
  const double sign = (inlet_face % 2 == 0) ? 1.0 : -1.0;
  const int axis = inlet_face / 2;
  const double iw = (double)(d.mx - 2), jw = (double)(d.my - 2);
  const double ic = 1.0 + iw / 2.0, jc = 1.0 + jw / 2.0;
  for (int64_t k = 1; k < d.mz - 1; ++k)
    for (int64_t j = 1; j < d.my - 1; ++j)
      for (int64_t i = 1; i < d.mx - 1; ++i) {
        double mag = 0.0;
        if (finit == 1) {
          mag = uin[axis];
        } else if (finit == 2) {
          const double in = (i - ic) / (iw / 2.0), jn = (j - jc) / (jw / 2.0);
          const double pi_ = 1.0 - in * in, pj = 1.0 - jn * jn;
          mag = uin[axis] * pi_ * pj;
          if (mag < 0.0) mag = 0.0;
        }
        const int64_t n = at(&d, i, j, k);
        double *u = ucont + 3 * n;
        u[0] = u[1] = u[2] = 0.0;
        if (mag != 0.0) {
          const double v = mag * sign * orientation;
          const double *m = (axis == 0 ? csi : axis == 1 ? eta : zet) + 3 * n;
          u[axis] = v * sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
        }
      }
}

/* v . m with the reference's summation order (BC_Handlers.c:345-372) */
static inline double dot3(const double *v, const double *m) { return v[0] * m[0] + v[1] * m[1] + v[2] * m[2]; }

/* One face of one handler.  Faces: 0 -Xi, 1 +Xi, 2 -Eta, 3 +Eta, 4 -Zeta,
 * 5 +Zeta (common.h:203-205).  Node plane p = 0 or m-1 along the face axis;
 * q = the neighbouring plane inside (1 or m-2), u = the contravariant plane
 * of the face (0 or m-2). */
static void apply_face(const dims_t *d, const ref_bc_t *bc, int face, const double *csi, const double *eta,
                       const double *zet, double *ucat, double *ucont) {
  const int axis = face / 2, pos = face % 2;
  const int64_t m[3] = {d->mx, d->my, d->mz};
  const int64_t p = pos ? m[axis] - 1 : 0;
  const int64_t q = pos ? p - 1 : p + 1;
  const int64_t u = pos ? p - 1 : p;
  const double *metric = axis == 0 ? csi : axis == 1 ? eta : zet;
  const int h = bc->handler[face];
  /* only the face plane: the loop bound of the face axis is pinned to p */
  const int64_t klo = axis == 2 ? p : 0, khi = axis == 2 ? p + 1 : d->mz;
  const int64_t jlo = axis == 1 ? p : 0, jhi = axis == 1 ? p + 1 : d->my;
  const int64_t ilo = axis == 0 ? p : 0, ihi = axis == 0 ? p + 1 : d->mx;
  for (int64_t k = klo; k < khi; ++k)
    for (int64_t j = jlo; j < jhi; ++j)
      for (int64_t i = ilo; i < ihi; ++i) {
        int64_t idx[3] = {i, j, k};
        const int64_t np = at(d, i, j, k);
        idx[axis] = q;
        const int64_t nq = at(d, idx[0], idx[1], idx[2]);
        idx[axis] = u;
        const int64_t nu = at(d, idx[0], idx[1], idx[2]);
        switch (h) {
          case REF_BC_WALL_NOSLIP: /* BC_Handlers.c:58-165 */
            ucont[3 * nu + axis] = 0.0;
            for (int c = 0; c < 3; ++c) ucat[3 * np + c] = -ucat[3 * nq + c];
            break;
          case REF_BC_NOGRAD_COPY_GHOST: /* BC_Handlers.c:500-590 */
            for (int c = 0; c < 3; ++c) ucat[3 * np + c] = ucat[3 * nq + c];
            if (!pos || m[axis] >= 3) ucont[3 * np + axis] = ucont[3 * nq + axis];
            break;
          case REF_BC_INLET_CONSTANT_VELOCITY: /* BC_Handlers.c:313-400 */
            for (int c = 0; c < 3; ++c) ucat[3 * np + c] = bc->vel[face][c];
            ucont[3 * nu + axis] = dot3(bc->vel[face], metric + 3 * nu);
            break;
          case REF_BC_INLET_PARABOLIC: /* BC_Handlers.c:695-795, k faces only */
            if (axis == 2) {
              const double iw = (double)(d->mx - 2), jw = (double)(d->my - 2);
              const double ic = 1.0 + iw / 2.0, jc = 1.0 + jw / 2.0;
              const double in = (i - ic) / (iw / 2.0), jn = (j - jc) / (jw / 2.0);
              const double pi_ = 1.0 - in * in, pj = 1.0 - jn * jn;
              double mag = bc->vmax[face] * pi_ * pj;
              if (mag < 0.0) mag = 0.0;
              const double *a = zet + 3 * np;
              const double area = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
              double dir[3] = {0.0, 0.0, 0.0};
              if (area > 1e-9) {
                dir[0] = a[0] / area;
                dir[1] = a[1] / area;
                dir[2] = a[2] / area;
              }
              double v[3] = {mag * dir[0], mag * dir[1], mag * dir[2]};
              for (int c = 0; c < 3; ++c) ucat[3 * np + c] = v[c];
              ucont[3 * nu + 2] = dot3(v, zet + 3 * nu);
            }
            break;
          default:
            break;
        }
      }
}

static int check_handlers(const ref_bc_t *bc) {
  for (int f = 0; f < 6; ++f) {
    const int h = bc->handler[f];
    if (h != REF_BC_WALL_NOSLIP && h != REF_BC_NOGRAD_COPY_GHOST && h != REF_BC_INLET_CONSTANT_VELOCITY &&
        h != REF_BC_INLET_PARABOLIC)
      return REF_E_ARG; /* BoundaryCondition_Create, Boundaries.c:402-466 */
  }
  return REF_OK;
}

/* BoundarySystem_Create's handler initialisation, Boundaries.c:467-565: the
 * two inlet handlers apply themselves once (BC_Handlers.c:238-277, 647-673),
 * faces in order 0..5. */
int ref_bc_initialize(int64_t IM, int64_t JM, int64_t KM, const ref_bc_t *bc, const double *csi,
                      const double *eta, const double *zet, double *ucat, double *ucont) {
  const int e = check_handlers(bc);
  if (e != REF_OK) return e;
  const dims_t d = {IM + 1, JM + 1, KM + 1};
  for (int f = 0; f < 6; ++f)
    if (bc->handler[f] == REF_BC_INLET_CONSTANT_VELOCITY || bc->handler[f] == REF_BC_INLET_PARABOLIC)
      apply_face(&d, bc, f, csi, eta, zet, ucat, ucont);
  return REF_OK;
}

/* BoundarySystem_ExecuteStep, Boundaries.c:566-646: every face's handler
 * applies, faces in order 0..5 (the PreStep hooks of these handlers do
 * nothing). */
int ref_bc_execute(int64_t IM, int64_t JM, int64_t KM, const ref_bc_t *bc, const double *csi,
                   const double *eta, const double *zet, double *ucat, double *ucont) {
  const int e = check_handlers(bc);
  if (e != REF_OK) return e;
  const dims_t d = {IM + 1, JM + 1, KM + 1};
  for (int f = 0; f < 6; ++f) apply_face(&d, bc, f, csi, eta, zet, ucat, ucont);
  return REF_OK;
}

/* Contra2Cart, setup.c:1294-1420: for interior cells 1 <= i <= mx-2 (likewise
 * j, k) solve M u = q by Cramer's rule, M's rows the metric vectors averaged
 * over the cell's two faces, q the averaged contravariant components;
 * |det M| < 1e-18 is an error (:1366-1368). */
int ref_contra2cart(int64_t IM, int64_t JM, int64_t KM, const double *ucont, const double *csi, const double *eta,
                    const double *zet, double *ucat) {
  const dims_t d = {IM + 1, JM + 1, KM + 1};
  int err = REF_OK;
#pragma omp parallel for schedule(static) reduction(min : err)
  for (int64_t k = 1; k < d.mz - 1; ++k)
    for (int64_t j = 1; j < d.my - 1; ++j)
      for (int64_t i = 1; i < d.mx - 1; ++i) {
        const int64_t n = at(&d, i, j, k), ni = at(&d, i - 1, j, k), nj = at(&d, i, j - 1, k),
                      nk = at(&d, i, j, k - 1);
        double M[3][3];
        for (int c = 0; c < 3; ++c) {
          M[0][c] = 0.5 * (csi[3 * ni + c] + csi[3 * n + c]);
          M[1][c] = 0.5 * (eta[3 * nj + c] + eta[3 * n + c]);
          M[2][c] = 0.5 * (zet[3 * nk + c] + zet[3 * n + c]);
        }
        const double q0 = 0.5 * (ucont[3 * ni + 0] + ucont[3 * n + 0]);
        const double q1 = 0.5 * (ucont[3 * nj + 1] + ucont[3 * n + 1]);
        const double q2 = 0.5 * (ucont[3 * nk + 2] + ucont[3 * n + 2]);
        const double det = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) -
                           M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                           M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
        if (fabs(det) < 1.0e-18) {
          err = REF_E_SINGULAR;
          continue;
        }
        const double r = 1.0 / det;
        const double d0 = q0 * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) - q1 * (M[0][1] * M[2][2] - M[0][2] * M[2][1]) +
                          q2 * (M[0][1] * M[1][2] - M[0][2] * M[1][1]);
        const double d1 = -q0 * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                          q1 * (M[0][0] * M[2][2] - M[0][2] * M[2][0]) - q2 * (M[0][0] * M[1][2] - M[0][2] * M[1][0]);
        const double d2 = q0 * (M[1][0] * M[2][1] - M[1][1] * M[2][0]) - q1 * (M[0][0] * M[2][1] - M[0][1] * M[2][0]) +
                          q2 * (M[0][0] * M[1][1] - M[0][1] * M[1][0]);
        ucat[3 * n + 0] = d0 * r;
        ucat[3 * n + 1] = d1 * r;
        ucat[3 * n + 2] = d2 * r;
      }
  return err;
}

/* SetEulerianFields without the restart branch, simulation.c:62-92: at
 * step 0 the initial interior ucont, then boundaries, Contra2Cart,
 * boundaries again. */
int ref_set_eulerian_fields(int64_t IM, int64_t JM, int64_t KM, const ref_bc_t *bc, const double *csi,
                            const double *eta, const double *zet, int64_t step, int finit, const double uin[3],
                            int orientation, double *ucat, double *ucont) {
  if (step == 0) ref_initial_ucont(IM, JM, KM, csi, eta, zet, finit, uin, ref_inlet_face(bc), orientation, ucont);
  int e = ref_bc_execute(IM, JM, KM, bc, csi, eta, zet, ucat, ucont);
  if (e != REF_OK) return e;
  e = ref_contra2cart(IM, JM, KM, ucont, csi, eta, zet, ucat);
  if (e != REF_OK) return e;
  return ref_bc_execute(IM, JM, KM, bc, csi, eta, zet, ucat, ucont);
}
