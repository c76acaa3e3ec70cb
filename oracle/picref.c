/*
 * picref.c -- CPU oracle for the PICurv per-step particle loop.
 *
 * TEST INFRASTRUCTURE ONLY (see picref.h).  Clean-room restatement of the
 * reference's arithmetic; every function cites the reference file:line it
 * follows.  Parity build: gcc -O2 -ffp-contract=off (no FMA, IEEE order).
 * The same source built with -O3 -march=native -fopenmp is the CPU baseline
 * timed beside the GPU; OpenMP pragmas only parallelise loops whose
 * iterations are independent, so both builds give the same results.
 */
#include "picref.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define REF_DISTANCE_THRESHOLD 1e-11 /* walkingsearch.c:11 */
#define REF_WEIGHT_CLAMP 1.0e-14     /* ParticleSwarm.c:5 */
#define REF_BBOX_TOLERANCE 1e-6      /* grid.c:5 */

static inline const double *node_at(const ref_grid_t *g, int64_t i, int64_t j, int64_t k) {
  return g->coor + 3 * ((k * (g->JM + 1) + j) * (g->IM + 1) + i);
}

/* walkingsearch.c:521-549 -- vertex order V0..V7 */
void ref_cell_vertices(const ref_grid_t *g, int64_t i, int64_t j, int64_t k, double V[8][3]) {
  static const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                {0, 1, 1}, {1, 1, 1}, {1, 0, 1}, {0, 0, 1}};
  for (int v = 0; v < 8; ++v) {
    const double *q = node_at(g, i + off[v][0], j + off[v][1], k + off[v][2]);
    V[v][0] = q[0];
    V[v][1] = q[1];
    V[v][2] = q[2];
  }
}

/* walkingsearch.c:25-59 -- mean of the 12 edge lengths, edges in the listed order */
double ref_cell_char_size(const double V[8][3]) {
  static const int E[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                               {6, 7}, {7, 4}, {0, 7}, {1, 6}, {2, 5}, {3, 4}};
  double total = 0.0;
  for (int e = 0; e < 12; ++e) {
    const double *A = V[E[e][0]], *B = V[E[e][1]];
    double dx = B[0] - A[0];
    double dy = B[1] - A[1];
    double dz = B[2] - A[2];
    total += sqrt(dx * dx + dy * dy + dz * dz);
  }
  return total / 12.0;
}

/* walkingsearch.c:102-227 -- signed distance to the plane through the face
 * centroid with normal (v2-v1)x(v4-v1), oriented away from the cell centroid
 * (absolute 1e-9 test on the unnormalised dot, :181), zeroed inside thr. */
int ref_signed_distance(const double v1[3], const double v2[3], const double v3[3],
                        const double v4[3], const double cc[3], const double p[3], double thr,
                        double *d) {
  double e1x = v2[0] - v1[0], e1y = v2[1] - v1[1], e1z = v2[2] - v1[2];
  double e2x = v4[0] - v1[0], e2y = v4[1] - v1[1], e2z = v4[2] - v1[2];
  double nx = e1y * e2z - e1z * e2y;
  double ny = e1z * e2x - e1x * e2z;
  double nz = e1x * e2y - e1y * e2x;
  double m = sqrt(nx * nx + ny * ny + nz * nz);
  if (m < 1.0e-12) return REF_E_DEGENERATE; /* :147-155 */
  double fcx = 0.25 * (v1[0] + v2[0] + v3[0] + v4[0]);
  double fcy = 0.25 * (v1[1] + v2[1] + v3[1] + v4[1]);
  double fcz = 0.25 * (v1[2] + v2[2] + v3[2] + v4[2]);
  double tx = cc[0] - fcx, ty = cc[1] - fcy, tz = cc[2] - fcz;
  double dot = nx * tx + ny * ty + nz * tz;
  if (dot > 1.0e-9) { /* :181-185 */
    nx = -nx;
    ny = -ny;
    nz = -nz;
  }
  nx /= m;
  ny /= m;
  nz /= m;
  double wx = fcx - p[0], wy = fcy - p[1], wz = fcz - p[2];
  double dd = wx * nx + wy * ny + wz * nz;
  if (fabs(dd) < thr) dd = 0.0; /* :219-222 */
  *d = dd;
  return REF_OK;
}

/* walkingsearch.c:271-389 -- centroid, then BACK, FRONT, BOTTOM, TOP, LEFT, RIGHT */
int ref_face_distances(const double V[8][3], const double p[3], double thr, double d[6]) {
  double cc[3] = {0.0, 0.0, 0.0};
  for (int v = 0; v < 8; ++v) {
    cc[0] += V[v][0];
    cc[1] += V[v][1];
    cc[2] += V[v][2];
  }
  cc[0] /= 8.0;
  cc[1] /= 8.0;
  cc[2] /= 8.0;
  int e;
  if ((e = ref_signed_distance(V[0], V[3], V[2], V[1], cc, p, thr, &d[REF_BACK]))) return e;
  if ((e = ref_signed_distance(V[4], V[7], V[6], V[5], cc, p, thr, &d[REF_FRONT]))) return e;
  if ((e = ref_signed_distance(V[0], V[1], V[6], V[7], cc, p, thr, &d[REF_BOTTOM]))) return e;
  if ((e = ref_signed_distance(V[3], V[4], V[5], V[2], cc, p, thr, &d[REF_TOP]))) return e;
  if ((e = ref_signed_distance(V[0], V[7], V[4], V[3], cc, p, thr, &d[REF_LEFT]))) return e;
  if ((e = ref_signed_distance(V[1], V[2], V[5], V[6], cc, p, thr, &d[REF_RIGHT]))) return e;
  return REF_OK;
}

/* walkingsearch.c:407-462 */
int ref_point_position(const double d[6]) {
  int inside = 1, zeros = 0;
  for (int f = 0; f < 6; ++f) {
    if (d[f] < 0.0) inside = 0;
    if (d[f] == 0.0) zeros++;
  }
  if (inside) return 0;
  if (zeros == 1) return 1;
  if (zeros == 2) return 2;
  if (zeros >= 3) return 3;
  return -1;
}

/* walkingsearch.c:811-859 */
int ref_evaluate_position(const double V[8][3], const double p[3], double d[6], int *position) {
  double h = ref_cell_char_size(V);
  double thr = REF_DISTANCE_THRESHOLD * h;
  int e = ref_face_distances(V, p, thr, d);
  if (e) return e;
  *position = ref_point_position(d);
  return REF_OK;
}

/* ParticleSwarm.c:1010-1047 */
void ref_update_weights(double d[6], double w[3]) {
  for (int f = 0; f < 6; ++f)
    if (d[f] <= REF_WEIGHT_CLAMP) d[f] = REF_WEIGHT_CLAMP;
  w[0] = d[REF_LEFT] / (d[REF_LEFT] + d[REF_RIGHT]);
  w[1] = d[REF_BOTTOM] / (d[REF_BOTTOM] + d[REF_TOP]);
  w[2] = d[REF_BACK] / (d[REF_FRONT] + d[REF_BACK]);
}

/* walkingsearch.c:882-959 */
void ref_step_indices(const double d[6], int64_t *i, int64_t *j, int64_t *k) {
  if (d[REF_FRONT] < 0.0)
    *k += 1;
  else if (d[REF_BACK] < 0.0)
    *k -= 1;
  if (d[REF_LEFT] < 0.0)
    *i -= 1;
  else if (d[REF_RIGHT] < 0.0)
    *i += 1;
  if (d[REF_BOTTOM] < 0.0)
    *j -= 1;
  else if (d[REF_TOP] < 0.0)
    *j += 1;
}

/* walkingsearch.c:1094-1242 with the start rule of :571-629 on one rank.
 * On one rank the ghosted cell-origin range (:650-708) is the whole grid, so
 * the hand-off branch never triggers, and FindOwnerOfCell (:1026-1071)
 * returns this rank for every in-range cell.  The stuck detector (:1149-1185)
 * cannot fire: a rejected cell always moves at least one index. */
int ref_locate_particle(const ref_grid_t *g, const double p[3], int64_t cell[3], double w[3],
                        int max_traversal, const int64_t *start_guess, int *steps) {
  int64_t i, j, k;
  if (cell[0] >= 0 && cell[1] >= 0 && cell[2] >= 0 && cell[0] < g->IM && cell[1] < g->JM &&
      cell[2] < g->KM) {
    i = cell[0];
    j = cell[1];
    k = cell[2];
  } else if (start_guess) {
    i = start_guess[0];
    j = start_guess[1];
    k = start_guess[2];
  } else {
    i = 0; /* (info.xs, info.ys, info.zs) on one rank */
    j = 0;
    k = 0;
  }
  int n = 0, concluded = 0;
  while (!concluded && n < max_traversal) {
    n++;
    if (i < 0 || i >= g->IM || j < 0 || j >= g->JM || k < 0 || k >= g->KM) {
      i = -1;
      break;
    }
    double V[8][3], d[6];
    int pos;
    ref_cell_vertices(g, i, j, k, V);
    int e = ref_evaluate_position(V, p, d, &pos);
    if (e) return e;
    if (pos >= 0) {
      concluded = 1;
      ref_update_weights(d, w);
    } else {
      ref_step_indices(d, &i, &j, &k);
    }
  }
  if (steps) *steps = n;
  if (i == -1 || !concluded) {
    cell[0] = cell[1] = cell[2] = -1;
    return REF_LOST;
  }
  cell[0] = i;
  cell[1] = j;
  cell[2] = k;
  return REF_ACTIVE_AND_LOCATED;
}

/* ParticleMotion.c:1931-2135 on one rank: skip ACTIVE_AND_LOCATED, walk the rest. */
int ref_locate_all(const ref_grid_t *g, int64_t n, const double *pos, int64_t *cell, double *w,
                   int64_t *status, int max_traversal, const int64_t *guess, int32_t *steps) {
  int err = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(min : err)
  for (int64_t p = 0; p < n; ++p) {
    if (status[p] == REF_ACTIVE_AND_LOCATED) {
      if (steps) steps[p] = 0;
      continue;
    }
    int s = 0;
    int r = ref_locate_particle(g, pos + 3 * p, cell + 3 * p, w + 3 * p, max_traversal,
                                guess ? guess + 3 * p : NULL, &s);
    if (r < 0) {
      err = r < err ? r : err;
      continue;
    }
    status[p] = r;
    if (steps) steps[p] = s;
  }
  return err;
}

/* interpolation.c:598-680: node (i,j,k) averages the valid cells among
 * (i-1..i, j-1..j, k-1..k), summed dk-outer, dj, di-inner, divided by the count. */
void ref_center_to_corner(int64_t IM, int64_t JM, int64_t KM, const double *ucat, double *corner) {
  const int64_t nx = IM + 1, ny = JM + 1;
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k <= KM; ++k)
    for (int64_t j = 0; j <= JM; ++j)
      for (int64_t i = 0; i <= IM; ++i) {
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        int64_t cnt = 0;
        for (int64_t dk = -1; dk <= 0; ++dk)
          for (int64_t dj = -1; dj <= 0; ++dj)
            for (int64_t di = -1; di <= 0; ++di) {
              int64_t ci = i + di, cj = j + dj, ck = k + dk;
              if (ci >= 0 && ci < IM && cj >= 0 && cj < JM && ck >= 0 && ck < KM) {
                const double *u = ucat + 3 * ((ck * ny + cj) * nx + ci);
                s0 += u[0];
                s1 += u[1];
                s2 += u[2];
                cnt++;
              }
            }
        double *o = corner + 3 * ((k * ny + j) * nx + i);
        if (cnt > 0) {
          o[0] = s0 / (double)cnt;
          o[1] = s1 / (double)cnt;
          o[2] = s2 / (double)cnt;
        } else {
          o[0] = o[1] = o[2] = 0.0;
        }
      }
}

/* interpolation.c:839-864 (weights) and 957-1042 (accumulate, normalise) */
void ref_trilinear(int64_t IM, int64_t JM, const double *corner, int64_t i, int64_t j, int64_t k,
                   double a1, double a2, double a3, double out[3]) {
  a1 = (0.0 < (1.0 < a1 ? 1.0 : a1)) ? (1.0 < a1 ? 1.0 : a1) : 0.0;
  a2 = (0.0 < (1.0 < a2 ? 1.0 : a2)) ? (1.0 < a2 ? 1.0 : a2) : 0.0;
  a3 = (0.0 < (1.0 < a3 ? 1.0 : a3)) ? (1.0 < a3 ? 1.0 : a3) : 0.0;
  const double oa1 = 1.0 - a1, oa2 = 1.0 - a2, oa3 = 1.0 - a3;
  double wc[8];
  wc[0] = oa1 * oa2 * oa3;
  wc[1] = a1 * oa2 * oa3;
  wc[2] = oa1 * a2 * oa3;
  wc[3] = a1 * a2 * oa3;
  wc[4] = oa1 * oa2 * a3;
  wc[5] = a1 * oa2 * a3;
  wc[6] = oa1 * a2 * a3;
  wc[7] = a1 * a2 * a3;
  static const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0},
                                {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
  const int64_t nx = IM + 1, ny = JM + 1;
  double sw = 0.0, ax = 0.0, ay = 0.0, az = 0.0;
  for (int c = 0; c < 8; ++c) {
    const double *f = corner + 3 * (((k + off[c][2]) * ny + (j + off[c][1])) * nx + (i + off[c][0]));
    ax += wc[c] * f[0];
    ay += wc[c] * f[1];
    az += wc[c] * f[2];
    sw += wc[c];
  }
  if (sw > 1.0e-14) {
    out[0] = ax / sw;
    out[1] = ay / sw;
    out[2] = az / sw;
  } else {
    out[0] = out[1] = out[2] = 0.0;
  }
}

/* interpolation.c:1341-1385 on one rank (gxs=0, gxm=IM+1): cell outside
 * [0,IM)x[0,JM)x[0,KM) -> velocity zero */
void ref_interp_all(int64_t IM, int64_t JM, int64_t KM, const double *corner, int64_t n,
                    const int64_t *cell, const double *w, double *vel) {
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < n; ++p) {
    int64_t i = cell[3 * p], j = cell[3 * p + 1], k = cell[3 * p + 2];
    if (i < 0 || i >= IM || j < 0 || j >= JM || k < 0 || k >= KM) {
      vel[3 * p] = vel[3 * p + 1] = vel[3 * p + 2] = 0.0;
      continue;
    }
    ref_trilinear(IM, JM, corner, i, j, k, w[3 * p], w[3 * p + 1], w[3 * p + 2], vel + 3 * p);
  }
}

/* ParticleMotion.c:19-29 (x += v*dt per component) looped as in :38-95 */
void ref_advect_euler(int64_t n, double *pos, const double *vel, double dt) {
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < 3 * n; ++p) pos[p] += vel[p] * dt;
}

/* ParticleMotion.c:2149-2176 */
void ref_reset_statuses(int64_t n, int64_t *status) {
  for (int64_t p = 0; p < n; ++p)
    if (status[p] == REF_ACTIVE_AND_LOCATED) status[p] = REF_NEEDS_LOCATION;
}

/* grid.c:762-883 on one rank: min/max over all nodes, padded by 1e-6 */
void ref_bbox(const ref_grid_t *g, double bmin[3], double bmax[3]) {
  for (int a = 0; a < 3; ++a) {
    bmin[a] = 1.7976931348623157e308;
    bmax[a] = -1.7976931348623157e308;
  }
  int64_t nn = (g->IM + 1) * (g->JM + 1) * (g->KM + 1);
  for (int64_t q = 0; q < nn; ++q)
    for (int a = 0; a < 3; ++a) {
      double x = g->coor[3 * q + a];
      if (x < bmin[a]) bmin[a] = x;
      if (x > bmax[a]) bmax[a] = x;
    }
  for (int a = 0; a < 3; ++a) {
    bmin[a] = bmin[a] - REF_BBOX_TOLERANCE;
    bmax[a] = bmax[a] + REF_BBOX_TOLERANCE;
  }
}

/* ParticleMotion.c:532-625 (LOST) then :283-383 with IsParticleInBox :106-110 */
int64_t ref_remove(int64_t n, double *pos, double *vel, int64_t *cell, double *w, int64_t *status,
                   int64_t *pid, double *P, const double bmin[3], const double bmax[3],
                   int64_t *removed_lost, int64_t *removed_oob) {
  int64_t out = 0, rl = 0, ro = 0;
  for (int64_t p = 0; p < n; ++p) {
    if (status[p] == REF_LOST) {
      rl++;
      continue;
    }
    const double *x = pos + 3 * p;
    int inbox = x[0] >= bmin[0] && x[0] <= bmax[0] && x[1] >= bmin[1] && x[1] <= bmax[1] &&
                x[2] >= bmin[2] && x[2] <= bmax[2];
    if (!inbox) {
      ro++;
      continue;
    }
    if (out != p) {
      memmove(pos + 3 * out, pos + 3 * p, 3 * sizeof(double));
      if (vel) memmove(vel + 3 * out, vel + 3 * p, 3 * sizeof(double));
      if (cell) memmove(cell + 3 * out, cell + 3 * p, 3 * sizeof(int64_t));
      if (w) memmove(w + 3 * out, w + 3 * p, 3 * sizeof(double));
      status[out] = status[p];
      if (pid) pid[out] = pid[p];
      if (P) P[out] = P[p];
    }
    out++;
  }
  if (removed_lost) *removed_lost = rl;
  if (removed_oob) *removed_oob = ro;
  return out;
}

/* ParticleMotion.c:975-1008: count[k][j][i] += 1 when (i,j,k) is in the owned
 * node range (0..IM on one rank); LOST particles carry cell -1 and are skipped. */
void ref_count_cells(int64_t IM, int64_t JM, int64_t KM, int64_t n, const int64_t *cell, double *count) {
  const int64_t nx = IM + 1, ny = JM + 1, nz = KM + 1;
  memset(count, 0, sizeof(double) * nx * ny * nz);
  for (int64_t p = 0; p < n; ++p) {
    int64_t i = cell[3 * p], j = cell[3 * p + 1], k = cell[3 * p + 2];
    if (i >= 0 && i < nx && j >= 0 && j < ny && k >= 0 && k < nz) count[(k * ny + j) * nx + i] += 1.0;
  }
}

/* interpolation.c:1879-1909 (sum into node slots, ghosted range 0..IM) and
 * 1989-2018 (avg = sum/count when count > 0.5, else 0) */
void ref_scatter_P(int64_t IM, int64_t JM, int64_t KM, int64_t n, const int64_t *cell,
                   const double *P, const int64_t *order, const double *count, double *avg) {
  const int64_t nx = IM + 1, ny = JM + 1, nz = KM + 1, nn = nx * ny * nz;
  double *sum = (double *)calloc((size_t)nn, sizeof(double));
  for (int64_t q = 0; q < n; ++q) {
    int64_t p = order ? order[q] : q;
    int64_t i = cell[3 * p], j = cell[3 * p + 1], k = cell[3 * p + 2];
    if (i >= 0 && i < nx && j >= 0 && j < ny && k >= 0 && k < nz) sum[(k * ny + j) * nx + i] += P[p];
  }
  for (int64_t q = 0; q < nn; ++q) avg[q] = count[q] > 0.5 ? sum[q] / count[q] : 0.0;
  free(sum);
}

/* ---- SPEC: face deposition ---- */
/* The particles of every cell in ascending PID order: a counting sort by cell
 * (stable in particle order), then an insertion sort of each cell's bucket by
 * PID.  out[start[c] .. start[c+1]) lists cell c's particles. */
static int64_t *cells_in_pid_order(int64_t IM, int64_t JM, int64_t KM, int64_t n, const int64_t *cell,
                                   const int64_t *pid, int64_t **start_out) {
  const int64_t ncell = IM * JM * KM;
  int64_t *start = (int64_t *)calloc((size_t)(ncell + 2), sizeof(int64_t));
  int64_t *list = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
  for (int64_t p = 0; p < n; ++p) {
    int64_t i = cell[3 * p], j = cell[3 * p + 1], k = cell[3 * p + 2];
    if (i < 0 || i >= IM || j < 0 || j >= JM || k < 0 || k >= KM) continue;
    start[(k * JM + j) * IM + i + 2]++;
  }
  for (int64_t c = 0; c < ncell; ++c) start[c + 2] += start[c + 1];
  for (int64_t p = 0; p < n; ++p) {
    int64_t i = cell[3 * p], j = cell[3 * p + 1], k = cell[3 * p + 2];
    if (i < 0 || i >= IM || j < 0 || j >= JM || k < 0 || k >= KM) continue;
    list[start[(k * JM + j) * IM + i + 1]++] = p;
  }
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < ncell; ++c)
    for (int64_t t = start[c] + 1; t < start[c + 1]; ++t) {
      const int64_t x = list[t];
      int64_t u = t - 1;
      while (u >= start[c] && pid[list[u]] > pid[x]) {
        list[u + 1] = list[u];
        --u;
      }
      list[u + 1] = x;
    }
  *start_out = start;
  return list;
}

void ref_deposit_faces(int64_t IM, int64_t JM, int64_t KM, int64_t n, const int64_t *cell,
                       const double *w, const double *vel, const int64_t *pid, double *iface,
                       double *jface, double *kface) {
  int64_t *start;
  int64_t *list = cells_in_pid_order(IM, JM, KM, n, cell, pid, &start);
  /* face f between lower cell cl and upper cell cu: the sum over cl's
   * particles (weight a) and the sum over cu's (weight 1-a), each from zero
   * in PID order, then lower + upper; a missing cell adds zero */
  for (int dir = 0; dir < 3; ++dir) {
    const int64_t fI = IM + (dir == 0), fJ = JM + (dir == 1), fK = KM + (dir == 2);
    double *out = dir == 0 ? iface : (dir == 1 ? jface : kface);
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < fK; ++k)
      for (int64_t j = 0; j < fJ; ++j)
        for (int64_t i = 0; i < fI; ++i) {
          double acc[2][4] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
          for (int side = 0; side < 2; ++side) {
            int64_t ci = i, cj = j, ck = k;
            if (side == 0) { /* lower cell */
              if (dir == 0) ci -= 1;
              if (dir == 1) cj -= 1;
              if (dir == 2) ck -= 1;
            }
            if (ci < 0 || ci >= IM || cj < 0 || cj >= JM || ck < 0 || ck >= KM) continue;
            const int64_t c = (ck * JM + cj) * IM + ci;
            double *s = acc[side];
            for (int64_t q = start[c]; q < start[c + 1]; ++q) {
              const int64_t p = list[q];
              const double a = w[3 * p + dir];
              const double wt = side == 0 ? a : 1.0 - a;
              s[0] += wt;
              s[1] += wt * vel[3 * p];
              s[2] += wt * vel[3 * p + 1];
              s[3] += wt * vel[3 * p + 2];
            }
          }
          double *o = out + 4 * ((k * fJ + j) * fI + i);
          for (int v = 0; v < 4; ++v) o[v] = acc[0][v] + acc[1][v];
        }
  }
  free(start);
  free(list);
}

/* SPEC: hexahedron volume as six tetrahedra around the diagonal V0-V5 (vertex
 * order walkingsearch.c:537-544); the tetrahedra (V0, Va, Vb, V5) for the
 * ring edges (a, b) in the order below, |det| summed, divided by 6. */
double ref_cell_volume(const double V[8][3]) {
  static const int T[6][2] = {{1, 2}, {1, 6}, {3, 2}, {3, 4}, {7, 6}, {7, 4}};
  double s = 0.0;
  for (int t = 0; t < 6; ++t) {
    const double *a = V[0], *b = V[T[t][0]], *c = V[T[t][1]], *d = V[5];
    const double x0 = b[0] - a[0], x1 = b[1] - a[1], x2 = b[2] - a[2];
    const double y0 = c[0] - a[0], y1 = c[1] - a[1], y2 = c[2] - a[2];
    const double z0 = d[0] - a[0], z1 = d[1] - a[1], z2 = d[2] - a[2];
    const double det = x0 * (y1 * z2 - y2 * z1) + x1 * (y2 * z0 - y0 * z2) + x2 * (y0 * z1 - y1 * z0);
    s += fabs(det);
  }
  return s / 6.0;
}

/* SPEC: face volume fraction = count * vp / V_dual, V_dual = half the summed
 * volumes of the face's cells (lower cell first; a boundary face has one). */
void ref_face_volfrac(int64_t IM, int64_t JM, int64_t KM, const double *coor, int dir, const double *face,
                      double vp, double *out) {
  const ref_grid_t g = {IM, JM, KM, coor};
  const int64_t fI = IM + (dir == 0), fJ = JM + (dir == 1), fK = KM + (dir == 2);
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < fK; ++k)
    for (int64_t j = 0; j < fJ; ++j)
      for (int64_t i = 0; i < fI; ++i) {
        double V[8][3], s = 0.0;
        const int64_t li = i - (dir == 0), lj = j - (dir == 1), lk = k - (dir == 2);
        if (li >= 0 && lj >= 0 && lk >= 0) {
          ref_cell_vertices(&g, li, lj, lk, V);
          s += ref_cell_volume(V);
        }
        if (i < IM && j < JM && k < KM) {
          ref_cell_vertices(&g, i, j, k, V);
          s += ref_cell_volume(V);
        }
        const int64_t f = (k * fJ + j) * fI + i;
        out[f] = (face[4 * f] * vp) / (0.5 * s);
      }
}

/* ---- step drivers ---- */
static int64_t finish_step(const ref_ctx_t *c, int64_t n, double *pos, double *vel, int64_t *cell,
                           double *w, int64_t *status, int64_t *pid, double *P, int64_t *removed) {
  ref_grid_t g = {c->IM, c->JM, c->KM, c->coor};
  double bmin[3], bmax[3];
  if (c->bbox) {
    for (int a = 0; a < 3; ++a) {
      bmin[a] = c->bbox[a];
      bmax[a] = c->bbox[3 + a];
    }
  } else {
    ref_bbox(&g, bmin, bmax);
  }
  int64_t rl = 0, ro = 0;
  n = ref_remove(n, pos, vel, cell, w, status, pid, P, bmin, bmax, &rl, &ro);
  if (removed) *removed = rl + ro;
  ref_interp_all(c->IM, c->JM, c->KM, c->corner, n, cell, w, vel);
  return n;
}

/* simulation.c:248-292: reset, advect (Euler), locate, remove, interpolate.
 * Deposition is left to the caller (ref_count_cells / ref_scatter_P / faces). */
int64_t ref_step_euler(const ref_ctx_t *c, int64_t n, double *pos, double *vel, int64_t *cell,
                       double *w, int64_t *status, int64_t *pid, double *P, double dt,
                       int64_t *removed) {
  ref_grid_t g = {c->IM, c->JM, c->KM, c->coor};
  ref_reset_statuses(n, status);
  ref_advect_euler(n, pos, vel, dt);
  int e = ref_locate_all(&g, n, pos, cell, w, status, c->max_traversal, NULL, NULL);
  if (e) return e;
  return finish_step(c, n, pos, vel, cell, w, status, pid, P, removed);
}

/* SPEC: Heun predictor-corrector in the same stage frame. */
int64_t ref_step_rk2(const ref_ctx_t *c, int64_t n, double *pos, double *vel, int64_t *cell,
                     double *w, int64_t *status, int64_t *pid, double *P, double dt,
                     int64_t *removed) {
  ref_grid_t g = {c->IM, c->JM, c->KM, c->coor};
  const double h = 0.5 * dt;
  int err = 0;
  ref_reset_statuses(n, status);
#pragma omp parallel for schedule(dynamic, 256) reduction(min : err)
  for (int64_t p = 0; p < n; ++p) {
    if (status[p] == REF_LOST) continue;
    double x0[3] = {pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]};
    double u0[3] = {vel[3 * p], vel[3 * p + 1], vel[3 * p + 2]};
    double xs[3] = {x0[0] + u0[0] * dt, x0[1] + u0[1] * dt, x0[2] + u0[2] * dt};
    int64_t *cl = cell + 3 * p;
    double *wp = w + 3 * p;
    int r = ref_locate_particle(&g, xs, cl, wp, c->max_traversal, NULL, NULL);
    if (r < 0) {
      err = r < err ? r : err;
      continue;
    }
    if (r == REF_LOST) {
      pos[3 * p] = xs[0];
      pos[3 * p + 1] = xs[1];
      pos[3 * p + 2] = xs[2];
      status[p] = REF_LOST;
      continue;
    }
    double us[3];
    ref_trilinear(c->IM, c->JM, c->corner, cl[0], cl[1], cl[2], wp[0], wp[1], wp[2], us);
    double x1[3] = {x0[0] + (u0[0] + us[0]) * h, x0[1] + (u0[1] + us[1]) * h,
                    x0[2] + (u0[2] + us[2]) * h};
    r = ref_locate_particle(&g, x1, cl, wp, c->max_traversal, NULL, NULL);
    if (r < 0) {
      err = r < err ? r : err;
      continue;
    }
    pos[3 * p] = x1[0];
    pos[3 * p + 1] = x1[1];
    pos[3 * p + 2] = x1[2];
    status[p] = r;
  }
  if (err) return err;
  return finish_step(c, n, pos, vel, cell, w, status, pid, P, removed);
}

/* ---- SPEC: MLP ---- */
static double bf16_round(double v) {
  /* round a binary32 value to bfloat16, nearest-even, returned as double */
  float f = (float)v;
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (double)f; /* inf/nan: unchanged */
  uint32_t lsb = (u >> 16) & 1u;
  u += 0x7fffu + lsb;
  u &= 0xffff0000u;
  memcpy(&f, &u, 4);
  return (double)f;
}
/* GELU z Phi(z) in its hard form: Phi replaced by its tangent at 0 clipped
   to [0, 1], the gate t = clamp(c z + 1/2, 0, 1) with c = the binary32 value
   of 1/sqrt(2 pi) = Phi'(0) (the hard-sigmoid construction of Howard et al.
   2019, "Searching for MobileNetV3", sec. 5.2).
   Derivative: t + c z where 0 < c z + 1/2 < 1, else t (0 or 1). */
static const double act_c = (double)0.3989422804014327f;
static double act_gate(double z) {
  const double u = z * act_c + 0.5;
  return u < 0.0 ? 0.0 : (u > 1.0 ? 1.0 : u);
}
static double act(double z) { return z * act_gate(z); }
static double act_d(double z) {
  const double u = z * act_c + 0.5;
  return (u > 0.0 && u < 1.0) ? u + act_c * z : act_gate(z);
}

int64_t ref_mlp_param_count(int H) { return (int64_t)H * 3 + H + H * H + H + 3 * H + 3; }

typedef struct {
  const double *W1, *b1, *W2, *b2, *W3, *b3;
} mlp_view_t;
static mlp_view_t mlp_view(int H, const double *pr) {
  mlp_view_t v;
  v.W1 = pr;
  v.b1 = v.W1 + 3 * H;
  v.W2 = v.b1 + H;
  v.b2 = v.W2 + H * H;
  v.W3 = v.b2 + H;
  v.b3 = v.W3 + 3 * H;
  return v;
}

static void mlp_fwd_one(int H, const double *Wr, mlp_view_t v, const double x[3], double *z1,
                        double *a1r, double *z2, double *a2r, double y[3], int rb) {
  (void)Wr;
  /* each sum starts from its bias; layer 1 adds x2, x1, x0 in that order */
  for (int o = 0; o < H; ++o) {
    double s = v.b1[o];
    for (int q = 2; q >= 0; --q) s += v.W1[3 * o + q] * x[q];
    z1[o] = s;
    double a = act(z1[o]);
    a1r[o] = rb ? bf16_round(a) : a;
  }
  for (int o = 0; o < H; ++o) {
    double s = v.b2[o];
    for (int q = 0; q < H; ++q) s += v.W2[H * o + q] * a1r[q];
    z2[o] = s;
    double a = act(z2[o]);
    a2r[o] = rb ? bf16_round(a) : a;
  }
  for (int o = 0; o < 3; ++o) {
    double s = v.b3[o];
    for (int q = 0; q < H; ++q) s += v.W3[H * o + q] * a2r[q];
    y[o] = s;
  }
}

static double *mlp_round_weights(int H, const double *params, int rb) {
  int64_t np = ref_mlp_param_count(H);
  double *r = (double *)malloc(sizeof(double) * (size_t)np);
  memcpy(r, params, sizeof(double) * (size_t)np);
  if (rb) {
    mlp_view_t v = mlp_view(H, r);
    double *W1 = (double *)v.W1, *W2 = (double *)v.W2, *W3 = (double *)v.W3;
    for (int q = 0; q < 3 * H; ++q) W1[q] = bf16_round(W1[q]);
    for (int q = 0; q < H * H; ++q) W2[q] = bf16_round(W2[q]);
    for (int q = 0; q < 3 * H; ++q) W3[q] = bf16_round(W3[q]);
  }
  return r;
}

void ref_mlp_forward(int H, const double *params, int64_t n, const double *x, double *y, int rb) {
  double *r = mlp_round_weights(H, params, rb);
  mlp_view_t v = mlp_view(H, r);
#pragma omp parallel
  {
    double *buf = (double *)malloc(sizeof(double) * 4 * (size_t)H);
#pragma omp for schedule(static)
    for (int64_t p = 0; p < n; ++p)
      mlp_fwd_one(H, r, v, x + 3 * p, buf, buf + H, buf + 2 * H, buf + 3 * H, y + 3 * p, rb);
    free(buf);
  }
  free(r);
}

double ref_mlp_loss_grad(int H, const double *params, int64_t n, const double *x, const double *t,
                         double *grad, int rb) {
  double *r = mlp_round_weights(H, params, rb);
  mlp_view_t v = mlp_view(H, r);
  int64_t np = ref_mlp_param_count(H);
  memset(grad, 0, sizeof(double) * (size_t)np);
  double *gW1 = grad, *gb1 = gW1 + 3 * H, *gW2 = gb1 + H, *gb2 = gW2 + H * H, *gW3 = gb2 + H,
         *gb3 = gW3 + 3 * H;
  double *z1 = (double *)malloc(sizeof(double) * 8 * (size_t)H);
  double *a1r = z1 + H, *z2 = a1r + H, *a2r = z2 + H, *g2 = a2r + H, *g2r = g2 + H, *g1 = g2r + H;
  double loss = 0.0;
  const double scale = 2.0 / (3.0 * (double)n);
  for (int64_t p = 0; p < n; ++p) {
    double y[3];
    mlp_fwd_one(H, r, v, x + 3 * p, z1, a1r, z2, a2r, y, rb);
    double g3[3];
    for (int o = 0; o < 3; ++o) {
      double e = y[o] - t[3 * p + o];
      loss += e * e;
      g3[o] = e * scale;
    }
    for (int o = 0; o < 3; ++o) {
      for (int q = 0; q < H; ++q) gW3[H * o + q] += g3[o] * a2r[q];
      gb3[o] += g3[o];
    }
    for (int q = 0; q < H; ++q) {
      double s = 0.0;
      for (int o = 0; o < 3; ++o) s += g3[o] * v.W3[H * o + q];
      g2[q] = s * act_d(z2[q]);
      g2r[q] = rb ? bf16_round(g2[q]) : g2[q];
    }
    for (int o = 0; o < H; ++o) {
      for (int q = 0; q < H; ++q) gW2[H * o + q] += g2r[o] * a1r[q];
      gb2[o] += g2[o];
    }
    for (int q = 0; q < H; ++q) {
      double s = 0.0;
      for (int o = 0; o < H; ++o) s += g2r[o] * v.W2[H * o + q];
      g1[q] = s * act_d(z1[q]);
    }
    for (int o = 0; o < H; ++o) {
      for (int q = 0; q < 3; ++q) gW1[3 * o + q] += g1[o] * x[3 * p + q];
      gb1[o] += g1[o];
    }
  }
  free(z1);
  free(r);
  return loss / (3.0 * (double)n);
}
