// walk.hpp -- device side of the walking search (gfx950).
//
// Two cell evaluators share one walk driver:
//   * eval_cell_exact   -- the reference arithmetic in FP64, operation for
//     operation (walkingsearch.c:25-59, 102-227, 271-389, 407-462).  The
//     library is compiled with -ffp-contract=off, so every a*b+c below is a
//     separately rounded multiply and add, as in the reference (built -O0,
//     no FMA, Makefile:15-16); sqrt and '/' are the correctly rounded IEEE
//     operations of the HIP runtime.
//   * eval_cell_fp32    -- the same quantities in FP32 on coordinates taken
//     relative to vertex 0, with an a-priori error bound per face.  Faces
//     whose sign, zero-band membership, orientation flip or degeneracy the
//     bound does not certify are recomputed by one out-of-line FP64 routine
//     (reference arithmetic, vertices re-read from memory), so the
//     classification -- hence the walk path and the owning cell -- is the
//     reference's, bit for bit.  Distances (and the weights made from them)
//     are FP32.
// Face vertex indices are compile-time constants (template parameters), so
// the per-lane vertex array is never indexed at run time.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace picgpu {

enum : int { F_LEFT = 0, F_RIGHT = 1, F_BOTTOM = 2, F_TOP = 3, F_FRONT = 4, F_BACK = 5 };
enum : int { ST_NEEDS = 0, ST_LOCATED = 1, ST_MIGRATING = 2, ST_LOST = 3, ST_UNINIT = 4 };

// Faces in the reference's evaluation order BACK, FRONT, BOTTOM, TOP, LEFT,
// RIGHT (walkingsearch.c:303-377): vertices (v1, v2, v3, v4) and the distance
// slot each fills.
template <int F>
struct FaceDef;
template <> struct FaceDef<0> { static constexpr int a = 0, b = 3, c = 2, d = 1, slot = F_BACK; };
template <> struct FaceDef<1> { static constexpr int a = 4, b = 7, c = 6, d = 5, slot = F_FRONT; };
template <> struct FaceDef<2> { static constexpr int a = 0, b = 1, c = 6, d = 7, slot = F_BOTTOM; };
template <> struct FaceDef<3> { static constexpr int a = 3, b = 4, c = 5, d = 2, slot = F_TOP; };
template <> struct FaceDef<4> { static constexpr int a = 0, b = 7, c = 4, d = 3, slot = F_LEFT; };
template <> struct FaceDef<5> { static constexpr int a = 1, b = 2, c = 5, d = 6, slot = F_RIGHT; };

// the same table for the out-of-line FP64 path (indexes memory, not registers)
__device__ __constant__ static const int8_t kFaceVerts[6][4] = {
    {0, 3, 2, 1}, {4, 7, 6, 5}, {0, 1, 6, 7}, {3, 4, 5, 2}, {0, 7, 4, 3}, {1, 2, 5, 6}};
__device__ __constant__ static const int8_t kFaceSlot[6] = {F_BACK, F_FRONT, F_BOTTOM, F_TOP, F_LEFT, F_RIGHT};

struct GridView {
  const double *coor;  // [(KM+1)(JM+1)(IM+1)][3]
  int IM, JM, KM;
  double rI = 0.0, rIJ = 0.0;  // 1/IM and 1/(IM*JM) in FP64 (cell decoding, kernels.hpp)
};

// node index of vertex v of cell (i,j,k) (vertex order walkingsearch.c:537-544)
__device__ __forceinline__ int64_t vertex_node(const GridView &g, int64_t base, int v) {
  const int64_t sy = g.IM + 1, sz = (int64_t)(g.JM + 1) * (g.IM + 1);
  const int dx = (v == 1 || v == 2 || v == 5 || v == 6) ? 1 : 0;
  const int dy = (v == 2 || v == 3 || v == 4 || v == 5) ? 1 : 0;
  const int dz = (v >= 4) ? 1 : 0;
  return base + dx + dy * sy + dz * sz;
}
__device__ __forceinline__ int64_t cell_base_node(const GridView &g, int i, int j, int k) {
  return (int64_t)k * (g.JM + 1) * (g.IM + 1) + (int64_t)j * (g.IM + 1) + i;
}

__device__ __forceinline__ void load_vertices(const GridView &g, int i, int j, int k, double V[8][3]) {
  const int64_t b = cell_base_node(g, i, j, k);
#pragma unroll
  for (int v = 0; v < 8; ++v) {
    const double *q = g.coor + 3 * vertex_node(g, b, v);
    V[v][0] = q[0];
    V[v][1] = q[1];
    V[v][2] = q[2];
  }
}

// walkingsearch.c:25-59
__device__ __forceinline__ double char_size_exact(const double V[8][3]) {
  constexpr int E[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                            {6, 7}, {7, 4}, {0, 7}, {1, 6}, {2, 5}, {3, 4}};
  double total = 0.0;
#pragma unroll
  for (int e = 0; e < 12; ++e) {
    double dx = V[E[e][1]][0] - V[E[e][0]][0];
    double dy = V[E[e][1]][1] - V[E[e][0]][1];
    double dz = V[E[e][1]][2] - V[E[e][0]][2];
    total += sqrt(dx * dx + dy * dy + dz * dz);
  }
  return total / 12.0;
}

// walkingsearch.c:287-295
__device__ __forceinline__ void centroid_exact(const double V[8][3], double cc[3]) {
  cc[0] = 0.0;
  cc[1] = 0.0;
  cc[2] = 0.0;
#pragma unroll
  for (int v = 0; v < 8; ++v) {
    cc[0] += V[v][0];
    cc[1] += V[v][1];
    cc[2] += V[v][2];
  }
  cc[0] /= 8.0;
  cc[1] /= 8.0;
  cc[2] /= 8.0;
}

// walkingsearch.c:102-227, the part that depends on the cell only: the unit
// normal n (oriented away from the centroid) and the face centre fc.  Returns
// false on a degenerate face.
__device__ __forceinline__ bool face_plane_exact(const double *v1, const double *v2, const double *v3,
                                                 const double *v4, const double cc[3], double n[3], double fc[3]) {
  double e1x = v2[0] - v1[0], e1y = v2[1] - v1[1], e1z = v2[2] - v1[2];
  double e2x = v4[0] - v1[0], e2y = v4[1] - v1[1], e2z = v4[2] - v1[2];
  double nx = e1y * e2z - e1z * e2y;
  double ny = e1z * e2x - e1x * e2z;
  double nz = e1x * e2y - e1y * e2x;
  double m = sqrt(nx * nx + ny * ny + nz * nz);
  if (m < 1.0e-12) return false;
  double fcx = 0.25 * (v1[0] + v2[0] + v3[0] + v4[0]);
  double fcy = 0.25 * (v1[1] + v2[1] + v3[1] + v4[1]);
  double fcz = 0.25 * (v1[2] + v2[2] + v3[2] + v4[2]);
  double tx = cc[0] - fcx, ty = cc[1] - fcy, tz = cc[2] - fcz;
  double dot = nx * tx + ny * ty + nz * tz;
  if (dot > 1.0e-9) {
    nx = -nx;
    ny = -ny;
    nz = -nz;
  }
  n[0] = nx / m;
  n[1] = ny / m;
  n[2] = nz / m;
  fc[0] = fcx;
  fc[1] = fcy;
  fc[2] = fcz;
  return true;
}

// walkingsearch.c:102-227, the part that depends on the point.
__device__ __forceinline__ double face_point_distance(const double n[3], const double fc[3], const double p[3],
                                                      double thr) {
  double wx = fc[0] - p[0], wy = fc[1] - p[1], wz = fc[2] - p[2];
  double dd = wx * n[0] + wy * n[1] + wz * n[2];
  if (fabs(dd) < thr) dd = 0.0;
  return dd;
}

__device__ __forceinline__ bool face_distance_exact(const double *v1, const double *v2, const double *v3,
                                                    const double *v4, const double cc[3], const double p[3],
                                                    double thr, double &d) {
  double n[3], fc[3];
  if (!face_plane_exact(v1, v2, v3, v4, cc, n, fc)) return false;
  d = face_point_distance(n, fc, p, thr);
  return true;
}

template <int F>
__device__ __forceinline__ bool face_exact(const double V[8][3], const double cc[3], const double p[3], double thr,
                                           double d[6]) {
  using D = FaceDef<F>;
  return face_distance_exact(V[D::a], V[D::b], V[D::c], V[D::d], cc, p, thr, d[D::slot]);
}

// walkingsearch.c:811-859 with :271-389.  Returns false on a degenerate face
// (the reference raises at the first degenerate face in evaluation order).
__device__ __forceinline__ bool eval_cell_exact(const double V[8][3], const double p[3], double thresh_scale,
                                                double d[6]) {
  const double thr = thresh_scale * char_size_exact(V);
  double cc[3];
  centroid_exact(V, cc);
  return face_exact<0>(V, cc, p, thr, d) && face_exact<1>(V, cc, p, thr, d) && face_exact<2>(V, cc, p, thr, d) &&
         face_exact<3>(V, cc, p, thr, d) && face_exact<4>(V, cc, p, thr, d) && face_exact<5>(V, cc, p, thr, d);
}

// The walk's acceptance.  The reference classifies a point as inside (0: no
// distance below zero), outside (-1: some below zero, none exactly zero) or
// on 1..3 faces (the number of zero distances, capped at 3)
// (walkingsearch.c:407-462) and accepts any classification >= 0
// (:1150-1160): inside, or some distance exactly zero.  That is a minimum and
// six equality tests; the count itself is never used.
template <typename R>
__device__ __forceinline__ bool point_accepted(const R d[6]) {
  const R mn = fmin(fmin(fmin(d[0], d[1]), fmin(d[2], d[3])), fmin(d[4], d[5]));
  const bool on_face = d[0] == (R)0 || d[1] == (R)0 || d[2] == (R)0 || d[3] == (R)0 || d[4] == (R)0 ||
                       d[5] == (R)0;
  return !(mn < (R)0) || on_face;
}

// walkingsearch.c:919-946
template <typename R>
__device__ __forceinline__ void step_indices(const R d[6], int &i, int &j, int &k) {
  if (d[F_FRONT] < (R)0)
    k += 1;
  else if (d[F_BACK] < (R)0)
    k -= 1;
  if (d[F_LEFT] < (R)0)
    i -= 1;
  else if (d[F_RIGHT] < (R)0)
    i += 1;
  if (d[F_BOTTOM] < (R)0)
    j -= 1;
  else if (d[F_TOP] < (R)0)
    j += 1;
}

// ParticleSwarm.c:1010-1047
template <typename R>
__device__ __forceinline__ void weights_from_distances(R d[6], R clamp, R w[3]) {
#pragma unroll
  for (int f = 0; f < 6; ++f)
    if (d[f] <= clamp) d[f] = clamp;
  if constexpr (sizeof(R) == 4) {
    // FP32 mode: hardware reciprocal (1 ulp), weights within a few ulp
    w[0] = d[F_LEFT] * __builtin_amdgcn_rcpf(d[F_LEFT] + d[F_RIGHT]);
    w[1] = d[F_BOTTOM] * __builtin_amdgcn_rcpf(d[F_BOTTOM] + d[F_TOP]);
    w[2] = d[F_BACK] * __builtin_amdgcn_rcpf(d[F_FRONT] + d[F_BACK]);
  } else {
    w[0] = d[F_LEFT] / (d[F_LEFT] + d[F_RIGHT]);
    w[1] = d[F_BOTTOM] / (d[F_BOTTOM] + d[F_TOP]);
    w[2] = d[F_BACK] / (d[F_FRONT] + d[F_BACK]);
  }
}

// ---------------------------------------------------------------------------
// FP32 filter.
//
// With u = 2^-24, Lc = max |component| of the vertex coordinates relative to
// V0 and Lw = max(Lc, max |component| of p - V0), a first-order rounding
// analysis of the FP32 evaluation (inputs rounded once to FP32, every
// product and sum rounded once, sqrt and reciprocal within 1 ulp as
// v_sqrt_f32 / v_rcp_f32 guarantee) bounds
//   |d32 - d|      <= u * Lw * (846 * Lc^2 / m + 21)
//   |dot32 - dot|  <= u * 600 * Lc^3
//   |m32 - m|      <= u * (84 * Lc^2 + 2 m)
// where d, dot, m are the exact values on the FP64 inputs; the FP64
// reference values differ from those by ~2^-29 of the same bounds.  The
// constants below (1040, 32, 1024, 128, 3) keep a margin over the analysis.
// Every edge has components <= 2 Lc, so the reference's zero band
// thr = 1e-11 * h is below 1e-11 * 6 Lc.  A face is certified when the
// decision the reference makes on it lies outside its bound.
// ---------------------------------------------------------------------------
template <int F>
__device__ __forceinline__ bool face_fp32(const float r[8][3], const float cc[3], const float q[3], float Lc2,
                                          float Lc3, float Lw, float thrU, float d[6]) {
  using D = FaceDef<F>;
  const float *v1 = r[D::a], *v2 = r[D::b], *v3 = r[D::c], *v4 = r[D::d];
  const float u = 5.9604645e-08f;  // 2^-24
  float e1x = v2[0] - v1[0], e1y = v2[1] - v1[1], e1z = v2[2] - v1[2];
  float e2x = v4[0] - v1[0], e2y = v4[1] - v1[1], e2z = v4[2] - v1[2];
  float nx = __builtin_fmaf(e1y, e2z, -e1z * e2y);
  float ny = __builtin_fmaf(e1z, e2x, -e1x * e2z);
  float nz = __builtin_fmaf(e1x, e2y, -e1y * e2x);
  // hardware sqrt and reciprocal (1 ulp each); the bounds below carry them
  float m = __builtin_amdgcn_sqrtf(__builtin_fmaf(nx, nx, __builtin_fmaf(ny, ny, nz * nz)));
  const float rm = __builtin_amdgcn_rcpf(m);
  float fcx = 0.25f * ((v1[0] + v2[0]) + (v3[0] + v4[0]));
  float fcy = 0.25f * ((v1[1] + v2[1]) + (v3[1] + v4[1]));
  float fcz = 0.25f * ((v1[2] + v2[2]) + (v3[2] + v4[2]));
  float dot = __builtin_fmaf(nx, cc[0] - fcx, __builtin_fmaf(ny, cc[1] - fcy, nz * (cc[2] - fcz)));
  float dn = __builtin_fmaf(nx, fcx - q[0], __builtin_fmaf(ny, fcy - q[1], nz * (fcz - q[2])));
  const float Em = u * (128.f * Lc2 + 3.f * m);
  const float Edot = u * 1024.f * Lc3;
  const float flip = (dot > 1.0e-9f) ? -1.f : 1.f;
  const float dd = flip * (dn * rm);
  const float Ed = u * Lw * (1040.f * Lc2 * rm + 32.f);
  d[D::slot] = dd;
  return (m - Em > 1.0e-12f * 1.001f) && (fabsf(dot - 1.0e-9f) > Edot + 1.0e-15f) && (fabsf(dd) > thrU + Ed);
}

// Out-of-line FP64 recomputation (reference arithmetic, vertices re-read
// from memory) for the rare uncertain faces.  Arguments and results are
// passed by value so that the caller's arrays stay in registers.
struct ExactCell {
  double thr, c0, c1, c2;  // zero band 1e-11*h and cell centroid
};
__device__ __noinline__ ExactCell exact_cell_info(const GridView g, int i, int j, int k, double thresh_scale) {
  double V[8][3];
  load_vertices(g, i, j, k, V);
  ExactCell e;
  e.thr = thresh_scale * char_size_exact(V);
  double cc[3];
  centroid_exact(V, cc);
  e.c0 = cc[0];
  e.c1 = cc[1];
  e.c2 = cc[2];
  return e;
}
struct ExactFace {
  double d;
  int ok;  // 0: degenerate face
};
__device__ __noinline__ ExactFace exact_face(const double *coor, int IM, int JM, int64_t base, int f, double p0,
                                             double p1, double p2, double thr, double c0, double c1, double c2) {
  const GridView g{coor, IM, JM, 0};
  const double *v1 = coor + 3 * vertex_node(g, base, kFaceVerts[f][0]);
  const double *v2 = coor + 3 * vertex_node(g, base, kFaceVerts[f][1]);
  const double *v3 = coor + 3 * vertex_node(g, base, kFaceVerts[f][2]);
  const double *v4 = coor + 3 * vertex_node(g, base, kFaceVerts[f][3]);
  const double cc[3] = {c0, c1, c2};
  const double p[3] = {p0, p1, p2};
  ExactFace r;
  r.d = 0.0;
  r.ok = face_distance_exact(v1, v2, v3, v4, cc, p, thr, r.d) ? 1 : 0;
  return r;
}

// store into slot kFaceSlot[f] with compile-time indices
__device__ __forceinline__ void set_face_slot(float d[6], int f, float v) {
  switch (f) {
    case 0: d[F_BACK] = v; break;
    case 1: d[F_FRONT] = v; break;
    case 2: d[F_BOTTOM] = v; break;
    case 3: d[F_TOP] = v; break;
    case 4: d[F_LEFT] = v; break;
    default: d[F_RIGHT] = v; break;
  }
}

struct FilterResult {
  int fallbacks;  // faces recomputed in FP64
  bool degenerate;
};

// Recomputes the faces flagged in need (bit f: face f in evaluation order) of
// cell (i,j,k) in FP64 with the reference arithmetic.
__device__ __forceinline__ void exact_fallback(const GridView &g, int i, int j, int k, const double p[3],
                                               double thresh_scale, unsigned need, float d[6], FilterResult &res) {
  const ExactCell e = exact_cell_info(g, i, j, k, thresh_scale);
  const int64_t base = cell_base_node(g, i, j, k);
#pragma unroll 1
  for (int f = 0; f < 6; ++f) {
    if (!((need >> f) & 1u)) continue;
    const ExactFace ef = exact_face(g.coor, g.IM, g.JM, base, f, p[0], p[1], p[2], e.thr, e.c0, e.c1, e.c2);
    if (!ef.ok) {
      res.degenerate = true;
      return;
    }
    float df = (float)ef.d;
    // keep the exact classification: a nonzero FP64 distance never maps to 0
    if (ef.d != 0.0 && df == 0.f) df = ef.d < 0.0 ? -1e-30f : 1e-30f;
    set_face_slot(d, f, df);
    res.fallbacks++;
  }
}

// V: the cell's vertices (already loaded); g, (i,j,k): for the FP64 fallback.
__device__ __forceinline__ FilterResult eval_cell_fp32_v(const double V[8][3], const GridView &g, int i, int j, int k,
                                                         const double p[3], double thresh_scale, float d[6]) {
  FilterResult res{0, false};
  float r[8][3];
  float q[3];
  {
#pragma unroll
    for (int v = 0; v < 8; ++v)
#pragma unroll
      for (int c = 0; c < 3; ++c) r[v][c] = (float)(V[v][c] - V[0][c]);
#pragma unroll
    for (int c = 0; c < 3; ++c) q[c] = (float)(p[c] - V[0][c]);
  }
  float Lc = 0.f;
#pragma unroll
  for (int v = 1; v < 8; ++v)
#pragma unroll
    for (int c = 0; c < 3; ++c) Lc = fmaxf(Lc, fabsf(r[v][c]));
  const float Lw = fmaxf(Lc, fmaxf(fabsf(q[0]), fmaxf(fabsf(q[1]), fabsf(q[2]))));
  const float Lc2 = Lc * Lc, Lc3 = Lc2 * Lc;
  const float thrU = (float)thresh_scale * 6.0f * Lc * 1.001f + 1e-30f;
  float cc[3];
#pragma unroll
  for (int c = 0; c < 3; ++c)
    cc[c] = (((r[0][c] + r[1][c]) + (r[2][c] + r[3][c])) + ((r[4][c] + r[5][c]) + (r[6][c] + r[7][c]))) * 0.125f;
  unsigned need = 0;
  need |= face_fp32<0>(r, cc, q, Lc2, Lc3, Lw, thrU, d) ? 0u : 1u;
  need |= face_fp32<1>(r, cc, q, Lc2, Lc3, Lw, thrU, d) ? 0u : 2u;
  need |= face_fp32<2>(r, cc, q, Lc2, Lc3, Lw, thrU, d) ? 0u : 4u;
  need |= face_fp32<3>(r, cc, q, Lc2, Lc3, Lw, thrU, d) ? 0u : 8u;
  need |= face_fp32<4>(r, cc, q, Lc2, Lc3, Lw, thrU, d) ? 0u : 16u;
  need |= face_fp32<5>(r, cc, q, Lc2, Lc3, Lw, thrU, d) ? 0u : 32u;
  if (need) exact_fallback(g, i, j, k, p, thresh_scale, need, d, res);
  return res;
}

__device__ __forceinline__ FilterResult eval_cell_fp32(const GridView &g, int i, int j, int k, const double p[3],
                                                       double thresh_scale, float d[6]) {
  double V[8][3];
  load_vertices(g, i, j, k, V);
  return eval_cell_fp32_v(V, g, i, j, k, p, thresh_scale, d);
}

// ---------------------------------------------------------------------------
// Face planes per cell.  The grid does not move during the particle steps, so
// the cell-only part of every walk evaluation (walkingsearch.c:25-59, 102-227,
// 287-295: unit normals, face centres, zero band) is computed once per grid,
// with the reference arithmetic, and stored per cell in the faces' evaluation
// order.  A walk evaluation is then six plane distances.
// ---------------------------------------------------------------------------

// FP64 record: the reference's own intermediate values, so that
// face_point_distance on them reproduces the reference distance bit for bit.
struct alignas(16) CellPlanes64 {
  double n[6][3];   // unit normal, evaluation order BACK FRONT BOTTOM TOP LEFT RIGHT
  double fc[6][3];  // face centre
  double thr;       // zero band thresh_scale * h; negative: the cell has a degenerate face
  double pad;
};
static_assert(sizeof(CellPlanes64) == 304, "CellPlanes64 layout");

// FP32 record: normal and offset relative to vertex 0 of the cell,
// D = n . (fc - V0) (FP64, rounded once), for the certified FP32 filter.
struct alignas(16) CellPlanes32 {
  float4 plane[6];  // (nx, ny, nz, D), evaluation order
  double v0[3];     // vertex 0 of the cell
  float thr;        // zero band rounded up; +inf: recompute every face in FP64
  float amax;       // max |D| over the faces
};
static_assert(sizeof(CellPlanes32) == 128, "CellPlanes32 layout");

template <int F>
__device__ __forceinline__ bool plane_exact(const double V[8][3], const double cc[3], CellPlanes64 &r) {
  using D = FaceDef<F>;
  return face_plane_exact(V[D::a], V[D::b], V[D::c], V[D::d], cc, r.n[F], r.fc[F]);
}

// walkingsearch.c:811-859 for cell (i,j,k) without the point
__device__ __forceinline__ void cell_planes_exact(const GridView &g, int i, int j, int k, double thresh_scale,
                                                  CellPlanes64 &r) {
  double V[8][3];
  load_vertices(g, i, j, k, V);
  const double thr = thresh_scale * char_size_exact(V);
  double cc[3];
  centroid_exact(V, cc);
  const bool ok = plane_exact<0>(V, cc, r) & plane_exact<1>(V, cc, r) & plane_exact<2>(V, cc, r) &
                  plane_exact<3>(V, cc, r) & plane_exact<4>(V, cc, r) & plane_exact<5>(V, cc, r);
  r.thr = ok ? thr : -1.0;
  r.pad = 0.0;
}

// FP64 evaluation on a record: the reference distances (walkingsearch.c:271-389).
template <int F>
__device__ __forceinline__ void plane_distance(const CellPlanes64 &r, const double p[3], double d[6]) {
  d[FaceDef<F>::slot] = face_point_distance(r.n[F], r.fc[F], p, r.thr);
}
__device__ __forceinline__ bool eval_planes_exact(const CellPlanes64 &r, const double p[3], double d[6]) {
  if (!(r.thr >= 0.0)) return false;
  plane_distance<0>(r, p, d);
  plane_distance<1>(r, p, d);
  plane_distance<2>(r, p, d);
  plane_distance<3>(r, p, d);
  plane_distance<4>(r, p, d);
  plane_distance<5>(r, p, d);
  return true;
}

// FP32 filter on a record.  With q = fl32(p - V0) (FP64 difference rounded
// once), Q = max |q_c|, u = 2^-24 and A = max |D|, the evaluation
//   d32 = D - fma(nx, qx, fma(ny, qy, nz * qz))
// differs from d* = (fc - p) . n (exact arithmetic on the reference's FP64 n
// and fc) by at most u * (2 A + 10.5 Q) (rounding of D, of n, of q, two
// roundings in the fma chain, one in the product and one in the difference);
// the reference's FP64 evaluation differs from d* by ~2^-29 of that.  A face
// is certified when |d32| > thr + u * (4 A + 16 Q): the reference's distance
// then has the sign of d32 and lies outside its zero band.  Other faces, and
// every face of a cell whose record holds thr = +inf (degenerate or extreme
// cells), are recomputed by the out-of-line FP64 routine.
template <int F>
__device__ __forceinline__ bool plane_fp32(const CellPlanes32 &r, const float q[3], float band, float d[6]) {
  const float4 pl = r.plane[F];
  const float dd = pl.w - __builtin_fmaf(pl.x, q[0], __builtin_fmaf(pl.y, q[1], pl.z * q[2]));
  d[FaceDef<F>::slot] = dd;
  return fabsf(dd) > band;
}

// FP32 distances; returns the faces (bit f: face f in evaluation order) that
// the bound does not certify.
__device__ __forceinline__ unsigned planes_fp32_distances(const CellPlanes32 &r, const double p[3], float d[6]) {
  const float q[3] = {(float)(p[0] - r.v0[0]), (float)(p[1] - r.v0[1]), (float)(p[2] - r.v0[2])};
  const float Q = fmaxf(fabsf(q[0]), fmaxf(fabsf(q[1]), fabsf(q[2])));
  const float u = 5.9604645e-08f;  // 2^-24
  const float band = r.thr + u * (4.f * r.amax + 16.f * Q);
  unsigned need = 0;
  need |= plane_fp32<0>(r, q, band, d) ? 0u : 1u;
  need |= plane_fp32<1>(r, q, band, d) ? 0u : 2u;
  need |= plane_fp32<2>(r, q, band, d) ? 0u : 4u;
  need |= plane_fp32<3>(r, q, band, d) ? 0u : 8u;
  need |= plane_fp32<4>(r, q, band, d) ? 0u : 16u;
  need |= plane_fp32<5>(r, q, band, d) ? 0u : 32u;
  return need;
}

__device__ __forceinline__ FilterResult eval_planes_fp32(const CellPlanes32 &r, const GridView &g, int i, int j, int k,
                                                         const double p[3], double thresh_scale, float d[6]) {
  FilterResult res{0, false};
  const unsigned need = planes_fp32_distances(r, p, d);
  if (need) exact_fallback(g, i, j, k, p, thresh_scale, need, d, res);
  return res;
}

// Distances of x to the faces of cell (i,j,k) from its plane record (planes
// points to CellPlanes64 records for FP64, CellPlanes32 for FP32).  False:
// the cell has a degenerate face.
template <typename R, bool EXACT>
__device__ __forceinline__ bool eval_cell_planes(const void *planes, const GridView &g, int i, int j, int k,
                                                 const double x[3], double thresh_scale, R d[6],
                                                 unsigned long long &fallbacks) {
  const int64_t c = (int64_t)i + (int64_t)g.IM * ((int64_t)j + (int64_t)g.JM * k);
  if constexpr (EXACT) {
    static_assert(sizeof(R) == sizeof(double), "exact evaluation is FP64");
    return eval_planes_exact(static_cast<const CellPlanes64 *>(planes)[c], x, d);
  } else {
    static_assert(sizeof(R) == sizeof(float), "filtered evaluation is FP32");
    FilterResult fr = eval_planes_fp32(static_cast<const CellPlanes32 *>(planes)[c], g, i, j, k, x, thresh_scale, d);
    fallbacks += fr.fallbacks;
    return !fr.degenerate;
  }
}

}  // namespace picgpu
