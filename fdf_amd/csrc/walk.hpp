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

// walkingsearch.c:102-227.  Returns false on a degenerate face.
__device__ __forceinline__ bool face_distance_exact(const double *v1, const double *v2, const double *v3,
                                                    const double *v4, const double cc[3], const double p[3],
                                                    double thr, double &d) {
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
  nx /= m;
  ny /= m;
  nz /= m;
  double wx = fcx - p[0], wy = fcy - p[1], wz = fcz - p[2];
  double dd = wx * nx + wy * ny + wz * nz;
  if (fabs(dd) < thr) dd = 0.0;
  d = dd;
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

// walkingsearch.c:407-462
template <typename R>
__device__ __forceinline__ int point_position(const R d[6]) {
  bool inside = true;
  int zeros = 0;
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    if (d[f] < (R)0) inside = false;
    if (d[f] == (R)0) zeros++;
  }
  if (inside) return 0;
  if (zeros == 0) return -1;
  return zeros >= 3 ? 3 : zeros;
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
  w[0] = d[F_LEFT] / (d[F_LEFT] + d[F_RIGHT]);
  w[1] = d[F_BOTTOM] / (d[F_BOTTOM] + d[F_TOP]);
  w[2] = d[F_BACK] / (d[F_FRONT] + d[F_BACK]);
}

// ---------------------------------------------------------------------------
// FP32 filter.
//
// With u = 2^-24, Lc = max |component| of the vertex coordinates relative to
// V0 and Lw = max(Lc, max |component| of p - V0), a first-order rounding
// analysis of the FP32 evaluation (inputs rounded once to FP32, every
// product, sum, sqrt and quotient rounded once) bounds
//   |d32 - d|      <= u * Lw * (846 * Lc^2 / m + 7)
//   |dot32 - dot|  <= u * 600 * Lc^3
//   |m32 - m|      <= u * (84 * Lc^2 + m)
// where d, dot, m are the exact values on the FP64 inputs; the FP64
// reference values differ from those by ~2^-29 of the same bounds.  The
// constants below (1024, 16, 1024, 128) keep a margin over the analysis.
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
  float m = sqrtf(__builtin_fmaf(nx, nx, __builtin_fmaf(ny, ny, nz * nz)));
  float fcx = 0.25f * ((v1[0] + v2[0]) + (v3[0] + v4[0]));
  float fcy = 0.25f * ((v1[1] + v2[1]) + (v3[1] + v4[1]));
  float fcz = 0.25f * ((v1[2] + v2[2]) + (v3[2] + v4[2]));
  float dot = __builtin_fmaf(nx, cc[0] - fcx, __builtin_fmaf(ny, cc[1] - fcy, nz * (cc[2] - fcz)));
  float dn = __builtin_fmaf(nx, fcx - q[0], __builtin_fmaf(ny, fcy - q[1], nz * (fcz - q[2])));
  const float Em = u * (128.f * Lc2 + 2.f * m);
  const float Edot = u * 1024.f * Lc3;
  const float flip = (dot > 1.0e-9f) ? -1.f : 1.f;
  const float dd = flip * (dn / m);
  const float Ed = u * Lw * (1024.f * Lc2 / m + 16.f);
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
  if (need) {
    const ExactCell e = exact_cell_info(g, i, j, k, thresh_scale);
    const int64_t base = cell_base_node(g, i, j, k);
#pragma unroll 1
    for (int f = 0; f < 6; ++f) {
      if (!((need >> f) & 1u)) continue;
      const ExactFace ef = exact_face(g.coor, g.IM, g.JM, base, f, p[0], p[1], p[2], e.thr, e.c0, e.c1, e.c2);
      if (!ef.ok) {
        res.degenerate = true;
        return res;
      }
      float df = (float)ef.d;
      // keep the exact classification: a nonzero FP64 distance never maps to 0
      if (ef.d != 0.0 && df == 0.f) df = ef.d < 0.0 ? -1e-30f : 1e-30f;
      set_face_slot(d, f, df);
      res.fallbacks++;
    }
  }
  return res;
}

__device__ __forceinline__ FilterResult eval_cell_fp32(const GridView &g, int i, int j, int k, const double p[3],
                                                       double thresh_scale, float d[6]) {
  double V[8][3];
  load_vertices(g, i, j, k, V);
  return eval_cell_fp32_v(V, g, i, j, k, p, thresh_scale, d);
}

}  // namespace picgpu
