// eulerian.hip -- Eulerian field preparation on gfx950: the stage that turns
// the contravariant field Ucont into the Cartesian Ucat the particle loop
// interpolates (SetEulerianFields, simulation.c:38-107), with the grid
// metrics it needs (Metric.c:195-622) and the boundary handlers
// (BC_Handlers.c, Boundaries.c).  FP64 throughout, compiled without FP
// contraction, so every value rounds as the reference's -O0 C does.
//
// All arrays are node-slot arrays [mz][my][mx] (mx = IM+1, i fastest), the
// reference's DMDA layout on one rank, where local and global vectors agree
// (DM_BOUNDARY_NONE, grid.c:59-60).
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <strings.h>

#include <algorithm>
#include <vector>

#include "picgpu_internal.hpp"

using namespace picgpu;

namespace {

constexpr int kEBlock = 256;
// Contra2Cart tile: 64 i x 8 j threads marching over kMarch k planes
constexpr int kTx = 64, kTy = 8, kMarch = 32;

inline int64_t eblocks(int64_t n) { return (n + kEBlock - 1) / kEBlock; }

struct NodeDims {
  int mx, my, mz;
  __host__ __device__ int64_t at(int i, int j, int k) const { return ((int64_t)k * my + j) * mx + i; }
  __host__ __device__ int64_t count() const { return (int64_t)mx * my * mz; }
};

struct V3 {
  double x, y, z;
};

__device__ __forceinline__ V3 ld3(const double *__restrict__ p, int64_t n) {
  return {p[3 * n], p[3 * n + 1], p[3 * n + 2]};
}
__device__ __forceinline__ void st3(double *__restrict__ p, int64_t n, V3 v) {
  p[3 * n] = v.x;
  p[3 * n + 1] = v.y;
  p[3 * n + 2] = v.z;
}
__device__ __forceinline__ double comp(V3 v, int c) { return c == 0 ? v.x : c == 1 ? v.y : v.z; }

// The plane a node's value comes from after one axis' boundary copies: plane
// 0 takes plane 1, then plane m-1 takes plane m-2 (which, for m = 2, already
// holds plane 1's value) (Metric.c:408-476, 566-612).  Never 0, at most m-1.
__device__ __forceinline__ int copy_src(int x, int m) {
  const int y = (x == m - 1) ? m - 2 : x;
  return y == 0 ? 1 : y;
}

// 0.5 * (a + b - c - e), summed left to right (Metric.c:333-338)
__device__ __forceinline__ V3 half_diff(V3 a, V3 b, V3 c, V3 e) {
  return {0.5 * (a.x + b.x - c.x - e.x), 0.5 * (a.y + b.y - c.y - e.y), 0.5 * (a.z + b.z - c.z - e.z)};
}
__device__ __forceinline__ V3 cross(V3 u, V3 v) {
  return {u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x};
}

// Face metric vectors at node (i,j,k): the cross product of the face's two
// mean edge vectors (Metric.c:330-404).  csi needs j,k >= 1; eta i,k >= 1;
// zet i,j >= 1.
__device__ V3 csi_at(const double *X, NodeDims d, int i, int j, int k) {
  const V3 deta = half_diff(ld3(X, d.at(i, j, k)), ld3(X, d.at(i, j, k - 1)), ld3(X, d.at(i, j - 1, k)),
                            ld3(X, d.at(i, j - 1, k - 1)));
  const V3 dzeta = half_diff(ld3(X, d.at(i, j - 1, k)), ld3(X, d.at(i, j, k)), ld3(X, d.at(i, j - 1, k - 1)),
                             ld3(X, d.at(i, j, k - 1)));
  return cross(deta, dzeta);
}
__device__ V3 eta_at(const double *X, NodeDims d, int i, int j, int k) {
  const V3 dxi = half_diff(ld3(X, d.at(i, j, k)), ld3(X, d.at(i, j, k - 1)), ld3(X, d.at(i - 1, j, k)),
                           ld3(X, d.at(i - 1, j, k - 1)));
  const V3 dzeta = half_diff(ld3(X, d.at(i, j, k)), ld3(X, d.at(i - 1, j, k)), ld3(X, d.at(i, j, k - 1)),
                             ld3(X, d.at(i - 1, j, k - 1)));
  return cross(dzeta, dxi);
}
__device__ V3 zet_at(const double *X, NodeDims d, int i, int j, int k) {
  const V3 dxi = half_diff(ld3(X, d.at(i, j, k)), ld3(X, d.at(i, j - 1, k)), ld3(X, d.at(i - 1, j, k)),
                           ld3(X, d.at(i - 1, j - 1, k)));
  const V3 deta = half_diff(ld3(X, d.at(i, j, k)), ld3(X, d.at(i - 1, j, k)), ld3(X, d.at(i, j - 1, k)),
                            ld3(X, d.at(i - 1, j - 1, k)));
  return cross(dxi, deta);
}

// Every node computes its final (post-copy) vectors directly at the source
// node of the composed copies, so the kernel reads only coordinates.
__global__ __launch_bounds__(kEBlock) void k_face_metrics(NodeDims d, const double *__restrict__ X,
                                                          double *__restrict__ csi, double *__restrict__ eta,
                                                          double *__restrict__ zet) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= d.count()) return;
  const int i = (int)(n % d.mx);
  const int64_t r = n / d.mx;
  const int j = (int)(r % d.my), k = (int)(r / d.my);
  st3(csi, n, csi_at(X, d, i, copy_src(j, d.my), copy_src(k, d.mz)));
  st3(eta, n, eta_at(X, d, copy_src(i, d.mx), j, copy_src(k, d.mz)));
  st3(zet, n, zet_at(X, d, copy_src(i, d.mx), copy_src(j, d.my), k));
}

// 0.25 * ((a+b+c+e) - (f+g+h+l)) (Metric.c:533-550)
__device__ __forceinline__ double quarter_diff(double a, double b, double c, double e, double f, double g, double h,
                                               double l) {
  return 0.25 * ((a + b + c + e) - (f + g + h + l));
}

// Jacobian determinant over the 8 nodes of the box below-left of node (i,j,k),
// i,j,k >= 1 (Metric.c:531-552)
__device__ double jacobian_det(const double *X, NodeDims d, int i, int j, int k) {
  const V3 p000 = ld3(X, d.at(i - 1, j - 1, k - 1)), p100 = ld3(X, d.at(i, j - 1, k - 1)),
           p010 = ld3(X, d.at(i - 1, j, k - 1)), p110 = ld3(X, d.at(i, j, k - 1)),
           p001 = ld3(X, d.at(i - 1, j - 1, k)), p101 = ld3(X, d.at(i, j - 1, k)), p011 = ld3(X, d.at(i - 1, j, k)),
           p111 = ld3(X, d.at(i, j, k));
  double dxi[3], deta[3], dzeta[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    dxi[c] = quarter_diff(comp(p111, c), comp(p101, c), comp(p110, c), comp(p100, c), comp(p011, c), comp(p001, c),
                          comp(p010, c), comp(p000, c));
    deta[c] = quarter_diff(comp(p111, c), comp(p011, c), comp(p110, c), comp(p010, c), comp(p101, c),
                           comp(p001, c), comp(p100, c), comp(p000, c));
    dzeta[c] = quarter_diff(comp(p111, c), comp(p101, c), comp(p011, c), comp(p001, c), comp(p110, c),
                            comp(p100, c), comp(p010, c), comp(p000, c));
  }
  return dxi[0] * (deta[1] * dzeta[2] - deta[2] * dzeta[1]) - dxi[1] * (deta[0] * dzeta[2] - deta[2] * dzeta[0]) +
         dxi[2] * (deta[0] * dzeta[1] - deta[1] * dzeta[0]);
}

// Aj = 1/det at the composed copy source; every node the reference computes
// (i,j,k >= 1) is also checked for |det| < 1e-12, reporting the smallest
// such node number.
__global__ __launch_bounds__(kEBlock) void k_jacobian_inverse(NodeDims d, const double *__restrict__ X,
                                                              double *__restrict__ aj,
                                                              unsigned long long *__restrict__ bad) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= d.count()) return;
  const int i = (int)(n % d.mx);
  const int64_t r = n / d.mx;
  const int j = (int)(r % d.my), k = (int)(r / d.my);
  const int si = copy_src(i, d.mx), sj = copy_src(j, d.my), sk = copy_src(k, d.mz);
  const bool own = i >= 1 && j >= 1 && k >= 1;
  const bool same = si == i && sj == j && sk == k;
  if (own && !same && fabs(jacobian_det(X, d, i, j, k)) < 1.0e-12) atomicMin(bad, (unsigned long long)n);
  const double det = jacobian_det(X, d, si, sj, sk);
  if (own && same && fabs(det) < 1.0e-12) atomicMin(bad, (unsigned long long)n);
  aj[n] = 1.0 / det;
}

// per-block min and max of v[0..n)
__global__ __launch_bounds__(kEBlock) void k_minmax(int64_t n, const double *__restrict__ v,
                                                    double *__restrict__ part) {
  __shared__ double smin[kEBlock], smax[kEBlock];
  double lo = INFINITY, hi = -INFINITY;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    const double x = v[t];
    lo = x < lo ? x : lo;
    hi = x > hi ? x : hi;
  }
  smin[threadIdx.x] = lo;
  smax[threadIdx.x] = hi;
  __syncthreads();
  for (int s = kEBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + s]);
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = smin[0];
    part[2 * blockIdx.x + 1] = smax[0];
  }
}

// x *= -1 over the four metric arrays (VecScale, Metric.c:236-239)
__global__ __launch_bounds__(kEBlock) void k_flip(int64_t nnode, double *__restrict__ csi, double *__restrict__ eta,
                                                  double *__restrict__ zet, double *__restrict__ aj) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < 3 * nnode) {
    csi[t] = csi[t] * -1.0;
    eta[t] = eta[t] * -1.0;
    zet[t] = zet[t] * -1.0;
  }
  if (t < nnode) aj[t] = aj[t] * -1.0;
}

__global__ void k_fill_u64(unsigned long long *p, int n, unsigned long long v) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = v;
}

template <typename T>
__global__ __launch_bounds__(kEBlock) void k_zero(int64_t n, T *__restrict__ p) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = (T)0;
}

// SetInitialInteriorField("Ucont") (AnalyticalSolution.c:1052-1168): every
// interior node gets zero except along the inlet axis, where the signed
// normal speed times the face area |metric| goes.
__global__ __launch_bounds__(kEBlock) void k_initial_ucont(NodeDims d, const double *__restrict__ csi,
                                                           const double *__restrict__ eta,
                                                           const double *__restrict__ zet, int finit, double u_axis,
                                                           int inlet_face, int orientation,
                                                           double *__restrict__ ucont) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= d.count()) return;
  const int i = (int)(n % d.mx);
  const int64_t r = n / d.mx;
  const int j = (int)(r % d.my), k = (int)(r / d.my);
  if (i <= 0 || i >= d.mx - 1 || j <= 0 || j >= d.my - 1 || k <= 0 || k >= d.mz - 1) return;
  const int axis = inlet_face / 2;
  double mag = 0.0;
  if (finit == 1) {
    mag = u_axis;
  } else if (finit == 2) {
    const double iw = (double)(d.mx - 2), jw = (double)(d.my - 2);
    const double ic = 1.0 + iw / 2.0, jc = 1.0 + jw / 2.0;
    const double in = ((double)i - ic) / (iw / 2.0), jn = ((double)j - jc) / (jw / 2.0);
    const double pi_ = 1.0 - in * in, pj = 1.0 - jn * jn;
    mag = u_axis * pi_ * pj;
    if (mag < 0.0) mag = 0.0;
  }
  V3 out = {0.0, 0.0, 0.0};
  if (mag != 0.0) {
    const double sign = (inlet_face % 2 == 0) ? 1.0 : -1.0;
    const double v = mag * sign * (double)orientation;
    const V3 m = ld3(axis == 0 ? csi : axis == 1 ? eta : zet, n);
    const double a = v * sqrt(m.x * m.x + m.y * m.y + m.z * m.z);
    if (axis == 0) out.x = a;
    else if (axis == 1) out.y = a;
    else out.z = a;
  }
  st3(ucont, n, out);
}

// the particle loop's cell-linear copy of Ucat, written beside the node array
template <typename R>
__device__ __forceinline__ void mirror_store(R *mirror, int IM, int JM, int KM, int i, int j, int k, V3 v) {
  if (!mirror || i >= IM || j >= JM || k >= KM) return;
  const int64_t c = ((int64_t)k * JM + j) * IM + i;
  mirror[3 * c] = (R)v.x;
  mirror[3 * c + 1] = (R)v.y;
  mirror[3 * c + 2] = (R)v.z;
}

struct FaceBC {
  int face, handler;
  double vel[3];
  double vmax;
};

// One boundary handler on one face: one thread per node of the face plane
// p (0 or m-1 along the face axis); q is the plane inside next to it and u
// the plane of the face's contravariant flux (0 or m-2).  Within a face no
// thread reads what another writes.
template <typename R>
__global__ __launch_bounds__(kEBlock) void k_bc_face(NodeDims d, FaceBC bc, const double *__restrict__ csi,
                                                     const double *__restrict__ eta, const double *__restrict__ zet,
                                                     double *__restrict__ ucat, double *__restrict__ ucont,
                                                     R *__restrict__ mirror, int IM, int JM, int KM) {
  const int axis = bc.face >> 1, pos = bc.face & 1;
  const int na = axis == 0 ? d.my : d.mx;
  const int nb = axis == 2 ? d.my : d.mz;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)na * nb) return;
  const int a = (int)(t % na), b = (int)(t / na);
  const int m = axis == 0 ? d.mx : axis == 1 ? d.my : d.mz;
  const int p = pos ? m - 1 : 0, q = pos ? p - 1 : p + 1, u = pos ? p - 1 : p;
  int i, j, k;
  if (axis == 0) {
    i = p;
    j = a;
    k = b;
  } else if (axis == 1) {
    i = a;
    j = p;
    k = b;
  } else {
    i = a;
    j = b;
    k = p;
  }
  const int64_t np = d.at(i, j, k);
  const int64_t stride = axis == 0 ? 1 : axis == 1 ? (int64_t)d.mx : (int64_t)d.mx * d.my;
  const int64_t nq = np + (q - p) * stride, nu = np + (u - p) * stride;
  const double *metric = axis == 0 ? csi : axis == 1 ? eta : zet;
  V3 out;
  switch (bc.handler) {
    case PICGPU_BC_WALL_NOSLIP: {  // BC_Handlers.c:58-165
      ucont[3 * nu + axis] = 0.0;
      const V3 in = ld3(ucat, nq);
      out = {-in.x, -in.y, -in.z};
      break;
    }
    case PICGPU_BC_NOGRAD_COPY_GHOST:  // BC_Handlers.c:500-590
      out = ld3(ucat, nq);
      if (!pos || m >= 3) ucont[3 * np + axis] = ucont[3 * nq + axis];
      break;
    case PICGPU_BC_INLET_CONSTANT_VELOCITY: {  // BC_Handlers.c:345-372
      out = {bc.vel[0], bc.vel[1], bc.vel[2]};
      const V3 mm = ld3(metric, nu);
      ucont[3 * nu + axis] = bc.vel[0] * mm.x + bc.vel[1] * mm.y + bc.vel[2] * mm.z;
      break;
    }
    case PICGPU_BC_INLET_PARABOLIC: {  // BC_Handlers.c:720-780, k faces only
      if (axis != 2) return;
      const double iw = (double)(d.mx - 2), jw = (double)(d.my - 2);
      const double ic = 1.0 + iw / 2.0, jc = 1.0 + jw / 2.0;
      const double in = ((double)i - ic) / (iw / 2.0), jn = ((double)j - jc) / (jw / 2.0);
      const double pi_ = 1.0 - in * in, pj = 1.0 - jn * jn;
      double mag = bc.vmax * pi_ * pj;
      if (mag < 0.0) mag = 0.0;
      const V3 ar = ld3(zet, np);
      const double area = sqrt(ar.x * ar.x + ar.y * ar.y + ar.z * ar.z);
      V3 dir = {0.0, 0.0, 0.0};
      if (area > 1e-9) dir = {ar.x / area, ar.y / area, ar.z / area};
      out = {mag * dir.x, mag * dir.y, mag * dir.z};
      const V3 zu = ld3(zet, nu);
      ucont[3 * nu + 2] = out.x * zu.x + out.y * zu.y + out.z * zu.z;
      break;
    }
    default:
      return;
  }
  st3(ucat, np, out);
  mirror_store(mirror, IM, JM, KM, i, j, k, out);
}

// The eight nodes of the box (i-1..i, j-1..j, k-1..k): a[dk][dj][di], index 1
// for the node's own coordinate, 0 for the one below.  The face vectors of
// cell slot (i,j,k) that Contra2Cart averages are all functions of them, by
// the expressions of csi_at / eta_at / zet_at.
struct Box {
  V3 a[2][2][2];
};
__device__ __forceinline__ V3 box_csi(const Box &b, int di) {
  return cross(half_diff(b.a[1][1][di], b.a[0][1][di], b.a[1][0][di], b.a[0][0][di]),
               half_diff(b.a[1][0][di], b.a[1][1][di], b.a[0][0][di], b.a[0][1][di]));
}
__device__ __forceinline__ V3 box_eta(const Box &b, int dj) {
  const V3 dxi = half_diff(b.a[1][dj][1], b.a[0][dj][1], b.a[1][dj][0], b.a[0][dj][0]);
  const V3 dzeta = half_diff(b.a[1][dj][1], b.a[1][dj][0], b.a[0][dj][1], b.a[0][dj][0]);
  return cross(dzeta, dxi);
}
__device__ __forceinline__ V3 box_zet(const Box &b, int dk) {
  const V3 dxi = half_diff(b.a[dk][1][1], b.a[dk][0][1], b.a[dk][1][0], b.a[dk][0][0]);
  const V3 deta = half_diff(b.a[dk][1][1], b.a[dk][1][0], b.a[dk][0][1], b.a[dk][0][0]);
  return cross(dxi, deta);
}
__device__ __forceinline__ V3 scale3(V3 v, double s) { return {v.x * s, v.y * s, v.z * s}; }

// Contra2Cart (setup.c:1334-1406) for interior cells 1 <= i <= mx-2 (likewise
// j, k).  A block is a kTx x kTy tile of (i, j) marching over kMarch k planes.
// Instead of reading the stored Csi/Eta/Zet (72 B per cell) it stages each
// plane of node coordinates ((kTx+1) x (kTy+1) nodes, 24 B per cell plus the halo) in
// LDS and recomputes the six face vectors of each cell with the expressions
// of k_face_metrics, times `flip` (1, or -1 after a left-handed grid was
// flipped), so they are the stored values bit for bit: for interior cells
// none of them is a boundary copy.  zet and Ucont.z of plane k-1 carry over
// in registers.  Cramer's rule in the reference's expression order;
// |det| < 1e-18 reports the node.
template <typename R>
__global__ __launch_bounds__(kTx *kTy) void k_contra2cart(NodeDims d, const double *__restrict__ X, double flip,
                                                          const double *__restrict__ ucont, double *__restrict__ ucat,
                                                          R *__restrict__ mirror, int IM, int JM, int KM,
                                                          unsigned long long *__restrict__ bad) {
  constexpr int RowD = (kTx + 1) * 3;        // doubles per staged node row
  constexpr int PlaneD = (kTy + 1) * RowD;   // doubles per staged plane
  __shared__ double tile[2][PlaneD];
  const int tx = (int)threadIdx.x, ty = (int)threadIdx.y, t = ty * kTx + tx;
  const int i0 = 1 + (int)blockIdx.x * kTx, j0 = 1 + (int)blockIdx.y * kTy;
  const int i = i0 + tx, j = j0 + ty;
  const int k0 = 1 + (int)blockIdx.z * kMarch;
  const int k1 = min(k0 + kMarch, d.mz - 1);
  const bool live = i <= d.mx - 2 && j <= d.my - 2;
  const int rmax = 3 * (d.mx - i0 + 1);  // staged doubles of a row that lie inside the grid
  // each thread stages kPer doubles of a plane: fetched into registers one
  // plane ahead (the loads fly while the current plane is computed), then
  // committed to LDS
  constexpr int kNT = kTx * kTy, kPer = (PlaneD + kNT - 1) / kNT;
  double pre[kPer];
  auto fetch = [&](int k) {
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      const int f = t + q * kNT;
      const int jj = f / RowD, r = f - jj * RowD;
      const int jn = j0 - 1 + jj;
      pre[q] = (f < PlaneD && r < rmax && jn <= d.my - 1) ? X[3 * d.at(i0 - 1, jn, k) + r] : 0.0;
    }
  };
  auto commit = [&](double *dst) {
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      const int f = t + q * kNT;
      if (f < PlaneD) dst[f] = pre[q];
    }
  };
  auto node = [&](const double *src, int dj, int di) -> V3 {
    const double *p = src + (ty + dj) * RowD + 3 * (tx + di);
    return {p[0], p[1], p[2]};
  };
  fetch(k0 - 1);
  commit(tile[(k0 - 1) & 1]);
  if (k0 < k1) fetch(k0);
  __syncthreads();
  Box b;
  V3 zprev = {0.0, 0.0, 0.0};
  double wprev = 0.0;
  const int64_t plane = (int64_t)d.mx * d.my;
  int64_t n = live ? d.at(i, j, k0 - 1) : 0;
  if (live) {
    const double *lo = tile[(k0 - 1) & 1];
#pragma unroll
    for (int dj = 0; dj < 2; ++dj)
#pragma unroll
      for (int di = 0; di < 2; ++di) b.a[1][dj][di] = node(lo, dj, di);
    zprev = scale3(box_zet(b, 1), flip);
    wprev = ucont[3 * n + 2];
  }
  for (int k = k0; k < k1; ++k) {
    const double *hi = tile[k & 1];
    commit(tile[k & 1]);  // overwrites plane k-2, last read before the previous barrier
    __syncthreads();
    if (k + 1 < k1) fetch(k + 1);
    if (live) {
      n += plane;
#pragma unroll
      for (int dj = 0; dj < 2; ++dj)
#pragma unroll
        for (int di = 0; di < 2; ++di) {
          b.a[0][dj][di] = b.a[1][dj][di];
          b.a[1][dj][di] = node(hi, dj, di);
        }
      const V3 cm = scale3(box_csi(b, 0), flip), c0 = scale3(box_csi(b, 1), flip);
      const V3 em = scale3(box_eta(b, 0), flip), e0 = scale3(box_eta(b, 1), flip);
      const V3 z0 = scale3(box_zet(b, 1), flip);
      const double um = ucont[3 * (n - 1)], vm = ucont[3 * (n - d.mx) + 1];
      const V3 u0 = ld3(ucont, n);
      const double m00 = 0.5 * (cm.x + c0.x), m01 = 0.5 * (cm.y + c0.y), m02 = 0.5 * (cm.z + c0.z);
      const double m10 = 0.5 * (em.x + e0.x), m11 = 0.5 * (em.y + e0.y), m12 = 0.5 * (em.z + e0.z);
      const double m20 = 0.5 * (zprev.x + z0.x), m21 = 0.5 * (zprev.y + z0.y), m22 = 0.5 * (zprev.z + z0.z);
      const double q0 = 0.5 * (um + u0.x), q1 = 0.5 * (vm + u0.y), q2 = 0.5 * (wprev + u0.z);
      zprev = z0;
      wprev = u0.z;
      const double det =
          m00 * (m11 * m22 - m12 * m21) - m01 * (m10 * m22 - m12 * m20) + m02 * (m10 * m21 - m11 * m20);
      if (fabs(det) < 1.0e-18) {
        atomicMin(bad, (unsigned long long)n);
      } else {
        const double rdet = 1.0 / det;
        const double d0 = q0 * (m11 * m22 - m12 * m21) - q1 * (m01 * m22 - m02 * m21) + q2 * (m01 * m12 - m02 * m11);
        const double d1 = -q0 * (m10 * m22 - m12 * m20) + q1 * (m00 * m22 - m02 * m20) - q2 * (m00 * m12 - m02 * m10);
        const double d2 = q0 * (m10 * m21 - m11 * m20) - q1 * (m00 * m21 - m01 * m20) + q2 * (m00 * m11 - m01 * m10);
        const V3 out = {d0 * rdet, d1 * rdet, d2 * rdet};
        st3(ucat, n, out);
        mirror_store(mirror, IM, JM, KM, i, j, k, out);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
NodeDims node_dims(const picgpu_ctx *c) { return {c->IM + 1, c->JM + 1, c->KM + 1}; }

int need_metrics(picgpu_ctx *c) {
  if (!c->have_grid) {
    set_error("no grid uploaded");
    return PICGPU_E_STATE;
  }
  if (!c->eul.have_metrics) {
    set_error("picgpu_compute_metrics has not run for this grid");
    return PICGPU_E_STATE;
  }
  return PICGPU_OK;
}

int flag_reset(picgpu_ctx *c) {
  PG_LAUNCH(c, "flag_fill", 1, 64, k_fill_u64, c->eul.flag.as<unsigned long long>(), 2, ~0ull);
  return PICGPU_OK;
}

int flag_read(picgpu_ctx *c, unsigned long long *v) {
  PG_HIP(hipMemcpyAsync(v, c->eul.flag.ptr, sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
  PG_HIP(hipStreamSynchronize(c->stream));
  return PICGPU_OK;
}

void node_ijk(const picgpu_ctx *c, unsigned long long n, int *i, int *j, int *k) {
  const NodeDims d = node_dims(c);
  *i = (int)(n % (unsigned long long)d.mx);
  const unsigned long long r = n / (unsigned long long)d.mx;
  *j = (int)(r % (unsigned long long)d.my);
  *k = (int)(r / (unsigned long long)d.my);
}

// Launches k_bc_face for every face, faces 0..5 (Boundaries.c:611-617), or
// only the inlet handlers (their Initialize, BC_Handlers.c:238-256, 647-664).
int run_boundaries(picgpu_ctx *c, bool inlets_only) {
  const NodeDims d = node_dims(c);
  for (int f = 0; f < 6; ++f) {
    const int h = c->eul.bc.handler[f];
    const bool inlet = h == PICGPU_BC_INLET_CONSTANT_VELOCITY || h == PICGPU_BC_INLET_PARABOLIC;
    if (inlets_only && !inlet) continue;
    FaceBC fb;
    fb.face = f;
    fb.handler = h;
    fb.vel[0] = c->eul.bc.velocity[f][0];
    fb.vel[1] = c->eul.bc.velocity[f][1];
    fb.vel[2] = c->eul.bc.velocity[f][2];
    fb.vmax = c->eul.bc.vmax[f];
    const int axis = f / 2;
    const int64_t nface = (int64_t)(axis == 0 ? d.my : d.mx) * (axis == 2 ? d.my : d.mz);
    int r = c->opt.precision == PICGPU_FP32
                ? [&]() -> int {
                    PG_LAUNCH(c, "bc_face", eblocks(nface), kEBlock, k_bc_face<float>, d, fb,
                              c->eul.csi.as<double>(), c->eul.eta.as<double>(), c->eul.zet.as<double>(),
                              c->eul.ucat.as<double>(), c->eul.ucont.as<double>(), c->ucat.as<float>(), c->IM,
                              c->JM, c->KM);
                    return PICGPU_OK;
                  }()
                : [&]() -> int {
                    PG_LAUNCH(c, "bc_face", eblocks(nface), kEBlock, k_bc_face<double>, d, fb,
                              c->eul.csi.as<double>(), c->eul.eta.as<double>(), c->eul.zet.as<double>(),
                              c->eul.ucat.as<double>(), c->eul.ucont.as<double>(), c->ucat.as<double>(), c->IM,
                              c->JM, c->KM);
                    return PICGPU_OK;
                  }();
    PG_CHECK(r);
  }
  return PICGPU_OK;
}

int launch_contra2cart(picgpu_ctx *c) {
  const NodeDims d = node_dims(c);
  if (d.mx < 3 || d.my < 3 || d.mz < 3) return PICGPU_OK;  // no interior cell
  const dim3 grid((unsigned)((d.mx - 2 + kTx - 1) / kTx), (unsigned)((d.my - 2 + kTy - 1) / kTy),
                  (unsigned)((d.mz - 2 + kMarch - 1) / kMarch));
  const dim3 block(kTx, kTy, 1);
  auto *bad = c->eul.flag.as<unsigned long long>();
  const double flip = c->eul.flipped ? -1.0 : 1.0;
  if (c->opt.precision == PICGPU_FP32) {
    PG_LAUNCH(c, "contra2cart", grid, block, k_contra2cart<float>, d, c->coor.as<double>(), flip,
              c->eul.ucont.as<double>(), c->eul.ucat.as<double>(), c->ucat.as<float>(), c->IM, c->JM, c->KM, bad);
  } else {
    PG_LAUNCH(c, "contra2cart", grid, block, k_contra2cart<double>, d, c->coor.as<double>(), flip,
              c->eul.ucont.as<double>(), c->eul.ucat.as<double>(), c->ucat.as<double>(), c->IM, c->JM, c->KM,
              bad);
  }
  return PICGPU_OK;
}

int check_contra2cart(picgpu_ctx *c) {
  unsigned long long first = 0;
  PG_CHECK(flag_read(c, &first));
  if (first != ~0ull) {
    int i, j, k;
    node_ijk(c, first, &i, &j, &k);
    set_error("Transformation matrix determinant is near zero at cell (%d,%d,%d)", i, j, k);
    return PICGPU_E_SINGULAR;
  }
  return PICGPU_OK;
}

int run_contra2cart(picgpu_ctx *c) {
  PG_CHECK(flag_reset(c));
  PG_CHECK(launch_contra2cart(c));
  return check_contra2cart(c);
}

int run_initial_ucont(picgpu_ctx *c, int finit, const double *uin) {
  int inlet = PICGPU_FACE_NEG_X;  // default inlet face (setup.c:136, io.c:2280)
  for (int f = 0; f < 6; ++f)
    if (c->eul.bc.type[f] == PICGPU_BCTYPE_INLET) {
      inlet = f;
      break;
    }
  const double u_axis = uin ? uin[inlet / 2] : 0.0;
  const NodeDims d = node_dims(c);
  PG_LAUNCH(c, "initial_ucont", eblocks(d.count()), kEBlock, k_initial_ucont, d, c->eul.csi.as<double>(),
            c->eul.eta.as<double>(), c->eul.zet.as<double>(), finit, u_axis, inlet, c->eul.orientation,
            c->eul.ucont.as<double>());
  return PICGPU_OK;
}

int need_bc(picgpu_ctx *c) {
  PG_CHECK(need_metrics(c));
  if (!c->eul.have_bc) {
    set_error("picgpu_boundary_create has not run");
    return PICGPU_E_STATE;
  }
  return PICGPU_OK;
}

// ValidateBCHandlerForBCType (io.c:2473-2488)
bool handler_valid_for_type(int type, int handler) {
  switch (type) {
    case PICGPU_BCTYPE_NOGRAD: return handler == PICGPU_BC_NOGRAD_COPY_GHOST;
    case PICGPU_BCTYPE_WALL: return handler == PICGPU_BC_WALL_NOSLIP || handler == PICGPU_BC_WALL_MOVING;
    case PICGPU_BCTYPE_INLET:
      return handler == PICGPU_BC_INLET_CONSTANT_VELOCITY || handler == PICGPU_BC_INLET_PARABOLIC;
    default: return true;
  }
}

bool handler_built(int handler) {  // BoundaryCondition_Create, Boundaries.c:417-433
  return handler == PICGPU_BC_NOGRAD_COPY_GHOST || handler == PICGPU_BC_WALL_NOSLIP ||
         handler == PICGPU_BC_INLET_CONSTANT_VELOCITY || handler == PICGPU_BC_INLET_PARABOLIC;
}

int lookup(const char *s, const char *const *names, const int *values, int n, int *out) {
  for (int t = 0; t < n; ++t)
    if (strcasecmp(s, names[t]) == 0) {
      *out = values[t];
      return 1;
    }
  return 0;
}

}  // namespace

namespace picgpu {

void eulerian_release(picgpu_ctx *c) {
  DevBuf *bufs[] = {&c->eul.csi, &c->eul.eta, &c->eul.zet, &c->eul.aj, &c->eul.ucont, &c->eul.ucat,
                    &c->eul.flag, &c->eul.part};
  for (DevBuf *b : bufs) b->release();
  c->eul.have_metrics = c->eul.have_bc = false;
}

int eulerian_field(picgpu_ctx *c, int which, const double **d_ptr, int64_t *count) {
  PG_CHECK(need_metrics(c));
  const int64_t nn = c->nnode;
  switch (which) {
    case PICGPU_FIELD_UCAT: *d_ptr = c->eul.ucat.as<double>(); *count = 3 * nn; break;
    case PICGPU_FIELD_UCONT: *d_ptr = c->eul.ucont.as<double>(); *count = 3 * nn; break;
    case PICGPU_FIELD_CSI: *d_ptr = c->eul.csi.as<double>(); *count = 3 * nn; break;
    case PICGPU_FIELD_ETA: *d_ptr = c->eul.eta.as<double>(); *count = 3 * nn; break;
    case PICGPU_FIELD_ZET: *d_ptr = c->eul.zet.as<double>(); *count = 3 * nn; break;
    case PICGPU_FIELD_AJ: *d_ptr = c->eul.aj.as<double>(); *count = nn; break;
    default: set_error("unknown field %d", which); return PICGPU_E_ARG;
  }
  return PICGPU_OK;
}

int eulerian_fields_begin(picgpu_ctx *c) {
  PG_CHECK(need_bc(c));
  return flag_reset(c);
}

int eulerian_fields_async(picgpu_ctx *c) {
  PG_CHECK(run_boundaries(c, false));
  PG_CHECK(launch_contra2cart(c));
  return run_boundaries(c, false);
}

int eulerian_fields_end(picgpu_ctx *c) { return check_contra2cart(c); }

int eulerian_take_ucat(picgpu_ctx *c, const double *ucat, hipMemcpyKind kind) {
  if (!c->eul.have_metrics) return PICGPU_OK;
  PG_HIP(hipMemcpyAsync(c->eul.ucat.ptr, ucat, sizeof(double) * 3 * (size_t)c->nnode, kind, c->stream));
  PG_HIP(hipStreamSynchronize(c->stream));
  return PICGPU_OK;
}

}  // namespace picgpu

extern "C" {

void picgpu_default_boundary_config(picgpu_boundary_config *bc) {
  if (!bc) return;
  memset(bc, 0, sizeof(*bc));
  for (int f = 0; f < 6; ++f) {
    bc->type[f] = PICGPU_BCTYPE_WALL;
    bc->handler[f] = PICGPU_BC_WALL_NOSLIP;
    bc->vmax[f] = 1.0;  // BC_Handlers.c:656
  }
}

// ParseAllBoundaryConditions (io.c:2515-2680) on one rank, with the handler
// parameters the two inlet handlers read at Initialize (vx/vy/vz,
// BC_Handlers.c:246-251; vmax, :658-660).  A later line for the same face
// replaces the earlier one, parameters included.
int picgpu_parse_boundary_file(const char *path, picgpu_boundary_config *bc) {
  if (!path || !bc) {
    set_error("parse_boundary_file: bad arguments");
    return PICGPU_E_ARG;
  }
  static const char *const faces[] = {"-Xi", "+Xi", "-Eta", "+Eta", "-Zeta", "+Zeta"};
  static const int face_v[] = {0, 1, 2, 3, 4, 5};
  static const char *const types[] = {"WALL", "SYMMETRY", "INLET", "OUTLET", "NOGRAD"};
  static const int type_v[] = {PICGPU_BCTYPE_WALL, PICGPU_BCTYPE_SYMMETRY, PICGPU_BCTYPE_INLET,
                               PICGPU_BCTYPE_OUTLET, PICGPU_BCTYPE_NOGRAD};
  static const char *const handlers[] = {"noslip", "constant_velocity", "conservation", "allcopy", "parabolic"};
  static const int handler_v[] = {PICGPU_BC_WALL_NOSLIP, PICGPU_BC_INLET_CONSTANT_VELOCITY,
                                  PICGPU_BC_OUTLET_CONSERVATION, PICGPU_BC_NOGRAD_COPY_GHOST,
                                  PICGPU_BC_INLET_PARABOLIC};
  picgpu_boundary_config out;
  picgpu_default_boundary_config(&out);
  FILE *fp = fopen(path, "r");
  if (!fp) {
    set_error("Could not open BCs file '%s'.", path);
    return PICGPU_E_ARG;
  }
  char line[1024];
  int rc = PICGPU_OK;
  while (rc == PICGPU_OK && fgets(line, sizeof(line), fp)) {
    char *s = line;
    while (*s == ' ' || *s == '\t' || *s == '\n' || *s == '\r' || *s == '\v' || *s == '\f') ++s;
    if (*s == '#' || *s == '\0') continue;
    char *save = nullptr;
    const char *fs = strtok_r(s, " \t\n\r", &save);
    const char *ts = strtok_r(nullptr, " \t\n\r", &save);
    const char *hs = strtok_r(nullptr, " \t\n\r", &save);
    if (!fs || !ts || !hs) continue;
    int f, t, h;
    if (!lookup(fs, faces, face_v, 6, &f)) {
      set_error("Unknown face specifier: %s", fs);
      rc = PICGPU_E_ARG;
    } else if (!lookup(ts, types, type_v, 5, &t)) {
      set_error("Unknown BC Type string: %s", ts);
      rc = PICGPU_E_ARG;
    } else if (!lookup(hs, handlers, handler_v, 5, &h)) {
      set_error("Unknown BC Handler string: %s", hs);
      rc = PICGPU_E_ARG;
    } else if (!handler_valid_for_type(t, h)) {
      set_error("Validation failed: Handler '%s' is not valid for Type '%s' on Face '%s'.", hs, ts, fs);
      rc = PICGPU_E_ARG;
    } else {
      out.type[f] = t;
      out.handler[f] = h;
      out.velocity[f][0] = out.velocity[f][1] = out.velocity[f][2] = 0.0;
      out.vmax[f] = 1.0;
      char *tok;
      while ((tok = strtok_r(nullptr, " \t\n\r", &save)) != nullptr) {
        char *eq = strchr(tok, '=');
        if (!eq) continue;
        *eq = '\0';
        const double v = strtod(eq + 1, nullptr);
        if (strcasecmp(tok, "vx") == 0) out.velocity[f][0] = v;
        else if (strcasecmp(tok, "vy") == 0) out.velocity[f][1] = v;
        else if (strcasecmp(tok, "vz") == 0) out.velocity[f][2] = v;
        else if (strcasecmp(tok, "vmax") == 0) out.vmax[f] = v;
      }
    }
  }
  fclose(fp);
  if (rc == PICGPU_OK) *bc = out;
  return rc;
}

int picgpu_compute_metrics(picgpu_ctx *c, int *orientation) {
  if (!c) return PICGPU_E_ARG;
  if (!c->have_grid) {
    set_error("no grid uploaded");
    return PICGPU_E_STATE;
  }
  PG_HIP(hipSetDevice(c->device));
  const NodeDims d = node_dims(c);
  const int64_t nn = d.count();
  auto &e = c->eul;
  e.have_metrics = e.have_bc = false;
  const size_t v3 = sizeof(double) * 3 * (size_t)nn;
  PG_CHECK(e.csi.ensure(v3));
  PG_CHECK(e.eta.ensure(v3));
  PG_CHECK(e.zet.ensure(v3));
  PG_CHECK(e.ucont.ensure(v3));
  PG_CHECK(e.ucat.ensure(v3));
  PG_CHECK(e.aj.ensure(sizeof(double) * (size_t)nn));
  PG_CHECK(e.flag.ensure(2 * sizeof(unsigned long long)));
  const int64_t nb = std::min<int64_t>(1024, eblocks(nn));
  PG_CHECK(e.part.ensure(sizeof(double) * 2 * (size_t)nb));
  PG_CHECK(c->ucat.ensure(c->rsize * 3 * (size_t)c->ncell));
  PG_CHECK(c->corner.ensure(c->rsize * 3 * (size_t)c->nnode));

  PG_LAUNCH(c, "face_metrics", eblocks(nn), kEBlock, k_face_metrics, d, c->coor.as<double>(), e.csi.as<double>(),
            e.eta.as<double>(), e.zet.as<double>());
  PG_CHECK(flag_reset(c));
  PG_LAUNCH(c, "jacobian_inverse", eblocks(nn), kEBlock, k_jacobian_inverse, d, c->coor.as<double>(),
            e.aj.as<double>(), e.flag.as<unsigned long long>());
  unsigned long long first = 0;
  PG_CHECK(flag_read(c, &first));
  if (first != ~0ull) {
    int i, j, k;
    node_ijk(c, first, &i, &j, &k);
    set_error("Jacobian is near zero at node (%d,%d,%d)", i, j, k);
    return PICGPU_E_SINGULAR;
  }
  // CheckAndFixGridOrientation (Metric.c:195-284)
  auto minmax = [&](double *lo, double *hi) -> int {
    PG_LAUNCH(c, "aj_minmax", nb, kEBlock, k_minmax, nn, e.aj.as<double>(), e.part.as<double>());
    std::vector<double> h(2 * (size_t)nb);
    PG_HIP(hipMemcpyAsync(h.data(), e.part.ptr, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
    PG_HIP(hipStreamSynchronize(c->stream));
    *lo = h[0];
    *hi = h[1];
    for (int64_t b = 1; b < nb; ++b) {
      *lo = std::min(*lo, h[2 * b]);
      *hi = std::max(*hi, h[2 * b + 1]);
    }
    return PICGPU_OK;
  };
  double lo, hi;
  PG_CHECK(minmax(&lo, &hi));
  if (lo < 0.0 && hi > 0.0) {
    set_error("Mixed Jacobian signs detected - grid is topologically inconsistent.");
    return PICGPU_E_GRID;
  }
  int o = 1;
  e.flipped = false;
  if (hi < 0.0) {
    o = -1;
    e.flipped = true;
    PG_LAUNCH(c, "metric_flip", eblocks(3 * nn), kEBlock, k_flip, nn, e.csi.as<double>(), e.eta.as<double>(),
              e.zet.as<double>(), e.aj.as<double>());
    PG_CHECK(minmax(&lo, &hi));
    if (lo <= 0.0) {
      set_error("Failed to flip grid orientation - Aj still non-positive.");
      return PICGPU_E_GRID;
    }
    if (lo != 0.0 && hi > 0.0) o = 1;  // Metric.c:262-263
  }
  e.orientation = o;
  // Ucat, Ucont and the particle loop's copy of Ucat start at zero (setup.c:250-268)
  PG_LAUNCH(c, "zero", eblocks(3 * nn), kEBlock, k_zero<double>, 3 * nn, e.ucont.as<double>());
  PG_LAUNCH(c, "zero", eblocks(3 * nn), kEBlock, k_zero<double>, 3 * nn, e.ucat.as<double>());
  if (c->opt.precision == PICGPU_FP32) {
    PG_LAUNCH(c, "zero", eblocks(3 * c->ncell), kEBlock, k_zero<float>, 3 * c->ncell, c->ucat.as<float>());
  } else {
    PG_LAUNCH(c, "zero", eblocks(3 * c->ncell), kEBlock, k_zero<double>, 3 * c->ncell, c->ucat.as<double>());
  }
  PG_HIP(hipStreamSynchronize(c->stream));
  c->have_ucat = true;
  e.have_metrics = true;
  if (orientation) *orientation = o;
  return PICGPU_OK;
}

int picgpu_boundary_create(picgpu_ctx *c, const picgpu_boundary_config *bc) {
  if (!c || !bc) return PICGPU_E_ARG;
  PG_CHECK(need_metrics(c));
  for (int f = 0; f < 6; ++f) {
    if (!handler_valid_for_type(bc->type[f], bc->handler[f])) {
      set_error("face %d: handler %d is not valid for type %d", f, bc->handler[f], bc->type[f]);
      return PICGPU_E_ARG;
    }
    if (!handler_built(bc->handler[f])) {
      set_error("Boundary handler type %d not recognized in factory.", bc->handler[f]);
      return PICGPU_E_UNSUPPORTED;
    }
  }
  PG_HIP(hipSetDevice(c->device));
  c->eul.bc = *bc;
  c->eul.have_bc = true;
  PG_CHECK(run_boundaries(c, true));
  PG_HIP(hipStreamSynchronize(c->stream));
  return PICGPU_OK;
}

int picgpu_upload_ucont(picgpu_ctx *c, const double *ucont) {
  if (!c || !ucont) return PICGPU_E_ARG;
  PG_CHECK(need_metrics(c));
  PG_HIP(hipSetDevice(c->device));
  PG_HIP(hipMemcpy(c->eul.ucont.ptr, ucont, sizeof(double) * 3 * (size_t)c->nnode, hipMemcpyHostToDevice));
  return PICGPU_OK;
}

int picgpu_upload_ucont_device(picgpu_ctx *c, const double *d_ucont) {
  if (!c || !d_ucont) return PICGPU_E_ARG;
  PG_CHECK(need_metrics(c));
  PG_HIP(hipSetDevice(c->device));
  PG_HIP(hipMemcpyAsync(c->eul.ucont.ptr, d_ucont, sizeof(double) * 3 * (size_t)c->nnode, hipMemcpyDeviceToDevice,
                        c->stream));
  PG_HIP(hipStreamSynchronize(c->stream));
  return PICGPU_OK;
}

int picgpu_set_initial_ucont(picgpu_ctx *c, int finit, const double uin[3]) {
  if (!c) return PICGPU_E_ARG;
  PG_CHECK(need_metrics(c));
  PG_HIP(hipSetDevice(c->device));
  PG_CHECK(run_initial_ucont(c, finit, uin));
  PG_HIP(hipStreamSynchronize(c->stream));
  return PICGPU_OK;
}

int picgpu_boundary_execute(picgpu_ctx *c) {
  if (!c) return PICGPU_E_ARG;
  PG_CHECK(need_bc(c));
  PG_HIP(hipSetDevice(c->device));
  PG_CHECK(run_boundaries(c, false));
  PG_HIP(hipStreamSynchronize(c->stream));
  return PICGPU_OK;
}

int picgpu_contra2cart(picgpu_ctx *c) {
  if (!c) return PICGPU_E_ARG;
  PG_CHECK(need_metrics(c));
  PG_HIP(hipSetDevice(c->device));
  return run_contra2cart(c);
}

int picgpu_set_eulerian_fields(picgpu_ctx *c, int64_t step, int64_t start_step, int finit, const double uin[3]) {
  if (!c) return PICGPU_E_ARG;
  PG_CHECK(need_bc(c));
  PG_HIP(hipSetDevice(c->device));
  if (step == start_step && start_step > 0) return PICGPU_OK;  // restart: fields were read (simulation.c:53-58)
  if (step == 0) PG_CHECK(run_initial_ucont(c, finit, uin));
  PG_CHECK(run_boundaries(c, false));
  PG_CHECK(run_contra2cart(c));
  PG_CHECK(run_boundaries(c, false));
  PG_HIP(hipStreamSynchronize(c->stream));
  return PICGPU_OK;
}

}  // extern "C"
