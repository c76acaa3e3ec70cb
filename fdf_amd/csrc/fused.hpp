// fused.hpp -- one-pass particle step for gfx950, tile-staged.
//
// k_step_tiled does, per particle and in one pass over the swarm, the stages
// the reference runs as separate sweeps (simulation.c:248-292): status reset,
// advection (forward Euler, ParticleMotion.c:19-29, or the Heun predictor /
// corrector), the walking search from the previous cell after each position
// update (walkingsearch.c:1094-1242), removal of LOST and out-of-box
// particles (ParticleMotion.c:532-625, 283-383) and trilinear interpolation
// of the corner velocity (interpolation.c:957-1042).
//
// Cells are grouped in 8x8x4 tiles; particles are kept in tile-major cell
// order (a counting sort by the tile-major cell key after every step).  One
// workgroup processes the particles of one tile and first stages the tile's
// corner velocities, with a one-cell halo, in LDS; interpolations that stay
// within the halo read LDS only.  Walk evaluations read each cell's
// precomputed face-plane record from global memory.  Each particle ends by
// taking a slot in the bucket of its new key, from which the next step's
// index is built.  Removed particles stay in the arrays, marked ST_REMOVED,
// until the swarm is compacted.
#pragma once
#include "kernels.hpp"

namespace picgpu {

constexpr int ST_REMOVED = 5;  // internal: removed, awaiting compaction

// tile geometry
constexpr int TX = 8, TY = 8, TZ = 4, TS = TX * TY * TZ;  // cells per tile
constexpr int BX = TX + 3, BY = TY + 3, BZ = TZ + 3, BN = BX * BY * BZ;  // staged nodes (one-cell halo)

struct TileMap {
  int NTI, NTJ, NTK;  // tiles per axis
};

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs on
// XCD b % 8), each with its own L2.  Tile of workgroup b: XCD x takes the x-th
// contiguous eighth of the tiles, in order, so that the particles a
// deposition workgroup reads from its lower neighbour tiles were read into
// the same L2 by the workgroups just before it.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_tile(int b, int nt) {
  const int q = nt / kXcds, r = nt % kXcds;
  const int x = b % kXcds, k = b / kXcds;
  return x * q + (x < r ? x : r) + k;
}

// Tile at position m of the brick order: bricks of kBrickI x kBrickJ x
// kBrickK tiles in raster order (i fastest), the tiles of a brick in raster
// order, partial bricks at the high ends of each axis.  Positions that are
// close hold tiles that are close in all three directions, so the halo
// nodes that neighbouring tiles stage in LDS are read into the L2 once.
constexpr int kBrickI = 8, kBrickJ = 4, kBrickK = 4;
__host__ __device__ __forceinline__ int brick_tile(int m, const TileMap &t) {
  const int slab = kBrickK * t.NTJ * t.NTI;  // tiles of a full layer of bricks
  const int bk = m / slab;
  int r = m - bk * slab;
  const int ek = t.NTK - bk * kBrickK < kBrickK ? t.NTK - bk * kBrickK : kBrickK;
  const int row = kBrickJ * t.NTI * ek;  // tiles of a full row of bricks in that layer
  const int bj = r / row;
  r -= bj * row;
  const int ej = t.NTJ - bj * kBrickJ < kBrickJ ? t.NTJ - bj * kBrickJ : kBrickJ;
  const int full = kBrickI * ej * ek;  // tiles of a full brick in that row
  const int bi = r / full;
  r -= bi * full;
  const int ei = t.NTI - bi * kBrickI < kBrickI ? t.NTI - bi * kBrickI : kBrickI;
  const int li = r % ei, lj = (r / ei) % ej, lk = r / (ei * ej);
  return ((bk * kBrickK + lk) * t.NTJ + bj * kBrickJ + lj) * t.NTI + bi * kBrickI + li;
}
// Tile of workgroup b of a tiled step launch (one workgroup per tile): each
// XCD takes a contiguous share of the brick order.
__device__ __forceinline__ int step_tile(int b, const TileMap &t) {
  return brick_tile(xcd_tile(b, t.NTI * t.NTJ * t.NTK), t);
}

__device__ __forceinline__ int32_t cell_key(const TileMap &t, int i, int j, int k) {
  const int ti = i / TX, tj = j / TY, tk = k / TZ;
  const int li = i - ti * TX, lj = j - tj * TY, lk = k - tk * TZ;
  return (((tk * t.NTJ + tj) * t.NTI + ti) * TS) + ((lk * TY + lj) * TX + li);
}

// Face deposition output of a tiled kernel.  A workgroup writes the faces
// whose two cells are both in its tile, and the faces at the domain's
// boundary; for a face between two tiles it writes its half (the lower
// cell's sum to hi, the upper cell's to lo), and k_face_seams adds the
// halves.  Seam faces of direction d lie on the planes m * T_d, m = 1 ..
// NT_d - 1 (coordinate < the cell count); seam entry (m, a, b) holds 4 reals:
//   i: ((m-1) KM + k) JM + j,  j: ((m-1) KM + k) IM + i,  k: ((m-1) JM + j) IM + i.
template <typename R>
struct FaceOut {
  R *face[3];     // i-faces [KM][JM][IM+1], j [KM][JM+1][IM], k [KM+1][JM][IM], 4 reals each
  R *lo[3], *hi[3];
};

// A face's four reals (count, momentum) as one vector store / load.
template <typename R>
struct Vec4;
template <>
struct Vec4<float> {
  using type = float __attribute__((ext_vector_type(4)));
};
template <>
struct Vec4<double> {
  using type = double __attribute__((ext_vector_type(4)));
};
template <typename R>
__device__ __forceinline__ void store4(R *dst, R a, R b, R c, R d) {
  using V = typename Vec4<R>::type;
  V v;
  v.x = a;
  v.y = b;
  v.z = c;
  v.w = d;
  *reinterpret_cast<V *>(dst) = v;
}
template <typename R>
__device__ __forceinline__ typename Vec4<R>::type load4(const R *src) {
  return *reinterpret_cast<const typename Vec4<R>::type *>(src);
}

__device__ __forceinline__ int64_t seam_entry(int d, int m, int i, int j, int k, int IM, int JM, int KM) {
  if (d == 0) return ((int64_t)(m - 1) * KM + k) * JM + j;
  if (d == 1) return ((int64_t)(m - 1) * KM + k) * IM + i;
  return ((int64_t)(m - 1) * JM + j) * IM + i;
}
__device__ __forceinline__ int64_t face_index(int d, int i, int j, int k, int IM, int JM) {
  if (d == 0) return ((int64_t)k * JM + j) * (IM + 1) + i;
  if (d == 1) return ((int64_t)k * (JM + 1) + j) * IM + i;
  return ((int64_t)k * JM + j) * IM + i;
}

template <typename R>
struct StepArgs {
  GridView g;
  double *pos;
  R *vel;
  R *w;
  int32_t *cell;
  int32_t *status;
  const R *corner;
  const void *planes;  // CellPlanes64 (FP64) or CellPlanes32 (FP32) per cell
  double dt;
  int max_traversal;
  double thresh_scale;
  R clamp;
  double bx0, by0, bz0, bx1, by1, bz1;
  TileMap tm;
  int32_t *count_next;  // [nkeys] histogram of the new keys (zeroed by the caller)
  int32_t *slot;        // [n] slot in the new key bucket, -1 if removed
  unsigned long long *counters;
  int32_t *defer_list;   // [n] particles left to k_step_deferred
  int32_t *defer_count;  // entries of defer_list (zeroed by the caller)
  unsigned long long *walk_hist;  // null, or [kCounterStripes][kWalkBins] walk-iteration histogram
  int all_located;  // 1 after the first step of a call: every particle of the index is
                    // ACTIVE_AND_LOCATED, statuses are neither read nor written
  // k_step_tiled<..., WU = false>: the start velocity u0 is read from
  // vel (u0_recompute 0: the first step of a call) or recomputed from the
  // particle's position and cell (1: every later step; vel and w are then not
  // kept between the steps of a call); with deposit 1 the workgroup also
  // deposits its particles' start state onto the faces, which is the
  // previous step's face deposition (FaceOut)
  int u0_recompute;
  int deposit;
  FaceOut<R> fo;
};

constexpr int kWalkBins = 64;
__device__ __forceinline__ int walk_bin(int n) { return n < kWalkBins - 1 ? n : kWalkBins - 1; }

// Corner velocities of cell (i,j,k) from global memory (order of
// interpolation.c:990-999).
template <typename R>
__device__ __forceinline__ void load_corners(const GridView &g, const R *__restrict__ corner, int i, int j, int k,
                                             R C[8][3]) {
  const int64_t sy = g.IM + 1, sz = (int64_t)(g.JM + 1) * (g.IM + 1);
  const int64_t b = (int64_t)k * sz + (int64_t)j * sy + i;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int64_t q = b + (c & 1) + ((c >> 1) & 1) * sy + ((c >> 2) & 1) * sz;
    C[c][0] = corner[3 * q];
    C[c][1] = corner[3 * q + 1];
    C[c][2] = corner[3 * q + 2];
  }
}

// pointers into LDS (address space 3), so that staged reads compile to ds_read
template <typename T>
using lds_ptr = __attribute__((address_space(3))) const T *;

// Precomputed face planes of the grid (one record per cell); corner
// velocities of a staged tile box from LDS (TILE), otherwise from global
// memory.  With DEFER, a cell evaluation that the FP32 filter does not certify
// is not recomputed in FP64 here: the walk reports EV_DEFER and the particle
// is stepped again by a kernel that carries the FP64 routine, which keeps
// that routine's registers out of the hot kernel.
template <typename R>
struct PlanesOf;
template <>
struct PlanesOf<double> {
  using type = CellPlanes64;
};
template <>
struct PlanesOf<float> {
  using type = CellPlanes32;
};

enum EvalOutcome : int { EV_OK = 0, EV_DEGENERATE = 1, EV_DEFER = 2 };

template <typename R, bool TILE, bool DEFER = false>
struct PlaneSrc {
  GridView g;
  const R *corner;
  const typename PlanesOf<R>::type *planes;
  lds_ptr<R> sk;   // TILE: [BZ][BY][BX][3]
  int i0, j0, k0;  // TILE: node origin of the staged box (the tile's first cell - 1)
  template <bool EXACT>
  __device__ __forceinline__ int eval(int i, int j, int k, const double x[3], double ts, R d[6],
                                      unsigned long long &fb) const {
    if constexpr (DEFER && !EXACT) {
      const int64_t c = (int64_t)i + (int64_t)g.IM * ((int64_t)j + (int64_t)g.JM * k);
      return planes_fp32_distances(planes[c], x, d) ? EV_DEFER : EV_OK;
    } else {
      return eval_cell_planes<R, EXACT>(planes, g, i, j, k, x, ts, d, fb) ? EV_OK : EV_DEGENERATE;
    }
  }
  // the same, given the cell's linear index lc = i + IM (j + JM k)
  template <bool EXACT>
  __device__ __forceinline__ int eval_lin(int i, int j, int k, int32_t lc, const double x[3], double ts, R d[6],
                                          unsigned long long &fb) const {
    if constexpr (DEFER && !EXACT) {
      return planes_fp32_distances(planes[lc], x, d) ? EV_DEFER : EV_OK;
    } else if constexpr (EXACT) {
      return eval_planes_exact(planes[lc], x, d) ? EV_OK : EV_DEGENERATE;
    } else {
      return eval_cell_planes<R, EXACT>(planes, g, i, j, k, x, ts, d, fb) ? EV_OK : EV_DEGENERATE;
    }
  }
  __device__ __forceinline__ void corners(int i, int j, int k, R C[8][3]) const {
    if constexpr (TILE) {
      if ((unsigned)(i - i0) < (unsigned)(BX - 1) && (unsigned)(j - j0) < (unsigned)(BY - 1) &&
          (unsigned)(k - k0) < (unsigned)(BZ - 1)) {
        const int b = ((k - k0) * BY + (j - j0)) * BX + (i - i0);
#pragma unroll
        for (int c = 0; c < 8; ++c) {
          lds_ptr<R> q = sk + 3 * (b + (c & 1) + ((c >> 1) & 1) * BX + ((c >> 2) & 1) * BX * BY);
          C[c][0] = q[0];
          C[c][1] = q[1];
          C[c][2] = q[2];
        }
        return;
      }
    }
    load_corners<R>(g, corner, i, j, k, C);
  }
};

// Plane records of every cell (one thread per cell).  FP64 records hold the
// reference's own normals, face centres and zero band; FP32 records hold the
// normals and offsets rounded once, relative to vertex 0.
__global__ __launch_bounds__(kBlock) void k_cell_planes64(GridView g, double thresh_scale,
                                                          CellPlanes64 *__restrict__ out) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= (int64_t)g.IM * g.JM * g.KM) return;
  int i, j, k;
  unpack_cell((int32_t)c, cell_idx(g), i, j, k);
  CellPlanes64 r;
  cell_planes_exact(g, i, j, k, thresh_scale, r);
  out[c] = r;
}

__global__ __launch_bounds__(kBlock) void k_cell_planes32(GridView g, double thresh_scale,
                                                          CellPlanes32 *__restrict__ out) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= (int64_t)g.IM * g.JM * g.KM) return;
  int i, j, k;
  unpack_cell((int32_t)c, cell_idx(g), i, j, k);
  CellPlanes64 r;
  cell_planes_exact(g, i, j, k, thresh_scale, r);
  const double *v0 = g.coor + 3 * cell_base_node(g, i, j, k);
  CellPlanes32 o;
  o.v0[0] = v0[0];
  o.v0[1] = v0[1];
  o.v0[2] = v0[2];
  double amax = 0.0;
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    const double D = r.n[f][0] * (r.fc[f][0] - v0[0]) + r.n[f][1] * (r.fc[f][1] - v0[1]) +
                     r.n[f][2] * (r.fc[f][2] - v0[2]);
    o.plane[f] = make_float4((float)r.n[f][0], (float)r.n[f][1], (float)r.n[f][2], (float)D);
    amax = fmax(amax, fabs(D));
  }
  float thr = (float)r.thr;
  if ((double)thr < r.thr) thr = nextafterf(thr, INFINITY);
  o.thr = r.thr >= 0.0 ? thr : INFINITY;
  o.amax = (float)amax * 1.0000005f;
  out[c] = o;
}

// One iteration of the walk of walkingsearch.c:1123-1212 for one lane:
// evaluate cell (ci,cj,ck) for x and either accept it (weights in wv), step
// to the next cell, or give up (outside the grid, MAX_TRAVERSAL reached, or a
// revisited cell: the reference would cycle until MAX_TRAVERSAL).
enum WalkOutcome : int { W_LOCATED = 0, W_LOST = 1, W_DEGENERATE = 2, W_CONTINUE = 3, W_DEFER = 4 };

template <typename R, bool EXACT, class Src>
__device__ __forceinline__ int walk_iter(const Src &src, const double x[3], int &ci, int &cj, int &ck, int &steps,
                                         CycleGuard &cyc, int max_traversal, double thresh_scale, R clamp, R wv[3],
                                         unsigned long long &fallbacks) {
  const GridView &g = src.g;
  if (steps >= max_traversal) return W_LOST;
  steps++;
  if (ci < 0 || ci >= g.IM || cj < 0 || cj >= g.JM || ck < 0 || ck >= g.KM) return W_LOST;
  if (cyc.revisit(pack_cell(ci, cj, ck, g.IM, g.JM))) {
    steps = max_traversal;
    return W_LOST;
  }
  R d[6];
  const int ev = src.template eval<EXACT>(ci, cj, ck, x, thresh_scale, d, fallbacks);
  if (ev == EV_DEGENERATE) return W_DEGENERATE;
  if (ev == EV_DEFER) return W_DEFER;
  if (point_accepted<R>(d)) {
    weights_from_distances<R>(d, clamp, wv);
    return W_LOCATED;
  }
  step_indices<R>(d, ci, cj, ck);
  return W_CONTINUE;
}

// walk_iter that also carries the cell's linear index lc = ci + IM (cj + JM
// ck): a step moves each of ci, cj, ck by at most one (step_indices), so lc
// follows with adds of 1, IM and IM JM instead of being recomputed with
// 64-bit multiplies for the record load and the cycle guard.  lc may leave
// [0, IM JM KM) only together with (ci, cj, ck), which the bounds test
// rejects before lc is used.
template <typename R, bool EXACT, class Src>
__device__ __forceinline__ int walk_iter_lin(const Src &src, const double x[3], int &ci, int &cj, int &ck,
                                             int32_t &lc, int &steps, CycleGuard &cyc, int max_traversal,
                                             double thresh_scale, R clamp, R wv[3], unsigned long long &fallbacks) {
  const GridView &g = src.g;
  if (steps >= max_traversal) return W_LOST;
  steps++;
  if (ci < 0 || ci >= g.IM || cj < 0 || cj >= g.JM || ck < 0 || ck >= g.KM) return W_LOST;
  if (cyc.revisit(lc)) {
    steps = max_traversal;
    return W_LOST;
  }
  R d[6];
  const int ev = src.template eval_lin<EXACT>(ci, cj, ck, lc, x, thresh_scale, d, fallbacks);
  if (ev == EV_DEGENERATE) return W_DEGENERATE;
  if (ev == EV_DEFER) return W_DEFER;
  if (point_accepted<R>(d)) {
    weights_from_distances<R>(d, clamp, wv);
    return W_LOCATED;
  }
  const int32_t sj = g.IM, sk = g.IM * g.JM;
  if (d[F_FRONT] < (R)0) {
    ck += 1;
    lc += sk;
  } else if (d[F_BACK] < (R)0) {
    ck -= 1;
    lc -= sk;
  }
  if (d[F_LEFT] < (R)0) {
    ci -= 1;
    lc -= 1;
  } else if (d[F_RIGHT] < (R)0) {
    ci += 1;
    lc += 1;
  }
  if (d[F_BOTTOM] < (R)0) {
    cj -= 1;
    lc -= sj;
  } else if (d[F_TOP] < (R)0) {
    cj += 1;
    lc += sj;
  }
  return W_CONTINUE;
}

// The whole walk from cell (ci,cj,ck) (negative -> (0,0,0),
// walkingsearch.c:611-616).  On success (ci,cj,ck) is the owning cell.
template <typename R, bool EXACT, class Src>
__device__ __forceinline__ int walk_one(const Src &src, const double x[3], int &ci, int &cj, int &ck,
                                        int max_traversal, double thresh_scale, R clamp, R wv[3], int &steps,
                                        unsigned long long &fallbacks) {
  if (ci < 0 || cj < 0 || ck < 0) {
    ci = 0;
    cj = 0;
    ck = 0;
  }
  steps = 0;
  CycleGuard cyc;
  for (;;) {
    const int r = walk_iter<R, EXACT>(src, x, ci, cj, ck, steps, cyc, max_traversal, thresh_scale, clamp, wv,
                                      fallbacks);
    if (r != W_CONTINUE) return r;
  }
}

template <typename R, class Src>
__device__ __forceinline__ void interp_at(const Src &src, int ci, int cj, int ck, const R wv[3], R u[3]) {
  R C[8][3];
  src.corners(ci, cj, ck, C);
  trilinear_vals<R>(C, wv[0], wv[1], wv[2], u);
}

struct StepTally {
  unsigned long long located = 0, lost = 0, oob = 0, iters = 0, fb = 0;
  int err = 0;
};

// A particle's state at the start of a step.
template <typename R>
struct PState {
  int32_t st, c0;
  double x0[3];
  R u0[3];
};

// read_vel false: u0 is left for the caller to recompute
template <typename R>
__device__ __forceinline__ void load_state(const StepArgs<R> &a, int64_t p, PState<R> &s, bool read_vel = true) {
  s.st = a.all_located ? ST_LOCATED : a.status[p];
  s.c0 = a.cell[p];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    s.x0[c] = a.pos[3 * p + c];
    s.u0[c] = read_vel ? a.vel[3 * p + c] : (R)0;
  }
}

// One particle's step (shared by the tiled and the generic kernel).
template <typename R, bool EXACT, bool RK2, class Src>
__device__ __forceinline__ void step_particle(const StepArgs<R> &a, const Src &src, int64_t p, const PState<R> &ps,
                                              StepTally &t) {
  const int32_t st = ps.st;
  if (st == ST_REMOVED) {
    a.slot[p] = -1;
    return;
  }
  const double *x0 = ps.x0;
  const R *u0 = ps.u0;
  int ci = -1, cj = -1, ck = -1;
  if (ps.c0 >= 0) unpack_cell(ps.c0, cell_idx(a.g), ci, cj, ck);
  double x1[3];
  R wv[3];  // set by a walk that locates
  int outcome;
  int s1 = 0, s2 = 0;
  if constexpr (RK2) {
    if (st == ST_LOST) {
      // LOST at step start: not advanced, removed
      x1[0] = x0[0];
      x1[1] = x0[1];
      x1[2] = x0[2];
      outcome = W_LOST;
    } else {
      double xs[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) xs[c] = x0[c] + (double)u0[c] * a.dt;
      outcome = walk_one<R, EXACT>(src, xs, ci, cj, ck, a.max_traversal, a.thresh_scale, a.clamp, wv, s1, t.fb);
      if (outcome == W_LOCATED) {
        R us[3];
        interp_at<R>(src, ci, cj, ck, wv, us);
        const double h = 0.5 * a.dt;
#pragma unroll
        for (int c = 0; c < 3; ++c) x1[c] = x0[c] + ((double)u0[c] + (double)us[c]) * h;
        outcome = walk_one<R, EXACT>(src, x1, ci, cj, ck, a.max_traversal, a.thresh_scale, a.clamp, wv, s2, t.fb);
      } else {
        x1[0] = xs[0];
        x1[1] = xs[1];
        x1[2] = xs[2];
      }
    }
  } else {
    // Euler: every particle is advanced, then located from its cell
#pragma unroll
    for (int c = 0; c < 3; ++c) x1[c] = x0[c] + (double)u0[c] * a.dt;
    outcome = walk_one<R, EXACT>(src, x1, ci, cj, ck, a.max_traversal, a.thresh_scale, a.clamp, wv, s1, t.fb);
  }
  t.iters += s1 + s2;
  if (a.walk_hist) atomicAdd(&a.walk_hist[(blockIdx.x & (kCounterStripes - 1)) * kWalkBins + walk_bin(s1 + s2)], 1ull);
  a.pos[3 * p] = x1[0];
  a.pos[3 * p + 1] = x1[1];
  a.pos[3 * p + 2] = x1[2];
  if (outcome == W_DEGENERATE) {
    t.err = 1;
    a.slot[p] = -1;
  } else if (outcome == W_LOST) {
    t.lost++;
    a.cell[p] = -1;
    a.status[p] = ST_REMOVED;
    a.slot[p] = -1;
  } else {
    const bool in = x1[0] >= a.bx0 && x1[0] <= a.bx1 && x1[1] >= a.by0 && x1[1] <= a.by1 && x1[2] >= a.bz0 &&
                    x1[2] <= a.bz1;
    a.cell[p] = pack_cell(ci, cj, ck, a.g.IM, a.g.JM);
    a.w[3 * p] = wv[0];
    a.w[3 * p + 1] = wv[1];
    a.w[3 * p + 2] = wv[2];
    if (in) {
      t.located++;
      R u1[3];
      interp_at<R>(src, ci, cj, ck, wv, u1);
      a.vel[3 * p] = u1[0];
      a.vel[3 * p + 1] = u1[1];
      a.vel[3 * p + 2] = u1[2];
      a.status[p] = ST_LOCATED;
      a.slot[p] = atomicAdd(&a.count_next[cell_key(a.tm, ci, cj, ck)], 1);
    } else {
      t.oob++;
      a.status[p] = ST_REMOVED;
      a.slot[p] = -1;
    }
  }
}

// one set of atomics per wavefront; every lane of the wave must call it.  A
// lane's particle counts fit 32 bits, so the wave sums them in 32 bits (one
// shuffle per step instead of two); walk iterations keep 64.
__device__ __forceinline__ void flush_tally(StepTally &t, unsigned long long *counters) {
  unsigned v_loc = (unsigned)t.located, v_lost = (unsigned)t.lost, v_oob = (unsigned)t.oob, v_fb = (unsigned)t.fb;
  unsigned long long v_it = t.iters;
  const bool any_err = __any(t.err != 0);
  for (int off = 32; off > 0; off >>= 1) {
    v_loc += __shfl_xor(v_loc, off);
    v_lost += __shfl_xor(v_lost, off);
    v_oob += __shfl_xor(v_oob, off);
    v_it += __shfl_xor(v_it, off);
    v_fb += __shfl_xor(v_fb, off);
  }
  if ((threadIdx.x & 63) == 0) {
    counters = counter_stripe(counters);
    if (v_loc) atomicAdd(&counters[C_LOCATED], (unsigned long long)v_loc);
    if (v_lost) atomicAdd(&counters[C_REMOVED_LOST], (unsigned long long)v_lost);
    if (v_oob) atomicAdd(&counters[C_REMOVED_OOB], (unsigned long long)v_oob);
    if (v_it) atomicAdd(&counters[C_WALK_ITERS], v_it);
    if (v_fb) atomicAdd(&counters[C_FALLBACKS], (unsigned long long)v_fb);
    if (any_err) atomicOr(&counters[C_ERROR], 1ull);
  }
}

// Slots taken in a tile's own LDS histogram carry this flag; the tile's base
// in each bucket (kbase) is added when the next index is built.
constexpr int32_t kLocalSlot = 0x40000000;

// Start states (weights and velocity) of the first DepCap<R> particles of a
// tile are kept in LDS for the tile's face sums; those of later particles
// go to the w and vel arrays.  The FP32 step kernel's registers hold it to
// four workgroups per CU, which leaves LDS for 864 entries (a tile holds
// about 512 particles at two per cell); the FP64 figure keeps three
// workgroups' LDS within the CU's 160 KiB.  The face sums' exchange (12
// values per cell) reuses the same LDS once the start states are summed.
template <typename R>
struct DepCap {
  static constexpr int value = sizeof(R) == 4 ? 864 : 608;
};
template <typename R, bool WU>
struct StageLen {
  static constexpr int dep = WU ? 1 : DepCap<R>::value;
  static constexpr int value = WU ? 6 : (6 * dep > 12 * TS ? 6 * dep : 12 * TS);
};
enum OutKind : int8_t {
  OUT_NONE = 0,   // deferred to k_step_deferred: nothing is written
  OUT_SLOT = 1,   // removed before the step: slot -1 only
  OUT_DEGEN = 2,  // degenerate face met: position, slot -1
  OUT_LOST = 3,   // LOST: position, no cell, removed
  OUT_OOB = 4,    // outside the bounding box: position, cell, weights, removed
  OUT_LOC = 5     // located: position, cell, weights, velocity, slot
};

// The staging loads of a tile (corner velocities of the staged box), held
// in registers between issue() and store() so that a kernel can issue its
// first particle loads in between.
template <typename R>
struct TileLoads {
  static constexpr int NU = (BN + kBlock - 1) / kBlock;  // staged nodes per thread
  R cv[NU][3];
  int64_t node[NU];
  template <class Src>
  __device__ __forceinline__ void issue(const StepArgs<R> &a, const Src &src) {
#pragma unroll
    for (int m = 0; m < NU; ++m) {
      const int u = threadIdx.x + m * kBlock;
      const int bx = u % BX, by = (u / BX) % BY, bz = u / (BX * BY);
      const int ni = src.i0 + bx, nj = src.j0 + by, nk = src.k0 + bz;
      node[m] = -1;
      if (u < BN && ni >= 0 && ni <= a.g.IM && nj >= 0 && nj <= a.g.JM && nk >= 0 && nk <= a.g.KM)
        node[m] = ((int64_t)nk * (a.g.JM + 1) + nj) * (a.g.IM + 1) + ni;
    }
#pragma unroll
    for (int m = 0; m < NU; ++m)
      if (node[m] >= 0) {
#pragma unroll
        for (int c = 0; c < 3; ++c) cv[m][c] = a.corner[3 * node[m] + c];
      }
  }
  __device__ __forceinline__ void store(R *sk) const {
#pragma unroll
    for (int m = 0; m < NU; ++m)
      if (node[m] >= 0) {
        const int u = threadIdx.x + m * kBlock;
#pragma unroll
        for (int c = 0; c < 3; ++c) sk[3 * u + c] = cv[m][c];
      }
  }
};

// A particle's result.  WU: weights and velocity are written (otherwise the
// next step recomputes them).  wslot false: the slot is written later.
template <typename R, bool WU>
__device__ __forceinline__ void emit_global(const StepArgs<R> &a, int64_t p, int kind, const double x[3],
                                            int32_t cell, const R w[3], const R u[3], int32_t slot,
                                            bool wslot = true) {
  if (kind == OUT_NONE) return;
  if (kind != OUT_SLOT) {
    a.pos[3 * p] = x[0];
    a.pos[3 * p + 1] = x[1];
    a.pos[3 * p + 2] = x[2];
  }
  if (kind >= OUT_LOST) {
    a.cell[p] = cell;
    if (!a.all_located) a.status[p] = kind == OUT_LOC ? ST_LOCATED : ST_REMOVED;
  }
  if (WU && kind >= OUT_OOB) {
    a.w[3 * p] = w[0];
    a.w[3 * p + 1] = w[1];
    a.w[3 * p + 2] = w[2];
  }
  if (WU && kind == OUT_LOC) {
    a.vel[3 * p] = u[0];
    a.vel[3 * p + 1] = u[1];
    a.vel[3 * p + 2] = u[2];
  }
  if (wslot) a.slot[p] = slot;
}

// Start state of a located particle at x0 in its cell (the record index lc,
// cell (ci, cj, ck)): the weights of x0 (ParticleSwarm.c:1010-1047) and the
// trilinear velocity (interpolation.c:957-1042), with the arithmetic the
// step that located it used, so the same bits.  Where the FP32 filter does
// not certify a face, the step that located the particle ran it through the
// FP64 routine (k_step_deferred or a list kernel), which left the weights
// and velocity in w and vel: they are read from there.
template <typename R, bool EXACT, class Src>
__device__ __forceinline__ void start_state(const StepArgs<R> &a, const Src &src, int64_t p, int32_t lc, int ci,
                                            int cj, int ck, const double x0[3], R w0[3], R u0[3]) {
  R d[6];
  bool certain = true;
  if constexpr (EXACT)
    eval_planes_exact(src.planes[lc], x0, d);
  else
    certain = planes_fp32_distances(src.planes[lc], x0, d) == 0u;
  if (certain) {
    weights_from_distances<R>(d, a.clamp, w0);
    interp_at<R>(src, ci, cj, ck, w0, u0);
  } else {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      w0[c] = a.w[3 * p + c];
      u0[c] = a.vel[3 * p + c];
    }
  }
}

// The tile's face deposition (no reference counterpart; the oracle's
// ref_deposit_faces).  Thread l sums the start states of cell l's particles,
// index positions kstart[t TS + l] .. kstart[t TS + l + 1] in PID order (from
// the LDS stage sd below DepCap, from w and vel beyond), once with weight
// 1 - a (the cell's lower faces) and once with weight a (its upper faces),
// per direction; a face is its lower cell's sum plus its upper cell's.  The
// lower neighbour's sums come through LDS (xch: 12 rows of TS, in the LDS of
// sd, which it overwrites) when it is in the tile; a face between two tiles
// gets its halves in the seam arrays (FaceOut).  Every thread of the
// workgroup calls it, after a barrier that follows the last write of a start
// state.
template <typename R, int CAPD>
__device__ __forceinline__ void tile_face_sums(const StepArgs<R> &a, int t, int32_t s,
                                               const int32_t *__restrict__ idx, const int32_t *__restrict__ kstart,
                                               R (*sd)[6], R (*xch)[TS]) {
  const int l = threadIdx.x;
  const int li = l % TX, lj = (l / TX) % TY, lk = l / (TX * TY);
  const int ti = t % a.tm.NTI, tj = (t / a.tm.NTI) % a.tm.NTJ, tk = t / (a.tm.NTI * a.tm.NTJ);
  const int IM = a.g.IM, JM = a.g.JM, KM = a.g.KM;
  const int i = ti * TX + li, j = tj * TY + lj, k = tk * TZ + lk;
  const bool own = i < IM && j < JM && k < KM;
  FaceAcc<R> hl[3][4], hh[3][4];
  if (own) {
    const int64_t key = (int64_t)t * TS + l;
    const int32_t b0 = kstart[key], b1 = kstart[key + 1];
    for (int32_t q = b0; q < b1; ++q) {
      R w[3], u[3];
      const int32_t lq = q - s;
      if (lq < CAPD) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          w[c] = sd[lq][c];
          u[c] = sd[lq][3 + c];
        }
      } else {
        const int64_t p = idx[q];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          w[c] = a.w[3 * p + c];
          u[c] = a.vel[3 * p + c];
        }
      }
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const R wa = w[d], wb = (R)1 - w[d];
        hh[d][0].add(wa);
        hh[d][1].add(wa * u[0]);
        hh[d][2].add(wa * u[1]);
        hh[d][3].add(wa * u[2]);
        hl[d][0].add(wb);
        hl[d][1].add(wb * u[0]);
        hl[d][2].add(wb * u[1]);
        hl[d][3].add(wb * u[2]);
      }
    }
  }
  const int g3[3] = {i, j, k}, l3[3] = {li, lj, lk}, dims[3] = {IM, JM, KM};
  constexpr int T3[3] = {TX, TY, TZ}, STR[3] = {1, TX, TX * TY};
  __syncthreads();  // every thread's reads of sd are over: xch takes its place
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int v = 0; v < 4; ++v) xch[4 * d + v][l] = hh[d][v].value();
  __syncthreads();
  if (!own) return;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    R *face = a.fo.face[d];
    const int64_t fl = face_index(d, i, j, k, IM, JM);
    if (l3[d] > 0) {  // lower neighbour in the tile
      const int m = l - STR[d];
      store4<R>(face + 4 * fl, xch[4 * d][m] + hl[d][0].value(), xch[4 * d + 1][m] + hl[d][1].value(),
                xch[4 * d + 2][m] + hl[d][2].value(), xch[4 * d + 3][m] + hl[d][3].value());
    } else if (g3[d] == 0) {  // the domain's first face: no lower cell
      store4<R>(face + 4 * fl, (R)0 + hl[d][0].value(), (R)0 + hl[d][1].value(), (R)0 + hl[d][2].value(),
                (R)0 + hl[d][3].value());
    } else {
      const int64_t se = seam_entry(d, g3[d] / T3[d], i, j, k, IM, JM, KM);
      store4<R>(a.fo.lo[d] + 4 * se, hl[d][0].value(), hl[d][1].value(), hl[d][2].value(), hl[d][3].value());
    }
    if (g3[d] + 1 == dims[d]) {  // the domain's last face: no upper cell
      const int64_t fu = face_index(d, i + (d == 0), j + (d == 1), k + (d == 2), IM, JM);
      store4<R>(face + 4 * fu, hh[d][0].value() + (R)0, hh[d][1].value() + (R)0, hh[d][2].value() + (R)0,
                hh[d][3].value() + (R)0);
    } else if (l3[d] == T3[d] - 1) {  // upper neighbour in the next tile
      const int64_t se = seam_entry(d, (g3[d] + 1) / T3[d], i, j, k, IM, JM, KM);
      store4<R>(a.fo.hi[d] + 4 * se, hh[d][0].value(), hh[d][1].value(), hh[d][2].value(), hh[d][3].value());
    }
  }
}

// Particles the fast pass hands to the general loop, per tile (index
// positions in LDS); more go to k_step_deferred.
constexpr int kSlowCap = 256;

// Where the fast pass leaves a particle whose walk left its cell: the walk
// in progress (predictor, or corrector after a predictor that stayed in the
// cell) evaluated the particle's own cell once, certified, and stepped to
// the neighbour (di, dj, dk) -- walk_iter_lin's step rule on the same
// distances.  The general loop continues that walk from the neighbour,
// with one iteration counted and the own cell in the cycle guard, so the
// walk is the one it would have run from the start.
constexpr int kHandCont = 0x80, kHandCorr = 0x40;
template <typename R>
__device__ __forceinline__ int hand_step(const R d[6]) {
  const int dk = d[F_FRONT] < (R)0 ? 1 : (d[F_BACK] < (R)0 ? -1 : 0);
  const int di = d[F_LEFT] < (R)0 ? -1 : (d[F_RIGHT] < (R)0 ? 1 : 0);
  const int dj = d[F_BOTTOM] < (R)0 ? -1 : (d[F_TOP] < (R)0 ? 1 : 0);
  return kHandCont | (di + 1) | ((dj + 1) << 2) | ((dk + 1) << 4);
}

// One workgroup per tile: stage the tile's corner velocities in LDS, then
// step the tile's particles idx[kstart[t*TS] .. kstart[(t+1)*TS]) in two
// passes.
//
// Fast pass (lane l takes index positions l, l + kBlock, ...): a particle
// whose walks each end in the cell they start from -- the predictor and the
// corrector both accept the particle's own cell at the first evaluation,
// every face certified in FP32 mode -- is stepped in straight-line code; one
// plane record serves the start state, the predictor and the corrector.  A
// walk that leaves the cell takes its step to the neighbour there too (one
// step per particle step; in FP32 the neighbour's record replaces the own
// one in the same registers), so that a particle that moves by one cell is
// finished with the others of its wavefront.
// A lane's next particle's state (and the index entry of the one after) is
// requested before the current particle is stepped, and the first ones with
// the staging loads.  Results are written as they come: the lanes of a
// wavefront finish fast-pass particles together, at consecutive index
// positions, so that their lines are written whole (a result written later,
// into a line already written back, costs a 32- or 64-byte write of its
// own).  With WU = false the step ends there: the next step recomputes the
// weights and velocity at the new position.  Every other particle (a walk
// that needs a second step, an uncertified face, a degenerate cell, a
// particle that is not located at the start of the step) is queued for the
// general loop.
//
// General loop (lanes take the queued particles): all walk iterations of a
// lane, across its particles, are one loop (predictor walk, corrector walk,
// next particle), so that the lanes of a wavefront evaluate cells together
// whatever the walk lengths of their particles; the next particle's state is
// requested one particle ahead.  Particles that stay in the tile count into
// an LDS histogram of the tile's TS keys, reserved in the global histogram
// once per key at the end; particles that leave it take a global slot.
//
// WU = false (every step of a call except with the per-step field
// preparation): weights and velocities are not written; each particle's
// start velocity is read from vel (a.u0_recompute 0: the first step of a
// call) or recomputed from its position and cell (1), which gives the bits
// the previous step computed.  With a.deposit the start states of the
// tile's particles -- the state the previous step left -- are also
// deposited onto the faces (tile_face_sums), so the previous step's
// deposition costs no second pass over the particles.
// WU = true: weights and velocities are written and the start velocity is
// read (the field may change between steps); deposition is a kernel of its own.
//
// The FP32 kernel is held to 128 VGPRs (four wavefronts per SIMD, which its
// LDS also allows).  The FP64 one is left to the compiler: held to three
// wavefronts, the occupancy it has anyway, its schedule came out 14 % slower.
template <typename R, bool EXACT, bool RK2, bool WU>
__global__ __launch_bounds__(kBlock, sizeof(R) == 4 ? 4 : 1) void k_step_tiled(
    StepArgs<R> a, const int32_t *__restrict__ idx, const int32_t *__restrict__ kstart,
    int32_t *__restrict__ kbase) {
  static_assert(TS == kBlock, "one thread per key of the tile");
  constexpr int CAPD = WU ? 1 : DepCap<R>::value;
  __shared__ R sk[BN * 3];
  __shared__ int32_t scount[TS];
  // the start states to deposit (sd), then the face-sum exchange (xch)
  __shared__ R sdx[StageLen<R, WU>::value];
  R(*const sd)[6] = reinterpret_cast<R(*)[6]>(sdx);
  R(*const xch)[TS] = reinterpret_cast<R(*)[TS]>(sdx);
  __shared__ int32_t shist[kWalkBins];
  __shared__ int32_t squeue[kSlowCap];
  // The hand-over of each queued particle (0: from its start) and, in FP32,
  // its predictor velocity, for a corrector walk continued.  FP64 continues
  // predictor walks only (RK2; Euler's one walk is a corrector walk): its
  // start states already take most of three workgroups' LDS.
  constexpr bool HAND_CORR = !EXACT || !RK2;
  __shared__ int8_t shand[kSlowCap];
  __shared__ R sus[!EXACT && RK2 ? kSlowCap : 1][3];
  __shared__ int32_t sqn;
  const int t = step_tile(blockIdx.x, a.tm);
  const int32_t s = kstart[(int64_t)t * TS], e = kstart[(int64_t)(t + 1) * TS];
  const bool dep = !WU && a.deposit;
  const bool recompute = !WU && a.u0_recompute;
  if (s == e && !dep) return;  // empty tile (uniform over the block)
  const int ti = t % a.tm.NTI, tj = (t / a.tm.NTI) % a.tm.NTJ, tk = t / (a.tm.NTI * a.tm.NTJ);
  PlaneSrc<R, true, !EXACT> src{a.g, a.corner, (const typename PlanesOf<R>::type *)a.planes, (lds_ptr<R>)sk,
                                ti * TX - 1, tj * TY - 1, tk * TZ - 1};
  const int32_t B = (int32_t)blockDim.x;

  // The first fast-pass particle of this lane and the one after it; their
  // loads are issued with the staging loads, before the first barrier.
  int32_t q = s + (int32_t)threadIdx.x;
  int32_t p = q < e ? idx[q] : -1;
  int32_t pn = q + B < e ? idx[q + B] : -1;
  PState<R> ps;
  if (s < e) {
    // every staging load is issued before the first particle's state, and
    // that before the first LDS store
    TileLoads<R> tl;
    tl.issue(a, src);
    if (p >= 0) load_state(a, p, ps, !recompute);
    tl.store(sk);
  }
  scount[threadIdx.x] = 0;
  if (threadIdx.x < kWalkBins) shist[threadIdx.x] = 0;
  if (threadIdx.x == 0) sqn = 0;
  __syncthreads();

  StepTally tally;
  const double h = 0.5 * a.dt;
  // the start state's deposition entry of index position q
  auto stage_dep = [&](int32_t qq, int64_t pp, const R w0[3], const R u0[3]) {
    const int32_t lq = qq - s;
    if (lq < CAPD) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        sd[lq][c] = w0[c];
        sd[lq][3 + c] = u0[c];
      }
    } else {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        a.w[3 * pp + c] = w0[c];
        a.vel[3 * pp + c] = u0[c];
      }
    }
  };

  // ---- fast pass ----
  int32_t gp = -1, gv = 0;  // a slot taken in the global histogram, stored at the next one (or the end)
  while (p >= 0) {
    // the lane's next particle but one, and the next particle's state
    const int32_t q2 = q + 2 * B;
    const int32_t p2 = q2 < e ? idx[q2] : -1;
    PState<R> psn;
    if (pn >= 0) load_state(a, pn, psn, !recompute);
    bool fast = ps.st == ST_LOCATED && ps.c0 >= 0;
    int hand = 0;               // how the general loop takes the particle over
    R us[3] = {0, 0, 0};        // predictor velocity (RK2)
    if (fast) {
      int32_t lc = ps.c0;  // the cell the walk in progress evaluates
      int ci, cj, ck;
      unpack_cell(lc, cell_idx(a.g), ci, cj, ck);
      // FP32: the record, read once into registers, serves every evaluation
      // in its cell; FP64: its 304 bytes are read from global memory where
      // they are used
      CellPlanes32 rec32;
      if constexpr (!EXACT) rec32 = src.planes[lc];
      auto dist = [&](const double x[3], R d[6]) -> bool {  // certified distances
        if constexpr (EXACT)
          return eval_planes_exact(src.planes[lc], x, d);
        else
          return planes_fp32_distances(rec32, x, d) == 0u;
      };
      if (recompute) {
        R d0[6], w0[3];
        if (dist(ps.x0, d0)) {
          weights_from_distances<R>(d0, a.clamp, w0);
          interp_at<R>(src, ci, cj, ck, w0, ps.u0);
        } else {
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            w0[c] = a.w[3 * p + c];
            ps.u0[c] = a.vel[3 * p + c];
          }
        }
        if (dep) stage_dep(q, p, w0, ps.u0);
      }
      double x1[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) x1[c] = ps.x0[c] + (double)ps.u0[c] * a.dt;
      R d[6];
      bool cert = dist(x1, d);
      fast = cert && point_accepted<R>(d);
      bool corr = !RK2;  // the walk in progress is the corrector (Euler: the only walk)
      // One step of the walk, taken here: the walk in progress evaluated its
      // cell (certified, not accepted); it steps to the neighbour by
      // walk_iter_lin's rule and evaluates it (iteration 2 of that walk), when
      // the neighbour is in the grid and MAX_TRAVERSAL allows a second
      // iteration.  One step per particle step: mv 1 after the predictor's,
      // 2 after the corrector's (Euler's walk), hs the step from the cell left.
      int mv = 0, hs = 0;
      auto step_once = [&]() {
        hs = hand_step<R>(d);
        const int di = (hs & 3) - 1, dj = ((hs >> 2) & 3) - 1, dk = ((hs >> 4) & 3) - 1;
        const int ni = ci + di, nj = cj + dj, nk = ck + dk;
        if (a.max_traversal >= 2 && ni >= 0 && ni < a.g.IM && nj >= 0 && nj < a.g.JM && nk >= 0 && nk < a.g.KM) {
          ci = ni;
          cj = nj;
          ck = nk;
          lc += di + dj * a.g.IM + dk * (a.g.IM * a.g.JM);
          if constexpr (!EXACT) rec32 = src.planes[lc];
          cert = dist(x1, d);
          fast = cert && point_accepted<R>(d);
          mv = corr ? 2 : 1;
        }
      };
      if (cert && !fast) step_once();
      if (RK2 && fast) {
        R ws[3];
        weights_from_distances<R>(d, a.clamp, ws);
        interp_at<R>(src, ci, cj, ck, ws, us);
#pragma unroll
        for (int c = 0; c < 3; ++c) x1[c] = ps.x0[c] + ((double)ps.u0[c] + (double)us[c]) * h;
        corr = true;
        cert = dist(x1, d);  // the corrector walk starts in the predictor's cell
        fast = cert && point_accepted<R>(d);
        if (cert && !fast && mv == 0) step_once();
      }
      if (!fast) {
        if (!cert)
          hand = EXACT ? 0 : -1;  // FP32: an uncertified face, for the FP64 routine of k_step_deferred
        else if (mv == 0)
          hand = HAND_CORR || !corr ? hand_step<R>(d) | (corr ? kHandCorr : 0) : 0;
        else if (mv == 1)
          hand = corr ? 0 : hs;  // predictor continued from the neighbour; a corrector that left it restarts
        else
          hand = HAND_CORR ? hs | kHandCorr : 0;  // corrector continued from the neighbour
      }
      if (fast) {
        R w1[3] = {0, 0, 0}, u1[3] = {0, 0, 0};
        if constexpr (WU) {
          weights_from_distances<R>(d, a.clamp, w1);
          interp_at<R>(src, ci, cj, ck, w1, u1);
        }
        const int it = (RK2 ? 2 : 1) + (mv != 0);
        tally.iters += it;
        if (a.walk_hist) atomicAdd(&shist[walk_bin(it)], 1);
        const bool in = x1[0] >= a.bx0 && x1[0] <= a.bx1 && x1[1] >= a.by0 && x1[1] <= a.by1 && x1[2] >= a.bz0 &&
                        x1[2] <= a.bz1;
        int32_t slot = -1;
        bool slot_later = false;
        if (in) {
          tally.located++;
          const int32_t key = cell_key(a.tm, ci, cj, ck);
          if (key / TS == t) {
            slot = atomicAdd(&scount[key - t * TS], 1) | kLocalSlot;
          } else {
            if (gp >= 0) a.slot[gp] = gv;
            gp = p;
            gv = atomicAdd(&a.count_next[key], 1);
            slot_later = true;
          }
        } else {
          tally.oob++;
        }
        emit_global<R, WU>(a, p, in ? OUT_LOC : OUT_OOB, x1, lc, w1, u1, slot, !slot_later);
      }
    }
    if (!fast) {
      const int r = hand < 0 ? kSlowCap : atomicAdd(&sqn, 1);
      if (r < kSlowCap) {
        squeue[r] = q;
        shand[r] = (int8_t)hand;
        if constexpr (!EXACT && RK2) {
#pragma unroll
          for (int c = 0; c < 3; ++c) sus[r][c] = us[c];
        }
      } else {
        // an uncertified face, or more than the queue holds: stepped by
        // k_step_deferred, which reads the start velocity from vel
        a.defer_list[atomicAdd(a.defer_count, 1)] = p;
        if (recompute) {
#pragma unroll
          for (int c = 0; c < 3; ++c) a.vel[3 * p + c] = ps.u0[c];
        }
      }
    }
    q += B;
    p = pn;
    pn = p2;
    ps = psn;
  }
  if (gp >= 0) a.slot[gp] = gv;
  __syncthreads();

  // ---- general loop over the queued particles ----
  const int32_t nq = sqn < kSlowCap ? sqn : kSlowCap;
  int32_t r = (int32_t)threadIdx.x;  // queue entry of the current particle
  auto qpos = [&](int32_t rr) -> int32_t { return rr < nq ? squeue[rr] : -1; };
  q = qpos(r);
  p = q >= 0 ? idx[q] : -1;
  int32_t qn = qpos(r + B);
  pn = qn >= 0 ? idx[qn] : -1;
  if (p >= 0) load_state(a, p, ps, !recompute);
  PState<R> pnext;  // state of particle pn, requested one particle ahead
  int32_t pnn = -1;
  // per-particle walk state; phase 0: no particle, 1: predictor walk (RK2),
  // 2: final walk, 3: LOST at step start (RK2; removed without a walk)
  int ph = 0;
  double xw[3] = {0, 0, 0};
  int ci = 0, cj = 0, ck = 0, steps = 0, stot = 0;
  int32_t lc = 0;  // ci + IM (cj + JM ck)
  CycleGuard cyc;
  R wv[3] = {0, 0, 0};
  int32_t fp = -1, fv = 0;  // a slot taken in the global histogram, stored at the next one (or the end)

  auto begin = [&]() {
    // request the next particle's state, then set up the walk of p
    if (pn >= 0) load_state(a, pn, pnext, !recompute);
    const int32_t qnn = qpos(r + 2 * B);
    pnn = qnn >= 0 ? idx[qnn] : -1;
    ci = cj = ck = 0;
    if (ps.c0 >= 0) unpack_cell(ps.c0, cell_idx(a.g), ci, cj, ck);
    lc = ps.c0 >= 0 ? ps.c0 : 0;
    steps = stot = 0;
    cyc = CycleGuard{};
    if (recompute) {
      // the start velocity, as the fast pass computed it (every particle of
      // the index is located)
      const int32_t lq = q - s;
      if (dep && lq < CAPD) {
#pragma unroll
        for (int c = 0; c < 3; ++c) ps.u0[c] = sd[lq][3 + c];
      } else {
        R w0[3];
        start_state<R, EXACT>(a, src, p, lc, ci, cj, ck, ps.x0, w0, ps.u0);
      }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) xw[c] = ps.x0[c] + (double)ps.u0[c] * a.dt;
    const int hc = (uint8_t)shand[r];
    if (ps.st == ST_REMOVED) {
      ph = 4;
    } else if (RK2 && ps.st == ST_LOST) {
      ph = 3;
    } else if (hc & kHandCont) {
      // the walk the fast pass began: own cell evaluated, step to the neighbour
      cyc.revisit(lc);
      steps = 1;
      const int di = (hc & 3) - 1, dj = ((hc >> 2) & 3) - 1, dk = ((hc >> 4) & 3) - 1;
      ci += di;
      cj += dj;
      ck += dk;
      lc += di + dj * a.g.IM + dk * (a.g.IM * a.g.JM);
      if (!EXACT && RK2 && (hc & kHandCorr)) {
        ph = 2;
        stot = 1;  // the predictor walk: one iteration, in the own cell
#pragma unroll
        for (int c = 0; c < 3; ++c) xw[c] = ps.x0[c] + ((double)ps.u0[c] + (double)sus[!EXACT && RK2 ? r : 0][c]) * h;
      } else {
        ph = RK2 ? 1 : 2;
      }
    } else {
      ph = RK2 ? 1 : 2;
    }
  };
  if (p >= 0) begin();

  while (ph != 0) {
    int rw;
    R ui[3] = {0, 0, 0};
    if (ph >= 3) {
      rw = W_LOST;
    } else {
      rw = walk_iter_lin<R, EXACT>(src, xw, ci, cj, ck, lc, steps, cyc, a.max_traversal, a.thresh_scale, a.clamp,
                                   wv, tally.fb);
      if (rw == W_CONTINUE) continue;
      if (rw == W_LOCATED) interp_at<R>(src, ci, cj, ck, wv, ui);
      if (RK2 && ph == 1 && rw == W_LOCATED) {
        // corrector position; the final walk starts from the predictor's cell
#pragma unroll
        for (int c = 0; c < 3; ++c) xw[c] = ps.x0[c] + ((double)ps.u0[c] + (double)ui[c]) * h;
        ph = 2;
        stot += steps;
        steps = 0;
        cyc = CycleGuard{};
        continue;
      }
    }
    // particle p is done: outcome rw at position xw (x0 when LOST at start)
    int kind;
    int32_t cell = -1, slot = -1;
    bool slot_later = false;
    if (rw == W_DEFER) {
      // stepped again from its start state by k_step_deferred, which reads
      // the start velocity from vel
      a.defer_list[atomicAdd(a.defer_count, 1)] = p;
      if (recompute) {
#pragma unroll
        for (int c = 0; c < 3; ++c) a.vel[3 * p + c] = ps.u0[c];
      }
      kind = OUT_NONE;
    } else if (ph == 4) {
      kind = OUT_SLOT;
    } else {
      if (ph == 3) {
#pragma unroll
        for (int c = 0; c < 3; ++c) xw[c] = ps.x0[c];
      }
      tally.iters += stot + steps;
      if (a.walk_hist) atomicAdd(&shist[walk_bin(stot + steps)], 1);
      if (rw == W_DEGENERATE) {
        tally.err = 1;
        kind = OUT_DEGEN;
      } else if (rw == W_LOST) {
        tally.lost++;
        kind = OUT_LOST;
      } else {
        const bool in = xw[0] >= a.bx0 && xw[0] <= a.bx1 && xw[1] >= a.by0 && xw[1] <= a.by1 && xw[2] >= a.bz0 &&
                        xw[2] <= a.bz1;
        cell = lc;
        if (in) {
          tally.located++;
          kind = OUT_LOC;
          const int32_t key = cell_key(a.tm, ci, cj, ck);
          if (key / TS == t) {
            slot = atomicAdd(&scount[key - t * TS], 1) | kLocalSlot;
          } else {
            if (fp >= 0) a.slot[fp] = fv;
            fp = p;
            fv = atomicAdd(&a.count_next[key], 1);
            slot_later = true;
          }
        } else {
          tally.oob++;
          kind = OUT_OOB;
        }
      }
    }
    emit_global<R, WU>(a, p, kind, xw, cell, wv, ui, slot, !slot_later);
    // next particle
    r += B;
    q = qn;
    qn = qpos(r + B);
    p = pn;
    pn = pnn;
    ph = 0;
    if (p >= 0) {
      ps = pnext;
      begin();
    }
  }
  if (fp >= 0) a.slot[fp] = fv;
  flush_tally(tally, a.counters);
  __syncthreads();
  const int32_t cnt = scount[threadIdx.x];
  if (cnt > 0) {
    const int64_t key = (int64_t)t * TS + threadIdx.x;
    kbase[key] = atomicAdd(&a.count_next[key], cnt);
  }
  if (a.walk_hist && threadIdx.x < kWalkBins && shist[threadIdx.x] > 0)
    atomicAdd(&a.walk_hist[(blockIdx.x & (kCounterStripes - 1)) * kWalkBins + threadIdx.x],
              (unsigned long long)shist[threadIdx.x]);
  if constexpr (!WU) {
    if (dep) tile_face_sums<R, CAPD>(a, t, s, idx, kstart, sd, xch);
  }
}

// The end of a step call (and the stage after any change of the swarm that
// the steps do not make): every particle's weights and velocity at its
// position and cell, written to w and vel, and, with a.deposit, the face
// deposition of that state.  One workgroup per tile, as k_step_tiled.
template <typename R, bool EXACT>
__global__ __launch_bounds__(kBlock) void k_step_finish(StepArgs<R> a, const int32_t *__restrict__ idx,
                                                        const int32_t *__restrict__ kstart) {
  constexpr int CAPD = DepCap<R>::value;
  __shared__ R sk[BN * 3];
  __shared__ R sdx[StageLen<R, false>::value];
  R(*const sd)[6] = reinterpret_cast<R(*)[6]>(sdx);
  R(*const xch)[TS] = reinterpret_cast<R(*)[TS]>(sdx);
  const int t = step_tile(blockIdx.x, a.tm);
  const int32_t s = kstart[(int64_t)t * TS], e = kstart[(int64_t)(t + 1) * TS];
  if (s == e && !a.deposit) return;
  const int ti = t % a.tm.NTI, tj = (t / a.tm.NTI) % a.tm.NTJ, tk = t / (a.tm.NTI * a.tm.NTJ);
  PlaneSrc<R, true, !EXACT> src{a.g, a.corner, (const typename PlanesOf<R>::type *)a.planes, (lds_ptr<R>)sk,
                                ti * TX - 1, tj * TY - 1, tk * TZ - 1};
  const int32_t B = (int32_t)blockDim.x;
  if (s < e) {
    int32_t q = s + (int32_t)threadIdx.x;
    int32_t p = q < e ? idx[q] : -1;
    int32_t pn = q + B < e ? idx[q + B] : -1;
    int32_t lc = 0;
    double x0[3];
    TileLoads<R> tl;
    tl.issue(a, src);
    if (p >= 0) {
      lc = a.cell[p];
#pragma unroll
      for (int c = 0; c < 3; ++c) x0[c] = a.pos[3 * p + c];
    }
    tl.store(sk);
    __syncthreads();
    while (p >= 0) {
      // the lane's next particle but one, and the next particle's state
      const int32_t q2 = q + 2 * B;
      const int32_t p2 = q2 < e ? idx[q2] : -1;
      int32_t lcn = 0;
      double xn[3];
      if (pn >= 0) {
        lcn = a.cell[pn];
#pragma unroll
        for (int c = 0; c < 3; ++c) xn[c] = a.pos[3 * pn + c];
      }
      int ci, cj, ck;
      unpack_cell(lc, cell_idx(a.g), ci, cj, ck);
      R w0[3], u0[3];
      start_state<R, EXACT>(a, src, p, lc, ci, cj, ck, x0, w0, u0);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        a.w[3 * (int64_t)p + c] = w0[c];
        a.vel[3 * (int64_t)p + c] = u0[c];
      }
      if (a.deposit && q - s < CAPD) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          sd[q - s][c] = w0[c];
          sd[q - s][3 + c] = u0[c];
        }
      }
      q += B;
      p = pn;
      pn = p2;
      lc = lcn;
#pragma unroll
      for (int c = 0; c < 3; ++c) x0[c] = xn[c];
    }
  }
  if (a.deposit) {
    __syncthreads();  // every start state staged (sd) or written (w, vel)
    tile_face_sums<R, CAPD>(a, t, s, idx, kstart, sd, xch);
  }
}

// Faces between two tiles: the lower cell's sum (hi) plus the upper cell's
// (lo).  Entries are numbered i-seams, then j-seams, then k-seams; there are
// fewer than half as many as cells, so 32-bit indices (and divisions) hold.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_face_seams(int IM, int JM, int KM, int32_t ni, int32_t nj, int32_t nk,
                                                       FaceOut<R> fo) {
  int32_t q = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  int d, i, j, k;
  if (q < ni) {
    d = 0;
    j = q % JM;
    const int32_t r = q / JM;
    k = r % KM;
    i = (r / KM + 1) * TX;
  } else if ((q -= ni) < nj) {
    d = 1;
    i = q % IM;
    const int32_t r = q / IM;
    k = r % KM;
    j = (r / KM + 1) * TY;
  } else if ((q -= nj) < nk) {
    d = 2;
    i = q % IM;
    const int32_t r = q / IM;
    j = r % JM;
    k = (r / JM + 1) * TZ;
  } else {
    return;
  }
  const int64_t f = face_index(d, i, j, k, IM, JM);
  R *face = fo.face[d];
  const R *lo = fo.lo[d] + 4 * (int64_t)q, *hi = fo.hi[d] + 4 * (int64_t)q;
  const auto vh = load4<R>(hi), vl = load4<R>(lo);
  store4<R>(face + 4 * f, vh.x + vl.x, vh.y + vl.y, vh.z + vl.z, vh.w + vl.w);
}

// Particles deferred by k_step_tiled, stepped from their start state with the
// FP64 routine for uncertified faces (a grid-stride loop over the list).
template <typename R, bool EXACT, bool RK2>
__global__ __launch_bounds__(kBlock) void k_step_deferred(StepArgs<R> a) {
  const int64_t nd = *a.defer_count;
  StepTally tally;
  const PlaneSrc<R, false> src{a.g, a.corner, (const typename PlanesOf<R>::type *)a.planes, nullptr, 0, 0, 0};
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nd; q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = a.defer_list[q];
    PState<R> ps;
    load_state(a, p, ps);
    step_particle<R, EXACT, RK2>(a, src, p, ps, tally);
  }
  flush_tally(tally, a.counters);
}

// Particles idx[q0 .. q1) from global memory (particles without a cell).
template <typename R, bool EXACT, bool RK2>
__global__ __launch_bounds__(kBlock) void k_step_list(StepArgs<R> a, const int32_t *__restrict__ idx, int64_t q0,
                                                      int64_t q1) {
  const int64_t q = q0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  StepTally tally;
  if (q < q1) {
    const int64_t p = idx[q];
    PState<R> ps;
    load_state(a, p, ps);
    const PlaneSrc<R, false> src{a.g, a.corner, (const typename PlanesOf<R>::type *)a.planes, nullptr, 0, 0, 0};
    step_particle<R, EXACT, RK2>(a, src, p, ps, tally);
  }
  flush_tally(tally, a.counters);
}

// Bucket of every particle for the index: its cell's key, nkeys for a live
// particle without a cell, none for a removed one.
__device__ __forceinline__ int32_t bucket_of(int32_t c, int32_t st, int IM, int JM, const TileMap &tm,
                                             int32_t nkeys) {
  if (c >= 0) {
    int i, j, k;
    unpack_cell(c, IM, JM, i, j, k);
    return cell_key(tm, i, j, k);
  }
  return st == ST_REMOVED ? -1 : nkeys;
}

// guess (may be null): the start cell of a particle without a cell, so that
// the index orders a fresh locate by where its walks start.
__global__ __launch_bounds__(kBlock) void k_key_histogram(int64_t n, const int32_t *__restrict__ cell,
                                                          const int32_t *__restrict__ status,
                                                          const int32_t *__restrict__ guess, int IM, int JM,
                                                          TileMap tm, int32_t nkeys, int32_t *__restrict__ count,
                                                          int32_t *__restrict__ slot) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int32_t c = cell[p];
  if (c < 0 && guess) c = guess[p];
  const int32_t b = bucket_of(c, status[p], IM, JM, tm, nkeys);
  slot[p] = b >= 0 ? atomicAdd(&count[b], 1) : -1;
}

// idx[kstart[bucket] + slot] = p
__global__ __launch_bounds__(kBlock) void k_key_scatter(int64_t n, const int32_t *__restrict__ cell,
                                                        const int32_t *__restrict__ status,
                                                        const int32_t *__restrict__ guess,
                                                        const int32_t *__restrict__ slot, int IM, int JM, TileMap tm,
                                                        int32_t nkeys, const int32_t *__restrict__ kstart,
                                                        int32_t *__restrict__ idx) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int32_t s = slot[p];
  if (s < 0) return;
  int32_t c = cell[p];
  if (c < 0 && guess) c = guess[p];
  const int32_t b = bucket_of(c, status[p], IM, JM, tm, nkeys);
  idx[kstart[b] + s] = (int32_t)p;
}

// count: n, or *n_dev when n_dev is given (a count still on the device)
__device__ __forceinline__ int64_t dev_count(int64_t n, const int32_t *n_dev) {
  return n_dev ? (int64_t)*n_dev : n;
}

// After a step: idx_next[kstart_next[key(cell[p])] + slot[p]] = p for the
// particles p = idx_cur[0..m) that survived it (slot >= 0; a slot in a
// tile's own histogram is offset by the tile's base kbase[key]).  Each thread
// takes kReindexPer entries, kBlock apart, and issues each round of its
// dependent loads for all of them together.
constexpr int kReindexPer = 4;
__global__ __launch_bounds__(kBlock) void k_key_reindex(int64_t m_max, const int32_t *__restrict__ m_dev,
                                                        const int32_t *__restrict__ idx_cur,
                                                        const int32_t *__restrict__ cell,
                                                        const int32_t *__restrict__ slot,
                                                        const int32_t *__restrict__ kbase, CellIdx cx, TileMap tm,
                                                        const int32_t *__restrict__ kstart,
                                                        int32_t *__restrict__ idx_next) {
  constexpr int NP = kReindexPer;
  const int64_t m = dev_count(m_max, m_dev);
  const int64_t q0 = (int64_t)blockIdx.x * (kBlock * NP) + threadIdx.x;
  int32_t p[NP], s[NP], c[NP], key[NP], base[NP];
#pragma unroll
  for (int u = 0; u < NP; ++u) p[u] = q0 + u * kBlock < m ? idx_cur[q0 + u * kBlock] : -1;
#pragma unroll
  for (int u = 0; u < NP; ++u) {
    s[u] = p[u] >= 0 ? slot[p[u]] : -1;
    c[u] = p[u] >= 0 ? cell[p[u]] : 0;
  }
#pragma unroll
  for (int u = 0; u < NP; ++u) {
    key[u] = 0;
    base[u] = 0;
    if (s[u] >= 0) {
      int i, j, k;
      unpack_cell(c[u], cx, i, j, k);
      key[u] = cell_key(tm, i, j, k);
      base[u] = kstart[key[u]];
      if (s[u] & kLocalSlot) base[u] += kbase[key[u]];
    }
  }
#pragma unroll
  for (int u = 0; u < NP; ++u)
    if (s[u] >= 0) idx_next[base[u] + (s[u] & (kLocalSlot - 1))] = p[u];
}

__global__ __launch_bounds__(kBlock) void k_iota(int64_t n, const int32_t *__restrict__ n_dev,
                                                 int32_t *__restrict__ idx) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < dev_count(n, n_dev)) idx[q] = (int32_t)q;
}

// Puts every bucket idx[kstart[key] .. kstart[key+1]) in ascending PID order,
// the summation order of deposition (and of the oracle).  Buckets of up to 8
// particles are ranked in registers; longer ones, rare at a few particles per
// cell, are insertion-sorted in place.  The PIDs are read as T: the 32-bit
// copy pid32 when every PID lies in [0, 2^31) (*wide == 0; the same order,
// half the bytes), the 64-bit pid otherwise.
template <typename T>
__device__ __forceinline__ void bucket_pid_order(int32_t s, int32_t e, int32_t *__restrict__ idx,
                                                 const T *__restrict__ pid) {
  const int32_t L = e - s;
  if (L <= 8) {
    int32_t q[8];
    T v[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) q[t] = t < L ? idx[s + t] : 0;
#pragma unroll
    for (int t = 0; t < 8; ++t) v[t] = t < L ? pid[q[t]] : (sizeof(T) == 4 ? (T)INT32_MAX : (T)INT64_MAX);
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      int r = 0;
#pragma unroll
      for (int b = 0; b < 8; ++b) r += (v[b] < v[a] || (v[b] == v[a] && b < a)) ? 1 : 0;
      if (a < L && r != a) idx[s + r] = q[a];  // entries already in place are not rewritten
    }
    return;
  }
  for (int32_t t = s + 1; t < e; ++t) {
    const int32_t x = idx[t];
    const T px = pid[x];
    int32_t u = t - 1;
    while (u >= s && pid[idx[u]] > px) {
      idx[u + 1] = idx[u];
      --u;
    }
    idx[u + 1] = x;
  }
}
__global__ __launch_bounds__(kBlock) void k_bucket_pid_order(int64_t nkeys, const int32_t *__restrict__ kstart,
                                                             int32_t *__restrict__ idx,
                                                             const int64_t *__restrict__ pid,
                                                             const int32_t *__restrict__ pid32,
                                                             const int32_t *__restrict__ wide) {
  const int64_t key = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (key >= nkeys) return;
  const int32_t s = kstart[key], e = kstart[key + 1];
  if (e - s < 2) return;
  if (*wide == 0)
    bucket_pid_order<int32_t>(s, e, idx, pid32);
  else
    bucket_pid_order<int64_t>(s, e, idx, pid);
}

// Face deposition (no reference counterpart).  A particle in cell (i,j,k)
// with weights (a1,a2,a3) adds (1-a1) to i-face i and a1 to i-face i+1, and
// likewise along j and k; faces sum count and momentum.  The lower face of
// cell c in direction d is the sum over the lower neighbour's particles
// (weight a_d) plus the sum over c's own (weight 1-a_d), each summed from
// zero in PID order (buckets ordered by k_bucket_pid_order).
//
// One workgroup per 8x8x4 tile of the extended range [0..IM]x[0..JM]x[0..KM].
// Pass 1: each thread runs its cell's particles once and leaves, per
// direction, its weight-a sums in LDS; 128 threads do the same, for one
// direction each, for the cells just below the tile's lower faces.  Pass 2:
// each thread sums its own particles with weight 1-a and adds the lower
// neighbour's sums.  A cell's first particles are read once, with
// independent loads, and kept in registers for both passes; the sums are
// plain sums (FaceAcc).  Neighbouring threads
// take neighbouring cells, so that each load instruction of a wavefront reads
// a few contiguous lines.  In FP32 the kernel is held to at least 6
// wavefronts per SIMD (at most 80 VGPRs; it takes 69, so 7 fit, with no
// spills); FP64 needs more registers (4 per SIMD).
template <typename R>
__global__ __launch_bounds__(kBlock, sizeof(R) == 4 ? 6 : 1) void k_deposit_tiles(
    int IM, int JM, int KM, TileMap tm, TileMap xm, const int32_t *__restrict__ kstart,
    const int32_t *__restrict__ idx, const R *__restrict__ w, const R *__restrict__ vel, R *__restrict__ fi,
    R *__restrict__ fj, R *__restrict__ fk) {
  static_assert(TS == kBlock, "one thread per cell of the tile");
  constexpr int NH = TY * TZ + TX * TZ + TX * TY;  // cells below the tile's i-, j- and k-faces
  __shared__ R us[3][4][TS];  // own cells: accumulator states
  __shared__ R hs[4][NH];     // cells below the tile
  const int b = xcd_tile(blockIdx.x, gridDim.x);
  const int bi = b % xm.NTI, bj = (b / xm.NTI) % xm.NTJ, bk = b / (xm.NTI * xm.NTJ);
  const int l = threadIdx.x;
  const int li = l % TX, lj = (l / TX) % TY, lk = l / (TX * TY);
  const int i = bi * TX + li, j = bj * TY + lj, k = bk * TZ + lk;
  const bool own = i < IM && j < JM && k < KM;
  // The first K particles of the own cell (and of the cell below, for the
  // threads that have one) are read once, in three rounds of independent
  // loads (bucket bounds, indices, particle data) shared by both cells, and
  // the own ones kept in registers for both passes; particles beyond the
  // first K are read in batches of K.
  constexpr int K = 3;  // 6 wavefronts per SIMD in FP32 (4 particles: 5; 2: no faster)
  // the cell below the tile's lower face that thread l < NH sums, direction hd
  int hi = -1, hj = -1, hk = -1, hd = 0;
  if (l < TY * TZ) {
    hd = 0, hi = bi * TX - 1, hj = bj * TY + l % TY, hk = bk * TZ + l / TY;
  } else if (l < TY * TZ + TX * TZ) {
    const int m = l - TY * TZ;
    hd = 1, hi = bi * TX + m % TX, hj = bj * TY - 1, hk = bk * TZ + m / TX;
  } else if (l < NH) {
    const int m = l - TY * TZ - TX * TZ;
    hd = 2, hi = bi * TX + m % TX, hj = bj * TY + m / TX, hk = bk * TZ - 1;
  }
  const bool halo = l < NH && hi >= 0 && hj >= 0 && hk >= 0 && hi < IM && hj < JM && hk < KM;
  int32_t s_own = 0, n_own = 0, sh = 0, nh = 0;
  if (own) {
    const int32_t key = cell_key(tm, i, j, k);
    s_own = kstart[key];
    n_own = kstart[key + 1] - s_own;
  }
  if (halo) {
    const int32_t key = cell_key(tm, hi, hj, hk);
    sh = kstart[key];
    nh = kstart[key + 1] - sh;
  }
  int32_t qo[K], qh[K];
#pragma unroll
  for (int u = 0; u < K; ++u) {
    qo[u] = u < n_own ? idx[s_own + u] : 0;
    qh[u] = u < nh ? idx[sh + u] : 0;
  }
  R cw[K][3], cv[K][3], hw[K], hv[K][3];
#pragma unroll
  for (int u = 0; u < K; ++u) {
    if (u < n_own) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        cw[u][c] = w[3 * (int64_t)qo[u] + c];
        cv[u][c] = vel[3 * (int64_t)qo[u] + c];
      }
    }
    if (u < nh) {
      hw[u] = w[3 * (int64_t)qh[u] + hd];
#pragma unroll
      for (int c = 0; c < 3; ++c) hv[u][c] = vel[3 * (int64_t)qh[u] + c];
    }
  }
  // batch of K particles at bucket positions t0 .. t0+K-1 of a bucket [s, s+n)
  auto load_batch = [&](int32_t s, int32_t n, int32_t t0, R bw[K][3], R bu[K][3]) {
    int32_t qq[K];
#pragma unroll
    for (int u = 0; u < K; ++u) qq[u] = t0 + u < n ? idx[s + t0 + u] : 0;
#pragma unroll
    for (int u = 0; u < K; ++u)
      if (t0 + u < n) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          bw[u][c] = w[3 * (int64_t)qq[u] + c];
          bu[u][c] = vel[3 * (int64_t)qq[u] + c];
        }
      }
  };
  // pass 1, the cell below the tile (direction hd)
  if (l < NH) {
    FaceAcc<R> acc[4];
    auto add_h = [&](R a, const R uu[3]) {
      acc[0].add(a);
      acc[1].add(a * uu[0]);
      acc[2].add(a * uu[1]);
      acc[3].add(a * uu[2]);
    };
#pragma unroll
    for (int u = 0; u < K; ++u)
      if (u < nh) add_h(hw[u], hv[u]);
    for (int32_t t0 = K; t0 < nh; t0 += K) {
      R bw[K][3], bu[K][3];
      load_batch(sh, nh, t0, bw, bu);
#pragma unroll
      for (int u = 0; u < K; ++u)
        if (t0 + u < nh) add_h(hd == 0 ? bw[u][0] : hd == 1 ? bw[u][1] : bw[u][2], bu[u]);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) hs[v][l] = acc[v].s;
  }
  // pass 1, own cell: states after the weight-a sums, all three directions
  {
    FaceAcc<R> acc[3][4];
    auto add_a = [&](const R wa[3], const R uu[3]) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const R a = wa[d];
        acc[d][0].add(a);
        acc[d][1].add(a * uu[0]);
        acc[d][2].add(a * uu[1]);
        acc[d][3].add(a * uu[2]);
      }
    };
#pragma unroll
    for (int u = 0; u < K; ++u)
      if (u < n_own) add_a(cw[u], cv[u]);
    for (int32_t t0 = K; t0 < n_own; t0 += K) {
      R bw[K][3], bu[K][3];
      load_batch(s_own, n_own, t0, bw, bu);
#pragma unroll
      for (int u = 0; u < K; ++u)
        if (t0 + u < n_own) add_a(bw[u], bu[u]);
    }
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int v = 0; v < 4; ++v) us[d][v][l] = acc[d][v].s;
  }
  __syncthreads();
  if (i > IM || j > JM || k > KM) return;
  // pass 2: lower faces, the own particles' sums plus the lower neighbour's
  const bool has[3] = {j < JM && k < KM, i < IM && k < KM, i < IM && j < JM};
  FaceAcc<R> acc[3][4];
  R lower[3][4];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const bool inside = d == 0 ? li > 0 : d == 1 ? lj > 0 : lk > 0;
    const int nl = l - (d == 0 ? 1 : d == 1 ? TX : TX * TY);  // lower neighbour in the tile
    const int h = d == 0 ? lj + TY * lk : d == 1 ? TY * TZ + li + TX * lk : TY * TZ + TX * TZ + li + TX * lj;
#pragma unroll
    for (int v = 0; v < 4; ++v) lower[d][v] = inside ? us[d][v][nl] : hs[v][h];
  }
  auto add_1ma = [&](const R wa[3], const R uu[3]) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const R a = (R)1 - wa[d];
      acc[d][0].add(a);
      acc[d][1].add(a * uu[0]);
      acc[d][2].add(a * uu[1]);
      acc[d][3].add(a * uu[2]);
    }
  };
#pragma unroll
  for (int u = 0; u < K; ++u)
    if (u < n_own) add_1ma(cw[u], cv[u]);
  for (int32_t t0 = K; t0 < n_own; t0 += K) {
    R bw[K][3], bu[K][3];
    load_batch(s_own, n_own, t0, bw, bu);
#pragma unroll
    for (int u = 0; u < K; ++u)
      if (t0 + u < n_own) add_1ma(bw[u], bu[u]);
  }
  R *out[3] = {fi + 4 * (((int64_t)k * JM + j) * (IM + 1) + i), fj + 4 * (((int64_t)k * (JM + 1) + j) * IM + i),
               fk + 4 * (((int64_t)k * JM + j) * IM + i)};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (!has[d]) continue;
#pragma unroll
    for (int v = 0; v < 4; ++v) out[d][v] = lower[d][v] + acc[d][v].value();
  }
}

// P average per cell through the keyed index (buckets in PID order)
template <typename R>
__global__ __launch_bounds__(kBlock) void k_scatter_P_keyed(int IM, int JM, int KM, TileMap tm,
                                                            const int32_t *__restrict__ kstart,
                                                            const int32_t *__restrict__ idx, const R *__restrict__ P,
                                                            R *__restrict__ avg) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= (int64_t)IM * JM * KM) return;
  int i, j, k;
  unpack_cell((int32_t)c, IM, JM, i, j, k);
  const int32_t key = cell_key(tm, i, j, k);
  const int32_t s = kstart[key], e = kstart[key + 1];
  Acc<R> acc;
  for (int32_t t = s; t < e; ++t) acc.add(P[idx[t]]);
  const R cnt = (R)(e - s);
  avg[c] = cnt > (R)0.5 ? acc.value() / cnt : (R)0;
}

// per-cell counts from the keyed histogram
__global__ __launch_bounds__(kBlock) void k_count_from_keys(int IM, int JM, int KM, TileMap tm,
                                                            const int32_t *__restrict__ kstart,
                                                            int32_t *__restrict__ out) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= (int64_t)IM * JM * KM) return;
  int i, j, k;
  unpack_cell((int32_t)c, IM, JM, i, j, k);
  const int32_t key = cell_key(tm, i, j, k);
  out[c] = kstart[key + 1] - kstart[key];
}

// The total and the largest of the per-cell counts, with the first node slot
// holding the largest (CalculateParticleCountPerCell's VecSum and VecMax,
// ParticleMotion.c:1021-1024; a cell (i,j,k) sits at node slot (i,j,k) and
// the other slots hold 0).  out[0] += the counts; out[1] = the maximum over
// cells of count << 33 | (2^33 - 1 - slot), so that the largest count wins
// and, among equal counts, the smallest slot (fewer than 2^33 slots, counts
// below 2^31).  Integer atomics: the result does not depend on their order.
__global__ __launch_bounds__(kBlock) void k_count_summary(int IM, int JM, int KM, const int32_t *__restrict__ count,
                                                          unsigned long long *__restrict__ out) {
  constexpr unsigned long long kSlotMask = (1ull << 33) - 1;
  const int64_t nc = (int64_t)IM * JM * KM;
  unsigned long long sum = 0, best = 0;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x) {
    const int32_t v = count[c];
    int i, j, k;
    unpack_cell((int32_t)c, IM, JM, i, j, k);
    const unsigned long long slot = ((unsigned long long)k * (JM + 1) + j) * (IM + 1) + i;
    const unsigned long long key = ((unsigned long long)(uint32_t)v << 33) | (kSlotMask - slot);
    sum += (unsigned long long)(uint32_t)v;
    best = key > best ? key : best;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    sum += __shfl_xor(sum, off);
    const unsigned long long o = __shfl_xor(best, off);
    best = o > best ? o : best;
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&out[0], sum);
    atomicMax(&out[1], best);
  }
}

// The two sums of LOG_INTERPOLATION_ERROR (logging.c:533-580) over the
// components q < 3n of the swarm's positions and velocities: the analytical
// velocity a of SetAnalyticalSolution (AnalyticalSolution.c:932-962) is the
// sine of each position component (finit 1) or (0, 0, 1) (finit 0);
// part[2b] = block b's sum of (a - v)^2, part[2b + 1] its sum of a^2.  Each
// thread strides over a fixed set of components and the block sums in a
// fixed tree, so that a grid of a given size gives the same sums every run.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_interp_error_sums(int64_t n, int finit, const double *__restrict__ pos,
                                                              const R *__restrict__ vel, double *__restrict__ part) {
  __shared__ double se[kBlock], sa[kBlock];
  double e = 0.0, aa = 0.0;
  for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < 3 * n; q += (int64_t)gridDim.x * kBlock) {
    const double a = finit == 1 ? sin(pos[q]) : (q % 3 == 2 ? 1.0 : 0.0);
    const double d = a - (double)vel[q];
    e += d * d;
    aa += a * a;
  }
  se[threadIdx.x] = e;
  sa[threadIdx.x] = aa;
  __syncthreads();
  for (int h = kBlock / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      se[threadIdx.x] += se[threadIdx.x + h];
      sa[threadIdx.x] += sa[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = se[0];
    part[2 * blockIdx.x + 1] = sa[0];
  }
}

// Physical reorder: dst[q] = src[idx[q]] for q < m.  Between the steps of one
// call (GATHER_STEP) statuses are not moved: every particle is
// ACTIVE_AND_LOCATED.  At the end of a call (GATHER_LAST) statuses are set
// to ACTIVE_AND_LOCATED.  Weights and velocities move when MOVE_VW is set:
// between the steps of a call they hold the start states that an FP32 step
// cannot recompute (uncertified faces) and the per-step field path reads;
// the FP64 steps of a call recompute every start state and leave them behind.
// Each thread moves kGatherPer particles, kBlock apart: all their loads are
// issued before the first store.
enum GatherMode : int { GATHER_ALL = 0, GATHER_STEP = 1, GATHER_LAST = 2 };
constexpr int kGatherPer = 4;
// (a template parameter: a run-time flag there made the FP32 gather 40 %
// slower, its loads no longer issued ahead of the stores).
template <typename R, bool MOVE_VW>
__global__ __launch_bounds__(kBlock) void k_gather_swarm(int64_t m_max, const int32_t *__restrict__ m_dev, int mode,
                                                         const int32_t *__restrict__ idx,
                                                         const double *__restrict__ pos, const R *__restrict__ vel,
                                                         const R *__restrict__ w, const int32_t *__restrict__ cell,
                                                         const int32_t *__restrict__ status,
                                                         const int64_t *__restrict__ pid, const R *__restrict__ P,
                                                         double *__restrict__ pos2, R *__restrict__ vel2,
                                                         R *__restrict__ w2, int32_t *__restrict__ cell2,
                                                         int32_t *__restrict__ status2, int64_t *__restrict__ pid2,
                                                         R *__restrict__ P2, const int32_t *__restrict__ pid32,
                                                         int32_t *__restrict__ pid32_2) {
  constexpr int NP = kGatherPer;
  const int64_t m = dev_count(m_max, m_dev);
  const int64_t q0 = (int64_t)blockIdx.x * (kBlock * NP) + threadIdx.x;
  int64_t p[NP];
#pragma unroll
  for (int u = 0; u < NP; ++u) p[u] = q0 + u * kBlock < m ? idx[q0 + u * kBlock] : -1;
  double x[NP][3];
  R v[NP][3], ww[NP][3], pp[NP];
  int32_t ce[NP], st[NP];
  int64_t id[NP];
#pragma unroll
  for (int u = 0; u < NP; ++u) {
    if (p[u] < 0) continue;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      x[u][c] = pos[3 * p[u] + c];
      if constexpr (MOVE_VW) {
        v[u][c] = vel[3 * p[u] + c];
        ww[u][c] = w[3 * p[u] + c];
      }
    }
    ce[u] = cell[p[u]];
    st[u] = mode == GATHER_ALL ? status[p[u]] : ST_LOCATED;
    id[u] = pid[p[u]];
    pp[u] = P[p[u]];
  }
#pragma unroll
  for (int u = 0; u < NP; ++u) {
    if (p[u] < 0) continue;
    const int64_t q = q0 + u * kBlock;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      pos2[3 * q + c] = x[u][c];
      if constexpr (MOVE_VW) {
        vel2[3 * q + c] = v[u][c];
        w2[3 * q + c] = ww[u][c];
      }
    }
    cell2[q] = ce[u];
    if (mode != GATHER_STEP) status2[q] = st[u];
    pid2[q] = id[u];
    P2[q] = pp[u];
    if (pid32) pid32_2[q] = (int32_t)id[u];  // the low 32 bits, as k_pid32 writes them
  }
}

// pid32[q] = the low 32 bits of pid[q]; *wide |= 1 when some PID lies
// outside [0, 2^31), where the 32-bit order is not the PID order.
__global__ __launch_bounds__(kBlock) void k_pid32(int64_t n, const int64_t *__restrict__ pid,
                                                  int32_t *__restrict__ pid32, int32_t *__restrict__ wide) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool out = false;
  if (q < n) {
    const int64_t v = pid[q];
    pid32[q] = (int32_t)v;
    out = v < 0 || v > (int64_t)INT32_MAX;
  }
  if (__ballot(out) != 0 && (threadIdx.x & 63) == 0) atomicOr(wide, 1);
}

}  // namespace picgpu
