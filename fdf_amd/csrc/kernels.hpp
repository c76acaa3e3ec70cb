// kernels.hpp -- gfx950 kernels of the per-step particle loop.
//
// One thread per particle (or node / cell / face).  Block size 256 = four
// 64-wide wavefronts.  All FP64 arithmetic that must match the reference is
// written in the reference's operation order; the build uses
// -ffp-contract=off, and FP32 code asks for fused multiply-adds explicitly.
// More than one source file includes this header: its kernels are templates
// or static.
#pragma once
#include "picgpu_internal.hpp"

namespace picgpu {

constexpr int kBlock = 256;

__device__ __forceinline__ void unpack_cell(int32_t c, int IM, int JM, int &i, int &j, int &k) {
  i = c % IM;
  int r = c / IM;
  j = r % JM;
  k = r / JM;
}

// The same decoding without integer division, given rI = fl(1/IM) and
// rIJ = fl(1/(IM*JM)) in FP64: (c + 0.5) * rIJ lies within (c/(IM*JM)) * 2^-51
// of (c + 0.5)/(IM*JM), which is at least 0.5/(IM*JM) from an integer, so the
// truncation is floor(c / (IM*JM)) for every c < 2^31 (likewise for j).
struct CellIdx {
  int IM, JM;
  double rI, rIJ;
};
__device__ __forceinline__ void unpack_cell(int32_t c, const CellIdx &x, int &i, int &j, int &k) {
  k = (int)(((double)c + 0.5) * x.rIJ);
  const int r = c - k * (x.IM * x.JM);
  j = (int)(((double)r + 0.5) * x.rI);
  i = r - j * x.IM;
}
__device__ __forceinline__ CellIdx cell_idx(const GridView &g) { return CellIdx{g.IM, g.JM, g.rI, g.rIJ}; }
__device__ __forceinline__ int32_t pack_cell(int i, int j, int k, int IM, int JM) {
  return (int32_t)(i + IM * (j + JM * k));
}

// Cycle detection for the walk (Brent).  For a fixed point the walk is a
// deterministic map cell -> next cell (walkingsearch.c:882-959), so a
// revisited cell means the walk is periodic: it will never accept a cell
// and the reference ends it LOST at MAX_TRAVERSAL (:1206-1212).  Warped
// k-faces leave thin gaps in which walks cycle among two or three cells;
// concluding them at the first revisit gives the reference's outcome
// without holding a 64-lane wavefront for up to 1000 iterations.
struct CycleGuard {
  int32_t saved = -1;
  int power = 1, lam = 0;
  __device__ __forceinline__ bool revisit(int32_t c) {
    if (c == saved) return true;
    if (++lam == power) {
      saved = c;
      power <<= 1;
      lam = 0;
    }
    return false;
  }
};

// ---------------------------------------------------------------------------
// ResetAllParticleStatuses (ParticleMotion.c:2149-2176) fused with
// UpdateAllParticlePositions (ParticleMotion.c:19-95): x += v*dt.
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(kBlock) void k_reset_advect(int64_t n, double *__restrict__ pos,
                                                         const R *__restrict__ vel,
                                                         int32_t *__restrict__ status, double dt,
                                                         int do_reset, int do_advect) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  if (do_reset && status[p] == ST_LOCATED) status[p] = ST_NEEDS;
  if (do_advect) {
#pragma unroll
    for (int c = 0; c < 3; ++c) pos[3 * p + c] += (double)vel[3 * p + c] * dt;
  }
}

// ---------------------------------------------------------------------------
// Walking search.  Each wavefront owns `chunk` consecutive particles and keeps
// its 64 lanes busy: a lane whose walk ends takes the next unassigned
// particle of the chunk (ballot + prefix count), so long and short walks are
// regrouped instead of idling until the wave's longest walk ends.
// ---------------------------------------------------------------------------
template <typename R>
struct LocateArgs {
  GridView g;
  const void *planes;  // CellPlanes64 (FP64) or CellPlanes32 (FP32) per cell
  int64_t n;
  int64_t chunk;
  const double *pos;
  int32_t *cell;
  R *w;
  int32_t *status;
  const int32_t *guess;  // may be null
  const int32_t *order;  // may be null: the q-th particle visited is order[q]
  int32_t *steps_out;    // may be null
  unsigned long long *walk_hist;  // null, or [kCounterStripes][64] walk-iteration histogram
  unsigned long long *counters;
  int max_traversal;
  double thresh_scale;
  R clamp;
  int skip_lost;
};

template <typename R, bool EXACT>
__global__ __launch_bounds__(kBlock) void k_locate(LocateArgs<R> a) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t base = wave * a.chunk;
  if (base >= a.n) return;
  const int64_t end = base + a.chunk < a.n ? base + a.chunk : a.n;
  int64_t cursor = base;
  int64_t p = -1;
  int ci = 0, cj = 0, ck = 0, steps = 0;
  CycleGuard cyc;
  double x[3] = {0.0, 0.0, 0.0};
  unsigned long long my_iters = 0, my_max = 0, my_fb = 0, my_loc = 0, my_lost = 0;
  int my_err = 0;
  const unsigned long long lane_mask_lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  while (true) {
    unsigned long long idle = __ballot(p < 0);
    if (idle != 0ull && cursor < end) {
      if (p < 0) {
        int64_t q = cursor + __popcll(idle & lane_mask_lt);
        if (q < end) {
          const int64_t pq = a.order ? (int64_t)a.order[q] : q;
          int32_t st = a.status[pq];
          bool skip = (st == ST_LOCATED) || (a.skip_lost && st == ST_LOST);
          if (!skip) {
            p = pq;
            x[0] = a.pos[3 * pq];
            x[1] = a.pos[3 * pq + 1];
            x[2] = a.pos[3 * pq + 2];
            int32_t c = a.cell[pq];
            if (c < 0 && a.guess) c = a.guess[pq];
            if (c >= 0) {
              unpack_cell(c, cell_idx(a.g), ci, cj, ck);
            } else {
              ci = 0;  // (info.xs, info.ys, info.zs) on one rank, walkingsearch.c:611-616
              cj = 0;
              ck = 0;
            }
            steps = 0;
            cyc = CycleGuard();
          } else if (a.steps_out) {
            a.steps_out[pq] = 0;
          }
        }
      }
      cursor += __popcll(idle);
    }
    if (__ballot(p >= 0) == 0ull) {
      if (cursor >= end) break;
      continue;
    }
    if (p < 0) continue;
    // ---- one iteration of the walk (walkingsearch.c:1123-1203) ----
    steps++;
    bool done = false, located = false;
    if (ci < 0 || ci >= a.g.IM || cj < 0 || cj >= a.g.JM || ck < 0 || ck >= a.g.KM) {
      done = true;  // walked outside the global domain: LOST (:1127-1132)
    } else if (cyc.revisit(pack_cell(ci, cj, ck, a.g.IM, a.g.JM))) {
      done = true;  // periodic walk: LOST at MAX_TRAVERSAL (:1206-1212)
      steps = a.max_traversal;
    } else {
      R d[6];
      const bool degenerate = !eval_cell_planes<R, EXACT>(a.planes, a.g, ci, cj, ck, x, a.thresh_scale, d, my_fb);
      if (degenerate) {
        my_err = 1;
        done = true;
      } else if (point_accepted<R>(d)) {
        R wv[3];
        weights_from_distances<R>(d, a.clamp, wv);
        a.w[3 * p] = wv[0];
        a.w[3 * p + 1] = wv[1];
        a.w[3 * p + 2] = wv[2];
        located = true;
        done = true;
      } else {
        step_indices<R>(d, ci, cj, ck);
        if (steps >= a.max_traversal) done = true;  // MAX_TRAVERSAL (:1206-1212)
      }
    }
    if (done) {
      if (located) {
        a.cell[p] = pack_cell(ci, cj, ck, a.g.IM, a.g.JM);
        a.status[p] = ST_LOCATED;
        my_loc++;
      } else if (!my_err) {
        a.cell[p] = -1;
        a.status[p] = ST_LOST;
        my_lost++;
      }
      if (a.steps_out) a.steps_out[p] = steps;
      if (a.walk_hist)
        atomicAdd(&a.walk_hist[(blockIdx.x & (kCounterStripes - 1)) * 64 + (steps < 63 ? steps : 63)], 1ull);
      my_iters += steps;
      my_max = my_max > (unsigned long long)steps ? my_max : (unsigned long long)steps;
      p = -1;
    }
  }
  // wave-level reduction of the statistics, one atomic per wave
  for (int off = 32; off > 0; off >>= 1) {
    my_iters += __shfl_xor(my_iters, off);
    unsigned long long om = __shfl_xor(my_max, off);
    my_max = my_max > om ? my_max : om;
    my_fb += __shfl_xor(my_fb, off);
    my_loc += __shfl_xor(my_loc, off);
    my_lost += __shfl_xor(my_lost, off);
    my_err |= __shfl_xor(my_err, off);
  }
  if (lane == 0) {
    unsigned long long *ctr = counter_stripe(a.counters);
    atomicAdd(&ctr[C_WALK_ITERS], my_iters);
    atomicMax(&ctr[C_WALK_MAX], my_max);
    atomicAdd(&ctr[C_FALLBACKS], my_fb);
    atomicAdd(&ctr[C_LOCATED], my_loc);
    atomicAdd(&ctr[C_LOST], my_lost);
    if (my_err) atomicOr(&ctr[C_ERROR], 1ull);
  }
}

// ---------------------------------------------------------------------------
// Centre -> corner (interpolation.c:598-680): node (i,j,k) = mean of the
// valid cells (i-1..i, j-1..j, k-1..k), summed dk, dj, di in that nesting.
// ucat is compact [KM][JM][IM][3].
//
// A thread takes the node column (i, j) over kC2CChunk consecutive k (grid:
// x over columns, y over chunks of k).  The four cells of plane k that node k
// sums last are the four that node k+1 sums first, so each node loads four
// cells, not eight; the next plane is requested while a node is summed.
// Neighbouring threads take neighbouring i, so loads and stores are
// contiguous.
// ---------------------------------------------------------------------------
constexpr int kC2CChunk = 32;

// the cells (i-1,j-1), (i,j-1), (i-1,j), (i,j) of plane ck (the (dj, di)
// order of the sum); v[c] false for a cell outside the grid
template <typename R>
__device__ __forceinline__ void c2c_plane(int IM, int JM, int KM, const R *__restrict__ ucat, int i, int j, int ck,
                                          R u[4][3], bool v[4]) {
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int ci = i - 1 + (c & 1), cj = j - 1 + (c >> 1);
    v[c] = ci >= 0 && ci < IM && cj >= 0 && cj < JM && ck >= 0 && ck < KM;
    if (v[c]) {
      const R *p = ucat + 3 * ((int64_t)ci + (int64_t)IM * (cj + (int64_t)JM * ck));
      u[c][0] = p[0];
      u[c][1] = p[1];
      u[c][2] = p[2];
    } else {
      u[c][0] = u[c][1] = u[c][2] = (R)0;
    }
  }
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_center_to_corner(int IM, int JM, int KM, const R *__restrict__ ucat,
                                                             R *__restrict__ corner) {
  const uint32_t col = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t nI = (uint32_t)IM + 1u;
  if (col >= nI * ((uint32_t)JM + 1u)) return;
  const int i = (int)(col % nI), j = (int)(col / nI);
  const int64_t plane = (int64_t)nI * (JM + 1);
  // chunks blockIdx.y, blockIdx.y + gridDim.y, ... (one per block unless the
  // grid's y limit binds)
  for (int64_t kc = (int64_t)blockIdx.y * kC2CChunk; kc < (int64_t)KM + 1; kc += (int64_t)gridDim.y * kC2CChunk) {
  const int k0 = (int)kc;
  const int k1 = k0 + kC2CChunk < KM + 1 ? k0 + kC2CChunk : KM + 1;
  R lo[4][3], hi[4][3];
  bool vlo[4], vhi[4];
  c2c_plane(IM, JM, KM, ucat, i, j, k0 - 1, lo, vlo);
  c2c_plane(IM, JM, KM, ucat, i, j, k0, hi, vhi);
  for (int k = k0; k < k1; ++k) {
    R nx[4][3];
    bool vnx[4];
    if (k + 1 < k1) c2c_plane(IM, JM, KM, ucat, i, j, k + 1, nx, vnx);
    R s0 = 0, s1 = 0, s2 = 0;
    int cnt = 0;
#pragma unroll
    for (int c = 0; c < 4; ++c)
      if (vlo[c]) {
        s0 += lo[c][0];
        s1 += lo[c][1];
        s2 += lo[c][2];
        cnt++;
      }
#pragma unroll
    for (int c = 0; c < 4; ++c)
      if (vhi[c]) {
        s0 += hi[c][0];
        s1 += hi[c][1];
        s2 += hi[c][2];
        cnt++;
      }
    R *o = corner + 3 * ((int64_t)k * plane + col);
    if (cnt > 0) {
      o[0] = s0 / (R)cnt;
      o[1] = s1 / (R)cnt;
      o[2] = s2 / (R)cnt;
    } else {
      o[0] = o[1] = o[2] = (R)0;
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      vlo[c] = vhi[c];
      vhi[c] = vnx[c];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        lo[c][d] = hi[c][d];
        hi[c][d] = nx[c][d];
      }
    }
  }
  }
}

// ---------------------------------------------------------------------------
// Trilinear interpolation (interpolation.c:839-864, 957-1042) per particle
// (interpolation.c:1341-1385): a cell outside the grid gives zero velocity.
// ---------------------------------------------------------------------------
// Trilinear interpolation from the corner values C (interpolation.c:957-1042
// with the weights of :839-864): clamp to [0,1], the eight corner weights in
// the order of :851-858 / 990-999, accumulate corner by corner, divide by the
// weight sum when it exceeds 1e-14.  FP64 mode: the reference operations.
// FP32 mode: fused multiply-adds and one hardware reciprocal.
template <typename R>
__device__ __forceinline__ void trilinear_vals(const R C[8][3], R a1, R a2, R a3, R out[3]) {
  a1 = ((R)0 < ((R)1 < a1 ? (R)1 : a1)) ? ((R)1 < a1 ? (R)1 : a1) : (R)0;
  a2 = ((R)0 < ((R)1 < a2 ? (R)1 : a2)) ? ((R)1 < a2 ? (R)1 : a2) : (R)0;
  a3 = ((R)0 < ((R)1 < a3 ? (R)1 : a3)) ? ((R)1 < a3 ? (R)1 : a3) : (R)0;
  const R oa1 = (R)1 - a1, oa2 = (R)1 - a2, oa3 = (R)1 - a3;
  R wc[8];
  wc[0] = oa1 * oa2 * oa3;
  wc[1] = a1 * oa2 * oa3;
  wc[2] = oa1 * a2 * oa3;
  wc[3] = a1 * a2 * oa3;
  wc[4] = oa1 * oa2 * a3;
  wc[5] = a1 * oa2 * a3;
  wc[6] = oa1 * a2 * a3;
  wc[7] = a1 * a2 * a3;
  R sw = 0, ax = 0, ay = 0, az = 0;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    if constexpr (sizeof(R) == 4) {
      ax = __builtin_fmaf(wc[c], C[c][0], ax);
      ay = __builtin_fmaf(wc[c], C[c][1], ay);
      az = __builtin_fmaf(wc[c], C[c][2], az);
    } else {
      ax += wc[c] * C[c][0];
      ay += wc[c] * C[c][1];
      az += wc[c] * C[c][2];
    }
    sw += wc[c];
  }
  if (sw > (R)1.0e-14) {
    if constexpr (sizeof(R) == 4) {
      const R rs = __builtin_amdgcn_rcpf(sw);
      out[0] = ax * rs;
      out[1] = ay * rs;
      out[2] = az * rs;
    } else {
      out[0] = ax / sw;
      out[1] = ay / sw;
      out[2] = az / sw;
    }
  } else {
    out[0] = out[1] = out[2] = (R)0;
  }
}

template <typename R>
__device__ __forceinline__ void trilinear(int IM, int JM, const R *__restrict__ corner, int i, int j, int k, R a1,
                                          R a2, R a3, R out[3]) {
  const int64_t sy = IM + 1, sz = (int64_t)(JM + 1) * (IM + 1);
  const int64_t b = (int64_t)k * sz + (int64_t)j * sy + i;
  R C[8][3];
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int64_t q = b + (c & 1) + ((c >> 1) & 1) * sy + ((c >> 2) & 1) * sz;
    C[c][0] = corner[3 * q];
    C[c][1] = corner[3 * q + 1];
    C[c][2] = corner[3 * q + 2];
  }
  trilinear_vals<R>(C, a1, a2, a3, out);
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_interp(int64_t n, int IM, int JM, int KM, const R *__restrict__ corner,
                                                   const int32_t *__restrict__ cell, const R *__restrict__ w,
                                                   R *__restrict__ vel) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int32_t c = cell[p];
  R out[3] = {(R)0, (R)0, (R)0};
  if (c >= 0) {
    int i, j, k;
    unpack_cell(c, IM, JM, i, j, k);
    trilinear<R>(IM, JM, corner, i, j, k, w[3 * p], w[3 * p + 1], w[3 * p + 2], out);
  }
  vel[3 * p] = out[0];
  vel[3 * p + 1] = out[1];
  vel[3 * p + 2] = out[2];
}

// ---------------------------------------------------------------------------
// Removal flags (ParticleMotion.c:532-625 LOST; :283-383 bounding box with
// IsParticleInBox :106-110).  keep[p] = 1 for survivors.
// ---------------------------------------------------------------------------
static __global__ __launch_bounds__(kBlock) void k_keep_flags(int64_t n, const double *__restrict__ pos,
                                                       const int32_t *__restrict__ status, int check_lost,
                                                       int check_box, double bx0, double by0, double bz0,
                                                       double bx1, double by1, double bz1,
                                                       int32_t *__restrict__ keep,
                                                       unsigned long long *counters) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int k = 0, rl = 0, ro = 0;
  if (p < n) {
    k = 1;
    if (check_lost && status[p] == ST_LOST) {
      k = 0;
      rl = 1;
    } else if (check_box) {
      const double x = pos[3 * p], y = pos[3 * p + 1], z = pos[3 * p + 2];
      bool in = x >= bx0 && x <= bx1 && y >= by0 && y <= by1 && z >= bz0 && z <= bz1;
      if (!in) {
        k = 0;
        ro = 1;
      }
    }
    keep[p] = k;
  }
  // wave totals
  unsigned long long bl = __ballot(rl != 0), bo = __ballot(ro != 0);
  if ((threadIdx.x & 63) == 0) {
    unsigned long long *ctr = counter_stripe(counters);
    if (bl) atomicAdd(&ctr[C_REMOVED_LOST], (unsigned long long)__popcll(bl));
    if (bo) atomicAdd(&ctr[C_REMOVED_OOB], (unsigned long long)__popcll(bo));
  }
}

// Moves every particle p with keep[p] to slot off[p] of the destination
// arrays (off = exclusive scan of keep), preserving order.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_compact(int64_t n, const int32_t *__restrict__ keep,
                                                    const int32_t *__restrict__ off, int64_t base, const double *pos,
                                                    const R *vel, const R *w, const int32_t *cell,
                                                    const int32_t *status, const int64_t *pid, const R *P,
                                                    double *pos2, R *vel2, R *w2, int32_t *cell2,
                                                    int32_t *status2, int64_t *pid2, R *P2) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n || !keep[p]) return;
  int64_t q = base + off[p];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    pos2[3 * q + c] = pos[3 * p + c];
    vel2[3 * q + c] = vel[3 * p + c];
    w2[3 * q + c] = w[3 * p + c];
  }
  cell2[q] = cell[p];
  status2[q] = status[p];
  pid2[q] = pid[p];
  P2[q] = P[p];
}

// ---------------------------------------------------------------------------
// Counting sort by owning cell: histogram with the slot of each particle,
// then a scatter to cell_start[c] + slot.  Particles with cell < 0 are
// dropped.  The slot order inside a cell is not deterministic; every
// consumer of a cell segment visits it in ascending PID order.
// ---------------------------------------------------------------------------
static __global__ __launch_bounds__(kBlock) void k_cell_histogram(int64_t n, const int32_t *__restrict__ cell,
                                                           int32_t *__restrict__ count, int32_t *__restrict__ rank) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int32_t c = cell[p];
  rank[p] = c >= 0 ? atomicAdd(&count[c], 1) : -1;
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_sort_scatter(int64_t n, const int32_t *__restrict__ rank,
                                                         const int32_t *__restrict__ start, const double *pos,
                                                         const R *vel, const R *w, const int32_t *cell,
                                                         const int32_t *status, const int64_t *pid, const R *P,
                                                         double *pos2, R *vel2, R *w2, int32_t *cell2,
                                                         int32_t *status2, int64_t *pid2, R *P2) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int32_t c = cell[p];
  if (c < 0) return;
  int64_t q = (int64_t)start[c] + rank[p];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    pos2[3 * q + k] = pos[3 * p + k];
    vel2[3 * q + k] = vel[3 * p + k];
    w2[3 * q + k] = w[3 * p + k];
  }
  cell2[q] = c;
  status2[q] = status[p];
  pid2[q] = pid[p];
  P2[q] = P[p];
}

// Visit a cell segment [s, e) in ascending PID order: returns the next index
// after PID `last` (PIDs are unique), or -1.
__device__ __forceinline__ int64_t next_by_pid(const int64_t *__restrict__ pid, int64_t s, int64_t e, int64_t last,
                                               int64_t &out_pid) {
  int64_t best = -1, bp = INT64_MAX;
  for (int64_t q = s; q < e; ++q) {
    int64_t v = pid[q];
    if (v > last && v < bp) {
      bp = v;
      best = q;
    }
  }
  out_pid = bp;
  return best;
}

// Neumaier-compensated accumulator of the P averages in FP32 (a cell's sum of
// equal P values stays exact at any count); plain sum in FP64.
template <typename R>
struct Acc {
  R s = 0, c = 0;
  __device__ __forceinline__ void add(R v) {
    if constexpr (sizeof(R) == 4) {
      // Neumaier: c collects the rounding error of every s + v, computed
      // exactly and without a branch by Knuth's TwoSum (the same value the
      // |s| >= |v| ordered form gives)
      const R t = s + v;
      const R bv = t - s;
      c += (s - (t - bv)) + (v - bv);
      s = t;
    } else {
      s += v;
    }
  }
  __device__ __forceinline__ R value() const { return s + c; }
};

// Accumulator of the face deposition: a plain sum in R, in the fixed order
// (one cell's particles in PID order).  A face sums a few dozen
// terms at most, so in FP32 its error stays within n * 2^-24 of the summed
// magnitudes, far inside the FP32 mode's 1e-5; a compensated sum costs seven
// instructions an addition instead of one.
template <typename R>
struct FaceAcc {
  R s = 0;
  __device__ __forceinline__ void add(R v) { s += v; }
  __device__ __forceinline__ R value() const { return s; }
};

// P average per cell: sum (interpolation.c:1838-1923) / count
// (interpolation.c:1943-2036), summed in PID order.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_scatter_P(int64_t ncell, const int32_t *__restrict__ start,
                                                      const int64_t *__restrict__ pid, const R *__restrict__ P,
                                                      R *__restrict__ avg) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  int64_t s = start[c], e = start[c + 1];
  Acc<R> acc;
  int64_t last = INT64_MIN, lp;
  for (int64_t t = s; t < e; ++t) {
    int64_t q = next_by_pid(pid, s, e, last, lp);
    acc.add(P[q]);
    last = lp;
  }
  R cnt = (R)(e - s);
  avg[c] = cnt > (R)0.5 ? acc.value() / cnt : (R)0;
}

// Face deposition (no reference counterpart).  A particle in cell (i,j,k)
// with weights (a1,a2,a3) adds (1-a1) to i-face i and a1 to i-face i+1, and
// likewise along j and k; the face sums count and momentum.  A face is the
// sum over its lower cell (weight a) plus the sum over its upper cell
// (weight 1-a), each summed from zero in PID order.  dir: 0 = i-faces
// [KM][JM][IM+1], 1 = j, 2 = k.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_deposit_faces(int dir, int IM, int JM, int KM,
                                                          const int32_t *__restrict__ start,
                                                          const int64_t *__restrict__ pid, const R *__restrict__ w,
                                                          const R *__restrict__ vel, R *__restrict__ out) {
  const int fI = IM + (dir == 0), fJ = JM + (dir == 1), fK = KM + (dir == 2);
  const int64_t nf = (int64_t)fI * fJ * fK;
  int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  int i = (int)(f % fI);
  int64_t r = f / fI;
  int j = (int)(r % fJ);
  int k = (int)(r / fJ);
  FaceAcc<R> acc[2][4];
#pragma unroll
  for (int side = 0; side < 2; ++side) {
    int ci = i, cj = j, ck = k;
    if (side == 0) {
      if (dir == 0) ci -= 1;
      if (dir == 1) cj -= 1;
      if (dir == 2) ck -= 1;
    }
    if (ci < 0 || ci >= IM || cj < 0 || cj >= JM || ck < 0 || ck >= KM) continue;
    int64_t c = (int64_t)ci + (int64_t)IM * (cj + (int64_t)JM * ck);
    int64_t s = start[c], e = start[c + 1];
    int64_t last = INT64_MIN, lp;
    for (int64_t t = s; t < e; ++t) {
      int64_t q = next_by_pid(pid, s, e, last, lp);
      last = lp;
      R a = w[3 * q + dir];
      R wt = side == 0 ? a : (R)1 - a;
      acc[side][0].add(wt);
      acc[side][1].add(wt * vel[3 * q]);
      acc[side][2].add(wt * vel[3 * q + 1]);
      acc[side][3].add(wt * vel[3 * q + 2]);
    }
  }
  R *o = out + 4 * f;
#pragma unroll
  for (int v = 0; v < 4; ++v) o[v] = acc[0][v].value() + acc[1][v].value();
}

// Hexahedron volume as six tetrahedra around the diagonal V0-V5.
__device__ __forceinline__ double cell_volume(const GridView &g, int i, int j, int k) {
  double V[8][3];
  load_vertices(g, i, j, k, V);
  const int T[6][2] = {{1, 2}, {1, 6}, {3, 2}, {3, 4}, {7, 6}, {7, 4}};
  double s = 0.0;
  for (int t = 0; t < 6; ++t) {
    const double *a = V[0], *b = V[T[t][0]], *c = V[T[t][1]], *d = V[5];
    double x0 = b[0] - a[0], x1 = b[1] - a[1], x2 = b[2] - a[2];
    double y0 = c[0] - a[0], y1 = c[1] - a[1], y2 = c[2] - a[2];
    double z0 = d[0] - a[0], z1 = d[1] - a[1], z2 = d[2] - a[2];
    double det = x0 * (y1 * z2 - y2 * z1) + x1 * (y2 * z0 - y0 * z2) + x2 * (y0 * z1 - y1 * z0);
    s += fabs(det);
  }
  return s / 6.0;
}

// Face volume fraction = count * V_p / V_dual with V_dual = half the summed
// volumes of the adjacent cells (lower cell first).
template <typename R>
__global__ __launch_bounds__(kBlock) void k_face_volfrac(int dir, GridView g, const R *__restrict__ face,
                                                         double vp, double *__restrict__ out) {
  const int IM = g.IM, JM = g.JM, KM = g.KM;
  const int fI = IM + (dir == 0), fJ = JM + (dir == 1), fK = KM + (dir == 2);
  const int64_t nf = (int64_t)fI * fJ * fK;
  int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  int i = (int)(f % fI);
  int64_t r = f / fI;
  int j = (int)(r % fJ);
  int k = (int)(r / fJ);
  double s = 0.0;
  int li = i - (dir == 0), lj = j - (dir == 1), lk = k - (dir == 2);
  if (li >= 0 && lj >= 0 && lk >= 0) s += cell_volume(g, li, lj, lk);
  if (i < IM && j < JM && k < KM) s += cell_volume(g, i, j, k);
  double vd = 0.5 * s;
  out[f] = ((double)face[4 * f] * vp) / vd;
}

// ---------------------------------------------------------------------------
// Device-side input preparation.
// ---------------------------------------------------------------------------
// per-block min/max of node coordinates: out[6*b .. 6*b+5] = (min xyz, max xyz)
static __global__ __launch_bounds__(kBlock) void k_bbox_partial(int64_t nnode, const double *__restrict__ coor,
                                                         double *__restrict__ out) {
  __shared__ double sh[6][kBlock];
  double lo[3] = {1.7976931348623157e308, 1.7976931348623157e308, 1.7976931348623157e308};
  double hi[3] = {-1.7976931348623157e308, -1.7976931348623157e308, -1.7976931348623157e308};
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nnode; q += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      double x = coor[3 * q + a];
      lo[a] = x < lo[a] ? x : lo[a];
      hi[a] = x > hi[a] ? x : hi[a];
    }
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    sh[a][threadIdx.x] = lo[a];
    sh[3 + a][threadIdx.x] = hi[a];
  }
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        double o = sh[a][threadIdx.x + s];
        sh[a][threadIdx.x] = o < sh[a][threadIdx.x] ? o : sh[a][threadIdx.x];
        double h = sh[3 + a][threadIdx.x + s];
        sh[3 + a][threadIdx.x] = h > sh[3 + a][threadIdx.x] ? h : sh[3 + a][threadIdx.x];
      }
    }
    __syncthreads();
  }
  if (threadIdx.x < 6) out[6 * blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}

// node-slot ucat (double) -> compact [KM][JM][IM][3] in the working precision
template <typename R>
__global__ __launch_bounds__(kBlock) void k_ucat_compact(int IM, int JM, int KM, const double *__restrict__ in,
                                                         R *__restrict__ out) {
  const int64_t ncell = (int64_t)IM * JM * KM;
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  int i = (int)(c % IM);
  int64_t r = c / IM;
  int j = (int)(r % JM);
  int k = (int)(r / JM);
  const double *u = in + 3 * (((int64_t)k * (JM + 1) + j) * (IM + 1) + i);
  out[3 * c] = (R)u[0];
  out[3 * c + 1] = (R)u[1];
  out[3 * c + 2] = (R)u[2];
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_init_particles(int64_t n, const double *__restrict__ pos_in,
                                                           int64_t pid_base, double *__restrict__ pos,
                                                           R *__restrict__ vel, R *__restrict__ w,
                                                           int32_t *__restrict__ cell, int32_t *__restrict__ status,
                                                           int64_t *__restrict__ pid, R *__restrict__ P) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    pos[3 * p + c] = pos_in[3 * p + c];
    vel[3 * p + c] = (R)0;
    w[3 * p + c] = (R)0;
  }
  cell[p] = -1;
  status[p] = ST_UNINIT;
  pid[p] = pid_base + p;
  P[p] = (R)101325.0;
}

static __global__ __launch_bounds__(kBlock) void k_keep_unlocated(int64_t n, const int32_t *__restrict__ cell,
                                                           int32_t *__restrict__ keep) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) keep[p] = cell[p] < 0 ? 1 : 0;
}

// ---------------------------------------------------------------------------
// Heun (RK2) predictor and corrector (no reference counterpart; the
// reference advects with forward Euler, ParticleMotion.c:19-95).
//   predictor: x* = x + u dt            (particles LOST at step start skip)
//   corrector: x1 = x + (u + u*) (dt/2) (particles located at x*)
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(kBlock) void k_rk2_predict(int64_t n, double *__restrict__ pos, const R *__restrict__ vel,
                                                        const int32_t *__restrict__ status, double *__restrict__ pos0,
                                                        R *__restrict__ vel0, double dt) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const bool lost = status[p] == ST_LOST;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double x = pos[3 * p + c];
    R u = vel[3 * p + c];
    pos0[3 * p + c] = x;
    vel0[3 * p + c] = u;
    if (!lost) pos[3 * p + c] = x + (double)u * dt;
  }
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_rk2_correct(int64_t n, double *__restrict__ pos, const R *__restrict__ vel,
                                                        int32_t *__restrict__ status, const double *__restrict__ pos0,
                                                        const R *__restrict__ vel0, double h) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  if (status[p] != ST_LOCATED) return;
#pragma unroll
  for (int c = 0; c < 3; ++c)
    pos[3 * p + c] = pos0[3 * p + c] + ((double)vel0[3 * p + c] + (double)vel[3 * p + c]) * h;
  status[p] = ST_NEEDS;
}

}  // namespace picgpu
