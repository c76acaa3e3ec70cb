// fused.hpp -- one-pass particle step for gfx950.
//
// k_step_fused does, per particle and in one pass over the swarm, the stages
// the reference runs as separate sweeps (simulation.c:248-292): status reset,
// advection (forward Euler, ParticleMotion.c:19-29, or the Heun predictor /
// corrector), the walking search from the previous cell after each position
// update (walkingsearch.c:1094-1242), removal of LOST and out-of-box
// particles (ParticleMotion.c:532-625, 283-383) and trilinear interpolation
// of the corner velocity (interpolation.c:957-1042).  It ends by taking a
// slot in its cell's histogram bucket, from which a counting sort builds the
// cell-ordered index used by deposition.  Removed particles stay in the
// arrays, marked ST_REMOVED, until the swarm is compacted.
#pragma once
#include "kernels.hpp"

namespace picgpu {

constexpr int ST_REMOVED = 5;  // internal: removed, awaiting compaction

template <typename R>
struct StepArgs {
  GridView g;
  int64_t n;
  double *pos;
  R *vel;
  R *w;
  int32_t *cell;
  int32_t *status;
  const R *corner;
  double dt;
  int max_traversal;
  double thresh_scale;
  R clamp;
  double bx0, by0, bz0, bx1, by1, bz1;
  int32_t *count;  // [ncell] histogram (zeroed by the caller)
  int32_t *slot;   // [n] slot in the cell bucket, -1 if removed
  unsigned long long *counters;
};

enum WalkOutcome : int { W_LOCATED = 0, W_LOST = 1, W_DEGENERATE = 2 };

// The walk of walkingsearch.c:1123-1212 from cell (ci,cj,ck) (negative ->
// (0,0,0), :611-616) for one lane.  On success (ci,cj,ck) is the owning cell
// and wv the weights.
template <typename R, bool EXACT>
__device__ __forceinline__ int walk_one(const GridView &g, const double x[3], int &ci, int &cj, int &ck,
                                        int max_traversal, double thresh_scale, R clamp, R wv[3], int &steps,
                                        unsigned long long &fallbacks) {
  if (ci < 0 || cj < 0 || ck < 0) {
    ci = 0;
    cj = 0;
    ck = 0;
  }
  steps = 0;
  CycleGuard cyc;
  while (steps < max_traversal) {
    steps++;
    if (ci < 0 || ci >= g.IM || cj < 0 || cj >= g.JM || ck < 0 || ck >= g.KM) return W_LOST;
    if (cyc.revisit(pack_cell(ci, cj, ck, g.IM, g.JM))) {
      steps = max_traversal;  // the reference walks the cycle until MAX_TRAVERSAL
      return W_LOST;
    }
    R d[6];
    if constexpr (EXACT) {
      double V[8][3];
      load_vertices(g, ci, cj, ck, V);
      double dd[6];
      if (!eval_cell_exact(V, x, thresh_scale, dd)) return W_DEGENERATE;
#pragma unroll
      for (int f = 0; f < 6; ++f) d[f] = dd[f];
    } else {
      float dd[6];
      FilterResult fr = eval_cell_fp32(g, ci, cj, ck, x, thresh_scale, dd);
      if (fr.degenerate) return W_DEGENERATE;
      fallbacks += fr.fallbacks;
#pragma unroll
      for (int f = 0; f < 6; ++f) d[f] = dd[f];
    }
    if (point_position<R>(d) >= 0) {
      weights_from_distances<R>(d, clamp, wv);
      return W_LOCATED;
    }
    step_indices<R>(d, ci, cj, ck);
  }
  return W_LOST;
}

template <typename R>
__device__ __forceinline__ void interp_at(const GridView &g, const R *__restrict__ corner, int ci, int cj, int ck,
                                          const R wv[3], R u[3]) {
  trilinear<R>(g.IM, g.JM, corner, ci, cj, ck, wv[0], wv[1], wv[2], u);
}

template <typename R, bool EXACT, bool RK2>
__global__ __launch_bounds__(kBlock) void k_step_fused(StepArgs<R> a) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int located = 0, lost = 0, oob = 0, err = 0, iters = 0;
  unsigned long long fb = 0;
  if (p < a.n) {
    const int32_t st = a.status[p];
    if (st == ST_REMOVED) {
      a.slot[p] = -1;
    } else {
      const double x0[3] = {a.pos[3 * p], a.pos[3 * p + 1], a.pos[3 * p + 2]};
      const R u0[3] = {a.vel[3 * p], a.vel[3 * p + 1], a.vel[3 * p + 2]};
      int ci = -1, cj = -1, ck = -1;
      const int32_t c0 = a.cell[p];
      if (c0 >= 0) unpack_cell(c0, a.g.IM, a.g.JM, ci, cj, ck);
      double x1[3];
      R wv[3] = {a.w[3 * p], a.w[3 * p + 1], a.w[3 * p + 2]};
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
          outcome = walk_one<R, EXACT>(a.g, xs, ci, cj, ck, a.max_traversal, a.thresh_scale, a.clamp, wv, s1, fb);
          if (outcome == W_LOCATED) {
            R us[3];
            interp_at<R>(a.g, a.corner, ci, cj, ck, wv, us);
            const double h = 0.5 * a.dt;
#pragma unroll
            for (int c = 0; c < 3; ++c) x1[c] = x0[c] + ((double)u0[c] + (double)us[c]) * h;
            outcome =
                walk_one<R, EXACT>(a.g, x1, ci, cj, ck, a.max_traversal, a.thresh_scale, a.clamp, wv, s2, fb);
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
        outcome = walk_one<R, EXACT>(a.g, x1, ci, cj, ck, a.max_traversal, a.thresh_scale, a.clamp, wv, s1, fb);
      }
      iters = s1 + s2;
      a.pos[3 * p] = x1[0];
      a.pos[3 * p + 1] = x1[1];
      a.pos[3 * p + 2] = x1[2];
      if (outcome == W_DEGENERATE) {
        err = 1;
        a.slot[p] = -1;
      } else if (outcome == W_LOST) {
        lost = 1;
        a.cell[p] = -1;
        a.status[p] = ST_REMOVED;
        a.slot[p] = -1;
      } else {
        const bool in = x1[0] >= a.bx0 && x1[0] <= a.bx1 && x1[1] >= a.by0 && x1[1] <= a.by1 && x1[2] >= a.bz0 &&
                        x1[2] <= a.bz1;
        const int32_t c1 = pack_cell(ci, cj, ck, a.g.IM, a.g.JM);
        a.cell[p] = c1;
        a.w[3 * p] = wv[0];
        a.w[3 * p + 1] = wv[1];
        a.w[3 * p + 2] = wv[2];
        if (in) {
          located = 1;
          R u1[3];
          interp_at<R>(a.g, a.corner, ci, cj, ck, wv, u1);
          a.vel[3 * p] = u1[0];
          a.vel[3 * p + 1] = u1[1];
          a.vel[3 * p + 2] = u1[2];
          a.status[p] = ST_LOCATED;
          a.slot[p] = atomicAdd(&a.count[c1], 1);
        } else {
          oob = 1;
          a.status[p] = ST_REMOVED;
          a.slot[p] = -1;
        }
      }
    }
  }
  // one set of atomics per wavefront
  unsigned long long v_loc = located, v_lost = lost, v_oob = oob, v_it = iters, v_fb = fb;
  int v_err = err;
  for (int off = 32; off > 0; off >>= 1) {
    v_loc += __shfl_xor(v_loc, off);
    v_lost += __shfl_xor(v_lost, off);
    v_oob += __shfl_xor(v_oob, off);
    v_it += __shfl_xor(v_it, off);
    v_fb += __shfl_xor(v_fb, off);
    v_err |= __shfl_xor(v_err, off);
  }
  if ((threadIdx.x & 63) == 0) {
    if (v_loc) atomicAdd(&a.counters[C_LOCATED], v_loc);
    if (v_lost) atomicAdd(&a.counters[C_REMOVED_LOST], v_lost);
    if (v_oob) atomicAdd(&a.counters[C_REMOVED_OOB], v_oob);
    if (v_it) atomicAdd(&a.counters[C_WALK_ITERS], v_it);
    if (v_fb) atomicAdd(&a.counters[C_FALLBACKS], v_fb);
    if (v_err) atomicOr(&a.counters[C_ERROR], 1ull);
  }
}

// idx[start[c] + slot[p]] = p for every live particle.
__global__ __launch_bounds__(kBlock) void k_index_scatter(int64_t n, const int32_t *__restrict__ cell,
                                                          const int32_t *__restrict__ slot,
                                                          const int32_t *__restrict__ start, int32_t *__restrict__ idx) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int32_t s = slot[p];
  if (s < 0) return;
  idx[start[cell[p]] + s] = (int32_t)p;
}

// Visit cell segment [s,e) of idx in ascending PID order.
__device__ __forceinline__ int64_t next_by_pid_idx(const int32_t *__restrict__ idx, const int64_t *__restrict__ pid,
                                                   int64_t s, int64_t e, int64_t last, int64_t &out_pid) {
  int64_t best = -1, bp = INT64_MAX;
  for (int64_t q = s; q < e; ++q) {
    int64_t v = pid[idx[q]];
    if (v > last && v < bp) {
      bp = v;
      best = idx[q];
    }
  }
  out_pid = bp;
  return best;
}

// Face deposition through the cell-ordered index (same definition and
// summation order as k_deposit_faces).
template <typename R>
__global__ __launch_bounds__(kBlock) void k_deposit_faces_idx(int dir, int IM, int JM, int KM,
                                                              const int32_t *__restrict__ start,
                                                              const int32_t *__restrict__ idx,
                                                              const int64_t *__restrict__ pid,
                                                              const R *__restrict__ w, const R *__restrict__ vel,
                                                              R *__restrict__ out) {
  const int fI = IM + (dir == 0), fJ = JM + (dir == 1), fK = KM + (dir == 2);
  const int64_t nf = (int64_t)fI * fJ * fK;
  int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  int i = (int)(f % fI);
  int64_t r = f / fI;
  int j = (int)(r % fJ);
  int k = (int)(r / fJ);
  Acc<R> a0, a1, a2, a3;
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
    if (e - s == 1) {
      int64_t q = idx[s];
      R av = w[3 * q + dir];
      R wt = side == 0 ? av : (R)1 - av;
      a0.add(wt);
      a1.add(wt * vel[3 * q]);
      a2.add(wt * vel[3 * q + 1]);
      a3.add(wt * vel[3 * q + 2]);
      continue;
    }
    int64_t last = INT64_MIN, lp;
    for (int64_t t = s; t < e; ++t) {
      int64_t q = next_by_pid_idx(idx, pid, s, e, last, lp);
      last = lp;
      R av = w[3 * q + dir];
      R wt = side == 0 ? av : (R)1 - av;
      a0.add(wt);
      a1.add(wt * vel[3 * q]);
      a2.add(wt * vel[3 * q + 1]);
      a3.add(wt * vel[3 * q + 2]);
    }
  }
  R *o = out + 4 * f;
  o[0] = a0.value();
  o[1] = a1.value();
  o[2] = a2.value();
  o[3] = a3.value();
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_scatter_P_idx(int64_t ncell, const int32_t *__restrict__ start,
                                                          const int32_t *__restrict__ idx,
                                                          const int64_t *__restrict__ pid, const R *__restrict__ P,
                                                          R *__restrict__ avg) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  int64_t s = start[c], e = start[c + 1];
  Acc<R> acc;
  int64_t last = INT64_MIN, lp;
  for (int64_t t = s; t < e; ++t) {
    int64_t q = next_by_pid_idx(idx, pid, s, e, last, lp);
    acc.add(P[q]);
    last = lp;
  }
  R cnt = (R)(e - s);
  avg[c] = cnt > (R)0.5 ? acc.value() / cnt : (R)0;
}

// Physical reorder: dst[q] = src[idx[q]] for q < m (cell order, live only).
template <typename R>
__global__ __launch_bounds__(kBlock) void k_gather_swarm(int64_t m, const int32_t *__restrict__ idx, const double *pos,
                                                         const R *vel, const R *w, const int32_t *cell,
                                                         const int32_t *status, const int64_t *pid, const R *P,
                                                         double *pos2, R *vel2, R *w2, int32_t *cell2,
                                                         int32_t *status2, int64_t *pid2, R *P2) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= m) return;
  int64_t p = idx[q];
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

}  // namespace picgpu
