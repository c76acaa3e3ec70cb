// picgpu.hip -- C-ABI of the gfx950 particle loop (include/picgpu.h).
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "fused.hpp"

using namespace picgpu;

namespace picgpu {

static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

hipEvent_t take_event(picgpu_ctx *c) {
  if (!c->event_pool.empty()) {
    hipEvent_t e = c->event_pool.back();
    c->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

void timer_begin(picgpu_ctx *c, const char *name, TimerRec &rec) {
  rec.a = rec.b = nullptr;
  if (!c->timing) return;
  if (!c->timing_only.empty() && strncmp(name, c->timing_only.c_str(), c->timing_only.size()) != 0) return;
  rec.name = name;
  rec.a = take_event(c);
  rec.b = take_event(c);
  if (rec.a) (void)hipEventRecord(rec.a, c->stream);
}

void timer_end(picgpu_ctx *c, TimerRec &rec) {
  if (!c->timing || !rec.a || !rec.b) return;
  (void)hipEventRecord(rec.b, c->stream);
  c->pending.push_back(rec);
}

int flush_timers(picgpu_ctx *c) {
  if (c->pending.empty()) return PICGPU_OK;
  PG_HIP(hipStreamSynchronize(c->stream));
  for (auto &r : c->pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
      auto &acc = c->timing_acc[r.name];
      acc.first += ms;
      acc.second += 1;
    }
    c->event_pool.push_back(r.a);
    c->event_pool.push_back(r.b);
  }
  c->pending.clear();
  return PICGPU_OK;
}

}  // namespace picgpu

static inline int64_t nblocks(int64_t n, int b = kBlock) { return (n + b - 1) / b; }

template <typename F>
static int dispatch_real(picgpu_ctx *c, F &&f) {
  if (c->opt.precision == PICGPU_FP32) return f((float *)nullptr);
  return f((double *)nullptr);
}

static constexpr size_t kCounterBytes = sizeof(unsigned long long) * C_NUM * kCounterStripes;
static constexpr size_t kHistBytes = sizeof(unsigned long long) * PICGPU_WALK_HIST_BINS * kCounterStripes;

// walk-iteration histogram of the next locate / fused step (collect_walk_stats)
static unsigned long long *hist_reset(picgpu_ctx *c) {
  if (!c->opt.collect_walk_stats) return nullptr;
  if (c->walk_hist.ensure(kHistBytes) != PICGPU_OK) return nullptr;
  if (hipMemsetAsync(c->walk_hist.ptr, 0, kHistBytes, c->stream) != hipSuccess) return nullptr;
  return c->walk_hist.as<unsigned long long>();
}

static int hist_read(picgpu_ctx *c) {
  if (!c->opt.collect_walk_stats || !c->walk_hist.ptr) return PICGPU_OK;
  std::vector<unsigned long long> h((size_t)PICGPU_WALK_HIST_BINS * kCounterStripes);
  PG_HIP(hipMemcpyAsync(h.data(), c->walk_hist.ptr, kHistBytes, hipMemcpyDeviceToHost, c->stream));
  PG_HIP(hipStreamSynchronize(c->stream));
  for (int b = 0; b < PICGPU_WALK_HIST_BINS; ++b) {
    int64_t v = 0;
    for (int s = 0; s < kCounterStripes; ++s) v += (int64_t)h[(size_t)s * PICGPU_WALK_HIST_BINS + b];
    c->stats.walk_hist[b] = v;
  }
  return PICGPU_OK;
}

static int counters_reset(picgpu_ctx *c) {
  PG_CHECK(c->counters.ensure(kCounterBytes));
  PG_HIP(hipMemsetAsync(c->counters.ptr, 0, kCounterBytes, c->stream));
  return PICGPU_OK;
}

// sums the stripes (maximum for C_WALK_MAX)
static int counters_read(picgpu_ctx *c, unsigned long long *h) {
  unsigned long long all[C_NUM * kCounterStripes];
  PG_HIP(hipMemcpyAsync(all, c->counters.ptr, kCounterBytes, hipMemcpyDeviceToHost, c->stream));
  PG_HIP(hipStreamSynchronize(c->stream));
  for (int k = 0; k < C_NUM; ++k) {
    unsigned long long v = 0;
    for (int s = 0; s < kCounterStripes; ++s) {
      const unsigned long long x = all[s * C_NUM + k];
      v = k == C_WALK_MAX ? (x > v ? x : v) : v + x;
    }
    h[k] = v;
  }
  return PICGPU_OK;
}

static int ensure_swarm(picgpu_ctx *c, SwarmBufs &s, int64_t n) {
  size_t m = (size_t)(n > 0 ? n : 1);
  PG_CHECK(s.pos.ensure(m * 3 * sizeof(double)));
  PG_CHECK(s.vel.ensure(m * 3 * c->rsize));
  PG_CHECK(s.w.ensure(m * 3 * c->rsize));
  PG_CHECK(s.cell.ensure(m * sizeof(int32_t)));
  PG_CHECK(s.status.ensure(m * sizeof(int32_t)));
  PG_CHECK(s.pid.ensure(m * sizeof(int64_t)));
  PG_CHECK(s.P.ensure(m * c->rsize));
  PG_CHECK(s.pid32.ensure(m * sizeof(int32_t)));
  return PICGPU_OK;
}

// Every move of the swarm arrays goes through here; only the key-order
// compaction of the fused steps moves pid32 along (and says so after the swap).
static void swap_swarm(picgpu_ctx *c) {
  std::swap(c->sw, c->alt);
  c->pid32_valid = false;
}

static int need_grid(picgpu_ctx *c) {
  if (!c->have_grid) {
    set_error("grid not uploaded");
    return PICGPU_E_STATE;
  }
  return PICGPU_OK;
}

static GridView grid_view(picgpu_ctx *c) {
  return GridView{c->coor.as<double>(), c->IM, c->JM, c->KM, c->IM > 0 ? 1.0 / c->IM : 0.0,
                  c->IM > 0 && c->JM > 0 ? 1.0 / ((double)c->IM * (double)c->JM) : 0.0};
}

// Padded bounding box of all nodes (grid.c:762-883 on one rank) and state reset
// after a grid upload.
static int finish_grid(picgpu_ctx *c) {
  const int64_t nb = std::min<int64_t>(1024, nblocks(c->nnode));
  DevBuf part;
  PG_CHECK(part.ensure(sizeof(double) * 6 * (size_t)nb));
  PG_LAUNCH(c, "bbox", nb, kBlock, k_bbox_partial, c->nnode, c->coor.as<double>(), part.as<double>());
  std::vector<double> h(6 * (size_t)nb);
  hipError_t e1 = hipMemcpyAsync(h.data(), part.ptr, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream);
  hipError_t e2 = hipStreamSynchronize(c->stream);
  part.release();
  if (e1 != hipSuccess || e2 != hipSuccess) {
    set_error("bounding box reduction failed");
    return PICGPU_E_HIP;
  }
  for (int a = 0; a < 3; ++a) {
    c->bmin[a] = 1.7976931348623157e308;
    c->bmax[a] = -1.7976931348623157e308;
  }
  for (int64_t b = 0; b < nb; ++b)
    for (int a = 0; a < 3; ++a) {
      c->bmin[a] = std::min(c->bmin[a], h[6 * b + a]);
      c->bmax[a] = std::max(c->bmax[a], h[6 * b + 3 + a]);
    }
  for (int a = 0; a < 3; ++a) {
    c->bmin[a] = c->bmin[a] - c->opt.bbox_tolerance;
    c->bmax[a] = c->bmax[a] + c->opt.bbox_tolerance;
  }
  c->have_grid = true;
  c->have_planes = false;
  c->have_ucat = false;
  c->eul.have_metrics = c->eul.have_bc = false;
  c->sorted = c->keyed = c->key_ordered = c->compact_pending = false;
  c->have_pavg = c->have_faces = false;
  return PICGPU_OK;
}

// ---------------------------------------------------------------------------
// lifetime
// ---------------------------------------------------------------------------
extern "C" {

void picgpu_default_options(picgpu_options *o) {
  if (!o) return;
  memset(o, 0, sizeof(*o));
  o->precision = PICGPU_FP64;
  o->use_mlp_guess = 0;
  o->max_traversal = 1000;
  o->distance_threshold = 1e-11;
  o->weight_clamp = 1.0e-14;
  o->bbox_tolerance = 1e-6;
  o->device = 0;
  o->collect_walk_stats = 0;
}

const char *picgpu_last_error(void) { return g_err; }

int picgpu_create(picgpu_ctx **out, const picgpu_options *opt) {
  if (!out) {
    set_error("null ctx pointer");
    return PICGPU_E_ARG;
  }
  *out = nullptr;
  picgpu_ctx *c = new picgpu_ctx();
  if (opt)
    c->opt = *opt;
  else
    picgpu_default_options(&c->opt);
  if (c->opt.precision != PICGPU_FP64 && c->opt.precision != PICGPU_FP32) {
    delete c;
    set_error("unknown precision mode %d", opt ? opt->precision : -1);
    return PICGPU_E_ARG;
  }
  if (c->opt.max_traversal <= 0) {
    delete c;
    set_error("max_traversal must be positive");
    return PICGPU_E_ARG;
  }
  c->rsize = c->opt.precision == PICGPU_FP32 ? sizeof(float) : sizeof(double);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    (void)hipGetLastError();
    delete c;
    set_error("no HIP device available");
    return PICGPU_E_HIP;
  }
  if (c->opt.device < 0 || c->opt.device >= ndev) {
    delete c;
    set_error("device %d out of range (%d devices)", c->opt.device, ndev);
    return PICGPU_E_ARG;
  }
  c->device = c->opt.device;
  if (hipSetDevice(c->device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    set_error("HIP stream creation failed");
    return PICGPU_E_HIP;
  }
  const int rs = noalloc_setup(c->device);
  if (rs != PICGPU_OK) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return rs;
  }
  *out = c;
  return PICGPU_OK;
}

int picgpu_destroy(picgpu_ctx *c) {
  if (!c) return PICGPU_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (auto &r : c->pending) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (auto e : c->event_pool) (void)hipEventDestroy(e);
  DevBuf *bufs[] = {&c->coor, &c->ucat, &c->corner, &c->rank, &c->keep, &c->walk_steps, &c->guess, &c->pos0,
                    &c->vel0, &c->cell_count, &c->cell_start, &c->pavg, &c->face[0], &c->face[1], &c->face[2],
                    &c->scan_tmp, &c->counters, &c->mlp_master, &c->mlp_bf16, &c->mlp_adam_m, &c->mlp_adam_v,
                    &c->mlp_grad, &c->mlp_partial, &c->mlp_x, &c->mlp_t, &c->mlp_y, &c->kcount,
                    &c->kstart[0], &c->kstart[1], &c->kidx[0], &c->kidx[1], &c->kslot, &c->kbase, &c->planes,
                    &c->defer, &c->walk_hist};
  for (DevBuf *b : bufs) b->release();
  c->step_counters.release();
  c->seams.release();
  eulerian_release(c);
  SwarmBufs *sws[] = {&c->sw, &c->alt};
  for (SwarmBufs *s : sws) {
    s->pos.release();
    s->vel.release();
    s->w.release();
    s->cell.release();
    s->status.release();
    s->pid.release();
    s->P.release();
    s->pid32.release();
  }
  c->pid_wide.release();
  (void)hipStreamDestroy(c->stream);
  delete c;
  return PICGPU_OK;
}

int picgpu_synchronize(picgpu_ctx *c) {
  if (!c) return PICGPU_E_ARG;
  PG_HIP(hipStreamSynchronize(c->stream));
  return PICGPU_OK;
}

// ---------------------------------------------------------------------------
// grid and Eulerian input
// ---------------------------------------------------------------------------
int picgpu_upload_grid(picgpu_ctx *c, int64_t IM, int64_t JM, int64_t KM, const double *coor) {
  if (!c || !coor || IM <= 0 || JM <= 0 || KM <= 0) {
    set_error("upload_grid: bad arguments");
    return PICGPU_E_ARG;
  }
  PG_HIP(hipSetDevice(c->device));
  if (!c->swarm_invalid) PG_CHECK(settle_swarm(c));  // the particles stay; the new grid drops their index
  if (IM * JM * KM >= (int64_t)INT32_MAX - 2 || (IM + 1) * (JM + 1) * (KM + 1) >= ((int64_t)1 << 40)) {
    set_error("upload_grid: grid too large for 32-bit packed cell indices");
    return PICGPU_E_UNSUPPORTED;
  }
  PG_HIP(hipSetDevice(c->device));
  c->IM = (int)IM;
  c->JM = (int)JM;
  c->KM = (int)KM;
  c->ncell = IM * JM * KM;
  c->nnode = (IM + 1) * (JM + 1) * (KM + 1);
  PG_CHECK(c->coor.ensure(sizeof(double) * 3 * (size_t)c->nnode));
  PG_HIP(hipMemcpy(c->coor.ptr, coor, sizeof(double) * 3 * (size_t)c->nnode, hipMemcpyHostToDevice));
  return finish_grid(c);
}

int picgpu_upload_grid_device(picgpu_ctx *c, int64_t IM, int64_t JM, int64_t KM, const double *d_coor) {
  if (!c || !d_coor || IM <= 0 || JM <= 0 || KM <= 0) {
    set_error("upload_grid_device: bad arguments");
    return PICGPU_E_ARG;
  }
  PG_HIP(hipSetDevice(c->device));
  if (!c->swarm_invalid) PG_CHECK(settle_swarm(c));  // the particles stay; the new grid drops their index
  if (IM * JM * KM >= (int64_t)INT32_MAX - 2) {
    set_error("upload_grid_device: grid too large for 32-bit packed cell indices");
    return PICGPU_E_UNSUPPORTED;
  }
  PG_HIP(hipSetDevice(c->device));
  c->IM = (int)IM;
  c->JM = (int)JM;
  c->KM = (int)KM;
  c->ncell = IM * JM * KM;
  c->nnode = (IM + 1) * (JM + 1) * (KM + 1);
  PG_CHECK(c->coor.ensure(sizeof(double) * 3 * (size_t)c->nnode));
  PG_HIP(hipMemcpyAsync(c->coor.ptr, d_coor, sizeof(double) * 3 * (size_t)c->nnode, hipMemcpyDeviceToDevice,
                        c->stream));
  return finish_grid(c);
}

int picgpu_upload_ucat_device(picgpu_ctx *c, const double *d_ucat) {
  if (!c || !d_ucat) return PICGPU_E_ARG;
  PG_CHECK(need_grid(c));
  PG_HIP(hipSetDevice(c->device));
  PG_CHECK(c->ucat.ensure(c->rsize * 3 * (size_t)c->ncell));
  PG_CHECK(c->corner.ensure(c->rsize * 3 * (size_t)c->nnode));
  int r = dispatch_real(c, [&](auto *tag) -> int {
    using R = std::remove_pointer_t<decltype(tag)>;
    PG_LAUNCH(c, "ucat_compact", nblocks(c->ncell), kBlock, k_ucat_compact<R>, c->IM, c->JM, c->KM, d_ucat,
              c->ucat.as<R>());
    return PICGPU_OK;
  });
  PG_CHECK(r);
  PG_HIP(hipStreamSynchronize(c->stream));
  c->have_ucat = true;
  return eulerian_take_ucat(c, d_ucat, hipMemcpyDeviceToDevice);
}

int picgpu_init_particles_device(picgpu_ctx *c, int64_t n, const double *d_pos, int64_t pid_base) {
  if (!c || n < 0 || (n > 0 && !d_pos)) return PICGPU_E_ARG;
  PG_CHECK(need_grid(c));
  if (n >= (int64_t)INT32_MAX) {
    set_error("init_particles_device: more than 2^31-1 particles");
    return PICGPU_E_UNSUPPORTED;
  }
  PG_HIP(hipSetDevice(c->device));
  PG_CHECK(ensure_swarm(c, c->sw, n));
  c->n = n;
  c->sorted = c->keyed = c->key_ordered = c->compact_pending = false;
  c->swarm_invalid = false;
  c->pid32_valid = false;
  if (n == 0) return PICGPU_OK;
  int r = dispatch_real(c, [&](auto *tag) -> int {
    using R = std::remove_pointer_t<decltype(tag)>;
    PG_LAUNCH(c, "init_particles", nblocks(n), kBlock, k_init_particles<R>, n, d_pos, pid_base,
              c->sw.pos.as<double>(), c->sw.vel.as<R>(), c->sw.w.as<R>(), c->sw.cell.as<int32_t>(),
              c->sw.status.as<int32_t>(), c->sw.pid.as<int64_t>(), c->sw.P.as<R>());
    return PICGPU_OK;
  });
  PG_CHECK(r);
  PG_HIP(hipStreamSynchronize(c->stream));
  return PICGPU_OK;
}

int picgpu_upload_ucat(picgpu_ctx *c, const double *ucat) {
  if (!c || !ucat) return PICGPU_E_ARG;
  PG_CHECK(need_grid(c));
  PG_HIP(hipSetDevice(c->device));
  const int64_t nx = c->IM + 1, ny = c->JM + 1;
  PG_CHECK(c->ucat.ensure(c->rsize * 3 * (size_t)c->ncell));
  PG_CHECK(c->corner.ensure(c->rsize * 3 * (size_t)c->nnode));
  return dispatch_real(c, [&](auto *tag) -> int {
    using R = std::remove_pointer_t<decltype(tag)>;
    std::vector<R> h((size_t)c->ncell * 3);
    for (int64_t k = 0; k < c->KM; ++k)
      for (int64_t j = 0; j < c->JM; ++j)
        for (int64_t i = 0; i < c->IM; ++i) {
          const double *u = ucat + 3 * ((k * ny + j) * nx + i);
          R *o = h.data() + 3 * ((k * c->JM + j) * c->IM + i);
          o[0] = (R)u[0];
          o[1] = (R)u[1];
          o[2] = (R)u[2];
        }
    PG_HIP(hipMemcpy(c->ucat.ptr, h.data(), sizeof(R) * h.size(), hipMemcpyHostToDevice));
    c->have_ucat = true;
    return eulerian_take_ucat(c, ucat, hipMemcpyHostToDevice);
  });
}

// ---------------------------------------------------------------------------
// swarm I/O
// ---------------------------------------------------------------------------
int picgpu_upload_particles(picgpu_ctx *c, int64_t n, const double *pos, const double *vel, const int64_t *cell,
                            const double *weight, const int64_t *status, const int64_t *pid, const double *P) {
  if (!c || n < 0 || (n > 0 && (!pos || !cell || !status || !pid))) {
    set_error("upload_particles: bad arguments");
    return PICGPU_E_ARG;
  }
  PG_CHECK(need_grid(c));
  if (n >= (int64_t)INT32_MAX) {
    set_error("upload_particles: more than 2^31-1 particles");
    return PICGPU_E_UNSUPPORTED;
  }
  PG_HIP(hipSetDevice(c->device));
  PG_CHECK(ensure_swarm(c, c->sw, n));
  c->n = n;
  c->sorted = c->keyed = c->key_ordered = c->compact_pending = false;
  c->swarm_invalid = false;
  c->pid32_valid = false;
  if (n == 0) return PICGPU_OK;
  PG_HIP(hipMemcpy(c->sw.pos.ptr, pos, sizeof(double) * 3 * (size_t)n, hipMemcpyHostToDevice));
  PG_HIP(hipMemcpy(c->sw.pid.ptr, pid, sizeof(int64_t) * (size_t)n, hipMemcpyHostToDevice));
  {
    std::vector<int32_t> hc((size_t)n), hs((size_t)n);
    for (int64_t p = 0; p < n; ++p) {
      int64_t i = cell[3 * p], j = cell[3 * p + 1], k = cell[3 * p + 2];
      bool ok = i >= 0 && j >= 0 && k >= 0 && i < c->IM && j < c->JM && k < c->KM;
      hc[p] = ok ? (int32_t)(i + (int64_t)c->IM * (j + (int64_t)c->JM * k)) : -1;
      hs[p] = (int32_t)status[p];
    }
    PG_HIP(hipMemcpy(c->sw.cell.ptr, hc.data(), sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice));
    PG_HIP(hipMemcpy(c->sw.status.ptr, hs.data(), sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice));
  }
  return dispatch_real(c, [&](auto *tag) -> int {
    using R = std::remove_pointer_t<decltype(tag)>;
    std::vector<R> h((size_t)n * 3);
    for (int64_t q = 0; q < 3 * n; ++q) h[q] = vel ? (R)vel[q] : (R)0;
    PG_HIP(hipMemcpy(c->sw.vel.ptr, h.data(), sizeof(R) * 3 * (size_t)n, hipMemcpyHostToDevice));
    for (int64_t q = 0; q < 3 * n; ++q) h[q] = weight ? (R)weight[q] : (R)0;
    PG_HIP(hipMemcpy(c->sw.w.ptr, h.data(), sizeof(R) * 3 * (size_t)n, hipMemcpyHostToDevice));
    for (int64_t q = 0; q < n; ++q) h[q] = P ? (R)P[q] : (R)101325.0;
    PG_HIP(hipMemcpy(c->sw.P.ptr, h.data(), sizeof(R) * (size_t)n, hipMemcpyHostToDevice));
    return PICGPU_OK;
  });
}

int picgpu_num_particles(picgpu_ctx *c, int64_t *n) {
  if (!c || !n) return PICGPU_E_ARG;
  *n = live_particles(c);
  return PICGPU_OK;
}

int picgpu_download_particles(picgpu_ctx *c, double *pos, double *vel, int64_t *cell, double *weight,
                              int64_t *status, int64_t *pid) {
  if (!c) return PICGPU_E_ARG;
  PG_HIP(hipSetDevice(c->device));
  PG_CHECK(settle_swarm(c));
  const int64_t n = c->n;
  PG_HIP(hipStreamSynchronize(c->stream));
  if (n == 0) return PICGPU_OK;
  if (pos) PG_HIP(hipMemcpy(pos, c->sw.pos.ptr, sizeof(double) * 3 * (size_t)n, hipMemcpyDeviceToHost));
  if (pid) PG_HIP(hipMemcpy(pid, c->sw.pid.ptr, sizeof(int64_t) * (size_t)n, hipMemcpyDeviceToHost));
  if (cell || status) {
    std::vector<int32_t> hc((size_t)n), hs((size_t)n);
    PG_HIP(hipMemcpy(hc.data(), c->sw.cell.ptr, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost));
    PG_HIP(hipMemcpy(hs.data(), c->sw.status.ptr, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost));
    for (int64_t p = 0; p < n; ++p) {
      if (cell) {
        int32_t v = hc[p];
        if (v < 0) {
          cell[3 * p] = cell[3 * p + 1] = cell[3 * p + 2] = -1;
        } else {
          cell[3 * p] = v % c->IM;
          cell[3 * p + 1] = (v / c->IM) % c->JM;
          cell[3 * p + 2] = v / ((int64_t)c->IM * c->JM);
        }
      }
      if (status) status[p] = hs[p];
    }
  }
  return dispatch_real(c, [&](auto *tag) -> int {
    using R = std::remove_pointer_t<decltype(tag)>;
    std::vector<R> h((size_t)n * 3);
    if (vel) {
      PG_HIP(hipMemcpy(h.data(), c->sw.vel.ptr, sizeof(R) * 3 * (size_t)n, hipMemcpyDeviceToHost));
      for (int64_t q = 0; q < 3 * n; ++q) vel[q] = (double)h[q];
    }
    if (weight) {
      PG_HIP(hipMemcpy(h.data(), c->sw.w.ptr, sizeof(R) * 3 * (size_t)n, hipMemcpyDeviceToHost));
      for (int64_t q = 0; q < 3 * n; ++q) weight[q] = (double)h[q];
    }
    return PICGPU_OK;
  });
}

int picgpu_download_walk_steps(picgpu_ctx *c, int32_t *steps) {
  if (!c || !steps) return PICGPU_E_ARG;
  PG_HIP(hipSetDevice(c->device));
  PG_CHECK(settle_swarm(c));
  if (!c->opt.collect_walk_stats || !c->walk_steps.ptr || c->walk_steps.bytes < sizeof(int32_t) * (size_t)c->n) {
    set_error("walk statistics not collected (collect_walk_stats=0 or no locate yet)");
    return PICGPU_E_STATE;
  }
  PG_HIP(hipStreamSynchronize(c->stream));
  if (c->n) PG_HIP(hipMemcpy(steps, c->walk_steps.ptr, sizeof(int32_t) * (size_t)c->n, hipMemcpyDeviceToHost));
  return PICGPU_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// stage implementations
// ---------------------------------------------------------------------------
static int do_reset_advect(picgpu_ctx *c, double dt, int reset, int advect) {
  PG_CHECK(settle_swarm(c));
  if (c->n == 0) return PICGPU_OK;
  return dispatch_real(c, [&](auto *tag) -> int {
    using R = std::remove_pointer_t<decltype(tag)>;
    PG_LAUNCH(c, "reset_advect", nblocks(c->n), kBlock, k_reset_advect<R>, c->n, c->sw.pos.as<double>(),
              c->sw.vel.as<R>(), c->sw.status.as<int32_t>(), dt, reset, advect);
    c->sorted = c->keyed = c->key_ordered = c->compact_pending = false;
    return PICGPU_OK;
  });
}

// Face-plane records of every cell, computed once per grid.
static int ensure_planes(picgpu_ctx *c) {
  if (c->have_planes) return PICGPU_OK;
  const size_t rec = c->opt.precision == PICGPU_FP32 ? sizeof(CellPlanes32) : sizeof(CellPlanes64);
  PG_CHECK(c->planes.ensure(rec * (size_t)(c->ncell > 0 ? c->ncell : 1)));
  if (c->ncell > 0) {
    if (c->opt.precision == PICGPU_FP32) {
      PG_LAUNCH(c, "cell_planes", nblocks(c->ncell), kBlock, k_cell_planes32, grid_view(c),
                c->opt.distance_threshold, c->planes.as<CellPlanes32>());
    } else {
      PG_LAUNCH(c, "cell_planes", nblocks(c->ncell), kBlock, k_cell_planes64, grid_view(c),
                c->opt.distance_threshold, c->planes.as<CellPlanes64>());
    }
  }
  c->have_planes = true;
  return PICGPU_OK;
}

static int build_key_index(picgpu_ctx *c, const int32_t *guess = nullptr);

// An FP64 locate of at least this many particles from guessed start cells
// visits them grouped by start cell (tile-major key), so that neighbouring
// lanes' walks read the same 304-byte face-plane records.  Each particle's
// walk, and so its result, is the same in any order, and the counters are
// integer sums.  Fresh positions in a 128x128x1024 tube: 9.1e8 -> 1.5e9
// locates/s in FP64; in FP32 (128-byte records) the scattered particle
// reads and writes cost more than the shared records save (3.7e9 -> 2.4e9),
// so FP32 visits in index order.
static constexpr int64_t kOrderedLocateMin = 1 << 20;

static int do_locate(picgpu_ctx *c, int skip_lost) {
  PG_CHECK(need_grid(c));
  PG_CHECK(settle_swarm(c));
  PG_CHECK(ensure_planes(c));
  PG_CHECK(counters_reset(c));
  const int64_t n = c->n;
  if (n > 0) {
    int32_t *guess = nullptr;
    if (c->opt.use_mlp_guess) {
      if (!c->have_mlp) {
        set_error("use_mlp_guess=1 but no MLP parameters are set");
        return PICGPU_E_STATE;
      }
      PG_CHECK(c->guess.ensure(sizeof(int32_t) * (size_t)n));
      PG_CHECK(mlp_guess_device(c, n, c->sw.pos.as<double>(), c->sw.cell.as<int32_t>(), c->guess.as<int32_t>()));
      guess = c->guess.as<int32_t>();
    }
    int32_t *steps = nullptr;
    if (c->opt.collect_walk_stats) {
      PG_CHECK(c->walk_steps.ensure(sizeof(int32_t) * (size_t)n));
      steps = c->walk_steps.as<int32_t>();
    }
    const int32_t *order = nullptr;
    if (guess && n >= kOrderedLocateMin && c->opt.precision == PICGPU_FP64) {
      // the key index buffers hold the order; the index is rebuilt before the
      // next fused step (marked stale first, in case the build fails part-way)
      c->keyed = c->key_ordered = false;
      PG_CHECK(build_key_index(c, guess));
      if (c->n_live == n) order = c->kidx[c->kcur].as<int32_t>();
    }
    unsigned long long *hist = hist_reset(c);
    const int64_t chunk = 256;  // particles per wavefront
    const int64_t waves = (n + chunk - 1) / chunk;
    const int64_t blocks = (waves * 64 + kBlock - 1) / kBlock;
    int r = dispatch_real(c, [&](auto *tag) -> int {
      using R = std::remove_pointer_t<decltype(tag)>;
      LocateArgs<R> a;
      a.g = grid_view(c);
      a.planes = c->planes.ptr;
      a.n = n;
      a.chunk = chunk;
      a.pos = c->sw.pos.as<double>();
      a.cell = c->sw.cell.as<int32_t>();
      a.w = c->sw.w.as<R>();
      a.status = c->sw.status.as<int32_t>();
      a.guess = guess;
      a.order = order;
      a.steps_out = steps;
      a.walk_hist = hist;
      a.counters = c->counters.as<unsigned long long>();
      a.max_traversal = c->opt.max_traversal;
      a.thresh_scale = c->opt.distance_threshold;
      a.clamp = (R)c->opt.weight_clamp;
      a.skip_lost = skip_lost;
      constexpr bool EX = sizeof(R) == sizeof(double);
      PG_LAUNCH(c, EX ? "locate_fp64" : "locate_fp32", blocks, kBlock, (k_locate<R, EX>), a);
      return PICGPU_OK;
    });
    PG_CHECK(r);
  }
  unsigned long long h[C_NUM];
  PG_CHECK(counters_read(c, h));
  if (h[C_ERROR]) {
    set_error("degenerate cell face (|n| < 1e-12) met during the walk");
    return PICGPU_E_DEGENERATE;
  }
  c->stats.walk_iterations = (int64_t)h[C_WALK_ITERS];
  c->stats.walk_max = (int64_t)h[C_WALK_MAX];
  c->stats.fp64_fallbacks = (int64_t)h[C_FALLBACKS];
  c->stats.located = (int64_t)h[C_LOCATED];
  c->stats.lost = (int64_t)h[C_LOST];
  if (n > 0) PG_CHECK(hist_read(c));
  c->sorted = c->keyed = c->key_ordered = c->compact_pending = false;
  return PICGPU_OK;
}

// Stable compaction of the swarm keeping particles by the flags.
static int do_remove(picgpu_ctx *c, int check_lost, int check_box, int64_t *removed_lost, int64_t *removed_oob) {
  PG_CHECK(settle_swarm(c));
  const int64_t n = c->n;
  if (removed_lost) *removed_lost = 0;
  if (removed_oob) *removed_oob = 0;
  if (n == 0) return PICGPU_OK;
  PG_CHECK(counters_reset(c));
  PG_CHECK(c->keep.ensure(sizeof(int32_t) * (size_t)(2 * n + 2)));
  int32_t *keep = c->keep.as<int32_t>();
  int32_t *off = keep + n + 1;
  PG_LAUNCH(c, "keep_flags", nblocks(n), kBlock, k_keep_flags, n, c->sw.pos.as<double>(), c->sw.status.as<int32_t>(),
            check_lost, check_box, c->bmin[0], c->bmin[1], c->bmin[2], c->bmax[0], c->bmax[1], c->bmax[2], keep,
            c->counters.as<unsigned long long>());
  PG_CHECK(scan_exclusive_i32(c, keep, off, n));
  int32_t total = 0;
  PG_HIP(hipMemcpyAsync(&total, off + n, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  unsigned long long h[C_NUM];
  PG_CHECK(counters_read(c, h));  // synchronises
  if (removed_lost) *removed_lost = (int64_t)h[C_REMOVED_LOST];
  if (removed_oob) *removed_oob = (int64_t)h[C_REMOVED_OOB];
  if (total == n) return PICGPU_OK;
  PG_CHECK(ensure_swarm(c, c->alt, n));
  int r = dispatch_real(c, [&](auto *tag) -> int {
    using R = std::remove_pointer_t<decltype(tag)>;
    PG_LAUNCH(c, "compact", nblocks(n), kBlock, k_compact<R>, n, keep, off, (int64_t)0, c->sw.pos.as<double>(), c->sw.vel.as<R>(),
              c->sw.w.as<R>(), c->sw.cell.as<int32_t>(), c->sw.status.as<int32_t>(), c->sw.pid.as<int64_t>(),
              c->sw.P.as<R>(), c->alt.pos.as<double>(), c->alt.vel.as<R>(), c->alt.w.as<R>(),
              c->alt.cell.as<int32_t>(), c->alt.status.as<int32_t>(), c->alt.pid.as<int64_t>(), c->alt.P.as<R>());
    return PICGPU_OK;
  });
  PG_CHECK(r);
  swap_swarm(c);
  c->n = total;
  c->sorted = c->keyed = c->key_ordered = c->compact_pending = false;
  return PICGPU_OK;
}

static int do_center_to_corner(picgpu_ctx *c) {
  if (!c->have_ucat) {
    set_error("ucat not uploaded");
    return PICGPU_E_STATE;
  }
  return dispatch_real(c, [&](auto *tag) -> int {
    using R = std::remove_pointer_t<decltype(tag)>;
    // node columns (i, j) x chunks of kC2CChunk planes
    const int64_t nch = (c->KM + 1 + kC2CChunk - 1) / kC2CChunk;
    const dim3 grid((unsigned)nblocks((int64_t)(c->IM + 1) * (c->JM + 1)), (unsigned)std::min<int64_t>(nch, 65535));
    PG_LAUNCH(c, "center_to_corner", grid, kBlock, k_center_to_corner<R>, c->IM, c->JM, c->KM, c->ucat.as<R>(),
              c->corner.as<R>());
    return PICGPU_OK;
  });
}

static int do_interp(picgpu_ctx *c) {
  PG_CHECK(settle_swarm(c));
  if (c->n == 0) return PICGPU_OK;
  return dispatch_real(c, [&](auto *tag) -> int {
    using R = std::remove_pointer_t<decltype(tag)>;
    PG_LAUNCH(c, "interp", nblocks(c->n), kBlock, k_interp<R>, c->n, c->IM, c->JM, c->KM, c->corner.as<R>(),
              c->sw.cell.as<int32_t>(), c->sw.w.as<R>(), c->sw.vel.as<R>());
    return PICGPU_OK;
  });
}

// Counting sort by cell; particles without a cell go to a last bucket.
static int do_sort(picgpu_ctx *c) {
  PG_CHECK(settle_swarm(c));
  const int64_t n = c->n;
  const int64_t nb = c->ncell + 1;  // + unlocated bucket
  PG_CHECK(c->cell_count.ensure(sizeof(int32_t) * (size_t)(nb + 1)));
  PG_CHECK(c->cell_start.ensure(sizeof(int32_t) * (size_t)(nb + 1)));
  PG_HIP(hipMemsetAsync(c->cell_count.ptr, 0, sizeof(int32_t) * (size_t)(nb + 1), c->stream));
  if (n > 0) {
    PG_CHECK(c->rank.ensure(sizeof(int32_t) * (size_t)n));
    PG_CHECK(ensure_swarm(c, c->alt, n));
    // unlocated particles are keyed to bucket ncell by a temporary cell view
    PG_LAUNCH(c, "cell_histogram", nblocks(n), kBlock, k_cell_histogram, n, c->sw.cell.as<int32_t>(),
              c->cell_count.as<int32_t>(), c->rank.as<int32_t>());
  }
  PG_CHECK(scan_exclusive_i32(c, c->cell_count.as<int32_t>(), c->cell_start.as<int32_t>(), c->ncell));
  if (n > 0) {
    int r = dispatch_real(c, [&](auto *tag) -> int {
      using R = std::remove_pointer_t<decltype(tag)>;
      PG_LAUNCH(c, "sort_scatter", nblocks(n), kBlock, k_sort_scatter<R>, n, c->rank.as<int32_t>(),
                c->cell_start.as<int32_t>(), c->sw.pos.as<double>(), c->sw.vel.as<R>(), c->sw.w.as<R>(),
                c->sw.cell.as<int32_t>(), c->sw.status.as<int32_t>(), c->sw.pid.as<int64_t>(), c->sw.P.as<R>(),
                c->alt.pos.as<double>(), c->alt.vel.as<R>(), c->alt.w.as<R>(), c->alt.cell.as<int32_t>(),
                c->alt.status.as<int32_t>(), c->alt.pid.as<int64_t>(), c->alt.P.as<R>());
      return PICGPU_OK;
    });
    PG_CHECK(r);
    // particles without a cell are appended after the located ones, in order
    int32_t located = 0;
    PG_HIP(hipMemcpyAsync(&located, c->cell_start.as<int32_t>() + c->ncell, sizeof(int32_t), hipMemcpyDeviceToHost,
                          c->stream));
    PG_HIP(hipStreamSynchronize(c->stream));
    if (located < n) {
      // stable compaction of the unlocated tail (rare: LOST particles kept by the caller)
      PG_CHECK(c->keep.ensure(sizeof(int32_t) * (size_t)(2 * n + 2)));
      int32_t *keep = c->keep.as<int32_t>();
      int32_t *off = keep + n + 1;
      PG_LAUNCH(c, "keep_unlocated", nblocks(n), kBlock, k_keep_unlocated, n, c->sw.cell.as<int32_t>(), keep);
      PG_CHECK(scan_exclusive_i32(c, keep, off, n));
      int r2 = dispatch_real(c, [&](auto *tag) -> int {
        using R = std::remove_pointer_t<decltype(tag)>;
        PG_LAUNCH(c, "compact_tail", nblocks(n), kBlock, k_compact<R>, n, keep, off, (int64_t)located,
                  c->sw.pos.as<double>(), c->sw.vel.as<R>(), c->sw.w.as<R>(), c->sw.cell.as<int32_t>(),
                  c->sw.status.as<int32_t>(), c->sw.pid.as<int64_t>(), c->sw.P.as<R>(), c->alt.pos.as<double>(),
                  c->alt.vel.as<R>(), c->alt.w.as<R>(), c->alt.cell.as<int32_t>(), c->alt.status.as<int32_t>(),
                  c->alt.pid.as<int64_t>(), c->alt.P.as<R>());
        return PICGPU_OK;
      });
      PG_CHECK(r2);
    }
    swap_swarm(c);
  }
  c->sorted = true;
  c->keyed = c->key_ordered = false;
  return PICGPU_OK;
}

static int do_scatter_P(picgpu_ctx *c) {
  PG_CHECK(c->pavg.ensure(c->rsize * (size_t)c->ncell));
  int r = dispatch_real(c, [&](auto *tag) -> int {
    using R = std::remove_pointer_t<decltype(tag)>;
    PG_LAUNCH(c, "scatter_P", nblocks(c->ncell), kBlock, k_scatter_P<R>, c->ncell, c->cell_start.as<int32_t>(),
              c->sw.pid.as<int64_t>(), c->sw.P.as<R>(), c->pavg.as<R>());
    return PICGPU_OK;
  });
  PG_CHECK(r);
  c->have_pavg = true;
  return PICGPU_OK;
}

static int64_t nface(picgpu_ctx *c, int dir) {
  return (int64_t)(c->IM + (dir == 0)) * (c->JM + (dir == 1)) * (c->KM + (dir == 2));
}

static int do_deposit_faces(picgpu_ctx *c) {
  for (int dir = 0; dir < 3; ++dir) PG_CHECK(c->face[dir].ensure(c->rsize * 4 * (size_t)nface(c, dir)));
  int r = dispatch_real(c, [&](auto *tag) -> int {
    using R = std::remove_pointer_t<decltype(tag)>;
    for (int dir = 0; dir < 3; ++dir) {
      PG_LAUNCH(c, "deposit_faces", nblocks(nface(c, dir)), kBlock, k_deposit_faces<R>, dir, c->IM, c->JM, c->KM,
                c->cell_start.as<int32_t>(), c->sw.pid.as<int64_t>(), c->sw.w.as<R>(), c->sw.vel.as<R>(),
                c->face[dir].as<R>());
    }
    return PICGPU_OK;
  });
  PG_CHECK(r);
  c->have_faces = true;
  return PICGPU_OK;
}

static int do_rk2(picgpu_ctx *c, double dt) {
  const int64_t n = c->n;
  PG_CHECK(c->pos0.ensure(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1)));
  PG_CHECK(c->vel0.ensure(c->rsize * 3 * (size_t)(n > 0 ? n : 1)));
  PG_CHECK(do_reset_advect(c, 0.0, 1, 0));
  if (n > 0) {
    int r = dispatch_real(c, [&](auto *tag) -> int {
      using R = std::remove_pointer_t<decltype(tag)>;
      PG_LAUNCH(c, "rk2_predict", nblocks(n), kBlock, k_rk2_predict<R>, n, c->sw.pos.as<double>(), c->sw.vel.as<R>(),
                c->sw.status.as<int32_t>(), c->pos0.as<double>(), c->vel0.as<R>(), dt);
      return PICGPU_OK;
    });
    PG_CHECK(r);
  }
  PG_CHECK(do_locate(c, 1));
  const int64_t pred_iters = c->stats.walk_iterations, pred_fallbacks = c->stats.fp64_fallbacks;
  PG_CHECK(do_interp(c));
  if (n > 0) {
    int r = dispatch_real(c, [&](auto *tag) -> int {
      using R = std::remove_pointer_t<decltype(tag)>;
      PG_LAUNCH(c, "rk2_correct", nblocks(n), kBlock, k_rk2_correct<R>, n, c->sw.pos.as<double>(), c->sw.vel.as<R>(),
                c->sw.status.as<int32_t>(), c->pos0.as<double>(), c->vel0.as<R>(), 0.5 * dt);
      return PICGPU_OK;
    });
    PG_CHECK(r);
  }
  PG_CHECK(do_locate(c, 1));
  // a step's walk statistics cover both of its walks, as the fused step's do
  c->stats.walk_iterations += pred_iters;
  c->stats.fp64_fallbacks += pred_fallbacks;
  return PICGPU_OK;
}

// ---------------------------------------------------------------------------
// fused step: one pass per particle, one workgroup per 8x8x4 tile of cells,
// particles visited through an index by tile-major cell key; each step
// builds the next index from the slots its particles took, and deposition
// reads that index.  The swarm is compacted into key order every
// kCompactEvery steps and at the end of picgpu_step.
// ---------------------------------------------------------------------------
static constexpr int kCompactEvery = 16;

static TileMap tile_map(const picgpu_ctx *c) {
  return TileMap{(c->IM + TX - 1) / TX, (c->JM + TY - 1) / TY, (c->KM + TZ - 1) / TZ};
}
static int64_t ntiles(const TileMap &t) { return (int64_t)t.NTI * t.NTJ * t.NTK; }

// Index of the current swarm in kidx[kcur] / kstart[kcur]: live particles by
// key, the live ones without a cell in the last bucket, removed ones left out.
// With guess, a particle without a cell is keyed by its guessed start cell
// instead (the order of a fresh locate; not an index of the swarm).
static int build_key_index(picgpu_ctx *c, const int32_t *guess) {
  const TileMap tm = tile_map(c);
  const int64_t nkeys = ntiles(tm) * TS;
  const int64_t n = c->n;
  const int cur = c->kcur;
  const size_t m = (size_t)(n > 0 ? n : 1);
  PG_CHECK(c->kcount.ensure(sizeof(int32_t) * (size_t)(nkeys + 2)));
  PG_CHECK(c->kstart[cur].ensure(sizeof(int32_t) * (size_t)(nkeys + 2)));
  PG_CHECK(c->kidx[cur].ensure(sizeof(int32_t) * m));
  PG_CHECK(c->kslot.ensure(sizeof(int32_t) * m));
  PG_HIP(hipMemsetAsync(c->kcount.ptr, 0, sizeof(int32_t) * (size_t)(nkeys + 1), c->stream));
  if (n > 0)
    PG_LAUNCH(c, "key_histogram", nblocks(n), kBlock, k_key_histogram, n, c->sw.cell.as<int32_t>(),
              c->sw.status.as<int32_t>(), guess, c->IM, c->JM, tm, (int32_t)nkeys, c->kcount.as<int32_t>(),
              c->kslot.as<int32_t>());
  PG_CHECK(scan_exclusive_i32(c, c->kcount.as<int32_t>(), c->kstart[cur].as<int32_t>(), nkeys + 1));
  if (n > 0)
    PG_LAUNCH(c, "key_scatter", nblocks(n), kBlock, k_key_scatter, n, c->sw.cell.as<int32_t>(),
              c->sw.status.as<int32_t>(), guess, c->kslot.as<int32_t>(), c->IM, c->JM, tm, (int32_t)nkeys,
              c->kstart[cur].as<int32_t>(), c->kidx[cur].as<int32_t>());
  int32_t h[2] = {0, 0};
  PG_HIP(hipMemcpyAsync(h, c->kstart[cur].as<int32_t>() + nkeys, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  PG_HIP(hipStreamSynchronize(c->stream));
  c->kbegin_unloc = h[0];
  c->kend_unloc = h[1];
  c->n_live = h[1];
  c->keyed = guess == nullptr;
  return PICGPU_OK;
}

// Moves the live particles into key order (the order of kidx[kcur]); the
// index becomes the identity.  live_dev: the live count when it is still on
// the device (live is then an upper bound, and c->n is set by the caller).
static int compact_by_key(picgpu_ctx *c, int64_t live, int mode, const int32_t *live_dev = nullptr,
                          bool move_vw = true) {
  PG_CHECK(ensure_swarm(c, c->alt, c->n));
  int32_t *idx = c->kidx[c->kcur].as<int32_t>();
  const bool move32 = c->pid32_valid;  // the 32-bit PID copy moves with the swarm
  if (live > 0) {
    int r = dispatch_real(c, [&](auto *tag) -> int {
      using R = std::remove_pointer_t<decltype(tag)>;
      auto gather = [&](auto vw_tag) -> int {
        PG_LAUNCH(c, "gather_swarm", nblocks(live, kBlock * kGatherPer), kBlock,
                  (k_gather_swarm<R, decltype(vw_tag)::value>), live, live_dev, mode, idx, c->sw.pos.as<double>(),
                  c->sw.vel.as<R>(), c->sw.w.as<R>(), c->sw.cell.as<int32_t>(), c->sw.status.as<int32_t>(),
                  c->sw.pid.as<int64_t>(), c->sw.P.as<R>(), c->alt.pos.as<double>(), c->alt.vel.as<R>(),
                  c->alt.w.as<R>(), c->alt.cell.as<int32_t>(), c->alt.status.as<int32_t>(), c->alt.pid.as<int64_t>(),
                  c->alt.P.as<R>(), move32 ? c->sw.pid32.as<int32_t>() : (const int32_t *)nullptr,
                  c->alt.pid32.as<int32_t>());
        return PICGPU_OK;
      };
      return move_vw ? gather(std::true_type{}) : gather(std::false_type{});
    });
    PG_CHECK(r);
    PG_LAUNCH(c, "iota", nblocks(live), kBlock, k_iota, live, live_dev, idx);
  }
  swap_swarm(c);
  c->pid32_valid = move32;
  if (!live_dev) c->n = live;
  c->key_ordered = true;
  return PICGPU_OK;
}

int picgpu::settle_swarm(picgpu_ctx *c) {
  if (c->swarm_invalid) {
    set_error("the particles are invalid after a step that met a degenerate face: upload, initialise or seed "
              "particles first");
    return PICGPU_E_STATE;
  }
  if (!c->compact_pending) return PICGPU_OK;
  c->compact_pending = false;
  PG_CHECK(compact_by_key(c, c->n_live, GATHER_LAST));
  c->steps_since_compact = 0;
  return PICGPU_OK;
}

int64_t picgpu::live_particles(const picgpu_ctx *c) { return c->compact_pending ? c->n_live : c->n; }

// Face arrays and the seam halves of the tiled face deposition.
static int face_out(picgpu_ctx *c, void *fo_out) {
  const TileMap tm = tile_map(c);
  const int64_t ni = (int64_t)(tm.NTI - 1) * c->JM * c->KM, nj = (int64_t)(tm.NTJ - 1) * c->IM * c->KM,
                nk = (int64_t)(tm.NTK - 1) * c->IM * c->JM;
  for (int dir = 0; dir < 3; ++dir) PG_CHECK(c->face[dir].ensure(c->rsize * 4 * (size_t)nface(c, dir)));
  PG_CHECK(c->seams.ensure(c->rsize * 4 * 2 * (size_t)(ni + nj + nk + 1)));
  return dispatch_real(c, [&](auto *tag) -> int {
    using R = std::remove_pointer_t<decltype(tag)>;
    FaceOut<R> &fo = *static_cast<FaceOut<R> *>(fo_out);
    R *base = c->seams.as<R>();
    const int64_t cnt[3] = {ni, nj, nk};
    for (int d = 0; d < 3; ++d) {
      fo.face[d] = c->face[d].as<R>();
      fo.lo[d] = base;
      fo.hi[d] = base + 4 * cnt[d];
      base += 8 * cnt[d];
    }
    return PICGPU_OK;
  });
}

static int launch_face_seams(picgpu_ctx *c, const void *fo_in) {
  const TileMap tm = tile_map(c);
  const int64_t ni = (int64_t)(tm.NTI - 1) * c->JM * c->KM, nj = (int64_t)(tm.NTJ - 1) * c->IM * c->KM,
                nk = (int64_t)(tm.NTK - 1) * c->IM * c->JM;
  if (ni + nj + nk == 0) return PICGPU_OK;
  return dispatch_real(c, [&](auto *tag) -> int {
    using R = std::remove_pointer_t<decltype(tag)>;
    const FaceOut<R> &fo = *static_cast<const FaceOut<R> *>(fo_in);
    PG_LAUNCH(c, "face_seams", nblocks(ni + nj + nk), kBlock, k_face_seams<R>, c->IM, c->JM, c->KM, (int32_t)ni,
              (int32_t)nj, (int32_t)nk, fo);
    return PICGPU_OK;
  });
}

template <typename R>
static void fill_step_args(picgpu_ctx *c, StepArgs<R> &a, double dt, unsigned long long *counters) {
  a.g = grid_view(c);
  a.pos = c->sw.pos.as<double>();
  a.vel = c->sw.vel.as<R>();
  a.w = c->sw.w.as<R>();
  a.cell = c->sw.cell.as<int32_t>();
  a.status = c->sw.status.as<int32_t>();
  a.corner = c->corner.as<R>();
  a.planes = c->planes.ptr;
  a.dt = dt;
  a.max_traversal = c->opt.max_traversal;
  a.thresh_scale = c->opt.distance_threshold;
  a.clamp = (R)c->opt.weight_clamp;
  a.bx0 = c->bmin[0];
  a.by0 = c->bmin[1];
  a.bz0 = c->bmin[2];
  a.bx1 = c->bmax[0];
  a.by1 = c->bmax[1];
  a.bz1 = c->bmax[2];
  a.tm = tile_map(c);
  a.count_next = c->kcount.as<int32_t>();
  a.slot = c->kslot.as<int32_t>();
  a.counters = counters;
  a.walk_hist = nullptr;
  a.all_located = 1;
  a.u0_recompute = 0;
  a.deposit = 0;
  a.defer_list = nullptr;
  a.defer_count = nullptr;
  for (int d = 0; d < 3; ++d) a.fo.face[d] = a.fo.lo[d] = a.fo.hi[d] = nullptr;
}

// Steps of one picgpu_step call are queued without waiting for the device:
// each step's statistics go to its own block of counters (kCounterStripes x
// C_NUM words), and the live count that sizes the next step's launches stays
// on the device (the total of the index, kstart[nkeys + 1]); the host reads
// the counters once per kStepBatch steps.
static constexpr int kStepBatch = 64;
static constexpr size_t kStepCounterWords = (size_t)kCounterStripes * C_NUM;

// The key index of the previous fused step stays valid until the swarm
// changes (every entry point that changes cells or statuses clears keyed), so
// a call that follows a step call starts from it.  Every particle of that
// index is ACTIVE_AND_LOCATED; if the previous call ended between
// compactions, the arrays still hold its removed particles (compact_pending)
// and the compaction schedule (every kCompactEvery steps) carries on.
//
// step: the step's number in the call; the first reads each particle's start
// velocity from vel, later ones recompute it (unless the field is prepared
// every step) and deposit the start state onto the faces, which is the
// previous step's deposition (finish_step_call deposits the last one).
// last: the call's last step, which compacts with weights and statuses if the
// schedule falls on it.  counters: this step's counter block.  *compacted:
// set when this step compacted the swarm.
static int do_step_fused(picgpu_ctx *c, double dt, unsigned flags, int step, bool last,
                         unsigned long long *counters, bool *compacted, int64_t *entries) {
  const bool rk2 = (flags & PICGPU_STEP_RK2) != 0;
  const bool wu = (flags & PICGPU_STEP_FIELDS) != 0;
  const bool faces = (flags & PICGPU_STEP_DEPOSIT_FACES) != 0;
  *compacted = false;
  PG_CHECK(ensure_planes(c));
  const bool fresh = !c->keyed;  // index built here: statuses still to be read
  if (fresh) {
    PG_CHECK(build_key_index(c));
    // a swarm in arbitrary order (fresh upload, stage-wise calls) is moved
    // into key order first, so that the tiles read their particles coalesced
    if (!c->key_ordered) PG_CHECK(compact_by_key(c, c->kend_unloc, GATHER_ALL));
  }
  const TileMap tm = tile_map(c);
  const int64_t nt = ntiles(tm), nkeys = nt * TS;
  const int64_t n = c->n;  // array length: an upper bound of the index entries
  // entries of the current index, when the host knows them (a fresh index,
  // or the first step of a batch)
  *entries = fresh ? c->kend_unloc : c->n_live;
  const int cur = c->kcur, nxt = cur ^ 1;
  PG_CHECK(c->kstart[nxt].ensure(sizeof(int32_t) * (size_t)(nkeys + 2)));
  PG_CHECK(c->kidx[nxt].ensure(sizeof(int32_t) * (size_t)(n > 0 ? n : 1)));
  PG_HIP(hipMemsetAsync(c->kcount.ptr, 0, sizeof(int32_t) * (size_t)(nkeys + 1), c->stream));
  // defer list [n] and its length
  PG_CHECK(c->defer.ensure(sizeof(int32_t) * (size_t)((n > 0 ? n : 1) + 1)));
  PG_HIP(hipMemsetAsync(c->defer.as<int32_t>() + (n > 0 ? n : 1), 0, sizeof(int32_t), c->stream));
  unsigned long long *hist = hist_reset(c);
  const int32_t *idx = c->kidx[cur].as<int32_t>();
  const int32_t *kstart = c->kstart[cur].as<int32_t>();
  const int32_t *m_dev = kstart + nkeys + 1;  // entries of the current index
  PG_CHECK(c->kbase.ensure(sizeof(int32_t) * (size_t)(nkeys > 0 ? nkeys : 1)));
  int32_t *kbase = c->kbase.as<int32_t>();
  const bool dep = !wu && faces && step > 0;
  alignas(16) unsigned char fo_buf[sizeof(FaceOut<double>)];
  if (dep) PG_CHECK(face_out(c, fo_buf));
  // an empty swarm still deposits (zero faces) when the step deposits
  if (n > 0 || dep) {
    int r = dispatch_real(c, [&](auto *tag) -> int {
      using R = std::remove_pointer_t<decltype(tag)>;
      StepArgs<R> a;
      fill_step_args<R>(c, a, dt, counters);
      a.walk_hist = hist;
      a.all_located = fresh ? 0 : 1;
      a.u0_recompute = !wu && step > 0;
      a.deposit = dep ? 1 : 0;
      if (dep) a.fo = *reinterpret_cast<const FaceOut<R> *>(fo_buf);
      a.defer_list = c->defer.as<int32_t>();
      a.defer_count = c->defer.as<int32_t>() + (n > 0 ? n : 1);
      constexpr bool EX = sizeof(R) == sizeof(double);
      const char *name = EX ? "step_fused_fp64" : "step_fused_fp32";
      const char *name_list = EX ? "step_list_fp64" : "step_list_fp32";
      const int64_t u0 = c->kbegin_unloc, u1 = c->kend_unloc;
      // deferred particles: at most the particles of the index; the kernel
      // reads their number on the device
      const int64_t dblocks = std::min<int64_t>(1024, nblocks(n));
      auto launch = [&](auto rk2_tag, auto wu_tag) -> int {
        constexpr bool RK = decltype(rk2_tag)::value, WU = decltype(wu_tag)::value;
        PG_LAUNCH(c, name, nt, kBlock, (k_step_tiled<R, EX, RK, WU>), a, idx, kstart, kbase);
        if (fresh && u1 > u0)
          PG_LAUNCH(c, name_list, nblocks(u1 - u0), kBlock, (k_step_list<R, EX, RK>), a, idx, u0, u1);
        if (n > 0) PG_LAUNCH(c, "step_deferred", dblocks, kBlock, (k_step_deferred<R, EX, RK>), a);
        return PICGPU_OK;
      };
      using T = std::true_type;
      using F = std::false_type;
      if (rk2) return wu ? launch(T{}, T{}) : launch(T{}, F{});
      return wu ? launch(F{}, T{}) : launch(F{}, F{});
    });
    PG_CHECK(r);
  }
  if (dep) {
    PG_CHECK(launch_face_seams(c, fo_buf));
    c->have_faces = true;
  }
  PG_CHECK(scan_exclusive_i32(c, c->kcount.as<int32_t>(), c->kstart[nxt].as<int32_t>(), nkeys + 1));
  if (n > 0)
    PG_LAUNCH(c, "key_reindex", nblocks(n, kBlock * kReindexPer), kBlock, k_key_reindex, n, m_dev, idx,
              c->sw.cell.as<int32_t>(),
              c->kslot.as<int32_t>(), kbase, CellIdx{c->IM, c->JM, grid_view(c).rI, grid_view(c).rIJ}, tm,
              c->kstart[nxt].as<int32_t>(), c->kidx[nxt].as<int32_t>());
  if (n > 0 && (flags & (PICGPU_STEP_DEPOSIT_CELLS | PICGPU_STEP_DEPOSIT_FACES))) {
    if (!c->pid32_valid) {
      // the 32-bit PID copy, made once after the swarm changed; compaction moves it
      PG_CHECK(c->pid_wide.ensure(sizeof(int32_t)));
      PG_HIP(hipMemsetAsync(c->pid_wide.ptr, 0, sizeof(int32_t), c->stream));
      PG_LAUNCH(c, "pid32", nblocks(n), kBlock, k_pid32, n, c->sw.pid.as<int64_t>(), c->sw.pid32.as<int32_t>(),
                c->pid_wide.as<int32_t>());
      c->pid32_valid = true;
    }
    PG_LAUNCH(c, "bucket_pid_order", nblocks(nkeys), kBlock, k_bucket_pid_order, nkeys, c->kstart[nxt].as<int32_t>(),
              c->kidx[nxt].as<int32_t>(), c->sw.pid.as<int64_t>(), c->sw.pid32.as<int32_t>(),
              c->pid_wide.as<int32_t>());
  }
  if (hist) PG_CHECK(hist_read(c));
  // every survivor of a fused step is located: the next index has no unlocated bucket
  c->kcur = nxt;
  c->kbegin_unloc = c->kend_unloc = -1;  // known after the call's counters are read
  const int32_t *nidx = c->kidx[nxt].as<int32_t>();
  const int32_t *nstart = c->kstart[nxt].as<int32_t>();
  int r = dispatch_real(c, [&](auto *tag) -> int {
    using R = std::remove_pointer_t<decltype(tag)>;
    if (flags & PICGPU_STEP_DEPOSIT_CELLS) {
      PG_CHECK(c->pavg.ensure(sizeof(R) * (size_t)c->ncell));
      PG_LAUNCH(c, "scatter_P", nblocks(c->ncell), kBlock, k_scatter_P_keyed<R>, c->IM, c->JM, c->KM, tm, nstart,
                nidx, c->sw.P.as<R>(), c->pavg.as<R>());
      c->have_pavg = true;
    }
    if (wu && faces) {
      // weights and velocities were written: the step's deposition is a pass of its own
      for (int dir = 0; dir < 3; ++dir) PG_CHECK(c->face[dir].ensure(sizeof(R) * 4 * (size_t)nface(c, dir)));
      const TileMap xm{c->IM / TX + 1, c->JM / TY + 1, c->KM / TZ + 1};  // tiles over [0..IM]x[0..JM]x[0..KM]
      PG_LAUNCH(c, "deposit_faces", ntiles(xm), kBlock, k_deposit_tiles<R>, c->IM, c->JM, c->KM, tm, xm, nstart, nidx,
                c->sw.w.as<R>(), c->sw.vel.as<R>(), c->face[0].as<R>(), c->face[1].as<R>(), c->face[2].as<R>());
      c->have_faces = true;
    }
    return PICGPU_OK;
  });
  PG_CHECK(r);
  c->steps_since_compact++;
  c->compact_pending = true;
  if (c->steps_since_compact >= kCompactEvery) {
    // every survivor of a fused step is located; at the call's last step the
    // weights and statuses move too; the live count is the new index's total
    // the FP64 steps recompute every start state: between the steps of a
    // call (and before the call's last pass writes them) weights and
    // velocities are not read, unless the per-step field path carries them
    const bool move_vw = wu || c->opt.precision == PICGPU_FP32;
    PG_CHECK(compact_by_key(c, n, last ? GATHER_LAST : GATHER_STEP, nstart + nkeys + 1, move_vw));
    c->steps_since_compact = 0;
    c->compact_pending = false;
    *compacted = true;
  }
  c->sorted = false;  // key order, not cell order
  return PICGPU_OK;
}

// After the last step of a call (weights and velocities not kept between
// steps): every particle's weights and velocity at its final position, and
// the last step's face deposition.
static int finish_step_call(picgpu_ctx *c, unsigned flags, unsigned long long *counters) {
  const bool faces = (flags & PICGPU_STEP_DEPOSIT_FACES) && !(flags & PICGPU_STEP_FACES_OF_START);
  const TileMap tm = tile_map(c);
  const int64_t nt = ntiles(tm);
  alignas(16) unsigned char fo_buf[sizeof(FaceOut<double>)];
  if (faces) PG_CHECK(face_out(c, fo_buf));
  int r = dispatch_real(c, [&](auto *tag) -> int {
    using R = std::remove_pointer_t<decltype(tag)>;
    StepArgs<R> a;
    fill_step_args<R>(c, a, 0.0, counters);
    a.deposit = faces ? 1 : 0;
    if (faces) a.fo = *reinterpret_cast<const FaceOut<R> *>(fo_buf);
    constexpr bool EX = sizeof(R) == sizeof(double);
    PG_LAUNCH(c, EX ? "step_finish_fp64" : "step_finish_fp32", nt, kBlock, (k_step_finish<R, EX>), a,
              c->kidx[c->kcur].as<int32_t>(), c->kstart[c->kcur].as<int32_t>());
    return PICGPU_OK;
  });
  PG_CHECK(r);
  if (faces) {
    PG_CHECK(launch_face_seams(c, fo_buf));
    c->have_faces = true;
  }
  return PICGPU_OK;
}

// Fused steps s0 .. s0 + nb - 1 of a call of nsteps: queued, then one read of
// their counters, which settles the live count, the array length after a
// compaction and the statistics.
static int step_batch(picgpu_ctx *c, double dt, unsigned flags, int s0, int nb, int nsteps) {
  const bool fields = (flags & PICGPU_STEP_FIELDS) != 0;
  PG_CHECK(c->step_counters.ensure(sizeof(unsigned long long) * kStepCounterWords * kStepBatch));
  PG_HIP(hipMemsetAsync(c->step_counters.ptr, 0, sizeof(unsigned long long) * kStepCounterWords * (size_t)nb,
                        c->stream));
  unsigned long long *ctr = c->step_counters.as<unsigned long long>();
  int64_t m0 = -1;  // entries of the index the batch starts from
  int compacted_at = -1;
  for (int b = 0; b < nb; ++b) {
    const int s = s0 + b;
    if (fields) PG_CHECK(eulerian_fields_async(c));
    PG_CHECK(do_center_to_corner(c));
    bool compacted = false;
    int64_t entries = 0;
    PG_CHECK(do_step_fused(c, dt, flags, s, s == nsteps - 1, ctr + (size_t)b * kStepCounterWords, &compacted,
                           &entries));
    if (b == 0) m0 = entries;
    if (compacted) compacted_at = b;
  }
  if (!fields && s0 + nb == nsteps) PG_CHECK(finish_step_call(c, flags, ctr + (size_t)(nb - 1) * kStepCounterWords));
  std::vector<unsigned long long> h(kStepCounterWords * (size_t)nb);
  PG_HIP(hipMemcpyAsync(h.data(), ctr, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost, c->stream));
  PG_HIP(hipStreamSynchronize(c->stream));
  auto sum = [&](int b, int k) {
    unsigned long long v = 0;
    for (int st = 0; st < kCounterStripes; ++st) v += h[(size_t)b * kStepCounterWords + (size_t)st * C_NUM + k];
    return (int64_t)v;
  };
  bool err = false;
  int64_t m = m0;
  for (int b = 0; b < nb; ++b) {
    c->stats.particle_steps += m;
    m = sum(b, C_LOCATED);
    err = err || sum(b, C_ERROR) != 0;
  }
  const int l = nb - 1;
  c->stats.located = m;
  c->stats.removed_lost = sum(l, C_REMOVED_LOST);
  c->stats.removed_out_of_bounds = sum(l, C_REMOVED_OOB);
  c->stats.lost = c->stats.removed_lost;
  c->stats.walk_iterations = sum(l, C_WALK_ITERS);
  c->stats.fp64_fallbacks = sum(l, C_FALLBACKS);
  if (err) {
    // a step stopped part-way: some particles carry that step's results, the
    // others the state it started from; later steps ran on that mix
    c->swarm_invalid = true;
    c->keyed = c->key_ordered = c->compact_pending = false;
    set_error("degenerate cell face (|n| < 1e-12) met during the walk");
    return PICGPU_E_DEGENERATE;
  }
  c->n_live = m;
  c->kbegin_unloc = c->kend_unloc = m;
  if (compacted_at >= 0) c->n = sum(compacted_at, C_LOCATED);
  return PICGPU_OK;
}

// ---------------------------------------------------------------------------
// stage-wise C-ABI
// ---------------------------------------------------------------------------
extern "C" {

int picgpu_reset_statuses(picgpu_ctx *c) {
  if (!c) return PICGPU_E_ARG;
  PG_HIP(hipSetDevice(c->device));
  return do_reset_advect(c, 0.0, 1, 0);
}

int picgpu_advect(picgpu_ctx *c, double dt) {
  if (!c) return PICGPU_E_ARG;
  PG_HIP(hipSetDevice(c->device));
  return do_reset_advect(c, dt, 0, 1);
}

int picgpu_locate(picgpu_ctx *c) {
  if (!c) return PICGPU_E_ARG;
  PG_HIP(hipSetDevice(c->device));
  return do_locate(c, 0);
}

int picgpu_remove_lost(picgpu_ctx *c, int64_t *removed) {
  if (!c) return PICGPU_E_ARG;
  PG_HIP(hipSetDevice(c->device));
  int64_t rl = 0;
  PG_CHECK(do_remove(c, 1, 0, &rl, nullptr));
  c->stats.removed_lost = rl;
  if (removed) *removed = rl;
  return PICGPU_OK;
}

int picgpu_remove_out_of_bounds(picgpu_ctx *c, int64_t *removed) {
  if (!c) return PICGPU_E_ARG;
  PG_HIP(hipSetDevice(c->device));
  PG_CHECK(need_grid(c));
  int64_t ro = 0;
  PG_CHECK(do_remove(c, 0, 1, nullptr, &ro));
  c->stats.removed_out_of_bounds = ro;
  if (removed) *removed = ro;
  return PICGPU_OK;
}

int picgpu_interpolate(picgpu_ctx *c) {
  if (!c) return PICGPU_E_ARG;
  PG_HIP(hipSetDevice(c->device));
  PG_CHECK(need_grid(c));
  PG_CHECK(do_center_to_corner(c));
  PG_CHECK(do_interp(c));
  PG_HIP(hipStreamSynchronize(c->stream));
  return PICGPU_OK;
}

int picgpu_scatter(picgpu_ctx *c) {
  if (!c) return PICGPU_E_ARG;
  PG_HIP(hipSetDevice(c->device));
  PG_CHECK(need_grid(c));
  PG_CHECK(do_sort(c));
  PG_CHECK(do_scatter_P(c));
  PG_HIP(hipStreamSynchronize(c->stream));
  return PICGPU_OK;
}

int picgpu_deposit_faces(picgpu_ctx *c, double particle_volume) {
  if (!c) return PICGPU_E_ARG;
  PG_HIP(hipSetDevice(c->device));
  PG_CHECK(need_grid(c));
  if (!c->sorted) PG_CHECK(do_sort(c));
  c->particle_volume = particle_volume;
  PG_CHECK(do_deposit_faces(c));
  PG_HIP(hipStreamSynchronize(c->stream));
  return PICGPU_OK;
}

int picgpu_set_particle_volume(picgpu_ctx *c, double particle_volume) {
  if (!c || !(particle_volume > 0.0 && particle_volume < HUGE_VAL)) {
    set_error("particle volume must be positive and finite");
    return PICGPU_E_ARG;
  }
  c->particle_volume = particle_volume;
  return PICGPU_OK;
}

int picgpu_step(picgpu_ctx *c, double dt, int nsteps, unsigned flags) {
  if (!c || nsteps < 0) return PICGPU_E_ARG;
  PG_HIP(hipSetDevice(c->device));
  PG_CHECK(need_grid(c));
  if (!c->have_ucat) {
    set_error("ucat not uploaded");
    return PICGPU_E_STATE;
  }
  if (c->swarm_invalid) return settle_swarm(c);  // reports that the particles are invalid
  if (flags & PICGPU_STEP_STAGED) PG_CHECK(settle_swarm(c));
  if (!c->compact_pending) c->steps_since_compact = 0;
  const bool fields = (flags & PICGPU_STEP_FIELDS) != 0;
  if (fields) PG_CHECK(eulerian_fields_begin(c));
  if (!(flags & PICGPU_STEP_STAGED)) {
    for (int s0 = 0; s0 < nsteps; s0 += kStepBatch)
      PG_CHECK(step_batch(c, dt, flags, s0, std::min(kStepBatch, nsteps - s0), nsteps));
  }
  for (int s = 0; s < nsteps && (flags & PICGPU_STEP_STAGED); ++s) {
    if (fields) PG_CHECK(eulerian_fields_async(c));
    c->stats.particle_steps += c->n;
    PG_CHECK(do_center_to_corner(c));
    if (flags & PICGPU_STEP_RK2) {
      PG_CHECK(do_rk2(c, dt));
    } else {
      PG_CHECK(do_reset_advect(c, dt, 1, 1));
      PG_CHECK(do_locate(c, 0));
    }
    int64_t rl = 0, ro = 0;
    PG_CHECK(do_remove(c, 1, 1, &rl, &ro));
    c->stats.removed_lost = rl;
    c->stats.removed_out_of_bounds = ro;
    PG_CHECK(do_interp(c));
    PG_CHECK(do_sort(c));
    if (flags & PICGPU_STEP_DEPOSIT_CELLS) PG_CHECK(do_scatter_P(c));
    if (flags & PICGPU_STEP_DEPOSIT_FACES) PG_CHECK(do_deposit_faces(c));
  }
  if (fields) PG_CHECK(eulerian_fields_end(c));
  PG_HIP(hipStreamSynchronize(c->stream));
  return PICGPU_OK;
}

// ---------------------------------------------------------------------------
// grid outputs
// ---------------------------------------------------------------------------
int picgpu_grid_field_size(picgpu_ctx *c, int which, int64_t *count) {
  if (!c || !count) return PICGPU_E_ARG;
  PG_CHECK(need_grid(c));
  switch (which) {
    case PICGPU_FIELD_PARTICLE_COUNT:
    case PICGPU_FIELD_P_AVERAGE: *count = c->nnode; break;
    case PICGPU_FIELD_CORNER_VELOCITY: *count = 3 * c->nnode; break;
    case PICGPU_FIELD_IFACE:
    case PICGPU_FIELD_JFACE:
    case PICGPU_FIELD_KFACE: *count = 4 * nface(c, which - PICGPU_FIELD_IFACE); break;
    case PICGPU_FIELD_IFACE_VOLFRAC:
    case PICGPU_FIELD_JFACE_VOLFRAC:
    case PICGPU_FIELD_KFACE_VOLFRAC: *count = nface(c, which - PICGPU_FIELD_IFACE_VOLFRAC); break;
    default: {
      const double *p = nullptr;
      return eulerian_field(c, which, &p, count);
    }
  }
  return PICGPU_OK;
}

// The per-cell counts in cell_count: from the stage-wise deposition, or from
// the bucket sizes of the key index after a fused step.
static int ensure_cell_counts(picgpu_ctx *c) {
  if (!c->sorted && !c->keyed) {
    set_error("ParticleCount needs picgpu_scatter/picgpu_deposit_faces/picgpu_step first");
    return PICGPU_E_STATE;
  }
  if (!c->sorted) {
    PG_CHECK(c->cell_count.ensure(sizeof(int32_t) * (size_t)(c->ncell + 1)));
    PG_LAUNCH(c, "count_from_keys", nblocks(c->ncell), kBlock, k_count_from_keys, c->IM, c->JM, c->KM,
              tile_map(c), c->kstart[c->kcur].as<int32_t>(), c->cell_count.as<int32_t>());
  }
  return PICGPU_OK;
}

int picgpu_particle_count_summary(picgpu_ctx *c, double *total, double *max_count, int64_t *max_index) {
  if (!c || !total || !max_count || !max_index) return PICGPU_E_ARG;
  PG_HIP(hipSetDevice(c->device));
  PG_CHECK(need_grid(c));
  PG_CHECK(ensure_cell_counts(c));
  PG_CHECK(c->count_summary.ensure(2 * sizeof(unsigned long long)));
  unsigned long long *d = c->count_summary.as<unsigned long long>();
  PG_HIP(hipMemsetAsync(d, 0, 2 * sizeof(unsigned long long), c->stream));
  PG_LAUNCH(c, "count_summary", std::min<int64_t>(nblocks(c->ncell), 2048), kBlock, k_count_summary, c->IM, c->JM,
            c->KM, c->cell_count.as<int32_t>(), d);
  unsigned long long h[2];
  PG_HIP(hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  PG_HIP(hipStreamSynchronize(c->stream));
  constexpr unsigned long long kSlotMask = (1ull << 33) - 1;
  *total = (double)h[0];
  *max_count = (double)(h[1] >> 33);
  *max_index = (int64_t)(kSlotMask - (h[1] & kSlotMask));
  return PICGPU_OK;
}

int picgpu_interpolation_error(picgpu_ctx *c, int finit, double *percent) {
  if (!c || !percent || (finit != 0 && finit != 1)) return PICGPU_E_ARG;
  PG_HIP(hipSetDevice(c->device));
  PG_CHECK(settle_swarm(c));
  const int64_t n = c->n;
  const int64_t nb = std::max<int64_t>(1, std::min<int64_t>(nblocks(3 * n), 1024));
  PG_CHECK(c->count_summary.ensure(sizeof(double) * 2 * (size_t)nb));
  double *part = c->count_summary.as<double>();
  int r = dispatch_real(c, [&](auto *tag) -> int {
    using R = std::remove_pointer_t<decltype(tag)>;
    PG_LAUNCH(c, "interp_error_sums", nb, kBlock, k_interp_error_sums<R>, n, finit, c->sw.pos.as<double>(),
              c->sw.vel.as<R>(), part);
    return PICGPU_OK;
  });
  PG_CHECK(r);
  std::vector<double> h(2 * (size_t)nb);
  PG_HIP(hipMemcpyAsync(h.data(), part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
  PG_HIP(hipStreamSynchronize(c->stream));
  double e = 0.0, a = 0.0;
  for (int64_t b = 0; b < nb; ++b) {
    e += h[2 * b];
    a += h[2 * b + 1];
  }
  *percent = 100.0 * (sqrt(e) / sqrt(a));  // 0/0 (no particles) gives NaN, as the reference's does
  return PICGPU_OK;
}

int picgpu_download_grid_field(picgpu_ctx *c, int which, double *out) {
  if (!c || !out) return PICGPU_E_ARG;
  PG_HIP(hipSetDevice(c->device));
  PG_CHECK(need_grid(c));
  PG_HIP(hipStreamSynchronize(c->stream));
  const int64_t nx = c->IM + 1, ny = c->JM + 1;
  if (which >= PICGPU_FIELD_UCAT && which <= PICGPU_FIELD_AJ) {
    const double *p = nullptr;
    int64_t count = 0;
    PG_CHECK(eulerian_field(c, which, &p, &count));
    PG_HIP(hipMemcpy(out, p, sizeof(double) * (size_t)count, hipMemcpyDeviceToHost));
    return PICGPU_OK;
  }
  if (which == PICGPU_FIELD_PARTICLE_COUNT) {
    PG_CHECK(ensure_cell_counts(c));
    PG_HIP(hipStreamSynchronize(c->stream));
    std::vector<int32_t> h((size_t)c->ncell);
    PG_HIP(hipMemcpy(h.data(), c->cell_count.ptr, sizeof(int32_t) * (size_t)c->ncell, hipMemcpyDeviceToHost));
    memset(out, 0, sizeof(double) * (size_t)c->nnode);
    for (int64_t k = 0; k < c->KM; ++k)
      for (int64_t j = 0; j < c->JM; ++j)
        for (int64_t i = 0; i < c->IM; ++i) out[(k * ny + j) * nx + i] = (double)h[(k * c->JM + j) * c->IM + i];
    return PICGPU_OK;
  }
  return dispatch_real(c, [&](auto *tag) -> int {
    using R = std::remove_pointer_t<decltype(tag)>;
    if (which == PICGPU_FIELD_P_AVERAGE) {
      if (!c->have_pavg) {
        set_error("P average not computed");
        return PICGPU_E_STATE;
      }
      std::vector<R> h((size_t)c->ncell);
      PG_HIP(hipMemcpy(h.data(), c->pavg.ptr, sizeof(R) * (size_t)c->ncell, hipMemcpyDeviceToHost));
      memset(out, 0, sizeof(double) * (size_t)c->nnode);
      for (int64_t k = 0; k < c->KM; ++k)
        for (int64_t j = 0; j < c->JM; ++j)
          for (int64_t i = 0; i < c->IM; ++i) out[(k * ny + j) * nx + i] = (double)h[(k * c->JM + j) * c->IM + i];
      return PICGPU_OK;
    }
    if (which == PICGPU_FIELD_CORNER_VELOCITY) {
      if (!c->have_ucat) return PICGPU_E_STATE;
      std::vector<R> h((size_t)c->nnode * 3);
      PG_HIP(hipMemcpy(h.data(), c->corner.ptr, sizeof(R) * h.size(), hipMemcpyDeviceToHost));
      for (size_t q = 0; q < h.size(); ++q) out[q] = (double)h[q];
      return PICGPU_OK;
    }
    if (which >= PICGPU_FIELD_IFACE && which <= PICGPU_FIELD_KFACE) {
      if (!c->have_faces) {
        set_error("faces not deposited");
        return PICGPU_E_STATE;
      }
      int dir = which - PICGPU_FIELD_IFACE;
      std::vector<R> h((size_t)nface(c, dir) * 4);
      PG_HIP(hipMemcpy(h.data(), c->face[dir].ptr, sizeof(R) * h.size(), hipMemcpyDeviceToHost));
      for (size_t q = 0; q < h.size(); ++q) out[q] = (double)h[q];
      return PICGPU_OK;
    }
    if (which >= PICGPU_FIELD_IFACE_VOLFRAC && which <= PICGPU_FIELD_KFACE_VOLFRAC) {
      if (!c->have_faces) {
        set_error("faces not deposited");
        return PICGPU_E_STATE;
      }
      int dir = which - PICGPU_FIELD_IFACE_VOLFRAC;
      int64_t nf = nface(c, dir);
      DevBuf tmp;
      PG_CHECK(tmp.ensure(sizeof(double) * (size_t)nf));
      {
        TimerRec _rec;
        timer_begin(c, "face_volfrac", _rec);
        hipLaunchKernelGGL(k_face_volfrac<R>, dim3(nblocks(nf)), dim3(kBlock), 0, c->stream, dir, grid_view(c),
                           c->face[dir].as<R>(), c->particle_volume, tmp.as<double>());
        timer_end(c, _rec);
      }
      hipError_t e1 = hipGetLastError();
      hipError_t e2 = hipMemcpyAsync(out, tmp.ptr, sizeof(double) * (size_t)nf, hipMemcpyDeviceToHost, c->stream);
      hipError_t e3 = hipStreamSynchronize(c->stream);
      tmp.release();
      if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
        set_error("face volume fraction failed");
        return PICGPU_E_HIP;
      }
      return PICGPU_OK;
    }
    set_error("unknown field %d", which);
    return PICGPU_E_ARG;
  });
}

// ---------------------------------------------------------------------------
// statistics and timing
// ---------------------------------------------------------------------------
int picgpu_get_stats(picgpu_ctx *c, picgpu_stats *st) {
  if (!c || !st) return PICGPU_E_ARG;
  *st = c->stats;
  st->n_particles = live_particles(c);
  return PICGPU_OK;
}

int picgpu_walk_stats_enable(picgpu_ctx *c, int on) {
  if (!c) return PICGPU_E_ARG;
  c->opt.collect_walk_stats = on ? 1 : 0;
  return PICGPU_OK;
}

int picgpu_timing_enable(picgpu_ctx *c, int on) {
  if (!c) return PICGPU_E_ARG;
  PG_CHECK(flush_timers(c));
  c->timing = on != 0;
  return PICGPU_OK;
}

int picgpu_timing_query(picgpu_ctx *c, const char *kernel, double *total_ms, int64_t *launches) {
  if (!c || !kernel) return PICGPU_E_ARG;
  PG_CHECK(flush_timers(c));
  auto it = c->timing_acc.find(kernel);
  if (total_ms) *total_ms = it == c->timing_acc.end() ? 0.0 : it->second.first;
  if (launches) *launches = it == c->timing_acc.end() ? 0 : it->second.second;
  return PICGPU_OK;
}

int picgpu_timing_reset(picgpu_ctx *c) {
  if (!c) return PICGPU_E_ARG;
  PG_CHECK(flush_timers(c));
  c->timing_acc.clear();
  return PICGPU_OK;
}

int picgpu_timing_select(picgpu_ctx *c, const char *kernel_prefix) {
  if (!c) return PICGPU_E_ARG;
  PG_CHECK(flush_timers(c));
  c->timing_only = kernel_prefix ? kernel_prefix : "";
  return PICGPU_OK;
}

}  // extern "C"
