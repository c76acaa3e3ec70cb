// scan.hip -- exclusive prefix sum of int32 on gfx950 (three-phase,
// recursive over block sums).  Used for cell offsets of the counting sort
// and for stream compaction of the swarm.
#include "picgpu_internal.hpp"

namespace picgpu {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;  // 2048

// inclusive scan across a 64-wide wavefront
__device__ __forceinline__ int32_t wave_inclusive(int32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    int32_t t = __shfl_up(v, off);
    if (lane >= off) v += t;
  }
  return v;
}

// Scans one tile per block: out = exclusive scan within the tile, sums[b] = tile total.
__global__ __launch_bounds__(kScanThreads) void k_scan_tile(const int32_t *__restrict__ in, int32_t *__restrict__ out,
                                                            int64_t n, int32_t *__restrict__ sums) {
  __shared__ int32_t wave_tot[kScanThreads / 64];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int32_t v[kScanItems];
  int32_t run = 0;
#pragma unroll
  for (int t = 0; t < kScanItems; ++t) {
    int64_t q = base + t;
    v[t] = q < n ? in[q] : 0;
    int32_t x = v[t];
    v[t] = run;
    run += x;
  }
  int32_t incl = wave_inclusive(run);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 63) wave_tot[wave] = incl;
  __syncthreads();
  int32_t woff = 0;
#pragma unroll
  for (int w = 0; w < kScanThreads / 64; ++w)
    if (w < wave) woff += wave_tot[w];
  const int32_t excl = woff + incl - run;
#pragma unroll
  for (int t = 0; t < kScanItems; ++t) {
    int64_t q = base + t;
    if (q < n) out[q] = v[t] + excl;
  }
  if (threadIdx.x == kScanThreads - 1) sums[blockIdx.x] = excl + run;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_add(int32_t *__restrict__ out, int64_t n,
                                                           const int32_t *__restrict__ offs) {
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  const int32_t o = offs[blockIdx.x];
#pragma unroll
  for (int t = 0; t < kScanItems; ++t) {
    int64_t q = base + t;
    if (q < n) out[q] += o;
  }
}

// out[0..n-1] = exclusive scan, tmp sized by caller via ctx->scan_tmp.
static int scan_rec(picgpu_ctx *c, const int32_t *in, int32_t *out, int64_t n, int32_t *tmp, int32_t *total_slot) {
  const int64_t nb = (n + kScanTile - 1) / kScanTile;
  int32_t *sums = tmp;
  int32_t *sums_scan = tmp + nb + 1;
  PG_LAUNCH(c, "scan_tile", nb, kScanThreads, k_scan_tile, in, out, n, sums);
  if (nb == 1) {
    // total = sums[0]
    PG_HIP(hipMemcpyAsync(total_slot, sums, sizeof(int32_t), hipMemcpyDeviceToDevice, c->stream));
    return PICGPU_OK;
  }
  PG_CHECK(scan_rec(c, sums, sums_scan, nb, tmp + 2 * (nb + 1), total_slot));
  PG_LAUNCH(c, "scan_add", nb, kScanThreads, k_scan_add, out, n, sums_scan);
  return PICGPU_OK;
}

int scan_exclusive_i32(picgpu_ctx *c, const int32_t *in, int32_t *out, int64_t n) {
  if (n <= 0) {
    PG_HIP(hipMemsetAsync(out, 0, sizeof(int32_t), c->stream));
    return PICGPU_OK;
  }
  // temporary: sum of level sizes (geometric) * 2
  int64_t need = 0, m = n;
  while (true) {
    int64_t nb = (m + kScanTile - 1) / kScanTile;
    need += 2 * (nb + 1);
    if (nb == 1) break;
    m = nb;
  }
  PG_CHECK(c->scan_tmp.ensure(sizeof(int32_t) * (size_t)(need + 16)));
  return scan_rec(c, in, out, n, c->scan_tmp.as<int32_t>(), out + n);
}

}  // namespace picgpu
