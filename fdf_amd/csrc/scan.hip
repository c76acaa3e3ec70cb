// scan.hip -- exclusive prefix sum of int32 on gfx950 (reduce, then scan).
// Used for the key offsets of the particle index, the cell offsets of the
// counting sort and the stream compaction of the swarm.
//
// Pass 1 sums each tile of kScanTile entries; the tile sums are scanned
// (recursively, in one workgroup once they fit one tile); pass 3 re-reads each
// tile and writes its exclusive scan from the tile's offset.  That reads the
// input twice and writes the output once: 12 bytes per entry.
#include "picgpu_internal.hpp"

namespace picgpu {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;  // 4096

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

// exclusive scan of one value per thread across the workgroup; *total = sum
__device__ __forceinline__ int32_t block_exclusive(int32_t v, int32_t *wave_tot, int32_t *total) {
  const int32_t incl = wave_inclusive(v);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 63) wave_tot[wave] = incl;
  __syncthreads();
  int32_t woff = 0, all = 0;
#pragma unroll
  for (int w = 0; w < kScanThreads / 64; ++w) {
    if (w < wave) woff += wave_tot[w];
    all += wave_tot[w];
  }
  *total = all;
  return woff + incl - v;
}

// entries [base, base + kScanItems) of thread threadIdx.x in tile b
__device__ __forceinline__ int64_t scan_base(int64_t b) {
  return b * kScanTile + (int64_t)threadIdx.x * kScanItems;
}

// pass 1: sums[b] = sum of tile b
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const int32_t *__restrict__ in, int64_t n,
                                                              int32_t *__restrict__ sums) {
  __shared__ int32_t wave_tot[kScanThreads / 64];
  const int64_t base = scan_base(blockIdx.x);
  int32_t run = 0;
  if (base + kScanItems <= n) {
#pragma unroll
    for (int t = 0; t < kScanItems; ++t) run += in[base + t];
  } else {
    for (int t = 0; t < kScanItems; ++t)
      if (base + t < n) run += in[base + t];
  }
  int32_t total;
  block_exclusive(run, wave_tot, &total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// pass 3 (and the single-tile case): out = offs[b] + exclusive scan within
// tile b; with total != null, *total = offs[b] + the tile's sum (last tile)
__global__ __launch_bounds__(kScanThreads) void k_scan_tile(const int32_t *__restrict__ in, int32_t *__restrict__ out,
                                                            int64_t n, const int32_t *__restrict__ offs,
                                                            int32_t *__restrict__ total) {
  __shared__ int32_t wave_tot[kScanThreads / 64];
  const int64_t base = scan_base(blockIdx.x);
  int32_t v[kScanItems];
  const bool full = base + kScanItems <= n;
  if (full) {
#pragma unroll
    for (int t = 0; t < kScanItems; ++t) v[t] = in[base + t];
  } else {
#pragma unroll
    for (int t = 0; t < kScanItems; ++t) v[t] = base + t < n ? in[base + t] : 0;
  }
  int32_t run = 0;
#pragma unroll
  for (int t = 0; t < kScanItems; ++t) {
    const int32_t x = v[t];
    v[t] = run;
    run += x;
  }
  int32_t tile_sum;
  const int32_t excl = block_exclusive(run, wave_tot, &tile_sum) + (offs ? offs[blockIdx.x] : 0);
  if (full) {
#pragma unroll
    for (int t = 0; t < kScanItems; ++t) out[base + t] = v[t] + excl;
  } else {
#pragma unroll
    for (int t = 0; t < kScanItems; ++t)
      if (base + t < n) out[base + t] = v[t] + excl;
  }
  if (total && threadIdx.x == 0 && blockIdx.x == gridDim.x - 1) *total = (offs ? offs[blockIdx.x] : 0) + tile_sum;
}

// out[0..n) = exclusive scan of in, *total = sum; tmp holds the tile sums of
// every level
static int scan_rec(picgpu_ctx *c, const int32_t *in, int32_t *out, int64_t n, int32_t *tmp, int32_t *total) {
  const int64_t nb = (n + kScanTile - 1) / kScanTile;
  if (nb == 1) {
    PG_LAUNCH(c, "scan_tile", 1, kScanThreads, k_scan_tile, in, out, n, (const int32_t *)nullptr, total);
    return PICGPU_OK;
  }
  int32_t *sums = tmp, *offs = tmp + nb;
  PG_LAUNCH(c, "scan_reduce", nb, kScanThreads, k_scan_reduce, in, n, sums);
  PG_CHECK(scan_rec(c, sums, offs, nb, tmp + 2 * nb, total));
  PG_LAUNCH(c, "scan_tile", nb, kScanThreads, k_scan_tile, in, out, n, (const int32_t *)offs, (int32_t *)nullptr);
  return PICGPU_OK;
}

// out[0..n] : exclusive scan of in[0..n), out[n] = total
int scan_exclusive_i32(picgpu_ctx *c, const int32_t *in, int32_t *out, int64_t n) {
  if (n <= 0) {
    PG_HIP(hipMemsetAsync(out, 0, sizeof(int32_t), c->stream));
    return PICGPU_OK;
  }
  // temporary: two arrays (sums, offsets) per level above the first
  int64_t need = 0, m = n;
  while (true) {
    const int64_t nb = (m + kScanTile - 1) / kScanTile;
    if (nb == 1) break;
    need += 2 * nb;
    m = nb;
  }
  PG_CHECK(c->scan_tmp.ensure(sizeof(int32_t) * (size_t)(need + 16)));
  return scan_rec(c, in, out, n, c->scan_tmp.as<int32_t>(), out + n);
}

}  // namespace picgpu
