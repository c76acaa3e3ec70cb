// mlp.hip -- the learned initial guess for the walking search (gfx950).
//
// A 3 -> 64 -> 64 -> 3 MLP with GELU in its hard (piecewise-linear) form,
// z Phi(z) with the normal CDF Phi replaced by its tangent at 0 clipped to
// [0, 1]: z clamp(z / sqrt(2 pi) + 1/2, 0, 1) (the construction of
// hard-sigmoid and hard-swish, Howard et al., "Searching for MobileNetV3",
// 2019, section 5.2, applied to GELU; the same function in training,
// inference and the oracle) maps normalised physical
// coordinates to the fractional logical position ((i+a1)/IM, (j+a2)/JM,
// (k+a3)/KM).  Weights are BF16, products accumulate in FP32.  In training
// layer 1 runs on the VALU (K = 3), layers 2 and 3 and the large backward
// products on MFMA (v_mfma_f32_32x32x8_bf16); inference runs every layer on
// MFMA (k_mlp_infer).  Training (forward, backward, Adam) runs on the
// GPU on (position, located cell) pairs produced by the walking search.  The
// guess only chooses where the walk starts (the slot of
// InitializeTraversalParameters' fallback start, walkingsearch.c:603-619);
// the walk itself is unchanged, so ownership stays the reference's.
//
// Rounding rule (the CPU oracle states the same): weights W1, W2, W3 and
// the matmul inputs a1 = act(z1), a2 = act(z2) and the backward g2 are
// rounded to BF16 (nearest even); x, biases, z, losses and every
// accumulation are FP32.  Biases are the first term of their sums: layer 1
// is fma(w0, x0, fma(w1, x1, fma(w2, x2, b))), and the products of layers 2
// and 3 accumulate onto the bias.
#include <math.h>
#include <string.h>

#include <vector>

#include "picgpu_internal.hpp"

namespace picgpu {

constexpr int H = PICGPU_MLP_HIDDEN;  // 64
constexpr int NP = H * 3 + H + H * H + H + 3 * H + 3;
constexpr int OFF_W1 = 0, OFF_B1 = 3 * H, OFF_W2 = OFF_B1 + H, OFF_B2 = OFF_W2 + H * H, OFF_W3 = OFF_B2 + H,
              OFF_B3 = OFF_W3 + 3 * H;
constexpr int TILE = 32;          // samples per wavefront
constexpr int MLP_THREADS = 64;   // one wavefront per block

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);  // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }

// Hard GELU z t(z), gate t = clamp(fma(z, c, 1/2), 0, 1) with
// c = fl32(1/sqrt(2 pi)) = Phi'(0): one fused multiply-add with the VOP3
// clamp and one multiply, no transcendental.  Derivative t + c z where the
// gate is strictly inside (0, 1), t (0 or 1) where it is clamped.
constexpr float kActC = 0.3989422804014327f;
__device__ __forceinline__ float act_gate(float z) {
  return __builtin_amdgcn_fmed3f(__builtin_fmaf(z, kActC, 0.5f), 0.0f, 1.0f);
}
__device__ __forceinline__ float act(float z) { return z * act_gate(z); }
__device__ __forceinline__ float act_d(float z) {
  const float u = __builtin_fmaf(z, kActC, 0.5f);
  return (u > 0.0f && u < 1.0f) ? __builtin_fmaf(kActC, z, u) : act_gate(z);
}

// C[32x32] += A[32xK] * B[Kx32]; A(m,k) = A[m*sAm + k*sAk], B(k,n) = B[k*sBk + n*sBn]
// (bf16 in LDS or global).  Operand layout of v_mfma_f32_32x32x8_bf16: lane l
// holds A(l%32, 4*(l/32)+v) and B(4*(l/32)+v, l%32), v = 0..3; the result
// register r of lane l is C(8*(r/4) + 4*(l/32) + r%4, l%32).
__device__ __forceinline__ void mm32(const uint16_t *A, int sAm, int sAk, const uint16_t *B, int sBk, int sBn, int K,
                                     f32x16 &acc) {
  const int l = threadIdx.x & 63;
  const int r = l & 31, h = l >> 5;
  for (int k0 = 0; k0 < K; k0 += 8) {
    bf16x4 a, b;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int k = k0 + 4 * h + v;
      a[v] = (short)A[r * sAm + k * sAk];
      b[v] = (short)B[k * sBk + r * sBn];
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
  }
}
__device__ __forceinline__ int res_row(int reg) { return 8 * (reg / 4) + 4 * ((threadIdx.x & 63) >> 5) + (reg % 4); }
__device__ __forceinline__ int res_col() { return threadIdx.x & 31; }

struct MlpWeights {
  const float *p;      // FP32 master parameters (biases are read from here)
  const uint16_t *wb;  // BF16 W1 [64][3], W2 [64][64], W3 [3][64] (same offsets as the master)
};

struct TileLds {
  float x[TILE][3];
  float z1[TILE][H];
  float z2[TILE][H];
  uint16_t a1[TILE][H];
  uint16_t a2[TILE][H];
  uint16_t g2[TILE][H];
  float y[TILE][3];
  float g3[TILE][3];
  float g1[TILE][H];
};

// Forward of one 32-sample tile held in L.x; fills z1, a1, z2, a2, y.
__device__ void tile_forward(const MlpWeights &Wt, TileLds &L) {
  const int l = threadIdx.x & 63;
  // layer 1 (VALU): lane l owns hidden unit q = l for all 32 samples
  {
    const int q = l;
    const float w0 = bf2f(Wt.wb[OFF_W1 + 3 * q]), w1 = bf2f(Wt.wb[OFF_W1 + 3 * q + 1]),
                w2 = bf2f(Wt.wb[OFF_W1 + 3 * q + 2]), b = Wt.p[OFF_B1 + q];
    for (int m = 0; m < TILE; ++m) {
      float z = __builtin_fmaf(w0, L.x[m][0], __builtin_fmaf(w1, L.x[m][1], __builtin_fmaf(w2, L.x[m][2], b)));
      L.z1[m][q] = z;
      L.a1[m][q] = f2bf(act(z));
    }
  }
  __syncthreads();
  // layer 2 (MFMA): Z2 = A1 (32x64) * W2^T, two 32-column tiles
  for (int t = 0; t < 2; ++t) {
    const int n = 32 * t + res_col();
    const float b = Wt.p[OFF_B2 + n];
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = b;  // the bias is the first term of each sum
    mm32(&L.a1[0][0], H, 1, Wt.wb + OFF_W2 + 32 * t * H, 1, H, H, acc);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int m = res_row(r);
      const float z = acc[r];
      L.z2[m][n] = z;
      L.a2[m][n] = f2bf(act(z));
    }
  }
  __syncthreads();
  // layer 3 (MFMA): Y = A2 (32x64) * W3^T, columns 0..2 of one tile
  {
    const int l3 = res_col();
    const int h = l >> 5;
    const float b3 = l3 < 3 ? Wt.p[OFF_B3 + l3] : 0.0f;
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = b3;
    for (int k0 = 0; k0 < H; k0 += 8) {
      bf16x4 a, b;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int k = k0 + 4 * h + v;
        a[v] = (short)L.a2[l3][k];
        b[v] = l3 < 3 ? (short)Wt.wb[OFF_W3 + l3 * H + k] : (short)0;
      }
      acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
    }
    if (l3 < 3) {
#pragma unroll
      for (int r = 0; r < 16; ++r) L.y[res_row(r)][l3] = acc[r];
    }
  }
  __syncthreads();
}

// ---------------------------------------------------------------- inference
// normalised input of a physical position
struct MlpNorm {
  double lo0, lo1, lo2, s0, s1, s2;
};
__device__ __forceinline__ float norm_axis(double x, double lo, double s) { return (float)((x - lo) * s - 1.0); }

// Inference of 32 samples per wavefront, every layer on MFMA; W1 and W2 in
// registers, W3 and the layer-2 biases in LDS.  The layers run transposed (Z1^T = W1 X^T, Z2^T = W2 A1^T,
// Y^T = W3 A2^T): lane l works on sample n = l%32 with h = l/32, and the
// result register r of a 32-unit column tile t is hidden unit
// 32t + 8(r/4) + 4h + r%4 of that sample, so a layer's results are already
// in B-operand registers of the next product.
//  - Layer 1: v_mfma_f32_32x32x2_f32 (FP32 operands, K = 2): two products per
//    column tile, k = (x0, x1) and (x2, 1), the 1 carrying the bias.
//  - Layers 2 and 3: gfx950's v_mfma_f32_32x32x16_bf16 (K = 16), four
//    products per column tile.  K block kb draws its B operand from the
//    registers 8(kb%2)..8(kb%2)+7 of column tile kb/2 of the previous layer,
//    so its k slot s stands for hidden unit infer_unit(kb, s); the weight
//    fragments are gathered in that order once per wavefront.  Layer 2
//    accumulates onto its bias, layer 3 onto its three.
// Sums are the network's; their rounding order is the hardware's (layer 1
// adds the bias with the products instead of first).  Per activation the
// VALU does the gate's clamped fma, a multiply and half a BF16 pair
// conversion.
__host__ __device__ constexpr int infer_unit(int kb, int s) {
  return 32 * (kb >> 1) + 8 * (2 * (kb & 1) + ((s & 7) >> 2)) + 4 * (s >> 3) + (s & 3);
}
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
// the K blocks of a product cover each hidden unit exactly once
constexpr bool infer_units_cover() {
  unsigned long long seen = 0;
  for (int kb = 0; kb < 4; ++kb)
    for (int s = 0; s < 16; ++s) seen |= 1ull << infer_unit(kb, s);
  return seen == ~0ull;
}
static_assert(infer_units_cover(), "K-block unit order");
// FP32 -> BF16 conversions ((__bf16)f) round to nearest even like f2bf; on
// gfx950 they are v_cvt_pk_bf16_f32, two values per instruction.

constexpr int INFER_WAVES = 4;      // wavefronts per block
constexpr int INFER_BLOCKS = 1024;  // 4 per CU: all resident at 4 wavefronts per SIMD (<= 128 VGPRs); they loop over 32-sample tiles

// GUESS: start cells from positions (cell < 0 or no cell array; others get
// -1, and tiles where nobody needs one skip the network).  Otherwise the
// network's outputs for normalised inputs xin (tests).
template <bool GUESS>
__global__ __launch_bounds__(64 * INFER_WAVES, 4) void k_mlp_infer(int64_t n, const double *__restrict__ pos,
                                                                const int32_t *__restrict__ cell,
                                                                const float *__restrict__ xin, MlpWeights Wt,
                                                                MlpNorm nm, int IM, int JM, int KM,
                                                                int32_t *__restrict__ guess, float *__restrict__ yout) {
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int r32 = l & 31, h = l >> 5;
  // layer 1 A operands: A(m, k) with m = r32 and k = h (first product) or 2 + h (second; k = 3 is the bias)
  float w1a[2][2];
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const int q = 32 * t + r32;
    w1a[t][0] = bf2f(Wt.wb[OFF_W1 + 3 * q + h]);
    w1a[t][1] = h == 0 ? bf2f(Wt.wb[OFF_W1 + 3 * q + 2]) : Wt.p[OFF_B1 + q];
  }
  // layer 2 A operands (registers) and layer 3 A operands (LDS, by K block and lane):
  // A(m, k) = W[32t + m][infer_unit(kb, k)], m = r32, k = 8h + j; layer 3 rows 3..31 are zero
  __shared__ bf16x8 w3s[4][64];
  __shared__ float4 b2s[H / 4];  // layer 2 biases
  bf16x8 w2a[2][4];
#pragma unroll
  for (int kb = 0; kb < 4; ++kb)
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int u = infer_unit(kb, 8 * h + j);
      w2a[0][kb][j] = __builtin_bit_cast(__bf16, Wt.wb[OFF_W2 + r32 * H + u]);
      w2a[1][kb][j] = __builtin_bit_cast(__bf16, Wt.wb[OFF_W2 + (32 + r32) * H + u]);
    }
  if (wv == 0) {
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) {
      bf16x8 f;
#pragma unroll
      for (int j = 0; j < 8; ++j)
        f[j] = __builtin_bit_cast(__bf16, r32 < 3 ? Wt.wb[OFF_W3 + r32 * H + infer_unit(kb, 8 * h + j)] : (uint16_t)0);
      w3s[kb][l] = f;
    }
    if (l < H / 4)
      b2s[l] = make_float4(Wt.p[OFF_B2 + 4 * l], Wt.p[OFF_B2 + 4 * l + 1], Wt.p[OFF_B2 + 4 * l + 2],
                           Wt.p[OFF_B2 + 4 * l + 3]);
  }
  __syncthreads();
  const float b30 = Wt.p[OFF_B3], b31 = Wt.p[OFF_B3 + 1], b32 = Wt.p[OFF_B3 + 2];

  const int64_t tiles = (n + 31) / 32;
  // tile is the same in every lane of a wave, so the loop and the branches on
  // __ballot below are wave-uniform (MFMA runs with all lanes)
  for (int64_t tile = (int64_t)blockIdx.x * INFER_WAVES + wv; tile < tiles;
       tile += (int64_t)gridDim.x * INFER_WAVES) {
    const int64_t p = tile * 32 + r32;
    const bool in = p < n;
    bool need = in;
    if (GUESS) need = in && (cell == nullptr || cell[p] < 0);
    if (__ballot(need) == 0) {
      if (GUESS && h == 0 && in) guess[p] = -1;
      continue;
    }
    float x0 = 0.f, x1 = 0.f, x2 = 0.f;
    if (in) {
      if (GUESS) {
        x0 = norm_axis(pos[3 * p], nm.lo0, nm.s0);
        x1 = norm_axis(pos[3 * p + 1], nm.lo1, nm.s1);
        x2 = norm_axis(pos[3 * p + 2], nm.lo2, nm.s2);
      } else {
        x0 = xin[3 * p];
        x1 = xin[3 * p + 1];
        x2 = xin[3 * p + 2];
      }
    }
    // layer 1: B(k, n) = x_k of sample n, B(3, n) = 1
    const float xb0 = h == 0 ? x0 : x1, xb1 = h == 0 ? x2 : 1.0f;
    bf16x8 a1[4];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      f32x16 z = {};
      z = __builtin_amdgcn_mfma_f32_32x32x2f32(w1a[t][0], xb0, z, 0, 0, 0);
      z = __builtin_amdgcn_mfma_f32_32x32x2f32(w1a[t][1], xb1, z, 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 16; ++r) a1[2 * t + r / 8][r % 8] = (__bf16)act(z[r]);
    }
    // layer 2
    bf16x8 a2[4];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      f32x16 acc;
#pragma unroll
      for (int g = 0; g < 4; ++g) {  // registers 4g..4g+3: units 32t + 8g + 4h + 0..3
        const float4 b = b2s[8 * t + 2 * g + h];
        acc[4 * g] = b.x;
        acc[4 * g + 1] = b.y;
        acc[4 * g + 2] = b.z;
        acc[4 * g + 3] = b.w;
      }
#pragma unroll
      for (int kb = 0; kb < 4; ++kb) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w2a[t][kb], a1[kb], acc, 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 16; ++r) a2[2 * t + r / 8][r % 8] = (__bf16)act(acc[r]);
    }
    // layer 3: rows 0..2 of the result (registers 0..2 of lanes 0..31) are the outputs
    f32x16 acc3 = {};
    acc3[0] = b30;
    acc3[1] = b31;
    acc3[2] = b32;
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) acc3 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w3s[kb][l], a2[kb], acc3, 0, 0, 0);
    if (h == 0 && in) {
      const float y0 = acc3[0], y1 = acc3[1], y2 = acc3[2];
      if (GUESS) {
        int32_t g = -1;
        if (need) {
          int i = (int)floorf(y0 * (float)IM), j = (int)floorf(y1 * (float)JM), k = (int)floorf(y2 * (float)KM);
          i = i < 0 ? 0 : (i >= IM ? IM - 1 : i);
          j = j < 0 ? 0 : (j >= JM ? JM - 1 : j);
          k = k < 0 ? 0 : (k >= KM ? KM - 1 : k);
          g = (int32_t)(i + IM * (j + JM * k));
        }
        guess[p] = g;
      } else {
        yout[3 * p] = y0;
        yout[3 * p + 1] = y1;
        yout[3 * p + 2] = y2;
      }
    }
  }
}

static int64_t infer_blocks(int64_t n) {
  const int64_t tiles = (n + 31) / 32;
  const int64_t b = (tiles + INFER_WAVES - 1) / INFER_WAVES;
  return b < INFER_BLOCKS ? b : INFER_BLOCKS;
}

// ---------------------------------------------------------------- training
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  // SplitMix64 finaliser (Steele, Lea, Flood 2014)
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Forward + backward of one 32-sample tile per block.  Rows are samples
// base..base+31 of x/t, or (when npairs > 0) random pairs drawn with the
// SplitMix64 hash of (seed, iteration, row).  Writes the tile's gradient
// (parameter layout) and its summed squared error to part[block].
__global__ __launch_bounds__(MLP_THREADS) void k_mlp_fwd_bwd(int64_t n, const float *__restrict__ x,
                                                             const float *__restrict__ t, int64_t npairs,
                                                             uint64_t seed, int iter, float scale, MlpWeights Wt,
                                                             float *__restrict__ part) {
  __shared__ TileLds L;
  __shared__ float tl[TILE][3];
  __shared__ int valid[TILE];
  const int l = threadIdx.x;
  const int64_t base = (int64_t)blockIdx.x * TILE;
  if (l < TILE) {
    int64_t s = base + l;
    int ok = s < n;
    if (npairs > 0) s = (int64_t)(mix64(seed ^ mix64(((uint64_t)iter << 32) + (uint64_t)(base + l))) % (uint64_t)npairs);
    valid[l] = ok;
    for (int c = 0; c < 3; ++c) {
      L.x[l][c] = ok ? x[3 * s + c] : 0.f;
      tl[l][c] = ok ? t[3 * s + c] : 0.f;
    }
  }
  __syncthreads();
  tile_forward(Wt, L);
  float *g = part + (int64_t)blockIdx.x * (NP + 1);
  // output error
  float sq = 0.f;
  if (l < TILE) {
    for (int c = 0; c < 3; ++c) {
      const float e = valid[l] ? L.y[l][c] - tl[l][c] : 0.f;
      sq = __builtin_fmaf(e, e, sq);
      L.g3[l][c] = e * scale;
    }
  }
  for (int off = 32; off > 0; off >>= 1) sq += __shfl_xor(sq, off);
  if (l == 0) g[NP] = sq;
  __syncthreads();
  // layer 3 gradients and dA2 -> g2 (lane n owns hidden unit n)
  {
    const int nn = l;
    float gw0 = 0.f, gw1 = 0.f, gw2 = 0.f, gb2 = 0.f;
    const float w30 = bf2f(Wt.wb[OFF_W3 + nn]), w31 = bf2f(Wt.wb[OFF_W3 + H + nn]),
                w32 = bf2f(Wt.wb[OFF_W3 + 2 * H + nn]);
    for (int m = 0; m < TILE; ++m) {
      const float a2 = bf2f(L.a2[m][nn]);
      gw0 = __builtin_fmaf(L.g3[m][0], a2, gw0);
      gw1 = __builtin_fmaf(L.g3[m][1], a2, gw1);
      gw2 = __builtin_fmaf(L.g3[m][2], a2, gw2);
      const float da2 = __builtin_fmaf(L.g3[m][0], w30, __builtin_fmaf(L.g3[m][1], w31, L.g3[m][2] * w32));
      const float g2 = da2 * act_d(L.z2[m][nn]);
      gb2 += g2;
      L.g2[m][nn] = f2bf(g2);
    }
    g[OFF_W3 + nn] = gw0;
    g[OFF_W3 + H + nn] = gw1;
    g[OFF_W3 + 2 * H + nn] = gw2;
    g[OFF_B2 + nn] = gb2;
    if (l < 3) {
      float s = 0.f;
      for (int m = 0; m < TILE; ++m) s += L.g3[m][l];
      g[OFF_B3 + l] = s;
    }
  }
  __syncthreads();
  // dW2[n][q] = sum_m g2[m][n] a1[m][q]  (M = n, N = q, K = m)
  for (int tn = 0; tn < 2; ++tn)
    for (int tq = 0; tq < 2; ++tq) {
      f32x16 acc = {};
      mm32(&L.g2[0][32 * tn], 1, H, &L.a1[0][32 * tq], H, 1, TILE, acc);
#pragma unroll
      for (int r = 0; r < 16; ++r) g[OFF_W2 + (32 * tn + res_row(r)) * H + 32 * tq + res_col()] = acc[r];
    }
  // dA1[m][q] = sum_n g2[m][n] W2[n][q]  (M = m, N = q, K = n) -> g1
  for (int tq = 0; tq < 2; ++tq) {
    f32x16 acc = {};
    mm32(&L.g2[0][0], H, 1, Wt.wb + OFF_W2 + 32 * tq, H, 1, H, acc);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int m = res_row(r), q = 32 * tq + res_col();
      L.g1[m][q] = acc[r] * act_d(L.z1[m][q]);
    }
  }
  __syncthreads();
  // layer 1 gradients (lane q owns hidden unit q)
  {
    const int q = l;
    float gx = 0.f, gy = 0.f, gz = 0.f, gb = 0.f;
    for (int m = 0; m < TILE; ++m) {
      const float g1 = L.g1[m][q];
      gx = __builtin_fmaf(g1, L.x[m][0], gx);
      gy = __builtin_fmaf(g1, L.x[m][1], gy);
      gz = __builtin_fmaf(g1, L.x[m][2], gz);
      gb += g1;
    }
    g[OFF_W1 + 3 * q] = gx;
    g[OFF_W1 + 3 * q + 1] = gy;
    g[OFF_W1 + 3 * q + 2] = gz;
    g[OFF_B1 + q] = gb;
  }
}

// grad[j] = sum over tiles of part[tile][j], j < NP + 1, in a fixed order:
// wavefront g of a block sums the tiles b = g (mod RED_GROUPS) in ascending
// order, then the RED_GROUPS partial sums are added in ascending g.  Each
// block covers 64 consecutive j, one per lane, so every load is a coalesced
// 256-byte row segment and RED_GROUPS loads are in flight per j.
constexpr int RED_GROUPS = 8;
__global__ __launch_bounds__(64 * RED_GROUPS) void k_mlp_reduce(int ntiles, const float *__restrict__ part,
                                                                float *__restrict__ grad) {
  __shared__ float ps[RED_GROUPS][64];
  const int lane = threadIdx.x & 63, grp = threadIdx.x >> 6;
  const int j = blockIdx.x * 64 + lane;
  float s = 0.f;
  if (j <= NP)
    for (int b = grp; b < ntiles; b += RED_GROUPS) s += part[(int64_t)b * (NP + 1) + j];
  ps[grp][lane] = s;
  __syncthreads();
  if (grp == 0 && j <= NP) {
    float t = ps[0][lane];
#pragma unroll
    for (int q = 1; q < RED_GROUPS; ++q) t += ps[q][lane];
    grad[j] = t;
  }
}

// Adam on the FP32 master, then refresh the BF16 weight copy.
__global__ __launch_bounds__(256) void k_mlp_adam(float *__restrict__ p, float *__restrict__ m, float *__restrict__ v,
                                                  const float *__restrict__ g, uint16_t *__restrict__ wb, float lr,
                                                  float b1, float b2, float eps, float bc1, float bc2) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= NP) return;
  const float gj = g[j];
  const float mj = b1 * m[j] + (1.f - b1) * gj;
  const float vj = b2 * v[j] + (1.f - b2) * gj * gj;
  m[j] = mj;
  v[j] = vj;
  const float pj = p[j] - lr * (mj / bc1) / (sqrtf(vj / bc2) + eps);
  p[j] = pj;
  wb[j] = f2bf(pj);
}

__global__ __launch_bounds__(256) void k_mlp_to_bf16(const float *__restrict__ p, uint16_t *__restrict__ wb) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < NP) wb[j] = f2bf(p[j]);
}

// training pairs from located particles (in swarm order): x normalised,
// t = fractional logical position
__global__ __launch_bounds__(256) void k_mlp_pair_flags(int64_t n, const int32_t *__restrict__ cell,
                                                        const int32_t *__restrict__ status, int32_t *__restrict__ flag) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) flag[p] = (cell[p] >= 0 && status[p] == ST_LOCATED) ? 1 : 0;
}

template <typename R>
__global__ __launch_bounds__(256) void k_mlp_pairs(int64_t n, const double *__restrict__ pos,
                                                   const int32_t *__restrict__ cell, const int32_t *__restrict__ flag,
                                                   const int32_t *__restrict__ off, const R *__restrict__ w, MlpNorm nm,
                                                   int IM, int JM, int KM, int64_t maxp, float *__restrict__ x,
                                                   float *__restrict__ t) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n || !flag[p]) return;
  const int64_t s = off[p];
  if (s >= maxp) return;
  const int32_t c = cell[p];
  const int i = c % IM, j = (c / IM) % JM, k = c / (IM * JM);
  x[3 * s] = norm_axis(pos[3 * p], nm.lo0, nm.s0);
  x[3 * s + 1] = norm_axis(pos[3 * p + 1], nm.lo1, nm.s1);
  x[3 * s + 2] = norm_axis(pos[3 * p + 2], nm.lo2, nm.s2);
  t[3 * s] = (float)(((double)i + (double)w[3 * p]) / IM);
  t[3 * s + 1] = (float)(((double)j + (double)w[3 * p + 1]) / JM);
  t[3 * s + 2] = (float)(((double)k + (double)w[3 * p + 2]) / KM);
}

static MlpWeights weights(picgpu_ctx *c) { return MlpWeights{c->mlp_master.as<float>(), c->mlp_bf16.as<uint16_t>()}; }

static MlpNorm norm_of(picgpu_ctx *c) {
  MlpNorm nm;
  double lo[3], s[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = c->mlp_lo[a];
    s[a] = c->mlp_scale[a];
  }
  nm.lo0 = lo[0];
  nm.lo1 = lo[1];
  nm.lo2 = lo[2];
  nm.s0 = s[0];
  nm.s1 = s[1];
  nm.s2 = s[2];
  return nm;
}

static int set_norm_from_grid(picgpu_ctx *c) {
  if (!c->have_grid) return PICGPU_OK;
  for (int a = 0; a < 3; ++a) {
    c->mlp_lo[a] = c->bmin[a];
    double ext = c->bmax[a] - c->bmin[a];
    c->mlp_scale[a] = ext > 0 ? 2.0 / ext : 1.0;
  }
  return PICGPU_OK;
}

static int ensure_mlp_buffers(picgpu_ctx *c) {
  PG_CHECK(c->mlp_master.ensure(sizeof(float) * NP));
  PG_CHECK(c->mlp_bf16.ensure(sizeof(uint16_t) * NP));
  PG_CHECK(c->mlp_adam_m.ensure(sizeof(float) * NP));
  PG_CHECK(c->mlp_adam_v.ensure(sizeof(float) * NP));
  PG_CHECK(c->mlp_grad.ensure(sizeof(float) * (NP + 1)));
  return PICGPU_OK;
}

int mlp_guess_device(picgpu_ctx *c, int64_t n, const double *d_pos, const int32_t *d_cell, int32_t *d_guess) {
  if (!c->have_mlp) {
    set_error("MLP parameters not set");
    return PICGPU_E_STATE;
  }
  if (n <= 0) return PICGPU_OK;
  PG_LAUNCH(c, "mlp_guess", infer_blocks(n), 64 * INFER_WAVES, k_mlp_infer<true>, n, d_pos, d_cell,
            (const float *)nullptr, weights(c), norm_of(c), c->IM, c->JM, c->KM, d_guess, (float *)nullptr);
  return PICGPU_OK;
}

static int upload_params(picgpu_ctx *c, const float *h) {
  PG_CHECK(ensure_mlp_buffers(c));
  PG_HIP(hipMemcpy(c->mlp_master.ptr, h, sizeof(float) * NP, hipMemcpyHostToDevice));
  PG_LAUNCH(c, "mlp_to_bf16", (NP + 255) / 256, 256, k_mlp_to_bf16, c->mlp_master.as<float>(),
            c->mlp_bf16.as<uint16_t>());
  PG_HIP(hipStreamSynchronize(c->stream));
  c->have_mlp = true;
  return PICGPU_OK;
}

// device forward+backward over host-visible device arrays x, t (n samples)
static int loss_grad_device(picgpu_ctx *c, int64_t n, const float *dx, const float *dt, int64_t npairs, uint64_t seed,
                            int iter, float *loss_host) {
  const int64_t tiles = (n + TILE - 1) / TILE;
  PG_CHECK(c->mlp_partial.ensure(sizeof(float) * (size_t)tiles * (NP + 1)));
  const float scale = 2.0f / (3.0f * (float)n);
  PG_LAUNCH(c, "mlp_fwd_bwd", tiles, MLP_THREADS, k_mlp_fwd_bwd, n, dx, dt, npairs, seed, iter, scale, weights(c),
            c->mlp_partial.as<float>());
  PG_LAUNCH(c, "mlp_reduce", (NP + 1 + 63) / 64, 64 * RED_GROUPS, k_mlp_reduce, (int)tiles, c->mlp_partial.as<float>(),
            c->mlp_grad.as<float>());
  if (loss_host) {
    float sq = 0.f;
    PG_HIP(hipMemcpyAsync(&sq, c->mlp_grad.as<float>() + NP, sizeof(float), hipMemcpyDeviceToHost, c->stream));
    PG_HIP(hipStreamSynchronize(c->stream));
    *loss_host = sq / (3.0f * (float)n);
  }
  return PICGPU_OK;
}

}  // namespace picgpu

using namespace picgpu;

extern "C" {

int64_t picgpu_mlp_param_count(void) { return NP; }

int picgpu_mlp_enable(picgpu_ctx *c, int on) {
  if (!c) return PICGPU_E_ARG;
  if (on && !c->have_mlp) {
    set_error("MLP parameters not set (picgpu_mlp_set_params or picgpu_mlp_train first)");
    return PICGPU_E_STATE;
  }
  c->opt.use_mlp_guess = on ? 1 : 0;
  return PICGPU_OK;
}

int picgpu_mlp_set_params(picgpu_ctx *c, const float *params) {
  if (!c || !params) return PICGPU_E_ARG;
  PG_HIP(hipSetDevice(c->device));
  PG_CHECK(set_norm_from_grid(c));
  return upload_params(c, params);
}

int picgpu_mlp_get_params(picgpu_ctx *c, float *params) {
  if (!c || !params) return PICGPU_E_ARG;
  if (!c->have_mlp) {
    set_error("MLP parameters not set");
    return PICGPU_E_STATE;
  }
  PG_HIP(hipSetDevice(c->device));
  PG_HIP(hipStreamSynchronize(c->stream));
  PG_HIP(hipMemcpy(params, c->mlp_master.ptr, sizeof(float) * NP, hipMemcpyDeviceToHost));
  return PICGPU_OK;
}

int picgpu_mlp_forward(picgpu_ctx *c, int64_t n, const float *x, float *y) {
  if (!c || n < 0 || (n > 0 && (!x || !y))) return PICGPU_E_ARG;
  if (!c->have_mlp) {
    set_error("MLP parameters not set");
    return PICGPU_E_STATE;
  }
  if (n == 0) return PICGPU_OK;
  PG_HIP(hipSetDevice(c->device));
  PG_CHECK(c->mlp_x.ensure(sizeof(float) * 3 * (size_t)n));
  PG_CHECK(c->mlp_y.ensure(sizeof(float) * 3 * (size_t)n));
  PG_HIP(hipMemcpy(c->mlp_x.ptr, x, sizeof(float) * 3 * (size_t)n, hipMemcpyHostToDevice));
  PG_LAUNCH(c, "mlp_forward", infer_blocks(n), 64 * INFER_WAVES, k_mlp_infer<false>, n, (const double *)nullptr,
            (const int32_t *)nullptr, c->mlp_x.as<float>(), weights(c), MlpNorm{}, 1, 1, 1, (int32_t *)nullptr,
            c->mlp_y.as<float>());
  PG_HIP(hipStreamSynchronize(c->stream));
  PG_HIP(hipMemcpy(y, c->mlp_y.ptr, sizeof(float) * 3 * (size_t)n, hipMemcpyDeviceToHost));
  return PICGPU_OK;
}

int picgpu_mlp_loss_grad(picgpu_ctx *c, int64_t n, const float *x, const float *t, float *loss, float *grad) {
  if (!c || n <= 0 || !x || !t) return PICGPU_E_ARG;
  if (!c->have_mlp) {
    set_error("MLP parameters not set");
    return PICGPU_E_STATE;
  }
  PG_HIP(hipSetDevice(c->device));
  PG_CHECK(ensure_mlp_buffers(c));
  PG_CHECK(c->mlp_x.ensure(sizeof(float) * 3 * (size_t)n));
  PG_CHECK(c->mlp_t.ensure(sizeof(float) * 3 * (size_t)n));
  PG_HIP(hipMemcpy(c->mlp_x.ptr, x, sizeof(float) * 3 * (size_t)n, hipMemcpyHostToDevice));
  PG_HIP(hipMemcpy(c->mlp_t.ptr, t, sizeof(float) * 3 * (size_t)n, hipMemcpyHostToDevice));
  float l = 0.f;
  PG_CHECK(loss_grad_device(c, n, c->mlp_x.as<float>(), c->mlp_t.as<float>(), 0, 0, 0, &l));
  if (loss) *loss = l;
  if (grad) PG_HIP(hipMemcpy(grad, c->mlp_grad.ptr, sizeof(float) * NP, hipMemcpyDeviceToHost));
  return PICGPU_OK;
}

int picgpu_mlp_train(picgpu_ctx *c, int64_t max_pairs, int batch, int iters, float lr, uint64_t seed,
                     float *final_loss) {
  if (!c || max_pairs <= 0 || batch <= 0 || iters < 0 || batch % TILE) {
    set_error("mlp_train: bad arguments (batch must be a positive multiple of 32)");
    return PICGPU_E_ARG;
  }
  PG_HIP(hipSetDevice(c->device));
  PG_CHECK(settle_swarm(c));
  if (!c->have_grid || c->n == 0) {
    set_error("mlp_train: needs a grid and a located swarm");
    return PICGPU_E_STATE;
  }
  PG_HIP(hipSetDevice(c->device));
  PG_CHECK(set_norm_from_grid(c));
  PG_CHECK(ensure_mlp_buffers(c));
  if (!c->have_mlp) {
    // deterministic initialisation: N(0, 1/fan_in) weights, zero biases
    std::vector<float> h(NP, 0.f);
    uint64_t s = seed ^ 0x5eedull;
    auto next = [&]() {
      s += 0x9E3779B97F4A7C15ull;
      uint64_t z = s;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      z ^= z >> 31;
      return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    auto gauss = [&]() {
      double u1 = next(), u2 = next();
      if (u1 < 1e-300) u1 = 1e-300;
      return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    };
    for (int q = 0; q < 3 * H; ++q) h[OFF_W1 + q] = (float)(gauss() * 1.0);
    for (int q = 0; q < H * H; ++q) h[OFF_W2 + q] = (float)(gauss() / sqrt((double)H));
    for (int q = 0; q < 3 * H; ++q) h[OFF_W3 + q] = (float)(gauss() / sqrt((double)H));
    for (int q = 0; q < 3; ++q) h[OFF_B3 + q] = 0.5f;
    PG_CHECK(upload_params(c, h.data()));
  }
  // training pairs from the located swarm
  PG_CHECK(c->mlp_x.ensure(sizeof(float) * 3 * (size_t)max_pairs));
  PG_CHECK(c->mlp_t.ensure(sizeof(float) * 3 * (size_t)max_pairs));
  const int64_t n = c->n;
  PG_CHECK(c->keep.ensure(sizeof(int32_t) * (size_t)(2 * n + 2)));
  int32_t *flag = c->keep.as<int32_t>();
  int32_t *off = flag + n + 1;
  PG_LAUNCH(c, "mlp_pair_flags", (n + 255) / 256, 256, k_mlp_pair_flags, n, c->sw.cell.as<int32_t>(),
            c->sw.status.as<int32_t>(), flag);
  PG_CHECK(scan_exclusive_i32(c, flag, off, n));
  if (c->opt.precision == PICGPU_FP32) {
    PG_LAUNCH(c, "mlp_pairs", (n + 255) / 256, 256, k_mlp_pairs<float>, n, c->sw.pos.as<double>(),
              c->sw.cell.as<int32_t>(), flag, off, c->sw.w.as<float>(), norm_of(c), c->IM, c->JM, c->KM, max_pairs,
              c->mlp_x.as<float>(), c->mlp_t.as<float>());
  } else {
    PG_LAUNCH(c, "mlp_pairs", (n + 255) / 256, 256, k_mlp_pairs<double>, n, c->sw.pos.as<double>(),
              c->sw.cell.as<int32_t>(), flag, off, c->sw.w.as<double>(), norm_of(c), c->IM, c->JM, c->KM, max_pairs,
              c->mlp_x.as<float>(), c->mlp_t.as<float>());
  }
  int32_t npairs = 0;
  PG_HIP(hipMemcpyAsync(&npairs, off + n, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  PG_HIP(hipStreamSynchronize(c->stream));
  if (npairs > max_pairs) npairs = (int32_t)max_pairs;
  if (npairs < TILE) {
    set_error("mlp_train: fewer than 32 located particles");
    return PICGPU_E_STATE;
  }
  PG_HIP(hipMemsetAsync(c->mlp_adam_m.ptr, 0, sizeof(float) * NP, c->stream));
  PG_HIP(hipMemsetAsync(c->mlp_adam_v.ptr, 0, sizeof(float) * NP, c->stream));
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  float bc1 = 1.f, bc2 = 1.f;
  for (int it = 0; it < iters; ++it) {
    PG_CHECK(loss_grad_device(c, batch, c->mlp_x.as<float>(), c->mlp_t.as<float>(), npairs, seed, it, nullptr));
    bc1 *= b1;
    bc2 *= b2;
    // cosine decay to lr/20
    const float lr_t = (float)(lr * (0.05 + 0.95 * 0.5 * (1.0 + cos(3.141592653589793 * it / (double)iters))));
    PG_LAUNCH(c, "mlp_adam", (NP + 255) / 256, 256, k_mlp_adam, c->mlp_master.as<float>(), c->mlp_adam_m.as<float>(),
              c->mlp_adam_v.as<float>(), c->mlp_grad.as<float>(), c->mlp_bf16.as<uint16_t>(), lr_t, b1, b2, eps,
              1.f - bc1, 1.f - bc2);
  }
  // loss over the first pairs (up to 64 K) with the trained weights
  const int64_t ne = npairs < 65536 ? ((int64_t)npairs / TILE) * TILE : 65536;
  float l = 0.f;
  PG_CHECK(loss_grad_device(c, ne, c->mlp_x.as<float>(), c->mlp_t.as<float>(), 0, 0, 0, &l));
  if (final_loss) *final_loss = l;
  return PICGPU_OK;
}

int picgpu_mlp_guess(picgpu_ctx *c, int64_t n, const double *pos, int64_t *cell) {
  if (!c || n < 0 || (n > 0 && (!pos || !cell))) return PICGPU_E_ARG;
  if (n == 0) return PICGPU_OK;
  PG_HIP(hipSetDevice(c->device));
  DevBuf dp, dg;
  PG_CHECK(dp.ensure(sizeof(double) * 3 * (size_t)n));
  PG_CHECK(dg.ensure(sizeof(int32_t) * (size_t)n));
  PG_HIP(hipMemcpy(dp.ptr, pos, sizeof(double) * 3 * (size_t)n, hipMemcpyHostToDevice));
  int r = mlp_guess_device(c, n, dp.as<double>(), nullptr, dg.as<int32_t>());
  if (r != PICGPU_OK) {
    dp.release();
    dg.release();
    return r;
  }
  std::vector<int32_t> h((size_t)n);
  PG_HIP(hipStreamSynchronize(c->stream));
  PG_HIP(hipMemcpy(h.data(), dg.ptr, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost));
  dp.release();
  dg.release();
  for (int64_t p = 0; p < n; ++p) {
    int32_t v = h[p];
    cell[3 * p] = v % c->IM;
    cell[3 * p + 1] = (v / c->IM) % c->JM;
    cell[3 * p + 2] = v / ((int64_t)c->IM * c->JM);
  }
  return PICGPU_OK;
}

}  // extern "C"
