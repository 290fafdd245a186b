// Learned offset proposal network on gfx950: BF16 matrix-core (MFMA) kernels
// with FP32 accumulation for training (forward, backward, Adam) and inference,
// and the C ABI declared in include/xcorr_cnn.h.
//
// Every convolution is an implicit GEMM on v_mfma_f32_16x16x32_bf16: a wave
// owns a tile of 16 pixels x 32 channels (two 16 x 16 accumulators); one K step
// of 32 is one 3 x 3 tap over the 32 input channels (conv1: all 9 taps x 2
// channels, zero-padded to 32).  Lane l of a wave supplies row (l % 16) of A
// and column (l % 16) of B with the 8 consecutive K values 8 (l / 16) ..
// 8 (l / 16) + 7, and holds D rows 4 (l / 16) .. + 3 of column l % 16.
// Activations are NHWC bf16 (32 channels = 64 contiguous bytes per pixel), so
// every A fragment is one 16-byte load.
#include "../../include/xcorr_cnn.h"
#include "../../include/xcorr_hip.h"

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

namespace xn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short bf16_t;

constexpr int kC = 32;                    // channels of every conv output
constexpr int kNParam = XCORR_CNN_NPARAM;
// parameter offsets (FP32 master vector)
constexpr int kW1 = 0, kB1 = kW1 + 9 * kC * 2;
constexpr int kW2 = kB1 + kC, kB2 = kW2 + 9 * kC * kC;
constexpr int kW3 = kB2 + kC, kB3 = kW3 + 9 * kC * kC;
constexpr int kW4 = kB3 + kC, kB4 = kW4 + 9 * kC * kC;
constexpr int kWh = kB4 + kC, kBh = kWh + 3 * kC;
static_assert(kBh + 3 == kNParam, "parameter layout");
// bf16 weight buffer: W1k[32 co][32 k] (k = 2 tap + ci, zero past 18), then for
// conv2..4 Wf[9][32 co][32 ci] (forward) and WT[9][32 ci][32 co] (backward data)
constexpr int kQ1 = 0, kQf2 = kQ1 + kC * 32, kQf3 = kQf2 + 9 * kC * kC, kQf4 = kQf3 + 9 * kC * kC;
constexpr int kQt2 = kQf4 + 9 * kC * kC, kQt3 = kQt2 + 9 * kC * kC, kQt4 = kQt3 + 9 * kC * kC;
constexpr int kQn = kQt4 + 9 * kC * kC;

__device__ __forceinline__ bf16_t f2bf(float f) {
  unsigned u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (bf16_t)0x7fc0;
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16_t h) { return __uint_as_float((unsigned)h << 16); }
__device__ __forceinline__ bf16x8 ld8(const bf16_t* p) { return __builtin_bit_cast(bf16x8, *(const uint4*)p); }
__device__ __forceinline__ bf16x8 zero8() { return __builtin_bit_cast(bf16x8, make_uint4(0u, 0u, 0u, 0u)); }
__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// Tile functions (shared by the layer kernels and the fused inference kernel)
// ---------------------------------------------------------------------------
// conv1's im2col for one lane: lane (m, g) supplies K values 8 g .. 8 g + 7,
// i.e. taps 4 g .. 4 g + 3 with both input channels of a tap adjacent (one
// dword of x0); taps past 8 (K >= 18) are zero.  Per tap: the input offset
// (ky - 1, kx - 1), or off.
struct Conv1Lane {
  int dy[4], dx[4];
  bool on[4];
};
__device__ __forceinline__ Conv1Lane conv1_lane(int lane) {
  Conv1Lane L;
#pragma unroll
  for (int jj = 0; jj < 4; jj++) {
    const int tap = 4 * (lane >> 4) + jj;
    L.on[jj] = tap < 9;
    L.dy[jj] = tap / 3 - 1;
    L.dx[jj] = tap - 3 * (tap / 3) - 1;
  }
  return L;
}

// conv1 (2 -> 32, stride 2, 64 x 64 -> 32 x 32): output pixels p0 .. p0 + 15 of
// one sample; xb = the sample's x0[64][64][2] (4-byte aligned).
__device__ __forceinline__ void conv1_tile(const bf16_t* xb, const bf16_t* W1k, int p0, int lane, const Conv1Lane& L,
                                           f32x4 (&acc)[2]) {
  const int m = lane & 15, g = lane >> 4;
  const int p = p0 + m, oy = p >> 5, ox = p & 31;
  unsigned av[4];
#pragma unroll
  for (int jj = 0; jj < 4; jj++) {
    const int iy = 2 * oy + L.dy[jj], ix = 2 * ox + L.dx[jj];
    av[jj] = (L.on[jj] && iy >= 0 && iy < 64 && ix >= 0 && ix < 64)
                 ? *(const unsigned*)(xb + (iy * 64 + ix) * 2) : 0u;
  }
  const bf16x8 a = __builtin_bit_cast(bf16x8, make_uint4(av[0], av[1], av[2], av[3]));
  acc[0] = mfma(a, ld8(W1k + m * 32 + 8 * g), f32x4{0.f, 0.f, 0.f, 0.f});
  acc[1] = mfma(a, ld8(W1k + (16 + m) * 32 + 8 * g), f32x4{0.f, 0.f, 0.f, 0.f});
}

// 32 -> 32 conv, HIN x HIN input, stride S: output pixels p0 .. p0 + 15.
template <int HIN, int S>
__device__ __forceinline__ void conv_tile(const bf16_t* xb, const bf16_t* Wf, int p0, int lane, f32x4 (&acc)[2]) {
  constexpr int HO = (HIN - 1) / S + 1;
  const int m = lane & 15, g = lane >> 4;
  const int p = p0 + m, oy = p / HO, ox = p - (p / HO) * HO;
  acc[0] = f32x4{0.f, 0.f, 0.f, 0.f};
  acc[1] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int tap = 0; tap < 9; tap++) {
    const int ky = tap / 3, kx = tap - 3 * (tap / 3);
    const int iy = oy * S + ky - 1, ix = ox * S + kx - 1;
    const bf16x8 a = (iy >= 0 && iy < HIN && ix >= 0 && ix < HIN) ? ld8(xb + (iy * HIN + ix) * kC + 8 * g) : zero8();
    acc[0] = mfma(a, ld8(Wf + (tap * kC + m) * kC + 8 * g), acc[0]);
    acc[1] = mfma(a, ld8(Wf + (tap * kC + 16 + m) * kC + 8 * g), acc[1]);
  }
}

// As conv_tile with the B fragments of all 9 taps already in registers
// (bw[tap][n-tile]); the same MFMA sequence on the same operands.
template <int HIN, int S>
__device__ __forceinline__ void conv_tile_w(const bf16_t* xb, const bf16x8 (&bw)[9][2], int p0, int lane,
                                            f32x4 (&acc)[2]) {
  constexpr int HO = (HIN - 1) / S + 1;
  const int m = lane & 15, g = lane >> 4;
  const int p = p0 + m, oy = p / HO, ox = p - (p / HO) * HO;
  acc[0] = f32x4{0.f, 0.f, 0.f, 0.f};
  acc[1] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int tap = 0; tap < 9; tap++) {
    const int ky = tap / 3, kx = tap - 3 * (tap / 3);
    const int iy = oy * S + ky - 1, ix = ox * S + kx - 1;
    const bf16x8 a = (iy >= 0 && iy < HIN && ix >= 0 && ix < HIN) ? ld8(xb + (iy * HIN + ix) * kC + 8 * g) : zero8();
    acc[0] = mfma(a, bw[tap][0], acc[0]);
    acc[1] = mfma(a, bw[tap][1], acc[1]);
  }
}
__device__ __forceinline__ void load_w(const bf16_t* Wf, int lane, bf16x8 (&bw)[9][2]) {
  const int m = lane & 15, g = lane >> 4;
#pragma unroll
  for (int tap = 0; tap < 9; tap++) {
    bw[tap][0] = ld8(Wf + (tap * kC + m) * kC + 8 * g);
    bw[tap][1] = ld8(Wf + (tap * kC + 16 + m) * kC + 8 * g);
  }
}

// bias + ReLU + bf16 store of a conv tile: y[(p0 + 4g + i) * 32 + co], lane
// (m, g) holding channel co = 16 t + m of pixels p0 + 4g + i.  Lanes m and m ^ 1
// swap one value (a DPP quad permutation) so that each stores a dword of two
// channels: even lanes channels (m, m + 1) of pixel i, odd lanes channels
// (m - 1, m) of pixel i + 1, for i = 0, 2.
__device__ __forceinline__ void store_tile(bf16_t* yb, const float* bias, int p0, int lane, const f32x4 (&acc)[2]) {
  const int m = lane & 15, g = lane >> 4;
  const bool odd = m & 1;
#pragma unroll
  for (int t = 0; t < 2; t++) {
    const float bv = bias[16 * t + m];
    unsigned h[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const float z = acc[t][i] + bv;
      h[i] = f2bf(z > 0.f ? z : 0.f);
    }
#pragma unroll
    for (int i = 0; i < 4; i += 2) {
      const unsigned send = odd ? h[i] : h[i + 1];
      const unsigned got = (unsigned)__builtin_amdgcn_mov_dpp((int)send, 0xB1, 0xf, 0xf, false);   // quad_perm 1,0,3,2
      const unsigned wv = odd ? (got | (h[i + 1] << 16)) : (h[i] | (got << 16));
      const int px = p0 + 4 * g + i + (odd ? 1 : 0), co = 16 * t + m - (odd ? 1 : 0);
      *(unsigned*)(yb + px * kC + co) = wv;
    }
  }
}

// Input gradient of a 32 -> 32 conv (HIN x HIN input, stride S) at input
// pixels q0 .. q0 + 15: dx[q][ci] = sum_{tap, co} W[tap][co][ci] dz[o(q, tap)][co].
template <int HIN, int S>
__device__ __forceinline__ void convT_tile(const bf16_t* dzb, const bf16_t* WT, int q0, int lane, f32x4 (&acc)[2]) {
  constexpr int HO = (HIN - 1) / S + 1;
  const int m = lane & 15, g = lane >> 4;
  const int q = q0 + m, iy = q / HIN, ix = q - (q / HIN) * HIN;
  acc[0] = f32x4{0.f, 0.f, 0.f, 0.f};
  acc[1] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int tap = 0; tap < 9; tap++) {
    const int ky = tap / 3, kx = tap - 3 * (tap / 3);
    const int ny = iy + 1 - ky, nx = ix + 1 - kx;
    const bool ok = ny >= 0 && nx >= 0 && (ny % S) == 0 && (nx % S) == 0 && ny / S < HO && nx / S < HO;
    const bf16x8 a = ok ? ld8(dzb + ((ny / S) * HO + nx / S) * kC + 8 * g) : zero8();
    acc[0] = mfma(a, ld8(WT + (tap * kC + m) * kC + 8 * g), acc[0]);
    acc[1] = mfma(a, ld8(WT + (tap * kC + 16 + m) * kC + 8 * g), acc[1]);
  }
}

// ---------------------------------------------------------------------------
// layer kernels (training path): one wave per 16-pixel tile, 4 waves per block
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_conv1_fwd(const bf16_t* __restrict__ x0, const bf16_t* __restrict__ q,
                                                   const float* __restrict__ prm, bf16_t* __restrict__ y, int n) {
  const int tile = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (tile >= n * 64) return;                 // 1024 pixels = 64 tiles per sample
  const int b = tile >> 6, p0 = (tile & 63) * 16;
  f32x4 acc[2];
  conv1_tile(x0 + (size_t)b * 64 * 64 * 2, q + kQ1, p0, lane, conv1_lane(lane), acc);
  store_tile(y + (size_t)b * 1024 * kC, prm + kB1, p0, lane, acc);
}

template <int HIN, int S>
__global__ void __launch_bounds__(256) k_conv_fwd(const bf16_t* __restrict__ x, const bf16_t* __restrict__ Wf,
                                                  const float* __restrict__ bias, bf16_t* __restrict__ y, int n) {
  constexpr int HO = (HIN - 1) / S + 1, TPS = HO * HO / 16;
  const int tile = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (tile >= n * TPS) return;
  const int b = tile / TPS, p0 = (tile - b * TPS) * 16;
  f32x4 acc[2];
  conv_tile<HIN, S>(x + (size_t)b * HIN * HIN * kC, Wf, p0, lane, acc);
  store_tile(y + (size_t)b * HO * HO * kC, bias, p0, lane, acc);
}

// dz_in = bf16(dx_in * (x_in > 0)) for a 32 -> 32 conv.
template <int HIN, int S>
__global__ void __launch_bounds__(256) k_conv_bwd_data(const bf16_t* __restrict__ dz, const bf16_t* __restrict__ xin,
                                                       const bf16_t* __restrict__ WT, bf16_t* __restrict__ dzin, int n) {
  constexpr int HO = (HIN - 1) / S + 1, TPS = HIN * HIN / 16;
  const int tile = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (tile >= n * TPS) return;
  const int b = tile / TPS, q0 = (tile - b * TPS) * 16;
  f32x4 acc[2];
  convT_tile<HIN, S>(dz + (size_t)b * HO * HO * kC, WT, q0, lane, acc);
  const int m = lane & 15, g = lane >> 4;
  const bf16_t* xb = xin + (size_t)b * HIN * HIN * kC;
  bf16_t* ob = dzin + (size_t)b * HIN * HIN * kC;
#pragma unroll
  for (int t = 0; t < 2; t++)
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int o = (q0 + 4 * g + i) * kC + 16 * t + m;
      ob[o] = f2bf(xb[o] != 0 ? acc[t][i] : 0.f);     // x > 0 exactly where its bf16 is non-zero (post-ReLU)
    }
}

// Weight and bias gradients of a conv: dW[tap][co][ci] += sum_px dz[px][co] x[in(px, tap)][ci].
// A block stages 32 output pixels at a time in LDS, transposed (channel-major),
// and accumulates all 36 (tap, co tile, ci tile) 16 x 16 tiles (conv1: 4 tiles
// over co x k), K = pixels.  Partial sums go to grad with FP32 atomics.
constexpr int kPxPitch = 40;   // bf16 per channel row of a staged 32-pixel chunk (16-byte aligned rows)

template <int CIN, int HIN, int S>
__global__ void __launch_bounds__(256) k_conv_bwd_w(const bf16_t* __restrict__ dz, const bf16_t* __restrict__ xin,
                                                    float* __restrict__ gW, float* __restrict__ gb, int n) {
  constexpr int HO = (HIN - 1) / S + 1;
  constexpr int NT = CIN == 2 ? 1 : 9;       // tap groups staged
  constexpr int TILES = CIN == 2 ? 4 : 36;
  constexpr int PER = TILES / 4;             // tiles per wave
  __shared__ __attribute__((aligned(16))) bf16_t dzT[kC * kPxPitch];
  __shared__ __attribute__((aligned(16))) bf16_t xT[NT * kC * kPxPitch];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, m = lane & 15, g = lane >> 4;
  const long total = (long)n * HO * HO;
  f32x4 acc[PER];
#pragma unroll
  for (int u = 0; u < PER; u++) acc[u] = f32x4{0.f, 0.f, 0.f, 0.f};
  float bsum = 0.f;
  for (long base = (long)blockIdx.x * 32; base < total; base += (long)gridDim.x * 32) {
    __syncthreads();
    for (int e = threadIdx.x; e < 32 * kC; e += 256) {
      const int px = e >> 5, co = e & 31;
      const long P = base + px;
      dzT[co * kPxPitch + px] = P < total ? dz[P * kC + co] : (bf16_t)0;
    }
    for (int e = threadIdx.x; e < NT * 32 * kC; e += 256) {
      const int grp = e >> 10, r = e & 1023, px = r >> 5, c = r & 31;
      const long P = base + px;
      bf16_t v = 0;
      if (P < total) {
        const int b = (int)(P / (HO * HO)), rr = (int)(P - (long)b * HO * HO), oy = rr / HO, ox = rr - (rr / HO) * HO;
        int tap, ci;
        if (CIN == 2) { tap = c >> 1; ci = c & 1; } else { tap = grp; ci = c; }
        if (CIN != 2 || c < 18) {
          const int ky = tap / 3, kx = tap - 3 * (tap / 3), iy = oy * S + ky - 1, ix = ox * S + kx - 1;
          if (iy >= 0 && iy < HIN && ix >= 0 && ix < HIN) v = xin[(((size_t)b * HIN + iy) * HIN + ix) * CIN + ci];
        }
      }
      xT[(grp * kC + c) * kPxPitch + px] = v;
    }
    __syncthreads();
    if (threadIdx.x < kC)
      for (int px = 0; px < 32; px++) bsum += bf2f(dzT[threadIdx.x * kPxPitch + px]);
#pragma unroll
    for (int u = 0; u < PER; u++) {
      const int t = wave * PER + u;                  // tile: (grp, co tile, c tile)
      const int grp = t >> 2, cot = (t >> 1) & 1, cit = t & 1;
#pragma unroll
      for (int ks = 0; ks < 32; ks += 32) {
        const bf16x8 a = ld8(dzT + (16 * cot + m) * kPxPitch + ks + 8 * g);
        const bf16x8 bb = ld8(xT + (grp * kC + 16 * cit + m) * kPxPitch + ks + 8 * g);
        acc[u] = mfma(a, bb, acc[u]);
      }
    }
  }
#pragma unroll
  for (int u = 0; u < PER; u++) {
    const int t = wave * PER + u;
    const int grp = t >> 2, cot = (t >> 1) & 1, cit = t & 1;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int co = 16 * cot + 4 * g + i, c = 16 * cit + m;
      if (CIN == 2) {
        if (c < 18) atomicAdd(&gW[(c >> 1) * kC * 2 + co * 2 + (c & 1)], acc[u][i]);
      } else {
        atomicAdd(&gW[(grp * kC + co) * kC + c], acc[u][i]);
      }
    }
  }
  if (threadIdx.x < kC) atomicAdd(&gb[threadIdx.x], bsum);
}

// Head: f = mean of x4 over 64 pixels, o = Wh f + bh, loss, and its gradient
// (head parameters by atomics, dz4 = bf16(df / 64 * (x4 > 0))).  One 64-thread
// block per sample.  tgt == nullptr: forward only.
__global__ void __launch_bounds__(64) k_head(const bf16_t* __restrict__ x4, const float* __restrict__ prm,
                                             const float* __restrict__ tgt, float* __restrict__ out,
                                             float* __restrict__ grad, float* __restrict__ loss,
                                             bf16_t* __restrict__ dz4, int n, float inv_n) {
  __shared__ float f[kC], d[3];
  const int b = blockIdx.x, t = threadIdx.x;
  const bf16_t* xb = x4 + (size_t)b * 64 * kC;
  if (t < kC) {
    float s = 0.f;
    for (int px = 0; px < 64; px++) s += bf2f(xb[px * kC + t]);
    f[t] = s * (1.0f / 64.0f);
  }
  __syncthreads();
  if (t == 0) {
    float o[3];
    for (int j = 0; j < 3; j++) {
      float s = prm[kBh + j];
      for (int c = 0; c < kC; c++) s += prm[kWh + j * kC + c] * f[c];
      o[j] = s;
      if (out) out[(size_t)b * 3 + j] = s;
    }
    if (tgt) {
      const float* tg = tgt + (size_t)b * 5;
      const float r0 = o[0] - tg[0], r1 = o[1] - tg[1], q = o[2], tc = tg[2], wr = tg[3], wc = tg[4];
      auto hub = [](float r) { const float a = fabsf(r); return a <= 1.f ? 0.5f * r * r : a - 0.5f; };
      const float bce = fmaxf(q, 0.f) - q * tc + log1pf(expf(-fabsf(q)));
      atomicAdd(loss, (wr * (hub(r0) + hub(r1)) + wc * bce) * inv_n);
      d[0] = wr * fminf(fmaxf(r0, -1.f), 1.f) * inv_n;
      d[1] = wr * fminf(fmaxf(r1, -1.f), 1.f) * inv_n;
      d[2] = wc * (1.f / (1.f + expf(-q)) - tc) * inv_n;
    }
  }
  __syncthreads();
  if (!tgt) return;
  for (int e = t; e < 3 * kC; e += 64) atomicAdd(&grad[kWh + e], d[e / kC] * f[e % kC]);
  if (t < 3) atomicAdd(&grad[kBh + t], d[t]);
  __shared__ float df[kC];
  if (t < kC) df[t] = (prm[kWh + t] * d[0] + prm[kWh + kC + t] * d[1] + prm[kWh + 2 * kC + t] * d[2]) * (1.0f / 64.0f);
  __syncthreads();
  bf16_t* ob = dz4 + (size_t)b * 64 * kC;
  for (int e = t; e < 64 * kC; e += 64) ob[e] = f2bf(xb[e] != 0 ? df[e & 31] : 0.f);
}

// bf16 weight forms from the FP32 master parameters.
__global__ void k_pack(const float* __restrict__ prm, bf16_t* __restrict__ q) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < kC * 32) {                         // W1k[co][k]
    const int co = e >> 5, k = e & 31;
    q[kQ1 + e] = k < 18 ? f2bf(prm[kW1 + (k >> 1) * kC * 2 + co * 2 + (k & 1)]) : (bf16_t)0;
  }
  if (e < 9 * kC * kC) {
    const int tap = e / (kC * kC), r = e % (kC * kC), a = r >> 5, c = r & 31;   // e = (tap, co=a, ci=c)
    const int offs[3] = {kW2, kW3, kW4}, qf[3] = {kQf2, kQf3, kQf4}, qt[3] = {kQt2, kQt3, kQt4};
    for (int l = 0; l < 3; l++) {
      const bf16_t v = f2bf(prm[offs[l] + e]);
      q[qf[l] + e] = v;                                   // Wf[tap][co][ci]
      q[qt[l] + (tap * kC + c) * kC + a] = v;             // WT[tap][ci][co]
    }
  }
}

__global__ void k_zero(float* __restrict__ p, int n) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) p[e] = 0.f;
}

__global__ void k_adam(float* __restrict__ p, float* __restrict__ m1, float* __restrict__ m2,
                       const float* __restrict__ g, int n, float lr, float bc1, float bc2) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float gv = g[e];
  const float a = 0.9f * m1[e] + 0.1f * gv;
  const float b = 0.999f * m2[e] + 0.001f * gv * gv;
  m1[e] = a;
  m2[e] = b;
  p[e] -= lr * (a / bc1) / (sqrtf(b / bc2) + 1e-8f);
}

// 4 x 4 box mean of the amplitude.
__global__ void k_box4(const void* __restrict__ img, int nx, int ny, int fmt, float* __restrict__ out, int dnx,
                       int dny) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)dnx * dny) return;
  const int y = (int)(e / dnx), x = (int)(e - (long)y * dnx);
  float s = 0.f;
  for (int r = 0; r < 4; r++) {
    const size_t row = (size_t)(4 * y + r) * nx + 4 * x;
    for (int c = 0; c < 4; c++) {
      float a;
      if (fmt == XCORR_HIP_CINT16) {
        const unsigned u = ((const unsigned*)img)[row + c];
        const float re = (float)(short)(u & 0xffffu), im = (float)(short)(u >> 16);
        a = sqrtf(__builtin_fmaf(re, re, im * im));
      } else {
        a = fabsf(((const float*)img)[row + c]);
      }
      s += a;
    }
  }
  out[e] = s * (1.0f / 16.0f);
}

// Normalised 64 x 64 crops of the master (channel 0, corner (xm, ym) of dsm)
// and the slave (channel 1, corner (xs, ys) of dss) into dst[(y * 64 + x) * 2 +
// ch] (4-byte aligned; block of 256 threads; red: 8 floats of shared scratch).
// Each channel is normalised by its own mean and standard deviation; the two
// channels' loads are in flight together, their reductions share barriers,
// and a pixel's two values go out as one dword.
__device__ void crop_norm2(const float* __restrict__ dsm, int dnx_m, int dny_m, int xm, int ym,
                           const float* __restrict__ dss, int dnx_s, int dny_s, int xs, int ys, bf16_t* dst,
                           float* red) {
  float v[16], w[16];
  float s0 = 0.f, s1 = 0.f;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    const int e = threadIdx.x + 256 * k, y = e >> 6, x = e & 63;
    const int gy = ym + y, gx = xm + x, hy = ys + y, hx = xs + x;
    v[k] = (gy >= 0 && gy < dny_m && gx >= 0 && gx < dnx_m) ? dsm[(size_t)gy * dnx_m + gx] : 0.f;
    w[k] = (hy >= 0 && hy < dny_s && hx >= 0 && hx < dnx_s) ? dss[(size_t)hy * dnx_s + hx] : 0.f;
  }
#pragma unroll
  for (int k = 0; k < 16; k++) {
    s0 += v[k];
    s1 += w[k];
  }
  auto bsum2 = [&](float a, float b) {
    for (int o = 32; o > 0; o >>= 1) {
      a += __shfl_xor(a, o);
      b += __shfl_xor(b, o);
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
      red[threadIdx.x >> 6] = a;
      red[4 + (threadIdx.x >> 6)] = b;
    }
    __syncthreads();
    return make_float2(red[0] + red[1] + red[2] + red[3], red[4] + red[5] + red[6] + red[7]);
  };
  const float2 sm = bsum2(s0, s1);
  const float mean0 = sm.x * (1.0f / 4096.0f), mean1 = sm.y * (1.0f / 4096.0f);
  float q0 = 0.f, q1 = 0.f;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    q0 += (v[k] - mean0) * (v[k] - mean0);
    q1 += (w[k] - mean1) * (w[k] - mean1);
  }
  const float2 sq = bsum2(q0, q1);
  const float den0 = fmaxf(sqrtf(sq.x * (1.0f / 4096.0f)), 1e-6f);
  const float den1 = fmaxf(sqrtf(sq.y * (1.0f / 4096.0f)), 1e-6f);
#pragma unroll
  for (int k = 0; k < 16; k++)
    *(unsigned*)(dst + (threadIdx.x + 256 * k) * 2) =
        (unsigned)f2bf((v[k] - mean0) / den0) | ((unsigned)f2bf((w[k] - mean1) / den1) << 16);
}

__global__ void __launch_bounds__(256) k_crops(const float* __restrict__ dsm, int dnx_m, int dny_m,
                                               const float* __restrict__ dss, int dnx_s, int dny_s,
                                               const int4* __restrict__ corners, bf16_t* __restrict__ x0) {
  __shared__ float red[8];
  const int b = blockIdx.x;
  const int4 c = corners[b];
  crop_norm2(dsm, dnx_m, dny_m, c.x, c.y, dss, dnx_s, dny_s, c.z, c.w, x0 + (size_t)b * 64 * 64 * 2, red);
}

// Fused inference: one block (4 waves) per sample, activations in LDS.
// R0 (16 KB): x0, then act2, then act4.  R1 (64 KB): the crop's reduction
// scratch, act1, act3, then the head's channel means.  80 KB in all, so two
// blocks share a CU.
template <bool CROP>
__global__ void __launch_bounds__(256) k_infer(const bf16_t* __restrict__ x0g, const float* __restrict__ dsm,
                                               int dnx_m, int dny_m, const float* __restrict__ dss, int dnx_s,
                                               int dny_s, const int4* __restrict__ corners,
                                               const bf16_t* __restrict__ q, const float* __restrict__ prm,
                                               float* __restrict__ out) {
  __shared__ __attribute__((aligned(16))) bf16_t R0[64 * 64 * 2];
  __shared__ __attribute__((aligned(16))) bf16_t R1[32 * 32 * kC];
  float* red = (float*)R1;                                  // crop_norm2 scratch, before act1
  float* fmean = (float*)R1;                                // after conv4 has read act3
  const int b = blockIdx.x, wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (CROP) {
    const int4 c = corners[b];
    crop_norm2(dsm, dnx_m, dny_m, c.x, c.y, dss, dnx_s, dny_s, c.z, c.w, R0, red);
  } else {
    const uint4* src = (const uint4*)(x0g + (size_t)b * 64 * 64 * 2);
    for (int e = threadIdx.x; e < 64 * 64 * 2 / 8; e += 256) ((uint4*)R0)[e] = src[e];
  }
  __syncthreads();
  f32x4 acc[2];
  const Conv1Lane L1 = conv1_lane(lane);
  for (int t = wave; t < 64; t += 4) {                      // conv1 -> R1 (32 x 32 x 32)
    conv1_tile(R0, q + kQ1, t * 16, lane, L1, acc);
    store_tile(R1, prm + kB1, t * 16, lane, acc);
  }
  bf16x8 bw[9][2];                                          // a layer's B fragments, loaded once per wave
  load_w(q + kQf2, lane, bw);
  __syncthreads();
  for (int t = wave; t < 16; t += 4) {                      // conv2 -> R0 (16 x 16 x 32)
    conv_tile_w<32, 2>(R1, bw, t * 16, lane, acc);
    store_tile(R0, prm + kB2, t * 16, lane, acc);
  }
  load_w(q + kQf3, lane, bw);
  __syncthreads();
  conv_tile_w<16, 2>(R0, bw, wave * 16, lane, acc);          // conv3 -> R1 (8 x 8 x 32)
  store_tile(R1, prm + kB3, wave * 16, lane, acc);
  load_w(q + kQf4, lane, bw);
  __syncthreads();
  conv_tile_w<8, 1>(R1, bw, wave * 16, lane, acc);           // conv4 -> R0 (8 x 8 x 32)
  store_tile(R0, prm + kB4, wave * 16, lane, acc);
  __syncthreads();
  // head, in k_head's order of operations (so both paths give the same bits)
  if (threadIdx.x < kC) {
    float s = 0.f;
    for (int px = 0; px < 64; px++) s += bf2f(R0[px * kC + threadIdx.x]);
    fmean[threadIdx.x] = s * (1.0f / 64.0f);
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int j = threadIdx.x;
    float s = prm[kBh + j];
    for (int c = 0; c < kC; c++) s += prm[kWh + j * kC + c] * fmean[c];
    out[(size_t)b * 3 + j] = s;
  }
}

}  // namespace xn

using namespace xn;

struct xcorr_cnn {
  int device = 0;
  hipStream_t stream = nullptr;   // the stream work is queued on: own_stream or the caller's
  hipStream_t own_stream = nullptr;
  int max_batch = 0;
  float* prm = nullptr;      // FP32 master parameters
  float* m1 = nullptr;       // Adam moments
  float* m2 = nullptr;
  float* grad = nullptr;     // gradient, then the loss (grad[kNParam])
  bf16_t* q = nullptr;       // bf16 weight forms
  bf16_t *a1 = nullptr, *a2 = nullptr, *a3 = nullptr, *a4 = nullptr;   // activations
  bf16_t *d1 = nullptr, *d2 = nullptr, *d3 = nullptr, *d4 = nullptr;   // pre-activation gradients
  float* hout = nullptr;     // head outputs of the training forward
  int4* corners = nullptr;
  int corners_n = 0;
  long step = 0;
  std::string err;
};

static std::string g_cnn_err;

static int cfail(xcorr_cnn* c, int code, const std::string& m) {
  if (c) c->err = m;
  else g_cnn_err = m;
  return code;
}

#define CNNCK(c, call)                                                                       \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess) return cfail(c, XCORR_HIP_EDEVICE, std::string(#call ": ") + hipGetErrorString(e_)); \
  } while (0)

static void cnn_free(xcorr_cnn* c) {
  void* ptrs[] = {c->prm, c->m1, c->m2, c->grad, c->q, c->a1, c->a2, c->a3, c->a4,
                  c->d1, c->d2, c->d3, c->d4, c->hout, c->corners};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  delete c;
}

extern "C" int xcorr_cnn_create(int device, int max_batch, xcorr_cnn** out) {
  if (!out || max_batch <= 0) return cfail(nullptr, XCORR_HIP_EINVAL, "cnn_create: bad arguments");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return cfail(nullptr, XCORR_HIP_EDEVICE, "no HIP device");
  if (device < 0 && hipGetDevice(&device) != hipSuccess) device = 0;
  if (device >= ndev) return cfail(nullptr, XCORR_HIP_EINVAL, "device index out of range");
  xcorr_cnn* c = new xcorr_cnn();
  c->device = device;
  c->max_batch = max_batch;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking);
  c->stream = c->own_stream;
  const size_t B = (size_t)max_batch;
  auto alloc = [&](void** p, size_t bytes) {
    if (e == hipSuccess) e = hipMalloc(p, bytes);
  };
  alloc((void**)&c->prm, kNParam * 4);
  alloc((void**)&c->m1, kNParam * 4);
  alloc((void**)&c->m2, kNParam * 4);
  alloc((void**)&c->grad, (kNParam + 1) * 4);
  alloc((void**)&c->q, kQn * 2);
  alloc((void**)&c->a1, B * 1024 * kC * 2);
  alloc((void**)&c->a2, B * 256 * kC * 2);
  alloc((void**)&c->a3, B * 64 * kC * 2);
  alloc((void**)&c->a4, B * 64 * kC * 2);
  alloc((void**)&c->d1, B * 1024 * kC * 2);
  alloc((void**)&c->d2, B * 256 * kC * 2);
  alloc((void**)&c->d3, B * 64 * kC * 2);
  alloc((void**)&c->d4, B * 64 * kC * 2);
  alloc((void**)&c->hout, B * 3 * 4);
  if (e != hipSuccess) {
    const std::string m = std::string("cnn_create: ") + hipGetErrorString(e);
    cnn_free(c);
    return cfail(nullptr, XCORR_HIP_ENOMEM, m);
  }
  *out = c;
  return 0;
}

extern "C" void xcorr_cnn_destroy(xcorr_cnn* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  cnn_free(c);
}

extern "C" const char* xcorr_cnn_last_error(const xcorr_cnn* c) { return c ? c->err.c_str() : g_cnn_err.c_str(); }
extern "C" void* xcorr_cnn_stream(xcorr_cnn* c) { return c ? (void*)c->stream : nullptr; }

extern "C" int xcorr_cnn_set_stream(xcorr_cnn* c, void* stream) {
  if (!c) return XCORR_HIP_EINVAL;
  c->stream = (hipStream_t)stream;          // NULL: the null stream
  return 0;
}

extern "C" int xcorr_cnn_sync(xcorr_cnn* c) {
  if (!c) return XCORR_HIP_EINVAL;
  CNNCK(c, hipStreamSynchronize(c->stream));
  return 0;
}

static int pack_weights(xcorr_cnn* c) {
  hipLaunchKernelGGL(k_pack, dim3((9 * kC * kC + 255) / 256), dim3(256), 0, c->stream, (const float*)c->prm, c->q);
  CNNCK(c, hipGetLastError());
  return 0;
}

extern "C" int xcorr_cnn_set_params(xcorr_cnn* c, const float* params) {
  if (!c || !params) return XCORR_HIP_EINVAL;
  CNNCK(c, hipSetDevice(c->device));
  CNNCK(c, hipMemcpyAsync(c->prm, params, kNParam * 4, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_zero, dim3((kNParam + 255) / 256), dim3(256), 0, c->stream, c->m1, kNParam);
  hipLaunchKernelGGL(k_zero, dim3((kNParam + 255) / 256), dim3(256), 0, c->stream, c->m2, kNParam);
  CNNCK(c, hipGetLastError());
  c->step = 0;
  const int rc = pack_weights(c);
  if (rc) return rc;
  CNNCK(c, hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int xcorr_cnn_get_params(xcorr_cnn* c, float* params) {
  if (!c || !params) return XCORR_HIP_EINVAL;
  CNNCK(c, hipSetDevice(c->device));
  CNNCK(c, hipMemcpyAsync(params, c->prm, kNParam * 4, hipMemcpyDeviceToHost, c->stream));
  CNNCK(c, hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int xcorr_cnn_box4(xcorr_cnn* c, const void* img, int nx, int ny, int format, float* out) {
  if (!c || !img || !out || nx < 4 || ny < 4 || (format != XCORR_HIP_CINT16 && format != XCORR_HIP_FLOAT32))
    return cfail(c, XCORR_HIP_EINVAL, "cnn_box4: bad arguments");
  CNNCK(c, hipSetDevice(c->device));
  const int dnx = nx / 4, dny = ny / 4;
  const long tot = (long)dnx * dny;
  hipLaunchKernelGGL(k_box4, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, c->stream, img, nx, ny, format, out,
                     dnx, dny);
  CNNCK(c, hipGetLastError());
  return 0;
}

static bool ds_ok(const float* p, int nx, int ny) { return p && nx > 0 && ny > 0; }

extern "C" int xcorr_cnn_crops(xcorr_cnn* c, const float* dsm, int dnx_m, int dny_m, const float* dss, int dnx_s,
                               int dny_s, const int32_t* corners, int n, uint16_t* x0) {
  if (!c || n < 0 || !ds_ok(dsm, dnx_m, dny_m) || !ds_ok(dss, dnx_s, dny_s) || (n > 0 && (!corners || !x0)))
    return cfail(c, XCORR_HIP_EINVAL, "cnn_crops: bad arguments");
  if (n == 0) return 0;
  CNNCK(c, hipSetDevice(c->device));
  hipLaunchKernelGGL(k_crops, dim3(n), dim3(256), 0, c->stream, dsm, dnx_m, dny_m, dss, dnx_s, dny_s,
                     (const int4*)corners, (bf16_t*)x0);
  CNNCK(c, hipGetLastError());
  return 0;
}

extern "C" int xcorr_cnn_infer(xcorr_cnn* c, const uint16_t* x0, int n, float* out) {
  if (!c || n < 0 || (n > 0 && (!x0 || !out))) return cfail(c, XCORR_HIP_EINVAL, "cnn_infer: bad arguments");
  if (n == 0) return 0;
  CNNCK(c, hipSetDevice(c->device));
  hipLaunchKernelGGL(k_infer<false>, dim3(n), dim3(256), 0, c->stream, (const bf16_t*)x0, (const float*)nullptr, 0, 0,
                     (const float*)nullptr, 0, 0, (const int4*)nullptr, (const bf16_t*)c->q, (const float*)c->prm, out);
  CNNCK(c, hipGetLastError());
  return 0;
}

extern "C" int xcorr_cnn_propose(xcorr_cnn* c, const float* dsm, int dnx_m, int dny_m, const float* dss, int dnx_s,
                                 int dny_s, const int32_t* corners, int n, float* out) {
  if (!c || n < 0 || !ds_ok(dsm, dnx_m, dny_m) || !ds_ok(dss, dnx_s, dny_s) || (n > 0 && (!corners || !out)))
    return cfail(c, XCORR_HIP_EINVAL, "cnn_propose: bad arguments");
  if (n == 0) return 0;
  CNNCK(c, hipSetDevice(c->device));
  hipLaunchKernelGGL(k_infer<true>, dim3(n), dim3(256), 0, c->stream, (const bf16_t*)nullptr, dsm, dnx_m, dny_m, dss,
                     dnx_s, dny_s, (const int4*)corners, (const bf16_t*)c->q, (const float*)c->prm, out);
  CNNCK(c, hipGetLastError());
  return 0;
}

// Training forward over nb <= max_batch samples into the activation buffers.
static int forward_layers(xcorr_cnn* c, const bf16_t* x0, int nb) {
  const dim3 blk(256);
  hipLaunchKernelGGL(k_conv1_fwd, dim3((nb * 64 + 3) / 4), blk, 0, c->stream, x0, (const bf16_t*)c->q,
                     (const float*)c->prm, c->a1, nb);
  hipLaunchKernelGGL((k_conv_fwd<32, 2>), dim3((nb * 16 + 3) / 4), blk, 0, c->stream, (const bf16_t*)c->a1,
                     (const bf16_t*)(c->q + kQf2), (const float*)(c->prm + kB2), c->a2, nb);
  hipLaunchKernelGGL((k_conv_fwd<16, 2>), dim3((nb * 4 + 3) / 4), blk, 0, c->stream, (const bf16_t*)c->a2,
                     (const bf16_t*)(c->q + kQf3), (const float*)(c->prm + kB3), c->a3, nb);
  hipLaunchKernelGGL((k_conv_fwd<8, 1>), dim3((nb * 4 + 3) / 4), blk, 0, c->stream, (const bf16_t*)c->a3,
                     (const bf16_t*)(c->q + kQf4), (const float*)(c->prm + kB4), c->a4, nb);
  CNNCK(c, hipGetLastError());
  return 0;
}

extern "C" int xcorr_cnn_forward(xcorr_cnn* c, const uint16_t* x0, int n, float* out) {
  if (!c || n < 0 || (n > 0 && (!x0 || !out))) return cfail(c, XCORR_HIP_EINVAL, "cnn_forward: bad arguments");
  CNNCK(c, hipSetDevice(c->device));
  for (int off = 0; off < n; off += c->max_batch) {
    const int nb = std::min(c->max_batch, n - off);
    int rc = forward_layers(c, (const bf16_t*)x0 + (size_t)off * 64 * 64 * 2, nb);
    if (rc) return rc;
    hipLaunchKernelGGL(k_head, dim3(nb), dim3(64), 0, c->stream, (const bf16_t*)c->a4, (const float*)c->prm,
                       (const float*)nullptr, out + (size_t)off * 3, (float*)nullptr, (float*)nullptr,
                       (bf16_t*)nullptr, nb, 1.0f / (float)nb);
    CNNCK(c, hipGetLastError());
  }
  return 0;
}

// Loss and gradient of one batch into c->grad (loss at grad[kNParam]).
static int backward(xcorr_cnn* c, const bf16_t* x0, const float* tgt, int nb, float* out) {
  const dim3 blk(256);
  hipLaunchKernelGGL(k_zero, dim3((kNParam + 1 + 255) / 256), blk, 0, c->stream, c->grad, kNParam + 1);
  int rc = forward_layers(c, x0, nb);
  if (rc) return rc;
  hipLaunchKernelGGL(k_head, dim3(nb), dim3(64), 0, c->stream, (const bf16_t*)c->a4, (const float*)c->prm, tgt,
                     out ? out : c->hout, c->grad, c->grad + kNParam, c->d4, nb, 1.0f / (float)nb);
  // weight-gradient blocks: enough to fill the chip, each over many 32-pixel chunks
  auto wblocks = [&](long px) { return (unsigned)std::max<long>(1, std::min<long>(1024, (px + 255) / 256)); };
  hipLaunchKernelGGL((k_conv_bwd_w<32, 8, 1>), dim3(wblocks((long)nb * 64)), blk, 0, c->stream, (const bf16_t*)c->d4,
                     (const bf16_t*)c->a3, c->grad + kW4, c->grad + kB4, nb);
  hipLaunchKernelGGL((k_conv_bwd_data<8, 1>), dim3((nb * 4 + 3) / 4), blk, 0, c->stream, (const bf16_t*)c->d4,
                     (const bf16_t*)c->a3, (const bf16_t*)(c->q + kQt4), c->d3, nb);
  hipLaunchKernelGGL((k_conv_bwd_w<32, 16, 2>), dim3(wblocks((long)nb * 64)), blk, 0, c->stream,
                     (const bf16_t*)c->d3, (const bf16_t*)c->a2, c->grad + kW3, c->grad + kB3, nb);
  hipLaunchKernelGGL((k_conv_bwd_data<16, 2>), dim3((nb * 16 + 3) / 4), blk, 0, c->stream, (const bf16_t*)c->d3,
                     (const bf16_t*)c->a2, (const bf16_t*)(c->q + kQt3), c->d2, nb);
  hipLaunchKernelGGL((k_conv_bwd_w<32, 32, 2>), dim3(wblocks((long)nb * 256)), blk, 0, c->stream,
                     (const bf16_t*)c->d2, (const bf16_t*)c->a1, c->grad + kW2, c->grad + kB2, nb);
  hipLaunchKernelGGL((k_conv_bwd_data<32, 2>), dim3((nb * 64 + 3) / 4), blk, 0, c->stream, (const bf16_t*)c->d2,
                     (const bf16_t*)c->a1, (const bf16_t*)(c->q + kQt2), c->d1, nb);
  hipLaunchKernelGGL((k_conv_bwd_w<2, 64, 2>), dim3(wblocks((long)nb * 1024)), blk, 0, c->stream,
                     (const bf16_t*)c->d1, x0, c->grad + kW1, c->grad + kB1, nb);
  CNNCK(c, hipGetLastError());
  return 0;
}

extern "C" int xcorr_cnn_gradient(xcorr_cnn* c, const uint16_t* x0, const float* tgt, int n, float* loss,
                                  float* grad, float* out) {
  if (!c || n <= 0 || n > c->max_batch || !x0 || !tgt) return cfail(c, XCORR_HIP_EINVAL, "cnn_gradient: bad arguments");
  CNNCK(c, hipSetDevice(c->device));
  int rc = backward(c, (const bf16_t*)x0, tgt, n, out);
  if (rc) return rc;
  std::vector<float> h(kNParam + 1);
  CNNCK(c, hipMemcpyAsync(h.data(), c->grad, (kNParam + 1) * 4, hipMemcpyDeviceToHost, c->stream));
  CNNCK(c, hipStreamSynchronize(c->stream));
  if (grad) memcpy(grad, h.data(), kNParam * 4);
  if (loss) *loss = h[kNParam];
  return 0;
}

extern "C" int xcorr_cnn_train_step(xcorr_cnn* c, const uint16_t* x0, const float* tgt, int n, float lr,
                                    float* loss) {
  if (!c || n <= 0 || n > c->max_batch || !x0 || !tgt || !(lr > 0.f))
    return cfail(c, XCORR_HIP_EINVAL, "cnn_train_step: bad arguments");
  CNNCK(c, hipSetDevice(c->device));
  int rc = backward(c, (const bf16_t*)x0, tgt, n, nullptr);
  if (rc) return rc;
  c->step++;
  const float bc1 = 1.0f - powf(0.9f, (float)c->step), bc2 = 1.0f - powf(0.999f, (float)c->step);
  hipLaunchKernelGGL(k_adam, dim3((kNParam + 255) / 256), dim3(256), 0, c->stream, c->prm, c->m1, c->m2,
                     (const float*)c->grad, kNParam, lr, bc1, bc2);
  CNNCK(c, hipGetLastError());
  rc = pack_weights(c);
  if (rc) return rc;
  if (loss) {
    CNNCK(c, hipMemcpyAsync(loss, c->grad + kNParam, 4, hipMemcpyDeviceToHost, c->stream));
    CNNCK(c, hipStreamSynchronize(c->stream));
  }
  return 0;
}
