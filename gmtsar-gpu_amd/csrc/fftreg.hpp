// Register-resident transforms of length 128 and 512 for gfx950.
//
// A transform is held by a group of G lanes: lane j keeps the V = N / G values
// with indices j + G r.  The passes are exactly the Stockham passes of the
// default radix plan of fft_spec.hpp (128: 16 x 8, 512: 8 x 8 x 8) -- same
// butterflies, same twiddles, same operation order -- so results are
// bit-identical to lds_fft_rows; only the data movement differs: one LDS
// transpose per pass boundary in a padded per-group tile.  A group never
// spans two waves, so the transposes synchronise the wave only.
//
//   N = 128: G = 8 lanes x 16 values; pass 1 radix 16 (one butterfly per
//            lane), pass 2 radix 8 (two butterflies per lane: j and j + 8).
//   N = 512: G = 64 lanes (one wave) x 8 values; three radix-8 passes.
#pragma once
#include "fft_spec.hpp"

namespace xc {

template <int N>
struct RegFft;

template <>
struct RegFft<128> {
  static constexpr int G = 8, V = 16, TW = 14;
  static constexpr int TILE = 8 * 17;           // float2 per group
  // pass-2 twiddles: w^(r j) for butterfly j (t[r - 1]) and w^(r (j + 8)) (t[7 + r - 1])
  __device__ __forceinline__ static void load_tw(const float2* __restrict__ tw, int j, float2 (&t)[TW]) {
#pragma unroll
    for (int r = 1; r < 8; r++) {
      t[r - 1] = tw[r * j];
      t[7 + r - 1] = tw[r * (j + 8)];
    }
  }
  // natural-order slot of element i in the tile (spectrum extraction, line buffers)
  __device__ __forceinline__ static int slot(int i) { return i + (i >> 4); }
  template <int DIR>
  __device__ __forceinline__ static void run(float2 (&v)[V], int j, float2* tile, const float2 (&t)[TW]) {
    Dft<16, DIR>::run(v);                        // y[16 j + r] = v[r]
#pragma unroll
    for (int r = 0; r < 16; r++) tile[j * 17 + r] = v[r];
    wave_sync();
    float2 u0[8], u1[8];
#pragma unroll
    for (int r = 0; r < 8; r++) {                // y[j + 16 r], y[j + 8 + 16 r]
      u0[r] = tile[r * 17 + j];
      u1[r] = tile[r * 17 + j + 8];
    }
    wave_sync();
#pragma unroll
    for (int r = 1; r < 8; r++) {
      float2 w0 = t[r - 1], w1 = t[7 + r - 1];
      if (DIR < 0) {
        w0.y = -w0.y;
        w1.y = -w1.y;
      }
      u0[r] = cmul(u0[r], w0);
      u1[r] = cmul(u1[r], w1);
    }
    Dft<8, DIR>::run(u0);                        // X[j + 16 r]
    Dft<8, DIR>::run(u1);                        // X[j + 8 + 16 r]
#pragma unroll
    for (int q = 0; q < 8; q++) {
      v[2 * q] = u0[q];
      v[2 * q + 1] = u1[q];
    }
  }
};

template <>
struct RegFft<512> {
  static constexpr int G = 64, V = 8, TW = 14;
  static constexpr int TILE = 64 * 9;
  // pass 2: w^(8 r (j & 7)) (t[r - 1]); pass 3: w^(r j) (t[7 + r - 1])
  __device__ __forceinline__ static void load_tw(const float2* __restrict__ tw, int j, float2 (&t)[TW]) {
#pragma unroll
    for (int r = 1; r < 8; r++) {
      t[r - 1] = tw[8 * r * (j & 7)];
      t[7 + r - 1] = tw[r * j];
    }
  }
  __device__ __forceinline__ static int slot(int i) { return i + (i >> 3); }
  template <int DIR>
  __device__ __forceinline__ static void run(float2 (&v)[V], int j, float2* tile, const float2 (&t)[TW]) {
    const int hi = j >> 3, k = j & 7;
    Dft<8, DIR>::run(v);                         // y[8 j + r]
#pragma unroll
    for (int r = 0; r < 8; r++) tile[j * 9 + r] = v[r];
    wave_sync();
#pragma unroll
    for (int r = 0; r < 8; r++) v[r] = tile[(hi + 8 * r) * 9 + k];   // y[j + 64 r]
    wave_sync();
#pragma unroll
    for (int r = 1; r < 8; r++) {
      float2 w = t[r - 1];
      if (DIR < 0) w.y = -w.y;
      v[r] = cmul(v[r], w);
    }
    Dft<8, DIR>::run(v);                         // z[64 hi + 8 r + k]
#pragma unroll
    for (int r = 0; r < 8; r++) tile[(8 * hi + r) * 9 + k] = v[r];
    wave_sync();
#pragma unroll
    for (int r = 0; r < 8; r++) v[r] = tile[(hi + 8 * r) * 9 + k];   // z[j + 64 r]
    wave_sync();
#pragma unroll
    for (int r = 1; r < 8; r++) {
      float2 w = t[7 + r - 1];
      if (DIR < 0) w.y = -w.y;
      v[r] = cmul(v[r], w);
    }
    Dft<8, DIR>::run(v);                         // X[j + 64 r]
  }
};

// The master-only spectrum of an unmasked row held as v[r] = Z[j + G r] (the
// transform of z = amp1 + i amp2 on the template columns): when `apply`,
// v[r] = extract_bin(Z[k], Z[N - k]) for every k.
template <int N>
__device__ __forceinline__ void extract_real(float2 (&v)[RegFft<N>::V], int j, float2* tile, bool apply) {
  constexpr int G = RegFft<N>::G, V = RegFft<N>::V;
#pragma unroll
  for (int r = 0; r < V; r++) tile[RegFft<N>::slot(j + G * r)] = v[r];
  wave_sync();
  if (apply) {
#pragma unroll
    for (int r = 0; r < V; r++) {
      const int k = j + G * r;
      v[r] = extract_bin(v[r], tile[RegFft<N>::slot((N - k) & (N - 1))]);
    }
  }
  wave_sync();
}

}  // namespace xc
