// Register-resident length-256 transforms for gfx950.
//
// A transform is held by 16 lanes of a wave (a "group"): lane j keeps the 16
// values with indices j + 16 r.  The two passes are the 16 x 16 Stockham plan of
// fft_spec.hpp (pass 1: radix 16, L = 1; pass 2: radix 16, L = 16, twiddle
// w^(r j)), so results are bit-identical to lds_fft_rows for n = 256; only the
// data movement differs: one transpose through a padded 16 x 17 LDS tile per
// group instead of a full LDS round trip per pass.  A group never spans two
// waves, so the transposes synchronise the wave only (wave_sync): groups run
// their transforms independently, also under divergent control flow.
#pragma once
#include "fft_spec.hpp"

namespace xc {

constexpr int kT256Pitch = 17;               // float2 per row of a group's transpose tile
constexpr int kT256Tile = 16 * kT256Pitch;   // float2 per group

// Twiddles of pass 2 for lane j: w256^(r j), r = 1..15, from the e^{+2 pi i m/256} table.
__device__ __forceinline__ void load_tw256(const float2* __restrict__ tw, int j, float2 (&t)[15]) {
#pragma unroll
  for (int r = 1; r < 16; r++) t[r - 1] = tw[r * j];
}

// In:  v[r] = x[j + 16 r] (lane j of the group).  Out: v[r] = X[j + 16 r].
// `tile` is the group's private 16 x 17 scratch.
template <int DIR>
__device__ __forceinline__ void fft256_group(float2 (&v)[16], int j, float2* tile, const float2 (&t)[15]) {
  Dft<16, DIR>::run(v);                       // pass 1 -> y[16 j + r] = v[r]
#pragma unroll
  for (int r = 0; r < 16; r++) tile[j * kT256Pitch + r] = v[r];
  wave_sync();
#pragma unroll
  for (int r = 0; r < 16; r++) v[r] = tile[r * kT256Pitch + j];   // y[j + 16 r]
  wave_sync();
#pragma unroll
  for (int r = 1; r < 16; r++) {
    float2 w = t[r - 1];
    if (DIR < 0) w.y = -w.y;
    v[r] = cmul(v[r], w);
  }
  Dft<16, DIR>::run(v);                       // pass 2 -> X[j + 16 r] = v[r]
}

// The master-only spectrum of an unmasked row held as v[r] = Z[j + 16 r], the
// transform of z = amp1 + i amp2 (template columns): when `apply`,
// v[r] = extract_bin(Z[k], Z[256 - k]) for every k = j + 16 r; `tile` is the
// group's 16 x 17 scratch.
__device__ __forceinline__ void extract_real256(float2 (&v)[16], int j, float2* tile, bool apply) {
#pragma unroll
  for (int r = 0; r < 16; r++) tile[r * kT256Pitch + j] = v[r];
  wave_sync();
  if (apply) {
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const int m = (256 - j - 16 * r) & 255;
      v[r] = extract_bin(v[r], tile[(m >> 4) * kT256Pitch + (m & 15)]);
    }
  }
  wave_sync();
}

// An offset of 0 the compiler cannot see through (an empty asm statement, no
// instruction): LDS reads indexed with it stay where they are written instead
// of being hoisted out of loops into registers.
__device__ __forceinline__ int opaque_zero() {
  int z = 0;
  asm volatile("" : "+v"(z));
  return z;
}

// As fft256_group, with the pass-2 twiddles read from LDS at the point of use:
// twl[15 j + r - 1] = w256^(r j).
template <int DIR>
__device__ __forceinline__ void fft256_group_lds(float2 (&v)[16], int j, float2* tile, const float2* twl) {
  Dft<16, DIR>::run(v);
#pragma unroll
  for (int r = 0; r < 16; r++) tile[j * kT256Pitch + r] = v[r];
  wave_sync();
#pragma unroll
  for (int r = 0; r < 16; r++) v[r] = tile[r * kT256Pitch + j];
  wave_sync();
  const float2* tj = twl + 15 * j + opaque_zero();
#pragma unroll
  for (int r = 1; r < 16; r++) {
    float2 w = tj[r - 1];
    if (DIR < 0) w.y = -w.y;
    v[r] = cmul(v[r], w);
  }
  Dft<16, DIR>::run(v);
}

}  // namespace xc
