// Register-resident length-256 transforms for gfx950.
//
// A transform is held by 16 lanes of a wave (a "group"): lane j keeps the 16
// values with indices j + 16 r.  The two passes are the 16 x 16 Stockham plan of
// fft_spec.hpp (pass 1: radix 16, L = 1; pass 2: radix 16, L = 16, twiddle
// w^(r j)), so results are bit-identical to lds_fft_rows for n = 256; only the
// data movement differs: one transpose through a padded 16 x 17 LDS tile per
// group instead of a full LDS round trip per pass.
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
// `tile` is the group's private 16 x 17 scratch.  All 256 threads of the block
// must call this together (it contains block barriers).
template <int DIR>
__device__ __forceinline__ void fft256_group(float2 (&v)[16], int j, float2* tile, const float2 (&t)[15]) {
  Dft<16, DIR>::run(v);                       // pass 1 -> y[16 j + r] = v[r]
#pragma unroll
  for (int r = 0; r < 16; r++) tile[j * kT256Pitch + r] = v[r];
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 16; r++) v[r] = tile[r * kT256Pitch + j];   // y[j + 16 r]
  __syncthreads();
#pragma unroll
  for (int r = 1; r < 16; r++) {
    float2 w = t[r - 1];
    if (DIR < 0) w.y = -w.y;
    v[r] = cmul(v[r], w);
  }
  Dft<16, DIR>::run(v);                       // pass 2 -> X[j + 16 r] = v[r]
}

}  // namespace xc
