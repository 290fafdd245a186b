// MI355X (gfx950) xcorr engine: kernels for the four stages of GMTSAR xcorr's
// offset-tracking path and the C ABI declared in include/xcorr_hip.h.
//
// Data flow per batch of B locations (all buffers device-resident):
//   k_range_ampl     stage (a)  window gather, range interpolation, amplitude,
//                               exact fixed-point patch sums
//   k_rowfft         stage (b)  demean/mask, z = c1 + i c2, FFT along x
//   k_colfft         stage (b)  FFT along y
//   k_product_colifft stage (b) C3 = C1 conj(C2) (-1)^(kx+ky), IFFT along y,
//                               keep the lag rows
//   k_rowifft_surf   stage (b)  IFFT along x, |c3| on the lag window, argmax
//   k_time_search    stage (c)  (-time/-time4) exact full search
//   k_finish         stage (c)+(d) NCC at the peak, high-res refinement, result
#include "../../include/xcorr_hip.h"
#include "fft_spec.hpp"
#include "fft256.hpp"
#include "fftreg.hpp"

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <type_traits>

#include <algorithm>
#include <climits>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

namespace xc {

constexpr int kThreads = 256;
constexpr int kMaxThreads = 512;  // generic LDS row/column kernels: launched with 64..512 threads
constexpr int kTile = 2048;     // complex elements per LDS FFT buffer (default tile)

struct DevParams {
  int NX, NY, N;
  int xs, ys, nx_corr, ny_corr, nxc, nyc;
  int ri, interp_flag, ifc, n2x, n2y, Lx, Ly;
  int corr_flag, format;
  int G1, G4, C3;                 // rows per block (k_range_ampl/k_rowfft, k_rowifft_surf), columns per block
  FftPlan plx, ply;               // radix plans of the generic LDS transforms along x (NX) and y (NY)
  int groups4;                    // k_rowifft_surf blocks per patch
  int tLR, tRC;                   // k_time_search: surface rows per block, template rows per LDS chunk
  int tw_nx, tw_ny, tw_lx, tw_ly, phi;   // offsets into the table buffer
  int e4;                         // high-res row factors e^{2 pi i k s / Ly}, [Ly/2 + 1][4] (k = 1..4)
  int mk_x, mk_y;                 // spectra of the 1-D mask boxes (npx, npy entries)
  int m_nx, m_ny, s_nx, s_ny;
  // lines held in device memory (all of each image, or a streamed band):
  // addresses are only ever formed for lines in [lo, hi)
  int m_lo, m_hi, s_lo, s_hi;
  // 1: the correlation surface is written whole (stage dumps); 0: the row
  // kernels that find the peak in one block write only its neighbourhood
  int keep_surf;
};

// --------------------------------------------------------------------------
// reductions
// --------------------------------------------------------------------------
__device__ __forceinline__ long long wave_sum_i64(long long v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
  for (int o = 32; o > 0; o >>= 1) {
    unsigned long long w = __shfl_xor(v, o);
    v = w > v ? w : v;
  }
  return v;
}
__device__ __forceinline__ double wave_sum_f64(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
// Block-wide exact int64 sum; result valid in every thread.
__device__ long long block_sum_i64(long long v, long long* red) {
  v = wave_sum_i64(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  long long t = 0;
  for (int w = 0; w < nw; w++) t += red[w];
  return t;
}
__device__ unsigned long long block_max_u64(unsigned long long v, unsigned long long* red) {
  v = wave_max_u64(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  unsigned long long t = 0;
  for (int w = 0; w < nw; w++) t = red[w] > t ? red[w] : t;
  return t;
}
__device__ double block_sum_f64(double v, double* red) {
  v = wave_sum_f64(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < nw; w++) t += red[w];
  return t;
}

// argmax key: first-index maximum of strictly positive values (the scan
// "if (v > best)" from best = 0, index 0).  Non-positive and NaN give 0.
__device__ __forceinline__ unsigned long long peak_key(float v, unsigned idx) {
  return v > 0.0f ? ((unsigned long long)__float_as_uint(v) << 32) | (unsigned long long)(~idx) : 0ull;
}
__device__ __forceinline__ void key_decode(unsigned long long key, float* v, unsigned* idx) {
  if (key == 0ull) { *v = 0.0f; *idx = 0u; return; }
  *v = __uint_as_float((unsigned)(key >> 32));
  *idx = ~(unsigned)(key & 0xffffffffull);
}

// trunc(a * 2^20) for 0 <= a < 2^43, exactly: x = h 2^32 + l with h = trunc(x 2^-32),
// l = x - h 2^32 (both steps exact), so trunc(x) = h 2^32 + trunc(l).
__device__ __forceinline__ long long fxp20(float a) {
  const float x = a * 1048576.0f;
  const float h = truncf(x * 2.3283064365386963e-10f);
  const float l = __builtin_fmaf(-h, 4294967296.0f, x);
  return (long long)(((unsigned long long)(unsigned)h << 32) + (unsigned long long)(unsigned)l);
}

// four floats at a 4-byte aligned address (one global_load_dwordx4)
typedef float f4a4 __attribute__((ext_vector_type(4), aligned(4)));

// A streaming store (nontemporal hint): data the next kernels read long after
// many other lines have passed through the cache.
__device__ __forceinline__ void st_nt(float2* p, float2 v) {
  typedef float f2v __attribute__((ext_vector_type(2)));
  __builtin_nontemporal_store(f2v{v.x, v.y}, (f2v*)p);
}

__device__ __forceinline__ double patch_mean(unsigned long long s, int N) {
  return (double)(long long)s * (1.0 / 1048576.0) / (double)N;
}

// Work item of this block for a grid whose size is a multiple of 8 (items
// beyond the real count are padding the caller skips).  Workgroups go to the
// 8 XCDs round-robin in blockIdx order; block i takes item
// (i mod 8) (grid / 8) + i / 8, so each XCD runs one contiguous range of
// items in order, and neighbouring items (patches of one strip, whose rows
// overlap) meet in one L2.
__device__ __forceinline__ int xcd_item() {
  return (int)(blockIdx.x & 7) * (int)(gridDim.x >> 3) + (int)(blockIdx.x >> 3);
}

// --------------------------------------------------------------------------
// stage (a): gather + range interpolation + amplitude + fixed-point sums
// xcorr.cu:237-292, fft_interpolate_routines.cu:76-113, conv_ampl_demean.cu:98-108, 703-726
// --------------------------------------------------------------------------
__global__ void __launch_bounds__(kMaxThreads) k_range_ampl(DevParams P, const void* __restrict__ master,
                                                         const void* __restrict__ slave,
                                                         const xcorr_hip_window* __restrict__ win,
                                                         float* __restrict__ amp,
                                                         unsigned long long* __restrict__ sums,
                                                         const float2* __restrict__ tabs) {
  extern __shared__ float2 smem[];
  __shared__ long long red[kMaxThreads / 64];
  const int NX = P.NX, G = P.G1, nel = G * NX;
  const int groups = P.NY / G;
  const int g = blockIdx.x % groups;
  const int rest = blockIdx.x / groups;
  const int img = rest & 1;
  const int b = rest >> 1;
  const xcorr_hip_window w = win[b];
  const int x0 = img ? w.sx0 : w.mx0;
  const int y0 = (img ? w.sy0 : w.my0) + g * G;
  const int inx = img ? P.s_nx : P.m_nx, iny = img ? P.s_ny : P.m_ny;
  const void* src = img ? slave : master;
  const int ri = P.ri;
  float2* A = smem;
  float2* B = A + nel;
  float2* Cb = B + nel;                       // only with ri > 2
  float* ab = (float*)(smem + (ri > 2 ? 3 : 2) * nel);

  for (int e = threadIdx.x; e < nel; e += blockDim.x) {
    const int row = e / NX, x = e - row * NX;
    const int gy = y0 + row, gx = x0 + x;
    float2 v = make_float2(0.0f, 0.0f);
    if (gy >= 0 && gy < iny && gx >= 0 && gx < inx) {
      const size_t o = (size_t)gy * inx + gx;
      if (P.format == XCORR_HIP_CINT16) {
        const unsigned u = ((const unsigned*)src)[o];
        v.x = (float)(short)(u & 0xffffu);
        v.y = (float)(short)(u >> 16);
      } else {
        v.x = ((const float*)src)[o];
      }
    }
    A[e] = v;
  }
  __syncthreads();
  if (ri == 1) {
    for (int e = threadIdx.x; e < nel; e += blockDim.x) ab[e] = ampl(A[e]);
  } else {
    for (int e = threadIdx.x; e < nel; e += blockDim.x) {
      const int row = e / NX, j = e - row * NX;
      const int t = NX / 2 + j;
      if (t % ri == 0) ab[e] = ampl(A[row * NX + t / ri]);
    }
    const float2* twn = tabs + P.tw_nx;
    float2* X = lds_fft_rows<-1>(A, B, NX, G, twn, P.plx);
    float2* other = (X == A) ? B : A;
    const float sc = 1.0f / (float)NX;
    for (int r = 1; r < ri; r++) {
      const float2* phi = tabs + P.phi + (size_t)(r - 1) * NX;
      float2* Y = other;
      float2* Ys = (ri > 2) ? Cb : X;         // scratch for the inverse passes
      for (int e = threadIdx.x; e < nel; e += blockDim.x) {
        const int k = e & (NX - 1);
        Y[e] = cmul(X[e], phi[k]);
      }
      __syncthreads();
      float2* Rr = lds_fft_rows<+1>(Y, Ys, NX, G, twn, P.plx);
      for (int e = threadIdx.x; e < nel; e += blockDim.x) {
        const int row = e / NX, j = e - row * NX;
        const int t = NX / 2 + j;
        if (t % ri == r) {
          float2 v = Rr[row * NX + t / ri];
          v.x = v.x * sc;
          v.y = v.y * sc;
          ab[e] = ampl(v);
        }
      }
      __syncthreads();
    }
  }
  __syncthreads();
  float* out = amp + ((size_t)(b * 2 + img) * P.NY + (size_t)g * G) * NX;
  long long s = 0;
  for (int e = threadIdx.x; e < nel; e += blockDim.x) {
    const float a = ab[e];
    out[e] = a;
    s += fxp20(a);
  }
  s = block_sum_i64(s, red);
  if (threadIdx.x == 0) atomicAdd(&sums[b * 2 + img], (unsigned long long)s);
}

// --------------------------------------------------------------------------
// 256-sample fast paths (npx = 256 for the row kernels, npy = 256 for the
// column kernels).  Same arithmetic as the generic kernels, register-resident
// transforms (fft256.hpp); one block per patch (or per 16-column tile) so the
// per-lane twiddles and phase factors are loaded once.
// --------------------------------------------------------------------------

// One line of stage (a) for npx = 256, ri = 2, held by a 16-lane group: gather,
// range interpolation, amplitudes.  On return ea[q] / oa[q] are the amplitudes of
// output columns 2q' - 128 and 2q' - 127, q' = j + 16 (q + 4): the original
// sample q' and the half-sample interpolant at q' + 1/2.
__device__ __forceinline__ void range_row256(const void* __restrict__ src, int format, int inx, int iny, bool inside,
                                             int gy, int x0, int j, float2* tile, const float2 (&t)[15],
                                             const float2 (&ph)[16], float (&ea)[8], float (&oa)[8]) {
  float2 v[16];
  if (format == XCORR_HIP_CINT16) {
    const unsigned* im = (const unsigned*)src;
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const int gx = x0 + j + 16 * r;
        unsigned u = 0u;
        if (inside || (gy >= 0 && gy < iny && gx >= 0 && gx < inx)) u = im[(size_t)gy * inx + gx];
        v[r] = make_float2((float)(short)(u & 0xffffu), (float)(short)(u >> 16));
      }
  } else {
    const float* im = (const float*)src;
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const int gx = x0 + j + 16 * r;
      float u = 0.0f;
      if (inside || (gy >= 0 && gy < iny && gx >= 0 && gx < inx)) u = im[(size_t)gy * inx + gx];
      v[r] = make_float2(u, 0.0f);
    }
  }
#pragma unroll
  for (int q = 0; q < 8; q++) ea[q] = ampl(v[q + 4]);
  fft256_group<-1>(v, j, tile, t);
#pragma unroll
  for (int r = 0; r < 16; r++) v[r] = cmul(v[r], ph[r]);
  fft256_group<+1>(v, j, tile, t);
  const float sc = 1.0f / 256.0f;
#pragma unroll
  for (int q = 0; q < 8; q++) {
    float2 o = v[q + 4];
    o.x = o.x * sc;
    o.y = o.y * sc;
    oa[q] = ampl(o);
  }
}

// Stage (a) for npx = 256, ri = 2 (time modes): one block per (patch, image),
// 16 groups of 16 lanes, one line per group and iteration.
__global__ void __launch_bounds__(kThreads) k_range256(DevParams P, const void* __restrict__ master,
                                                       const void* __restrict__ slave,
                                                       const xcorr_hip_window* __restrict__ win,
                                                       float* __restrict__ amp,
                                                       unsigned long long* __restrict__ sums,
                                                       const float2* __restrict__ tabs) {
  __shared__ float2 tiles[16 * kT256Tile];
  __shared__ long long red[kThreads / 64];
  const int g = threadIdx.x >> 4, j = threadIdx.x & 15;
  const int img = blockIdx.x & 1, b = blockIdx.x >> 1;
  const xcorr_hip_window w = win[b];
  const int x0 = img ? w.sx0 : w.mx0, y0 = img ? w.sy0 : w.my0;
  const int inx = img ? P.s_nx : P.m_nx, iny = img ? P.s_ny : P.m_ny;
  const void* src = img ? slave : master;
  float2* tile = tiles + g * kT256Tile;
  float2 t[15];
  load_tw256(tabs + P.tw_nx, j, t);
  float2 ph[16];
#pragma unroll
  for (int r = 0; r < 16; r++) ph[r] = tabs[P.phi + j + 16 * r];
  const bool inside = x0 >= 0 && x0 + 256 <= inx && y0 >= 0 && y0 + P.NY <= iny;
  float2* outb = (float2*)(amp + (size_t)(2 * b + img) * P.N);
  long long s = 0;
  for (int row = g; row < P.NY; row += 16) {
    float ea[8], oa[8];
    range_row256(src, P.format, inx, iny, inside, y0 + row, x0, j, tile, t, ph, ea, oa);
#pragma unroll
    for (int q = 0; q < 8; q++) {
      outb[(size_t)row * 128 + j + 16 * q] = make_float2(ea[q], oa[q]);
      s += fxp20(ea[q]) + fxp20(oa[q]);
    }
  }
  s = block_sum_i64(s, red);
  if (threadIdx.x == 0) sums[2 * b + img] = (unsigned long long)s;
}

// Stage (a) and the row transform of stage (b) for npx = 256, ri = 2, frequency
// mode: one block per patch.  For each line: both images' range interpolation
// and amplitudes, then z = amp1 + i amp2 mask and its FFT along x (the means
// come off later, in the spectrum).  Amplitudes are kept for the NCC: the
// master's everywhere, the slave's on the template.
__global__ void __launch_bounds__(kThreads, 3) k_stageA256(DevParams P, const void* __restrict__ master,
                                                        const void* __restrict__ slave,
                                                        const xcorr_hip_window* __restrict__ win,
                                                        float* __restrict__ amp,
                                                        unsigned long long* __restrict__ sums,
                                                        float2* __restrict__ Z, const float2* __restrict__ tabs) {
  __shared__ float2 tiles[16 * kT256Tile];
  __shared__ float2 mrow[16 * 128];                // each group's master amplitude line (column pairs)
  __shared__ long long red[kThreads / 64];
  const int g = threadIdx.x >> 4, j = threadIdx.x & 15;
  const int b = blockIdx.x;
  const xcorr_hip_window w = win[b];
  float2* tile = tiles + g * kT256Tile;
  float2* mr = mrow + g * 128;
  float2 t[15];
  load_tw256(tabs + P.tw_nx, j, t);
  float2 ph[16];
#pragma unroll
  for (int r = 0; r < 16; r++) ph[r] = tabs[P.phi + j + 16 * r];
  const int NY = P.NY;
  const bool in_m = w.mx0 >= 0 && w.mx0 + 256 <= P.m_nx && w.my0 >= 0 && w.my0 + NY <= P.m_ny;
  const bool in_s = w.sx0 >= 0 && w.sx0 + 256 <= P.s_nx && w.sy0 >= 0 && w.sy0 + NY <= P.s_ny;
  float2* out1 = (float2*)(amp + (size_t)(2 * b) * P.N);
  float2* out2 = (float2*)(amp + (size_t)(2 * b + 1) * P.N);
  float2* zb = Z + (size_t)b * P.N;
  long long s1 = 0, s2 = 0;
  for (int row = g; row < NY; row += 16) {
    {
      float ea[8], oa[8];
      range_row256(master, P.format, P.m_nx, P.m_ny, in_m, w.my0 + row, w.mx0, j, tile, t, ph, ea, oa);
#pragma unroll
      for (int q = 0; q < 8; q++) {
        const float2 a = make_float2(ea[q], oa[q]);
        out1[(size_t)row * 128 + j + 16 * q] = a;
        mr[j + 16 * q] = a;
        s1 += fxp20(ea[q]) + fxp20(oa[q]);
      }
    }
    {
      float ea[8], oa[8];
      range_row256(slave, P.format, P.s_nx, P.s_ny, in_s, w.sy0 + row, w.sx0, j, tile, t, ph, ea, oa);
      const bool trow = row >= P.ys && row < P.ys + P.ny_corr;
#pragma unroll
      for (int q = 0; q < 8; q++) {
        const float2 a = make_float2(ea[q], oa[q]);
        if (trow && q >= 2 && q < 6) out2[(size_t)row * 128 + j + 16 * q] = a;
        tile[j + 16 * q] = a;                      // the tile doubles as the slave line buffer
        s2 += fxp20(ea[q]) + fxp20(oa[q]);
      }
    }
    wave_sync();                                   // the group's two lines are in LDS
    const float* m1 = (const float*)mr;
    const float* a2 = (const float*)tile;
    const bool in_y = row >= P.ys && row < NY - P.ys;
    float2 v[16];
#pragma unroll
    for (int r = 0; r < 16; r++) {              // every row carries the slave's template columns
      const int x = j + 16 * r;
      v[r] = make_float2(m1[x], (x >= P.xs && x < 256 - P.xs) ? a2[x] : 0.0f);
    }
    wave_sync();                                   // read before the transform reuses the tile
    fft256_group<-1>(v, j, tile, t);
    extract_real256(v, j, tile, !in_y);           // unmasked rows: the master's spectrum
#pragma unroll
    for (int r = 0; r < 16; r++) zb[(size_t)row * 256 + j + 16 * r] = v[r];
  }
  s1 = block_sum_i64(s1, red);
  s2 = block_sum_i64(s2, red);
  if (threadIdx.x == 0) {
    sums[2 * b] = (unsigned long long)s1;
    sums[2 * b + 1] = (unsigned long long)s2;
  }
}

// --------------------------------------------------------------------------
// Register-resident stage (a) + row transform for npx = npy = N in {128, 512},
// ri = 2, frequency mode (the arithmetic of k_range_ampl + k_rowfft).
// --------------------------------------------------------------------------
// One line of stage (a) held by a group of G lanes (v[r] = line sample
// x0 + j + G r): ea[q] / oa[q] are the amplitudes of output columns 2c and
// 2c + 1, c = j + G q (q < V / 2): the original sample c + N / 4 and the
// half-sample interpolant after it.
// sph: the phase factors e^{2 pi i k~ / 2N} in LDS, read at the point of use.
template <int N>
__device__ __forceinline__ void range_row(const void* __restrict__ src, int format, int inx, int iny, bool inside,
                                          int gy, int x0, int j, float2* tile,
                                          const float2 (&t)[RegFft<N>::TW], const float2* sph,
                                          float (&ea)[RegFft<N>::V / 2], float (&oa)[RegFft<N>::V / 2]) {
  constexpr int G = RegFft<N>::G, V = RegFft<N>::V;
  float2 v[V];
  if (format == XCORR_HIP_CINT16) {
    const unsigned* im = (const unsigned*)src;
#pragma unroll
    for (int r = 0; r < V; r++) {
      const int gx = x0 + j + G * r;
      unsigned u = 0u;
      if (inside || (gy >= 0 && gy < iny && gx >= 0 && gx < inx)) u = im[(size_t)gy * inx + gx];
      v[r] = make_float2((float)(short)(u & 0xffffu), (float)(short)(u >> 16));
    }
  } else {
    const float* im = (const float*)src;
#pragma unroll
    for (int r = 0; r < V; r++) {
      const int gx = x0 + j + G * r;
      float u = 0.0f;
      if (inside || (gy >= 0 && gy < iny && gx >= 0 && gx < inx)) u = im[(size_t)gy * inx + gx];
      v[r] = make_float2(u, 0.0f);
    }
  }
#pragma unroll
  for (int q = 0; q < V / 2; q++) ea[q] = ampl(v[q + V / 4]);
  RegFft<N>::template run<-1>(v, j, tile, t);
  const float2* phj = sph + j + opaque_zero();
#pragma unroll
  for (int r = 0; r < V; r++) v[r] = cmul(v[r], phj[G * r]);
  RegFft<N>::template run<+1>(v, j, tile, t);
  const float sc = 1.0f / (float)N;
#pragma unroll
  for (int q = 0; q < V / 2; q++) {
    float2 o = v[q + V / 4];
    o.x = o.x * sc;
    o.y = o.y * sc;
    oa[q] = ampl(o);
  }
}

// One block per patch, 256 / G groups, one line per group and iteration: both
// images' range interpolation and amplitudes (kept for the NCC: the master's
// everywhere, the slave's on the template), exact fixed-point sums, then
// z = amp1 + i amp2 mask and its FFT along x (the means come off in the
// spectrum, demean_row_bin); unmasked rows keep the master's spectrum extracted from the packed transform.
template <int N, bool ROWFFT>
__global__ void __launch_bounds__(kThreads) k_stageA(DevParams P, const void* __restrict__ master,
                                                     const void* __restrict__ slave,
                                                     const xcorr_hip_window* __restrict__ win,
                                                     float* __restrict__ amp, unsigned long long* __restrict__ sums,
                                                     float2* __restrict__ Z, const float2* __restrict__ tabs) {
  using F = RegFft<N>;
  constexpr int G = F::G, V = F::V, GPB = kThreads / G;
  __shared__ float2 tiles[GPB * F::TILE];
  __shared__ float2 mrow[GPB * (N / 2)];           // each group's master amplitude line (column pairs)
  __shared__ float2 sph[N];                         // range-interpolation phase factors
  __shared__ long long red[kThreads / 64];
  const int g = threadIdx.x / G, j = threadIdx.x % G;
  const int b = blockIdx.x;
  const xcorr_hip_window w = win[b];
  float2* tile = tiles + g * F::TILE;
  float2* mr = mrow + g * (N / 2);
  float2 t[F::TW];
  F::load_tw(tabs + P.tw_nx, j, t);
  for (int k = threadIdx.x; k < N; k += kThreads) sph[k] = tabs[P.phi + k];
  __syncthreads();
  const bool in_m = w.mx0 >= 0 && w.mx0 + N <= P.m_nx && w.my0 >= 0 && w.my0 + N <= P.m_ny;
  const bool in_s = w.sx0 >= 0 && w.sx0 + N <= P.s_nx && w.sy0 >= 0 && w.sy0 + N <= P.s_ny;
  float2* out1 = (float2*)(amp + (size_t)(2 * b) * P.N);
  float2* out2 = (float2*)(amp + (size_t)(2 * b + 1) * P.N);
  float2* zb = ROWFFT ? Z + (size_t)b * P.N : nullptr;
  const int c_lo = P.xs / 2, c_hi = (P.xs + P.nx_corr) / 2;   // template column pairs
  long long s1 = 0, s2 = 0;
  for (int row = g; row < N; row += GPB) {        // groups never wait for other waves inside the loop
    {
      float ea[V / 2], oa[V / 2];
      range_row<N>(master, P.format, P.m_nx, P.m_ny, in_m, w.my0 + row, w.mx0, j, tile, t, sph, ea, oa);
#pragma unroll
      for (int q = 0; q < V / 2; q++) {
        const float2 a = make_float2(ea[q], oa[q]);
        out1[(size_t)row * (N / 2) + j + G * q] = a;
        if (ROWFFT) mr[j + G * q] = a;
        s1 += fxp20(ea[q]) + fxp20(oa[q]);
      }
    }
    {
      float ea[V / 2], oa[V / 2];
      range_row<N>(slave, P.format, P.s_nx, P.s_ny, in_s, w.sy0 + row, w.sx0, j, tile, t, sph, ea, oa);
      const bool trow = row >= P.ys && row < P.ys + P.ny_corr;
#pragma unroll
      for (int q = 0; q < V / 2; q++) {
        const float2 a = make_float2(ea[q], oa[q]);
        const int c = j + G * q;
        if (trow && c >= c_lo && c < c_hi) out2[(size_t)row * (N / 2) + c] = a;
        if (ROWFFT) tile[c] = a;                   // the tile doubles as the slave line buffer
        s2 += fxp20(ea[q]) + fxp20(oa[q]);
      }
    }
    if (!ROWFFT) continue;                         // time modes: amplitudes and sums only
    wave_sync();                                   // the group's line buffers are its own
    const float* m1 = (const float*)mr;
    const float* a2 = (const float*)tile;
    const bool in_y = row >= P.ys && row < N - P.ys;
    float2 v[V];
#pragma unroll
    for (int r = 0; r < V; r++) {                  // every row carries the slave's template columns
      const int x = j + G * r;
      v[r] = make_float2(m1[x], (x >= P.xs && x < N - P.xs) ? a2[x] : 0.0f);
    }
    wave_sync();
    F::template run<-1>(v, j, tile, t);
    extract_real<N>(v, j, tile, !in_y);            // unmasked rows: the master's spectrum
#pragma unroll
    for (int r = 0; r < V; r++) zb[(size_t)row * N + j + G * r] = v[r];
  }
  s1 = block_sum_i64(s1, red);
  s2 = block_sum_i64(s2, red);
  if (threadIdx.x == 0) {
    sums[2 * b] = (unsigned long long)s1;
    sums[2 * b + 1] = (unsigned long long)s2;
  }
}

// --------------------------------------------------------------------------
// Strips (square windows of 128, 256, 512 samples, ri = 2; frequency mode,
// and -time on 128 x 128 without the row transforms).  Windows that share the
// master column, the slave column and the slave line offset see the same
// range-interpolated lines, so a strip computes each line once: amplitudes,
// their exact fixed-point sum, and the row transforms of z = amp1 + i amp2
// mask for the two mask states a line can have inside a patch (masked rows
// [ys, npy - ys), unmasked rows elsewhere).  A patch is a run of npy strip
// lines starting at its row base.  The arithmetic of every value is that of
// the per-patch path.
// --------------------------------------------------------------------------
constexpr int kStripBlockLines = 128;     // strip lines per block (16 groups x 8; 64 and 256 measured slower)

// Raw samples x0 + j + 16 r of line gy (one dword per sample in both formats),
// zero outside the image.  Branch-free: the address is clamped into the image
// (so every load is in bounds) and the value is zeroed afterwards, which lets
// all 16 loads issue back to back.
// EDGE false: columns [x0, x0 + 256) lie in the image.  Lines [lo, hi) are
// held in device memory; the clamped address always falls among them.
template <bool EDGE>
__device__ __forceinline__ void load_line256(const unsigned* __restrict__ im, int inx, int iny, int lo, int hi, int gy,
                                             int x0, int j, unsigned (&u)[16]) {
  const bool oky = (unsigned)gy < (unsigned)iny;
  const unsigned* line = im + (ptrdiff_t)min(max(gy, lo), hi - 1) * inx;
  if (!EDGE) {
    const unsigned* p = line + x0 + j;
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const unsigned w = p[16 * r];
      u[r] = oky ? w : 0u;
    }
  } else {
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const int gx = x0 + j + 16 * r;
      const unsigned w = line[min(max(gx, 0), inx - 1)];
      u[r] = (oky && (unsigned)gx < (unsigned)inx) ? w : 0u;
    }
  }
}

// Range interpolation of one held line (range_row256 after the gather), with
// twiddles and phase factors from LDS.  sph holds phi[0..128]; the rest follow
// from phi[k] = conj(phi[256 - k]) (exact: the table entries are built from
// the same angles).
template <int FMT>
__device__ __forceinline__ void interp_line256(const unsigned (&u)[16], int j, float2* tile, const float2* twl,
                                               const float2* sph, float (&ea)[8], float (&oa)[8]) {
  float2 v[16];
#pragma unroll
  for (int r = 0; r < 16; r++)
    v[r] = FMT == XCORR_HIP_CINT16 ? make_float2((float)(short)(u[r] & 0xffffu), (float)(short)(u[r] >> 16))
                                   : make_float2(__uint_as_float(u[r]), 0.0f);
#pragma unroll
  for (int q = 0; q < 8; q++) ea[q] = ampl(v[q + 4]);
  fft256_group_lds<-1>(v, j, tile, twl);
  sph += opaque_zero();
#pragma unroll
  for (int r = 0; r < 16; r++) {
    float2 ph;
    if (r < 8) {
      ph = sph[j + 16 * r];
    } else if (r == 8) {
      ph = sph[128 - j];
      if (j != 0) ph.y = -ph.y;
    } else {
      ph = sph[256 - j - 16 * r];
      ph.y = -ph.y;
    }
    v[r] = cmul(v[r], ph);
  }
  fft256_group_lds<+1>(v, j, tile, twl);
  const float sc = 1.0f / 256.0f;
#pragma unroll
  for (int q = 0; q < 8; q++) {
    float2 o = v[q + 4];
    o.x = o.x * sc;
    o.y = o.y * sc;
    oa[q] = ampl(o);
  }
}

struct StripDesc {
  int mx0, my0, sx0, sy0;                 // master / slave column and line of strip line 0
  int line0, nlines;                      // first global line (a multiple of kStripBlockLines), count
  int pad0, pad1;
};

// needM[l] / needU[l] = 1 when some patch of the batch holds line l as a masked / unmasked row.
__global__ void k_need_clear(unsigned char* __restrict__ needM, unsigned char* __restrict__ needU, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    needM[i] = 0;
    needU[i] = 0;
  }
}
__global__ void k_need_mark(DevParams P, const int* __restrict__ rb, unsigned char* __restrict__ needM,
                            unsigned char* __restrict__ needU, int nb) {
  const int b = blockIdx.x;
  if (b >= nb) return;
  for (int y = threadIdx.x; y < P.NY; y += blockDim.x) {
    const size_t l = (size_t)rb[b] + y;
    if (y >= P.ys && y < P.NY - P.ys) needM[l] = 1;
    else needU[l] = 1;
  }
}

// A 64-bit value moved between lanes by DPP control CTRL (rows in ROWS);
// lanes whose source lies outside their row of 16 get 0.
template <int CTRL, int ROWS>
__device__ __forceinline__ long long dpp_i64(long long v) {
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)v, CTRL, ROWS, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(v >> 32), CTRL, ROWS, 0xf, true);
  return (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  return __builtin_bit_cast(double, dpp_i64<CTRL, 0xf>(__builtin_bit_cast(long long, v)));
}

// Sums over a group of G consecutive lanes (G a power of two <= 64); the total
// is valid in the group's last lane.  A group of 16 is one DPP row: an
// inclusive scan by row shifts of 1, 2, 4 and 8 (register moves, no LDS round
// trip) leaves the total in lane 15; other G take an xor butterfly (every
// lane gets the total).  The summands are integers whose partial sums are
// exact, so the order does not change the result.
template <int G>
__device__ __forceinline__ long long group_sum_i64(long long v) {
  if (G == 16) {
    v += dpp_i64<0x111, 0xf>(v);                   // row_shr:1
    v += dpp_i64<0x112, 0xf>(v);                   // row_shr:2
    v += dpp_i64<0x114, 0xf>(v);                   // row_shr:4
    v += dpp_i64<0x118, 0xf>(v);                   // row_shr:8
    return v;
  }
  for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
template <int G>
__device__ __forceinline__ double group_sum_f64(double v) {
  if (G == 16) {
    v += dpp_f64<0x111>(v);
    v += dpp_f64<0x112>(v);
    v += dpp_f64<0x114>(v);
    v += dpp_f64<0x118>(v);
    return v;
  }
  for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Exact fixed-point sum of the amplitudes a group of G lanes holds for one
// line, as sum_k trunc(a_k 2^20).  Complex int16 samples have |x| < 46341;
// their half-sample interpolants exceed that by at most the Lebesgue constant
// of trigonometric interpolation ((2/pi) ln N + 1.1 < 5.1 for N <= 512), so
// a < 2^18, every term is an integer below 2^38 and a line sum of up to 512
// terms stays below 2^53: FP64 accumulation is exact in any order.  Real
// float samples are unbounded and take the int64 route.
template <int FMT, int G = 16>
struct LineSum {
  double d = 0.0;
  long long i = 0;
  __device__ __forceinline__ void add(float a) {
    if (FMT == XCORR_HIP_CINT16) d += (double)truncf(a * 1048576.0f);
    else i += fxp20(a);
  }
  __device__ __forceinline__ long long group_total() {
    if (FMT == XCORR_HIP_CINT16) return (long long)group_sum_f64<G>(d);
    return group_sum_i64<G>(i);
  }
};

// One block per kStripBlockLines strip lines; 16 groups of 16 lanes, one line
// per group and iteration.  Out: amp1 (all columns), amp2 (template columns of
// masked lines), lsum (per-line fixed-point sums), Zm (the row spectrum of
// z = amp1 + i amp2 on the template columns, where some patch uses the line).
// Work item b < blk_n (blocks in xcd_item order) covers kStripBlockLines
// lines from global line blk[2 (b + blk_off) + 1] of strip blk[2 (b + blk_off)].
// EDGE: the strips' windows may cross the image sides.  Without ROWFFT (the
// time modes) amplitudes and line sums only.
template <int FMT, bool EDGE, bool ROWFFT = true>
__global__ void __launch_bounds__(kThreads, EDGE ? 2 : 4) k_stripA256(DevParams P, const void* __restrict__ master,
                                                        const void* __restrict__ slave,
                                                        const StripDesc* __restrict__ strips,
                                                        const int* __restrict__ blk, int blk_off, int blk_n,
                                                        const unsigned char* __restrict__ needM,
                                                        const unsigned char* __restrict__ needU,
                                                        float* __restrict__ amp1, float* __restrict__ amp2,
                                                        long long* __restrict__ lsum, float2* __restrict__ Zm,
                                                        const float2* __restrict__ tabs) {
  // one shared array: 16 transpose tiles | pass-2 twiddles w256^(r j)
  // [16][15] | phase factors phi[0..128] (37.8 KB: 4 blocks per CU)
  __shared__ float2 lds[16 * kT256Tile + 240 + 129];
  const int g = threadIdx.x >> 4, j = threadIdx.x & 15;
  const int wi = xcd_item();
  if (wi >= blk_n) return;
  const int bb = wi + blk_off;
  const StripDesc S = strips[blk[2 * bb]];
  float2* tile = lds + g * kT256Tile;
  float2* twl = lds + 16 * kT256Tile;
  float2* sph = twl + 240;
  for (int k = threadIdx.x; k < 240; k += kThreads) twl[k] = tabs[P.tw_nx + (k % 15 + 1) * (k / 15)];
  for (int k = threadIdx.x; k <= 128; k += kThreads) sph[k] = tabs[P.phi + k];
  const unsigned* im_m = (const unsigned*)master;
  const unsigned* im_s = (const unsigned*)slave;
  const int first = blk[2 * bb + 1] - S.line0;                   // strip-relative line of this block
  const int nit = min(kStripBlockLines, S.nlines - first + 15) >> 4;
  // Both lines of an iteration are loaded together, before the master
  // transforms (4 waves per SIMD hide the rest of the latency).
  __syncthreads();
  for (int it = 0; it < nit; it++) {
    const int i = first + 16 * it + g;                            // strip-relative line of this group
    const bool act = i < S.nlines;
    const size_t l = (size_t)(S.line0 + i);
    const bool wm = act && needM[l], wu = act && needU[l];
    unsigned um[16], us[16];
    load_line256<EDGE>(im_m, P.m_nx, P.m_ny, P.m_lo, P.m_hi, S.my0 + i, S.mx0, j, um);
    load_line256<EDGE>(im_s, P.s_nx, P.s_ny, P.s_lo, P.s_hi, S.sy0 + i, S.sx0, j, us);
    {
      float ea[8], oa[8];
      interp_line256<FMT>(um, j, tile, twl, sph, ea, oa);
      LineSum<FMT> s;
      float2* o = (float2*)(amp1 + l * 256);
#pragma unroll
      for (int q = 0; q < 8; q++) {
        const float2 a = make_float2(ea[q], oa[q]);
        if (act) st_nt(o + j + 16 * q, a);
        if (ROWFFT) tile[j + 16 * q] = a;          // through the tile into column order
        s.add(ea[q]);
        s.add(oa[q]);
      }
      const long long st = s.group_total();
      if (act && j == 15) lsum[2 * l] = st;
    }
    float m1v[16];                              // master amplitude of column j + 16 r
    wave_sync();
    if (ROWFFT) {
      const float* m1 = (const float*)tile;
#pragma unroll
      for (int r = 0; r < 16; r++) m1v[r] = m1[j + 16 * r];
      wave_sync();                              // read before the slave's transforms reuse the tile
    }
    {
      float ea[8], oa[8];
      interp_line256<FMT>(us, j, tile, twl, sph, ea, oa);
      LineSum<FMT> s;
      float2* o = (float2*)(amp2 + l * 256);
#pragma unroll
      for (int q = 0; q < 8; q++) {
        const float2 a = make_float2(ea[q], oa[q]);
        if (wm && q >= 2 && q < 6) st_nt(o + j + 16 * q, a);
        if (ROWFFT) tile[j + 16 * q] = a;          // the tile doubles as the slave line buffer
        s.add(ea[q]);
        s.add(oa[q]);
      }
      const long long st = s.group_total();
      if (act && j == 15) lsum[2 * l + 1] = st;
    }
    wave_sync();                                // the slave line is in LDS
    if (!ROWFFT) continue;                      // time modes: amplitudes and sums only
    const float* a2 = (const float*)tile;
    float2 v[16];
#pragma unroll
    for (int r = 0; r < 16; r++) {              // npx = 256, xs = 64: x in [xs, npx - xs) iff r in [4, 12)
      const int x = j + 16 * r;
      v[r] = make_float2(m1v[r], (r >= 4 && r < 12) ? a2[x] : 0.0f);
    }
    wave_sync();                                // read before the transform reuses the tile
    // one transform serves both uses of the line: masked rows take it as is,
    // unmasked rows extract the master's spectrum from it (k_colpass256);
    // wm / wu are uniform over the group's 16 lanes (one line per group)
    if (wm || wu) {
      fft256_group_lds<-1>(v, j, tile, twl);
#pragma unroll
      for (int r = 0; r < 16; r++) st_nt(Zm + l * 256 + j + 16 * r, v[r]);
    }
    wave_sync();                                // the next line's writes into the tile follow
  }
}

// Raw samples x0 + j + G r (r < V) of line gy for a group of G lanes, as
// load_line256: the address is clamped into the held lines [lo, hi) and the
// image columns, and samples outside the image are zeroed afterwards.
template <int G, int V, bool EDGE>
__device__ __forceinline__ void load_line_g(const unsigned* __restrict__ im, int inx, int iny, int lo, int hi, int gy,
                                            int x0, int j, unsigned (&u)[V]) {
  const bool oky = (unsigned)gy < (unsigned)iny;
  const unsigned* line = im + (ptrdiff_t)min(max(gy, lo), hi - 1) * inx;
  if (!EDGE) {
    const unsigned* q = line + x0 + j;
#pragma unroll
    for (int r = 0; r < V; r++) {
      const unsigned w = q[G * r];
      u[r] = oky ? w : 0u;
    }
  } else {
#pragma unroll
    for (int r = 0; r < V; r++) {
      const int gx = x0 + j + G * r;
      const unsigned w = line[min(max(gx, 0), inx - 1)];
      u[r] = (oky && (unsigned)gx < (unsigned)inx) ? w : 0u;
    }
  }
}

// Range interpolation of one held line for npx = N in {128, 512} (range_row
// after the gather): amplitudes ea / oa of output columns 2c, 2c + 1,
// c = j + G q.
template <int N, int FMT>
__device__ __forceinline__ void interp_line(const unsigned (&u)[RegFft<N>::V], int j, float2* tile,
                                            const float2 (&t)[RegFft<N>::TW], const float2* sph,
                                            float (&ea)[RegFft<N>::V / 2], float (&oa)[RegFft<N>::V / 2]) {
  using F = RegFft<N>;
  constexpr int G = F::G, V = F::V;
  float2 v[V];
#pragma unroll
  for (int r = 0; r < V; r++)
    v[r] = FMT == XCORR_HIP_CINT16 ? make_float2((float)(short)(u[r] & 0xffffu), (float)(short)(u[r] >> 16))
                                   : make_float2(__uint_as_float(u[r]), 0.0f);
#pragma unroll
  for (int q = 0; q < V / 2; q++) ea[q] = ampl(v[q + V / 4]);
  F::template run<-1>(v, j, tile, t);
  const float2* phj = sph + j + opaque_zero();
#pragma unroll
  for (int r = 0; r < V; r++) v[r] = cmul(v[r], phj[G * r]);
  F::template run<+1>(v, j, tile, t);
  const float sc = 1.0f / (float)N;
#pragma unroll
  for (int q = 0; q < V / 2; q++) {
    float2 o = v[q + V / 4];
    o.x = o.x * sc;
    o.y = o.y * sc;
    oa[q] = ampl(o);
  }
}

// Strip lines for npx = npy = N in {128, 512} (the outputs of k_stripA256;
// without ROWFFT, for the time modes, amplitudes and line sums only):
// work item b < blk_n (xcd_item order) covers kStripBlockLines lines from
// global line blk[2 (b + blk_off) + 1] of strip blk[2 (b + blk_off)], one line
// per group of G lanes and iteration;
// the groups never wait for each other.  Zm holds the row spectrum of
// z = amp1 + i amp2 (template columns) of every line some patch uses.
template <int N, int FMT, bool EDGE, bool ROWFFT = true>
__global__ void __launch_bounds__(kThreads) k_stripA(DevParams P, const void* __restrict__ master,
                                                     const void* __restrict__ slave,
                                                     const StripDesc* __restrict__ strips,
                                                     const int* __restrict__ blk, int blk_off, int blk_n,
                                                     const unsigned char* __restrict__ needM,
                                                     const unsigned char* __restrict__ needU,
                                                     float* __restrict__ amp1, float* __restrict__ amp2,
                                                     long long* __restrict__ lsum, float2* __restrict__ Zm,
                                                     const float2* __restrict__ tabs) {
  using F = RegFft<N>;
  constexpr int G = F::G, V = F::V, GPB = kThreads / G;
  __shared__ float2 tiles[GPB * F::TILE];
  __shared__ float2 mrow[GPB * (N / 2)];           // each group's master amplitude line (column pairs)
  __shared__ float2 sph[N];                         // range-interpolation phase factors
  const int g = threadIdx.x / G, j = threadIdx.x % G;
  const int wi = xcd_item();
  if (wi >= blk_n) return;
  const int bb = wi + blk_off;
  const StripDesc S = strips[blk[2 * bb]];
  float2* tile = tiles + g * F::TILE;
  float2* mr = mrow + g * (N / 2);
  float2 t[F::TW];
  F::load_tw(tabs + P.tw_nx, j, t);
  for (int k = threadIdx.x; k < N; k += kThreads) sph[k] = tabs[P.phi + k];
  __syncthreads();
  const unsigned* im_m = (const unsigned*)master;
  const unsigned* im_s = (const unsigned*)slave;
  const int first = blk[2 * bb + 1] - S.line0;                   // strip-relative line of this block
  const int end = min(first + kStripBlockLines, S.nlines);
  const int c_lo = P.xs / 2, c_hi = (P.xs + P.nx_corr) / 2;     // template column pairs
  for (int i = first + g; i < end; i += GPB) {
    const size_t l = (size_t)(S.line0 + i);
    const bool wm = needM[l], wu = needU[l];                      // uniform over the group
    unsigned u[V];
    float ea[V / 2], oa[V / 2];
    load_line_g<G, V, EDGE>(im_m, P.m_nx, P.m_ny, P.m_lo, P.m_hi, S.my0 + i, S.mx0, j, u);
    interp_line<N, FMT>(u, j, tile, t, sph, ea, oa);
    {
      LineSum<FMT, G> s;
      float2* o = (float2*)(amp1 + l * N);
#pragma unroll
      for (int q = 0; q < V / 2; q++) {
        const float2 a = make_float2(ea[q], oa[q]);
        o[j + G * q] = a;
        if (ROWFFT) mr[j + G * q] = a;
        s.add(ea[q]);
        s.add(oa[q]);
      }
      const long long st = s.group_total();
      if (j == G - 1) lsum[2 * l] = st;
    }
    load_line_g<G, V, EDGE>(im_s, P.s_nx, P.s_ny, P.s_lo, P.s_hi, S.sy0 + i, S.sx0, j, u);
    interp_line<N, FMT>(u, j, tile, t, sph, ea, oa);
    {
      LineSum<FMT, G> s;
      float2* o = (float2*)(amp2 + l * N);
#pragma unroll
      for (int q = 0; q < V / 2; q++) {
        const float2 a = make_float2(ea[q], oa[q]);
        const int c = j + G * q;
        if (wm && c >= c_lo && c < c_hi) o[c] = a;
        if (ROWFFT) tile[c] = a;                   // the tile doubles as the slave line buffer
        s.add(ea[q]);
        s.add(oa[q]);
      }
      const long long st = s.group_total();
      if (j == G - 1) lsum[2 * l + 1] = st;
    }
    if (!ROWFFT) continue;                         // time modes: amplitudes and sums only
    wave_sync();                                   // the group's two lines are in LDS
    const float* m1 = (const float*)mr;
    const float* a2 = (const float*)tile;
    float2 v[V];
    if (wm || wu) {                                // one transform for both uses (k_colpass extracts)
#pragma unroll
      for (int r = 0; r < V; r++) {
        const int x = j + G * r;
        v[r] = make_float2(m1[x], (x >= P.xs && x < N - P.xs) ? a2[x] : 0.0f);
      }
      wave_sync();                                 // read before the transform reuses the tile
      F::template run<-1>(v, j, tile, t);
#pragma unroll
      for (int r = 0; r < V; r++) Zm[l * N + j + G * r] = v[r];
    }
    wave_sync();                                   // the next line rewrites mr and the tile
  }
}

// Patch sums from the per-line sums: one wave per patch, exact in int64.
__global__ void k_patch_sums(DevParams P, const int* __restrict__ rb, const long long* __restrict__ lsum,
                             unsigned long long* __restrict__ sums, int nb) {
  const int b = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (b >= nb) return;
  long long s1 = 0, s2 = 0;
  for (int y = lane; y < P.NY; y += 64) {
    const size_t l = (size_t)rb[b] + y;
    s1 += lsum[2 * l];
    s2 += lsum[2 * l + 1];
  }
  s1 = wave_sum_i64(s1);
  s2 = wave_sum_i64(s2);
  if (lane == 0) {
    sums[2 * b] = (unsigned long long)s1;
    sums[2 * b + 1] = (unsigned long long)s2;
  }
}

// Stage (b) rows for npx = 256: one block per patch.
__global__ void __launch_bounds__(kThreads) k_rowfft256(DevParams P, const float* __restrict__ amp,
                                                        const unsigned long long* __restrict__ sums,
                                                        float2* __restrict__ Z, const float2* __restrict__ tabs) {
  __shared__ float2 tiles[16 * kT256Tile];
  const int g = threadIdx.x >> 4, j = threadIdx.x & 15;
  const int b = blockIdx.x;
  const float* a1 = amp + (size_t)(2 * b) * P.N;
  const float* a2 = a1 + P.N;
  float2* zb = Z + (size_t)b * P.N;
  float2* tile = tiles + g * kT256Tile;
  float2 t[15];
  load_tw256(tabs + P.tw_nx, j, t);
  for (int row = g; row < P.NY; row += 16) {
    const bool in_y = row >= P.ys && row < P.NY - P.ys;
    float2 v[16];
#pragma unroll
    for (int r = 0; r < 16; r++) {              // every row carries the slave's template columns
      const int x = j + 16 * r;
      v[r] = make_float2(a1[row * 256 + x], (x >= P.xs && x < 256 - P.xs) ? a2[row * 256 + x] : 0.0f);
    }
    fft256_group<-1>(v, j, tile, t);
    extract_real256(v, j, tile, !in_y);           // unmasked rows: the master's spectrum
#pragma unroll
    for (int r = 0; r < 16; r++) zb[(size_t)row * 256 + j + 16 * r] = v[r];
  }
}

// C3 at one bin from A = Z[ky][kx] and Zm = Z[-ky][-kx] (Z = FFT2(c1 + i c2)):
// C1 = (A + conj(Zm)) / 2, conj(C2) = i conj(A - conj(Zm)) / 2 written as the
// swapped pair (d.y, d.x) / 2, C3 = C1 conj(C2) (-1)^(kx+ky).  The halvings are
// applied as one factor of +-1/4 after the complex product: scaling by a power
// of two commutes with every rounding of cmul (the values are far from the
// overflow and subnormal ranges), so the result is bit-identical to halving
// the operands first.
__device__ __forceinline__ float2 c3_bin(float2 a, float2 zm, int kx, int ky) {
  const float2 bb = make_float2(zm.x, -zm.y);
  const float2 c1 = make_float2(a.x + bb.x, a.y + bb.y);
  const float2 d = csub(a, bb);
  const float2 p = cmul(c1, make_float2(d.y, d.x));
  const float f = ((kx + ky) & 1) ? -0.25f : 0.25f;
  return make_float2(p.x * f, p.y * f);
}

// c3_bin without its final factor (-1)^(kx+ky) / 4 (the RAW spectra of the
// strip column pass).  Both move into the inverse transforms exactly: a sign
// (-1)^k on the inputs of the 16 x 16 transforms is a shift of their outputs by
// N/2 (every butterfly sees the negated values of the shifted ones, and IEEE
// negation and round-to-nearest commute), so the kept lag rows and columns are
// read N/2 further on, and 1/4 joins the surface scale 1/N (a power of two,
// which commutes with every rounding away from the subnormal range).
__device__ __forceinline__ float2 c3_raw(float2 a, float2 zm) {
  const float2 bb = make_float2(zm.x, -zm.y);
  const float2 c1 = make_float2(a.x + bb.x, a.y + bb.y);
  const float2 d = csub(a, bb);
  return cmul(c1, make_float2(d.y, d.x));
}

// Demean in the row spectrum: the row transform of z = amp1 + i amp2 mask minus
// that of m1 + i m2 mask, i.e. Zr - m1 npx delta(kx) - i m2 [y in mask] Mx[kx].
__device__ __forceinline__ float2 demean_row_bin(float2 v, int kx, bool ym, float m1NX, float m2f, float2 mx) {
  float re = v.x, im = v.y;
  if (kx == 0) re = re - m1NX;
  if (ym) {
    re = __builtin_fmaf(m2f, mx.y, re);
    im = __builtin_fmaf(-m2f, mx.x, im);
  }
  return make_float2(re, im);
}

// Stage (b) columns, fused, for npy = 256 (ny_corr/2 = 64, nyc = 128): FFT along y,
// C3 = C1 conj(C2) (-1)^(kx+ky), IFFT along y, keep lag rows 64..191.  Block p
// owns columns T = [16p, 16p+16) and their mirrors M = {npx - 16p - g} (with
// npx/2 in place of the duplicate column 0 for p = 0), so every column is
// transformed once and C3 needs no second pass over Z.  Requires npx >= 32.
// Row spectra come from Z[b] (rb null) or, for strips, from line rb[b] + y of
// Zm (the packed transform; unmasked rows extract the master's spectrum).
// STRIPS: rows from strip lines (rb), else from Z[b] (rb unused).
template <bool STRIPS>
__global__ void __launch_bounds__(kThreads) k_colpass256(DevParams P, const float2* __restrict__ Z,
                                                         const int* __restrict__ rb,
                                                         const unsigned long long* __restrict__ sums,
                                                         float2* __restrict__ D, const float2* __restrict__ tabs,
                                                         int nwork) {
  __shared__ float2 T[256 * kT256Pitch];
  __shared__ float2 sx[256];                                 // spectrum of column 0 (self-mirrored)
  __shared__ float2 sxm[256];                                // spectrum of column npx/2 (p == 0)
  __shared__ float2 twl[240];                                // pass-2 twiddles w256^(r j), twl[15 j + r - 1]
  const int g = threadIdx.x >> 4, j = threadIdx.x & 15;
  const int NX = P.NX;
  const int pairs = NX >> 5;
  for (int k = threadIdx.x; k < 240; k += kThreads) twl[k] = tabs[P.tw_ny + (k % 15 + 1) * (k / 15)];
  // work items (patch, column tile) in xcd_item order keep the tiles of a
  // patch, whose mirrored column segments straddle each other's cache lines,
  // and the next patches of a strip, which share rows, in one L2
  const int wi = xcd_item();
  if (wi >= nwork) return;
  const int p = wi % pairs, b = wi / pairs;
  const size_t row0 = STRIPS ? (size_t)rb[b] : (size_t)b * 256;
  // npy = 256 gives ys = 64: row y0 + 16 k (y0 < 16) is masked iff 4 <= k < 12
  const float2* z0 = Z + (row0 + (threadIdx.x >> 4)) * NX;
  const int kxT = 16 * p + g;
  const int kxM = (p == 0 && g == 0) ? NX / 2 : NX - 16 * p - g;
  const bool self = (p == 0 && g == 0);                      // columns 0 and npx/2 mirror onto themselves
  const float m1NX = (float)patch_mean(sums[2 * b], P.N) * (float)NX;
  const float m2f = (float)patch_mean(sums[2 * b + 1], P.N);
  float2 vT[16], vM[16];
  // Thread (y0, c) loads column c of the T and M tiles at rows y0 + 16 k: all
  // 32 loads are in flight together, then the demeaned values go through LDS.
  {
    const int c = threadIdx.x & 15, y0 = threadIdx.x >> 4;
    const int kxt = 16 * p + c;
    const int kxm = (p == 0 && c == 0) ? NX / 2 : NX - 16 * p - c;
    const float2 mkt = tabs[P.mk_x + kxt], mkm = tabs[P.mk_x + kxm];
    const bool self = p == 0 && c == 0;
    float2 rT[16], rM[16];
#pragma unroll
    for (int k = 0; k < 16; k++) {
      const float2* zr = z0 + (size_t)(16 * k) * NX;
      rT[k] = zr[kxt];
      // rows y0 + 16 k outside [64, 192) are unmasked.  Per patch their
      // spectrum is Hermitian (the producer extracted it), so column npx - kx
      // is conj(column kx) (npx/2 is stored); strip lines hold the packed
      // transform, from which the master's spectrum is extracted here
      if ((k >= 4 && k < 12) || STRIPS) {
        rM[k] = zr[kxm];
      } else {
        const float2 w = zr[self ? kxm : kxt];   // branch-free; a cache hit for kxt
        rM[k] = self ? w : make_float2(w.x, -w.y);
      }
    }
    if (STRIPS)
#pragma unroll
      for (int k = 0; k < 16; k++)
        if (!(k >= 4 && k < 12)) {
          const float2 a = rT[k], q = rM[k];      // Z[kx], Z[npx - kx] (self: Z[0], Z[npx/2])
          rT[k] = extract_bin(a, self ? a : q);
          rM[k] = extract_bin(q, self ? q : a);
        }
#pragma unroll
    for (int k = 0; k < 16; k++) {
      const int y = y0 + 16 * k;
      T[y * kT256Pitch + c] = demean_row_bin(rT[k], kxt, k >= 4 && k < 12, m1NX, m2f, mkt);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 16; r++) vT[r] = T[(j + 16 * r) * kT256Pitch + g];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; k++) {
      const int y = y0 + 16 * k;
      T[y * kT256Pitch + c] = demean_row_bin(rM[k], kxm, k >= 4 && k < 12, m1NX, m2f, mkm);
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 16; r++) vM[r] = T[(j + 16 * r) * kT256Pitch + g];
  __syncthreads();
  fft256_group_lds<-1>(vT, j, T + g * kT256Tile, twl);
  fft256_group_lds<-1>(vM, j, T + g * kT256Tile, twl);
  __syncthreads();                                           // every group is done with its tile
  // T products need the M spectra at the mirrored rows
#pragma unroll
  for (int r = 0; r < 16; r++) T[(j + 16 * r) * kT256Pitch + g] = vM[r];
  if (self)
#pragma unroll
    for (int r = 0; r < 16; r++) {
      sx[j + 16 * r] = vT[r];
      sxm[j + 16 * r] = vM[r];
    }
  __syncthreads();
  float2 pT[16];
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const int ky = j + 16 * r;
    const int my = (256 - ky) & 255;
    const float2 zm = self ? sx[my] : T[my * kT256Pitch + g];
    pT[r] = STRIPS ? c3_raw(vT[r], zm) : c3_bin(vT[r], zm, kxT, ky);
  }
  __syncthreads();
  fft256_group_lds<+1>(pT, j, T + g * kT256Tile, twl);
  __syncthreads();                                           // every group is done with its tile
  // lag rows 64..191: r = 4..11, or r = 12..15, 0..3 for the RAW spectra
  constexpr int R0 = STRIPS ? 12 : 4;
#pragma unroll
  for (int q = 0; q < 8; q++) T[(j + 16 * q) * kT256Pitch + g] = pT[(q + R0) & 15];
  __syncthreads();
  float2* db = D + (size_t)b * 128 * NX;
  for (int e = threadIdx.x; e < 128 * 16; e += kThreads)
    db[(size_t)(e >> 4) * NX + 16 * p + (e & 15)] = T[(e >> 4) * kT256Pitch + (e & 15)];
  // Only columns 0..npx/2 are needed downstream (the rest are conjugate
  // mirrors); of the M columns that is npx/2 alone, held by group 0 of the
  // patch's first tile.  Its product and inverse transform come last, once
  // every wave has stored its rows, so waves 1..3 finish instead of waiting.
  if (p == 0) {
    __syncthreads();                                         // T is free for the group's tile
    if (self) {
      float2 u[16];
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const int ky = j + 16 * r;
        u[r] = STRIPS ? c3_raw(sxm[ky], sxm[(256 - ky) & 255]) : c3_bin(sxm[ky], sxm[(256 - ky) & 255], kxM, ky);
      }
      fft256_group_lds<+1>(u, j, T + g * kT256Tile, twl);
      // strips: lag row y's value of column npx/2 is kept in row 0 at column
      // npx/2 + y (unused there), so the row pass reads every lag row as its
      // first npx/2 columns, whole cache lines, and these stores fill lines
#pragma unroll
      for (int q = 0; q < 8; q++) {
        const int y = j + 16 * q;
        db[STRIPS ? (size_t)(NX / 2 + y) : (size_t)y * NX + NX / 2] = u[(q + R0) & 15];
      }
    }
  }
}

// The 8 x 8 neighbourhood of the surface peak that k_finish reads (the flat
// index k = (ip - 4 + iy) nxc + jp - 4 + ix of d_cppow_strip,
// conv_ampl_demean.cu:75-95; 0 outside the surface), stored as 64 values at
// the start of the patch's surface slot when the surface is not kept whole:
// value (iy, ix) at iy * 8 + ix.  Surface row y holds the entries of rows
// y_lo <= y <= y_hi.
struct NbrWindow {
  int k0, ktot, lg, y_lo, y_hi;
  __device__ void zero_outside(float* sb) const {   // thread e < 64: entry e when it lies outside the surface
    const int e = threadIdx.x;
    const int k = k0 + ((e >> 3) << lg) + (e & 7);
    if (k < 0 || k >= ktot) sb[e] = 0.0f;
  }
  __device__ void put(float* sb, int y, int col, float v) const {   // surface value (y, col)
    const int d = (y << lg) + col - k0;
    const int iy = d >> lg, ix = d & ((1 << lg) - 1);
    if (d >= 0 && iy < 8 && ix < 8) sb[iy * 8 + ix] = v;
  }
};
__device__ inline NbrWindow nbr_window(unsigned long long best, int nxc, int nyc) {
  float vb;
  unsigned ib;
  key_decode(best, &vb, &ib);
  NbrWindow w;
  w.lg = __builtin_ctz(nxc);                   // nxc is a power of two >= 8
  const int ip = (int)(ib >> w.lg), jp = (int)(ib & (nxc - 1));
  w.k0 = ((ip - 4) << w.lg) + jp - 4;
  w.ktot = nyc << w.lg;
  // floor divisions of k0 >= -4 nxc - 4 by nxc, kept non-negative
  w.y_lo = max(0, ((w.k0 + (8 << w.lg)) >> w.lg) - 8);
  w.y_hi = min(nyc - 1, ((w.k0 + (7 << w.lg) + 7 + (8 << w.lg)) >> w.lg) - 8);
  return w;
}

// Stage (b) row IFFT + surface + argmax for npx = 256 (nx_corr/2 = 64, nxc = 128):
// one block per patch, so the peak key and the sum are written directly.  The
// kept rows are Hermitian (their transforms are real), so rows r and
// r + nyc/2 go through one inverse transform as ra + i rb; |c3| is |Re| of
// the result for row r and |Im| for row r + nyc/2.  With keep the surface is
// written whole; otherwise only the 8 x 8 neighbourhood of the peak that
// k_finish reads goes out (surface_nbr).
template <bool RAW>
__global__ void __launch_bounds__(kThreads) k_rowifft_surf256(DevParams P, const float2* __restrict__ D,
                                                              float* __restrict__ surf,
                                                              unsigned long long* __restrict__ peak,
                                                              double* __restrict__ partial,
                                                              const float2* __restrict__ tabs, int keep) {
  __shared__ float2 tiles[16 * kT256Tile];
  __shared__ float2 twl[240];                   // pass-2 twiddles w256^(r j), twl[15 j + r - 1]
  __shared__ unsigned long long redk[kThreads / 64];
  __shared__ double redd[kThreads / 64];
  const int g = threadIdx.x >> 4, j = threadIdx.x & 15;
  const int b = blockIdx.x;
  const float2* db = D + (size_t)b * P.nyc * 256;
  float* sb = surf + (size_t)b * P.nyc * 128;
  float2* tile = tiles + g * kT256Tile;
  for (int k = threadIdx.x; k < 240; k += kThreads) twl[k] = tabs[P.tw_nx + (k % 15 + 1) * (k / 15)];
  __syncthreads();
  // RAW lag rows (k_colpass256 on strips): the factor 1/4 joins the scale and
  // the kept columns 64..191 are outputs r = 12..15, 0..3 instead of 4..11
  const float sc = (RAW ? 0.25f : 1.0f) / (float)P.N;
  constexpr int R0 = RAW ? 12 : 4;
  const int hp = P.nyc >> 1;
  unsigned long long best = 0ull;
  double sum = 0.0;
  const int iters = (hp + 15) >> 4;
  for (int it = 0; it < iters; it++) {
    const int row = it * 16 + g;                 // rows row and row + hp (mirrored bins from 0 .. 128)
    const bool active = row < hp;
    // branch-free loads (an inactive group reads a valid row and discards
    // it), so all 32 are in flight together
    const int rowa = min(row, hp - 1);
    const float2* ra = db + (size_t)rowa * 256;
    const float2* rc = ra + (size_t)hp * 256;
    float2 a[16], c[16];
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const int x = j + 16 * r;
      const int xm = x <= 128 ? x : 256 - x;
      // RAW (strip) lag rows keep column 128 in row 0 at column 128 + row
      const bool nyq = RAW && r == 8 && j == 0;
      a[r] = *(nyq ? db + 128 + rowa : ra + xm);
      c[r] = *(nyq ? db + 128 + rowa + hp : rc + xm);
    }
    float2 v[16];
#pragma unroll
    for (int r = 0; r < 16; r++) {
      float2 ar = a[r], cr = c[r];
      if (j + 16 * r > 128) {                    // upper half: conjugate mirrors
        ar.y = -ar.y;
        cr.y = -cr.y;
      }
      v[r] = active ? make_float2(ar.x - cr.y, ar.y + cr.x) : make_float2(0.0f, 0.0f);
    }
    fft256_group_lds<+1>(v, j, tile, twl);
    if (!active) continue;
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const float ca = fabsf(v[(q + R0) & 15].x * sc), cb = fabsf(v[(q + R0) & 15].y * sc);
      const unsigned ia = (unsigned)(row * 128 + j + 16 * q), ib = ia + (unsigned)hp * 128;
      if (keep) {
        sb[ia] = ca;
        sb[ib] = cb;
      }
      const unsigned long long ka = peak_key(ca, ia), kb = peak_key(cb, ib);
      best = ka > best ? ka : best;
      best = kb > best ? kb : best;
      sum += (double)ca;
      sum += (double)cb;
    }
  }
  best = block_max_u64(best, redk);
  sum = block_sum_f64(sum, redd);
  if (threadIdx.x == 0) {
    peak[b] = best;
    partial[b] = sum;
  }
  if (keep) return;
  // The neighbourhood: surface rows y_lo .. y_hi (at most 10), one per group,
  // transformed again with the same arithmetic; a wave with no row left stops.
  const NbrWindow nw = nbr_window(best, 128, P.nyc);
  if (threadIdx.x < 64) nw.zero_outside(sb);
  if (nw.y_lo + (g & ~3) > nw.y_hi) return;
  const int y = min(nw.y_lo + g, nw.y_hi);       // spare groups of a wave redo the last row
  const int pr = y < hp ? y : y - hp;
  const float2* ra = db + (size_t)pr * 256;
  const float2* rc = ra + (size_t)hp * 256;
  float2 v[16];
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const int x = j + 16 * r;
    const int xm = x <= 128 ? x : 256 - x;
    const bool nyq = RAW && r == 8 && j == 0;
    float2 ar = *(nyq ? db + 128 + pr : ra + xm), cr = *(nyq ? db + 128 + pr + hp : rc + xm);
    if (x > 128) {
      ar.y = -ar.y;
      cr.y = -cr.y;
    }
    v[r] = make_float2(ar.x - cr.y, ar.y + cr.x);
  }
  fft256_group_lds<+1>(v, j, tile, twl);
  if (nw.y_lo + g <= nw.y_hi)
#pragma unroll
    for (int q = 0; q < 8; q++)
      nw.put(sb, y, j + 16 * q, fabsf((y < hp ? v[(q + R0) & 15].x : v[(q + R0) & 15].y) * sc));
}

// Stage (b) columns for npx = npy = N in {128, 512}: FFT along y, C3 = C1
// conj(C2) (-1)^(kx+ky), IFFT along y, keep lag rows N/4 .. 3N/4 - 1 (the
// arithmetic of k_colfft + k_product_colifft).  Block p of a patch owns the TC
// columns T = [TC p, TC p + TC) and their mirrors M = {N - kx} (N/2 in place of
// the duplicate of column 0), each group CPG of each; only columns 0 .. N/2 of
// C3 are inverse-transformed (C3 is Hermitian in this arithmetic).
template <int N>
struct ColCfg {
  static constexpr int GPB = kThreads / RegFft<N>::G;
  static constexpr int CPG = GPB >= 8 ? 1 : 8 / GPB;   // columns per group (512: 2)
  static constexpr int TC = GPB * CPG;                 // T columns per block (128: 32, 512: 8)
  static constexpr int PITCH = TC + 1;
};

// Row spectra come from Z[b] (rb null; unmasked rows already hold the master's
// spectrum) or, for strips, from line rb[b] + y of Zm, the packed transform,
// from which unmasked rows extract the master's spectrum.
// Work items w = (patch w / tiles, column tile w % tiles), w < nwork.
// Workgroups go to the 8 XCDs round-robin in blockIdx order; block i takes
// item (i mod 8) (grid / 8) + i / 8 (the grid is a multiple of 8), so each XCD
// runs a contiguous range of items in order: the tiles of a patch, whose
// column segments share cache lines, and neighbouring patches of a strip,
// which share lines, meet in one L2.
template <int N>
__global__ void __launch_bounds__(kThreads) k_colpass(DevParams P, const float2* __restrict__ Z,
                                                      const int* __restrict__ rb,
                                                      const unsigned long long* __restrict__ sums,
                                                      float2* __restrict__ D, const float2* __restrict__ tabs,
                                                      int nwork) {
  using F = RegFft<N>;
  using C = ColCfg<N>;
  constexpr int G = F::G, V = F::V, CPG = C::CPG, TC = C::TC, PITCH = C::PITCH;
  // the groups' transform tiles share the column tile's memory: they are live
  // only between block barriers that separate them from every use of T
  constexpr int TSZ = N * PITCH > C::GPB * F::TILE ? N * PITCH : C::GPB * F::TILE;
  __shared__ float2 T[TSZ];
  __shared__ float2 sx[N];
  float2* const ftiles = T;
  const int tiles = N / (2 * TC);
  const int wi = xcd_item();
  if (wi >= nwork) return;
  const int p = wi % tiles, b = wi / tiles;
  const int g = threadIdx.x / G, j = threadIdx.x % G;
  float2* ftile = ftiles + g * F::TILE;
  float2 t[F::TW];
  F::load_tw(tabs + P.tw_ny, j, t);
  const float m1NX = (float)patch_mean(sums[2 * b], P.N) * (float)N;
  const float m2f = (float)patch_mean(sums[2 * b + 1], P.N);
  const size_t row0 = rb ? (size_t)rb[b] : (size_t)b * N;
  auto kx_of = [&](int i, bool mirror) { return !mirror ? TC * p + i : ((p == 0 && i == 0) ? N / 2 : N - TC * p - i); };
  float2 vT[CPG][V], vM[CPG][V];
  // Thread (y0, i) loads columns i of the T and M tiles at rows y0 + LS k:
  // all 2 LK loads are in flight together (the unmasked strip rows extract the
  // master's spectrum from the pair in registers), then the demeaned values go
  // through LDS.
  static_assert(kThreads % TC == 0 && (N * TC) % kThreads == 0, "column tile");
  constexpr int LS = kThreads / TC, LK = N * TC / kThreads;
  const int i0 = threadIdx.x % TC, y0 = threadIdx.x / TC;
  // square windows: ys = N / 4, so row y0 + LS k (y0 < LS) is masked iff
  // LK / 4 <= k < 3 LK / 4 (known at compile time)
  static_assert(LS * LK == N, "column tile rows");
  {
    const float2* z0 = Z + (row0 + y0) * N;
    const int kxt = kx_of(i0, false), kxm = kx_of(i0, true);   // mirror bins of each other
    const bool self = p == 0 && i0 == 0;                        // columns 0 and N/2 mirror onto themselves
    const float2 mkt = tabs[P.mk_x + kxt], mkm = tabs[P.mk_x + kxm];
    float2 rT[LK], rM[LK];
#pragma unroll
    for (int k = 0; k < LK; k++) {
      rT[k] = z0[(size_t)(LS * k) * N + kxt];
      rM[k] = z0[(size_t)(LS * k) * N + kxm];
    }
    // unmasked strip rows hold the packed transform: extract the master's spectrum
    if (rb)
#pragma unroll
      for (int k = 0; k < LK; k++)
        if (!(k >= LK / 4 && k < 3 * LK / 4)) {
          const float2 a = rT[k], q = rM[k];                    // Z[kx], Z[N - kx] (self: Z[0], Z[N/2])
          rT[k] = extract_bin(a, self ? a : q);
          rM[k] = extract_bin(q, self ? q : a);
        }
#pragma unroll
    for (int k = 0; k < LK; k++)
      T[(y0 + LS * k) * PITCH + i0] = demean_row_bin(rT[k], kxt, k >= LK / 4 && k < 3 * LK / 4, m1NX, m2f, mkt);
    __syncthreads();
#pragma unroll
    for (int c = 0; c < CPG; c++)
#pragma unroll
      for (int r = 0; r < V; r++) vT[c][r] = T[(j + G * r) * PITCH + g * CPG + c];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < LK; k++)
      T[(y0 + LS * k) * PITCH + i0] = demean_row_bin(rM[k], kxm, k >= LK / 4 && k < 3 * LK / 4, m1NX, m2f, mkm);
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < CPG; c++)
#pragma unroll
    for (int r = 0; r < V; r++) vM[c][r] = T[(j + G * r) * PITCH + g * CPG + c];
  __syncthreads();
#pragma unroll
  for (int c = 0; c < CPG; c++) {
    F::template run<-1>(vT[c], j, ftile, t);
    F::template run<-1>(vM[c], j, ftile, t);
  }
  __syncthreads();                                 // transform tiles (aliasing T) are done
  const bool holder = p == 0 && g == 0;           // holds column 0 (T) and N/2 (M): self-mirrored
  // M spectra in natural row order for the T products; column 0's own spectrum in sx
#pragma unroll
  for (int c = 0; c < CPG; c++)
#pragma unroll
    for (int r = 0; r < V; r++) T[(j + G * r) * PITCH + g * CPG + c] = vM[c][r];
  if (holder)
#pragma unroll
    for (int r = 0; r < V; r++) sx[j + G * r] = vT[0][r];
  __syncthreads();
  float2 pT[CPG][V];
#pragma unroll
  for (int c = 0; c < CPG; c++) {
    const int i = g * CPG + c, kx = TC * p + i;
    const bool self = holder && c == 0;
#pragma unroll
    for (int r = 0; r < V; r++) {
      const int ky = j + G * r, my = (N - ky) & (N - 1);
      const float2 zm = self ? sx[my] : T[my * PITCH + i];
      pT[c][r] = c3_bin(vT[c][r], zm, kx, ky);
    }
  }
  if (p == 0) {                                    // the product of column N/2 from its own spectrum
    __syncthreads();
    if (holder)
#pragma unroll
      for (int r = 0; r < V; r++) sx[j + G * r] = vM[0][r];
    __syncthreads();
    if (holder)
#pragma unroll
      for (int r = 0; r < V; r++) {
        const int ky = j + G * r;
        vM[0][r] = c3_bin(vM[0][r], sx[(N - ky) & (N - 1)], N / 2, ky);
      }
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < CPG; c++) F::template run<+1>(pT[c], j, ftile, t);
  if (holder) F::template run<+1>(vM[0], j, ftile, t);   // column N/2 alone; the other groups skip it
  __syncthreads();                                 // transform tiles (aliasing T) are done
  // lag rows N/4 .. 3N/4 - 1 (r = V/4 .. 3V/4 - 1) through LDS, written as rows of D
#pragma unroll
  for (int c = 0; c < CPG; c++)
#pragma unroll
    for (int q = 0; q < V / 2; q++) T[(j + G * q) * PITCH + g * CPG + c] = pT[c][q + V / 4];
  if (holder)
#pragma unroll
    for (int q = 0; q < V / 2; q++) sx[j + G * q] = vM[0][q + V / 4];
  __syncthreads();
  float2* db = D + (size_t)b * (N / 2) * N;
  for (int e = threadIdx.x; e < (N / 2) * TC; e += kThreads) {
    const int iy = e / TC, i = e - (e / TC) * TC;
    db[(size_t)iy * N + TC * p + i] = T[iy * PITCH + i];
  }
  if (p == 0)
    for (int e = threadIdx.x; e < N / 2; e += kThreads) db[(size_t)e * N + N / 2] = sx[e];
}

// Stage (b) row IFFT + surface + argmax for npx = npy = N in {128, 512}: one
// block per patch (the arithmetic of k_rowifft_surf, rows paired as in
// k_rowifft_surf256; the surface sum is one partial per patch).
template <int N>
__global__ void __launch_bounds__(kThreads) k_rowifft_surf_reg(DevParams P, const float2* __restrict__ D,
                                                               float* __restrict__ surf,
                                                               unsigned long long* __restrict__ peak,
                                                               double* __restrict__ partial,
                                                               const float2* __restrict__ tabs, int keep) {
  using F = RegFft<N>;
  constexpr int G = F::G, V = F::V, GPB = kThreads / G, NYC = N / 2;
  __shared__ float2 tiles[GPB * F::TILE];
  __shared__ unsigned long long redk[kThreads / 64];
  __shared__ double redd[kThreads / 64];
  const int g = threadIdx.x / G, j = threadIdx.x % G;
  const int b = blockIdx.x;
  const float2* db = D + (size_t)b * NYC * N;
  float* sb = surf + (size_t)b * NYC * NYC;
  float2* tile = tiles + g * F::TILE;
  float2 t[F::TW];
  F::load_tw(tabs + P.tw_nx, j, t);
  const float sc = 1.0f / (float)P.N;
  unsigned long long best = 0ull;
  double sum = 0.0;
  constexpr int HP = NYC / 2;                      // rows row and row + HP share one transform
  // one pass (N = 128: a row pair per group): the lanes keep their surface
  // values for the neighbourhood instead of transforming its rows again
  constexpr bool ONE = HP == GPB;
  float ka_v[ONE ? V / 2 : 1], kb_v[ONE ? V / 2 : 1];
  for (int row0 = 0; row0 < HP; row0 += GPB) {    // uniform trip count (HP is a multiple of GPB)
    const int row = row0 + g;
    float2 v[V];
#pragma unroll
    for (int r = 0; r < V; r++) {
      const int x = j + G * r;
      const int xm = x <= N / 2 ? x : N - x;
      float2 a = db[(size_t)row * N + xm], c = db[(size_t)(row + HP) * N + xm];
      if (x > N / 2) {                              // upper half: conjugate mirrors
        a.y = -a.y;
        c.y = -c.y;
      }
      v[r] = make_float2(a.x - c.y, a.y + c.x);
    }
    F::template run<+1>(v, j, tile, t);
#pragma unroll
    for (int q = 0; q < V / 2; q++) {
      const float ca = fabsf(v[q + V / 4].x * sc), cb = fabsf(v[q + V / 4].y * sc);
      const unsigned ia = (unsigned)(row * NYC + j + G * q), ib = ia + (unsigned)(HP * NYC);
      if (keep) {
        sb[ia] = ca;
        sb[ib] = cb;
      }
      if (ONE) {
        ka_v[ONE ? q : 0] = ca;
        kb_v[ONE ? q : 0] = cb;
      }
      const unsigned long long ka = peak_key(ca, ia), kb = peak_key(cb, ib);
      best = ka > best ? ka : best;
      best = kb > best ? kb : best;
      sum += (double)ca;
      sum += (double)cb;
    }
  }
  best = block_max_u64(best, redk);
  sum = block_sum_f64(sum, redd);
  if (threadIdx.x == 0) {
    peak[b] = best;
    partial[b] = sum;
  }
  if (keep) return;
  // the peak's neighbourhood (NbrWindow): from the values the lanes hold, or
  // its surface rows transformed again
  const NbrWindow nw = nbr_window(best, NYC, NYC);
  if (threadIdx.x < 64) nw.zero_outside(sb);
  if (ONE) {
#pragma unroll
    for (int q = 0; q < (ONE ? V / 2 : 0); q++) {
      nw.put(sb, g, j + G * q, ka_v[q]);
      nw.put(sb, g + HP, j + G * q, kb_v[q]);
    }
    return;
  }
  for (int yb = nw.y_lo; yb <= nw.y_hi; yb += GPB) {   // uniform over the block
    const int y = min(yb + g, nw.y_hi);
    const int pr = y < HP ? y : y - HP;
    float2 v[V];
#pragma unroll
    for (int r = 0; r < V; r++) {
      const int x = j + G * r;
      const int xm = x <= N / 2 ? x : N - x;
      float2 a = db[(size_t)pr * N + xm], c = db[(size_t)(pr + HP) * N + xm];
      if (x > N / 2) {
        a.y = -a.y;
        c.y = -c.y;
      }
      v[r] = make_float2(a.x - c.y, a.y + c.x);
    }
    F::template run<+1>(v, j, tile, t);
    if (yb + g <= nw.y_hi)
#pragma unroll
      for (int q = 0; q < V / 2; q++)
        nw.put(sb, y, j + G * q, fabsf((y < HP ? v[q + V / 4].x : v[q + V / 4].y) * sc));
  }
}

// --------------------------------------------------------------------------
// stage (b): demean, mask, pack z = c1 + i c2, FFT along x
// conv_ampl_demean.cu:111-139, xcorr.cu:385-401, do_freq_xcorr.cu:90-101
// --------------------------------------------------------------------------
__global__ void __launch_bounds__(kMaxThreads) k_rowfft(DevParams P, const float* __restrict__ amp,
                                                     const unsigned long long* __restrict__ sums,
                                                     float2* __restrict__ Z, const float2* __restrict__ tabs) {
  extern __shared__ float2 smem[];
  const int NX = P.NX, NY = P.NY, G = P.G1, nel = G * NX;
  const int groups = NY / G;
  const int g = blockIdx.x % groups, b = blockIdx.x / groups;
  const float* a1 = amp + (size_t)(2 * b) * P.N + (size_t)g * G * NX;
  const float* a2 = a1 + P.N;
  float2* A = smem;
  float2* Bf = A + nel;
  // raw amplitudes; the means come off in the row spectrum (demean_row_bin)
  for (int e = threadIdx.x; e < nel; e += blockDim.x) {    // every row carries the slave's template columns
    const int x = e % NX;
    A[e] = make_float2(a1[e], (x >= P.xs && x < NX - P.xs) ? a2[e] : 0.0f);
  }
  __syncthreads();
  float2* R = lds_fft_rows<-1>(A, Bf, NX, G, tabs + P.tw_nx, P.plx);
  float2* out = Z + (size_t)b * P.N + (size_t)g * G * NX;
  for (int e = threadIdx.x; e < nel; e += blockDim.x) {
    const int row = e / NX, x = e - row * NX;
    const int y = g * G + row;
    float2 v = R[e];
    if (!(y >= P.ys && y < NY - P.ys))            // unmasked rows: the master's spectrum
      v = extract_bin(v, R[row * NX + ((NX - x) & (NX - 1))]);
    out[e] = v;
  }
}

// FFT along y, in place on Z (columns tile per block).
__global__ void __launch_bounds__(kMaxThreads) k_colfft(DevParams P, float2* __restrict__ Z,
                                                     const unsigned long long* __restrict__ sums,
                                                     const float2* __restrict__ tabs) {
  extern __shared__ float2 smem[];
  const int NX = P.NX, NY = P.NY, C = P.C3, nel = C * NY;
  const int tiles = NX / C;
  const int ct = blockIdx.x % tiles, b = blockIdx.x / tiles;
  float2* zb = Z + (size_t)b * P.N + ct * C;
  float2* A = smem;
  float2* Bf = A + nel;
  const float m1NX = (float)patch_mean(sums[2 * b], P.N) * (float)NX;
  const float m2f = (float)patch_mean(sums[2 * b + 1], P.N);
  for (int e = threadIdx.x; e < nel; e += blockDim.x) {
    const int y = e / C, c = e - y * C;
    const int kx = ct * C + c;
    A[c * NY + y] = demean_row_bin(zb[(size_t)y * NX + c], kx, y >= P.ys && y < NY - P.ys, m1NX, m2f,
                                   tabs[P.mk_x + kx]);
  }
  __syncthreads();
  float2* R = lds_fft_rows<-1>(A, Bf, NY, C, tabs + P.tw_ny, P.ply);
  for (int e = threadIdx.x; e < nel; e += blockDim.x) {
    const int y = e / C, c = e - y * C;
    zb[(size_t)y * NX + c] = R[c * NY + y];
  }
}

// C3 = C1 conj(C2) (-1)^(kx+ky) with C1, C2 split from Z = FFT2(c1 + i c2);
// IFFT along y; keep rows [ny_corr/2, ny_corr/2 + nyc).   do_freq_xcorr.cu:46-64, 103-113
__global__ void __launch_bounds__(kMaxThreads) k_product_colifft(DevParams P, const float2* __restrict__ Z,
                                                              float2* __restrict__ D,
                                                              const float2* __restrict__ tabs) {
  extern __shared__ float2 smem[];
  const int NX = P.NX, NY = P.NY, C = P.C3, nel = C * NY;
  const int tiles = NX / C;
  const int ct = blockIdx.x % tiles, b = blockIdx.x / tiles;
  const float2* zb = Z + (size_t)b * P.N;
  float2* A = smem;
  float2* M = A + nel;
  float2* O = M + nel;
  for (int e = threadIdx.x; e < nel; e += blockDim.x) {
    const int y = e / C, c = e - y * C;
    const int kx = ct * C + c;
    const int mx = (NX - kx) & (NX - 1);
    A[c * NY + y] = zb[(size_t)y * NX + kx];
    M[c * NY + y] = zb[(size_t)y * NX + mx];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < nel; e += blockDim.x) {
    const int c = e / NY, ky = e - c * NY;
    const int kx = ct * C + c;
    const float2 zm = M[c * NY + ((NY - ky) & (NY - 1))];
    O[e] = c3_bin(A[e], zm, kx, ky);
  }
  __syncthreads();
  float2* R = lds_fft_rows<+1>(O, A, NY, C, tabs + P.tw_ny, P.ply);
  const int r0 = P.ny_corr / 2;
  float2* db = D + (size_t)b * P.nyc * NX + ct * C;
  for (int e = threadIdx.x; e < P.nyc * C; e += blockDim.x) {
    const int iy = e / C, c = e - iy * C;
    db[(size_t)iy * NX + c] = R[c * NY + r0 + iy];
  }
}

// IFFT along x of the kept rows, |c3|/(NX NY) on the lag window, argmax and
// partial sums.   do_freq_xcorr.cu:67-84, 113-114; conv_ampl_demean.cu:434-554
__global__ void __launch_bounds__(kThreads) k_rowifft_surf(DevParams P, const float2* __restrict__ D,
                                                           float* __restrict__ surf,
                                                           unsigned long long* __restrict__ peak,
                                                           double* __restrict__ partial,
                                                           const float2* __restrict__ tabs) {
  extern __shared__ float2 smem[];
  __shared__ unsigned long long redk[kThreads / 64];
  __shared__ double redd[kThreads / 64];
  const int NX = P.NX, G = P.G4, nxc = P.nxc;
  const int groups = P.groups4;
  const int g = blockIdx.x % groups, b = blockIdx.x / groups;
  // G surface rows per block as G / 2 pairs (r, r + nyc/2), r from a0: the
  // kept rows are Hermitian, so a pair goes through one transform as ra + i rb
  const int hp = P.nyc >> 1, G2 = G >> 1, nel2 = G2 * NX, a0 = g * G2;
  const float2* db = D + (size_t)b * P.nyc * NX;
  float2* A = smem;
  float2* Bf = A + nel2;
  // columns above npx/2 are the conjugate mirrors of the lower half (C3 is Hermitian)
  for (int e = threadIdx.x; e < nel2; e += kThreads) {
    const int row = e / NX, x = e - row * NX;
    const int xm = x <= NX / 2 ? x : NX - x;
    float2 a = db[(size_t)(a0 + row) * NX + xm], c = db[(size_t)(a0 + row + hp) * NX + xm];
    if (x > NX / 2) {
      a.y = -a.y;
      c.y = -c.y;
    }
    A[e] = make_float2(a.x - c.y, a.y + c.x);
  }
  __syncthreads();
  float2* R = lds_fft_rows<+1>(A, Bf, NX, G2, tabs + P.tw_nx, P.plx);
  const int q0 = P.nx_corr / 2;
  const float sc = 1.0f / (float)P.N;
  float* sb = surf + (size_t)b * P.nyc * nxc;
  unsigned long long best = 0ull;
  double sum = 0.0;
  for (int e = threadIdx.x; e < G2 * nxc; e += kThreads) {
    const int row = e / nxc, ix = e - row * nxc;
    const float2 v = R[row * NX + q0 + ix];
    const float ca = fabsf(v.x * sc), cb = fabsf(v.y * sc);
    const unsigned ia = (unsigned)((a0 + row) * nxc + ix), ib = ia + (unsigned)(hp * nxc);
    sb[ia] = ca;
    sb[ib] = cb;
    const unsigned long long ka = peak_key(ca, ia), kb = peak_key(cb, ib);
    best = ka > best ? ka : best;
    best = kb > best ? kb : best;
    sum += (double)ca;
    sum += (double)cb;
  }
  best = block_max_u64(best, redk);
  sum = block_sum_f64(sum, redd);
  if (threadIdx.x == 0) {
    atomicMax(&peak[b], best);
    partial[(size_t)b * groups + g] = sum;
  }
}

// --------------------------------------------------------------------------
// stages (c) + (d): NCC at the peak, high-res refinement, results
// do_freq_xcorr.cu:150-200, do_time_int_xcorr.cu:71-100, highres_corr.cu:23-79,
// conv_ampl_demean.cu:75-95, fft_interpolate_routines.cu:118-147
// --------------------------------------------------------------------------
// Amplitudes of patch b: rows y of A1 / A2 at (rb[b] + y) npx (strips) or
// b pstride + y npx.  The result goes to out[perm[b]] (or out[b]).
// Patches are mapped to blocks by xcd_item: neighbours in a strip, whose
// amplitude rows overlap, meet in one L2.
__global__ void __launch_bounds__(kThreads) k_finish(DevParams P, const float* __restrict__ A1,
                                                     const float* __restrict__ A2, const int* __restrict__ rb,
                                                     long long pstride, const int* __restrict__ perm,
                                                     const unsigned long long* __restrict__ sums,
                                                     const float* __restrict__ surf,
                                                     const double* __restrict__ surfd,
                                                     const unsigned long long* __restrict__ peak,
                                                     const double* __restrict__ partial,
                                                     const int* __restrict__ tpeak,
                                                     const double* __restrict__ tcave,
                                                     xcorr_hip_result* __restrict__ out,
                                                     const float2* __restrict__ tabs, int nb, int nbr) {
  extern __shared__ float2 smem[];
  __shared__ double redd[kThreads / 64];
  __shared__ unsigned long long redk[kThreads / 64];
  __shared__ float2 sE[512];                 // e^{+2 pi i m/Lx} (m < Lx)
  __shared__ float md[64];
  __shared__ float2 Xr[64];
  __shared__ double s_msc;
  const int b = xcd_item();
  if (b >= nb) return;
  const int NX = P.NX, nxc = P.nxc, nyc = P.nyc;
  const bool freq = P.corr_flag == XCORR_HIP_FREQ;
  xcorr_hip_result res;
  int ip, jp;
  float cmax;
  if (freq) {
    float v;
    unsigned idx;
    key_decode(peak[b], &v, &idx);
    cmax = v;
    ip = (int)idx / nxc;
    jp = (int)idx % nxc;
    double s = 0.0;
    for (int g = 0; g < P.groups4; g++) s += partial[(size_t)b * P.groups4 + g];
    res.cave = s / (double)(nxc * nyc);
    const int ipeak = ip - nyc / 2, jpeak = jp - nxc / 2;
    const float m1f = (float)patch_mean(sums[2 * b], P.N);
    const float m2f = (float)patch_mean(sums[2 * b + 1], P.N);
    const size_t pb = rb ? (size_t)rb[b] * NX : (size_t)b * pstride;
    const float* a1 = A1 + pb;
    const float* a2 = A2 + pb;
    // exact integer sums in FP64: |products| < 2^35 and every partial sum < 2^53,
    // so the accumulation order cannot change them
    double num = 0.0, d1 = 0.0, d2 = 0.0;
    const int ncx = P.nx_corr, nt = P.nx_corr * P.ny_corr;
    const float* arow = a1 + (size_t)(P.ys + ipeak) * NX + P.xs + jpeak;
    const float* brow = a2 + (size_t)P.ys * NX + P.xs;
    const int lg = __builtin_ctz(ncx);          // nx_corr is a power of two
    // rows of nx_corr >= 4 values: quads of consecutive values (one 16-byte
    // load each, 4-byte aligned), 8 quads of each window per thread in flight
    const int nq = ncx >= 4 ? nt >> 2 : 0;
    for (int q0 = 0; q0 < nq; q0 += 8 * kThreads) {
      f4a4 ra[8], rb8[8];
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const int q = q0 + k * kThreads + threadIdx.x;
        ra[k] = f4a4{0.0f, 0.0f, 0.0f, 0.0f};
        rb8[k] = f4a4{0.0f, 0.0f, 0.0f, 0.0f};
        if (q < nq) {
          const int e = q << 2;
          const size_t o = (size_t)(e >> lg) * NX + (e & (ncx - 1));
          ra[k] = *(const f4a4*)(arow + o);
          rb8[k] = *(const f4a4*)(brow + o);
        }
      }
#pragma unroll
      for (int k = 0; k < 8; k++) {
        if (q0 + k * kThreads + (int)threadIdx.x < nq) {
#pragma unroll
          for (int c = 0; c < 4; c++) {
            const double av = (double)(int)(ra[k][c] - m1f);
            const double bv = (double)(int)(rb8[k][c] - m2f);
            num = __builtin_fma(av, bv, num);
            d1 = __builtin_fma(av, av, d1);
            d2 = __builtin_fma(bv, bv, d2);
          }
        }
      }
    }
    for (int e0 = nq > 0 ? nt : 0; e0 < nt; e0 += 8 * kThreads) {   // nx_corr 2: 8 loads per thread in flight
      float ra[8], rb8[8];
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const int e = e0 + k * kThreads + threadIdx.x;
        ra[k] = 0.0f;
        rb8[k] = 0.0f;
        if (e < nt) {
          const size_t o = (size_t)(e >> lg) * NX + (e & (ncx - 1));
          ra[k] = arow[o];
          rb8[k] = brow[o];
        }
      }
#pragma unroll
      for (int k = 0; k < 8; k++) {
        if (e0 + k * kThreads + (int)threadIdx.x < nt) {
          const double av = (double)(int)(ra[k] - m1f);
          const double bv = (double)(int)(rb8[k] - m2f);
          num = __builtin_fma(av, bv, num);
          d1 = __builtin_fma(av, av, d1);
          d2 = __builtin_fma(bv, bv, d2);
        }
      }
    }
    num = block_sum_f64(num, redd);
    d1 = block_sum_f64(d1, redd);
    d2 = block_sum_f64(d2, redd);
    const double den = sqrt(1.0 * d1 * d2);
    const double gamma = (den == 0.0) ? 0.0 : 100.0 * fabs(num / den);
    if (threadIdx.x == 0) s_msc = (double)cmax / gamma;
    res.xoff = (float)(-jpeak);
    res.yoff = (float)(-ipeak);
    res.corr = (float)gamma;
  } else {
    ip = tpeak[4 * b + 0];
    jp = tpeak[4 * b + 1];
    cmax = __int_as_float(tpeak[4 * b + 2]);
    res.cave = tcave[b];
    res.xoff = (float)(-(jp - nxc / 2));
    res.yoff = (float)(-(ip - nyc / 2));
    res.corr = cmax;
  }
  __syncthreads();
  res.ip = ip;
  res.jp = jp;
  res.cmax = cmax;
  res.xfrac = 0.0f;
  res.yfrac = 0.0f;
  res.hip = -1;
  res.hjp = -1;

  if (P.interp_flag) {
    const int Lx = P.Lx, Ly = P.Ly;
    float2* rows = smem;                 // 8 x Lx
    float2* Xc = rows + 8 * Lx;          // 8 x Lx: column t's spectrum at Xc[k Lx + t] (consecutive t, no bank conflicts)
    if (threadIdx.x < 64) {
      const int iy = threadIdx.x >> 3, ix = threadIdx.x & 7;
      const long k = (long)(ip - 4 + iy) * nxc + (jp - 4 + ix);
      float q = 0.0f;
      if (k >= 0 && k < (long)nxc * nyc) {
        // nbr: the row kernel (the fused -time search) left only this
        // neighbourhood, at the start of the slot (NbrWindow)
        const size_t e = (size_t)b * nxc * nyc + (nbr ? threadIdx.x : k);
        const float v = freq ? (float)((double)surf[e] / s_msc) : (float)surfd[e];
        q = sqrtf(sqrtf(v));
      }
      md[threadIdx.x] = q;
    }
    __syncthreads();
    if (threadIdx.x < 8) {
      float2 v[8];
#pragma unroll
      for (int k = 0; k < 8; k++) v[k] = make_float2(md[threadIdx.x * 8 + k], 0.0f);
      Dft<8, -1>::run(v);
#pragma unroll
      for (int k = 0; k < 8; k++) Xr[threadIdx.x * 8 + k] = v[k];
    }
    __syncthreads();
    const float2* sEx = sE;
    for (int m = threadIdx.x; m < Lx; m += kThreads) sE[m] = tabs[P.tw_lx + m];
    __syncthreads();
    const float sx = 1.0f / 8.0f, sy = 1.0f / 8.0f;
    for (int e = threadIdx.x; e < 8 * Lx; e += kThreads) {
      const int iy = e / Lx, t = e - iy * Lx;
      float2 acc = make_float2(0.0f, 0.0f);
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const int kt = k < 4 ? k : k - 8;
        const int m = (kt * t) & (Lx - 1);
        const float2 ex = sEx[m];
        const float2 X = Xr[iy * 8 + k];
        acc.x = __builtin_fmaf(X.x, ex.x, __builtin_fmaf(-X.y, ex.y, acc.x));
        acc.y = __builtin_fmaf(X.x, ex.y, __builtin_fmaf(X.y, ex.x, acc.y));
      }
      rows[iy * Lx + t] = make_float2(acc.x * sx, acc.y * sx);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < Lx; t += kThreads) {
      float2 v[8];
#pragma unroll
      for (int k = 0; k < 8; k++) v[k] = rows[k * Lx + t];
      Dft<8, -1>::run(v);
#pragma unroll
      for (int k = 0; k < 8; k++) Xc[k * Lx + t] = v[k];
    }
    __syncthreads();
    // The column pass forms the real part only, with the bins k and 8 - k
    // paired (P_k = X[k].re + X[8-k].re, Q_k = X[k].im - X[8-k].im) and rows
    // s and Ly - s from the same two sums: a = X[0].re + sum_k cos P_k +
    // cos_4 X[4].re, c = sum_k sin Q_k - sin_4 X[4].im (k = 1, 2, 3 in order),
    // v(s) = (a - c) / 8, v(Ly - s) = (a + c) / 8.  E4[s][k - 1] =
    // e^{2 pi i k s / Ly} (k = 1..4, s = 0..Ly/2) comes from the table in
    // global memory.  Each thread owns one column t (its paired spectrum in
    // registers) and walks s = 0..Ly/2; when Lx is a multiple of 64 the lanes
    // of a wave share s and the factors are scalar loads.  The maximum is over
    // peak keys (the largest value, then the smallest index; values <= 0 and
    // NaN never win), so the order of the rows does not matter.
    const float2* E4 = tabs + P.e4;
    unsigned long long bestk = 0ull;
    const int per = kThreads >= Lx ? kThreads / Lx : 1;
    const int hL = Ly / 2;
    auto scan = [&](int s0, int t) {
      float2 X[8];
#pragma unroll
      for (int k = 0; k < 8; k++) X[k] = Xc[k * Lx + t];
      float pk[3], qk[3];
#pragma unroll
      for (int k = 1; k < 4; k++) {
        pk[k - 1] = X[k].x + X[8 - k].x;
        qk[k - 1] = X[k].y - X[8 - k].y;
      }
      for (int s = s0; s <= hL; s += per) {
        const float2* es = E4 + s * 4;
        float a = X[0].x, c = 0.0f;
#pragma unroll
        for (int k = 0; k < 3; k++) {
          const float2 ey = es[k];
          a = __builtin_fmaf(ey.x, pk[k], a);
          c = __builtin_fmaf(ey.y, qk[k], c);
        }
        const float2 e4 = es[3];
        a = __builtin_fmaf(e4.x, X[4].x, a);
        c = __builtin_fmaf(-e4.y, X[4].y, c);
        const unsigned long long k1 = peak_key((a - c) * sy, (unsigned)(s * Lx + t));
        bestk = k1 > bestk ? k1 : bestk;
        if (s > 0 && s < hL) {
          const unsigned long long k2 = peak_key((a + c) * sy, (unsigned)((Ly - s) * Lx + t));
          bestk = k2 > bestk ? k2 : bestk;
        }
      }
    };
    for (int t0 = 0; t0 < Lx; t0 += kThreads) {
      const int t = t0 + (per > 1 ? (int)threadIdx.x % Lx : (int)threadIdx.x);
      if (t >= Lx) continue;
      const int s0 = per > 1 ? (int)threadIdx.x / Lx : 0;
      if (per == 1 || Lx % 64 == 0) scan(__builtin_amdgcn_readfirstlane(s0), t);   // s uniform over the wave
      else scan(s0, t);
    }
    unsigned long long best = block_max_u64(bestk, redk);
    float bv;
    unsigned bidx;
    key_decode(best, &bv, &bidx);
    const int hp = (int)bidx / Lx, hq = (int)bidx % Lx;
    const float half = (float)Ly / 2.0f;
    res.xfrac = ((float)hq - half) / (float)P.ifc;
    res.yfrac = ((float)hp - half) / (float)P.ifc;
    res.hip = hp;
    res.hjp = hq;
  }
  if (threadIdx.x == 0) out[perm ? perm[b] : b] = res;
}

// --------------------------------------------------------------------------
// stage (c), time domain: full search over all (2 ysearch) x (2 xsearch) lags
// do_time_int_xcorr.cu:30-68 (calc_time_corr), 103-144 (calc_time_corr_hat),
// 146-189 (do_time_corr).  Sums are exact: FP64 accumulation of integer
// products for -time (|partial sums| < 2^53), 128-bit integers for -time4.
// --------------------------------------------------------------------------
typedef unsigned __int128 u128;

// correctly rounded u64 -> double: hi*2^32 is exact, one rounding in the add
__device__ __forceinline__ double u64_to_double(unsigned long long t) {
  return (double)(unsigned)(t >> 32) * 4294967296.0 + (double)(unsigned)(t & 0xffffffffull);
}
// correctly rounded u128 -> double: top 64 bits with a sticky bit
__device__ __forceinline__ double u128_to_double(u128 x) {
  const unsigned long long hi = (unsigned long long)(x >> 64);
  if (hi == 0ull) return u64_to_double((unsigned long long)x);
  const int shift = 64 - __clzll((long long)hi);
  unsigned long long t = (unsigned long long)(x >> shift);
  const u128 mask = (((u128)1) << shift) - 1;
  if (x & mask) t |= 1ull;
  return ldexp(u64_to_double(t), shift);
}

template <int MODE>
__global__ void __launch_bounds__(kThreads) k_time_search(DevParams P, const float* __restrict__ amp,
                                                          const unsigned long long* __restrict__ sums,
                                                          double* __restrict__ surfd) {
  extern __shared__ int ism[];
  __shared__ long long redl[kThreads / 64];
  __shared__ u128 red128[kThreads];
  const int NX = P.NX, nxc = P.nxc, LR = P.tLR, RC = P.tRC, ncx = P.nx_corr, ncy = P.ny_corr;
  const int nblk = P.nyc / LR;
  const int blk = blockIdx.x % nblk, b = blockIdx.x / nblk;
  const int ic0 = blk * LR;                       // first surface row of this block
  const float m1f = (float)patch_mean(sums[2 * b], P.N);
  const float m2f = (float)patch_mean(sums[2 * b + 1], P.N);
  const float* a1 = amp + (size_t)(2 * b) * P.N;
  const float* a2 = a1 + P.N;
  int* Mb = ism;                                  // (RC + LR - 1) x NX master ints
  int* Tb = ism + (RC + LR - 1) * NX;             // RC x nx_corr template ints
  const int nlag = LR * nxc;
  double ab[2] = {0.0, 0.0}, aa[2] = {0.0, 0.0};
  u128 n4[2] = {0, 0}, a4[2] = {0, 0};
  long long b2 = 0;
  u128 b4 = 0;
  for (int r0 = 0; r0 < ncy; r0 += RC) {
    const int rc = min(RC, ncy - r0);
    __syncthreads();
    const int mrows = rc + LR - 1;
    const size_t mrow0 = (size_t)(ic0 + r0);      // = ysearch + ioff0 + r0 with ioff0 = ic0 - ysearch
    for (int e = threadIdx.x; e < mrows * NX; e += kThreads) {
      const int rr = e / NX, x = e - rr * NX;
      Mb[e] = (int)(a1[(mrow0 + rr) * NX + x] - m1f);
    }
    for (int e = threadIdx.x; e < rc * ncx; e += kThreads) {
      const int rr = e / ncx, c = e - rr * ncx;
      const int v = (int)(a2[(size_t)(P.ys + r0 + rr) * NX + P.xs + c] - m2f);
      Tb[e] = v;
      const long long v2 = (long long)v * v;
      if (MODE == 0) b2 += v2;
      else b4 += (u128)(unsigned long long)v2 * (unsigned long long)v2;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const int t = threadIdx.x + q * kThreads;
      if (t >= nlag) break;
      const int lr = t / nxc, jc = t - lr * nxc;
      for (int rr = 0; rr < rc; rr++) {
        const int* arow = Mb + (lr + rr) * NX + jc;
        const int* brow = Tb + rr * ncx;
        if (MODE == 0) {
          double s_ab = ab[q], s_aa = aa[q];
          for (int c = 0; c < ncx; c++) {
            const double av = (double)arow[c], bv = (double)brow[c];
            s_ab = __builtin_fma(av, bv, s_ab);
            s_aa = __builtin_fma(av, av, s_aa);
          }
          ab[q] = s_ab;
          aa[q] = s_aa;
        } else {
          u128 s_n = n4[q], s_a = a4[q];
          for (int c = 0; c < ncx; c++) {
            const long long av = arow[c], bv = brow[c];
            const unsigned long long ua = (unsigned long long)(av * av), ub = (unsigned long long)(bv * bv);
            s_n += (u128)ua * ub;
            s_a += (u128)ua * ua;
          }
          n4[q] = s_n;
          a4[q] = s_a;
        }
      }
    }
  }
  double d2d;
  if (MODE == 0) {
    d2d = (double)block_sum_i64(b2, redl);
  } else {
    __syncthreads();
    red128[threadIdx.x] = b4;
    __syncthreads();
    u128 t4 = 0;
    for (int i = 0; i < kThreads; i++) t4 += red128[i];
    d2d = u128_to_double(t4);
  }
  double* sd = surfd + (size_t)b * nxc * P.nyc;
#pragma unroll
  for (int q = 0; q < 2; q++) {
    const int t = threadIdx.x + q * kThreads;
    if (t >= nlag) break;
    const int lr = t / nxc, jc = t - lr * nxc;
    double g;
    if (MODE == 0) {
      const double den = sqrt(1.0 * aa[q] * d2d);
      g = (den == 0.0) ? 0.0 : 100.0 * fabs(ab[q] / den);
    } else {
      const double den = sqrt(u128_to_double(a4[q]) * d2d);
      const double gg = (den == 0.0) ? 0.0 : fabs(u128_to_double(n4[q]) / den);
      g = (gg <= 0.5) ? 0.0 : 100.0 * sqrt(gg * 2.0 - 1.0);
    }
    sd[(size_t)(ic0 + lr) * nxc + jc] = g;
  }
}

// -time search with register blocking: each thread owns TL consecutive lags
// (jc0 .. jc0 + TL - 1) of one surface row; master and template rows are held
// as doubles in LDS (converted once), so the inner loop is one broadcast
// template load, one master load and TL + 1 FMAs per template sample: the TL
// products and the master energy of lag jc0, from which the energies of the
// other lags follow by exact sliding updates.  Every partial sum is an integer
// below 2^53 (|a|, |b| < 2^17, at most 2^16 terms), so the values equal
// k_time_search<0>'s.
template <int TL>
__global__ void __launch_bounds__(kThreads) k_time_search_blk(DevParams P, const float* __restrict__ amp,
                                                              const unsigned long long* __restrict__ sums,
                                                              double* __restrict__ surfd, int LR, int RC) {
  extern __shared__ double dsm[];
  __shared__ long long redl[kThreads / 64];
  const int NX = P.NX, nxc = P.nxc, ncx = P.nx_corr, ncy = P.ny_corr;
  const int qg = nxc / TL;                        // lag groups per surface row
  const int nblk = P.nyc / LR;
  const int blk = blockIdx.x % nblk, b = blockIdx.x / nblk;
  const int ic0 = blk * LR;
  const float m1f = (float)patch_mean(sums[2 * b], P.N);
  const float m2f = (float)patch_mean(sums[2 * b + 1], P.N);
  const float* a1 = amp + (size_t)(2 * b) * P.N;
  const float* a2 = a1 + P.N;
  double* Mb = dsm;                               // (RC + LR - 1) x NX master values
  double* Tb = dsm + (size_t)(RC + LR - 1) * NX;  // RC x nx_corr template values
  const int t = threadIdx.x;
  const bool act = t < LR * qg;
  const int lr = act ? t / qg : 0, jc0 = act ? TL * (t - (t / qg) * qg) : 0;
  double ab[TL], aa[TL];
#pragma unroll
  for (int k = 0; k < TL; k++) {
    ab[k] = 0.0;
    aa[k] = 0.0;
  }
  long long b2 = 0;
  for (int r0 = 0; r0 < ncy; r0 += RC) {
    const int rc = min(RC, ncy - r0);
    __syncthreads();
    const int mrows = rc + LR - 1;
    const size_t mrow0 = (size_t)(ic0 + r0);
    for (int e = threadIdx.x; e < mrows * NX; e += kThreads) {
      const int rr = e / NX, x = e - rr * NX;
      Mb[e] = (double)(int)(a1[(mrow0 + rr) * NX + x] - m1f);
    }
    for (int e = threadIdx.x; e < rc * ncx; e += kThreads) {
      const int rr = e / ncx, c = e - rr * ncx;
      const int v = (int)(a2[(size_t)(P.ys + r0 + rr) * NX + P.xs + c] - m2f);
      Tb[e] = (double)v;
      b2 += (long long)v * v;
    }
    __syncthreads();
    if (act) {
      for (int rr = 0; rr < rc; rr++) {
        const double* arow = Mb + (lr + rr) * NX + jc0;
        const double* brow = Tb + rr * ncx;
        double x[TL];
#pragma unroll
        for (int k = 0; k < TL; k++) x[k] = arow[k];
        double e0 = 0.0;
        for (int c = 0; c < ncx; c++) {
          const double bv = brow[c];
#pragma unroll
          for (int k = 0; k < TL; k++) ab[k] = __builtin_fma(x[k], bv, ab[k]);
          e0 = __builtin_fma(x[0], x[0], e0);
#pragma unroll
          for (int k = 0; k + 1 < TL; k++) x[k] = x[k + 1];
          x[TL - 1] = arow[c + TL];               // jc0 + ncx + TL - 1 <= NX - 1: in the row
        }
        aa[0] += e0;
#pragma unroll
        for (int k = 1; k < TL; k++) {
          e0 = e0 - arow[k - 1] * arow[k - 1] + arow[ncx + k - 1] * arow[ncx + k - 1];
          aa[k] += e0;
        }
      }
    }
  }
  const double d2d = (double)block_sum_i64(b2, redl);
  if (!act) return;
  double* sd = surfd + (size_t)b * nxc * P.nyc + (size_t)(ic0 + lr) * nxc + jc0;
#pragma unroll
  for (int k = 0; k < TL; k++) {
    const double den = sqrt(1.0 * aa[k] * d2d);
    sd[k] = (den == 0.0) ? 0.0 : 100.0 * fabs(ab[k] / den);
  }
}

// --------------------------------------------------------------------------
// -time search on the int8 matrix cores for xsearch = ysearch = 32 (128 x 128
// windows, 64 x 64 template, 64 x 64 lags).  The integer values a (window)
// and b (template), |a|, |b| < 2^20, are written as three signed 7-bit limbs,
// a = l2 2^14 + l1 2^7 + l0 with l0, l1 in [-64, 63]; every limb product is
// below 2^12 in magnitude, a lag sums 4096 of them, and at most three limb
// pairs share a weight 2^(7 w), so the five per-weight int32 accumulators
// cannot overflow and sum_w acc_w 2^(7 w) (int64) is the exact correlation.
// For master row r the lags (dy, dx) get sum_x A[r][x + dx] B[r - dy][x]: a
// Toeplitz (dx, x) times template (x, dy) product, one 16 x 16 x 64 MFMA per
// (dx tile, dy tile, limb pair).  A patch with a value outside the limb range
// takes an exact int64 loop instead.
// --------------------------------------------------------------------------
typedef int v4i_t __attribute__((ext_vector_type(4)));

// Inclusive prefix sum of an int64 over the 64 lanes of a wave with DPP lane
// moves (row shifts by 1, 2, 4, 8 within rows of 16, then the row broadcasts
// of lanes 15 and 31); exact in any order.
__device__ __forceinline__ long long wave_incl_scan_i64(long long x) {
  x += dpp_i64<0x111, 0xf>(x);                   // row_shr:1
  x += dpp_i64<0x112, 0xf>(x);                   // row_shr:2
  x += dpp_i64<0x114, 0xf>(x);                   // row_shr:4
  x += dpp_i64<0x118, 0xf>(x);                   // row_shr:8
  x += dpp_i64<0x142, 0xa>(x);                   // row_bcast:15 into rows 1 and 3
  x += dpp_i64<0x143, 0xc>(x);                   // row_bcast:31 into rows 2 and 3
  return x;
}
__device__ __forceinline__ long long wave_lane63_i64(long long x) {
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)x, 63);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(x >> 32), 63);
  return (long long)(((unsigned long long)hi << 32) | lo);
}

// The square of a window integer for -time4, below 2^40 (six unsigned 7-bit
// limbs) for |v| < 2^20; a magnitude beyond is clamped to 2^20, whose square
// kSqLimit flags the patch for the exact loop.
constexpr unsigned long long kSqLimit = 1ull << 40;
__device__ __forceinline__ unsigned long long sq_clamped(int v) {
  const unsigned a = (unsigned)(v < 0 ? -v : v);
  const unsigned long long c = a > (1u << 20) - 1 ? (1u << 20) : a;
  return c * c;
}

// The power POW of a window integer for the master energies: a^2 in lo (POW
// 2), or the low and high 40 bits of a^4 < 2^80 (POW 4, |a| < 2^20; a patch
// with a value beyond takes the search's exact loop, which does not use these
// energies), so that sums of 128 parts stay far below 2^63.
template <int POW>
__device__ __forceinline__ void pow_parts(float f, float mf, long long& lo, long long& hi) {
  const int v = (int)(f - mf);
  if (POW == 2) {
    lo = (long long)v * v;
    hi = 0;
  } else {
    const unsigned long long sq = sq_clamped(v);
    const u128 q = (u128)sq * sq;
    lo = (long long)(unsigned long long)(q & ((1ull << 40) - 1));
    hi = (long long)(unsigned long long)(q >> 40);
  }
}

// Exact master energies sum_{y, x < 64} A[y + dy][x + dx]^2 of every lag of a
// 128 x 128 window (xsearch = ysearch = 32; int64 sliding sums), as doubles
// (exact: < 2^53).  One block per patch.  (-time4's fourth powers need
// 128-bit sums: k_time_energy_rows<4, 128>.)
__global__ void __launch_bounds__(kThreads) k_time_energy32(DevParams P, const float* __restrict__ A1,
                                                            const int* __restrict__ rb,
                                                            const unsigned long long* __restrict__ sums,
                                                            double* __restrict__ aa, int nb) {
  __shared__ long long E[127 * 64];
  const int b = xcd_item();                     // neighbours in a strip share an XCD (and its L2)
  if (b >= nb) return;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const float m1f = (float)patch_mean(sums[2 * b], P.N);
  const float* a1 = rb ? A1 + (size_t)rb[b] * 128 : A1 + (size_t)(2 * b) * P.N;
  // row energies E[r][dx] = sum_{x < 64} A[r][x + dx]^2, one row per wave and
  // step: lane l holds A[r][l] and A[r][64 + l], and
  // E[r][l] = (sum_{x >= l} A[r][x]^2) + (sum_{x < l} A[r][64 + x]^2),
  // a suffix and an exclusive prefix sum over the wave (exact in int64)
  static_assert(kThreads == 256, "four waves, 32 rows each");
  float f0[32], f1[32];                           // all loads of the wave's rows in flight together
#pragma unroll
  for (int k = 0; k < 32; k++) {
    const int r = min(wave + 4 * k, 126);
    f0[k] = a1[r * 128 + lane];
    f1[k] = a1[r * 128 + 64 + lane];
  }
#pragma unroll
  for (int k = 0; k < 32; k++) {
    const int r = wave + 4 * k;
    if (r >= 127) break;
    long long q0, q1, unused;
    pow_parts<2>(f0[k], m1f, q0, unused);
    pow_parts<2>(f1[k], m1f, q1, unused);
    const long long p0 = wave_incl_scan_i64(q0), p1 = wave_incl_scan_i64(q1);
    const long long tot0 = wave_lane63_i64(p0);
    E[r * 64 + lane] = (tot0 - p0 + q0) + (p1 - q1);
  }
  __syncthreads();
  // aa[dy][dx] = sum_{y < 64} E[dy + y][dx]: thread (dx, d0) slides down lags d0 .. d0 + 15
  const int dx = lane, d0 = 16 * wave;
  long long s = 0;
  for (int y = 0; y < 64; y++) s += E[(d0 + y) * 64 + dx];
  double* ob = aa + (size_t)b * 64 * 64;
  ob[d0 * 64 + dx] = (double)s;
  for (int dy = d0 + 1; dy < d0 + 16; dy++) {
    s += E[(dy + 63) * 64 + dx] - E[(dy - 1) * 64 + dx];
    ob[dy * 64 + dx] = (double)s;
  }
}

__device__ __forceinline__ void limbs7(int v, signed char* l) {
  const int l0 = ((v + 64) & 127) - 64;
  const int v1 = (v - l0) >> 7;
  const int l1 = ((v1 + 64) & 127) - 64;
  l[0] = (signed char)l0;
  l[1] = (signed char)l1;
  l[2] = (signed char)((v1 - l1) >> 7);
}

// 16 bytes of a limb row starting at byte offset o (any alignment): the five
// dwords that cover them and one byte-align per output dword
// (v_alignbyte_b32: bytes sh .. sh + 3 of the pair {hi, lo}).  The callers'
// offsets differ by the lane within each group of four lanes (o & 3 = lane &
// 3, the same dword-aligned start), so the four lanes share the five dwords:
// lane k of the quad reads dword k, all read dword 4, and DPP quad broadcasts
// hand each lane the other three (all lanes active).
__device__ __forceinline__ v4i_t ld16_unaligned(const signed char* row, int o) {
  const unsigned* wp = (const unsigned*)(row + (o & ~3));
  const unsigned sh = (unsigned)(o & 3);
  const unsigned mine = wp[threadIdx.x & 3];
  unsigned w[5];
  w[0] = (unsigned)__builtin_amdgcn_mov_dpp((int)mine, 0x00, 0xf, 0xf, false);   // quad_perm 0,0,0,0
  w[1] = (unsigned)__builtin_amdgcn_mov_dpp((int)mine, 0x55, 0xf, 0xf, false);   // quad_perm 1,1,1,1
  w[2] = (unsigned)__builtin_amdgcn_mov_dpp((int)mine, 0xAA, 0xf, 0xf, false);   // quad_perm 2,2,2,2
  w[3] = (unsigned)__builtin_amdgcn_mov_dpp((int)mine, 0xFF, 0xf, 0xf, false);   // quad_perm 3,3,3,3
  w[4] = wp[4];
  v4i_t r;
#pragma unroll
  for (int k = 0; k < 4; k++) r[k] = (int)__builtin_amdgcn_alignbyte(w[k + 1], w[k], sh);
  return r;
}

constexpr int kLimbPitch = 144;                   // bytes per limb row of the window (128 + 16 for the tail loads)
constexpr int kBPad = 32, kBRows = 128;           // template limb rows: 32 zero rows, 64 rows, 32 zero rows
// Byte offset of 16-byte chunk c of template limb row `row`: the chunks of a
// row are rotated by row / 4, so the 16 lanes of a group, which read chunk c
// of 16 consecutive rows, hit 16 different 16-byte bank groups.
__device__ __forceinline__ int bl_chunk(int row, int c) { return 16 * ((c + (row >> 2)) & 3); }

// The wave's 2 x 2 tiles of 16 x 16 lags (dx0 + 16 i + .., dy0 + 16 j + ..):
// for every master row r that meets them, the Toeplitz rows A[r][x + dx]
// times the template rows, one 16 x 16 x 64 int8 product per limb pair
// (ka < KA, kb < KB), accumulated in acc[i][j][ka + kb].
template <int KA, int KB, bool DB = (KA * KB <= 16), int NS>
__device__ __forceinline__ void i8_lag_tiles(const signed char* Al, const signed char* Bl, int dx0, int dy0, int m,
                                             int g, v4i_t (&acc)[2][2][NS]) {
  static_assert(KA + KB - 1 <= NS, "one accumulator per limb weight");
  auto load = [&](int r, v4i_t (&A)[2][KA], v4i_t (&B)[2][KB]) {
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int k = 0; k < KA; k++)                // Toeplitz rows: A[r][x + dx], dx = dx0 + 16 i + m, x = 16 g ..
        A[i][k] = ld16_unaligned(Al + (k * 128 + r) * kLimbPitch, dx0 + 16 * i + m + 16 * g);
#pragma unroll
    for (int j = 0; j < 2; j++) {
      // template row of lag row dy = dy0 + 16 j + m: y = r - dy lies in [-31, 94];
      // the rows outside [0, 64) are the zero rows around the template
      const int y = r - (dy0 + 16 * j + m);
#pragma unroll
      for (int k = 0; k < KB; k++) {
        const int row = k * kBRows + kBPad + y;
        B[j][k] = *(const v4i_t*)(Bl + row * 64 + bl_chunk(row, g));
      }
    }
  };
  // lag-row tile j meets master rows dy0 + 16 j .. dy0 + 16 j + 78 only: the
  // products of a row outside that range are zero for every lane and skipped
  // (the bounds are wave-uniform)
  auto mma = [&](int r, const v4i_t (&A)[2][KA], const v4i_t (&B)[2][KB]) {
#pragma unroll
    for (int j = 0; j < 2; j++) {
      if (r < dy0 + 16 * j || r >= dy0 + 16 * j + 79) continue;
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int ka = 0; ka < KA; ka++)
#pragma unroll
          for (int kb = 0; kb < KB; kb++)
            acc[i][j][ka + kb] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[i][ka], B[j][kb], acc[i][j][ka + kb], 0, 0, 0);
    }
  };
  // master rows r = dy0 .. dy0 + 94 meet the wave's lag rows; the operands of
  // row r + 1 load while the products of row r run (with more than 16 limb
  // products per tile and row, one operand set: the products cover the loads)
  const int r_end = dy0 + 32 + 63;
  if constexpr (!DB) {
    for (int r = dy0; r < r_end; r++) {
      v4i_t A0[2][KA], B0[2][KB];
      load(r, A0, B0);
      mma(r, A0, B0);
    }
  } else {
    v4i_t A0[2][KA], B0[2][KB], A1[2][KA], B1[2][KB];
    load(dy0, A0, B0);
    int r = dy0;
    for (; r + 1 < r_end; r += 2) {
      load(r + 1, A1, B1);
      mma(r, A0, B0);
      if (r + 2 < r_end) load(r + 2, A0, B0);
      mma(r + 1, A1, B1);
    }
    if (r < r_end) mma(r, A0, B0);
  }
}

// The window integers (int)(a - m1) of a 128 x 128 block (rows pa floats
// apart from a) and the template integers (int)(t - m2) of a 64 x 64 block
// (rows pb floats apart from t), as NL = 3 signed 7-bit limbs each (SQ, for
// -time4: their squares, below 2^40, as six unsigned 7-bit limbs, of which the
// first NL = 4 are written here and limbs 4 and 5 by pack_i8_sq_high), into
// the LDS layouts i8_lag_tiles reads: Al, NL x 128 limb rows of kLimbPitch
// bytes (the last 16 zero); Bl, NL x kBRows rows of 64 bytes with kBPad zero
// rows on either side of the 64 template rows.  big: a value outside the limb
// range (|v| >= 2^20); hia / hib for the window / the template: limb 2 nonzero
// somewhere (SQ: bit 0, limb 3 nonzero, a square >= 2^21; bit 1, limbs 4 and 5
// in use, a square >= 2^28); e2 += the template's sum of squares (SQ: of
// fourth powers, 128-bit).  Four columns per thread and
// step: the float4 loads of the window (16 per thread, in NH groups) and of the
// template (4, with the first group) are issued before their first use; each
// limb row takes one dword per 4 columns.  NH = 2 halves the registers the
// loads hold, for callers that keep accumulators live across blocks.
template <int NH = 1, bool SQ = false>
__device__ __forceinline__ void pack_i8_limbs(const float* __restrict__ a1, int pa, const float* __restrict__ t,
                                              int pb, float m1f, float m2f, signed char* Al, signed char* Bl,
                                              int& big, int& hia, int& hib,
                                              typename std::conditional<SQ, u128, long long>::type& e2s) {
  static_assert(kThreads == 256, "16 window and 4 template float4 per thread");
  constexpr int NL = SQ ? 4 : 3;
  using E_t = typename std::conditional<SQ, u128, long long>::type;
  auto pack4 = [&](const float4 f, float mf, unsigned (&w)[NL], int& hi, E_t* e2) {
    const int v[4] = {(int)(f.x - mf), (int)(f.y - mf), (int)(f.z - mf), (int)(f.w - mf)};
#pragma unroll
    for (int k = 0; k < NL; k++) w[k] = 0u;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      if constexpr (SQ) {
        const unsigned long long sq = sq_clamped(v[q]);   // < 2^40: limbs (sq >> 7 k) & 127
        big |= sq >= kSqLimit;
        if (e2) *e2 += (u128)sq * sq;
        hi |= ((sq >> 21) ? 1 : 0) | ((sq >> 28) ? 2 : 0);
#pragma unroll
        for (int k = 0; k < 4; k++) w[k] |= (unsigned)((sq >> (7 * k)) & 127u) << (8 * q);
      } else {
        big |= (v[q] >= (1 << 20) || v[q] <= -(1 << 20));
        if (e2) *e2 += (long long)v[q] * v[q];
        signed char l[3];
        limbs7(v[q], l);
        hi |= l[2];
#pragma unroll
        for (int k = 0; k < 3; k++) w[k] |= (unsigned)(unsigned char)l[k] << (8 * q);
      }
    }
  };
  static_assert(NH == 1 || NH == 2, "window loads in one or two groups");
  constexpr int KW = 16 / NH;
  float4 fb[4];
#pragma unroll
  for (int hh = 0; hh < NH; hh++) {
    float4 fa[KW];
#pragma unroll
    for (int k = 0; k < KW; k++) {
      const int idx = threadIdx.x + kThreads * (hh * KW + k);   // row idx / 32, columns 4 (idx % 32) ..
      fa[k] = ((const float4*)(a1 + (size_t)(idx >> 5) * pa))[idx & 31];
    }
    if (hh == 0)
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int idx = threadIdx.x + kThreads * k;   // template row idx / 16, columns 4 (idx % 16) ..
        fb[k] = ((const float4*)(t + (size_t)(idx >> 4) * pb))[idx & 15];
      }
#pragma unroll
    for (int k = 0; k < KW; k++) {
      const int idx = threadIdx.x + kThreads * (hh * KW + k), r = idx >> 5, x = 4 * (idx & 31);
      unsigned w[NL];
      pack4(fa[k], m1f, w, hia, nullptr);
#pragma unroll
      for (int l = 0; l < NL; l++) *(unsigned*)(Al + (l * 128 + r) * kLimbPitch + x) = w[l];
    }
  }
  for (int e = threadIdx.x; e < NL * 128 * 4; e += kThreads)     // columns 128 .. 143 of every limb row
    *(unsigned*)(Al + (e >> 2) * kLimbPitch + 128 + 4 * (e & 3)) = 0u;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int idx = threadIdx.x + kThreads * k, y = idx >> 4, x = 4 * (idx & 15);
    unsigned w[NL];
    pack4(fb[k], m2f, w, hib, &e2s);
#pragma unroll
    for (int l = 0; l < NL; l++) {
      const int row = l * kBRows + kBPad + y;
      *(unsigned*)(Bl + row * 64 + bl_chunk(row, x >> 4) + (x & 15)) = w[l];
    }
  }
  for (int e = threadIdx.x; e < NL * 64 * 4; e += kThreads) {      // the zero rows, 16 bytes a store
    const int l = e / 256, q = e % 256, row = q < 128 ? q >> 2 : kBPad + 64 + ((q - 128) >> 2);
    *(uint4*)(Bl + (l * kBRows + row) * 64 + 16 * (q & 3)) = make_uint4(0u, 0u, 0u, 0u);
  }
}

// Limbs 4 and 5 of the squares (bits 28 .. 41) of the window and template
// blocks of pack_i8_limbs<NH, true>, for a block pair in which a square
// reaches 2^28: the same layouts, planes 4 and 5 (their tails and zero rows
// too).
__device__ __forceinline__ void pack_i8_sq_high(const float* __restrict__ a1, int pa, const float* __restrict__ t,
                                                int pb, float m1f, float m2f, signed char* Al, signed char* Bl) {
  auto high4 = [](const float4 f, float mf, unsigned (&w)[2]) {
    const int v[4] = {(int)(f.x - mf), (int)(f.y - mf), (int)(f.z - mf), (int)(f.w - mf)};
    w[0] = w[1] = 0u;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const unsigned long long sq = sq_clamped(v[q]);
      w[0] |= (unsigned)((sq >> 28) & 127u) << (8 * q);
      w[1] |= (unsigned)((sq >> 35) & 127u) << (8 * q);
    }
  };
#pragma unroll 4
  for (int k = 0; k < 16; k++) {
    const int idx = threadIdx.x + kThreads * k, r = idx >> 5, x = 4 * (idx & 31);   // row idx / 32, columns 4 (idx % 32) ..
    unsigned w[2];
    high4(((const float4*)(a1 + (size_t)r * pa))[idx & 31], m1f, w);
#pragma unroll
    for (int l = 0; l < 2; l++) *(unsigned*)(Al + ((4 + l) * 128 + r) * kLimbPitch + x) = w[l];
  }
  for (int e = threadIdx.x; e < 2 * 128 * 4; e += kThreads)      // columns 128 .. 143 of the two planes
    *(unsigned*)(Al + (4 * 128 + (e >> 2)) * kLimbPitch + 128 + 4 * (e & 3)) = 0u;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int idx = threadIdx.x + kThreads * k, y = idx >> 4, x = 4 * (idx & 15);
    unsigned w[2];
    high4(((const float4*)(t + (size_t)y * pb))[idx & 15], m2f, w);
#pragma unroll
    for (int l = 0; l < 2; l++) {
      const int row = (4 + l) * kBRows + kBPad + y;
      *(unsigned*)(Bl + row * 64 + bl_chunk(row, x >> 4) + (x & 15)) = w[l];
    }
  }
  for (int e = threadIdx.x; e < 2 * 64 * 4; e += kThreads) {       // the zero rows of the two planes
    const int l = 4 + e / 256, q = e % 256, row = q < 128 ? q >> 2 : kBPad + 64 + ((q - 128) >> 2);
    *(uint4*)(Bl + (l * kBRows + row) * 64 + 16 * (q & 3)) = make_uint4(0u, 0u, 0u, 0u);
  }
}

__global__ void __launch_bounds__(kThreads, 2) k_time_search_i8(DevParams P, const float* __restrict__ A1,
                                                             const float* __restrict__ A2,
                                                             const int* __restrict__ rb,
                                                             const unsigned long long* __restrict__ sums,
                                                             const double* __restrict__ aa,
                                                             double* __restrict__ surfd, int* __restrict__ tpeak,
                                                             double* __restrict__ tcave, int nb) {
  __shared__ __attribute__((aligned(16))) signed char Al[3 * 128 * kLimbPitch];
  __shared__ __attribute__((aligned(16))) signed char Bl[3 * kBRows * 64];
  __shared__ long long redl[kThreads / 64];
  __shared__ int sbig, shi[2];                    // values beyond the limbs; limb 2 nonzero (master, template)
  const int b = xcd_item();                     // neighbours in a strip share an XCD (and its L2)
  if (b >= nb) return;
  const float m1f = (float)patch_mean(sums[2 * b], P.N);
  const float m2f = (float)patch_mean(sums[2 * b + 1], P.N);
  // amplitudes of patch b: rows of 128 from row base rb[b] of the strip lines, or per patch
  const float* a1 = rb ? A1 + (size_t)rb[b] * 128 : A1 + (size_t)(2 * b) * P.N;
  const float* a2 = rb ? A2 + (size_t)rb[b] * 128 : a1 + P.N;
  if (threadIdx.x == 0) {
    sbig = 0;
    shi[0] = 0;
    shi[1] = 0;
  }
  __syncthreads();
  int big = 0, hia = 0, hib = 0;
  long long b2 = 0;
  pack_i8_limbs(a1, 128, a2 + (size_t)P.ys * 128 + P.xs, 128, m1f, m2f, Al, Bl, big, hia, hib, b2);
  if (big) atomicOr(&sbig, 1);
  if (hia) atomicOr(&shi[0], 1);
  if (hib) atomicOr(&shi[1], 1);
  const double d2d = (double)block_sum_i64(b2, redl);     // (its barriers also publish the limbs)
  double* sd = surfd + (size_t)b * 64 * 64;
  const double* ab_aa = aa + (size_t)b * 64 * 64;
  // the surface is staged in LDS (over the window limbs, padded 1 per 16) for the peak scan
  double* gs = (double*)Al;
  if (sbig) {                                     // exact int64 fallback (values beyond the limb range)
    __syncthreads();
    for (int lag = threadIdx.x; lag < 64 * 64; lag += kThreads) {
      const int dy = lag >> 6, dx = lag & 63;
      long long num = 0;
      for (int y = 0; y < 64; y++)
        for (int x = 0; x < 64; x++) {
          const long long av = (int)(a1[(y + dy) * 128 + x + dx] - m1f);
          const long long bv = (int)(a2[(size_t)(P.ys + y) * 128 + P.xs + x] - m2f);
          num += av * bv;
        }
      const double den = sqrt(1.0 * ab_aa[lag] * d2d);
      const double gv = (den == 0.0) ? 0.0 : 100.0 * fabs((double)num / den);
      if (P.keep_surf) sd[lag] = gv;
      gs[lag + (lag >> 4)] = gv;
    }
  } else {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, m = lane & 15, g = lane >> 4;
  // the wave's 2 x 2 tiles of 16 x 16 lags (wave-uniform corners)
  const int dx0 = __builtin_amdgcn_readfirstlane(32 * (wave & 1));
  const int dy0 = __builtin_amdgcn_readfirstlane(32 * (wave >> 1));
  v4i_t acc[2][2][5];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int w = 0; w < 5; w++) acc[i][j][w] = v4i_t{0, 0, 0, 0};
  // limb 2 is zero throughout a window whose values all lie in [-8256, 8127]:
  // its products are skipped (9, 6 or 4 limb products per term instead of 9)
  if (shi[0] && shi[1]) i8_lag_tiles<3, 3>(Al, Bl, dx0, dy0, m, g, acc);
  else if (shi[0]) i8_lag_tiles<3, 2>(Al, Bl, dx0, dy0, m, g, acc);
  else if (shi[1]) i8_lag_tiles<2, 3>(Al, Bl, dx0, dy0, m, g, acc);
  else i8_lag_tiles<2, 2>(Al, Bl, dx0, dy0, m, g, acc);
  __syncthreads();                                // every wave is done with the limbs
  // lane holds lags dx = dx0 + 16 i + 4 g + q, dy = dy0 + 16 j + m
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        long long num = 0;
#pragma unroll
        for (int w = 0; w < 5; w++) num += (long long)acc[i][j][w][q] << (7 * w);
        const int dx = dx0 + 16 * i + 4 * g + q, dy = dy0 + 16 * j + m;
        const int lag = dy * 64 + dx;
        const double den = sqrt(1.0 * ab_aa[lag] * d2d);
        const double gv = (den == 0.0) ? 0.0 : 100.0 * fabs((double)num / den);
        if (P.keep_surf) sd[lag] = gv;
        gs[lag + (lag >> 4)] = gv;
      }
  }
  __syncthreads();
  // Peak with do_time_corr's scan semantics (do_time_int_xcorr.cu:162-180), as
  // k_time_peak: thread t scans lags 16 t .. 16 t + 15 knowing the maximum of
  // all earlier lags (an exclusive prefix maximum over the threads).
  const int t = threadIdx.x;
  const double* gc = gs + 17 * t;
  double mx = -1.0, sm = 0.0;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    mx = gc[k] > mx ? gc[k] : mx;
    sm += gc[k];
  }
  __shared__ double wmax[kThreads / 64];
  double inc = mx;                                // inclusive prefix maximum within the wave
  for (int o = 1; o < 64; o <<= 1) {
    const double v = __shfl_up(inc, o);
    if ((t & 63) >= o) inc = v > inc ? v : inc;
  }
  if ((t & 63) == 63) wmax[t >> 6] = inc;
  __syncthreads();
  double pm = __shfl_up(inc, 1);
  if ((t & 63) == 0) pm = -1.0;
  for (int u = 0; u < (t >> 6); u++) pm = wmax[u] > pm ? wmax[u] : pm;
  long long last = -1;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    const double v = gc[k];
    if (v > (double)(float)pm) last = 16 * t + k;
    pm = v > pm ? v : pm;
  }
  __shared__ unsigned long long redk[kThreads / 64];
  const unsigned long long key = block_max_u64((unsigned long long)(last + 1), redk);
  const double gmax = wmax[0] > wmax[1] ? (wmax[0] > wmax[2] ? (wmax[0] > wmax[3] ? wmax[0] : wmax[3])
                                                             : (wmax[2] > wmax[3] ? wmax[2] : wmax[3]))
                                        : (wmax[1] > wmax[2] ? (wmax[1] > wmax[3] ? wmax[1] : wmax[3])
                                                             : (wmax[2] > wmax[3] ? wmax[2] : wmax[3]));
  const double ssum = block_sum_f64(sm, (double*)redk);
  // Without keep, the slot gets only the 8 x 8 neighbourhood of the peak that
  // k_finish reads (flat index k = (ip - 4 + iy) 64 + jp - 4 + ix, as
  // d_cppow_strip, conv_ampl_demean.cu:75-95), entry (iy, ix) at iy * 8 + ix.
  if (!P.keep_surf && t < 64) {
    const int idx = (int)key - 1;
    const int k = (idx / 64 - 4 + (t >> 3)) * 64 + idx % 64 - 4 + (t & 7);
    sd[t] = (k >= 0 && k < 64 * 64) ? gs[k + (k >> 4)] : 0.0;
  }
  if (t == 0) {
    const int idx = (int)key - 1;
    tpeak[4 * b + 0] = idx / 64;
    tpeak[4 * b + 1] = idx % 64;
    tpeak[4 * b + 2] = __float_as_int((float)gmax);
    tpeak[4 * b + 3] = 0;
    tcave[b] = ssum / 4096.0;
  }
}

// Exact master energies sum_{y, x < NL} A[y + dy][x + dx]^POW of the NL x NL
// lags of a W x W window, NL = W / 2 (W = 256: xsearch = ysearch = 64; W = 128:
// 32), as doubles (POW 2: the value 1.0 * aa takes in the NCC; POW 4: the
// 128-bit sum rounded to nearest).  One block per patch.  Row energies
// E[r][dx] = sum_{x < NL} A[r][x + dx]^POW (POW 4: in two 40-bit parts) of 32
// rows at a time go to LDS: lane l of a wave holds A[r][64 s + l], s < W / 64,
// and E[r][l] = suffix of segment 0 from l + the whole segments between +
// prefix of the next segment below l (W = 256: E[r][l] from segments 0..2,
// E[r][64 + l] from segments 1..3).  Then threads 0..NL - 1 run down one lag
// column each: with S(r) the sum of the first r row energies, S(r) for r < NL
// is parked in the output (POW 4: its high 64 bits in hi, the patch's surface
// slot, which the search overwrites later) and aa[dy] = S(dy + NL) - S(dy)
// overwrites it (int64 or 128-bit throughout, exact).
template <int POW, int W>
__global__ void __launch_bounds__(kThreads) k_time_energy_rows(DevParams P, const float* __restrict__ amp,
                                                               const int* __restrict__ rb,
                                                               const unsigned long long* __restrict__ sums,
                                                               double* __restrict__ aa, double* __restrict__ hi,
                                                               int nb) {
  static_assert(W == 128 || W == 256, "128 x 128 or 256 x 256 windows");
  constexpr int NL = W / 2, SG = W / 64;          // lags per side, 64-sample segments per row
  constexpr int NP = POW == 2 ? 1 : 2;           // parts per row energy
  __shared__ long long E[NP * 32 * NL];
  const int b = xcd_item();
  if (b >= nb) return;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const float m1f = (float)patch_mean(sums[2 * b], P.N);
  // amplitudes of patch b: rows of W from row base rb[b] of the strip lines, or per patch
  const float* a1 = rb ? amp + (size_t)rb[b] * W : amp + (size_t)(2 * b) * P.N;
  double* ob = aa + (size_t)b * NL * NL;
  long long* sb = (long long*)ob;
  unsigned long long* hb = POW == 2 ? nullptr : (unsigned long long*)(hi + (size_t)b * NL * NL);
  const int col = threadIdx.x & (NL - 1);
  using S_t = typename std::conditional<POW == 2, long long, u128>::type;
  S_t S = 0;
  for (int ch = 0; ch < W / 32; ch++) {
    float f[8][SG];                               // rows 32 ch + wave + 4 k: all loads in flight together
#pragma unroll
    for (int k = 0; k < 8; k++)
#pragma unroll
      for (int sg = 0; sg < SG; sg++) f[k][sg] = a1[(32 * ch + wave + 4 * k) * W + 64 * sg + lane];
    // E[r][l] (and E[r][64 + l]) of one part (lane l holds segment sg's q[sg])
    auto row_e = [&](const long long (&q)[SG], long long* e) {
      long long pf[SG], tt[SG];
#pragma unroll
      for (int sg = 0; sg < SG; sg++) {
        pf[sg] = wave_incl_scan_i64(q[sg]);
        tt[sg] = wave_lane63_i64(pf[sg]);
      }
      if constexpr (W == 256) {
        e[lane] = (tt[0] - pf[0] + q[0]) + tt[1] + (pf[2] - q[2]);
        e[64 + lane] = (tt[1] - pf[1] + q[1]) + tt[2] + (pf[SG - 1] - q[SG - 1]);
      } else {
        e[lane] = (tt[0] - pf[0] + q[0]) + (pf[1] - q[1]);
      }
    };
#pragma unroll
    for (int k = 0; k < 8; k++) {
      long long ql[SG], qh[SG];
#pragma unroll
      for (int sg = 0; sg < SG; sg++) pow_parts<POW>(f[k][sg], m1f, ql[sg], qh[sg]);
      const int rr = wave + 4 * k;
      row_e(ql, E + rr * NL);
      if (NP == 2) row_e(qh, E + 32 * NL + rr * NL);
    }
    __syncthreads();
    if (threadIdx.x < NL) {
      // rows r >= NL (chunks NL / 32 ..) subtract S(r - NL), parked NL rows
      // earlier: the parked sums are loaded G rows at a time
      constexpr int G = POW == 2 ? 16 : 8;
      const bool late = ch >= NL / 32;
      for (int r0 = 0; r0 < 32; r0 += G) {
        unsigned long long pl[G], ph[G];
        if (late) {
#pragma unroll
          for (int q = 0; q < G; q++) {
            const int o = ((32 * ch + r0 + q) & (NL - 1)) * NL + col;
            pl[q] = (unsigned long long)sb[o];
            ph[q] = POW == 2 ? 0ull : hb[o];
          }
        }
#pragma unroll
        for (int q = 0; q < G; q++) {
          const int rr = r0 + q;                  // S holds S(32 ch + rr)
          const int o = ((32 * ch + rr) & (NL - 1)) * NL + col;
          if constexpr (POW == 2) {
            if (!late) sb[o] = S;
            else ob[o] = (double)(S - (long long)pl[q]);
            S += E[rr * NL + col];
          } else {
            if (!late) {
              sb[o] = (long long)(unsigned long long)S;
              hb[o] = (unsigned long long)(S >> 64);
            } else {
              ob[o] = u128_to_double(S - (((u128)ph[q] << 64) | pl[q]));
            }
            S += ((u128)(unsigned long long)E[32 * NL + rr * NL + col] << 40) +
                 (unsigned long long)E[rr * NL + col];
          }
        }
      }
    }
    __syncthreads();
  }
}

// -time search on the int8 matrix cores for xsearch = ysearch = 64 (256 x 256
// windows, 128 x 128 template, 128 x 128 lags).  Block (patch b, quadrant q)
// computes the 64 x 64 lags dy = dyq + 0..63, dx = dxq + 0..63, and
// num(dy, dx) = sum over the template blocks (h, k) < 2 of a sub-search of
// k_time_search_i8's shape: the 128 x 128 window block at rows dyq + 64 h,
// columns dxq + 64 k against the 64 x 64 template block at (64 h, 64 k).  The
// four sub-searches add into the same int32 accumulators, each of which sums
// at most 3 x 4 x 4096 limb products of magnitude <= 2^12: still exact.  The
// surface goes to surfd for the peak scan (k_time_peak4k); a patch with a
// value outside the limb range takes the exact int64 loop.
__global__ void __launch_bounds__(kThreads, 2) k_time_search_i8q(DevParams P, const float* __restrict__ A1,
                                                              const float* __restrict__ A2,
                                                              const int* __restrict__ rb,
                                                              const unsigned long long* __restrict__ sums,
                                                              const double* __restrict__ aa,
                                                              double* __restrict__ surfd, int nb) {
  __shared__ __attribute__((aligned(16))) signed char Al[3 * 128 * kLimbPitch];
  __shared__ __attribute__((aligned(16))) signed char Bl[3 * kBRows * 64];
  __shared__ long long redl[kThreads / 64];
  __shared__ int sbig[4], shi[4][2];              // per template block: values beyond the limbs; limb 2 nonzero
  const int item = xcd_item();                    // the four quadrants of a patch are consecutive items
  if (item >= 4 * nb) return;
  const int b = item >> 2, q = item & 3;
  const int dyq = 64 * (q >> 1), dxq = 64 * (q & 1);
  const float m1f = (float)patch_mean(sums[2 * b], P.N);
  const float m2f = (float)patch_mean(sums[2 * b + 1], P.N);
  // amplitudes of patch b: rows of 256 from row base rb[b] of the strip lines, or per patch
  const float* a1 = rb ? A1 + (size_t)rb[b] * 256 : A1 + (size_t)(2 * b) * P.N;
  const float* a2 = rb ? A2 + (size_t)rb[b] * 256 : a1 + P.N;
  if (threadIdx.x < 4) {
    sbig[threadIdx.x] = 0;
    shi[threadIdx.x][0] = 0;
    shi[threadIdx.x][1] = 0;
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, m = lane & 15, g = lane >> 4;
  const int dx0 = __builtin_amdgcn_readfirstlane(32 * (wave & 1));
  const int dy0 = __builtin_amdgcn_readfirstlane(32 * (wave >> 1));
  v4i_t acc[2][2][5];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int w = 0; w < 5; w++) acc[i][j][w] = v4i_t{0, 0, 0, 0};
  long long b2 = 0;
  bool anybig = false;
  for (int ph = 0; ph < 4; ph++) {
    const int h = ph >> 1, k = ph & 1;
    __syncthreads();                              // the flags are cleared; the last block's limbs are read
    int big = 0, hia = 0, hib = 0;
    pack_i8_limbs<2>(a1 + (size_t)(dyq + 64 * h) * 256 + dxq + 64 * k, 256,
                  a2 + (size_t)(P.ys + 64 * h) * 256 + P.xs + 64 * k, 256, m1f, m2f, Al, Bl, big, hia, hib, b2);
    if (big) atomicOr(&sbig[ph], 1);
    if (hia) atomicOr(&shi[ph][0], 1);
    if (hib) atomicOr(&shi[ph][1], 1);
    __syncthreads();                              // the limbs and flags are published
    anybig = anybig || sbig[ph];
    if (anybig) continue;
    // limb 2 is zero throughout a block whose values all lie in [-8256, 8127]
    if (shi[ph][0] && shi[ph][1]) i8_lag_tiles<3, 3, false>(Al, Bl, dx0, dy0, m, g, acc);
    else if (shi[ph][0]) i8_lag_tiles<3, 2, false>(Al, Bl, dx0, dy0, m, g, acc);
    else if (shi[ph][1]) i8_lag_tiles<2, 3, false>(Al, Bl, dx0, dy0, m, g, acc);
    else i8_lag_tiles<2, 2>(Al, Bl, dx0, dy0, m, g, acc);
  }
  const double d2d = (double)block_sum_i64(b2, redl);
  double* sd = surfd + (size_t)b * 128 * 128;
  const double* ab_aa = aa + (size_t)b * 128 * 128;
  if (anybig) {                                   // exact int64 loop (values beyond the limb range)
    for (int l = threadIdx.x; l < 64 * 64; l += kThreads) {
      const int dy = dyq + (l >> 6), dx = dxq + (l & 63);
      long long num = 0;
      for (int y = 0; y < 128; y++)
        for (int x = 0; x < 128; x++) {
          const long long av = (int)(a1[(size_t)(y + dy) * 256 + x + dx] - m1f);
          const long long bv = (int)(a2[(size_t)(P.ys + y) * 256 + P.xs + x] - m2f);
          num += av * bv;
        }
      const int lag = dy * 128 + dx;
      const double den = sqrt(1.0 * ab_aa[lag] * d2d);
      sd[lag] = (den == 0.0) ? 0.0 : 100.0 * fabs((double)num / den);
    }
    return;
  }
  // lane holds lags dx = dxq + dx0 + 16 i + 4 g + q, dy = dyq + dy0 + 16 j + m
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int qq = 0; qq < 4; qq++) {
        long long num = 0;
#pragma unroll
        for (int w = 0; w < 5; w++) num += (long long)acc[i][j][w][qq] << (7 * w);
        const int dx = dxq + dx0 + 16 * i + 4 * g + qq, dy = dyq + dy0 + 16 * j + m;
        const int lag = dy * 128 + dx;
        const double den = sqrt(1.0 * ab_aa[lag] * d2d);
        sd[lag] = (den == 0.0) ? 0.0 : 100.0 * fabs((double)num / den);
      }
}

// -time4 (Gatelli, do_time_int_xcorr.cu:103-144) on the int8 matrix cores for
// xsearch = ysearch = XS = 32 (128 x 128 windows, one block per patch) or 64
// (256 x 256 windows: four lag quadrants of four 64 x 64 template blocks each,
// as k_time_search_i8q).  The numerator sum a^2 b^2 of every lag: with the
// squares (below 2^40 for |a|, |b| < 2^20) written as six unsigned 7-bit
// limbs, the limb products of weight 2^(7 s), s = ka + kb <= 10, add into
// eleven int32 accumulators, each of which sums at most 6 limb pairs x 4
// blocks x 4096 products below 2^14 (< 2^31); sum_s acc_s 2^(7 s) (128-bit)
// is the exact numerator.  The master energies sum a^4 come from
// k_time_energy_rows<4, W>, the template's sum b^4 from the packing
// (128-bit).  Per block pair only the limbs in use are multiplied:
// limbs 4 and 5 (and their second packing pass) where a square reaches 2^28
// (|v| > 16383), else limb 3 where one reaches 2^21 (|v| > 1448), else limbs
// 0 .. 2.  A patch with a value beyond 2^20 (float input) takes an exact
// 128-bit loop (k_time_search<1>'s sums).  The surface goes to surfd for the
// peak scan (k_time_peak4k).
template <int XS>
__global__ void __launch_bounds__(kThreads, 1) k_time4_search_i8(DevParams P, const float* __restrict__ A1,
                                                              const float* __restrict__ A2,
                                                              const int* __restrict__ rb,
                                                              const unsigned long long* __restrict__ sums,
                                                              const double* __restrict__ aa,
                                                              double* __restrict__ surfd, int nb) {
  static_assert(XS == 32 || XS == 64, "128 x 128 or 256 x 256 windows");
  constexpr int W = 4 * XS, NL = 2 * XS, NQ = XS == 64 ? 4 : 1;   // window side, lags (= template side), quadrants
  __shared__ __attribute__((aligned(16))) signed char Al[6 * 128 * kLimbPitch];
  __shared__ __attribute__((aligned(16))) signed char Bl[6 * kBRows * 64];
  __shared__ int sbig[NQ], shi[NQ][2];            // per template block: values beyond the limbs; limbs in use
  const int item = xcd_item();                    // the quadrants of a patch are consecutive items
  if (item >= NQ * nb) return;
  const int b = item / NQ, q = item % NQ;
  const int dyq = 64 * (q >> 1), dxq = 64 * (q & 1);
  const float m1f = (float)patch_mean(sums[2 * b], P.N);
  const float m2f = (float)patch_mean(sums[2 * b + 1], P.N);
  // amplitudes of patch b: rows of W from row base rb[b] of the strip lines, or per patch
  const float* a1 = rb ? A1 + (size_t)rb[b] * W : A1 + (size_t)(2 * b) * P.N;
  const float* a2 = rb ? A2 + (size_t)rb[b] * W : a1 + P.N;
  if (threadIdx.x < NQ) {
    sbig[threadIdx.x] = 0;
    shi[threadIdx.x][0] = 0;
    shi[threadIdx.x][1] = 0;
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, m = lane & 15, g = lane >> 4;
  const int dx0 = __builtin_amdgcn_readfirstlane(32 * (wave & 1));
  const int dy0 = __builtin_amdgcn_readfirstlane(32 * (wave >> 1));
  v4i_t acc[2][2][11];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int w = 0; w < 11; w++) acc[i][j][w] = v4i_t{0, 0, 0, 0};
  u128 b4 = 0;
  bool anybig = false;
#pragma unroll 1
  for (int ph = 0; ph < NQ; ph++) {
    const int h = ph >> 1, k = ph & 1;
    __syncthreads();                              // the flags are cleared; the last block's limbs are read
    int big = 0, hia = 0, hib = 0;
    const float* wb = a1 + (size_t)(dyq + 64 * h) * W + dxq + 64 * k;
    const float* tb = a2 + (size_t)(P.ys + 64 * h) * W + P.xs + 64 * k;
    pack_i8_limbs<2, true>(wb, W, tb, W, m1f, m2f, Al, Bl, big, hia, hib, b4);
    if (big) atomicOr(&sbig[ph], 1);
    if (hia) atomicOr(&shi[ph][0], hia);
    if (hib) atomicOr(&shi[ph][1], hib);
    __syncthreads();                              // the limbs and flags are published
    anybig = anybig || sbig[ph];
    if (anybig) continue;
    const int fa = shi[ph][0], fb = shi[ph][1];
    if ((fa | fb) & 2) {                          // a square >= 2^28: all six limbs
      pack_i8_sq_high(wb, W, tb, W, m1f, m2f, Al, Bl);
      __syncthreads();
      i8_lag_tiles<6, 6>(Al, Bl, dx0, dy0, m, g, acc);
    } else if (fa && fb) {
      i8_lag_tiles<4, 4>(Al, Bl, dx0, dy0, m, g, acc);
    } else if (fa) {
      i8_lag_tiles<4, 3>(Al, Bl, dx0, dy0, m, g, acc);
    } else if (fb) {
      i8_lag_tiles<3, 4>(Al, Bl, dx0, dy0, m, g, acc);
    } else {
      i8_lag_tiles<3, 3>(Al, Bl, dx0, dy0, m, g, acc);
    }
  }
  // the template's sum of b^4 over the block (every thread gets it): partial
  // sums through LDS (over the window limbs, free once every wave is done)
  u128* red = (u128*)Al;
  auto block_sum_u128 = [&](u128 v) -> u128 {
    __syncthreads();
    red[threadIdx.x] = v;
    __syncthreads();
    if (threadIdx.x < 16) {
      u128 t = 0;
      for (int i = 0; i < 16; i++) t += red[16 * threadIdx.x + i];
      red[kThreads + threadIdx.x] = t;
    }
    __syncthreads();
    u128 t = 0;
    for (int i = 0; i < 16; i++) t += red[kThreads + i];
    return t;
  };
  double* sd = surfd + (size_t)b * NL * NL;
  const double* ab_aa = aa + (size_t)b * NL * NL;
  if (anybig) {                                   // exact 128-bit loop (values beyond the limb range)
    u128 t4 = 0;
    for (int e = threadIdx.x; e < NL * NL; e += kThreads) {
      const long long bv = (int)(a2[(size_t)(P.ys + e / NL) * W + P.xs + e % NL] - m2f);
      const unsigned long long ub = (unsigned long long)(bv * bv);
      t4 += (u128)ub * ub;
    }
    const double d2d = u128_to_double(block_sum_u128(t4));
    for (int l = threadIdx.x; l < 64 * 64; l += kThreads) {
      const int dy = dyq + (l >> 6), dx = dxq + (l & 63);
      u128 n4 = 0, a4 = 0;
      for (int y = 0; y < NL; y++)
        for (int x = 0; x < NL; x++) {
          const long long av = (int)(a1[(size_t)(y + dy) * W + x + dx] - m1f);
          const long long bv = (int)(a2[(size_t)(P.ys + y) * W + P.xs + x] - m2f);
          const unsigned long long ua = (unsigned long long)(av * av), ub = (unsigned long long)(bv * bv);
          n4 += (u128)ua * ub;
          a4 += (u128)ua * ua;
        }
      const double den = sqrt(u128_to_double(a4) * d2d);
      const double gg = (den == 0.0) ? 0.0 : fabs(u128_to_double(n4) / den);
      sd[dy * NL + dx] = (gg <= 0.5) ? 0.0 : 100.0 * sqrt(gg * 2.0 - 1.0);
    }
    return;
  }
  const double d2d = u128_to_double(block_sum_u128(b4));
  // lane holds lags dx = dxq + dx0 + 16 i + 4 g + q, dy = dyq + dy0 + 16 j + m
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int qq = 0; qq < 4; qq++) {
        u128 num = 0;
#pragma unroll
        for (int w = 0; w < 11; w++) num += (u128)(unsigned)acc[i][j][w][qq] << (7 * w);
        const int dx = dxq + dx0 + 16 * i + 4 * g + qq, dy = dyq + dy0 + 16 * j + m;
        const int lag = dy * NL + dx;
        const double den = sqrt(ab_aa[lag] * d2d);
        const double gg = (den == 0.0) ? 0.0 : fabs(u128_to_double(num) / den);
        sd[lag] = (gg <= 0.5) ? 0.0 : 100.0 * sqrt(gg * 2.0 - 1.0);
      }
}

// Peak of the time-domain surface with do_time_corr's scan semantics
// (do_time_int_xcorr.cu:162-180): m = -1; for each lag in row-major order,
// if |g| > m then m = (float)|g| and the lag becomes the peak.  Equivalently the
// peak is the last lag whose g exceeds (float)(max of all earlier g); the
// reported correlation is (float)(max g).
__global__ void __launch_bounds__(kThreads) k_time_peak(DevParams P, const double* __restrict__ surfd,
                                                        int* __restrict__ tpeak, double* __restrict__ tcave) {
  __shared__ double smax[kThreads];
  __shared__ double ssum[kThreads];
  __shared__ unsigned long long redk[kThreads / 64];
  const int b = blockIdx.x;
  const int total = P.nxc * P.nyc;
  const int chunk = (total + kThreads - 1) / kThreads;
  const int lo = min(total, (int)threadIdx.x * chunk), hi = min(total, lo + chunk);
  const double* g = surfd + (size_t)b * total;
  double mx = -1.0, sm = 0.0;
  for (int i = lo; i < hi; i++) {
    const double v = fabs(g[i]);
    mx = v > mx ? v : mx;
    sm += g[i];
  }
  smax[threadIdx.x] = mx;
  ssum[threadIdx.x] = sm;
  __syncthreads();
  double pm = -1.0;
  for (int u = 0; u < (int)threadIdx.x; u++) pm = smax[u] > pm ? smax[u] : pm;
  long long last = -1;
  for (int i = lo; i < hi; i++) {
    const double v = fabs(g[i]);
    if (v > (double)(float)pm) last = i;
    pm = v > pm ? v : pm;
  }
  const unsigned long long key = block_max_u64((unsigned long long)(last + 1), redk);
  if (threadIdx.x == 0) {
    double gmax = -1.0, s = 0.0;
    for (int u = 0; u < kThreads; u++) {
      gmax = smax[u] > gmax ? smax[u] : gmax;
      s += ssum[u];
    }
    const int idx = (int)key - 1;
    tpeak[4 * b + 0] = idx / P.nxc;
    tpeak[4 * b + 1] = idx % P.nxc;
    tpeak[4 * b + 2] = __float_as_int((float)gmax);
    tpeak[4 * b + 3] = 0;
    tcave[b] = s / (double)total;
  }
}

// k_time_peak for surfaces of a multiple of 4096 lags (-time / -time4 with
// nxc nyc >= 4096): the surface passes through LDS 4096 lags at a time with
// coalesced loads (padded 1 per 16), and thread t scans lags 16 t .. 16 t + 15
// of each chunk knowing the maximum of every earlier lag -- a wave prefix
// maximum by lane shifts, the earlier waves' maxima and the earlier chunks'
// carry -- as k_time_search_i8's fused scan.  Same peak, same (float) maximum;
// the mean sums in another order.
__global__ void __launch_bounds__(kThreads) k_time_peak4k(DevParams P, const double* __restrict__ surfd,
                                                          int* __restrict__ tpeak, double* __restrict__ tcave) {
  __shared__ double gs[4096 + 256];
  __shared__ double wmax[kThreads / 64];
  __shared__ unsigned long long redk[kThreads / 64];
  const int b = blockIdx.x, t = threadIdx.x, lane = t & 63;
  const int total = P.nxc * P.nyc;
  const double* g = surfd + (size_t)b * total;
  const double* gc = gs + 17 * t;
  double carry = -1.0, sm = 0.0;                  // carry: the maximum of every earlier chunk
  long long last = -1;
  for (int c0 = 0; c0 < total; c0 += 4096) {
    __syncthreads();                              // the last chunk's values and maxima are read
#pragma unroll
    for (int k = 0; k < 16; k++) {
      const int e = t + kThreads * k;
      gs[e + (e >> 4)] = g[c0 + e];
    }
    __syncthreads();
    double mx = -1.0;
#pragma unroll
    for (int k = 0; k < 16; k++) {
      mx = fabs(gc[k]) > mx ? fabs(gc[k]) : mx;
      sm += gc[k];
    }
    double inc = mx;                              // inclusive prefix maximum within the wave
    for (int o = 1; o < 64; o <<= 1) {
      const double v = __shfl_up(inc, o);
      if (lane >= o) inc = v > inc ? v : inc;
    }
    if (lane == 63) wmax[t >> 6] = inc;
    __syncthreads();
    double pm = __shfl_up(inc, 1);
    if (lane == 0) pm = -1.0;
    for (int u = 0; u < (t >> 6); u++) pm = wmax[u] > pm ? wmax[u] : pm;
    pm = carry > pm ? carry : pm;
#pragma unroll
    for (int k = 0; k < 16; k++) {
      const double v = fabs(gc[k]);
      if (v > (double)(float)pm) last = c0 + 16 * t + k;
      pm = v > pm ? v : pm;
    }
    for (int u = 0; u < kThreads / 64; u++) carry = wmax[u] > carry ? wmax[u] : carry;
  }
  const unsigned long long key = block_max_u64((unsigned long long)(last + 1), redk);
  const double ssum = block_sum_f64(sm, (double*)redk);
  if (t == 0) {
    const int idx = (int)key - 1;
    tpeak[4 * b + 0] = idx / P.nxc;
    tpeak[4 * b + 1] = idx % P.nxc;
    tpeak[4 * b + 2] = __float_as_int((float)carry);
    tpeak[4 * b + 3] = 0;
    tcave[b] = ssum / (double)total;
  }
}

__global__ void k_init(unsigned long long* __restrict__ sums, unsigned long long* __restrict__ peak, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 2 * n) sums[i] = 0ull;
  if (i < n) peak[i] = 0ull;
}

// --------------------------------------------------------------------------
// host-side tables (same construction as the CPU restatement's twiddles)
// --------------------------------------------------------------------------
static void twiddle64(int n, long m, double* c, double* s) {
  long mm = m % n;
  if (mm < 0) mm += n;
  if (n < 4) {
    *c = (mm == 0) ? 1.0 : -1.0;
    *s = 0.0;
    return;
  }
  const long q = (4 * mm) / n;
  const long r = mm - q * (n / 4);
  double c0, s0;
  if (r == 0) {
    c0 = 1.0; s0 = 0.0;
  } else if (8 * r == n) {
    c0 = s0 = 0.70710678118654752440;
  } else if (8 * r < n) {
    const double t = 2.0 * M_PI * (double)r / (double)n;
    c0 = cos(t); s0 = sin(t);
  } else {
    const double t = 2.0 * M_PI * (double)(n / 4 - r) / (double)n;
    c0 = sin(t); s0 = cos(t);
  }
  switch (q) {
    case 0: *c = c0; *s = s0; break;
    case 1: *c = -s0; *s = c0; break;
    case 2: *c = -c0; *s = -s0; break;
    default: *c = s0; *s = -c0; break;
  }
}
static float2 twf(int n, long m) {
  double c, s;
  twiddle64(n, m, &c, &s);
  return make_float2((float)c, (float)s);
}

}  // namespace xc

using namespace xc;

// --------------------------------------------------------------------------
// context
// --------------------------------------------------------------------------
struct KernelTimer {
  const char* name;
  float ms;
  int launches;
};

// Scratch of the host-side window sort (strip_order), kept by a context across calls.
struct OrderScratch {
  std::vector<unsigned long long> key;
  std::vector<int> tmp, cnt;
};

struct xcorr_hip_ctx {
  xcorr_hip_params p{};
  DevParams dp{};
  int device = 0;
  hipStream_t stream = nullptr;
  int max_batch = 0;
  const void* d_master = nullptr;
  const void* d_slave = nullptr;
  void* own_master = nullptr;
  void* own_slave = nullptr;
  bool images = false;
  float* amp = nullptr;
  unsigned long long* sums = nullptr;
  float2* Z = nullptr;
  float2* D = nullptr;
  float* surf = nullptr;
  double* surfd = nullptr;
  unsigned long long* peak = nullptr;
  double* partial = nullptr;
  int* tpeak = nullptr;
  double* tcave = nullptr;
  xcorr_hip_window* d_win = nullptr;
  xcorr_hip_result* d_out = nullptr;
  float2* tabs = nullptr;
  // strip mode (npx = npy = 256, ri = 2, frequency): per-line buffers
  bool strips = false;
  int max_lines = 0;                      // line capacity of a batch (multiple of kStripBlockLines)
  float* amp1s = nullptr;
  float* amp2s = nullptr;
  long long* lsum = nullptr;
  size_t z_elems = 0;                     // capacity of Z (float2)
  unsigned char* need = nullptr;          // needM [max_lines] then needU [max_lines]
  int* d_plan = nullptr;                  // rb | perm | blk_strip | strips, per batch
  void* h_stage[2] = {nullptr, nullptr};  // pinned staging of the plan, two batches in flight
  hipEvent_t ev_stage[2] = {nullptr, nullptr};
  bool stage_used[2] = {false, false};
  int stage_next = 0;
  size_t plan_bytes = 0;
  std::vector<int> order;                 // windows sorted by strip key
  OrderScratch order_scratch;
  std::vector<xcorr_hip_window> hwin;     // host copy of device windows
  std::string err;
  std::vector<KernelTimer> timers;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<int> last_perm, last_rb;    // last batch: location index and row base of each slot
  // streamed images: copy stream, two band buffers per image, pinned chunks
  hipStream_t cstream = nullptr;
  void* band[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};   // [image][buffer]
  size_t band_bytes[2] = {0, 0};
  void* chunk[2] = {nullptr, nullptr};
  size_t chunk_bytes = 0;
  hipEvent_t ev_chunk[2] = {nullptr, nullptr}, ev_loaded[2] = {nullptr, nullptr}, ev_done[2] = {nullptr, nullptr};
  bool chunk_used[2] = {false, false};
  int chunk_next = 0;
  xcorr_hip_result* d_res = nullptr;
  size_t d_res_n = 0;
  // tuning (xcorr_hip_set_tuning / xcorr_hip_autotune): none of these change a result bit
  int tune_batch = 0;                     // locations per launch (0: max_batch)
  int nt_rows = kThreads;                 // threads of the generic row / column kernels
  bool reg_paths = true;                  // register-resident kernels where the shape has them
  double* tcaa = nullptr;                 // -time / -time4, xsearch = ysearch = 32 or 64: master energies of every lag
  // strip batches alternate between this context and a second one (own
  // buffers and stream, created on the first call with more than one batch),
  // so the kernels of one batch overlap those of the next
  xcorr_hip_ctx* twin = nullptr;
  bool twin_failed = false;               // allocation failed: one context, no overlap
  bool one_context = false;               // never use a twin (contexts of a multi-pair graph)
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  xcorr_hip_result* h_res = nullptr;      // pinned host results of one batch (host-output calls)
  const xcorr_hip_ctx* last_buf = nullptr;   // context whose buffers hold the last batch
};

static std::string g_create_err;

static int fail(xcorr_hip_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  else g_create_err = msg;
  return code;
}

// Device buffers of one batch of the per-patch kernels (indexed by the
// patch's slot in the batch; no stride depends on the batch capacity).  A
// context's own buffers, or ones carved from the static array of
// xcorr_hip_correlate_noalloc.
struct PatchBufs {
  float* amp;                   // [B][2][N] amplitudes
  unsigned long long* sums;     // [B][2] exact patch sums
  float2* Z;                    // [B][N] row / 2-D spectra
  float2* D;                    // [B][nyc][NX] lag rows
  float* surf;                  // [B][nyc][nxc] correlation surface (frequency)
  double* surfd;                // [B][nyc][nxc] correlation surface (time modes)
  unsigned long long* peak;     // [B] peak keys
  double* partial;              // [B][max(groups4, nyc/G4)] surface partial sums
  int* tpeak;                   // [B][4] time-mode peaks
  double* tcave;                // [B] time-mode surface means
  double* tcaa;                 // [B][nyc][nxc] -time / -time4 on 128 x 128 and 256 x 256 windows: master energies of every lag
};

static PatchBufs ctx_bufs(const xcorr_hip_ctx* c) {
  return PatchBufs{c->amp, c->sums, c->Z, c->D, c->surf, c->surfd, c->peak, c->partial, c->tpeak, c->tcave, c->tcaa};
}

#define HIPCK(ctx, call)                                                                  \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return fail(ctx, XCORR_HIP_EDEVICE, std::string(#call ": ") + hipGetErrorString(e_)); \
  } while (0)

static bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }
static bool same_plan(const FftPlan& a, const FftPlan& b) {
  if (a.np != b.np) return false;
  for (int i = 0; i < a.np; i++)
    if (a.R[i] != b.R[i]) return false;
  return true;
}
static long long round_up(long long v, long long m) { return (v + m - 1) / m * m; }

extern "C" int xcorr_hip_abi_version(void) { return XCORR_HIP_ABI_VERSION; }

extern "C" void xcorr_hip_default_params(xcorr_hip_params* p, int xsearch, int ysearch) {
  memset(p, 0, sizeof(*p));
  p->xsearch = xsearch;
  p->ysearch = ysearch;
  p->nx_corr = 2 * xsearch;
  p->ny_corr = 2 * ysearch;
  p->npx = p->nx_corr + 2 * xsearch;
  p->npy = p->ny_corr + 2 * ysearch;
  p->nxc = 2 * xsearch;
  p->nyc = 2 * ysearch;
  p->ri = 2;
  p->interp_flag = 1;
  p->interp_factor = 16;
  p->n2x = 8;
  p->n2y = 8;
  p->corr_flag = XCORR_HIP_FREQ;
  p->format = XCORR_HIP_CINT16;
}

static int validate(const xcorr_hip_params* p, std::string* why) {
  auto bad = [&](const char* m) { *why = m; return XCORR_HIP_EINVAL; };
  if (!is_pow2(p->xsearch) || !is_pow2(p->ysearch) || p->xsearch < 4 || p->ysearch < 4 ||
      p->xsearch > 256 || p->ysearch > 256)
    return bad("xsearch/ysearch must be powers of two in [4, 256]");
  if (p->nx_corr != 2 * p->xsearch || p->ny_corr != 2 * p->ysearch || p->npx != 4 * p->xsearch ||
      p->npy != 4 * p->ysearch || p->nxc != 2 * p->xsearch || p->nyc != 2 * p->ysearch)
    return bad("derived sizes must follow parse_xcorr_input.c (nx_corr = nxc = 2*xsearch, npx = 4*xsearch)");
  if (!is_pow2(p->ri) || p->ri > 16) return bad("ri must be 1 or a power of two <= 16");
  if (p->interp_flag && (!is_pow2(p->interp_factor) || p->interp_factor > 32))
    return bad("interp_factor must be a power of two <= 32");
  if (p->n2x != 8 || p->n2y != 8) return bad("n2x/n2y must be 8");
  if (p->corr_flag != XCORR_HIP_FREQ && p->corr_flag != XCORR_HIP_TIME && p->corr_flag != XCORR_HIP_TIME4)
    return bad("corr_flag must be 0, 1 or 2");
  if (p->format != XCORR_HIP_CINT16 && p->format != XCORR_HIP_FLOAT32) return bad("format must be 0 or 1");
  return 0;
}

static void free_scratch(xcorr_hip_ctx* c) {
  void* ptrs[] = {c->amp, c->sums, c->Z, c->D, c->surf, c->surfd, c->peak, c->partial, c->tpeak,
                  c->tcave, c->d_win, c->d_out, c->tabs, c->amp1s, c->amp2s, c->lsum, c->need, c->d_plan,
                  c->tcaa};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  for (int i = 0; i < 2; i++) {
    if (c->h_stage[i]) (void)hipHostFree(c->h_stage[i]);
    if (c->ev_stage[i]) (void)hipEventDestroy(c->ev_stage[i]);
    c->h_stage[i] = nullptr;
    c->ev_stage[i] = nullptr;
  }
  c->amp = nullptr; c->sums = nullptr; c->Z = nullptr; c->D = nullptr; c->surf = nullptr;
  c->surfd = nullptr; c->peak = nullptr; c->partial = nullptr; c->tpeak = nullptr; c->tcave = nullptr;
  c->d_win = nullptr; c->d_out = nullptr; c->tabs = nullptr; c->amp1s = nullptr; c->amp2s = nullptr;
  c->lsum = nullptr; c->need = nullptr; c->d_plan = nullptr; c->tcaa = nullptr;
  c->z_elems = 0;
}

// Frequency-mode square 128 / 512 windows with ri 2 and the default radix
// plans run the register-resident kernels (k_stageA, k_colpass,
// k_rowifft_surf_reg).
static bool reg_shape(const xcorr_hip_ctx* c) {
  const DevParams& d = c->dp;
  return c->reg_paths && d.corr_flag == XCORR_HIP_FREQ && d.ri == 2 && d.NX == d.NY &&
         (d.NX == 128 || d.NX == 512) && same_plan(d.plx, plan_for(d.NX)) && same_plan(d.ply, plan_for(d.NY));
}
// Surface partial sums per patch: one for the one-block-per-surface row
// kernels (npx = 256 with the default plan, the register kernels), else one
// per k_rowifft_surf block.
static int groups4_for(const xcorr_hip_ctx* c) {
  const DevParams& d = c->dp;
  if ((d.NX == 256 && same_plan(d.plx, plan_for(256))) || reg_shape(c)) return 1;
  return d.nyc / d.G4;
}

// Strip contexts (square 128 / 256 / 512 windows, ri 2, frequency mode) run
// the strips with the default radix plans; a caller-set plan runs the
// per-patch kernels, whose buffers are allocated when such a plan is first set.
// -time and -time4 on 128 x 128 and 256 x 256 windows (xsearch = ysearch = 32
// or 64) run on the int8 matrix cores, with the master energies of every lag
// in a buffer of their own (tcaa).
static bool time_i8_shape(int corr_flag, int xs, int ys) {
  return (corr_flag == XCORR_HIP_TIME || corr_flag == XCORR_HIP_TIME4) && (xs == 32 || xs == 64) && ys == xs;
}

static bool use_strips(const xcorr_hip_ctx* c) {
  const DevParams& d = c->dp;
  return c->strips && same_plan(d.plx, plan_for(d.NX)) && same_plan(d.ply, plan_for(d.NY));
}
static int ensure_patch_buffers(xcorr_hip_ctx* c) {
  const size_t B = (size_t)c->max_batch, N = (size_t)c->dp.N;
  if (!c->amp && hipMalloc((void**)&c->amp, B * 2 * N * sizeof(float)) != hipSuccess) return XCORR_HIP_ENOMEM;
  if (c->dp.corr_flag == XCORR_HIP_FREQ && c->z_elems < B * N) {
    float2* z = nullptr;
    if (hipMalloc((void**)&z, B * N * sizeof(float2)) != hipSuccess) return XCORR_HIP_ENOMEM;
    if (hipStreamSynchronize(c->stream) != hipSuccess) return XCORR_HIP_EDEVICE;
    (void)hipFree(c->Z);
    c->Z = z;
    c->z_elems = B * N;
  }
  return 0;
}

// Plan layout (ints): rb[max_batch] | perm[max_batch] | blk[2 max_lines / 64] | StripDesc[max_batch]
static size_t plan_ints(const xcorr_hip_ctx* c) {
  return 2 * (size_t)c->max_batch + 2 * (size_t)c->max_lines / kStripBlockLines +
         (size_t)c->max_batch * (sizeof(StripDesc) / sizeof(int));
}

extern "C" int xcorr_hip_create(const xcorr_hip_params* p, int max_batch, int device, xcorr_hip_ctx** out) {
  if (!p || !out) return fail(nullptr, XCORR_HIP_EINVAL, "null argument");
  *out = nullptr;
  std::string why;
  if (validate(p, &why)) return fail(nullptr, XCORR_HIP_EINVAL, why);
  xcorr_hip_ctx* c = new xcorr_hip_ctx();
  c->p = *p;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    delete c;
    return fail(nullptr, XCORR_HIP_EDEVICE, "no HIP device");
  }
  if (device < 0) {
    if (hipGetDevice(&device) != hipSuccess) device = 0;
  }
  if (device >= ndev) {
    delete c;
    return fail(nullptr, XCORR_HIP_EINVAL, "device index out of range");
  }
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return fail(nullptr, XCORR_HIP_EDEVICE, "stream creation failed");
  }
  DevParams& d = c->dp;
  d.NX = p->npx; d.NY = p->npy; d.N = d.NX * d.NY;
  d.xs = p->xsearch; d.ys = p->ysearch; d.nx_corr = p->nx_corr; d.ny_corr = p->ny_corr;
  d.nxc = p->nxc; d.nyc = p->nyc; d.ri = p->ri; d.interp_flag = p->interp_flag;
  d.ifc = p->interp_factor; d.n2x = p->n2x; d.n2y = p->n2y;
  d.Lx = p->n2x * p->interp_factor; d.Ly = p->n2y * p->interp_factor;
  d.corr_flag = p->corr_flag; d.format = p->format;
  d.plx = plan_for(d.NX);
  d.ply = plan_for(d.NY);
  d.G1 = std::min(d.NY, std::max(1, kTile / d.NX));
  d.G4 = std::min(d.nyc, std::max(2, kTile / d.NX));   // k_rowifft_surf pairs its rows
  d.groups4 = groups4_for(c);                      // needs the plans and G4
  d.C3 = std::min(d.NX, std::max(1, kTile / d.NY));
  // time search: LR surface rows x nxc lags per block (<= 2 lags per thread),
  // template rows staged RC at a time within 64 KB of LDS
  d.tLR = std::min(d.nyc, std::max(1, kThreads / d.nxc));
  {
    const int budget = 65536 / 4 - (d.tLR - 1) * d.NX;
    d.tRC = std::max(1, std::min(16, budget / (d.NX + d.nx_corr)));
  }

  // tables: e^{+2 pi i m/n} for n = NX, NY, Lx, Ly; range-interp phases
  std::vector<float2> tab;
  auto add_table = [&](int n) {
    int off = (int)tab.size();
    for (int m = 0; m < n; m++) tab.push_back(twf(n, m));
    return off;
  };
  d.tw_nx = add_table(d.NX);
  d.tw_ny = add_table(d.NY);
  d.tw_lx = add_table(d.Lx);
  d.tw_ly = add_table(d.Ly);
  auto add_mask = [&](int n, int sx) {   // spectrum of the box [sx, n - sx), summed in double
    int off = (int)tab.size();
    for (int k = 0; k < n; k++) {
      double re = 0.0, im = 0.0;
      for (int x = sx; x < n - sx; x++) {
        double cc, ss;
        twiddle64(n, (long)k * x, &cc, &ss);
        re += cc;
        im -= ss;
      }
      tab.push_back(make_float2((float)re, (float)im));
    }
    return off;
  };
  d.mk_x = add_mask(d.NX, d.xs);
  d.mk_y = add_mask(d.NY, d.ys);
  d.e4 = (int)tab.size();                 // E4[s][k - 1] = e^{2 pi i k s / Ly}, k = 1..4, s = 0..Ly/2
  for (int s = 0; s <= d.Ly / 2; s++)
    for (int k = 1; k <= 4; k++) tab.push_back(twf(d.Ly, (long)((k * s) & (d.Ly - 1))));
  d.phi = (int)tab.size();
  const int M = d.ri * d.NX;
  // the strip kernel stores phi[0..npx/2] only and rebuilds phi[k] = conj(phi[npx - k])
  bool phi_sym = true;
  for (int k = d.NX / 2 + 1; k < d.NX && d.ri == 2; k++) {
    const float2 a = twf(M, k - d.NX), b = twf(M, d.NX - k);
    phi_sym = phi_sym && a.x == b.x && a.y == -b.y;
  }
  for (int r = 1; r < d.ri; r++)
    for (int k = 0; k < d.NX; k++) {
      const long kt = (k < d.NX / 2) ? k : k - d.NX;
      tab.push_back(twf(M, kt * r));
    }
  if (tab.empty()) tab.push_back(make_float2(0.f, 0.f));

  if (max_batch <= 0) {
    // scratch per location (both paths allocate at most about this much):
    // amplitudes / strip lines, spectra, lag rows, surface
    const size_t per = (size_t)d.N * 4 * (2 + 4) + (size_t)d.nyc * d.NX * 8 + (size_t)d.nxc * d.nyc * 8 + 256;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = (size_t)8 << 30;
    const size_t budget = std::min(free_b / 4, (size_t)24 << 30);
    max_batch = (int)std::max<size_t>(64, std::min<size_t>(8192, budget / per));
  }
  c->max_batch = max_batch;
  const size_t B = (size_t)max_batch;
  const size_t N = (size_t)d.N;
  const size_t ns = (size_t)d.nxc * d.nyc;
  hipError_t e = hipSuccess;
  auto alloc = [&](void** ptr, size_t bytes) {
    if (e == hipSuccess) e = hipMalloc(ptr, bytes ? bytes : 16);
  };
  // strips: square 128 / 256 / 512 windows in frequency mode, and -time /
  // -time4 on 128 x 128 and 256 x 256 windows (xs = ys = 32 or 64, the int8
  // searches); ri 2
  c->strips = d.NX == d.NY && d.ri == 2 && phi_sym &&
              ((p->corr_flag == XCORR_HIP_FREQ && (d.NX == 128 || d.NX == 256 || d.NX == 512)) ||
               time_i8_shape(p->corr_flag, p->xsearch, p->ysearch));
  if (c->strips) {
    // half the lines of unshared windows: overlapping windows need far fewer
    c->max_lines = std::max(max_batch * (d.NY / 2), d.NY + kStripBlockLines);
    c->max_lines = (c->max_lines + kStripBlockLines - 1) / kStripBlockLines * kStripBlockLines;
    const size_t L = (size_t)c->max_lines;
    alloc((void**)&c->amp1s, L * d.NX * sizeof(float));
    alloc((void**)&c->amp2s, L * d.NX * sizeof(float));
    alloc((void**)&c->lsum, L * 2 * sizeof(long long));
    alloc((void**)&c->need, L * 2);
    c->plan_bytes = plan_ints(c) * sizeof(int);
    alloc((void**)&c->d_plan, c->plan_bytes);
    for (int i = 0; i < 2; i++) {
      if (e == hipSuccess) e = hipHostMalloc(&c->h_stage[i], c->plan_bytes, hipHostMallocDefault);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_stage[i], hipEventDisableTiming);
    }
  } else {
    alloc((void**)&c->amp, B * 2 * N * sizeof(float));
  }
  alloc((void**)&c->sums, B * 2 * sizeof(unsigned long long));
  if (p->corr_flag == XCORR_HIP_FREQ) {
    c->z_elems = c->strips ? (size_t)c->max_lines * d.NX : B * N;
    alloc((void**)&c->Z, c->z_elems * sizeof(float2));
    alloc((void**)&c->D, B * d.nyc * d.NX * sizeof(float2));
    alloc((void**)&c->surf, B * ns * sizeof(float));
    // room for the generic row path's groups too (a tuned radix plan can select it for npx = 256)
    alloc((void**)&c->partial, B * std::max(d.groups4, d.nyc / d.G4) * sizeof(double));
  } else {
    alloc((void**)&c->surfd, B * ns * sizeof(double));
    if (time_i8_shape(p->corr_flag, p->xsearch, p->ysearch)) alloc((void**)&c->tcaa, B * ns * sizeof(double));
    alloc((void**)&c->tpeak, B * 4 * sizeof(int));
    alloc((void**)&c->tcave, B * sizeof(double));
  }
  alloc((void**)&c->peak, B * sizeof(unsigned long long));
  alloc((void**)&c->d_win, B * sizeof(xcorr_hip_window));
  alloc((void**)&c->d_out, B * sizeof(xcorr_hip_result));
  alloc((void**)&c->tabs, tab.size() * sizeof(float2));
  if (e == hipSuccess) e = hipMemcpy(c->tabs, tab.data(), tab.size() * sizeof(float2), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipEventCreate(&c->ev0);
  if (e == hipSuccess) e = hipEventCreate(&c->ev1);
  if (e != hipSuccess) {
    std::string m = std::string("device allocation failed: ") + hipGetErrorString(e);
    free_scratch(c);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return fail(nullptr, XCORR_HIP_ENOMEM, m);
  }
  *out = c;
  return 0;
}

static void free_stream(xcorr_hip_ctx* c) {
  if (c->cstream) (void)hipStreamSynchronize(c->cstream);
  for (int i = 0; i < 2; i++) {
    for (int b = 0; b < 2; b++)
      if (c->band[i][b]) (void)hipFree(c->band[i][b]);
    if (c->chunk[i]) (void)hipHostFree(c->chunk[i]);
    hipEvent_t* evs[3] = {&c->ev_chunk[i], &c->ev_loaded[i], &c->ev_done[i]};
    for (hipEvent_t* e : evs)
      if (*e) (void)hipEventDestroy(*e);
  }
  if (c->d_res) (void)hipFree(c->d_res);
  if (c->cstream) (void)hipStreamDestroy(c->cstream);
}

extern "C" void xcorr_hip_destroy(xcorr_hip_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->twin) xcorr_hip_destroy(c->twin);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->h_res) (void)hipHostFree(c->h_res);
  free_stream(c);
  free_scratch(c);
  if (c->own_master) (void)hipFree(c->own_master);
  if (c->own_slave) (void)hipFree(c->own_slave);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

extern "C" const char* xcorr_hip_last_error(const xcorr_hip_ctx* c) {
  return c ? c->err.c_str() : g_create_err.c_str();
}

extern "C" void* xcorr_hip_stream(xcorr_hip_ctx* c) { return c ? (void*)c->stream : nullptr; }

extern "C" int xcorr_hip_sync(xcorr_hip_ctx* c) {
  if (!c) return XCORR_HIP_EINVAL;
  HIPCK(c, hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int xcorr_hip_set_images(xcorr_hip_ctx* c, const void* master, int m_nx, int m_ny, const void* slave,
                                    int s_nx, int s_ny, unsigned flags) {
  if (!c) return XCORR_HIP_EINVAL;
  if (!master || !slave || m_nx <= 0 || m_ny <= 0 || s_nx <= 0 || s_ny <= 0)
    return fail(c, XCORR_HIP_EINVAL, "set_images: bad image arguments");
  HIPCK(c, hipSetDevice(c->device));
  if (c->own_master) { (void)hipFree(c->own_master); c->own_master = nullptr; }
  if (c->own_slave) { (void)hipFree(c->own_slave); c->own_slave = nullptr; }
  const size_t es = c->p.format == XCORR_HIP_CINT16 ? 4 : 4;
  if (flags & XCORR_HIP_DEVICE_PTRS) {
    c->d_master = master;
    c->d_slave = slave;
  } else {
    const size_t mb = (size_t)m_nx * m_ny * es, sb = (size_t)s_nx * s_ny * es;
    if (hipMalloc(&c->own_master, mb) != hipSuccess || hipMalloc(&c->own_slave, sb) != hipSuccess)
      return fail(c, XCORR_HIP_ENOMEM, "set_images: device allocation failed");
    HIPCK(c, hipMemcpyAsync(c->own_master, master, mb, hipMemcpyHostToDevice, c->stream));
    HIPCK(c, hipMemcpyAsync(c->own_slave, slave, sb, hipMemcpyHostToDevice, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    c->d_master = c->own_master;
    c->d_slave = c->own_slave;
  }
  c->dp.m_nx = m_nx; c->dp.m_ny = m_ny; c->dp.s_nx = s_nx; c->dp.s_ny = s_ny;
  c->dp.m_lo = 0; c->dp.m_hi = m_ny; c->dp.s_lo = 0; c->dp.s_hi = s_ny;
  c->images = true;
  return 0;
}

namespace {
struct Launch {
  xcorr_hip_ctx* c;
  bool timing;
  int slot = 0;
  hipEvent_t evs[16] = {};
  int nev = 0;
};
}  // namespace

static int timer_index(xcorr_hip_ctx* c, const char* name) {
  for (size_t i = 0; i < c->timers.size(); i++)
    if (!strcmp(c->timers[i].name, name)) return (int)i;
  c->timers.push_back({name, 0.0f, 0});
  return (int)c->timers.size() - 1;
}

// Launch helper: records events around the kernel when timing is on.
template <typename F>
static int launch_timed(xcorr_hip_ctx* c, bool timing, const char* name, F&& f) {
  if (!timing) {
    f();
    HIPCK(c, hipGetLastError());
    return 0;
  }
  HIPCK(c, hipEventRecord(c->ev0, c->stream));
  f();
  HIPCK(c, hipGetLastError());
  HIPCK(c, hipEventRecord(c->ev1, c->stream));
  HIPCK(c, hipEventSynchronize(c->ev1));
  float ms = 0.0f;
  HIPCK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
  const int i = timer_index(c, name);
  c->timers[i].ms += ms;
  c->timers[i].launches += 1;
  return 0;
}

static int run_batch(xcorr_hip_ctx* c, const PatchBufs& b, const xcorr_hip_window* d_win, int nb,
                     xcorr_hip_result* d_out, bool timing);
static int run_strips(xcorr_hip_ctx* c, const xcorr_hip_window* w, int n, xcorr_hip_result* out, bool dev_out,
                      bool timing, bool serial);

extern "C" int xcorr_hip_correlate(xcorr_hip_ctx* c, const xcorr_hip_window* win, int n, xcorr_hip_result* out,
                                   unsigned flags) {
  if (!c) return XCORR_HIP_EINVAL;
  if (n < 0 || (n > 0 && (!win || !out))) return fail(c, XCORR_HIP_EINVAL, "correlate: bad arguments");
  if (!c->images) return fail(c, XCORR_HIP_ESTATE, "correlate: no images set");
  HIPCK(c, hipSetDevice(c->device));
  c->dp.keep_surf = (flags & XCORR_HIP_KEEP_SURFACE) ? 1 : 0;
  const bool dev_out = (flags & XCORR_HIP_DEVICE_PTRS) != 0;
  const bool dev_win = dev_out && !(flags & XCORR_HIP_HOST_WINDOWS);
  const bool timing = (flags & XCORR_HIP_TIMING) != 0;
  if (timing)
    for (auto& t : c->timers) { t.ms = 0.0f; t.launches = 0; }
  if (use_strips(c)) {
    const xcorr_hip_window* hw = win;
    if (dev_win && n > 0) {     // the plan is built on the host
      c->hwin.resize(n);
      HIPCK(c, hipMemcpyAsync(c->hwin.data(), win, (size_t)n * sizeof(xcorr_hip_window), hipMemcpyDeviceToHost,
                              c->stream));
      HIPCK(c, hipStreamSynchronize(c->stream));
      hw = c->hwin.data();
    }
    const int rc = run_strips(c, hw, n, out, dev_out, timing, timing || (flags & XCORR_HIP_SERIAL) != 0);
    if (rc) return rc;
  } else {
    const int bcap = c->tune_batch > 0 ? std::min(c->tune_batch, c->max_batch) : c->max_batch;
    for (int off = 0; off < n; off += bcap) {
      const int nb = std::min(bcap, n - off);
      const xcorr_hip_window* dw = dev_win ? win + off : c->d_win;
      xcorr_hip_result* dout = dev_out ? out + off : c->d_out;
      if (!dev_win)
        HIPCK(c, hipMemcpyAsync(c->d_win, win + off, nb * sizeof(xcorr_hip_window), hipMemcpyHostToDevice,
                                c->stream));
      int rc = run_batch(c, ctx_bufs(c), dw, nb, dout, timing);
      if (rc) return rc;
      if (!dev_out) {
        HIPCK(c, hipMemcpyAsync(out + off, c->d_out, nb * sizeof(xcorr_hip_result), hipMemcpyDeviceToHost, c->stream));
        HIPCK(c, hipStreamSynchronize(c->stream));
      }
      c->last_perm.resize(nb);
      for (int i = 0; i < nb; i++) c->last_perm[i] = off + i;
      c->last_rb.clear();
    }
  }
  if (!(flags & XCORR_HIP_NO_SYNC)) HIPCK(c, hipStreamSynchronize(c->stream));
  return 0;
}

// Strip mode.  Windows are sorted by (master column, slave column, slave line
// offset, master line); a batch takes them in that order and merges each run
// of overlapping windows with one key into a strip.  Per batch the plan holds
// the row base and location index of every patch, the strip of every line
// block and the strips themselves (layout of plan_ints).
struct StripBatch {
  int nb, ns, nl, nblk, nblk_in;
};

// Sort key k of a window, least significant first: master line, slave line
// offset, slave column, master column.
static long long order_key(const xcorr_hip_window& q, int k) {
  switch (k) {
    case 0: return q.my0;
    case 1: return (long long)q.sy0 - q.my0;
    case 2: return q.sx0;
    default: return q.mx0;
  }
}

// Window indices in strip order: by (mx0, sx0, sy0 - my0, my0), equal keys in
// index order.  A least-significant-digit radix sort (16-bit digits of each
// key minus its minimum, a key whose values are all equal skipped) is stable,
// so it gives exactly that order; on the 40,000 windows of a 100 x 400 grid it
// takes a quarter of a comparison sort's time, which is host latency before the
// first batch can start.  The scratch vectors (OrderScratch) are the caller's,
// reused across calls.
static void strip_order(const xcorr_hip_window* w, int n, std::vector<int>& ord, OrderScratch& sc) {
  constexpr int kBits = 16, kDigits = 1 << kBits;
  ord.resize(n);
  for (int i = 0; i < n; i++) ord[i] = i;
  if (n < 2) return;
  sc.key.resize(n);
  sc.tmp.resize(n);
  sc.cnt.resize(kDigits);
  for (int k = 0; k < 4; k++) {
    long long lo = order_key(w[0], k), hi = lo;
    for (int i = 1; i < n; i++) {
      const long long v = order_key(w[i], k);
      lo = std::min(lo, v);
      hi = std::max(hi, v);
    }
    for (int i = 0; i < n; i++) sc.key[i] = (unsigned long long)(order_key(w[i], k) - lo);
    for (int sh = 0; ((unsigned long long)(hi - lo) >> sh) != 0; sh += kBits) {
      std::fill(sc.cnt.begin(), sc.cnt.end(), 0);
      for (int i = 0; i < n; i++) sc.cnt[(sc.key[ord[i]] >> sh) & (kDigits - 1)]++;
      for (int d = 0, acc = 0; d < kDigits; d++) {
        const int t = sc.cnt[d];
        sc.cnt[d] = acc;
        acc += t;
      }
      for (int i = 0; i < n; i++) sc.tmp[sc.cnt[(sc.key[ord[i]] >> sh) & (kDigits - 1)]++] = ord[i];
      ord.swap(sc.tmp);
    }
  }
}

// Line blocks bs[lo .. hi) in order of their first master line, then of their
// strip's master column: blocks of neighbouring strips that read the same
// image lines come next to each other, and with the kernels' xcd_item order
// run on one XCD, so those lines are read from HBM once into its L2.
static void order_line_blocks(const StripDesc* st, int* bs, int lo, int hi) {
  struct E {
    long long row;
    int mx0, k, l0;
  };
  std::vector<E> v;
  v.reserve(hi - lo);
  for (int e = lo; e < hi; e++) {
    const StripDesc& S = st[bs[2 * e]];
    v.push_back({(long long)S.my0 + (bs[2 * e + 1] - S.line0), S.mx0, bs[2 * e], bs[2 * e + 1]});
  }
  std::stable_sort(v.begin(), v.end(),
                   [](const E& a, const E& b) { return a.row != b.row ? a.row < b.row : a.mx0 < b.mx0; });
  for (int e = lo; e < hi; e++) {
    bs[2 * e] = v[e - lo].k;
    bs[2 * e + 1] = v[e - lo].l0;
  }
}

// Plans the next batch (at most `cap` windows) from ord[*pos] on into h
// (plan_ints(c) ints); advances *pos.  Returns false if a window does not fit.
static bool plan_strip_batch(const xcorr_hip_ctx* c, const DevParams& d, const xcorr_hip_window* w,
                             const std::vector<int>& ord, int* pos, int n, int cap, int* h, StripBatch* sb) {
  const int NY = d.NY, mb = c->max_batch, ML = c->max_lines, blkcap = 2 * (ML / kStripBlockLines);
  const long long pad_ny = round_up(NY, kStripBlockLines);
  int* rb = h;
  int* perm = h + mb;
  int* bs = h + 2 * mb;
  StripDesc* st = (StripDesc*)(h + 2 * mb + blkcap);
  int nb = 0, ns = 0;
  long long L = 0, cur_end = 0;
  while (*pos < n && nb < cap) {
    const xcorr_hip_window& q = w[ord[*pos]];
    bool ext = false;
    if (ns > 0) {
      const StripDesc& S = st[ns - 1];
      ext = S.mx0 == q.mx0 && S.sx0 == q.sx0 && (long long)S.sy0 - S.my0 == (long long)q.sy0 - q.my0 &&
            (long long)q.my0 <= cur_end;
    }
    if (ext) {
      StripDesc& S = st[ns - 1];
      const long long end = std::max(cur_end, (long long)q.my0 + NY);
      const long long lines = S.line0 + round_up(end - S.my0, kStripBlockLines);
      if (lines > ML) break;
      cur_end = end;
      S.nlines = (int)(end - S.my0);
      L = lines;
      rb[nb] = S.line0 + (q.my0 - S.my0);
    } else {
      if (L + pad_ny > ML) break;
      StripDesc& S = st[ns++];
      S.mx0 = q.mx0; S.my0 = q.my0; S.sx0 = q.sx0; S.sy0 = q.sy0;
      S.line0 = (int)L; S.nlines = NY; S.pad0 = S.pad1 = 0;
      cur_end = (long long)q.my0 + NY;
      rb[nb] = (int)L;
      L += pad_ny;
    }
    perm[nb] = ord[*pos];
    nb++;
    (*pos)++;
  }
  if (nb == 0) return false;
  // line blocks: strips whose windows lie inside both images first, then the rest
  int nblk = 0, nblk_in = 0;
  for (int pass = 0; pass < 2; pass++) {
    for (int k = 0; k < ns; k++) {
      const StripDesc& S = st[k];
      const bool in = S.mx0 >= 0 && S.mx0 + d.NX <= d.m_nx && S.sx0 >= 0 && S.sx0 + d.NX <= d.s_nx;
      if (in != (pass == 0)) continue;
      for (int l0 = S.line0; l0 < S.line0 + S.nlines; l0 += kStripBlockLines) {
        bs[2 * nblk] = k;
        bs[2 * nblk + 1] = l0;
        nblk++;
      }
    }
    if (pass == 0) nblk_in = nblk;
  }
  order_line_blocks(st, bs, 0, nblk_in);
  order_line_blocks(st, bs, nblk_in, nblk);
  sb->nb = nb;
  sb->ns = ns;
  sb->nl = (int)L;
  sb->nblk = nblk;
  sb->nblk_in = nblk_in;
  return true;
}

static size_t strip_plan_bytes(const xcorr_hip_ctx* c, const StripBatch& sb) {
  return (size_t)(2 * c->max_batch + 2 * (c->max_lines / kStripBlockLines)) * sizeof(int) +
         (size_t)sb.ns * sizeof(StripDesc);
}

// k_colpass<N> over nb patches (rows of Z per patch, or strip lines from the
// row bases rb) on stream st.
template <int N>
static void launch_colpass(xcorr_hip_ctx* c, const PatchBufs& b, const DevParams& d, hipStream_t st, const int* rb,
                           int nb) {
  const int nwork = nb * (N / (2 * ColCfg<N>::TC));
  hipLaunchKernelGGL(k_colpass<N>, dim3((nwork + 7) & ~7), dim3(kThreads), 0, st, d, (const float2*)b.Z,
                     rb, (const unsigned long long*)b.sums, b.D, (const float2*)c->tabs, nwork);
}

// The fused -time search of 128 x 128 windows leaves only the peak's
// neighbourhood in the surface slot (unless the surface is kept).
static bool time_nbr(const DevParams& d, const PatchBufs& b) {
  return d.corr_flag == XCORR_HIP_TIME && d.xs == 32 && d.ys == 32 && b.tcaa;
}

// The -time / -time4 search of 128 x 128 and 256 x 256 windows (xs = ys = 32
// or 64, time_i8_shape) on the int8 matrix cores: sliding energies, then the
// search (-time on 128 x 128 windows with its peak scan), then, for the
// others, the peak scan.  Amplitudes per patch (rb null: A1 = A2 = amp) or on
// strip lines from row bases rb.
static void launch_time_i8(const PatchBufs& b, const DevParams& d, hipStream_t st, const float* A1, const float* A2,
                           const int* rb, int nb) {
  const dim3 g1((nb + 7) & ~7), g4((4 * nb + 7) & ~7), blk(kThreads);
  const unsigned long long* sums = b.sums;
  if (d.corr_flag == XCORR_HIP_TIME && d.xs == 32) {
    hipLaunchKernelGGL(k_time_energy32, g1, blk, 0, st, d, A1, rb, sums, b.tcaa, nb);
    hipLaunchKernelGGL(k_time_search_i8, g1, blk, 0, st, d, A1, A2, rb, sums, (const double*)b.tcaa, b.surfd,
                       b.tpeak, b.tcave, nb);
    return;
  }
  if (d.corr_flag == XCORR_HIP_TIME) {
    hipLaunchKernelGGL((k_time_energy_rows<2, 256>), g1, blk, 0, st, d, A1, rb, sums, b.tcaa, (double*)nullptr, nb);
    hipLaunchKernelGGL(k_time_search_i8q, g4, blk, 0, st, d, A1, A2, rb, sums, (const double*)b.tcaa, b.surfd, nb);
  } else if (d.xs == 32) {
    hipLaunchKernelGGL((k_time_energy_rows<4, 128>), g1, blk, 0, st, d, A1, rb, sums, b.tcaa, b.surfd, nb);
    hipLaunchKernelGGL(k_time4_search_i8<32>, g1, blk, 0, st, d, A1, A2, rb, sums, (const double*)b.tcaa, b.surfd,
                       nb);
  } else {
    hipLaunchKernelGGL((k_time_energy_rows<4, 256>), g1, blk, 0, st, d, A1, rb, sums, b.tcaa, b.surfd, nb);
    hipLaunchKernelGGL(k_time4_search_i8<64>, g4, blk, 0, st, d, A1, A2, rb, sums, (const double*)b.tcaa, b.surfd,
                       nb);
  }
  hipLaunchKernelGGL(k_time_peak4k, dim3(nb), blk, 0, st, d, (const double*)b.surfd, b.tpeak, b.tcave);
}

// Kernels of one strip batch whose plan is in device memory at d_plan, on
// stream st with images (master, slave) and image dims in d.  Results go to
// out[perm[b]] (dev_out) or to c->d_out[b].
static int launch_strip_batch(xcorr_hip_ctx* c, const DevParams& d, const void* master, const void* slave,
                              hipStream_t st, const int* d_plan, const StripBatch& sb, xcorr_hip_result* out,
                              bool dev_out, bool timing) {
  const int mb = c->max_batch, ML = c->max_lines, blkcap = 2 * (ML / kStripBlockLines);
  const int nb = sb.nb, nl = sb.nl, nblk = sb.nblk, nblk_in = sb.nblk_in;
  const int* d_rb = d_plan;
  const int* d_perm = d_plan + mb;
  const int* d_bs = d_plan + 2 * mb;
  const StripDesc* d_st = (const StripDesc*)(d_plan + 2 * mb + blkcap);
  unsigned char* needM = c->need;
  unsigned char* needU = c->need + ML;
  int rc = launch_timed(c, timing, "k_need", [&] {
    hipLaunchKernelGGL(k_need_clear, dim3((nl + 255) / 256), dim3(256), 0, st, needM, needU, nl);
    hipLaunchKernelGGL(k_need_mark, dim3(nb), dim3(256), 0, st, d, d_rb, needM, needU, nb);
  });
  if (rc) return rc;
  const bool ci = d.format == XCORR_HIP_CINT16;
  const bool freq = d.corr_flag == XCORR_HIP_FREQ;    // else -time / -time4 on 128 x 128 or 256 x 256 windows
  rc = launch_timed(c, timing, d.NX == 256 ? "k_stripA256" : "k_stripA", [&] {
    using KA = void (*)(DevParams, const void*, const void*, const StripDesc*, const int*, int, int,
                        const unsigned char*, const unsigned char*, float*, float*, long long*, float2*,
                        const float2*);
    KA kin, kedge;
    if (d.NX == 256 && freq) {
      kin = ci ? k_stripA256<XCORR_HIP_CINT16, false> : k_stripA256<XCORR_HIP_FLOAT32, false>;
      kedge = ci ? k_stripA256<XCORR_HIP_CINT16, true> : k_stripA256<XCORR_HIP_FLOAT32, true>;
    } else if (d.NX == 256) {
      kin = ci ? k_stripA256<XCORR_HIP_CINT16, false, false> : k_stripA256<XCORR_HIP_FLOAT32, false, false>;
      kedge = ci ? k_stripA256<XCORR_HIP_CINT16, true, false> : k_stripA256<XCORR_HIP_FLOAT32, true, false>;
    } else if (d.NX == 128 && !freq) {
      kin = ci ? k_stripA<128, XCORR_HIP_CINT16, false, false> : k_stripA<128, XCORR_HIP_FLOAT32, false, false>;
      kedge = ci ? k_stripA<128, XCORR_HIP_CINT16, true, false> : k_stripA<128, XCORR_HIP_FLOAT32, true, false>;
    } else if (d.NX == 128) {
      kin = ci ? k_stripA<128, XCORR_HIP_CINT16, false> : k_stripA<128, XCORR_HIP_FLOAT32, false>;
      kedge = ci ? k_stripA<128, XCORR_HIP_CINT16, true> : k_stripA<128, XCORR_HIP_FLOAT32, true>;
    } else {
      kin = ci ? k_stripA<512, XCORR_HIP_CINT16, false> : k_stripA<512, XCORR_HIP_FLOAT32, false>;
      kedge = ci ? k_stripA<512, XCORR_HIP_CINT16, true> : k_stripA<512, XCORR_HIP_FLOAT32, true>;
    }
    if (nblk_in > 0)
      hipLaunchKernelGGL(kin, dim3((nblk_in + 7) & ~7), dim3(kThreads), 0, st, d, master, slave, d_st, d_bs, 0,
                         nblk_in, (const unsigned char*)needM, (const unsigned char*)needU, c->amp1s, c->amp2s,
                         c->lsum, c->Z, (const float2*)c->tabs);
    if (nblk > nblk_in)
      hipLaunchKernelGGL(kedge, dim3((nblk - nblk_in + 7) & ~7), dim3(kThreads), 0, st, d, master, slave, d_st, d_bs,
                         nblk_in, nblk - nblk_in, (const unsigned char*)needM, (const unsigned char*)needU, c->amp1s,
                         c->amp2s, c->lsum, c->Z, (const float2*)c->tabs);
  });
  if (rc) return rc;
  rc = launch_timed(c, timing, "k_patch_sums", [&] {
    hipLaunchKernelGGL(k_patch_sums, dim3((nb + 3) / 4), dim3(256), 0, st, d, d_rb, (const long long*)c->lsum,
                       c->sums, nb);
  });
  if (rc) return rc;
  if (!freq) {
    rc = launch_timed(c, timing, d.corr_flag == XCORR_HIP_TIME4 ? "k_time_search4" : "k_time_search", [&] {
      launch_time_i8(ctx_bufs(c), d, st, (const float*)c->amp1s, (const float*)c->amp2s, d_rb, nb);
    });
  } else {
  rc = launch_timed(c, timing, d.NX == 256 ? "k_colpass256" : "k_colpass", [&] {
    if (d.NX == 256)
      hipLaunchKernelGGL(k_colpass256<true>, dim3((nb * 8 + 7) & ~7), dim3(kThreads), 0, st, d, (const float2*)c->Z, d_rb,
                         (const unsigned long long*)c->sums, c->D, (const float2*)c->tabs, nb * 8);
    else if (d.NX == 128)
      launch_colpass<128>(c, ctx_bufs(c), d, st, d_rb, nb);
    else
      launch_colpass<512>(c, ctx_bufs(c), d, st, d_rb, nb);
  });
  if (rc) return rc;
  rc = launch_timed(c, timing, d.NX == 256 ? "k_rowifft_surf256" : "k_rowifft_surf_reg", [&] {
    if (d.NX == 256)
      hipLaunchKernelGGL(k_rowifft_surf256<true>, dim3(nb), dim3(kThreads), 0, st, d, (const float2*)c->D, c->surf,
                         c->peak, c->partial, (const float2*)c->tabs, d.keep_surf);
    else if (d.NX == 128)
      hipLaunchKernelGGL(k_rowifft_surf_reg<128>, dim3(nb), dim3(kThreads), 0, st, d, (const float2*)c->D, c->surf,
                         c->peak, c->partial, (const float2*)c->tabs, d.keep_surf);
    else
      hipLaunchKernelGGL(k_rowifft_surf_reg<512>, dim3(nb), dim3(kThreads), 0, st, d, (const float2*)c->D, c->surf,
                         c->peak, c->partial, (const float2*)c->tabs, d.keep_surf);
  });
  }
  if (rc) return rc;
  const size_t sh5 = d.interp_flag ? (size_t)d.Lx * 16 * sizeof(float2) : 16;
  return launch_timed(c, timing, "k_finish", [&] {
    hipLaunchKernelGGL(k_finish, dim3((nb + 7) & ~7), dim3(kThreads), sh5, st, d, (const float*)c->amp1s,
                       (const float*)c->amp2s, d_rb, 0LL, dev_out ? d_perm : (const int*)nullptr,
                       (const unsigned long long*)c->sums, (const float*)c->surf, (const double*)c->surfd,
                       (const unsigned long long*)c->peak, (const double*)c->partial, (const int*)c->tpeak,
                       (const double*)c->tcave, dev_out ? out : c->d_out, (const float2*)c->tabs, nb,
                       (int)((freq || time_nbr(d, ctx_bufs(c))) && !d.keep_surf));
  });
}

// The second batch context of a strip context (same parameters and batch
// size: its own line buffers, spectra, surfaces, plan buffer and stream; the
// images and tables it uses are this context's).  Without the memory for it
// the strips run on one context.
static xcorr_hip_ctx* ensure_twin(xcorr_hip_ctx* c) {
  if (c->twin || c->twin_failed) return c->twin;
  xcorr_hip_ctx* t = nullptr;
  if (xcorr_hip_create(&c->p, c->max_batch, c->device, &t) != 0) {
    c->twin_failed = true;
    g_create_err.clear();
  }
  (void)hipSetDevice(c->device);
  c->twin = t;
  return t;
}

// Strip batches in plan order.  Unless serial (kernel timing, whose events
// need one stream, or XCORR_HIP_SERIAL), batch k runs on context k mod 2 -- this one or its twin, each
// on its own stream -- so one batch's kernels overlap the next one's.  The
// twin's stream first waits for the work queued on this context's stream
// before the call, and this stream waits for the twin's at the end, so callers
// see one ordered stream.  Host results of a batch are scattered once the
// next batch is queued (or before its context is used again).  *twin is set
// once the twin's stream has forked from this one; run_strips joins it on
// every exit.
static int run_strips_body(xcorr_hip_ctx* c, const xcorr_hip_window* w, int n, xcorr_hip_result* out, bool dev_out,
                           bool timing, bool serial, xcorr_hip_ctx** twin) {
  strip_order(w, n, c->order, c->order_scratch);
  int cap = std::min(c->max_batch, c->tune_batch > 0 ? c->tune_batch : c->max_batch);
  // the two contexts take alternate batches: an even number of equal batches
  // keeps them busy to the end (40,000 windows at 8192: six of 6667, not five
  // with the last alone, 1.5 % faster)
  if (!serial && !c->one_context && n > cap) {
    int nbat = (n + cap - 1) / cap;
    nbat += nbat & 1;
    cap = (n + nbat - 1) / nbat;
  }
  if (!serial) {
    if (!c->ev_fork) HIPCK(c, hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    if (!c->ev_join) HIPCK(c, hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
    HIPCK(c, hipEventRecord(c->ev_fork, c->stream));
  }
  xcorr_hip_ctx* x2 = nullptr;                   // the twin, once this call has used it
  xcorr_hip_ctx* pend = nullptr;                 // context whose batch's host results are not scattered yet
  std::vector<int> pend_perm;
  auto scatter = [&]() -> int {
    if (!pend) return 0;
    HIPCK(c, hipStreamSynchronize(pend->stream));
    for (size_t i = 0; i < pend_perm.size(); i++) out[pend_perm[i]] = pend->h_res[i];
    pend = nullptr;
    return 0;
  };
  for (int pos = 0, k = 0; pos < n; k++) {
    xcorr_hip_ctx* x = c;
    if (!serial && !c->one_context && (k & 1)) {
      if (!x2) {
        x2 = ensure_twin(c);
        if (x2) {
          HIPCK(c, hipStreamWaitEvent(x2->stream, c->ev_fork, 0));
          *twin = x2;
        }
      }
      if (x2) x = x2;
    }
    const int si = c->stage_next;
    c->stage_next ^= 1;
    if (c->stage_used[si]) HIPCK(c, hipEventSynchronize(c->ev_stage[si]));
    int* h = (int*)c->h_stage[si];
    StripBatch sb;
    if (!plan_strip_batch(c, c->dp, w, c->order, &pos, n, cap, h, &sb))
      return fail(c, XCORR_HIP_EINVAL, "correlate: a window does not fit the line capacity");
    HIPCK(c, hipMemcpyAsync(x->d_plan, h, strip_plan_bytes(c, sb), hipMemcpyHostToDevice, x->stream));
    HIPCK(c, hipEventRecord(c->ev_stage[si], x->stream));
    c->stage_used[si] = true;
    c->last_perm.assign(h + c->max_batch, h + c->max_batch + sb.nb);
    c->last_rb.assign(h, h + sb.nb);
    c->last_buf = x;
    const int rc = launch_strip_batch(x, c->dp, c->d_master, c->d_slave, x->stream, x->d_plan, sb, out, dev_out,
                                      timing);
    if (rc) {
      if (x != c) c->err = x->err;
      return rc;
    }
    if (!dev_out) {
      if (pend == x) {                           // one context: its pinned results are about to be reused
        const int rs = scatter();
        if (rs) return rs;
      }
      if (!x->h_res)
        HIPCK(c, hipHostMalloc((void**)&x->h_res, (size_t)c->max_batch * sizeof(xcorr_hip_result),
                               hipHostMallocDefault));
      HIPCK(c, hipMemcpyAsync(x->h_res, x->d_out, sb.nb * sizeof(xcorr_hip_result), hipMemcpyDeviceToHost,
                              x->stream));
      const int rs = scatter();
      if (rs) return rs;
      pend = x;
      pend_perm = c->last_perm;
    }
  }
  return scatter();
}

// The stream of this context waits for the twin's work on every exit, so
// xcorr_hip_sync and the caller's stream order cover every batch that was
// queued; after a failure the twin's queued kernels (which still write the
// caller's results and read the images) are also waited for before returning.
static int run_strips(xcorr_hip_ctx* c, const xcorr_hip_window* w, int n, xcorr_hip_result* out, bool dev_out,
                      bool timing, bool serial) {
  xcorr_hip_ctx* x2 = nullptr;
  const int rc = run_strips_body(c, w, n, out, dev_out, timing, serial, &x2);
  if (!x2) return rc;
  if (rc) {
    (void)hipStreamSynchronize(x2->stream);
    if (hipEventRecord(c->ev_join, x2->stream) == hipSuccess) (void)hipStreamWaitEvent(c->stream, c->ev_join, 0);
    return rc;
  }
  HIPCK(c, hipEventRecord(c->ev_join, x2->stream));
  HIPCK(c, hipStreamWaitEvent(c->stream, c->ev_join, 0));
  return 0;
}

static int time_search(xcorr_hip_ctx* c, const PatchBufs& b, int nb, bool timing);

static int run_batch(xcorr_hip_ctx* c, const PatchBufs& b, const xcorr_hip_window* d_win, int nb,
                     xcorr_hip_result* d_out, bool timing) {
  const DevParams& d = c->dp;
  const int rc0 = launch_timed(c, timing, "k_init", [&] {
    hipLaunchKernelGGL(k_init, dim3((2 * nb + 255) / 256), dim3(256), 0, c->stream, b.sums, b.peak, nb);
  });
  if (rc0) return rc0;
  const size_t nel1 = (size_t)d.G1 * d.NX;
  const size_t sh1 = nel1 * sizeof(float2) * (d.ri > 2 ? 3 : 2) + nel1 * sizeof(float);
  // the register-resident 256-point kernels run the default radix-16 x 16 plan
  const bool row256 = d.NX == 256 && same_plan(d.plx, plan_for(256));
  const bool col256 = d.NY == 256 && same_plan(d.ply, plan_for(256));
  const bool freq = d.corr_flag == XCORR_HIP_FREQ;
  const bool fusedA = row256 && d.ri == 2 && freq;
  // register-resident stage (a) + row transform for square 128 / 512 windows
  const bool regA = reg_shape(c);
  // time modes on square 128 / 512 windows: the same register-resident stage A without the row transform
  const bool regT = !freq && c->reg_paths && d.ri == 2 && d.NX == d.NY && (d.NX == 128 || d.NX == 512) &&
                    same_plan(d.plx, plan_for(d.NX));
  int rc;
  if (regA)
    rc = launch_timed(c, timing, "k_stageA", [&] {
      if (d.NX == 128)
        hipLaunchKernelGGL((k_stageA<128, true>), dim3(nb), dim3(kThreads), 0, c->stream, d, c->d_master, c->d_slave,
                           d_win, b.amp, b.sums, b.Z, (const float2*)c->tabs);
      else
        hipLaunchKernelGGL((k_stageA<512, true>), dim3(nb), dim3(kThreads), 0, c->stream, d, c->d_master, c->d_slave,
                           d_win, b.amp, b.sums, b.Z, (const float2*)c->tabs);
    });
  else if (regT)
    rc = launch_timed(c, timing, "k_stageA", [&] {
      if (d.NX == 128)
        hipLaunchKernelGGL((k_stageA<128, false>), dim3(nb), dim3(kThreads), 0, c->stream, d, c->d_master,
                           c->d_slave, d_win, b.amp, b.sums, (float2*)nullptr, (const float2*)c->tabs);
      else
        hipLaunchKernelGGL((k_stageA<512, false>), dim3(nb), dim3(kThreads), 0, c->stream, d, c->d_master,
                           c->d_slave, d_win, b.amp, b.sums, (float2*)nullptr, (const float2*)c->tabs);
    });
  else if (fusedA)
    rc = launch_timed(c, timing, "k_stageA256", [&] {
      hipLaunchKernelGGL(k_stageA256, dim3(nb), dim3(kThreads), 0, c->stream, d, c->d_master, c->d_slave, d_win,
                         b.amp, b.sums, b.Z, (const float2*)c->tabs);
    });
  else if (row256 && d.ri == 2)
    rc = launch_timed(c, timing, "k_range256", [&] {
      hipLaunchKernelGGL(k_range256, dim3(nb * 2), dim3(kThreads), 0, c->stream, d, c->d_master, c->d_slave, d_win,
                         b.amp, b.sums, (const float2*)c->tabs);
    });
  else
    rc = launch_timed(c, timing, "k_range_ampl", [&] {
      hipLaunchKernelGGL(k_range_ampl, dim3(nb * 2 * (d.NY / d.G1)), dim3(c->nt_rows), sh1, c->stream, d, c->d_master,
                         c->d_slave, d_win, b.amp, b.sums, (const float2*)c->tabs);
    });
  if (rc) return rc;
  if (freq) {
    if (fusedA || regA)
      rc = 0;
    else if (row256)
      rc = launch_timed(c, timing, "k_rowfft256", [&] {
        hipLaunchKernelGGL(k_rowfft256, dim3(nb), dim3(kThreads), 0, c->stream, d, (const float*)b.amp,
                           (const unsigned long long*)b.sums, b.Z, (const float2*)c->tabs);
      });
    else
      rc = launch_timed(c, timing, "k_rowfft", [&] {
        hipLaunchKernelGGL(k_rowfft, dim3(nb * (d.NY / d.G1)), dim3(c->nt_rows), nel1 * sizeof(float2) * 2, c->stream,
                           d, (const float*)b.amp, (const unsigned long long*)b.sums, b.Z, (const float2*)c->tabs);
      });
    if (rc) return rc;
    const size_t nel3 = (size_t)d.C3 * d.NY;
    if (regA) {
      rc = launch_timed(c, timing, "k_colpass", [&] {
        if (d.NX == 128)
          launch_colpass<128>(c, b, d, c->stream, nullptr, nb);
        else
          launch_colpass<512>(c, b, d, c->stream, nullptr, nb);
      });
    } else if (col256 && d.NX >= 32) {
      rc = launch_timed(c, timing, "k_colpass256", [&] {
        hipLaunchKernelGGL(k_colpass256<false>, dim3((nb * (d.NX / 32) + 7) & ~7), dim3(kThreads), 0, c->stream, d,
                           (const float2*)b.Z, (const int*)nullptr,
                           (const unsigned long long*)b.sums, b.D, (const float2*)c->tabs, nb * (d.NX / 32));
      });
    } else {
      rc = launch_timed(c, timing, "k_colfft", [&] {
        hipLaunchKernelGGL(k_colfft, dim3(nb * (d.NX / d.C3)), dim3(c->nt_rows), nel3 * sizeof(float2) * 2,
                           c->stream, d, b.Z, (const unsigned long long*)b.sums, (const float2*)c->tabs);
      });
      if (rc) return rc;
      rc = launch_timed(c, timing, "k_product_colifft", [&] {
        hipLaunchKernelGGL(k_product_colifft, dim3(nb * (d.NX / d.C3)), dim3(c->nt_rows), nel3 * sizeof(float2) * 3,
                           c->stream, d, (const float2*)b.Z, b.D, (const float2*)c->tabs);
      });
    }
    if (rc) return rc;
    const size_t nel4 = (size_t)d.G4 * d.NX;
    if (regA)
      rc = launch_timed(c, timing, "k_rowifft_surf_reg", [&] {
        if (d.NX == 128)
          hipLaunchKernelGGL(k_rowifft_surf_reg<128>, dim3(nb), dim3(kThreads), 0, c->stream, d, (const float2*)b.D,
                             b.surf, b.peak, b.partial, (const float2*)c->tabs, d.keep_surf);
        else
          hipLaunchKernelGGL(k_rowifft_surf_reg<512>, dim3(nb), dim3(kThreads), 0, c->stream, d, (const float2*)b.D,
                             b.surf, b.peak, b.partial, (const float2*)c->tabs, d.keep_surf);
      });
    else if (row256)
      rc = launch_timed(c, timing, "k_rowifft_surf256", [&] {
        hipLaunchKernelGGL(k_rowifft_surf256<false>, dim3(nb), dim3(kThreads), 0, c->stream, d, (const float2*)b.D,
                           b.surf, b.peak, b.partial, (const float2*)c->tabs, d.keep_surf);
      });
    else
      rc = launch_timed(c, timing, "k_rowifft_surf", [&] {
        hipLaunchKernelGGL(k_rowifft_surf, dim3(nb * d.groups4), dim3(kThreads), nel4 * sizeof(float2) * 2,
                           c->stream, d, (const float2*)b.D, b.surf, b.peak, b.partial, (const float2*)c->tabs);
      });
    if (rc) return rc;
  } else {
    rc = time_search(c, b, nb, timing);
    if (rc) return rc;
  }
  const size_t sh5 = d.interp_flag ? (size_t)d.Lx * 16 * sizeof(float2) : 16;
  rc = launch_timed(c, timing, "k_finish", [&] {
    hipLaunchKernelGGL(k_finish, dim3((nb + 7) & ~7), dim3(kThreads), sh5, c->stream, d, (const float*)b.amp,
                       (const float*)b.amp + d.N, (const int*)nullptr, (long long)2 * d.N, (const int*)nullptr,
                       (const unsigned long long*)b.sums, (const float*)b.surf, (const double*)b.surfd,
                       (const unsigned long long*)b.peak, (const double*)b.partial, (const int*)b.tpeak,
                       (const double*)b.tcave, d_out, (const float2*)c->tabs, nb,
                       (int)(((freq && (regA || row256)) || time_nbr(d, b)) && !d.keep_surf));
  });
  return rc;
}

static int time_search(xcorr_hip_ctx* c, const PatchBufs& b, int nb, bool timing) {
  const DevParams& d = c->dp;
  const size_t sh = ((size_t)(d.tRC + d.tLR - 1) * d.NX + (size_t)d.tRC * d.nx_corr) * sizeof(int);
  const dim3 grid(nb * (d.nyc / d.tLR));
  int rc;
  // blocked search: LR surface rows x nxc lags per block, 4 lags per thread; RC
  // template rows per LDS chunk (doubles, within 64 KB)
  constexpr int TL = 8;                           // lags per thread
  const int LRb = std::max(1, std::min(d.nyc, kThreads * TL / d.nxc));
  const int RCb = std::min(8, (5120 - (LRb - 1) * d.NX) / (d.NX + d.nx_corr));   // <= 40 KB of LDS
  // 128 x 128 and 256 x 256 windows: the matrix-core searches, peak scan included
  if (time_i8_shape(d.corr_flag, d.xs, d.ys) && b.tcaa)
    return launch_timed(c, timing, d.corr_flag == XCORR_HIP_TIME4 ? "k_time_search4" : "k_time_search", [&] {
      launch_time_i8(b, d, c->stream, (const float*)b.amp, (const float*)b.amp, nullptr, nb);
    });
  if (d.corr_flag == XCORR_HIP_TIME && d.nxc % TL == 0 && RCb >= 1)
    rc = launch_timed(c, timing, "k_time_search", [&] {
      const size_t shb = ((size_t)(RCb + LRb - 1) * d.NX + (size_t)RCb * d.nx_corr) * sizeof(double);
      hipLaunchKernelGGL(k_time_search_blk<TL>, dim3(nb * (d.nyc / LRb)), dim3(kThreads), shb, c->stream, d,
                         (const float*)b.amp, (const unsigned long long*)b.sums, b.surfd, LRb, RCb);
    });
  else if (d.corr_flag == XCORR_HIP_TIME)
    rc = launch_timed(c, timing, "k_time_search", [&] {
      hipLaunchKernelGGL(k_time_search<0>, grid, dim3(kThreads), sh, c->stream, d, (const float*)b.amp,
                         (const unsigned long long*)b.sums, b.surfd);
    });
  else
    rc = launch_timed(c, timing, "k_time_search4", [&] {
      hipLaunchKernelGGL(k_time_search<1>, grid, dim3(kThreads), sh, c->stream, d, (const float*)b.amp,
                         (const unsigned long long*)b.sums, b.surfd);
    });
  if (rc) return rc;
  return launch_timed(c, timing, "k_time_peak", [&] {
    if ((d.nxc * d.nyc) % 4096 == 0)
      hipLaunchKernelGGL(k_time_peak4k, dim3(nb), dim3(kThreads), 0, c->stream, d, (const double*)b.surfd, b.tpeak,
                         b.tcave);
    else
      hipLaunchKernelGGL(k_time_peak, dim3(nb), dim3(kThreads), 0, c->stream, d, (const double*)b.surfd, b.tpeak,
                         b.tcave);
  });
}

// --------------------------------------------------------------------------
// Streamed images: windows are grouped into bands of lines; the lines of band
// k + 1 are read (on this thread, through the caller's reader) and copied on
// a second stream while band k is correlated.
// --------------------------------------------------------------------------
extern "C" int xcorr_hip_correlate_stream(xcorr_hip_ctx* c, xcorr_hip_line_reader reader, void* user, int m_nx,
                                          int m_ny, int s_nx, int s_ny, int band_lines,
                                          const xcorr_hip_window* win, int n, xcorr_hip_result* out) {
  if (!c) return XCORR_HIP_EINVAL;
  if (!reader || m_nx <= 0 || m_ny <= 0 || s_nx <= 0 || s_ny <= 0 || n < 0 || (n > 0 && (!win || !out)))
    return fail(c, XCORR_HIP_EINVAL, "correlate_stream: bad arguments");
  if (n == 0) return 0;
  HIPCK(c, hipSetDevice(c->device));
  const int NY = c->dp.NY;
  long long C = band_lines;
  if (C <= 0) {                                   // about 1.5 GB of band buffers per image
    const long long per_line = 4LL * std::max(m_nx, s_nx);
    C = std::max<long long>(2LL * NY, std::min<long long>(16384, ((long long)768 << 20) / per_line));
  }
  // lines each window needs (master and slave lines that exist)
  std::vector<int> lo(n), hi(n), ord(n);
  for (int i = 0; i < n; i++) {
    long long a = LLONG_MAX, b = LLONG_MIN;
    const long long ma = std::max<long long>(win[i].my0, 0), mb = std::min<long long>((long long)win[i].my0 + NY, m_ny);
    const long long sa = std::max<long long>(win[i].sy0, 0), sb = std::min<long long>((long long)win[i].sy0 + NY, s_ny);
    if (ma < mb) { a = std::min(a, ma); b = std::max(b, mb); }
    if (sa < sb) { a = std::min(a, sa); b = std::max(b, sb); }
    if (a > b) a = b = 0;                         // no lines in either image
    if (b - a > C) return fail(c, XCORR_HIP_EINVAL, "correlate_stream: a window spans more lines than a band");
    lo[i] = (int)a;
    hi[i] = (int)b;
    ord[i] = i;
  }
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return lo[x] < lo[y]; });
  struct Band { int b0, b1, begin, end; };
  std::vector<Band> bands;
  for (int i = 0; i < n;) {
    Band B{lo[ord[i]], lo[ord[i]], i, i};
    while (i < n && (long long)hi[ord[i]] <= (long long)B.b0 + C) {
      B.b1 = std::max(B.b1, hi[ord[i]]);
      i++;
    }
    B.end = i;
    bands.push_back(B);
  }
  // buffers (kept for later calls)
  const size_t need[2] = {(size_t)C * m_nx * 4, (size_t)C * s_nx * 4};
  for (int img = 0; img < 2; img++)
    if (c->band_bytes[img] < need[img]) {
      for (int b = 0; b < 2; b++) {
        if (c->band[img][b]) (void)hipFree(c->band[img][b]);
        c->band[img][b] = nullptr;
      }
      c->band_bytes[img] = 0;
      for (int b = 0; b < 2; b++)
        if (hipMalloc(&c->band[img][b], need[img]) != hipSuccess)
          return fail(c, XCORR_HIP_ENOMEM, "correlate_stream: band buffer allocation failed");
      c->band_bytes[img] = need[img];
    }
  const int K = (int)std::min<long long>(C, 256);          // lines per pinned chunk
  const size_t cb = (size_t)K * std::max(m_nx, s_nx) * 4;
  if (c->chunk_bytes < cb) {
    if (c->cstream) HIPCK(c, hipStreamSynchronize(c->cstream));
    for (int i = 0; i < 2; i++) {
      if (c->chunk[i]) (void)hipHostFree(c->chunk[i]);
      c->chunk[i] = nullptr;
      c->chunk_used[i] = false;
    }
    c->chunk_bytes = 0;
    for (int i = 0; i < 2; i++)
      if (hipHostMalloc(&c->chunk[i], cb, hipHostMallocDefault) != hipSuccess)
        return fail(c, XCORR_HIP_ENOMEM, "correlate_stream: pinned staging allocation failed");
    c->chunk_bytes = cb;
  }
  if (!c->cstream) HIPCK(c, hipStreamCreateWithFlags(&c->cstream, hipStreamNonBlocking));
  for (int i = 0; i < 2; i++) {
    if (!c->ev_chunk[i]) HIPCK(c, hipEventCreateWithFlags(&c->ev_chunk[i], hipEventDisableTiming));
    if (!c->ev_loaded[i]) HIPCK(c, hipEventCreateWithFlags(&c->ev_loaded[i], hipEventDisableTiming));
    if (!c->ev_done[i]) HIPCK(c, hipEventCreateWithFlags(&c->ev_done[i], hipEventDisableTiming));
  }
  if (c->d_res_n < (size_t)n) {
    HIPCK(c, hipStreamSynchronize(c->stream));
    if (c->d_res) (void)hipFree(c->d_res);
    c->d_res = nullptr;
    c->d_res_n = 0;
    if (hipMalloc((void**)&c->d_res, (size_t)n * sizeof(xcorr_hip_result)) != hipSuccess)
      return fail(c, XCORR_HIP_ENOMEM, "correlate_stream: result buffer allocation failed");
    c->d_res_n = n;
  }
  const int nx_of[2] = {m_nx, s_nx}, ny_of[2] = {m_ny, s_ny};
  auto load = [&](int k) -> int {
    const Band& B = bands[k];
    const int buf = k & 1;
    if (k >= 2) HIPCK(c, hipStreamWaitEvent(c->cstream, c->ev_done[buf], 0));   // band k - 2 is done with it
    for (int img = 0; img < 2; img++) {
      const int nx = nx_of[img];
      const int a = std::max(B.b0, 0), b = std::min(B.b1, ny_of[img]);
      for (int l = a; l < b; l += K) {
        const int cnt = std::min(K, b - l);
        const int slot = c->chunk_next;                 // the chunk copied two chunks ago
        c->chunk_next ^= 1;
        if (c->chunk_used[slot]) HIPCK(c, hipEventSynchronize(c->ev_chunk[slot]));
        if (reader(user, img, l, cnt, c->chunk[slot]))
          return fail(c, XCORR_HIP_EINVAL, "correlate_stream: the line reader failed");
        char* dst = (char*)c->band[img][buf] + (size_t)(l - B.b0) * nx * 4;
        HIPCK(c, hipMemcpyAsync(dst, c->chunk[slot], (size_t)cnt * nx * 4, hipMemcpyHostToDevice, c->cstream));
        HIPCK(c, hipEventRecord(c->ev_chunk[slot], c->cstream));
        c->chunk_used[slot] = true;
      }
    }
    HIPCK(c, hipEventRecord(c->ev_loaded[buf], c->cstream));
    return 0;
  };
  // the images set with set_images come back afterwards
  const void* keep_m = c->d_master;
  const void* keep_s = c->d_slave;
  const bool keep_images = c->images;
  const DevParams keep_dp = c->dp;
  std::vector<xcorr_hip_window> sub(n);
  for (int i = 0; i < n; i++) sub[i] = win[ord[i]];
  int rc = load(0);
  for (size_t k = 0; !rc && k < bands.size(); k++) {
    const Band& B = bands[k];
    const int buf = (int)(k & 1);
    rc = -1;
    HIPCK(c, hipStreamWaitEvent(c->stream, c->ev_loaded[buf], 0));
    // device pointers at line 0 of each image, so that line y sits at im + y nx
    c->d_master = (const char*)c->band[0][buf] - (ptrdiff_t)B.b0 * m_nx * 4;
    c->d_slave = (const char*)c->band[1][buf] - (ptrdiff_t)B.b0 * s_nx * 4;
    c->dp.m_nx = m_nx; c->dp.m_ny = m_ny; c->dp.s_nx = s_nx; c->dp.s_ny = s_ny;
    c->dp.m_lo = c->dp.s_lo = B.b0;
    c->dp.m_hi = c->dp.s_hi = (int)(B.b0 + C);
    c->images = true;
    rc = xcorr_hip_correlate(c, sub.data() + B.begin, B.end - B.begin, c->d_res + B.begin,
                             XCORR_HIP_DEVICE_PTRS | XCORR_HIP_HOST_WINDOWS | XCORR_HIP_NO_SYNC);
    if (rc) break;
    HIPCK(c, hipEventRecord(c->ev_done[buf], c->stream));
    if (k + 1 < bands.size()) rc = load((int)k + 1);
  }
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->cstream);
  c->d_master = keep_m;
  c->d_slave = keep_s;
  c->images = keep_images;
  c->dp = keep_dp;
  if (rc) return rc;
  if (e != hipSuccess) return fail(c, XCORR_HIP_EDEVICE, std::string("correlate_stream: ") + hipGetErrorString(e));
  std::vector<xcorr_hip_result> tmp(n);
  HIPCK(c, hipMemcpy(tmp.data(), c->d_res, (size_t)n * sizeof(xcorr_hip_result), hipMemcpyDeviceToHost));
  for (int i = 0; i < n; i++) out[ord[i]] = tmp[i];
  // dumps refer to indices of this call
  for (int& q : c->last_perm) q = ord[q + (int)bands.back().begin];
  return 0;
}

// --------------------------------------------------------------------------
// Tuning.  Batch size, LDS tiles (rows / columns per block) and threads per
// block of the generic LDS kernels never change a result bit: every sum they
// form is exact or keeps its order.  A radix plan other than the default is a
// different FP32 rounding of the transforms of that length.
// --------------------------------------------------------------------------
// --------------------------------------------------------------------------
// xcorr_hip_correlate_noalloc: the per-patch kernels on buffers carved from
// one static device array per GPU, for callers that cannot allocate device
// memory while they run.
// --------------------------------------------------------------------------
constexpr size_t kNoAllocBytes = XCORR_HIP_NOALLOC_SCRATCH;
constexpr int kNoAllocDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kNoAllocDevices];   // device address of each GPU's copy of `scratch`

// Address of the current device's copy of `scratch`.  In host code the name
// `scratch` is a host-side stand-in, never a device address: the address is
// looked up once per device (with that device current) and only this pointer
// is handed to kernels.
static int noalloc_scratch(xcorr_hip_ctx* c, unsigned char** out) {
  if (c->device < 0 || c->device >= kNoAllocDevices)
    return fail(c, XCORR_HIP_EINVAL, "correlate_noalloc: device index beyond 63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[c->device]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(c, XCORR_HIP_EDEVICE, std::string("correlate_noalloc: hipGetSymbolAddress(scratch): ") +
                                            hipGetErrorString(e));
    g_scratch_ptr[c->device] = ptr;
  }
  *out = (unsigned char*)g_scratch_ptr[c->device];
  return 0;
}

// The per-patch buffers of nb windows laid out one after another from base,
// each on a 256-byte boundary (the same per-patch strides as a context's own
// buffers).  Returns the bytes they take; with base null only the size.
static size_t carve_patch_bufs(const DevParams& d, bool i8, size_t nb, unsigned char* base, PatchBufs* b) {
  size_t off = 0;
  auto take = [&](size_t bytes) -> void* {
    void* ptr = base ? (void*)(base + off) : nullptr;
    off += (size_t)round_up((long long)bytes, 256);
    return ptr;
  };
  *b = PatchBufs{};
  const size_t N = (size_t)d.N, ns = (size_t)d.nxc * d.nyc;
  b->amp = (float*)take(nb * 2 * N * sizeof(float));
  b->sums = (unsigned long long*)take(nb * 2 * sizeof(unsigned long long));
  b->peak = (unsigned long long*)take(nb * sizeof(unsigned long long));
  if (d.corr_flag == XCORR_HIP_FREQ) {
    b->Z = (float2*)take(nb * N * sizeof(float2));
    b->D = (float2*)take(nb * d.nyc * d.NX * sizeof(float2));
    b->surf = (float*)take(nb * ns * sizeof(float));
    b->partial = (double*)take(nb * std::max(d.groups4, d.nyc / d.G4) * sizeof(double));
  } else {
    b->surfd = (double*)take(nb * ns * sizeof(double));
    if (i8) b->tcaa = (double*)take(nb * ns * sizeof(double));
    b->tpeak = (int*)take(nb * 4 * sizeof(int));
    b->tcave = (double*)take(nb * sizeof(double));
  }
  return off;
}

extern "C" int xcorr_hip_correlate_noalloc(xcorr_hip_ctx* c, const xcorr_hip_window* d_win, int n,
                                           xcorr_hip_result* d_out, unsigned flags) {
  if (!c) return XCORR_HIP_EINVAL;
  if (n < 0 || (n > 0 && (!d_win || !d_out))) return fail(c, XCORR_HIP_EINVAL, "correlate_noalloc: bad arguments");
  if (flags & ~XCORR_HIP_NO_SYNC)
    return fail(c, XCORR_HIP_EINVAL, "correlate_noalloc: the only accepted flag is XCORR_HIP_NO_SYNC");
  if (!c->images) return fail(c, XCORR_HIP_ESTATE, "correlate_noalloc: no images set");
  HIPCK(c, hipSetDevice(c->device));
  c->dp.keep_surf = 0;
  if (n == 0) return 0;
  unsigned char* base = nullptr;
  const int rs = noalloc_scratch(c, &base);
  if (rs) return rs;
  const DevParams& d = c->dp;
  const bool i8 = time_i8_shape(d.corr_flag, d.xs, d.ys);        // energies too
  PatchBufs b;
  const size_t one = carve_patch_bufs(d, i8, 1, nullptr, &b);
  if (one > kNoAllocBytes)
    return fail(c, XCORR_HIP_EINVAL, "correlate_noalloc: the buffers of one " + std::to_string(d.NX) + " x " +
                                         std::to_string(d.NY) + " window (" + std::to_string(one) +
                                         " bytes) exceed the static array (" + std::to_string(kNoAllocBytes) +
                                         " bytes)");
  size_t cap = kNoAllocBytes / one;              // windows per internal batch
  while (cap > 1 && carve_patch_bufs(d, i8, cap, nullptr, &b) > kNoAllocBytes) cap--;
  carve_patch_bufs(d, i8, cap, base, &b);
  for (int off = 0; off < n; off += (int)cap) {
    const int nb = (int)std::min<size_t>(cap, (size_t)(n - off));
    const int rc = run_batch(c, b, d_win + off, nb, d_out + off, false);
    if (rc) return rc;
  }
  if (!(flags & XCORR_HIP_NO_SYNC)) HIPCK(c, hipStreamSynchronize(c->stream));
  return 0;
}

static size_t lds_rows(const DevParams& d, int G) {     // k_range_ampl (the larger of the two row kernels)
  const size_t nel = (size_t)G * d.NX;
  return std::max(nel * sizeof(float2) * (d.ri > 2 ? 3 : 2) + nel * sizeof(float), nel * sizeof(float2) * 2);
}
static size_t lds_cols(const DevParams& d, int C) { return (size_t)C * d.NY * sizeof(float2) * 3; }
constexpr size_t kMaxDynLds = 65536;

static void plan_to_ints(const FftPlan& p, int32_t* R) {
  for (int i = 0; i < 4; i++) R[i] = i < p.np ? p.R[i] : 0;
}
static FftPlan ints_to_plan(const int32_t* R) {
  FftPlan p{0, {0, 0, 0, 0}};
  while (p.np < 4 && R[p.np] != 0) {
    p.R[p.np] = R[p.np];
    p.np++;
  }
  return p;
}

extern "C" int xcorr_hip_get_tuning(xcorr_hip_ctx* c, xcorr_hip_tuning* t) {
  if (!c || !t) return XCORR_HIP_EINVAL;
  memset(t, 0, sizeof(*t));
  t->batch = c->tune_batch > 0 ? std::min(c->tune_batch, c->max_batch) : c->max_batch;
  t->rows_per_block = c->dp.G1;
  t->cols_per_block = c->dp.C3;
  t->threads = c->nt_rows;
  plan_to_ints(c->dp.plx, t->radix_x);
  plan_to_ints(c->dp.ply, t->radix_y);
  return 0;
}

extern "C" int xcorr_hip_set_tuning(xcorr_hip_ctx* c, const xcorr_hip_tuning* t) {
  if (!c || !t) return XCORR_HIP_EINVAL;
  DevParams& d = c->dp;
  if (t->batch < 1 || t->batch > c->max_batch) return fail(c, XCORR_HIP_EINVAL, "set_tuning: batch out of range");
  if (t->threads < 64 || t->threads > kMaxThreads || t->threads % 64)
    return fail(c, XCORR_HIP_EINVAL, "set_tuning: threads must be a multiple of 64 in [64, 512]");
  const int G = t->rows_per_block, C = t->cols_per_block;
  if (!is_pow2(G) || G > d.NY || lds_rows(d, G) > kMaxDynLds)
    return fail(c, XCORR_HIP_EINVAL, "set_tuning: rows_per_block must be a power of two <= npy within 64 KB of LDS");
  if (!is_pow2(C) || C > d.NX || lds_cols(d, C) > kMaxDynLds)
    return fail(c, XCORR_HIP_EINVAL, "set_tuning: cols_per_block must be a power of two <= npx within 64 KB of LDS");
  const FftPlan px = ints_to_plan(t->radix_x), py = ints_to_plan(t->radix_y);
  if (!plan_valid(px, d.NX) || !plan_valid(py, d.NY))
    return fail(c, XCORR_HIP_EINVAL, "set_tuning: a radix plan (2/4/8/16, up to 4 passes) must multiply to its length");
  if (d.NX == d.NY && !same_plan(px, py))
    return fail(c, XCORR_HIP_EINVAL, "set_tuning: transforms of equal length share one radix plan");
  if (c->strips && (!same_plan(px, plan_for(d.NX)) || !same_plan(py, plan_for(d.NY))) && ensure_patch_buffers(c))
    return fail(c, XCORR_HIP_ENOMEM, "set_tuning: per-patch buffers for a non-default radix plan");
  c->tune_batch = t->batch;
  c->nt_rows = t->threads;
  d.G1 = G;
  d.C3 = C;
  d.plx = px;
  d.ply = py;
  d.groups4 = groups4_for(c);
  return 0;
}

// Radix plans for length n: non-increasing and non-decreasing radix sequences
// of 2/4/8/16, at most one radix 2, at most 4 passes.
static void plan_candidates(int n, std::vector<FftPlan>& out) {
  out.clear();
  const int rad[4] = {16, 8, 4, 2};
  std::vector<FftPlan> all;
  FftPlan cur{0, {0, 0, 0, 0}};
  std::function<void(int, int, int)> rec = [&](int rem, int maxi, int twos) {
    if (rem == 1) {
      if (cur.np > 0) all.push_back(cur);
      return;
    }
    if (cur.np == 4) return;
    for (int i = maxi; i < 4; i++) {
      const int r = rad[i];
      if (rem % r) continue;
      if (r == 2 && twos) continue;
      cur.R[cur.np++] = r;
      rec(rem / r, i, twos + (r == 2));
      cur.np--;
    }
  };
  rec(n, 0, 0);
  for (const FftPlan& p : all) {
    out.push_back(p);
    FftPlan q = p;
    std::reverse(q.R, q.R + q.np);
    if (!same_plan(p, q)) out.push_back(q);
  }
}

extern "C" int xcorr_hip_autotune(xcorr_hip_ctx* c, const xcorr_hip_window* win, int n, unsigned flags,
                                  xcorr_hip_tuning* out) {
  if (!c) return XCORR_HIP_EINVAL;
  if (!win || n <= 0) return fail(c, XCORR_HIP_EINVAL, "autotune: no sample windows");
  if (!c->images) return fail(c, XCORR_HIP_ESTATE, "autotune: no images set");
  HIPCK(c, hipSetDevice(c->device));
  const DevParams& d = c->dp;
  n = std::min(n, 2 * c->max_batch);
  if (c->d_res_n < (size_t)n) {
    HIPCK(c, hipStreamSynchronize(c->stream));
    if (c->d_res) (void)hipFree(c->d_res);
    c->d_res = nullptr;
    c->d_res_n = 0;
    if (hipMalloc((void**)&c->d_res, (size_t)n * sizeof(xcorr_hip_result)) != hipSuccess)
      return fail(c, XCORR_HIP_ENOMEM, "autotune: result buffer allocation failed");
    c->d_res_n = n;
  }
  hipEvent_t e0, e1;
  HIPCK(c, hipEventCreate(&e0));
  HIPCK(c, hipEventCreate(&e1));
  int timed = 0;
  // time of one pass over the sample: best of three after one untimed run
  auto measure = [&](const xcorr_hip_tuning& t, float* ms) -> int {
    int rc = xcorr_hip_set_tuning(c, &t);
    if (rc) return rc;
    float best = 1e30f;
    for (int r = 0; r < 4; r++) {
      HIPCK(c, hipEventRecord(e0, c->stream));
      rc = xcorr_hip_correlate(c, win, n, c->d_res, XCORR_HIP_DEVICE_PTRS | XCORR_HIP_HOST_WINDOWS | XCORR_HIP_NO_SYNC);
      if (rc) return rc;
      HIPCK(c, hipEventRecord(e1, c->stream));
      HIPCK(c, hipEventSynchronize(e1));
      float ms1 = 0.0f;
      HIPCK(c, hipEventElapsedTime(&ms1, e0, e1));
      if (r > 0) best = std::min(best, ms1);
    }
    timed++;
    *ms = best;
    return 0;
  };
  xcorr_hip_tuning best;
  xcorr_hip_get_tuning(c, &best);
  float best_ms = 0.0f;
  int rc = measure(best, &best_ms);
  auto consider = [&](const xcorr_hip_tuning& t) -> int {
    float ms = 0.0f;
    const int r = measure(t, &ms);
    if (r == XCORR_HIP_EINVAL) return 0;      // not a valid setting for this shape
    if (r) return r;
    if (ms < best_ms) {
      best_ms = ms;
      best = t;
    }
    return 0;
  };
  // coordinate descent: threads, row tile, column tile, radix plans, batch.
  // Strip contexts start from the strips (default plans); a radix plan
  // candidate runs the per-patch kernels, and only if one of them wins are the
  // launch settings of those kernels tuned.
  auto tune_launch = [&]() {
    if (!rc)
      for (int th : {128, 256, 512}) {
        xcorr_hip_tuning t = best;
        t.threads = th;
        if (th != best.threads && (rc = consider(t))) break;
      }
    if (!rc)
      for (int G = 1; G <= d.NY && !rc; G *= 2) {
        xcorr_hip_tuning t = best;
        t.rows_per_block = G;
        const size_t nel = (size_t)G * d.NX;
        if (G != best.rows_per_block && nel >= 256 && lds_rows(d, G) <= kMaxDynLds) rc = consider(t);
      }
    if (!rc)
      for (int C = 1; C <= d.NX && !rc; C *= 2) {
        xcorr_hip_tuning t = best;
        t.cols_per_block = C;
        const size_t nel = (size_t)C * d.NY;
        if (C != best.cols_per_block && nel >= 256 && lds_cols(d, C) <= kMaxDynLds) rc = consider(t);
      }
  };
  auto tune_radix = [&]() {
    if (rc || d.corr_flag != XCORR_HIP_FREQ || !(flags & XCORR_HIP_TUNE_RADIX)) return;
    std::vector<FftPlan> cand;
    plan_candidates(d.NX, cand);
    for (const FftPlan& pl : cand) {
      if (rc) break;
      xcorr_hip_tuning t = best;
      plan_to_ints(pl, t.radix_x);
      if (d.NX == d.NY) plan_to_ints(pl, t.radix_y);
      if (!same_plan(pl, ints_to_plan(best.radix_x))) rc = consider(t);
    }
    if (d.NX != d.NY) {
      plan_candidates(d.NY, cand);
      for (const FftPlan& pl : cand) {
        if (rc) break;
        xcorr_hip_tuning t = best;
        plan_to_ints(pl, t.radix_y);
        if (!same_plan(pl, ints_to_plan(best.radix_y))) rc = consider(t);
      }
    }
  };
  if (!c->strips) {
    tune_launch();
    tune_radix();
  } else {
    tune_radix();
    if (!same_plan(ints_to_plan(best.radix_x), plan_for(d.NX)) || !same_plan(ints_to_plan(best.radix_y), plan_for(d.NY)))
      tune_launch();
  }
  if (!rc)
    for (int b = c->max_batch / 2; b >= 512 && !rc; b /= 2) {
      xcorr_hip_tuning t = best;
      t.batch = b;
      rc = consider(t);
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  rc = xcorr_hip_set_tuning(c, &best);
  if (rc) return rc;
  best.ms_per_batch = best_ms * (float)std::min(best.batch, n) / (float)n;
  best.candidates = timed;
  if (out) *out = best;
  return 0;
}

// --------------------------------------------------------------------------
// Multi-pair workloads: many SLC pairs with their own patch shapes, captured
// once into one HIP graph.  Pairs with equal parameters form a class that
// shares one context (scratch, tables, tuning); each class is one branch of
// the graph, so the branches of different shapes run concurrently.  Every
// plan (strip plans, device windows) is built and uploaded at create, so a
// run is a single graph launch with no host work.
// --------------------------------------------------------------------------
struct MultiBatch {
  int pair;                               // pair index
  int off;                                // strips: plan offset (ints); generic: index in the class's device windows
  int first;                              // generic: index of the batch's first window within its pair
  int nb;                                 // windows in this batch
  StripBatch sb;                          // strips only
};

struct xcorr_hip_multi {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<xcorr_hip_pair> pairs;
  std::vector<int> pair_class;
  std::vector<xcorr_hip_ctx*> ctx;        // per class
  std::vector<xcorr_hip_tuning> tuning;   // per class
  std::vector<void*> dbuf;                // per class: strip plans or windows
  std::vector<std::vector<MultiBatch>> batches;
  std::vector<hipEvent_t> events;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::string err;
};

static std::string g_multi_err;

static int mfail(xcorr_hip_multi* m, int code, const std::string& msg) {
  if (m) m->err = msg;
  else g_multi_err = msg;
  return code;
}

static void multi_free(xcorr_hip_multi* m) {
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  if (m->exec) (void)hipGraphExecDestroy(m->exec);
  if (m->graph) (void)hipGraphDestroy(m->graph);
  for (hipEvent_t e : m->events) (void)hipEventDestroy(e);
  for (void* q : m->dbuf)
    if (q) (void)hipFree(q);
  for (xcorr_hip_ctx* c : m->ctx) xcorr_hip_destroy(c);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

// Point a context at a pair's device images.
static void bind_pair(xcorr_hip_ctx* c, const xcorr_hip_pair& q) {
  c->d_master = q.master;
  c->d_slave = q.slave;
  c->dp.m_nx = q.m_nx; c->dp.m_ny = q.m_ny; c->dp.s_nx = q.s_nx; c->dp.s_ny = q.s_ny;
  c->dp.m_lo = 0; c->dp.m_hi = q.m_ny; c->dp.s_lo = 0; c->dp.s_hi = q.s_ny;
  c->images = true;
}

#define MHIPCK(m, call)                                                                     \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return mfail(m, XCORR_HIP_EDEVICE, std::string(#call ": ") + hipGetErrorString(e_)); \
  } while (0)

static int multi_build(xcorr_hip_multi* m, unsigned flags) {
  const int np = (int)m->pairs.size();
  // classes of equal parameters
  std::vector<xcorr_hip_params> cls;
  for (int i = 0; i < np; i++) {
    int k = -1;
    for (size_t j = 0; j < cls.size(); j++)
      if (!memcmp(&cls[j], &m->pairs[i].params, sizeof(xcorr_hip_params))) k = (int)j;
    if (k < 0) {
      k = (int)cls.size();
      cls.push_back(m->pairs[i].params);
    }
    m->pair_class.push_back(k);
  }
  const int nc = (int)cls.size();
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = (size_t)8 << 30;
  const size_t budget = std::min<size_t>((size_t)12 << 30, free_b / (2 * (size_t)std::max(nc, 1)));
  m->ctx.assign(nc, nullptr);
  m->tuning.assign(nc, xcorr_hip_tuning{});
  m->dbuf.assign(nc, nullptr);
  m->batches.assign(nc, {});
  for (int k = 0; k < nc; k++) {
    const xcorr_hip_params& p = cls[k];
    int biggest = -1, nmax = 0;
    for (int i = 0; i < np; i++)
      if (m->pair_class[i] == k && m->pairs[i].n > nmax) {
        nmax = m->pairs[i].n;
        biggest = i;
      }
    const size_t N = (size_t)p.npx * p.npy;
    const size_t per = N * 4 * (2 + 4) + (size_t)p.nyc * p.npx * 8 + (size_t)p.nxc * p.nyc * 8 + 256;
    int mb = (int)std::max<size_t>(256, std::min<size_t>(8192, budget / per));
    mb = std::max(64, std::min(mb, std::max(nmax, 64)));
    int rc = xcorr_hip_create(&p, mb, m->device, &m->ctx[k]);
    if (rc) return mfail(m, rc, std::string("multi: context for a pair class: ") + xcorr_hip_last_error(nullptr));
    xcorr_hip_ctx* c = m->ctx[k];
    c->one_context = true;         // the graph runs each class on one stream: tune it that way
    if ((flags & XCORR_HIP_TUNE) && biggest >= 0) {
      bind_pair(c, m->pairs[biggest]);
      rc = xcorr_hip_autotune(c, m->pairs[biggest].win, m->pairs[biggest].n, flags, &m->tuning[k]);
      if (rc) return mfail(m, rc, std::string("multi: autotune: ") + c->err);
    } else {
      xcorr_hip_get_tuning(c, &m->tuning[k]);
    }
    const int cap = m->tuning[k].batch;
    // plans / windows of every pair of the class, uploaded once
    std::vector<int> host;
    std::vector<xcorr_hip_window> hw;
    for (int i = 0; i < np; i++) {
      if (m->pair_class[i] != k || m->pairs[i].n == 0) continue;
      const xcorr_hip_pair& q = m->pairs[i];
      if (use_strips(c)) {
        bind_pair(c, q);
        std::vector<int> ord;
        strip_order(q.win, q.n, ord, c->order_scratch);
        const size_t pi = plan_ints(c);
        for (int pos = 0; pos < q.n;) {
          MultiBatch mbt{};
          mbt.pair = i;
          mbt.off = (int)host.size();
          host.resize(host.size() + pi, 0);
          if (!plan_strip_batch(c, c->dp, q.win, ord, &pos, q.n, cap, host.data() + mbt.off, &mbt.sb))
            return mfail(m, XCORR_HIP_EINVAL, "multi: a window does not fit the line capacity");
          mbt.nb = mbt.sb.nb;
          m->batches[k].push_back(mbt);
        }
      } else {
        const int base = (int)hw.size();
        hw.insert(hw.end(), q.win, q.win + q.n);
        for (int off = 0; off < q.n; off += cap) {
          MultiBatch mbt{};
          mbt.pair = i;
          mbt.off = base + off;
          mbt.first = off;
          mbt.nb = std::min(cap, q.n - off);
          m->batches[k].push_back(mbt);
        }
      }
    }
    const size_t bytes = use_strips(c) ? host.size() * sizeof(int) : hw.size() * sizeof(xcorr_hip_window);
    if (bytes) {
      if (hipMalloc(&m->dbuf[k], bytes) != hipSuccess) return mfail(m, XCORR_HIP_ENOMEM, "multi: plan allocation failed");
      MHIPCK(m, hipMemcpy(m->dbuf[k], use_strips(c) ? (const void*)host.data() : (const void*)hw.data(), bytes,
                          hipMemcpyHostToDevice));
    }
  }
  // capture: one branch per class, forked from and joined into m->stream
  hipEvent_t fork;
  MHIPCK(m, hipEventCreateWithFlags(&fork, hipEventDisableTiming));
  m->events.push_back(fork);
  for (int k = 0; k < nc; k++) {
    hipEvent_t e;
    MHIPCK(m, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    m->events.push_back(e);
  }
  MHIPCK(m, hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
  int rc = 0;
  hipError_t e = hipEventRecord(fork, m->stream);
  for (int k = 0; k < nc && e == hipSuccess && !rc; k++) {
    xcorr_hip_ctx* c = m->ctx[k];
    e = hipStreamWaitEvent(c->stream, fork, 0);
    for (const MultiBatch& b : m->batches[k]) {
      if (e != hipSuccess || rc) break;
      const xcorr_hip_pair& q = m->pairs[b.pair];
      bind_pair(c, q);
      if (use_strips(c))
        rc = launch_strip_batch(c, c->dp, q.master, q.slave, c->stream, (const int*)m->dbuf[k] + b.off, b.sb, q.out,
                                true, false);
      else
        rc = run_batch(c, ctx_bufs(c), (const xcorr_hip_window*)m->dbuf[k] + b.off, b.nb, q.out + b.first, false);
    }
    if (e == hipSuccess && !rc) e = hipEventRecord(m->events[1 + k], c->stream);
    if (e == hipSuccess && !rc) e = hipStreamWaitEvent(m->stream, m->events[1 + k], 0);
  }
  hipGraph_t g = nullptr;
  const hipError_t ee = hipStreamEndCapture(m->stream, &g);
  m->graph = g;
  if (rc) return mfail(m, rc, std::string("multi: capture: ") + m->ctx[0]->err);
  if (e != hipSuccess) return mfail(m, XCORR_HIP_EDEVICE, std::string("multi: capture: ") + hipGetErrorString(e));
  if (ee != hipSuccess) return mfail(m, XCORR_HIP_EDEVICE, std::string("multi: end capture: ") + hipGetErrorString(ee));
  MHIPCK(m, hipGraphInstantiate(&m->exec, m->graph, nullptr, nullptr, 0));
  return 0;
}

extern "C" int xcorr_hip_multi_create(const xcorr_hip_pair* pairs, int npairs, int device, unsigned flags,
                                      xcorr_hip_multi** out) {
  if (!out || npairs < 0 || (npairs > 0 && !pairs)) return mfail(nullptr, XCORR_HIP_EINVAL, "multi_create: bad arguments");
  *out = nullptr;
  for (int i = 0; i < npairs; i++) {
    const xcorr_hip_pair& q = pairs[i];
    std::string why;
    if (validate(&q.params, &why)) return mfail(nullptr, XCORR_HIP_EINVAL, "multi_create: pair " + std::to_string(i) + ": " + why);
    if (!q.master || !q.slave || q.m_nx <= 0 || q.m_ny <= 0 || q.s_nx <= 0 || q.s_ny <= 0 || q.n < 0 ||
        (q.n > 0 && (!q.win || !q.out)))
      return mfail(nullptr, XCORR_HIP_EINVAL, "multi_create: pair " + std::to_string(i) + ": bad images or windows");
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return mfail(nullptr, XCORR_HIP_EDEVICE, "no HIP device");
  if (device < 0 && hipGetDevice(&device) != hipSuccess) device = 0;
  if (device >= ndev) return mfail(nullptr, XCORR_HIP_EINVAL, "device index out of range");
  xcorr_hip_multi* m = new xcorr_hip_multi();
  m->device = device;
  m->pairs.assign(pairs, pairs + npairs);
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) {
    multi_free(m);
    return mfail(nullptr, XCORR_HIP_EDEVICE, "multi_create: stream creation failed");
  }
  const int rc = multi_build(m, flags);
  if (rc) {
    const std::string msg = m->err;
    multi_free(m);
    return mfail(nullptr, rc, msg);
  }
  *out = m;
  return 0;
}

extern "C" int xcorr_hip_multi_run(xcorr_hip_multi* m, unsigned flags) {
  if (!m) return XCORR_HIP_EINVAL;
  MHIPCK(m, hipSetDevice(m->device));
  MHIPCK(m, hipGraphLaunch(m->exec, m->stream));
  if (!(flags & XCORR_HIP_NO_SYNC)) MHIPCK(m, hipStreamSynchronize(m->stream));
  return 0;
}

extern "C" int xcorr_hip_multi_sync(xcorr_hip_multi* m) {
  if (!m) return XCORR_HIP_EINVAL;
  MHIPCK(m, hipStreamSynchronize(m->stream));
  return 0;
}

extern "C" void* xcorr_hip_multi_stream(xcorr_hip_multi* m) { return m ? (void*)m->stream : nullptr; }

extern "C" int xcorr_hip_multi_classes(xcorr_hip_multi* m, xcorr_hip_params* params, xcorr_hip_tuning* tuning,
                                       int* pair_class, int cap) {
  if (!m) return XCORR_HIP_EINVAL;
  const int nc = (int)m->ctx.size();
  for (int k = 0; k < nc && k < cap; k++) {
    if (params) params[k] = m->ctx[k]->p;
    if (tuning) tuning[k] = m->tuning[k];
  }
  if (pair_class)
    for (size_t i = 0; i < m->pair_class.size(); i++) pair_class[i] = m->pair_class[i];
  return nc;
}

extern "C" void xcorr_hip_multi_destroy(xcorr_hip_multi* m) {
  if (!m) return;
  (void)hipSetDevice(m->device);
  multi_free(m);
}

extern "C" const char* xcorr_hip_multi_last_error(const xcorr_hip_multi* m) {
  return m ? m->err.c_str() : g_multi_err.c_str();
}

extern "C" int xcorr_hip_kernel_times(xcorr_hip_ctx* c, const char** names, float* ms, int* launches, int cap) {
  if (!c) return XCORR_HIP_EINVAL;
  const int n = (int)c->timers.size();
  for (int i = 0; i < n && i < cap; i++) {
    if (names) names[i] = c->timers[i].name;
    if (ms) ms[i] = c->timers[i].ms;
    if (launches) launches[i] = c->timers[i].launches;
  }
  return n;
}

extern "C" long xcorr_hip_dump(xcorr_hip_ctx* c, int what, int slot, void* host, size_t bytes) {
  if (!c || !host) return XCORR_HIP_EINVAL;
  const int loc = slot;
  slot = -1;
  for (size_t i = 0; i < c->last_perm.size(); i++)
    if (c->last_perm[i] == loc) slot = (int)i;
  if (slot < 0) return fail(c, XCORR_HIP_EINVAL, "dump: location not in the last batch");
  HIPCK(c, hipSetDevice(c->device));
  HIPCK(c, hipStreamSynchronize(c->stream));
  const DevParams& d = c->dp;
  // strip batches alternate between two contexts: the last one's buffers
  const xcorr_hip_ctx* bf = (c->last_buf && !c->last_rb.empty()) ? c->last_buf : c;
  const void* src = nullptr;
  size_t nbytes = 0;
  std::vector<unsigned long long> tmp;
  double means[2];
  switch (what) {
    case 0:
    case 1:
      if (!c->last_rb.empty())   // npy consecutive strip lines from the row base
        src = (what ? bf->amp2s : bf->amp1s) + (size_t)c->last_rb[slot] * d.NX;
      else
        src = c->amp + ((size_t)slot * 2 + what) * d.N;
      nbytes = (size_t)d.N * sizeof(float);
      break;
    case 2:
      if (d.corr_flag == XCORR_HIP_FREQ) {
        if (!d.keep_surf)
          return fail(c, XCORR_HIP_ESTATE, "dump: the correlation surface is kept only by correlate calls with "
                                           "XCORR_HIP_KEEP_SURFACE");
        src = bf->surf + (size_t)slot * d.nxc * d.nyc;
        nbytes = (size_t)d.nxc * d.nyc * sizeof(float);
      } else {
        src = bf->surfd + (size_t)slot * d.nxc * d.nyc;
        nbytes = (size_t)d.nxc * d.nyc * sizeof(double);
      }
      break;
    case 3: {
      unsigned long long s[2];
      HIPCK(c, hipMemcpy(s, bf->sums + 2 * slot, sizeof(s), hipMemcpyDeviceToHost));
      for (int i = 0; i < 2; i++) means[i] = (double)(long long)s[i] * (1.0 / 1048576.0) / (double)d.N;
      if (bytes < sizeof(means)) return fail(c, XCORR_HIP_EINVAL, "dump: buffer too small");
      memcpy(host, means, sizeof(means));
      return (long)sizeof(means);
    }
    default:
      return fail(c, XCORR_HIP_EINVAL, "dump: unknown buffer");
  }
  if (bytes < nbytes) return fail(c, XCORR_HIP_EINVAL, "dump: buffer too small");
  HIPCK(c, hipMemcpy(host, src, nbytes, hipMemcpyDeviceToHost));
  return (long)nbytes;
}
