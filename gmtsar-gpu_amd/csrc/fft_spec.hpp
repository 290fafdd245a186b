// FP32 arithmetic building blocks of the xcorr path on gfx950.
//
// Every operation here is written out in a fixed order so that results are
// reproducible bit for bit: the file is compiled with -ffp-contract=off and
// the only fused multiply-adds are the explicit fmaf calls.
//
//   complex multiply   re = fma(a.re, b.re, -(a.im*b.im)), im = fma(a.re, b.im, a.im*b.re)
//   FFT                Stockham autosort passes, radix plan per length (plan_for
//                      by default; the generic LDS kernels take any valid plan),
//                      DFT-2/4/8/16 butterflies, twiddles e^{+2 pi i m/n} from a
//                      host-built table (negated imaginary part for forward)
//   amplitude          sqrtf(fma(re, re, im*im))
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace xc {

struct FftPlan {
  int np;      // number of passes
  int R[4];    // radix of each pass
};

// Radix plan per transform length (lengths 1..4096, powers of two).
__host__ __device__ inline FftPlan plan_for(int n) {
  FftPlan p{0, {0, 0, 0, 0}};
  switch (n) {
    case 1: break;
    case 2: p.np = 1; p.R[0] = 2; break;
    case 4: p.np = 1; p.R[0] = 4; break;
    case 8: p.np = 1; p.R[0] = 8; break;
    case 16: p.np = 1; p.R[0] = 16; break;
    case 32: p.np = 2; p.R[0] = 8; p.R[1] = 4; break;
    case 64: p.np = 2; p.R[0] = 8; p.R[1] = 8; break;
    case 128: p.np = 2; p.R[0] = 16; p.R[1] = 8; break;
    case 256: p.np = 2; p.R[0] = 16; p.R[1] = 16; break;
    case 512: p.np = 3; p.R[0] = 8; p.R[1] = 8; p.R[2] = 8; break;
    case 1024: p.np = 3; p.R[0] = 16; p.R[1] = 16; p.R[2] = 4; break;
    case 2048: p.np = 3; p.R[0] = 16; p.R[1] = 16; p.R[2] = 8; break;
    case 4096: p.np = 3; p.R[0] = 16; p.R[1] = 16; p.R[2] = 16; break;
    default: p.np = -1; break;
  }
  return p;
}

// Barrier for LDS data exchanged among the lanes of one wave (every transform
// group of the register-resident kernels lies within one wave): completes the
// wave's LDS writes and keeps the compiler from moving LDS accesses across it.
// Other waves are not waited for.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
// The spectrum of the real part of a row transformed as z = re + i im, at bin
// k, from a = Z[k] and q = Z[(n - k) mod n]: ((a.re + q.re) / 2, (a.im - q.im) / 2).
// Exactly Hermitian: the value at n - k is the conjugate of the value at k.
__device__ __forceinline__ float2 extract_bin(float2 a, float2 q) {
  return make_float2((a.x + q.x) * 0.5f, (a.y - q.y) * 0.5f);
}
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  float t1 = a.y * b.y;
  float t2 = a.y * b.x;
  return make_float2(__builtin_fmaf(a.x, b.x, -t1), __builtin_fmaf(a.x, b.y, t2));
}
template <int DIR>
__device__ __forceinline__ float2 mul_w4(float2 a) {
  return DIR < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}
#define XC_SQRT1_2F 0.70710678118654752440f
#define XC_C16 0.92387953251128675613f
#define XC_S16 0.38268343236508977173f
template <int DIR>
__device__ __forceinline__ float2 mul_w8_1(float2 a) {
  return DIR < 0 ? make_float2(XC_SQRT1_2F * (a.x + a.y), XC_SQRT1_2F * (a.y - a.x))
                 : make_float2(XC_SQRT1_2F * (a.x - a.y), XC_SQRT1_2F * (a.x + a.y));
}
template <int DIR>
__device__ __forceinline__ float2 mul_w8_3(float2 a) {
  return DIR < 0 ? make_float2(XC_SQRT1_2F * (a.y - a.x), -(XC_SQRT1_2F * (a.x + a.y)))
                 : make_float2(-(XC_SQRT1_2F * (a.x + a.y)), XC_SQRT1_2F * (a.x - a.y));
}

template <int DIR>
__device__ __forceinline__ void dft4(float2& v0, float2& v1, float2& v2, float2& v3) {
  float2 t0 = cadd(v0, v2), t1 = csub(v0, v2);
  float2 t2 = cadd(v1, v3), t3 = mul_w4<DIR>(csub(v1, v3));
  v0 = cadd(t0, t2);
  v2 = csub(t0, t2);
  v1 = cadd(t1, t3);
  v3 = csub(t1, t3);
}

template <int R, int DIR>
struct Dft;

template <int DIR>
struct Dft<2, DIR> {
  __device__ __forceinline__ static void run(float2* v) {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};
template <int DIR>
struct Dft<4, DIR> {
  __device__ __forceinline__ static void run(float2* v) { dft4<DIR>(v[0], v[1], v[2], v[3]); }
};
template <int DIR>
struct Dft<8, DIR> {
  __device__ __forceinline__ static void run(float2* v) {
    float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    float2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<DIR>(e0, e1, e2, e3);
    dft4<DIR>(o0, o1, o2, o3);
    o1 = mul_w8_1<DIR>(o1);
    o2 = mul_w4<DIR>(o2);
    o3 = mul_w8_3<DIR>(o3);
    v[0] = cadd(e0, o0); v[4] = csub(e0, o0);
    v[1] = cadd(e1, o1); v[5] = csub(e1, o1);
    v[2] = cadd(e2, o2); v[6] = csub(e2, o2);
    v[3] = cadd(e3, o3); v[7] = csub(e3, o3);
  }
};
template <int DIR>
struct Dft<16, DIR> {
  __device__ __forceinline__ static void run(float2* v) {
    float2 f[4][4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      f[j][0] = v[j]; f[j][1] = v[j + 4]; f[j][2] = v[j + 8]; f[j][3] = v[j + 12];
      dft4<DIR>(f[j][0], f[j][1], f[j][2], f[j][3]);
    }
    const float2 w1 = make_float2(XC_C16, DIR < 0 ? -XC_S16 : XC_S16);
    const float2 w3 = make_float2(XC_S16, DIR < 0 ? -XC_C16 : XC_C16);
    const float2 w9 = make_float2(-XC_C16, DIR < 0 ? XC_S16 : -XC_S16);
    f[1][1] = cmul(f[1][1], w1);
    f[1][2] = mul_w8_1<DIR>(f[1][2]);
    f[1][3] = cmul(f[1][3], w3);
    f[2][1] = mul_w8_1<DIR>(f[2][1]);
    f[2][2] = mul_w4<DIR>(f[2][2]);
    f[2][3] = mul_w8_3<DIR>(f[2][3]);
    f[3][1] = cmul(f[3][1], w3);
    f[3][2] = mul_w8_3<DIR>(f[3][2]);
    f[3][3] = cmul(f[3][3], w9);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      float2 a = f[0][q], b = f[1][q], c = f[2][q], d = f[3][q];
      dft4<DIR>(a, b, c, d);
      v[q] = a; v[q + 4] = b; v[q + 8] = c; v[q + 12] = d;
    }
  }
};

__device__ __forceinline__ float ampl(float2 v) { return sqrtf(__builtin_fmaf(v.x, v.x, v.y * v.y)); }

// One Stockham pass over `nrows` independent length-n rows held in LDS.
// L = product of the radices already applied; tw = e^{+2 pi i m/n}, m < n.
template <int R, int DIR>
__device__ __forceinline__ void lds_pass(const float2* __restrict__ src, float2* __restrict__ dst, int n, int L,
                                         int nrows, const float2* __restrict__ tw) {
  const int nb = n / R;
  const int total = nrows * nb;
  const int tstep = n / (L * R);
  for (int t = threadIdx.x; t < total; t += blockDim.x) {
    const int row = t / nb;
    const int j = t - row * nb;
    const int k = j & (L - 1);
    const float2* s = src + row * n;
    float2* d = dst + row * n;
    float2 v[R];
#pragma unroll
    for (int r = 0; r < R; r++) v[r] = s[j + r * nb];
    if (L > 1) {
#pragma unroll
      for (int r = 1; r < R; r++) {
        float2 w = tw[r * k * tstep];
        if (DIR < 0) w.y = -w.y;
        v[r] = cmul(v[r], w);
      }
    }
    Dft<R, DIR>::run(v);
    const int base = (j - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; r++) d[base + r * L] = v[r];
  }
}

// FFT of `nrows` rows of length n held in LDS buffer a (b is scratch of the
// same size).  Returns the buffer that holds the result.  Callers must
// __syncthreads() before (data written) -- this function syncs between and
// after the passes.
template <int DIR>
__device__ float2* lds_fft_rows(float2* a, float2* b, int n, int nrows, const float2* __restrict__ tw,
                                const FftPlan& pl) {
  float2* src = a;
  float2* dst = b;
  int L = 1;
  for (int p = 0; p < pl.np; p++) {
    switch (pl.R[p]) {
      case 2: lds_pass<2, DIR>(src, dst, n, L, nrows, tw); break;
      case 4: lds_pass<4, DIR>(src, dst, n, L, nrows, tw); break;
      case 8: lds_pass<8, DIR>(src, dst, n, L, nrows, tw); break;
      default: lds_pass<16, DIR>(src, dst, n, L, nrows, tw); break;
    }
    L *= pl.R[p];
    float2* t = src;
    src = dst;
    dst = t;
    __syncthreads();
  }
  return src;
}
template <int DIR>
__device__ float2* lds_fft_rows(float2* a, float2* b, int n, int nrows, const float2* __restrict__ tw) {
  return lds_fft_rows<DIR>(a, b, n, nrows, tw, plan_for(n));
}

// A radix plan is valid for n when its radices (2, 4, 8 or 16) multiply to n.
__host__ __device__ inline bool plan_valid(const FftPlan& p, int n) {
  if (p.np < 0 || p.np > 4) return false;
  long prod = 1;
  for (int i = 0; i < p.np; i++) {
    if (p.R[i] != 2 && p.R[i] != 4 && p.R[i] != 8 && p.R[i] != 16) return false;
    prod *= p.R[i];
  }
  return prod == n;
}

}  // namespace xc
