/*
 * TEST INFRASTRUCTURE ONLY -- FP32 "sequential" restatement of the xcorr hot
 * path (see xcorr_oracle.h).  Every floating-point operation below follows a
 * fixed specification that the HIP kernels implement as well:
 *
 *   complex multiply   re = fma(a.re, b.re, -(a.im*b.im)), im = fma(a.re, b.im, a.im*b.re)
 *   FFT                Stockham autosort, radix plan per length (fft_plan),
 *                      DFT-2/4/8/16 butterflies below, twiddles from one
 *                      table (xo_twiddle64 rounded to float)
 *   amplitude          sqrtf(fma(re, re, im*im))
 *   patch mean         exact fixed-point sum of amplitudes (2^-20 units)
 *   row spectra        every row transforms z = amp1 + i amp2 (amp2 on the
 *                      template columns only); rows outside the mask take the
 *                      master's spectrum out of it, X[k] = ((Z[k].re +
 *                      Z[-k].re) / 2, (Z[k].im - Z[-k].im) / 2); the means
 *                      come off in this spectrum
 *   |c3|               the kept lag rows iy and iy + nyc/2 share one inverse
 *                      row transform as ra + i rb: |Re| * 1/N and |Im| * 1/N
 *   high-res           direct 8-term sums per 1-D pass; the column pass pairs
 *                      the bins k and 8 - k and the rows s and L - s
 *                      (xo32_interp2d)
 *
 * Compile with -ffp-contract=off so that only the explicit fmaf calls fuse.
 */
#include "xcorr_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef xo_c32 c32;

/* ------------------------------------------------------------------------ */
/* twiddles: e^{+2 pi i m / n} with exact values on the axes and diagonals  */
/* ------------------------------------------------------------------------ */
void xo_twiddle64(int n, long m, double *c, double *s) {
  long mm = m % n;
  if (mm < 0) mm += n;
  if (n < 4) { /* n = 1 or 2 */
    *c = (mm == 0) ? 1.0 : -1.0;
    *s = 0.0;
    return;
  }
  long q = (4 * mm) / n;          /* quadrant */
  long r = mm - q * (n / 4);      /* 0 <= r < n/4 */
  double c0, s0;
  if (r == 0) {
    c0 = 1.0; s0 = 0.0;
  } else if (8 * r == n) {
    c0 = s0 = 0.70710678118654752440;
  } else if (8 * r < n) {
    double t = 2.0 * M_PI * (double)r / (double)n;
    c0 = cos(t); s0 = sin(t);
  } else {
    double t = 2.0 * M_PI * (double)(n / 4 - r) / (double)n;
    c0 = sin(t); s0 = cos(t);
  }
  switch (q) {
    case 0: *c = c0;  *s = s0;  break;
    case 1: *c = -s0; *s = c0;  break;
    case 2: *c = -c0; *s = -s0; break;
    default: *c = s0; *s = -c0; break;
  }
}

static c32 twf(int n, long m, int dir) {
  double c, s;
  xo_twiddle64(n, m, &c, &s);
  c32 w = {(float)c, (float)s};
  if (dir < 0) w.im = -w.im;
  return w;
}

/* ------------------------------------------------------------------------ */
/* arithmetic primitives                                                    */
/* ------------------------------------------------------------------------ */
static inline c32 cadd(c32 a, c32 b) { c32 r = {a.re + b.re, a.im + b.im}; return r; }
static inline c32 csub(c32 a, c32 b) { c32 r = {a.re - b.re, a.im - b.im}; return r; }
static inline c32 cmul(c32 a, c32 b) {
  float t1 = a.im * b.im;
  float t2 = a.im * b.re;
  c32 r = {fmaf(a.re, b.re, -t1), fmaf(a.re, b.im, t2)};
  return r;
}
/* multiply by w4 = -i (dir<0) or +i (dir>0) */
static inline c32 mul_w4(c32 a, int dir) {
  c32 r;
  if (dir < 0) { r.re = a.im; r.im = -a.re; }
  else { r.re = -a.im; r.im = a.re; }
  return r;
}
#define XO_SQRT1_2F 0.70710678118654752440f
/* multiply by w8^1 */
static inline c32 mul_w8_1(c32 a, int dir) {
  c32 r;
  if (dir < 0) { r.re = XO_SQRT1_2F * (a.re + a.im); r.im = XO_SQRT1_2F * (a.im - a.re); }
  else { r.re = XO_SQRT1_2F * (a.re - a.im); r.im = XO_SQRT1_2F * (a.re + a.im); }
  return r;
}
/* multiply by w8^3 */
static inline c32 mul_w8_3(c32 a, int dir) {
  c32 r;
  if (dir < 0) { r.re = XO_SQRT1_2F * (a.im - a.re); r.im = -(XO_SQRT1_2F * (a.re + a.im)); }
  else { r.re = -(XO_SQRT1_2F * (a.re + a.im)); r.im = XO_SQRT1_2F * (a.re - a.im); }
  return r;
}

#define XO_C16 0.92387953251128675613f /* cos(pi/8) */
#define XO_S16 0.38268343236508977173f /* sin(pi/8) */

static inline void dft2(c32 *v) {
  c32 a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

static inline void dft4(c32 *v0, c32 *v1, c32 *v2, c32 *v3, int dir) {
  c32 t0 = cadd(*v0, *v2), t1 = csub(*v0, *v2);
  c32 t2 = cadd(*v1, *v3), t3 = mul_w4(csub(*v1, *v3), dir);
  *v0 = cadd(t0, t2);
  *v2 = csub(t0, t2);
  *v1 = cadd(t1, t3);
  *v3 = csub(t1, t3);
}

static inline void dft8(c32 *v, int dir) {
  c32 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  c32 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4(&e0, &e1, &e2, &e3, dir);
  dft4(&o0, &o1, &o2, &o3, dir);
  o1 = mul_w8_1(o1, dir);
  o2 = mul_w4(o2, dir);
  o3 = mul_w8_3(o3, dir);
  v[0] = cadd(e0, o0); v[4] = csub(e0, o0);
  v[1] = cadd(e1, o1); v[5] = csub(e1, o1);
  v[2] = cadd(e2, o2); v[6] = csub(e2, o2);
  v[3] = cadd(e3, o3); v[7] = csub(e3, o3);
}

static inline void dft16(c32 *v, int dir) {
  c32 f[4][4];
  for (int j = 0; j < 4; j++) {
    f[j][0] = v[j]; f[j][1] = v[j + 4]; f[j][2] = v[j + 8]; f[j][3] = v[j + 12];
    dft4(&f[j][0], &f[j][1], &f[j][2], &f[j][3], dir);
  }
  c32 w1 = {XO_C16, dir < 0 ? -XO_S16 : XO_S16};   /* w16^1 */
  c32 w3 = {XO_S16, dir < 0 ? -XO_C16 : XO_C16};   /* w16^3 */
  c32 w9 = {-XO_C16, dir < 0 ? XO_S16 : -XO_S16};  /* w16^9 */
  f[1][1] = cmul(f[1][1], w1);
  f[1][2] = mul_w8_1(f[1][2], dir);
  f[1][3] = cmul(f[1][3], w3);
  f[2][1] = mul_w8_1(f[2][1], dir);
  f[2][2] = mul_w4(f[2][2], dir);
  f[2][3] = mul_w8_3(f[2][3], dir);
  f[3][1] = cmul(f[3][1], w3);
  f[3][2] = mul_w8_3(f[3][2], dir);
  f[3][3] = cmul(f[3][3], w9);
  for (int q = 0; q < 4; q++) {
    c32 a = f[0][q], b = f[1][q], c = f[2][q], d = f[3][q];
    dft4(&a, &b, &c, &d, dir);
    v[q] = a; v[q + 4] = b; v[q + 8] = c; v[q + 12] = d;
  }
}

static inline void dftR(c32 *v, int R, int dir) {
  switch (R) {
    case 2: dft2(v); break;
    case 4: dft4(&v[0], &v[1], &v[2], &v[3], dir); break;
    case 8: dft8(v, dir); break;
    case 16: dft16(v, dir); break;
    default: break;
  }
}

/* Plans set with xo32_set_fft_plan (index log2 n); np 0 = none set. */
static int g_plan_np[13];
static int g_plan_R[13][4];

static int ilog2(int n) {
  int l = 0;
  while ((1 << l) < n) l++;
  return (1 << l) == n ? l : -1;
}

int xo32_set_fft_plan(int n, int np, const int *R) {
  int l = ilog2(n);
  if (l < 0 || l > 12) return -1;
  if (np <= 0) {
    g_plan_np[l] = 0;
    return 0;
  }
  if (np > 4) return -1;
  long prod = 1;
  for (int i = 0; i < np; i++) {
    if (R[i] != 2 && R[i] != 4 && R[i] != 8 && R[i] != 16) return -1;
    prod *= R[i];
  }
  if (prod != n) return -1;
  for (int i = 0; i < np; i++) g_plan_R[l][i] = R[i];
  g_plan_np[l] = np;
  return 0;
}

/* radix plan per length (the HIP kernels' default table), unless one was set */
static int fft_plan(int n, int *R) {
  int l = ilog2(n);
  if (l >= 0 && l <= 12 && g_plan_np[l] > 0) {
    for (int i = 0; i < g_plan_np[l]; i++) R[i] = g_plan_R[l][i];
    return g_plan_np[l];
  }
  switch (n) {
    case 1: return 0;
    case 2: R[0] = 2; return 1;
    case 4: R[0] = 4; return 1;
    case 8: R[0] = 8; return 1;
    case 16: R[0] = 16; return 1;
    case 32: R[0] = 8; R[1] = 4; return 2;
    case 64: R[0] = 8; R[1] = 8; return 2;
    case 128: R[0] = 16; R[1] = 8; return 2;
    case 256: R[0] = 16; R[1] = 16; return 2;
    case 512: R[0] = 8; R[1] = 8; R[2] = 8; return 3;
    case 1024: R[0] = 16; R[1] = 16; R[2] = 4; return 3;
    case 2048: R[0] = 16; R[1] = 16; R[2] = 8; return 3;
    case 4096: R[0] = 16; R[1] = 16; R[2] = 16; return 3;
    default: return -1;
  }
}

/* One Stockham pass: radix R, L = product of the radices already applied. */
static void fft_pass(const c32 *src, c32 *dst, int n, int R, int L, int dir) {
  int nb = n / R;
  int tstep = n / (L * R);
  for (int j = 0; j < nb; j++) {
    int k = j % L;
    c32 v[16];
    for (int r = 0; r < R; r++) v[r] = src[j + r * nb];
    if (L > 1)
      for (int r = 1; r < R; r++) v[r] = cmul(v[r], twf(n, (long)r * k * tstep, dir));
    dftR(v, R, dir);
    for (int r = 0; r < R; r++) dst[(j - k) * R + k + r * L] = v[r];
  }
}

int xo32_fft(c32 *x, int n, int dir) {
  int R[4];
  int np = fft_plan(n, R);
  if (np < 0) return -1;
  if (np == 0) return 0;
  c32 *tmp = (c32 *)malloc(sizeof(c32) * n);
  if (!tmp) return -2;
  c32 *src = x, *dst = tmp;
  int L = 1;
  for (int p = 0; p < np; p++) {
    fft_pass(src, dst, n, R[p], L, dir);
    L *= R[p];
    c32 *t = src; src = dst; dst = t;
  }
  if (src != x) memcpy(x, src, sizeof(c32) * n);
  free(tmp);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* stage (a): range interpolation + amplitude                               */
/* xcorr.cu:237-250, fft_interpolate_routines.cu:76-113, conv_ampl_demean.cu:98-108 */
/* ------------------------------------------------------------------------ */
static inline float ampl32(c32 v) { return sqrtf(fmaf(v.re, v.re, v.im * v.im)); }

/*
 * amp[j], j in [0,n), is |interp[n/2 + j]| where interp is the ri-fold
 * zero-pad interpolation of the row (length M = ri*n, Nyquist bin not split,
 * fft_interpolate_routines.cu:96-97).  Output t = ri*q + r: r == 0 is the
 * original sample x[q]; r >= 1 is (1/n) * IFFT_n(X * phi_r)[q] with
 * phi_r[k] = e^{+2 pi i ktilde r / M}, ktilde = k (k < n/2) or k - n.
 */
int xo32_range_interp_row(const c32 *x, int n, int ri, float *amp) {
  if (ri <= 1) {
    for (int j = 0; j < n; j++) amp[j] = ampl32(x[j]);
    return 0;
  }
  int M = ri * n;
  c32 *X = (c32 *)malloc(sizeof(c32) * n);
  c32 *Y = (c32 *)malloc(sizeof(c32) * n);
  if (!X || !Y) { free(X); free(Y); return -2; }
  memcpy(X, x, sizeof(c32) * n);
  if (xo32_fft(X, n, -1)) { free(X); free(Y); return -1; }
  float scale = 1.0f / (float)n;
  for (int j = 0; j < n; j++) {
    int t = n / 2 + j;
    if (t % ri == 0) amp[j] = ampl32(x[t / ri]);
  }
  for (int r = 1; r < ri; r++) {
    for (int k = 0; k < n; k++) {
      long kt = (k < n / 2) ? k : k - n;
      Y[k] = cmul(X[k], twf(M, kt * r, +1));
    }
    xo32_fft(Y, n, +1);
    for (int j = 0; j < n; j++) {
      int t = n / 2 + j;
      if (t % ri == r) {
        c32 v = Y[t / ri];
        v.re = v.re * scale;
        v.im = v.im * scale;
        amp[j] = ampl32(v);
      }
    }
  }
  free(X);
  free(Y);
  return 0;
}

/* Spectrum of the 1-D mask box [s, n - s) (make_mask, xcorr.cu:385-401):
 * M[k] = sum_{x=s}^{n-s-1} e^{-2 pi i k x/n}, summed in double in increasing x,
 * rounded to float.  M[-k] == conj(M[k]) exactly. */
void xo_mask_spectrum(int n, int s, c32 *m) {
  for (int k = 0; k < n; k++) {
    double re = 0.0, im = 0.0;
    for (int x = s; x < n - s; x++) {
      double c, sn;
      xo_twiddle64(n, (long)k * x, &c, &sn);
      re += c;
      im -= sn;
    }
    m[k].re = (float)re;
    m[k].im = (float)im;
  }
}

/* exact, order-free patch sum: each amplitude truncated to 2^-20 units */
static inline int64_t fxp20(float a) { return (int64_t)(a * 1048576.0f); }

/* ------------------------------------------------------------------------ */
/* one patch                                                                */
/* ------------------------------------------------------------------------ */
static int check_params(const xo_params *p) {
  int R[4];
  if (fft_plan(p->npx, R) < 0 || fft_plan(p->npy, R) < 0) return -1;
  if (p->ri < 1 || (p->ri & (p->ri - 1))) return -1;
  if (p->nxc > p->npx || p->nyc > p->npy) return -1;
  if (p->interp_flag) {
    if (fft_plan(p->n2x, R) < 0 || fft_plan(p->n2y, R) < 0) return -1;
    if (p->interp_factor < 1) return -1;
  }
  return 0;
}

/* exact int64 normalised correlation, do_time_int_xcorr.cu:30-68 / 71-100 */
static double ncc_int(const int32_t *i1, const int32_t *i2, const xo_params *p, int ioff, int joff) {
  int64_t num = 0, d1 = 0, d2 = 0;
  int NX = p->npx;
  for (int r = 0; r < p->ny_corr; r++) {
    const int32_t *ar = i1 + (size_t)(p->ysearch + r + ioff) * NX + p->xsearch + joff;
    const int32_t *br = i2 + (size_t)(p->ysearch + r) * NX + p->xsearch;
    for (int c = 0; c < p->nx_corr; c++) {
      int64_t a = ar[c], b = br[c];
      num += a * b;
      d1 += a * a;
      d2 += b * b;
    }
  }
  double den = sqrt(1.0 * (double)d1 * (double)d2);
  if (den == 0.0) return 0.0;
  return 100.0 * fabs((double)num / den);
}

/* 4th-power (Gatelli) correlation, do_time_int_xcorr.cu:103-144, exact sums */
static double ncc_int4(const int32_t *i1, const int32_t *i2, const xo_params *p, int ioff, int joff) {
  __int128 num = 0, d1 = 0, d2 = 0;
  int NX = p->npx;
  for (int r = 0; r < p->ny_corr; r++) {
    const int32_t *ar = i1 + (size_t)(p->ysearch + r + ioff) * NX + p->xsearch + joff;
    const int32_t *br = i2 + (size_t)(p->ysearch + r) * NX + p->xsearch;
    for (int c = 0; c < p->nx_corr; c++) {
      int64_t a = ar[c], b = br[c];
      __int128 a2 = (__int128)(a * a), b2 = (__int128)(b * b);
      num += a2 * b2;
      d1 += a2 * a2;
      d2 += b2 * b2;
    }
  }
  double den = sqrt((double)d1 * (double)d2);
  double g = (den == 0.0) ? 0.0 : fabs((double)num / den);
  if (g <= 0.5) return 0.0;
  return 100.0 * sqrt(g * 2.0 - 1.0);
}

int xo32_patch(const xo_params *p, const c32 *mwin, const c32 *swin, xo_result *out, xo_dump32 *dump) {
  if (check_params(p)) return -1;
  const int NX = p->npx, NY = p->npy, N = NX * NY;
  const int nxc = p->nxc, nyc = p->nyc;
  int rc = 0;
  double *tsurf = NULL;
  float *amp1 = (float *)malloc(sizeof(float) * N);
  float *amp2 = (float *)malloc(sizeof(float) * N);
  int32_t *i1 = (int32_t *)malloc(sizeof(int32_t) * N);
  int32_t *i2 = (int32_t *)malloc(sizeof(int32_t) * N);
  c32 *z = (c32 *)malloc(sizeof(c32) * N);
  c32 *col = (c32 *)malloc(sizeof(c32) * (NX > NY ? NX : NY));
  c32 *c3 = (c32 *)malloc(sizeof(c32) * N);
  float *surf = (float *)malloc(sizeof(float) * nxc * nyc);
  if (!amp1 || !amp2 || !i1 || !i2 || !z || !col || !c3 || !surf) { rc = -2; goto done; }

  /* (a) range interpolation, amplitude, mean, demean, mask, int cast
   *     xcorr.cu:261-292, conv_ampl_demean.cu:98-139, 703-726 */
  double mean[2];
  for (int img = 0; img < 2; img++) {
    const c32 *w = img ? swin : mwin;
    float *amp = img ? amp2 : amp1;
    int64_t s = 0;
    for (int y = 0; y < NY; y++) {
      if (xo32_range_interp_row(w + (size_t)y * NX, NX, p->ri, amp + (size_t)y * NX)) { rc = -3; goto done; }
      for (int x = 0; x < NX; x++) s += fxp20(amp[(size_t)y * NX + x]);
    }
    mean[img] = (double)s * (1.0 / 1048576.0) / (double)N;
  }
  float m1f = (float)mean[0], m2f = (float)mean[1];
  for (int y = 0; y < NY; y++) {
    int inmask_y = (y >= p->ysearch) && (y < NY - p->ysearch);
    for (int x = 0; x < NX; x++) {
      size_t o = (size_t)y * NX + x;
      float c1 = amp1[o] - m1f;
      float c2 = amp2[o] - m2f;
      int inmask = inmask_y && (x >= p->xsearch) && (x < NX - p->xsearch); /* make_mask xcorr.cu:385-401 */
      if (!inmask) c2 = 0.0f;
      i1[o] = (int32_t)c1;
      i2[o] = (int32_t)c2;
      /* the transform takes the raw amplitudes; the means come off in the
       * spectrum (mask_spectrum / spectral split below).  Every row carries
       * the slave on the template columns: rows outside the mask take the
       * master's spectrum out of the packed transform (below) */
      z[o].re = amp1[o];
      z[o].im = (x >= p->xsearch && x < NX - p->xsearch) ? amp2[o] : 0.0f;
    }
  }

  int ip = 0, jp = 0;
  float cmax = 0.0f;
  double cave = 0.0, gamma = 0.0, msc = 1.0;
  int surf_is_time = 0;

  if (p->corr_flag == 2) {
    /* (b) frequency correlation: do_freq_xcorr.cu:90-200 with c1 + i*c2 packed */
    c32 *mx = (c32 *)malloc(sizeof(c32) * NX);
    c32 *zr = (c32 *)malloc(sizeof(c32) * NX);
    if (!mx || !zr) { free(mx); free(zr); rc = -2; goto done; }
    xo_mask_spectrum(NX, p->xsearch, mx);
    for (int y = 0; y < NY; y++) {
      c32 *row = z + (size_t)y * NX;
      xo32_fft(row, NX, -1);
      /* rows outside the mask: the spectrum of the real part (the master) of
       * the packed row, X[k] = ((Z[k].re + Z[-k].re) / 2, (Z[k].im - Z[-k].im) / 2)
       * (exactly Hermitian) */
      if (!((y >= p->ysearch) && (y < NY - p->ysearch))) {
        memcpy(zr, row, sizeof(c32) * NX);
        for (int kx = 0; kx < NX; kx++) {
          const c32 a = zr[kx], q = zr[(NX - kx) % NX];
          row[kx].re = (a.re + q.re) * 0.5f;
          row[kx].im = (a.im - q.im) * 0.5f;
        }
      }
    }
    free(zr);
    /* demean in the row spectrum: z - m1 - i m2 mask has row transform
     * Zr - m1 NX delta(kx) - i m2 [y in mask] Mx[kx] */
    {
      const float m1NX = m1f * (float)NX;
      for (int y = 0; y < NY; y++) {
        const int ym = (y >= p->ysearch) && (y < NY - p->ysearch);
        for (int kx = 0; kx < NX; kx++) {
          c32 *v = &z[(size_t)y * NX + kx];
          float re = v->re, im = v->im;
          if (kx == 0) re = re - m1NX;
          if (ym) {
            re = fmaf(m2f, mx[kx].im, re);
            im = fmaf(-m2f, mx[kx].re, im);
          }
          v->re = re;
          v->im = im;
        }
      }
    }
    for (int x = 0; x < NX; x++) {
      for (int y = 0; y < NY; y++) col[y] = z[(size_t)y * NX + x];
      xo32_fft(col, NY, -1);
      for (int y = 0; y < NY; y++) z[(size_t)y * NX + x] = col[y];
    }
    for (int ky = 0; ky < NY; ky++) {
      for (int kx = 0; kx < NX; kx++) {
        c32 A = z[(size_t)ky * NX + kx];
        c32 Zm = z[(size_t)((NY - ky) % NY) * NX + (NX - kx) % NX];
        c32 B = {Zm.re, -Zm.im};
        c32 C1 = {(A.re + B.re) * 0.5f, (A.im + B.im) * 0.5f};
        c32 D = csub(A, B);
        c32 C2c = {0.5f * D.im, 0.5f * D.re}; /* conj(C2), C2 = (A-B)/(2i) */
        c32 P = cmul(C1, C2c);
        if ((kx + ky) & 1) { P.re = -P.re; P.im = -P.im; }
        c3[(size_t)ky * NX + kx] = P;
      }
    }
    free(mx);
    const int r0 = p->ny_corr / 2, q0 = p->nx_corr / 2;
    /* C3 is exactly Hermitian (C3[-k] == conj(C3[k]) in this arithmetic), so
     * only columns 0..NX/2 are inverse-transformed along y; the others are
     * their conjugate mirrors */
    for (int x = 0; x <= NX / 2; x++) {
      for (int y = 0; y < NY; y++) col[y] = c3[(size_t)y * NX + x];
      xo32_fft(col, NY, +1);
      for (int y = 0; y < NY; y++) c3[(size_t)y * NX + x] = col[y];
    }
    for (int x = NX / 2 + 1; x < NX; x++)
      for (int y = 0; y < NY; y++) {
        c32 v = c3[(size_t)y * NX + (NX - x)];
        v.im = -v.im;
        c3[(size_t)y * NX + x] = v;
      }
    /* the kept rows are Hermitian, so their inverse transforms are real: rows
     * iy and iy + nyc/2 go through one transform as ra + i rb, and
     * |c3| is |Re| of the result for row iy and |Im| for row iy + nyc/2
     * (d_fillcorrmat's hypot of a real value, do_freq_xcorr.cu:26-29, 67-84) */
    const float sc = 1.0f / (float)N;
    const int hp = nyc / 2;
    c32 *u = (c32 *)malloc(sizeof(c32) * NX);
    if (!u) { rc = -2; goto done; }
    for (int iy = 0; iy < hp; iy++) {
      const c32 *ra = c3 + (size_t)(iy + r0) * NX, *rb = c3 + (size_t)(iy + hp + r0) * NX;
      for (int x = 0; x < NX; x++) {
        u[x].re = ra[x].re - rb[x].im;
        u[x].im = ra[x].im + rb[x].re;
      }
      xo32_fft(u, NX, +1);
      for (int ix = 0; ix < nxc; ix++) {
        c32 v = u[ix + q0];
        surf[(size_t)iy * nxc + ix] = fabsf(v.re * sc);
        surf[(size_t)(iy + hp) * nxc + ix] = fabsf(v.im * sc);
      }
    }
    free(u);
    /* argmax: strict '>' from 0 in row-major order (conv_ampl_demean.cu:434-554, 642-675) */
    int idx = 0;
    float best = 0.0f;
    double sum = 0.0;
    for (int i = 0; i < nxc * nyc; i++) {
      float v = surf[i];
      sum += (double)v;
      if (v > best) { best = v; idx = i; }
    }
    cmax = best;
    cave = sum / (double)(nxc * nyc);
    ip = idx / nxc;
    jp = idx % nxc;
    int ipeak = ip - nyc / 2, jpeak = jp - nxc / 2;
    gamma = ncc_int(i1, i2, p, ipeak, jpeak);   /* do_time_int_xcorr.cu:71-100 */
    msc = (double)cmax / gamma;
    out->xoff = (float)(-jpeak);
    out->yoff = (float)(-ipeak);
    out->corr = (float)gamma;
  } else {
    /* time-domain full search, do_time_int_xcorr.cu:146-189 */
    surf_is_time = 1;
    tsurf = (double *)malloc(sizeof(double) * nxc * nyc);
    if (!tsurf) { rc = -2; goto done; }
    float max_corr = -1.0f;
    int ipk = -9999, jpk = -9999;
    for (int ioff = -p->ysearch; ioff < p->ysearch; ioff++) {
      for (int joff = -p->xsearch; joff < p->xsearch; joff++) {
        int ic = ioff + p->ysearch, jc = joff + p->xsearch;
        double g = (p->corr_flag == 0) ? ncc_int(i1, i2, p, ioff, joff) : ncc_int4(i1, i2, p, ioff, joff);
        tsurf[(size_t)ic * nxc + jc] = g;
        if (fabs(g) > (double)max_corr) {
          max_corr = (float)fabs(g);
          jpk = joff;
          ipk = ioff;
        }
      }
    }
    double sum = 0.0;
    for (int i = 0; i < nxc * nyc; i++) { surf[i] = (float)tsurf[i]; sum += tsurf[i]; }
    cave = sum / (double)(nxc * nyc);
    ip = ipk + nyc / 2;
    jp = jpk + nxc / 2;
    cmax = max_corr;
    out->xoff = (float)(-jpk);
    out->yoff = (float)(-ipk);
    out->corr = max_corr;
  }
  out->ip = ip;
  out->jp = jp;
  out->cmax = cmax;
  out->cave = cave;
  out->xfrac = 0.0f;
  out->yfrac = 0.0f;
  out->hip = -1;
  out->hjp = -1;

  /* (d) high-res refinement: highres_corr.cu:23-79, conv_ampl_demean.cu:75-95,
   *     fft_interpolate_routines.cu:118-147, 1-D passes evaluated directly */
  if (p->interp_flag) {
    const int n2x = p->n2x, n2y = p->n2y, ifc = p->interp_factor;
    const int Lx = n2x * ifc, Ly = n2y * ifc;
    float *md = (float *)malloc(sizeof(float) * n2x * n2y);
    float *hres = (float *)malloc(sizeof(float) * Lx * Ly);
    if (!md || !hres) { free(md); free(hres); rc = -2; goto done; }
    for (int iy = 0; iy < n2y; iy++) {
      for (int ix = 0; ix < n2x; ix++) {
        long k = (long)(ip - n2y / 2 + iy) * nxc + (jp - n2x / 2 + ix);
        float q = 0.0f;
        if (k >= 0 && k < (long)nxc * nyc) {
          float v = surf_is_time ? (float)tsurf[k] : (float)((double)surf[k] / msc);
          q = sqrtf(sqrtf(v));
        }
        md[iy * n2x + ix] = q;
      }
    }
    if (xo32_interp2d(md, n2x, n2y, ifc, hres)) { free(md); free(hres); rc = -2; goto done; }
    int idx = 0;
    float best = 0.0f;
    for (int i = 0; i < Lx * Ly; i++)
      if (hres[i] > best) { best = hres[i]; idx = i; }
    int hp = idx / Lx, hq = idx % Lx;
    float half = (float)Ly / 2.0f;
    out->xfrac = ((float)hq - half) / (float)ifc;
    out->yfrac = ((float)hp - half) / (float)ifc;
    out->hip = hp;
    out->hjp = hq;
    if (dump && dump->md) memcpy(dump->md, md, sizeof(float) * n2x * n2y);
    if (dump && dump->hres) memcpy(dump->hres, hres, sizeof(float) * Lx * Ly);
    free(md); free(hres);
  }

  if (dump) {
    if (dump->amp1) memcpy(dump->amp1, amp1, sizeof(float) * N);
    if (dump->amp2) memcpy(dump->amp2, amp2, sizeof(float) * N);
    if (dump->mean) { dump->mean[0] = mean[0]; dump->mean[1] = mean[1]; }
    if (dump->surf) memcpy(dump->surf, surf, sizeof(float) * nxc * nyc);
  }
done:
  free(amp1); free(amp2); free(i1); free(i2); free(z); free(col); free(c3); free(surf); free(tsurf);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* 2-D zero-pad interpolation of a real n2y x n2x neighbourhood by ifc      */
/* (fft_interpolate_routines.cu:76-147: rows, then columns, Nyquist bin at  */
/* -n/2, each pass * 1/n); returns the real part.  Each 1-D pass is the    */
/* n-term sum sum_k X[k] e^{2 pi i ktilde t / L} / n, evaluated directly.  */
/* The column pass forms only the real part, with the bins k and n - k     */
/* paired: Re(X[k] e^{i k phi} + X[n-k] e^{-i k phi}) = cos(k phi) P_k -     */
/* sin(k phi) Q_k (P_k = X[k].re + X[n-k].re, Q_k = X[k].im - X[n-k].im),   */
/* and rows s and L - s (phi -> -phi) from the same two sums:              */
/* a = X[0].re + sum_k cos P_k + cos_{n/2} X[n/2].re,                      */
/* c = sum_k sin Q_k - sin_{n/2} X[n/2].im (k = 1 .. n/2 - 1, in order),   */
/* v(s) = (a - c) / n and v(L - s) = (a + c) / n, with the factors of row   */
/* s (0 <= s <= L/2).                                                        */
/* ------------------------------------------------------------------------ */
int xo32_interp2d(const float *md, int n2x, int n2y, int ifc, float *hres) {
  const int Lx = n2x * ifc, Ly = n2y * ifc;
  c32 *rows = (c32 *)malloc(sizeof(c32) * n2y * Lx);
  c32 *X = (c32 *)malloc(sizeof(c32) * (n2x > n2y ? n2x : n2y));
  if (!rows || !X) { free(rows); free(X); return -2; }
  const float sx = 1.0f / (float)n2x, sy = 1.0f / (float)n2y;
  for (int iy = 0; iy < n2y; iy++) {
    for (int k = 0; k < n2x; k++) { X[k].re = md[iy * n2x + k]; X[k].im = 0.0f; }
    xo32_fft(X, n2x, -1);
    for (int t = 0; t < Lx; t++) {
      c32 acc = {0.0f, 0.0f};
      for (int k = 0; k < n2x; k++) {
        long kt = (k < n2x / 2) ? k : k - n2x;
        long m = (kt * t) % Lx;
        if (m < 0) m += Lx;
        c32 e = twf(Lx, m, +1);
        acc.re = fmaf(X[k].re, e.re, fmaf(-X[k].im, e.im, acc.re));
        acc.im = fmaf(X[k].re, e.im, fmaf(X[k].im, e.re, acc.im));
      }
      acc.re = acc.re * sx;
      acc.im = acc.im * sx;
      rows[(size_t)iy * Lx + t] = acc;
    }
  }
  const int h = n2y / 2;
  for (int t = 0; t < Lx; t++) {
    for (int k = 0; k < n2y; k++) X[k] = rows[(size_t)k * Lx + t];
    xo32_fft(X, n2y, -1);
    for (int s = 0; s <= Ly / 2; s++) {
      float a = X[0].re, c = 0.0f;
      for (int k = 1; k < h; k++) {
        c32 e = twf(Ly, ((long)k * s) % Ly, +1);
        a = fmaf(e.re, X[k].re + X[n2y - k].re, a);
        c = fmaf(e.im, X[k].im - X[n2y - k].im, c);
      }
      c32 e = twf(Ly, ((long)h * s) % Ly, +1);
      a = fmaf(e.re, X[h].re, a);
      c = fmaf(-e.im, X[h].im, c);
      hres[(size_t)s * Lx + t] = (a - c) * sy;
      if (s > 0 && s < Ly / 2) hres[(size_t)(Ly - s) * Lx + t] = (a + c) * sy;
    }
  }
  free(rows);
  free(X);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* window gather + batch driver (read_xcorr_data.c:52-85, xcorr.cu:330-339) */
/* ------------------------------------------------------------------------ */
void xo_gather_window(int format, const void *img, int nx, int ny, int x0, int y0, int NX, int NY, c32 *w) {
  for (int y = 0; y < NY; y++) {
    int yy = y0 + y;
    for (int x = 0; x < NX; x++) {
      int xx = x0 + x;
      c32 v = {0.0f, 0.0f};
      if (yy >= 0 && yy < ny && xx >= 0 && xx < nx) {
        size_t o = (size_t)yy * nx + xx;
        if (format == 0) {
          const int16_t *s = (const int16_t *)img;
          v.re = (float)s[2 * o];
          v.im = (float)s[2 * o + 1];
        } else {
          v.re = ((const float *)img)[o];
        }
      }
      w[(size_t)y * NX + x] = v;
    }
  }
}

int xo32_run(const xo_params *p, int format, const void *master, int m_nx, int m_ny,
             const void *slave, int s_nx, int s_ny, const int *locs, int n, xo_result *out, int nthreads) {
  if (check_params(p)) return -1;
  int err = 0;
  const int NX = p->npx, NY = p->npy;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    c32 *mw = (c32 *)malloc(sizeof(c32) * NX * NY);
    c32 *sw = (c32 *)malloc(sizeof(c32) * NX * NY);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
    for (int i = 0; i < n; i++) {
      if (!mw || !sw) { err = -2; continue; }
      const int *l = locs + 4 * i;
      xo_gather_window(format, master, m_nx, m_ny, l[0], l[1], NX, NY, mw);
      xo_gather_window(format, slave, s_nx, s_ny, l[2], l[3], NX, NY, sw);
      int rc = xo32_patch(p, mw, sw, out + i, NULL);
      if (rc) err = rc;
    }
    free(mw);
    free(sw);
  }
  return err;
}
