/*
 * TEST INFRASTRUCTURE ONLY -- FP64 textbook restatement of the xcorr hot path
 * (see xcorr_oracle.h).  Written the plain way, following the reference's
 * data flow one to one:
 *
 *   range interpolation   FFT_N, bins [0,N/2) -> [0,N/2) and [N/2,N) -> [M-N/2,M)
 *                         of a zeroed length-M array, IFFT_M, * 1/N, keep
 *                         columns [N/2, 3N/2)   (xcorr.cu:237-250,
 *                         fft_interpolate_routines.cu:76-113)
 *   amplitude / mean      hypot, plain double sum        (conv_ampl_demean.cu:98-139, 703-726)
 *   correlation           C1 = FFT2(c1), C2 = FFT2(c2), C3 = C1 conj(C2) (-1)^(kx+ky),
 *                         c3 = IFFT2(C3)/(N M), |c3| on the lag window
 *                         (do_freq_xcorr.cu:67-117)
 *   high-res              corr^(1/4) neighbourhood, 1-D zero-pad interpolation of
 *                         rows then of columns (fft_interpolate_routines.cu:118-147),
 *                         argmax of the real part (conv_ampl_demean.cu:227-346, 730-756)
 *
 * The FFT is an independent iterative radix-2 transform.
 */
#include "xcorr_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef xo_c64 c64;
void xo_gather_window(int format, const void *img, int nx, int ny, int x0, int y0, int NX, int NY, xo_c32 *w);

static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

int xo64_fft(c64 *x, int n, int dir) {
  if (!is_pow2(n)) return -1;
  for (int i = 1, j = 0; i < n; i++) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) { c64 t = x[i]; x[i] = x[j]; x[j] = t; }
  }
  for (int len = 2; len <= n; len <<= 1) {
    int half = len / 2;
    for (int k = 0; k < half; k++) {
      double c, s;
      xo_twiddle64(len, k, &c, &s);
      if (dir < 0) s = -s;
      for (int i = 0; i < n; i += len) {
        c64 u = x[i + k], v = x[i + k + half];
        c64 t = {v.re * c - v.im * s, v.re * s + v.im * c};
        x[i + k].re = u.re + t.re;
        x[i + k].im = u.im + t.im;
        x[i + k + half].re = u.re - t.re;
        x[i + k + half].im = u.im - t.im;
      }
    }
  }
  return 0;
}

/* zero-pad interpolation of one length-n array into length ifactor*n */
static int interp1d(const c64 *in, int n, int ifactor, c64 *outM) {
  int M = ifactor * n;
  c64 *X = (c64 *)malloc(sizeof(c64) * n);
  if (!X) return -2;
  memcpy(X, in, sizeof(c64) * n);
  xo64_fft(X, n, -1);
  memset(outM, 0, sizeof(c64) * M);
  int N2 = n / 2;
  for (int k = 0; k < N2; k++) outM[k] = X[k];
  for (int k = N2; k < n; k++) outM[(2 * ifactor - 1) * N2 + (k - N2)] = X[k];
  xo64_fft(outM, M, +1);
  double nm = (double)ifactor * 2.0 / (2.0 * M);
  for (int t = 0; t < M; t++) { outM[t].re *= nm; outM[t].im *= nm; }
  free(X);
  return 0;
}

int xo64_range_interp_row(const c64 *x, int n, int ri, double *amp) {
  if (ri <= 1) {
    for (int j = 0; j < n; j++) amp[j] = hypot(x[j].re, x[j].im);
    return 0;
  }
  c64 *w = (c64 *)malloc(sizeof(c64) * n * ri);
  if (!w) return -2;
  interp1d(x, n, ri, w);
  for (int j = 0; j < n; j++) amp[j] = hypot(w[n / 2 + j].re, w[n / 2 + j].im);
  free(w);
  return 0;
}

/* 2-D zero-pad interpolation by ifc: 1-D interpolation of the rows, then of
 * the columns of the complex result (fft_interpolate_routines.cu:118-147);
 * returns the real part */
int xo64_interp2d(const double *md, int n2x, int n2y, int ifc, double *hres) {
  const int Lx = n2x * ifc, Ly = n2y * ifc;
  c64 *row = (c64 *)malloc(sizeof(c64) * n2x);
  c64 *t1 = (c64 *)malloc(sizeof(c64) * n2y * Lx);
  c64 *colv = (c64 *)malloc(sizeof(c64) * n2y);
  c64 *colo = (c64 *)malloc(sizeof(c64) * Ly);
  if (!row || !t1 || !colv || !colo) { free(row); free(t1); free(colv); free(colo); return -2; }
  for (int iy = 0; iy < n2y; iy++) {
    for (int k = 0; k < n2x; k++) { row[k].re = md[iy * n2x + k]; row[k].im = 0.0; }
    interp1d(row, n2x, ifc, t1 + (size_t)iy * Lx);
  }
  for (int t = 0; t < Lx; t++) {
    for (int k = 0; k < n2y; k++) colv[k] = t1[(size_t)k * Lx + t];
    interp1d(colv, n2y, ifc, colo);
    for (int s = 0; s < Ly; s++) hres[(size_t)s * Lx + t] = colo[s].re;
  }
  free(row); free(t1); free(colv); free(colo);
  return 0;
}

static void fft2(c64 *a, int NX, int NY, int dir, c64 *col) {
  for (int y = 0; y < NY; y++) xo64_fft(a + (size_t)y * NX, NX, dir);
  for (int x = 0; x < NX; x++) {
    for (int y = 0; y < NY; y++) col[y] = a[(size_t)y * NX + x];
    xo64_fft(col, NY, dir);
    for (int y = 0; y < NY; y++) a[(size_t)y * NX + x] = col[y];
  }
}

static double ncc64(const int32_t *i1, const int32_t *i2, const xo_params *p, int ioff, int joff, int four) {
  long double num = 0, d1 = 0, d2 = 0;
  int NX = p->npx;
  for (int r = 0; r < p->ny_corr; r++)
    for (int c = 0; c < p->nx_corr; c++) {
      long double a = i1[(size_t)(p->ysearch + r + ioff) * NX + p->xsearch + joff + c];
      long double b = i2[(size_t)(p->ysearch + r) * NX + p->xsearch + c];
      if (four) { num += a * a * b * b; d1 += a * a * a * a; d2 += b * b * b * b; }
      else { num += a * b; d1 += a * a; d2 += b * b; }
    }
  long double den = sqrtl(d1 * d2);
  if (den == 0) return 0.0;
  double g = (double)fabsl(num / den);
  if (!four) return 100.0 * g;
  if (g <= 0.5) return 0.0;
  return 100.0 * sqrt(2.0 * g - 1.0);
}

int xo64_patch(const xo_params *p, const xo_c32 *mwin, const xo_c32 *swin, xo_result *out, xo_dump64 *dump) {
  const int NX = p->npx, NY = p->npy, N = NX * NY;
  const int nxc = p->nxc, nyc = p->nyc;
  if (!is_pow2(NX) || !is_pow2(NY) || p->ri < 1) return -1;
  int rc = 0;
  double *amp[2] = {(double *)malloc(sizeof(double) * N), (double *)malloc(sizeof(double) * N)};
  c64 *c[2] = {(c64 *)malloc(sizeof(c64) * N), (c64 *)malloc(sizeof(c64) * N)};
  int32_t *iv[2] = {(int32_t *)malloc(sizeof(int32_t) * N), (int32_t *)malloc(sizeof(int32_t) * N)};
  c64 *row = (c64 *)malloc(sizeof(c64) * (NX > NY ? NX : NY));
  double *surf = (double *)malloc(sizeof(double) * nxc * nyc);
  if (!amp[0] || !amp[1] || !c[0] || !c[1] || !iv[0] || !iv[1] || !row || !surf) { rc = -2; goto done; }

  double mean[2];
  for (int img = 0; img < 2; img++) {
    const xo_c32 *w = img ? swin : mwin;
    double s = 0.0;
    for (int y = 0; y < NY; y++) {
      for (int x = 0; x < NX; x++) { row[x].re = w[(size_t)y * NX + x].re; row[x].im = w[(size_t)y * NX + x].im; }
      xo64_range_interp_row(row, NX, p->ri, amp[img] + (size_t)y * NX);
      for (int x = 0; x < NX; x++) s += amp[img][(size_t)y * NX + x];
    }
    mean[img] = s / (double)N;
    for (int y = 0; y < NY; y++)
      for (int x = 0; x < NX; x++) {
        size_t o = (size_t)y * NX + x;
        double v = amp[img][o] - mean[img];
        if (img == 1) {
          int in = y >= p->ysearch && y < NY - p->ysearch && x >= p->xsearch && x < NX - p->xsearch;
          if (!in) v = 0.0;
        }
        c[img][o].re = v;
        c[img][o].im = 0.0;
        iv[img][o] = (int32_t)v;
      }
  }

  int ip, jp;
  double cmax, cave, gamma, msc = 1.0;
  if (p->corr_flag == 2) {
    fft2(c[0], NX, NY, -1, row);
    fft2(c[1], NX, NY, -1, row);
    for (int ky = 0; ky < NY; ky++)
      for (int kx = 0; kx < NX; kx++) {
        size_t o = (size_t)ky * NX + kx;
        c64 a = c[0][o], b = c[1][o];
        c64 pr = {a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im};
        if ((kx + ky) & 1) { pr.re = -pr.re; pr.im = -pr.im; }
        c[0][o] = pr;
      }
    fft2(c[0], NX, NY, +1, row);
    const int r0 = p->ny_corr / 2, q0 = p->nx_corr / 2;
    for (int iy = 0; iy < nyc; iy++)
      for (int ix = 0; ix < nxc; ix++) {
        c64 v = c[0][(size_t)(iy + r0) * NX + ix + q0];
        surf[(size_t)iy * nxc + ix] = hypot(v.re / N, v.im / N);
      }
    int idx = 0;
    double best = 0.0, sum = 0.0;
    for (int i = 0; i < nxc * nyc; i++) {
      sum += surf[i];
      if (surf[i] > best) { best = surf[i]; idx = i; }
    }
    cmax = best;
    cave = sum / (nxc * nyc);
    ip = idx / nxc;
    jp = idx % nxc;
    gamma = ncc64(iv[0], iv[1], p, ip - nyc / 2, jp - nxc / 2, 0);
    msc = cmax / gamma;
    if (dump && dump->surf) memcpy(dump->surf, surf, sizeof(double) * nxc * nyc);
    for (int i = 0; i < nxc * nyc; i++) surf[i] /= msc;
    out->xoff = (float)(-(jp - nxc / 2));
    out->yoff = (float)(-(ip - nyc / 2));
    out->corr = (float)gamma;
  } else {
    double best = -1.0, sum = 0.0;
    int ipk = 0, jpk = 0;
    for (int ic = 0; ic < nyc; ic++)
      for (int jc = 0; jc < nxc; jc++) {
        double g = ncc64(iv[0], iv[1], p, ic - p->ysearch, jc - p->xsearch, p->corr_flag == 1);
        surf[(size_t)ic * nxc + jc] = g;
        sum += g;
        if (fabs(g) > best) { best = fabs(g); ipk = ic; jpk = jc; }
      }
    ip = ipk; jp = jpk; cmax = best; cave = sum / (nxc * nyc);
    if (dump && dump->surf) memcpy(dump->surf, surf, sizeof(double) * nxc * nyc);
    out->xoff = (float)(-(jp - nxc / 2));
    out->yoff = (float)(-(ip - nyc / 2));
    out->corr = (float)best;
  }
  out->ip = ip;
  out->jp = jp;
  out->cmax = (float)cmax;
  out->cave = cave;
  out->xfrac = out->yfrac = 0.0f;
  out->hip = out->hjp = -1;

  if (p->interp_flag) {
    const int n2x = p->n2x, n2y = p->n2y, ifc = p->interp_factor;
    const int Lx = n2x * ifc, Ly = n2y * ifc;
    double *md = (double *)calloc((size_t)n2x * n2y, sizeof(double));
    double *hres = (double *)malloc(sizeof(double) * Lx * Ly);
    if (!md || !hres) { free(md); free(hres); rc = -2; goto done; }
    for (int iy = 0; iy < n2y; iy++)
      for (int ix = 0; ix < n2x; ix++) {
        long k = (long)(ip - n2y / 2 + iy) * nxc + (jp - n2x / 2 + ix);
        if (k >= 0 && k < (long)nxc * nyc) md[iy * n2x + ix] = pow(surf[k], 0.25);
      }
    xo64_interp2d(md, n2x, n2y, ifc, hres);
    int idx = 0;
    double best = 0.0;
    for (int i = 0; i < Lx * Ly; i++)
      if (hres[i] > best) { best = hres[i]; idx = i; }
    int hp = idx / Lx, hq = idx % Lx;
    out->xfrac = (float)((hq - Ly / 2.0) / ifc);
    out->yfrac = (float)((hp - Ly / 2.0) / ifc);
    out->hip = hp;
    out->hjp = hq;
    if (dump && dump->md) memcpy(dump->md, md, sizeof(double) * n2x * n2y);
    if (dump && dump->hres) memcpy(dump->hres, hres, sizeof(double) * Lx * Ly);
    free(md); free(hres);
  }
  if (dump) {
    if (dump->amp1) memcpy(dump->amp1, amp[0], sizeof(double) * N);
    if (dump->amp2) memcpy(dump->amp2, amp[1], sizeof(double) * N);
    if (dump->mean) { dump->mean[0] = mean[0]; dump->mean[1] = mean[1]; }
  }
done:
  free(amp[0]); free(amp[1]); free(c[0]); free(c[1]); free(iv[0]); free(iv[1]); free(row); free(surf);
  return rc;
}

int xo64_run(const xo_params *p, int format, const void *master, int m_nx, int m_ny,
             const void *slave, int s_nx, int s_ny, const int *locs, int n, xo_result *out, int nthreads) {
  int err = 0;
  const int NX = p->npx, NY = p->npy;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    xo_c32 *mw = (xo_c32 *)malloc(sizeof(xo_c32) * NX * NY);
    xo_c32 *sw = (xo_c32 *)malloc(sizeof(xo_c32) * NX * NY);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
    for (int i = 0; i < n; i++) {
      if (!mw || !sw) { err = -2; continue; }
      const int *l = locs + 4 * i;
      xo_gather_window(format, master, m_nx, m_ny, l[0], l[1], NX, NY, mw);
      xo_gather_window(format, slave, s_nx, s_ny, l[2], l[3], NX, NY, sw);
      int rc = xo64_patch(p, mw, sw, out + i, NULL);
      if (rc) err = rc;
    }
    free(mw);
    free(sw);
  }
  return err;
}
