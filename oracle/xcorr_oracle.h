/*
 * TEST INFRASTRUCTURE ONLY.  CPU restatement of GMTSAR xcorr's offset-tracking
 * hot path, used as the checker for the HIP implementation.  Nothing shipped
 * links this code; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg load it.
 *
 * Two restatements live here:
 *
 *   xo32_*  FP32 "sequential" path.  Follows the reference stage by stage
 *           (xcorr.cu:237-292, do_freq_xcorr.cu:90-200,
 *           do_time_int_xcorr.cu:30-189, highres_corr.cu:23-79,
 *           fft_interpolate_routines.cu:76-147) with one fixed FP32
 *           arithmetic specification (FFT radix plan, twiddle tables, the
 *           order of every add/mul/fma).  The HIP kernels implement the same
 *           specification, so this path is their bit-exact comparator.
 *
 *   xo64_*  FP64 textbook path.  Same stages written the plain way (zero-pad
 *           interpolation through a full-length FFT, two separate 2-D FFTs,
 *           full inverse, hypot, pow) in double precision with an
 *           independent radix-2 FFT.  It is the tolerance reference for
 *           sub-pixel offsets and correlation values.
 *
 * Parity pins (tests/test_oracle.py): FFTs against numpy.fft and a direct
 * DFT; range interpolation against its closed forms (original samples at
 * multiples of ri, complex exponentials reproduced exactly); the frequency
 * correlation surface against a direct (brute-force) cross-correlation;
 * the normalised correlation against an int64 brute force; high-res
 * interpolation against bandlimited cosines; whole patches against
 * Fourier-shifted synthetic pairs with known offsets; boundary functions
 * (locations, PRM parsing, output lines) against the reference's own host C
 * code built under oracle/_ref.
 */
#ifndef XCORR_ORACLE_H
#define XCORR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { float re, im; } xo_c32;
typedef struct { double re, im; } xo_c64;

/* Per-run parameters, named as in struct xcorr (xcorr_d.h:23-77). */
typedef struct {
  int npx, npy;          /* patch window (4*xsearch, 4*ysearch) */
  int xsearch, ysearch;  /* search half-widths */
  int nx_corr, ny_corr;  /* template size (2*xsearch, 2*ysearch) */
  int nxc, nyc;          /* correlation surface size (2*xsearch, 2*ysearch) */
  int ri;                /* range interpolation factor (1 = none) */
  int interp_flag;       /* high-res refinement on/off */
  int interp_factor;     /* high-res factor */
  int n2x, n2y;          /* high-res neighbourhood (8x8) */
  int corr_flag;         /* 0 time, 1 time4 (Gatelli), 2 frequency */
} xo_params;

/* Per-location result (the fields of struct locs, xcorr_d.h:10-21, plus the
 * intermediate peak data a caller may want to check). */
typedef struct {
  double cave;           /* surface mean */
  float xoff, yoff;      /* integer-lag offsets: -jpeak, -ipeak */
  float corr;            /* normalised correlation at the peak (percent) */
  float xfrac, yfrac;    /* high-res sub-lag refinement */
  float cmax;            /* surface value at the peak */
  int ip, jp;            /* peak row/col in the nyc x nxc surface */
  int hip, hjp;          /* high-res peak row/col */
} xo_result;

/* Optional stage dumps (any pointer may be NULL). */
typedef struct {
  float *amp1, *amp2;    /* npy*npx amplitudes after range interpolation */
  double *mean;          /* [2] patch means */
  float *surf;           /* nyc*nxc correlation surface (freq) or gamma (time) */
  float *md;             /* n2y*n2x corr^(1/4) neighbourhood */
  float *hres;           /* (n2y*ifc)*(n2x*ifc) real part of the refined surface */
} xo_dump32;

typedef struct {
  double *amp1, *amp2;
  double *mean;
  double *surf;
  double *md;
  double *hres;
} xo_dump64;

/* --- spec building blocks (exported for unit tests) --- */
void xo_twiddle64(int n, long m, double *c, double *s);       /* e^{+2 pi i m/n} */
int  xo32_fft(xo_c32 *x, int n, int dir);                      /* in place; dir -1 fwd, +1 inv (unnormalised) */
/* Radix plan of every n-point FP32 transform from now on (R[0..np-1] in
 * {2, 4, 8, 16}, product n, np <= 4); np <= 0 restores the default plan.
 * Follows the HIP engine's tuned plans (xcorr_hip_set_tuning). */
int  xo32_set_fft_plan(int n, int np, const int *R);
int  xo64_fft(xo_c64 *x, int n, int dir);
int  xo32_range_interp_row(const xo_c32 *x, int n, int ri, float *amp);
int  xo64_range_interp_row(const xo_c64 *x, int n, int ri, double *amp);
/* high-res: real part of the ifc-fold 2-D zero-pad interpolation of a real
 * n2y x n2x neighbourhood, (n2y*ifc) x (n2x*ifc) row-major */
int  xo32_interp2d(const float *md, int n2x, int n2y, int ifc, float *hres);
int  xo64_interp2d(const double *md, int n2x, int n2y, int ifc, double *hres);

/* --- one patch: windows are npy*npx complex samples (row-major) --- */
int xo32_patch(const xo_params *p, const xo_c32 *mwin, const xo_c32 *swin,
               xo_result *out, xo_dump32 *dump);
int xo64_patch(const xo_params *p, const xo_c32 *mwin, const xo_c32 *swin,
               xo_result *out, xo_dump64 *dump);

/* --- many patches straight from the SLC images (window gather included) ---
 * format 0: images are complex int16 (I,Q interleaved), format 1: real float.
 * locs holds 4 ints per location: master x0, y0 and slave x0, y0 of the window's
 * top-left corner.  Samples outside an image read as zero.  nthreads <= 0 uses
 * the OpenMP default. */
int xo32_run(const xo_params *p, int format,
             const void *master, int m_nx, int m_ny,
             const void *slave, int s_nx, int s_ny,
             const int *locs, int n, xo_result *out, int nthreads);
int xo64_run(const xo_params *p, int format,
             const void *master, int m_nx, int m_ny,
             const void *slave, int s_nx, int s_ny,
             const int *locs, int n, xo_result *out, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
