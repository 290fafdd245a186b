/*
 * xcorr_hip.h -- C ABI of the MI355X xcorr offset-tracking engine.
 *
 * Plain C types only (no HIP or torch types in the signatures).  A context owns
 * the device scratch, the FFT tables and one HIP stream; it is not thread-safe,
 * several contexts may coexist.  Every entry point returns 0 on success or a
 * negative XCORR_HIP_E* code; xcorr_hip_last_error() gives the message.
 *
 * Reference interface each entry point replaces (diego-romano/gmtsar-gpu):
 *   xcorr_hip_create        d_allocate_arrays + make_mask          xcorr.cu:65-119, 312-318, 385-401
 *   xcorr_hip_set_images    read_xcorr_data + patch H2D copies     read_xcorr_data.c:52-85, xcorr.cu:326-340
 *   xcorr_hip_correlate     per-row stage sequence                 xcorr.cu:342-352 (d_assign_values
 *                           261-292, d_do_freq_corr_strip do_freq_xcorr.cu:123-200,
 *                           d_do_highres_corr highres_corr.cu:23-79) and the CPU time
 *                           search do_time_corr do_time_int_xcorr.cu:146-189 (-time/-time4)
 *   xcorr_hip_destroy       d_free_arrays                          xcorr.cu:122-158
 *   xcorr_hip_last_error    gErrCk / gfftErrCk abort messages      gpu_err.h:13-32
 */
#ifndef XCORR_HIP_H
#define XCORR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XCORR_HIP_ABI_VERSION 1

/* error codes */
#define XCORR_HIP_OK 0
#define XCORR_HIP_EINVAL (-1)   /* bad parameter or size */
#define XCORR_HIP_ENOMEM (-2)   /* device allocation failed */
#define XCORR_HIP_EDEVICE (-3)  /* HIP runtime / kernel error */
#define XCORR_HIP_ESTATE (-4)   /* call out of order (e.g. no images set) */

/* correlation kinds (struct xcorr corr_flag, xcorr_d.h:25) */
#define XCORR_HIP_TIME 0        /* -time: normalised int correlation, full search */
#define XCORR_HIP_TIME4 1       /* -time4: 4th-power (Gatelli) correlation, full search */
#define XCORR_HIP_FREQ 2        /* default: FFT correlation + NCC at the peak */

/* image formats (struct xcorr format, xcorr_d.h:24) */
#define XCORR_HIP_CINT16 0      /* complex int16, I/Q interleaved, row-major lines */
#define XCORR_HIP_FLOAT32 1     /* real float32 (-real) */

/* flags */
#define XCORR_HIP_DEVICE_PTRS 1u   /* image (set_images) or window/result (correlate) pointers are device memory */
#define XCORR_HIP_NO_SYNC 2u       /* correlate: return without waiting for the stream (device results only) */
#define XCORR_HIP_TIMING 4u        /* correlate: record per-kernel HIP events (xcorr_hip_kernel_times) */
#define XCORR_HIP_HOST_WINDOWS 8u  /* correlate with DEVICE_PTRS: windows are host memory, results device memory */

/* Run parameters, named after struct xcorr (xcorr_d.h:23-77).  Sizes follow
 * parse_xcorr_input.c:139-158: nx_corr = nxc = 2*xsearch, npx = 4*xsearch. */
typedef struct xcorr_hip_params {
  int32_t xsearch, ysearch;       /* search half-widths, powers of two in [4, 256] */
  int32_t nx_corr, ny_corr;       /* template size */
  int32_t npx, npy;               /* window size */
  int32_t nxc, nyc;               /* correlation surface size */
  int32_t ri;                     /* range interpolation factor: 1 (-norange) or a power of two <= 16 */
  int32_t interp_flag;            /* 1: high-res sub-pixel refinement, 0: -nointerp */
  int32_t interp_factor;          /* high-res factor, power of two in [1, 32] */
  int32_t n2x, n2y;               /* high-res neighbourhood, 8 x 8 */
  int32_t corr_flag;              /* XCORR_HIP_TIME / TIME4 / FREQ */
  int32_t format;                 /* XCORR_HIP_CINT16 / XCORR_HIP_FLOAT32 */
} xcorr_hip_params;

/* One location: top-left corners of the master and slave windows, in image
 * samples (x = range column, y = azimuth line).  The reference derives them as
 * mx0 = x - npx/2, my0 = y - npy/2, sx0 = mx0 + rshift,
 * sy0 = my0 + ashift + (int)(y * astretcha)   (xcorr.cu:332-333, read_xcorr_data.c:68-75).
 * Samples outside an image read as zero. */
typedef struct xcorr_hip_window {
  int32_t mx0, my0, sx0, sy0;
} xcorr_hip_window;

/* Per-location result (struct locs fields, xcorr_d.h:10-21, plus peak data).
 * The caller composes the reference's output (print_results.c:37-38):
 *   dx = xoff/ri - xfrac/ri + rshift,  dy = yoff - yfrac + ashift + ishft. */
typedef struct xcorr_hip_result {
  double cave;                    /* mean of the correlation surface */
  float xoff, yoff;               /* integer lag offsets (-jpeak, -ipeak) */
  float corr;                     /* normalised correlation at the peak, percent */
  float xfrac, yfrac;             /* high-res refinement, lag units (0 with -nointerp) */
  float cmax;                     /* surface value at the peak */
  int32_t ip, jp;                 /* peak row/col in the nyc x nxc surface */
  int32_t hip, hjp;               /* high-res peak row/col (-1 with -nointerp) */
} xcorr_hip_result;

typedef struct xcorr_hip_ctx xcorr_hip_ctx;

/* Fill the derived sizes of *p from xsearch/ysearch and defaults matching
 * set_defaults (parse_xcorr_input.c:38-88): ri 2, interp 16, n2 8, freq, cint16. */
void xcorr_hip_default_params(xcorr_hip_params *p, int xsearch, int ysearch);

/* Allocate a context for up to max_batch locations per batch (max_batch <= 0
 * picks up to 8192 from the window size and a quarter of the free device
 * memory).  device < 0 uses the current device. */
int xcorr_hip_create(const xcorr_hip_params *p, int max_batch, int device, xcorr_hip_ctx **out);

/* Make the master/slave images available to the context.  With
 * XCORR_HIP_DEVICE_PTRS the pointers are device memory the caller keeps alive;
 * otherwise the images are copied to device memory owned by the context.
 * Samples per line: m_nx / s_nx, lines: m_ny / s_ny. */
int xcorr_hip_set_images(xcorr_hip_ctx *ctx, const void *master, int m_nx, int m_ny,
                         const void *slave, int s_nx, int s_ny, unsigned flags);

/* Correlate n locations (any n >= 0; batches of max_batch internally).
 * win[n] in, out[n] out (host memory unless XCORR_HIP_DEVICE_PTRS; with
 * XCORR_HIP_HOST_WINDOWS as well, windows are host memory and results device
 * memory).  out[i] always belongs to win[i].  For xsearch = ysearch = 64, ri 2,
 * frequency mode, windows that share mx0, sx0 and sy0 - my0 and overlap in
 * lines are computed from shared range-interpolated lines; the windows are then
 * read on the host (device windows are copied back first). */
int xcorr_hip_correlate(xcorr_hip_ctx *ctx, const xcorr_hip_window *win, int n,
                        xcorr_hip_result *out, unsigned flags);

/* Streamed images (pairs larger than device memory, or file reads overlapped
 * with the correlation).  reader(user, image, first, count, dst) copies lines
 * [first, first + count) of image 0 (master, m_nx samples per line) or image 1
 * (slave, s_nx samples per line) into dst: row-major, 4 bytes per sample, in
 * the context's format; lines past the end of a file read as zero.  It returns
 * 0 on success.  Lines are requested in increasing order per band. */
typedef int (*xcorr_hip_line_reader)(void *user, int image, int first, int count, void *dst);

/* Correlate n locations (host windows, host results) of a pair read through
 * `reader`.  Windows are grouped into bands of band_lines lines (<= 0: about
 * 768 MB of lines per image); while one band is correlated the next one is
 * read and copied to the device on a second stream.  A window whose master and
 * slave lines together span more than band_lines lines is rejected.  Replaces
 * the per-row reads read_xcorr_data.c:52-85 and the patch H2D copies
 * xcorr.cu:326-340.  Images set with xcorr_hip_set_images are kept. */
int xcorr_hip_correlate_stream(xcorr_hip_ctx *ctx, xcorr_hip_line_reader reader, void *user,
                               int m_nx, int m_ny, int s_nx, int s_ny, int band_lines,
                               const xcorr_hip_window *win, int n, xcorr_hip_result *out);

/* Wait for all work queued on the context's stream. */
int xcorr_hip_sync(xcorr_hip_ctx *ctx);

/* The hipStream_t (as void*) the context launches on. */
void *xcorr_hip_stream(xcorr_hip_ctx *ctx);

/* Per-kernel device time of the last correlate call made with XCORR_HIP_TIMING:
 * fills names[i] (static strings) and ms[i] (summed over batches) for up to cap
 * kernels; returns the number of kernels. */
int xcorr_hip_kernel_times(xcorr_hip_ctx *ctx, const char **names, float *ms, int *launches, int cap);

/* Copy an internal stage buffer of location `slot` (its index in the last
 * correlate call; it must belong to that call's last batch) to host memory,
 * for stage-by-stage parity tests.  what: 0 amp master (npy*npx f32),
 * 1 amp slave (valid on the template window; elsewhere only in some modes),
 * 2 correlation surface (nyc*nxc f32; f64 in time modes),
 * 3 patch means (2 f64).  Returns bytes copied or a negative code. */
long xcorr_hip_dump(xcorr_hip_ctx *ctx, int what, int slot, void *host, size_t bytes);

void xcorr_hip_destroy(xcorr_hip_ctx *ctx);

/* Message of the last failure on ctx (or of the last failed create when ctx is NULL). */
const char *xcorr_hip_last_error(const xcorr_hip_ctx *ctx);

int xcorr_hip_abi_version(void);

#ifdef __cplusplus
}
#endif
#endif
