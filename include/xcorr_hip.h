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
 * Tuning and multi-pair entry points have no counterpart in the reference
 * (each says so below).
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
#define XCORR_HIP_SERIAL 64u       /* correlate: strip batches one after another on the context's stream (no second
                                      context; kernel durations then are those of a kernel alone on the device) */
#define XCORR_HIP_KEEP_SURFACE 128u  /* correlate: write every correlation surface whole, for xcorr_hip_dump (by
                                        default the kernels keep only each peak's 8 x 8 neighbourhood) */

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
 * read on the host (device windows are copied back first).  On those shapes a
 * call with more than one batch runs alternate batches on a second context
 * the first such call allocates (the same buffers again: the two batches
 * overlap on the device); its work is ordered on xcorr_hip_stream(ctx) like
 * the rest.  If that allocation fails the batches run one after another. */
int xcorr_hip_correlate(xcorr_hip_ctx *ctx, const xcorr_hip_window *win, int n,
                        xcorr_hip_result *out, unsigned flags);

/* Bytes of the static device array xcorr_hip_correlate_noalloc works in (one
 * per GPU the library runs on, held until the process ends). */
#define XCORR_HIP_NOALLOC_SCRATCH (64u << 20)

/* Correlate n locations without allocating anything.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory while they
 * run -- code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  The results equal xcorr_hip_correlate's bit for bit.
 * Windows (d_win) and results (d_out) are device memory; the images must be
 * set.  Every per-batch buffer (amplitudes, spectra, lag rows, surfaces,
 * sums, peaks) is carved from one static device array of
 * XCORR_HIP_NOALLOC_SCRATCH bytes per GPU; the call loops over internal
 * batches of as many windows as fit in it (256 x 256 windows: 48), using the
 * per-patch kernels (no strip sharing, no second context, no host staging).
 * A window shape whose buffers exceed the array is rejected with
 * XCORR_HIP_EINVAL.  No hipMalloc, hipFree, memset node or allocating routine
 * is called, and with XCORR_HIP_NO_SYNC (the only accepted flag) nothing waits:
 * the kernels are queued on xcorr_hip_stream(ctx).  All calls on one GPU share
 * its array, which keeps the last call's contents: calls that overlap in time
 * on one GPU (other contexts, streams or threads) race and must be serialised
 * by the caller. */
int xcorr_hip_correlate_noalloc(xcorr_hip_ctx *ctx, const xcorr_hip_window *d_win, int n,
                                xcorr_hip_result *d_out, unsigned flags);

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

/* Launch settings of a context.  batch, rows_per_block, cols_per_block and
 * threads never change a result bit (every sum they touch is exact or keeps
 * its order).  A radix plan other than the default (npx/npy = 128: 16 x 8,
 * 256: 16 x 16, 512: 8 x 8 x 8, ...) is a different FP32 rounding of the
 * transforms of that length; the CPU restatement accepts the same plan.
 * Radix plans apply to the generic LDS kernels.  Square 128 / 256 / 512
 * windows (ri 2, frequency mode; -time on 128 x 128) run shared strip lines
 * with the default plans; setting another plan on such a context switches it
 * to the per-patch kernels (their buffers are allocated at that call).
 * No counterpart in the reference (cuFFT picks its own plans, xcorr.cu:161-230
 * sizes batches but is unused). */
typedef struct xcorr_hip_tuning {
  int32_t batch;            /* locations per launch, 1..max_batch */
  int32_t rows_per_block;   /* generic row kernels: window rows per block (LDS tile), power of two */
  int32_t cols_per_block;   /* generic column kernels: columns per block (LDS tile), power of two */
  int32_t threads;          /* generic row / column kernels: threads per block, 64..512, multiple of 64 */
  int32_t radix_x[4];       /* radix plan (2/4/8/16 per pass, 0-terminated) of the npx-point transforms */
  int32_t radix_y[4];       /* the same for npy (equal lengths share one plan) */
  float ms_per_batch;       /* autotune: measured time of one batch with these settings */
  int32_t candidates;       /* autotune: settings timed */
} xcorr_hip_tuning;

#define XCORR_HIP_TUNE 16u        /* multi_create: autotune each pair class */
#define XCORR_HIP_TUNE_RADIX 32u  /* autotune / multi_create: also try radix plans (changes FP32 rounding) */

int xcorr_hip_get_tuning(xcorr_hip_ctx *ctx, xcorr_hip_tuning *t);
int xcorr_hip_set_tuning(xcorr_hip_ctx *ctx, const xcorr_hip_tuning *t);

/* Time candidate settings on up to 2*max_batch of the given host windows (the
 * images must be set) by coordinate descent -- threads, row tile, column
 * tile, radix plans (with XCORR_HIP_TUNE_RADIX), batch -- keep the fastest,
 * apply it and report it in *out (may be NULL). */
int xcorr_hip_autotune(xcorr_hip_ctx *ctx, const xcorr_hip_window *win, int n, unsigned flags,
                       xcorr_hip_tuning *out);

/* Multi-pair workloads: many SLC pairs, each with its own parameters (patch
 * size, interpolation factor, time or frequency correlation), images and
 * windows, run as ONE HIP graph launch.  Pairs with equal parameters form a
 * class sharing one context; classes are concurrent branches of the graph.
 * Everything a run needs (strip plans, device windows, tuned settings) is
 * prepared at create, so a run does no host work.  Images and results are
 * device memory the caller keeps alive; windows are host memory read at
 * create.  No counterpart in the reference: it replaces running the
 * per-row loop of do_correlation (xcorr.cu:299-378) once per pair. */
typedef struct xcorr_hip_pair {
  xcorr_hip_params params;
  const void *master;       /* device memory, m_ny lines of m_nx samples */
  int32_t m_nx, m_ny;
  const void *slave;        /* device memory, s_ny lines of s_nx samples */
  int32_t s_nx, s_ny;
  const xcorr_hip_window *win;   /* host memory, n windows */
  int32_t n;
  xcorr_hip_result *out;    /* device memory, n results (out[i] belongs to win[i]) */
} xcorr_hip_pair;

typedef struct xcorr_hip_multi xcorr_hip_multi;

/* flags: XCORR_HIP_TUNE, XCORR_HIP_TUNE_RADIX. */
int xcorr_hip_multi_create(const xcorr_hip_pair *pairs, int npairs, int device, unsigned flags,
                           xcorr_hip_multi **out);
/* One graph launch over every pair; XCORR_HIP_NO_SYNC returns at once. */
int xcorr_hip_multi_run(xcorr_hip_multi *m, unsigned flags);
int xcorr_hip_multi_sync(xcorr_hip_multi *m);
void *xcorr_hip_multi_stream(xcorr_hip_multi *m);
/* Classes of the workload: params[k], tuning[k] (k < cap) and pair_class[i]
 * for every pair (any pointer may be NULL).  Returns the number of classes. */
int xcorr_hip_multi_classes(xcorr_hip_multi *m, xcorr_hip_params *params, xcorr_hip_tuning *tuning,
                            int *pair_class, int cap);
void xcorr_hip_multi_destroy(xcorr_hip_multi *m);
const char *xcorr_hip_multi_last_error(const xcorr_hip_multi *m);

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
 * 2 correlation surface (nyc*nxc f32, only after a call with XCORR_HIP_KEEP_SURFACE; f64 in time modes),
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
