// Drop-in `xcorr` for MI355X: same command line, PRM/SLC inputs and output file
// as GMTSAR's xcorr (xcorr.cu:437-490), with the correlation work done by the
// HIP engine (include/xcorr_hip.h).
#include <fcntl.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include <chrono>
#include <vector>

#include "../../include/xcorr_hip.h"
#include "xcorr_host.h"

static const char* USAGE =
    "xcorr [GMT5SAR] - Compute 2-D cross-correlation of two images\n\n"
    "\nUsage: xcorr master.PRM slave.PRM [-time] [-real] [-freq] [-nx n] [-ny n] [-xsearch xs] [-ysearch ys]\n"
    "master.PRM     \tPRM file for reference image\n"
    "slave.PRM     \t \tPRM file of secondary image\n"
    "-time      \t\tuse time cross-correlation\n"
    "-time4      \t\tuse 4th power time cross-correlation\n"
    "-freq      \t\tuse frequency cross-correlation (default)\n"
    "-real      \t\tread float numbers instead of complex numbers\n"
    "-noshift  \t\tignore ashift and rshift in prm file (set to 0)\n"
    "-nx  nx    \t\tnumber of locations in x (range) direction (int)\n"
    "-ny  ny    \t\tnumber of locations in y (azimuth) direction (int)\n"
    "-nointerp     \t\tdo not interpolate correlation function\n"
    "-range_interp ri  \tinterpolate range by ri (power of two) [default: 2]\n"
    "-norange     \t\tdo not range interpolate \n"
    "-xsearch xs\t\tsearch window size in x (range) direction (int power of 2 [32 64 128 256])\n"
    "-ysearch ys\t\tsearch window size in y (azimuth) direction (int power of 2 [32 64 128 256])\n"
    "-interp  factor    \tinterpolate correlation function by factor (int) [default, 16]\n"
    "-v\t\t\tverbose\n"
    "output: \n freq_xcorr.dat (default) \n time_xcorr.dat (if -time option))\n"
    "\nuse fitoffset.csh to convert output to PRM format\n"
    "\nExample:\n"
    "xcorr IMG-HH-ALPSRP075880660-H1.0__A.PRM IMG-HH-ALPSRP129560660-H1.0__A.PRM -nx 20 -ny 50 \n";

// The SLC pair, read line band by line band (read_complex_short2 /
// read_real_float2, read_xcorr_data.c:11-49, read whole lines the same way).
struct SlcPair {
  int fd[2];
  int nx[2];
};

static int slc_reader(void* user, int image, int first, int count, void* dst) {
  const SlcPair* p = (const SlcPair*)user;
  const size_t bytes = (size_t)count * p->nx[image] * 4;
  off_t off = (off_t)first * p->nx[image] * 4;
  size_t got = 0;
  while (got < bytes) {
    const ssize_t r = pread(p->fd[image], (char*)dst + got, bytes - got, off + (off_t)got);
    if (r < 0) return -1;
    if (r == 0) break;                 // past the end of the file: zeros
    got += (size_t)r;
  }
  if (got < bytes) memset((char*)dst + got, 0, bytes - got);
  return 0;
}

int main(int argc, char** argv) {
  xh_config xc;
  if (argc < 3) {
    fprintf(stderr, " %s %s \n", USAGE, "");
    return 1;
  }
  xh_set_defaults(&xc);
  int rc = xh_parse_command_line(argc, argv, &xc, USAGE);
  if (rc) return rc;
  char err[1024];
  if (xh_handle_prm(argv[1], argv[2], &xc, err, sizeof(err))) {
    fputs(err, stderr);
    return 1;
  }
  if (xc.corr_flag == 0) strcpy(xc.filename, "time_xcorr.dat");
  if (xc.corr_flag == 1) strcpy(xc.filename, "time_xcorr_Gatelli.dat");
  if (xc.corr_flag == 2) strcpy(xc.filename, "freq_xcorr.dat");
  FILE* out = fopen(xc.filename, "w");
  if (!out) {
    fprintf(stderr, " %s %s \n", "Can't open output file", xc.filename);
    return 1;
  }
  const int nlocs = xh_get_locations(&xc, nullptr, 0);
  std::vector<int32_t> xy((size_t)2 * (nlocs > 0 ? nlocs : 1));
  xh_get_locations(&xc, xy.data(), nlocs);

  const auto t0 = std::chrono::steady_clock::now();
  xcorr_hip_params p;
  xcorr_hip_default_params(&p, xc.xsearch, xc.ysearch);
  p.ri = xc.ri;
  p.interp_flag = xc.interp_flag;
  p.interp_factor = xc.interp_factor;
  p.corr_flag = xc.corr_flag;
  p.format = xc.format;
  xcorr_hip_ctx* ctx = nullptr;
  rc = xcorr_hip_create(&p, 0, -1, &ctx);
  if (rc) {
    fprintf(stderr, " %s %s \n", "xcorr engine:", xcorr_hip_last_error(nullptr));
    return 1;
  }
  SlcPair pair;
  pair.fd[0] = open(xc.data1_name, O_RDONLY);
  pair.fd[1] = open(xc.data2_name, O_RDONLY);
  pair.nx[0] = xc.m_nx;
  pair.nx[1] = xc.s_nx;
  if (pair.fd[0] < 0 || pair.fd[1] < 0) {
    fprintf(stderr, " %s %s \n", "Cannot open SLC_file", pair.fd[0] < 0 ? xc.data1_name : xc.data2_name);
    return 1;
  }
  std::vector<xcorr_hip_window> win((size_t)(nlocs > 0 ? nlocs : 1));
  for (int i = 0; i < nlocs; i++)
    xh_window(&xc, xy[2 * i], xy[2 * i + 1], &win[i].mx0, &win[i].my0, &win[i].sx0, &win[i].sy0);
  std::vector<xcorr_hip_result> res((size_t)(nlocs > 0 ? nlocs : 1));
  // lines per band (tuning; the default sizes bands from the line length)
  const char* bl = getenv("XCORR_HIP_BAND_LINES");
  const int band_lines = bl ? atoi(bl) : 0;
  rc = xcorr_hip_correlate_stream(ctx, slc_reader, &pair, xc.m_nx, xc.m_ny, xc.s_nx, xc.s_ny, band_lines, win.data(),
                                  nlocs, res.data());
  close(pair.fd[0]);
  close(pair.fd[1]);
  if (rc) {
    fprintf(stderr, " %s %s \n", "xcorr engine:", xcorr_hip_last_error(ctx));
    xcorr_hip_destroy(ctx);
    return 1;
  }
  char line[256];
  for (int i = 0; i < nlocs; i++) {
    const xcorr_hip_result& r = res[i];
    const int x = xy[2 * i], y = xy[2 * i + 1];
    xh_format_result(&xc, x, y, r.xoff, r.yoff, r.xfrac, r.yfrac, r.corr, line, sizeof(line));
    fputs(line, out);
    if (xc.verbose) {
      const int ishft = (int)(y * xc.astretcha);
      const float xo = (r.xoff / (float)xc.ri) - (r.xfrac / (float)xc.ri) + xc.x_offset;
      const float yo = r.yoff - r.yfrac + xc.y_offset + ishft;
      fprintf(stdout, " location %d (%3d,%3d) interpolation (range %d corr %d) correlation %6.2f offset (%6.3f,%6.3f) \n",
              i, x, y, xc.ri, xc.interp_flag ? xc.interp_factor : 1, r.corr, xo, yo);
    }
  }
  fclose(out);
  xcorr_hip_destroy(ctx);
  const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  fprintf(stdout, " elapsed time: %lf \n", el);
  return 0;
}
