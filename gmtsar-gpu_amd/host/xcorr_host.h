// Host side of the drop-in `xcorr`: command line, PRM files, location grid,
// SLC reading and the output lines, as GMTSAR's xcorr defines them.  Plain C
// ABI so the CLI, tests and other hosts can share it.
#ifndef XCORR_HOST_H
#define XCORR_HOST_H

#include <stddef.h>
#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

// Run configuration: the fields of struct xcorr (xcorr_d.h:23-77) that the
// path uses.
typedef struct xh_config {
  int format;          // 0 complex int16, 1 real float (-real)
  int corr_flag;       // 0 -time, 1 -time4, 2 -freq (default)
  int offset_flag;     // -noshift
  int interp_flag;     // 1 unless -nointerp
  int interp_factor;   // -interp
  int nx_corr, ny_corr;
  int xsearch, ysearch;
  int m_nx, m_ny, s_nx, s_ny;
  int nxc, nyc, npx, npy;
  int nxl, nyl;
  int x_offset, y_offset;   // rshift, ashift of the aligned PRM
  int nlocs, x_inc, y_inc;
  int ri;
  int n2x, n2y;
  int verbose;
  double astretcha;    // (prf2 - prf1) / prf1
  char data1_name[256];
  char data2_name[256];
  char filename[128];
} xh_config;

// The PRM keys xcorr reads (sio_struct.c:132-223).
typedef struct xh_prm {
  char SLC_file[256];
  int num_rng_bins;
  int num_patches;
  int num_valid_az;
  int rshift, ashift;
  double prf;
} xh_prm;

// set_defaults (parse_xcorr_input.c:38-88).
void xh_set_defaults(xh_config *xc);

// parse_command_line (parse_xcorr_input.c:93-183).  Returns 0, or the process
// exit code the reference would end with (1) after printing its message to
// stderr.
int xh_parse_command_line(int argc, char **argv, xh_config *xc, const char *usage);

// get_sio_struct subset (sio_struct.c:132-223): "key = value" lines.
// Returns 0 or -1 when the file cannot be opened.
int xh_read_prm(const char *path, xh_prm *prm);

// handle_prm (parse_xcorr_input.c:185-238): master and aligned PRMs into xc.
// Returns 0 or 1 (die) with the message in err.
int xh_handle_prm(const char *master_prm, const char *slave_prm, xh_config *xc, char *err, size_t errlen);

// get_locations (get_locations.c:14-45): writes nlocs (x, y) pairs; returns nlocs.
int xh_get_locations(xh_config *xc, int32_t *xy, int cap);

// Window corners of location (x, y): xcorr.cu:332-333, read_xcorr_data.c:68-75.
void xh_window(const xh_config *xc, int x, int y, int32_t *mx0, int32_t *my0, int32_t *sx0, int32_t *sy0);

// Output line of print_results (print_results.c:12-46); returns its length.
int xh_format_result(const xh_config *xc, int x, int y, float xoff, float yoff, float xfrac, float yfrac,
                     float corr, char *buf, size_t buflen);

// Read a whole SLC (lines x samples, 4 bytes per sample) into buf; missing
// lines read as zero.  Returns 0 or -1.
int xh_read_slc(const char *path, int nx, int ny, void *buf);

#ifdef __cplusplus
}
#endif
#endif
