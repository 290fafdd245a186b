/*
 * TEST INFRASTRUCTURE ONLY.  Driver around the reference's own host C library
 * (parse_xcorr_input.c, sio_struct.c, get_locations.c, read_xcorr_data.c,
 * print_results.c, utils.c compiled from /root/reference into oracle/_ref/ by
 * the Makefile's `ref` target).  Tests use it to pin the boundary functions of
 * the drop-in xcorr host: option/PRM handling, the location grid, the window
 * reads and the output line format.
 *
 *   ref_host locs  master.PRM slave.PRM [xcorr options]
 *       prints "P m_nx m_ny s_nx s_ny x_offset y_offset astretcha npx npy nxc nyc
 *       nx_corr ny_corr xsearch ysearch ri interp_flag interp_factor corr_flag
 *       nxl nyl format" then "L x y" per location
 *   ref_host read  master.PRM slave.PRM iloc [xcorr options]
 *       prints the npy x npx master and slave windows of location iloc as
 *       "M y x re im" / "S y x re im" lines (xcorr.cu:330-337 window placement)
 *   ref_host print ri interp_flag interp_factor x_offset y_offset astretcha
 *       reads "x y xoff yoff xfrac yfrac corr" lines on stdin and writes
 *       print_results() lines to stdout
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gmtsar_d.h"
#include "xcorr.h"

void set_defaults(struct xcorr *);
void parse_command_line(int, char **, struct xcorr *, int *, int *, const char *);
void handle_prm(char **, struct xcorr *, int);
void get_locations(struct xcorr *);
void read_xcorr_data(struct xcorr, int);
void print_results(struct xcorr, int);

static const char *USAGE_TXT = "ref_host usage\n";

static void setup(struct xcorr *xc, int argc, char **argv) {
  int nfiles = 2, input_flag = 0;
  memset(xc, 0, sizeof(*xc));
  xc->interp_flag = 0;
  xc->corr_flag = 2;
  xc->offset_flag = 0;
  set_defaults(xc);
  parse_command_line(argc, argv, xc, &nfiles, &input_flag, USAGE_TXT);
  if (input_flag == 0) handle_prm(argv, xc, nfiles);
  get_locations(xc);
}

int main(int argc, char **argv) {
  struct xcorr xc;
  verbose = 0;
  debug = 0;
  if (argc < 2) return 2;
  if (!strcmp(argv[1], "locs") && argc >= 4) {
    setup(&xc, argc - 1, argv + 1);
    printf("P %d %d %d %d %d %d %.17g %d %d %d %d %d %d %d %d %d %d %d %d %d %d %d\n", xc.m_nx, xc.m_ny, xc.s_nx,
           xc.s_ny, xc.x_offset, xc.y_offset, xc.astretcha, xc.npx, xc.npy, xc.nxc, xc.nyc, xc.nx_corr,
           xc.ny_corr, xc.xsearch, xc.ysearch, xc.ri, xc.interp_flag, xc.interp_factor, xc.corr_flag, xc.nxl,
           xc.nyl, xc.format);
    for (int i = 0; i < xc.nlocs; i++) printf("L %d %d\n", xc.loc[i].x, xc.loc[i].y);
    return 0;
  }
  if (!strcmp(argv[1], "read") && argc >= 5) {
    int iloc = atoi(argv[4]);
    /* shift the iloc argument out so parse_command_line sees the options */
    char **av = (char **)malloc(sizeof(char *) * argc);
    av[0] = argv[1];
    av[1] = argv[2];
    av[2] = argv[3];
    for (int i = 5; i < argc; i++) av[i - 2] = argv[i];
    setup(&xc, argc - 2, av);
    xc.d1 = (struct FCOMPLEX *)calloc((size_t)xc.m_nx * xc.npy, sizeof(struct FCOMPLEX));
    xc.d2 = (struct FCOMPLEX *)calloc((size_t)xc.s_nx * xc.npy, sizeof(struct FCOMPLEX));
    read_xcorr_data(xc, iloc);
    int mx = xc.loc[iloc].x - xc.npx / 2;
    int sx = xc.loc[iloc].x + xc.x_offset - xc.npx / 2;
    for (int y = 0; y < xc.npy; y++)
      for (int x = 0; x < xc.npx; x++) {
        struct FCOMPLEX a = xc.d1[(size_t)y * xc.m_nx + mx + x];
        struct FCOMPLEX b = xc.d2[(size_t)y * xc.s_nx + sx + x];
        printf("M %d %d %.9g %.9g\nS %d %d %.9g %.9g\n", y, x, a.r, a.i, y, x, b.r, b.i);
      }
    return 0;
  }
  if (!strcmp(argv[1], "print") && argc >= 8) {
    memset(&xc, 0, sizeof(xc));
    xc.ri = atoi(argv[2]);
    xc.interp_flag = atoi(argv[3]);
    xc.interp_factor = atoi(argv[4]);
    xc.x_offset = atoi(argv[5]);
    xc.y_offset = atoi(argv[6]);
    xc.astretcha = atof(argv[7]);
    xc.file = stdout;
    xc.loc = (struct locs *)calloc(1, sizeof(struct locs));
    int x, y;
    float xoff, yoff, xfrac, yfrac, corr;
    while (scanf("%d %d %f %f %f %f %f", &x, &y, &xoff, &yoff, &xfrac, &yfrac, &corr) == 7) {
      xc.loc[0].x = x;
      xc.loc[0].y = y;
      xc.loc[0].xoff = xoff;
      xc.loc[0].yoff = yoff;
      xc.loc[0].xfrac = xfrac;
      xc.loc[0].yfrac = yfrac;
      xc.loc[0].corr = corr;
      print_results(xc, 0);
    }
    return 0;
  }
  return 2;
}
