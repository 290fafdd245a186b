// Host side of the drop-in xcorr: see xcorr_host.h.  Messages and formats follow
// the reference so that scripts reading xcorr's output keep working.
#include "xcorr_host.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <string>

extern "C" void xh_set_defaults(xh_config* xc) {
  // parse_xcorr_input.c:38-88 (both corr_flag branches set the same values);
  // xcorr.cu:448-449 sets corr_flag = 2 before it.
  memset(xc, 0, sizeof(*xc));
  xc->corr_flag = 2;
  xc->format = 0;
  xc->ri = 2;
  xc->nxl = 16;
  xc->nyl = 32;
  xc->x_offset = 0;
  xc->y_offset = 0;
  xc->nx_corr = 128;
  xc->ny_corr = 128;
  xc->xsearch = 64;
  xc->ysearch = 64;
  xc->n2x = 8;
  xc->n2y = 8;
  xc->npx = xc->nx_corr + 2 * xc->xsearch;
  xc->npy = xc->ny_corr + 2 * xc->ysearch;
  xc->nxc = 2 * xc->xsearch;
  xc->nyc = 2 * xc->ysearch;
  xc->astretcha = 0.0;
  xc->interp_flag = 1;
  xc->interp_factor = 16;
  xc->offset_flag = 0;
  strcpy(xc->filename, "freq_xcorr.dat");
}

static int die(const char* s1, const char* s2) {
  // utils.c:39-43
  fprintf(stderr, " %s %s \n", s1, s2);
  return 1;
}

extern "C" int xh_parse_command_line(int na, char** a, xh_config* xc, const char* usage) {
  // parse_xcorr_input.c:93-183
  for (int n = 3; n < na; n++) {
    if (!strcmp(a[n], "-freq")) {
      xc->corr_flag = 2;
      fprintf(stderr, " using frequency cross correlation\n");
    } else if (!strcmp(a[n], "-time")) {
      xc->corr_flag = 0;
      fprintf(stderr, " using time cross correlation\n");
    } else if (!strcmp(a[n], "-time4")) {
      xc->corr_flag = 1;
      fprintf(stderr, " using 4th power time cross correlation\n");
    } else if (!strcmp(a[n], "-real")) {
      xc->format = 1;
      fprintf(stderr, " no interpolation\n");
    } else if (!strcmp(a[n], "-nointerp")) {
      xc->interp_flag = 0;
      fprintf(stderr, " no interpolation\n");
    } else if (!strcmp(a[n], "-noshift")) {
      xc->offset_flag = 1;
      fprintf(stderr, " ignoring shift in prm\n");
    } else if (!strcmp(a[n], "-interp")) {
      if (++n == na) return die(" no option after -interp!\n", "");
      xc->interp_flag = 1;
      xc->interp_factor = atoi(a[n]);
      fprintf(stderr, " setting interpolation factor to %d\n", xc->interp_factor);
    } else if (!strcmp(a[n], "-range_interp")) {
      if (++n == na) return die(" no option after -range_interp!\n", "");
      xc->ri = atoi(a[n]);
      fprintf(stderr, " setting range interpolation factor to %d\n", xc->ri);
    } else if (!strcmp(a[n], "-nx")) {
      if (++n == na) return die(" no option after -nx!\n", "");
      xc->nxl = atoi(a[n]);
      fprintf(stderr, " setting nx to %d\n", xc->nxl);
    } else if (!strcmp(a[n], "-ny")) {
      if (++n == na) return die(" no option after -ny!\n", "");
      xc->nyl = atoi(a[n]);
      fprintf(stderr, " setting ny to %d\n", xc->nyl);
    } else if (!strcmp(a[n], "-xsearch")) {
      if (++n == na) return die(" no option after -xsearch!\n", "");
      xc->xsearch = atoi(a[n]);
      xc->nx_corr = 2 * xc->xsearch;
      xc->npx = xc->nx_corr + 2 * xc->xsearch;
      xc->nxc = 2 * xc->xsearch;
      if ((xc->xsearch - 1) & xc->xsearch) return die(" xsearch needs to be power of 2! (32 64 128 256) \n", "");
      fprintf(stderr, " setting xsearch to %d\n", xc->xsearch);
      fprintf(stderr, " setting nx_corr to %d\n", xc->nx_corr);
    } else if (!strcmp(a[n], "-ysearch")) {
      if (++n == na) return die(" no option after -ysearch!\n", "");
      xc->ysearch = atoi(a[n]);
      xc->ny_corr = 2 * xc->ysearch;
      xc->npy = xc->ny_corr + 2 * xc->ysearch;
      xc->nyc = 2 * xc->ysearch;
      if ((xc->ysearch - 1) & xc->ysearch) return die(" ysearch needs to be power of 2! (32 64 128 256) \n", "");
      fprintf(stderr, " setting ysearch to %d\n", xc->ysearch);
      fprintf(stderr, " setting ny_corr to %d\n", xc->ny_corr);
    } else if (!strcmp(a[n], "-v")) {
      xc->verbose = 1;
      fprintf(stderr, " verbose output \n");
    } else if (!strcmp(a[n], "-norange")) {
      xc->ri = 1;
      fprintf(stderr, " no range interpolation \n");
    } else {
      // The reference routes every other "-..." word to its unfinished -input
      // branch (parse_xcorr_input.c:165-174), which skips the PRM files; this
      // host reports it like any unknown word instead.
      fprintf(stderr, " %s *** option not recognized ***\n\n", a[n]);
      fprintf(stderr, " %s someone made a mistake!\n\n", a[n]);
      fprintf(stderr, " %s I think it was you! \n\n", a[n]);
      fprintf(stderr, "%s", usage ? usage : "");
      return 1;
    }
  }
  return 0;
}

static std::string trim(const std::string& s) {
  size_t a = s.find_first_not_of(" \t\r\n");
  if (a == std::string::npos) return "";
  size_t b = s.find_last_not_of(" \t\r\n");
  return s.substr(a, b - a + 1);
}

extern "C" int xh_read_prm(const char* path, xh_prm* prm) {
  // sio_struct.c:132-223, keys used by xcorr; xshift/yshift are the old names
  // of rshift/ashift (sio_struct.c:173-175).
  memset(prm, 0, sizeof(*prm));
  FILE* f = fopen(path, "r");
  if (!f) return -1;
  char line[1024];
  while (fgets(line, sizeof(line), f)) {
    std::string s(line);
    size_t eq = s.find('=');
    if (eq == std::string::npos) continue;
    std::string name = trim(s.substr(0, eq));
    std::string value = trim(s.substr(eq + 1));
    size_t sp = value.find_first_of(" \t");
    if (sp != std::string::npos) value = value.substr(0, sp);
    if (name.find_first_of(" \t") != std::string::npos) name = name.substr(name.find_last_of(" \t") + 1);
    if (name == "SLC_file") {
      snprintf(prm->SLC_file, sizeof(prm->SLC_file), "%s", value.c_str());
    } else if (name == "num_rng_bins") {
      prm->num_rng_bins = atoi(value.c_str());
    } else if (name == "num_patches") {
      prm->num_patches = atoi(value.c_str());
    } else if (name == "num_valid_az") {
      prm->num_valid_az = atoi(value.c_str());
    } else if (name == "rshift" || name == "xshift") {
      prm->rshift = atoi(value.c_str());
    } else if (name == "ashift" || name == "yshift") {
      prm->ashift = atoi(value.c_str());
    } else if (name == "PRF") {
      prm->prf = atof(value.c_str());
    }
  }
  fclose(f);
  return 0;
}

extern "C" int xh_handle_prm(const char* mprm, const char* sprm, xh_config* xc, char* err, size_t errlen) {
  // parse_xcorr_input.c:185-238
  xh_prm r[2];
  const char* names[2] = {mprm, sprm};
  for (int i = 0; i < 2; i++) {
    if (xh_read_prm(names[i], &r[i])) {
      snprintf(err, errlen, " %s %s \n", "Can't open prmfile ", names[i]);
      return 1;
    }
    if (i == 0) {
      snprintf(xc->data1_name, sizeof(xc->data1_name), "%s", r[0].SLC_file);
      FILE* f = fopen(xc->data1_name, "r");
      if (!f) {
        snprintf(err, errlen, " %s %s \n", "Cannot open SLC_file", xc->data1_name);
        return 1;
      }
      fclose(f);
      xc->m_nx = r[0].num_rng_bins;
      xc->m_ny = r[0].num_patches * r[0].num_valid_az;
    } else {
      snprintf(xc->data2_name, sizeof(xc->data2_name), "%s", r[1].SLC_file);
      FILE* f = fopen(xc->data2_name, "r");
      if (!f) {
        snprintf(err, errlen, " %s %s \n", "Cannot open SLC_file", xc->data2_name);
        return 1;
      }
      fclose(f);
      xc->s_nx = r[1].num_rng_bins;
      xc->s_ny = r[1].num_patches * r[1].num_valid_az;
      xc->x_offset = r[1].rshift;
      xc->y_offset = r[1].ashift;
      xc->astretcha = (r[0].prf > 0) ? (r[1].prf - r[0].prf) / r[0].prf : 0.0;
    }
    if (xc->offset_flag == 1) {
      xc->x_offset = xc->y_offset = 0;
      fprintf(stderr, " setting ashift and rshift to zero\n");
    }
  }
  fprintf(stderr, " %d %d %d %d %d %d %f\n", xc->m_nx, xc->m_ny, xc->s_nx, xc->s_ny, xc->x_offset, xc->y_offset,
          xc->astretcha);
  return 0;
}

extern "C" int xh_get_locations(xh_config* xc, int32_t* xy, int cap) {
  // get_locations.c:14-45
  xc->x_inc = (xc->m_nx - 2 * (xc->xsearch + xc->nx_corr)) / (xc->nxl + 3);
  xc->y_inc = (xc->m_ny - 2 * (xc->ysearch + xc->ny_corr)) / (xc->nyl + 1);
  int n = 0;
  for (int j = 1; j <= xc->nyl; j++)
    for (int i = 2; i <= xc->nxl + 1; i++) {
      if (xy && n < cap) {
        xy[2 * n] = xc->npx + i * xc->x_inc;
        xy[2 * n + 1] = xc->npy + j * xc->y_inc;
      }
      n++;
    }
  xc->nlocs = n;
  fprintf(stderr, " locations  n %d nx %d nyl %d nxl %d x_inc %d y_inc %d\n", n, xc->m_nx, xc->nyl, xc->nxl,
          xc->x_inc, xc->y_inc);
  return n;
}

extern "C" void xh_window(const xh_config* xc, int x, int y, int32_t* mx0, int32_t* my0, int32_t* sx0,
                          int32_t* sy0) {
  const int ishft = (int)(y * xc->astretcha);
  *mx0 = x - xc->npx / 2;
  *my0 = y - xc->npy / 2;
  *sx0 = x + xc->x_offset - xc->npx / 2;
  *sy0 = y + xc->y_offset + ishft - xc->npy / 2;
}

extern "C" int xh_format_result(const xh_config* xc, int x, int y, float xoff, float yoff, float xfrac, float yfrac,
                                float corr, char* buf, size_t buflen) {
  // print_results.c:12-46 (float arithmetic as written there)
  const int ishft = (int)(y * xc->astretcha);
  xoff = (xoff / (float)xc->ri) - (xfrac / (float)xc->ri) + xc->x_offset;
  yoff = yoff - yfrac + xc->y_offset + ishft;
  return snprintf(buf, buflen, " %d %6.3f %d %6.3f %6.2f \n", x, xoff, y, yoff, corr);
}

extern "C" int xh_read_slc(const char* path, int nx, int ny, void* buf) {
  // read_complex_short2 / read_real_float2 (read_xcorr_data.c:11-49) read whole
  // lines; here the whole file is read once.  Lines past the end read as zero.
  const size_t bytes = (size_t)nx * (size_t)ny * 4;
  memset(buf, 0, bytes);
  FILE* f = fopen(path, "rb");
  if (!f) return -1;
  size_t got = fread(buf, 1, bytes, f);
  (void)got;
  fclose(f);
  return 0;
}
