// lsm_files.cpp -- host side of the file-level entry points: Grid
// construction and its binary file, the option-file parser, the log-file
// summaries, and drainTop / imbibeTop.  Every file format and text layout
// here is the reference's, so files and logs interchange with it.
#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <string>
#include "lsmpqs_hip.h"

namespace {

// Ghost layers and fill-box rule per accuracy level (reference:
// src/geometry/qss_grid.c:47,789-1036).  The fill box of each axis is
// [D_lo + lo_add, D_hi - hi_sub] where D is the central-difference box of the
// given level (D1 = [1, N-2], D2 = [2, N-3], D3 = [3, N-4] on N voxels).
struct AccuracyRule {
  int ghost, level, lo_add, hi_sub;
};
const AccuracyRule kRules[4] = {
    {2, 1, 0, 1},  // LOW: HJ ENO1
    {3, 2, 1, 1},  // MEDIUM: HJ ENO2
    {5, 3, 1, 2},  // HIGH: HJ ENO3
    {4, 1, 2, 3},  // VERY_HIGH: HJ WENO5
};

const AccuracyRule& rule_for(int accuracy) {
  return (accuracy >= LOW && accuracy <= VERY_HIGH) ? kRules[accuracy] : kRules[MEDIUM];
}

void set_index_limits(Grid* g, int accuracy) {
  const AccuracyRule& r = rule_for(accuracy);
  int* gb[3][2] = {{&g->ilo_gb, &g->ihi_gb}, {&g->jlo_gb, &g->jhi_gb}, {&g->klo_gb, &g->khi_gb}};
  int* d1[3][2] = {{&g->ilo_D1_fb, &g->ihi_D1_fb}, {&g->jlo_D1_fb, &g->jhi_D1_fb},
                   {&g->klo_D1_fb, &g->khi_D1_fb}};
  int* d2[3][2] = {{&g->ilo_D2_fb, &g->ihi_D2_fb}, {&g->jlo_D2_fb, &g->jhi_D2_fb},
                   {&g->klo_D2_fb, &g->khi_D2_fb}};
  int* d3[3][2] = {{&g->ilo_D3_fb, &g->ihi_D3_fb}, {&g->jlo_D3_fb, &g->jhi_D3_fb},
                   {&g->klo_D3_fb, &g->khi_D3_fb}};
  int* fb[3][2] = {{&g->ilo_fb, &g->ihi_fb}, {&g->jlo_fb, &g->jhi_fb}, {&g->klo_fb, &g->khi_fb}};
  for (int a = 0; a < 3; ++a) {
    const int n = g->grid_dims_ghostbox[a];
    *gb[a][0] = 0; *gb[a][1] = n - 1;
    *d1[a][0] = 1; *d1[a][1] = n - 2;
    *d2[a][0] = 2; *d2[a][1] = n - 3;
    *d3[a][0] = 3; *d3[a][1] = n - 4;
    *fb[a][0] = r.level + r.lo_add;
    *fb[a][1] = n - 1 - r.level - r.hi_sub;
  }
  g->mark_gb = 120;
  g->mark_D1 = 119;
  g->mark_D2 = 118;
  g->mark_D3 = 117;
  g->num_nb_levels = r.ghost;
  g->mark_fb = (unsigned char)(120 - r.ghost);
  const bool wide = accuracy > MEDIUM;
  g->beta = (wide ? 3 : 2) * g->dx[0];
  g->gamma = (wide ? 6 : 4) * g->dx[0];
}

// Shared body of the createGridSet* constructors: dims from extents (or the
// reverse), ghost layers, 2D placeholders for the third axis.
Grid* make_grid(int num_dims, const QSSLIB_REAL* dxs, const int* dims, const QSSLIB_REAL* x_lo,
                const QSSLIB_REAL* x_hi, int accuracy) {
  Grid* g = (Grid*)calloc(1, sizeof(Grid));
  if (!g) return nullptr;
  if (num_dims == 2 || num_dims == 3) {
    g->num_dims = num_dims;
  } else {
    printf("\nInvalid grid dimension %d, set to default 2.\n", num_dims);
    g->num_dims = 2;
  }
  const int ghost = rule_for(accuracy).ghost;
  if (dxs) for (int a = 0; a < 3; ++a) g->dx[a] = dxs[a];
  g->num_gridpts = 1;
  for (int a = 0; a < g->num_dims; ++a) {
    g->x_lo[a] = x_lo[a];
    if (dims) {
      g->grid_dims[a] = dims[a];
      g->x_hi[a] = x_hi[a];
      g->dx[a] = (g->x_hi[a] - g->x_lo[a]) / g->grid_dims[a];
    } else {
      QSSLIB_REAL span = x_hi[a] - x_lo[a];
      g->grid_dims[a] = (int)rint(span / g->dx[a]);
      g->x_hi[a] = x_lo[a] + g->grid_dims[a] * g->dx[a];
    }
    g->grid_dims_ghostbox[a] = g->grid_dims[a] + 2 * ghost;
    g->x_lo_ghostbox[a] = g->x_lo[a] - ghost * g->dx[a];
    g->x_hi_ghostbox[a] = g->x_hi[a] + ghost * g->dx[a];
    g->num_gridpts *= g->grid_dims_ghostbox[a];
  }
  if (g->num_dims == 2) {
    g->x_lo[2] = g->x_hi[2] = 0;
    g->x_lo_ghostbox[2] = g->x_hi_ghostbox[2] = 0;
    g->grid_dims[2] = g->grid_dims_ghostbox[2] = 1;
    g->dx[2] = 0;
  }
  set_index_limits(g, accuracy);
  return g;
}

// Fields of the grid binary file in file order (reference:
// src/geometry/qss_grid.c:614-787): a flat sequence with no header or padding.
template <typename F>
void grid_fields(Grid* g, F&& io) {
  io(&g->num_dims, sizeof(int), 1);
  io(g->x_lo, sizeof(QSSLIB_REAL), 3);
  io(g->x_hi, sizeof(QSSLIB_REAL), 3);
  io(g->x_lo_ghostbox, sizeof(QSSLIB_REAL), 3);
  io(g->x_hi_ghostbox, sizeof(QSSLIB_REAL), 3);
  io(g->grid_dims, sizeof(int), 3);
  io(g->grid_dims_ghostbox, sizeof(int), 3);
  io(g->dx, sizeof(QSSLIB_REAL), 3);
  io(&g->num_gridpts, sizeof(int), 1);
  int* box[] = {&g->ilo_gb,    &g->ihi_gb,    &g->jlo_gb,    &g->jhi_gb,    &g->klo_gb,
                &g->khi_gb,    &g->ilo_fb,    &g->ihi_fb,    &g->jlo_fb,    &g->jhi_fb,
                &g->klo_fb,    &g->khi_fb,    &g->ilo_D1_fb, &g->ihi_D1_fb, &g->jlo_D1_fb,
                &g->jhi_D1_fb, &g->klo_D1_fb, &g->khi_D1_fb, &g->ilo_D2_fb, &g->ihi_D2_fb,
                &g->jlo_D2_fb, &g->jhi_D2_fb, &g->klo_D2_fb, &g->khi_D2_fb, &g->ilo_D3_fb,
                &g->ihi_D3_fb, &g->jlo_D3_fb, &g->jhi_D3_fb, &g->klo_D3_fb, &g->khi_D3_fb,
                &g->num_nb_levels};
  for (int* v : box) io(v, sizeof(int), 1);
  unsigned char* marks[] = {&g->mark_gb, &g->mark_D1, &g->mark_D2, &g->mark_D3, &g->mark_fb};
  for (unsigned char* m : marks) io(m, 1, 1);
  io(&g->beta, sizeof(QSSLIB_REAL), 1);
  io(&g->gamma, sizeof(QSSLIB_REAL), 1);
}

// The reference decompresses a .gz file in place before reading it and
// compresses it again afterwards (src/utilities/qss_file.c:13-64).
std::string gunzip_in_place(const char* name, bool* gz) {
  size_t len = strlen(name);
  *gz = len > 3 && !strcmp(name + len - 3, ".gz");
  std::string base(name, *gz ? len - 3 : len);
  if (*gz) {
    std::string cmd = "gunzip -f '" + std::string(name) + "'";
    if (system(cmd.c_str()) != 0) fprintf(stderr, "lsmpqs_hip: gunzip failed for %s\n", name);
  }
  return base;
}

void gzip_in_place(const std::string& base, bool gz) {
  if (!gz) return;
  std::string cmd = "gzip -f '" + base + "'";
  if (system(cmd.c_str()) != 0) fprintf(stderr, "lsmpqs_hip: gzip failed for %s\n", base.c_str());
}

// ---------------------------------------------------------------------------
// Option file: one "key value" per line; the key is recognised from its first
// letters exactly as the reference's parser does (src/utilities/qss_options.c:118-494),
// including lines read 79 characters at a time.
double num_after_key(const char* s) {
  double v = 0;
  sscanf(s, "%*s %lf ", &v);
  return v;
}
int int_after_key(const char* s, int fallback) {
  int v = fallback;
  sscanf(s, "%*s %d ", &v);
  return v;
}

// flag options accept 0 or 1; anything else keeps the current value
void set_flag(int* field, const char* s, const char* key) {
  int v = int_after_key(s, -12345);
  if (v == 0 || v == 1) *field = v;
  else printf("\nIncorrect %s option %d, set to default.\n", key, v);
}

void parse_option_line(Options* o, const char* line) {
  int n = 0;
  while (isspace((unsigned char)line[n])) n++;
  const char* s = line + n;
  const int c = tolower((unsigned char)s[0]);
  auto at = [&](int k) { return (char)tolower((unsigned char)s[k]); };
  static const char* kAccuracy[] = {"low", "medium", "high", "very_high"};
  switch (c) {
  case 'a':
    if (at(1) == 'c') {
      char word[100] = "";
      sscanf(s, "accuracy %99s", word);
      bool found = false;
      for (int i = 0; i < 4; ++i)
        if (!strcmp(word, kAccuracy[i])) {
          found = true;
          o->accuracy_id = (QSSLIB_SPATIAL_DERIVATIVE_ACCURACY_TYPE)i;
          snprintf(o->accuracy, sizeof o->accuracy, "%s", kAccuracy[i]);
        }
      if (!found) printf("\nAccuracy type %s not found, set to default.", word);
    } else if (at(2) == 'a') {
      o->amax = (QSSLIB_REAL)num_after_key(s);
    } else if (at(2) == 'i') {
      o->amin = (QSSLIB_REAL)num_after_key(s);
    } else {
      o->a = (QSSLIB_REAL)num_after_key(s);
    }
    break;
  case 'b':
    o->b = (QSSLIB_REAL)num_after_key(s);
    break;
  case 'c':
    if (at(1) == 'm') {
      o->cmax = (QSSLIB_REAL)num_after_key(s);
    } else if (at(5) == 'p') {
      set_flag(&o->checkpoint, s, "checkpoint");
    } else if (at(6) == 'c') {
      o->check_connectivity = int_after_key(s, o->check_connectivity);
    } else if (at(1) == 'e') {
      o->center_inlet = int_after_key(s, o->center_inlet);
    }
    break;
  case 'd':
    if (at(1) == 'x') o->dx = (QSSLIB_REAL)num_after_key(s);
    else if (s[3] == 'r') set_flag(&o->do_reinit, s, "do_reinit");
    else if (s[3] == 'm') set_flag(&o->do_mask, s, "do_mask");
    else if (at(1) == 'c') o->dc = (QSSLIB_REAL)num_after_key(s);
    break;
  case 'e':
    if (at(1) == 'p' && at(4) == 'c') {
      double v = num_after_key(s);
      if (v > 0) o->eps_coefficient = (QSSLIB_REAL)v;
      else printf("\nIncorrect eps_coefficient option %g, set to default.\n", v);
    } else if (at(1) == 'r' && at(10) == 'z') {
      o->err_check_zone = (QSSLIB_REAL)num_after_key(s);
    } else if (at(1) == 'p' && at(4) == 's') {
      o->eps_stop = (QSSLIB_REAL)num_after_key(s);
    } else if (at(1) == 'x' && at(2) == 't') {
      int v = int_after_key(s, 0);
      if (v < 0 || v > 3) {
        printf("\nIncorrect extrapol_id option %d, set to default.\n", v);
        v = 0;
      }
      o->extrapol_id = v;
      o->extrapol_func = (void (*)())signedLinearExtrapolationBCqss;
    }
    break;
  case 'i':
    o->init_step = int_after_key(s, o->init_step);
    break;
  case 'm':
    if (s[1] == 'a') set_flag(&o->mask_reinit, s, "mask_reinit");
    break;
  case 'n':
    set_flag(&o->narrow_band, s, "narrow_band");
    break;
  case 'o':
    if (s[1] == 'u') {
      char word[100] = "";
      sscanf(s, "outfile %99s", word);
      snprintf(o->outfile, sizeof o->outfile, "%s", word);
      // path = the outfile name up to and including its last '/', else "./"
      const char* slash = strrchr(o->outfile, '/');
      if (slash) snprintf(o->path, sizeof o->path, "%.*s", (int)(slash - o->outfile + 1), o->outfile);
      else snprintf(o->path, sizeof o->path, "./");
    } else if (s[1] == 'v') {
      o->overlap = (QSSLIB_REAL)num_after_key(s);
    }
    break;
  case 'r':
    if (s[1] == 'e') {
      int v = int_after_key(s, -12345);
      if (v == 0 || v == 1) {
        o->reservoir_inlet = v;
      } else {
        printf("\nIncorrect reservoir_inlet option %d, set to default.\n", v);
        o->reservoir_inlet = 0;
      }
    }
    break;
  case 's':
    if (at(1) == 'a') set_flag(&o->save_data, s, "save_data");
    else if (at(5) == 't') set_flag(&o->stop_touch, s, "stop_touch");
    else if (at(1) == 'u') set_flag(&o->subcell_fix, s, "subcell_fix");
    else if (at(1) == 'p') o->order_space_accur = int_after_key(s, o->order_space_accur);
    break;
  case 't':
    if (at(1) == 'm') o->tmax = (QSSLIB_REAL)num_after_key(s);
    else if (at(1) == 'p') o->tplot = (QSSLIB_REAL)num_after_key(s);
    else if (at(1) == 'i') o->order_time_accur = int_after_key(s, o->order_time_accur);
    else if (at(1) == 'h') o->theta = (QSSLIB_REAL)num_after_key(s) * 3.14159 / 180;
    break;
  case 'v':
    if (at(9) == 'm' && at(10) == 'a') o->vol_frac_max = (QSSLIB_REAL)num_after_key(s);
    else if (at(9) == 'm' && at(10) == 'i') o->vol_frac_min = (QSSLIB_REAL)num_after_key(s);
    break;
  case 'u':
    if (at(4) == 'v') o->use_var_theta = int_after_key(s, o->use_var_theta);
    else if (at(4) == 's') o->use_satn_stop = int_after_key(s, o->use_satn_stop);
    break;
  case 'p': {
    int v = int_after_key(s, -12345);
    if (v == 0 || v == 1) {
      o->print_details = v;
    } else {
      printf("\nIncorrect print_details option %d, set to default.\n", v);
      o->print_details = 1;
    }
    break;
  }
  default:
    break;
  }
}

// drainTop / imbibeTop body (reference: src/level_set_evolution/drain_top.c:30-138,
// imbibe_top.c:32-140)
int run_top(Options* o, char* fdata, char* fgrid, char* fmask, bool drain) {
  time_t t0;
  time(&t0);
  if (!o || !fdata || !fgrid || !fmask) {
    // the reference leaves the grid unset on this path; here it is an error
    fprintf(stderr, "lsmpqs_hip: %s needs data, grid and mask file names\n",
            drain ? "drainTop" : "imbibeTop");
    return 1;
  }
  QSS_DataArrays* p = allocateQSSDataArrays();
  int n1[3] = {0, 0, 0}, n2[3] = {0, 0, 0};
  p->phi = readDataArrayQSS(n1, fdata);
  p->mask = readDataArrayQSS(n2, fmask);
  Grid* g = readGridFromBinaryFile(fgrid);
  if (!p->phi || !p->mask || !g) {
    fprintf(stderr, "lsmpqs_hip: could not read the input files\n");
    if (!p->phi) p->phi = QSSLIB_SERIAL_dummy_pointer;
    if (!p->mask) p->mask = QSSLIB_SERIAL_dummy_pointer;
    destroyQSSDataArrays(p);
    destroyGrid(g);
    return 1;
  }
  allocateMemoryForQSSDataArrays(p, g);
  for (int i = 0; i < 3; ++i) {
    if (n1[i] != g->grid_dims_ghostbox[i] || n2[i] != g->grid_dims_ghostbox[i]) {
      printf("\nInput data dimensions n1[%d]=%d, n2[%d]=%d,", i, n1[i], i, n2[i]);
      printf(" (grid->grid_dims_ghostbox)[%d]=%d", i, g->grid_dims_ghostbox[i]);
      printf(" don't match\n");
      printf("\nTerminating...");
      destroyQSSDataArrays(p);
      destroyGrid(g);
      return 1;
    }
  }
  if (g->num_dims != 3 && g->num_dims != 2) {
    fprintf(stderr, "lsmpqs_hip: grids must be 2D or 3D (num_dims = %d)\n", g->num_dims);
    destroyQSSDataArrays(p);
    destroyGrid(g);
    return 1;
  }
  FILE* fp = fopen(o->outfile, "a+");
  if (!fp) {
    fprintf(stderr, "lsmpqs_hip: cannot open log file %s\n", o->outfile);
    destroyQSSDataArrays(p);
    destroyGrid(g);
    return 1;
  }
  if (o->print_details) {
    fprintf(fp, "Time Start %s\n", ctime(&t0));
    fprintf(fp, drain ? "Level Set Method simulation\n"
                      : "Constant Curvature Model Level Set Method simulation\n");
    printOptions(o, fp);
    printGrid(g, fp);
  }
  fflush(fp);
  if (g->num_dims == 3) {
    if (drain) drain_model3d(o, p, g, fp);
    else imbibe_model3d(o, p, g, fp);
  } else {
    if (drain) drain_model2d(o, p, g, fp);
    else imbibe_model2d(o, p, g, fp);
  }
  int rc = *lsmpqs_hip_last_error() ? 1 : 0;
  if (o->save_data) {
    std::string name = std::string(o->path) + "/data_final";
    writeDataArrayQSS(p->phi, g, &name[0], 1);
  }
  destroyQSSDataArrays(p);
  destroyGrid(g);
  fclose(fp);
  return rc;
}

}  // namespace

extern "C" {

Grid* createGridSetDx(int num_dims, QSSLIB_REAL dx, QSSLIB_REAL* x_lo, QSSLIB_REAL* x_hi,
                      QSSLIB_SPATIAL_DERIVATIVE_ACCURACY_TYPE accuracy) {
  QSSLIB_REAL d[3] = {dx, dx, dx};
  return make_grid(num_dims, d, nullptr, x_lo, x_hi, accuracy);
}

Grid* createGridSetDxDyDz(int num_dims, QSSLIB_REAL dx, QSSLIB_REAL dy, QSSLIB_REAL dz,
                          QSSLIB_REAL* x_lo, QSSLIB_REAL* x_hi,
                          QSSLIB_SPATIAL_DERIVATIVE_ACCURACY_TYPE accuracy) {
  QSSLIB_REAL d[3] = {dx, dy, dz};
  return make_grid(num_dims, d, nullptr, x_lo, x_hi, accuracy);
}

Grid* createGridSetGridDims(int num_dims, int* grid_dims, QSSLIB_REAL* x_lo, QSSLIB_REAL* x_hi,
                            QSSLIB_SPATIAL_DERIVATIVE_ACCURACY_TYPE accuracy) {
  return make_grid(num_dims, nullptr, grid_dims, x_lo, x_hi, accuracy);
}

void destroyGrid(Grid* g) { free(g); }

void writeGridToBinaryFile(Grid* g, char* file_name, int zip_status) {
  FILE* fp = fopen(file_name, "wb");
  if (!fp) {
    printf("\nCould not open file %s", file_name);
    return;
  }
  grid_fields(g, [&](void* v, size_t sz, size_t n) { fwrite(v, sz, n, fp); });
  fclose(fp);
  gzip_in_place(file_name, zip_status == 1);
}

Grid* readGridFromBinaryFile(char* file_name) {
  bool gz = false;
  std::string base = gunzip_in_place(file_name, &gz);
  Grid* g = (Grid*)calloc(1, sizeof(Grid));
  FILE* fp = fopen(base.c_str(), "rb");
  if (!fp) {
    printf("\nCould not open file %s", base.c_str());
    return g;
  }
  bool ok = true;
  grid_fields(g, [&](void* v, size_t sz, size_t n) { ok = ok && fread(v, sz, n, fp) == n; });
  fclose(fp);
  gzip_in_place(base, gz);
  if (!ok) fprintf(stderr, "lsmpqs_hip: grid file %s is truncated\n", file_name);
  return g;
}

// reference: src/geometry/qss_grid.c:323-440 (same text, line for line)
void printGrid(Grid* g, FILE* fp) {
  fprintf(fp, "Number of dimensions: %d\n", g->num_dims);
  const bool two = g->num_dims == 2;
  if (two) {
    fprintf(fp, "Geometric limits (without ghostcells): [%g,%g] x [%g,%g]\n", (double)g->x_lo[0],
            (double)g->x_hi[0], (double)g->x_lo[1], (double)g->x_hi[1]);
    fprintf(fp, "Geometric limits (with ghostcells): [%g,%g] x [%g,%g]\n",
            (double)g->x_lo_ghostbox[0], (double)g->x_hi_ghostbox[0], (double)g->x_lo_ghostbox[1],
            (double)g->x_hi_ghostbox[1]);
    fprintf(fp, "Grid size (without ghostcells): %d x %d\n", g->grid_dims[0], g->grid_dims[1]);
    fprintf(fp, "Grid size (with ghostcells): %d x %d\n", g->grid_dims_ghostbox[0],
            g->grid_dims_ghostbox[1]);
    fprintf(fp, "Grid spacing: dx = %g, dy = %g \n", (double)g->dx[0], (double)g->dx[1]);
  } else {
    fprintf(fp, "Geometric limits (without ghostcells): [%g,%g] x [%g,%g] x [%g,%g]\n",
            (double)g->x_lo[0], (double)g->x_hi[0], (double)g->x_lo[1], (double)g->x_hi[1],
            (double)g->x_lo[2], (double)g->x_hi[2]);
    fprintf(fp, "Geometric limits (with ghostcells): [%g,%g] x [%g,%g] x [%g,%g]\n",
            (double)g->x_lo_ghostbox[0], (double)g->x_hi_ghostbox[0], (double)g->x_lo_ghostbox[1],
            (double)g->x_hi_ghostbox[1], (double)g->x_lo_ghostbox[2], (double)g->x_hi_ghostbox[2]);
    fprintf(fp, "Grid size (without ghostcells): %d x %d x %d\n", g->grid_dims[0], g->grid_dims[1],
            g->grid_dims[2]);
    fprintf(fp, "Grid size (with ghostcells): %d x %d x %d\n", g->grid_dims_ghostbox[0],
            g->grid_dims_ghostbox[1], g->grid_dims_ghostbox[2]);
    fprintf(fp, "Grid spacing: dx = %g, dy = %g, dz = %g \n", (double)g->dx[0], (double)g->dx[1],
            (double)g->dx[2]);
  }
  fprintf(fp, "Total number of grid points: %d\n", g->num_gridpts);
  struct Box {
    const char* label;
    int lo[3], hi[3];
  } boxes[] = {
      {"Ghost box index limits", {g->ilo_gb, g->jlo_gb, g->klo_gb}, {g->ihi_gb, g->jhi_gb, g->khi_gb}},
      {"Fill box index limits", {g->ilo_fb, g->jlo_fb, g->klo_fb}, {g->ihi_fb, g->jhi_fb, g->khi_fb}},
      {"1st order differences fill box", {g->ilo_D1_fb, g->jlo_D1_fb, g->klo_D1_fb},
       {g->ihi_D1_fb, g->jhi_D1_fb, g->khi_D1_fb}},
      {"2nd order differences fill box", {g->ilo_D2_fb, g->jlo_D2_fb, g->klo_D2_fb},
       {g->ihi_D2_fb, g->jhi_D2_fb, g->khi_D2_fb}},
      {"3rd order differences fill box", {g->ilo_D3_fb, g->jlo_D3_fb, g->klo_D3_fb},
       {g->ihi_D3_fb, g->jhi_D3_fb, g->khi_D3_fb}},
  };
  for (const Box& b : boxes) {
    if (two)
      fprintf(fp, "%s: [%d,%d] x [%d,%d]\n", b.label, b.lo[0], b.hi[0], b.lo[1], b.hi[1]);
    else
      fprintf(fp, "%s: [%d,%d] x [%d,%d] x [%d,%d]\n", b.label, b.lo[0], b.hi[0], b.lo[1], b.hi[1],
              b.lo[2], b.hi[2]);
  }
  fprintf(fp, "Number of narrow band levels (local method) %d\n", g->num_nb_levels);
  fprintf(fp, "Boundary layer marks (local method) gb %u D1 %u D2 %u D3 %u fb %u\n", g->mark_gb,
          g->mark_D1, g->mark_D2, g->mark_D3, g->mark_fb);
  fprintf(fp, "Narrow band width (local method) beta(inner) %g gamma(outer) %g\n", (double)g->beta,
          (double)g->gamma);
}

Options* createOptionsFromInputFile(char* file_name) {
  Options* o = createOptionsDefault();
  FILE* fp = fopen(file_name, "r");
  if (!fp) {
    printf("\nCouldnt open file %s", file_name);
    return o;
  }
  char line[100];
  while (fgets(line, 80, fp)) parse_option_line(o, line);
  fclose(fp);
  return o;
}

// reference: src/utilities/qss_options.c:497-541 (same text, line for line)
void printOptions(Options* o, FILE* fp) {
  const char* rule = "  -----------------------------------------------------------------------------\n";
  auto section = [&](const char* title) {
    fputs(rule, fp);
    fprintf(fp, "  %s: \n", title);
    fputs(rule, fp);
  };
  section("Geometry options");
  fprintf(fp, "  dx                   %8g  [ grid spacing ]\n", (double)o->dx);
  fprintf(fp, "  mask_reinit          %8d  [ reinitialize mask or not ]\n", o->mask_reinit);
  fprintf(fp, "  reservoir_inlet      %8d  [ impose reservoir at inlet or not ]\n", o->reservoir_inlet);
  section("Running options");
  fprintf(fp, "  a                    %8g  [ Initial normal velocity (or cap pressure) ]\n", (double)o->a);
  fprintf(fp, "  b                    %8g  [ Curvature coefficient   ]\n", (double)o->b);
  fprintf(fp, "  dc                   %8g  [ curvature increment/decrement ]\n", (double)o->dc);
  fprintf(fp, "   space accuracy       %8d [ accuracy options: low, medium, high, very_high ]\n",
          o->order_space_accur);
  fprintf(fp, "   time accuracy        %8d  [ accuracy options: low, medium, high, very_high ]\n",
          o->order_time_accur);
  fprintf(fp, "  theta                %8g  [ contact angle, in radians ]\n", (double)o->theta);
  fprintf(fp, "  tplot                %8g  [ intervals for reinit, error ]\n", (double)o->tplot);
  fprintf(fp, "  init_step            %8d  [ Initial step (if restarting a stopped simulation) ]\n",
          o->init_step);
  fprintf(fp, "  overlap              %8g  [ overlap between pore space and mask allowed ]\n",
          (double)o->overlap);
  fprintf(fp, "  eps_coefficient      %8g  [ numerical width of smearing for Heaviside ]\n",
          (double)o->eps_coefficient);
  fprintf(fp, "  check_connectivity   %8d  [ check for trapped components or not ]\n",
          o->check_connectivity);
  fprintf(fp, "  checkpoint           %8d  [ output checkpoint data or not ]\n", o->checkpoint);
  fprintf(fp, "  err_check_zone       %8g  [ distance from zero ls to check error ]\n",
          (double)o->err_check_zone);
  fprintf(fp, "  do_reinit            %8d  [ reinitilize periodically (1) or not (0)]\n", o->do_reinit);
  fprintf(fp, "  do_mask              %8d  [ impose mask (1) or not (0)]\n", o->do_mask);
  fprintf(fp, "  print_details        %8d  [ print details (1) or not (0)   ]\n", o->print_details);
  fprintf(fp, "  center_inlet         %8d  [ Inlet at center or not \n", o->center_inlet);
  section("Stopping options");
  fprintf(fp, "  tmax                 %8g  [ max running time allowed ]\n", (double)o->tmax);
  fprintf(fp, "  eps_stop             %8g  [ criteria for stopping ]\n", (double)o->eps_stop);
  fprintf(fp, "  tmax_reinit          %8g  [ max reinit time allowed ]\n", (double)o->tmax_r);
  fprintf(fp, "  cmax                 %8g  [ maximum curvature ]\n", (double)o->cmax);
  fprintf(fp, "  amax                 %8g  [ maximum a, for drainage ]\n", (double)o->amax);
  fprintf(fp, "  amin                 %8g  [ minimum a, for imbibition ]\n", (double)o->amin);
  fprintf(fp, "  vol_frac_max         %8g  [ maximum nw fraction allowed ]\n", (double)o->vol_frac_max);
  fprintf(fp, "  stop_touch           %8d  [ stop the simulation when opp. bdry. is touched ]\n",
          o->stop_touch);
  fprintf(fp, "  save_data            %8d  [ save data (1) or not (0)       ]\n", o->save_data);
  fprintf(fp, "  outfile              %8s  [ output file name ]\n", o->outfile);
}

int drainTop(Options* options, char* fname_data_in, char* fname_grid_in, char* fname_mask_in) {
  return run_top(options, fname_data_in, fname_grid_in, fname_mask_in, true);
}

int imbibeTop(Options* options, char* fname_data_in, char* fname_grid_in, char* fname_mask_in) {
  return run_top(options, fname_data_in, fname_grid_in, fname_mask_in, false);
}

}  // extern "C"
