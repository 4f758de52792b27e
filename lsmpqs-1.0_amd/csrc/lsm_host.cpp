// lsm_host.cpp -- host-side records and files of the C ABI: Options
// defaults, QSS_DataArrays allocation, headered array files.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <string>
#include "lsm_gz.h"
#include "lsmpqs_hip.h"

extern "C" {

// reference: src/utilities/qss_options.c:42-117
Options* createOptionsDefault(void) {
  Options* o = (Options*)calloc(1, sizeof(Options));
  if (!o) return nullptr;
  snprintf(o->outfile, sizeof o->outfile, "out_file");
  snprintf(o->path, sizeof o->path, "./");
  o->dx = (QSSLIB_REAL)0.04;
  o->tmax = 150.0;
  o->tplot = 0.5;
  snprintf(o->accuracy, sizeof o->accuracy, "medium");
  o->accuracy_id = MEDIUM;
  o->a = (QSSLIB_REAL)0.15;
  o->b = (QSSLIB_REAL)0.04;
  o->save_data = 1;
  o->do_reinit = 1;
  o->do_mask = 1;
  o->eps_stop = (QSSLIB_REAL)0.014;
  o->narrow_band = 0;
  o->overlap = 0;
  o->init_step = 0;
  o->print_details = 1;
  o->tmax_r = 10 * o->dx;
  o->subcell_fix = 1;
  o->dc = (QSSLIB_REAL)0.1;
  o->cmax = (QSSLIB_REAL)(1.0 / (o->dx / 2.0));
  o->amax = 10;
  o->amin = (QSSLIB_REAL)0.05;
  o->vol_frac_max = (QSSLIB_REAL)0.97;
  o->vol_frac_min = 0;
  o->mask_reinit = 1;
  o->stop_touch = 1;
  o->eps_coefficient = (QSSLIB_REAL)1.5;
  o->extrapol_id = 0;
  o->extrapol_func = (void (*)())signedLinearExtrapolationBCqss;
  o->checkpoint = 0;
  o->reservoir_inlet = 1;
  o->theta = 0;
  o->C = 1;
  o->check_connectivity = 1;
  o->err_check_zone = 20 * o->dx;
  o->use_var_theta = 0;
  o->order_space_accur = 3;
  o->order_time_accur = 3;
  o->use_satn_stop = 1;
  o->conserve_imbibe = 0;
  o->center_inlet = 0;
  return o;
}

void destroyOptions(Options* o) { free(o); }

// reference: src/utilities/qss_data_arrays.c:27-66
QSS_DataArrays* allocateQSSDataArrays(void) {
  QSS_DataArrays* p = (QSS_DataArrays*)malloc(sizeof(QSS_DataArrays));
  if (!p) return nullptr;
  QSSLIB_REAL** r[] = {&p->phi, &p->phi_next, &p->phi_extra, &p->phi_prev, &p->mask, &p->lse_rhs,
                       &p->normal_velocity, &p->curvature_coeff, &p->external_velocity_x,
                       &p->external_velocity_y, &p->external_velocity_z, &p->mask_x, &p->mask_y,
                       &p->mask_z, &p->scratch1, &p->scratch2, &p->mask_w, &p->mask_nw,
                       &p->mask_disconn_init, &p->theta, &p->overlap};
  for (auto q : r) *q = QSSLIB_SERIAL_dummy_pointer;
  p->connectivity = QSSLIB_SERIAL_dummy_pointer_int;
  p->phi_bin = QSSLIB_SERIAL_dummy_pointer_uchar;
  return p;
}

// reference: src/utilities/qss_data_arrays.c:116-209 (dummy pointer => allocate)
void allocateMemoryForQSSDataArrays(QSS_DataArrays* p, Grid* g) {
  size_t n = (size_t)g->num_gridpts;
  QSSLIB_REAL** zeroed[] = {&p->phi, &p->phi_next, &p->phi_extra, &p->phi_prev, &p->mask};
  for (auto q : zeroed)
    if (*q == QSSLIB_SERIAL_dummy_pointer) *q = (QSSLIB_REAL*)calloc(n, sizeof(QSSLIB_REAL));
  QSSLIB_REAL** rest[] = {&p->lse_rhs, &p->normal_velocity, &p->curvature_coeff,
                          &p->external_velocity_x, &p->external_velocity_y,
                          &p->external_velocity_z, &p->mask_x, &p->mask_y, &p->mask_z,
                          &p->scratch1, &p->scratch2, &p->theta, &p->overlap, &p->mask_w,
                          &p->mask_nw, &p->mask_disconn_init};
  for (auto q : rest)
    if (*q == QSSLIB_SERIAL_dummy_pointer) *q = (QSSLIB_REAL*)malloc(n * sizeof(QSSLIB_REAL));
  if (p->connectivity == QSSLIB_SERIAL_dummy_pointer_int) p->connectivity = (int*)malloc(n * sizeof(int));
  if (p->phi_bin == QSSLIB_SERIAL_dummy_pointer_uchar) p->phi_bin = (unsigned char*)malloc(n);
}

static void free_real(QSSLIB_REAL* q) {
  if (q && q != QSSLIB_SERIAL_dummy_pointer) free(q);
}

void destroyQSSDataArrays(QSS_DataArrays* p) {
  if (!p) return;
  lsmpqs_release(p);
  QSSLIB_REAL* r[] = {p->phi, p->phi_next, p->phi_extra, p->phi_prev, p->mask, p->lse_rhs,
                      p->normal_velocity, p->curvature_coeff, p->external_velocity_x,
                      p->external_velocity_y, p->external_velocity_z, p->mask_x, p->mask_y,
                      p->mask_z, p->scratch1, p->scratch2, p->mask_w, p->mask_nw,
                      p->mask_disconn_init, p->theta, p->overlap};
  for (auto q : r) free_real(q);
  if (p->connectivity && p->connectivity != QSSLIB_SERIAL_dummy_pointer_int) free(p->connectivity);
  if (p->phi_bin && p->phi_bin != QSSLIB_SERIAL_dummy_pointer_uchar) free(p->phi_bin);
  free(p);
}

// Headered array file: int[3] ghost-box dims, then QSSLIB_REAL[N]; zip_status
// 1 compresses with gzip afterwards (reference: src/utilities/qss_data_arrays.c:214-228,
// src/utilities/qss_file.c:50-64)
// The gzip file is written directly, deflated on several threads
// (lsm::write_array_gz), and the uncompressed name removed, which leaves the
// files the reference's write + `gzip -f` leaves.
void writeDataArrayQSS(QSSLIB_REAL* data, Grid* g, char* file_name, int zip_status) {
  if (zip_status == 1) {
    if (!lsm::write_array_gz(std::string(file_name) + ".gz", g->grid_dims_ghostbox, data,
                             sizeof(QSSLIB_REAL) * (size_t)g->num_gridpts))
      fprintf(stderr, "lsmpqs_hip: cannot write %s.gz\n", file_name);
    remove(file_name);
    return;
  }
  FILE* fp = fopen(file_name, "wb");
  if (!fp) return;
  fwrite(g->grid_dims_ghostbox, sizeof(int), 3, fp);
  fwrite(data, sizeof(QSSLIB_REAL), (size_t)g->num_gridpts, fp);
  fclose(fp);
}

// Headerless array file: QSSLIB_REAL[N] over the ghost box; zip_status 1
// gzips it (reference: src/utilities/qss_data_arrays.c:231-242)
void writeDataArrayRaw(QSSLIB_REAL* data, Grid* g, char* file_name, int zip_status) {
  const size_t bytes = sizeof(QSSLIB_REAL) * (size_t)g->num_gridpts;
  if (zip_status == 1) {
    if (!lsm::gz_write(std::string(file_name) + ".gz", {{data, bytes}}, 0))
      fprintf(stderr, "lsmpqs_hip: cannot write %s.gz\n", file_name);
    remove(file_name);
    return;
  }
  FILE* fp = fopen(file_name, "wb");
  if (!fp) return;
  fwrite(data, 1, bytes, fp);
  fclose(fp);
}

// reference: src/utilities/qss_data_arrays.c:347-382 (a .gz name is gunzipped
// in place and compressed again afterwards)
QSSLIB_REAL* readDataArrayQSS(int* dims, char* file_name) {
  char base[512];
  size_t len = strlen(file_name);
  int gz = len > 3 && !strcmp(file_name + len - 3, ".gz");
  snprintf(base, sizeof base, "%.*s", (int)(gz ? len - 3 : len), file_name);
  char cmd[1200];
  if (gz) {
    snprintf(cmd, sizeof cmd, "gunzip -f '%s'", file_name);
    if (system(cmd) != 0) return nullptr;
  }
  QSSLIB_REAL* data = nullptr;
  FILE* fp = fopen(base, "rb");
  if (fp) {
    if (fread(dims, sizeof(int), 3, fp) == 3) {
      size_t n = (size_t)dims[0] * dims[1] * dims[2];
      data = (QSSLIB_REAL*)malloc(n * sizeof(QSSLIB_REAL));
      if (data && fread(data, sizeof(QSSLIB_REAL), n, fp) != n) {
        free(data);
        data = nullptr;
      }
    }
    fclose(fp);
  } else {
    printf("\nCould not open file %s", file_name);
  }
  if (gz) {
    snprintf(cmd, sizeof cmd, "gzip -f '%s'", base);
    if (system(cmd) != 0) fprintf(stderr, "lsmpqs_hip: gzip failed for %s\n", base);
  }
  return data;
}

}  // extern "C"
