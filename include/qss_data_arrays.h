/*
 * qss_data_arrays.h -- the QSS_DataArrays record of the LSMPQS C ABI.
 *
 * Every pointer is a HOST array over the ghost box (Grid::num_gridpts
 * entries).  Field order follows the reference so that a record filled by
 * reference-side code is accepted unchanged (reference:
 * src/utilities/qss_data_arrays.h:24-73).
 *
 * Ownership follows allocateMemoryForQSSDataArrays (reference:
 * src/utilities/qss_data_arrays.c:116-209): a field equal to
 * QSSLIB_SERIAL_dummy_pointer ((T*)-1) is allocated by the call, NULL is
 * left alone, any other pointer is caller-owned.
 *
 * The GPU library keeps device mirrors of these arrays inside a session keyed
 * to the record; the host arrays are authoritative on entry to and exit from
 * every entry point.
 */
#ifndef LSMPQS_QSS_DATA_ARRAYS_H
#define LSMPQS_QSS_DATA_ARRAYS_H

#include "QSSLIB_config.h"
#include "qss_grid.h"

#ifdef __cplusplus
extern "C" {
#endif

#define QSSLIB_SERIAL_dummy_pointer        ((QSSLIB_REAL*)(-1))
#define QSSLIB_SERIAL_dummy_pointer_int    ((int*)(-1))
#define QSSLIB_SERIAL_dummy_pointer_uchar  ((unsigned char*)(-1))

typedef struct _QSS_DataArrays {
  QSSLIB_REAL *phi, *phi_next, *phi_extra;   /* level set, stage buffer, reservoir box */
  QSSLIB_REAL *phi_prev;                     /* phi at the start of an outer step */
  QSSLIB_REAL *mask;                         /* solid signed distance, > 0 in solid */
  QSSLIB_REAL *lse_rhs;                      /* right-hand side of the level-set equation */
  QSSLIB_REAL *normal_velocity;              /* per-voxel a (var path) */
  QSSLIB_REAL *curvature_coeff;              /* per-voxel b (var path) */
  QSSLIB_REAL *external_velocity_x;          /* advective V (var path) */
  QSSLIB_REAL *external_velocity_y;
  QSSLIB_REAL *external_velocity_z;
  QSSLIB_REAL *mask_x, *mask_y, *mask_z;     /* central gradient of mask */
  QSSLIB_REAL *scratch1, *scratch2;          /* stage buffers */
  int         *connectivity;                 /* component labels (0 = background) */
  unsigned char *phi_bin;                    /* 1 where the phase is present */
  QSSLIB_REAL *mask_w, *mask_nw, *mask_disconn_init;  /* trapped-phase masks */
  QSSLIB_REAL *theta;                        /* per-voxel contact angle (radians) */
  QSSLIB_REAL *overlap;                      /* per-voxel overlap (var theta) */
} QSS_DataArrays;

QSS_DataArrays *allocateQSSDataArrays(void);
void allocateMemoryForQSSDataArrays(QSS_DataArrays *p, Grid *g);
void destroyQSSDataArrays(QSS_DataArrays *p);

#ifdef __cplusplus
}
#endif

#endif
