/*
 * qss_grid.h -- the Grid record of the LSMPQS C ABI.
 *
 * Field order and types follow the reference struct so that a Grid built by
 * reference-side code (readGridFromBinaryFile, createGridSetDx) can be passed
 * to this library unchanged (reference: src/geometry/qss_grid.h:29-81).
 *
 * Index convention: arrays cover the ghost box, column-major with x fastest,
 * idx = i + j*grid_dims_ghostbox[0] + k*grid_dims_ghostbox[0]*grid_dims_ghostbox[1]
 * (reference: src/geometry/qss_initialization3d.c:136-141).  The fill box
 * [ilo_fb..ihi_fb] x [jlo_fb..jhi_fb] x [klo_fb..khi_fb] holds the evolved
 * voxels; the rest are ghost layers set by boundary extrapolation.
 */
#ifndef LSMPQS_QSS_GRID_H
#define LSMPQS_QSS_GRID_H

#include "QSSLIB_config.h"

#ifdef __cplusplus
extern "C" {
#endif

/* spatial accuracy level (reference: src/geometry/qss_grid.h:92-93);
   the number of ghost layers per level is {2,3,5,4} (qss_grid.c:47) */
typedef enum { LOW = 0, MEDIUM = 1, HIGH = 2, VERY_HIGH = 3 }
  QSSLIB_SPATIAL_DERIVATIVE_ACCURACY_TYPE;

typedef struct _Grid {
  int          num_dims;               /* 2 or 3 */
  QSSLIB_REAL  x_lo[3];                /* physical extent of the fill box */
  QSSLIB_REAL  x_hi[3];
  QSSLIB_REAL  x_lo_ghostbox[3];       /* physical extent including ghosts */
  QSSLIB_REAL  x_hi_ghostbox[3];
  int          grid_dims[3];           /* fill-box voxels per axis */
  int          grid_dims_ghostbox[3];  /* ghost-box voxels per axis */
  QSSLIB_REAL  dx[3];                  /* spacing per axis */
  int          num_gridpts;            /* product of grid_dims_ghostbox */
  int ilo_gb, ihi_gb, jlo_gb, jhi_gb, klo_gb, khi_gb;   /* ghost box */
  int ilo_fb, ihi_fb, jlo_fb, jhi_fb, klo_fb, khi_fb;   /* fill box */
  int ilo_D1_fb, ihi_D1_fb, jlo_D1_fb, jhi_D1_fb, klo_D1_fb, khi_D1_fb;
  int ilo_D2_fb, ihi_D2_fb, jlo_D2_fb, jhi_D2_fb, klo_D2_fb, khi_D2_fb;
  int ilo_D3_fb, ihi_D3_fb, jlo_D3_fb, jhi_D3_fb, klo_D3_fb, khi_D3_fb;
  int          num_nb_levels;          /* narrow-band levels (unused here) */
  unsigned char mark_gb, mark_D1, mark_D2, mark_D3, mark_fb;
  QSSLIB_REAL  beta, gamma;            /* narrow-band widths (unused here) */
} Grid;

#ifdef __cplusplus
}
#endif

#endif
