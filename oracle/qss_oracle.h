/*
 * qss_oracle.h -- CPU restatement of the LSMPQS 3D hot path (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY: called by tests/, the smoke check and the
 * CPU-baseline leg of bench.py as a checker; never by the product library.
 *
 * Kernel-level routines keep the reference's Fortran by-reference argument
 * lists (see qss_oracle_kernels.c); driver-level routines keep the C
 * signatures of the reference and add one trailing argument, nthreads: the
 * number of OpenMP z-slabs the reference would split the fill box into, on
 * which its time step depends (reference:
 * src/level_set_evolution/constCurvModel3d.c:105-108,139-149).
 */
#ifndef LSMPQS_QSS_ORACLE_H
#define LSMPQS_QSS_ORACLE_H

#include <stdio.h>
#include "QSSLIB_config.h"
#include "qss_grid.h"
#include "qss_options.h"
#include "qss_data_arrays.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef QSSLIB_REAL orc_real;

/* kernel level (qss_oracle_kernels.c); argument lists as in the reference */
#define ORC_B6 const int *, const int *, const int *, const int *, const int *, const int *
void orc_getrhs3d2ndconst(const orc_real *, const orc_real *, const orc_real *, orc_real *,
                          orc_real *, ORC_B6, ORC_B6, const orc_real *, const orc_real *,
                          const orc_real *);
void orc_getrhs3d3rdconst(const orc_real *, const orc_real *, const orc_real *, orc_real *,
                          orc_real *, ORC_B6, ORC_B6, const orc_real *, const orc_real *,
                          const orc_real *);
void orc_getrhs3d2nd(const orc_real *, const orc_real *, const orc_real *, const orc_real *,
                     const orc_real *, const orc_real *, orc_real *, orc_real *, ORC_B6, ORC_B6,
                     const orc_real *, const orc_real *, const orc_real *);
void orc_getrhs3d3rd(const orc_real *, const orc_real *, const orc_real *, const orc_real *,
                     const orc_real *, const orc_real *, orc_real *, orc_real *, ORC_B6, ORC_B6,
                     const orc_real *, const orc_real *, const orc_real *);
void orc_getrhs3d5th(const orc_real *, const orc_real *, const orc_real *, const orc_real *,
                     const orc_real *, const orc_real *, orc_real *, orc_real *, ORC_B6, ORC_B6,
                     const orc_real *, const orc_real *, const orc_real *);
void orc_qss3drk1step(orc_real *, ORC_B6, const orc_real *, ORC_B6, const orc_real *, ORC_B6,
                      ORC_B6, const orc_real *);
void orc_qss3dtvdrk2stage2(orc_real *, ORC_B6, const orc_real *, ORC_B6, const orc_real *,
                           ORC_B6, const orc_real *, ORC_B6, ORC_B6, const orc_real *);
void orc_qss3dtvdrk3stage3(orc_real *, ORC_B6, const orc_real *, ORC_B6, const orc_real *,
                           ORC_B6, const orc_real *, ORC_B6, ORC_B6, const orc_real *);
void orc_signedlinearextrapolation3d(orc_real *, ORC_B6, ORC_B6, const int *);
void orc_qss3dmaxnormdifflocal(orc_real *, const orc_real *, ORC_B6, const orc_real *, ORC_B6,
                               ORC_B6, const orc_real *);
void orc_imposemaskparallel(orc_real *, const orc_real *, const orc_real *, const orc_real *,
                            ORC_B6, const int *, const int *);
void orc_qss3dvolumeregionphilessthanzero(orc_real *, const orc_real *, ORC_B6, ORC_B6,
                                          const orc_real *, const orc_real *, const orc_real *,
                                          const orc_real *);
void orc_qssheaviside(orc_real *, const orc_real *, const orc_real *);
void orc_imposetrapvelocity3d(const orc_real *, const orc_real *, orc_real *, orc_real *,
                              orc_real *, orc_real *, orc_real *, ORC_B6, ORC_B6, const orc_real *);
void orc_getcentralgrad3dorder2(orc_real *, orc_real *, orc_real *, ORC_B6, const orc_real *,
                                ORC_B6, ORC_B6, const orc_real *, const orc_real *,
                                const orc_real *);
void orc_setvarnorm3d(const orc_real *, const orc_real *, const orc_real *, const orc_real *,
                      orc_real *, const orc_real *, const orc_real *, ORC_B6, ORC_B6,
                      const orc_real *);
void orc_setvarcurvadv3d(const orc_real *, const orc_real *, const orc_real *, const orc_real *,
                         orc_real *, orc_real *, orc_real *, orc_real *, orc_real *, orc_real *,
                         const orc_real *, ORC_B6, ORC_B6, const orc_real *);
void orc_qss3dcomputedistanceforsubcellfix(orc_real *, const orc_real *, ORC_B6, ORC_B6,
                                           const orc_real *, const orc_real *, const orc_real *);
void orc_qss3dcomputereinitializationeqnrhssubcellfixorder1(
    orc_real *, const orc_real *, ORC_B6, const orc_real *, const orc_real *, ORC_B6,
    const orc_real *, const orc_real *, const orc_real *);

void orc_computereinitializationeqnrhs3d(orc_real *, const orc_real *, const orc_real *, ORC_B6,
                                         ORC_B6, const orc_real *, const orc_real *,
                                         const orc_real *, const int *, const orc_real *);

/* driver level (qss_oracle_driver.c) */
void orc_signedLinearExtrapolationBCqss(orc_real *phi, Grid *g, int bdry_location_idx);
void orc_reinitializeSubcellFix3d(orc_real *data, Grid *g, Options *o);
/* smoothed-sign reinitialization; sixth: value of the reference's undefined
   ENO3 coefficient (see orc_computereinitializationeqnrhs3d).  Pinned against
   the reference build only on voxels whose ENO3 takes no branch using it; on
   the others the reference's result is undefined behaviour (parity with the
   reference unpinned there; closed forms pin the value 1/6). */
void orc_reinitializeMedium3d(orc_real *data, Grid *g, Options *o, orc_real sixth);
void orc_findConnectivity3d(QSS_DataArrays *p, Grid *g);
void orc_getMainInd_new(Options *o, QSS_DataArrays *p, Grid *g);
void orc_trapComponents_mask(QSS_DataArrays *p, Grid *g, Options *o);
void orc_setThetaOverlap(Options *o, Grid *g);
int  orc_rhs_scheme(Options *o);
orc_real orc_constCurvModel3d(Options *o, QSS_DataArrays *p, Grid *g, FILE *fp, int nthreads);
orc_real orc_constCurvModel3dNoVar(Options *o, QSS_DataArrays *p, Grid *g, FILE *fp,
                                   orc_real a, int nthreads);
void orc_drain_model3d(Options *o, QSS_DataArrays *p, Grid *g, FILE *fp, int nthreads);
void orc_imbibe_model3d(Options *o, QSS_DataArrays *p, Grid *g, FILE *fp, int nthreads);
void orc_model_setup3d(Options *o, QSS_DataArrays *p, Grid *g, int drain);
/* pre- and post-processing computations (reference: src/applications/
 * createMaskUBC.c, createInitFromMask.c, makeDiffSteps.c) */
void orc_mask_from_image(const unsigned char *image, int nx_vox, int ny_vox, int nz_vox,
                         Grid *g, Options *o, orc_real *mask);
void orc_init_from_mask(const orc_real *mask, Grid *g, Options *o, orc_real *data);
int orc_diff_steps(const orc_real *a, const orc_real *b, Grid *g, orc_real *diff,
                   orc_real *connected_only);

void orc_createSphere(orc_real *phi, orc_real cx, orc_real cy, orc_real cz, orc_real r,
                      int inside_flag, Grid *g);
void orc_createBox(orc_real *phi, orc_real x0, orc_real y0, orc_real z0,
                   orc_real lx, orc_real ly, orc_real lz, int inside_flag, Grid *g);

/* per-outer-step record written by the drivers, for tests */
typedef struct {
  int outer_steps;        /* outer steps taken by the last constCurvModel3d call */
  int inner_steps;        /* inner RK steps summed over that call */
  double last_dt;         /* last inner time step */
  double satn;            /* non-wetting saturation after the last outer step */
  double max_abs_err;     /* steady-state measure after the last outer step */
  int stale_norm_voxels;  /* voxels whose normal term reused a previous value */
} orc_stats;
void orc_get_stats(orc_stats *s);

/* Test switches (no counterpart in the reference).  orc_set_quirks: bits 1
   (ENO3 x line reads phi(i+3,j,j)), 2 (WENO5 normal term with + sign), 4
   (|vel_n| < 1e-5 reuses the previous voxel's normal term), all set by
   default; a cleared bit gives the textbook scheme.  orc_set_connectivity:
   26 (default, the reference's rule) or 6 (face neighbours).  Both return
   the previous value. */
unsigned orc_set_quirks(unsigned q);
int orc_set_connectivity(int c);

#ifdef __cplusplus
}
#endif

#endif
