/*
 * lsmpqs_hip.h -- the C ABI of the MI355X LSMPQS level-set library.
 *
 * liblsmpqs_hip_f32.so / liblsmpqs_hip_f64.so export the entry points below
 * under the reference's own names, so C code written against the reference's
 * headers links against this library unchanged (same QSSLIB_REAL).  Each
 * declaration cites the reference interface it replaces.  All arrays are
 * HOST arrays over the ghost box; the library stages them to the GPU, runs
 * there, and writes back what the reference would have modified.  Scratch
 * arrays of QSS_DataArrays (lse_rhs, scratch1/2, phi_next) are not written.
 *
 * Errors: the reference has no error codes on this path.  Here
 * constCurvModel3d / constCurvModel3dNoVar return -2 on a device or input
 * error (the reference's -1 keeps its meaning: saturation limit reached);
 * void entry points leave their outputs untouched on error.  The message of
 * the calling thread's last error is lsmpqs_hip_last_error().
 *
 * Entry points are not re-entrant for one Options or QSS_DataArrays record
 * (the reference mutates both); separate records may be used from separate
 * threads.
 */
#ifndef LSMPQS_HIP_H
#define LSMPQS_HIP_H

#include <stdio.h>
#include "QSSLIB_config.h"
#include "qss_grid.h"
#include "qss_options.h"
#include "qss_data_arrays.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- evolution (reference: src/level_set_evolution/constCurvModel3d.h:29-30) ----
 * Outer/inner pseudo-time loop at fixed pressure.  Variable-coefficient path
 * (reads normal_velocity, curvature_coeff, external_velocity_*; kernel order
 * from order_space_accur: 2 ENO2, 3/4 ENO3, 5 WENO5).  Returns the pseudo
 * time reached, -1 when the saturation limit is hit, -2 on error. */
QSSLIB_REAL constCurvModel3d(Options *options, QSS_DataArrays *p, Grid *g, FILE *fp_out);
/* Constant-coefficient path (a and options->b; ENO2 for order 2, ENO3 otherwise). */
QSSLIB_REAL constCurvModel3dNoVar(Options *options, QSS_DataArrays *p, Grid *g, FILE *fp_out,
                                  QSSLIB_REAL a);

/* ---- file-level drivers (reference: src/level_set_evolution/drain_top.h:28,
 *      imbibe_top.h:28) ----
 * Read phi (data), mask and grid files (.gz allowed), run drain_model3d /
 * imbibe_model3d with the log in options->outfile, write
 * options->path/data_final.gz when save_data.  Returns 0, or 1 on an input
 * error, a mismatch of the array and grid dimensions or a device error
 * (message in lsmpqs_hip_last_error()).  2D grids run drain_model2d /
 * imbibe_model2d, as the reference does (drain_top.c:116-119). */
int drainTop(Options *options, char *fname_data_in, char *fname_grid_in, char *fname_mask_in);
int imbibeTop(Options *options, char *fname_data_in, char *fname_grid_in, char *fname_mask_in);

/* ---- model drivers (reference: src/level_set_evolution/drain_model3d.h:31,
 *      imbibe_model3d.h:32) ----
 * Curvature schedule a += b*dc up to amax (drain) / down to amin (imbibe).
 * Writes data_step_N.gz in the working directory after each curvature step,
 * as the reference does, unless LSMPQS_NO_STEP_FILES=1. */
void drain_model3d(Options *options, QSS_DataArrays *p, Grid *g, FILE *fp_out);
void imbibe_model3d(Options *options, QSS_DataArrays *p, Grid *g, FILE *fp_out);

/* ---- 2D models (reference: src/level_set_evolution/drain_model2d.h,
 *      imbibe_model2d.h, constCurvModel2d.h) ----
 * The same schedule on a 2D grid (grid_dims_ghostbox = {nx, ny, 1}), with
 * the reference's 2D discretization: HJ-ENO2 / HJ-ENO3 (no WENO5 in 2D),
 * the 2D subcell-fix reinitialization and 8-connected trapping.  Host arrays
 * are copied to the device on entry and back on return.  use_var_theta is
 * rejected (error message set).  constCurvModel2d / constCurvModel2dNoVar
 * return -2 on an error, else what the reference returns (the elapsed time,
 * or -1 when the saturation limit or the vanishing volume stops the run). */
void drain_model2d(Options *options, QSS_DataArrays *p, Grid *g, FILE *fp_out);
void imbibe_model2d(Options *options, QSS_DataArrays *p, Grid *g, FILE *fp_out);
QSSLIB_REAL constCurvModel2d(Options *options, QSS_DataArrays *p, Grid *g, FILE *fp_out);
QSSLIB_REAL constCurvModel2dNoVar(Options *options, QSS_DataArrays *p, Grid *g, FILE *fp_out,
                                  QSSLIB_REAL a);
/* reference: src/reinitialization/reinitialize_top.h (reinitialize_top.c:123-203) */
void reinitializeSubcellFix2d(QSSLIB_REAL *data, Grid *g, Options *o);
/* labels of phi_bin (8-neighbourhood, largest linear index, 0 on background)
   into p->connectivity (reference: src/connectivity/connectivity.h:22) */
void findConnectivity2d(QSS_DataArrays *p, Grid *g);
/* trapComponents_mask, getMainInd_new, setThetaOverlap,
   initializeDisconnectedMasks and signedLinearExtrapolationBCqss (index 9
   only; other indices set an error) take 2D grids too, with the reference's
   2D rules. */

/* ---- kernel level, Fortran calling convention (every argument by reference;
 *      arrays over ghost-box bounds gb = ilo_gb..khi_gb with a fill box fb).
 *      Reference symbols of the same names: src/spatial_derivatives/
 *      qss_spatial_derivatives3d.f (getRHS3d2ndConst :302, getRHS3d3rdConst :572,
 *      getRHS3d2nd :1020, getRHS3d3rd :1335, getRHS3d5th :1827,
 *      setVarNorm3d :13, setVarCurvAdv3d :178, getCentralGrad3dOrder2 :2987,
 *      signedLinearExtrapolation3d :3055), src/time_integration/
 *      qss_tvd_runge_kutta3d.f (:2, :61, :131), src/utilities/qss_util3d.f
 *      (qss3dMaxNormDiffLocal :65, imposeMaskParallel :130,
 *      qss3dVolumeRegionPhiLessThanZero :288), src/reinitialization/
 *      qss_reinitialization3d.f (:476, :585).  Each call stages its host
 *      arrays on the GPU and back; arrays of one call must share the ghost-box
 *      bounds.  Errors set lsmpqs_hip_last_error(). ---- */
#define LSMPQS_GB6 const int *, const int *, const int *, const int *, const int *, const int *
void getrhs3d2ndconst_(const QSSLIB_REAL *phi, const QSSLIB_REAL *vel_n, const QSSLIB_REAL *b,
                       QSSLIB_REAL *max_H_over_dX, QSSLIB_REAL *lse_rhs, LSMPQS_GB6, LSMPQS_GB6,
                       const QSSLIB_REAL *dx, const QSSLIB_REAL *dy, const QSSLIB_REAL *dz);
void getrhs3d3rdconst_(const QSSLIB_REAL *phi, const QSSLIB_REAL *vel_n, const QSSLIB_REAL *b,
                       QSSLIB_REAL *max_H_over_dX, QSSLIB_REAL *lse_rhs, LSMPQS_GB6, LSMPQS_GB6,
                       const QSSLIB_REAL *dx, const QSSLIB_REAL *dy, const QSSLIB_REAL *dz);
void getrhs3d2nd_(const QSSLIB_REAL *phi, const QSSLIB_REAL *vel_n, const QSSLIB_REAL *b,
                  const QSSLIB_REAL *vel_x, const QSSLIB_REAL *vel_y, const QSSLIB_REAL *vel_z,
                  QSSLIB_REAL *max_H_over_dX, QSSLIB_REAL *lse_rhs, LSMPQS_GB6, LSMPQS_GB6,
                  const QSSLIB_REAL *dx, const QSSLIB_REAL *dy, const QSSLIB_REAL *dz);
void getrhs3d3rd_(const QSSLIB_REAL *phi, const QSSLIB_REAL *vel_n, const QSSLIB_REAL *b,
                  const QSSLIB_REAL *vel_x, const QSSLIB_REAL *vel_y, const QSSLIB_REAL *vel_z,
                  QSSLIB_REAL *max_H_over_dX, QSSLIB_REAL *lse_rhs, LSMPQS_GB6, LSMPQS_GB6,
                  const QSSLIB_REAL *dx, const QSSLIB_REAL *dy, const QSSLIB_REAL *dz);
void getrhs3d5th_(const QSSLIB_REAL *phi, const QSSLIB_REAL *vel_n, const QSSLIB_REAL *b,
                  const QSSLIB_REAL *vel_x, const QSSLIB_REAL *vel_y, const QSSLIB_REAL *vel_z,
                  QSSLIB_REAL *max_H_over_dX, QSSLIB_REAL *lse_rhs, LSMPQS_GB6, LSMPQS_GB6,
                  const QSSLIB_REAL *dx, const QSSLIB_REAL *dy, const QSSLIB_REAL *dz);
void qss3drk1step_(QSSLIB_REAL *u_next, LSMPQS_GB6, const QSSLIB_REAL *u_cur, LSMPQS_GB6,
                   const QSSLIB_REAL *rhs, LSMPQS_GB6, LSMPQS_GB6, const QSSLIB_REAL *dt);
void qss3dtvdrk2stage2_(QSSLIB_REAL *u_stage2, LSMPQS_GB6, const QSSLIB_REAL *u_stage1,
                        LSMPQS_GB6, const QSSLIB_REAL *u_cur, LSMPQS_GB6,
                        const QSSLIB_REAL *rhs, LSMPQS_GB6, LSMPQS_GB6, const QSSLIB_REAL *dt);
void qss3dtvdrk3stage3_(QSSLIB_REAL *u_next, LSMPQS_GB6, const QSSLIB_REAL *u_stage2,
                        LSMPQS_GB6, const QSSLIB_REAL *u_cur, LSMPQS_GB6,
                        const QSSLIB_REAL *rhs, LSMPQS_GB6, LSMPQS_GB6, const QSSLIB_REAL *dt);
/* one face (0..5 = x-, x+, y-, y+, z-, z+) over the whole ghost range of the other axes */
void signedlinearextrapolation3d_(QSSLIB_REAL *phi, LSMPQS_GB6, LSMPQS_GB6,
                                  const int *bdry_location_idx);
void imposemaskparallel_(QSSLIB_REAL *phi_masked, const QSSLIB_REAL *mask,
                         const QSSLIB_REAL *phi, const QSSLIB_REAL *overlap, LSMPQS_GB6,
                         const int *cur_klo, const int *cur_khi);
/* FP64 accumulation in a fixed tree order (the reference accumulates in
   QSSLIB_REAL in loop order) */
void qss3dvolumeregionphilessthanzero_(QSSLIB_REAL *volume, const QSSLIB_REAL *phi, LSMPQS_GB6,
                                       LSMPQS_GB6, const QSSLIB_REAL *dx, const QSSLIB_REAL *dy,
                                       const QSSLIB_REAL *dz, const QSSLIB_REAL *epsilon);
void qss3dmaxnormdifflocal_(QSSLIB_REAL *max_norm, const QSSLIB_REAL *field1, LSMPQS_GB6,
                            const QSSLIB_REAL *field2, LSMPQS_GB6, LSMPQS_GB6,
                            const QSSLIB_REAL *zone);
void getcentralgrad3dorder2_(QSSLIB_REAL *gx, QSSLIB_REAL *gy, QSSLIB_REAL *gz, LSMPQS_GB6,
                             const QSSLIB_REAL *phi, LSMPQS_GB6, LSMPQS_GB6,
                             const QSSLIB_REAL *dx, const QSSLIB_REAL *dy, const QSSLIB_REAL *dz);
void setvarnorm3d_(const QSSLIB_REAL *mask, const QSSLIB_REAL *mask_x, const QSSLIB_REAL *mask_y,
                   const QSSLIB_REAL *mask_z, QSSLIB_REAL *var_a, const QSSLIB_REAL *a0,
                   const QSSLIB_REAL *ca, LSMPQS_GB6, LSMPQS_GB6, const QSSLIB_REAL *dx);
void setvarcurvadv3d_(const QSSLIB_REAL *mask, const QSSLIB_REAL *mask_x,
                      const QSSLIB_REAL *mask_y, const QSSLIB_REAL *mask_z, QSSLIB_REAL *var_b,
                      QSSLIB_REAL *vel_x, QSSLIB_REAL *vel_y, QSSLIB_REAL *vel_z,
                      QSSLIB_REAL *b_max_over_dx, QSSLIB_REAL *max_U_over_dx,
                      const QSSLIB_REAL *const_b, LSMPQS_GB6, LSMPQS_GB6, const QSSLIB_REAL *dx);
void qss3dcomputedistanceforsubcellfix_(QSSLIB_REAL *distance0, const QSSLIB_REAL *phi0,
                                        LSMPQS_GB6, LSMPQS_GB6, const QSSLIB_REAL *dx,
                                        const QSSLIB_REAL *dy, const QSSLIB_REAL *dz);
void qss3dcomputereinitializationeqnrhssubcellfixorder1_(
    QSSLIB_REAL *reinit_rhs, const QSSLIB_REAL *phi, LSMPQS_GB6, const QSSLIB_REAL *phi0,
    const QSSLIB_REAL *distance0, LSMPQS_GB6, const QSSLIB_REAL *dx, const QSSLIB_REAL *dy,
    const QSSLIB_REAL *dz);

/* ---- kernel level, 2D (Fortran calling convention, 4-int ghost-box and
 *      fill-box bounds; reference: src/spatial_derivatives/
 *      qss_spatial_derivatives2d.f qss2dGetRHS2nd / 2ndconst / 3rd / 3rdconst
 *      and qss2dSignedLinearExtrapolation, src/time_integration/
 *      qss_tvd_runge_kutta2d.f).  The constant forms add no advective term
 *      where the reference adds an unassigned one. ---- */
#define LSMPQS_GB4 const int *, const int *, const int *, const int *
void qss2dgetrhs2nd_(const QSSLIB_REAL *phi, const QSSLIB_REAL *vel_n, const QSSLIB_REAL *b,
                     const QSSLIB_REAL *vel_x, const QSSLIB_REAL *vel_y,
                     QSSLIB_REAL *max_H_over_dX, QSSLIB_REAL *lse_rhs, LSMPQS_GB4, LSMPQS_GB4,
                     const QSSLIB_REAL *dx, const QSSLIB_REAL *dy);
void qss2dgetrhs3rd_(const QSSLIB_REAL *phi, const QSSLIB_REAL *vel_n, const QSSLIB_REAL *b,
                     const QSSLIB_REAL *vel_x, const QSSLIB_REAL *vel_y,
                     QSSLIB_REAL *max_H_over_dX, QSSLIB_REAL *lse_rhs, LSMPQS_GB4, LSMPQS_GB4,
                     const QSSLIB_REAL *dx, const QSSLIB_REAL *dy);
void qss2dgetrhs2ndconst_(const QSSLIB_REAL *phi, const QSSLIB_REAL *vel_n, const QSSLIB_REAL *b,
                          QSSLIB_REAL *max_H_over_dX, QSSLIB_REAL *lse_rhs, LSMPQS_GB4,
                          LSMPQS_GB4, const QSSLIB_REAL *dx, const QSSLIB_REAL *dy);
void qss2dgetrhs3rdconst_(const QSSLIB_REAL *phi, const QSSLIB_REAL *vel_n, const QSSLIB_REAL *b,
                          QSSLIB_REAL *max_H_over_dX, QSSLIB_REAL *lse_rhs, LSMPQS_GB4,
                          LSMPQS_GB4, const QSSLIB_REAL *dx, const QSSLIB_REAL *dy);
void qss2drk1step_(QSSLIB_REAL *u_next, LSMPQS_GB4, const QSSLIB_REAL *u_cur, LSMPQS_GB4,
                   const QSSLIB_REAL *rhs, LSMPQS_GB4, LSMPQS_GB4, const QSSLIB_REAL *dt);
void qss2dtvdrk2stage2_(QSSLIB_REAL *u_stage2, LSMPQS_GB4, const QSSLIB_REAL *u_stage1,
                        LSMPQS_GB4, const QSSLIB_REAL *u_cur, LSMPQS_GB4,
                        const QSSLIB_REAL *rhs, LSMPQS_GB4, LSMPQS_GB4, const QSSLIB_REAL *dt);
void qss2dtvdrk3stage3_(QSSLIB_REAL *u_next, LSMPQS_GB4, const QSSLIB_REAL *u_stage2,
                        LSMPQS_GB4, const QSSLIB_REAL *u_cur, LSMPQS_GB4,
                        const QSSLIB_REAL *rhs, LSMPQS_GB4, LSMPQS_GB4, const QSSLIB_REAL *dt);
/* one edge (0..3 = x-, x+, y-, y+) over the whole ghost range of the other axis */
void qss2dsignedlinearextrapolation_(QSSLIB_REAL *phi, LSMPQS_GB4, LSMPQS_GB4,
                                     const int *bdry_location_idx);

/* ---- trapping (reference: src/connectivity/connectivity.h:23-27) ---- */
void trapComponents_mask(QSS_DataArrays *p, Grid *g, Options *o);
/* labels of phi_bin (26-neighbourhood, each component labelled with its
   largest linear index, 0 on background) into p->connectivity */
void findConnectivity3d(QSS_DataArrays *p, Grid *g);
void getMainInd_new(Options *o, QSS_DataArrays *p, Grid *g);

/* ---- reinitialization (reference: src/reinitialization/reinitialize_top.h:28) ---- */
void reinitializeSubcellFix3d(QSSLIB_REAL *data, Grid *g, Options *o);
/* smoothed-sign reinitialization, TVD "RK2" until options->tmax_r
   (reference: src/reinitialization/reinitialize_top.h:26, reinitialize_top.c:205-295).
   Its HJ-ENO3 uses 1/6 where the reference's coefficient `sixth` is never
   assigned (qss_reinitialization3d.f:109-110); on voxels that take none of
   those branches the result is the reference's. */
void reinitializeMedium3d(QSSLIB_REAL *data, Grid *g, Options *o);

/* ---- setup and boundary helpers (reference: src/utilities/qss_general_util.h:28-37) ---- */
void signedLinearExtrapolationBCqss(QSSLIB_REAL *phi, Grid *g, int bdry_location_idx);
void initializeDisconnectedMasks(QSSLIB_REAL *data, Grid *g);
void createReservoirInlet3d(QSS_DataArrays *p, Grid *g);
void setSpaceDerivFunc(Grid *g, Options *options);
void setThetaOverlap(QSS_DataArrays *p, Grid *g, Options *options);

/* ---- records and files (reference: src/utilities/qss_options.h:213-272,
 *      src/utilities/qss_data_arrays.h:84-180) ---- */
Options *createOptionsDefault(void);
void destroyOptions(Options *options);
void writeDataArrayQSS(QSSLIB_REAL *data, Grid *grid, char *file_name, int zip_status);
/* headerless QSSLIB_REAL[num_gridpts] (reference: src/utilities/qss_data_arrays.h:151) */
void writeDataArrayRaw(QSSLIB_REAL *data, Grid *grid, char *file_name, int zip_status);
QSSLIB_REAL *readDataArrayQSS(int *grid_dims_ghostbox, char *file_name);
/* option file "key value" lines (reference: src/utilities/qss_options.h:227,
   qss_options.c:118-494; keys recognised by their first letters as there) */
Options *createOptionsFromInputFile(char *file_name);
void printOptions(Options *options, FILE *fp);

/* ---- grids (reference: src/geometry/qss_grid.h:107-255) ----
 * accuracy LOW / MEDIUM / HIGH / VERY_HIGH adds 2 / 3 / 5 / 4 ghost layers
 * and sets the ENO1 / ENO2 / ENO3 / WENO5 fill box.  The binary file is the
 * reference's field-by-field record (zip_status 1: gzip afterwards). */
Grid *createGridSetDx(int num_dims, QSSLIB_REAL dx, QSSLIB_REAL *x_lo, QSSLIB_REAL *x_hi,
                      QSSLIB_SPATIAL_DERIVATIVE_ACCURACY_TYPE accuracy);
Grid *createGridSetDxDyDz(int num_dims, QSSLIB_REAL dx, QSSLIB_REAL dy, QSSLIB_REAL dz,
                          QSSLIB_REAL *x_lo, QSSLIB_REAL *x_hi,
                          QSSLIB_SPATIAL_DERIVATIVE_ACCURACY_TYPE accuracy);
Grid *createGridSetGridDims(int num_dims, int *grid_dims, QSSLIB_REAL *x_lo, QSSLIB_REAL *x_hi,
                            QSSLIB_SPATIAL_DERIVATIVE_ACCURACY_TYPE accuracy);
void destroyGrid(Grid *grid);
void printGrid(Grid *grid, FILE *fp);
void writeGridToBinaryFile(Grid *grid, char *file_name, int zip_status);
Grid *readGridFromBinaryFile(char *file_name);

/* ---- extensions with no counterpart in the reference ---- */

/* message of the calling thread's last error ("" if none) */
const char *lsmpqs_hip_last_error(void);

/* Number of z-slabs whose last voxels set the time step.  The reference's
   dt uses H/dx at the last fill voxel of each OpenMP thread's slab, so it
   depends on the thread count; pass the OMP_NUM_THREADS a reference run used.
   0 (default): LSMPQS_DT_SLABS, else OMP_NUM_THREADS, else the CPU count. */
int lsmpqs_set_dt_slabs(int nslabs);

/* Reference-behaviour switches, all set by default; returns the old value.
   1: ENO3 x line reads phi(i+3,j,j); 2: WENO5 normal term with + sign;
   4: |vel_n| < 1e-5 reuses the previous voxel's normal term.
   8 (not set by default): the FP32 library evaluates every expression in the
   reference's order with IEEE divisions, instead of its fast WENO5 (one
   reciprocal per reconstruction, shared z faces) and FMA-contracted curvature. */
unsigned lsmpqs_set_quirks(unsigned mask);

/* Batched drain_model3d (drain != 0) / imbibe_model3d of n independent boxes
   (subvolumes of any size, each with its own QSS_DataArrays, Grid and
   Options, hence its own curvature schedule).  No counterpart in the
   reference, which runs one box per process.  Each box runs the single-box
   driver on a host thread of its own; the boxes form up to four groups of
   consecutive indices (LSMPQS_BATCH_GROUPS, default 4), each with its own
   stream, and the inner time steps of all boxes of a group waiting for one
   run together: a batched time-step probe, per RK stage one
   persistent stage kernel fed by a work queue of 64 x 8 x zc voxel bricks of
   all boxes, and a batched ghost extrapolation, captured once as a HIP graph
   and replayed.  zc and the number of persistent workgroups are picked on
   the first step by timing.  The device work of the outer steps (trapping,
   reinitialization, merge and floor, reductions) of the boxes waiting for
   one likewise runs as one launch per pass over all of them, when trapping
   is on and the reinitialization is a single iteration.  Results equal the
   boxes' individual runs bit for bit.  All boxes must share the spatial scheme, the time order and the
   contact-angle path (theta > 0 or not).  fp_out: one log per box or NULL;
   no data_step_N files are written.  Returns 0, or -2 with the message in
   lsmpqs_hip_last_error(). */
typedef struct {
  int boxes;
  long long batched_steps;    /* batched inner steps (graph launches) */
  long long box_steps;        /* inner steps summed over the boxes */
  int brick_depth;            /* z planes per brick chosen by the tuner */
  int workgroups;             /* persistent workgroups chosen by the tuner */
  double batched_ms;          /* GPU time of the batched steps (sum over the
                                 rounds; the groups' rounds may overlap) */
  double autotune_best_ms;    /* fastest first-stage launch seen while tuning */
  long long batched_outer;    /* batched outer steps (launch sequences) */
  long long box_outer;        /* outer steps they served, summed over the boxes */
  double batched_outer_ms;    /* GPU time of the batched outer steps (likewise) */
} lsmpqs_batch_stats;
int lsmpqs_batch_model3d(int n, Options **options, QSS_DataArrays **p, Grid **g, FILE **fp_out,
                         int drain, lsmpqs_batch_stats *stats);

/* Gradient of a saturation loss through the level-set time steps.  No
   counterpart in the reference (which has no derivatives); the forward map is
   the reference's variable-coefficient constCurvModel3d inner step
   (src/level_set_evolution/constCurvModel3d.c:127-230) with
   setVarNorm3dTheta's per-voxel contact angle
   (src/spatial_derivatives/qss_spatial_derivatives3d.f:94-170; the
   reference's drain_model3d.c:124-131 never reaches it, its test
   ~(use_var_theta) being nonzero for 0 and 1) at a fixed time step: nsteps TVD-RK steps of order options->order_time_accur with
   the scheme options->order_space_accur (2, 3 or 5), the library's ghost
   extrapolation and the imposition max(phi, mask - overlap) after each
   step and the cap phi = min(phi, phi_cap) (no counterpart in the
   reference, whose inner steps are few between reinitializations: over
   many steps the signed extrapolation grows phi exponentially where the
   mask's advection velocity enters the box through a face, in solid away
   from the interface; the default 6 dx, twice the WENO5 stencil reach,
   leaves the stencils of the |phi| < 3 dx band alone), then
     S_w  = 1 - V(phi < 0) / V(mask < 0)   (smoothed Heaviside, width
                                            options->eps_coefficient dx)
     loss = (S_w - target_sw)^2.
   Reference quirks are off in this map (normal term -vel_n |grad phi| for
   every scheme, no stale normal terms, no phi(i+3,j,j) read); the FP32
   library uses its fast WENO5 and curvature arithmetic.  theta: per-voxel
   contact angle (radians) on the ghost box; p->phi the initial state (ghost
   values used as given), p->mask the pore mask.  On return p->phi holds the
   final state, p->normal_velocity (when not NULL) the per-voxel vel_n of the
   forward map on the fill box, dloss_dtheta the gradient on the fill box and 0 on ghosts,
   args the loss, S_w, the time step and d loss / d b (dloss_dtheta NULL:
   loss only, no reverse pass, dloss_db = 0).  The
   reverse pass transposes the linearisation of every stage with the
   forward's discrete choices (ENO stencils, upwind sides, Godunov branch,
   extrapolation signs, imposition) held fixed.  With a gradient requested
   and every step's state and three stage values in half the free device
   memory, the forward keeps them all and nothing is recomputed; otherwise
   (or with LSMPQS_ADJOINT_STORE=segment, or =none) states are stored every
   checkpoint_every steps and recomputed in between.  Returns 0, or -2 with
   the message in lsmpqs_hip_last_error(). */
typedef struct {
  int nsteps;                 /* in: time steps, >= 1 */
  int checkpoint_every;       /* in: steps between stored states, >= 1 */
  double dt;                  /* in: fixed time step; <= 0: CFL rule on the initial state */
  double a0;                  /* in: pressure coefficient */
  double b;                   /* in: curvature coefficient */
  double target_sw;           /* in: target wetting saturation */
  double phi_cap;             /* in: cap on phi after every step; <= 0: 6 dx */
  double loss;                /* out */
  double sw;                  /* out: final wetting saturation */
  double dt_used;             /* out */
  double dloss_db;            /* out: d loss / d b */
  double cap_used;            /* out: the cap applied */
} lsmpqs_adjoint_args;
int lsmpqs_adjoint_run(Options *options, QSS_DataArrays *p, Grid *g, const QSSLIB_REAL *theta,
                       QSSLIB_REAL *dloss_dtheta, lsmpqs_adjoint_args *args);

/* ---- pre- and post-processing (the computations of the reference's
 *      src/applications programs; files stay with the caller) ---- */
/* createMaskUBC (src/applications/createMaskUBC.c:54-116): a segmented image
   (nx*ny*nz bytes, x fastest, 255 = grain) becomes the grid createMaskUBC
   builds (createGridSetGridDims, x_lo = 3 options->dx, accuracy
   options->accuracy_id, which must give 3 ghost layers) and the mask +dx in
   grain, -dx in pore, 0 on ghosts, reinitialized by the subcell-fix iteration
   on the GPU until options->tmax_r.  *grid_out (destroyGrid) and *mask_out
   (free) are allocated here.  Returns 0 or -2. */
int lsmpqs_mask_from_image(const unsigned char *image, int nx, int ny, int nz, Options *options,
                           Grid **grid_out, QSSLIB_REAL **mask_out);
/* createInitFromMask (src/applications/createInitFromMask.c:75-100): the
   non-wetting sphere (radius 0.7 L centred 0.49 L before the x-low face; with
   options->center_inlet radius 90 dx at the centre, x centre taken from the y
   bounds as the reference does), then max(mask, phi), into data (ghost box).
   Returns 0 or -2. */
int lsmpqs_init_from_mask(const QSSLIB_REAL *mask, Grid *g, Options *options, QSSLIB_REAL *data);
/* makeDiffSteps (src/applications/makeDiffSteps.c:70-140): diff =
   max(-step_a, step_b - dx/2) over the ghost box; its negative voxels plus
   the inlet/outlet ghost slabs are labelled on the GPU (26-neighbourhood); if
   the inlet and outlet voxels at mid j, k share a component, connected_only
   (may be NULL) receives diff with the sign flipped where diff <= 0 outside
   it, and 1 is returned ("percolates"); 0 when they do not, -2 on error. */
int lsmpqs_diff_steps(const QSSLIB_REAL *step_a, const QSSLIB_REAL *step_b, Grid *g,
                      QSSLIB_REAL *diff, QSSLIB_REAL *connected_only);

/* Allocation-free volume of the region phi < 0.  No counterpart in the
   reference: it is for callers that cannot allocate device memory (code
   under HIP graph capture, tight serving loops, harnesses that forbid
   hipMalloc).  The same FP64 sum as qss3dvolumeregionphilessthanzero_
   (smoothed Heaviside of width eps, dV = dx dy dz) over DEVICE arrays:
   phi is a device array over the ghost box gb = {ilo, ihi, jlo, jhi, klo,
   khi} with fill box fb (same convention as the Fortran-ABI routine), and
   *volume is a device scalar, written asynchronously on `stream` (a
   hipStream_t; NULL: the null stream).  The call allocates nothing: its
   reduction partials and result live in a library-owned 32.8 KB device
   array, one per GPU, held until the process ends; the result does not
   depend on the box size, which may be anything up to 2^31 ghost voxels.
   Calls that overlap on one GPU (concurrent streams or host threads) race
   on that array.  Returns 0, or -2 with the message in
   lsmpqs_hip_last_error(). */
int lsmpqs_volume_noalloc(QSSLIB_REAL *volume, const QSSLIB_REAL *phi, const int *gb,
                          const int *fb, QSSLIB_REAL dx, QSSLIB_REAL dy, QSSLIB_REAL dz,
                          QSSLIB_REAL eps, void *stream);

/* free the device state kept for record p */
void lsmpqs_release(QSS_DataArrays *p);

/* Benchmark / test hook: run exactly nsteps inner time steps (probe, RK
   stages, ghost extrapolation, mask imposition) on the device state of p.
   upload != 0 first copies phi, mask, mask_w, mask_nw and (var_path) the
   coefficient arrays from the host.  Times with HIP events on the library's
   stream. */
typedef struct {
  int profile;                /* in: also time each stage kernel */
  double elapsed_ms;          /* out: event-to-event time of the nsteps */
  double stage_ms[3];         /* out: summed stage-kernel time, stages 1..3 */
  long long stage_launches[3];
  long long inner_steps;
  double last_dt;
  int nstale;                 /* voxels reusing a previous normal term */
  long long outer_steps;      /* lsmpqs_session_stats: outer steps since the last model call */
  double inner_loop_ms;       /* lsmpqs_session_stats: GPU time inside inner loops */
} lsmpqs_timing;
int lsmpqs_bench_inner_steps(Options *o, QSS_DataArrays *p, Grid *g, int nsteps, int var_path,
                             QSSLIB_REAL a, int upload, lsmpqs_timing *out);
/* Run the set-up half of drain_model3d (drain != 0) or imbibe_model3d on
   p->phi and p->mask, plus the coefficients and trapped masks of the first
   curvature step (a = options->a), leaving the state on the device for
   lsmpqs_bench_inner_steps(..., upload = 0, ...). */
int lsmpqs_model_prepare(Options *o, QSS_DataArrays *p, Grid *g, int drain);
/* counters of the last drain/imbibe/constCurvModel3d call on record p
   (inner_steps, outer_steps, inner_loop_ms) */
int lsmpqs_session_stats(QSS_DataArrays *p, Grid *g, lsmpqs_timing *out);
/* copy one device field ("phi", "mask_w", "mask_nw") back to p */
int lsmpqs_download(QSS_DataArrays *p, Grid *g, const char *field);

#ifdef __cplusplus
}
#endif

#endif
