/*
 * qss_options.h -- the Options record of the LSMPQS C ABI.
 *
 * Field order and types follow the reference struct (reference:
 * src/utilities/qss_options.h:34-168) so records created by reference-side
 * code (createOptionsDefault, createOptionsFromInputFile) are accepted as is.
 * The three function-pointer slots keep their places; this library ignores
 * their values and selects its own kernels from order_space_accur, theta and
 * use_var_theta exactly as setSpaceDerivFunc does
 * (reference: src/utilities/qss_general_util.c:82-143).
 *
 * The model drivers mutate the record (b_max_over_dx, max_U_over_dx,
 * overlap, conserve_imbibe, phi_w_ind, phi_nw_ind, space_deriv_func), as the
 * reference does, so one record must not be shared by concurrent calls.
 */
#ifndef LSMPQS_QSS_OPTIONS_H
#define LSMPQS_QSS_OPTIONS_H

#include "QSSLIB_config.h"
#include "qss_grid.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct _Options {
  char   outfile[256];          /* log file name */
  char   path[256];             /* output directory */
  QSSLIB_REAL dx;               /* nominal spacing (the Grid's dx is used) */
  QSSLIB_REAL tmax;             /* pseudo-time limit per curvature step */
  QSSLIB_REAL tmax_r;           /* pseudo-time of one reinitialization */
  QSSLIB_REAL tplot;            /* pseudo-time of one outer step */
  char   accuracy[20];
  QSSLIB_SPATIAL_DERIVATIVE_ACCURACY_TYPE accuracy_id;
  QSSLIB_REAL a;                /* pressure (normal speed) term */
  QSSLIB_REAL b;                /* curvature coefficient */
  int    save_data;
  int    do_reinit;
  int    do_mask;
  QSSLIB_REAL eps_stop;         /* steady-state threshold on max|dphi| */
  int    subcell_fix;
  void  (*reinit_func)();
  QSSLIB_REAL dc;               /* curvature increment per step */
  QSSLIB_REAL cmax;
  QSSLIB_REAL amax;             /* drainage stops once a >= amax */
  QSSLIB_REAL amin;             /* imbibition stops once a <= amin */
  QSSLIB_REAL vol_frac_max;     /* stop when the NW saturation exceeds this */
  QSSLIB_REAL vol_frac_min;
  int    mask_reinit;           /* reinitialize the mask once at setup */
  int    narrow_band;
  int    init_step;             /* number of the first data_step_N file - 1 */
  QSSLIB_REAL overlap;          /* pore/grain overlap set from theta */
  int    stop_touch;
  QSSLIB_REAL eps_coefficient;  /* Heaviside width in units of dx */
  int    extrapol_id;
  void  (*extrapol_func)();
  void  (*space_deriv_func)();
  int    order_space_accur;     /* 2 ENO2, 3/4 ENO3, 5 WENO5 (theta > 0) */
  int    order_time_accur;      /* 1, 2 or 3 Runge-Kutta stages */
  int    checkpoint;
  int    reservoir_inlet;
  QSSLIB_REAL theta;            /* contact angle in radians */
  QSSLIB_REAL C;
  QSSLIB_REAL b_max_over_dx, max_U_over_dx;  /* time-step bounds, set by the model */
  int    check_connectivity;    /* trapping on (1) or off (0) */
  int    phi_w_ind;             /* linear index of the wetting outlet voxel */
  int    phi_nw_ind;            /* linear index of the non-wetting inlet voxel */
  QSSLIB_REAL err_check_zone;   /* band |phi| < this used by the steady-state test */
  int    use_var_theta;
  QSSLIB_REAL theta_1;
  QSSLIB_REAL theta_2;
  int    print_details;
  int    use_satn_stop;         /* stop an outer loop once saturation stalls */
  int    conserve_imbibe;       /* freeze trapped wetting phase (imbibition) */
  int    center_inlet;
} Options;

#ifdef __cplusplus
}
#endif

#endif
