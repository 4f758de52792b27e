/*
 * QSSLIB_config.h -- compile-time precision of the LSMPQS C ABI.
 *
 * The reference selects the floating type of every array and scalar of its
 * C API with the QSSLIB_REAL macro (reference: config/QSSLIB_config.h:15-21,
 * default float).  This library is built twice, once per value, and the two
 * shared objects export the same entry-point names:
 *
 *   liblsmpqs_hip_f32.so   QSSLIB_REAL = float   (FP32 storage and stencils)
 *   liblsmpqs_hip_f64.so   QSSLIB_REAL = double  (FP64 end to end)
 *
 * A caller compiles against these headers with the matching -DQSSLIB_REAL.
 * The index-range macros below are the ones the reference's drivers use when
 * they pass a Grid to the Fortran-ABI kernels (reference:
 * config/QSSLIB_config.h:43-45).
 */
#ifndef LSMPQS_QSSLIB_CONFIG_H
#define LSMPQS_QSSLIB_CONFIG_H

#include <float.h>

#ifndef QSSLIB_REAL
#define QSSLIB_REAL float
#endif

#ifndef LSMLIB_REAL
#define LSMLIB_REAL QSSLIB_REAL
#endif

/* threshold below which |vel_n| and |grad phi|^2 count as zero in the kernels */
#ifndef QSSLIB_ZERO_TOL
#define QSSLIB_ZERO_TOL 1.e-5
#endif

#define GB_DIMS  &(g->ilo_gb),&(g->ihi_gb),&(g->jlo_gb),&(g->jhi_gb),&(g->klo_gb),&(g->khi_gb)
#define FB_DIMS  &(g->ilo_fb),&(g->ihi_fb),&(g->jlo_fb),&(g->jhi_fb),&(g->klo_fb),&(g->khi_fb)

#endif
