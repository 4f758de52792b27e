/*
 * qss_oracle_kernels.c -- CPU restatement of the LSMPQS 3D numerical kernels.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle: tests/, the smoke check
 * and the CPU-baseline leg of bench.py call it as a checker.  The product
 * (the HIP library under lsmpqs-1.0_amd/) never links or calls it.
 *
 * Every routine restates one reference kernel with the reference's argument
 * list (Fortran by-reference ABI: every argument is a pointer, arrays are
 * column-major over an explicit ghost box) and the reference's order of
 * floating-point operations, so that the FP64 build matches the reference's
 * FP64 build bit for bit.  Compiled twice: -DQSSLIB_REAL=double and =float.
 * Names carry an orc_ prefix and no trailing underscore so both libraries can
 * be loaded side by side.
 */
#include <math.h>
#include <stdlib.h>
#include "qss_oracle.h"

typedef QSSLIB_REAL R;

/* square root in the working precision: the reference's Fortran takes the
   square root of a default real in that real's precision, so the FP32 build
   must not round through double */
static R wsqrt(R x) { return sizeof(R) == 4 ? (R)sqrtf((float)x) : (R)sqrt((double)x); }

/* ghost-box addressing shared by all kernels */
typedef struct { int ilo, jlo, klo, nx, nxy; long n; } box_t;

static box_t mkbox(const int *ilo, const int *ihi, const int *jlo,
                   const int *jhi, const int *klo, const int *khi)
{
  box_t b;
  b.ilo = *ilo; b.jlo = *jlo; b.klo = *klo;
  b.nx = *ihi - *ilo + 1;
  b.nxy = b.nx * (*jhi - *jlo + 1);
  b.n = (long)b.nxy * (*khi - *klo + 1);
  return b;
}
#define AT(a, bx, i, j, k) \
  (a)[((i) - (bx).ilo) + (long)((j) - (bx).jlo) * (bx).nx + (long)((k) - (bx).klo) * (bx).nxy]

static R rmax(R a, R b) { return a > b ? a : b; }
static R rmin(R a, R b) { return a < b ? a : b; }
static R rmax3(R a, R b, R c) { return rmax(rmax(a, b), c); }

/* ------------------------------------------------------------------------ */
/* One-sided derivatives along one axis.  p[0..6] = phi at offsets -3..+3.   */

/* HJ-ENO2 (reference: src/spatial_derivatives/qss_spatial_derivatives3d.f:386-456) */
static void eno2_1d(const R *p, R inv_d, R *plus, R *minus)
{
  const R half = (R)0.5;
  R D1 = p[3] - p[2];
  R D1p = p[4] - p[3];
  R D2 = p[4] + p[2] - 2 * p[3];
  R D2p = p[5] + p[3] - 2 * p[4];
  R D2m = p[1] + p[3] - 2 * p[2];
  if (fabs(D2) < fabs(D2p)) *plus = (D1p - half * D2) * inv_d;
  else                      *plus = (D1p - half * D2p) * inv_d;
  if (fabs(D2m) < fabs(D2)) *minus = (D1 + half * D2m) * inv_d;
  else                      *minus = (D1 + half * D2) * inv_d;
}

/* HJ-ENO3 (reference: qss_spatial_derivatives3d.f:672-745).  inv_plus and
   inv_minus are separate because the reference scales y- and z-minus by
   1/dx (qss_spatial_derivatives3d.f:823,900). */
static void eno3_1d(const R *p, R inv_plus, R inv_minus, R *plus, R *minus)
{
  const R half = (R)0.5, third = (R)(1.0 / 3.0), sixth = (R)(1.0 / 6.0);
  R D1 = p[3] - p[2], D1p = p[4] - p[3], D1m = p[2] - p[1];
  R D1mm = p[1] - p[0], D1pp = p[5] - p[4], D1ppp = p[6] - p[5];
  R D2 = D1p - D1, D2p = D1pp - D1p, D2m = D1 - D1m;
  R D2pp = D1ppp - D1pp, D2mm = D1m - D1mm;
  R D3 = D2 - D2m, D3p = D2p - D2, D3pp = D2pp - D2p, D3m = D2m - D2mm;
  R v;
  v = D1p;
  if (fabs(D2) < fabs(D2p)) {
    v = v - half * D2;
    if (fabs(D3) < fabs(D3p)) v = v - sixth * D3; else v = v - sixth * D3p;
  } else {
    v = v - half * D2p;
    if (fabs(D3p) < fabs(D3pp)) v = v + third * D3p; else v = v + third * D3pp;
  }
  *plus = v * inv_plus;
  v = D1;
  if (fabs(D2m) < fabs(D2)) {
    v = v + half * D2m;
    if (fabs(D3m) < fabs(D3)) v = v + third * D3m; else v = v + third * D3;
  } else {
    v = v + half * D2;
    if (fabs(D3) < fabs(D3p)) v = v - sixth * D3; else v = v - sixth * D3p;
  }
  *minus = v * inv_minus;
}

/* The same ENO3 with the 1/6 coefficient as a parameter: the reference's
   smoothed-sign reinitialization declares `sixth` without a value
   (qss_reinitialization3d.f:109-110), so what it multiplies is not fixed by
   the source. */
static void eno3_1d_sixth(const R *p, R inv_plus, R inv_minus, R sixth, R *plus, R *minus)
{
  const R half = (R)0.5, third = (R)(1.0 / 3.0);
  R D1 = p[3] - p[2], D1p = p[4] - p[3], D1m = p[2] - p[1];
  R D1mm = p[1] - p[0], D1pp = p[5] - p[4], D1ppp = p[6] - p[5];
  R D2 = D1p - D1, D2p = D1pp - D1p, D2m = D1 - D1m;
  R D2pp = D1ppp - D1pp, D2mm = D1m - D1mm;
  R D3 = D2 - D2m, D3p = D2p - D2, D3pp = D2pp - D2p, D3m = D2m - D2mm;
  R v;
  v = D1p;
  if (fabs(D2) < fabs(D2p)) {
    v = v - half * D2;
    if (fabs(D3) < fabs(D3p)) v = v - sixth * D3; else v = v - sixth * D3p;
  } else {
    v = v - half * D2p;
    if (fabs(D3p) < fabs(D3pp)) v = v + third * D3p; else v = v + third * D3pp;
  }
  *plus = v * inv_plus;
  v = D1;
  if (fabs(D2m) < fabs(D2)) {
    v = v + half * D2m;
    if (fabs(D3m) < fabs(D3)) v = v + third * D3m; else v = v + third * D3;
  } else {
    v = v + half * D2;
    if (fabs(D3) < fabs(D3p)) v = v - sixth * D3; else v = v - sixth * D3p;
  }
  *minus = v * inv_minus;
}

/* HJ-WENO5 with weights 0.1/0.6/0.3 (reference: qss_spatial_derivatives3d.f:1960-2035) */
static R weno5_combine(R v1, R v2, R v3, R v4, R v5)
{
  const R one_third = (R)(1.0 / 3.0), seven_sixths = (R)(7.0 / 6.0);
  const R eleven_sixths = (R)(11.0 / 6.0), one_sixth = (R)(1.0 / 6.0);
  const R five_sixths = (R)(5.0 / 6.0), thirteen_twelfths = (R)(13.0 / 12.0);
  const R one_fourth = (R)0.25, tiny = (R)1.0e-35;
  R eps = (R)1e-6 * rmax(rmax(rmax(rmax(v1 * v1, v2 * v2), v3 * v3), v4 * v4), v5 * v5) + tiny;
  R p1 = one_third * v1 - seven_sixths * v2 + eleven_sixths * v3;
  R p2 = -one_sixth * v2 + five_sixths * v3 + one_third * v4;
  R p3 = one_third * v3 + five_sixths * v4 - one_sixth * v5;
  R t, S1, S2, S3, a1, a2, a3, inv;
  t = v1 - (R)2 * v2 + v3; S1 = thirteen_twelfths * (t * t);
  t = v1 - (R)4 * v2 + (R)3 * v3; S1 = S1 + one_fourth * (t * t);
  t = v2 - (R)2 * v3 + v4; S2 = thirteen_twelfths * (t * t);
  t = v2 - v4; S2 = S2 + one_fourth * (t * t);
  t = v3 - (R)2 * v4 + v5; S3 = thirteen_twelfths * (t * t);
  t = (R)3 * v3 - (R)4 * v4 + v5; S3 = S3 + one_fourth * (t * t);
  t = S1 + eps; a1 = (R)0.1 / (t * t);
  t = S2 + eps; a2 = (R)0.6 / (t * t);
  t = S3 + eps; a3 = (R)0.3 / (t * t);
  inv = (R)1 / (a1 + a2 + a3);
  a1 = a1 * inv; a2 = a2 * inv; a3 = a3 * inv;
  return a1 * p1 + a2 * p2 + a3 * p3;
}

static void weno5_1d(const R *p, R inv_d, R *plus, R *minus)
{
  R D1 = p[3] - p[2], D1p = p[4] - p[3], D1m = p[2] - p[1];
  R D1pp = p[5] - p[4], D1ppp = p[6] - p[5], D1mm = p[1] - p[0];
  *plus = weno5_combine(D1ppp * inv_d, D1pp * inv_d, D1p * inv_d, D1 * inv_d, D1m * inv_d);
  *minus = weno5_combine(D1mm * inv_d, D1m * inv_d, D1 * inv_d, D1p * inv_d, D1pp * inv_d);
}

/* ------------------------------------------------------------------------ */
/* Shared per-voxel pieces of the RHS kernels.                               */

/* Godunov |grad phi|^2 for the sign of vel_n (qss_spatial_derivatives3d.f:460-480) */
static R godunov_sq(R vn, R xp, R xm, R yp, R ym, R zp, R zm)
{
  R a, b, s;
  if (vn > 0) {
    a = rmax(xm, 0); a = a * a; b = rmin(xp, 0); b = b * b; s = rmax(a, b);
    a = rmax(ym, 0); a = a * a; b = rmin(yp, 0); b = b * b; s = s + rmax(a, b);
    a = rmax(zm, 0); a = a * a; b = rmin(zp, 0); b = b * b; s = s + rmax(a, b);
  } else {
    a = rmin(xm, 0); a = a * a; b = rmax(xp, 0); b = b * b; s = rmax(a, b);
    a = rmin(ym, 0); a = a * a; b = rmax(yp, 0); b = b * b; s = s + rmax(a, b);
    a = rmin(zm, 0); a = a * a; b = rmax(zp, 0); b = b * b; s = s + rmax(a, b);
  }
  return s;
}

/* H/dx of the CFL bound (qss_spatial_derivatives3d.f:488-513) */
static R h_over_dx(R vn, R xp, R xm, R yp, R ym, R zp, R zm,
                   R dx, R dy, R dz, R inv_dx, R inv_dy, R inv_dz)
{
  R mdx = rmax3(dx, dy, dz);
  R max_dx_sq = mdx * mdx;
  R xc = rmax(fabs(xp), fabs(xm));
  R yc = rmax(fabs(yp), fabs(ym));
  R zc = rmax(fabs(zp), fabs(zm));
  R ng = wsqrt(xc * xc + yc * yc + zc * zc + max_dx_sq);
  return fabs(vn) / ng * (xc * inv_dx + yc * inv_dy + zc * inv_dz);
}

/* b-free curvature term kappa*|grad phi| from wide central differences
   (qss_spatial_derivatives3d.f:518-556) */
static R curvature_term(const R *phi, box_t bx, int i, int j, int k,
                        R fx, R fy, R fz)
{
  R px = (AT(phi, bx, i + 1, j, k) - AT(phi, bx, i - 1, j, k)) * fx;
  R py = (AT(phi, bx, i, j + 1, k) - AT(phi, bx, i, j - 1, k)) * fy;
  R pz = (AT(phi, bx, i, j, k + 1) - AT(phi, bx, i, j, k - 1)) * fz;
  R c = AT(phi, bx, i, j, k);
  R pxx = (AT(phi, bx, i + 2, j, k) + AT(phi, bx, i - 2, j, k) - 2 * c) * fx * fx;
  R pzz = (AT(phi, bx, i, j, k + 2) + AT(phi, bx, i, j, k - 2) - 2 * c) * fz * fz;
  R pyy = (AT(phi, bx, i, j + 2, k) + AT(phi, bx, i, j - 2, k) - 2 * c) * fy * fy;
  R pxy = (AT(phi, bx, i + 1, j + 1, k) + AT(phi, bx, i - 1, j - 1, k)
           - AT(phi, bx, i - 1, j + 1, k) - AT(phi, bx, i + 1, j - 1, k)) * fx * fy;
  R pxz = (AT(phi, bx, i + 1, j, k + 1) + AT(phi, bx, i - 1, j, k - 1)
           - AT(phi, bx, i - 1, j, k + 1) - AT(phi, bx, i + 1, j, k - 1)) * fx * fz;
  R pyz = (AT(phi, bx, i, j + 1, k + 1) + AT(phi, bx, i, j - 1, k - 1)
           - AT(phi, bx, i, j - 1, k + 1) - AT(phi, bx, i, j + 1, k - 1)) * fy * fz;
  R g2 = px * px + py * py + pz * pz;
  R curv;
  if (g2 < (R)1.e-5) return 0;
  curv = pxx * py * py + pyy * px * px - 2 * pxy * px * py
       + pxx * pz * pz + pzz * px * px - 2 * pxz * px * pz
       + pyy * pz * pz + pzz * py * py - 2 * pyz * py * pz;
  return curv / g2;
}

enum { SCHEME_ENO2 = 2, SCHEME_ENO3 = 3, SCHEME_WENO5 = 5 };

/* Reference-behaviour switches, all set by default (the reference's results):
   1  the ENO3 x line reads phi(i+3,j,j)          (qss_spatial_derivatives3d.f:677,1454)
   2  the WENO5 normal term enters with + sign     (:2276; - for ENO2/ENO3, :484)
   4  where |vel_n| < 1e-5 the normal term of the previous voxel of the loop
      is reused                                    (:459-486)
   With a bit cleared the scheme is the textbook one: phi(i+3,j,k), -vel_n
   |grad phi|_G for every scheme, and no normal term where |vel_n| < 1e-5. */
static unsigned orc_quirk_bits = 7u;

unsigned orc_set_quirks(unsigned q)
{
  unsigned old = orc_quirk_bits;
  orc_quirk_bits = q & 7u;
  return old;
}

/* Gather the 7-point line along each axis and apply the scheme.  The ENO3
   x-line reads phi(i+3,j,j) in place of phi(i+3,j,k), as the reference does
   (qss_spatial_derivatives3d.f:677,1454). */
static void upwind_derivs(const R *phi, box_t bx, int i, int j, int k, int scheme,
                          R inv_dx, R inv_dy, R inv_dz, R d[6])
{
  R lx[7], ly[7], lz[7];
  int o;
  for (o = -3; o <= 3; o++) {
    int io = i + o, jo = j + o, ko = k + o;
    if (scheme == SCHEME_ENO2 && (o == -3 || o == 3)) {
      lx[o + 3] = ly[o + 3] = lz[o + 3] = 0;
      continue;
    }
    lx[o + 3] = AT(phi, bx, io, j, k);
    ly[o + 3] = AT(phi, bx, i, jo, k);
    lz[o + 3] = AT(phi, bx, i, j, ko);
  }
  if (scheme == SCHEME_ENO3 && (orc_quirk_bits & 1u)) {
    /* linear address of phi(i+3,j,j); outside the array (possible only when
       the j range exceeds the k range) the reference reads undefined memory,
       the oracle reads 0 */
    long a = (i + 3 - bx.ilo) + (long)(j - bx.jlo) * bx.nx + (long)(j - bx.klo) * bx.nxy;
    lx[6] = (a >= 0 && a < bx.n) ? phi[a] : 0;
  }
  if (scheme == SCHEME_ENO2) {
    eno2_1d(lx, inv_dx, &d[0], &d[1]);
    eno2_1d(ly, inv_dy, &d[2], &d[3]);
    eno2_1d(lz, inv_dz, &d[4], &d[5]);
  } else if (scheme == SCHEME_ENO3) {
    eno3_1d(lx, inv_dx, inv_dx, &d[0], &d[1]);
    eno3_1d(ly, inv_dy, inv_dx, &d[2], &d[3]);
    eno3_1d(lz, inv_dz, inv_dx, &d[4], &d[5]);
  } else {
    weno5_1d(lx, inv_dx, &d[0], &d[1]);
    weno5_1d(ly, inv_dy, &d[2], &d[3]);
    weno5_1d(lz, inv_dz, &d[4], &d[5]);
  }
}

/* Common body of the five getRHS3d* kernels.  vel_n/b are scalars (const
   path) when vn_arr/b_arr are NULL, per-voxel otherwise; vx/vy/vz are NULL on
   the const path.  norm_sign is -1 for ENO2/ENO3 and +1 for WENO5
   (qss_spatial_derivatives3d.f:484 vs :2276).  temp_norm persists across
   voxels of one call, including its stale reuse where |vel_n| < 1e-5
   (qss_spatial_derivatives3d.f:459-486); its value before the first
   assignment is undefined in the reference and 0 here. */
static void rhs_body(const R *phi, const R *vn_arr, R vn_s, const R *b_arr, R b_s,
                     const R *vx, const R *vy, const R *vz, R *max_H, R *rhs,
                     box_t bx, int ilo, int ihi, int jlo, int jhi, int klo, int khi,
                     R dx, R dy, R dz, int scheme, R norm_sign)
{
  R fx = (R)0.5 / dx, fy = (R)0.5 / dy, fz = (R)0.5 / dz;
  R inv_dx = (R)1.0 / dx, inv_dy = (R)1.0 / dy, inv_dz = (R)1.0 / dz;
  R temp_norm = 0;
  int i, j, k;
  for (k = klo; k <= khi; k++)
    for (j = jlo; j <= jhi; j++)
      for (i = ilo; i <= ihi; i++) {
        R d[6], vn, bb, temp_curv, temp_vel = 0, H;
        upwind_derivs(phi, bx, i, j, k, scheme, inv_dx, inv_dy, inv_dz, d);
        vn = vn_arr ? AT(vn_arr, bx, i, j, k) : vn_s;
        if (fabs(vn) >= (R)1.e-5) {
          R s = godunov_sq(vn, d[0], d[1], d[2], d[3], d[4], d[5]);
          temp_norm = norm_sign * vn * wsqrt(s);
        } else if (!(orc_quirk_bits & 4u)) {
          temp_norm = 0;
        }
        H = h_over_dx(vn, d[0], d[1], d[2], d[3], d[4], d[5], dx, dy, dz,
                      inv_dx, inv_dy, inv_dz);
        *max_H = -1;
        if (H > *max_H) *max_H = H;
        if (vx) {
          R ax = AT(vx, bx, i, j, k) >= 0 ? d[1] : d[0];
          R ay = AT(vy, bx, i, j, k) >= 0 ? d[3] : d[2];
          R az = AT(vz, bx, i, j, k) >= 0 ? d[5] : d[4];
          temp_vel = -(AT(vx, bx, i, j, k) * ax + AT(vy, bx, i, j, k) * ay
                       + AT(vz, bx, i, j, k) * az);
        }
        bb = b_arr ? AT(b_arr, bx, i, j, k) : b_s;
        temp_curv = bb * curvature_term(phi, bx, i, j, k, fx, fy, fz);
        if (vx) AT(rhs, bx, i, j, k) = temp_norm + temp_curv + temp_vel;
        else    AT(rhs, bx, i, j, k) = temp_norm + temp_curv;
      }
}

#define GBARGS const int *ilo_gb, const int *ihi_gb, const int *jlo_gb, \
               const int *jhi_gb, const int *klo_gb, const int *khi_gb
#define FBARGS const int *ilo_fb, const int *ihi_fb, const int *jlo_fb, \
               const int *jhi_fb, const int *klo_fb, const int *khi_fb
#define GB mkbox(ilo_gb, ihi_gb, jlo_gb, jhi_gb, klo_gb, khi_gb)
#define FB *ilo_fb, *ihi_fb, *jlo_fb, *jhi_fb, *klo_fb, *khi_fb

/* reference: qss_spatial_derivatives3d.f:302-567 */
void orc_getrhs3d2ndconst(const R *phi, const R *vel_n, const R *b, R *max_H, R *rhs,
                          GBARGS, FBARGS, const R *dx, const R *dy, const R *dz)
{
  rhs_body(phi, NULL, *vel_n, NULL, *b, NULL, NULL, NULL, max_H, rhs, GB, FB,
           *dx, *dy, *dz, SCHEME_ENO2, -1);
}

/* reference: qss_spatial_derivatives3d.f:572-1011 */
void orc_getrhs3d3rdconst(const R *phi, const R *vel_n, const R *b, R *max_H, R *rhs,
                          GBARGS, FBARGS, const R *dx, const R *dy, const R *dz)
{
  rhs_body(phi, NULL, *vel_n, NULL, *b, NULL, NULL, NULL, max_H, rhs, GB, FB,
           *dx, *dy, *dz, SCHEME_ENO3, -1);
}

/* reference: qss_spatial_derivatives3d.f:1020-1326 */
void orc_getrhs3d2nd(const R *phi, const R *vel_n, const R *b, const R *vx,
                     const R *vy, const R *vz, R *max_H, R *rhs,
                     GBARGS, FBARGS, const R *dx, const R *dy, const R *dz)
{
  rhs_body(phi, vel_n, 0, b, 0, vx, vy, vz, max_H, rhs, GB, FB,
           *dx, *dy, *dz, SCHEME_ENO2, -1);
}

/* reference: qss_spatial_derivatives3d.f:1335-1816 */
void orc_getrhs3d3rd(const R *phi, const R *vel_n, const R *b, const R *vx,
                     const R *vy, const R *vz, R *max_H, R *rhs,
                     GBARGS, FBARGS, const R *dx, const R *dy, const R *dz)
{
  rhs_body(phi, vel_n, 0, b, 0, vx, vy, vz, max_H, rhs, GB, FB,
           *dx, *dy, *dz, SCHEME_ENO3, -1);
}

/* reference: qss_spatial_derivatives3d.f:1827-2386 (normal term with + sign, :2276) */
void orc_getrhs3d5th(const R *phi, const R *vel_n, const R *b, const R *vx,
                     const R *vy, const R *vz, R *max_H, R *rhs,
                     GBARGS, FBARGS, const R *dx, const R *dy, const R *dz)
{
  rhs_body(phi, vel_n, 0, b, 0, vx, vy, vz, max_H, rhs, GB, FB,
           *dx, *dy, *dz, SCHEME_WENO5, (orc_quirk_bits & 2u) ? +1 : -1);
}

/* ------------------------------------------------------------------------ */
/* TVD Runge-Kutta stages (reference: src/time_integration/qss_tvd_runge_kutta3d.f) */

/* :2-56 */
void orc_qss3drk1step(R *u_next, GBARGS, const R *u_cur,
                      const int *a1, const int *a2, const int *a3, const int *a4, const int *a5, const int *a6,
                      const R *rhs,
                      const int *b1, const int *b2, const int *b3, const int *b4, const int *b5, const int *b6,
                      FBARGS, const R *dt)
{
  box_t bx = GB;
  int i, j, k;
  (void)a1; (void)a2; (void)a3; (void)a4; (void)a5; (void)a6;
  (void)b1; (void)b2; (void)b3; (void)b4; (void)b5; (void)b6;
  for (k = *klo_fb; k <= *khi_fb; k++)
    for (j = *jlo_fb; j <= *jhi_fb; j++)
      for (i = *ilo_fb; i <= *ihi_fb; i++)
        AT(u_next, bx, i, j, k) = AT(u_cur, bx, i, j, k) + *dt * AT(rhs, bx, i, j, k);
}

/* :61-126 ("TVD-RK2 stage 2" uses the 3/4, 1/4 weights) */
void orc_qss3dtvdrk2stage2(R *u2, GBARGS, const R *u1,
                           const int *a1, const int *a2, const int *a3, const int *a4, const int *a5, const int *a6,
                           const R *u,
                           const int *b1, const int *b2, const int *b3, const int *b4, const int *b5, const int *b6,
                           const R *rhs,
                           const int *c1, const int *c2, const int *c3, const int *c4, const int *c5, const int *c6,
                           FBARGS, const R *dt)
{
  box_t bx = GB;
  int i, j, k;
  (void)a1; (void)a2; (void)a3; (void)a4; (void)a5; (void)a6;
  (void)b1; (void)b2; (void)b3; (void)b4; (void)b5; (void)b6;
  (void)c1; (void)c2; (void)c3; (void)c4; (void)c5; (void)c6;
  for (k = *klo_fb; k <= *khi_fb; k++)
    for (j = *jlo_fb; j <= *jhi_fb; j++)
      for (i = *ilo_fb; i <= *ihi_fb; i++)
        AT(u2, bx, i, j, k) = (R)0.75 * AT(u, bx, i, j, k)
                            + (R)0.25 * (AT(u1, bx, i, j, k) + *dt * AT(rhs, bx, i, j, k));
}

/* :131-200 */
void orc_qss3dtvdrk3stage3(R *u_next, GBARGS, const R *u2,
                           const int *a1, const int *a2, const int *a3, const int *a4, const int *a5, const int *a6,
                           const R *u,
                           const int *b1, const int *b2, const int *b3, const int *b4, const int *b5, const int *b6,
                           const R *rhs,
                           const int *c1, const int *c2, const int *c3, const int *c4, const int *c5, const int *c6,
                           FBARGS, const R *dt)
{
  box_t bx = GB;
  const R one_third = (R)(1.0 / 3.0), two_thirds = (R)(2.0 / 3.0);
  int i, j, k;
  (void)a1; (void)a2; (void)a3; (void)a4; (void)a5; (void)a6;
  (void)b1; (void)b2; (void)b3; (void)b4; (void)b5; (void)b6;
  (void)c1; (void)c2; (void)c3; (void)c4; (void)c5; (void)c6;
  for (k = *klo_fb; k <= *khi_fb; k++)
    for (j = *jlo_fb; j <= *jhi_fb; j++)
      for (i = *ilo_fb; i <= *ihi_fb; i++)
        AT(u_next, bx, i, j, k) = one_third * AT(u, bx, i, j, k)
                                + two_thirds * (AT(u2, bx, i, j, k) + *dt * AT(rhs, bx, i, j, k));
}

/* ------------------------------------------------------------------------ */
/* Signed linear extrapolation into the ghost layers of one face
   (reference: qss_spatial_derivatives3d.f:3055-3191).  Fortran SIGN(1,x)
   is copysign. */
void orc_signedlinearextrapolation3d(R *phi, GBARGS, FBARGS, const int *face)
{
  box_t bx = GB;
  int i, j, k;
  R s, slope;
  switch (*face) {
  case 0:
    for (k = *klo_gb; k <= *khi_gb; k++)
      for (j = *jlo_gb; j <= *jhi_gb; j++) {
        R c = AT(phi, bx, *ilo_fb, j, k);
        s = copysign((R)1, c);
        slope = s * fabs(c - AT(phi, bx, *ilo_fb + 1, j, k));
        for (i = *ilo_gb; i <= *ilo_fb - 1; i++)
          AT(phi, bx, i, j, k) = c + slope * (R)(*ilo_fb - i);
      }
    break;
  case 1:
    for (k = *klo_gb; k <= *khi_gb; k++)
      for (j = *jlo_gb; j <= *jhi_gb; j++) {
        R c = AT(phi, bx, *ihi_fb, j, k);
        s = copysign((R)1, c);
        slope = s * fabs(c - AT(phi, bx, *ihi_fb - 1, j, k));
        for (i = *ihi_fb + 1; i <= *ihi_gb; i++)
          AT(phi, bx, i, j, k) = c + slope * (R)(i - *ihi_fb);
      }
    break;
  case 2:
    for (i = *ilo_gb; i <= *ihi_gb; i++)
      for (k = *klo_gb; k <= *khi_gb; k++) {
        R c = AT(phi, bx, i, *jlo_fb, k);
        s = copysign((R)1, c);
        slope = s * fabs(c - AT(phi, bx, i, *jlo_fb + 1, k));
        for (j = *jlo_gb; j <= *jlo_fb - 1; j++)
          AT(phi, bx, i, j, k) = c + slope * (R)(*jlo_fb - j);
      }
    break;
  case 3:
    for (i = *ilo_gb; i <= *ihi_gb; i++)
      for (k = *klo_gb; k <= *khi_gb; k++) {
        R c = AT(phi, bx, i, *jhi_fb, k);
        s = copysign((R)1, c);
        slope = s * fabs(c - AT(phi, bx, i, *jhi_fb - 1, k));
        for (j = *jhi_fb + 1; j <= *jhi_gb; j++)
          AT(phi, bx, i, j, k) = c + slope * (R)(j - *jhi_fb);
      }
    break;
  case 4:
    for (i = *ilo_gb; i <= *ihi_gb; i++)
      for (j = *jlo_gb; j <= *jhi_gb; j++) {
        R c = AT(phi, bx, i, j, *klo_fb);
        s = copysign((R)1, c);
        slope = s * fabs(c - AT(phi, bx, i, j, *klo_fb + 1));
        for (k = *klo_gb; k <= *klo_fb - 1; k++)
          AT(phi, bx, i, j, k) = c + slope * (R)(*klo_fb - k);
      }
    break;
  case 5:
    for (i = *ilo_gb; i <= *ihi_gb; i++)
      for (j = *jlo_gb; j <= *jhi_gb; j++) {
        R c = AT(phi, bx, i, j, *khi_fb);
        s = copysign((R)1, c);
        slope = s * fabs(c - AT(phi, bx, i, j, *khi_fb - 1));
        for (k = *khi_fb + 1; k <= *khi_gb; k++)
          AT(phi, bx, i, j, k) = c + slope * (R)(k - *khi_fb);
      }
    break;
  default:
    break;
  }
}

/* ------------------------------------------------------------------------ */
/* Utilities (reference: src/utilities/qss_util3d.f) */

/* :65-125; returns -1 when no voxel of the box lies in the band */
void orc_qss3dmaxnormdifflocal(R *out, const R *f1, GBARGS, const R *f2,
                               const int *a1, const int *a2, const int *a3, const int *a4, const int *a5, const int *a6,
                               const int *ilo, const int *ihi, const int *jlo,
                               const int *jhi, const int *klo, const int *khi,
                               const R *zone)
{
  box_t bx = GB;
  int i, j, k;
  (void)a1; (void)a2; (void)a3; (void)a4; (void)a5; (void)a6;
  *out = -1;
  for (k = *klo; k <= *khi; k++)
    for (j = *jlo; j <= *jhi; j++)
      for (i = *ilo; i <= *ihi; i++)
        if (fabs(AT(f1, bx, i, j, k)) < *zone) {
          R d = fabs(AT(f1, bx, i, j, k) - AT(f2, bx, i, j, k));
          if (d > *out) *out = d;
        }
}

/* :130-175; phi_masked = max(phi, mask - overlap) over planes cur_klo..cur_khi */
void orc_imposemaskparallel(R *phi_masked, const R *mask, const R *phi, const R *overlap,
                            GBARGS, const int *cur_klo, const int *cur_khi)
{
  box_t bx = GB;
  int i, j, k;
  for (k = *cur_klo; k <= *cur_khi; k++)
    for (j = *jlo_gb; j <= *jhi_gb; j++)
      for (i = *ilo_gb; i <= *ihi_gb; i++) {
        R m = AT(mask, bx, i, j, k) - *overlap;
        AT(phi_masked, bx, i, j, k) = (m > AT(phi, bx, i, j, k)) ? m : AT(phi, bx, i, j, k);
      }
}

/* :288-358; smoothed volume of {phi < 0}, accumulated in R in loop order */
void orc_qss3dvolumeregionphilessthanzero(R *volume, const R *phi, GBARGS,
                                          const int *ilo, const int *ihi, const int *jlo,
                                          const int *jhi, const int *klo, const int *khi,
                                          const R *dx, const R *dy, const R *dz, const R *epsilon)
{
  box_t bx = GB;
  const R pi = (R)3.14159265358979323846, one_over_pi = (R)0.31830988618379;
  R dV = *dx * *dy * *dz;
  int i, j, k;
  *volume = 0;
  for (k = *klo; k <= *khi; k++)
    for (j = *jlo; j <= *jhi; j++)
      for (i = *ilo; i <= *ihi; i++) {
        R p = AT(phi, bx, i, j, k);
        R q = p / *epsilon;
        if (p < -*epsilon) *volume = *volume + dV;
        else if (p < *epsilon)
          *volume = *volume + (R)0.5 * ((R)1 - q - one_over_pi * sin(pi * q)) * dV;
      }
}

static R basic_heaviside(R x) { return x < 0 ? 0 : (x > 0 ? 1 : (R)0.5); }

/* :450-470 */
void orc_qssheaviside(R *val, const R *x, const R *eps)
{
  const R pi = (R)3.14159265358979323846, one_over_pi = (R)0.31830988618379;
  R v1 = basic_heaviside(*x + *eps), v2 = basic_heaviside(*eps - *x);
  R v3 = basic_heaviside(*x - *eps);
  *val = ((R)0.5 + *x / (2 * *eps) + ((R)0.5 * one_over_pi) * sin((pi * *x) / *eps)) * v1 * v2 + v3;
}

/* :499-581; zero the coefficients where phi_w > 1.5 dx */
void orc_imposetrapvelocity3d(const R *phi_w, const R *phi_nw, R *vn, R *b, R *vx, R *vy, R *vz,
                              GBARGS, FBARGS, const R *dx)
{
  box_t bx = GB;
  R eps_var = (R)1.5 * *dx;
  int i, j, k;
  (void)phi_nw;
  for (k = *klo_fb; k <= *khi_fb; k++)
    for (j = *jlo_fb; j <= *jhi_fb; j++)
      for (i = *ilo_fb; i <= *ihi_fb; i++)
        if (AT(phi_w, bx, i, j, k) > eps_var) {
          AT(vn, bx, i, j, k) = 0; AT(b, bx, i, j, k) = 0;
          AT(vx, bx, i, j, k) = 0; AT(vy, bx, i, j, k) = 0; AT(vz, bx, i, j, k) = 0;
        }
}

/* ------------------------------------------------------------------------ */
/* Coefficient fields (reference: qss_spatial_derivatives3d.f) */

/* :2987-3048 */
void orc_getcentralgrad3dorder2(R *gx, R *gy, R *gz, GBARGS, const R *phi,
                                const int *a1, const int *a2, const int *a3, const int *a4, const int *a5, const int *a6,
                                FBARGS, const R *dx, const R *dy, const R *dz)
{
  box_t bx = GB;
  R fx = (R)0.5 / *dx, fy = (R)0.5 / *dy, fz = (R)0.5 / *dz;
  int i, j, k;
  (void)a1; (void)a2; (void)a3; (void)a4; (void)a5; (void)a6;
  for (k = *klo_fb; k <= *khi_fb; k++)
    for (j = *jlo_fb; j <= *jhi_fb; j++)
      for (i = *ilo_fb; i <= *ihi_fb; i++) {
        AT(gx, bx, i, j, k) = (AT(phi, bx, i + 1, j, k) - AT(phi, bx, i - 1, j, k)) * fx;
        AT(gy, bx, i, j, k) = (AT(phi, bx, i, j + 1, k) - AT(phi, bx, i, j - 1, k)) * fy;
        AT(gz, bx, i, j, k) = (AT(phi, bx, i, j, k + 1) - AT(phi, bx, i, j, k - 1)) * fz;
      }
}

/* h1, h2, |grad mask| and s shared by setVarNorm3d and setVarCurvAdv3d */
static void coeff_common(const R *mask, const R *mx, const R *my, const R *mz, box_t bx,
                         int i, int j, int k, R dx, R *h1, R *h2, R *absgrad, R *s)
{
  R eps_var = (R)1.5 * dx;
  R m = AT(mask, bx, i, j, k), x, negm = -m;
  orc_qssheaviside(h1, &negm, &eps_var);
  orc_qssheaviside(h2, &m, &eps_var);
  *absgrad = AT(mx, bx, i, j, k) * AT(mx, bx, i, j, k) + AT(my, bx, i, j, k) * AT(my, bx, i, j, k)
           + AT(mz, bx, i, j, k) * AT(mz, bx, i, j, k);
  *absgrad = wsqrt(*absgrad);
  x = m * m + *absgrad * dx * *absgrad * dx;
  x = wsqrt(x);
  *s = m / x;
}

/* :13-87; var_a = h1 a0 - s h2 cos(theta) |grad mask| */
void orc_setvarnorm3d(const R *mask, const R *mx, const R *my, const R *mz, R *var_a,
                      const R *a0, const R *ca, GBARGS, FBARGS, const R *dx)
{
  box_t bx = GB;
  int i, j, k;
  for (k = *klo_fb; k <= *khi_fb; k++)
    for (j = *jlo_fb; j <= *jhi_fb; j++)
      for (i = *ilo_fb; i <= *ihi_fb; i++) {
        R h1, h2, ag, s;
        coeff_common(mask, mx, my, mz, bx, i, j, k, *dx, &h1, &h2, &ag, &s);
        AT(var_a, bx, i, j, k) = h1 * *a0 - s * h2 * (R)1 * cos(*ca) * ag;
      }
}

/* :178-298; var_b = h1 b, V = s grad(mask) h2, and the two time-step bounds */
void orc_setvarcurvadv3d(const R *mask, const R *mx, const R *my, const R *mz, R *var_b,
                         R *vx, R *vy, R *vz, R *b_max_over_dx, R *max_U_over_dx,
                         const R *const_b, GBARGS, FBARGS, const R *dx)
{
  box_t bx = GB;
  R inv = (R)1.0 / *dx, b_max = -1, umax = -1;
  int i, j, k;
  for (k = *klo_fb; k <= *khi_fb; k++)
    for (j = *jlo_fb; j <= *jhi_fb; j++)
      for (i = *ilo_fb; i <= *ihi_fb; i++) {
        R h1, h2, ag, s, u;
        coeff_common(mask, mx, my, mz, bx, i, j, k, *dx, &h1, &h2, &ag, &s);
        AT(var_b, bx, i, j, k) = h1 * *const_b;
        if (AT(var_b, bx, i, j, k) > b_max) b_max = AT(var_b, bx, i, j, k);
        AT(vx, bx, i, j, k) = s * (R)1 * AT(mx, bx, i, j, k) * h2;
        AT(vy, bx, i, j, k) = s * (R)1 * AT(my, bx, i, j, k) * h2;
        AT(vz, bx, i, j, k) = s * (R)1 * AT(mz, bx, i, j, k) * h2;
        u = fabs(AT(vx, bx, i, j, k)) * inv + fabs(AT(vy, bx, i, j, k)) * inv
          + fabs(AT(vz, bx, i, j, k)) * inv;
        if (u > umax) umax = u;
      }
  *max_U_over_dx = umax;
  *b_max_over_dx = 2 * b_max * (inv * inv + inv * inv + inv * inv);
}

/* ------------------------------------------------------------------------ */
/* Subcell-fix reinitialization (reference: src/reinitialization/qss_reinitialization3d.f) */

static int sign_change(R a, R c) { return (a > 0 && c <= 0) || (a <= 0 && c > 0); }

/* :476-579; Russo-Smereka distance near the interface, -1000*max_dx elsewhere */
void orc_qss3dcomputedistanceforsubcellfix(R *dist0, const R *phi0, GBARGS, FBARGS,
                                           const R *dx, const R *dy, const R *dz)
{
  box_t bx = GB;
  R flag = (R)-1000.0 * rmax3(*dx, *dy, *dz);
  int i, j, k;
  for (k = *klo_fb; k <= *khi_fb; k++)
    for (j = *jlo_fb; j <= *jhi_fb; j++)
      for (i = *ilo_fb; i <= *ihi_fb; i++) {
        R c = AT(phi0, bx, i, j, k);
        R xp = AT(phi0, bx, i + 1, j, k), xm = AT(phi0, bx, i - 1, j, k);
        R yp = AT(phi0, bx, i, j + 1, k), ym = AT(phi0, bx, i, j - 1, k);
        R zp = AT(phi0, bx, i, j, k + 1), zm = AT(phi0, bx, i, j, k - 1);
        int near = sign_change(xp, c) || sign_change(xm, c) || sign_change(yp, c)
                || sign_change(ym, c) || sign_change(zp, c) || sign_change(zm, c);
        if (!near) {
          AT(dist0, bx, i, j, k) = flag;
        } else {
          R delta = rmax(rmax(rmax(fabs(xp - xm) / (R)2, fabs(xp - c)), fabs(c - xm)), (R)1.e-5);
          delta = rmax(rmax(rmax(fabs(yp - ym) / (R)2, fabs(yp - c)), fabs(c - ym)), delta);
          delta = rmax(rmax(rmax(fabs(zp - zm) / (R)2, fabs(zp - c)), fabs(c - zm)), delta);
          AT(dist0, bx, i, j, k) = c / delta;
        }
      }
}

/* :585-746; first-order Godunov away from the interface, subcell fix near it */
void orc_qss3dcomputereinitializationeqnrhssubcellfixorder1(
    R *rhs, const R *phi, GBARGS, const R *phi0, const R *dist0, FBARGS,
    const R *dx, const R *dy, const R *dz)
{
  box_t bx = GB;
  R max_dx = rmax3(*dx, *dy, *dz);
  R flag = (R)-1000.0 * max_dx;
  R inv_dx = (R)1.0 / *dx, inv_dy = (R)1.0 / *dy, inv_dz = (R)1.0 / *dz;
  int i, j, k, d;
  for (k = *klo_fb; k <= *khi_fb; k++)
    for (j = *jlo_fb; j <= *jhi_fb; j++)
      for (i = *ilo_fb; i <= *ihi_fb; i++) {
        R c = AT(phi, bx, i, j, k);
        R gp[3], gm[3], star[3], sgn0;
        R p0 = AT(phi0, bx, i, j, k);
        gp[0] = (AT(phi, bx, i + 1, j, k) - c) * inv_dx;
        gm[0] = (c - AT(phi, bx, i - 1, j, k)) * inv_dx;
        gp[1] = (AT(phi, bx, i, j + 1, k) - c) * inv_dy;
        gm[1] = (c - AT(phi, bx, i, j - 1, k)) * inv_dy;
        gp[2] = (AT(phi, bx, i, j, k + 1) - c) * inv_dz;
        gm[2] = (c - AT(phi, bx, i, j, k - 1)) * inv_dz;
        sgn0 = fabs(p0) > (R)1.e-5 ? copysign((R)1, p0) : (R)0;
        if (AT(dist0, bx, i, j, k) == flag) {
          for (d = 0; d < 3; d++) {
            if (c > 0) { gp[d] = rmax(-gp[d], 0); gm[d] = rmax(gm[d], 0); }
            else       { gp[d] = rmax(gp[d], 0);  gm[d] = rmax(-gm[d], 0); }
            star[d] = rmax(gp[d], gm[d]);
          }
          if (fabs(c) > (R)1.e-5) {
            R n2 = star[0] * star[0] + star[1] * star[1] + star[2] * star[2];
            AT(rhs, bx, i, j, k) = sgn0 - sgn0 * wsqrt(n2);
          } else {
            AT(rhs, bx, i, j, k) = 0;
          }
        } else {
          AT(rhs, bx, i, j, k) = AT(dist0, bx, i, j, k) - sgn0 * fabs(c) / max_dx;
        }
      }
}

/* ------------------------------------------------------------------------ */
/* Smoothed-sign reinitialization right-hand side
   (reference: src/reinitialization/qss_reinitialization3d.f:39-470):
   HJ-ENO3 one-sided derivatives (minus sides of y and z scaled by 1/dx, as
   there), Godunov selection by the sign of phi (or phi0), then
   S (1 - |grad phi|) with S = phi / sqrt(phi^2 + |grad phi|^2 dx_max^2)
   (with phi0: phi0 / sqrt(phi0^2 + dx_max^2)); 0 where |phi| < 1e-5.
   `sixth` is the value the reference's undefined coefficient takes. */
void orc_computereinitializationeqnrhs3d(R *rhs, const R *phi, const R *phi0, GBARGS, FBARGS,
                                         const R *dx, const R *dy, const R *dz,
                                         const int *use_phi0_for_sgn, const R *sixth)
{
  box_t bx = GB;
  R inv_dx = (R)1.0 / *dx, inv_dy = (R)1.0 / *dy, inv_dz = (R)1.0 / *dz;
  R dx_sq = rmax3(*dx, *dy, *dz);
  dx_sq = dx_sq * dx_sq;
  int i, j, k, o, d;
  for (k = *klo_fb; k <= *khi_fb; k++)
    for (j = *jlo_fb; j <= *jhi_fb; j++)
      for (i = *ilo_fb; i <= *ihi_fb; i++) {
        R lx[7], ly[7], lz[7], gp[3], gm[3], star[3];
        for (o = 0; o < 7; o++) {
          lx[o] = AT(phi, bx, i + o - 3, j, k);
          ly[o] = AT(phi, bx, i, j + o - 3, k);
          lz[o] = AT(phi, bx, i, j, k + o - 3);
        }
        eno3_1d_sixth(lx, inv_dx, inv_dx, *sixth, &gp[0], &gm[0]);
        eno3_1d_sixth(ly, inv_dy, inv_dx, *sixth, &gp[1], &gm[1]);
        eno3_1d_sixth(lz, inv_dz, inv_dx, *sixth, &gp[2], &gm[2]);
        const int use0 = *use_phi0_for_sgn == 1;
        R c = use0 ? AT(phi0, bx, i, j, k) : AT(phi, bx, i, j, k);
        for (d = 0; d < 3; d++) {
          if (c > 0) { gp[d] = rmax(-gp[d], 0); gm[d] = rmax(gm[d], 0); }
          else       { gp[d] = rmax(gp[d], 0);  gm[d] = rmax(-gm[d], 0); }
          star[d] = rmax(gp[d], gm[d]);
        }
        if (fabs(c) >= (R)1.e-5) {
          R n2 = star[0] * star[0] + star[1] * star[1] + star[2] * star[2];
          R sgn = use0 ? c / wsqrt(c * c + dx_sq) : c / wsqrt(c * c + n2 * dx_sq);
          AT(rhs, bx, i, j, k) = sgn * ((R)1.0 - wsqrt(n2));
        } else {
          AT(rhs, bx, i, j, k) = 0;
        }
      }
}
