/*
 * qss_oracle_driver.c -- CPU restatement of the LSMPQS 3D evolution drivers.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle): tests/, the smoke check and the
 * CPU-baseline leg of bench.py call it as a checker; the product never does.
 *
 * The reference runs the inner loop of constCurvModel3d as an OpenMP team
 * that splits the fill box into nthreads z-slabs.  Every slab's results are
 * separated by barriers, so the only observable effects of the split are
 * (1) the time step, which uses H/dx at the last voxel of each slab, and
 * (2) the reuse of the previous voxel's normal term where |vel_n| < 1e-5,
 * which restarts at each slab.  This restatement runs the slabs one after the
 * other and reproduces both.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include "qss_oracle.h"

typedef QSSLIB_REAL R;

static orc_stats g_stats;
void orc_get_stats(orc_stats *s) { *s = g_stats; }

#define GB6(g) &(g)->ilo_gb, &(g)->ihi_gb, &(g)->jlo_gb, &(g)->jhi_gb, &(g)->klo_gb, &(g)->khi_gb
#define FB6(g) &(g)->ilo_fb, &(g)->ihi_fb, &(g)->jlo_fb, &(g)->jhi_fb, &(g)->klo_fb, &(g)->khi_fb

static long npts(const Grid *g) { return (long)g->num_gridpts; }
static R *ralloc(const Grid *g) { return (R *)calloc((size_t)npts(g), sizeof(R)); }

/* reference: src/utilities/qss_general_util.c:422-706 (index 9 = six faces in order) */
void orc_signedLinearExtrapolationBCqss(R *phi, Grid *g, int idx)
{
  int f;
  if (idx == 9) {
    for (f = 0; f < 6; f++) orc_signedlinearextrapolation3d(phi, GB6(g), FB6(g), &f);
  } else if (idx >= 0 && idx < 6) {
    orc_signedlinearextrapolation3d(phi, GB6(g), FB6(g), &idx);
  }
}

/* reference: src/reinitialization/reinitialize_top.c:297-377 */
void orc_reinitializeSubcellFix3d(R *data, Grid *g, Options *o)
{
  long n = npts(g), l;
  R *dist0 = ralloc(g), *copy = ralloc(g), *rhs = ralloc(g), *next = ralloc(g);
  R t_r = 0, dt_r = (R)0.5 * g->dx[0];
  memcpy(copy, data, (size_t)n * sizeof(R));
  orc_qss3dcomputedistanceforsubcellfix(dist0, copy, GB6(g), FB6(g), &g->dx[0], &g->dx[1], &g->dx[2]);
  while (t_r < o->tmax_r) {
    orc_qss3dcomputereinitializationeqnrhssubcellfixorder1(rhs, data, GB6(g), copy, dist0, FB6(g),
                                                           &g->dx[0], &g->dx[1], &g->dx[2]);
    orc_qss3drk1step(next, GB6(g), data, GB6(g), rhs, GB6(g), FB6(g), &dt_r);
    orc_signedLinearExtrapolationBCqss(next, g, 9);
    for (l = 0; l < n; l++) data[l] = next[l];
    t_r = t_r + dt_r;
  }
  free(dist0); free(copy); free(rhs); free(next);
}

/* reference: src/reinitialization/reinitialize_top.c:205-295 -- TVD "RK2"
   (stage 2 = 3/4 u + 1/4 (u1 + dt L)) of the smoothed-sign equation with
   dt_r = 0.5 dx until tmax_r, both stages followed by the all-face BC */
void orc_reinitializeMedium3d(R *data, Grid *g, Options *o, R sixth)
{
  long n = npts(g), l;
  R *data0 = ralloc(g), *rhs = ralloc(g), *next = ralloc(g), *stage1 = ralloc(g);
  R t_r = 0, dt_r = (R)0.5 * g->dx[0], tmax_r = o->tmax_r;
  int use0 = 0;
  memcpy(data0, data, (size_t)n * sizeof(R));
  while (t_r < tmax_r) {
    orc_computereinitializationeqnrhs3d(rhs, data, data0, GB6(g), FB6(g), &g->dx[0], &g->dx[1],
                                        &g->dx[2], &use0, &sixth);
    orc_qss3drk1step(stage1, GB6(g), data, GB6(g), rhs, GB6(g), FB6(g), &dt_r);
    orc_signedLinearExtrapolationBCqss(stage1, g, 9);
    orc_computereinitializationeqnrhs3d(rhs, stage1, data0, GB6(g), FB6(g), &g->dx[0], &g->dx[1],
                                        &g->dx[2], &use0, &sixth);
    orc_qss3dtvdrk2stage2(next, GB6(g), stage1, GB6(g), data, GB6(g), rhs, GB6(g), FB6(g), &dt_r);
    orc_signedLinearExtrapolationBCqss(next, g, 9);
    for (l = 0; l < n; l++) data[l] = next[l];
    t_r = t_r + dt_r;
  }
  free(data0); free(rhs); free(next); free(stage1);
}

/* Union-find over the whole ghost box with the 26-neighbourhood; every
   component is labelled with its largest linear index, which is the root the
   reference's union rule produces (reference:
   src/connectivity/connectivity.c:48-83,148-217).  Background gets 0. */
static int uf_find(int *parent, int i)
{
  int r = i, nx;
  while (parent[r] != r) r = parent[r];
  while (parent[i] != r) { nx = parent[i]; parent[i] = r; i = nx; }
  return r;
}

/* 26 (the reference's rule, connectivity.c:159-196) or 6 (face neighbours
   only); set by orc_set_connectivity */
static int orc_conn = 26;

int orc_set_connectivity(int c)
{
  int old = orc_conn;
  orc_conn = (c == 6) ? 6 : 26;
  return old;
}

void orc_findConnectivity3d(QSS_DataArrays *p, Grid *g)
{
  int nx = g->grid_dims_ghostbox[0], ny = g->grid_dims_ghostbox[1], nz = g->grid_dims_ghostbox[2];
  long n = npts(g);
  int *parent = (int *)malloc((size_t)n * sizeof(int));
  int i, j, k, di, dj, dk;
  long l;
  for (l = 0; l < n; l++) parent[l] = (int)l;
  for (k = 0; k < nz; k++)
    for (j = 0; j < ny; j++)
      for (i = 0; i < nx; i++) {
        int a = i + j * nx + k * nx * ny;
        if (!p->phi_bin[a]) continue;
        for (dk = -1; dk <= 1; dk++)
          for (dj = -1; dj <= 1; dj++)
            for (di = -1; di <= 1; di++) {
              int i2 = i + di, j2 = j + dj, k2 = k + dk, b, ra, rb;
              if (!(di || dj || dk)) continue;
              if (orc_conn == 6 && (di != 0) + (dj != 0) + (dk != 0) != 1) continue;
              if (i2 < 0 || j2 < 0 || k2 < 0 || i2 >= nx || j2 >= ny || k2 >= nz) continue;
              b = i2 + j2 * nx + k2 * nx * ny;
              if (!p->phi_bin[b]) continue;
              ra = uf_find(parent, a); rb = uf_find(parent, b);
              if (ra < rb) parent[ra] = rb; else if (rb < ra) parent[rb] = ra;
            }
      }
  for (l = 0; l < n; l++) p->connectivity[l] = p->phi_bin[l] ? uf_find(parent, (int)l) : 0;
  free(parent);
}

/* reference: src/connectivity/connectivity.c:219-296 (3D branches) */
void orc_getMainInd_new(Options *o, QSS_DataArrays *p, Grid *g)
{
  int nx = g->grid_dims_ghostbox[0], nxy = nx * g->grid_dims_ghostbox[1];
  int xi, yi, zi;
  (void)p;
  if (o->center_inlet) {
    xi = (int)(0.5 * (g->ilo_fb + g->ihi_fb));
  } else {
    xi = g->ilo_fb - 1;
  }
  yi = (int)(0.5 * (g->jlo_fb + g->jhi_fb));
  zi = (int)(0.5 * (g->klo_fb + g->khi_fb));
  o->phi_nw_ind = xi + yi * nx + zi * nxy;
  xi = g->ihi_fb + 1;
  o->phi_w_ind = xi + yi * nx + zi * nxy;
}

/* reference: src/connectivity/trapComponents.c:156-375 (3D branches) */
void orc_trapComponents_mask(QSS_DataArrays *p, Grid *g, Options *o)
{
  long n = npts(g), l;
  int nx = g->grid_dims_ghostbox[0], nxy = nx * g->grid_dims_ghostbox[1];
  int i, j, k, main_val;
  unsigned char *disc = (unsigned char *)malloc((size_t)n);
  R fill = -20 * g->dx[0];

  for (l = 0; l < n; l++) if (p->mask_nw[l] > 0) p->phi[l] = -p->mask_nw[l];
  for (l = 0; l < n; l++) p->phi_bin[l] = p->phi[l] < 0 ? 1 : 0;
  if (o->center_inlet) {
    k = (int)(0.5 * (g->klo_fb + g->khi_fb));
    i = (int)(0.5 * (g->ilo_fb + g->ihi_fb));
    for (j = g->jlo_gb; j <= g->jhi_gb; j++) p->phi_bin[i + j * nx + k * nxy] = 1;
  } else {
    i = g->ilo_fb - 1;
    for (k = g->klo_fb; k <= g->khi_fb; k++)
      for (j = g->jlo_fb; j <= g->jhi_fb; j++) p->phi_bin[i + j * nx + k * nxy] = 1;
  }
  orc_findConnectivity3d(p, g);
  main_val = p->connectivity[o->phi_nw_ind];
  for (l = 0; l < n; l++)
    disc[l] = (p->connectivity[l] == main_val || p->connectivity[l] == 0) ? 0 : 1;
  for (l = 0; l < n; l++) p->mask_nw[l] = fill;
  for (l = 0; l < n; l++) if (disc[l] == 1 && p->phi[l] < 0) p->mask_nw[l] = -p->phi[l];
  for (l = 0; l < n; l++) p->connectivity[l] = 0;

  for (l = 0; l < n; l++) {
    R s = -p->phi[l];
    s = (p->mask[l] > s) ? p->mask[l] : s;
    s = (p->mask_nw[l] > s) ? p->mask_nw[l] : s;
    p->scratch1[l] = s;
    p->phi_bin[l] = s < 0 ? 1 : 0;
  }
  if (o->center_inlet) {
    for (j = g->jlo_gb; j <= g->jhi_gb; j++)
      for (i = g->ilo_gb; i <= g->ihi_gb; i++) {
        p->phi_bin[i + j * nx + (g->klo_fb - 1) * nxy] = 1;
        p->phi_bin[i + j * nx + (g->khi_fb + 1) * nxy] = 1;
      }
    for (j = g->jlo_gb; j <= g->jhi_gb; j++)
      for (k = g->klo_gb; k <= g->khi_gb; k++) {
        p->phi_bin[(g->ilo_fb - 1) + j * nx + k * nxy] = 1;
        p->phi_bin[(g->ihi_fb + 1) + j * nx + k * nxy] = 1;
      }
  } else {
    i = g->ihi_fb + 1;
    for (k = g->klo_fb; k <= g->khi_fb; k++)
      for (j = g->jlo_fb; j <= g->jhi_fb; j++) p->phi_bin[i + j * nx + k * nxy] = 1;
  }
  orc_findConnectivity3d(p, g);
  main_val = p->connectivity[o->phi_w_ind];
  for (l = 0; l < n; l++)
    disc[l] = (p->connectivity[l] == main_val || p->connectivity[l] == 0) ? 0 : 1;
  for (l = 0; l < n; l++) p->mask_w[l] = fill;
  for (l = 0; l < n; l++) if (disc[l] == 1 && p->scratch1[l] < 0) p->mask_w[l] = -p->scratch1[l];
  for (l = 0; l < n; l++) p->scratch1[l] = 0;
  free(disc);
}

/* Constant-theta overlap (reference: src/utilities/qss_general_util.c:150-174) */
void orc_setThetaOverlap(Options *o, Grid *g)
{
  R theta = o->theta;
  R rad = (R)(3.14159 / 180);
  if (theta <= 20 * rad) o->overlap = 0;
  else if (theta <= 40 * rad) o->overlap = (R)(0.3 * g->dx[0]);
  else if (theta <= 50 * rad) o->overlap = (R)(0.5 * g->dx[0]);
  else o->overlap = (R)(1 * g->dx[0]);
}

/* RHS kernel choice for 3D (reference: src/utilities/qss_general_util.c:115-141).
   Returns 2, 3 or 5 for the scheme, negated for the constant-coefficient
   kernels. */
int orc_rhs_scheme(Options *o)
{
  int cst = (o->theta == 0) && (~o->use_var_theta != 0);
  switch (o->order_space_accur) {
  case 2: return cst ? -2 : 2;
  case 3: case 4: return cst ? -3 : 3;
  case 5: return cst ? -3 : 5;
  default: return 3;
  }
}

typedef void (*rhs_var_fn)(const R *, const R *, const R *, const R *, const R *, const R *,
                           R *, R *, ORC_B6, ORC_B6, const R *, const R *, const R *);
typedef void (*rhs_cst_fn)(const R *, const R *, const R *, R *, R *, ORC_B6, ORC_B6,
                           const R *, const R *, const R *);

/* one RHS sweep over all slabs; returns the time-step H/dx (max over slab ends) */
static R rhs_all_slabs(Options *o, QSS_DataArrays *p, Grid *g, const R *phi, int scheme,
                       R a_const, int nthreads)
{
  int t, nslices = g->khi_fb - g->klo_fb + 1;
  R cur = -1;
  for (t = 0; t < nthreads; t++) {
    int klo = g->klo_fb + nslices * t / nthreads;
    int khi = g->klo_fb + nslices * (t + 1) / nthreads - 1;
    R mh = -1;
    if (khi > g->khi_fb) khi = g->khi_fb;
    if (klo > khi) continue;
    if (scheme < 0) {
      rhs_cst_fn f = scheme == -2 ? orc_getrhs3d2ndconst : orc_getrhs3d3rdconst;
      f(phi, &a_const, &o->b, &mh, p->lse_rhs, GB6(g), &g->ilo_fb, &g->ihi_fb, &g->jlo_fb,
        &g->jhi_fb, &klo, &khi, &g->dx[0], &g->dx[1], &g->dx[2]);
    } else {
      rhs_var_fn f = scheme == 2 ? orc_getrhs3d2nd : (scheme == 3 ? orc_getrhs3d3rd : orc_getrhs3d5th);
      f(phi, p->normal_velocity, p->curvature_coeff, p->external_velocity_x,
        p->external_velocity_y, p->external_velocity_z, &mh, p->lse_rhs, GB6(g), &g->ilo_fb,
        &g->ihi_fb, &g->jlo_fb, &g->jhi_fb, &klo, &khi, &g->dx[0], &g->dx[1], &g->dx[2]);
    }
    if (mh > cur) cur = mh;
  }
  return cur;
}

static void impose_max(R *phi, const R *m, const R *src, R overlap, long n)
{
  long l;
  for (l = 0; l < n; l++) {
    R v = m[l] - overlap;
    phi[l] = (v > src[l]) ? v : src[l];
  }
}

static int count_stale(QSS_DataArrays *p, Grid *g, int scheme, R a_const)
{
  int nx = g->grid_dims_ghostbox[0], nxy = nx * g->grid_dims_ghostbox[1], i, j, k, c = 0;
  if (scheme < 0) return fabs(a_const) < (R)1.e-5 ? -1 : 0;
  for (k = g->klo_fb; k <= g->khi_fb; k++)
    for (j = g->jlo_fb; j <= g->jhi_fb; j++)
      for (i = g->ilo_fb; i <= g->ihi_fb; i++)
        if (fabs(p->normal_velocity[i + j * nx + k * nxy]) < (R)1.e-5) c++;
  return c;
}

/* Outer/inner loop of the evolution (reference:
   src/level_set_evolution/constCurvModel3d.c:50-334 for the variable-coefficient
   path, :336-610 for the constant one).  scheme < 0 selects the latter. */
static R ccm_core(Options *o, QSS_DataArrays *p, Grid *g, FILE *fp, R a_const, int nthreads,
                  int var_path)
{
  long n = npts(g), l;
  int scheme = orc_rhs_scheme(o);
  R dt = 0, dt_sub, t = 0, cfl = (R)0.5;
  R max_abs_err = 100, vol_phi = 100, vol_pore, satn, satn_old;
  R vsmall = g->dx[0] * g->dx[1] * g->dx[2];
  R eps = o->eps_coefficient * g->dx[0];
  R max_satn = o->vol_frac_max, zero = 0;
  R *stage = var_path ? p->scratch2 : p->scratch1;

  memset(&g_stats, 0, sizeof(g_stats));
  g_stats.stale_norm_voxels = count_stale(p, g, scheme, a_const);
  orc_qss3dvolumeregionphilessthanzero(&vol_pore, p->mask, GB6(g), FB6(g), &g->dx[0], &g->dx[1],
                                       &g->dx[2], &eps);
  orc_qss3dvolumeregionphilessthanzero(&vol_phi, p->phi, GB6(g), FB6(g), &g->dx[0], &g->dx[1],
                                       &g->dx[2], &eps);
  satn = vol_phi / vol_pore;

  while ((t < o->tmax) && (max_abs_err > o->eps_stop)) {
    g_stats.outer_steps++;
    dt_sub = 0;
    memcpy(p->phi_prev, p->phi, (size_t)n * sizeof(R));
    while (dt_sub < o->tplot) {
      R H = rhs_all_slabs(o, p, g, p->phi, scheme, a_const, nthreads);
      dt = cfl / (H + o->b_max_over_dx + o->max_U_over_dx);
      orc_qss3drk1step(p->phi_next, GB6(g), p->phi, GB6(g), p->lse_rhs, GB6(g), FB6(g), &dt);
      orc_signedLinearExtrapolationBCqss(p->phi_next, g, 9);
      if (o->order_time_accur >= 2) {
        rhs_all_slabs(o, p, g, p->phi_next, scheme, a_const, nthreads);
        orc_qss3dtvdrk2stage2(stage, GB6(g), p->phi_next, GB6(g), p->phi, GB6(g), p->lse_rhs,
                              GB6(g), FB6(g), &dt);
        memcpy(p->phi_next, stage, (size_t)n * sizeof(R));
        orc_signedLinearExtrapolationBCqss(p->phi_next, g, 9);
      }
      if (o->order_time_accur >= 3) {
        rhs_all_slabs(o, p, g, p->phi_next, scheme, a_const, nthreads);
        orc_qss3dtvdrk3stage3(p->scratch2, GB6(g), p->phi_next, GB6(g), p->phi, GB6(g),
                              p->lse_rhs, GB6(g), FB6(g), &dt);
        memcpy(p->phi_next, p->scratch2, (size_t)n * sizeof(R));
        orc_signedLinearExtrapolationBCqss(p->phi_next, g, 9);
      }
      for (l = 0; l < n; l++) { p->lse_rhs[l] = 0; p->scratch1[l] = 0; p->scratch2[l] = 0; }
      dt_sub += dt;
      g_stats.inner_steps++;
      impose_max(p->phi, p->mask, p->phi_next, o->overlap, n);
      if (o->check_connectivity) {
        impose_max(p->phi, p->mask_w, p->phi, zero, n);
        impose_max(p->phi, p->mask_nw, p->phi, zero, n);
      }
    }
    if (fp) fprintf(fp, "%d threads: Level set time = %lf\n", nthreads, 0.0);
    t += dt_sub;
    if (var_path) impose_max(p->phi, p->mask, p->phi, zero, n);
    if (o->check_connectivity) {
      orc_trapComponents_mask(p, g, o);
      orc_reinitializeSubcellFix3d(p->mask_w, g, o);
      orc_reinitializeSubcellFix3d(p->mask_nw, g, o);
      if (var_path && o->conserve_imbibe)
        orc_imposetrapvelocity3d(p->mask_nw, p->mask_w, p->normal_velocity, p->curvature_coeff,
                                 p->external_velocity_x, p->external_velocity_y,
                                 p->external_velocity_z, GB6(g), FB6(g), &g->dx[0]);
      for (l = 0; l < n; l++) if (p->mask_nw[l] > 0) p->phi[l] = -p->mask_nw[l];
    }
    orc_reinitializeSubcellFix3d(p->phi, g, o);
    orc_qss3dmaxnormdifflocal(&max_abs_err, p->phi, GB6(g), p->phi_prev, GB6(g), FB6(g),
                              &o->err_check_zone);
    satn_old = satn;
    orc_qss3dvolumeregionphilessthanzero(&vol_phi, p->phi, GB6(g), FB6(g), &g->dx[0], &g->dx[1],
                                         &g->dx[2], &eps);
    satn = vol_phi / vol_pore;
    g_stats.last_dt = dt;
    g_stats.satn = satn;
    g_stats.max_abs_err = max_abs_err;
    if ((satn > max_satn) || (vol_phi <= vsmall)) return -1;
    if (fp) {
      fprintf(fp, "t = %4.1f\t", t);
      fprintf(fp, "max_abs_err = %4.3f,\t", max_abs_err);
      fprintf(fp, "nw phase satn pct = %4.3f\n", satn * 100);
    }
    if (o->use_satn_stop) {
      if (fabsf((float)(100 * (satn - satn_old))) < 0.001) break;
    }
  }
  return t;
}

R orc_constCurvModel3d(Options *o, QSS_DataArrays *p, Grid *g, FILE *fp, int nthreads)
{
  return ccm_core(o, p, g, fp, 0, nthreads, 1);
}

R orc_constCurvModel3dNoVar(Options *o, QSS_DataArrays *p, Grid *g, FILE *fp, R a, int nthreads)
{
  return ccm_core(o, p, g, fp, a, nthreads, 0);
}

/* Analytic initialisers (reference: src/geometry/qss_initialization3d.c:34-166,440-522) */
void orc_createSphere(R *phi, R cx, R cy, R cz, R r, int inside_flag, Grid *g)
{
  int nx = g->grid_dims_ghostbox[0], ny = g->grid_dims_ghostbox[1], nz = g->grid_dims_ghostbox[2];
  int i, j, k;
  for (k = 0; k < nz; k++)
    for (j = 0; j < ny; j++)
      for (i = 0; i < nx; i++) {
        R x = g->x_lo_ghostbox[0] + g->dx[0] * i;
        R y = g->x_lo_ghostbox[1] + g->dx[1] * j;
        R z = g->x_lo_ghostbox[2] + g->dx[2] * k;
        R d = (R)(sqrt((x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz)) - r);
        R mx = -FLT_MAX;
        if (inside_flag >= 0) d = -d;
        if (d > mx) mx = d;
        phi[i + j * nx + (long)k * nx * ny] = mx;
      }
}

void orc_createBox(R *phi, R x0, R y0, R z0, R lx, R ly, R lz, int inside_flag, Grid *g)
{
  int nx = g->grid_dims_ghostbox[0], ny = g->grid_dims_ghostbox[1], nz = g->grid_dims_ghostbox[2];
  R px[6], py[6], pz[6], nxv[6], nyv[6], nzv[6];
  int i, j, k, l;
  for (l = 0; l < 3; l++) {
    px[l] = x0; py[l] = y0; pz[l] = z0;
    nxv[l] = l == 0 ? -1 : 0; nyv[l] = l == 1 ? -1 : 0; nzv[l] = l == 2 ? -1 : 0;
  }
  for (l = 3; l < 6; l++) {
    px[l] = x0 + lx; py[l] = y0 + ly; pz[l] = z0 + lz;
    nxv[l] = l == 3 ? 1 : 0; nyv[l] = l == 4 ? 1 : 0; nzv[l] = l == 5 ? 1 : 0;
  }
  for (k = 0; k < nz; k++)
    for (j = 0; j < ny; j++)
      for (i = 0; i < nx; i++) {
        R x = g->x_lo_ghostbox[0] + g->dx[0] * i;
        R y = g->x_lo_ghostbox[1] + g->dx[1] * j;
        R z = g->x_lo_ghostbox[2] + g->dx[2] * k;
        R mx = -FLT_MAX;
        for (l = 0; l < 6; l++) {
          R dot = (x - px[l]) * nxv[l] + (y - py[l]) * nyv[l] + (z - pz[l]) * nzv[l];
          R nrm = nxv[l] * nxv[l] + nyv[l] * nyv[l] + nzv[l] * nzv[l];
          R d;
          nrm = (R)sqrt(nrm);
          d = dot / nrm;
          if (d > mx) mx = d;
        }
        if (inside_flag >= 1) mx = -mx;
        phi[i + j * nx + (long)k * nx * ny] = mx;
      }
}

/* Setup shared by drain_model3d and imbibe_model3d before the curvature loop
   (reference: src/level_set_evolution/drain_model3d.c:56-100,
   imbibe_model3d.c:56-97). */
void orc_model_setup3d(Options *o, QSS_DataArrays *p, Grid *g, int drain)
{
  long n = npts(g), l;
  R cx = (R)(0.5 * (g->x_lo[0] + g->x_hi[0]));
  R cy = (R)(0.5 * (g->x_lo[1] + g->x_hi[1]));
  R cz = (R)(0.5 * (g->x_lo[2] + g->x_hi[2]));
  R rad = 2 * (g->x_hi[0] - g->x_lo[0]);
  orc_createSphere(p->mask_w, cx, cy, cz, rad, -1, g);
  orc_createSphere(p->mask_nw, cx, cy, cz, rad, -1, g);
  if (p->mask_disconn_init) memcpy(p->mask_disconn_init, p->mask_nw, (size_t)n * sizeof(R));
  orc_createBox(p->phi_extra, g->x_lo_ghostbox[0], g->x_lo_ghostbox[1], g->x_lo_ghostbox[2],
                3 * g->dx[0], g->x_hi_ghostbox[1] - g->x_lo_ghostbox[1],
                g->x_hi_ghostbox[2] - g->x_lo_ghostbox[2], -1, g);
  if (drain && o->mask_reinit) orc_reinitializeSubcellFix3d(p->mask, g, o);
  orc_getcentralgrad3dorder2(p->mask_x, p->mask_y, p->mask_z, GB6(g), p->mask, GB6(g), FB6(g),
                             &g->dx[0], &g->dx[1], &g->dx[2]);
  orc_setThetaOverlap(o, g);
  if (o->reservoir_inlet) {
    for (l = 0; l < n; l++) p->phi[l] = (p->phi[l] < p->phi_extra[l]) ? p->phi[l] : p->phi_extra[l];
    impose_max(p->phi, p->mask, p->phi, 0, n);
  }
  orc_getMainInd_new(o, p, g);
}

static void model_loop(Options *o, QSS_DataArrays *p, Grid *g, FILE *fp, int nthreads, int drain)
{
  long n = npts(g), l;
  R a0 = o->a, da = o->b * o->dc;
  int var_path = (o->theta > 0) || o->use_var_theta;
  int step = o->init_step + 1;
  orc_model_setup3d(o, p, g, drain);
  while (drain ? (a0 < o->amax) : (a0 > o->amin)) {
    R t_step;
    if (fp) fprintf(fp, "\nStep %d: a = %f\n", step, (double)a0);
    if (var_path) {
      orc_setvarcurvadv3d(p->mask, p->mask_x, p->mask_y, p->mask_z, p->curvature_coeff,
                          p->external_velocity_x, p->external_velocity_y, p->external_velocity_z,
                          &o->b_max_over_dx, &o->max_U_over_dx, &o->b, GB6(g), FB6(g), &g->dx[0]);
    } else {
      R ix = 1 / g->dx[0], iy = 1 / g->dx[1], iz = 1 / g->dx[2];
      o->b_max_over_dx = 2 * o->b * (ix * ix + iy * iy + iz * iz);
      o->max_U_over_dx = 0;
    }
    if (!drain) o->conserve_imbibe = 1;
    if (var_path)
      orc_setvarnorm3d(p->mask, p->mask_x, p->mask_y, p->mask_z, p->normal_velocity, &a0,
                       &o->theta, GB6(g), FB6(g), &g->dx[0]);
    if (o->check_connectivity) orc_trapComponents_mask(p, g, o);
    t_step = var_path ? orc_constCurvModel3d(o, p, g, fp, nthreads)
                      : orc_constCurvModel3dNoVar(o, p, g, fp, a0, nthreads);
    if (t_step == -1) break;
    if (drain) a0 += da; else a0 -= da;
    if (o->reservoir_inlet)
      for (l = 0; l < n; l++) p->phi[l] = (p->phi[l] < p->phi_extra[l]) ? p->phi[l] : p->phi_extra[l];
    impose_max(p->phi, p->mask, p->phi, 0, n);
    step++;
  }
}

/* reference: src/level_set_evolution/drain_model3d.c:49-165 (file output omitted) */
void orc_drain_model3d(Options *o, QSS_DataArrays *p, Grid *g, FILE *fp, int nthreads)
{
  model_loop(o, p, g, fp, nthreads, 1);
}

/* reference: src/level_set_evolution/imbibe_model3d.c:49-161 (file output omitted) */
void orc_imbibe_model3d(Options *o, QSS_DataArrays *p, Grid *g, FILE *fp, int nthreads)
{
  model_loop(o, p, g, fp, nthreads, 0);
}

/* ------------------------------------------------------------------------
 * Pre- and post-processing applications (their computations, without the
 * file handling of their main programs).
 */

/* createMaskUBC's mask: +dx on image voxels equal to 255 (grain), -dx on
 * the others (pore), 0 on ghosts, then the subcell-fix reinitialization;
 * g is the grid createMaskUBC builds (createGridSetGridDims with x_lo = 3 dx).
 * reference: src/applications/createMaskUBC.c:87-116,
 * src/utilities/qss_general_util.c:261-306 */
void orc_mask_from_image(const unsigned char *image, int nx_vox, int ny_vox, int nz_vox,
                         Grid *g, Options *o, R *mask)
{
  int nx_gb = g->grid_dims_ghostbox[0], nxy_gb = nx_gb * g->grid_dims_ghostbox[1];
  int i, j, k;
  long l;
  (void)nz_vox;
  for (l = 0; l < npts(g); l++) mask[l] = 0;
  for (k = g->klo_fb; k <= g->khi_fb; k++)
    for (j = g->jlo_fb; j <= g->jhi_fb; j++)
      for (i = g->ilo_fb; i <= g->ihi_fb; i++) {
        long idx = (long)(i - 3) + (long)(j - 3) * nx_vox + (long)(k - 3) * nx_vox * ny_vox;
        long gb = i + (long)j * nx_gb + (long)k * nxy_gb;
        mask[gb] = image[idx] == 255 ? o->dx : -o->dx;
      }
  orc_reinitializeSubcellFix3d(mask, g, o);
}

/* createInitFromMask's initial phi: the non-wetting sphere (radius 0.7 L
 * centred 0.49 L before the x-low face, or radius 90 dx at the centre with
 * center_inlet, whose x centre the reference takes from the y bounds),
 * then max(mask, phi).  reference: src/applications/createInitFromMask.c:75-100 */
void orc_init_from_mask(const R *mask, Grid *g, Options *o, R *data)
{
  R radius, cx, cy, cz;
  long l;
  if (o->center_inlet) {
    radius = 90 * g->dx[0];
    cy = 0.5 * (g->x_lo[1] + g->x_hi[1]);
    cx = 0.5 * (g->x_lo[1] + g->x_hi[1]);
    cz = 0.5 * (g->x_lo[2] + g->x_hi[2]);
  } else {
    radius = 0.7 * (g->x_hi[0] - g->x_lo[0]);
    cy = 0.5 * (g->x_lo[1] + g->x_hi[1]);
    cx = (g->x_lo[0]) - radius * 0.7;
    cz = 0.5 * (g->x_lo[2] + g->x_hi[2]);
  }
  orc_createSphere(data, cx, cy, cz, radius, -1, g);
  for (l = 0; l < npts(g); l++) data[l] = (mask[l] > data[l]) ? mask[l] : data[l];
}

/* makeDiffSteps' percolation lens of two saturation steps a < b:
 * diff = max(-a, b - dx/2); the voxels with diff < 0, plus the inlet and
 * outlet ghost slabs, are labelled (26-neighbourhood); if the inlet and
 * outlet voxels at mid j, k share a label, connected_only = diff with the
 * sign flipped where diff <= 0 outside that component, and 1 is returned,
 * else 0.  reference: src/applications/makeDiffSteps.c:70-140 */
int orc_diff_steps(const R *a, const R *b, Grid *g, R *diff, R *connected_only)
{
  long n = npts(g), l;
  int nx_gb = g->grid_dims_ghostbox[0], nxy_gb = nx_gb * g->grid_dims_ghostbox[1];
  int i, j, k, inlet, outlet, main_val, perc;
  QSS_DataArrays q;
  memset(&q, 0, sizeof q);
  q.phi_bin = (unsigned char *)malloc((size_t)n);
  q.connectivity = (int *)malloc((size_t)n * sizeof(int));
  for (l = 0; l < n; l++) {
    R m = -a[l];
    R p = b[l];
    p -= 0.5 * g->dx[0];
    diff[l] = (m > p) ? m : p;
    q.phi_bin[l] = diff[l] < 0 ? 1 : 0;
  }
  for (k = g->klo_fb; k <= g->khi_fb; k++)
    for (j = g->jlo_fb; j <= g->jhi_fb; j++) {
      q.phi_bin[(g->ilo_fb - 1) + (long)j * nx_gb + (long)k * nxy_gb] = 1;
      q.phi_bin[(g->ihi_fb + 1) + (long)j * nx_gb + (long)k * nxy_gb] = 1;
    }
  orc_findConnectivity3d(&q, g);
  j = 0.5 * (g->jlo_fb + g->jhi_fb);
  k = 0.5 * (g->klo_fb + g->khi_fb);
  inlet = (g->ilo_fb - 1) + j * nx_gb + k * nxy_gb;
  outlet = (g->ihi_fb + 1) + j * nx_gb + k * nxy_gb;
  perc = q.connectivity[inlet] == q.connectivity[outlet];
  if (perc) {
    main_val = q.connectivity[inlet];
    for (l = 0; l < n; l++) {
      R v = diff[l];
      if ((q.connectivity[l] != main_val) & (v <= 0)) v *= -1;
      connected_only[l] = v;
    }
  }
  free(q.phi_bin);
  free(q.connectivity);
  (void)i;
  return perc;
}
