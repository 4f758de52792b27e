// lsm_model2d.hip -- the two-dimensional drainage and imbibition models on
// the GPU (reference: src/level_set_evolution/drain_model2d.c,
// imbibe_model2d.c, constCurvModel2d.c; src/spatial_derivatives/
// qss_spatial_derivatives2d.f; src/reinitialization/qss_reinitialization2d.f,
// reinitialize_top.c:123-203; src/utilities/qss_util2d.f).
//
// A 2D grid holds a few million voxels at most, so every kernel here is a
// plain one-voxel-per-lane kernel over the fill box.  The pieces whose 2D
// form equals the 3D one on a box of one plane (boundary extrapolation,
// elementwise imposition and merge, the FP64 volume and max-norm reductions,
// the labelling, whose 26-neighbourhood within one plane is the reference's
// 8-neighbourhood with the same largest-index labels) run the 3D kernels on
// that box.  The arithmetic of each kernel follows the reference's 2D
// expression order; the reference's 2D quirks are kept (listed at each
// kernel).
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <stdexcept>
#include <string>
#include <vector>
#include "lsm_kernels.h"
#include "lsm_session.h"
#include "lsmpqs_hip.h"

namespace lsm {
namespace {

#define X2CHECK(x)                                                               \
  do {                                                                           \
    hipError_t e_ = (x);                                                         \
    if (e_ != hipSuccess)                                                        \
      throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

// fill box of a 2D grid and the one-plane 3D box the shared kernels take
struct Box2 {
  int nx, ny, ilo, ihi, jlo, jhi;
  __host__ __device__ int at(int i, int j) const { return i + nx * j; }
  __host__ __device__ int nfill() const { return (ihi - ilo + 1) * (jhi - jlo + 1); }
  Box box3() const {
    Box b;
    b.nx = nx; b.ny = ny; b.nz = 1;
    b.ilo = ilo; b.ihi = ihi; b.jlo = jlo; b.jhi = jhi; b.klo = 0; b.khi = 0;
    return b;
  }
};

Box2 box2_of(const Grid* g) {
  Box2 b;
  b.nx = g->grid_dims_ghostbox[0];
  b.ny = g->grid_dims_ghostbox[1];
  b.ilo = g->ilo_fb; b.ihi = g->ihi_fb; b.jlo = g->jlo_fb; b.jhi = g->jhi_fb;
  return b;
}

int blocks2(int64_t n) {
  int64_t b = (n + 255) / 256;
  return (int)(b > 16384 ? 16384 : (b < 1 ? 1 : b));
}

__device__ __forceinline__ void fill_ij(const Box2& b, int f, int& i, int& j) {
  const int w = b.ihi - b.ilo + 1;
  j = b.jlo + f / w;
  i = b.ilo + (f - (f / w) * w);
}

// ---------------------------------------------------------------------------
// Right-hand side + TVD-RK stage (qss_spatial_derivatives2d.f:248-1400 with
// qss_tvd_runge_kutta2d.f).  S = 2: HJ-ENO2 (qss2dGetRHS2nd[const]); S = 3:
// HJ-ENO3 (qss2dgetRHS3rd[Const], whose phi_y^- is scaled by 1/dx).  VAR:
// per-voxel vel_n, b and advection velocity; otherwise the scalars a and b
// and no advective term (the reference's const subroutines add an
// unassigned temp_vel there, qss_spatial_derivatives2d.f:74,215; it is taken
// as 0).  Where |vel_n| < 1e-5 the
// reference leaves temp_norm unassigned and so reuses the value of the last
// voxel of the same slab (j-slabs of the OpenMP split, i fastest) that did
// assign it, or an unset value when none did (taken as 0).

struct Rhs2Args {
  Box2 b;
  const real* in;   // stage input (ghosts valid)
  const real* un;   // u^n (stages 2 and 3)
  real* out;        // stage result (fill box)
  const real *vn, *bb, *vx, *vy;  // variable coefficients (VAR)
  real a_const, b_const;
  real dx, dy;
  const real* dt;
  int nslabs;
  const int* stale_src;  // VAR: the voxel whose normal term each voxel reuses, or -1 (stale_src2d)
};

template <int S>
__device__ __forceinline__ void derivs2d(const real* in, const Box2& b, int i, int j, real inv_dx,
                                         real inv_dy, real* d) {
  real lx[7], ly[7];
#pragma unroll
  for (int o = 0; o < 7; ++o) {
    lx[o] = (S == 3 || (o >= 1 && o <= 5)) ? in[b.at(i - 3 + o, j)] : real(0);
    ly[o] = (S == 3 || (o >= 1 && o <= 5)) ? in[b.at(i, j - 3 + o)] : real(0);
  }
  if (S == 2) {
    eno2(lx, inv_dx, d[0], d[1]);
    eno2(ly, inv_dy, d[2], d[3]);
  } else {
    eno3(lx, inv_dx, inv_dx, d[0], d[1]);
    eno3(ly, inv_dy, inv_dx, d[2], d[3]);
  }
}

// Godunov |grad phi| term, the reference's two-branch form (d = {x+, x-, y+, y-})
__device__ __forceinline__ real norm2d(real vn, const real* d) {
  real g2;
  if (vn > real(0)) {
    real x1 = tmax(d[1], real(0)), x2 = tmin(d[0], real(0));
    real y1 = tmax(d[3], real(0)), y2 = tmin(d[2], real(0));
    g2 = tmax(x1 * x1, x2 * x2) + tmax(y1 * y1, y2 * y2);
  } else {
    real x1 = tmin(d[1], real(0)), x2 = tmax(d[0], real(0));
    real y1 = tmin(d[3], real(0)), y2 = tmax(d[2], real(0));
    g2 = tmax(x1 * x1, x2 * x2) + tmax(y1 * y1, y2 * y2);
  }
  return -vn * sqrt(g2);
}

// slab of fill row j among n slabs: [lo, hi]
__device__ __forceinline__ void slab_of(const Box2& b, int nslabs, int j, int& lo, int& hi) {
  const int nsl = b.jhi - b.jlo + 1;
  int t = (int)(((int64_t)(j - b.jlo + 1) * nslabs - 1) / nsl);
  while (t > 0 && b.jlo + (int)((int64_t)nsl * t / nslabs) > j) --t;
  while (t + 1 < nslabs && b.jlo + (int)((int64_t)nsl * (t + 1) / nslabs) <= j) ++t;
  lo = b.jlo + (int)((int64_t)nsl * t / nslabs);
  hi = b.jlo + (int)((int64_t)nsl * (t + 1) / nslabs) - 1;
}

// Where the reference leaves temp_norm unassigned (|vel_n| < 1e-5) a voxel
// takes the normal term of the last voxel of its j-slab, in loop order (j,
// then i), that assigned it (qss_spatial_derivatives2d.f); stale_src2d lists
// that voxel for every fill voxel once per change of vel_n, in O(n).
template <int S, bool VAR>
__device__ real stale_norm2d(const Rhs2Args& a, int i, int j) {
  if (!VAR) return real(0);  // a scalar |a| < 1e-5 leaves every voxel unassigned
  const Box2& b = a.b;
  const int src = a.stale_src[b.at(i, j)];
  if (src < 0) return real(0);
  const int jj = src / b.nx, ii = src - jj * b.nx;
  const real inv_dx = real(1.0) / a.dx, inv_dy = real(1.0) / a.dy;
  real d[4];
  derivs2d<S>(a.in, b, ii, jj, inv_dx, inv_dy, d);
  return norm2d(a.vn[src], d);
}

// last voxel of each fill row with |vel_n| >= 1e-5, or -1 (one thread per row)
__global__ void stale_rows2d_kernel(Box2 b, const real* vn, int* row_last) {
  const int j = b.jlo + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (j > b.jhi) return;
  int last = -1;
  for (int i = b.ilo; i <= b.ihi; ++i)
    if (fabs(vn[b.at(i, j)]) >= real(1.e-5)) last = b.at(i, j);
  row_last[j - b.jlo] = last;
}
// src[v]: the last voxel before v in its slab's loop order with |vel_n| >=
// 1e-5 (-1 if none, and for those voxels themselves); one thread per row
__global__ void stale_src2d_kernel(Box2 b, const real* vn, int nslabs, const int* row_last,
                                   int* src) {
  const int j = b.jlo + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (j > b.jhi) return;
  int lo, hi;
  slab_of(b, nslabs, j, lo, hi);
  int carry = -1;
  for (int jj = j - 1; jj >= lo && carry < 0; --jj) carry = row_last[jj - b.jlo];
  for (int i = b.ilo; i <= b.ihi; ++i) {
    const int idx = b.at(i, j);
    if (fabs(vn[idx]) >= real(1.e-5)) {
      src[idx] = -1;
      carry = idx;
    } else {
      src[idx] = carry;
    }
  }
}
void launch_stale_src2d(const Box2& b, const real* vn, int nslabs, int* row_last, int* src,
                        hipStream_t st) {
  const int rows = b.jhi - b.jlo + 1;
  const dim3 g((rows + 63) / 64), t(64);
  hipLaunchKernelGGL(stale_rows2d_kernel, g, t, 0, st, b, vn, row_last);
  hipLaunchKernelGGL(stale_src2d_kernel, g, t, 0, st, b, vn, nslabs, (const int*)row_last, src);
}

template <int S, bool VAR, int STAGE>
__global__ void rhs2d_kernel(Rhs2Args a) {
  const Box2 b = a.b;
  const int n = b.nfill();
  const real inv_dx = real(1.0) / a.dx, inv_dy = real(1.0) / a.dy;
  const real fx = real(0.5) / a.dx, fy = real(0.5) / a.dy;
  const real dt = STAGE ? *a.dt : real(0);
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < n; f += gridDim.x * blockDim.x) {
    int i, j;
    fill_ij(b, f, i, j);
    const int idx = b.at(i, j);
    const real* P = a.in;
    real d[4];
    derivs2d<S>(P, b, i, j, inv_dx, inv_dy, d);
    const real vn = VAR ? a.vn[idx] : a.a_const;
    real temp_norm;
    if (fabs(vn) >= real(1.e-5)) temp_norm = norm2d(vn, d);
    else temp_norm = stale_norm2d<S, VAR>(a, i, j);
    real temp_vel = real(0);
    if (VAR) {
      const real vxv = a.vx[idx], vyv = a.vy[idx];
      const real ax = vxv >= real(0) ? d[1] : d[0];
      const real ay = vyv >= real(0) ? d[3] : d[2];
      temp_vel = -(vxv * ax + vyv * ay);
    }
    // curvature term (qss_spatial_derivatives2d.f:445-466)
    const real c = P[idx];
    const real px = (P[idx + 1] - P[idx - 1]) * fx;
    const real py = (P[idx + b.nx] - P[idx - b.nx]) * fy;
    const real pxx = (P[idx + 2] + P[idx - 2] - real(2) * c) * fx * fx;
    const real pyy = (P[idx + 2 * b.nx] + P[idx - 2 * b.nx] - real(2) * c) * fy * fy;
    const real pxy = (P[idx + 1 + b.nx] + P[idx - 1 - b.nx] - P[idx - 1 + b.nx] - P[idx + 1 - b.nx])
                     * fx * fy;
    const real g2 = px * px + py * py;
    real curv = real(0);
    if (!(g2 < real(1.e-5))) curv = (pxx * py * py + pyy * px * px - real(2) * pxy * px * py) / g2;
    const real temp_curv = (VAR ? a.bb[idx] : a.b_const) * curv;
    const real rhs = temp_norm + temp_curv + temp_vel;
    real val;
    if (STAGE == 0) val = rhs;  // the right-hand side alone
    else if (STAGE == 1) val = c + dt * rhs;
    else if (STAGE == 2) val = real(0.75) * a.un[idx] + real(0.25) * (c + dt * rhs);
    else val = real(1.0 / 3.0) * a.un[idx] + real(2.0 / 3.0) * (c + dt * rhs);
    a.out[idx] = val;
  }
}

// dt from H/dx at the last fill voxel (ihi, slab's last row) of each j-slab
// (the reference resets max_H_over_dX inside its voxel loop), CFL 0.5:
// dt = 0.5 / (max H + b_max/dx + U_max/dx)
template <int S, bool VAR>
__device__ real h2d_at(const Rhs2Args& a, int i, int j) {
  const Box2& b = a.b;
  const real inv_dx = real(1.0) / a.dx, inv_dy = real(1.0) / a.dy;
  real d[4];
  derivs2d<S>(a.in, b, i, j, inv_dx, inv_dy, d);
  const real vn = VAR ? a.vn[b.at(i, j)] : a.a_const;
  const real mdx = tmax(a.dx, a.dy);
  const real max_dx_sq = mdx * mdx;
  const real xc = tmax(fabs(d[0]), fabs(d[1])), yc = tmax(fabs(d[2]), fabs(d[3]));
  const real ng = sqrt(xc * xc + yc * yc + max_dx_sq);
  return fabs(vn) / ng * (xc * inv_dx + yc * inv_dy);
}

// H/dx at the last fill voxel (ihi, jhi), what one call of the reference's
// 2D right-hand side returns
template <int S, bool VAR>
__global__ void h2d_kernel(Rhs2Args a, real* h) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *h = h2d_at<S, VAR>(a, a.b.ihi, a.b.jhi);
}

template <int S, bool VAR>
__global__ void probe2d_kernel(Rhs2Args a, real bmax_over_dx, real umax_over_dx, real* dt_out) {
  __shared__ real hs[1024];
  const Box2 b = a.b;
  const int t = threadIdx.x;
  const int nsl = b.jhi - b.jlo + 1;
  real h = real(-1);
  if (t < a.nslabs) {
    int jhi = b.jlo + (int)((int64_t)nsl * (t + 1) / a.nslabs) - 1;
    int jlo = b.jlo + (int)((int64_t)nsl * t / a.nslabs);
    if (jhi > b.jhi) jhi = b.jhi;
    if (jlo <= jhi) h = h2d_at<S, VAR>(a, b.ihi, jhi);
  }
  hs[t] = h;
  __syncthreads();
  if (t == 0) {
    real cur = real(-1);
    for (int s = 0; s < a.nslabs; ++s)
      if (hs[s] > cur) cur = hs[s];
    *dt_out = real(0.5) / (cur + bmax_over_dx + umax_over_dx);
  }
}

// ---------------------------------------------------------------------------
// Subcell-fix reinitialization in 2D (qss_reinitialization2d.f:550-768,
// reinitialize_top.c:123-203).  Unlike the 3D form: a voxel is next to the
// interface when phi0 times a face neighbour's phi0 is negative, delta takes
// the full central difference |phi0(+1) - phi0(-1)|, and D0 = 2 phi0 / delta.

__global__ void reinit2d_distance_kernel(Box2 b, real* d0, const real* p0, real dx, real dy) {
  const int n = b.nfill();
  const real flag = real(-1000.0) * tmax(dx, dy);
  const real zero_tol = real(1.e-5);
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < n; f += gridDim.x * blockDim.x) {
    int i, j;
    fill_ij(b, f, i, j);
    const int idx = b.at(i, j);
    const real c = p0[idx];
    const real xp = p0[idx + 1], xm = p0[idx - 1], yp = p0[idx + b.nx], ym = p0[idx - b.nx];
    const bool near = (xp * c < real(0)) || (xm * c < real(0)) || (yp * c < real(0)) ||
                      (ym * c < real(0));
    if (!near) {
      d0[idx] = flag;
    } else {
      real delta = tmax(tmax(tmax(fabs(xp - xm), fabs(xp - c)), fabs(c - xm)), zero_tol);
      delta = tmax(tmax(tmax(fabs(yp - ym), fabs(yp - c)), fabs(c - ym)), delta);
      d0[idx] = real(2) * c / delta;
    }
  }
}

__global__ void reinit2d_step_kernel(Box2 b, real* next, const real* data, const real* p0,
                                     const real* d0, real dt_r, real dx, real dy) {
  const int n = b.nfill();
  const real max_dx = tmax(dx, dy);
  const real flag = real(-1000.0) * max_dx;
  const real inv_dx = real(1.0) / dx, inv_dy = real(1.0) / dy;
  const real zero_tol = real(1.e-5);
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < n; f += gridDim.x * blockDim.x) {
    int i, j;
    fill_ij(b, f, i, j);
    const int idx = b.at(i, j);
    const real c = data[idx];
    real gp[2] = {(data[idx + 1] - c) * inv_dx, (data[idx + b.nx] - c) * inv_dy};
    real gm[2] = {(c - data[idx - 1]) * inv_dx, (c - data[idx - b.nx]) * inv_dy};
    const real q0 = p0[idx];
    const real sgn0 = fabs(q0) > zero_tol ? copysign(real(1), q0) : real(0);
    const real dd = d0[idx];
    real rhs;
    if (dd == flag) {
      real st[2];
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        real pp, mm;
        if (c > real(0)) { pp = tmax(-gp[k], real(0)); mm = tmax(gm[k], real(0)); }
        else { pp = tmax(gp[k], real(0)); mm = tmax(-gm[k], real(0)); }
        st[k] = tmax(pp, mm);
      }
      if (fabs(c) > zero_tol) rhs = sgn0 - sgn0 * sqrt(st[0] * st[0] + st[1] * st[1]);
      else rhs = real(0);
    } else {
      rhs = dd - sgn0 * fabs(c) / max_dx;
    }
    next[idx] = c + dt_r * rhs;
  }
}

// ---------------------------------------------------------------------------
// Coefficient fields (qss_spatial_derivatives2d.f:14-247, 1408-1460) and the
// trapped-phase velocity scaling (qss_util2d.f:355-410).

__global__ void grad2d_kernel(Box2 b, real* gx, real* gy, const real* m, real dx, real dy) {
  const int n = b.nfill();
  const real fx = real(0.5) / dx, fy = real(0.5) / dy;
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < n; f += gridDim.x * blockDim.x) {
    int i, j;
    fill_ij(b, f, i, j);
    const int idx = b.at(i, j);
    gx[idx] = (m[idx + 1] - m[idx - 1]) * fx;
    gy[idx] = (m[idx + b.nx] - m[idx - b.nx]) * fy;
  }
}

// var_a = H(-mask) a0 - s H(mask) cos(theta) |grad mask|, s = mask / sqrt(mask^2 + |grad mask|^2 dx^2)
__global__ void varnorm2d_kernel(Box2 b, real* va, const real* m, const real* mx, const real* my,
                                 real a0, real ca, real dx) {
  const int n = b.nfill();
  const real eps_var = real(1.5) * dx, C2 = real(1);
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < n; f += gridDim.x * blockDim.x) {
    int i, j;
    fill_ij(b, f, i, j);
    const int idx = b.at(i, j);
    const real h1 = heaviside(-m[idx], eps_var), h2 = heaviside(m[idx], eps_var);
    real ag = mx[idx] * mx[idx] + my[idx] * my[idx];
    ag = sqrt(ag);
    real x = m[idx] * m[idx] + ag * dx * ag * dx;
    x = sqrt(x);
    const real s = m[idx] / x;
    va[idx] = h1 * a0 - s * h2 * C2 * cos(ca) * ag;
  }
}

// var_b = H(-mask) b, V = s grad(mask) H(mask); the U bound uses 1/dx for both
// axes, as the reference does
__global__ void varcurvadv2d_kernel(Box2 b, real* vb, real* vx, real* vy, const real* m,
                                    const real* mx, const real* my, real const_b, real dx) {
  const int n = b.nfill();
  const real eps_var = real(1.5) * dx, C2 = real(1);
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < n; f += gridDim.x * blockDim.x) {
    int i, j;
    fill_ij(b, f, i, j);
    const int idx = b.at(i, j);
    const real h1 = heaviside(-m[idx], eps_var), h2 = heaviside(m[idx], eps_var);
    real ag = mx[idx] * mx[idx] + my[idx] * my[idx];
    ag = sqrt(ag);
    real x = m[idx] * m[idx] + ag * dx * ag * dx;
    x = sqrt(x);
    const real s = m[idx] / x;
    vb[idx] = h1 * const_b;
    vx[idx] = s * C2 * mx[idx] * h2;
    vy[idx] = s * C2 * my[idx] * h2;
  }
}

// vel_n, b, V scaled by H(-mask_w) H(-mask_nw)
__global__ void trapvel2d_kernel(Box2 b, const real* mask_w, const real* mask_nw, real* vn,
                                 real* vb, real* vx, real* vy, real dx) {
  const int n = b.nfill();
  const real eps_var = real(1.5) * dx;
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < n; f += gridDim.x * blockDim.x) {
    int i, j;
    fill_ij(b, f, i, j);
    const int idx = b.at(i, j);
    const real h1 = heaviside(-mask_w[idx], eps_var), h2 = heaviside(-mask_nw[idx], eps_var);
    vn[idx] = h1 * h2 * vn[idx];
    vb[idx] = h1 * h2 * vb[idx];
    vx[idx] = h1 * h2 * vx[idx];
    vy[idx] = h1 * h2 * vy[idx];
  }
}

// initializeDisconnectedMasks in 2D: a circle of radius 2 (x_hi - x_lo)
// about the domain centre, negative inside (qss_general_util.c:711-719,
// qss_initialization2d.c:127-178)
__global__ void circle2d_kernel(Box2 b, real* f, real x0, real y0, real dx, real dy, real cx,
                                real cy, real r) {
  const int n = b.nx * b.ny;
  for (int l = blockIdx.x * blockDim.x + threadIdx.x; l < n; l += gridDim.x * blockDim.x) {
    const int j = l / b.nx, i = l - j * b.nx;
    const real x = x0 + dx * i, y = y0 + dy * j;
    // the reference calls the double sqrt and subtracts in double
    const real s2 = (x - cx) * (x - cx) + (y - cy) * (y - cy);
    f[l] = (real)(sqrt((double)s2) - (double)r);
  }
}

// Kernel-level pieces (qss_tvd_runge_kutta2d.f; the one-edge signed linear
// extrapolation of qss_spatial_derivatives2d.f): the fill box of b only.
template <int STAGE>
__global__ void rk2d_kernel(Box2 b, real* out, const real* u1, const real* un, const real* rhs,
                            real dt) {
  const int n = b.nfill();
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < n; f += gridDim.x * blockDim.x) {
    int i, j;
    fill_ij(b, f, i, j);
    const int idx = b.at(i, j);
    if (STAGE == 1) out[idx] = u1[idx] + dt * rhs[idx];
    else if (STAGE == 2) out[idx] = real(0.75) * un[idx] + real(0.25) * (u1[idx] + dt * rhs[idx]);
    else out[idx] = real(1.0 / 3.0) * un[idx] + real(2.0 / 3.0) * (u1[idx] + dt * rhs[idx]);
  }
}

// edge 0..3 = x-, x+, y-, y+: the ghost layers beyond the fill box on that
// edge, over the whole ghost range of the other axis
__global__ void edge2d_kernel(Box2 b, real* phi, int edge) {
  const int ax = edge >> 1, hi = edge & 1;
  const int cols = ax == 0 ? b.ny : b.nx;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < cols; c += gridDim.x * blockDim.x) {
    auto at = [&](int w) { return ax == 0 ? b.at(w, c) : b.at(c, w); };
    const int lo_f = ax == 0 ? b.ilo : b.jlo, hi_f = ax == 0 ? b.ihi : b.jhi;
    const int nax = ax == 0 ? b.nx : b.ny;
    const int e = hi ? hi_f : lo_f, e2 = hi ? hi_f - 1 : lo_f + 1;
    const real pe = phi[at(e)];
    const real slope = copysign(real(1), pe) * fabs(pe - phi[at(e2)]);
    if (hi) {
      for (int w = hi_f + 1; w < nax; ++w) phi[at(w)] = pe + slope * real(w - hi_f);
    } else {
      for (int w = 0; w < lo_f; ++w) phi[at(w)] = pe + slope * real(lo_f - w);
    }
  }
}

// ---------------------------------------------------------------------------
// Host side

struct Run2 {
  Box2 b;
  Box b3;
  int64_t n;
  hipStream_t st = 0;
  std::vector<void*> owned;
  real *phi = nullptr, *next = nullptr, *tmp = nullptr, *prev = nullptr, *mask = nullptr;
  real *mask_w = nullptr, *mask_nw = nullptr, *vn = nullptr, *bb = nullptr, *vx = nullptr,
       *vy = nullptr, *mx = nullptr, *my = nullptr, *s = nullptr, *d0 = nullptr, *p0 = nullptr,
       *zero = nullptr;
  real* dt = nullptr;
  double *part = nullptr, *red = nullptr;
  uint8_t* bin = nullptr;
  int *labels = nullptr, *troot = nullptr, *roots = nullptr;
  int *stale_src = nullptr, *stale_rows = nullptr;  // stale_src2d of vn
  int64_t roots_cap = 0;
  real dx, dy;
  int nslabs = 1;

  template <typename P>
  P* alloc(int64_t count) {
    void* p = nullptr;
    X2CHECK(hipMalloc(&p, (size_t)(count > 0 ? count : 1) * sizeof(P)));
    owned.push_back(p);
    return (P*)p;
  }
  explicit Run2(const Grid* g) {
    b = box2_of(g);
    b3 = b.box3();
    n = (int64_t)b.nx * b.ny;
    dx = g->dx[0];
    dy = g->dx[1];
    const int nsl = b.jhi - b.jlo + 1;
    nslabs = dt_slab_count(nsl);
    real** f[] = {&phi, &next, &tmp, &prev, &mask, &mask_w, &mask_nw, &vn, &bb, &vx, &vy,
                  &mx, &my, &s, &d0, &p0, &zero};
    for (real** q : f) {
      *q = alloc<real>(n);
      X2CHECK(launch_fill(*q, real(0), n, st));
    }
    dt = alloc<real>(1);
    part = alloc<double>(reduce_partials());
    red = alloc<double>(4);
    bin = alloc<uint8_t>(n);
    labels = alloc<int>(n);
    troot = alloc<int>(n);
    roots_cap = n / 16 + 1;
    roots = alloc<int>(roots_cap + 1);
    stale_src = alloc<int>(n);
    stale_rows = alloc<int>(b.jhi - b.jlo + 1);
  }
  // the normal-term sources of the voxels with |vn| < 1e-5, after vn changed
  void stale() {
    launch_stale_src2d(b, vn, nslabs, stale_rows, stale_src, st);
    X2CHECK(hipGetLastError());
  }
  ~Run2() {
    for (void* p : owned) (void)hipFree(p);
  }
  void up(real* d, const real* h) {
    if (h) X2CHECK(hipMemcpy(d, h, n * sizeof(real), hipMemcpyHostToDevice));
  }
  void down(real* h, const real* d) {
    if (h) X2CHECK(hipMemcpy(h, d, n * sizeof(real), hipMemcpyDeviceToHost));
  }
  void bc(real* f) {
    BcArgs ba{};
    ba.bx = b3; ba.src = f; ba.dst = f;
    X2CHECK(launch_bc(ba, false, st));
  }
  void reinit(real*& f, real tmax_r) {
    real t_r = 0, dt_r = real(0.5) * dx;
    X2CHECK(hipMemcpyAsync(p0, f, n * sizeof(real), hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(reinit2d_distance_kernel, dim3(blocks2(b.nfill())), dim3(256), 0, st, b,
                       d0, p0, dx, dy);
    X2CHECK(hipGetLastError());
    while (t_r < tmax_r) {
      hipLaunchKernelGGL(reinit2d_step_kernel, dim3(blocks2(b.nfill())), dim3(256), 0, st, b,
                         tmp, f, p0, d0, dt_r, dx, dy);
      X2CHECK(hipGetLastError());
      bc(tmp);
      std::swap(f, tmp);
      t_r = t_r + dt_r;
    }
  }
  // out[0] = smoothed volume of {f < 0} (dV = dx dy), out[1] = local max-norm vs f2
  void reduce(const real* f, const real* f2, real zone, real eps, double out[2]) {
    X2CHECK(launch_outer_reduce(b3, f, f2, zone, eps, dx * dy, part, red, st));
    X2CHECK(hipMemcpyAsync(out, red, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    X2CHECK(hipStreamSynchronize(st));
  }
};

// main component indices (connectivity.c:219-249)
void main_ind2(Options* o, const Grid* g) {
  const int nx = g->grid_dims_ghostbox[0];
  const int ymid = (int)(0.5 * (g->jlo_fb + g->jhi_fb));
  if (o->center_inlet) {
    const int xc = (int)(0.5 * (g->ilo_fb + g->ihi_fb));
    o->phi_nw_ind = xc + ymid * nx;
    o->phi_w_ind = (g->ilo_fb - 1) + ymid * nx;
  } else {
    o->phi_nw_ind = (g->ilo_fb - 1) + ymid * nx;
    o->phi_w_ind = (g->ilo_fb + 1) + ymid * nx;
  }
}

void trap2(Run2& r, const Options* o, const Grid* g, bool full_labels = false) {
  TrapArgs ta{};
  ta.bx = r.b3;
  ta.phi = r.phi; ta.mask = r.mask; ta.mask_nw = r.mask_nw; ta.mask_w = r.mask_w;
  ta.s = r.s; ta.bin = r.bin; ta.labels = r.labels; ta.troot = r.troot;
  ta.roots = r.roots; ta.roots_cap = r.roots_cap; ta.full_labels = full_labels;
  ta.nw_ind = o->phi_nw_ind; ta.w_ind = o->phi_w_ind;
  ta.center_inlet = o->center_inlet;
  ta.fill = -20 * g->dx[0];
  ta.dims2 = 1;
  if (o->phi_nw_ind < 0 || o->phi_nw_ind >= r.n || o->phi_w_ind < 0 || o->phi_w_ind >= r.n)
    throw std::runtime_error("phi_nw_ind / phi_w_ind outside the ghost box");
  X2CHECK(launch_trap_components(ta, 26, r.st));
}

template <int S, bool VAR>
void launch_stage2(Run2& r, const Rhs2Args& a, int stage) {
  dim3 grid(blocks2(r.b.nfill())), block(256);
  if (stage == 1) hipLaunchKernelGGL((rhs2d_kernel<S, VAR, 1>), grid, block, 0, r.st, a);
  else if (stage == 2) hipLaunchKernelGGL((rhs2d_kernel<S, VAR, 2>), grid, block, 0, r.st, a);
  else hipLaunchKernelGGL((rhs2d_kernel<S, VAR, 3>), grid, block, 0, r.st, a);
  X2CHECK(hipGetLastError());
}

template <int S, bool VAR>
real inner_step2(Run2& r, Rhs2Args a, int order, real bmax, real umax) {
  a.in = r.phi;
  hipLaunchKernelGGL((probe2d_kernel<S, VAR>), dim3(1), dim3(((a.nslabs + 63) / 64) * 64), 0, r.st,
                     a, bmax, umax, r.dt);
  X2CHECK(hipGetLastError());
  a.dt = r.dt;
  a.un = r.phi;
  a.out = r.next;
  launch_stage2<S, VAR>(r, a, 1);
  r.bc(r.next);
  if (order >= 2) {
    a.in = r.next; a.out = r.tmp;
    launch_stage2<S, VAR>(r, a, 2);
    std::swap(r.next, r.tmp);
    r.bc(r.next);
  }
  if (order >= 3) {
    a.in = r.next; a.out = r.tmp;
    launch_stage2<S, VAR>(r, a, 3);
    std::swap(r.next, r.tmp);
    r.bc(r.next);
  }
  real dt = 0;
  X2CHECK(hipMemcpyAsync(&dt, r.dt, sizeof(real), hipMemcpyDeviceToHost, r.st));
  X2CHECK(hipStreamSynchronize(r.st));
  return dt;
}

// Outer loop of constCurvModel2d (VAR) / constCurvModel2dNoVar
// (constCurvModel2d.c:46-346, 348-640)
real ccm2(Run2& r, Options* o, const Grid* g, FILE* fp, bool var, real a_const) {
  const real vsmall = g->dx[0] * g->dx[1];
  const real eps = o->eps_coefficient * g->dx[0];
  const real max_satn = o->vol_frac_max;
  double out[2];
  r.reduce(r.mask, nullptr, 0, eps, out);
  const real vol_pore = (real)out[0];
  r.reduce(r.phi, nullptr, 0, eps, out);
  real vol_phi = (real)out[0];
  real satn = vol_phi / vol_pore, satn_old;
  real max_abs_err = 100, t = 0;
  int order = o->order_time_accur;
  if (order > 3) order = 3;
  int scheme = o->order_space_accur <= 2 ? 2 : 3;
  Rhs2Args a{};
  a.b = r.b; a.vn = r.vn; a.bb = r.bb; a.vx = r.vx; a.vy = r.vy;
  a.a_const = a_const; a.b_const = o->b; a.dx = r.dx; a.dy = r.dy; a.nslabs = r.nslabs;
  a.stale_src = r.stale_src;
  if (var) r.stale();
  while ((t < o->tmax) && (max_abs_err > o->eps_stop) &&
         (var || ((vol_phi > vsmall) && (satn < max_satn)))) {
    X2CHECK(hipMemcpyAsync(r.prev, r.phi, r.n * sizeof(real), hipMemcpyDeviceToDevice, r.st));
    real dt_sub = 0;
    while (dt_sub < o->tplot) {
      real dt;
      if (scheme == 2) dt = var ? inner_step2<2, true>(r, a, order, o->b_max_over_dx, o->max_U_over_dx)
                                : inner_step2<2, false>(r, a, order, o->b_max_over_dx, o->max_U_over_dx);
      else dt = var ? inner_step2<3, true>(r, a, order, o->b_max_over_dx, o->max_U_over_dx)
                    : inner_step2<3, false>(r, a, order, o->b_max_over_dx, o->max_U_over_dx);
      if (!(dt > real(0)) || !isfinite((double)dt))
        throw std::runtime_error("non-positive or non-finite time step");
      dt_sub += dt;
      // imposition over the whole box, then the boundary condition
      X2CHECK(launch_impose(r.phi, r.mask, r.next, o->overlap, r.n, r.st));
      if (o->check_connectivity) {
        X2CHECK(launch_impose(r.phi, r.mask_w, r.phi, real(0), r.n, r.st));
        X2CHECK(launch_impose(r.phi, r.mask_nw, r.phi, real(0), r.n, r.st));
      }
      r.bc(r.phi);
    }
    if (fp) fprintf(fp, "1 GPU: Level set time = %lf\n", 0.0);
    t += dt_sub;
    if (var) X2CHECK(launch_impose(r.phi, r.mask, r.phi, real(0), r.n, r.st));
    if (o->check_connectivity) {
      trap2(r, o, g);
      r.reinit(r.mask_w, o->tmax_r);
      r.reinit(r.mask_nw, o->tmax_r);
      if (var && o->conserve_imbibe) {
        hipLaunchKernelGGL(trapvel2d_kernel, dim3(blocks2(r.b.nfill())), dim3(256), 0, r.st, r.b,
                           r.mask_w, r.mask_nw, r.vn, r.bb, r.vx, r.vy, g->dx[0]);
        X2CHECK(hipGetLastError());
        r.stale();
      }
      X2CHECK(launch_merge(r.phi, r.mask_nw, r.n, r.st));
    }
    if (fp) fprintf(fp, "Reinitializing....");
    r.reinit(r.phi, o->tmax_r);
    if (fp) fprintf(fp, "Reinitialized\n");
    r.reduce(r.phi, r.prev, o->err_check_zone, eps, out);
    vol_phi = (real)out[0];
    max_abs_err = (real)out[1];
    satn_old = satn;
    satn = vol_phi / vol_pore;
    if ((satn > max_satn) || (vol_phi <= vsmall)) return -1;
    if (fp) {
      fprintf(fp, "connectivity time = %lf\n", 0.0);
      fprintf(fp, "t = %f\t", (double)t);
      fprintf(fp, "max_abs_err = %4.3f,\t", (double)max_abs_err);
      fprintf(fp, "nw phase satn pct = %4.3f\n", (double)(100 * satn));
      fflush(fp);
    }
    printf("t = %4.1f\t", (double)t);
    printf("max_abs_err = %4.3f,\t", (double)max_abs_err);
    printf("nw phase satn pct = %4.3f\n", (double)(100 * satn));
    if (o->use_satn_stop) {
      if (fabsf((float)(100 * (satn - satn_old))) < 0.001) break;
    }
  }
  return t;
}

void update_coeffs2(Run2& r, Options* o, const Grid* g, real a0, bool var) {
  const int nf = r.b.nfill();
  if (var) {
    hipLaunchKernelGGL(varcurvadv2d_kernel, dim3(blocks2(nf)), dim3(256), 0, r.st, r.b, r.bb, r.vx,
                       r.vy, r.mask, r.mx, r.my, o->b, g->dx[0]);
    X2CHECK(hipGetLastError());
    // b_max and U_max = max |vx|/dx + |vy|/dx over the fill box (a max is
    // exact in any order); the 3D reduction with a zero third component
    // adds +0 to each U
    const real inv_dx = real(1.0) / g->dx[0], inv_dy = real(1.0) / g->dx[0];
    double m[2];
    X2CHECK(launch_coeff_reduce(r.b3, r.bb, r.vx, r.vy, r.zero, inv_dx, r.part, r.red, r.st));
    X2CHECK(hipMemcpyAsync(m, r.red, 2 * sizeof(double), hipMemcpyDeviceToHost, r.st));
    X2CHECK(hipStreamSynchronize(r.st));
    const real bmax = (real)m[0], umax = (real)m[1];
    o->max_U_over_dx = umax;
    o->b_max_over_dx = 2 * bmax * (inv_dx * inv_dx + inv_dy * inv_dy);
    hipLaunchKernelGGL(varnorm2d_kernel, dim3(blocks2(nf)), dim3(256), 0, r.st, r.b, r.vn, r.mask,
                       r.mx, r.my, a0, o->theta, g->dx[0]);
    X2CHECK(hipGetLastError());
  } else {
    const real inv_dx = 1 / g->dx[0], inv_dy = 1 / g->dx[1];
    o->b_max_over_dx = 2 * o->b * (inv_dx * inv_dx + inv_dy * inv_dy);
    o->max_U_over_dx = 0;
  }
}

void up_all(Run2& r, const QSS_DataArrays* p) {
  r.up(r.phi, p->phi);
  r.up(r.mask, p->mask);
  r.up(r.mask_w, p->mask_w);
  r.up(r.mask_nw, p->mask_nw);
  r.up(r.vn, p->normal_velocity);
  r.up(r.bb, p->curvature_coeff);
  r.up(r.vx, p->external_velocity_x);
  r.up(r.vy, p->external_velocity_y);
  r.up(r.mx, p->mask_x);
  r.up(r.my, p->mask_y);
}

void down_all(Run2& r, QSS_DataArrays* p) {
  X2CHECK(hipStreamSynchronize(r.st));
  r.down(p->phi, r.phi);
  r.down(p->mask, r.mask);
  r.down(p->mask_w, r.mask_w);
  r.down(p->mask_nw, r.mask_nw);
  r.down(p->normal_velocity, r.vn);
  r.down(p->curvature_coeff, r.bb);
  r.down(p->external_velocity_x, r.vx);
  r.down(p->external_velocity_y, r.vy);
  r.down(p->mask_x, r.mx);
  r.down(p->mask_y, r.my);
}

void check2d(const Grid* g) {
  if (!g || g->num_dims != 2) throw std::runtime_error("the 2D models need a 2D grid");
  if ((int64_t)g->grid_dims_ghostbox[0] * g->grid_dims_ghostbox[1] >= ((int64_t)1 << 31))
    throw std::runtime_error("2D ghost box exceeds 2^31 voxels");
}

// drain_model2d / imbibe_model2d (drain_model2d.c:42-149, imbibe_model2d.c:42-145)
void model2d(Options* o, QSS_DataArrays* p, Grid* g, FILE* fp, bool drain) {
  check2d(g);
  if (o->use_var_theta)
    throw std::runtime_error("use_var_theta (theta.gz) is not supported by the 2D models");
  Run2 r(g);
  // drain_model2d / imbibe_model2d read only phi and mask: the trapped masks,
  // the mask gradient and the coefficient fields are set before first use
  r.up(r.phi, p->phi);
  r.up(r.mask, p->mask);
  const bool var = o->theta > 0;
  real a0 = o->a;
  const real da = o->b * o->dc;
  int step = o->init_step + 1;
  // disconnected-component masks: the domain-centred circle of radius 2 (x_hi - x_lo)
  const real cx = (real)(0.5 * (g->x_lo[0] + g->x_hi[0])), cy = (real)(0.5 * (g->x_lo[1] + g->x_hi[1]));
  const real rad = 2 * (g->x_hi[0] - g->x_lo[0]);
  for (real* m : {r.mask_w, r.mask_nw}) {
    hipLaunchKernelGGL(circle2d_kernel, dim3(blocks2(r.n)), dim3(256), 0, r.st, r.b, m,
                       g->x_lo_ghostbox[0], g->x_lo_ghostbox[1], g->dx[0], g->dx[1], cx, cy, rad);
    X2CHECK(hipGetLastError());
  }
  if (p->mask_disconn_init) r.down(p->mask_disconn_init, r.mask_nw);
  if (o->mask_reinit) r.reinit(r.mask, o->tmax_r);
  hipLaunchKernelGGL(grad2d_kernel, dim3(blocks2(r.b.nfill())), dim3(256), 0, r.st, r.b, r.mx, r.my,
                     r.mask, g->dx[0], g->dx[1]);
  X2CHECK(hipGetLastError());
  // setSpaceDerivFunc and setThetaOverlap in 2D (qss_general_util.c:84-112,196-201)
  o->overlap = (o->theta <= 20 * (real)(3.14159 / 180)) ? real(0) : 3 * g->dx[0];
  main_ind2(o, g);
  const bool files = !(getenv("LSMPQS_NO_STEP_FILES") && atoi(getenv("LSMPQS_NO_STEP_FILES")));
  std::vector<real> host_phi;
  while (drain ? (a0 < o->amax) : (a0 > o->amin)) {
    printf("\nStep %d: a = %f\n", step, (double)a0);
    if (fp) fprintf(fp, "\nStep %d: a = %f\n", step, (double)a0);
    update_coeffs2(r, o, g, a0, var);
    if (!drain) o->conserve_imbibe = 1;
    if (o->check_connectivity) trap2(r, o, g);
    const real t_step = ccm2(r, o, g, fp, var, a0);
    if (t_step == -1) break;
    if (drain) a0 += da;
    else a0 -= da;
    if (files) {
      host_phi.resize(r.n);
      X2CHECK(hipStreamSynchronize(r.st));
      r.down(host_phi.data(), r.phi);
      char fname[64];
      snprintf(fname, sizeof fname, "data_step_%d", step);
      writeDataArrayQSS(host_phi.data(), g, fname, 1);
    }
    step += 1;
  }
  // the fields the models write (theta = 0 leaves the coefficient fields alone)
  X2CHECK(hipStreamSynchronize(r.st));
  r.down(p->phi, r.phi);
  r.down(p->mask, r.mask);
  r.down(p->mask_w, r.mask_w);
  r.down(p->mask_nw, r.mask_nw);
  r.down(p->mask_x, r.mx);
  r.down(p->mask_y, r.my);
  if (var) {
    r.down(p->normal_velocity, r.vn);
    r.down(p->curvature_coeff, r.bb);
    r.down(p->external_velocity_x, r.vx);
    r.down(p->external_velocity_y, r.vy);
  }
}

void labels_down(Run2& r, int* conn) {
  std::vector<int> lab((size_t)r.n);
  X2CHECK(hipMemcpyAsync(lab.data(), r.labels, (size_t)r.n * sizeof(int), hipMemcpyDeviceToHost,
                         r.st));
  X2CHECK(hipStreamSynchronize(r.st));
  for (int64_t l = 0; l < r.n; ++l) conn[l] = lab[l] >= 0 ? lab[l] : 0;
}

// ghost-box bounds (4 ints) + fill box (4 ints) -> a Box2 indexed from the
// ghost box's lower corner; the fill box plus a halo of h must lie inside
Box2 kbox(const int* ilo_gb, const int* ihi_gb, const int* jlo_gb, const int* jhi_gb,
          const int* ilo_fb, const int* ihi_fb, const int* jlo_fb, const int* jhi_fb, int h) {
  Box2 b;
  b.nx = *ihi_gb - *ilo_gb + 1;
  b.ny = *jhi_gb - *jlo_gb + 1;
  b.ilo = *ilo_fb - *ilo_gb; b.ihi = *ihi_fb - *ilo_gb;
  b.jlo = *jlo_fb - *jlo_gb; b.jhi = *jhi_fb - *jlo_gb;
  if (b.nx <= 0 || b.ny <= 0) throw std::runtime_error("empty ghost box");
  if ((int64_t)b.nx * b.ny >= ((int64_t)1 << 31))
    throw std::runtime_error("2D ghost box exceeds 2^31 voxels");
  if (b.ilo > b.ihi || b.jlo > b.jhi) throw std::runtime_error("empty fill box");
  if (b.ilo - h < 0 || b.jlo - h < 0 || b.ihi + h >= b.nx || b.jhi + h >= b.ny)
    throw std::runtime_error("fill box plus stencil reaches outside the ghost box");
  return b;
}

void same_gb2(const Box2& b, const int* i0, const int* i1, const int* j0, const int* j1,
              const int* ilo_gb, const int* jlo_gb) {
  if (*i0 != *ilo_gb || *j0 != *jlo_gb || *i1 - *i0 + 1 != b.nx || *j1 - *j0 + 1 != b.ny)
    throw std::runtime_error("arrays of one call must share the ghost-box bounds");
}

struct Dev2 {
  std::vector<void*> p;
  ~Dev2() {
    for (void* q : p) (void)hipFree(q);
  }
  real* put(const real* h, int64_t n) {
    void* d = nullptr;
    X2CHECK(hipMalloc(&d, (size_t)(n > 0 ? n : 1) * sizeof(real)));
    p.push_back(d);
    if (h) X2CHECK(hipMemcpy(d, h, (size_t)n * sizeof(real), hipMemcpyHostToDevice));
    return (real*)d;
  }
  int* ints(int64_t n) {
    void* d = nullptr;
    X2CHECK(hipMalloc(&d, (size_t)(n > 0 ? n : 1) * sizeof(int)));
    p.push_back(d);
    return (int*)d;
  }
  static void get(real* h, const real* d, int64_t n) {
    X2CHECK(hipDeviceSynchronize());
    X2CHECK(hipMemcpy(h, d, (size_t)n * sizeof(real), hipMemcpyDeviceToHost));
  }
};

// one call of qss2dGetRHS{2nd,3rd}[const] over the given fill box (one slab)
template <int S, bool VAR>
void rhs2d_call(const real* phi, const real* vn, const real* bb, const real* vx, const real* vy,
                real* max_H, real* lse_rhs, const Box2& b, real dx, real dy, real a0, real b0) {
  const int64_t n = (int64_t)b.nx * b.ny;
  Dev2 dv;
  Rhs2Args a{};
  a.b = b;
  a.in = dv.put(phi, n);
  if (VAR) {
    a.vn = dv.put(vn, n);
    a.bb = dv.put(bb, n);
    a.vx = dv.put(vx, n);
    a.vy = dv.put(vy, n);
  }
  a.a_const = a0;
  a.b_const = b0;
  a.dx = dx;
  a.dy = dy;
  a.nslabs = 1;
  if (VAR) {
    int* src = dv.ints(n);
    launch_stale_src2d(b, a.vn, 1, dv.ints(b.jhi - b.jlo + 1), src, 0);
    X2CHECK(hipGetLastError());
    a.stale_src = src;
  }
  a.out = dv.put(lse_rhs, n);
  real* h = dv.put(nullptr, 1);
  hipLaunchKernelGGL((rhs2d_kernel<S, VAR, 0>), dim3(blocks2(b.nfill())), dim3(256), 0, 0, a);
  X2CHECK(hipGetLastError());
  hipLaunchKernelGGL((h2d_kernel<S, VAR>), dim3(1), dim3(64), 0, 0, a, h);
  X2CHECK(hipGetLastError());
  Dev2::get(lse_rhs, a.out, n);
  Dev2::get(max_H, h, 1);
}

template <int STAGE>
void rk2d_call(real* out, const real* u1, const real* un, const real* rhs, const Box2& b, real dt) {
  const int64_t n = (int64_t)b.nx * b.ny;
  Dev2 dv;
  real* d_out = dv.put(out, n);
  const real* d_u1 = dv.put(u1, n);
  const real* d_un = STAGE == 1 ? nullptr : dv.put(un, n);
  const real* d_rhs = dv.put(rhs, n);
  hipLaunchKernelGGL((rk2d_kernel<STAGE>), dim3(blocks2(b.nfill())), dim3(256), 0, 0, b, d_out,
                     d_u1, d_un, d_rhs, dt);
  X2CHECK(hipGetLastError());
  Dev2::get(out, d_out, n);
}

}  // namespace

// 2D forms of the shared helper entry points (called from their C entry
// points when num_dims == 2)

// signedLinearExtrapolationBCqss, all four edges (index 9) composed in the
// reference's order (qss_general_util.c:428-520)
void api2d_bc(real* phi, const Grid* g, int bdry_location_idx) {
  check2d(g);
  if (bdry_location_idx != 9)
    throw std::runtime_error("2D boundary conditions: only bdry_location_idx 9 (all edges) is built");
  Run2 r(g);
  r.up(r.phi, phi);
  r.bc(r.phi);
  X2CHECK(hipStreamSynchronize(r.st));
  r.down(phi, r.phi);
}

// trapComponents_mask on a 2D grid (trapComponents.c:156-375, 2D branches)
void api2d_trap(QSS_DataArrays* p, Grid* g, Options* o) {
  check2d(g);
  Run2 r(g);
  r.up(r.phi, p->phi);
  r.up(r.mask, p->mask);
  r.up(r.mask_nw, p->mask_nw);
  trap2(r, o, g, p->connectivity != nullptr);
  X2CHECK(hipStreamSynchronize(r.st));
  r.down(p->phi, r.phi);
  r.down(p->mask_nw, r.mask_nw);
  r.down(p->mask_w, r.mask_w);
  if (p->scratch1)
    for (int64_t l = 0; l < r.n; ++l) p->scratch1[l] = 0;
  if (p->phi_bin)
    X2CHECK(hipMemcpy(p->phi_bin, r.bin, (size_t)r.n, hipMemcpyDeviceToHost));
  if (p->connectivity) labels_down(r, p->connectivity);
}

// findConnectivity2d: 8-connected labels of phi_bin over the ghost box, each
// component labelled with its largest linear index, 0 on background
// (connectivity.c:86-146)
void api2d_conn(QSS_DataArrays* p, Grid* g) {
  check2d(g);
  Run2 r(g);
  X2CHECK(hipMemcpy(r.bin, p->phi_bin, (size_t)r.n, hipMemcpyHostToDevice));
  X2CHECK(launch_label_components(r.b3, r.bin, r.labels, r.troot, 26, r.st, r.roots, r.roots_cap,
                                  true));
  labels_down(r, p->connectivity);
}

// initializeDisconnectedMasks on a 2D grid (qss_general_util.c:711-719)
void api2d_disconn(real* data, const Grid* g) {
  check2d(g);
  Run2 r(g);
  const real cx = (real)(0.5 * (g->x_lo[0] + g->x_hi[0])), cy = (real)(0.5 * (g->x_lo[1] + g->x_hi[1]));
  const real rad = 2 * (g->x_hi[0] - g->x_lo[0]);
  hipLaunchKernelGGL(circle2d_kernel, dim3(blocks2(r.n)), dim3(256), 0, r.st, r.b, r.phi,
                     g->x_lo_ghostbox[0], g->x_lo_ghostbox[1], g->dx[0], g->dx[1], cx, cy, rad);
  X2CHECK(hipGetLastError());
  X2CHECK(hipStreamSynchronize(r.st));
  r.down(data, r.phi);
}

void api2d_main_ind(Options* o, const Grid* g) { main_ind2(o, g); }

// setThetaOverlap's 2D rule (qss_general_util.c:200-205)
void api2d_overlap(Options* o, const Grid* g) {
  o->overlap = (o->theta <= 20 * (real)(3.14159 / 180)) ? real(0) : 3 * g->dx[0];
}

}  // namespace lsm

using namespace lsm;

extern "C" {

void drain_model2d(Options* options, QSS_DataArrays* p, Grid* g, FILE* fp_out) {
  set_error("");
  try {
    model2d(options, p, g, fp_out, true);
  } catch (const std::exception& e) {
    set_error(e.what());
  }
}

void imbibe_model2d(Options* options, QSS_DataArrays* p, Grid* g, FILE* fp_out) {
  set_error("");
  try {
    model2d(options, p, g, fp_out, false);
  } catch (const std::exception& e) {
    set_error(e.what());
  }
}

static QSSLIB_REAL ccm2_entry(Options* o, QSS_DataArrays* p, Grid* g, FILE* fp, bool var,
                              QSSLIB_REAL a) {
  set_error("");
  try {
    check2d(g);
    Run2 r(g);
    up_all(r, p);
    // phi_nw_ind / phi_w_ind as the caller set them (getMainInd_new), as in the reference
    QSSLIB_REAL t = ccm2(r, o, g, fp, var, a);
    down_all(r, p);
    return t;
  } catch (const std::exception& e) {
    set_error(e.what());
    return (QSSLIB_REAL)-2;
  }
}

QSSLIB_REAL constCurvModel2d(Options* options, QSS_DataArrays* p, Grid* g, FILE* fp_out) {
  return ccm2_entry(options, p, g, fp_out, true, 0);
}

QSSLIB_REAL constCurvModel2dNoVar(Options* options, QSS_DataArrays* p, Grid* g, FILE* fp_out,
                                  QSSLIB_REAL a) {
  return ccm2_entry(options, p, g, fp_out, false, a);
}

// ---- kernel level, 2D (Fortran calling convention; reference:
// qss_spatial_derivatives2d.f qss2dGetRHS2nd / 2ndconst / 3rd / 3rdconst and
// the one-edge extrapolation, qss_tvd_runge_kutta2d.f) ----
#define GB4 const int *ilo_gb, const int *ihi_gb, const int *jlo_gb, const int *jhi_gb
#define FB4 const int *ilo_fb, const int *ihi_fb, const int *jlo_fb, const int *jhi_fb
#define GB4N(p) const int *p##i0, const int *p##i1, const int *p##j0, const int *p##j1
#define SAME2(b, p) same_gb2(b, p##i0, p##i1, p##j0, p##j1, ilo_gb, jlo_gb)
#define TRY2 \
  set_error("");  \
  try {
#define CATCH2                      \
  }                                 \
  catch (const std::exception& e) { \
    set_error(e.what());            \
  }

void qss2dgetrhs2nd_(const QSSLIB_REAL* phi, const QSSLIB_REAL* vel_n, const QSSLIB_REAL* b,
                     const QSSLIB_REAL* vx, const QSSLIB_REAL* vy, QSSLIB_REAL* max_H,
                     QSSLIB_REAL* lse_rhs, GB4, FB4, const QSSLIB_REAL* dx, const QSSLIB_REAL* dy) {
  TRY2 rhs2d_call<2, true>(phi, vel_n, b, vx, vy, max_H, lse_rhs,
                           kbox(ilo_gb, ihi_gb, jlo_gb, jhi_gb, ilo_fb, ihi_fb, jlo_fb, jhi_fb, 2),
                           *dx, *dy, 0, 0);
  CATCH2
}

void qss2dgetrhs3rd_(const QSSLIB_REAL* phi, const QSSLIB_REAL* vel_n, const QSSLIB_REAL* b,
                     const QSSLIB_REAL* vx, const QSSLIB_REAL* vy, QSSLIB_REAL* max_H,
                     QSSLIB_REAL* lse_rhs, GB4, FB4, const QSSLIB_REAL* dx, const QSSLIB_REAL* dy) {
  TRY2 rhs2d_call<3, true>(phi, vel_n, b, vx, vy, max_H, lse_rhs,
                           kbox(ilo_gb, ihi_gb, jlo_gb, jhi_gb, ilo_fb, ihi_fb, jlo_fb, jhi_fb, 3),
                           *dx, *dy, 0, 0);
  CATCH2
}

void qss2dgetrhs2ndconst_(const QSSLIB_REAL* phi, const QSSLIB_REAL* vel_n, const QSSLIB_REAL* b,
                          QSSLIB_REAL* max_H, QSSLIB_REAL* lse_rhs, GB4, FB4,
                          const QSSLIB_REAL* dx, const QSSLIB_REAL* dy) {
  TRY2 rhs2d_call<2, false>(phi, nullptr, nullptr, nullptr, nullptr, max_H, lse_rhs,
                            kbox(ilo_gb, ihi_gb, jlo_gb, jhi_gb, ilo_fb, ihi_fb, jlo_fb, jhi_fb, 2),
                            *dx, *dy, *vel_n, *b);
  CATCH2
}

void qss2dgetrhs3rdconst_(const QSSLIB_REAL* phi, const QSSLIB_REAL* vel_n, const QSSLIB_REAL* b,
                          QSSLIB_REAL* max_H, QSSLIB_REAL* lse_rhs, GB4, FB4,
                          const QSSLIB_REAL* dx, const QSSLIB_REAL* dy) {
  TRY2 rhs2d_call<3, false>(phi, nullptr, nullptr, nullptr, nullptr, max_H, lse_rhs,
                            kbox(ilo_gb, ihi_gb, jlo_gb, jhi_gb, ilo_fb, ihi_fb, jlo_fb, jhi_fb, 3),
                            *dx, *dy, *vel_n, *b);
  CATCH2
}

void qss2drk1step_(QSSLIB_REAL* u_next, GB4, const QSSLIB_REAL* u_cur, GB4N(a),
                   const QSSLIB_REAL* rhs, GB4N(c), FB4, const QSSLIB_REAL* dt) {
  TRY2 Box2 bx = kbox(ilo_gb, ihi_gb, jlo_gb, jhi_gb, ilo_fb, ihi_fb, jlo_fb, jhi_fb, 0);
  SAME2(bx, a);
  SAME2(bx, c);
  rk2d_call<1>(u_next, u_cur, nullptr, rhs, bx, *dt);
  CATCH2
}

void qss2dtvdrk2stage2_(QSSLIB_REAL* u_stage2, GB4, const QSSLIB_REAL* u_stage1, GB4N(a),
                        const QSSLIB_REAL* u_cur, GB4N(c), const QSSLIB_REAL* rhs, GB4N(e), FB4,
                        const QSSLIB_REAL* dt) {
  TRY2 Box2 bx = kbox(ilo_gb, ihi_gb, jlo_gb, jhi_gb, ilo_fb, ihi_fb, jlo_fb, jhi_fb, 0);
  SAME2(bx, a);
  SAME2(bx, c);
  SAME2(bx, e);
  rk2d_call<2>(u_stage2, u_stage1, u_cur, rhs, bx, *dt);
  CATCH2
}

void qss2dtvdrk3stage3_(QSSLIB_REAL* u_next, GB4, const QSSLIB_REAL* u_stage2, GB4N(a),
                        const QSSLIB_REAL* u_cur, GB4N(c), const QSSLIB_REAL* rhs, GB4N(e), FB4,
                        const QSSLIB_REAL* dt) {
  TRY2 Box2 bx = kbox(ilo_gb, ihi_gb, jlo_gb, jhi_gb, ilo_fb, ihi_fb, jlo_fb, jhi_fb, 0);
  SAME2(bx, a);
  SAME2(bx, c);
  SAME2(bx, e);
  rk2d_call<3>(u_next, u_stage2, u_cur, rhs, bx, *dt);
  CATCH2
}

void qss2dsignedlinearextrapolation_(QSSLIB_REAL* phi, GB4, FB4, const int* bdry_location_idx) {
  TRY2 const int edge = *bdry_location_idx;
  if (edge < 0 || edge > 3) return;  // the reference's routine does nothing for other indices
  Box2 bx = kbox(ilo_gb, ihi_gb, jlo_gb, jhi_gb, ilo_fb, ihi_fb, jlo_fb, jhi_fb, 0);
  if ((edge >> 1) == 0 ? bx.ihi <= bx.ilo : bx.jhi <= bx.jlo)
    throw std::runtime_error("extrapolation needs two fill voxels along the edge normal");
  const int64_t n = (int64_t)bx.nx * bx.ny;
  Dev2 dv;
  real* d = dv.put(phi, n);
  hipLaunchKernelGGL(edge2d_kernel, dim3(blocks2((edge >> 1) == 0 ? bx.ny : bx.nx)), dim3(256), 0,
                     0, bx, d, edge);
  X2CHECK(hipGetLastError());
  Dev2::get(phi, d, n);
  CATCH2
}

void findConnectivity2d(QSS_DataArrays* p, Grid* g) {
  set_error("");
  try {
    api2d_conn(p, g);
  } catch (const std::exception& e) {
    set_error(e.what());
  }
}

void reinitializeSubcellFix2d(QSSLIB_REAL* data, Grid* g, Options* o) {
  set_error("");
  try {
    check2d(g);
    Run2 r(g);
    r.up(r.phi, data);
    r.reinit(r.phi, o->tmax_r);
    X2CHECK(hipStreamSynchronize(r.st));
    r.down(data, r.phi);
  } catch (const std::exception& e) {
    set_error(e.what());
  }
}

}  // extern "C"
