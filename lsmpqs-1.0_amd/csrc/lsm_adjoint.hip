// lsm_adjoint.hip -- gradient of a saturation loss with respect to a
// per-voxel contact-angle field and the curvature coefficient b, by a
// hand-written reverse pass through the level-set time steps.
//
// Forward model: nsteps TVD-RK steps of the variable-coefficient equation
// (pressure a0 with contact angle theta(x), curvature b, advection from the
// mask) with the library's own stage and ghost-extrapolation kernels at a
// fixed time step, the mask imposition after each step, and
// loss = (S_w - target)^2 with S_w = 1 - V(phi < 0) / V(mask < 0) from the
// reference's smoothed Heaviside.  The reverse pass is the transpose of the
// derivative of exactly that map, with every discrete choice of the forward
// frozen: ENO stencils, Godunov and upwind selections, the |grad phi|^2 cut
// of the curvature, the signs of the ghost extrapolation and the imposition
// max.  WENO weights are differentiated exactly.  Local Jacobians of the
// scheme come from the forward's own templated functions evaluated on dual
// numbers (7 partials per axis line); curvature partials are written out.
// Transposed stencils scatter with atomics.  The forward keeps every step's
// state and stage values when they fit in half the free device memory;
// otherwise states are checkpointed every k steps and recomputed between
// checkpoints.
#include <math.h>
#include <stdio.h>
#include <algorithm>
#include <chrono>
#include <stdexcept>
#include <string>
#include <vector>
#include "lsm_adjoint.h"
#include "lsm_batch.h"
#include "lsm_coeff.h"
#include "lsm_xfer.h"
#include "lsmpqs_hip.h"

namespace lsm {

// ---------------------------------------------------------------------------
// Dual numbers: value and N partial derivatives.  Comparisons use the value,
// so branches follow the forward exactly.  Own namespace: the overloads are
// found by argument-dependent lookup from the scheme's templates and do not
// hide the real-valued fabs/sqrt/fmax elsewhere.
namespace ad {
template <int N>
struct Dual {
  real v;
  real d[N];
  __device__ Dual() : v(0) {
#pragma unroll
    for (int q = 0; q < N; ++q) d[q] = 0;
  }
  __device__ Dual(real x) : v(x) {
#pragma unroll
    for (int q = 0; q < N; ++q) d[q] = 0;
  }
};
template <int N> __device__ Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r(a.v + b.v);
#pragma unroll
  for (int q = 0; q < N; ++q) r.d[q] = a.d[q] + b.d[q];
  return r;
}
template <int N> __device__ Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r(a.v - b.v);
#pragma unroll
  for (int q = 0; q < N; ++q) r.d[q] = a.d[q] - b.d[q];
  return r;
}
template <int N> __device__ Dual<N> operator-(const Dual<N>& a) {
  Dual<N> r(-a.v);
#pragma unroll
  for (int q = 0; q < N; ++q) r.d[q] = -a.d[q];
  return r;
}
template <int N> __device__ Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r(a.v * b.v);
#pragma unroll
  for (int q = 0; q < N; ++q) r.d[q] = a.d[q] * b.v + a.v * b.d[q];
  return r;
}
template <int N> __device__ Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r(a.v / b.v);
#pragma unroll
  for (int q = 0; q < N; ++q) r.d[q] = (a.d[q] - r.v * b.d[q]) / b.v;
  return r;
}
template <int N> __device__ bool operator<(const Dual<N>& a, const Dual<N>& b) { return a.v < b.v; }
template <int N> __device__ bool operator>(const Dual<N>& a, const Dual<N>& b) { return a.v > b.v; }
template <int N> __device__ Dual<N> fabs(const Dual<N>& a) { return a.v < real(0) ? -a : a; }
template <int N> __device__ Dual<N> fmax(const Dual<N>& a, const Dual<N>& b) { return a.v >= b.v ? a : b; }
template <int N> __device__ Dual<N> sqrt(const Dual<N>& a) {
  Dual<N> r(::sqrt(a.v));
  const real f = a.v > real(0) ? real(0.5) / r.v : real(0);
#pragma unroll
  for (int q = 0; q < N; ++q) r.d[q] = a.d[q] * f;
  return r;
}
// the fast reconstruction's hardware reciprocal has the derivative of 1/x
template <int N> __device__ Dual<N> fast_rcp(const Dual<N>& a) { return Dual<N>(real(1)) / a; }
}  // namespace ad
using ad::Dual;

// ---------------------------------------------------------------------------
// Transposed right-hand side of one stage: for every fill voxel i,
// adj[j] += coef lam_i dL_i/dphi_j over its stencil, g_vn[i] += coef lam_i
// dL_i/dvel_n, g_b[i] += coef lam_i dL_i/db_i (b_i the curvature coefficient
// field).  L is the stage kernel's exact-order right-hand side.
struct AdjRhsArgs {
  Box bx;
  const real* phi;
  const real* lam;
  real coef;
  real* adj;
  real* g_vn;
  real* g_b;
  const real *vn, *b, *vx, *vy, *vz;
  real dx, dy, dz, norm_sign;
};

// stencil offsets of the curvature terms, in the order of curv_partials' w[]
__constant__ int kCurvOff[25][3] = {
    {0, 0, 0},                                                  // c
    {-2, 0, 0}, {-1, 0, 0}, {1, 0, 0}, {2, 0, 0},               // x line
    {0, -2, 0}, {0, -1, 0}, {0, 1, 0}, {0, 2, 0},               // y line
    {0, 0, -2}, {0, 0, -1}, {0, 0, 1}, {0, 0, 2},               // z line
    {1, 1, 0}, {-1, -1, 0}, {-1, 1, 0}, {1, -1, 0},             // xy
    {1, 0, 1}, {-1, 0, -1}, {-1, 0, 1}, {1, 0, -1},             // xz
    {0, 1, 1}, {0, -1, -1}, {0, -1, 1}, {0, 1, -1},             // yz
};

// value and partial derivatives of the curvature term (the stage kernel's
// curvature(), qss_spatial_derivatives3d.f:518-556) w.r.t. its 25 inputs
__device__ real curv_partials(const real* v, real fx, real fy, real fz, real* w) {
  const real c = v[0];
  const real xm2 = v[1], xm1 = v[2], xp1 = v[3], xp2 = v[4];
  const real ym2 = v[5], ym1 = v[6], yp1 = v[7], yp2 = v[8];
  const real zm2 = v[9], zm1 = v[10], zp1 = v[11], zp2 = v[12];
  const real* xy = v + 13;
  const real* xz = v + 17;
  const real* yz = v + 21;
  const real curv = curvature(c, xm2, xm1, xp1, xp2, ym2, ym1, yp1, yp2, zm2, zm1, zp1, zp2, xy, xz,
                              yz, fx, fy, fz);
#pragma unroll
  for (int q = 0; q < 25; ++q) w[q] = 0;
  const real px = (xp1 - xm1) * fx, py = (yp1 - ym1) * fy, pz = (zp1 - zm1) * fz;
  const real g2 = px * px + py * py + pz * pz;
  if (g2 < real(1.e-5)) return curv;
  const real pxx = (xp2 + xm2 - real(2) * c) * fx * fx;
  const real pyy = (yp2 + ym2 - real(2) * c) * fy * fy;
  const real pzz = (zp2 + zm2 - real(2) * c) * fz * fz;
  const real pxy = (xy[0] + xy[1] - xy[2] - xy[3]) * fx * fy;
  const real pxz = (xz[0] + xz[1] - xz[2] - xz[3]) * fx * fz;
  const real pyz = (yz[0] + yz[1] - yz[2] - yz[3]) * fy * fz;
  const real ig = real(1) / g2;
  // d curv / d(first, second derivatives)
  const real npx = real(2) * (pyy * px - pxy * py + pzz * px - pxz * pz);
  const real npy = real(2) * (pxx * py - pxy * px + pzz * py - pyz * pz);
  const real npz = real(2) * (pxx * pz - pxz * px + pyy * pz - pyz * py);
  const real cpx = (npx - curv * real(2) * px) * ig;
  const real cpy = (npy - curv * real(2) * py) * ig;
  const real cpz = (npz - curv * real(2) * pz) * ig;
  const real cxx = (py * py + pz * pz) * ig, cyy = (px * px + pz * pz) * ig;
  const real czz = (px * px + py * py) * ig;
  const real cxy = real(-2) * px * py * ig, cxz = real(-2) * px * pz * ig;
  const real cyz = real(-2) * py * pz * ig;
  // chain to the stencil values
  w[0] = real(-2) * (cxx * fx * fx + cyy * fy * fy + czz * fz * fz);
  w[1] = cxx * fx * fx; w[4] = cxx * fx * fx;
  w[2] = -cpx * fx;     w[3] = cpx * fx;
  w[5] = cyy * fy * fy; w[8] = cyy * fy * fy;
  w[6] = -cpy * fy;     w[7] = cpy * fy;
  w[9] = czz * fz * fz; w[12] = czz * fz * fz;
  w[10] = -cpz * fz;    w[11] = cpz * fz;
  w[13] = cxy * fx * fy; w[14] = cxy * fx * fy; w[15] = -cxy * fx * fy; w[16] = -cxy * fx * fy;
  w[17] = cxz * fx * fz; w[18] = cxz * fx * fz; w[19] = -cxz * fx * fz; w[20] = -cxz * fx * fz;
  w[21] = cyz * fy * fz; w[22] = cyz * fy * fz; w[23] = -cyz * fy * fz; w[24] = -cyz * fy * fz;
  return curv;
}

// One-sided derivatives of one axis line l[0..6] = phi(-3..+3) as the stage
// kernel forms them: FW (single precision, WENO5) the fast face
// reconstructions, phi^- from the window phi(-3..+2) and phi^+ from
// phi(-2..+3); otherwise the scheme's reference-order formulas.
template <typename T, int S, bool FW>
__device__ __forceinline__ void line_derivs(const T* l, real h, real inv_axis, real inv_x, T& plus,
                                            T& minus) {
  if constexpr (FW && S == WENO5) {
    const FaceScale<real> f = face_scale(h);
    FaceScale<T> sc;
    sc.inv = T(f.inv);
    sc.thr = T(f.thr);
    sc.c1 = T(f.c1);
    sc.c2 = T(f.c2);
    minus = weno5_face(l, sc).minus_right;
    plus = weno5_face(l + 1, sc).plus_left;
  } else {
    axis_derivs<T, S>(l, T(inv_axis), T(inv_x), plus, minus);
  }
}

// w[o] = fp d plus/d phi_o + fm d minus/d phi_o along one axis line (stride
// st through voxel idx), o = 0..6 for phi(-3..+3); zero when fp = fm = 0
template <int S, bool FW>
__device__ __forceinline__ void axis_weights(const real* phi, int64_t idx, int64_t st, real h,
                                             real inv_axis, real inv_x, real fp, real fm,
                                             real* w) {
#pragma unroll
  for (int o = 0; o < 7; ++o) w[o] = real(0);
  if (fp == real(0) && fm == real(0)) return;
  if constexpr (FW && S == WENO5) {
    // reverse mode through the two face reconstructions
    real l[7];
#pragma unroll
    for (int o = 0; o < 7; ++o) l[o] = phi[idx + (o - 3) * st];
    const FaceScale<real> sc = face_scale(h);
    weno5_face_adj(l, sc, fm, real(0), w);      // phi^- from phi(-3..+2)
    weno5_face_adj(l + 1, sc, real(0), fp, w + 1);  // phi^+ from phi(-2..+3)
    return;
  }
  if constexpr (!FW && S == WENO5) {
    // reverse mode through the reference-order reconstruction
    real l[7];
#pragma unroll
    for (int o = 0; o < 7; ++o) l[o] = phi[idx + (o - 3) * st];
    weno5_line_adj(l, inv_axis, fp, fm, w);
    return;
  }
  Dual<7> l[7];
#pragma unroll
  for (int o = 0; o < 7; ++o) {
    l[o] = Dual<7>(phi[idx + (o - 3) * st]);
    l[o].d[o] = 1;
  }
  Dual<7> plus, minus;
  line_derivs<Dual<7>, S, FW>(l, h, inv_axis, inv_x, plus, minus);
#pragma unroll
  for (int o = 0; o < 7; ++o) w[o] = fp * plus.d[o] + fm * minus.d[o];
}

// d/dd_q of the Godunov |grad phi| = sqrt(godunov_sq(vn, d)) (lsm_device.h),
// with the forward's selections: per axis x = max(s d+, -s d-, 0) with
// s = sign(vel_n), the first operand on ties, so |grad phi| = sqrt(x^2 + y^2
// + z^2) depends on one of the two one-sided derivatives of an axis (or
// neither), with derivative +-s x / |grad phi|.  Returns |grad phi|.
__device__ __forceinline__ real godunov_partials(real vn, const real* d, real* g) {
  const real s = vn > real(0) ? real(1) : real(-1);
  real x[3];
  int pick[3];  // 1: d[2a+1] selected, 0: d[2a], -1: neither
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const real hi = s * d[2 * a + 1], lo = -(s * d[2 * a]);
    const bool first = hi >= lo;
    const real v = first ? hi : lo;
    x[a] = v >= real(0) ? v : real(0);
    pick[a] = v >= real(0) ? (first ? 1 : 0) : -1;
  }
  const real g2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
  const real gn = sqrt(g2);
  const real f = g2 > real(0) ? real(1) / gn : real(0);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const real w = s * x[a] * f;
    g[2 * a] = pick[a] == 0 ? -w : real(0);
    g[2 * a + 1] = pick[a] == 1 ? w : real(0);
  }
  return gn;
}

// Transposed right-hand side, one thread per fill voxel with lam != 0 (x
// fastest, blocks of 256; in a loop over several voxels per thread the
// compiler overlapped consecutive voxels' loads and needed 172 VGPRs, two
// waves per SIMD, instead of 78, six).  The curvature stencil's axis points
// are added to the axis lines' weights and the centre's contributions are
// summed before the atomic adds: 31 targets (6 per axis line, the centre,
// the 12 cross points) instead of 21 + 25.  Each axis line's weights are
// added as soon as they are formed.
template <int S, bool FW>
__global__ __launch_bounds__(256) void adj_rhs_kernel(AdjRhsArgs a) {
  const Box bx = a.bx;
  const int nxf = bx.ihi - bx.ilo + 1, nyf = bx.jhi - bx.jlo + 1;
  const int64_t f = (int64_t)blockIdx.x * 256 + threadIdx.y * ROW_TX + threadIdx.x;
  if (f >= (int64_t)nxf * nyf * (bx.khi - bx.klo + 1)) return;
  const int64_t row = f / nxf;
  const int64_t idx = bx.at(bx.ilo + (int)(f - row * nxf), bx.jlo + (int)(row % nyf),
                            bx.klo + (int)(row / nyf));
  const real ix = real(1.0) / a.dx, iy = real(1.0) / a.dy, iz = real(1.0) / a.dz;
  const int64_t sy = bx.nx, sz = (int64_t)bx.nx * bx.ny;
  {
    const real lam = a.lam[idx] * a.coef;
    if (lam == real(0)) return;
    // the six one-sided derivatives, as the stage kernel forms them
    real d[6];
    {
      real v[7];
#pragma unroll
      for (int o = 0; o < 7; ++o) v[o] = a.phi[idx + (o - 3)];
      line_derivs<real, S, FW>(v, a.dx, ix, ix, d[0], d[1]);
#pragma unroll
      for (int o = 0; o < 7; ++o) v[o] = a.phi[idx + (o - 3) * sy];
      line_derivs<real, S, FW>(v, a.dy, iy, ix, d[2], d[3]);
#pragma unroll
      for (int o = 0; o < 7; ++o) v[o] = a.phi[idx + (o - 3) * sz];
      line_derivs<real, S, FW>(v, a.dz, iz, ix, d[4], d[5]);
    }
    // partials of the right-hand side w.r.t. them and vel_n
    const real vn = a.vn[idx];
    const real vxv = a.vx[idx], vyv = a.vy[idx], vzv = a.vz[idx];
    real fd[6] = {0, 0, 0, 0, 0, 0};
    real fvn = 0;
    if (fabs(vn) >= real(1.e-5)) {
      real g[6];
      const real gn = godunov_partials(vn, d, g);
#pragma unroll
      for (int q = 0; q < 6; ++q) fd[q] = a.norm_sign * vn * g[q];
      fvn = a.norm_sign * gn;
    }
    fd[vxv >= real(0) ? 1 : 0] -= vxv;
    fd[vyv >= real(0) ? 3 : 2] -= vyv;
    fd[vzv >= real(0) ? 5 : 4] -= vzv;
    a.g_vn[idx] += lam * fvn;
    // curvature term: partials w.r.t. its 25 stencil values (kCurvOff order);
    // the cross points' contributions are added here
    const real bb = a.b[idx];
    real cw[25];
    {
      const real fx = real(0.5) / a.dx, fy = real(0.5) / a.dy, fz = real(0.5) / a.dz;
      auto off = [&](int q) {  // block-uniform offset of stencil point q from idx
        return (int64_t)kCurvOff[q][0] + kCurvOff[q][1] * sy + kCurvOff[q][2] * sz;
      };
      const real* pb = a.phi + idx;
      real cv[25];
#pragma unroll
      for (int q = 0; q < 25; ++q) cv[q] = pb[off(q)];
      const real curv = curv_partials(cv, fx, fy, fz, cw);
      a.g_b[idx] += lam * curv;
#pragma unroll
      for (int q = 0; q < 25; ++q) cw[q] *= bb;
      real* ab = a.adj + idx;
#pragma unroll
      for (int q = 13; q < 25; ++q)
        if (cw[q] != real(0)) atomicAdd(ab + off(q), lam * cw[q]);
    }
    // one-sided derivatives, one axis line at a time, with the curvature
    // stencil's points on that line (offsets -2, -1, +1, +2)
    real cen = cw[0];
    real w[7];
    auto line = [&](int64_t st, int c) {
      w[1] += cw[c]; w[2] += cw[c + 1]; w[4] += cw[c + 2]; w[5] += cw[c + 3];
      cen += w[3];
#pragma unroll
      for (int o = 0; o < 7; ++o)
        if (o != 3 && w[o] != real(0)) atomicAdd(&a.adj[idx + (o - 3) * st], lam * w[o]);
    };
    axis_weights<S, FW>(a.phi, idx, 1, a.dx, ix, ix, fd[0], fd[1], w);
    line(1, 1);
    axis_weights<S, FW>(a.phi, idx, sy, a.dy, iy, ix, fd[2], fd[3], w);
    line(sy, 5);
    axis_weights<S, FW>(a.phi, idx, sz, a.dz, iz, ix, fd[4], fd[5], w);
    line(sz, 9);
    if (cen != real(0)) atomicAdd(&a.adj[idx], lam * cen);
  }
}

// ---------------------------------------------------------------------------
// Transposed ghost extrapolation.  A ghost value is a linear combination of
// at most 8 fill values once the signs the forward took are fixed:
// extrap(c, c2, d) = c + s t (c - c2) d with s = sign(c), t = sign(c - c2).
struct Lin {
  real v;
  int n;
  int64_t idx[8];
  real w[8];
};
__device__ Lin lin_fetch(const real* f, const Box& bx, int i, int j, int k) {
  Lin r;
  r.idx[0] = bx.at(i, j, k);
  r.v = f[r.idx[0]];
  r.w[0] = 1;
  r.n = 1;
  return r;
}
__device__ Lin lin_extrap(const Lin& c, const Lin& c2, int dist) {
  const real s = copysign(real(1), c.v);
  const real diff = c.v - c2.v;
  const real t = diff >= real(0) ? real(1) : real(-1);
  Lin r;
  r.v = c.v + s * fabs(diff) * real(dist);
  const real kk = s * t * real(dist);
  r.n = 0;
  for (int q = 0; q < c.n; ++q) { r.idx[r.n] = c.idx[q]; r.w[r.n++] = c.w[q] * (real(1) + kk); }
  for (int q = 0; q < c2.n; ++q) { r.idx[r.n] = c2.idx[q]; r.w[r.n++] = -c2.w[q] * kk; }
  return r;
}
__device__ Lin lin_bc_x(const real* f, const Box& bx, int i, int j, int k) {
  if (i < bx.ilo) return lin_extrap(lin_fetch(f, bx, bx.ilo, j, k), lin_fetch(f, bx, bx.ilo + 1, j, k), bx.ilo - i);
  if (i > bx.ihi) return lin_extrap(lin_fetch(f, bx, bx.ihi, j, k), lin_fetch(f, bx, bx.ihi - 1, j, k), i - bx.ihi);
  return lin_fetch(f, bx, i, j, k);
}
__device__ Lin lin_bc_y(const real* f, const Box& bx, int i, int j, int k) {
  if (j < bx.jlo) return lin_extrap(lin_bc_x(f, bx, i, bx.jlo, k), lin_bc_x(f, bx, i, bx.jlo + 1, k), bx.jlo - j);
  if (j > bx.jhi) return lin_extrap(lin_bc_x(f, bx, i, bx.jhi, k), lin_bc_x(f, bx, i, bx.jhi - 1, k), j - bx.jhi);
  return lin_bc_x(f, bx, i, j, k);
}
__device__ Lin lin_bc_z(const real* f, const Box& bx, int i, int j, int k) {
  if (k < bx.klo) return lin_extrap(lin_bc_y(f, bx, i, j, bx.klo), lin_bc_y(f, bx, i, j, bx.klo + 1), bx.klo - k);
  if (k > bx.khi) return lin_extrap(lin_bc_y(f, bx, i, j, bx.khi), lin_bc_y(f, bx, i, j, bx.khi - 1), k - bx.khi);
  return lin_bc_y(f, bx, i, j, k);
}

// lam's ghost entries move into its fill entries through the extrapolation
// of `src`; with FINAL only where the imposition kept the extrapolated value
template <bool FINAL>
__global__ void adj_bc_kernel(Box bx, real* lam, const real* src, const real* floor) {
  int64_t nA, nB, nC;
  ghost_regions(bx, nA, nB, nC);
  const int64_t total = nA + nB + nC;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
       g += (int64_t)gridDim.x * blockDim.x) {
    int i, j, k;
    ghost_ijk(bx, g, nA, nB, i, j, k);
    const int64_t idx = bx.at(i, j, k);
    const real lg = lam[idx];
    lam[idx] = 0;
    if (lg == real(0)) continue;
    Lin L = lin_bc_z(src, bx, i, j, k);
    if (FINAL && floor[idx] > L.v) continue;
    for (int q = 0; q < L.n; ++q) atomicAdd(&lam[L.idx[q]], lg * L.w[q]);
  }
}

// ---------------------------------------------------------------------------
// Elementwise pieces over the fill box.
// One RK stage's combination transposed, over the whole ghost box row by
// row: a += alpha lk and b += beta lk on the fill box, with "=" instead of
// "+=" for an accumulator's first term of the step (its ghost entries are
// then zeroed too), which replaces zeroing it first; a may be null.
__global__ void adj_combine_kernel(Box bx, real* a, real alpha, int a_assign, real* b, real beta,
                                   int b_assign, const real* lk) {
  for_rows<false>(bx, [&](int64_t idx, int i, int j, int k) {
    const real v = bx.in_fill(i, j, k) ? lk[idx] : real(0);
    if (a) {
      const real x = alpha * v;
      a[idx] = a_assign ? x : a[idx] + x;
    }
    const real y = beta * v;
    b[idx] = b_assign ? y : b[idx] + y;
  });
}
// The last stage's cap and imposition transposed, over the whole ghost box
// row by row: lam = 0 where the cap was taken, and on the fill box where
// the imposition took the floor (result equal to the floor).
__global__ void adj_cap_impose_kernel(Box bx, real* lam, const real* out, real cap,
                                      const real* floor) {
  for_rows<false>(bx, [&](int64_t idx, int i, int j, int k) {
    const real o = out[idx];
    if (o == cap || (bx.in_fill(i, j, k) && o == floor[idx])) lam[idx] = real(0);
  });
}
// Cap after every step, u = min(u, cap) over the ghost box (its transpose,
// lam = 0 where the cap was taken, is in adj_cap_impose_kernel).  Without reinitialization
// between steps, the signed extrapolation at a face where the mask's
// advection velocity points into the box feeds a downwind difference, and
// phi grows exponentially there (in solid, away from the interface).  A cap
// of twice the WENO5 stencil reach leaves every stencil of the |phi| < 3dx
// band of a distance-like field untouched and bounds that growth.
__global__ void adj_cap_kernel(real* u, real cap, int64_t n) {
  for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < n;
       l += (int64_t)gridDim.x * blockDim.x) {
    const real v = u[l];
    u[l] = v > cap ? cap : v;
  }
}
// d loss / d phi on the fill box: 2 (S_w - target) dV H'(phi) / V_pore, with
// H the smoothed volume fraction of the volume reduction
// (reference: src/utilities/qss_util3d.f:288-358)
__global__ void adj_loss_kernel(Box bx, real* lam, const real* phi, real scale, real eps) {
  FillIter it(bx);
  // the volume reduction's constants: d/dx [0.5 (1 - q - (1/pi) sin(pi q))], q = x/eps
  const real pi = real(3.14159265358979323846), one_over_pi = real(0.31830988618379);
  for (int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; f < it.n;
       f += (int64_t)gridDim.x * blockDim.x) {
    int i, j, k;
    it.ijk(f, i, j, k);
    const int64_t idx = bx.at(i, j, k);
    const real x = phi[idx];
    real h = real(0);
    if (x >= -eps && x < eps)
      h = (real(1) + one_over_pi * pi * cos(pi * (x / eps))) / (real(2) * eps);
    lam[idx] = scale * h;
  }
}
// vel_n = h1 a0 - s h2 cos(theta) |grad mask| per voxel (the reference's
// setVarNorm3dTheta, qss_spatial_derivatives3d.f:88-170) and its theta
// derivative s h2 sin(theta) |grad mask|
__global__ void adj_var_norm_theta_kernel(Box bx, real* vn, real* dvn, const real* mask,
                                          const real* gx, const real* gy, const real* gz,
                                          const real* theta, real a0, real dx) {
  FillIter it(bx);
  for (int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; f < it.n;
       f += (int64_t)gridDim.x * blockDim.x) {
    int i, j, k;
    it.ijk(f, i, j, k);
    const int64_t idx = bx.at(i, j, k);
    real h1, h2, ag, s;
    coeff_common(mask[idx], gx[idx], gy[idx], gz[idx], dx, h1, h2, ag, s);
    const real th = theta[idx];
    vn[idx] = h1 * a0 - s * h2 * real(1) * cos(th) * ag;
    dvn[idx] = s * h2 * sin(th) * ag;
  }
}
// d loss / d theta = g_vn dvn; partial sums of g_b h1 per block (FP64)
__global__ void adj_finish_kernel(Box bx, real* dtheta, const real* g_vn, const real* dvn,
                                  const real* g_b, const real* mask, const real* gx,
                                  const real* gy, const real* gz, real dx, double* part) {
  __shared__ double sh[256];
  FillIter it(bx);
  double acc = 0;
  for (int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; f < it.n;
       f += (int64_t)gridDim.x * blockDim.x) {
    int i, j, k;
    it.ijk(f, i, j, k);
    const int64_t idx = bx.at(i, j, k);
    dtheta[idx] = g_vn[idx] * dvn[idx];
    real h1, h2, ag, s;
    coeff_common(mask[idx], gx[idx], gy[idx], gz[idx], dx, h1, h2, ag, s);
    acc += (double)g_b[idx] * (double)h1;
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// ---------------------------------------------------------------------------
// Host driver.
namespace {

#define ACHECK(x)                                                                \
  do {                                                                           \
    hipError_t e_ = (x);                                                         \
    if (e_ != hipSuccess)                                                        \
      throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

constexpr int kBlocks = 2048;
// Single precision runs the stage kernels' fast WENO5 and curvature (the
// reference-order single-precision WENO5 forms 0/0 on flat windows); the
// reverse pass differentiates the same reconstruction.
constexpr bool kFast = sizeof(real) == 4;

struct Buffers {
  std::vector<real*> all;
  real* get(int64_t n) {
    real* p = nullptr;
    ACHECK(hipMalloc((void**)&p, (size_t)n * sizeof(real)));
    all.push_back(p);
    return p;
  }
  // count fields of n values in one allocation (each 256-byte aligned)
  std::vector<real*> slab(int count, int64_t n) {
    const int64_t stride = (n + 63) / 64 * 64;
    real* p = get(stride * count);
    std::vector<real*> f(count);
    for (int q = 0; q < count; ++q) f[q] = p + stride * q;
    return f;
  }
  void release() {
    for (real* p : all) (void)hipFree(p);
    all.clear();
  }
  ~Buffers() { release(); }
};

struct Adjoint {
  Session* s;
  StepCfg c;
  int nstage;
  real dt, cap;
  real *u1, *u2, *u3, *pre;  // stage results of one step
  real *lam, *lamn, *lams[2];
  real *g_vn, *g_b, *dvn;

  void zero(real* f) { ACHECK(launch_fill(f, real(0), s->n, s->st)); }
  void copy(real* d, const real* src) {
    ACHECK(hipMemcpyAsync(d, src, (size_t)s->n * sizeof(real), hipMemcpyDeviceToDevice, s->st));
  }

  // stage inputs/outputs of a step from u: stage k reads in[k], writes out[k]
  StepArgs args(const real* u, real* out) {
    StepArgs a = make_step_args(s, c);
    for (int k = 0; k < 3; ++k) {
      a.stage[k].fast = kFast ? 1 : 0;
      a.stage[k].nstale = 0;
      a.stage[k].eno3_quirk = 0;
      a.stage[k].un = u;
    }
    const real* ins[3] = {u, u1, u2};
    real* outs[3] = {u1, u2, out};
    if (nstage == 1) outs[0] = out;
    if (nstage == 2) outs[1] = out;
    for (int k = 0; k < nstage; ++k) {
      a.stage[k].in = ins[k];
      a.stage[k].out = outs[k];
      a.stage[k].pre = k == nstage - 1 ? pre : nullptr;
      a.bc[k].src = k == nstage - 1 ? pre : outs[k];
      a.bc[k].dst = outs[k];
      a.bc[k].floor = s->floor;
    }
    return a;
  }

  void forward_step(const real* u, real* out) {
    StepArgs a = args(u, out);
    for (int k = 0; k < nstage; ++k) {
      const bool fin = k == nstage - 1;
      ACHECK(launch_stage(a.stage[k], c.scheme, true, k + 1, fin, s->st));
      ACHECK(launch_bc(a.bc[k], fin, s->st));
    }
    hipLaunchKernelGGL(adj_cap_kernel, dim3(kBlocks), dim3(256), 0, s->st, out, cap, s->n);
    ACHECK(hipGetLastError());
  }

  void rhs_transpose(const real* phi, const real* l, real coef, real* adj) {
    AdjRhsArgs r{};
    r.bx = s->bx; r.phi = phi; r.lam = l; r.coef = coef; r.adj = adj;
    r.g_vn = g_vn; r.g_b = g_b;
    r.vn = s->vn; r.b = s->b; r.vx = s->vx; r.vy = s->vy; r.vz = s->vz;
    r.dx = c.dx; r.dy = c.dy; r.dz = c.dz; r.norm_sign = c.norm_sign;
    const dim3 b(ROW_TX, ROW_TY);
    const Box& fb = s->bx;
    const int64_t nf = (int64_t)(fb.ihi - fb.ilo + 1) * (fb.jhi - fb.jlo + 1) * (fb.khi - fb.klo + 1);
    const dim3 gv((unsigned)((nf + 255) / 256));  // one fill voxel per thread
    switch (c.scheme) {
    case 2: hipLaunchKernelGGL((adj_rhs_kernel<ENO2, false>), gv, b, 0, s->st, r); break;
    case 3: hipLaunchKernelGGL((adj_rhs_kernel<ENO3, false>), gv, b, 0, s->st, r); break;
    default:
      if (kFast) hipLaunchKernelGGL((adj_rhs_kernel<WENO5, true>), gv, b, 0, s->st, r);
      else hipLaunchKernelGGL((adj_rhs_kernel<WENO5, false>), gv, b, 0, s->st, r);
      break;
    }
    ACHECK(hipGetLastError());
  }
  void bc_transpose(real* l, const real* src, bool fin) {
    if (fin) hipLaunchKernelGGL(adj_bc_kernel<true>, dim3(kBlocks), dim3(256), 0, s->st, s->bx, l, src, s->floor);
    else hipLaunchKernelGGL(adj_bc_kernel<false>, dim3(kBlocks), dim3(256), 0, s->st, s->bx, l, src, s->floor);
    ACHECK(hipGetLastError());
  }

  // lam: adjoint of u_{n+1} in, adjoint of u_n out (u: state u_n).  The
  // step's stage values (u1, u2, u3 and the pre-imposition face layers) are
  // recomputed first unless the caller points those members at stored ones.
  void reverse_step(const real* u, bool recompute = true) {
    if (recompute) forward_step(u, u3);
    static const real alpha3[3] = {0, real(0.75), real(1.0 / 3.0)};
    static const real beta3[3] = {1, real(0.25), real(2.0 / 3.0)};
    static const real alpha2[2] = {0, real(0.75)};
    static const real beta2[2] = {1, real(0.25)};
    const real* alpha = nstage == 3 ? alpha3 : alpha2;
    const real* beta = nstage == 3 ? beta3 : beta2;
    const real* ins[3] = {u, u1, u2};
    const dim3 rg(row_blocks<false>(s->bx, 16384)), rb(ROW_TX, ROW_TY);  // for_rows' launch shape
    // cap and imposition of the last stage, then its ghost values
    hipLaunchKernelGGL(adj_cap_impose_kernel, rg, rb, 0, s->st, s->bx, lam, u3, cap, s->floor);
    ACHECK(hipGetLastError());
    bc_transpose(lam, pre, true);
    // lamn and lams[] start from their first term of this step
    bool set_n = false, set_l[2] = {false, false};
    for (int k = nstage - 1; k >= 0; --k) {
      real* lk = k == nstage - 1 ? lam : lams[k];
      real* prev = k == 0 ? lamn : lams[k - 1];
      real* a = alpha[k] != real(0) ? lamn : nullptr;  // never prev: alpha[0] = 0
      const bool a_assign = a && !set_n;
      if (a) set_n = true;
      bool& pset = k == 0 ? set_n : set_l[k - 1];
      const bool b_assign = !pset;
      pset = true;
      hipLaunchKernelGGL(adj_combine_kernel, rg, rb, 0, s->st, s->bx, a, alpha[k], a_assign ? 1 : 0,
                         prev, beta[k], b_assign ? 1 : 0, (const real*)lk);
      ACHECK(hipGetLastError());
      rhs_transpose(ins[k], lk, beta[k] * dt, prev);
      if (k > 0) bc_transpose(prev, ins[k], false);
    }
    std::swap(lam, lamn);
  }
};

}  // namespace
}  // namespace lsm

using namespace lsm;

extern "C" int lsmpqs_adjoint_run(Options* o, QSS_DataArrays* p, Grid* g, const QSSLIB_REAL* theta,
                                  QSSLIB_REAL* dloss_dtheta, lsmpqs_adjoint_args* args) {
  set_error("");
  try {
    if (!o || !p || !g || !theta || !args || !p->phi || !p->mask)
      throw std::runtime_error("lsmpqs_adjoint_run: options, arrays (phi, mask), grid, theta and args required");
    if (args->nsteps < 1 || args->checkpoint_every < 1)
      throw std::runtime_error("lsmpqs_adjoint_run: nsteps and checkpoint_every must be >= 1");
    // what the reverse pass keeps: all, segment or none (described at the forward)
    const char* sm = getenv("LSMPQS_ADJOINT_STORE");
    const std::string store_mode = sm && *sm ? sm : "all";
    if (store_mode != "all" && store_mode != "segment" && store_mode != "none")
      throw std::runtime_error("lsmpqs_adjoint_run: LSMPQS_ADJOINT_STORE must be all, segment or none");
    // LSMPQS_ADJOINT_TRACE=1: host-time breakdown on stderr
    const char* te = getenv("LSMPQS_ADJOINT_TRACE");
    const bool tr = te && *te && atoi(te) != 0;
    const auto t_start = std::chrono::steady_clock::now();
    std::vector<std::pair<const char*, double>> marks;
    auto mark = [&](const char* what, hipStream_t st) {
      if (!tr) return;
      (void)hipStreamSynchronize(st);
      marks.emplace_back(what, std::chrono::duration<double>(std::chrono::steady_clock::now() -
                                                              t_start).count());
    };
    Session* s = session_for(p, g);
    ensure_var(s);
    ensure_grad(s);
    const int64_t n = s->n;
    const size_t bytes = (size_t)n * sizeof(real);
    Buffers buf;
    Adjoint A{};
    A.s = s;
    // one allocation for the fixed fields (an allocation costs about as much
    // as a few time steps at 128^3)
    const std::vector<real*> fx = buf.slab(13, n);
    A.u1 = fx[0]; A.u2 = fx[1]; A.u3 = fx[2]; A.pre = fx[3];
    A.lam = fx[4]; A.lamn = fx[5]; A.lams[0] = fx[6]; A.lams[1] = fx[7];
    A.g_vn = fx[8]; A.g_b = fx[9]; A.dvn = fx[10];
    real* d_theta = fx[11];
    real* u = fx[12];
    ACHECK(hipMemcpyAsync(u, p->phi, bytes, hipMemcpyHostToDevice, s->st));
    ACHECK(hipMemcpyAsync(s->mask, p->mask, bytes, hipMemcpyHostToDevice, s->st));
    ACHECK(hipMemcpyAsync(d_theta, theta, bytes, hipMemcpyHostToDevice, s->st));
    for (real* f : {A.pre, A.u1, A.u2, A.u3, A.g_vn, A.g_b, A.dvn}) A.zero(f);
    const Box& bx = s->bx;
    const real dx = g->dx[0];
    // coefficients: b, V from the mask; vel_n from the theta field
    ACHECK(launch_central_grad(bx, s->gx, s->gy, s->gz, s->mask, g->dx[0], g->dx[1], g->dx[2], s->st));
    ACHECK(launch_var_curv_adv(bx, s->b, s->vx, s->vy, s->vz, s->mask, s->gx, s->gy, s->gz,
                               (real)args->b, dx, s->st));
    hipLaunchKernelGGL(adj_var_norm_theta_kernel, dim3(kBlocks), dim3(256), 0, s->st, bx, s->vn,
                       A.dvn, s->mask, s->gx, s->gy, s->gz, d_theta, (real)args->a0, dx);
    ACHECK(hipGetLastError());
    ACHECK(launch_floor(s->floor, s->mask, s->mask_w, s->mask_nw, o->overlap, 0, n, s->st));
    // step configuration: the library's variable-coefficient stages, exact order
    StepCfg& c = A.c;
    c.scheme = o->order_space_accur == 2 ? 2 : (o->order_space_accur == 5 ? 5 : 3);
    c.var = true;
    c.order = o->order_time_accur < 1 ? 1 : (o->order_time_accur > 3 ? 3 : o->order_time_accur);
    c.a_const = (real)args->a0;
    c.b_const = (real)args->b;
    c.norm_sign = real(-1);
    c.overlap = o->overlap;
    c.conn = 0;
    c.slabs = dt_slabs(g);
    c.quirk_eno3 = 0;
    c.dx = g->dx[0]; c.dy = g->dx[1]; c.dz = g->dx[2];
    A.nstage = c.order == 1 ? 1 : (c.order == 2 ? 2 : 3);
    s->nstale = 0;
    // fixed time step: given, or the reference's rule on the initial state
    real dt = (real)args->dt;
    if (!(dt > 0)) {
      const real inv = real(1.0) / dx;
      ACHECK(launch_coeff_reduce(bx, s->b, s->vx, s->vy, s->vz, inv, s->part, s->red, s->st));
      ACHECK(hipMemcpyAsync(s->h_red, s->red, 2 * sizeof(double), hipMemcpyDeviceToHost, s->st));
      ACHECK(hipStreamSynchronize(s->st));
      c.bmax = 2 * (real)s->h_red[0] * (inv * inv + inv * inv + inv * inv);
      c.umax = (real)s->h_red[1];
      ACHECK(hipMemcpyAsync(s->phi, u, bytes, hipMemcpyDeviceToDevice, s->st));
      StepArgs a = make_step_args(s, c);
      ACHECK(launch_probe(a.probe, c.scheme, s->st));
      ACHECK(hipMemcpyAsync(s->h_dt, s->dt, sizeof(real), hipMemcpyDeviceToHost, s->st));
      ACHECK(hipStreamSynchronize(s->st));
      dt = *s->h_dt;
    }
    if (!(dt > 0) || !isfinite((double)dt)) throw std::runtime_error("invalid time step");
    A.dt = dt;
    A.cap = args->phi_cap > 0 ? (real)args->phi_cap : real(6) * g->dx[0];
    args->cap_used = (double)A.cap;
    ACHECK(hipMemcpyAsync(s->dt, &dt, sizeof(real), hipMemcpyHostToDevice, s->st));
    mark("set-up", s->st);
    // What the reverse pass keeps (LSMPQS_ADJOINT_STORE):
    //   all (default): with a gradient requested and the whole trajectory
    //     (4 fields per step: the state after it and its three stage values)
    //     in half the free device memory, the forward keeps every step and the
    //     reverse pass recomputes nothing (checkpoint_every is then unused);
    //   segment, or all when the trajectory does not fit: states every
    //     checkpoint_every steps, each segment recomputed once, keeping its
    //     stage values when they fit;
    //   none: states every checkpoint_every steps, each segment recomputed,
    //     and every step once more in the reverse pass.
    // The kernels are deterministic, so every mode replays the same values.
    const int N = args->nsteps, K = std::min(args->checkpoint_every, args->nsteps);
    size_t mem_free = 0, mem_total = 0;
    const bool whole = dloss_dtheta && store_mode == "all" &&
                       hipMemGetInfo(&mem_free, &mem_total) == hipSuccess &&
                       (size_t)4 * (size_t)N * bytes < mem_free / 2;
    // whole trajectory: state q in wst[q] (wst[0] = u), stage values of step q
    // in wu1[q], wu2[q], wpre[q]
    std::vector<real*> wst, wu1, wu2, wpre;
    const int nseg = whole ? 0 : (N + K - 1) / K;
    std::vector<real*> ck(nseg), seg(whole ? 0 : K);
    real* un = nullptr;
    if (whole) {
      const std::vector<real*> f = buf.slab(4 * N, n);
      wst.push_back(u);
      wst.insert(wst.end(), f.begin(), f.begin() + N);
      wu1.assign(f.begin() + N, f.begin() + 2 * N);
      wu2.assign(f.begin() + 2 * N, f.begin() + 3 * N);
      wpre.assign(f.begin() + 3 * N, f.end());
    } else {
      const std::vector<real*> f = buf.slab(nseg + K + 1, n);
      for (int m = 0; m < nseg; ++m) ck[m] = f[m];
      for (int q = 0; q < K; ++q) seg[q] = f[nseg + q];
      un = f[nseg + K];
    }
    real* const own[4] = {A.u1, A.u2, A.u3, A.pre};
    auto use_whole = [&](int q) {
      A.u1 = wu1[q]; A.u2 = wu2[q]; A.pre = wpre[q]; A.u3 = wst[q + 1];
    };
    if (whole) {
      for (int q = 0; q < N; ++q) {
        use_whole(q);
        A.forward_step(wst[q], wst[q + 1]);
      }
      u = wst[N];
    } else {
      for (int step = 0; step < N; ++step) {
        if (step % K == 0) A.copy(ck[step / K], u);
        A.forward_step(u, un);
        std::swap(u, un);
      }
    }
    // loss
    const real eps = o->eps_coefficient * dx;
    const real dV = g->dx[0] * g->dx[1] * g->dx[2];
    double vol[2];
    const real* fields[2] = {u, s->mask};
    for (int q = 0; q < 2; ++q) {
      ACHECK(launch_outer_reduce(bx, fields[q], nullptr, 0, eps, dV, s->part, s->red, s->st));
      ACHECK(hipMemcpyAsync(s->h_red, s->red, sizeof(double), hipMemcpyDeviceToHost, s->st));
      ACHECK(hipStreamSynchronize(s->st));
      vol[q] = s->h_red[0];
    }
    const double sw = 1.0 - vol[0] / vol[1];
    const double r = sw - args->target_sw;
    args->sw = sw;
    args->loss = r * r;
    args->dt_used = (double)dt;
    ACHECK(hipMemcpyAsync(p->phi, u, bytes, hipMemcpyDeviceToHost, s->st));
    if (p->normal_velocity)  // the per-voxel vel_n of the forward map
      ACHECK(hipMemcpyAsync(p->normal_velocity, s->vn, bytes, hipMemcpyDeviceToHost, s->st));
    if (!dloss_dtheta) {  // loss only
      args->dloss_db = 0;
      ACHECK(hipStreamSynchronize(s->st));
      return 0;
    }
    mark("forward and loss", s->st);
    // reverse
    A.zero(A.lam);
    hipLaunchKernelGGL(adj_loss_kernel, dim3(kBlocks), dim3(256), 0, s->st, bx, A.lam, u,
                       (real)(2.0 * r * (double)dV / vol[1]), eps);
    ACHECK(hipGetLastError());
    // The segment's recompute keeps every step's stage values when they fit
    // in half the free device memory (3 fields per step and one for the last
    // output), so that the reverse steps need not recompute them: the
    // forward then runs twice in all instead of three times.  The stage
    // kernels are deterministic, so the stored values are the recomputed
    // ones bit for bit.
    const bool store = !whole && store_mode != "none" &&
                       hipMemGetInfo(&mem_free, &mem_total) == hipSuccess &&
                       (size_t)(3 * K + 1) * bytes < mem_free / 2;
    std::vector<real*> su1, su2, spre;
    real* slast = nullptr;
    if (store) {
      const std::vector<real*> f = buf.slab(3 * K + 1, n);
      su1.assign(f.begin(), f.begin() + K);
      su2.assign(f.begin() + K, f.begin() + 2 * K);
      spre.assign(f.begin() + 2 * K, f.begin() + 3 * K);
      slast = f[3 * K];
    }
    for (int q = whole ? N - 1 : -1; q >= 0; --q) {
      use_whole(q);
      A.reverse_step(wst[q], false);
    }
    for (int m = nseg - 1; m >= 0; --m) {
      const int s0 = m * K, s1 = std::min(N, s0 + K), len = s1 - s0;
      A.copy(seg[0], ck[m]);
      if (store) {
        auto use = [&](int q) {
          A.u1 = su1[q]; A.u2 = su2[q]; A.pre = spre[q];
          A.u3 = q + 1 < len ? seg[q + 1] : slast;
        };
        for (int q = 0; q < len; ++q) {
          use(q);
          A.forward_step(seg[q], A.u3);
        }
        for (int q = len - 1; q >= 0; --q) {
          use(q);
          A.reverse_step(seg[q], false);
        }
      } else {
        for (int step = s0 + 1; step < s1; ++step) A.forward_step(seg[step - 1 - s0], seg[step - s0]);
        for (int step = s1 - 1; step >= s0; --step) A.reverse_step(seg[step - s0]);
      }
    }
    A.u1 = own[0]; A.u2 = own[1]; A.u3 = own[2]; A.pre = own[3];
    mark("reverse", s->st);
    // theta and b gradients
    double* part = nullptr;
    ACHECK(hipMalloc((void**)&part, sizeof(double) * kBlocks));
    A.zero(d_theta);  // ghost entries carry no gradient (theta was staged here)
    hipLaunchKernelGGL(adj_finish_kernel, dim3(kBlocks), dim3(256), 0, s->st, bx, d_theta, A.g_vn,
                       A.dvn, A.g_b, s->mask, s->gx, s->gy, s->gz, dx, part);
    hipError_t le = hipGetLastError();
    std::vector<double> h(kBlocks);
    hipError_t ce = hipMemcpyAsync(h.data(), part, sizeof(double) * kBlocks, hipMemcpyDeviceToHost, s->st);
    hipError_t se = hipStreamSynchronize(s->st);
    (void)hipFree(part);
    ACHECK(le);
    ACHECK(ce);
    ACHECK(se);
    double db = 0;
    for (double v : h) db += v;
    args->dloss_db = db;
    ACHECK(copy_to_host(dloss_dtheta, d_theta, bytes, s->st));
    mark("gradients to the host", s->st);
    buf.release();
    mark("device buffers freed", s->st);
    if (tr) {
      double prev = 0;
      for (const auto& m : marks) {
        fprintf(stderr, "lsmpqs adjoint trace: %s %.4f s\n", m.first, m.second - prev);
        prev = m.second;
      }
    }
  } catch (const std::exception& e) {
    set_error(e.what());
    fprintf(stderr, "lsmpqs_hip: %s\n", e.what());
    return -2;
  }
  return 0;
}
