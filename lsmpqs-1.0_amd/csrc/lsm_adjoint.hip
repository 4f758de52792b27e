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
// Transposed stencils scatter with atomics.  States are checkpointed every
// k steps and recomputed between checkpoints.
#include <math.h>
#include <stdio.h>
#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>
#include "lsm_adjoint.h"
#include "lsm_batch.h"
#include "lsm_coeff.h"
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

// One thread per fill voxel with lam != 0, x rows per wave (for_rows).  The
// weights of the three axis lines and of the curvature stencil are summed
// per target voxel before the atomic adds: 31 targets (the x line, the y and
// z lines without the centre, the 12 cross points) instead of 21 + 25.
template <int S, bool FW>
__global__ __launch_bounds__(256) void adj_rhs_kernel(AdjRhsArgs a) {
  const Box bx = a.bx;
  const real fx = real(0.5) / a.dx, fy = real(0.5) / a.dy, fz = real(0.5) / a.dz;
  const real ix = real(1.0) / a.dx, iy = real(1.0) / a.dy, iz = real(1.0) / a.dz;
  const int64_t sy = bx.nx, sz = (int64_t)bx.nx * bx.ny;
  for_rows<true>(bx, [&](int64_t idx, int i, int j, int k) {
    const real lam = a.lam[idx] * a.coef;
    if (lam == real(0)) return;
    // the six one-sided derivatives, as the stage kernel forms them
    real lx[7], ly[7], lz[7], d[6];
#pragma unroll
    for (int o = 0; o < 7; ++o) {
      lx[o] = a.phi[idx + (o - 3)];
      ly[o] = a.phi[idx + (o - 3) * sy];
      lz[o] = a.phi[idx + (o - 3) * sz];
    }
    line_derivs<real, S, FW>(lx, a.dx, ix, ix, d[0], d[1]);
    line_derivs<real, S, FW>(ly, a.dy, iy, ix, d[2], d[3]);
    line_derivs<real, S, FW>(lz, a.dz, iz, ix, d[4], d[5]);
    // partials of the right-hand side w.r.t. them and vel_n
    const real vn = a.vn[idx], bb = a.b[idx];
    const real vxv = a.vx[idx], vyv = a.vy[idx], vzv = a.vz[idx];
    real fd[6] = {0, 0, 0, 0, 0, 0};
    real fvn = 0;
    if (fabs(vn) >= real(1.e-5)) {
      Dual<6> D[6];
#pragma unroll
      for (int q = 0; q < 6; ++q) { D[q] = Dual<6>(d[q]); D[q].d[q] = 1; }
      Dual<6> G = sqrt(godunov_sq(Dual<6>(vn), D));
#pragma unroll
      for (int q = 0; q < 6; ++q) fd[q] = a.norm_sign * vn * G.d[q];
      fvn = a.norm_sign * G.v;
    }
    fd[vxv >= real(0) ? 1 : 0] -= vxv;
    fd[vyv >= real(0) ? 3 : 2] -= vyv;
    fd[vzv >= real(0) ? 5 : 4] -= vzv;
    // curvature term
    real cv[25], cw[25];
#pragma unroll
    for (int q = 0; q < 25; ++q)
      cv[q] = a.phi[bx.at(i + kCurvOff[q][0], j + kCurvOff[q][1], k + kCurvOff[q][2])];
    const real curv = curv_partials(cv, fx, fy, fz, cw);
    a.g_vn[idx] += lam * fvn;
    a.g_b[idx] += lam * curv;
    // one-sided derivatives, one axis line at a time, then the curvature
    // stencil's centre and axis points onto the lines (kCurvOff order)
    real wx[7], wy[7], wz[7];
    axis_weights<S, FW>(a.phi, idx, 1, a.dx, ix, ix, fd[0], fd[1], wx);
    axis_weights<S, FW>(a.phi, idx, sy, a.dy, iy, ix, fd[2], fd[3], wy);
    axis_weights<S, FW>(a.phi, idx, sz, a.dz, iz, ix, fd[4], fd[5], wz);
    wx[3] += wy[3] + wz[3] + bb * cw[0];
    wx[1] += bb * cw[1]; wx[2] += bb * cw[2]; wx[4] += bb * cw[3]; wx[5] += bb * cw[4];
    wy[1] += bb * cw[5]; wy[2] += bb * cw[6]; wy[4] += bb * cw[7]; wy[5] += bb * cw[8];
    wz[1] += bb * cw[9]; wz[2] += bb * cw[10]; wz[4] += bb * cw[11]; wz[5] += bb * cw[12];
#pragma unroll
    for (int o = 0; o < 7; ++o) {
      if (wx[o] != real(0)) atomicAdd(&a.adj[idx + (o - 3)], lam * wx[o]);
      if (o != 3 && wy[o] != real(0)) atomicAdd(&a.adj[idx + (o - 3) * sy], lam * wy[o]);
      if (o != 3 && wz[o] != real(0)) atomicAdd(&a.adj[idx + (o - 3) * sz], lam * wz[o]);
    }
#pragma unroll
    for (int q = 13; q < 25; ++q) {
      const real w = bb * cw[q];
      if (w != real(0))
        atomicAdd(&a.adj[bx.at(i + kCurvOff[q][0], j + kCurvOff[q][1], k + kCurvOff[q][2])], lam * w);
    }
  });
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
__global__ void adj_axpy_kernel(Box bx, real* dst, const real* src, real alpha) {
  FillIter it(bx);
  for (int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; f < it.n;
       f += (int64_t)gridDim.x * blockDim.x) {
    int i, j, k;
    it.ijk(f, i, j, k);
    const int64_t idx = bx.at(i, j, k);
    dst[idx] += alpha * src[idx];
  }
}
// lam = 0 where the imposition took the floor (result equal to the floor)
__global__ void adj_impose_kernel(Box bx, real* lam, const real* out, const real* floor) {
  FillIter it(bx);
  for (int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; f < it.n;
       f += (int64_t)gridDim.x * blockDim.x) {
    int i, j, k;
    it.ijk(f, i, j, k);
    const int64_t idx = bx.at(i, j, k);
    if (out[idx] == floor[idx]) lam[idx] = 0;
  }
}
// Cap after every step, u = min(u, cap) over the ghost box, and its
// transpose (lam = 0 where the cap was taken).  Without reinitialization
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
__global__ void adj_cap_rev_kernel(real* lam, const real* out, real cap, int64_t n) {
  for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < n;
       l += (int64_t)gridDim.x * blockDim.x)
    if (out[l] == cap) lam[l] = 0;
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
  ~Buffers() {
    for (real* p : all) (void)hipFree(p);
  }
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
    const dim3 g(row_blocks<true>(s->bx, 16384)), b(ROW_TX, ROW_TY);  // for_rows' launch shape
    switch (c.scheme) {
    case 2: hipLaunchKernelGGL((adj_rhs_kernel<ENO2, false>), g, b, 0, s->st, r); break;
    case 3: hipLaunchKernelGGL((adj_rhs_kernel<ENO3, false>), g, b, 0, s->st, r); break;
    default:
      if (kFast) hipLaunchKernelGGL((adj_rhs_kernel<WENO5, true>), g, b, 0, s->st, r);
      else hipLaunchKernelGGL((adj_rhs_kernel<WENO5, false>), g, b, 0, s->st, r);
      break;
    }
    ACHECK(hipGetLastError());
  }
  void bc_transpose(real* l, const real* src, bool fin) {
    if (fin) hipLaunchKernelGGL(adj_bc_kernel<true>, dim3(kBlocks), dim3(256), 0, s->st, s->bx, l, src, s->floor);
    else hipLaunchKernelGGL(adj_bc_kernel<false>, dim3(kBlocks), dim3(256), 0, s->st, s->bx, l, src, s->floor);
    ACHECK(hipGetLastError());
  }
  void axpy(real* d, const real* src, real alpha) {
    hipLaunchKernelGGL(adj_axpy_kernel, dim3(kBlocks), dim3(256), 0, s->st, s->bx, d, src, alpha);
    ACHECK(hipGetLastError());
  }

  // lam: adjoint of u_{n+1} in, adjoint of u_n out (u: state u_n)
  void reverse_step(const real* u) {
    forward_step(u, u3);  // u1, u2, u3 and the pre-imposition face layers of this step
    static const real alpha3[3] = {0, real(0.75), real(1.0 / 3.0)};
    static const real beta3[3] = {1, real(0.25), real(2.0 / 3.0)};
    static const real alpha2[2] = {0, real(0.75)};
    static const real beta2[2] = {1, real(0.25)};
    const real* alpha = nstage == 3 ? alpha3 : alpha2;
    const real* beta = nstage == 3 ? beta3 : beta2;
    const real* ins[3] = {u, u1, u2};
    zero(lamn);
    for (real* l : lams) zero(l);
    // imposition of the last stage, then its ghost values
    hipLaunchKernelGGL(adj_cap_rev_kernel, dim3(kBlocks), dim3(256), 0, s->st, lam, u3, cap, s->n);
    hipLaunchKernelGGL(adj_impose_kernel, dim3(kBlocks), dim3(256), 0, s->st, s->bx, lam, u3, s->floor);
    ACHECK(hipGetLastError());
    bc_transpose(lam, pre, true);
    for (int k = nstage - 1; k >= 0; --k) {
      real* lk = k == nstage - 1 ? lam : lams[k];
      real* prev = k == 0 ? lamn : lams[k - 1];
      if (alpha[k] != real(0)) axpy(lamn, lk, alpha[k]);
      axpy(prev, lk, beta[k]);
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
    Session* s = session_for(p, g);
    ensure_var(s);
    ensure_grad(s);
    const int64_t n = s->n;
    const size_t bytes = (size_t)n * sizeof(real);
    Buffers buf;
    Adjoint A{};
    A.s = s;
    A.u1 = buf.get(n); A.u2 = buf.get(n); A.u3 = buf.get(n); A.pre = buf.get(n);
    A.lam = buf.get(n); A.lamn = buf.get(n); A.lams[0] = buf.get(n); A.lams[1] = buf.get(n);
    A.g_vn = buf.get(n); A.g_b = buf.get(n); A.dvn = buf.get(n);
    real* d_theta = buf.get(n);
    real* u = buf.get(n);
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
    // forward with checkpoints every k steps
    const int N = args->nsteps, K = std::min(args->checkpoint_every, args->nsteps);
    const int nseg = (N + K - 1) / K;
    std::vector<real*> ck(nseg), seg(K);
    for (int m = 0; m < nseg; ++m) ck[m] = buf.get(n);
    for (int q = 0; q < K; ++q) seg[q] = buf.get(n);
    real* un = buf.get(n);
    for (int step = 0; step < N; ++step) {
      if (step % K == 0) A.copy(ck[step / K], u);
      A.forward_step(u, un);
      std::swap(u, un);
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
    // reverse
    A.zero(A.lam);
    hipLaunchKernelGGL(adj_loss_kernel, dim3(kBlocks), dim3(256), 0, s->st, bx, A.lam, u,
                       (real)(2.0 * r * (double)dV / vol[1]), eps);
    ACHECK(hipGetLastError());
    for (int m = nseg - 1; m >= 0; --m) {
      const int s0 = m * K, s1 = std::min(N, s0 + K);
      A.copy(seg[0], ck[m]);
      for (int step = s0 + 1; step < s1; ++step) A.forward_step(seg[step - 1 - s0], seg[step - s0]);
      for (int step = s1 - 1; step >= s0; --step) A.reverse_step(seg[step - s0]);
    }
    // theta and b gradients
    double* part = nullptr;
    ACHECK(hipMalloc((void**)&part, sizeof(double) * kBlocks));
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
    {
      // ghost entries carry no gradient
      std::vector<real> tmp(n);
      ACHECK(hipMemcpy(tmp.data(), d_theta, bytes, hipMemcpyDeviceToHost));
      const int nx = g->grid_dims_ghostbox[0], ny = g->grid_dims_ghostbox[1];
      for (int64_t l = 0; l < n; ++l) {
        int i = (int)(l % nx), j = (int)((l / nx) % ny), k = (int)(l / ((int64_t)nx * ny));
        bool fillv = i >= g->ilo_fb && i <= g->ihi_fb && j >= g->jlo_fb && j <= g->jhi_fb &&
                     k >= g->klo_fb && k <= g->khi_fb;
        dloss_dtheta[l] = fillv ? tmp[l] : (QSSLIB_REAL)0;
      }
    }
    ACHECK(hipStreamSynchronize(s->st));
  } catch (const std::exception& e) {
    set_error(e.what());
    fprintf(stderr, "lsmpqs_hip: %s\n", e.what());
    return -2;
  }
  return 0;
}
