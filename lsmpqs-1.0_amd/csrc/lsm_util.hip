// lsm_util.hip -- per-outer-step kernels: elementwise mask algebra, FP64
// reductions (saturation volume, steady-state measure, time-step bounds),
// subcell-fix reinitialization, the variable-coefficient fields and the
// analytic initialisers.  All are HBM-streaming kernels: grid-stride loops,
// one voxel per lane per iteration, x-fastest so every wave reads 256/512
// contiguous bytes.
#include <float.h>
#include <algorithm>
#include "lsm_coeff.h"
#include "lsm_kernels.h"

namespace lsm {

static int stream_blocks(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > 16384) b = 16384;
  if (b < 1) b = 1;
  return (int)b;
}

// ---------------------------------------------------------------------------
// Elementwise mask algebra (reference: src/utilities/qss_macros.h:79,192,239;
// src/utilities/qss_util3d.f:130-175)

__global__ void impose_kernel(real* dst, const real* m, const real* src, real ov, int64_t n) {
  for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < n;
       l += (int64_t)gridDim.x * blockDim.x) {
    real v = m[l] - ov, s = src[l];
    dst[l] = (v > s) ? v : s;
  }
}
hipError_t launch_impose(real* dst, const real* m, const real* src, real ov, int64_t n,
                         hipStream_t st) {
  hipLaunchKernelGGL(impose_kernel, dim3(stream_blocks(n)), dim3(256), 0, st, dst, m, src, ov, n);
  return hipGetLastError();
}

__global__ void floor_kernel(real* f, const real* m, const real* mw, const real* mnw, real ov,
                             int conn, int64_t n) {
  for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < n;
       l += (int64_t)gridDim.x * blockDim.x) {
    real v = m[l] - ov;
    if (conn) {
      real a = mw[l];
      v = (a > v) ? a : v;
      a = mnw[l];
      v = (a > v) ? a : v;
    }
    f[l] = v;
  }
}
hipError_t launch_floor(real* floor, const real* m, const real* mw, const real* mnw, real overlap,
                        int conn, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL(floor_kernel, dim3(stream_blocks(n)), dim3(256), 0, st, floor, m, mw, mnw,
                     overlap, conn, n);
  return hipGetLastError();
}

__global__ void merge_kernel(real* phi, const real* m, int64_t n) {
  for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < n;
       l += (int64_t)gridDim.x * blockDim.x) {
    real v = m[l];
    if (v > real(0)) phi[l] = -v;
  }
}
hipError_t launch_merge(real* phi, const real* m, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL(merge_kernel, dim3(stream_blocks(n)), dim3(256), 0, st, phi, m, n);
  return hipGetLastError();
}

__global__ void min_kernel(real* phi, const real* e, int64_t n) {
  for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < n;
       l += (int64_t)gridDim.x * blockDim.x) {
    real a = phi[l], b = e[l];
    phi[l] = (a < b) ? a : b;
  }
}
hipError_t launch_min(real* phi, const real* e, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL(min_kernel, dim3(stream_blocks(n)), dim3(256), 0, st, phi, e, n);
  return hipGetLastError();
}

__global__ void fill_kernel(real* dst, real v, int64_t n) {
  for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < n;
       l += (int64_t)gridDim.x * blockDim.x)
    dst[l] = v;
}
hipError_t launch_fill(real* dst, real v, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL(fill_kernel, dim3(stream_blocks(n)), dim3(256), 0, st, dst, v, n);
  return hipGetLastError();
}

constexpr int RED_BLOCKS = REDUCE_BLOCKS;
int reduce_partials() { return 2 * RED_BLOCKS; }

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}

// Smoothed volume of {f1 < 0} (reference: src/utilities/qss_util3d.f:288-358)
// accumulated in FP64, and the local max-norm difference
// (reference: src/utilities/qss_util3d.f:65-125).
__global__ __launch_bounds__(256) void outer_reduce_kernel(Box bx, const real* f1, const real* f2,
                                                           real zone, real eps, real dV,
                                                           double* part) {
  const real pi = real(3.14159265358979323846), one_over_pi = real(0.31830988618379);
  double vol = 0.0;
  real mx = real(-1);
  for_rows<true>(bx, [&](int64_t idx, int i, int j, int k) {
    (void)i; (void)j; (void)k;
    real p = f1[idx];
    real q = p / eps;
    if (p < -eps) vol += (double)dV;
    else if (p < eps) vol += (double)(real(0.5) * (real(1) - q - one_over_pi * sin(pi * q)) * dV);
    if (f2 && fabs(p) < zone) {
      real d = fabs(p - f2[idx]);
      if (d > mx) mx = d;
    }
  });
  __shared__ double s_sum[ROW_TY], s_max[ROW_TY];
  double ws = wave_sum(vol), wm = wave_max((double)mx);
  const int lane = threadIdx.x, wid = threadIdx.y;  // 64 x 4 blocks: one wave per y
  if (lane == 0) { s_sum[wid] = ws; s_max[wid] = wm; }
  __syncthreads();
  if (lane == 0 && wid == 0) {
    double a = 0.0, m = -1.0;
    for (int w = 0; w < ROW_TY; ++w) { a += s_sum[w]; m = fmax(m, s_max[w]); }
    part[blockIdx.x] = a;
    part[RED_BLOCKS + blockIdx.x] = m;
  }
}

__device__ __forceinline__ void finish_reduce_body(const double* part, int nb, double* out,
                                                   int mode) {
  // mode 0: out[0] = sum(part[0..nb)), out[1] = max(part[RB..RB+nb))
  // mode 1: out[0] = max(part[0..nb)), out[1] = max(part[RB..RB+nb))
  double a = mode == 0 ? 0.0 : -DBL_MAX, m = -DBL_MAX;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    a = mode == 0 ? a + part[b] : fmax(a, part[b]);
    m = fmax(m, part[RED_BLOCKS + b]);
  }
  a = mode == 0 ? wave_sum(a) : wave_max(a);
  m = wave_max(m);
  __shared__ double sa[4], sm[4];
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) { sa[wid] = a; sm[wid] = m; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double x = sa[0], y = sm[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
      x = mode == 0 ? x + sa[w] : fmax(x, sa[w]);
      y = fmax(y, sm[w]);
    }
    out[0] = x;
    out[1] = y;
  }
}
__global__ __launch_bounds__(256) void finish_reduce_kernel(const double* part, int nb, double* out,
                                                            int mode) {
  finish_reduce_body(part, nb, out, mode);
}

hipError_t launch_outer_reduce(const Box& bx, const real* f1, const real* f2, real zone, real eps,
                               real dV, double* part, double* out, hipStream_t st) {
  const int nb = row_blocks<true>(bx, RED_BLOCKS);
  hipLaunchKernelGGL(outer_reduce_kernel, dim3(nb), dim3(ROW_TX, ROW_TY), 0, st, bx, f1, f2, zone,
                     eps, dV, part);
  hipLaunchKernelGGL(finish_reduce_kernel, dim3(1), dim3(256), 0, st, part, nb, out, 0);
  return hipGetLastError();
}

// b_max and U_max of setVarCurvAdv3d (qss_spatial_derivatives3d.f:261-296)
__global__ __launch_bounds__(256) void coeff_reduce_kernel(Box bx, const real* b, const real* vx,
                                                           const real* vy, const real* vz,
                                                           real inv, double* part) {
  real bm = real(-1), um = real(-1);
  for_rows<true>(bx, [&](int64_t idx, int i, int j, int k) {
    (void)i; (void)j; (void)k;
    real bv = b[idx];
    if (bv > bm) bm = bv;
    real u = fabs(vx[idx]) * inv + fabs(vy[idx]) * inv + fabs(vz[idx]) * inv;
    if (u > um) um = u;
  });
  __shared__ double s0[ROW_TY], s1[ROW_TY];
  double a = wave_max((double)bm), m = wave_max((double)um);
  const int lane = threadIdx.x, wid = threadIdx.y;  // 64 x 4 blocks: one wave per y
  if (lane == 0) { s0[wid] = a; s1[wid] = m; }
  __syncthreads();
  if (lane == 0 && wid == 0) {
    double x = s0[0], y = s1[0];
    for (int w = 1; w < ROW_TY; ++w) { x = fmax(x, s0[w]); y = fmax(y, s1[w]); }
    part[blockIdx.x] = x;
    part[RED_BLOCKS + blockIdx.x] = y;
  }
}

hipError_t launch_coeff_reduce(const Box& bx, const real* b, const real* vx, const real* vy,
                               const real* vz, real inv, double* part, double* out,
                               hipStream_t st) {
  const int nb = row_blocks<true>(bx, RED_BLOCKS);
  hipLaunchKernelGGL(coeff_reduce_kernel, dim3(nb), dim3(ROW_TX, ROW_TY), 0, st, bx, b, vx, vy, vz,
                     inv, part);
  hipLaunchKernelGGL(finish_reduce_kernel, dim3(1), dim3(256), 0, st, part, nb, out, 1);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Subcell-fix reinitialization (reference:
// src/reinitialization/qss_reinitialization3d.f:476-746)

// Whether a face neighbour lies across zero from c, the reference's test
// of the six neighbours one by one ((a > 0 and c <= 0) or (a <= 0 and c > 0),
// :518-570): with c > 0 some neighbour is <= 0, i.e. their minimum is; with
// c <= 0 some neighbour is > 0, i.e. their maximum is; a NaN c or neighbour
// makes no crossing either way.  Two min3/max3 chains instead of twelve
// comparisons and their selects.
__device__ __forceinline__ bool crosses_zero(real c, real xp, real xm, real yp, real ym, real zp,
                                             real zm) {
  const real lo = fmin(fmin(fmin(xp, xm), fmin(yp, ym)), fmin(zp, zm));
  const real hi = fmax(fmax(fmax(xp, xm), fmax(yp, ym)), fmax(zp, zm));
  return (c > real(0)) ? (lo <= real(0)) : ((c <= real(0)) & (hi > real(0)));
}

// Quotient a / b: IEEE division, or in the fast single-precision mode a
// times the hardware reciprocal (1 ulp; b is never denormal here: delta >= 1e-5
// and max dx are the divisors)
template <bool FAST>
__device__ __forceinline__ real rdiv(real a, real b) {
  if constexpr (FAST && sizeof(real) == 4) return a * fast_rcp(b);
  else return a / b;
}
template <bool FAST>
__device__ __forceinline__ real rsqrt_(real x) {
  if constexpr (FAST && sizeof(real) == 4) return __builtin_amdgcn_sqrtf(x);
  else return sqrt(x);
}

// D0 of one voxel from phi0 and its six neighbours (:518-570): phi0 / delta
// next to the interface (a sign change to a face neighbour), else the flag
template <bool FAST>
__device__ __forceinline__ real reinit_d0(real c, real xp, real xm, real yp, real ym, real zp,
                                          real zm, real flag) {
  if (!crosses_zero(c, xp, xm, yp, ym, zp, zm)) return flag;
  real delta = tmax(tmax(tmax(fabs(xp - xm) / real(2), fabs(xp - c)), fabs(c - xm)), real(1.e-5));
  delta = tmax(tmax(tmax(fabs(yp - ym) / real(2), fabs(yp - c)), fabs(c - ym)), delta);
  delta = tmax(tmax(tmax(fabs(zp - zm) / real(2), fabs(zp - c)), fabs(c - zm)), delta);
  return rdiv<FAST>(c, delta);
}

// First-order subcell-fix right-hand side (:665-737): sgn(phi0)(1 - |grad phi|_G)
// with one-sided differences gp (forward), gm (backward), or, next to the
// interface (d0 != flag), d0 - sgn(phi0) |phi| / max dx
template <bool FAST>
__device__ __forceinline__ real reinit_rhs(real c, const real* gp, const real* gm, real p0,
                                           real d0, real flag, real max_dx) {
  real sgn0 = fabs(p0) > real(1.e-5) ? copysign(real(1), p0) : real(0);
  if (d0 != flag) return d0 - rdiv<FAST>(sgn0 * fabs(c), max_dx);
  // per axis the reference's max(max(-gp, 0), max(gm, 0)) for c > 0 and
  // max(max(gp, 0), max(-gm, 0)) otherwise, as max(max(-s gp, s gm), 0) with
  // s = +-1 (the same values, no branch on c)
  const real sc = c > real(0) ? real(1) : real(-1);
  real st[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) st[d] = tmax(tmax(-(sc * gp[d]), sc * gm[d]), real(0));
  if (!(fabs(c) > real(1.e-5))) return real(0);
  real n2 = st[0] * st[0] + st[1] * st[1] + st[2] * st[2];
  return sgn0 - sgn0 * rsqrt_<FAST>(n2);
}

// The kernels below take FAST (single precision with fast_rcp and the
// hardware square root, the model drivers' FP32 arithmetic) as a template
// flag; the exported kernel-level routines run them with FAST = false.
template <bool FAST>
__global__ void reinit_distance_kernel(Box bx, real* dist0, const real* phi0, real dx, real dy,
                                       real dz) {
  real flag = real(-1000.0) * tmax(tmax(dx, dy), dz);
  for_rows<true>(bx, [&](int64_t idx, int i, int j, int k) {
    (void)i; (void)j; (void)k;
    int64_t sy = bx.nx, sz = (int64_t)bx.nx * bx.ny;
    dist0[idx] = reinit_d0<FAST>(phi0[idx], phi0[idx + 1], phi0[idx - 1], phi0[idx + sy],
                                 phi0[idx - sy], phi0[idx + sz], phi0[idx - sz], flag);
  });
}
hipError_t launch_reinit_distance(const Box& bx, real* dist0, const real* phi0, real dx, real dy,
                                  real dz, hipStream_t st, bool fast) {
  const dim3 g(row_blocks<true>(bx, 16384)), b(ROW_TX, ROW_TY);
  if (fast) hipLaunchKernelGGL(reinit_distance_kernel<true>, g, b, 0, st, bx, dist0, phi0, dx, dy, dz);
  else hipLaunchKernelGGL(reinit_distance_kernel<false>, g, b, 0, st, bx, dist0, phi0, dx, dy, dz);
  return hipGetLastError();
}

// one forward-Euler step of the reinitialization equation (RK1 fused with
// the first-order subcell-fix right-hand side)
template <bool FAST>
__global__ void reinit_step_kernel(Box bx, real* next, const real* data, const real* phi0,
                                   const real* dist0, real dt_r, real dx, real dy, real dz) {
  real max_dx = tmax(tmax(dx, dy), dz);
  real flag = real(-1000.0) * max_dx;
  real inv_dx = real(1.0) / dx, inv_dy = real(1.0) / dy, inv_dz = real(1.0) / dz;
  int64_t sy = bx.nx, sz = (int64_t)bx.nx * bx.ny;
  for_rows<true>(bx, [&](int64_t idx, int i, int j, int k) {
    (void)i; (void)j; (void)k;
    real c = data[idx];
    real gp[3], gm[3];
    gp[0] = (data[idx + 1] - c) * inv_dx;
    gm[0] = (c - data[idx - 1]) * inv_dx;
    gp[1] = (data[idx + sy] - c) * inv_dy;
    gm[1] = (c - data[idx - sy]) * inv_dy;
    gp[2] = (data[idx + sz] - c) * inv_dz;
    gm[2] = (c - data[idx - sz]) * inv_dz;
    next[idx] = c + dt_r * reinit_rhs<FAST>(c, gp, gm, phi0[idx], dist0[idx], flag, max_dx);
  });
}
hipError_t launch_reinit_step(const Box& bx, real* next, const real* data, const real* phi0,
                              const real* dist0, real dt_r, real dx, real dy, real dz,
                              hipStream_t st, bool fast) {
  const dim3 g(row_blocks<true>(bx, 16384)), b(ROW_TX, ROW_TY);
  if (fast)
    hipLaunchKernelGGL(reinit_step_kernel<true>, g, b, 0, st, bx, next, data, phi0, dist0, dt_r,
                       dx, dy, dz);
  else
    hipLaunchKernelGGL(reinit_step_kernel<false>, g, b, 0, st, bx, next, data, phi0, dist0, dt_r,
                       dx, dy, dz);
  return hipGetLastError();
}

// the right-hand side alone (qss3dComputeReinitializationEqnRHSSubcellFixOrder1)
__global__ void reinit_rhs_kernel(Box bx, real* rhs, const real* data, const real* phi0,
                                  const real* dist0, real dx, real dy, real dz) {
  real max_dx = tmax(tmax(dx, dy), dz);
  real flag = real(-1000.0) * max_dx;
  real inv_dx = real(1.0) / dx, inv_dy = real(1.0) / dy, inv_dz = real(1.0) / dz;
  int64_t sy = bx.nx, sz = (int64_t)bx.nx * bx.ny;
  for_rows<true>(bx, [&](int64_t idx, int i, int j, int k) {
    (void)i; (void)j; (void)k;
    real c = data[idx];
    real gp[3], gm[3];
    gp[0] = (data[idx + 1] - c) * inv_dx;
    gm[0] = (c - data[idx - 1]) * inv_dx;
    gp[1] = (data[idx + sy] - c) * inv_dy;
    gm[1] = (c - data[idx - sy]) * inv_dy;
    gp[2] = (data[idx + sz] - c) * inv_dz;
    gm[2] = (c - data[idx - sz]) * inv_dz;
    rhs[idx] = reinit_rhs<false>(c, gp, gm, phi0[idx], dist0[idx], flag, max_dx);
  });
}
hipError_t launch_reinit_rhs(const Box& bx, real* rhs, const real* data, const real* phi0,
                             const real* dist0, real dx, real dy, real dz, hipStream_t st) {
  hipLaunchKernelGGL(reinit_rhs_kernel, dim3(row_blocks<true>(bx, 16384)), dim3(ROW_TX, ROW_TY), 0, st, bx, rhs,
                     data, phi0, dist0, dx, dy, dz);
  return hipGetLastError();
}

// The first step from phi0 itself (data = phi0): D0 and the step from the
// same seven values, with no copy of phi0 and no D0 field.  Bit for bit the
// distance kernel followed by the step kernel on that first iteration.
// the seven values of a voxel's stencil
struct Stencil7 {
  real c, xp, xm, yp, ym, zp, zm;
};
__device__ __forceinline__ Stencil7 stencil7(const real* __restrict__ f, int64_t idx, int64_t sy,
                                             int64_t sz) {
  return Stencil7{f[idx], f[idx + 1], f[idx - 1], f[idx + sy], f[idx - sy], f[idx + sz],
                  f[idx - sz]};
}
// one subcell-fix step of phi0 itself at one voxel (the first iteration)
template <bool FAST>
__device__ __forceinline__ real reinit_first_value(const Stencil7& v, real dt_r, real inv_dx,
                                                   real inv_dy, real inv_dz, real flag,
                                                   real max_dx) {
  const real c = v.c;
  real d0 = reinit_d0<FAST>(c, v.xp, v.xm, v.yp, v.ym, v.zp, v.zm, flag);
  real gp[3] = {(v.xp - c) * inv_dx, (v.yp - c) * inv_dy, (v.zp - c) * inv_dz};
  real gm[3] = {(c - v.xm) * inv_dx, (c - v.ym) * inv_dy, (c - v.zm) * inv_dz};
  return c + dt_r * reinit_rhs<FAST>(c, gp, gm, c, d0, flag, max_dx);
}
constexpr int REINIT_U = 4;  // voxels per lane and pass of the first-step kernels

template <bool FAST>
__device__ __forceinline__ void reinit_first_body(const Box& bx, real* __restrict__ next,
                                                  const real* __restrict__ phi0, real dt_r, real dx,
                                                  real dy, real dz, int lb, int lg) {
  real max_dx = tmax(tmax(dx, dy), dz);
  real flag = real(-1000.0) * max_dx;
  real inv_dx = real(1.0) / dx, inv_dy = real(1.0) / dy, inv_dz = real(1.0) / dz;
  int64_t sy = bx.nx, sz = (int64_t)bx.nx * bx.ny;
  Stencil7 v[REINIT_U];
  for_rows_u<true, REINIT_U>(
      bx, lb, lg, [&](int u, int64_t idx) { v[u] = stencil7(phi0, idx, sy, sz); },
      [&](int u, int64_t idx, int, int, int) {
        next[idx] = reinit_first_value<FAST>(v[u], dt_r, inv_dx, inv_dy, inv_dz, flag, max_dx);
      });
}
template <bool FAST>
__global__ void reinit_first_step_kernel(Box bx, real* __restrict__ next,
                                         const real* __restrict__ phi0, real dt_r, real dx,
                                         real dy, real dz) {
  reinit_first_body<FAST>(bx, next, phi0, dt_r, dx, dy, dz, blockIdx.x, gridDim.x);
}
// reinit_first_step_kernel fused with the outer reductions of its result:
// per-block FP64 partials of the smoothed volume of {next < 0} and of the
// local max-norm |next - prev| over |next| < zone, in the same order as
// outer_reduce_kernel (same launch shape), so the sums are the same bits.
template <bool FAST>
__device__ __forceinline__ void reinit_first_reduce_body(
    const Box& bx, real* __restrict__ next, const real* __restrict__ phi0, real dt_r, real dx,
    real dy, real dz, const real* __restrict__ prev, real zone, real eps, real dV, double* part,
    int lb, int lg) {
  const real pi = real(3.14159265358979323846), one_over_pi = real(0.31830988618379);
  real max_dx = tmax(tmax(dx, dy), dz);
  real flag = real(-1000.0) * max_dx;
  real inv_dx = real(1.0) / dx, inv_dy = real(1.0) / dy, inv_dz = real(1.0) / dz;
  int64_t sy = bx.nx, sz = (int64_t)bx.nx * bx.ny;
  double vol = 0.0;
  real mx = real(-1);
  Stencil7 v[REINIT_U];
  real pv[REINIT_U];
  // the voxels of a lane are visited in the order of for_rows (u runs along
  // the row), so the per-lane sums are added in the same order
  for_rows_u<true, REINIT_U>(
      bx, lb, lg,
      [&](int u, int64_t idx) {
        v[u] = stencil7(phi0, idx, sy, sz);
        pv[u] = prev[idx];
      },
      [&](int u, int64_t idx, int, int, int) {
        const real p = reinit_first_value<FAST>(v[u], dt_r, inv_dx, inv_dy, inv_dz, flag, max_dx);
        next[idx] = p;
        real q = p / eps;
        if (p < -eps) vol += (double)dV;
        else if (p < eps)
          vol += (double)(real(0.5) * (real(1) - q - one_over_pi * sin(pi * q)) * dV);
        if (fabs(p) < zone) {
          real d = fabs(p - pv[u]);
          if (d > mx) mx = d;
        }
      });
  __shared__ double s_sum[ROW_TY], s_max[ROW_TY];
  double ws = wave_sum(vol), wm = wave_max((double)mx);
  const int lane = threadIdx.x, wid = threadIdx.y;
  if (lane == 0) { s_sum[wid] = ws; s_max[wid] = wm; }
  __syncthreads();
  if (lane == 0 && wid == 0) {
    double a = 0.0, m = -1.0;
    for (int w = 0; w < ROW_TY; ++w) { a += s_sum[w]; m = fmax(m, s_max[w]); }
    part[lb] = a;
    part[RED_BLOCKS + lb] = m;
  }
}
template <bool FAST>
__global__ __launch_bounds__(256) void reinit_first_step_reduce_kernel(
    Box bx, real* __restrict__ next, const real* __restrict__ phi0, real dt_r, real dx, real dy,
    real dz, const real* __restrict__ prev, real zone, real eps, real dV, double* part) {
  reinit_first_reduce_body<FAST>(bx, next, phi0, dt_r, dx, dy, dz, prev, zone, eps, dV, part,
                                 blockIdx.x, gridDim.x);
}
hipError_t launch_reinit_first_step_reduce(const Box& bx, real* next, const real* phi0, real dt_r,
                                           real dx, real dy, real dz, const real* prev, real zone,
                                           real eps, real dV, double* part, double* out,
                                           hipStream_t st, bool fast) {
  const int nb = row_blocks<true>(bx, RED_BLOCKS);
  if (fast)
    hipLaunchKernelGGL(reinit_first_step_reduce_kernel<true>, dim3(nb), dim3(ROW_TX, ROW_TY), 0,
                       st, bx, next, phi0, dt_r, dx, dy, dz, prev, zone, eps, dV, part);
  else
    hipLaunchKernelGGL(reinit_first_step_reduce_kernel<false>, dim3(nb), dim3(ROW_TX, ROW_TY), 0,
                       st, bx, next, phi0, dt_r, dx, dy, dz, prev, zone, eps, dV, part);
  hipLaunchKernelGGL(finish_reduce_kernel, dim3(1), dim3(256), 0, st, part, nb, out, 0);
  return hipGetLastError();
}

// merge and imposition floor in one pass: phi = -mask_nw where mask_nw > 0
// (merge_kernel) and floor = max(mask - overlap, mask_w, mask_nw) (floor_kernel).
// Four voxels per thread and iteration (a stride apart), all their loads
// issued before any value is used.
__device__ __forceinline__ void merge_floor_body(real* phi, real* f, const real* m, const real* mw,
                                                 const real* mnw, real ov, int64_t n, int lb,
                                                 int lg) {
  constexpr int U = 4;
  const int64_t stride = (int64_t)lg * blockDim.x;
  for (int64_t l0 = (int64_t)lb * blockDim.x + threadIdx.x; l0 < n; l0 += U * stride) {
    real b[U], mm[U], a[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t l = min(l0 + u * stride, n - 1);
      b[u] = mnw[l];
      mm[u] = m[l];
      a[u] = mw[l];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t l = l0 + u * stride;
      if (l >= n) break;
      if (b[u] > real(0)) phi[l] = -b[u];
      real v = mm[u] - ov;
      v = (a[u] > v) ? a[u] : v;
      v = (b[u] > v) ? b[u] : v;
      f[l] = v;
    }
  }
}
__global__ void merge_floor_kernel(real* phi, real* f, const real* m, const real* mw,
                                   const real* mnw, real ov, int64_t n) {
  merge_floor_body(phi, f, m, mw, mnw, ov, n, blockIdx.x, gridDim.x);
}
hipError_t launch_merge_floor(real* phi, real* floor, const real* m, const real* mw,
                              const real* mnw, real overlap, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL(merge_floor_kernel, dim3(stream_blocks(n)), dim3(256), 0, st, phi, floor, m,
                     mw, mnw, overlap, n);
  return hipGetLastError();
}

hipError_t launch_reinit_first_step(const Box& bx, real* next, const real* phi0, real dt_r,
                                    real dx, real dy, real dz, hipStream_t st, bool fast) {
  const dim3 g(row_blocks<true>(bx, 16384)), b(ROW_TX, ROW_TY);
  if (fast)
    hipLaunchKernelGGL(reinit_first_step_kernel<true>, g, b, 0, st, bx, next, phi0, dt_r, dx, dy, dz);
  else
    hipLaunchKernelGGL(reinit_first_step_kernel<false>, g, b, 0, st, bx, next, phi0, dt_r, dx, dy,
                       dz);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Coefficient fields (reference: qss_spatial_derivatives3d.f:13-87,178-298,2987-3048)

__global__ void central_grad_kernel(Box bx, real* gx, real* gy, real* gz, const real* f, real fx,
                                    real fy, real fz) {
  int64_t sy = bx.nx, sz = (int64_t)bx.nx * bx.ny;
  for_rows<true>(bx, [&](int64_t idx, int i, int j, int k) {
    (void)i; (void)j; (void)k;
    gx[idx] = (f[idx + 1] - f[idx - 1]) * fx;
    gy[idx] = (f[idx + sy] - f[idx - sy]) * fy;
    gz[idx] = (f[idx + sz] - f[idx - sz]) * fz;
  });
}
hipError_t launch_central_grad(const Box& bx, real* gx, real* gy, real* gz, const real* f,
                               real dx, real dy, real dz, hipStream_t st) {
  hipLaunchKernelGGL(central_grad_kernel, dim3(row_blocks<true>(bx, 16384)), dim3(ROW_TX, ROW_TY), 0, st, bx, gx, gy,
                     gz, f, real(0.5) / dx, real(0.5) / dy, real(0.5) / dz);
  return hipGetLastError();
}


__global__ void var_norm_kernel(Box bx, real* vn, const real* mask, const real* gx, const real* gy,
                                const real* gz, real a0, real cos_theta, real dx) {
  for_rows<true>(bx, [&](int64_t idx, int i, int j, int k) {
    (void)i; (void)j; (void)k;
    real h1, h2, ag, s;
    coeff_common(mask[idx], gx[idx], gy[idx], gz[idx], dx, h1, h2, ag, s);
    vn[idx] = h1 * a0 - s * h2 * real(1) * cos_theta * ag;
  });
}
hipError_t launch_var_norm(const Box& bx, real* vn, const real* mask, const real* gx,
                           const real* gy, const real* gz, real a0, real cos_theta, real dx,
                           hipStream_t st) {
  hipLaunchKernelGGL(var_norm_kernel, dim3(row_blocks<true>(bx, 16384)), dim3(ROW_TX, ROW_TY), 0, st, bx, vn, mask,
                     gx, gy, gz, a0, cos_theta, dx);
  return hipGetLastError();
}

__global__ void var_curv_adv_kernel(Box bx, real* b, real* vx, real* vy, real* vz,
                                    const real* mask, const real* gx, const real* gy,
                                    const real* gz, real const_b, real dx) {
  for_rows<true>(bx, [&](int64_t idx, int i, int j, int k) {
    (void)i; (void)j; (void)k;
    real h1, h2, ag, s;
    real mxv = gx[idx], myv = gy[idx], mzv = gz[idx];
    coeff_common(mask[idx], mxv, myv, mzv, dx, h1, h2, ag, s);
    b[idx] = h1 * const_b;
    vx[idx] = s * real(1) * mxv * h2;
    vy[idx] = s * real(1) * myv * h2;
    vz[idx] = s * real(1) * mzv * h2;
  });
}
hipError_t launch_var_curv_adv(const Box& bx, real* b, real* vx, real* vy, real* vz,
                               const real* mask, const real* gx, const real* gy, const real* gz,
                               real const_b, real dx, hipStream_t st) {
  hipLaunchKernelGGL(var_curv_adv_kernel, dim3(row_blocks<true>(bx, 16384)), dim3(ROW_TX, ROW_TY), 0, st, bx, b, vx,
                     vy, vz, mask, gx, gy, gz, const_b, dx);
  return hipGetLastError();
}

// reference: src/utilities/qss_util3d.f:499-581
__device__ __forceinline__ void trap_velocity_body(const Box& bx, const real* phi_w, real* vn,
                                                   real* b, real* vx, real* vy, real* vz,
                                                   real eps_var, int lb, int lg) {
  for_rows_b<true>(bx, lb, lg, [&](int64_t idx, int i, int j, int k) {
    (void)i; (void)j; (void)k;
    if (phi_w[idx] > eps_var) {
      vn[idx] = real(0); b[idx] = real(0);
      vx[idx] = real(0); vy[idx] = real(0); vz[idx] = real(0);
    }
  });
}
__global__ void trap_velocity_kernel(Box bx, const real* phi_w, real* vn, real* b, real* vx,
                                     real* vy, real* vz, real eps_var) {
  trap_velocity_body(bx, phi_w, vn, b, vx, vy, vz, eps_var, blockIdx.x, gridDim.x);
}
hipError_t launch_trap_velocity(const Box& bx, const real* phi_w, real* vn, real* b, real* vx,
                                real* vy, real* vz, real dx, hipStream_t st) {
  hipLaunchKernelGGL(trap_velocity_kernel, dim3(row_blocks<true>(bx, 16384)), dim3(ROW_TX, ROW_TY), 0, st, bx, phi_w,
                     vn, b, vx, vy, vz, real(1.5) * dx);
  return hipGetLastError();
}

__global__ void collect_stale_kernel(Box bx, const real* vn, int64_t* list, int* count, int cap) {
  for_rows<true>(bx, [&](int64_t idx, int i, int j, int k) {
    (void)i; (void)j; (void)k;
    if (fabs(vn[idx]) < real(1.e-5)) {
      int pos = atomicAdd(count, 1);
      if (pos < cap) list[pos] = idx;
    }
  });
}
hipError_t launch_collect_stale(const Box& bx, const real* vn, int64_t* list, int* count, int cap,
                                hipStream_t st) {
  hipLaunchKernelGGL(collect_stale_kernel, dim3(row_blocks<true>(bx, 16384)), dim3(ROW_TX, ROW_TY), 0, st, bx, vn,
                     list, count, cap);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Analytic signed distances (reference: src/geometry/qss_initialization3d.c:34-166,440-522).
// The reference's C computes the sphere distance as sqrt() of the
// QSSLIB_REAL sum in double, minus the radius in double, then stores it.
struct Coords {
  real xlo[3], dx[3];
};

__global__ void sphere_kernel(real* phi, Box bx, Coords c, real cx, real cy, real cz, real r,
                              int inside) {
  int64_t n = bx.size();
  for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < n;
       l += (int64_t)gridDim.x * blockDim.x) {
    int64_t t = l / bx.nx;
    int i = (int)(l - t * bx.nx), j = (int)(t % bx.ny), k = (int)(t / bx.ny);
    real x = c.xlo[0] + c.dx[0] * i;
    real y = c.xlo[1] + c.dx[1] * j;
    real z = c.xlo[2] + c.dx[2] * k;
    real s = (x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz);
    real d = (real)(sqrt((double)s) - (double)r);
    if (inside >= 0) d = -d;
    real m = -FLT_MAX;
    if (d > m) m = d;
    phi[l] = m;
  }
}
hipError_t launch_sphere(real* phi, const Box& bx, const real* xlo, const real* dx, real cx,
                         real cy, real cz, real r, int inside, hipStream_t st) {
  Coords c;
  for (int a = 0; a < 3; ++a) { c.xlo[a] = xlo[a]; c.dx[a] = dx[a]; }
  hipLaunchKernelGGL(sphere_kernel, dim3(stream_blocks(bx.size())), dim3(256), 0, st, phi, bx, c,
                     cx, cy, cz, r, inside);
  return hipGetLastError();
}

__global__ void box_kernel(real* phi, Box bx, Coords c, Coords corner_side, int inside) {
  int64_t n = bx.size();
  for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < n;
       l += (int64_t)gridDim.x * blockDim.x) {
    int64_t t = l / bx.nx;
    int i = (int)(l - t * bx.nx), j = (int)(t % bx.ny), k = (int)(t / bx.ny);
    real p[3] = {c.xlo[0] + c.dx[0] * i, c.xlo[1] + c.dx[1] * j, c.xlo[2] + c.dx[2] * k};
    real m = -FLT_MAX;
    for (int l6 = 0; l6 < 6; ++l6) {
      int ax = l6 % 3;
      real nrm = l6 < 3 ? real(-1) : real(1);
      real pt = l6 < 3 ? corner_side.xlo[ax] : corner_side.xlo[ax] + corner_side.dx[ax];
      real nv[3] = {real(0), real(0), real(0)};
      nv[ax] = nrm;
      real pts[3] = {l6 < 3 ? corner_side.xlo[0] : corner_side.xlo[0] + corner_side.dx[0],
                     l6 < 3 ? corner_side.xlo[1] : corner_side.xlo[1] + corner_side.dx[1],
                     l6 < 3 ? corner_side.xlo[2] : corner_side.xlo[2] + corner_side.dx[2]};
      (void)pt;
      real dot = (p[0] - pts[0]) * nv[0] + (p[1] - pts[1]) * nv[1] + (p[2] - pts[2]) * nv[2];
      real nn = nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2];
      nn = (real)sqrt((double)nn);
      real d = dot / nn;
      if (d > m) m = d;
    }
    if (inside >= 1) m = -m;
    phi[l] = m;
  }
}
hipError_t launch_box(real* phi, const Box& bx, const real* xlo, const real* dx,
                      const real* corner, const real* side, int inside, hipStream_t st) {
  Coords c, cs;
  for (int a = 0; a < 3; ++a) {
    c.xlo[a] = xlo[a]; c.dx[a] = dx[a];
    cs.xlo[a] = corner[a]; cs.dx[a] = side[a];
  }
  hipLaunchKernelGGL(box_kernel, dim3(stream_blocks(bx.size())), dim3(256), 0, st, phi, bx, c, cs,
                     inside);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Reinitialization, merge and floor, trapped-phase velocities and reductions
// of many boxes at once (batched runs; Multi: block -> box and its block).

int reinit_pass_blocks(const Box& bx) { return row_blocks<true>(bx, 16384); }
int reduce_pass_blocks(const Box& bx) { return row_blocks<true>(bx, RED_BLOCKS); }
int stream_pass_blocks(int64_t n) { return stream_blocks(n); }

template <bool FAST>
__global__ void reinit_first_multi(const OuterJob* J, Multi m, int pass) {
  int lb, lg;
  const int r = multi_box(m, blockIdx.x, lb, lg);
  const OuterJob& j = J[r];
  reinit_first_body<FAST>(j.ta.bx, j.rout[pass], j.rin[pass], j.dt_r, j.dx, j.dy, j.dz, lb, lg);
}
template <bool FAST>
__global__ __launch_bounds__(256) void reinit_reduce_multi(const OuterJob* J, Multi m) {
  int lb, lg;
  const int r = multi_box(m, blockIdx.x, lb, lg);
  const OuterJob& j = J[r];
  reinit_first_reduce_body<FAST>(j.ta.bx, j.rout[2], j.rin[2], j.dt_r, j.dx, j.dy, j.dz, j.prev,
                                 j.zone, j.eps, j.dV, j.part, lb, lg);
}
// one block per box, over the box's reduce_pass_blocks partials
__global__ __launch_bounds__(256) void finish_reduce_multi(const OuterJob* J, Multi m) {
  int lb, lg;
  const int r = multi_box(m, blockIdx.x, lb, lg);
  finish_reduce_body(J[r].part, J[r].nred, J[r].red, 0);
}
__global__ void merge_floor_multi(const OuterJob* J, Multi m) {
  int lb, lg;
  const int r = multi_box(m, blockIdx.x, lb, lg);
  const OuterJob& j = J[r];
  merge_floor_body(j.ta.phi, j.floor, j.ta.mask, j.rout[0], j.rout[1], j.overlap, j.ta.bx.size(),
                   lb, lg);
}
__global__ void trap_velocity_multi(const OuterJob* J, Multi m) {
  int lb, lg;
  const int r = multi_box(m, blockIdx.x, lb, lg);
  const OuterJob& j = J[r];
  if (j.vn) trap_velocity_body(j.ta.bx, j.rout[1], j.vn, j.b, j.vx, j.vy, j.vz, j.eps_var, lb, lg);
}
__global__ void gather_red_kernel(const OuterJob* J, int n, double* out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) {
    out[2 * r] = J[r].red[0];
    out[2 * r + 1] = J[r].red[1];
  }
}

hipError_t launch_reinit_multi(const OuterJob* J, int pass, const Multi& rows, bool fast,
                               hipStream_t st) {
  if (fast)
    hipLaunchKernelGGL(reinit_first_multi<true>, dim3(rows.total), dim3(ROW_TX, ROW_TY), 0, st, J,
                       rows, pass);
  else
    hipLaunchKernelGGL(reinit_first_multi<false>, dim3(rows.total), dim3(ROW_TX, ROW_TY), 0, st, J,
                       rows, pass);
  return hipGetLastError();
}
hipError_t launch_reinit_reduce_multi(const OuterJob* J, const Multi& rows, const Multi& one,
                                      bool fast, hipStream_t st) {
  if (fast)
    hipLaunchKernelGGL(reinit_reduce_multi<true>, dim3(rows.total), dim3(ROW_TX, ROW_TY), 0, st, J,
                       rows);
  else
    hipLaunchKernelGGL(reinit_reduce_multi<false>, dim3(rows.total), dim3(ROW_TX, ROW_TY), 0, st,
                       J, rows);
  hipLaunchKernelGGL(finish_reduce_multi, dim3(one.total), dim3(256), 0, st, J, one);
  return hipGetLastError();
}
hipError_t launch_merge_floor_multi(const OuterJob* J, const Multi& stream, hipStream_t st) {
  hipLaunchKernelGGL(merge_floor_multi, dim3(stream.total), dim3(256), 0, st, J, stream);
  return hipGetLastError();
}
hipError_t launch_trap_velocity_multi(const OuterJob* J, const Multi& rows, hipStream_t st) {
  hipLaunchKernelGGL(trap_velocity_multi, dim3(rows.total), dim3(ROW_TX, ROW_TY), 0, st, J, rows);
  return hipGetLastError();
}
hipError_t launch_gather_red(const OuterJob* J, int n, double* out, hipStream_t st) {
  hipLaunchKernelGGL(gather_red_kernel, dim3((n + 63) / 64), dim3(64), 0, st, J, n, out);
  return hipGetLastError();
}

}  // namespace lsm
