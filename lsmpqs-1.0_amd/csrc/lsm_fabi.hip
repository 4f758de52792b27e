// lsm_fabi.hip -- the reference's kernel-level entry points (Fortran calling
// convention: every argument by reference, arrays over explicit ghost-box
// bounds with a fill-box subrange) on the GPU.  Each call copies its host
// arrays to the device, runs one kernel over the fill box and copies back
// what the routine writes.  Reference routines:
//   src/spatial_derivatives/qss_spatial_derivatives3d.f (getRHS3d*, the
//     coefficient routines, getCentralGrad3dOrder2, signedLinearExtrapolation3d),
//   src/time_integration/qss_tvd_runge_kutta3d.f,
//   src/utilities/qss_util3d.f (imposeMaskParallel, the volume and the local
//     max-norm),
//   src/reinitialization/qss_reinitialization3d.f (the subcell-fix pieces).
// These are for callers that drive the reference's own loops; the model
// entry points do not use them.  Several arrays of one call must share the
// same ghost-box bounds (the reference always passes one GB_DIMS).
#include <math.h>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>
#include "lsm_kernels.h"
#include "lsm_session.h"
#include "lsmpqs_hip.h"

namespace lsm {
namespace {

#define FCHECK(x)                                                                \
  do {                                                                           \
    hipError_t e_ = (x);                                                         \
    if (e_ != hipSuccess)                                                        \
      throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

// ghost-box bounds (6 ints) + fill box (6 ints) -> a Box indexed from the
// ghost box's lower corner
struct FBox {
  Box b;
  int glo[3];
  int64_t n() const { return b.size(); }
};

FBox fbox(const int* ilo_gb, const int* ihi_gb, const int* jlo_gb, const int* jhi_gb,
          const int* klo_gb, const int* khi_gb, const int* ilo_fb, const int* ihi_fb,
          const int* jlo_fb, const int* jhi_fb, const int* klo_fb, const int* khi_fb) {
  FBox f;
  f.glo[0] = *ilo_gb; f.glo[1] = *jlo_gb; f.glo[2] = *klo_gb;
  f.b.nx = *ihi_gb - *ilo_gb + 1;
  f.b.ny = *jhi_gb - *jlo_gb + 1;
  f.b.nz = *khi_gb - *klo_gb + 1;
  f.b.ilo = *ilo_fb - *ilo_gb; f.b.ihi = *ihi_fb - *ilo_gb;
  f.b.jlo = *jlo_fb - *jlo_gb; f.b.jhi = *jhi_fb - *jlo_gb;
  f.b.klo = *klo_fb - *klo_gb; f.b.khi = *khi_fb - *klo_gb;
  if (f.b.nx <= 0 || f.b.ny <= 0 || f.b.nz <= 0) throw std::runtime_error("empty ghost box");
  if ((int64_t)f.b.nx * f.b.ny * f.b.nz >= ((int64_t)1 << 31))
    throw std::runtime_error("ghost box exceeds 2^31 voxels");
  return f;
}

// the fill box with a halo of h voxels must lie inside the ghost box
void need_halo(const FBox& f, int h) {
  const Box& b = f.b;
  if (b.ilo > b.ihi || b.jlo > b.jhi || b.klo > b.khi) return;  // nothing to do
  if (b.ilo - h < 0 || b.jlo - h < 0 || b.klo - h < 0 || b.ihi + h >= b.nx ||
      b.jhi + h >= b.ny || b.khi + h >= b.nz)
    throw std::runtime_error("fill box plus stencil reaches outside the ghost box");
}

void same_gb(const FBox& a, const int* i0, const int* i1, const int* j0, const int* j1,
             const int* k0, const int* k1) {
  if (*i0 != a.glo[0] || *j0 != a.glo[1] || *k0 != a.glo[2] || *i1 - *i0 + 1 != a.b.nx ||
      *j1 - *j0 + 1 != a.b.ny || *k1 - *k0 + 1 != a.b.nz)
    throw std::runtime_error("arrays of one call must share the ghost-box bounds");
}

__host__ __device__ int64_t nfill(const Box& b) {
  if (b.ilo > b.ihi || b.jlo > b.jhi || b.klo > b.khi) return 0;
  return (int64_t)(b.ihi - b.ilo + 1) * (b.jhi - b.jlo + 1) * (b.khi - b.klo + 1);
}

int blocks(int64_t n) {
  int64_t g = (n + 255) / 256;
  return (int)(g < 1 ? 1 : (g > 16384 ? 16384 : g));
}

__device__ __forceinline__ void fill_ijk(const Box& b, int64_t f, int& i, int& j, int& k) {
  const int64_t w = b.ihi - b.ilo + 1, h = b.jhi - b.jlo + 1;
  const int64_t r = f / w;
  i = b.ilo + (int)(f - r * w);
  j = b.jlo + (int)(r % h);
  k = b.klo + (int)(r / h);
}

// device copies of host arrays; each is freed with the call
struct Dev {
  std::vector<void*> p;
  ~Dev() {
    for (void* q : p) (void)hipFree(q);
  }
  template <typename T>
  T* put(const T* h, int64_t n) {
    void* d = nullptr;
    FCHECK(hipMalloc(&d, (size_t)(n > 0 ? n : 1) * sizeof(T)));
    p.push_back(d);
    if (h) FCHECK(hipMemcpy(d, h, (size_t)n * sizeof(T), hipMemcpyHostToDevice));
    return (T*)d;
  }
};

template <typename T>
void get(T* h, const T* d, int64_t n) {
  FCHECK(hipDeviceSynchronize());
  FCHECK(hipMemcpy(h, d, (size_t)n * sizeof(T), hipMemcpyDeviceToHost));
}

// ---------------------------------------------------------------------------
// Right-hand sides (getRHS3d2ndConst :302-567, getRHS3d3rdConst :572-1011,
// getRHS3d2nd :1020-1326, getRHS3d3rd :1335-1816, getRHS3d5th :1827-2386).

struct RhsArgs {
  Box b;
  const real *phi, *vn, *bb, *vx, *vy, *vz;  // vn..vz NULL on the constant path
  real a, bconst;
  real* rhs;
  real dx, dy, dz;
  int qk;          // k offset of the ENO3 x-stencil read phi(i+3, j, j)
  real norm_sign;  // -1, or +1 for the WENO5 routine
  real* hmax;      // H/dx at the last fill voxel
};

template <int S>
__device__ void derivs_at(const RhsArgs& a, int i, int j, int k, real* d) {
  const Box& b = a.b;
  const int64_t idx = b.at(i, j, k), sy = b.nx, sz = (int64_t)b.nx * b.ny;
  real lx[7], ly[7], lz[7];
#pragma unroll
  for (int o = 0; o < 7; ++o) {
    const bool use = S != ENO2 || (o >= 1 && o <= 5);
    lx[o] = use ? a.phi[idx + (o - 3)] : real(0);
    ly[o] = use ? a.phi[idx + (o - 3) * sy] : real(0);
    lz[o] = use ? a.phi[idx + (o - 3) * sz] : real(0);
  }
  // the reference's ENO3 x stencil reads phi(i+3, j, j) for its last point
  if (S == ENO3) lx[6] = a.phi[b.at(i + 3, j, j + a.qk)];
  const real inv_dx = real(1.0) / a.dx, inv_dy = real(1.0) / a.dy, inv_dz = real(1.0) / a.dz;
  axis_derivs<real, S>(lx, inv_dx, inv_dx, d[0], d[1]);
  axis_derivs<real, S>(ly, inv_dy, inv_dx, d[2], d[3]);
  axis_derivs<real, S>(lz, inv_dz, inv_dx, d[4], d[5]);
}

template <int S>
__device__ __forceinline__ real norm_term(const RhsArgs& a, real vn, const real* d) {
  return a.norm_sign * vn * sqrt(godunov_sq(vn, d));
}

// The reference assigns temp_norm only where |vel_n| >= 1e-5; elsewhere the
// value of the last earlier voxel of the same call (loop order, i fastest)
// that assigned it carries over, or an unassigned value when none did
// (taken as 0).
template <int S, bool VAR>
__device__ real stale_norm(const RhsArgs& a, int i, int j, int k) {
  if (!VAR) return real(0);
  const Box& b = a.b;
  while (true) {
    if (i > b.ilo) --i;
    else if (j > b.jlo) { --j; i = b.ihi; }
    else if (k > b.klo) { --k; j = b.jhi; i = b.ihi; }
    else return real(0);
    const real v = a.vn[b.at(i, j, k)];
    if (fabs(v) >= real(1.e-5)) {
      real d[6];
      derivs_at<S>(a, i, j, k, d);
      return norm_term<S>(a, v, d);
    }
  }
}

template <int S, bool VAR>
__global__ void rhs3d_kernel(RhsArgs a) {
  const Box b = a.b;
  const int64_t n = nfill(b);
  const int64_t sy = b.nx, sz = (int64_t)b.nx * b.ny;
  const real fx = real(0.5) / a.dx, fy = real(0.5) / a.dy, fz = real(0.5) / a.dz;
  for (int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; f < n;
       f += (int64_t)gridDim.x * blockDim.x) {
    int i, j, k;
    fill_ijk(b, f, i, j, k);
    const int64_t idx = b.at(i, j, k);
    real d[6];
    derivs_at<S>(a, i, j, k, d);
    const real vn = VAR ? a.vn[idx] : a.a;
    real temp_norm;
    if (fabs(vn) >= real(1.e-5)) temp_norm = norm_term<S>(a, vn, d);
    else temp_norm = stale_norm<S, VAR>(a, i, j, k);
    const real* P = a.phi;
    real xy[4] = {P[idx + sy + 1], P[idx - sy - 1], P[idx + sy - 1], P[idx - sy + 1]};
    real xz[4] = {P[idx + sz + 1], P[idx - sz - 1], P[idx + sz - 1], P[idx - sz + 1]};
    real yz[4] = {P[idx + sz + sy], P[idx - sz - sy], P[idx + sz - sy], P[idx - sz + sy]};
    const real curv =
        curvature(P[idx], P[idx - 2], P[idx - 1], P[idx + 1], P[idx + 2], P[idx - 2 * sy],
                  P[idx - sy], P[idx + sy], P[idx + 2 * sy], P[idx - 2 * sz], P[idx - sz],
                  P[idx + sz], P[idx + 2 * sz], xy, xz, yz, fx, fy, fz);
    real rhs;
    if (VAR) {
      const real vxv = a.vx[idx], vyv = a.vy[idx], vzv = a.vz[idx];
      const real ax = vxv >= real(0) ? d[1] : d[0];
      const real ay = vyv >= real(0) ? d[3] : d[2];
      const real az = vzv >= real(0) ? d[5] : d[4];
      const real temp_vel = -(vxv * ax + vyv * ay + vzv * az);
      rhs = temp_norm + a.bb[idx] * curv + temp_vel;
    } else {
      rhs = temp_norm + a.bconst * curv;
    }
    a.rhs[idx] = rhs;
    if (i == b.ihi && j == b.jhi && k == b.khi)
      *a.hmax = h_over_dx(vn, d, a.dx, a.dy, a.dz);
  }
}

template <int S, bool VAR>
void rhs_call(const real* phi, const real* vn, const real* bb, const real* vx, const real* vy,
              const real* vz, real* max_H, real* lse_rhs, const FBox& f, real dx, real dy, real dz,
              real a0, real b0) {
  need_halo(f, S == ENO2 ? 2 : 3);
  const int64_t n = f.n();
  RhsArgs a{};
  a.b = f.b;
  a.dx = dx; a.dy = dy; a.dz = dz;
  a.norm_sign = (S == WENO5) ? real(1) : real(-1);
  a.qk = f.glo[1] - f.glo[2];
  if (S == ENO3 && nfill(f.b) > 0) {
    const int k0 = f.b.jlo + a.qk, k1 = f.b.jhi + a.qk;
    if (k0 < 0 || k1 >= f.b.nz)
      throw std::runtime_error(
          "ENO3 x stencil reads phi(i+3, j, j) outside the ghost box (the reference reads out "
          "of bounds for this box)");
  }
  Dev dv;
  a.phi = dv.put(phi, n);
  if (VAR) {
    a.vn = dv.put(vn, n);
    a.bb = dv.put(bb, n);
    a.vx = dv.put(vx, n);
    a.vy = dv.put(vy, n);
    a.vz = dv.put(vz, n);
  } else {
    a.a = a0;
    a.bconst = b0;
  }
  a.rhs = dv.put(lse_rhs, n);
  real h0 = -1;
  a.hmax = dv.put(&h0, 1);
  const int64_t nf = nfill(f.b);
  if (nf > 0) {
    hipLaunchKernelGGL((rhs3d_kernel<S, VAR>), dim3(blocks(nf)), dim3(256), 0, 0, a);
    FCHECK(hipGetLastError());
  }
  get(lse_rhs, a.rhs, n);
  get(max_H, a.hmax, 1);
}

// ---------------------------------------------------------------------------
// TVD Runge-Kutta pieces (qss_tvd_runge_kutta3d.f:2-56, 61-126, 131-200)

template <int STAGE>
__global__ void rk_kernel(Box b, real* out, const real* u1, const real* un, const real* rhs,
                          real dt) {
  const int64_t n = nfill(b);
  for (int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; f < n;
       f += (int64_t)gridDim.x * blockDim.x) {
    int i, j, k;
    fill_ijk(b, f, i, j, k);
    const int64_t idx = b.at(i, j, k);
    if (STAGE == 1) out[idx] = u1[idx] + dt * rhs[idx];
    else if (STAGE == 2) out[idx] = real(0.75) * un[idx] + real(0.25) * (u1[idx] + dt * rhs[idx]);
    else out[idx] = real(1.0 / 3.0) * un[idx] + real(2.0 / 3.0) * (u1[idx] + dt * rhs[idx]);
  }
}

template <int STAGE>
void rk_call(real* out, const real* u1, const real* un, const real* rhs, const FBox& f, real dt) {
  const int64_t n = f.n();
  Dev dv;
  real* d_out = dv.put(out, n);
  const real* d_u1 = dv.put(u1, n);
  const real* d_un = STAGE == 1 ? nullptr : dv.put(un, n);
  const real* d_rhs = dv.put(rhs, n);
  const int64_t nf = nfill(f.b);
  if (nf > 0) {
    hipLaunchKernelGGL((rk_kernel<STAGE>), dim3(blocks(nf)), dim3(256), 0, 0, f.b, d_out, d_u1,
                       d_un, d_rhs, dt);
    FCHECK(hipGetLastError());
  }
  get(out, d_out, n);
}

// ---------------------------------------------------------------------------
// One face of the signed linear extrapolation (qss_spatial_derivatives3d.f:
// 3055-3191): the ghost layers beyond the fill box on that face, over the
// whole ghost range of the other two axes.

__global__ void face_kernel(Box b, real* phi, int face) {
  const int ax = face >> 1, hi = face & 1;
  const int n0 = ax == 0 ? b.ny : b.nx, n1 = ax == 2 ? b.ny : b.nz;  // the other two axes
  const int lo_f = ax == 0 ? b.ilo : (ax == 1 ? b.jlo : b.klo);
  const int hi_f = ax == 0 ? b.ihi : (ax == 1 ? b.jhi : b.khi);
  const int nax = ax == 0 ? b.nx : (ax == 1 ? b.ny : b.nz);
  const int64_t cols = (int64_t)n0 * n1;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < cols;
       c += (int64_t)gridDim.x * blockDim.x) {
    const int u = (int)(c % n0), v = (int)(c / n0);
    auto at = [&](int w) -> int64_t {
      if (ax == 0) return b.at(w, u, v);
      if (ax == 1) return b.at(u, w, v);
      return b.at(u, v, w);
    };
    const int e = hi ? hi_f : lo_f, e2 = hi ? hi_f - 1 : lo_f + 1;
    const real pe = phi[at(e)];
    const real s = copysign(real(1), pe);
    const real slope = s * fabs(pe - phi[at(e2)]);
    if (hi) {
      for (int w = hi_f + 1; w < nax; ++w) phi[at(w)] = pe + slope * real(w - hi_f);
    } else {
      for (int w = 0; w < lo_f; ++w) phi[at(w)] = pe + slope * real(lo_f - w);
    }
  }
}

// ---------------------------------------------------------------------------
// imposeMaskParallel (qss_util3d.f:130-175): planes cur_klo..cur_khi, all of
// each plane's ghost box

__global__ void impose_planes_kernel(real* dst, const real* m, const real* src, real ov,
                                     int64_t off, int64_t n) {
  for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < n;
       l += (int64_t)gridDim.x * blockDim.x) {
    const real mm = m[off + l] - ov, v = src[off + l];
    dst[off + l] = mm > v ? mm : v;
  }
}

// ---------------------------------------------------------------------------
// Scratch of the allocation-free volume entry point: the per-block partial
// sums and maxima of the reduction (2 REDUCE_BLOCKS doubles) and its FP64
// result (2 doubles), 32.8 KB in all.  One copy per GPU, held until the
// process ends.  In host code `scratch` names a host-side stand-in, not the
// device array: host code uses only the address hipGetSymbolAddress returns
// for the current device (scratch_ptr()).
constexpr int SCRATCH_DOUBLES = 2 * REDUCE_BLOCKS + 4;
__device__ __align__(256) double scratch[SCRATCH_DOUBLES];

constexpr int MAX_DEVICES = 64;

// Device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept.
double* scratch_ptr() {
  static std::mutex mu;
  static double* ptr[MAX_DEVICES] = {};
  int dev = 0;
  FCHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) throw std::runtime_error("device ordinal out of range");
  std::lock_guard<std::mutex> lock(mu);
  if (!ptr[dev]) {
    void* p = nullptr;
    FCHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    ptr[dev] = (double*)p;
  }
  return ptr[dev];
}

// dst[0] = src[0] rounded to QSSLIB_REAL, or 0 for an empty fill box
__global__ void store_real_kernel(const double* src, real* dst, int empty) {
  if (threadIdx.x == 0 && blockIdx.x == 0) dst[0] = empty ? real(0) : (real)src[0];
}

}  // namespace
}  // namespace lsm

using namespace lsm;

#define GB6 const int *ilo_gb, const int *ihi_gb, const int *jlo_gb, const int *jhi_gb, \
            const int *klo_gb, const int *khi_gb
#define FB6 const int *ilo_fb, const int *ihi_fb, const int *jlo_fb, const int *jhi_fb, \
            const int *klo_fb, const int *khi_fb
#define FBOX() fbox(ilo_gb, ihi_gb, jlo_gb, jhi_gb, klo_gb, khi_gb, ilo_fb, ihi_fb, jlo_fb, \
                    jhi_fb, klo_fb, khi_fb)
#define FTRY \
  set_error("");  \
  try {
#define FCATCH                            \
  }                                       \
  catch (const std::exception& e) {       \
    set_error(e.what());                  \
  }

extern "C" {

void getrhs3d2ndconst_(const QSSLIB_REAL* phi, const QSSLIB_REAL* vel_n, const QSSLIB_REAL* b,
                       QSSLIB_REAL* max_H, QSSLIB_REAL* lse_rhs, GB6, FB6, const QSSLIB_REAL* dx,
                       const QSSLIB_REAL* dy, const QSSLIB_REAL* dz) {
  FTRY rhs_call<ENO2, false>(phi, nullptr, nullptr, nullptr, nullptr, nullptr, max_H, lse_rhs,
                             FBOX(), *dx, *dy, *dz, *vel_n, *b);
  FCATCH
}

void getrhs3d3rdconst_(const QSSLIB_REAL* phi, const QSSLIB_REAL* vel_n, const QSSLIB_REAL* b,
                       QSSLIB_REAL* max_H, QSSLIB_REAL* lse_rhs, GB6, FB6, const QSSLIB_REAL* dx,
                       const QSSLIB_REAL* dy, const QSSLIB_REAL* dz) {
  FTRY rhs_call<ENO3, false>(phi, nullptr, nullptr, nullptr, nullptr, nullptr, max_H, lse_rhs,
                             FBOX(), *dx, *dy, *dz, *vel_n, *b);
  FCATCH
}

#define VAR_RHS(NAME, SCHEME)                                                                  \
  void NAME(const QSSLIB_REAL* phi, const QSSLIB_REAL* vel_n, const QSSLIB_REAL* b,             \
            const QSSLIB_REAL* vx, const QSSLIB_REAL* vy, const QSSLIB_REAL* vz,                \
            QSSLIB_REAL* max_H, QSSLIB_REAL* lse_rhs, GB6, FB6, const QSSLIB_REAL* dx,          \
            const QSSLIB_REAL* dy, const QSSLIB_REAL* dz) {                                     \
    FTRY rhs_call<SCHEME, true>(phi, vel_n, b, vx, vy, vz, max_H, lse_rhs, FBOX(), *dx, *dy,   \
                                *dz, 0, 0);                                                     \
    FCATCH                                                                                      \
  }
VAR_RHS(getrhs3d2nd_, ENO2)
VAR_RHS(getrhs3d3rd_, ENO3)
VAR_RHS(getrhs3d5th_, WENO5)

#define GB6N(p) const int *p##i0, const int *p##i1, const int *p##j0, const int *p##j1, \
                const int *p##k0, const int *p##k1
#define SAME(p) same_gb(f, p##i0, p##i1, p##j0, p##j1, p##k0, p##k1)

void qss3drk1step_(QSSLIB_REAL* u_next, GB6, const QSSLIB_REAL* u_cur, GB6N(a),
                   const QSSLIB_REAL* rhs, GB6N(b), FB6, const QSSLIB_REAL* dt) {
  FTRY FBox f = FBOX();
  SAME(a);
  SAME(b);
  rk_call<1>(u_next, u_cur, nullptr, rhs, f, *dt);
  FCATCH
}

void qss3dtvdrk2stage2_(QSSLIB_REAL* u_stage2, GB6, const QSSLIB_REAL* u_stage1, GB6N(a),
                        const QSSLIB_REAL* u_cur, GB6N(b), const QSSLIB_REAL* rhs, GB6N(c), FB6,
                        const QSSLIB_REAL* dt) {
  FTRY FBox f = FBOX();
  SAME(a);
  SAME(b);
  SAME(c);
  rk_call<2>(u_stage2, u_stage1, u_cur, rhs, f, *dt);
  FCATCH
}

void qss3dtvdrk3stage3_(QSSLIB_REAL* u_next, GB6, const QSSLIB_REAL* u_stage2, GB6N(a),
                        const QSSLIB_REAL* u_cur, GB6N(b), const QSSLIB_REAL* rhs, GB6N(c), FB6,
                        const QSSLIB_REAL* dt) {
  FTRY FBox f = FBOX();
  SAME(a);
  SAME(b);
  SAME(c);
  rk_call<3>(u_next, u_stage2, u_cur, rhs, f, *dt);
  FCATCH
}

void signedlinearextrapolation3d_(QSSLIB_REAL* phi, GB6, FB6, const int* bdry_location_idx) {
  FTRY FBox f = FBOX();
  const int face = *bdry_location_idx;
  if (face < 0 || face > 5) return;  // the reference's routine does nothing for other indices
  need_halo(f, 0);
  const Box& b = f.b;
  if ((face >> 1) == 0 ? b.ihi <= b.ilo : ((face >> 1) == 1 ? b.jhi <= b.jlo : b.khi <= b.klo))
    throw std::runtime_error("extrapolation needs two fill voxels along the face normal");
  Dev dv;
  real* d = dv.put(phi, f.n());
  const int64_t cols = (face >> 1) == 0 ? (int64_t)b.ny * b.nz
                                        : ((face >> 1) == 1 ? (int64_t)b.nx * b.nz
                                                            : (int64_t)b.nx * b.ny);
  hipLaunchKernelGGL(face_kernel, dim3(blocks(cols)), dim3(256), 0, 0, b, d, face);
  FCHECK(hipGetLastError());
  get(phi, d, f.n());
  FCATCH
}

void imposemaskparallel_(QSSLIB_REAL* phi_masked, const QSSLIB_REAL* mask, const QSSLIB_REAL* phi,
                         const QSSLIB_REAL* overlap, GB6, const int* cur_klo, const int* cur_khi) {
  FTRY FBox f = fbox(ilo_gb, ihi_gb, jlo_gb, jhi_gb, klo_gb, khi_gb, ilo_gb, ihi_gb, jlo_gb,
                     jhi_gb, klo_gb, khi_gb);
  const int k0 = *cur_klo - *klo_gb, k1 = *cur_khi - *klo_gb;
  if (k1 < k0) return;
  if (k0 < 0 || k1 >= f.b.nz) throw std::runtime_error("plane range outside the ghost box");
  const int64_t plane = (int64_t)f.b.nx * f.b.ny, n = f.n();
  Dev dv;
  real* d_dst = dv.put(phi_masked, n);
  const real* d_m = dv.put(mask, n);
  const real* d_src = dv.put(phi, n);
  hipLaunchKernelGGL(impose_planes_kernel, dim3(blocks(plane * (k1 - k0 + 1))), dim3(256), 0, 0,
                     d_dst, d_m, d_src, *overlap, plane * k0, plane * (k1 - k0 + 1));
  FCHECK(hipGetLastError());
  get(phi_masked, d_dst, n);
  FCATCH
}

// The two reductions accumulate in FP64 in a fixed tree order; the
// reference accumulates in QSSLIB_REAL in loop order, so the FP64 results
// agree to rounding and the FP32 library is the more accurate.
void qss3dvolumeregionphilessthanzero_(QSSLIB_REAL* volume, const QSSLIB_REAL* phi, GB6, FB6,
                                       const QSSLIB_REAL* dx, const QSSLIB_REAL* dy,
                                       const QSSLIB_REAL* dz, const QSSLIB_REAL* eps) {
  FTRY FBox f = FBOX();
  need_halo(f, 0);
  Dev dv;
  const real* d = dv.put(phi, f.n());
  double* part = dv.put<double>(nullptr, reduce_partials());
  double* out = dv.put<double>(nullptr, 4);
  FCHECK(launch_outer_reduce(f.b, d, nullptr, 0, *eps, (*dx) * (*dy) * (*dz), part, out, 0));
  double h[2];
  get(h, out, 2);
  *volume = (QSSLIB_REAL)h[0];
  FCATCH
}

int lsmpqs_volume_noalloc(QSSLIB_REAL* volume, const QSSLIB_REAL* phi, const int* gb,
                          const int* fb, QSSLIB_REAL dx, QSSLIB_REAL dy, QSSLIB_REAL dz,
                          QSSLIB_REAL eps, void* stream) {
  FTRY
  if (!volume || !phi || !gb || !fb) throw std::runtime_error("lsmpqs_volume_noalloc: NULL argument");
  FBox f = fbox(gb + 0, gb + 1, gb + 2, gb + 3, gb + 4, gb + 5, fb + 0, fb + 1, fb + 2, fb + 3,
                fb + 4, fb + 5);
  need_halo(f, 0);
  if (reduce_partials() + 2 > SCRATCH_DOUBLES)
    throw std::runtime_error("lsmpqs_volume_noalloc: reduction needs more than its " +
                             std::to_string(SCRATCH_DOUBLES * sizeof(double)) + "-byte scratch");
  double* part = scratch_ptr();
  double* out = part + reduce_partials();
  hipStream_t st = (hipStream_t)stream;
  const int empty = nfill(f.b) == 0;
  if (!empty) FCHECK(launch_outer_reduce(f.b, phi, nullptr, 0, eps, dx * dy * dz, part, out, st));
  hipLaunchKernelGGL(store_real_kernel, dim3(1), dim3(64), 0, st, out, volume, empty);
  FCHECK(hipGetLastError());
  return 0;
  }
  catch (const std::exception& e) {
    set_error(e.what());
    return -2;
  }
}

void qss3dmaxnormdifflocal_(QSSLIB_REAL* max_norm, const QSSLIB_REAL* field1, GB6,
                            const QSSLIB_REAL* field2, GB6N(a), FB6, const QSSLIB_REAL* zone) {
  FTRY FBox f = FBOX();
  SAME(a);
  need_halo(f, 0);
  Dev dv;
  const real* d1 = dv.put(field1, f.n());
  const real* d2 = dv.put(field2, f.n());
  double* part = dv.put<double>(nullptr, reduce_partials());
  double* out = dv.put<double>(nullptr, 4);
  FCHECK(launch_outer_reduce(f.b, d1, d2, *zone, real(1), real(1), part, out, 0));
  double h[2];
  get(h, out, 2);
  *max_norm = (QSSLIB_REAL)h[1];
  FCATCH
}

void getcentralgrad3dorder2_(QSSLIB_REAL* gx, QSSLIB_REAL* gy, QSSLIB_REAL* gz, GB6,
                             const QSSLIB_REAL* phi, GB6N(a), FB6, const QSSLIB_REAL* dx,
                             const QSSLIB_REAL* dy, const QSSLIB_REAL* dz) {
  FTRY FBox f = FBOX();
  SAME(a);
  need_halo(f, 1);
  const int64_t n = f.n();
  Dev dv;
  real* dgx = dv.put(gx, n);
  real* dgy = dv.put(gy, n);
  real* dgz = dv.put(gz, n);
  const real* dp = dv.put(phi, n);
  FCHECK(launch_central_grad(f.b, dgx, dgy, dgz, dp, *dx, *dy, *dz, 0));
  get(gx, dgx, n);
  get(gy, dgy, n);
  get(gz, dgz, n);
  FCATCH
}

void setvarnorm3d_(const QSSLIB_REAL* mask, const QSSLIB_REAL* mx, const QSSLIB_REAL* my,
                   const QSSLIB_REAL* mz, QSSLIB_REAL* var_a, const QSSLIB_REAL* a0,
                   const QSSLIB_REAL* ca, GB6, FB6, const QSSLIB_REAL* dx) {
  FTRY FBox f = FBOX();
  need_halo(f, 0);
  const int64_t n = f.n();
  Dev dv;
  const real* dm = dv.put(mask, n);
  const real* dx_ = dv.put(mx, n);
  const real* dy_ = dv.put(my, n);
  const real* dz_ = dv.put(mz, n);
  real* dva = dv.put(var_a, n);
  FCHECK(launch_var_norm(f.b, dva, dm, dx_, dy_, dz_, *a0, (real)cos(*ca), *dx, 0));
  get(var_a, dva, n);
  FCATCH
}

void setvarcurvadv3d_(const QSSLIB_REAL* mask, const QSSLIB_REAL* mx, const QSSLIB_REAL* my,
                      const QSSLIB_REAL* mz, QSSLIB_REAL* var_b, QSSLIB_REAL* vx, QSSLIB_REAL* vy,
                      QSSLIB_REAL* vz, QSSLIB_REAL* b_max_over_dx, QSSLIB_REAL* max_U_over_dx,
                      const QSSLIB_REAL* const_b, GB6, FB6, const QSSLIB_REAL* dx) {
  FTRY FBox f = FBOX();
  need_halo(f, 0);
  const int64_t n = f.n();
  Dev dv;
  const real* dm = dv.put(mask, n);
  const real* dx_ = dv.put(mx, n);
  const real* dy_ = dv.put(my, n);
  const real* dz_ = dv.put(mz, n);
  real* db = dv.put(var_b, n);
  real* dvx = dv.put(vx, n);
  real* dvy = dv.put(vy, n);
  real* dvz = dv.put(vz, n);
  FCHECK(launch_var_curv_adv(f.b, db, dvx, dvy, dvz, dm, dx_, dy_, dz_, *const_b, *dx, 0));
  const real inv = real(1.0) / *dx;
  double* part = dv.put<double>(nullptr, reduce_partials());
  double* out = dv.put<double>(nullptr, 4);
  FCHECK(launch_coeff_reduce(f.b, db, dvx, dvy, dvz, inv, part, out, 0));
  double h[2];
  get(h, out, 2);
  get(var_b, db, n);
  get(vx, dvx, n);
  get(vy, dvy, n);
  get(vz, dvz, n);
  const real bmax = (real)h[0];
  *max_U_over_dx = (real)h[1];
  *b_max_over_dx = 2 * bmax * (inv * inv + inv * inv + inv * inv);
  FCATCH
}

void qss3dcomputedistanceforsubcellfix_(QSSLIB_REAL* distance0, const QSSLIB_REAL* phi0, GB6,
                                        FB6, const QSSLIB_REAL* dx, const QSSLIB_REAL* dy,
                                        const QSSLIB_REAL* dz) {
  FTRY FBox f = FBOX();
  need_halo(f, 1);
  const int64_t n = f.n();
  Dev dv;
  real* dd = dv.put(distance0, n);
  const real* dp = dv.put(phi0, n);
  FCHECK(launch_reinit_distance(f.b, dd, dp, *dx, *dy, *dz, 0));
  get(distance0, dd, n);
  FCATCH
}

void qss3dcomputereinitializationeqnrhssubcellfixorder1_(
    QSSLIB_REAL* reinit_rhs, const QSSLIB_REAL* phi, GB6, const QSSLIB_REAL* phi0,
    const QSSLIB_REAL* distance0, GB6N(a), const QSSLIB_REAL* dx, const QSSLIB_REAL* dy,
    const QSSLIB_REAL* dz) {
  // the second bounds group is the fill box, as in the reference's call
  // (reinitialize_top.c:341-344)
  FTRY FBox f = fbox(ilo_gb, ihi_gb, jlo_gb, jhi_gb, klo_gb, khi_gb, ai0, ai1, aj0, aj1, ak0,
                     ak1);
  need_halo(f, 1);
  const int64_t n = f.n();
  Dev dv;
  real* dr = dv.put(reinit_rhs, n);
  const real* dp = dv.put(phi, n);
  const real* dp0 = dv.put(phi0, n);
  const real* dd = dv.put(distance0, n);
  FCHECK(launch_reinit_rhs(f.b, dr, dp, dp0, dd, *dx, *dy, *dz, 0));
  get(reinit_rhs, dr, n);
  FCATCH
}

}  // extern "C"
