// lsm_stage.hip -- the hot kernels of one level-set time step on gfx950:
// the fused right-hand side + TVD Runge-Kutta stage, the time-step probe and
// the composed signed-linear-extrapolation boundary condition.
//
// Stage kernel layout: a 64 x 8 thread block owns a 64 x 8 column of fill
// voxels and marches it along z.  The z line of every voxel (k-3 .. k+3)
// lives in registers, one plane ahead prefetched.  Four z-planes of the
// input field, each with a 3-voxel halo in x and y, rotate through LDS: the
// tile interior of plane k+1 is written from the register line, only the
// 468-voxel halo ring is read from memory (each thread owns one fixed halo
// slot), and with four slots one barrier per plane separates writing plane
// k+1 from reading planes k-1..k+1.  The right-hand side never touches HBM:
// it is combined with u^n and dt into the stage result in registers, and the
// last stage also applies the mask imposition, so a full step writes only
// the stage buffers and phi.
//
// Single-precision WENO5 fast path: every x, y and z face of the tile is
// reconstructed once (both one-sided values of a face share the smoothness
// indicators).  Each thread takes the x and y faces on its low side and
// keeps their phi^- values; the phi^+ values go through LDS to the voxel on
// the other side.  Threads 0..71 also take the tile's far y row (64 faces)
// and far x column (8 faces).  z faces are carried in registers from one
// plane to the next.
#include <algorithm>
#include <type_traits>
#include "lsm_kernels.h"

namespace lsm {

constexpr int TX = 64;
constexpr int TY = 8;
constexpr int PX = TX + 6;
constexpr int PY = TY + 6;
constexpr int NSLOT = 4;
constexpr int NHALO = PX * PY - TX * TY;  // 468
static_assert(NHALO <= TX * TY, "one halo voxel per thread");

// phi(i+3,j,j): the reference's ENO3 x line uses the k index j
// (qss_spatial_derivatives3d.f:677,1454); outside the array it reads 0.
__device__ __forceinline__ real eno3_quirk_read(const real* f, const Box& bx, int i, int j) {
  int64_t a = (int64_t)(i + 3) + (int64_t)bx.nx * ((int64_t)j + (int64_t)bx.ny * j);
  return (a >= 0 && a < bx.size()) ? f[a] : real(0);
}

// The six one-sided derivatives at (i,j,k) read straight from global memory
// (used off the fast path: the time-step probe and stale normal terms).
template <int S>
__device__ void derivs_global(const real* f, const Box& bx, int i, int j, int k, real dx, real dy,
                              real dz, int quirk, real* d) {
  real lx[7], ly[7], lz[7];
#pragma unroll
  for (int o = -3; o <= 3; ++o) {
    bool need = !(S == ENO2 && (o == -3 || o == 3));
    lx[o + 3] = need ? f[bx.at(i + o, j, k)] : real(0);
    ly[o + 3] = need ? f[bx.at(i, j + o, k)] : real(0);
    lz[o + 3] = need ? f[bx.at(i, j, k + o)] : real(0);
  }
  if (S == ENO3 && quirk) lx[6] = eno3_quirk_read(f, bx, i, j);
  real inv_dx = real(1.0) / dx, inv_dy = real(1.0) / dy, inv_dz = real(1.0) / dz;
  axis_derivs<real, S>(lx, inv_dx, inv_dx, d[0], d[1]);
  axis_derivs<real, S>(ly, inv_dy, inv_dx, d[2], d[3]);
  axis_derivs<real, S>(lz, inv_dz, inv_dx, d[4], d[5]);
}

// Normal term a voxel with |vel_n| < 1e-5 inherits: the reference leaves
// temp_norm unassigned there and so reuses the value of the last voxel of the
// same slab, in loop order, that did assign it
// (qss_spatial_derivatives3d.f:459-486).
template <int S>
__device__ real stale_norm(const StageArgs& a, int64_t idx) {
  int lo = 0, hi = a.nstale - 1;
  while (lo <= hi) {
    int mid = (lo + hi) >> 1;
    int64_t v = a.stale_idx[mid];
    if (v == idx) {
      int64_t src = a.stale_src[mid];
      if (src < 0) return real(0);
      int k = (int)(src / ((int64_t)a.bx.nx * a.bx.ny));
      int64_t r = src - (int64_t)k * a.bx.nx * a.bx.ny;
      int j = (int)(r / a.bx.nx), i = (int)(r - (int64_t)j * a.bx.nx);
      real d[6];
      derivs_global<S>(a.in, a.bx, i, j, k, a.dx, a.dy, a.dz, a.eno3_quirk, d);
      real vs = a.vn[src];
      return a.norm_sign * vs * sqrt(godunov_sq(vs, d));
    }
    if (v < idx) lo = mid + 1; else hi = mid - 1;
  }
  return real(0);
}

// LDS position (row * PX + column) of the halo voxel thread `tid` stages,
// and that voxel's offset within a z-plane; lds = -1 when the thread stages
// none, off = -1 when the voxel lies outside the ghost box (staged as 0).
struct HaloSlot {
  int lds, off;
};
__device__ __forceinline__ HaloSlot halo_slot(const Box& bx, int i0, int j0, int tid) {
  HaloSlot h{-1, -1};
  if (tid >= NHALO) return h;
  int r, c;
  if (tid < 3 * PX) {  // rows 0..2
    r = tid / PX;
    c = tid - r * PX;
  } else if (tid < 6 * PX) {  // rows TY+3 .. TY+5
    int e = tid - 3 * PX;
    r = TY + 3 + e / PX;
    c = e - (e / PX) * PX;
  } else {  // rows 3 .. TY+2, columns 0..2 and TX+3..TX+5
    int e = tid - 6 * PX;
    r = 3 + e / 6;
    int q = e - (e / 6) * 6;
    c = q < 3 ? q : TX + q;
  }
  h.lds = r * PX + c;
  int gi = i0 - 3 + c, gj = j0 - 3 + r;
  if (gi >= 0 && gi < bx.nx && gj >= 0 && gj < bx.ny) h.off = gi + bx.nx * gj;
  return h;
}

// Pointwise inputs of one voxel: coefficients (variable path), u^n (stages
// 2 and 3), imposition floor (last stage).
template <typename T>
struct PointIn {
  T vn = 0, b = 0, vx = 0, vy = 0, vz = 0, un = 0, floor = 0;
  template <bool VAR, int STAGE, bool FINAL>
  __device__ __forceinline__ void load(const StageArgs& a, int64_t idx) {
    if (VAR) { vn = a.vn[idx]; b = a.b[idx]; vx = a.vx[idx]; vy = a.vy[idx]; vz = a.vz[idx]; }
    if (STAGE > 1) un = a.un[idx];
    if (FINAL) floor = a.floor[idx];
  }
};

// FAST (single precision only): WENO5 faces with one reciprocal per
// reconstruction shared between neighbours, FMA contraction in the
// curvature, hardware square root.  Otherwise every expression follows the
// reference's order.
// One brick: the 64 x 8 column `tile` (x fastest, then y, then z chunk) of
// the decomposition in a.tiles_*.  Called by the whole block; returns at the
// same point for every thread.
template <int S, bool VAR, int STAGE, bool FINAL, bool FAST>
__device__ __forceinline__ void stage_tile(const StageArgs& a, int tile) {
  constexpr bool FW = FAST && S == WENO5;
  constexpr bool FE = FAST && S != WENO5;  // fast ENO2 / ENO3 (eno*_fast_raw)
  __shared__ real pl[NSLOT][PY * PX];
  __shared__ real xp[FW ? 2 : 1][FW ? TY : 1][TX + 1];  // phi_x^+ by face, double-buffered
  __shared__ real yp[FW ? 2 : 1][FW ? TY + 1 : 1][TX];  // phi_y^+ by face
  const Box bx = a.bx;
  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * TX + tx;
  const int tcol = tile % a.tiles_x, trow = (tile / a.tiles_x) % a.tiles_y;
  const int tzc = tile / (a.tiles_x * a.tiles_y);
  const int i0 = bx.ilo + tcol * TX, j0 = bx.jlo + trow * TY;
  const int i = i0 + tx, j = j0 + ty;
  const int zc = (bx.khi - bx.klo + a.tiles_z) / a.tiles_z;
  const int k0 = bx.klo + tzc * zc;
  const int k1 = min(k0 + zc - 1, bx.khi);
  if (k0 > k1) return;
  const bool active = (i <= bx.ihi) && (j <= bx.jhi);
  const bool inbox = (i < bx.nx) && (j < bx.ny);
  const int64_t plane = (int64_t)bx.nx * bx.ny;
  const real* __restrict__ in = a.in;
  const HaloSlot hs = halo_slot(bx, i0, j0, tid);
  const int me = (ty + 3) * PX + tx + 3;  // this thread's voxel in an LDS plane

  const real inv_dx = real(1.0) / a.dx, inv_dy = real(1.0) / a.dy, inv_dz = real(1.0) / a.dz;
  const real fx = real(0.5) / a.dx, fy = real(0.5) / a.dy, fz = real(0.5) / a.dz;
  // cubic voxels: the fast curvature factors out f^2 = (0.5/dx)^2
  const bool iso = a.dx == a.dy && a.dy == a.dz;
  const real f2 = fx * fx, curv_thr = real(1.e-5) / f2;
  const FaceScale<real> sx = face_scale(a.dx), sy = face_scale(a.dy), sz = face_scale(a.dz);
  const real dt = *a.dt;

  // z line k-3..k+3 of this thread's column; wnext = plane k+4 in flight
  const real* col = in + (inbox ? (int64_t)i + (int64_t)bx.nx * j : 0);
  real w[7];
#pragma unroll
  for (int o = 0; o < 7; ++o) w[o] = real(0);
  real wnext = real(0);
  if (inbox) {
#pragma unroll
    for (int o = 0; o < 6; ++o) w[o] = col[(int64_t)(k0 - 3 + o) * plane];
    wnext = col[(int64_t)(k0 + 3) * plane];
  }
  // halo voxel of plane k+1 in flight; a halo voxel outside the ghost box
  // (partial tiles) is read by no active voxel's stencil, so its thread
  // reads voxel 0's column instead
  const real* hp = in + (hs.off >= 0 ? hs.off : 0);
  auto halo = [&](int k) { return hs.off >= 0 ? hp[(int64_t)k * plane] : real(0); };
  if (hs.lds >= 0) {
    pl[0][hs.lds] = halo(k0 - 1);
    pl[1][hs.lds] = halo(k0);
  }
  pl[0][me] = w[2];
  pl[1][me] = w[3];
  real hnext = hs.lds >= 0 ? halo(k0 + 1) : real(0);

  real zcarry = real(0);  // phi_z^-(k) from the face k-1/2 (FW only)
  if constexpr (FW) zcarry = weno5_face_raw(&w[0], sz).minus_right;
  __syncthreads();

  // Plane k+1 into the LDS ring (read after this plane's barrier), the z
  // line's plane k+3, and the next loads in flight: plane k+4 of the z line
  // and the halo voxel of plane k+2.  Stage 1 does this at the start of the
  // plane with its loads guarded.  Stages 2 and 3, which also load u^n (and
  // the floor), do it after the x and y faces with every load unguarded
  // (inactive lanes of partial tiles read clamped, in-box addresses and
  // discard the values; plane k+4 is clamped to the box on the last planes,
  // where it is not used): the compiler then neither folds the u^n product
  // into the load block, which made each plane wait for its loads at once,
  // nor drains the whole load queue, the previous plane's store included, at
  // the top of each plane (512^3 FP32: stages 2 and 3 1.45 / 1.54 ms ->
  // 1.13 / 1.28 ms; stage 1 is slower in that form).
  constexpr bool EARLY_RING = STAGE == 1;
  auto ring_next = [&](int k, int sp) {
    w[6] = wnext;
    wnext = col[(int64_t)min(k + 4, bx.nz - 1) * plane];
    pl[sp][me] = w[4];
    if (hs.lds >= 0) pl[sp][hs.lds] = hnext;
    hnext = hp[(int64_t)(k + 2) * plane];  // with hs.off < 0 only inactive voxels read it
  };
  int64_t idx = bx.at(i, j, k0);
  int64_t pidx = bx.at(min(i, bx.ihi), min(j, bx.jhi), k0);
  for (int k = k0; k <= k1; ++k, idx += plane, pidx += plane) {
    const int sm = (k - k0) & 3, sc = (k - k0 + 1) & 3, sp = (k - k0 + 2) & 3;
    const int buf = (k - k0) & 1;
    // ---- phase A: start this voxel's loads, x/y faces of plane k, stage
    // plane k+1, z face of plane k
    if constexpr (EARLY_RING) {
      w[6] = wnext;
      if (inbox && k + 4 < bx.nz) wnext = col[(int64_t)(k + 4) * plane];
      pl[sp][me] = w[4];
      if (hs.lds >= 0) {
        pl[sp][hs.lds] = hnext;
        hnext = halo(k + 2);
      }
    }
    PointIn<real> cur;  // in flight during the face computations
    if (!EARLY_RING || active) cur.template load<VAR, STAGE, FINAL>(a, EARLY_RING ? idx : pidx);
    real d[6];
    real qx[6], qy[6];
    if constexpr (FW) {
      const real* P = pl[sc];
#pragma unroll
      for (int o = 0; o < 6; ++o) {
        qx[o] = P[(ty + 3) * PX + tx + o];   // phi(i-3+o, j)
        qy[o] = P[(ty + o) * PX + tx + 3];   // phi(i, j-3+o)
      }
      // faces times dx; phase B applies sign(vel_n)/dx
      Weno5Face<real> f = weno5_face_raw(qx, sx);
      d[1] = f.minus_right;
      xp[buf][ty][tx] = f.plus_left;
      f = weno5_face_raw(qy, sy);
      d[3] = f.minus_right;
      yp[buf][ty][tx] = f.plus_left;
      if (tid < TX + TY) {  // far y row and far x column of the tile
        const bool yrow = tid < TX;
        const real* q0 = yrow ? P + TY * PX + tid + 3 : P + (tid - TX + 3) * PX + TX;
        const int st = yrow ? PX : 1;
        real q[6];
#pragma unroll
        for (int o = 0; o < 6; ++o) q[o] = q0[o * st];
        f = weno5_face_raw(q, yrow ? sy : sx);
        real* dst = yrow ? &yp[buf][TY][tid] : &xp[buf][tid - TX][TX];
        *dst = f.plus_left;
      }
    }
    if constexpr (!EARLY_RING) ring_next(k, sp);
    if constexpr (FW) {
      Weno5Face<real> f = weno5_face_raw(&w[1], sz);
      d[4] = f.plus_left;
      d[5] = zcarry;
      zcarry = f.minus_right;
    }
    __syncthreads();
    // ---- phase B: right-hand side and stage update of voxel (i, j, k)
    if (active) {
      const real vn = VAR ? cur.vn : a.a_const, bv = VAR ? cur.b : a.b_const;
      const real vxv = cur.vx, vyv = cur.vy, vzv = cur.vz, unv = cur.un, flo = cur.floor;
      const real* P = pl[sc];
      real lx[7], ly[7];
      if constexpr (FW) {
        d[0] = xp[buf][ty][tx + 1];
        d[2] = yp[buf][ty + 1][tx];
#pragma unroll
        for (int o = 0; o < 6; ++o) { lx[o] = qx[o]; ly[o] = qy[o]; }
        lx[6] = ly[6] = real(0);  // not used by the curvature
      } else {
#pragma unroll
        for (int o = 0; o < 7; ++o) {
          lx[o] = P[me - 3 + o];
          ly[o] = P[me + (o - 3) * PX];
        }
        if (S == ENO3 && a.eno3_quirk) lx[6] = eno3_quirk_read(in, bx, i, j);
        if constexpr (FE) {  // unscaled; sign(vel_n)/dx applied below
          if constexpr (S == ENO2) {
            eno2_fast_raw(lx, d[0], d[1]);
            eno2_fast_raw(ly, d[2], d[3]);
            eno2_fast_raw(w, d[4], d[5]);
          } else {
            eno3_fast_raw(lx, d[0], d[1]);
            eno3_fast_raw(ly, d[2], d[3]);
            eno3_fast_raw(w, d[4], d[5]);
          }
        } else {
          axis_derivs<real, S>(lx, inv_dx, inv_dx, d[0], d[1]);
          axis_derivs<real, S>(ly, inv_dy, inv_dx, d[2], d[3]);
          axis_derivs<real, S>(w, inv_dz, inv_dx, d[4], d[5]);
        }
      }

      real adv_sign = real(-1);
      if constexpr (FW || FE) {
        // one-sided derivatives times sign(vel_n)/dx, so that the Godunov
        // selection needs no per-axis sign products; the reference scales
        // the ENO3 y- and z- minus sides by 1/dx (qss_spatial_derivatives3d.f:823,900)
        const real sg = vn > real(0) ? real(1) : real(-1);
        const real gx = sg * inv_dx, gy = sg * inv_dy, gz = sg * inv_dz;
        const real gym = S == ENO3 ? gx : gy, gzm = S == ENO3 ? gx : gz;
        d[0] *= gx; d[1] *= gx;
        d[2] *= gy; d[3] *= gym;
        d[4] *= gz; d[5] *= gzm;
        adv_sign = -sg;
      }
      real temp_norm;
      if (fabs(vn) >= real(1.e-5)) {
        real g2 = (FW || FE) ? godunov_sq_signed(d) : godunov_sq(vn, d);
        real gn;
        if constexpr (FAST) gn = __builtin_amdgcn_sqrtf(g2);
        else gn = sqrt(g2);
        temp_norm = a.norm_sign * vn * gn;
      } else {
        temp_norm = (VAR && a.nstale > 0) ? stale_norm<S>(a, idx) : real(0);
      }
      // On the contact-angle path the curvature coefficient is h1 b, exactly
      // 0 deeper than 1.5 dx inside the solid, where the curvature term
      // b kappa |grad phi| is 0 whatever kappa is (up to the sign of a zero
      // term): a wave whose voxels all have b = 0 skips it.
      real curv = real(0);
      if (!VAR || bv != real(0)) {
        const real* Pp = pl[sp];
        const real* Pm = pl[sm];
        real xy[4] = {P[me + PX + 1], P[me - PX - 1], P[me + PX - 1], P[me - PX + 1]};
        real xz[4] = {Pp[me + 1], Pm[me - 1], Pp[me - 1], Pm[me + 1]};
        real yz[4] = {Pp[me + PX], Pm[me - PX], Pp[me - PX], Pm[me + PX]};
        if constexpr (FAST) {
          if (iso)
            curv = curvature_iso(lx[3], lx[1], lx[2], lx[4], lx[5], ly[1], ly[2], ly[4], ly[5],
                                 w[1], w[2], w[4], w[5], xy, xz, yz, f2, curv_thr);
          else
            curv = curvature_fast(lx[3], lx[1], lx[2], lx[4], lx[5], ly[1], ly[2], ly[4], ly[5],
                                  w[1], w[2], w[4], w[5], xy, xz, yz, fx, fy, fz);
        } else {
          curv = curvature(lx[3], lx[1], lx[2], lx[4], lx[5], ly[1], ly[2], ly[4], ly[5], w[1],
                           w[2], w[4], w[5], xy, xz, yz, fx, fy, fz);
        }
      }
      real rhs;
      if (VAR) {
        real ax = vxv >= real(0) ? d[1] : d[0];
        real ay = vyv >= real(0) ? d[3] : d[2];
        real az = vzv >= real(0) ? d[5] : d[4];
        if constexpr (FAST) {
          rhs = var_rhs_fast(temp_norm, bv, curv, vxv, vyv, vzv, ax, ay, az, adv_sign);
        } else {
          real temp_vel = -(vxv * ax + vyv * ay + vzv * az);
          rhs = temp_norm + bv * curv + temp_vel;
        }
      } else {
        rhs = temp_norm + bv * curv;
      }
      real val;
      if constexpr (FAST) {
        val = rk_value_fast<STAGE>(lx[3], unv, dt, rhs);
      } else if (STAGE == 1) {
        val = lx[3] + dt * rhs;
      } else if (STAGE == 2) {
        val = real(0.75) * unv + real(0.25) * (lx[3] + dt * rhs);
      } else {
        val = real(1.0 / 3.0) * unv + real(2.0 / 3.0) * (lx[3] + dt * rhs);
      }
      if (FINAL) {
        bool face = (i <= bx.ilo + 1) || (i >= bx.ihi - 1) || (j <= bx.jlo + 1) ||
                    (j >= bx.jhi - 1) || (k <= bx.klo + 1) || (k >= bx.khi - 1);
        if (face) a.pre[idx] = val;
        val = (flo > val) ? flo : val;
      }
      a.out[idx] = val;
    }
#pragma unroll
    for (int o = 0; o < 6; ++o) w[o] = w[o + 1];
  }
}

template <int S, bool VAR, int STAGE, bool FINAL, bool FAST>
__global__ __launch_bounds__(TX * TY) void stage_kernel(StageArgs a) {
  stage_tile<S, VAR, STAGE, FINAL, FAST>(a, xcd_tile(blockIdx.x, a.tiles_x * a.tiles_y * a.tiles_z));
}

template <int S, bool VAR, int STAGE, bool FAST>
static hipError_t launch_sf(const StageArgs& a, bool fin, dim3 g, dim3 b, hipStream_t st) {
  if (fin) hipLaunchKernelGGL((stage_kernel<S, VAR, STAGE, true, FAST>), g, b, 0, st, a);
  else hipLaunchKernelGGL((stage_kernel<S, VAR, STAGE, false, FAST>), g, b, 0, st, a);
  return hipGetLastError();
}

template <int S, bool VAR, bool FAST>
static hipError_t launch_s(const StageArgs& a, int stage, bool fin, dim3 g, dim3 b, hipStream_t st) {
  switch (stage) {
  case 1: return launch_sf<S, VAR, 1, FAST>(a, fin, g, b, st);
  case 2: return launch_sf<S, VAR, 2, FAST>(a, fin, g, b, st);
  default: return launch_sf<S, VAR, 3, FAST>(a, fin, g, b, st);
  }
}

template <bool FAST>
static hipError_t launch_scheme(const StageArgs& a, int scheme, bool var, int stage, bool fin,
                                dim3 grid, dim3 block, hipStream_t st) {
  switch (scheme) {
  case 2: return var ? launch_s<ENO2, true, FAST>(a, stage, fin, grid, block, st)
                     : launch_s<ENO2, false, FAST>(a, stage, fin, grid, block, st);
  case 3: return var ? launch_s<ENO3, true, FAST>(a, stage, fin, grid, block, st)
                     : launch_s<ENO3, false, FAST>(a, stage, fin, grid, block, st);
  case 5: return var ? launch_s<WENO5, true, FAST>(a, stage, fin, grid, block, st)
                     : launch_s<WENO5, false, FAST>(a, stage, fin, grid, block, st);
  default: return hipErrorInvalidValue;
  }
}

template <typename R>
static hipError_t pick_arith(const StageArgs& a, int scheme, bool var, int stage, bool fin,
                             dim3 grid, dim3 block, hipStream_t st) {
  if constexpr (std::is_same<R, float>::value) {
    if (a.fast) return launch_scheme<true>(a, scheme, var, stage, fin, grid, block, st);
  }
  return launch_scheme<false>(a, scheme, var, stage, fin, grid, block, st);
}

hipError_t launch_stage(const StageArgs& a, int scheme, bool var, int stage, bool fin,
                        hipStream_t st) {
  const Box& bx = a.bx;
  int nxf = bx.ihi - bx.ilo + 1, nyf = bx.jhi - bx.jlo + 1, nzf = bx.khi - bx.klo + 1;
  int gx = (nxf + TX - 1) / TX, gy = (nyf + TY - 1) / TY;
  // enough z-chunks to give every CU several blocks, each chunk >= 8 planes
  int cols = gx * gy;
  int gz = (2048 + cols - 1) / cols;
  gz = max(1, min(gz, (nzf + 7) / 8));
  StageArgs t = a;
  t.tiles_x = gx; t.tiles_y = gy; t.tiles_z = gz;
  dim3 grid(gx * gy * gz), block(TX, TY, 1);
  return pick_arith<real>(t, scheme, var, stage, fin, grid, block, st);
}

// ---------------------------------------------------------------------------
// Smoothed-sign reinitialization, one fused right-hand side + RK stage
// (right-hand side: qss_reinitialization3d.f:39-470 with use_phi0_for_sgn = 0;
// stages: reinitialize_top.c:205-295).  HJ-ENO3 one-sided derivatives (the
// minus sides of y and z scaled by 1/dx, as there; the 1/6 coefficient the
// reference leaves undefined is 1/6), Godunov selection by the sign of phi,
// rhs = S (1 - |grad phi|) with S = phi / sqrt(phi^2 + |grad phi|^2 dx_max^2),
// 0 where |phi| < 1e-5.  Stage 1: out = in + dt rhs; stage 2:
// out = 3/4 un + 1/4 (in + dt rhs).  Same tile walk as the stage kernel:
// 64 x 8 columns marched along z, z line in registers, four LDS planes.
template <int STAGE>
__global__ __launch_bounds__(TX * TY) void reinit_smooth_kernel(ReinitArgs a) {
  __shared__ real pl[NSLOT][PY * PX];
  const Box bx = a.bx;
  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * TX + tx;
  const int tile = xcd_tile(blockIdx.x, a.tiles_x * a.tiles_y * a.tiles_z);
  const int tcol = tile % a.tiles_x, trow = (tile / a.tiles_x) % a.tiles_y;
  const int tzc = tile / (a.tiles_x * a.tiles_y);
  const int i0 = bx.ilo + tcol * TX, j0 = bx.jlo + trow * TY;
  const int i = i0 + tx, j = j0 + ty;
  const int zc = (bx.khi - bx.klo + a.tiles_z) / a.tiles_z;
  const int k0 = bx.klo + tzc * zc;
  const int k1 = min(k0 + zc - 1, bx.khi);
  if (k0 > k1) return;
  const bool active = (i <= bx.ihi) && (j <= bx.jhi);
  const bool inbox = (i < bx.nx) && (j < bx.ny);
  const int64_t plane = (int64_t)bx.nx * bx.ny;
  const real* __restrict__ in = a.in;
  const HaloSlot hs = halo_slot(bx, i0, j0, tid);
  const int me = (ty + 3) * PX + tx + 3;
  const real inv_dx = real(1.0) / a.dx, inv_dy = real(1.0) / a.dy, inv_dz = real(1.0) / a.dz;
  real dx_sq = tmax(tmax(a.dx, a.dy), a.dz);
  dx_sq = dx_sq * dx_sq;

  const real* col = in + (inbox ? (int64_t)i + (int64_t)bx.nx * j : 0);
  real w[7];
#pragma unroll
  for (int o = 0; o < 7; ++o) w[o] = real(0);
  real wnext = real(0);
  if (inbox) {
#pragma unroll
    for (int o = 0; o < 6; ++o) w[o] = col[(int64_t)(k0 - 3 + o) * plane];
    wnext = col[(int64_t)(k0 + 3) * plane];
  }
  const real* hp = hs.off >= 0 ? in + hs.off : nullptr;
  auto halo = [&](int k) { return hp ? hp[(int64_t)k * plane] : real(0); };
  if (hs.lds >= 0) {
    pl[0][hs.lds] = halo(k0 - 1);
    pl[1][hs.lds] = halo(k0);
  }
  pl[0][me] = w[2];
  pl[1][me] = w[3];
  real hnext = hs.lds >= 0 ? halo(k0 + 1) : real(0);
  __syncthreads();

  int64_t idx = bx.at(i, j, k0);
  for (int k = k0; k <= k1; ++k, idx += plane) {
    const int sc = (k - k0 + 1) & 3, sp = (k - k0 + 2) & 3;
    w[6] = wnext;
    if (inbox && k + 4 < bx.nz) wnext = col[(int64_t)(k + 4) * plane];
    pl[sp][me] = w[4];
    if (hs.lds >= 0) {
      pl[sp][hs.lds] = hnext;
      hnext = halo(k + 2);
    }
    real unv = 0;
    if (active && STAGE == 2) unv = a.un[idx];
    __syncthreads();
    if (active) {
      const real* P = pl[sc];
      real lx[7], ly[7];
#pragma unroll
      for (int o = 0; o < 7; ++o) {
        lx[o] = P[me - 3 + o];
        ly[o] = P[me + (o - 3) * PX];
      }
      real gp[3], gm[3];
      eno3(lx, inv_dx, inv_dx, gp[0], gm[0]);
      eno3(ly, inv_dy, inv_dx, gp[1], gm[1]);
      eno3(w, inv_dz, inv_dx, gp[2], gm[2]);
      const real c = lx[3];
      real star[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        real p_, m_;
        if (c > real(0)) { p_ = tmax(-gp[d], real(0)); m_ = tmax(gm[d], real(0)); }
        else { p_ = tmax(gp[d], real(0)); m_ = tmax(-gm[d], real(0)); }
        star[d] = tmax(p_, m_);
      }
      real rhs = real(0);
      if (fabs(c) >= real(1.e-5)) {
        real n2 = star[0] * star[0] + star[1] * star[1] + star[2] * star[2];
        real sgn = c / sqrt(c * c + n2 * dx_sq);
        rhs = sgn * (real(1.0) - sqrt(n2));
      }
      real val;
      if (STAGE == 1) val = c + a.dt * rhs;
      else val = real(0.75) * unv + real(0.25) * (c + a.dt * rhs);
      a.out[idx] = val;
    }
#pragma unroll
    for (int o = 0; o < 6; ++o) w[o] = w[o + 1];
  }
}

hipError_t launch_reinit_smooth(const ReinitArgs& a, int stage, hipStream_t st) {
  const Box& bx = a.bx;
  int nxf = bx.ihi - bx.ilo + 1, nyf = bx.jhi - bx.jlo + 1, nzf = bx.khi - bx.klo + 1;
  int gx = (nxf + TX - 1) / TX, gy = (nyf + TY - 1) / TY;
  int cols = gx * gy;
  int gz = max(1, min((2048 + cols - 1) / cols, (nzf + 7) / 8));
  ReinitArgs t = a;
  t.tiles_x = gx; t.tiles_y = gy; t.tiles_z = gz;
  dim3 grid(gx * gy * gz), block(TX, TY, 1);
  if (stage == 1) hipLaunchKernelGGL(reinit_smooth_kernel<1>, grid, block, 0, st, t);
  else hipLaunchKernelGGL(reinit_smooth_kernel<2>, grid, block, 0, st, t);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Time-step probe: H/dx at (ihi_fb, jhi_fb, last plane of each slab), the
// voxels whose value the reference's per-thread max_H_over_dX ends up holding
// (qss_spatial_derivatives3d.f:490; constCurvModel3d.c:105-108,139-149).
// H/dx over the slab-end voxels reduced by the block; returns dt in thread 0
template <int S>
__device__ real probe_block(const ProbeArgs& a) {
  __shared__ real hs[1024];
  const Box bx = a.bx;
  int t = threadIdx.x;
  real h = real(-1);
  int nsl = bx.khi - bx.klo + 1;
  if (t < a.nslabs) {
    int klo = bx.klo + (int)((int64_t)nsl * t / a.nslabs);
    int khi = bx.klo + (int)((int64_t)nsl * (t + 1) / a.nslabs) - 1;
    if (khi > bx.khi) khi = bx.khi;
    if (klo <= khi) {
      real d[6];
      derivs_global<S>(a.phi, bx, bx.ihi, bx.jhi, khi, a.dx, a.dy, a.dz, a.eno3_quirk, d);
      real vn = a.vn ? a.vn[bx.at(bx.ihi, bx.jhi, khi)] : a.a_const;
      h = h_over_dx(vn, d, a.dx, a.dy, a.dz);
    }
  }
  hs[t] = h;
  __syncthreads();
  real dt = real(0);
  if (t == 0) {
    real cur = real(-1);
    for (int s = 0; s < a.nslabs; ++s)
      if (hs[s] > cur) cur = hs[s];
    dt = real(0.5) / (cur + a.bmax_over_dx + a.umax_over_dx);
  }
  return dt;
}

template <int S>
__global__ void probe_kernel(ProbeArgs a) {
  real dt = probe_block<S>(a);
  if (threadIdx.x == 0) *a.dt_out = dt;
}

hipError_t launch_probe(const ProbeArgs& a, int scheme, hipStream_t st) {
  if (a.nslabs < 1 || a.nslabs > 1024) return hipErrorInvalidValue;
  int nt = ((a.nslabs + 63) / 64) * 64;
  switch (scheme) {
  case 2: hipLaunchKernelGGL(probe_kernel<ENO2>, dim3(1), dim3(nt), 0, st, a); break;
  case 3: hipLaunchKernelGGL(probe_kernel<ENO3>, dim3(1), dim3(nt), 0, st, a); break;
  case 5: hipLaunchKernelGGL(probe_kernel<WENO5>, dim3(1), dim3(nt), 0, st, a); break;
  default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Boundary condition.  The reference extrapolates face by face (x-, x+, y-,
// y+, z-, z+), each pass covering the whole extent of the other two axes
// (qss_spatial_derivatives3d.f:3055-3191, qss_general_util.c:630-698).  The
// value a ghost voxel ends with therefore depends only on fill-box values:
// z ghosts extrapolate the (x,y)-completed planes klo/klo+1, y ghosts the
// x-completed rows jlo/jlo+1, x ghosts the fill values ilo/ilo+1.  Each
// ghost voxel evaluates that composition directly, so one launch with no
// ordering replaces the six passes.
__device__ __forceinline__ real extrap(real c, real c2, int dist) {
  real s = copysign(real(1), c);
  real slope = s * fabs(c - c2);
  return c + slope * real(dist);
}

__device__ __forceinline__ real bc_x(const real* f, const Box& bx, int i, int j, int k) {
  if (i < bx.ilo) return extrap(f[bx.at(bx.ilo, j, k)], f[bx.at(bx.ilo + 1, j, k)], bx.ilo - i);
  if (i > bx.ihi) return extrap(f[bx.at(bx.ihi, j, k)], f[bx.at(bx.ihi - 1, j, k)], i - bx.ihi);
  return f[bx.at(i, j, k)];
}
__device__ __forceinline__ real bc_y(const real* f, const Box& bx, int i, int j, int k) {
  if (j < bx.jlo) return extrap(bc_x(f, bx, i, bx.jlo, k), bc_x(f, bx, i, bx.jlo + 1, k), bx.jlo - j);
  if (j > bx.jhi) return extrap(bc_x(f, bx, i, bx.jhi, k), bc_x(f, bx, i, bx.jhi - 1, k), j - bx.jhi);
  return bc_x(f, bx, i, j, k);
}
__device__ __forceinline__ real bc_z(const real* f, const Box& bx, int i, int j, int k) {
  if (k < bx.klo) return extrap(bc_y(f, bx, i, j, bx.klo), bc_y(f, bx, i, j, bx.klo + 1), bx.klo - k);
  if (k > bx.khi) return extrap(bc_y(f, bx, i, j, bx.khi), bc_y(f, bx, i, j, bx.khi - 1), k - bx.khi);
  return bc_y(f, bx, i, j, k);
}

// ghost voxel g of the three regions (z ghost planes, then y ghost rows of
// the fill planes, then x ghost columns of the fill rows)
template <bool FINAL>
__device__ __forceinline__ void bc_one(const BcArgs& a, int64_t g, int64_t nA, int64_t nB) {
  const Box bx = a.bx;
  int i, j, k;
  ghost_ijk(bx, g, nA, nB, i, j, k);
  real v = bc_z(a.src, bx, i, j, k);
  int64_t idx = bx.at(i, j, k);
  if (FINAL) {
    real f = a.floor[idx];
    v = (f > v) ? f : v;
  }
  a.dst[idx] = v;
}

template <bool FINAL>
__global__ void bc_kernel(BcArgs a, int64_t nA, int64_t nB, int64_t nC) {
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < nA + nB + nC; g += stride)
    bc_one<FINAL>(a, g, nA, nB);
}

int64_t ghost_count(const Box& bx) {
  int64_t nA, nB, nC;
  ghost_regions(bx, nA, nB, nC);
  return nA + nB + nC;
}

int bc_pass_blocks(const Box& bx) {
  int64_t nA, nB, nC;
  ghost_regions(bx, nA, nB, nC);
  int64_t tot = nA + nB + nC;
  int blocks = (int)std::min<int64_t>((tot + 255) / 256, 8192);
  return blocks < 1 ? 1 : blocks;
}

hipError_t launch_bc(const BcArgs& a, bool final_mask, hipStream_t st) {
  int64_t nA, nB, nC;
  ghost_regions(a.bx, nA, nB, nC);
  const int blocks = bc_pass_blocks(a.bx);
  if (final_mask) hipLaunchKernelGGL(bc_kernel<true>, dim3(blocks), dim3(256), 0, st, a, nA, nB, nC);
  else hipLaunchKernelGGL(bc_kernel<false>, dim3(blocks), dim3(256), 0, st, a, nA, nB, nC);
  return hipGetLastError();
}

// ghost extrapolation of the reinitialized field rout[pass] of many boxes
// (batched outer steps; Multi: block -> box and its block there)
__global__ void bc_multi_kernel(const OuterJob* J, Multi m, int pass) {
  int lb, lg;
  const int r = multi_box(m, blockIdx.x, lb, lg);
  BcArgs a{};
  a.bx = J[r].ta.bx;
  a.src = J[r].rout[pass];
  a.dst = J[r].rout[pass];
  int64_t nA, nB, nC;
  ghost_regions(a.bx, nA, nB, nC);
  const int64_t stride = (int64_t)lg * blockDim.x;
  for (int64_t g = (int64_t)lb * blockDim.x + threadIdx.x; g < nA + nB + nC; g += stride)
    bc_one<false>(a, g, nA, nB);
}
hipError_t launch_bc_multi(const OuterJob* J, int pass, const Multi& ghosts, hipStream_t st) {
  hipLaunchKernelGGL(bc_multi_kernel, dim3(ghosts.total), dim3(256), 0, st, J, ghosts, pass);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Batched inner steps (records, counts and prefix sums in device memory).

// largest r < n with prefix[r] <= b
template <typename I>
__device__ __forceinline__ int find_rec(const I* prefix, int n, I b) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__global__ void batch_reset_kernel(int* counters, int n) {
  if ((int)threadIdx.x < n) counters[threadIdx.x] = 0;
}
hipError_t launch_batch_reset(int* counters, int ncounters, hipStream_t st) {
  hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(64), 0, st, counters, ncounters);
  return hipGetLastError();
}

template <int S>
__global__ void probe_batch_kernel(BatchView v) {
  const int r = blockIdx.x;
  if (r >= v.count[0]) return;
  const ProbeArgs a = v.recs[r].probe;
  real dt = probe_block<S>(a);
  if (threadIdx.x == 0) {
    *a.dt_out = dt;
    v.dts[r] = dt;
  }
}
hipError_t launch_probe_batch(const BatchView& v, int scheme, int nmax, hipStream_t st) {
  dim3 grid(nmax), block(1024);
  switch (scheme) {
  case 2: hipLaunchKernelGGL(probe_batch_kernel<ENO2>, grid, block, 0, st, v); break;
  case 3: hipLaunchKernelGGL(probe_batch_kernel<ENO3>, grid, block, 0, st, v); break;
  case 5: hipLaunchKernelGGL(probe_batch_kernel<WENO5>, grid, block, 0, st, v); break;
  default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Persistent stage kernel: each workgroup takes the next brick of the queue
// (bricks of all records in record order) until the queue is empty.
template <int S, bool VAR, int STAGE, bool FINAL, bool FAST>
__global__ __launch_bounds__(TX * TY) void stage_batch_kernel(BatchView v, int si) {
  __shared__ int next;
  const int n = v.count[0], total = v.brick_prefix[n];
  const int tid = threadIdx.y * TX + threadIdx.x;
  for (;;) {
    if (tid == 0) next = atomicAdd(&v.counters[si], 1);
    __syncthreads();
    const int b = next;
    __syncthreads();
    if (b >= total) break;
    const int r = find_rec(v.brick_prefix, n, b);
    const StageArgs a = v.recs[r].stage[si];
    stage_tile<S, VAR, STAGE, FINAL, FAST>(a, b - v.brick_prefix[r]);
  }
}

template <int S, bool VAR, bool FAST>
static hipError_t stage_batch_s(const BatchView& v, int si, int stage, bool fin, dim3 g, hipStream_t st) {
  dim3 b(TX, TY, 1);
#define LSM_SB(ST, FI) hipLaunchKernelGGL((stage_batch_kernel<S, VAR, ST, FI, FAST>), g, b, 0, st, v, si)
  switch (stage * 2 + (fin ? 1 : 0)) {
  case 2: LSM_SB(1, false); break;
  case 3: LSM_SB(1, true); break;
  case 4: LSM_SB(2, false); break;
  case 5: LSM_SB(2, true); break;
  case 6: LSM_SB(3, false); break;
  default: LSM_SB(3, true); break;
  }
#undef LSM_SB
  return hipGetLastError();
}

template <bool FAST>
static hipError_t stage_batch_f(const BatchView& v, int si, int scheme, bool var, int stage,
                                bool fin, dim3 g, hipStream_t st) {
  switch (scheme) {
  case 2: return var ? stage_batch_s<ENO2, true, FAST>(v, si, stage, fin, g, st)
                     : stage_batch_s<ENO2, false, FAST>(v, si, stage, fin, g, st);
  case 3: return var ? stage_batch_s<ENO3, true, FAST>(v, si, stage, fin, g, st)
                     : stage_batch_s<ENO3, false, FAST>(v, si, stage, fin, g, st);
  case 5: return var ? stage_batch_s<WENO5, true, FAST>(v, si, stage, fin, g, st)
                     : stage_batch_s<WENO5, false, FAST>(v, si, stage, fin, g, st);
  default: return hipErrorInvalidValue;
  }
}

hipError_t launch_stage_batch(const BatchView& v, int si, int scheme, bool var, int stage,
                              bool fin, bool fast, int grid, hipStream_t st) {
  if (grid < 1) return hipErrorInvalidValue;
  if constexpr (std::is_same<real, float>::value) {
    if (fast) return stage_batch_f<true>(v, si, scheme, var, stage, fin, dim3(grid), st);
  }
  return stage_batch_f<false>(v, si, scheme, var, stage, fin, dim3(grid), st);
}

int stage_bricks(StageArgs& a, int zc) {
  const Box& bx = a.bx;
  int nxf = bx.ihi - bx.ilo + 1, nyf = bx.jhi - bx.jlo + 1, nzf = bx.khi - bx.klo + 1;
  a.tiles_x = (nxf + TX - 1) / TX;
  a.tiles_y = (nyf + TY - 1) / TY;
  a.tiles_z = std::max(1, (nzf + zc - 1) / zc);
  return a.tiles_x * a.tiles_y * a.tiles_z;
}

// Each block takes a contiguous chunk of the ghost voxels of all boxes; a
// thread finds the box of its first voxel by binary search and then steps
// forward through the prefix sums as its voxels advance.
template <bool FINAL>
__global__ void bc_batch_kernel(BatchView v, int si) {
  const int n = v.count[0];
  const int64_t total = v.ghost_prefix[n];
  const int64_t chunk = (total + gridDim.x - 1) / gridDim.x;
  const int64_t g0 = (int64_t)blockIdx.x * chunk;
  const int64_t g1 = min(g0 + chunk, total);
  int r = -1;
  int64_t nA = 0, nB = 0, nC = 0;
  for (int64_t g = g0 + threadIdx.x; g < g1; g += blockDim.x) {
    if (r < 0) r = find_rec(v.ghost_prefix, n, g);
    else
      while (v.ghost_prefix[r + 1] <= g) ++r;
    const BcArgs& a = v.recs[r].bc[si];
    ghost_regions(a.bx, nA, nB, nC);
    bc_one<FINAL>(a, g - v.ghost_prefix[r], nA, nB);
  }
}
hipError_t launch_bc_batch(const BatchView& v, int si, bool final_mask, int grid, hipStream_t st) {
  if (final_mask) hipLaunchKernelGGL(bc_batch_kernel<true>, dim3(grid), dim3(256), 0, st, v, si);
  else hipLaunchKernelGGL(bc_batch_kernel<false>, dim3(grid), dim3(256), 0, st, v, si);
  return hipGetLastError();
}

}  // namespace lsm
