// lsm_device.h -- per-voxel numerics of the LSMPQS 3D hot path, as device
// functions templated on the storage/compute type T (float or double).
//
// Every expression keeps the reference's operation order and the library is
// compiled without FMA contraction, so the double instantiation reproduces
// the reference's FP64 results bit for bit wherever only + - * / sqrt are
// involved.  Citations are to src/spatial_derivatives/qss_spatial_derivatives3d.f
// unless stated otherwise.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lsm {

enum Scheme { ENO2 = 2, ENO3 = 3, WENO5 = 5 };

template <typename T> __host__ __device__ __forceinline__ T tmax(T a, T b) { return a > b ? a : b; }
template <typename T> __host__ __device__ __forceinline__ T tmin(T a, T b) { return a < b ? a : b; }
template <typename T> __host__ __device__ __forceinline__ T tabs(T a) { return fabs(a); }

// Geometry of one ghost box and its fill box; passed by value to kernels.
struct Box {
  int nx, ny, nz;             // ghost-box extents
  int ilo, ihi, jlo, jhi, klo, khi;  // fill box (inclusive)
  __host__ __device__ int64_t at(int i, int j, int k) const {
    return (int64_t)i + (int64_t)nx * ((int64_t)j + (int64_t)ny * k);
  }
  __host__ __device__ int64_t size() const { return (int64_t)nx * ny * nz; }
  __host__ __device__ bool in_fill(int i, int j, int k) const {
    return i >= ilo && i <= ihi && j >= jlo && j <= jhi && k >= klo && k <= khi;
  }
};

// Workgroups are handed to the eight XCDs round-robin by linear id, and each
// XCD has its own L2.  Block b is given tile number xcd * share + b / 8 (the
// shares differ by at most one when the count is not a multiple of 8), so an
// XCD works through a contiguous run of tiles, x fastest, and the halo rows
// and columns neighbouring tiles share are fetched into one L2 instead of
// several.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  const int q = nb >> 3, r = nb & 7, xcd = b & 7, slot = b >> 3;
  return xcd < r ? xcd * (q + 1) + slot : r * (q + 1) + (xcd - r) * q + slot;
}

// HJ-ENO2 one-sided derivatives from p[0..6] = phi(-3..+3) (:386-456)
template <typename T>
__device__ __forceinline__ void eno2(const T* p, T inv_d, T& plus, T& minus) {
  const T half = T(0.5);
  T D1 = p[3] - p[2];
  T D1p = p[4] - p[3];
  T D2 = p[4] + p[2] - T(2) * p[3];
  T D2p = p[5] + p[3] - T(2) * p[4];
  T D2m = p[1] + p[3] - T(2) * p[2];
  plus = (tabs(D2) < tabs(D2p)) ? (D1p - half * D2) * inv_d : (D1p - half * D2p) * inv_d;
  minus = (tabs(D2m) < tabs(D2)) ? (D1 + half * D2m) * inv_d : (D1 + half * D2) * inv_d;
}

// HJ-ENO3 (:672-745); the minus side is scaled by inv_minus, which the
// reference sets to 1/dx on every axis (:823,900)
template <typename T>
__device__ __forceinline__ void eno3(const T* p, T inv_plus, T inv_minus, T& plus, T& minus) {
  const T half = T(0.5), third = T(1.0 / 3.0), sixth = T(1.0 / 6.0);
  T D1 = p[3] - p[2], D1p = p[4] - p[3], D1m = p[2] - p[1];
  T D1mm = p[1] - p[0], D1pp = p[5] - p[4], D1ppp = p[6] - p[5];
  T D2 = D1p - D1, D2p = D1pp - D1p, D2m = D1 - D1m;
  T D2pp = D1ppp - D1pp, D2mm = D1m - D1mm;
  T D3 = D2 - D2m, D3p = D2p - D2, D3pp = D2pp - D2p, D3m = D2m - D2mm;
  T v = D1p;
  if (tabs(D2) < tabs(D2p)) {
    v = v - half * D2;
    v = (tabs(D3) < tabs(D3p)) ? v - sixth * D3 : v - sixth * D3p;
  } else {
    v = v - half * D2p;
    v = (tabs(D3p) < tabs(D3pp)) ? v + third * D3p : v + third * D3pp;
  }
  plus = v * inv_plus;
  v = D1;
  if (tabs(D2m) < tabs(D2)) {
    v = v + half * D2m;
    v = (tabs(D3m) < tabs(D3)) ? v + third * D3m : v + third * D3;
  } else {
    v = v + half * D2;
    v = (tabs(D3) < tabs(D3p)) ? v - sixth * D3 : v - sixth * D3p;
  }
  minus = v * inv_minus;
}

// Fast single-precision HJ-ENO2 / HJ-ENO3: the same stencil choices as eno2
// / eno3, with FMA contraction, selections instead of branches, the
// comparisons the two sides share evaluated once, and the one-sided values
// left unscaled (times dx): the caller applies sign(vel_n)/dx per component,
// as for the fast WENO5.  Rounding differs from the reference's expression
// order by a few ulp, so a stencil choice at a near-tie can differ.
template <typename T>
__host__ __device__ __forceinline__ void eno2_fast_raw(const T* p, T& plus, T& minus) {
#pragma clang fp contract(fast)
  const T D1 = p[3] - p[2], D1p = p[4] - p[3], D1m = p[2] - p[1], D1pp = p[5] - p[4];
  const T D2 = D1p - D1, D2p = D1pp - D1p, D2m = D1 - D1m;
  plus = T(-0.5) * (fabs(D2) < fabs(D2p) ? D2 : D2p) + D1p;
  minus = T(0.5) * (fabs(D2m) < fabs(D2) ? D2m : D2) + D1;
}

template <typename T>
__host__ __device__ __forceinline__ void eno3_fast_raw(const T* p, T& plus, T& minus) {
#pragma clang fp contract(fast)
  const T D1 = p[3] - p[2], D1p = p[4] - p[3], D1m = p[2] - p[1];
  const T D1mm = p[1] - p[0], D1pp = p[5] - p[4], D1ppp = p[6] - p[5];
  const T D2 = D1p - D1, D2p = D1pp - D1p, D2m = D1 - D1m;
  const T D2pp = D1ppp - D1pp, D2mm = D1m - D1mm;
  const T D3 = D2 - D2m, D3p = D2p - D2, D3pp = D2pp - D2p, D3m = D2m - D2mm;
  const bool pl = fabs(D2) < fabs(D2p);    // plus side: stencil widens to the left
  const bool mi = fabs(D2m) < fabs(D2);    // minus side: stencil widens to the left
  const bool c0 = fabs(D3) < fabs(D3p);    // shared by both sides
  const T s3p = pl ? (c0 ? D3 : D3p) : (fabs(D3p) < fabs(D3pp) ? D3p : D3pp);
  const T s3m = mi ? (fabs(D3m) < fabs(D3) ? D3m : D3) : (c0 ? D3 : D3p);
  plus = (pl ? T(-1.0 / 6.0) : T(1.0 / 3.0)) * s3p + (T(-0.5) * (pl ? D2 : D2p) + D1p);
  minus = (mi ? T(1.0 / 3.0) : T(-1.0 / 6.0)) * s3m + (T(0.5) * (mi ? D2m : D2) + D1);
}

// HJ-WENO5 combination with linear weights 0.1/0.6/0.3 (:1960-2035)
template <typename T>
__host__ __device__ __forceinline__ T weno5_combine(T v1, T v2, T v3, T v4, T v5) {
  const T one_third = T(1.0 / 3.0), seven_sixths = T(7.0 / 6.0);
  const T eleven_sixths = T(11.0 / 6.0), one_sixth = T(1.0 / 6.0);
  const T five_sixths = T(5.0 / 6.0), thirteen_twelfths = T(13.0 / 12.0);
  const T one_fourth = T(0.25), tiny = T(1.0e-35);
  T eps = T(1e-6) * tmax(tmax(tmax(tmax(v1 * v1, v2 * v2), v3 * v3), v4 * v4), v5 * v5) + tiny;
  T p1 = one_third * v1 - seven_sixths * v2 + eleven_sixths * v3;
  T p2 = -one_sixth * v2 + five_sixths * v3 + one_third * v4;
  T p3 = one_third * v3 + five_sixths * v4 - one_sixth * v5;
  T t, S1, S2, S3;
  t = v1 - T(2) * v2 + v3; S1 = thirteen_twelfths * (t * t);
  t = v1 - T(4) * v2 + T(3) * v3; S1 = S1 + one_fourth * (t * t);
  t = v2 - T(2) * v3 + v4; S2 = thirteen_twelfths * (t * t);
  t = v2 - v4; S2 = S2 + one_fourth * (t * t);
  t = v3 - T(2) * v4 + v5; S3 = thirteen_twelfths * (t * t);
  t = T(3) * v3 - T(4) * v4 + v5; S3 = S3 + one_fourth * (t * t);
  T a1, a2, a3;
  t = S1 + eps; a1 = T(0.1) / (t * t);
  t = S2 + eps; a2 = T(0.6) / (t * t);
  t = S3 + eps; a3 = T(0.3) / (t * t);
  T inv = T(1) / (a1 + a2 + a3);
  a1 = a1 * inv; a2 = a2 * inv; a3 = a3 * inv;
  return a1 * p1 + a2 * p2 + a3 * p3;
}

template <typename T>
__host__ __device__ __forceinline__ void weno5(const T* p, T inv_d, T& plus, T& minus) {
  T D1 = p[3] - p[2], D1p = p[4] - p[3], D1m = p[2] - p[1];
  T D1pp = p[5] - p[4], D1ppp = p[6] - p[5], D1mm = p[1] - p[0];
  plus = weno5_combine(D1ppp * inv_d, D1pp * inv_d, D1p * inv_d, D1 * inv_d, D1m * inv_d);
  minus = weno5_combine(D1mm * inv_d, D1m * inv_d, D1 * inv_d, D1p * inv_d, D1pp * inv_d);
}

// ---------------------------------------------------------------------------
// Fast single-precision HJ-WENO5.  Same scheme and weights, rearranged for
// throughput: with q_k = (S_k + eps)^2 the normalised weights are
// w_k = c_k prod_{m != k} q_m / sum, so one hardware reciprocal replaces the
// four IEEE divisions, and FMA contraction is allowed.  The weights depend
// only on ratios, so the smoothness indicators are formed from the raw
// differences divided by their largest magnitude m, with eps divided by
// (m/dx)^2 (1e-6 + 1e-35 dx^2/m^2): every q_k then lies in [1e-12, 2e3] and
// no product underflows into the denormal range, where the hardware
// reciprocal returns infinity.  The candidates are formed from the raw
// differences and scaled by 1/dx once.  Rounding differs from the
// reference's expression order by a few ulp.  Below m = 1e-18 dx the
// reconstruction is the middle candidate (the reference's float build forms
// 0/0 when all differences vanish).
template <typename T>
struct Weno5Face {
  T minus_right;  // phi^-(i+1)
  T plus_left;    // phi^+(i)
};

// Per-axis constants of the fast reconstruction.
template <typename T>
struct FaceScale {
  T inv;  // 1/dx
  T thr;  // 1e-18 dx: flat-window bound on the largest raw difference
  T c1;   // sqrt(1e-35) dx: eps / (m/dx)^2 = 1e-6 + (c1/m)^2
  T c2;   // 2 c1
};
template <typename T>
__host__ __device__ inline FaceScale<T> face_scale(T dx) {
  FaceScale<T> s;
  s.inv = T(1) / dx;
  s.thr = T(1e-18) * dx;
  s.c1 = T(3.16227766016837933e-18) * dx;
  s.c2 = T(6.32455532033675866e-18) * dx;
  return s;
}

// v_rcp_f32 (1 ulp, denormal inputs read as zero); the host version, used
// only by host-side tests of these formulas, models the same flush
__host__ __device__ __forceinline__ float fast_rcp(float x) {
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_amdgcn_rcpf(x);
#else
  return (x != 0.0f && fabsf(x) < 1.17549435e-38f) ? copysignf(__builtin_inff(), x) : 1.0f / x;
#endif
}
__host__ __device__ __forceinline__ double fast_rcp(double x) { return 1.0 / x; }

// Both one-sided values that share the face between voxels i and i+1:
// q[0..5] = phi(i-2..i+3).  phi^-(i+1) uses the differences r0..r4 in order,
// phi^+(i) the same differences reversed, so the smoothness indicators are
// computed once (the reversed window's S1 and S3 are the window's S3 and S1).  The combination is written as
// p2 + w1 (p1 - p2) + w3 (p3 - p2), where p1 - p2 = g1/3 and p3 - p2 = g2/6
// with the third differences g1 = r0 - 3r1 + 3r2 - r3, g2 = r1 - 3r2 + 3r3 - r4
// (Jiang and Peng, SIAM J. Sci. Comput. 21 (2000), eq. 2.10); the reversed
// window has g1' = -g2 and g2' = -g1.
//
// weno5_face_raw returns both values times dx (the caller applies 1/dx,
// folded into another product).  All weight factors are carried at a
// common scale, which the normalised weights do not see: the smoothness
// indicators as 4 S_k = 13/3 t_k^2 + u_k^2 with eps times 4, so q_k is 16
// times the q_k above (then in [1.6e-11, 3.2e4]), and the weight sums and
// numerators times 10.  With A = q23 g1 and B = q12 g2 the numerators are
// A/3 + B/2 = (A + B)/3 + B/6 and -(A/2 + B/3) = -((A + B)/3 + A/6), the sums
// q23 + 3 q12 + 6 q13 and q12 + 3 q23 + 6 q13.  The middle candidates are
// r2 + (D2 + 2 D3)/6 and r2 - (2 D2 + D3)/6 with the second differences
// D_k = r_k - r_{k-1}.
template <typename T>
__host__ __device__ __forceinline__ Weno5Face<T> weno5_face_raw(const T* q, const FaceScale<T>& sc) {
#pragma clang fp contract(fast)
  T r0 = q[1] - q[0], r1 = q[2] - q[1], r2 = q[3] - q[2], r3 = q[4] - q[3], r4 = q[5] - q[4];
  T D1 = r1 - r0, D2 = r2 - r1, D3 = r3 - r2, D4 = r4 - r3;
  T u2 = D2 + D3;
  Weno5Face<T> f;
  f.minus_right = (T(1) / T(6)) * (u2 + D3) + r2;
  f.plus_left = (T(-1) / T(6)) * (u2 + D2) + r2;
  T m = fmax(fmax(fmax(fabs(r0), fabs(r1)), fabs(r2)), fmax(fabs(r3), fabs(r4)));
  if (!(m > sc.thr)) return f;
  T t1 = D2 - D1, t2 = D3 - D2, t3 = D4 - D3;         // e1-2e2+e3, e2-2e3+e4, e3-2e4+e5
  T u1 = T(3) * D2 - D1, u3 = T(3) * D3 - D4;         // e1-4e2+3e3, 3e3-4e4+e5
  T S1 = (T(13) / T(3) * t1) * t1 + u1 * u1;
  T S2 = (T(13) / T(3) * t2) * t2 + u2 * u2;
  T S3 = (T(13) / T(3) * t3) * t3 + u3 * u3;
  T im = fast_rcp(m);
  T im2 = im * im;
  T ce = sc.c2 * im;
  T eps = ce * ce + T(4e-6);
  T q1 = S1 * im2 + eps; q1 = q1 * q1;
  T q2 = S2 * im2 + eps; q2 = q2 * q2;
  T q3 = S3 * im2 + eps; q3 = q3 * q3;
  T q12 = q1 * q2, q13 = q1 * q3, q23 = q2 * q3;
  T A = q23 * (t1 - t2);  // g1 = r0 - 3r1 + 3r2 - r3
  T B = q12 * (t2 - t3);  // g2 = r1 - 3r2 + 3r3 - r4
  T base = T(6) * q13 + (q12 + q23);
  T sum_m = T(2) * q12 + base;
  T sum_p = T(2) * q23 + base;
  T s3 = (A + B) * (T(1) / T(3));
  T num_m = (T(1) / T(6)) * B + s3;
  T num_p = (T(1) / T(6)) * A + s3;
  f.minus_right = num_m * fast_rcp(sum_m) + f.minus_right;
  f.plus_left = -(num_p * fast_rcp(sum_p)) + f.plus_left;
  return f;
}

// weno5_face_raw scaled by 1/dx
template <typename T>
__host__ __device__ __forceinline__ Weno5Face<T> weno5_face(const T* q, const FaceScale<T>& sc) {
  Weno5Face<T> f = weno5_face_raw(q, sc);
  f.minus_right = f.minus_right * sc.inv;
  f.plus_left = f.plus_left * sc.inv;
  return f;
}

// kappa |grad phi| with FMA contraction and a reciprocal (fast T path)
template <typename T>
__device__ __forceinline__ T curvature_fast(T c, T xm2, T xm1, T xp1, T xp2, T ym2, T ym1, T yp1,
                                            T yp2, T zm2, T zm1, T zp1, T zp2, const T* xy,
                                            const T* xz, const T* yz, T fx, T fy, T fz) {
#pragma clang fp contract(fast)
  T px = (xp1 - xm1) * fx, py = (yp1 - ym1) * fy, pz = (zp1 - zm1) * fz;
  T pxx = (xp2 + xm2 - T(2.0) * c) * (fx * fx);
  T pzz = (zp2 + zm2 - T(2.0) * c) * (fz * fz);
  T pyy = (yp2 + ym2 - T(2.0) * c) * (fy * fy);
  T pxy = (xy[0] + xy[1] - xy[2] - xy[3]) * (fx * fy);
  T pxz = (xz[0] + xz[1] - xz[2] - xz[3]) * (fx * fz);
  T pyz = (yz[0] + yz[1] - yz[2] - yz[3]) * (fy * fz);
  T g2 = px * px + py * py + pz * pz;
  if (g2 < 1.e-5f) return T(0.0);
  T curv = pxx * (py * py + pz * pz) + pyy * (px * px + pz * pz) + pzz * (px * px + py * py)
             - T(2.0) * (pxy * px * py + pxz * px * pz + pyz * py * pz);
  return curv * fast_rcp(g2);
}

// curvature_fast on a grid with dx = dy = dz (f = fx = fy = fz): the
// expression is homogeneous, so the differences are taken unscaled, the
// common factor f^2 is applied once, and the |grad phi|^2 < 1e-5 test reads
// g2 < thr with thr = 1e-5 / f^2 on the unscaled sum
template <typename T>
__device__ __forceinline__ T curvature_iso(T c, T xm2, T xm1, T xp1, T xp2, T ym2, T ym1, T yp1,
                                           T yp2, T zm2, T zm1, T zp1, T zp2, const T* xy,
                                           const T* xz, const T* yz, T f2, T thr) {
#pragma clang fp contract(fast)
  T px = xp1 - xm1, py = yp1 - ym1, pz = zp1 - zm1;
  T pxx = (xp2 + xm2) - T(2.0) * c;
  T pzz = (zp2 + zm2) - T(2.0) * c;
  T pyy = (yp2 + ym2) - T(2.0) * c;
  T pxy = (xy[0] + xy[1]) - (xy[2] + xy[3]);
  T pxz = (xz[0] + xz[1]) - (xz[2] + xz[3]);
  T pyz = (yz[0] + yz[1]) - (yz[2] + yz[3]);
  T sx = px * px, sy = py * py, sz = pz * pz;
  T g2 = sx + sy + sz;
  if (g2 < thr) return T(0.0);
  T curv = pxx * (sy + sz) + pyy * (sx + sz) + pzz * (sx + sy)
             - T(2.0) * (pxy * px * py + pxz * px * pz + pyz * py * pz);
  return (f2 * curv) * fast_rcp(g2);
}

// Upwind derivative pair along one axis with the scheme's scalings.
template <typename T, int S>
__device__ __forceinline__ void axis_derivs(const T* p, T inv_axis, T inv_x, T& plus, T& minus) {
  if (S == ENO2) eno2(p, inv_axis, plus, minus);
  else if (S == ENO3) eno3(p, inv_axis, inv_x, plus, minus);
  else weno5(p, inv_axis, plus, minus);
}

// Godunov |grad phi|^2 selected by the sign of vel_n (:460-480).  Per axis
// the reference takes max(max(d-,0)^2, min(d+,0)^2) for vel_n > 0 and
// max(min(d-,0)^2, max(d+,0)^2) otherwise; since rounding of a square is
// monotone this equals max(s d-, -s d+, 0)^2 with s = sign(vel_n), bit for
// bit, without a branch on the sign.
template <typename T>
__device__ __forceinline__ T godunov_sq(T vn, const T* d) {
  const T s = vn > T(0) ? T(1) : T(-1);
  T x = fmax(fmax(s * d[1], -(s * d[0])), T(0));
  T y = fmax(fmax(s * d[3], -(s * d[2])), T(0));
  T z = fmax(fmax(s * d[5], -(s * d[4])), T(0));
  return x * x + y * y + z * z;
}

// godunov_sq of one-sided derivatives already multiplied by s = sign(vel_n):
// per axis max(d-, -d+, 0)^2
template <typename T>
__device__ __forceinline__ T godunov_sq_signed(const T* d) {
#pragma clang fp contract(fast)
  T x = fmax(fmax(d[1], -d[0]), T(0));
  T y = fmax(fmax(d[3], -d[2]), T(0));
  T z = fmax(fmax(d[5], -d[4]), T(0));
  return x * x + y * y + z * z;
}

// Right-hand side of the variable-coefficient path (fast T path):
// temp_norm + b kappa|grad phi| + adv_sign (V . upwind grad phi), with
// adv_sign = -1, or -sign(vel_n) when the derivatives carry that sign
template <typename T>
__device__ __forceinline__ T var_rhs_fast(T temp_norm, T b, T curv, T vx, T vy, T vz, T ax, T ay,
                                          T az, T adv_sign) {
#pragma clang fp contract(fast)
  return (b * curv + temp_norm) + adv_sign * (vx * ax + vy * ay + vz * az);
}

// TVD-RK stage value from the stage input c, u^n and the right-hand side
// (fast T path): RK1 c + dt L; "RK2 stage 2" and RK3 stage 2
// 3/4 u^n + 1/4 (c + dt L); RK3 stage 3 1/3 u^n + 2/3 (c + dt L)
template <int STAGE, typename T>
__device__ __forceinline__ T rk_value_fast(T c, T un, T dt, T rhs) {
#pragma clang fp contract(fast)
  T v = dt * rhs + c;
  if (STAGE == 1) return v;
  if (STAGE == 2) return T(0.75) * un + T(0.25) * v;
  return T(1.0 / 3.0) * un + T(2.0 / 3.0) * v;
}

// H/dx of the CFL bound (:488-513); d = {x+, x-, y+, y-, z+, z-}
template <typename T>
__device__ __forceinline__ T h_over_dx(T vn, const T* d, T dx, T dy, T dz) {
  T mdx = tmax(tmax(dx, dy), dz);
  T max_dx_sq = mdx * mdx;
  T inv_dx = T(1.0) / dx, inv_dy = T(1.0) / dy, inv_dz = T(1.0) / dz;
  T xc = tmax(tabs(d[0]), tabs(d[1]));
  T yc = tmax(tabs(d[2]), tabs(d[3]));
  T zc = tmax(tabs(d[4]), tabs(d[5]));
  T ng = sqrt(xc * xc + yc * yc + zc * zc + max_dx_sq);
  return tabs(vn) / ng * (xc * inv_dx + yc * inv_dy + zc * inv_dz);
}

// kappa |grad phi| from wide central differences (:518-556).  c = phi(i,j,k);
// xm2..xp2 the x line, likewise y, z; cross = {(+1,+1),(-1,-1),(-1,+1),(+1,-1)}
// in the xy, xz and yz planes.
template <typename T>
__device__ __forceinline__ T curvature(T c, T xm2, T xm1, T xp1, T xp2, T ym2, T ym1, T yp1, T yp2,
                                       T zm2, T zm1, T zp1, T zp2, const T* xy, const T* xz,
                                       const T* yz, T fx, T fy, T fz) {
  T px = (xp1 - xm1) * fx;
  T py = (yp1 - ym1) * fy;
  T pz = (zp1 - zm1) * fz;
  T pxx = (xp2 + xm2 - T(2) * c) * fx * fx;
  T pzz = (zp2 + zm2 - T(2) * c) * fz * fz;
  T pyy = (yp2 + ym2 - T(2) * c) * fy * fy;
  T pxy = (xy[0] + xy[1] - xy[2] - xy[3]) * fx * fy;
  T pxz = (xz[0] + xz[1] - xz[2] - xz[3]) * fx * fz;
  T pyz = (yz[0] + yz[1] - yz[2] - yz[3]) * fy * fz;
  T g2 = px * px + py * py + pz * pz;
  if (g2 < T(1.e-5)) return T(0);
  T curv = pxx * py * py + pyy * px * px - T(2) * pxy * px * py
         + pxx * pz * pz + pzz * px * px - T(2) * pxz * px * pz
         + pyy * pz * pz + pzz * py * py - T(2) * pyz * py * pz;
  return curv / g2;
}

// Smoothed Heaviside (reference: src/utilities/qss_util3d.f:450-491)
template <typename T>
__device__ __forceinline__ T heaviside(T x, T eps) {
  const T pi = T(3.14159265358979323846), one_over_pi = T(0.31830988618379);
  auto basic = [](T v) { return v < T(0) ? T(0) : (v > T(0) ? T(1) : T(0.5)); };
  T v1 = basic(x + eps), v2 = basic(eps - x), v3 = basic(x - eps);
  return (T(0.5) + x / (T(2) * eps) + (T(0.5) * one_over_pi) * sin((pi * x) / eps)) * v1 * v2 + v3;
}

}  // namespace lsm
