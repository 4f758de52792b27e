// lsm_adjoint.h -- reverse-mode derivatives of the WENO5 reconstructions of
// lsm_device.h (the fast face form weno5_face and the reference-order
// weno5), for the adjoint's transposed right-hand side.  Host-callable so
// they can be checked on the CPU.
#pragma once
#include <math.h>
#include "lsm_device.h"

namespace lsm {

// dq[0..5] += sm * d(minus_right)/dq + sp * d(plus_left)/dq for the window
// q[0..5] = phi(i-2..i+3), with the forward's discrete choices (flat-window
// test, the largest-difference operand) taken as the forward takes them.
// The forward is restated in working precision; rounding differences from
// the forward's contracted arithmetic do not matter to a derivative, so
// divisions by constants are products with their reciprocals and the three
// divisions by data are reciprocals (fast_rcp: the hardware reciprocal in
// single precision).
template <typename T>
__host__ __device__ inline void weno5_face_adj(const T* q, const FaceScale<T>& sc, T sm, T sp,
                                               T* dq) {
  const T r[5] = {q[1] - q[0], q[2] - q[1], q[3] - q[2], q[4] - q[3], q[5] - q[4]};
  T dr[5] = {0, 0, 0, 0, 0};
  const T dA_m = sm * sc.inv, dA_p = sp * sc.inv;  // adjoints of the bracketed sums
  T dpm2 = dA_m, dpp2 = dA_p;
  // m = fmax(fmax(fmax(|r0|, |r1|), fmax(|r2|, |r3|)), |r4|), first operand on ties
  T a01 = fabs(r[0]) >= fabs(r[1]) ? fabs(r[0]) : fabs(r[1]);
  int j01 = fabs(r[0]) >= fabs(r[1]) ? 0 : 1;
  T a23 = fabs(r[2]) >= fabs(r[3]) ? fabs(r[2]) : fabs(r[3]);
  int j23 = fabs(r[2]) >= fabs(r[3]) ? 2 : 3;
  T a03 = a01 >= a23 ? a01 : a23;
  int j03 = a01 >= a23 ? j01 : j23;
  const T m = a03 >= fabs(r[4]) ? a03 : fabs(r[4]);
  const int jm = a03 >= fabs(r[4]) ? j03 : 4;
  const T i6 = T(1.0 / 6.0), i3 = T(1.0 / 3.0), i20 = T(1.0 / 20.0), i30 = T(1.0 / 30.0);
  if (m > sc.thr) {
    const T D1 = r[1] - r[0], D2 = r[2] - r[1], D3 = r[3] - r[2], D4 = r[4] - r[3];
    const T t1 = D2 - D1, t2 = D3 - D2, t3 = D4 - D3;
    const T u1 = T(3) * D2 - D1, u2 = D2 + D3, u3 = T(3) * D3 - D4;
    const T c13 = T(13) / T(12);
    const T S1 = c13 * (t1 * t1) + T(0.25) * (u1 * u1);
    const T S2 = c13 * (t2 * t2) + T(0.25) * (u2 * u2);
    const T S3 = c13 * (t3 * t3) + T(0.25) * (u3 * u3);
    const T im = fast_rcp(m), im2 = im * im, ce = sc.c1 * im;
    const T eps = T(1e-6) + ce * ce;
    const T a1 = S1 * im2 + eps, a2 = S2 * im2 + eps, a3 = S3 * im2 + eps;
    const T Q1 = a1 * a1, Q2 = a2 * a2, Q3 = a3 * a3;
    const T q12 = Q1 * Q2, q13 = Q1 * Q3, q23 = Q2 * Q3;
    const T g1 = t1 - t2, g2 = t2 - t3;
    const T U = q23 * g1, V = q12 * g2;
    const T mid = T(0.6) * q13;
    const T sum_m = T(0.1) * q23 + T(0.3) * q12 + mid;
    const T sum_p = T(0.1) * q12 + T(0.3) * q23 + mid;
    const T num_m = U * i30 + V * i20;
    const T num_p = -(U * i20 + V * i30);
    const T rm = fast_rcp(sum_m), rp = fast_rcp(sum_p);
    // outputs: (pm2 + num_m rm) inv, (pp2 + num_p rp) inv
    const T dnum_m = dA_m * rm, dsum_m = -dA_m * num_m * rm * rm;
    const T dnum_p = dA_p * rp, dsum_p = -dA_p * num_p * rp * rp;
    const T dU = dnum_m * i30 - dnum_p * i20;
    const T dV = dnum_m * i20 - dnum_p * i30;
    T dq23 = T(0.1) * dsum_m + T(0.3) * dsum_p + dU * g1;
    T dq12 = T(0.3) * dsum_m + T(0.1) * dsum_p + dV * g2;
    const T dq13 = T(0.6) * (dsum_m + dsum_p);
    const T dg1 = dU * q23, dg2 = dV * q12;
    const T dQ1 = dq12 * Q2 + dq13 * Q3;
    const T dQ2 = dq12 * Q1 + dq23 * Q3;
    const T dQ3 = dq13 * Q1 + dq23 * Q2;
    const T da1 = T(2) * a1 * dQ1, da2 = T(2) * a2 * dQ2, da3 = T(2) * a3 * dQ3;
    const T dS1 = da1 * im2, dS2 = da2 * im2, dS3 = da3 * im2;
    const T dim2 = da1 * S1 + da2 * S2 + da3 * S3;
    const T deps = da1 + da2 + da3;
    const T dim = (T(2) * ce * deps) * sc.c1 + T(2) * im * dim2;
    const T dm = -dim * im * im;
    const T dt1 = T(2) * c13 * t1 * dS1 + dg1;
    const T dt2 = T(2) * c13 * t2 * dS2 - dg1 + dg2;
    const T dt3 = T(2) * c13 * t3 * dS3 - dg2;
    const T du1 = T(0.5) * u1 * dS1, du2 = T(0.5) * u2 * dS2, du3 = T(0.5) * u3 * dS3;
    const T dD1 = -dt1 - du1;
    const T dD2 = dt1 - dt2 + T(3) * du1 + du2;
    const T dD3 = dt2 - dt3 + du2 + T(3) * du3;
    const T dD4 = dt3 - du3;
    dr[0] += -dD1;
    dr[1] += dD1 - dD2;
    dr[2] += dD2 - dD3;
    dr[3] += dD3 - dD4;
    dr[4] += dD4;
#pragma unroll
    for (int k = 0; k < 5; ++k)
      if (k == jm) dr[k] += dm * (r[k] < T(0) ? T(-1) : T(1));
  }
  // middle candidates: pm2 = -r1/6 + 5 r2/6 + r3/3, pp2 = -r3/6 + 5 r2/6 + r1/3
  dr[1] += -dpm2 * i6 + dpp2 * i3;
  dr[2] += T(5) / T(6) * (dpm2 + dpp2);
  dr[3] += dpm2 * i3 - dpp2 * i6;
  dq[0] -= dr[0];
  dq[1] += dr[0] - dr[1];
  dq[2] += dr[1] - dr[2];
  dq[3] += dr[2] - dr[3];
  dq[4] += dr[3] - dr[4];
  dq[5] += dr[4];
}

// dv[0..4] += s * d weno5_combine(v)/dv (the reference-order HJ-WENO5
// combination, lsm_device.h), with the largest v_k^2 in the epsilon taken as
// the forward takes it.
template <typename T>
__host__ __device__ inline void weno5_combine_adj(const T* v, T s, T* dv) {
  const T c13 = T(13.0 / 12.0), tiny = T(1.0e-35);
  // eps = 1e-6 max v_k^2 + 1e-35; tmax(a, b) = a > b ? a : b keeps the later
  // of equal operands
  T mx = v[0] * v[0];
  int jm = 0;
  for (int k = 1; k < 5; ++k) {
    const T sq = v[k] * v[k];
    if (!(mx > sq)) { mx = sq; jm = k; }
  }
  const T eps = T(1e-6) * mx + tiny;
  const T p1 = T(1.0 / 3.0) * v[0] - T(7.0 / 6.0) * v[1] + T(11.0 / 6.0) * v[2];
  const T p2 = -T(1.0 / 6.0) * v[1] + T(5.0 / 6.0) * v[2] + T(1.0 / 3.0) * v[3];
  const T p3 = T(1.0 / 3.0) * v[2] + T(5.0 / 6.0) * v[3] - T(1.0 / 6.0) * v[4];
  const T t1 = v[0] - T(2) * v[1] + v[2], u1 = v[0] - T(4) * v[1] + T(3) * v[2];
  const T t2 = v[1] - T(2) * v[2] + v[3], u2 = v[1] - v[3];
  const T t3 = v[2] - T(2) * v[3] + v[4], u3 = T(3) * v[2] - T(4) * v[3] + v[4];
  const T S1 = c13 * (t1 * t1) + T(0.25) * (u1 * u1);
  const T S2 = c13 * (t2 * t2) + T(0.25) * (u2 * u2);
  const T S3 = c13 * (t3 * t3) + T(0.25) * (u3 * u3);
  // four reciprocals instead of the thirteen divisions of the formulas below
  const T e1 = S1 + eps, e2 = S2 + eps, e3 = S3 + eps;
  const T ie1 = T(1) / e1, ie2 = T(1) / e2, ie3 = T(1) / e3;
  const T a1 = T(0.1) * (ie1 * ie1), a2 = T(0.6) * (ie2 * ie2), a3 = T(0.3) * (ie3 * ie3);
  const T A = a1 + a2 + a3;
  const T iA = T(1) / A, sA = s * iA;
  const T out = (a1 * p1 + a2 * p2 + a3 * p3) * iA;
  // out = sum a_k p_k / A: d/dp_k = a_k / A, d/da_k = (p_k - out) / A
  const T dp1 = sA * a1, dp2 = sA * a2, dp3 = sA * a3;
  const T da1 = sA * (p1 - out), da2 = sA * (p2 - out), da3 = sA * (p3 - out);
  // a_k = c_k / e_k^2: d/de_k = -2 a_k / e_k
  const T de1 = -T(2) * da1 * a1 * ie1, de2 = -T(2) * da2 * a2 * ie2, de3 = -T(2) * da3 * a3 * ie3;
  const T deps = de1 + de2 + de3;
  const T dt1 = T(2) * c13 * t1 * de1, du1 = T(0.5) * u1 * de1;
  const T dt2 = T(2) * c13 * t2 * de2, du2 = T(0.5) * u2 * de2;
  const T dt3 = T(2) * c13 * t3 * de3, du3 = T(0.5) * u3 * de3;
  dv[0] += T(1.0 / 3.0) * dp1 + dt1 + du1;
  dv[1] += -T(7.0 / 6.0) * dp1 - T(1.0 / 6.0) * dp2 - T(2) * dt1 - T(4) * du1 + dt2 + du2;
  dv[2] += T(11.0 / 6.0) * dp1 + T(5.0 / 6.0) * dp2 + T(1.0 / 3.0) * dp3 + dt1 + T(3) * du1 -
           T(2) * dt2 + dt3 + T(3) * du3;
  dv[3] += T(1.0 / 3.0) * dp2 + T(5.0 / 6.0) * dp3 + dt2 - du2 - T(2) * dt3 - T(4) * du3;
  dv[4] += -T(1.0 / 6.0) * dp3 + dt3 + du3;
#pragma unroll
  for (int k = 0; k < 5; ++k)
    if (k == jm) dv[k] += deps * T(1e-6) * T(2) * v[k];
}

// dp[0..6] += sp * d plus/dp + sm * d minus/dp for weno5(p, inv_d) of
// lsm_device.h (p = phi(-3..+3))
template <typename T>
__host__ __device__ inline void weno5_line_adj(const T* p, T inv_d, T sp, T sm, T* dp) {
  const T D[6] = {p[1] - p[0], p[2] - p[1], p[3] - p[2], p[4] - p[3], p[5] - p[4], p[6] - p[5]};
  // D1mm, D1m, D1, D1p, D1pp, D1ppp = D[0..5]
  T dD[6] = {0, 0, 0, 0, 0, 0};
  // plus = combine(D1ppp, D1pp, D1p, D1, D1m) * inv_d each
  const T vp[5] = {D[5] * inv_d, D[4] * inv_d, D[3] * inv_d, D[2] * inv_d, D[1] * inv_d};
  T dvp[5] = {0, 0, 0, 0, 0};
  weno5_combine_adj(vp, sp, dvp);
  dD[5] += dvp[0] * inv_d; dD[4] += dvp[1] * inv_d; dD[3] += dvp[2] * inv_d;
  dD[2] += dvp[3] * inv_d; dD[1] += dvp[4] * inv_d;
  // minus = combine(D1mm, D1m, D1, D1p, D1pp)
  const T vm[5] = {D[0] * inv_d, D[1] * inv_d, D[2] * inv_d, D[3] * inv_d, D[4] * inv_d};
  T dvm[5] = {0, 0, 0, 0, 0};
  weno5_combine_adj(vm, sm, dvm);
  for (int k = 0; k < 5; ++k) dD[k] += dvm[k] * inv_d;
  for (int k = 0; k < 6; ++k) {
    dp[k] -= dD[k];
    dp[k + 1] += dD[k];
  }
}

}  // namespace lsm
