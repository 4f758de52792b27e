// Host-side check of weno5_face_adj (lsm_adjoint.h), the reverse-mode
// derivative of the fast WENO5 face reconstruction, against central
// differences of weno5_face itself in double precision, and of its single-
// precision instantiation against its double one on the same (float)
// inputs.  Windows have differences
// of magnitude 1e-3 .. 1e2 (no switch of the largest difference or of the
// flat-window test within the difference step).  For random seeds (sm, sp)
// prints "<worst relative error FD vs adjoint, double> <worst relative
// error float vs double adjoint> <flat-window error>".
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <random>
#include "lsm_adjoint.h"

using lsm::FaceScale;
using lsm::face_scale;
using lsm::weno5_face;
using lsm::weno5_face_adj;

static double objective(const double* q, const FaceScale<double>& sc, double sm, double sp) {
  auto f = weno5_face(q, sc);
  return sm * f.minus_right + sp * f.plus_left;
}

int main() {
  std::mt19937 g(2024);
  std::uniform_real_distribution<double> u(-1, 1), e(-3, 2);
  double worst_fd = 0, worst_f32 = 0;
  for (double dx : {1.0, 0.04}) {
    const FaceScale<double> sc = face_scale(dx);
    const FaceScale<float> scf = face_scale((float)dx);
    for (int t = 0; t < 20000; ++t) {
      double q[6], scale = std::pow(10.0, e(g));
      q[0] = u(g);
      for (int i = 1; i < 6; ++i) q[i] = q[i - 1] + scale * u(g);
      for (int i = 0; i < 6; ++i) q[i] = (double)(float)q[i];  // inputs both precisions hold
      const double sm = u(g), sp = u(g);
      double dq[6] = {0, 0, 0, 0, 0, 0};
      weno5_face_adj(q, sc, sm, sp, dq);
      // central differences; skip windows whose largest |difference| is
      // within the step of another (the derivative switches there)
      double r[5], sorted[5];
      for (int i = 0; i < 5; ++i) r[i] = std::fabs(q[i + 1] - q[i]);
      std::copy(r, r + 5, sorted);
      std::sort(sorted, sorted + 5);
      const double h = 1e-7 * scale;
      if (sorted[4] - sorted[3] < 10 * h) continue;
      double nrm = 0, err = 0;
      for (int i = 0; i < 6; ++i) {
        double qp[6], qm[6];
        std::copy(q, q + 6, qp);
        std::copy(q, q + 6, qm);
        qp[i] += h;
        qm[i] -= h;
        const double fd = (objective(qp, sc, sm, sp) - objective(qm, sc, sm, sp)) / (2 * h);
        err = std::max(err, std::fabs(fd - dq[i]));
        nrm = std::max(nrm, std::fabs(dq[i]));
      }
      if (nrm > 0) worst_fd = std::max(worst_fd, err / nrm);
      float qf[6], dqf[6] = {0, 0, 0, 0, 0, 0};
      for (int i = 0; i < 6; ++i) qf[i] = (float)q[i];
      weno5_face_adj(qf, scf, (float)sm, (float)sp, dqf);
      double ef = 0;
      for (int i = 0; i < 6; ++i) ef = std::max(ef, std::fabs(dqf[i] - dq[i]));
      if (nrm > 0) worst_f32 = std::max(worst_f32, ef / nrm);
    }
  }
  // flat window: the reconstruction is the middle candidate, a linear map
  const FaceScale<double> sc = face_scale(1.0);
  double q[6] = {0.5, 0.5, 0.5, 0.5, 0.5, 0.5}, dq[6] = {0, 0, 0, 0, 0, 0};
  weno5_face_adj(q, sc, 1.0, 0.0, dq);  // pm2 = -r1/6 + 5 r2/6 + r3/3
  const double want[6] = {0, 1.0 / 6, -1.0 / 6 - 5.0 / 6, 5.0 / 6 - 1.0 / 3, 1.0 / 3, 0};
  double flat = 0;
  for (int i = 0; i < 6; ++i) flat = std::max(flat, std::fabs(dq[i] - want[i]));
  // reference-order WENO5 of a 7-point line (both one-sided values)
  double worst_line = 0;
  for (int t = 0; t < 20000; ++t) {
    double p[7], scale = std::pow(10.0, e(g));
    p[0] = u(g);
    for (int i = 1; i < 7; ++i) p[i] = p[i - 1] + scale * u(g);
    const double inv = t % 2 ? 1.0 : 25.0, sp = u(g), sm = u(g);
    double dp[7] = {0, 0, 0, 0, 0, 0, 0};
    lsm::weno5_line_adj(p, inv, sp, sm, dp);
    // skip lines whose largest v^2 in either window is within the step of another
    double d[6];
    for (int i = 0; i < 6; ++i) d[i] = std::fabs(p[i + 1] - p[i]);
    double a[5] = {d[1], d[2], d[3], d[4], d[5]}, b[5] = {d[0], d[1], d[2], d[3], d[4]};
    std::sort(a, a + 5);
    std::sort(b, b + 5);
    const double h = 1e-7 * scale;
    if (a[4] - a[3] < 10 * h || b[4] - b[3] < 10 * h) continue;
    double err = 0, nrm = 0;
    for (int i = 0; i < 7; ++i) {
      double pp[7], pm[7], fp, fm, gp, gm;
      std::copy(p, p + 7, pp);
      std::copy(p, p + 7, pm);
      pp[i] += h;
      pm[i] -= h;
      lsm::weno5(pp, inv, fp, fm);
      lsm::weno5(pm, inv, gp, gm);
      const double fd = (sp * (fp - gp) + sm * (fm - gm)) / (2 * h);
      err = std::max(err, std::fabs(fd - dp[i]));
      nrm = std::max(nrm, std::fabs(dp[i]));
    }
    if (nrm > 0) worst_line = std::max(worst_line, err / nrm);
  }
  printf("%.3e %.3e %.3e %.3e\n", worst_fd, worst_f32, flat, worst_line);
  return 0;
}
