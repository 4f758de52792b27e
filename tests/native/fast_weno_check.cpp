// Host-side check of the single-precision fast HJ-WENO5 face reconstruction
// of lsm_device.h: over windows whose differences span 1e-40 .. 1e3 it must
// return finite values and agree with the textbook double-precision WENO5
// (weights 0.1/0.6/0.3, eps = 1e-6 max v^2 + 1e-35, v the differences
// divided by dx) to float accuracy, for dx = 1 and dx = 0.01.
// Prints "<non-finite count> <worst error / max|difference|>".
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <random>
#include "lsm_device.h"

static double weno_ref(double v1, double v2, double v3, double v4, double v5) {
  double eps = 1e-6 * std::max({v1 * v1, v2 * v2, v3 * v3, v4 * v4, v5 * v5}) + 1e-35;
  double p1 = v1 / 3 - 7 * v2 / 6 + 11 * v3 / 6, p2 = -v2 / 6 + 5 * v3 / 6 + v4 / 3;
  double p3 = v3 / 3 + 5 * v4 / 6 - v5 / 6;
  double t1 = v1 - 2 * v2 + v3, u1 = v1 - 4 * v2 + 3 * v3;
  double t2 = v2 - 2 * v3 + v4, u2 = v2 - v4;
  double t3 = v3 - 2 * v4 + v5, u3 = 3 * v3 - 4 * v4 + v5;
  double S1 = 13. / 12 * t1 * t1 + .25 * u1 * u1, S2 = 13. / 12 * t2 * t2 + .25 * u2 * u2;
  double S3 = 13. / 12 * t3 * t3 + .25 * u3 * u3;
  double a1 = .1 / ((S1 + eps) * (S1 + eps)), a2 = .6 / ((S2 + eps) * (S2 + eps));
  double a3 = .3 / ((S3 + eps) * (S3 + eps));
  return (a1 * p1 + a2 * p2 + a3 * p3) / (a1 + a2 + a3);
}

int main() {
  std::mt19937 g(12345);
  std::uniform_real_distribution<float> u(-1, 1), e(-40, 3);
  long bad = 0;
  double worst = 0;
  for (int it = 0; it < 2000000; ++it) {
    float sc = std::pow(10.0f, e(g));
    float q[6];
    q[0] = u(g);
    for (int m = 1; m < 6; ++m) q[m] = q[m - 1] + sc * u(g);
    const float dx = (it & 1) ? 1.0f : 0.01f;
    lsm::Weno5Face<float> f = lsm::weno5_face(q, lsm::face_scale(dx));
    if (!std::isfinite(f.minus_right) || !std::isfinite(f.plus_left)) { ++bad; continue; }
    double d[5];
    double scale = 0;
    for (int m = 0; m < 5; ++m) {
      d[m] = ((double)q[m + 1] - (double)q[m]) / dx;
      scale = std::max(scale, std::fabs(d[m]));
    }
    if (scale < 1e-17) continue;
    double mr = weno_ref(d[0], d[1], d[2], d[3], d[4]), pl = weno_ref(d[4], d[3], d[2], d[1], d[0]);
    worst = std::max({worst, std::fabs(f.minus_right - mr) / scale, std::fabs(f.plus_left - pl) / scale});
  }
  printf("%ld %.6e\n", bad, worst);
  return 0;
}
