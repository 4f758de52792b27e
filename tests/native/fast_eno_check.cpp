// Host-side check of the single-precision fast HJ-ENO2 / HJ-ENO3 one-sided
// derivatives of lsm_device.h (eno2_fast_raw, eno3_fast_raw) against the
// Osher-Shu ENO construction written here from its definition (Newton
// divided differences, the stencil grown towards the smaller difference at
// each level; Osher and Shu, SIAM J. Numer. Anal. 28 (1991) 907-922, as in
// Osher and Fedkiw, Level Set Methods, sec. 3.3): the double instantiation
// must agree to rounding, the float one to float accuracy wherever no
// stencil choice is a near-tie.  Prints "<worst double> <worst float>
// <float samples compared>" (errors relative to the largest difference).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <random>
#include "lsm_device.h"

// one-sided derivative at point 3 of p[0..6] (unit spacing): side +1 takes
// the first difference to the right, -1 to the left; order 2 or 3; near_tie
// is set when a comparison was within rel of a tie
static double eno_os(const double* p, int side, int order, bool& near_tie, double rel) {
  auto d1 = [&](int m) { return p[m + 1] - p[m]; };            // at m + 1/2
  auto d2 = [&](int m) { return (p[m + 1] - 2 * p[m] + p[m - 1]) / 2; };  // at m
  auto d3 = [&](int m) { return (d2(m + 1) - d2(m)) / 3; };     // at m + 1/2
  const int i = 3;
  int k = side > 0 ? i : i - 1;
  double v = d1(k);
  // second level: D2 at k or k + 1
  double a = d2(k), b = d2(k + 1);
  if (std::fabs(std::fabs(a) - std::fabs(b)) <= rel * (std::fabs(a) + std::fabs(b))) near_tie = true;
  double c;
  int ks;
  if (std::fabs(a) < std::fabs(b)) { c = a; ks = k - 1; } else { c = b; ks = k; }
  v += c * (2 * (i - k) - 1);
  if (order == 3) {
    double a3 = d3(ks), b3 = d3(ks + 1);
    if (std::fabs(std::fabs(a3) - std::fabs(b3)) <= rel * (std::fabs(a3) + std::fabs(b3)))
      near_tie = true;
    double cs = std::fabs(a3) < std::fabs(b3) ? a3 : b3;
    const int r = i - ks;
    v += cs * (3 * r * r - 6 * r + 2);
  }
  return v;
}

int main() {
  std::mt19937 g(2024);
  std::uniform_real_distribution<double> u(-1, 1), e(-6, 3);
  double worst64 = 0, worst32 = 0;
  long n32 = 0;
  for (int it = 0; it < 1000000; ++it) {
    const double sc = std::pow(10.0, e(g));
    double p[7];
    float pf[7];
    p[0] = u(g);
    for (int m = 1; m < 7; ++m) p[m] = p[m - 1] + sc * (0.3 + u(g));
    for (int m = 0; m < 7; ++m) pf[m] = (float)p[m];
    for (int m = 0; m < 7; ++m) p[m] = pf[m];   // the same data in both precisions
    double scale = 0;
    for (int m = 0; m < 6; ++m) scale = std::max(scale, std::fabs(p[m + 1] - p[m]));
    for (int order = 2; order <= 3; ++order) {
      bool tie64 = false, tie32 = false;
      const double rp = eno_os(p, +1, order, tie64, 1e-12), rm = eno_os(p, -1, order, tie64, 1e-12);
      double dp, dm;
      float fp, fm;
      if (order == 2) {
        lsm::eno2_fast_raw(p, dp, dm);
        lsm::eno2_fast_raw(pf, fp, fm);
      } else {
        lsm::eno3_fast_raw(p, dp, dm);
        lsm::eno3_fast_raw(pf, fp, fm);
      }
      if (!tie64) worst64 = std::max({worst64, std::fabs(dp - rp) / scale, std::fabs(dm - rm) / scale});
      eno_os(p, +1, order, tie32, 1e-4);
      eno_os(p, -1, order, tie32, 1e-4);
      if (!tie32) {
        worst32 = std::max({worst32, std::fabs(fp - rp) / scale, std::fabs(fm - rm) / scale});
        ++n32;
      }
    }
  }
  printf("%.6e %.6e %ld\n", worst64, worst32, n32);
  return 0;
}
