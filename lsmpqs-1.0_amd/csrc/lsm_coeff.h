// lsm_coeff.h -- fill-box iteration and the mask-derived coefficient terms
// shared by the per-outer-step kernels and the adjoint kernels.
#pragma once
#include "lsm_kernels.h"

namespace lsm {

// Linear index over the fill box -> (i, j, k), x fastest.
struct FillIter {
  Box bx;
  int nxf, nyf;
  int64_t n;
  __host__ __device__ FillIter(const Box& b) : bx(b) {
    nxf = b.ihi - b.ilo + 1;
    nyf = b.jhi - b.jlo + 1;
    n = (int64_t)nxf * nyf * (b.khi - b.klo + 1);
  }
  __device__ void ijk(int64_t f, int& i, int& j, int& k) const {
    int64_t t = f / nxf;
    i = bx.ilo + (int)(f - t * nxf);
    j = bx.jlo + (int)(t % nyf);
    k = bx.klo + (int)(t / nyf);
  }
};

// Heaviside weights, |grad mask| and the smoothed sign of the mask of one
// voxel (reference: qss_spatial_derivatives3d.f:13-87,178-298).
__device__ __forceinline__ void coeff_common(real m, real mx, real my, real mz, real dx, real& h1,
                                             real& h2, real& ag, real& s) {
  real eps_var = real(1.5) * dx;
  h1 = heaviside(-m, eps_var);
  h2 = heaviside(m, eps_var);
  ag = mx * mx + my * my + mz * mz;
  ag = sqrt(ag);
  real x = m * m + ag * dx * ag * dx;
  x = sqrt(x);
  s = m / x;
}

}  // namespace lsm
