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

// Row-wise iteration for grid-stride kernels launched with 64 x 4 blocks:
// each wave takes whole x rows (j, k) of the box, b * 4 + ty, then
// + gridDim.x * 4, ..., and its lanes the row's voxels 64 apart, so every
// wave reads contiguous memory and the row's (j, k) costs one 32-bit
// division per row instead of 64-bit divisions per voxel.  b is the block's
// XCD-contiguous number (xcd_tile): the blocks of one XCD take neighbouring
// rows and planes, so a stencil's j +- 1 and k +- 1 rows are read through
// one L2.  FILL: the fill box; otherwise the whole ghost box.
// f(idx, i, j, k).
constexpr int ROW_TX = 64, ROW_TY = 4;

template <bool FILL>
__host__ __device__ inline void row_extent(const Box& b, int& i0, int& nxr, int& j0, int& nyr,
                                           int& k0, int& nzr) {
  if (FILL) {
    i0 = b.ilo; nxr = b.ihi - b.ilo + 1;
    j0 = b.jlo; nyr = b.jhi - b.jlo + 1;
    k0 = b.klo; nzr = b.khi - b.klo + 1;
  } else {
    i0 = 0; nxr = b.nx; j0 = 0; nyr = b.ny; k0 = 0; nzr = b.nz;
  }
  if (nxr <= 0 || nyr <= 0 || nzr <= 0) nxr = nyr = nzr = 0;
}

// rows of the box (0 when it is empty)
template <bool FILL>
inline int64_t box_rows(const Box& b) {
  int i0, nxr, j0, nyr, k0, nzr;
  row_extent<FILL>(b, i0, nxr, j0, nyr, k0, nzr);
  return (int64_t)nyr * nzr;
}

// blocks of a row-wise launch: one block per 4 rows, at most cap
template <bool FILL>
inline int row_blocks(const Box& b, int cap) {
  int64_t g = (box_rows<FILL>(b) + ROW_TY - 1) / ROW_TY;
  return (int)(g < 1 ? 1 : (g > cap ? cap : g));
}

// lb, lg: the block's number and the block count of the launch over this box
// (blockIdx.x and gridDim.x, or a box's share of a launch over many boxes)
template <bool FILL, typename F>
__device__ __forceinline__ void for_rows_b(const Box& b, int lb, int lg, F&& f) {
  int i0, nxr, j0, nyr, k0, nzr;
  row_extent<FILL>(b, i0, nxr, j0, nyr, k0, nzr);
  const unsigned nrows = (unsigned)nyr * (unsigned)nzr;
  const int64_t plane = (int64_t)b.nx * b.ny;
  const unsigned blk = (unsigned)xcd_tile(lb, lg);
  for (unsigned r = blk * ROW_TY + threadIdx.y; r < nrows; r += (unsigned)lg * ROW_TY) {
    const unsigned kk = r / (unsigned)nyr, jj = r - kk * (unsigned)nyr;
    const int j = j0 + (int)jj, k = k0 + (int)kk;
    const int64_t base = (int64_t)k * plane + (int64_t)j * b.nx;
    for (int x = threadIdx.x; x < nxr; x += ROW_TX) {
      const int i = i0 + x;
      f(base + i, i, j, k);
    }
  }
}
template <bool FILL, typename F>
__device__ __forceinline__ void for_rows(const Box& b, F&& f) {
  for_rows_b<FILL>(b, (int)blockIdx.x, (int)gridDim.x, f);
}

// for_rows with U voxels of its row per lane and pass (64 apart): load(u,
// idx) for all U first, at indices clamped to the row so that every load is
// in bounds, then use(u, idx, i, j, k) for those inside the row.  A lane then
// has U voxels' loads in flight at once instead of one, which latency-bound
// stencil passes need to approach the HBM rate.
template <bool FILL, int U, typename Ld, typename F>
__device__ __forceinline__ void for_rows_u(const Box& b, int lb, int lg, Ld&& load, F&& use) {
  int i0, nxr, j0, nyr, k0, nzr;
  row_extent<FILL>(b, i0, nxr, j0, nyr, k0, nzr);
  const unsigned nrows = (unsigned)nyr * (unsigned)nzr;
  const int64_t plane = (int64_t)b.nx * b.ny;
  const unsigned blk = (unsigned)xcd_tile(lb, lg);
  for (unsigned r = blk * ROW_TY + threadIdx.y; r < nrows; r += (unsigned)lg * ROW_TY) {
    const unsigned kk = r / (unsigned)nyr, jj = r - kk * (unsigned)nyr;
    const int j = j0 + (int)jj, k = k0 + (int)kk;
    const int64_t base = (int64_t)k * plane + (int64_t)j * b.nx + i0;
    for (int x0 = threadIdx.x; x0 < nxr; x0 += ROW_TX * U) {
#pragma unroll
      for (int u = 0; u < U; ++u) load(u, base + min(x0 + u * ROW_TX, nxr - 1));
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int x = x0 + u * ROW_TX;
        if (x < nxr) use(u, base + x, i0 + x, j, k);
      }
    }
  }
}

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
