// lsm_apps.hip -- the computations of the reference's pre- and post-processing
// programs on the GPU: the level-set mask of a segmented image
// (createMaskUBC), the initial non-wetting phase (createInitFromMask) and the
// percolating lens between two saturation steps (makeDiffSteps).  File
// handling stays with the caller (writeGridToBinaryFile, writeDataArrayQSS,
// writeDataArrayRaw).
#include <stdio.h>
#include <stdlib.h>
#include <stdexcept>
#include <string>
#include "lsm_coeff.h"
#include "lsm_session.h"
#include "lsmpqs_hip.h"

namespace lsm {
namespace {

#define XCHECK(x)                                                                \
  do {                                                                           \
    hipError_t e_ = (x);                                                         \
    if (e_ != hipSuccess)                                                        \
      throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

// device buffers freed on scope exit
struct DevBuf {
  void* p = nullptr;
  explicit DevBuf(size_t bytes) { XCHECK(hipMalloc(&p, bytes ? bytes : 1)); }
  ~DevBuf() { (void)hipFree(p); }
  template <typename T> T* as() const { return (T*)p; }
};

int blocks_for(int64_t n) {
  int64_t b = (n + 255) / 256;
  return (int)(b > 16384 ? 16384 : (b < 1 ? 1 : b));
}

// mask = +dx where the image holds 255 (grain), -dx elsewhere, over the fill
// box; the image is the fill box, x fastest (reference:
// src/applications/createMaskUBC.c:100-113)
__global__ void image_mask_kernel(Box bx, real* mask, const unsigned char* image, real dx) {
  FillIter it(bx);
  for (int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; f < it.n;
       f += (int64_t)gridDim.x * blockDim.x) {
    int i, j, k;
    it.ijk(f, i, j, k);
    mask[bx.at(i, j, k)] = image[f] == 255 ? dx : -dx;
  }
}

// data = max(mask, data) (IMPOSE_MASK, reference: src/utilities/qss_macros.h:79-86)
__global__ void max_mask_kernel(real* data, const real* mask, int64_t n) {
  for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < n;
       l += (int64_t)gridDim.x * blockDim.x) {
    const real m = mask[l], v = data[l];
    data[l] = (m > v) ? m : v;
  }
}

// diff = max(-a, b - dx/2); bin = diff < 0, with the inlet (ilo-1) and
// outlet (ihi+1) ghost slabs of the fill rows set to 1 (reference:
// src/applications/makeDiffSteps.c:84-120)
__global__ void diff_kernel(Box bx, real* diff, uint8_t* bin, const real* a, const real* b,
                            real dx) {
  const int64_t n = bx.size();
  for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < n;
       l += (int64_t)gridDim.x * blockDim.x) {
    const real m = -a[l];
    real p = b[l];
    p -= real(0.5) * dx;
    const real d = (m > p) ? m : p;
    diff[l] = d;
    const int64_t t = l / bx.nx;
    const int i = (int)(l - t * bx.nx), j = (int)(t % bx.ny), k = (int)(t / bx.ny);
    const bool slab = (i == bx.ilo - 1 || i == bx.ihi + 1) && j >= bx.jlo && j <= bx.jhi &&
                      k >= bx.klo && k <= bx.khi;
    bin[l] = (slab || d < real(0)) ? 1 : 0;
  }
}

// out = diff with the sign flipped where diff <= 0 outside the component
// `main_label` (labels: root index on foreground, -1 on background, which
// the reference's connectivity array holds as 0)
__global__ void lens_kernel(real* out, const real* diff, const int* labels, int main_label,
                            int64_t n) {
  for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < n;
       l += (int64_t)gridDim.x * blockDim.x) {
    const int lab = labels[l] >= 0 ? labels[l] : 0;
    real v = diff[l];
    if ((lab != main_label) & (v <= real(0))) v *= real(-1);
    out[l] = v;
  }
}

}  // namespace
}  // namespace lsm

using namespace lsm;

extern "C" int lsmpqs_mask_from_image(const unsigned char* image, int nx, int ny, int nz,
                                      Options* o, Grid** grid_out, QSSLIB_REAL** mask_out) {
  set_error("");
  Grid* g = nullptr;
  QSSLIB_REAL* mask = nullptr;
  try {
    if (!image || !o || !grid_out || !mask_out)
      throw std::runtime_error("lsmpqs_mask_from_image: image, options and outputs required");
    if (nx < 4 || ny < 4 || nz < 4)
      throw std::runtime_error("lsmpqs_mask_from_image: 3D images of at least 4 voxels per axis");
    const int num_gc = 3;
    int n[3] = {nx, ny, nz};
    QSSLIB_REAL x_lo[3], x_hi[3];
    for (int a = 0; a < 3; ++a) {
      x_lo[a] = 0 + num_gc * o->dx;
      x_hi[a] = x_lo[a] + n[a] * o->dx;
    }
    g = createGridSetGridDims(3, n, x_lo, x_hi,
                              (QSSLIB_SPATIAL_DERIVATIVE_ACCURACY_TYPE)o->accuracy_id);
    if (!g) throw std::runtime_error("lsmpqs_mask_from_image: grid creation failed");
    const Box bx = box_of(g);
    if (bx.ilo != num_gc || bx.jlo != num_gc || bx.klo != num_gc ||
        bx.ihi - bx.ilo + 1 != nx || bx.jhi - bx.jlo + 1 != ny || bx.khi - bx.klo + 1 != nz)
      throw std::runtime_error(
          "lsmpqs_mask_from_image: the image maps onto a fill box starting at 3 ghost layers "
          "(options->accuracy_id must give them)");
    const int64_t npts = bx.size(), nimg = (int64_t)nx * ny * nz;
    mask = (QSSLIB_REAL*)calloc((size_t)npts, sizeof(QSSLIB_REAL));
    if (!mask) throw std::runtime_error("lsmpqs_mask_from_image: out of host memory");
    {
      DevBuf d_img(nimg), d_mask(npts * sizeof(real));
      XCHECK(hipMemcpy(d_img.p, image, nimg, hipMemcpyHostToDevice));
      XCHECK(hipMemset(d_mask.p, 0, npts * sizeof(real)));
      hipLaunchKernelGGL(image_mask_kernel, dim3(blocks_for(nimg)), dim3(256), 0, 0, bx,
                         d_mask.as<real>(), d_img.as<const unsigned char>(), o->dx);
      XCHECK(hipGetLastError());
      XCHECK(hipMemcpy(mask, d_mask.p, npts * sizeof(real), hipMemcpyDeviceToHost));
    }
    // subcell-fix reinitialization on the GPU (reference: qss_general_util.c:261-306
    // runs the same iteration as reinitializeSubcellFix3d)
    reinitializeSubcellFix3d(mask, g, o);
    const std::string err = last_error();
    session_release((const QSS_DataArrays*)mask);
    if (!err.empty()) throw std::runtime_error(err);
    *grid_out = g;
    *mask_out = mask;
    return 0;
  } catch (const std::exception& e) {
    set_error(e.what());
    fprintf(stderr, "lsmpqs_hip: %s\n", e.what());
    free(mask);
    if (g) destroyGrid(g);
    return -2;
  }
}

extern "C" int lsmpqs_init_from_mask(const QSSLIB_REAL* mask, Grid* g, Options* o,
                                     QSSLIB_REAL* data) {
  set_error("");
  try {
    if (!mask || !g || !o || !data)
      throw std::runtime_error("lsmpqs_init_from_mask: mask, grid, options and output required");
    if (g->num_dims != 3) throw std::runtime_error("lsmpqs_init_from_mask: 3D grids only");
    QSSLIB_REAL radius, cx, cy, cz;
    if (o->center_inlet) {  // the x centre from the y bounds, as the reference has it
      radius = 90 * g->dx[0];
      cy = 0.5 * (g->x_lo[1] + g->x_hi[1]);
      cx = 0.5 * (g->x_lo[1] + g->x_hi[1]);
      cz = 0.5 * (g->x_lo[2] + g->x_hi[2]);
    } else {
      radius = 0.7 * (g->x_hi[0] - g->x_lo[0]);
      cy = 0.5 * (g->x_lo[1] + g->x_hi[1]);
      cx = (g->x_lo[0]) - radius * 0.7;
      cz = 0.5 * (g->x_lo[2] + g->x_hi[2]);
    }
    const Box bx = box_of(g);
    const int64_t n = bx.size();
    DevBuf d_mask(n * sizeof(real)), d_data(n * sizeof(real));
    XCHECK(hipMemcpy(d_mask.p, mask, n * sizeof(real), hipMemcpyHostToDevice));
    XCHECK(launch_sphere(d_data.as<real>(), bx, g->x_lo_ghostbox, g->dx, cx, cy, cz, radius, -1, 0));
    hipLaunchKernelGGL(max_mask_kernel, dim3(blocks_for(n)), dim3(256), 0, 0, d_data.as<real>(),
                       d_mask.as<const real>(), n);
    XCHECK(hipGetLastError());
    XCHECK(hipMemcpy(data, d_data.p, n * sizeof(real), hipMemcpyDeviceToHost));
    return 0;
  } catch (const std::exception& e) {
    set_error(e.what());
    fprintf(stderr, "lsmpqs_hip: %s\n", e.what());
    return -2;
  }
}

extern "C" int lsmpqs_diff_steps(const QSSLIB_REAL* step_a, const QSSLIB_REAL* step_b, Grid* g,
                                 QSSLIB_REAL* diff, QSSLIB_REAL* connected_only) {
  set_error("");
  try {
    if (!step_a || !step_b || !g || !diff)
      throw std::runtime_error("lsmpqs_diff_steps: two steps, a grid and the diff output required");
    if (g->num_dims != 3) throw std::runtime_error("lsmpqs_diff_steps: 3D grids only");
    const Box bx = box_of(g);
    const int64_t n = bx.size();
    if (n >= ((int64_t)1 << 31))
      throw std::runtime_error("lsmpqs_diff_steps: ghost box exceeds 2^31 voxels (int labels)");
    if (bx.ilo < 1 || bx.ihi > bx.nx - 2)
      throw std::runtime_error("lsmpqs_diff_steps: needs a ghost layer on both x faces");
    DevBuf d_a(n * sizeof(real)), d_b(n * sizeof(real)), d_diff(n * sizeof(real));
    DevBuf d_bin(n), d_lab(n * sizeof(int)), d_root(n * sizeof(int));
    XCHECK(hipMemcpy(d_a.p, step_a, n * sizeof(real), hipMemcpyHostToDevice));
    XCHECK(hipMemcpy(d_b.p, step_b, n * sizeof(real), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(diff_kernel, dim3(blocks_for(n)), dim3(256), 0, 0, bx, d_diff.as<real>(),
                       d_bin.as<uint8_t>(), d_a.as<const real>(), d_b.as<const real>(), g->dx[0]);
    XCHECK(hipGetLastError());
    XCHECK(launch_label_components(bx, d_bin.as<const uint8_t>(), d_lab.as<int>(),
                                   d_root.as<int>(), 26, 0));
    XCHECK(hipMemcpy(diff, d_diff.p, n * sizeof(real), hipMemcpyDeviceToHost));
    // inlet and outlet voxels at mid j, k (reference: makeDiffSteps.c:122-129)
    const int j = (int)(0.5 * (bx.jlo + bx.jhi)), k = (int)(0.5 * (bx.klo + bx.khi));
    int lab[2];
    XCHECK(hipMemcpy(&lab[0], d_lab.as<int>() + bx.at(bx.ilo - 1, j, k), sizeof(int),
                     hipMemcpyDeviceToHost));
    XCHECK(hipMemcpy(&lab[1], d_lab.as<int>() + bx.at(bx.ihi + 1, j, k), sizeof(int),
                     hipMemcpyDeviceToHost));
    const int in_val = lab[0] >= 0 ? lab[0] : 0, out_val = lab[1] >= 0 ? lab[1] : 0;
    if (in_val != out_val) return 0;
    if (connected_only) {
      hipLaunchKernelGGL(lens_kernel, dim3(blocks_for(n)), dim3(256), 0, 0, d_a.as<real>(),
                         d_diff.as<const real>(), d_lab.as<const int>(), in_val, n);
      XCHECK(hipGetLastError());
      XCHECK(hipMemcpy(connected_only, d_a.p, n * sizeof(real), hipMemcpyDeviceToHost));
    }
    return 1;
  } catch (const std::exception& e) {
    set_error(e.what());
    fprintf(stderr, "lsmpqs_hip: %s\n", e.what());
    return -2;
  }
}
