// lsm_ccl.hip -- connected-component labelling on the GPU and the
// trapped-phase mask rebuild.
//
// Labelling is a concurrent union-find in two levels: inside 16x8x4 tiles in
// LDS, then across tile boundaries in global memory.  Each voxel hooks itself
// to its foreground neighbours (13 of the 26, the other 13 hook to it) by
// linking the smaller root under the larger one with atomicCAS; a final pass
// compresses every path.  Linking only ever points a root at a larger index,
// so each component ends rooted at its largest linear index, which is exactly
// the label the reference's sequential union-find produces
// (reference: src/connectivity/connectivity.c:48-83,148-217).  Labels are
// therefore bit-identical to the reference's, not just the same partition.
#include "lsm_kernels.h"

namespace lsm {

__device__ __forceinline__ int uf_find(int* L, int x) {
  volatile int* V = L;
  int p = V[x];
  while (p != x) {
    int gp = V[p];
    if (gp != p) V[x] = gp;  // path halving; gp is an ancestor of x
    x = p;
    p = gp;
  }
  return x;
}

__device__ __forceinline__ void uf_unite(int* L, int a, int b) {
  while (true) {
    a = uf_find(L, a);
    b = uf_find(L, b);
    if (a == b) return;
    if (a > b) { int t = a; a = b; b = t; }
    int old = atomicCAS(&L[a], a, b);
    if (old == a) return;
    a = old;
  }
}

// Pass 1: union-find inside one 16 x 8 x 4 tile in LDS.  Local indices are
// ordered like global linear indices (x fastest, then y, then z), so the
// local root -- the largest local index of the tile component -- is also its
// largest global index.  Every voxel leaves pointing straight at that root.
constexpr int CX = 16, CY = 8, CZ = 4, CN = CX * CY * CZ;

__device__ __forceinline__ int lds_find(volatile int* p, int x) {
  int q = p[x];
  while (q != x) { x = q; q = p[x]; }
  return x;
}

__global__ __launch_bounds__(CN) void ccl_local_kernel(Box bx, const uint8_t* bin, int* L, int* T,
                                                     int conn) {
  __shared__ int p[CN];
  int tx = threadIdx.x, ty = threadIdx.y, tz = threadIdx.z;
  int l = tx + CX * (ty + CY * tz);
  int i = blockIdx.x * CX + tx, j = blockIdx.y * CY + ty, k = blockIdx.z * CZ + tz;
  bool inside = i < bx.nx && j < bx.ny && k < bx.nz;
  int64_t g = inside ? bx.at(i, j, k) : 0;
  bool fg = inside && bin[g];
  p[l] = fg ? l : -1;
  __syncthreads();
  if (fg) {
    for (int dk = 0; dk <= 1; ++dk)
      for (int dj = (dk ? -1 : 0); dj <= 1; ++dj)
        for (int di = ((dk || dj) ? -1 : 1); di <= 1; ++di) {
          if (conn == 6 && (abs(di) + abs(dj) + abs(dk)) != 1) continue;
          int x2 = tx + di, y2 = ty + dj, z2 = tz + dk;
          if (x2 < 0 || x2 >= CX || y2 < 0 || y2 >= CY || z2 >= CZ) continue;
          int m = x2 + CX * (y2 + CY * z2);
          if (p[m] < 0) continue;
          int a = l, b = m;
          while (true) {
            a = lds_find(p, a);
            b = lds_find(p, b);
            if (a == b) break;
            if (a > b) { int t = a; a = b; b = t; }
            int old = atomicCAS(&p[a], a, b);
            if (old == a) break;
            a = old;
          }
        }
  }
  __syncthreads();
  if (inside) {
    if (fg) {
      int r = lds_find(p, l);
      int rx = r % CX, ry = (r / CX) % CY, rz = r / (CX * CY);
      int root = (int)bx.at(blockIdx.x * CX + rx, blockIdx.y * CY + ry, blockIdx.z * CZ + rz);
      L[g] = root;
      T[g] = root;
    } else {
      L[g] = -1;
      T[g] = -1;
    }
  }
}

// Pass 2: unite across tile boundaries only (the forward half of the
// neighbourhood, as in pass 1).  T holds every voxel's tile root and does not
// change; unions start from tile roots, so only root entries of L move.  A
// union (l, m) is skipped when the voxel before l along x (or, failing that,
// along y) lies in the same tiles on both sides and joins the same pair of
// tile roots: that predecessor's union, or its own predecessor's, performs
// it.  This removes the repeated unions of one root pair along a tile face.
__device__ __forceinline__ bool same_pair(const uint8_t* bin, const int* T, int64_t l, int64_t m,
                                          int64_t e, int a, int b) {
  return bin[l - e] && bin[m - e] && T[l - e] == a && T[m - e] == b;
}

__global__ void ccl_border_kernel(Box bx, const uint8_t* bin, int* L, const int* T, int conn) {
  int64_t n = bx.size();
  for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < n;
       l += (int64_t)gridDim.x * blockDim.x) {
    if (!bin[l]) continue;
    int64_t t = l / bx.nx;
    int i = (int)(l - t * bx.nx), j = (int)(t % bx.ny), k = (int)(t / bx.ny);
    int ti = i / CX, tj = j / CY, tk = k / CZ;
    bool edge = (i % CX == 0) || (i % CX == CX - 1) || (j % CY == 0) || (j % CY == CY - 1) ||
                (k % CZ == CZ - 1);
    if (!edge) continue;
    int a0 = T[l];
    for (int dk = 0; dk <= 1; ++dk)
      for (int dj = (dk ? -1 : 0); dj <= 1; ++dj)
        for (int di = ((dk || dj) ? -1 : 1); di <= 1; ++di) {
          if (conn == 6 && (abs(di) + abs(dj) + abs(dk)) != 1) continue;
          int i2 = i + di, j2 = j + dj, k2 = k + dk;
          if (i2 < 0 || i2 >= bx.nx || j2 < 0 || j2 >= bx.ny || k2 >= bx.nz) continue;
          if (i2 / CX == ti && j2 / CY == tj && k2 / CZ == tk) continue;
          int64_t m = bx.at(i2, j2, k2);
          if (!bin[m]) continue;
          int b0 = T[m];
          if (a0 == b0) continue;
          if (i % CX != 0 && i2 % CX != 0 && same_pair(bin, T, l, m, 1, a0, b0)) continue;
          if (j % CY != 0 && j2 % CY != 0 && same_pair(bin, T, l, m, bx.nx, a0, b0)) continue;
          uf_unite(L, a0, b0);
        }
  }
}

__global__ void ccl_compress_kernel(int* L, int64_t n) {
  for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < n;
       l += (int64_t)gridDim.x * blockDim.x) {
    if (L[l] >= 0) L[l] = uf_find(L, (int)l);
  }
}

static int grid_for(int64_t n) {
  int64_t b = (n + 255) / 256;
  return (int)(b > 16384 ? 16384 : (b < 1 ? 1 : b));
}

// labels: root (largest linear index of the component) on foreground, -1 on background
hipError_t launch_label_components(const Box& bx, const uint8_t* bin, int* L, int* T, int conn,
                                   hipStream_t st) {
  int64_t n = bx.size();
  dim3 tiles((bx.nx + CX - 1) / CX, (bx.ny + CY - 1) / CY, (bx.nz + CZ - 1) / CZ);
  hipLaunchKernelGGL(ccl_local_kernel, tiles, dim3(CX, CY, CZ), 0, st, bx, bin, L, T, conn);
  hipLaunchKernelGGL(ccl_border_kernel, dim3(grid_for(n)), dim3(256), 0, st, bx, bin, L,
                     (const int*)T, conn);
  hipLaunchKernelGGL(ccl_compress_kernel, dim3(grid_for(n)), dim3(256), 0, st, L, n);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Trapped-phase masks (reference: src/connectivity/trapComponents.c:156-375).
// The reference's connectivity value is the root label on foreground voxels
// and 0 elsewhere; a voxel is "disconnected" when that value is neither the
// value at the inlet/outlet index nor 0.

__device__ __forceinline__ int conn_value(const int* L, int64_t idx) {
  int v = L[idx];
  return v >= 0 ? v : 0;
}

// phi = merge(phi, mask_nw); bin = phi < 0, plus the inlet plane forced to 1
__global__ void trap_nw_bin_kernel(TrapArgs a) {
  const Box bx = a.bx;
  int64_t n = bx.size();
  int ci = (int)(0.5 * (bx.ilo + bx.ihi)), ck = (int)(0.5 * (bx.klo + bx.khi));
  for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < n;
       l += (int64_t)gridDim.x * blockDim.x) {
    real mn = a.mask_nw[l];
    real p = a.phi[l];
    if (mn > real(0)) { p = -mn; a.phi[l] = p; }
    uint8_t b = p < real(0) ? 1 : 0;
    int64_t t = l / bx.nx;
    int i = (int)(l - t * bx.nx), j = (int)(t % bx.ny), k = (int)(t / bx.ny);
    if (a.center_inlet) {
      if (i == ci && k == ck) b = 1;
    } else if (i == bx.ilo - 1 && j >= bx.jlo && j <= bx.jhi && k >= bx.klo && k <= bx.khi) {
      b = 1;
    }
    a.bin[l] = b;
  }
}

// mask_nw from the non-wetting labels; s = max(-phi, mask, mask_nw);
// bin = s < 0, plus the outlet plane(s) forced to 1
__global__ void trap_w_bin_kernel(TrapArgs a) {
  const Box bx = a.bx;
  int64_t n = bx.size();
  int main_val = conn_value(a.labels, a.nw_ind);
  for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < n;
       l += (int64_t)gridDim.x * blockDim.x) {
    int cv = conn_value(a.labels, l);
    bool disc = !(cv == main_val || cv == 0);
    real p = a.phi[l];
    real mn = (disc && p < real(0)) ? -p : a.fill;
    a.mask_nw[l] = mn;
    real s = -p;
    real m = a.mask[l];
    s = (m > s) ? m : s;
    s = (mn > s) ? mn : s;
    a.s[l] = s;
    uint8_t b = s < real(0) ? 1 : 0;
    int64_t t = l / bx.nx;
    int i = (int)(l - t * bx.nx), j = (int)(t % bx.ny), k = (int)(t / bx.ny);
    if (a.center_inlet) {
      if (k == bx.klo - 1 || k == bx.khi + 1 || i == bx.ilo - 1 || i == bx.ihi + 1) b = 1;
    } else if (i == bx.ihi + 1 && j >= bx.jlo && j <= bx.jhi && k >= bx.klo && k <= bx.khi) {
      b = 1;
    }
    a.bin[l] = b;
  }
}

__global__ void trap_w_mask_kernel(TrapArgs a) {
  int64_t n = a.bx.size();
  int main_val = conn_value(a.labels, a.w_ind);
  for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < n;
       l += (int64_t)gridDim.x * blockDim.x) {
    int cv = conn_value(a.labels, l);
    bool disc = !(cv == main_val || cv == 0);
    real s = a.s[l];
    a.mask_w[l] = (disc && s < real(0)) ? -s : a.fill;
  }
}

hipError_t launch_trap_components(const TrapArgs& a, int conn, hipStream_t st) {
  int64_t n = a.bx.size();
  hipLaunchKernelGGL(trap_nw_bin_kernel, dim3(grid_for(n)), dim3(256), 0, st, a);
  hipError_t e = launch_label_components(a.bx, a.bin, a.labels, a.troot, conn, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(trap_w_bin_kernel, dim3(grid_for(n)), dim3(256), 0, st, a);
  e = launch_label_components(a.bx, a.bin, a.labels, a.troot, conn, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(trap_w_mask_kernel, dim3(grid_for(n)), dim3(256), 0, st, a);
  return hipGetLastError();
}

}  // namespace lsm
