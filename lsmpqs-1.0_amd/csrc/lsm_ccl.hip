// lsm_ccl.hip -- connected-component labelling on the GPU and the
// trapped-phase mask rebuild.
//
// Labelling is a concurrent union-find in two levels: inside 64x16x16 tiles
// in LDS, then across tile boundaries in global memory (three launches, one
// per axis of the tile grid).  Each voxel hooks itself
// to its foreground neighbours (13 of the 26, the other 13 hook to it) by
// linking the smaller root under the larger one with atomicCAS; a final pass
// compresses every path.  Linking only ever points a root at a larger index,
// so each component ends rooted at its largest linear index, which is exactly
// the label the reference's sequential union-find produces
// (reference: src/connectivity/connectivity.c:48-83,148-217).  Labels are
// therefore bit-identical to the reference's, not just the same partition.
#include <algorithm>
#include <cstdint>

#include "lsm_coeff.h"
#include "lsm_kernels.h"

namespace lsm {

// Entries that other threads change concurrently are read and written with
// relaxed atomics: every access goes to memory, and the accesses keep their
// LDS or global address space (volatile ones would be issued as flat
// instructions).
__device__ __forceinline__ int ld_dev(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_dev(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int ld_wg(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void st_wg(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ int uf_find(int* L, int x) {
  int p = ld_dev(&L[x]);
  while (p != x) {
    int gp = ld_dev(&L[p]);
    if (gp != p) st_dev(&L[x], gp);  // path halving; gp is an ancestor of x
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

// Pass 1: union-find inside one 64 x 16 x 16 tile in LDS (64 KB).  A block
// of 64 x 16 threads owns the tile (one wave per x row); each thread takes
// the 16 voxels of its z column.  Local indices are ordered like global linear indices (x fastest,
// then y, then z), so the local root -- the largest local index of the tile
// component -- is also its largest global index.  Every voxel leaves
// pointing straight at that root, in L (which pass 2 links further) and in T
// (which stays: the tile root of every voxel).
constexpr int CX = 64, CY = 16, CZ = 16, CN = CX * CY * CZ;
static_assert(CX == 64, "one wave per tile row");

// root of x with path halving (each visited entry is re-pointed at its
// grandparent, an ancestor, so concurrent finds and links stay consistent)
__device__ __forceinline__ int lds_find(int* p, int x) {
  int q = ld_wg(&p[x]);
  while (q != x) {
    int g = ld_wg(&p[q]);
    if (g != q) st_wg(&p[x], g);
    x = q;
    q = g;
  }
  return x;
}

__device__ __forceinline__ void lds_unite(int* p, int a, int b) {
  while (true) {
    a = lds_find(p, a);
    b = lds_find(p, b);
    if (a == b) return;
    if (a > b) { int t = a; a = b; b = t; }
    int old = atomicCAS(&p[a], a, b);
    if (old == a) return;
    a = old;
  }
}

// unions of voxel column (tx, ty) with its neighbours one step along
// (di, dj, dk), at the planes z of the bit mask `need`
__device__ __forceinline__ void unite_planes(int* p, unsigned need, int tx, int ty, int di, int dj,
                                             int dk) {
  while (need) {
    const int z = __builtin_ctz(need);
    need &= need - 1;
    const int l = tx + CX * (ty + CY * z);
    lds_unite(p, l, l + di + CX * (dj + CY * dk));
  }
}

// roots: [count, tile root, tile root, ...] with room for `cap` roots; every
// tile root is appended (when the list is full, the count still grows and
// the root-jump passes scan T instead)
// With all = false only the tile roots' entries of L are written (they are
// the only entries read before a full labelling's resolve pass).
// (cx, cy, cz): the tile's position in the tile grid
__device__ __forceinline__ void ccl_local_tile(const Box& bx, const uint8_t* bin, int* L, int* T,
                                               int conn, int* roots, int cap, bool all, int cx,
                                               int cy, int cz) {
  __shared__ int p[CN];
  // bit z of cm[ty + 1][tx + 1] is voxel (tx, ty, z) of the tile; the frame
  // of one column around the tile reads as background
  __shared__ unsigned short cm[CY + 2][CX + 2];
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int i = cx * CX + tx, j = cy * CY + ty, kb = cz * CZ;
  const bool col = i < bx.nx && j < bx.ny;
  const int plane = bx.nx * bx.ny;
  const int g0 = i + bx.nx * j;
  unsigned fgm = 0;  // bit z: voxel (i, j, kb + z) is foreground
#pragma unroll
  for (int z = 0; z < CZ; ++z) {
    const int l = tx + CX * (ty + CY * z);
    const bool fg = col && kb + z < bx.nz && bin[g0 + (kb + z) * plane];
    fgm |= fg ? (1u << z) : 0u;
    // x runs: every foreground voxel starts out pointing at the last voxel
    // of its run of foreground voxels along the tile row
    const unsigned long long row = __ballot(fg);
    const unsigned long long ones = ~(row >> tx);  // first zero bit at or after this voxel
    const int run = ones ? __ffsll((long long)ones) - 1 : CX - tx;
    p[l] = fg ? l + run - 1 : -1;
  }
  cm[ty + 1][tx + 1] = (unsigned short)fgm;
  for (int e = ty * CX + tx; e < (CY + 2) * (CX + 2); e += CX * CY) {
    const int r = e / (CX + 2), c = e - r * (CX + 2);
    if (r == 0 || r == CY + 1 || c == 0 || c == CX + 1) cm[r][c] = 0;
  }
  __syncthreads();
  // Unions with the forward neighbours of the 26- (6-) neighbourhood, all 16
  // planes of the column at once as bit masks.  The +x neighbour is already
  // joined by the run.  A union (l, l + (di, dj, dk)) with di != 0 is implied
  // when l + di is foreground (that voxel is joined to l by the run and to
  // the target by the pair with di = 0); one with di = 0, dj != 0, dk = 1 is
  // implied when l + dj (y) is foreground.  A union (l, m) with di >= 0 is
  // implied when l - 1 and m - 1 are both foreground: voxel l - 1 takes the
  // same-direction pair (l - 1, m - 1) (or skips it by these rules), and the
  // runs join l - 1 to l and m - 1 to m.  (With di = -1 the first rule
  // already covers it.)  Inside a run only the first voxel facing a run of
  // the neighbouring row then unites with it.
  unsigned C[3][3];  // C[a + 1][b + 1]: foreground planes of column (tx + a, ty + b)
#pragma unroll
  for (int b = 0; b < 3; ++b)
#pragma unroll
    for (int a = 0; a < 3; ++a) C[a][b] = cm[ty + b][tx + a];
  const unsigned xm = C[0][1], xp = C[2][1];
#pragma unroll
  for (int dk = 0; dk <= 1; ++dk)
#pragma unroll
    for (int dj = (dk ? -1 : 0); dj <= 1; ++dj)
#pragma unroll
      for (int di = ((dk || dj) ? -1 : 1); di <= 1; ++di) {
        if (dk == 0 && dj == 0) continue;  // +x: the run
        const bool face = (abs(di) + abs(dj) + abs(dk)) == 1;
        if (conn == 6 && !face) continue;
        unsigned need = fgm & (C[di + 1][dj + 1] >> dk);
        if (conn != 6) {
          if (di == 1) need &= ~xp;
          if (di == -1) need &= ~xm;
          if (di == 0 && dk == 1 && dj != 0) need &= ~C[1][dj + 1];
        }
        if (di >= 0) need &= ~(xm & (C[di][dj + 1] >> dk));
        unite_planes(p, need, tx, ty, di, dj, dk);
      }
  __syncthreads();
  if (!col) return;
  for (int z = 0; z < CZ && kb + z < bx.nz; ++z) {
    const int l = tx + CX * (ty + CY * z);
    const int g = g0 + (kb + z) * plane;
    int root = -1;
    if (fgm >> z & 1u) {
      int r = lds_find(p, l);
      int rx = r % CX, ry = (r / CX) % CY, rz = r / (CX * CY);
      root = (cx * CX + rx) + bx.nx * ((cy * CY + ry) + bx.ny * (kb + rz));
      if (r == l && roots) {
        const int q = atomicAdd(&roots[0], 1);
        if (q < cap) roots[1 + q] = g;
      }
    }
    if (all || root == g) L[g] = root;
    T[g] = root;
  }
}
__global__ __launch_bounds__(CX * CY) void ccl_local_kernel(Box bx, const uint8_t* bin, int* L,
                                                          int* T, int conn, int* roots, int cap,
                                                          bool all) {
  ccl_local_tile(bx, bin, L, T, conn, roots, cap, all, blockIdx.x, blockIdx.y, blockIdx.z);
}

// Pass 2: unions across tile boundaries.  A pair of neighbouring foreground
// voxels (l, m) in different tiles, with tile offset o = tile(m) - tile(l),
// is united by l when o is positive in (z, y, x) lexicographic order and by
// m otherwise, so every cross-tile pair is taken exactly once, and the voxel
// taking it lies on the +x, +y or +z face of its tile.  The pairs are taken
// in three launches by offset class -- o = (0, 0, 1) from the +x faces, then
// o = (0, 1, *) from the +y faces, then o = (1, *, *) from the +z faces --
// with a pass between them that points every tile root straight at its
// current root.  Each launch then only builds chains along one axis of the
// tile grid (at most its tile count long), where one launch over all pairs
// built chains across the whole grid, and every hop of a find is a
// device-coherent load.  One block per tile, one thread per face voxel.
// Each thread reads the tile roots of its candidate neighbours in one batch
// (T = -1 marks background), and the tile-root pairs go through a block-wide
// hash set in LDS, so each distinct pair of tile roots is united once per
// tile.  Unions start from the tile roots in T, so only root entries of L
// move.
constexpr int PAIR_SET = 2048;                  // slots of the per-tile pair set
constexpr unsigned long long NO_PAIR = ~0ull;

// true when (a, b) was not in the set yet (or the set is full: unite anyway)
__device__ __forceinline__ bool pair_insert(unsigned long long* set, int a, int b) {
  const unsigned long long key = ((unsigned long long)(unsigned)a << 32) | (unsigned)b;
  unsigned h = ((unsigned)a * 0x9E3779B1u) ^ ((unsigned)b * 0x85EBCA77u);
  for (int probe = 0; probe < PAIR_SET; ++probe) {
    unsigned slot = (h + probe) & (PAIR_SET - 1);
    unsigned long long old = atomicCAS(&set[slot], NO_PAIR, key);
    if (old == NO_PAIR) return true;
    if (old == key) return false;
  }
  return true;
}

// face voxels of one tile for offset class AX (0: +x face, 1: +y, 2: +z)
template <int AX>
constexpr int face_count() {
  return AX == 0 ? CY * CZ : (AX == 1 ? CX * CZ : CX * CY);
}

// face job e of tile blockIdx, offset class AX
template <int AX>
__device__ __forceinline__ void border_job(const Box& bx, int* L, const int* T, int conn,
                                           unsigned long long* set, int plane, int e, int cx,
                                           int cy, int cz) {
  int lx, ly, lz;
  if (AX == 0) { lx = CX - 1; ly = e % CY; lz = e / CY; }
  else if (AX == 1) { lx = e % CX; ly = CY - 1; lz = e / CX; }
  else { lx = e % CX; ly = e / CX; lz = CZ - 1; }
  const int i = cx * CX + lx, j = cy * CY + ly, k = cz * CZ + lz;
  if (i >= bx.nx || j >= bx.ny || k >= bx.nz) return;
  const int l = i + bx.nx * j + plane * k;
  const int a0 = T[l];
  if (a0 < 0) return;
  int nb[26];
#pragma unroll
  for (int q = 0; q < 26; ++q) {
    const int o = q < 13 ? q : q + 1;  // skip the centre of the 3x3x3 cube
    const int di = o % 3 - 1, dj = (o / 3) % 3 - 1, dk = o / 9 - 1;
    nb[q] = -1;
    if (conn == 6 && (abs(di) + abs(dj) + abs(dk)) != 1) continue;
    const int oz = (lz + dk) >= CZ ? 1 : ((lz + dk) < 0 ? -1 : 0);
    const int oy = (ly + dj) >= CY ? 1 : ((ly + dj) < 0 ? -1 : 0);
    const int ox = (lx + di) >= CX ? 1 : ((lx + di) < 0 ? -1 : 0);
    const bool cls = AX == 2 ? oz > 0 : (AX == 1 ? (oz == 0 && oy > 0)
                                                 : (oz == 0 && oy == 0 && ox > 0));
    if (!cls) continue;
    const int i2 = i + di, j2 = j + dj, k2 = k + dk;
    if (i2 < 0 || i2 >= bx.nx || j2 < 0 || j2 >= bx.ny || k2 < 0 || k2 >= bx.nz) continue;
    nb[q] = T[i2 + bx.nx * j2 + plane * k2];
  }
  // z faces: unite from the roots the tile roots pointed at after the last
  // root-jump pass (each an ancestor of its tile root, so a valid start for
  // the union).  The pairs of one tile face then collapse in the pair set to
  // the few pairs of components that meet there, instead of one union per
  // pair of tile-local components, and equal roots need no union: on the
  // last pass, where the largest components meet, every tile of a z layer
  // otherwise walks the same few root entries (512^3 pack: 320 -> 128 us;
  // the x and y passes, with mostly distinct roots, take longer with the
  // extra loads).
  const int ra = AX == 2 ? ld_dev(&L[a0]) : a0;
  int rb[26];
#pragma unroll
  for (int q = 0; q < 26; ++q)
    rb[q] = (nb[q] >= 0 && nb[q] != a0) ? (AX == 2 ? ld_dev(&L[nb[q]]) : nb[q]) : -1;
  int last = -1;
#pragma unroll
  for (int q = 0; q < 26; ++q) {
    const int b = rb[q];
    if (b < 0 || b == ra || b == last) continue;
    last = b;
    if (pair_insert(set, ra, b)) uf_unite(L, ra, b);
  }
}

// Threads per border block by axis; each thread takes face voxels a block
// stride apart.  A block's time is mostly the latency of its loads and
// union-find hops: for the x and y faces 256-thread blocks (eight tiles in
// flight per CU, 16 KB of LDS each) finish sooner than 1024-thread ones (two
// tiles per CU; 512^3 pack, y faces 131 -> 93 us), but the z faces, where
// the largest components meet, took 320 / 454 / 584 us with 1024- / 512- /
// 256-thread blocks (more tiles linking the same few roots at once).
template <int AX>
constexpr int border_threads() {
  return AX == 2 ? 1024 : 256;
}

template <int AX>
__device__ __forceinline__ void ccl_border_tile(const Box& bx, int* L, const int* T, int conn,
                                                int cx, int cy, int cz) {
  __shared__ unsigned long long set[PAIR_SET];
  for (int e = threadIdx.x; e < PAIR_SET; e += blockDim.x) set[e] = NO_PAIR;
  __syncthreads();
  const int plane = bx.nx * bx.ny;
  for (int e = threadIdx.x; e < face_count<AX>(); e += blockDim.x)
    border_job<AX>(bx, L, T, conn, set, plane, e, cx, cy, cz);
}
template <int AX>
__global__ __launch_bounds__(1024) void ccl_border_kernel(Box bx, int* L, const int* T, int conn) {
  ccl_border_tile<AX>(bx, L, T, conn, blockIdx.x, blockIdx.y, blockIdx.z);
}

// After each border launch: every tile root (T[v] == v) is pointed straight
// at its current root, from the root list of the local pass, or by a scan of
// T when there is no list or it overflowed.  Only tile-root entries are
// written, each by its own thread; a walk reads entries that no thread
// writes meanwhile except to another ancestor of the same root.
__device__ __forceinline__ void jump_root(int* L, int v) {
  int r = ld_dev(&L[v]);
  for (int q = ld_dev(&L[r]); q != r; q = ld_dev(&L[r])) r = q;
  st_dev(&L[v], r);
}
__device__ __forceinline__ void ccl_jump_body(int* L, const int* T, int n, const int* roots,
                                              int cap, int lb, int lg) {
  const int stride = lg * blockDim.x, first = lb * blockDim.x + threadIdx.x;
  const int count = roots ? roots[0] : 0;
  if (roots && count <= cap) {
    for (int q = first; q < count; q += stride) jump_root(L, roots[1 + q]);
  } else {
    for (int v = first; v < n; v += stride)
      if (T[v] == v) jump_root(L, v);
  }
}
__global__ void ccl_jump_kernel(int* L, const int* T, int n, const int* roots, int cap) {
  ccl_jump_body(L, T, n, roots, cap, blockIdx.x, gridDim.x);
}

// Full labels once every tile root points at its component's root: each
// voxel takes its tile root's entry (tile-root entries stay as they are).
__global__ void ccl_resolve_kernel(int* L, const int* T, int n) {
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    const int t = T[v];
    if (t >= 0 && t != v) L[v] = L[t];
  }
}

__global__ void ccl_reset_kernel(int* roots) { roots[0] = 0; }

constexpr int UNR = 4;  // voxels per thread and iteration of the trapping passes

static int grid_for(int64_t n) {
  int64_t b = (n + 255) / 256;
  return (int)(b > 16384 ? 16384 : (b < 1 ? 1 : b));
}
// the trapping passes take UNR voxels per thread and iteration
static int trap_grid(int64_t n) {
  int64_t b = (n + 256 * UNR - 1) / (256 * UNR);
  return (int)(b > 16384 ? 16384 : (b < 1 ? 1 : b));
}

// labels: root (largest linear index of the component) on foreground, -1 on
// background; with full = false only the tile roots' entries hold it (the
// label of voxel v is then L[T[v]]).
hipError_t launch_label_components(const Box& bx, const uint8_t* bin, int* L, int* T, int conn,
                                   hipStream_t st, int* roots, int64_t roots_cap, bool full) {
  int64_t n = bx.size();
  if (n >= ((int64_t)1 << 31)) return hipErrorInvalidValue;
  const int cap = roots ? (int)std::min<int64_t>(roots_cap, INT32_MAX - 1) : 0;
  dim3 tiles((bx.nx + CX - 1) / CX, (bx.ny + CY - 1) / CY, (bx.nz + CZ - 1) / CZ);
  const int gj = roots ? (int)std::min<int64_t>(std::max<int64_t>((cap + 255) / 256, 1), 16384)
                       : grid_for(n);
  if (roots) hipLaunchKernelGGL(ccl_reset_kernel, dim3(1), dim3(1), 0, st, roots);
  hipLaunchKernelGGL(ccl_local_kernel, tiles, dim3(CX, CY, 1), 0, st, bx, bin, L, T, conn, roots,
                     cap, full);
  hipLaunchKernelGGL(ccl_border_kernel<0>, tiles, dim3(border_threads<0>()), 0, st, bx, L,
                     (const int*)T, conn);
  hipLaunchKernelGGL(ccl_jump_kernel, dim3(gj), dim3(256), 0, st, L, (const int*)T, (int)n,
                     (const int*)roots, cap);
  hipLaunchKernelGGL(ccl_border_kernel<1>, tiles, dim3(border_threads<1>()), 0, st, bx, L,
                     (const int*)T, conn);
  hipLaunchKernelGGL(ccl_jump_kernel, dim3(gj), dim3(256), 0, st, L, (const int*)T, (int)n,
                     (const int*)roots, cap);
  hipLaunchKernelGGL(ccl_border_kernel<2>, tiles, dim3(border_threads<2>()), 0, st, bx, L,
                     (const int*)T, conn);
  hipLaunchKernelGGL(ccl_jump_kernel, dim3(gj), dim3(256), 0, st, L, (const int*)T, (int)n,
                     (const int*)roots, cap);
  if (full) hipLaunchKernelGGL(ccl_resolve_kernel, dim3(grid_for(n)), dim3(256), 0, st, L,
                               (const int*)T, (int)n);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Trapped-phase masks (reference: src/connectivity/trapComponents.c:156-375;
// the 2D inlet/outlet rules at :196-211 and :297-337).
// The reference's connectivity value is the root label on foreground voxels
// and 0 elsewhere; a voxel is "disconnected" when that value is neither the
// value at the inlet/outlet index nor 0.

// through the tile root: L holds the label at tile roots (and, after a full
// labelling, everywhere)
__device__ __forceinline__ int conn_value(const int* L, const int* T, int64_t idx) {
  int t = T[idx];
  return t >= 0 ? L[t] : 0;
}

// The trapping passes take UNR voxels per thread and iteration (stride
// apart), with all their loads issued before any value is used, so that a
// lane has UNR voxels' loads in flight instead of one (the two wetting-phase
// passes below read a label through the voxel's tile root, two dependent
// loads).

// phi = merge(phi, mask_nw); bin = phi < 0 (force_bin_kernel then sets the
// inlet voxels)
__device__ __forceinline__ void trap_nw_bin_body(const TrapArgs& a, int lb, int lg) {
  const int64_t n = a.bx.size();
  real* __restrict__ phi = a.phi;
  const real* __restrict__ im = a.impose_mask;
  const real* __restrict__ mnw = a.mask_nw;
  uint8_t* __restrict__ bin = a.bin;
  const int64_t stride = (int64_t)lg * blockDim.x;
  for (int64_t l0 = (int64_t)lb * blockDim.x + threadIdx.x; l0 < n; l0 += UNR * stride) {
    real p[UNR], m[UNR], mn[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const int64_t l = min(l0 + u * stride, n - 1);
      p[u] = phi[l];
      m[u] = im ? im[l] : real(0);
      mn[u] = mnw[l];
    }
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const int64_t l = l0 + u * stride;
      if (l >= n) break;
      real v = p[u];
      if (im && m[u] > v) v = m[u];  // phi = max(phi, mask) first (the var path's imposition)
      if (mn[u] > real(0)) v = -mn[u];
      // stored only where it changed: most voxels keep phi, and waves with
      // no change then write nothing
      if ((im || mn[u] > real(0)) && v != p[u]) phi[l] = v;
      bin[l] = v < real(0) ? 1 : 0;
    }
  }
}
__global__ void trap_nw_bin_kernel(TrapArgs a) { trap_nw_bin_body(a, blockIdx.x, gridDim.x); }

// mask_nw from the non-wetting labels; s = max(-phi, mask, mask_nw);
// bin = s < 0 (force_bin_kernel then sets the outlet voxels)
__device__ __forceinline__ void trap_w_bin_body(const TrapArgs& a, int lb, int lg) {
  const int64_t n = a.bx.size();
  const int* __restrict__ T = a.troot;
  const int* __restrict__ L = a.labels;
  const real* __restrict__ phi = a.phi;
  const real* __restrict__ mask = a.mask;
  real* __restrict__ mnw = a.mask_nw;
  real* __restrict__ sw = a.s;
  uint8_t* __restrict__ bin = a.bin;
  const int main_val = conn_value(L, T, a.nw_ind);
  const int64_t stride = (int64_t)lg * blockDim.x;
  for (int64_t l0 = (int64_t)lb * blockDim.x + threadIdx.x; l0 < n; l0 += UNR * stride) {
    int t[UNR];
    real p[UNR], m[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const int64_t l = l0 + u * stride;
      const bool in = l < n;
      p[u] = in ? phi[l] : real(0);
      m[u] = in ? mask[l] : real(0);
    }
    // the label is needed only inside the non-wetting phase (phi < 0), a
    // small part of the box for most of a drainage: elsewhere T is not read
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const int64_t l = l0 + u * stride;
      t[u] = (l < n && p[u] < real(0)) ? T[l] : -1;
    }
    int cv[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) cv[u] = t[u] >= 0 ? L[t[u]] : 0;
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const int64_t l = l0 + u * stride;
      if (l >= n) break;
      const bool disc = !(cv[u] == main_val || cv[u] == 0);
      const real mn = (disc && p[u] < real(0)) ? -p[u] : a.fill;
      mnw[l] = mn;
      real v = -p[u];
      v = (m[u] > v) ? m[u] : v;
      v = (mn > v) ? mn : v;
      sw[l] = v;
      bin[l] = v < real(0) ? 1 : 0;
    }
  }
}
__global__ void trap_w_bin_kernel(TrapArgs a) { trap_w_bin_body(a, blockIdx.x, gridDim.x); }

void ForceSet::add(const Box& bx, int i0, int i1, int j0, int j1, int k0, int k1) {
  i0 = std::max(i0, 0); j0 = std::max(j0, 0); k0 = std::max(k0, 0);
  i1 = std::min(i1, bx.nx - 1); j1 = std::min(j1, bx.ny - 1); k1 = std::min(k1, bx.nz - 1);
  if (i0 > i1 || j0 > j1 || k0 > k1) return;
  lo[n][0] = i0; lo[n][1] = j0; lo[n][2] = k0;
  hi[n][0] = i1; hi[n][1] = j1; hi[n][2] = k1;
  ++n;
}

// non-wetting phase: the inlet (the ghost plane i = ilo - 1 over the fill
// box's j, k; 2D the whole ghost column; with a centre inlet the line
// i = ci, k = ck, in 2D the point i = ci, j = cj)
ForceSet trap_nw_force(const TrapArgs& a) {
  const Box& b = a.bx;
  const int ci = (int)(0.5 * (b.ilo + b.ihi)), cj = (int)(0.5 * (b.jlo + b.jhi));
  const int ck = (int)(0.5 * (b.klo + b.khi));
  const int all = 1 << 30;
  ForceSet f{};
  if (a.dims2) {
    if (a.center_inlet) f.add(b, ci, ci, cj, cj, 0, all);
    else f.add(b, b.ilo - 1, b.ilo - 1, 0, all, 0, all);
  } else if (a.center_inlet) {
    f.add(b, ci, ci, 0, all, ck, ck);
  } else {
    f.add(b, b.ilo - 1, b.ilo - 1, b.jlo, b.jhi, b.klo, b.khi);
  }
  return f;
}

// wetting phase: the outlet (the ghost plane i = ihi + 1 over the fill box's
// j, k; with a centre inlet the ghost planes k = klo - 1, k = khi + 1,
// i = ilo - 1, i = ihi + 1; 2D the fill column i = ilo + 1, or with a centre
// inlet the four ghost lines)
ForceSet trap_w_force(const TrapArgs& a) {
  const Box& b = a.bx;
  const int all = 1 << 30;
  ForceSet f{};
  if (a.dims2) {
    if (a.center_inlet) {
      f.add(b, b.ilo - 1, b.ilo - 1, 0, all, 0, all);
      f.add(b, b.ihi + 1, b.ihi + 1, 0, all, 0, all);
      f.add(b, 0, all, b.jlo - 1, b.jlo - 1, 0, all);
      f.add(b, 0, all, b.jhi + 1, b.jhi + 1, 0, all);
    } else {
      f.add(b, b.ilo + 1, b.ilo + 1, b.jlo, b.jhi, 0, all);
    }
  } else if (a.center_inlet) {
    f.add(b, 0, all, 0, all, b.klo - 1, b.klo - 1);
    f.add(b, 0, all, 0, all, b.khi + 1, b.khi + 1);
    f.add(b, b.ilo - 1, b.ilo - 1, 0, all, 0, all);
    f.add(b, b.ihi + 1, b.ihi + 1, 0, all, 0, all);
  } else {
    f.add(b, b.ihi + 1, b.ihi + 1, b.jlo, b.jhi, b.klo, b.khi);
  }
  return f;
}

__device__ __forceinline__ void force_bin_body(uint8_t* bin, const Box& bx, const ForceSet& f,
                                               int lb, int lg) {
  for (int q = 0; q < f.n; ++q) {
    const unsigned nx = (unsigned)(f.hi[q][0] - f.lo[q][0] + 1);
    const unsigned ny = (unsigned)(f.hi[q][1] - f.lo[q][1] + 1);
    const unsigned cnt = nx * ny * (unsigned)(f.hi[q][2] - f.lo[q][2] + 1);
    for (unsigned t = lb * blockDim.x + threadIdx.x; t < cnt; t += lg * blockDim.x) {
      const unsigned r = t / nx, kk = r / ny;
      bin[bx.at(f.lo[q][0] + (int)(t - r * nx), f.lo[q][1] + (int)(r - kk * ny),
                f.lo[q][2] + (int)kk)] = 1;
    }
  }
}
__global__ void force_bin_kernel(uint8_t* bin, Box bx, ForceSet f) {
  force_bin_body(bin, bx, f, blockIdx.x, gridDim.x);
}

__device__ __forceinline__ void trap_w_mask_body(const TrapArgs& a, int lb, int lg) {
  const int64_t n = a.bx.size();
  const int* __restrict__ T = a.troot;
  const int* __restrict__ L = a.labels;
  const real* __restrict__ sw = a.s;
  real* __restrict__ mw = a.mask_w;
  const int main_val = conn_value(L, T, a.w_ind);
  const int64_t stride = (int64_t)lg * blockDim.x;
  for (int64_t l0 = (int64_t)lb * blockDim.x + threadIdx.x; l0 < n; l0 += UNR * stride) {
    int t[UNR];
    real v[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const int64_t l = l0 + u * stride;
      v[u] = l < n ? sw[l] : real(0);
    }
    // the label is needed only inside the wetting phase (s < 0): the solid
    // and the non-wetting phase do not read T
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const int64_t l = l0 + u * stride;
      t[u] = (l < n && v[u] < real(0)) ? T[l] : -1;
    }
    int cv[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) cv[u] = t[u] >= 0 ? L[t[u]] : 0;
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const int64_t l = l0 + u * stride;
      if (l >= n) break;
      const bool disc = !(cv[u] == main_val || cv[u] == 0);
      mw[l] = (disc && v[u] < real(0)) ? -v[u] : a.fill;
    }
  }
}
__global__ void trap_w_mask_kernel(TrapArgs a) { trap_w_mask_body(a, blockIdx.x, gridDim.x); }

hipError_t launch_trap_components(const TrapArgs& a, int conn, hipStream_t st) {
  int64_t n = a.bx.size();
  hipLaunchKernelGGL(trap_nw_bin_kernel, dim3(trap_grid(n)), dim3(256), 0, st, a);
  hipLaunchKernelGGL(force_bin_kernel, dim3(256), dim3(256), 0, st, a.bin, a.bx, trap_nw_force(a));
  hipError_t e = launch_label_components(a.bx, a.bin, a.labels, a.troot, conn, st, a.roots,
                                         a.roots_cap, false);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(trap_w_bin_kernel, dim3(trap_grid(n)), dim3(256), 0, st, a);
  hipLaunchKernelGGL(force_bin_kernel, dim3(256), dim3(256), 0, st, a.bin, a.bx, trap_w_force(a));
  e = launch_label_components(a.bx, a.bin, a.labels, a.troot, conn, st, a.roots, a.roots_cap,
                              a.full_labels);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(trap_w_mask_kernel, dim3(trap_grid(n)), dim3(256), 0, st, a);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Trapping of many boxes at once (batched runs): the same passes, each one
// launch over every box (Multi: block -> box and its block there).

int trap_pass_blocks(const Box& bx) { return trap_grid(bx.size()); }
int force_pass_blocks() { return 64; }
int label_tile_blocks(const Box& bx) {
  return ((bx.nx + CX - 1) / CX) * ((bx.ny + CY - 1) / CY) * ((bx.nz + CZ - 1) / CZ);
}
int label_jump_blocks(int64_t roots_cap) {
  const int cap = (int)std::min<int64_t>(roots_cap, INT32_MAX - 1);
  return (int)std::min<int64_t>(std::max<int64_t>((cap + 255) / 256, 1), 16384);
}

// tile t of box bx -> its position in the tile grid
__device__ __forceinline__ void tile_pos(const Box& bx, int t, int& cx, int& cy, int& cz) {
  const int tx = (bx.nx + CX - 1) / CX, ty = (bx.ny + CY - 1) / CY;
  cx = t % tx;
  cy = (t / tx) % ty;
  cz = t / (tx * ty);
}

__global__ void trap_nw_bin_multi(const OuterJob* J, Multi m) {
  int lb, lg;
  const int r = multi_box(m, blockIdx.x, lb, lg);
  trap_nw_bin_body(J[r].ta, lb, lg);
}
__global__ void trap_w_bin_multi(const OuterJob* J, Multi m) {
  int lb, lg;
  const int r = multi_box(m, blockIdx.x, lb, lg);
  trap_w_bin_body(J[r].ta, lb, lg);
}
__global__ void trap_w_mask_multi(const OuterJob* J, Multi m) {
  int lb, lg;
  const int r = multi_box(m, blockIdx.x, lb, lg);
  trap_w_mask_body(J[r].ta, lb, lg);
}
__global__ void force_bin_multi(const OuterJob* J, Multi m, int wet) {
  int lb, lg;
  const int r = multi_box(m, blockIdx.x, lb, lg);
  force_bin_body(J[r].ta.bin, J[r].ta.bx, wet ? J[r].fw : J[r].fnw, lb, lg);
}
// one block per box
__global__ void ccl_reset_multi(const OuterJob* J, Multi m) {
  int lb, lg;
  const int r = multi_box(m, blockIdx.x, lb, lg);
  if (threadIdx.x == 0 && J[r].ta.roots) J[r].ta.roots[0] = 0;
}
__global__ __launch_bounds__(CX * CY) void ccl_local_multi(const OuterJob* J, Multi m, int conn) {
  int lb, lg, cx, cy, cz;
  const int r = multi_box(m, blockIdx.x, lb, lg);
  const TrapArgs& a = J[r].ta;
  tile_pos(a.bx, lb, cx, cy, cz);
  const int cap = a.roots ? (int)min(a.roots_cap, (int64_t)(INT32_MAX - 1)) : 0;
  ccl_local_tile(a.bx, a.bin, a.labels, a.troot, conn, a.roots, cap, false, cx, cy, cz);
}
template <int AX>
__global__ __launch_bounds__(1024) void ccl_border_multi(const OuterJob* J, Multi m, int conn) {
  int lb, lg, cx, cy, cz;
  const int r = multi_box(m, blockIdx.x, lb, lg);
  const TrapArgs& a = J[r].ta;
  tile_pos(a.bx, lb, cx, cy, cz);
  ccl_border_tile<AX>(a.bx, a.labels, a.troot, conn, cx, cy, cz);
}
__global__ void ccl_jump_multi(const OuterJob* J, Multi m) {
  int lb, lg;
  const int r = multi_box(m, blockIdx.x, lb, lg);
  const TrapArgs& a = J[r].ta;
  const int cap = a.roots ? (int)min(a.roots_cap, (int64_t)(INT32_MAX - 1)) : 0;
  ccl_jump_body(a.labels, a.troot, (int)a.bx.size(), a.roots, cap, lb, lg);
}

// the labelling passes of launch_label_components (full = false)
static void label_multi(const OuterJob* J, const Multi& one, const Multi& tiles, const Multi& jump,
                        int conn, hipStream_t st) {
  hipLaunchKernelGGL(ccl_reset_multi, dim3(one.total), dim3(64), 0, st, J, one);
  hipLaunchKernelGGL(ccl_local_multi, dim3(tiles.total), dim3(CX, CY, 1), 0, st, J, tiles, conn);
  hipLaunchKernelGGL(ccl_border_multi<0>, dim3(tiles.total), dim3(border_threads<0>()), 0, st,
                     J, tiles, conn);
  hipLaunchKernelGGL(ccl_jump_multi, dim3(jump.total), dim3(256), 0, st, J, jump);
  hipLaunchKernelGGL(ccl_border_multi<1>, dim3(tiles.total), dim3(border_threads<1>()), 0, st,
                     J, tiles, conn);
  hipLaunchKernelGGL(ccl_jump_multi, dim3(jump.total), dim3(256), 0, st, J, jump);
  hipLaunchKernelGGL(ccl_border_multi<2>, dim3(tiles.total), dim3(border_threads<2>()), 0, st,
                     J, tiles, conn);
  hipLaunchKernelGGL(ccl_jump_multi, dim3(jump.total), dim3(256), 0, st, J, jump);
}

hipError_t launch_trap_multi(const OuterJob* J, const Multi& trap, const Multi& force,
                             const Multi& one, const Multi& tiles, const Multi& jump, int conn,
                             hipStream_t st) {
  hipLaunchKernelGGL(trap_nw_bin_multi, dim3(trap.total), dim3(256), 0, st, J, trap);
  hipLaunchKernelGGL(force_bin_multi, dim3(force.total), dim3(256), 0, st, J, force, 0);
  label_multi(J, one, tiles, jump, conn, st);
  hipLaunchKernelGGL(trap_w_bin_multi, dim3(trap.total), dim3(256), 0, st, J, trap);
  hipLaunchKernelGGL(force_bin_multi, dim3(force.total), dim3(256), 0, st, J, force, 1);
  label_multi(J, one, tiles, jump, conn, st);
  hipLaunchKernelGGL(trap_w_mask_multi, dim3(trap.total), dim3(256), 0, st, J, trap);
  return hipGetLastError();
}

}  // namespace lsm
