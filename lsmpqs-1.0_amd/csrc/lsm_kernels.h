// lsm_kernels.h -- host-side launchers of the gfx950 kernels.  One build of
// the library serves one QSSLIB_REAL; every launcher enqueues on `st` and
// returns the HIP status of the launch.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "QSSLIB_config.h"
#include "lsm_device.h"

namespace lsm {

typedef QSSLIB_REAL real;

// Inputs of one fused RHS + Runge-Kutta stage over the fill box.
struct StageArgs {
  Box bx;
  const real* in;        // field the right-hand side is taken of (ghosts valid)
  const real* un;        // u^n, read pointwise by stages 2 and 3
  real* out;             // stage result (fill box)
  real* pre;             // final stage: pre-imposition values of the two layers next to each face
  const real* dt;        // device scalar time step
  const real *vn, *b, *vx, *vy, *vz;   // variable coefficients (var path)
  real a_const, b_const;               // constant coefficients (const path)
  real dx, dy, dz;
  real norm_sign;        // -1, or +1 for the reference's WENO5 normal term
  int eno3_quirk;        // read phi(i+3,j,j) in the ENO3 x line, as the reference does
  const real* floor;     // final stage: phi = max(phi, floor), see launch_floor
  const int64_t* stale_idx;  // sorted voxels with |vel_n| < 1e-5 (var path)
  const int64_t* stale_src;  // voxel whose normal term each of them reuses, or -1
  int nstale;
  int fast;              // single precision: fast WENO5/curvature arithmetic
  int tiles_x, tiles_y, tiles_z;  // launch decomposition, set by launch_stage
};

// scheme: 2, 3 or 5; var: variable coefficients; stage: 1, 2 or 3;
// final: apply the mask imposition and write the face layers to `pre`.
hipError_t launch_stage(const StageArgs& a, int scheme, bool var, int stage, bool final_stage,
                        hipStream_t st);

// Time step from H/dx at the last fill voxel of each of `nslabs` z-slabs:
// dt = 0.5 / (max H + b_max_over_dx + max_U_over_dx), written to *dt_out.
struct ProbeArgs {
  Box bx;
  const real* phi;
  const real* vn;        // NULL on the const path
  real a_const;
  real dx, dy, dz;
  int eno3_quirk;
  int nslabs;
  real bmax_over_dx, umax_over_dx;
  real* dt_out;
};
hipError_t launch_probe(const ProbeArgs& a, int scheme, hipStream_t st);

// Signed linear extrapolation of every ghost voxel from the fill box, composed
// in the reference's face order x-, x+, y-, y+, z-, z+.  With final_mask set,
// reads `src` fill values, writes dst ghosts = max(BC, floor).
struct BcArgs {
  Box bx;
  const real* src;
  real* dst;
  const real* floor;
};
hipError_t launch_bc(const BcArgs& a, bool final_mask, hipStream_t st);

// Batched inner steps of many boxes.  One record per box; the launches read
// the record count and prefix sums from device memory, so one captured graph
// serves every step whatever the set of boxes.
struct BatchRec {
  ProbeArgs probe;
  StageArgs stage[3];
  BcArgs bc[3];
};
struct BatchView {
  const BatchRec* recs;
  const int* count;              // [0]: number of records n
  const int* brick_prefix;       // n+1 entries: first brick of each record
  const int64_t* ghost_prefix;   // n+1 entries: first ghost voxel of each record
  int* counters;                 // work-queue counters, one per stage launch
  real* dts;                     // dt of each record (written by the probe)
};
hipError_t launch_batch_reset(int* counters, int ncounters, hipStream_t st);
hipError_t launch_probe_batch(const BatchView& v, int scheme, int nmax, hipStream_t st);
// persistent: `grid` workgroups take bricks from counters[si] until none is left
hipError_t launch_stage_batch(const BatchView& v, int si, int scheme, bool var, int stage,
                              bool final_stage, bool fast, int grid, hipStream_t st);
hipError_t launch_bc_batch(const BatchView& v, int si, bool final_mask, int grid, hipStream_t st);
// bricks and ghost voxels of one box for a z-chunk length zc (sets a.tiles_*)
int stage_bricks(StageArgs& a, int zc);
int64_t ghost_count(const Box& bx);

// Ghost voxel g of a box -> (i, j, k), in three regions: the z ghost planes
// (nA voxels), the y ghost rows of the fill planes (nB), the x ghost columns
// of the fill rows.
__host__ __device__ inline void ghost_ijk(const Box& bx, int64_t g, int64_t nA, int64_t nB, int& i,
                                          int& j, int& k) {
  // 32-bit divisions: a ghost box holds fewer than 2^31 voxels
  const unsigned nx = (unsigned)bx.nx, ny = (unsigned)bx.ny;
  const unsigned nxf = (unsigned)(bx.ihi - bx.ilo + 1), nyf = (unsigned)(bx.jhi - bx.jlo + 1);
  const unsigned gxn = nx - nxf, gyn = ny - nyf;
  if (g < nA) {
    const unsigned u = (unsigned)g, t = u / nx, kk = t / ny;
    i = (int)(u - t * nx);
    j = (int)(t - kk * ny);
    k = (int)kk < bx.klo ? (int)kk : bx.khi + 1 + ((int)kk - bx.klo);
  } else if (g < nA + nB) {
    const unsigned h = (unsigned)(g - nA), t = h / nx, kk = t / gyn;
    i = (int)(h - t * nx);
    const int jj = (int)(t - kk * gyn);
    k = bx.klo + (int)kk;
    j = jj < bx.jlo ? jj : bx.jhi + 1 + (jj - bx.jlo);
  } else {
    const unsigned h = (unsigned)(g - nA - nB), t = h / gxn, kk = t / nyf;
    const int ii = (int)(h - t * gxn);
    j = bx.jlo + (int)(t - kk * nyf);
    k = bx.klo + (int)kk;
    i = ii < bx.ilo ? ii : bx.ihi + 1 + (ii - bx.ilo);
  }
}

__host__ __device__ inline void ghost_regions(const Box& bx, int64_t& nA, int64_t& nB,
                                              int64_t& nC) {
  int nxf = bx.ihi - bx.ilo + 1, nyf = bx.jhi - bx.jlo + 1, nzf = bx.khi - bx.klo + 1;
  nA = (int64_t)(bx.nz - nzf) * bx.nx * bx.ny;
  nB = (int64_t)nzf * (bx.ny - nyf) * bx.nx;
  nC = (int64_t)nzf * nyf * (bx.nx - nxf);
}

// Elementwise over the whole ghost box.
hipError_t launch_impose(real* dst, const real* m, const real* src, real overlap, int64_t n,
                         hipStream_t st);                     // dst = max(src, m - overlap)
hipError_t launch_merge(real* phi, const real* m, int64_t n, hipStream_t st);  // phi = m>0 ? -m : phi
hipError_t launch_min(real* phi, const real* e, int64_t n, hipStream_t st);    // phi = min(phi, e)
hipError_t launch_fill(real* dst, real v, int64_t n, hipStream_t st);
// Imposition floor of the last stage: floor = mask - overlap, and with the
// trapped-phase masks max(mask - overlap, mask_w, mask_nw).  max(max(max(v,
// m - ov), mw), mnw) == max(v, floor) exactly, signed zeros included (both
// keep the first maximal operand in the order v, m - ov, mw, mnw).
hipError_t launch_floor(real* floor, const real* m, const real* mw, const real* mnw, real overlap,
                        int conn, int64_t n, hipStream_t st);

// FP64 reductions over the fill box.  out[0] = smoothed volume of {f1 < 0}
// (width eps), out[1] = max |f1 - f2| over |f1| < zone (-1 if none; skipped
// when f2 is NULL).  partials needs reduce_partials() = 2 REDUCE_BLOCKS
// doubles whatever the box size (at most REDUCE_BLOCKS blocks per pass).
constexpr int REDUCE_BLOCKS = 2048;
int reduce_partials();
hipError_t launch_outer_reduce(const Box& bx, const real* f1, const real* f2, real zone, real eps,
                               real dV, double* partials, double* out, hipStream_t st);
// max of a fill-box field (the reference's b_max / U_max scans); writes out[0]
hipError_t launch_coeff_reduce(const Box& bx, const real* b, const real* vx, const real* vy,
                               const real* vz, real inv_dx, double* partials, double* out,
                               hipStream_t st);

// Smoothed-sign reinitialization: stage 1 out = in + dt L(in), stage 2
// out = 3/4 un + 1/4 (in + dt L(in)), fill box only.
struct ReinitArgs {
  Box bx;
  const real* in;
  const real* un;
  real* out;
  real dt, dx, dy, dz;
  int tiles_x, tiles_y, tiles_z;  // set by the launcher
};
hipError_t launch_reinit_smooth(const ReinitArgs& a, int stage, hipStream_t st);

// Subcell-fix reinitialization pieces.
// fast: single-precision fast arithmetic (hardware reciprocal and square
// root), which the model drivers use in the FP32 library
hipError_t launch_reinit_distance(const Box& bx, real* dist0, const real* phi0, real dx, real dy,
                                  real dz, hipStream_t st, bool fast = false);
hipError_t launch_reinit_step(const Box& bx, real* next, const real* data, const real* phi0,
                              const real* dist0, real dt_r, real dx, real dy, real dz,
                              hipStream_t st, bool fast = false);
// the subcell-fix right-hand side alone, fill box
hipError_t launch_reinit_rhs(const Box& bx, real* rhs, const real* data, const real* phi0,
                             const real* dist0, real dx, real dy, real dz, hipStream_t st);
// the first step with data = phi0, D0 formed on the fly (no phi0 copy, no D0 field)
hipError_t launch_reinit_first_step(const Box& bx, real* next, const real* phi0, real dt_r,
                                    real dx, real dy, real dz, hipStream_t st, bool fast = false);

// launch_reinit_first_step plus the outer reductions of its result (as
// launch_outer_reduce(bx, next, prev, zone, eps, dV, ...), same sums)
hipError_t launch_reinit_first_step_reduce(const Box& bx, real* next, const real* phi0, real dt_r,
                                           real dx, real dy, real dz, const real* prev, real zone,
                                           real eps, real dV, double* partials, double* out,
                                           hipStream_t st, bool fast = false);
// phi = -mask_nw where mask_nw > 0, and the imposition floor with the
// trapped-phase masks (launch_merge then launch_floor with conn = 1), one pass
hipError_t launch_merge_floor(real* phi, real* floor, const real* m, const real* mw,
                              const real* mnw, real overlap, int64_t n, hipStream_t st);

// Coefficient fields of the variable-coefficient path.
hipError_t launch_central_grad(const Box& bx, real* gx, real* gy, real* gz, const real* f,
                               real dx, real dy, real dz, hipStream_t st);
hipError_t launch_var_norm(const Box& bx, real* vn, const real* mask, const real* gx,
                           const real* gy, const real* gz, real a0, real cos_theta, real dx,
                           hipStream_t st);
hipError_t launch_var_curv_adv(const Box& bx, real* b, real* vx, real* vy, real* vz,
                               const real* mask, const real* gx, const real* gy, const real* gz,
                               real const_b, real dx, hipStream_t st);
hipError_t launch_trap_velocity(const Box& bx, const real* phi_w, real* vn, real* b, real* vx,
                                real* vy, real* vz, real dx, hipStream_t st);
// voxels with |vn| < 1e-5: appends fill-box indices to list (capacity cap), counts in *count
hipError_t launch_collect_stale(const Box& bx, const real* vn, int64_t* list, int* count, int cap,
                                hipStream_t st);

// Analytic initialisers over the ghost box.
hipError_t launch_sphere(real* phi, const Box& bx, const real* xlo, const real* dx, real cx,
                         real cy, real cz, real r, int inside_flag, hipStream_t st);
hipError_t launch_box(real* phi, const Box& bx, const real* xlo, const real* dx, const real* corner,
                      const real* side, int inside_flag, hipStream_t st);

// Connected components (26-neighbourhood) and trapped-phase masks.
struct TrapArgs {
  Box bx;
  real* phi;
  const real* mask;
  real* mask_nw;
  real* mask_w;
  real* s;               // wetting-phase field max(-phi, mask, mask_nw)
  uint8_t* bin;
  int* labels;
  int* troot;            // tile roots (labelling scratch)
  int* roots;            // tile-root list (labelling scratch) or NULL
  int64_t roots_cap;     // room for roots in that list
  bool full_labels;      // leave the wetting-phase labels resolved at every voxel
  int nw_ind, w_ind;
  int center_inlet;
  real fill;             // -20 dx
  int dims2;             // 2D grid on a one-plane box: the 2D inlet/outlet rules
  const real* impose_mask;  // not NULL: phi = max(phi, impose_mask) first, in the same pass
};
hipError_t launch_trap_components(const TrapArgs& a, int connectivity, hipStream_t st);
// labels: largest linear index of the component on foreground, -1 on
// background; troot: scratch of the same size (each voxel's tile root);
// roots: optional list [count, roots...] with room for roots_cap tile roots
// (NULL: the passes between the border launches scan troot).  full = false
// leaves the label only at the tile roots: the label of v is labels[troot[v]].
hipError_t launch_label_components(const Box& bx, const uint8_t* bin, int* labels, int* troot,
                                   int connectivity, hipStream_t st, int* roots = nullptr,
                                   int64_t roots_cap = 0, bool full = true);

// The voxels the reference forces into the binary field before labelling:
// up to four boxes of the ghost box, inclusive bounds.
struct ForceSet {
  int n;
  int lo[4][3], hi[4][3];
  void add(const Box& bx, int i0, int i1, int j0, int j1, int k0, int k1);
};
ForceSet trap_nw_force(const TrapArgs& a);  // the inlet voxels of the non-wetting labelling
ForceSet trap_w_force(const TrapArgs& a);   // the outlet voxels of the wetting labelling

// ---------------------------------------------------------------------------
// One outer step of many boxes at once (batched runs).  Each launch covers
// every box: block b serves box r, prefix[r] <= b < prefix[r + 1], as block
// b - prefix[r] of prefix[r + 1] - prefix[r], and each box gets the grid its
// single-box launch uses, so every voxel sees the single-box arithmetic and
// the reductions add in the same order.
struct Multi {
  const int* prefix;  // device, n + 1 entries
  int n;
  int total;          // prefix[n]: blocks of the launch
};
__device__ __forceinline__ int multi_box(const Multi& m, int b, int& lb, int& lg) {
  int lo = 0, hi = m.n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (m.prefix[mid] <= b) lo = mid; else hi = mid - 1;
  }
  lb = b - m.prefix[lo];
  lg = m.prefix[lo + 1] - m.prefix[lo];
  return lo;
}

// Everything one box's outer step needs (constCurvModel3d.c:258-292 with
// trapping and a single subcell-fix reinitialization iteration).  The three
// reinitializations read rin[p] and write rout[p] (mask_w, mask_nw, phi).
struct OuterJob {
  TrapArgs ta;
  ForceSet fnw, fw;
  real* rin[3];
  real* rout[3];
  real dt_r, dx, dy, dz;
  real* floor;             // merge and floor: phi = ta.phi, masks rout[0], rout[1]
  real overlap;
  const real* prev;        // reductions of rout[2] against prev
  real zone, eps, dV;
  double* part;
  int nred;                // partials of that reduction (reduce_pass_blocks)
  double* red;
  real *vn, *b, *vx, *vy, *vz;  // trapped-phase velocities from rout[1], or vn = NULL
  real eps_var;
};
// Block counts per box of each launch kind (the single-box launchers' grids).
int trap_pass_blocks(const Box& bx);      // trapping passes over the ghost box
int force_pass_blocks();                  // forced inlet/outlet voxels
int label_tile_blocks(const Box& bx);     // labelling tiles
int label_jump_blocks(int64_t roots_cap); // root-jump passes
int reinit_pass_blocks(const Box& bx);    // reinitialization pass
int reduce_pass_blocks(const Box& bx);    // reinitialization pass with the outer reductions
int stream_pass_blocks(int64_t n);        // merge and floor
int bc_pass_blocks(const Box& bx);        // ghost extrapolation
// The launch sequences (jobs: device copies of the records).
hipError_t launch_trap_multi(const OuterJob* jobs, const Multi& trap, const Multi& force,
                             const Multi& one, const Multi& tiles, const Multi& jump, int conn,
                             hipStream_t st);
hipError_t launch_reinit_multi(const OuterJob* jobs, int pass, const Multi& rows, bool fast,
                               hipStream_t st);
hipError_t launch_reinit_reduce_multi(const OuterJob* jobs, const Multi& rows, const Multi& one,
                                      bool fast, hipStream_t st);
hipError_t launch_merge_floor_multi(const OuterJob* jobs, const Multi& stream, hipStream_t st);
hipError_t launch_trap_velocity_multi(const OuterJob* jobs, const Multi& rows, hipStream_t st);
hipError_t launch_bc_multi(const OuterJob* jobs, int pass, const Multi& ghosts, hipStream_t st);
// out[2r], out[2r + 1] = jobs[r].red[0], [1]
hipError_t launch_gather_red(const OuterJob* jobs, int n, double* out, hipStream_t st);

}  // namespace lsm
