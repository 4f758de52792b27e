// lsm_driver.hip -- the level-set evolution drivers on one GPU, behind the
// reference's C entry points.
//
// Every entry point copies the host arrays it reads to the device, runs the
// whole computation device-resident on one stream, and copies back what the
// reference would have modified.  Control decisions that need scalars (the
// inner pseudo-time loop, the outer stopping tests) read back one or two
// values per step through pinned memory; the inner loop is pipelined so the
// GPU never waits for the host (see inner_loop()).
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <chrono>
#include <stdexcept>
#include <vector>
#include <memory>
#include "lsm_batch.h"
#include "lsm_xfer.h"
#include "lsm_gz.h"
#include "lsm_session.h"
#include "lsmpqs_hip.h"

namespace lsm {

#define DCHECK(x)                                                                \
  do {                                                                           \
    hipError_t e_ = (x);                                                         \
    if (e_ != hipSuccess)                                                        \
      throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

static size_t bytes(const Session* s) { return (size_t)s->n * sizeof(real); }

// whole fields between the caller's host arrays and the session, in stream
// order; both return once the copy is complete (copy_to_device and
// copy_to_host stage large fields through pinned buffers)
static void up(Session* s, real* d, const real* h) {
  if (!h) throw std::runtime_error("required host array is NULL");
  DCHECK(copy_to_device(d, h, bytes(s), s->st));
}
static void down(Session* s, real* h, const real* d) {
  if (h) DCHECK(copy_to_host(h, d, bytes(s), s->st));
}
static void sync(Session* s) { DCHECK(hipStreamSynchronize(s->st)); }

// Which right-hand-side kernel the reference's dispatcher picks
// (reference: src/utilities/qss_general_util.c:115-141).  Returns the scheme
// order; *cst says whether the constant-coefficient variant is selected.
static int rhs_scheme(const Options* o, bool var_path) {
  switch (o->order_space_accur) {
  case 2: return 2;
  case 3: case 4: return 3;
  case 5: return var_path ? 5 : 3;
  default: return 3;
  }
}

// ---------------------------------------------------------------------------
// Outer reductions requested from reinit_dev: volume of {f < 0} and the
// local max-norm against prev, into s->red (see launch_outer_reduce)
struct ReduceReq {
  const real* prev;
  real zone, eps, dV;
};

// Reinitialization with the subcell fix (reference:
// src/reinitialization/reinitialize_top.c:297-377).  `f` is replaced by the
// result; A, B and S serve as next-iterate, phi0 copy and distance buffers.
// With `red`, the outer reductions of the result are left in s->red (fused
// into the single-iteration pass when there is one).
// fast: the FP32 library's fast arithmetic (the model drivers; off with the
// exact-arithmetic switch); the exported reinitializeSubcellFix3d is exact.
static bool fast_fp32() { return sizeof(real) == 4 && !(quirks() & EXACT_FP32); }
static void reinit_dev(Session* s, real*& f, const Grid* g, const Options* o,
                       const ReduceReq* red = nullptr, bool fast = false) {
  const Box& bx = s->bx;
  real t_r = 0, dt_r = real(0.5) * g->dx[0];
  auto reduce_after = [&]() {
    if (red)
      DCHECK(launch_outer_reduce(bx, f, red->prev, red->zone, red->eps, red->dV, s->part, s->red,
                                 s->st));
  };
  if (!(t_r < o->tmax_r)) {
    reduce_after();
    return;
  }
  if (!(dt_r < o->tmax_r)) {
    // one iteration (tmax_r = 0.4 at dx = 1): data = phi0, fused first step
    if (red)
      DCHECK(launch_reinit_first_step_reduce(bx, s->A, f, dt_r, g->dx[0], g->dx[1], g->dx[2],
                                             red->prev, red->zone, red->eps, red->dV, s->part,
                                             s->red, s->st, fast));
    else
      DCHECK(launch_reinit_first_step(bx, s->A, f, dt_r, g->dx[0], g->dx[1], g->dx[2], s->st,
                                      fast));
    BcArgs ba{};
    ba.bx = bx; ba.src = s->A; ba.dst = s->A;
    DCHECK(launch_bc(ba, false, s->st));
    std::swap(f, s->A);
    return;
  }
  real* phi0 = s->B;
  real* dist0 = s->S;
  DCHECK(hipMemcpyAsync(phi0, f, bytes(s), hipMemcpyDeviceToDevice, s->st));
  DCHECK(launch_reinit_distance(bx, dist0, phi0, g->dx[0], g->dx[1], g->dx[2], s->st, fast));
  while (t_r < o->tmax_r) {
    DCHECK(launch_reinit_step(bx, s->A, f, phi0, dist0, dt_r, g->dx[0], g->dx[1], g->dx[2], s->st,
                              fast));
    BcArgs ba{};
    ba.bx = bx; ba.src = s->A; ba.dst = s->A;
    DCHECK(launch_bc(ba, false, s->st));
    std::swap(f, s->A);
    t_r = t_r + dt_r;
  }
  reduce_after();
}

// Smoothed-sign reinitialization of f (reference:
// src/reinitialization/reinitialize_top.c:205-295): "RK2" with dt_r = 0.5 dx
// until tmax_r, the all-face BC after each stage.  Uses A (stage 1), B (next).
static void reinit_medium_dev(Session* s, real*& f, const Grid* g, const Options* o) {
  ReinitArgs ra{};
  ra.bx = s->bx;
  ra.dx = g->dx[0]; ra.dy = g->dx[1]; ra.dz = g->dx[2];
  real t_r = 0, dt_r = real(0.5) * g->dx[0], tmax_r = o->tmax_r;
  ra.dt = dt_r;
  BcArgs ba{};
  ba.bx = s->bx;
  while (t_r < tmax_r) {
    ra.in = f; ra.un = f; ra.out = s->A;
    DCHECK(launch_reinit_smooth(ra, 1, s->st));
    ba.src = s->A; ba.dst = s->A;
    DCHECK(launch_bc(ba, false, s->st));
    ra.in = s->A; ra.un = f; ra.out = s->B;
    DCHECK(launch_reinit_smooth(ra, 2, s->st));
    ba.src = s->B; ba.dst = s->B;
    DCHECK(launch_bc(ba, false, s->st));
    std::swap(f, s->B);
    t_r = t_r + dt_r;
  }
}

// Trapped-phase masks (reference: src/connectivity/trapComponents.c:156-375)
// full_labels: leave the wetting-phase labels resolved at every voxel (for
// callers that read them back); otherwise only tile roots carry them
// impose: phi = max(phi, impose) first, fused into the first trapping pass
static TrapArgs trap_args(Session* s, const Grid* g, const Options* o, bool full_labels,
                          const real* impose) {
  ensure_labels(s);
  TrapArgs ta{};
  ta.impose_mask = impose;
  ta.bx = s->bx;
  ta.phi = s->phi; ta.mask = s->mask; ta.mask_nw = s->mask_nw; ta.mask_w = s->mask_w;
  ta.s = s->S; ta.bin = s->bin; ta.labels = s->labels; ta.troot = s->troot;
  ta.roots = s->roots; ta.roots_cap = s->roots_cap; ta.full_labels = full_labels;
  ta.nw_ind = o->phi_nw_ind; ta.w_ind = o->phi_w_ind;
  ta.center_inlet = o->center_inlet;
  ta.fill = -20 * g->dx[0];
  if (o->phi_nw_ind < 0 || o->phi_nw_ind >= s->n || o->phi_w_ind < 0 || o->phi_w_ind >= s->n)
    throw std::runtime_error("phi_nw_ind / phi_w_ind outside the ghost box (call getMainInd_new)");
  return ta;
}
// 26, or 6 with LSMPQS_CONNECTIVITY=6
static int label_connectivity() {
  int conn = getenv("LSMPQS_CONNECTIVITY") ? atoi(getenv("LSMPQS_CONNECTIVITY")) : 26;
  return conn == 6 ? 6 : 26;
}
static void trap_dev(Session* s, const Grid* g, const Options* o, bool full_labels = false,
                     const real* impose = nullptr) {
  const TrapArgs ta = trap_args(s, g, o, full_labels, impose);
  DCHECK(launch_trap_components(ta, label_connectivity(), s->st));
}

// The outer step's device work after the inner steps with trapping and a
// single subcell-fix reinitialization iteration (dt_r = dx / 2 >= tmax_r > 0),
// as a record a batched run executes together with other boxes': trapping,
// reinitialization of mask_w, mask_nw (into A, then mask_w's old buffer),
// trapped-phase velocities, merge and floor, reinitialization of phi with
// the outer reductions (into mask_nw's old buffer).  outer_rotate() applies
// the resulting buffer rotation to the session afterwards.
static bool outer_batchable(const Options* o, const Grid* g) {
  const real dt_r = real(0.5) * g->dx[0];
  return o->check_connectivity && real(0) < o->tmax_r && !(dt_r < o->tmax_r);
}
static OuterJob outer_job(Session* s, const Grid* g, const Options* o, bool var, real eps,
                          real dV) {
  OuterJob j{};
  j.ta = trap_args(s, g, o, false, var ? s->mask : nullptr);
  j.fnw = trap_nw_force(j.ta);
  j.fw = trap_w_force(j.ta);
  j.rin[0] = s->mask_w;  j.rout[0] = s->A;
  j.rin[1] = s->mask_nw; j.rout[1] = s->mask_w;
  j.rin[2] = s->phi;     j.rout[2] = s->mask_nw;
  j.dt_r = real(0.5) * g->dx[0];
  j.dx = g->dx[0]; j.dy = g->dx[1]; j.dz = g->dx[2];
  j.floor = s->floor;
  j.overlap = o->overlap;
  j.prev = s->phi_prev;
  j.zone = o->err_check_zone; j.eps = eps; j.dV = dV;
  j.part = s->part;
  j.nred = reduce_pass_blocks(s->bx);
  j.red = s->red;
  if (var && o->conserve_imbibe) {
    j.vn = s->vn; j.b = s->b; j.vx = s->vx; j.vy = s->vy; j.vz = s->vz;
  }
  j.eps_var = real(1.5) * g->dx[0];
  return j;
}
static void outer_rotate(Session* s) {
  real* a0 = s->A;
  s->A = s->phi;
  s->phi = s->mask_nw;
  s->mask_nw = s->mask_w;
  s->mask_w = a0;
}

// Sorted list of fill voxels with |vel_n| < 1e-5 and, for each, the voxel
// whose normal term the reference's loop reuses: the closest preceding voxel
// of the same z-slab, in (k, j, i) loop order, with |vel_n| >= 1e-5.
static void build_stale(Session* s, const Grid* g) {
  s->nstale = 0;
  if (!(quirks() & QUIRK_STALE_NORM)) return;
  const Box& bx = s->bx;
  for (int attempt = 0; attempt < 2; ++attempt) {
    if (s->stale_cap == 0 || attempt == 1) {
      int cap = attempt == 0 ? (1 << 16) : *s->h_count;
      if (s->stale_list) { (void)hipFree(s->stale_list); (void)hipFree(s->stale_idx); (void)hipFree(s->stale_src); }
      DCHECK(hipMalloc((void**)&s->stale_list, (size_t)cap * sizeof(int64_t)));
      DCHECK(hipMalloc((void**)&s->stale_idx, (size_t)cap * sizeof(int64_t)));
      DCHECK(hipMalloc((void**)&s->stale_src, (size_t)cap * sizeof(int64_t)));
      s->stale_cap = cap;
    }
    DCHECK(hipMemsetAsync(s->count, 0, sizeof(int), s->st));
    DCHECK(launch_collect_stale(bx, s->vn, s->stale_list, s->count, s->stale_cap, s->st));
    DCHECK(hipMemcpyAsync(s->h_count, s->count, sizeof(int), hipMemcpyDeviceToHost, s->st));
    sync(s);
    if (*s->h_count <= s->stale_cap) break;
  }
  int cnt = *s->h_count;
  if (cnt == 0) return;
  std::vector<int64_t> idx(cnt), src(cnt);
  DCHECK(hipMemcpyAsync(idx.data(), s->stale_list, (size_t)cnt * sizeof(int64_t),
                        hipMemcpyDeviceToHost, s->st));
  sync(s);
  std::sort(idx.begin(), idx.end());
  int T = dt_slabs(g);
  int nsl = bx.khi - bx.klo + 1;
  int64_t plane = (int64_t)bx.nx * bx.ny;
  for (int e = 0; e < cnt; ++e) {
    int64_t v = idx[e];
    int k = (int)(v / plane);
    int64_t r = v - k * plane;
    int j = (int)(r / bx.nx), i = (int)(r - (int64_t)j * bx.nx);
    // first plane of the slab containing k
    int t = (int)(((int64_t)(k - bx.klo + 1) * T - 1) / nsl);
    while (t > 0 && bx.klo + (int)((int64_t)nsl * t / T) > k) --t;
    while (t + 1 < T && bx.klo + (int)((int64_t)nsl * (t + 1) / T) <= k) ++t;
    int slab_klo = bx.klo + (int)((int64_t)nsl * t / T);
    int64_t prev;
    if (i > bx.ilo) prev = bx.at(i - 1, j, k);
    else if (j > bx.jlo) prev = bx.at(bx.ihi, j - 1, k);
    else if (k > slab_klo) prev = bx.at(bx.ihi, bx.jhi, k - 1);
    else prev = -1;
    if (prev >= 0 && e > 0 && idx[e - 1] == prev) src[e] = src[e - 1];
    else src[e] = prev;
  }
  DCHECK(hipMemcpyAsync(s->stale_idx, idx.data(), (size_t)cnt * sizeof(int64_t),
                        hipMemcpyHostToDevice, s->st));
  DCHECK(hipMemcpyAsync(s->stale_src, src.data(), (size_t)cnt * sizeof(int64_t),
                        hipMemcpyHostToDevice, s->st));
  sync(s);
  s->nstale = cnt;
}

// ---------------------------------------------------------------------------
// One inner time step: dt probe, the RK stages (each a fused RHS+stage
// kernel followed by the ghost extrapolation), mask imposition fused into the
// last stage.  Returns nothing; dt lands in s->dt (device) and s->h_dt
// (pinned, valid after ev_dt).

static void stage_timed(Session* s, const StageArgs& a, const StepCfg& c, int stage, bool fin) {
  if (s->profile) DCHECK(hipEventRecord(s->stage_ev[2 * stage], s->st));
  DCHECK(launch_stage(a, c.scheme, c.var, stage, fin, s->st));
  if (s->profile) {
    DCHECK(hipEventRecord(s->stage_ev[2 * stage + 1], s->st));
  }
}

StepArgs make_step_args(Session* s, const StepCfg& c) {
  const Box& bx = s->bx;
  StepArgs a{};
  ProbeArgs& pa = a.probe;
  pa.bx = bx; pa.phi = s->phi; pa.vn = c.var ? s->vn : nullptr; pa.a_const = c.a_const;
  pa.dx = c.dx; pa.dy = c.dy; pa.dz = c.dz; pa.eno3_quirk = c.quirk_eno3; pa.nslabs = c.slabs;
  pa.bmax_over_dx = c.bmax; pa.umax_over_dx = c.umax; pa.dt_out = s->dt;

  StageArgs sa{};
  sa.bx = bx; sa.dt = s->dt;
  sa.vn = s->vn; sa.b = s->b; sa.vx = s->vx; sa.vy = s->vy; sa.vz = s->vz;
  sa.a_const = c.a_const; sa.b_const = c.b_const;
  sa.dx = c.dx; sa.dy = c.dy; sa.dz = c.dz;
  sa.norm_sign = c.norm_sign; sa.eno3_quirk = c.quirk_eno3;
  sa.floor = s->floor;
  sa.stale_idx = s->stale_idx; sa.stale_src = s->stale_src;
  sa.nstale = (c.var ? s->nstale : 0);
  sa.fast = (sizeof(real) == 4 && !(quirks() & EXACT_FP32)) ? 1 : 0;
  BcArgs bc{};
  bc.bx = bx;
  a.order = c.order;
  for (int i = 0; i < 3; ++i) { a.stage[i] = sa; a.bc[i] = bc; }
  if (c.order <= 1) {
    // phi_new = max(phi + dt L(phi), floor) into B; face layers into A
    a.stage[0].in = s->phi; a.stage[0].un = s->phi; a.stage[0].out = s->B; a.stage[0].pre = s->A;
    a.bc[0].src = s->A; a.bc[0].dst = s->B; a.bc[0].floor = s->floor;
    return a;
  }
  a.stage[0].in = s->phi; a.stage[0].un = s->phi; a.stage[0].out = s->A; a.stage[0].pre = nullptr;
  a.bc[0].src = s->A; a.bc[0].dst = s->A;
  // the new state goes to phi (in place, pointwise) or, with s->redirect, to
  // phi_prev's buffer
  real* res = s->redirect ? s->phi_prev : s->phi;
  if (c.order == 2) {
    a.stage[1].in = s->A; a.stage[1].un = s->phi; a.stage[1].out = res; a.stage[1].pre = s->B;
    a.bc[1].src = s->B; a.bc[1].dst = res; a.bc[1].floor = s->floor;
  } else {
    a.stage[1].in = s->A; a.stage[1].un = s->phi; a.stage[1].out = s->B; a.stage[1].pre = nullptr;
    a.bc[1].src = s->B; a.bc[1].dst = s->B;
    a.stage[2].in = s->B; a.stage[2].un = s->phi; a.stage[2].out = res; a.stage[2].pre = s->A;
    a.bc[2].src = s->A; a.bc[2].dst = res; a.bc[2].floor = s->floor;
  }
  return a;
}

// Inner steps replayed as HIP graphs (LSMPQS_GRAPHS=1).  A step is two
// graphs: the time-step probe with the read-back of dt into pinned memory,
// then the RK stages with their ghost extrapolations; the host records ev_dt
// between them, so it learns dt while the stages run, as with direct
// launches.  A graph pair is keyed by the step's kernel arguments (buffers,
// coefficients, scheme); the session's buffers rotate between a few
// arrangements (phi and phi_prev swap every outer step, the order-1 result
// lands in B), so a session keeps a few pairs and replaces the least
// recently used.  No memset is recorded.  Off by default: on MI355X with
// ROCm 7 the two graph launches of a step cost more host time than its eight
// direct launches (64^3 ENO2 drainage, 844 steps: 55 ms of inner-loop time
// with graphs against 38 ms without), and the direct launches already keep
// the GPU busy while the host waits for dt.
struct StepGraph {
  StepArgs key;
  int scheme = 0;
  bool var = false;
  hipGraphExec_t probe = nullptr, stages = nullptr;
  uint64_t used = 0;
};
struct GraphCache {
  static constexpr int kMax = 8;
  std::vector<StepGraph> g;
  uint64_t clock = 0;
};

static void destroy_graph(StepGraph& e) {
  if (e.probe) (void)hipGraphExecDestroy(e.probe);
  if (e.stages) (void)hipGraphExecDestroy(e.stages);
  e.probe = e.stages = nullptr;
}

void graph_cache_free(GraphCache* c) {
  for (StepGraph& e : c->g) destroy_graph(e);
  delete c;
}

static bool graphs_enabled() {
  const char* e = getenv("LSMPQS_GRAPHS");
  return e && *e && atoi(e) != 0;
}

static void enqueue_probe(Session* s, const StepArgs& a, const StepCfg& c) {
  DCHECK(launch_probe(a.probe, c.scheme, s->st));
  DCHECK(hipMemcpyAsync(s->h_dt, s->dt, sizeof(real), hipMemcpyDeviceToHost, s->st));
}

static void enqueue_stages(Session* s, const StepArgs& a, const StepCfg& c) {
  const int nst = c.order <= 1 ? 1 : (c.order == 2 ? 2 : 3);
  for (int i = 0; i < nst; ++i) {
    const bool fin = i == nst - 1;
    stage_timed(s, a.stage[i], c, i + 1, fin);
    DCHECK(launch_bc(a.bc[i], fin, s->st));
  }
}

// capture what `body` enqueues on s->st into an executable graph
template <typename F>
static hipGraphExec_t capture_on(Session* s, F&& body) {
  DCHECK(hipStreamBeginCapture(s->st, hipStreamCaptureModeThreadLocal));
  hipGraph_t graph = nullptr;
  try {
    body();
  } catch (...) {
    (void)hipStreamEndCapture(s->st, &graph);
    if (graph) (void)hipGraphDestroy(graph);
    throw;
  }
  DCHECK(hipStreamEndCapture(s->st, &graph));
  hipGraphExec_t exec = nullptr;
  hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  DCHECK(e);
  return exec;
}

static StepGraph& step_graph(Session* s, const StepArgs& a, const StepCfg& c) {
  if (!s->graphs) s->graphs = new GraphCache();
  GraphCache& gc = *s->graphs;
  for (StepGraph& e : gc.g)
    if (e.scheme == c.scheme && e.var == c.var && !memcmp(&e.key, &a, sizeof(StepArgs))) {
      e.used = ++gc.clock;
      return e;
    }
  StepGraph* slot;
  if ((int)gc.g.size() < GraphCache::kMax) {
    gc.g.emplace_back();
    slot = &gc.g.back();
  } else {
    slot = &gc.g[0];
    for (StepGraph& e : gc.g)
      if (e.used < slot->used) slot = &e;
    destroy_graph(*slot);
  }
  slot->key = a;
  slot->scheme = c.scheme;
  slot->var = c.var;
  slot->probe = capture_on(s, [&] { enqueue_probe(s, a, c); });
  slot->stages = capture_on(s, [&] { enqueue_stages(s, a, c); });
  slot->used = ++gc.clock;
  return *slot;
}

static void inner_step(Session* s, const StepCfg& c) {
  StepArgs a = make_step_args(s, c);
  if (!s->profile && graphs_enabled()) {
    StepGraph& g = step_graph(s, a, c);
    DCHECK(hipGraphLaunch(g.probe, s->st));
    DCHECK(hipEventRecord(s->ev_dt, s->st));
    DCHECK(hipGraphLaunch(g.stages, s->st));
  } else {
    enqueue_probe(s, a, c);
    DCHECK(hipEventRecord(s->ev_dt, s->st));
    enqueue_stages(s, a, c);
  }
  if (c.order <= 1) std::swap(s->phi, s->B);
}

// Imposition floor from the current masks; called whenever mask, mask_w,
// mask_nw or the overlap may have changed before inner steps run.
static void update_floor(Session* s, const StepCfg& c) {
  DCHECK(launch_floor(s->floor, s->mask, s->mask_w, s->mask_nw, c.overlap, c.conn, s->n, s->st));
}

static StepCfg make_cfg(Session* s, const Options* o, const Grid* g, bool var, real a_const) {
  StepCfg c{};
  c.scheme = rhs_scheme(o, var);
  c.var = var;
  c.order = o->order_time_accur < 1 ? 1 : (o->order_time_accur > 3 ? 3 : o->order_time_accur);
  c.a_const = a_const; c.b_const = o->b;
  c.norm_sign = (var && c.scheme == 5 && (quirks() & QUIRK_WENO5_SIGN)) ? real(1) : real(-1);
  c.bmax = o->b_max_over_dx; c.umax = o->max_U_over_dx; c.overlap = o->overlap;
  c.conn = o->check_connectivity ? 1 : 0;
  c.slabs = dt_slabs(g);
  c.quirk_eno3 = (quirks() & QUIRK_ENO3_READ) ? 1 : 0;
  c.dx = g->dx[0]; c.dy = g->dx[1]; c.dz = g->dx[2];
  (void)s;
  return c;
}

static void profile_collect(Session* s, int order) {
  if (!s->profile) return;
  sync(s);
  for (int st = 1; st <= (order < 1 ? 1 : (order > 3 ? 3 : order)); ++st) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, s->stage_ev[2 * st], s->stage_ev[2 * st + 1]) == hipSuccess) {
      s->tm.stage_ms[st] += ms;
      s->tm.stage_launches[st] += 1;
    }
  }
}

// The pseudo-time loop `while (dt_sub < tplot)` (constCurvModel3d.c:127-244).
// dt of step n is known right after its probe kernel, so the host waits on
// that event while the GPU is still running step n's stages, then enqueues
// step n+1 (or stops).  Returns dt_sub; *last_dt gets the final dt.
static real inner_loop(Session* s, const StepCfg& c, real tplot, real* last_dt) {
  real dt_sub = 0, dt = 0;
  BatchHook* hook = batch_hook();
  while (dt_sub < tplot) {
    if (hook) {
      dt = hook->step(s, c);
      if (c.order <= 1) std::swap(s->phi, s->B);
    } else {
      inner_step(s, c);
      DCHECK(hipEventSynchronize(s->ev_dt));
      dt = *s->h_dt;
    }
    if (s->redirect) {  // the step wrote phi_prev's buffer: it is the new phi
      std::swap(s->phi, s->phi_prev);
      s->redirect = false;
    }
    if (!(dt > real(0)) || !isfinite((double)dt))
      throw std::runtime_error("non-positive or non-finite time step");
    profile_collect(s, c.order);
    dt_sub += dt;
    s->tm.inner_steps++;
  }
  if (hook) hook->away();
  if (last_dt) *last_dt = dt;
  return dt_sub;
}

// outer reductions: out[0] = volume of {f < 0}, out[1] = local max-norm diff vs f2
static void read_reduce(Session* s) {
  DCHECK(hipMemcpyAsync(s->h_red, s->red, 2 * sizeof(double), hipMemcpyDeviceToHost, s->st));
  sync(s);
}
static void outer_reduce(Session* s, const real* f, const real* f2, real zone, real eps, real dV) {
  DCHECK(launch_outer_reduce(s->bx, f, f2, zone, eps, dV, s->part, s->red, s->st));
  read_reduce(s);
}

// Outer loop of constCurvModel3d / constCurvModel3dNoVar
// (reference: src/level_set_evolution/constCurvModel3d.c:50-334, 336-610).
static real ccm_dev(Session* s, Options* o, const Grid* g, FILE* fp, bool var, real a_const) {
  const Box& bx = s->bx;
  BatchHook* hook = batch_hook();
  real vsmall = g->dx[0] * g->dx[1] * g->dx[2];
  real eps = o->eps_coefficient * g->dx[0];
  real dV = g->dx[0] * g->dx[1] * g->dx[2];
  double max_abs_err = 100, vol_phi, vol_pore, satn, satn_old;
  real t = 0;
  s->redirect = false;
  StepCfg c = make_cfg(s, o, g, var, a_const);
  if (s->profile && s->stage_ev.empty()) {
    s->stage_ev.resize(8);
    for (auto& e : s->stage_ev) DCHECK(hipEventCreate(&e));
  }
  update_floor(s, c);
  outer_reduce(s, s->mask, nullptr, 0, eps, dV);
  vol_pore = s->h_red[0];
  outer_reduce(s, s->phi, nullptr, 0, eps, dV);
  vol_phi = s->h_red[0];
  satn = vol_phi / vol_pore;
  (void)bx;
  std::unique_ptr<StepWriter> ckpt;
  while ((t < o->tmax) && (max_abs_err > (double)o->eps_stop)) {
    s->tm.outer_steps++;
    // phi_prev = phi: the first inner step writes its result to phi_prev's
    // buffer and the two are swapped (RK1 copies)
    if (c.order >= 2) s->redirect = true;
    else DCHECK(hipMemcpyAsync(s->phi_prev, s->phi, bytes(s), hipMemcpyDeviceToDevice, s->st));
    // A batched box's inner steps run on the batch's stream, not on s->st:
    // their time is the host time of the inner loop, and the box's stream
    // is synchronized only when the batch takes its next request.
    const auto t_in = std::chrono::steady_clock::now();
    if (!hook) DCHECK(hipEventRecord(s->ev_a, s->st));
    real last_dt = 0;
    real dt_sub = inner_loop(s, c, o->tplot, &last_dt);
    if (s->redirect) {  // no inner step ran
      s->redirect = false;
      DCHECK(hipMemcpyAsync(s->phi_prev, s->phi, bytes(s), hipMemcpyDeviceToDevice, s->st));
    }
    float ms = 0;
    if (!hook) {
      DCHECK(hipEventRecord(s->ev_b, s->st));
      sync(s);
      (void)hipEventElapsedTime(&ms, s->ev_a, s->ev_b);
    } else {
      ms = (float)(1e3 * std::chrono::duration<double>(std::chrono::steady_clock::now() - t_in)
                             .count());
    }
    s->tm.inner_ms += ms;
    if (fp) fprintf(fp, "1 GPU: Level set time = %lf\n", ms * 1e-3);
    const auto t3 = std::chrono::steady_clock::now();
    t += dt_sub;
    bool batched = false;
    if (hook && outer_batchable(o, g)) {
      // the batch runs this box's outer-step device work with the other
      // boxes' and leaves the reductions in s->h_red
      batched = hook->outer(s, outer_job(s, g, o, var, eps, dV), fast_fp32(),
                            label_connectivity());
      if (batched) {
        outer_rotate(s);
        if (var && o->conserve_imbibe) build_stale(s, g);
      }
    }
    if (batched) {
      // done
    } else if (var && !o->check_connectivity) {
      DCHECK(launch_impose(s->phi, s->mask, s->phi, 0, s->n, s->st));
    }
    if (!batched && o->check_connectivity) {
      trap_dev(s, g, o, false, var ? s->mask : nullptr);
      reinit_dev(s, s->mask_w, g, o, nullptr, fast_fp32());
      reinit_dev(s, s->mask_nw, g, o, nullptr, fast_fp32());
      if (var && o->conserve_imbibe) {
        DCHECK(launch_trap_velocity(s->bx, s->mask_nw, s->vn, s->b, s->vx, s->vy, s->vz,
                                    g->dx[0], s->st));
        build_stale(s, g);
      }
      DCHECK(launch_merge_floor(s->phi, s->floor, s->mask, s->mask_w, s->mask_nw, c.overlap,
                                s->n, s->st));
    }
    if (fp) fprintf(fp, "Reinitializing....");
    if (!batched) {
      const ReduceReq rq{s->phi_prev, o->err_check_zone, eps, dV};
      reinit_dev(s, s->phi, g, o, &rq, fast_fp32());
    }
    if (fp) fprintf(fp, "Reinitialized\n");
    if (!batched) read_reduce(s);
    vol_phi = s->h_red[0];
    max_abs_err = s->h_red[1];
    satn_old = satn;
    satn = vol_phi / vol_pore;
    if ((satn > (double)o->vol_frac_max) || (vol_phi <= (double)vsmall)) {
      if (ckpt) ckpt->finish();
      return -1;
    }
    if (o->checkpoint) {
      // restart files of this outer step (gz, working directory), written by a
      // worker thread from snapshots while the next outer step runs; a later
      // step's files replace them in order
      if (!ckpt) ckpt.reset(new StepWriter(bytes(s), g->grid_dims_ghostbox, 4));
      ckpt->submit(s->st, s->phi, "checkpoint_phi");
      ckpt->submit(s->st, s->phi_prev, "checkpoint_phi_prev");
      ckpt->submit(s->st, s->mask_w, "mask_w");
      ckpt->submit(s->st, s->mask_nw, "mask_nw");
    }
    const double conn_s =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t3).count();
    if (fp) {
      fprintf(fp, "connectivity time = %lf\n", conn_s);
      fprintf(fp, "t = %4.1f\t", (double)t);
      fprintf(fp, "max_abs_err = %4.3f,\t", max_abs_err);
      fprintf(fp, "nw phase satn pct = %4.3f\n", satn * 100);
      fflush(fp);
    }
    if (o->use_satn_stop) {
      if (fabsf((float)(100 * (satn - satn_old))) < 0.001) break;
    }
  }
  if (ckpt) ckpt->finish();
  return t;
}

// ---------------------------------------------------------------------------
// Model drivers (reference: src/level_set_evolution/drain_model3d.c:49-165,
// imbibe_model3d.c:49-161)

static void set_theta_overlap(Options* o, const Grid* g) {
  real theta = o->theta;
  real rad = (real)(3.14159 / 180);
  if (theta <= 20 * rad) o->overlap = 0;
  else if (theta <= 40 * rad) o->overlap = (real)(0.3 * g->dx[0]);
  else if (theta <= 50 * rad) o->overlap = (real)(0.5 * g->dx[0]);
  else o->overlap = (real)(1 * g->dx[0]);
}

static void main_ind(Options* o, const Grid* g) {
  int nx = g->grid_dims_ghostbox[0], nxy = nx * g->grid_dims_ghostbox[1];
  int xi = o->center_inlet ? (int)(0.5 * (g->ilo_fb + g->ihi_fb)) : g->ilo_fb - 1;
  int yi = (int)(0.5 * (g->jlo_fb + g->jhi_fb));
  int zi = (int)(0.5 * (g->klo_fb + g->khi_fb));
  o->phi_nw_ind = xi + yi * nx + zi * nxy;
  o->phi_w_ind = (g->ihi_fb + 1) + yi * nx + zi * nxy;
}

static void disconnected_init(Session* s, real* f, const Grid* g) {
  real cx = (real)(0.5 * (g->x_lo[0] + g->x_hi[0]));
  real cy = (real)(0.5 * (g->x_lo[1] + g->x_hi[1]));
  real cz = (real)(0.5 * (g->x_lo[2] + g->x_hi[2]));
  real r = 2 * (g->x_hi[0] - g->x_lo[0]);
  DCHECK(launch_sphere(f, s->bx, g->x_lo_ghostbox, g->dx, cx, cy, cz, r, -1, s->st));
}

static void reservoir_box(Session* s, const Grid* g) {
  real corner[3] = {g->x_lo_ghostbox[0], g->x_lo_ghostbox[1], g->x_lo_ghostbox[2]};
  real side[3] = {3 * g->dx[0], g->x_hi_ghostbox[1] - g->x_lo_ghostbox[1],
                  g->x_hi_ghostbox[2] - g->x_lo_ghostbox[2]};
  DCHECK(launch_box(s->phi_extra, s->bx, g->x_lo_ghostbox, g->dx, corner, side, -1, s->st));
}

BatchHook*& batch_hook() {
  static thread_local BatchHook* h = nullptr;
  return h;
}

// data_step_N files of the model drivers: off with LSMPQS_NO_STEP_FILES=1
// and in batched runs (the boxes would share the file names)
static bool step_files_enabled() {
  if (batch_hook()) return false;
  const char* e = getenv("LSMPQS_NO_STEP_FILES");
  return !(e && *e && atoi(e) != 0);
}

// Setup before the curvature loop (drain_model3d.c:56-100): trapped masks as
// one huge sphere, reservoir box, mask reinitialization (drainage only), mask
// gradient, overlap, inlet imposition, inlet/outlet indices.
static void model_setup(Session* s, Options* o, QSS_DataArrays* p, Grid* g, bool drain, bool var) {
  ensure_grad(s);
  if (var) ensure_var(s);
  up(s, s->phi, p->phi);
  up(s, s->mask, p->mask);
  disconnected_init(s, s->mask_w, g);
  disconnected_init(s, s->mask_nw, g);
  reservoir_box(s, g);
  if (drain && o->mask_reinit) reinit_dev(s, s->mask, g, o, nullptr, fast_fp32());
  DCHECK(launch_central_grad(s->bx, s->gx, s->gy, s->gz, s->mask, g->dx[0], g->dx[1], g->dx[2],
                             s->st));
  set_theta_overlap(o, g);
  if (o->reservoir_inlet) {
    DCHECK(launch_min(s->phi, s->phi_extra, s->n, s->st));
    DCHECK(launch_impose(s->phi, s->mask, s->phi, 0, s->n, s->st));
  }
  main_ind(o, g);
  if (p->mask_disconn_init) down(s, p->mask_disconn_init, s->mask_nw);
  sync(s);
}

// Coefficients of one curvature step (drain_model3d.c:108-131): the
// time-step bounds, and on the variable path b, V and a = a0 from the mask.
static void model_coeffs(Session* s, Options* o, const Grid* g, bool var, bool drain, real a0) {
  if (var) {
    DCHECK(launch_var_curv_adv(s->bx, s->b, s->vx, s->vy, s->vz, s->mask, s->gx, s->gy, s->gz,
                               o->b, g->dx[0], s->st));
    real inv = real(1.0) / g->dx[0];
    DCHECK(launch_coeff_reduce(s->bx, s->b, s->vx, s->vy, s->vz, inv, s->part, s->red, s->st));
    DCHECK(hipMemcpyAsync(s->h_red, s->red, 2 * sizeof(double), hipMemcpyDeviceToHost, s->st));
    sync(s);
    real bmax = (real)s->h_red[0], umax = (real)s->h_red[1];
    o->max_U_over_dx = umax;
    o->b_max_over_dx = 2 * bmax * (inv * inv + inv * inv + inv * inv);
  } else {
    real ix = 1 / g->dx[0], iy = 1 / g->dx[1], iz = 1 / g->dx[2];
    o->b_max_over_dx = 2 * o->b * (ix * ix + iy * iy + iz * iz);
    o->max_U_over_dx = 0;
  }
  if (!drain) o->conserve_imbibe = 1;
  if (var) {
    real ct = (real)cos((double)o->theta);
    DCHECK(launch_var_norm(s->bx, s->vn, s->mask, s->gx, s->gy, s->gz, a0, ct, g->dx[0], s->st));
    build_stale(s, g);
  }
}

void model_run(Options* o, QSS_DataArrays* p, Grid* g, FILE* fp, bool drain) {
  BatchHook* hook = batch_hook();
  Session* s = session_for(p, g);
  if (hook) hook->mark(0);
  s->tm = Timers();
  bool var = (o->theta > 0) || o->use_var_theta;
  model_setup(s, o, p, g, drain, var);
  if (hook) hook->mark(1);
  real a0 = o->a, da = o->b * o->dc;
  int step = o->init_step + 1;
  std::unique_ptr<StepWriter> writer;
  while (drain ? (a0 < o->amax) : (a0 > o->amin)) {
    printf("\nStep %d: a = %f\n", step, (double)a0);
    if (fp) fprintf(fp, "\nStep %d: a = %f\n", step, (double)a0);
    model_coeffs(s, o, g, var, drain, a0);
    if (hook) hook->mark(2);
    if (o->check_connectivity) trap_dev(s, g, o);
    if (hook) hook->mark(3);
    real t_step = ccm_dev(s, o, g, fp, var, a0);
    if (t_step == -1) break;
    if (drain) a0 += da; else a0 -= da;
    if (o->reservoir_inlet) DCHECK(launch_min(s->phi, s->phi_extra, s->n, s->st));
    DCHECK(launch_impose(s->phi, s->mask, s->phi, 0, s->n, s->st));
    if (step_files_enabled()) {
      // written by a worker thread from a snapshot while the next step runs
      char fname[256];
      snprintf(fname, sizeof fname, "data_step_%d", step);
      if (!writer) writer.reset(new StepWriter(bytes(s), g->grid_dims_ghostbox));
      writer->submit(s->st, s->phi, fname);
    }
    step += 1;
  }
  if (writer) writer->finish();
  down(s, p->phi, s->phi);
  down(s, p->mask, s->mask);
  down(s, p->mask_w, s->mask_w);
  down(s, p->mask_nw, s->mask_nw);
  down(s, p->phi_prev, s->phi_prev);
  down(s, p->phi_extra, s->phi_extra);
  down(s, p->mask_x, s->gx);
  down(s, p->mask_y, s->gy);
  down(s, p->mask_z, s->gz);
  if (var) {
    down(s, p->normal_velocity, s->vn);
    down(s, p->curvature_coeff, s->b);
    down(s, p->external_velocity_x, s->vx);
    down(s, p->external_velocity_y, s->vy);
    down(s, p->external_velocity_z, s->vz);
  }
  sync(s);
}

}  // namespace lsm

using namespace lsm;

// ===========================================================================
// C ABI
// ===========================================================================

extern "C" {

const char* lsmpqs_hip_last_error(void) { return lsm::last_error(); }

int lsmpqs_set_dt_slabs(int n) {
  lsm::set_slabs(n);
  return n;
}

unsigned lsmpqs_set_quirks(unsigned q) {
  unsigned old = lsm::quirks();
  lsm::set_quirks(q);
  return old;
}

void lsmpqs_release(QSS_DataArrays* p) { lsm::session_release(p); }

#define API_TRY try {
#define API_CATCH(ret)                      \
  }                                         \
  catch (const std::exception& e) {         \
    lsm::set_error(e.what());               \
    fprintf(stderr, "lsmpqs_hip: %s\n", e.what()); \
    return ret;                             \
  }

QSSLIB_REAL constCurvModel3d(Options* o, QSS_DataArrays* p, Grid* g, FILE* fp) {
  API_TRY
  lsm::set_error("");
  Session* s = session_for(p, g);
  s->tm = Timers();
  ensure_var(s);
  up(s, s->phi, p->phi);
  up(s, s->mask, p->mask);
  if (o->check_connectivity) {
    up(s, s->mask_w, p->mask_w);
    up(s, s->mask_nw, p->mask_nw);
  }
  up(s, s->vn, p->normal_velocity);
  up(s, s->b, p->curvature_coeff);
  up(s, s->vx, p->external_velocity_x);
  up(s, s->vy, p->external_velocity_y);
  up(s, s->vz, p->external_velocity_z);
  sync(s);
  build_stale(s, g);
  real t = ccm_dev(s, o, g, fp, true, 0);
  down(s, p->phi, s->phi);
  down(s, p->phi_prev, s->phi_prev);
  if (o->check_connectivity) {
    down(s, p->mask_w, s->mask_w);
    down(s, p->mask_nw, s->mask_nw);
    if (o->conserve_imbibe) {
      down(s, p->normal_velocity, s->vn);
      down(s, p->curvature_coeff, s->b);
      down(s, p->external_velocity_x, s->vx);
      down(s, p->external_velocity_y, s->vy);
      down(s, p->external_velocity_z, s->vz);
    }
  }
  sync(s);
  return t;
  API_CATCH(-2)
}

QSSLIB_REAL constCurvModel3dNoVar(Options* o, QSS_DataArrays* p, Grid* g, FILE* fp,
                                  QSSLIB_REAL a) {
  API_TRY
  lsm::set_error("");
  Session* s = session_for(p, g);
  s->tm = Timers();
  up(s, s->phi, p->phi);
  up(s, s->mask, p->mask);
  if (o->check_connectivity) {
    up(s, s->mask_w, p->mask_w);
    up(s, s->mask_nw, p->mask_nw);
  }
  real t = ccm_dev(s, o, g, fp, false, a);
  down(s, p->phi, s->phi);
  down(s, p->phi_prev, s->phi_prev);
  if (o->check_connectivity) {
    down(s, p->mask_w, s->mask_w);
    down(s, p->mask_nw, s->mask_nw);
  }
  sync(s);
  return t;
  API_CATCH(-2)
}

void drain_model3d(Options* o, QSS_DataArrays* p, Grid* g, FILE* fp) {
  API_TRY
  lsm::set_error("");
  model_run(o, p, g, fp, true);
  API_CATCH()
}

void imbibe_model3d(Options* o, QSS_DataArrays* p, Grid* g, FILE* fp) {
  API_TRY
  lsm::set_error("");
  model_run(o, p, g, fp, false);
  API_CATCH()
}

void trapComponents_mask(QSS_DataArrays* p, Grid* g, Options* o) {
  API_TRY
  lsm::set_error("");
  if (g->num_dims == 2) {
    api2d_trap(p, g, o);
    return;
  }
  Session* s = session_for(p, g);
  up(s, s->phi, p->phi);
  up(s, s->mask, p->mask);
  up(s, s->mask_nw, p->mask_nw);
  trap_dev(s, g, o, p->connectivity != nullptr);
  down(s, p->phi, s->phi);
  down(s, p->mask_nw, s->mask_nw);
  down(s, p->mask_w, s->mask_w);
  if (p->scratch1) {
    DCHECK(hipMemsetAsync(s->S, 0, bytes(s), s->st));
    down(s, p->scratch1, s->S);
  }
  std::vector<int> lab;
  if (p->connectivity || p->phi_bin) {
    lab.resize((size_t)s->n);
    DCHECK(hipMemcpyAsync(lab.data(), s->labels, (size_t)s->n * sizeof(int), hipMemcpyDeviceToHost,
                          s->st));
    if (p->phi_bin)
      DCHECK(hipMemcpyAsync(p->phi_bin, s->bin, (size_t)s->n, hipMemcpyDeviceToHost, s->st));
  }
  sync(s);
  if (p->connectivity)
    for (int64_t l = 0; l < s->n; ++l) p->connectivity[l] = lab[l] >= 0 ? lab[l] : 0;
  API_CATCH()
}

void findConnectivity3d(QSS_DataArrays* p, Grid* g) {
  API_TRY
  lsm::set_error("");
  Session* s = session_for(p, g);
  ensure_labels(s);
  DCHECK(hipMemcpyAsync(s->bin, p->phi_bin, (size_t)s->n, hipMemcpyHostToDevice, s->st));
  int conn = getenv("LSMPQS_CONNECTIVITY") ? atoi(getenv("LSMPQS_CONNECTIVITY")) : 26;
  DCHECK(launch_label_components(s->bx, s->bin, s->labels, s->troot, conn == 6 ? 6 : 26, s->st,
                                 s->roots, s->roots_cap, true));
  std::vector<int> lab((size_t)s->n);
  DCHECK(hipMemcpyAsync(lab.data(), s->labels, (size_t)s->n * sizeof(int), hipMemcpyDeviceToHost,
                        s->st));
  sync(s);
  for (int64_t l = 0; l < s->n; ++l) p->connectivity[l] = lab[l] >= 0 ? lab[l] : 0;
  API_CATCH()
}

void reinitializeSubcellFix3d(QSSLIB_REAL* data, Grid* g, Options* o) {
  API_TRY
  lsm::set_error("");
  // a private session keyed to the data pointer
  Session* s = session_for((const QSS_DataArrays*)data, g);
  up(s, s->D, data);
  reinit_dev(s, s->D, g, o);
  down(s, data, s->D);
  sync(s);
  API_CATCH()
}

void reinitializeMedium3d(QSSLIB_REAL* data, Grid* g, Options* o) {
  API_TRY
  lsm::set_error("");
  Session* s = session_for((const QSS_DataArrays*)data, g);
  up(s, s->D, data);
  reinit_medium_dev(s, s->D, g, o);
  down(s, data, s->D);
  sync(s);
  API_CATCH()
}

void signedLinearExtrapolationBCqss(QSSLIB_REAL* phi, Grid* g, int bdry_location_idx) {
  API_TRY
  lsm::set_error("");
  if (g->num_dims == 2) {
    api2d_bc(phi, g, bdry_location_idx);
    return;
  }
  if (bdry_location_idx != 9) return;  // the reference's 3D dispatcher handles only 9
  Session* s = session_for((const QSS_DataArrays*)phi, g);
  up(s, s->D, phi);
  BcArgs ba{};
  ba.bx = s->bx; ba.src = s->D; ba.dst = s->D;
  DCHECK(launch_bc(ba, false, s->st));
  down(s, phi, s->D);
  sync(s);
  API_CATCH()
}

void getMainInd_new(Options* o, QSS_DataArrays* p, Grid* g) {
  (void)p;
  if (g->num_dims == 2) api2d_main_ind(o, g);
  else main_ind(o, g);
}

void setThetaOverlap(QSS_DataArrays* p, Grid* g, Options* o) {
  (void)p;
  if (g->num_dims == 3) set_theta_overlap(o, g);
  else if (g->num_dims == 2) api2d_overlap(o, g);
}

void setSpaceDerivFunc(Grid* g, Options* o) {
  // kernels are chosen from order_space_accur and theta at run time; the
  // slot is cleared so no stale pointer into another library survives
  (void)g;
  o->space_deriv_func = nullptr;
}

void createReservoirInlet3d(QSS_DataArrays* p, Grid* g) {
  API_TRY
  Session* s = session_for(p, g);
  ensure_grad(s);
  reservoir_box(s, g);
  down(s, p->phi_extra, s->phi_extra);
  sync(s);
  API_CATCH()
}

void initializeDisconnectedMasks(QSSLIB_REAL* data, Grid* g) {
  API_TRY
  if (g->num_dims == 2) {
    api2d_disconn(data, g);
    return;
  }
  Session* s = session_for((const QSS_DataArrays*)data, g);
  disconnected_init(s, s->D, g);
  down(s, data, s->D);
  sync(s);
  API_CATCH()
}

int lsmpqs_bench_inner_steps(Options* o, QSS_DataArrays* p, Grid* g, int nsteps, int var_path,
                             QSSLIB_REAL a, int upload, lsmpqs_timing* out) {
  API_TRY
  lsm::set_error("");
  Session* s = session_for(p, g);
  bool var = var_path != 0;
  if (upload) {
    if (var) ensure_var(s);
    up(s, s->phi, p->phi);
    up(s, s->mask, p->mask);
    up(s, s->mask_w, p->mask_w);
    up(s, s->mask_nw, p->mask_nw);
    if (var) {
      up(s, s->vn, p->normal_velocity);
      up(s, s->b, p->curvature_coeff);
      up(s, s->vx, p->external_velocity_x);
      up(s, s->vy, p->external_velocity_y);
      up(s, s->vz, p->external_velocity_z);
    }
    sync(s);
    if (var) build_stale(s, g);
  }
  s->profile = out && out->profile;
  if (s->profile && s->stage_ev.empty()) {
    s->stage_ev.resize(8);
    for (auto& e : s->stage_ev) DCHECK(hipEventCreate(&e));
  }
  s->tm = Timers();
  s->redirect = false;
  StepCfg c = make_cfg(s, o, g, var, a);
  update_floor(s, c);
  DCHECK(hipEventRecord(s->ev_a, s->st));
  for (int n = 0; n < nsteps; ++n) {
    inner_step(s, c);
    profile_collect(s, c.order);
  }
  DCHECK(hipEventRecord(s->ev_b, s->st));
  sync(s);
  float ms = 0;
  DCHECK(hipEventElapsedTime(&ms, s->ev_a, s->ev_b));
  if (out) {
    out->elapsed_ms = ms;
    for (int st = 0; st < 3; ++st) {
      out->stage_ms[st] = s->tm.stage_ms[st + 1];
      out->stage_launches[st] = s->tm.stage_launches[st + 1];
    }
    out->inner_steps = nsteps;
    out->last_dt = (double)*s->h_dt;
    out->nstale = s->nstale;
  }
  s->profile = false;
  return 0;
  API_CATCH(-1)
}

int lsmpqs_model_prepare(Options* o, QSS_DataArrays* p, Grid* g, int drain) {
  API_TRY
  lsm::set_error("");
  Session* s = session_for(p, g);
  bool var = (o->theta > 0) || o->use_var_theta;
  model_setup(s, o, p, g, drain != 0, var);
  model_coeffs(s, o, g, var, drain != 0, o->a);
  if (o->check_connectivity) trap_dev(s, g, o);
  sync(s);
  return 0;
  API_CATCH(-1)
}

int lsmpqs_session_stats(QSS_DataArrays* p, Grid* g, lsmpqs_timing* out) {
  API_TRY
  Session* s = session_for(p, g);
  out->inner_steps = s->tm.inner_steps;
  out->outer_steps = s->tm.outer_steps;
  out->inner_loop_ms = s->tm.inner_ms;
  return 0;
  API_CATCH(-1)
}

int lsmpqs_download(QSS_DataArrays* p, Grid* g, const char* field) {
  API_TRY
  Session* s = session_for(p, g);
  if (!strcmp(field, "phi")) down(s, p->phi, s->phi);
  else if (!strcmp(field, "mask_w")) down(s, p->mask_w, s->mask_w);
  else if (!strcmp(field, "mask_nw")) down(s, p->mask_nw, s->mask_nw);
  else return -1;
  sync(s);
  return 0;
  API_CATCH(-1)
}

}  // extern "C"
