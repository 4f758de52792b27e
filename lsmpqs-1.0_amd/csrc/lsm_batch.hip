// lsm_batch.hip -- many independent boxes (subvolumes of uneven size, each
// with its own options and curvature schedule) drained or imbibed together.
//
// Each box runs the ordinary model driver on a host thread of its own, with
// its own device session and stream, so the schedule, the stopping tests
// and the host logic are the single-box code.  Its inner time steps and the
// device work of its outer steps are handed to the scheduler on the calling
// thread instead: once every live box of a group (half of the boxes) is
// waiting, one batched round runs for them -- for inner steps a batched
// time-step probe, then per RK stage one persistent stage kernel that takes
// 64 x 8 x zc bricks of all boxes from a work queue, and one batched ghost
// extrapolation; for outer steps one launch per pass over all the boxes.
// The inner sequence is captured once per group as a HIP graph; each step
// only uploads the per-box records (pointers, geometry, brick and ghost
// prefix sums) and replays the graph.
// The brick depth zc and the number of persistent workgroups are picked on
// the first step by timing the first stage over the actual boxes.  Every
// voxel sees exactly the arithmetic of the single-box path, so a batched run
// reproduces the boxes' individual runs bit for bit.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>
#include "lsm_batch.h"
#include "lsmpqs_hip.h"

namespace lsm {
namespace {

#define BCHECK(x)                                                                \
  do {                                                                           \
    hipError_t e_ = (x);                                                         \
    if (e_ != hipSuccess)                                                        \
      throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

constexpr int kCounters = 4;

struct Request {
  Session* s;
  StepCfg c;
  real dt = 0;
  bool done = false;  // under m; the box waits on cv
  std::mutex m;
  std::condition_variable cv;
  const OuterJob* job = nullptr;  // an outer step instead of an inner step
  bool fast = false;
  int conn = 26;
};

constexpr int kTables = 9;  // block-prefix tables of a batched outer step

class Batch final : public BatchHook {
 public:
  // The boxes form ngroups groups of consecutive indices.  A round serves the
  // parked requests of one group once each of its live boxes is parked or
  // away; each group's rounds run on its own stream, so that one group's
  // boxes wake up, run their host logic and park again while other groups'
  // rounds run on the GPU (with one group, every round waited for that
  // turnaround with the GPU idle), and the groups' kernels overlap.
  Batch(int nbox, int ngroups) : nmax_(nbox) {
    ng_ = std::max(1, std::min({ngroups, kMaxGroups, nbox}));
    for (int i = 0; i < nbox; ++i) live_[group_of(i, nbox, ng_)] += 1;
  }
  ~Batch() override { release(); }
  static int group_of(int box, int nbox, int ng) { return (int)((int64_t)box * ng / nbox); }
  int groups() const { return ng_; }

  // Box threads run their own work (set-up, outer steps) only while holding
  // one of `slots` run slots: a box takes a slot when it starts and after
  // each batched step, and gives it back when it parks for a step or
  // finishes, so at most `slots` boxes issue runtime calls from the host at
  // a time (the other live boxes are parked or waiting for a slot).  A box
  // holding a slot never waits for another box, so every slot comes back.
  // Without the limit after set-up, the ~30 boxes a round wakes spent 0.6 -
  // 1.4 ms each in their few runtime calls (64 subvolumes of a 768^3 pack,
  // 16 host cores) instead of 0.1 ms.
  void set_slots(int n) { slots_ = n < 1 ? 1 : n; }
  void acquire_slot() {
    const double t0 = now_s();
    std::unique_lock<std::mutex> lk(slot_mu_);
    cv_slot_.wait(lk, [&] { return slots_ > 0; });
    slots_ -= 1;
    holding() = true;
    trace().slot_s += now_s() - t0;
  }
  static bool& holding() {
    static thread_local bool h = false;
    return h;
  }
  static int& group() {  // the calling box thread's group
    static thread_local int g = 0;
    return g;
  }

  // Host-time breakdown of one box thread (LSMPQS_BATCH_TRACE=1 prints the
  // sums and maxima over the boxes to stderr after the run).
  struct BoxTrace {
    double start = 0, first_step = -1, last_step = 0, end = 0;
    double marks[4] = {-1, -1, -1, -1};  // first time each set-up mark was passed
    double away_since = -1, away_s = 0;  // outer-step work (away() to the next step)
    long long aways = 0;
    double slot_s = 0, parked_s = 0;
    long long steps = 0;
    // per served request: round served to wake-up, wake-up to the next
    // request (host work)
    double wake_s = 0, host_s = 0, slot_at = -1;
  };
  static double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
  }
  static BoxTrace& trace() {
    static thread_local BoxTrace t;
    return t;
  }
  std::vector<BoxTrace> traces;  // filled by box_finished
  double sched_launch_s = 0, sched_wait_s = 0, first_wait_s = -1;
  void release_slot() {
    if (!holding()) return;
    holding() = false;
    {
      std::lock_guard<std::mutex> lk(slot_mu_);
      slots_ += 1;
    }
    cv_slot_.notify_one();
  }

  // box thread: one inner step of box s
  real step(Session* s, const StepCfg& c) override {
    Request r{s, c};
    park(r);
    return r.dt;
  }
  // box thread: the device work of one outer step of box s
  bool outer(Session* s, const OuterJob& j, bool fast, int conn) override {
    Request r{s, StepCfg{}};
    r.job = &j;
    r.fast = fast;
    r.conn = conn;
    park(r);
    return true;
  }

  // box thread: wait until a round has served r
  void park(Request& r) {
    {
      BoxTrace& t = trace();
      if (t.slot_at >= 0) t.host_s += now_s() - t.slot_at;
    }
    // the box's own stream is idle between batched requests except after
    // set-up and curvature-step work; a synchronization costs a marker and
    // its wait, a query only reads the stream's state
    const hipError_t q = hipStreamQuery(r.s->st);
    if (q == hipErrorNotReady) BCHECK(hipStreamSynchronize(r.s->st));
    else BCHECK(q);
    release_slot();
    const int g = group();
    {
      std::unique_lock<std::mutex> lk(mu_);
      if (box_away()) {
        box_away() = false;
        away_[g] -= 1;
        BoxTrace& t = trace();
        if (t.away_since >= 0) { t.away_s += now_s() - t.away_since; t.aways += 1; t.away_since = -1; }
      }
      if (!failure_.empty()) throw std::runtime_error(failure_);
      parked_[g].push_back(&r);
      if (ready_locked(g)) cv_sched_.notify_one();
    }
    // each request is woken on its own (no herd on mu_ after a round)
    BoxTrace& tr = trace();
    const double t0 = now_s();
    if (tr.first_step < 0) tr.first_step = t0;
    {
      std::unique_lock<std::mutex> lk(r.m);
      r.cv.wait(lk, [&] { return r.done; });
    }
    tr.last_step = now_s();
    tr.parked_s += tr.last_step - t0;
    tr.steps += 1;
    tr.wake_s += tr.last_step - done_at_.load(std::memory_order_relaxed);
    acquire_slot();
    tr.slot_at = now_s();
    std::lock_guard<std::mutex> lk(mu_);
    if (!failure_.empty()) throw std::runtime_error(failure_);
  }

  void mark(int phase) override {
    BoxTrace& t = trace();
    if (phase >= 0 && phase < 4 && t.marks[phase] < 0) t.marks[phase] = now_s();
  }

  // box thread: outer-step work until the next step()
  void away() override {
    const int g = group();
    std::lock_guard<std::mutex> lk(mu_);
    if (!box_away()) {
      box_away() = true;
      away_[g] += 1;
      trace().away_since = now_s();
    }
    if (ready_locked(g)) cv_sched_.notify_one();
  }

  void box_finished() {
    release_slot();
    const int g = group();
    std::lock_guard<std::mutex> lk(mu_);
    trace().end = now_s();
    traces.push_back(trace());
    if (box_away()) {
      box_away() = false;
      away_[g] -= 1;
    }
    live_[g] -= 1;
    if (ready_locked(g) || live_total() == 0) cv_sched_.notify_one();
  }

  // (mu_ held) a round of group g can start: every live box of g is parked
  // or away, and one is parked.  Box threads wake the scheduler only then,
  // not at every park.
  bool ready_locked(int g) const {
    return !parked_[g].empty() && (int)parked_[g].size() + away_[g] == live_[g];
  }
  int live_total() const {
    int t = 0;
    for (int g = 0; g < ng_; ++g) t += live_[g];
    return t;
  }

  // calling thread: serve rounds until every box has finished.  A round
  // runs the requests of every parked box of a group once each live box of
  // that group is parked or away (in outer-step work, which then overlaps
  // the round on the GPU).  Each group has its own stream, so one group's
  // round runs while the other's boxes turn around, and the two rounds'
  // kernels may overlap on the GPU (each fills the other's tails); a
  // round's results are delivered once its stream has passed its end event.
  // Which round serves a box's request changes nothing in its arithmetic.
  void run() {
    std::deque<int> inflight;  // lanes with a round on the stream, oldest first
    const bool sync = trace_on();  // per-part timing needs one round at a time
    for (;;) {
      int g = -1;
      std::vector<Request*> round;
      const double w0 = now_s();
      {
        std::unique_lock<std::mutex> lk(mu_);
        auto pick = [&] {
          for (int q = 0; q < ng_; ++q) {
            const int c = (next_ + q) % ng_;
            if (!lanes_[c].busy && ready_locked(c)) return c;
          }
          return -1;
        };
        if (inflight.empty()) {
          cv_sched_.wait(lk, [&] { return live_total() == 0 || pick() >= 0; });
          if (live_total() == 0) break;
        } else if ((int)inflight.size() < ng_) {
          // the other group's boxes are in their host turnaround: take their
          // round as soon as they are parked, else look at the stream again
          cv_sched_.wait_for(lk, std::chrono::microseconds(40), [&] { return pick() >= 0; });
        }
        g = pick();
        if (g >= 0) {
          round.swap(parked_[g]);
          lanes_[g].busy = true;
          next_ = (g + 1) % ng_;
        }
      }
      if (g >= 0) {
        const double w1 = now_s();
        sched_wait_s += w1 - w0;
        if (first_wait_s < 0) first_wait_s = w1 - w0;
        Lane& L = lanes_[g];
        L.round = std::move(round);
        try {
          launch(L);
          sched_launch_s += now_s() - w1;
        } catch (const std::exception& e) {
          {
            std::lock_guard<std::mutex> lk(mu_);
            failure_ = std::string("batched step: ") + e.what();
          }
          L.failed = true;
        }
        inflight.push_back(g);
        if (!sync && !L.failed) continue;  // perhaps the other group is ready too
      }
      // deliver every round whose stream has passed its end event (the
      // groups' streams are independent, so in any order); with both groups
      // in flight and neither done, or one round at a time, wait for one
      if (inflight.empty()) continue;
      bool any = false;
      for (size_t q = 0; q < inflight.size();) {
        Lane& O = lanes_[inflight[q]];
        bool done = O.failed;
        if (!done) {
          const hipError_t e = hipEventQuery(O.ev_done);
          if (e == hipSuccess) done = true;
          else if (e != hipErrorNotReady) { note_failure(e, O); done = true; }
        }
        if (done) {
          inflight.erase(inflight.begin() + q);
          finish(O);
          any = true;
        } else {
          ++q;
        }
      }
      if (!any && sync) {
        Lane& O = lanes_[inflight.front()];
        const hipError_t e = hipEventSynchronize(O.ev_done);
        if (e != hipSuccess) note_failure(e, O);
        inflight.pop_front();
        finish(O);
      } else if (!any && (int)inflight.size() == ng_) {
        std::this_thread::sleep_for(std::chrono::microseconds(20));
      }
    }
  }

  lsmpqs_batch_stats stats{};

 private:
  // One group's buffers: the per-box records and prefix tables of its inner
  // steps (with their captured graph), the jobs and tables of its outer
  // steps, the pinned read-backs, and the round it has in flight.
  struct Lane {
    BatchView view{};
    BatchRec* d_recs = nullptr;
    int *d_count = nullptr, *d_brick = nullptr, *d_counters = nullptr;
    int64_t* d_ghost = nullptr;
    real* d_dts = nullptr;
    BatchRec* h_recs = nullptr;
    int *h_count = nullptr, *h_brick = nullptr;
    int64_t* h_ghost = nullptr;
    real* h_dts = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    OuterJob *d_jobs = nullptr, *h_jobs = nullptr;
    int *d_tab = nullptr, *h_tab = nullptr;
    double *d_redout = nullptr, *h_redout = nullptr;
    hipStream_t st = nullptr;  // the group's own stream (default priority)
    hipEvent_t ev_in0 = nullptr, ev_in1 = nullptr, ev_out0 = nullptr, ev_out1 = nullptr,
               ev_done = nullptr;
    std::vector<Request*> round, inner, outer;
    bool busy = false, failed = false;
  };
  static constexpr int kMaxGroups = 4;
  Lane lanes_[kMaxGroups];
  int ng_ = 1;
  int next_ = 0;  // lane looked at first

  void note_failure(hipError_t e, Lane& L) {
    std::lock_guard<std::mutex> lk(mu_);
    failure_ = std::string("batched step: ") + hipGetErrorString(e);
    L.failed = true;
  }

  void ensure_stream() {
    if (ev0_) return;
    BCHECK(hipEventCreate(&ev0_));
    BCHECK(hipEventCreate(&ev1_));
    for (int q = 0; q < ng_; ++q) {
      Lane& L = lanes_[q];
      BCHECK(hipStreamCreateWithFlags(&L.st, hipStreamNonBlocking));
      BCHECK(hipEventCreate(&L.ev_in0));
      BCHECK(hipEventCreate(&L.ev_in1));
      BCHECK(hipEventCreate(&L.ev_out0));
      BCHECK(hipEventCreate(&L.ev_out1));
      BCHECK(hipEventCreateWithFlags(&L.ev_done, hipEventDisableTiming));
    }
  }

  void configure(const Request& first) {
    scheme_ = first.c.scheme;
    var_ = first.c.var;
    order_ = first.c.order;
    fast_ = sizeof(real) == 4 && !(quirks() & EXACT_FP32);
    nstage_ = order_ <= 1 ? 1 : (order_ == 2 ? 2 : 3);
    ensure_stream();
    for (int q = 0; q < ng_; ++q) {
      Lane& L = lanes_[q];
      BCHECK(hipMalloc((void**)&L.d_recs, sizeof(BatchRec) * nmax_));
      BCHECK(hipMalloc((void**)&L.d_count, sizeof(int) * 4));
      BCHECK(hipMalloc((void**)&L.d_brick, sizeof(int) * (nmax_ + 1)));
      BCHECK(hipMalloc((void**)&L.d_ghost, sizeof(int64_t) * (nmax_ + 1)));
      BCHECK(hipMalloc((void**)&L.d_counters, sizeof(int) * kCounters));
      BCHECK(hipMalloc((void**)&L.d_dts, sizeof(real) * nmax_));
      BCHECK(hipHostMalloc((void**)&L.h_recs, sizeof(BatchRec) * nmax_, hipHostMallocDefault));
      BCHECK(hipHostMalloc((void**)&L.h_count, sizeof(int) * 4, hipHostMallocDefault));
      BCHECK(hipHostMalloc((void**)&L.h_brick, sizeof(int) * (nmax_ + 1), hipHostMallocDefault));
      BCHECK(hipHostMalloc((void**)&L.h_ghost, sizeof(int64_t) * (nmax_ + 1), hipHostMallocDefault));
      BCHECK(hipHostMalloc((void**)&L.h_dts, sizeof(real) * nmax_, hipHostMallocDefault));
      L.view.recs = L.d_recs;
      L.view.count = L.d_count;
      L.view.brick_prefix = L.d_brick;
      L.view.ghost_prefix = L.d_ghost;
      L.view.counters = L.d_counters;
      L.view.dts = L.d_dts;
    }
    int dev = 0, cus = 256;
    BCHECK(hipGetDevice(&dev));
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      cus = 256;
    cus_ = cus;
    bc_grid_ = 4 * cus_;
  }

  void release() {
    // rounds still on the streams (after a failure) finish before their
    // buffers go
    for (Lane& L : lanes_)
      if (L.st) (void)hipStreamSynchronize(L.st);
    for (Lane& L : lanes_) {
      if (L.exec) (void)hipGraphExecDestroy(L.exec);
      if (L.graph) (void)hipGraphDestroy(L.graph);
      void* d[] = {L.d_recs, L.d_count, L.d_brick, L.d_ghost, L.d_counters, L.d_dts, L.d_jobs,
                   L.d_tab, L.d_redout};
      for (void* p : d)
        if (p) (void)hipFree(p);
      void* h[] = {L.h_recs, L.h_count, L.h_brick, L.h_ghost, L.h_dts, L.h_jobs, L.h_tab,
                   L.h_redout};
      for (void* p : h)
        if (p) (void)hipHostFree(p);
      hipEvent_t ev[] = {L.ev_in0, L.ev_in1, L.ev_out0, L.ev_out1, L.ev_done};
      for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
      if (L.st) (void)hipStreamDestroy(L.st);
    }
    for (hipEvent_t e : pass_ev_)
      if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : inner_ev_)
      if (e) (void)hipEventDestroy(e);
    if (ev0_) (void)hipEventDestroy(ev0_);
    if (ev1_) (void)hipEventDestroy(ev1_);
  }

  // records of the round's boxes for brick depth zc, copied to the device
  void upload(Lane& L, const std::vector<Request*>& round, int zc) {
    const int n = (int)round.size();
    L.h_count[0] = n;
    L.h_brick[0] = 0;
    L.h_ghost[0] = 0;
    for (int i = 0; i < n; ++i) {
      StepArgs a = make_step_args(round[i]->s, round[i]->c);
      BatchRec& r = L.h_recs[i];
      r.probe = a.probe;
      int bricks = 0;
      for (int k = 0; k < 3; ++k) {
        r.stage[k] = a.stage[k];
        r.bc[k] = a.bc[k];
        bricks = stage_bricks(r.stage[k], zc);
      }
      L.h_brick[i + 1] = L.h_brick[i] + bricks;
      L.h_ghost[i + 1] = L.h_ghost[i] + ghost_count(a.probe.bx);
    }
    BCHECK(hipMemcpyAsync(L.d_recs, L.h_recs, sizeof(BatchRec) * n, hipMemcpyHostToDevice, L.st));
    BCHECK(hipMemcpyAsync(L.d_count, L.h_count, sizeof(int), hipMemcpyHostToDevice, L.st));
    BCHECK(hipMemcpyAsync(L.d_brick, L.h_brick, sizeof(int) * (n + 1), hipMemcpyHostToDevice, L.st));
    BCHECK(hipMemcpyAsync(L.d_ghost, L.h_ghost, sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice,
                          L.st));
  }

  // timed: events between the parts (host-time tracing; not captured)
  void enqueue_step(Lane& L, bool timed = false) {
    if (timed) mark_inner(0, L.st);
    BCHECK(launch_batch_reset(L.d_counters, kCounters, L.st));
    BCHECK(launch_probe_batch(L.view, scheme_, nmax_, L.st));
    BCHECK(hipMemcpyAsync(L.h_dts, L.d_dts, sizeof(real) * nmax_, hipMemcpyDeviceToHost, L.st));
    for (int k = 0; k < nstage_; ++k) {
      const bool fin = k == nstage_ - 1;
      if (timed) mark_inner(1 + 2 * k, L.st);
      BCHECK(launch_stage_batch(L.view, k, scheme_, var_, k + 1, fin, fast_, grid_, L.st));
      if (timed) mark_inner(2 + 2 * k, L.st);
      BCHECK(launch_bc_batch(L.view, k, fin, bc_grid_, L.st));
    }
    if (timed) mark_inner(1 + 2 * nstage_, L.st);
  }
  void mark_inner(int q, hipStream_t st) {
    if (!inner_ev_[q]) BCHECK(hipEventCreate(&inner_ev_[q]));
    BCHECK(hipEventRecord(inner_ev_[q], st));
  }

  // brick depth and workgroup count: fastest first stage over this round
  void autotune(Lane& L, const std::vector<Request*>& round) {
    const int depths[] = {8, 16, 32, 64};
    const int per_cu[] = {1, 2, 4};
    float best = 1e30f;
    upload(L, round, depths[0]);
    BCHECK(launch_probe_batch(L.view, scheme_, nmax_, L.st));  // stage 1 reads this step's dt
    for (int zc : depths) {
      upload(L, round, zc);
      for (int pc : per_cu) {
        const int grid = pc * cus_;
        float t = 1e30f;
        for (int rep = 0; rep < 2; ++rep) {
          BCHECK(launch_batch_reset(L.d_counters, kCounters, L.st));
          BCHECK(hipEventRecord(ev0_, L.st));
          BCHECK(launch_stage_batch(L.view, 0, scheme_, var_, 1, nstage_ == 1, fast_, grid, L.st));
          BCHECK(hipEventRecord(ev1_, L.st));
          BCHECK(hipEventSynchronize(ev1_));
          float ms = 0;
          BCHECK(hipEventElapsedTime(&ms, ev0_, ev1_));
          t = std::min(t, ms);
        }
        if (t < best) {
          best = t;
          zc_ = zc;
          grid_ = grid;
        }
      }
    }
    // order 1 wrote the imposed result into B; the real step recomputes it
    stats.brick_depth = zc_;
    stats.workgroups = grid_;
    stats.autotune_best_ms = best;
    ratio_ = (double)bricks_of(round, zc_) / grid_;
  }

  // bricks of the round's boxes at depth zc (one stage)
  static int bricks_of(const std::vector<Request*>& round, int zc) {
    int tot = 0;
    for (const Request* r : round) {
      StageArgs a{};
      a.bx = r->s->bx;
      tot += stage_bricks(a, zc);
    }
    return tot;
  }
  // Brick depth of a later round: the deepest whose bricks per persistent
  // workgroup are at least the tuned round's, so that a round of fewer or
  // smaller boxes keeps the queue as well balanced over the workgroups (the
  // graph's workgroup count stays fixed; the depth only changes the brick
  // table, and no voxel's arithmetic).
  int round_depth(const std::vector<Request*>& round) const {
    for (int zc : {64, 32, 16})
      if (zc <= zc_ && bricks_of(round, zc) >= ratio_ * grid_) return zc;
    return 8;
  }

  void capture(Lane& L) {
    BCHECK(hipStreamSynchronize(L.st));  // once per group: its uploads and tuning done
    BCHECK(hipStreamBeginCapture(L.st, hipStreamCaptureModeThreadLocal));
    try {
      enqueue_step(L);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(L.st, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    BCHECK(hipStreamEndCapture(L.st, &L.graph));
    BCHECK(hipGraphInstantiate(&L.exec, L.graph, nullptr, nullptr, 0));
  }

  // enqueue the round of lane L (outer steps, then inner steps) and the
  // event that marks its end; finish(L) delivers the results
  void launch(Lane& L) {
    L.inner.clear();
    L.outer.clear();
    L.failed = false;
    for (Request* r : L.round) (r->job ? L.outer : L.inner).push_back(r);
    ensure_stream();
    if (!L.outer.empty()) launch_outer(L);
    if (!L.inner.empty()) launch_inner(L);
    BCHECK(hipEventRecord(L.ev_done, L.st));
  }

  // results of lane L's round (its stream work is complete, or it failed)
  // to the boxes, which are then woken
  void finish(Lane& L) {
    if (!L.failed) {
      try {
        collect(L);
      } catch (const std::exception& e) {
        std::lock_guard<std::mutex> lk(mu_);
        failure_ = std::string("batched step: ") + e.what();
      }
    }
    {
      std::lock_guard<std::mutex> lk(mu_);
      L.busy = false;
    }
    done_at_.store(now_s(), std::memory_order_relaxed);
    // notify under the request's lock: once it is released the box may
    // return and its request (on the box's stack) is gone
    for (Request* r : L.round) {
      std::lock_guard<std::mutex> g(r->m);
      r->done = true;
      r->cv.notify_one();
    }
    L.round.clear();
  }
  void collect(Lane& L) {
    const bool tr = trace_on();
    if (!L.outer.empty()) {
      const int n = (int)L.outer.size();
      float ms = 0;
      BCHECK(hipEventElapsedTime(&ms, L.ev_out0, L.ev_out1));
      stats.batched_outer += 1;
      stats.box_outer += n;
      stats.batched_outer_ms += ms;
      if (tr)
        for (int q = 0; q < 4; ++q) {
          float m = 0;
          BCHECK(hipEventElapsedTime(&m, pass_ev_[q], pass_ev_[q + 1]));
          pass_ms[q] += m;
        }
      for (int i = 0; i < n; ++i) {
        L.outer[i]->s->h_red[0] = L.h_redout[2 * i];
        L.outer[i]->s->h_red[1] = L.h_redout[2 * i + 1];
      }
    }
    if (!L.inner.empty()) {
      float ms = 0;
      BCHECK(hipEventElapsedTime(&ms, L.ev_in0, L.ev_in1));
      if (tr)
        for (int q = 0; q < 1 + 2 * nstage_; ++q) {
          float m = 0;
          BCHECK(hipEventElapsedTime(&m, inner_ev_[q], inner_ev_[q + 1]));
          inner_ms[q] += m;
        }
      stats.batched_steps += 1;
      stats.batched_ms += ms;
      stats.box_steps += (long long)L.inner.size();
      for (size_t i = 0; i < L.inner.size(); ++i) L.inner[i]->dt = L.h_dts[i];
    }
  }

  // One outer step of each box of the lane's round: every pass one launch
  // over all of them, in the single-box order (lsm_driver.hip: trap_dev,
  // reinit_dev of mask_w and mask_nw, trapped-phase velocities, merge and
  // floor, reinit_dev of phi with the outer reductions), then the
  // reductions read back.
  void launch_outer(Lane& L) {
    const std::vector<Request*>& reqs = L.outer;
    const int n = (int)reqs.size();
    const int conn = reqs[0]->conn;
    const bool fast = reqs[0]->fast;
    for (Request* r : reqs)
      if (r->conn != conn || r->fast != fast)
        throw std::runtime_error("batched outer steps need one connectivity and arithmetic mode");
    if (!L.d_jobs) {
      BCHECK(hipMalloc((void**)&L.d_jobs, sizeof(OuterJob) * nmax_));
      BCHECK(hipMalloc((void**)&L.d_tab, sizeof(int) * kTables * (nmax_ + 1)));
      BCHECK(hipMalloc((void**)&L.d_redout, sizeof(double) * 2 * nmax_));
      BCHECK(hipHostMalloc((void**)&L.h_jobs, sizeof(OuterJob) * nmax_, hipHostMallocDefault));
      BCHECK(hipHostMalloc((void**)&L.h_tab, sizeof(int) * kTables * (nmax_ + 1),
                           hipHostMallocDefault));
      BCHECK(hipHostMalloc((void**)&L.h_redout, sizeof(double) * 2 * nmax_, hipHostMallocDefault));
    }
    if (n > nmax_) throw std::runtime_error("more outer-step requests than boxes");
    bool velocities = false;
    for (int i = 0; i < n; ++i) {
      L.h_jobs[i] = *reqs[i]->job;
      velocities = velocities || L.h_jobs[i].vn != nullptr;
    }
    // per launch kind: a prefix table of the boxes' single-box grids
    auto table = [&](int k, auto count) {
      int* p = L.h_tab + k * (nmax_ + 1);
      p[0] = 0;
      for (int i = 0; i < n; ++i) p[i + 1] = p[i] + count(L.h_jobs[i]);
      return Multi{L.d_tab + k * (nmax_ + 1), n, p[n]};
    };
    const Multi trap = table(0, [](const OuterJob& j) { return trap_pass_blocks(j.ta.bx); });
    const Multi force = table(1, [](const OuterJob&) { return force_pass_blocks(); });
    const Multi one = table(2, [](const OuterJob&) { return 1; });
    const Multi tiles = table(3, [](const OuterJob& j) { return label_tile_blocks(j.ta.bx); });
    const Multi jump = table(4, [](const OuterJob& j) { return label_jump_blocks(j.ta.roots_cap); });
    const Multi rows = table(5, [](const OuterJob& j) { return reinit_pass_blocks(j.ta.bx); });
    const Multi red = table(6, [](const OuterJob& j) { return j.nred; });
    const Multi stream = table(7, [](const OuterJob& j) { return stream_pass_blocks(j.ta.bx.size()); });
    const Multi ghosts = table(8, [](const OuterJob& j) { return bc_pass_blocks(j.ta.bx); });
    OuterJob* jobs = L.d_jobs;
    BCHECK(hipMemcpyAsync(jobs, L.h_jobs, sizeof(OuterJob) * n, hipMemcpyHostToDevice, L.st));
    BCHECK(hipMemcpyAsync(L.d_tab, L.h_tab, sizeof(int) * kTables * (nmax_ + 1),
                          hipMemcpyHostToDevice, L.st));
    BCHECK(hipEventRecord(L.ev_out0, L.st));
    const bool tr = trace_on();
    if (tr) mark_pass(0, L.st);
    BCHECK(launch_trap_multi(jobs, trap, force, one, tiles, jump, conn, L.st));
    if (tr) mark_pass(1, L.st);
    for (int pass = 0; pass < 2; ++pass) {
      BCHECK(launch_reinit_multi(jobs, pass, rows, fast, L.st));
      BCHECK(launch_bc_multi(jobs, pass, ghosts, L.st));
    }
    if (tr) mark_pass(2, L.st);
    if (velocities) BCHECK(launch_trap_velocity_multi(jobs, rows, L.st));
    BCHECK(launch_merge_floor_multi(jobs, stream, L.st));
    if (tr) mark_pass(3, L.st);
    BCHECK(launch_reinit_reduce_multi(jobs, red, one, fast, L.st));
    BCHECK(launch_bc_multi(jobs, 2, ghosts, L.st));
    BCHECK(launch_gather_red(jobs, n, L.d_redout, L.st));
    if (tr) mark_pass(4, L.st);
    BCHECK(hipEventRecord(L.ev_out1, L.st));
    BCHECK(hipMemcpyAsync(L.h_redout, L.d_redout, sizeof(double) * 2 * n, hipMemcpyDeviceToHost,
                          L.st));
  }

  // per-pass timing of the batched outer steps (LSMPQS_BATCH_TRACE=1)
  static bool trace_on() {
    const char* e = getenv("LSMPQS_BATCH_TRACE");
    return e && *e && atoi(e) != 0;
  }
  void mark_pass(int q, hipStream_t st) {
    if (!pass_ev_[q]) BCHECK(hipEventCreate(&pass_ev_[q]));
    BCHECK(hipEventRecord(pass_ev_[q], st));
  }

 public:
  double pass_ms[4] = {0, 0, 0, 0};  // trapping, mask reinitializations, merge, phi reinitialization
  double inner_ms[7] = {0, 0, 0, 0, 0, 0, 0};  // probe, then per stage: stage kernel, ghost extrapolation

 private:
  hipEvent_t pass_ev_[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  hipEvent_t inner_ev_[8] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};

  void launch_inner(Lane& L) {
    const std::vector<Request*>& round = L.inner;
    if (!configured_) {
      configure(*round[0]);
      configured_ = true;
    }
    for (Request* r : round)
      if (r->c.scheme != scheme_ || r->c.var != var_ || r->c.order != order_)
        throw std::runtime_error(
            "all boxes of a batch need the same spatial scheme, time order and contact-angle path");
    if (zc_ == 0) autotune(L, round);
    const int zc = round_depth(round);
    upload(L, round, zc);
    if (!L.exec) capture(L);
    BCHECK(hipEventRecord(L.ev_in0, L.st));
    if (trace_on()) enqueue_step(L, true);  // the same launches, uncaptured, with events between them
    else BCHECK(hipGraphLaunch(L.exec, L.st));
    BCHECK(hipEventRecord(L.ev_in1, L.st));
  }

  std::mutex mu_, slot_mu_;
  std::condition_variable cv_sched_, cv_slot_;
  int slots_ = 1 << 30;
  std::vector<Request*> parked_[kMaxGroups];
  int live_[kMaxGroups] = {0, 0, 0, 0};
  int away_[kMaxGroups] = {0, 0, 0, 0};  // live boxes between leaving an inner loop and their next request
  static bool& box_away() {
    static thread_local bool a = false;
    return a;
  }
  std::string failure_;
  std::atomic<double> done_at_{0};  // when the last round's requests were marked done

  int nmax_;
  int scheme_ = 0, order_ = 0, nstage_ = 0;
  bool var_ = false, fast_ = false;
  int cus_ = 256, zc_ = 0, grid_ = 0, bc_grid_ = 1024;
  double ratio_ = 0;  // bricks per persistent workgroup of the tuned round
  bool configured_ = false;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;  // autotuning
};

}  // namespace
}  // namespace lsm

extern "C" int lsmpqs_batch_model3d(int n, Options** options, QSS_DataArrays** p, Grid** g,
                                    FILE** fp_out, int drain, lsmpqs_batch_stats* stats) {
  using namespace lsm;
  set_error("");
  if (n < 1 || !options || !p || !g) {
    set_error("lsmpqs_batch_model3d: n >= 1 records with options, arrays and grids required");
    return -2;
  }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < i; ++j)
      if (p[i] == p[j] || options[i] == options[j]) {
        set_error("lsmpqs_batch_model3d: every box needs its own QSS_DataArrays and Options");
        return -2;
      }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    set_error("lsmpqs_batch_model3d: no HIP device");
    return -2;
  }
  // box groups, each with its own stream and rounds (LSMPQS_BATCH_GROUPS,
  // 1 - 4, default 4: 64 boxes of a 768^3 pack take 1.96 s with one group,
  // 1.70 - 1.75 s with two, 1.66 - 1.71 s with three, 1.65 - 1.70 s with four)
  int groups = 4;
  {
    const char* e = getenv("LSMPQS_BATCH_GROUPS");
    if (e && *e) groups = atoi(e);
  }
  Batch batch(n, groups);
  // boxes doing host-driven work at once (LSMPQS_BATCH_THREADS, default 16,
  // a GPU's share of host cores; 0: all).  64 boxes of a 768^3 pack: the
  // batched drainage takes 2.09 s with 64 active box threads, 2.10 s with
  // 16, 2.12 s with 8 and 2.38 s with 4 -- the limit keeps the host threads
  // that contend for the runtime at the cores there are, at no cost.
  {
    const char* e = getenv("LSMPQS_BATCH_THREADS");
    int k = (e && *e) ? atoi(e) : 16;
    batch.set_slots(k > 0 ? k : n);
  }
  std::vector<std::string> errors(n);
  std::vector<std::thread> threads;
  threads.reserve(n);
  for (int i = 0; i < n; ++i) {
    threads.emplace_back([&, i] {
      Batch::group() = Batch::group_of(i, n, batch.groups());
      Batch::trace().start = Batch::now_s();
      batch.acquire_slot();
      try {
        if (hipSetDevice(dev) != hipSuccess) throw std::runtime_error("hipSetDevice failed");
        batch_hook() = &batch;
        model_run(options[i], p[i], g[i], fp_out ? fp_out[i] : nullptr, drain != 0);
      } catch (const std::exception& e) {
        errors[i] = e.what();
      }
      batch_hook() = nullptr;
      batch.box_finished();
    });
  }
  const double t_run = Batch::now_s();
  batch.run();
  for (auto& t : threads) t.join();
  if (getenv("LSMPQS_BATCH_TRACE") && atoi(getenv("LSMPQS_BATCH_TRACE")) != 0) {
    // per box: set-up (start to first step), parked (waiting for batched
    // steps), own work between steps, tail (last step to end), slot waits
    double su = 0, pk = 0, own = 0, tl = 0, sl = 0, msu = 0, mtl = 0, mend = 0;
    double ph[5] = {0, 0, 0, 0, 0};  // session, model set-up, coefficients, first trapping, to step
    double aw = 0, wk = 0, hs = 0;
    long long naw = 0, nst = 0;
    for (const auto& t : batch.traces) {
      double prev = t.start;
      for (int q = 0; q < 4; ++q)
        if (t.marks[q] >= 0) { ph[q] += t.marks[q] - prev; prev = t.marks[q]; }
      if (t.first_step >= 0) ph[4] += t.first_step - prev;
      aw += t.away_s;
      naw += t.aways;
      wk += t.wake_s; hs += t.host_s; nst += t.steps;
      const double setup = (t.first_step < 0 ? t.end : t.first_step) - t.start;
      const double tail = t.end - (t.first_step < 0 ? t.end : t.last_step);
      su += setup; pk += t.parked_s; tl += tail; sl += t.slot_s;
      own += (t.end - t.start) - setup - tail - t.parked_s;
      msu = std::max(msu, setup); mtl = std::max(mtl, tail); mend = std::max(mend, t.end - t_run);
    }
    fprintf(stderr,
            "lsmpqs batch trace: %d boxes, wall %.3f s; scheduler waiting %.3f s (%.3f before the "
            "first round), in steps %.3f s; per box sums: set-up %.3f (max %.3f), parked %.3f, own "
            "work %.3f, tail %.3f (max %.3f), slot waits %.3f s\n",
            n, mend, batch.sched_wait_s, batch.first_wait_s, batch.sched_launch_s, su, msu, pk,
            own, tl, mtl, sl);
    fprintf(stderr,
            "lsmpqs batch trace: set-up sums: session %.3f, model set-up %.3f, coefficients "
            "%.3f, first trapping %.3f, to the first step %.3f s; outer-step work %.3f s in %lld "
            "stretches\n",
            ph[0], ph[1], ph[2], ph[3], ph[4], aw, naw);
    fprintf(stderr,
            "lsmpqs batch trace: per served request: wake-up %.1f us, host work to the next "
            "request %.1f us (%lld requests)\n",
            1e6 * wk / std::max(1LL, nst), 1e6 * hs / std::max(1LL, nst), nst);
    fprintf(stderr,
            "lsmpqs batch trace: batched outer steps: trapping %.3f, mask reinitializations "
            "%.3f, merge and floor %.3f, phi reinitialization and reductions %.3f s\n",
            batch.pass_ms[0] * 1e-3, batch.pass_ms[1] * 1e-3, batch.pass_ms[2] * 1e-3,
            batch.pass_ms[3] * 1e-3);
    fprintf(stderr,
            "lsmpqs batch trace: batched inner steps (uncaptured): probe %.3f, stages %.3f / "
            "%.3f / %.3f, ghost extrapolations %.3f / %.3f / %.3f s\n",
            batch.inner_ms[0] * 1e-3, batch.inner_ms[1] * 1e-3, batch.inner_ms[3] * 1e-3,
            batch.inner_ms[5] * 1e-3, batch.inner_ms[2] * 1e-3, batch.inner_ms[4] * 1e-3,
            batch.inner_ms[6] * 1e-3);
  }
  if (stats) {
    *stats = batch.stats;
    stats->boxes = n;
  }
  for (int i = 0; i < n; ++i)
    if (!errors[i].empty()) {
      std::string m = "box " + std::to_string(i) + ": " + errors[i];
      set_error(m);
      fprintf(stderr, "lsmpqs_hip: %s\n", m.c_str());
      return -2;
    }
  return 0;
}
