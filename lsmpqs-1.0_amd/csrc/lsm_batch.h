// lsm_batch.h -- what the model driver and the batched runtime share: the
// configuration of one inner time step, the kernel arguments it expands to,
// and the hook through which a box thread of a batched run hands its inner
// steps to the batch scheduler.
#pragma once
#include "lsm_session.h"

namespace lsm {

// Scheme and coefficients of the inner steps of one constCurvModel3d call.
struct StepCfg {
  int scheme;
  bool var;
  int order;
  real a_const, b_const;
  real norm_sign;
  real bmax, umax, overlap;
  int conn, slabs, quirk_eno3;
  real dx, dy, dz;
};

// Kernel arguments of one inner step of one box: the time-step probe, the
// stages 1..order and the ghost extrapolation after each (the last one with
// the imposition floor).  For order 1 the result lands in B, and the caller
// swaps phi and B afterwards.
struct StepArgs {
  ProbeArgs probe;
  StageArgs stage[3];
  BcArgs bc[3];
  int order;
};
StepArgs make_step_args(Session* s, const StepCfg& c);

// Set by the batched runtime in each box thread: the inner loop then asks
// the scheduler for every inner step instead of launching it itself.
struct BatchHook {
  virtual ~BatchHook() {}
  // runs one inner step of box s (after the box's own stream is idle) and
  // returns its dt; the order-1 phi/B swap is left to the caller
  virtual real step(Session* s, const StepCfg& c) = 0;
  // the box has left its inner loop for outer-step work (trapping,
  // reinitialization, the next curvature step); batched steps of the other
  // boxes need not wait for it until it asks for its next step
  virtual void away() {}
  // runs the device work of one outer step of box s (after the box's own
  // stream is idle) with the other boxes' and leaves its reductions in
  // s->h_red; false: not taken, the caller runs it itself
  virtual bool outer(Session* s, const OuterJob& j, bool fast, int conn) {
    (void)s; (void)j; (void)fast; (void)conn;
    return false;
  }
  // progress marks of the box's set-up (host-time tracing only)
  virtual void mark(int phase) { (void)phase; }
};
BatchHook*& batch_hook();  // thread-local, nullptr outside box threads

// drain_model3d (drain) / imbibe_model3d body on the calling thread
void model_run(Options* o, QSS_DataArrays* p, Grid* g, FILE* fp, bool drain);

}  // namespace lsm
