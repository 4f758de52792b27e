// lsm_session.h -- device-resident state of one QSS_DataArrays record.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>
#include "qss_data_arrays.h"
#include "qss_grid.h"
#include "qss_options.h"
#include "lsm_kernels.h"

namespace lsm {

// Reference-behaviour switches (all on by default: results match the reference).
enum Quirk : unsigned {
  QUIRK_ENO3_READ = 1u,    // ENO3 x line reads phi(i+3,j,j)
  QUIRK_WENO5_SIGN = 2u,   // WENO5 normal term enters with + sign
  QUIRK_STALE_NORM = 4u,   // |vel_n| < 1e-5 reuses the previous voxel's normal term
  EXACT_FP32 = 8u,         // single precision keeps the reference's expression order
};

struct Timers {
  double stage_ms[4] = {0, 0, 0, 0};  // summed stage-kernel time per stage (1..3), [0] unused
  int64_t stage_launches[4] = {0, 0, 0, 0};
  double inner_ms = 0;                 // wall time of inner steps (event to event)
  int64_t inner_steps = 0;
  int64_t outer_steps = 0;
};

struct GraphCache;  // replayed inner-step graphs of a session (lsm_driver.hip)
void graph_cache_free(GraphCache* c);

struct Session {
  const QSS_DataArrays* key = nullptr;
  Box bx{};
  int64_t n = 0;
  hipStream_t st = nullptr;
  // device fields
  real *phi = nullptr, *A = nullptr, *B = nullptr, *S = nullptr, *D = nullptr, *phi_prev = nullptr;
  real *mask = nullptr, *mask_w = nullptr, *mask_nw = nullptr, *phi_extra = nullptr;
  real* floor = nullptr;  // imposition floor of the last RK stage (launch_floor)
  real *vn = nullptr, *b = nullptr, *vx = nullptr, *vy = nullptr, *vz = nullptr;
  real *gx = nullptr, *gy = nullptr, *gz = nullptr;
  int* labels = nullptr;
  int* troot = nullptr;
  int* roots = nullptr;   // tile-root list of the labelling: [count, roots...]
  int64_t roots_cap = 0;
  uint8_t* bin = nullptr;
  double *part = nullptr, *red = nullptr;
  real* dt = nullptr;
  int* count = nullptr;
  int64_t *stale_idx = nullptr, *stale_src = nullptr, *stale_list = nullptr;
  int stale_cap = 0, nstale = 0;
  // pinned host staging
  real* h_dt = nullptr;
  double* h_red = nullptr;
  int* h_count = nullptr;
  hipEvent_t ev_dt = nullptr, ev_a = nullptr, ev_b = nullptr;
  std::vector<hipEvent_t> stage_ev;  // timing pairs when profiling is on
  bool profile = false;
  // the next inner step writes its result into phi_prev's buffer, after
  // which phi and phi_prev are swapped: phi_prev keeps the state before the
  // step without a copy (RK orders 2 and 3, whose last stage can write to a
  // buffer other than its u^n)
  bool redirect = false;
  GraphCache* graphs = nullptr;
  Timers tm;
  // device memory: one allocation per group of fields (the fields above
  // point into these), freed with the session
  std::vector<void*> slabs;
};

// error state of the calling thread
void set_error(const std::string& msg);
const char* last_error();

Session* session_for(const QSS_DataArrays* p, const Grid* g);  // creates or resizes
void session_release(const QSS_DataArrays* p);
Box box_of(const Grid* g);

unsigned quirks();
void set_quirks(unsigned q);
void set_slabs(int s);
int dt_slabs(const Grid* g);
int dt_slab_count(int planes);  // the same for a fill box of that many slab planes

// 2D forms of the shared helper entry points (lsm_model2d.hip); they throw
// std::runtime_error on a device or input error
void api2d_bc(real* phi, const Grid* g, int bdry_location_idx);
void api2d_trap(QSS_DataArrays* p, Grid* g, Options* o);
void api2d_conn(QSS_DataArrays* p, Grid* g);
void api2d_disconn(real* data, const Grid* g);
void api2d_main_ind(Options* o, const Grid* g);
void api2d_overlap(Options* o, const Grid* g);

void ensure_var(Session* s);     // vn, b, vx, vy, vz
void ensure_grad(Session* s);    // mask gradient and reservoir box
void ensure_labels(Session* s);  // component labels and binary field

}  // namespace lsm
