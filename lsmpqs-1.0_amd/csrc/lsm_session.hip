// lsm_session.hip -- device buffers of a QSS_DataArrays record, error state
// and run-time switches.
#include <stdlib.h>
#include <string.h>
#include <unistd.h>
#include <map>
#include <mutex>
#include <stdexcept>
#include "lsm_session.h"

namespace lsm {

static thread_local std::string g_err;
void set_error(const std::string& m) { g_err = m; }
const char* last_error() { return g_err.c_str(); }

static std::mutex g_mu;
static std::map<const QSS_DataArrays*, Session*> g_sessions;

static unsigned g_quirks = QUIRK_ENO3_READ | QUIRK_WENO5_SIGN | QUIRK_STALE_NORM;
static int g_slabs = 0;  // 0: derive from the environment

unsigned quirks() { return g_quirks; }
void set_quirks(unsigned q) { g_quirks = q; }
void set_slabs(int s) { g_slabs = s; }

// Number of slabs the reference's OpenMP team would use: LSMPQS_DT_SLABS,
// else OMP_NUM_THREADS, else the online CPU count, clamped to the number of
// fill-box planes (z-planes in 3D, j-rows in 2D).
int dt_slabs(const Grid* g) { return dt_slab_count(g->khi_fb - g->klo_fb + 1); }

int dt_slab_count(int nsl) {
  int s = g_slabs;
  if (s <= 0) {
    const char* e = getenv("LSMPQS_DT_SLABS");
    if (!e || !*e) e = getenv("OMP_NUM_THREADS");
    if (e && *e) s = atoi(e);
    if (s <= 0) s = (int)sysconf(_SC_NPROCESSORS_ONLN);
  }
  if (s > nsl) s = nsl;
  if (s > 1024) s = 1024;
  if (s < 1) s = 1;
  return s;
}

Box box_of(const Grid* g) {
  Box b;
  b.nx = g->grid_dims_ghostbox[0];
  b.ny = g->grid_dims_ghostbox[1];
  b.nz = g->grid_dims_ghostbox[2];
  b.ilo = g->ilo_fb; b.ihi = g->ihi_fb;
  b.jlo = g->jlo_fb; b.jhi = g->jhi_fb;
  b.klo = g->klo_fb; b.khi = g->khi_fb;
  return b;
}

#define SCHECK(x)                                                                \
  do {                                                                           \
    hipError_t e_ = (x);                                                         \
    if (e_ != hipSuccess)                                                        \
      throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename P>
static void dalloc(P*& p, int64_t count) {
  if (!p) SCHECK(hipMalloc((void**)&p, (size_t)count * sizeof(P)));
}

static void free_all(Session* s) {
  if (s->graphs) {
    graph_cache_free(s->graphs);
    s->graphs = nullptr;
  }
  void* bufs[] = {s->phi, s->A, s->B, s->S, s->D, s->phi_prev, s->mask, s->mask_w, s->mask_nw,
                  s->floor, s->phi_extra, s->vn, s->b, s->vx, s->vy, s->vz, s->gx, s->gy, s->gz,
                  s->labels, s->troot, s->roots, s->bin, s->part, s->red, s->dt, s->count,
                  s->stale_idx, s->stale_src, s->stale_list};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  if (s->h_dt) (void)hipHostFree(s->h_dt);
  if (s->h_red) (void)hipHostFree(s->h_red);
  if (s->h_count) (void)hipHostFree(s->h_count);
  if (s->ev_dt) (void)hipEventDestroy(s->ev_dt);
  if (s->ev_a) (void)hipEventDestroy(s->ev_a);
  if (s->ev_b) (void)hipEventDestroy(s->ev_b);
  for (hipEvent_t e : s->stage_ev) (void)hipEventDestroy(e);
  if (s->st) (void)hipStreamDestroy(s->st);
}

Session* session_for(const QSS_DataArrays* p, const Grid* g) {
  std::lock_guard<std::mutex> lk(g_mu);
  Box bx = box_of(g);
  auto it = g_sessions.find(p);
  if (it != g_sessions.end()) {
    Session* s = it->second;
    if (s->bx.nx == bx.nx && s->bx.ny == bx.ny && s->bx.nz == bx.nz) {
      s->bx = bx;
      return s;
    }
    free_all(s);
    delete s;
    g_sessions.erase(it);
  }
  if (bx.ilo < 3 || bx.jlo < 3 || bx.klo < 3 || bx.ihi > bx.nx - 4 || bx.jhi > bx.ny - 4 ||
      bx.khi > bx.nz - 4 || bx.ihi - bx.ilo < 3 || bx.jhi - bx.jlo < 3 || bx.khi - bx.klo < 3)
    throw std::runtime_error("grid needs at least 3 ghost layers and 4 fill voxels per axis");
  if ((int64_t)bx.nx * bx.ny * bx.nz >= ((int64_t)1 << 31))
    throw std::runtime_error("ghost box exceeds 2^31 voxels (int labels, as in the reference)");
  Session* s = new Session();
  try {
    s->key = p;
    s->bx = bx;
    s->n = bx.size();
    SCHECK(hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking));
    int64_t n = s->n;
    dalloc(s->phi, n); dalloc(s->A, n); dalloc(s->B, n); dalloc(s->S, n); dalloc(s->D, n);
    dalloc(s->phi_prev, n); dalloc(s->mask, n); dalloc(s->mask_w, n); dalloc(s->mask_nw, n);
    dalloc(s->floor, n);
    dalloc(s->part, reduce_partials()); dalloc(s->red, 4); dalloc(s->dt, 4); dalloc(s->count, 4);
    SCHECK(hipHostMalloc((void**)&s->h_dt, 4 * sizeof(real), hipHostMallocDefault));
    SCHECK(hipHostMalloc((void**)&s->h_red, 4 * sizeof(double), hipHostMallocDefault));
    SCHECK(hipHostMalloc((void**)&s->h_count, 4 * sizeof(int), hipHostMallocDefault));
    SCHECK(hipEventCreate(&s->ev_dt));
    SCHECK(hipEventCreate(&s->ev_a));
    SCHECK(hipEventCreate(&s->ev_b));
    // zero every field so ghost/unused entries are defined
    real* f[] = {s->phi, s->A, s->B, s->S, s->D, s->phi_prev, s->mask, s->mask_w, s->mask_nw,
                 s->floor};
    for (real* q : f) SCHECK(launch_fill(q, real(0), n, s->st));
    SCHECK(hipStreamSynchronize(s->st));
  } catch (...) {
    free_all(s);
    delete s;
    throw;
  }
  g_sessions[p] = s;
  return s;
}

// lazily allocated fields of the variable-coefficient path and labelling
void ensure_var(Session* s) {
  int64_t n = s->n;
  bool fresh = !s->vn;
  dalloc(s->vn, n); dalloc(s->b, n); dalloc(s->vx, n); dalloc(s->vy, n); dalloc(s->vz, n);
  if (fresh) {
    real* f[] = {s->vn, s->b, s->vx, s->vy, s->vz};
    for (real* q : f) SCHECK(launch_fill(q, real(0), n, s->st));
  }
}
void ensure_grad(Session* s) {
  int64_t n = s->n;
  bool fresh = !s->gx;
  dalloc(s->gx, n); dalloc(s->gy, n); dalloc(s->gz, n); dalloc(s->phi_extra, n);
  if (fresh) {
    real* f[] = {s->gx, s->gy, s->gz, s->phi_extra};
    for (real* q : f) SCHECK(launch_fill(q, real(0), n, s->st));
  }
}
void ensure_labels(Session* s) {
  dalloc(s->labels, s->n);
  dalloc(s->troot, s->n);
  dalloc(s->bin, s->n);
  // room for n/16 tile roots (a 64x16x16 tile holds 16384 voxels); a labelling
  // with more falls back to scanning troot
  s->roots_cap = s->n / 16 + 1;
  dalloc(s->roots, s->roots_cap + 1);
}

void session_release(const QSS_DataArrays* p) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_sessions.find(p);
  if (it == g_sessions.end()) return;
  (void)hipStreamSynchronize(it->second->st);
  free_all(it->second);
  delete it->second;
  g_sessions.erase(it);
}

}  // namespace lsm
