// lsm_session.hip -- device buffers of a QSS_DataArrays record, error state
// and run-time switches.
#include <stdlib.h>
#include <string.h>
#include <unistd.h>
#include <map>
#include <initializer_list>
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

// One device allocation for a group of fields, each 256-byte aligned: a
// session makes four such allocations (core, variable coefficients, mask
// gradient, labelling) instead of one per field, so that many boxes set up
// at once (batched runs) do not queue on the allocator.
struct SlabPart {
  void** p;
  size_t bytes;
};
static void slab_alloc(Session* s, std::initializer_list<SlabPart> parts) {
  size_t tot = 0;
  for (const SlabPart& q : parts) tot += (q.bytes + 255) & ~(size_t)255;
  char* base = nullptr;
  SCHECK(hipMalloc((void**)&base, tot));
  s->slabs.push_back(base);
  size_t off = 0;
  for (const SlabPart& q : parts) {
    *q.p = base + off;
    off += (q.bytes + 255) & ~(size_t)255;
  }
}
template <typename P>
static SlabPart part(P*& p, int64_t count) {
  return SlabPart{(void**)&p, (size_t)count * sizeof(P)};
}

static void free_all(Session* s) {
  if (s->graphs) {
    graph_cache_free(s->graphs);
    s->graphs = nullptr;
  }
  void* bufs[] = {s->stale_idx, s->stale_src, s->stale_list};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  for (void* p : s->slabs) (void)hipFree(p);
  s->slabs.clear();
  if (s->h_dt) (void)hipHostFree(s->h_dt);
  if (s->ev_dt) (void)hipEventDestroy(s->ev_dt);
  if (s->ev_a) (void)hipEventDestroy(s->ev_a);
  if (s->ev_b) (void)hipEventDestroy(s->ev_b);
  for (hipEvent_t e : s->stage_ev) (void)hipEventDestroy(e);
  if (s->st) (void)hipStreamDestroy(s->st);
}

// A new session's buffers, allocated and zeroed outside the session-map
// lock (other threads keep creating and finding their own sessions).
static Session* session_new(const QSS_DataArrays* p, const Box& bx) {
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
    slab_alloc(s, {part(s->phi, n), part(s->A, n), part(s->B, n), part(s->S, n), part(s->D, n),
                   part(s->phi_prev, n), part(s->mask, n), part(s->mask_w, n),
                   part(s->mask_nw, n), part(s->floor, n), part(s->part, reduce_partials()),
                   part(s->red, 4), part(s->dt, 4), part(s->count, 4)});
    // pinned read-back slots: dt, the two reductions, a counter
    char* h = nullptr;
    SCHECK(hipHostMalloc((void**)&h, 256, hipHostMallocDefault));
    s->h_dt = (real*)h;
    s->h_red = (double*)(h + 64);
    s->h_count = (int*)(h + 128);
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
  return s;
}

Session* session_for(const QSS_DataArrays* p, const Grid* g) {
  Box bx = box_of(g);
  Session* stale = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_sessions.find(p);
    if (it != g_sessions.end()) {
      Session* s = it->second;
      if (s->bx.nx == bx.nx && s->bx.ny == bx.ny && s->bx.nz == bx.nz) {
        s->bx = bx;
        return s;
      }
      stale = s;
      g_sessions.erase(it);
    }
  }
  if (stale) {
    (void)hipStreamSynchronize(stale->st);
    free_all(stale);
    delete stale;
  }
  Session* s = session_new(p, bx);
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_sessions.find(p);
  if (it != g_sessions.end()) {  // another thread made one for the same record meanwhile
    free_all(s);
    delete s;
    return it->second;
  }
  g_sessions[p] = s;
  return s;
}

// lazily allocated fields of the variable-coefficient path and labelling
void ensure_var(Session* s) {
  if (s->vn) return;
  int64_t n = s->n;
  slab_alloc(s, {part(s->vn, n), part(s->b, n), part(s->vx, n), part(s->vy, n), part(s->vz, n)});
  real* f[] = {s->vn, s->b, s->vx, s->vy, s->vz};
  for (real* q : f) SCHECK(launch_fill(q, real(0), n, s->st));
}
void ensure_grad(Session* s) {
  if (s->gx) return;
  int64_t n = s->n;
  slab_alloc(s, {part(s->gx, n), part(s->gy, n), part(s->gz, n), part(s->phi_extra, n)});
  real* f[] = {s->gx, s->gy, s->gz, s->phi_extra};
  for (real* q : f) SCHECK(launch_fill(q, real(0), n, s->st));
}
void ensure_labels(Session* s) {
  if (s->labels) return;
  // room for n/16 tile roots (a 64x16x16 tile holds 16384 voxels); a labelling
  // with more falls back to scanning troot
  s->roots_cap = s->n / 16 + 1;
  slab_alloc(s, {part(s->labels, s->n), part(s->troot, s->n), part(s->bin, s->n),
                 part(s->roots, s->roots_cap + 1)});
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
