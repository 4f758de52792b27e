// lsm_xfer.cpp -- staged host <-> device copies (see lsm_xfer.h).
#include "lsm_xfer.h"

#include <string.h>
#include <algorithm>
#include <mutex>
#include <thread>
#include <vector>

namespace lsm {
namespace {

constexpr size_t CHUNK = size_t(32) << 20;  // bytes per staging buffer
constexpr int NBUF = 4;                     // staging buffers in flight
constexpr int NTHR = 8;                     // host threads per chunk copy

struct Pool {
  std::mutex mu;  // one staged copy at a time
  char* buf[NBUF] = {nullptr, nullptr, nullptr, nullptr};
  hipEvent_t ev[NBUF] = {nullptr, nullptr, nullptr, nullptr};
  bool ready = false;
  hipError_t init() {
    if (ready) return hipSuccess;
    for (int i = 0; i < NBUF; ++i) {
      hipError_t e = hipHostMalloc((void**)&buf[i], CHUNK, hipHostMallocDefault);
      if (e != hipSuccess) return e;
      e = hipEventCreateWithFlags(&ev[i], hipEventDisableTiming);
      if (e != hipSuccess) return e;
    }
    ready = true;
    return hipSuccess;
  }
};
Pool& pool() {
  static Pool* p = new Pool();  // kept to process end: pinned buffers may outlive static destructors' order
  return *p;
}

// memcpy of len bytes split over NTHR threads
void host_copy(char* dst, const char* src, size_t len) {
  const size_t per = (len + NTHR - 1) / NTHR;
  std::thread th[NTHR];
  int used = 0;
  for (int t = 0; t < NTHR; ++t) {
    const size_t a = t * per, b = std::min(len, a + per);
    if (a >= b) break;
    th[used++] = std::thread([=] { memcpy(dst + a, src + a, b - a); });
  }
  for (int t = 0; t < used; ++t) th[t].join();
}

#define XRET(x)                      \
  do {                               \
    hipError_t e_ = (x);             \
    if (e_ != hipSuccess) return e_; \
  } while (0)

}  // namespace

hipError_t copy_to_host(void* h, const void* d, size_t bytes, hipStream_t st) {
  if (bytes < XFER_MIN) {
    XRET(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, st));
    return hipStreamSynchronize(st);
  }
  Pool& p = pool();
  std::lock_guard<std::mutex> lk(p.mu);
  XRET(p.init());
  const size_t nch = (bytes + CHUNK - 1) / CHUNK;
  char* dst = (char*)h;
  const char* src = (const char*)d;
  // chunk c goes to buffer c % NBUF; chunk c - NBUF is copied out of that
  // buffer on the host before chunk c is issued into it
  for (size_t c = 0; c < nch + NBUF; ++c) {
    if (c >= (size_t)NBUF) {
      const size_t k = c - NBUF;
      const int b = (int)(k % NBUF);
      XRET(hipEventSynchronize(p.ev[b]));
      const size_t off = k * CHUNK;
      host_copy(dst + off, p.buf[b], std::min(CHUNK, bytes - off));
    }
    if (c < nch) {
      const int b = (int)(c % NBUF);
      const size_t off = c * CHUNK;
      XRET(hipMemcpyAsync(p.buf[b], src + off, std::min(CHUNK, bytes - off),
                          hipMemcpyDeviceToHost, st));
      XRET(hipEventRecord(p.ev[b], st));
    }
  }
  return hipSuccess;
}

hipError_t copy_to_device(void* d, const void* h, size_t bytes, hipStream_t st) {
  if (bytes < XFER_MIN) {
    XRET(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, st));
    return hipStreamSynchronize(st);
  }
  Pool& p = pool();
  std::lock_guard<std::mutex> lk(p.mu);
  XRET(p.init());
  const size_t nch = (bytes + CHUNK - 1) / CHUNK;
  char* dst = (char*)d;
  const char* src = (const char*)h;
  // buffer c % NBUF is refilled once its previous upload (chunk c - NBUF) is done
  for (size_t c = 0; c < nch; ++c) {
    const int b = (int)(c % NBUF);
    if (c >= (size_t)NBUF) XRET(hipEventSynchronize(p.ev[b]));
    const size_t off = c * CHUNK, len = std::min(CHUNK, bytes - off);
    host_copy(p.buf[b], src + off, len);
    XRET(hipMemcpyAsync(dst + off, p.buf[b], len, hipMemcpyHostToDevice, st));
    XRET(hipEventRecord(p.ev[b], st));
  }
  return hipStreamSynchronize(st);
}

}  // namespace lsm
