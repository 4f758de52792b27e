// lsm_gz.cpp -- multi-threaded gzip array files and the asynchronous step-file
// writer (see lsm_gz.h).
#include "lsm_gz.h"

#include <stdio.h>
#include <zlib.h>
#include <algorithm>
#include <atomic>
#include <stdexcept>

namespace lsm {

namespace {

constexpr size_t kChunk = size_t(8) << 20;  // uncompressed bytes per gzip member

// one gzip member (header, deflate stream, CRC32 and size trailer) of `in`
bool gz_member(const unsigned char* in, size_t n, std::vector<unsigned char>& out) {
  z_stream z{};
  if (deflateInit2(&z, Z_DEFAULT_COMPRESSION, Z_DEFLATED, 15 + 16, 8, Z_DEFAULT_STRATEGY) != Z_OK)
    return false;
  out.resize(deflateBound(&z, (uLong)n) + 64);
  z.next_in = const_cast<unsigned char*>(in);
  z.avail_in = (uInt)n;
  z.next_out = out.data();
  z.avail_out = (uInt)out.size();
  const int rc = deflate(&z, Z_FINISH);
  const size_t used = out.size() - z.avail_out;
  deflateEnd(&z);
  if (rc != Z_STREAM_END) return false;
  out.resize(used);
  return true;
}

int default_threads() {
  unsigned h = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(16u, h ? h : 4u));
}

}  // namespace

bool gz_write(const std::string& path, const std::vector<GzPart>& parts, int threads) {
  struct Piece {
    const unsigned char* p;
    size_t n;
  };
  std::vector<Piece> pieces;
  for (const GzPart& part : parts)
    for (size_t off = 0; off < part.bytes; off += kChunk)
      pieces.push_back({(const unsigned char*)part.data + off, std::min(kChunk, part.bytes - off)});
  std::vector<std::vector<unsigned char>> out(pieces.size());
  std::atomic<size_t> next{0};
  std::atomic<bool> ok{true};
  auto work = [&] {
    for (size_t i = next++; i < pieces.size(); i = next++)
      if (!gz_member(pieces[i].p, pieces[i].n, out[i])) ok = false;
  };
  const int nt = std::max(1, std::min<int>(threads > 0 ? threads : default_threads(),
                                           (int)pieces.size()));
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back(work);
  work();
  for (auto& t : pool) t.join();
  if (!ok) return false;
  FILE* fp = fopen(path.c_str(), "wb");
  if (!fp) return false;
  bool good = true;
  for (const auto& m : out)
    if (!m.empty() && fwrite(m.data(), 1, m.size(), fp) != m.size()) good = false;
  if (pieces.empty()) {  // an empty member keeps the file a valid gzip stream
    std::vector<unsigned char> m;
    good = gz_member(nullptr, 0, m) && fwrite(m.data(), 1, m.size(), fp) == m.size();
  }
  return fclose(fp) == 0 && good;
}

bool write_array_gz(const std::string& path, const int dims[3], const void* data, size_t bytes) {
  return gz_write(path, {{dims, 3 * sizeof(int)}, {data, bytes}}, 0);
}

StepWriter::StepWriter(size_t bytes, const int dims[3], int slots)
    : bytes_(bytes), nslots_(slots < 1 ? 1 : (slots > MAX_SLOTS ? MAX_SLOTS : slots)) {
  for (int a = 0; a < 3; ++a) dims_[a] = dims[a];
  for (int k = 0; k < nslots_; ++k) {
    Slot& s = slots_[k];
    if (hipHostMalloc(&s.host, bytes_, hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&s.copied, hipEventDisableTiming) != hipSuccess) {
      for (Slot& t : slots_) {
        if (t.host) (void)hipHostFree(t.host);
        if (t.copied) (void)hipEventDestroy(t.copied);
      }
      throw std::runtime_error("step-file writer: pinned snapshot buffers not available");
    }
  }
  worker_ = std::thread([this] { run(); });
}

StepWriter::~StepWriter() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  cv_work_.notify_all();
  if (worker_.joinable()) worker_.join();
  for (Slot& s : slots_) {
    if (s.copied) (void)hipEventDestroy(s.copied);
    if (s.host) (void)hipHostFree(s.host);
  }
}

StepWriter::Slot* StepWriter::free_slot() {
  for (int k = 0; k < nslots_; ++k)
    if (!slots_[k].busy) return &slots_[k];
  return nullptr;
}

void StepWriter::submit(hipStream_t st, const void* device_field, const std::string& name) {
  std::unique_lock<std::mutex> lk(mu_);
  cv_free_.wait(lk, [&] { return free_slot() != nullptr || !error_.empty(); });
  if (!error_.empty()) throw std::runtime_error(error_);
  Slot& s = *free_slot();
  const int k = (int)(&s - slots_);
  if (hipMemcpyAsync(s.host, device_field, bytes_, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipEventRecord(s.copied, st) != hipSuccess)
    throw std::runtime_error("step-file writer: snapshot copy failed for " + name);
  s.busy = true;
  ++pending_;
  jobs_.push_back({k, name});
  cv_work_.notify_one();
}

void StepWriter::finish() {
  std::unique_lock<std::mutex> lk(mu_);
  cv_free_.wait(lk, [&] { return pending_ == 0; });
  if (!error_.empty()) throw std::runtime_error(error_);
}

void StepWriter::run() {
  for (;;) {
    Job job;
    {
      std::unique_lock<std::mutex> lk(mu_);
      cv_work_.wait(lk, [&] { return stop_ || !jobs_.empty(); });
      if (jobs_.empty()) return;
      job = jobs_.front();
      jobs_.pop_front();
    }
    Slot& s = slots_[job.slot];
    std::string err;
    if (hipEventSynchronize(s.copied) != hipSuccess)
      err = "step-file writer: snapshot copy failed for " + job.name;
    else if (!write_array_gz(job.name + ".gz", dims_, s.host, bytes_))
      err = "step-file writer: cannot write " + job.name + ".gz";
    {
      std::lock_guard<std::mutex> lk(mu_);
      s.busy = false;
      --pending_;
      if (!err.empty() && error_.empty()) error_ = err;
    }
    cv_free_.notify_all();
  }
}

}  // namespace lsm
