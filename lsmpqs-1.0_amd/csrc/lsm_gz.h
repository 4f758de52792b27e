// lsm_gz.h -- array files: gzip written in process on several threads, and an
// asynchronous writer of per-step snapshots of a device field.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace lsm {

// Writes `path` as a gzip file whose decompressed content is the
// concatenation of the parts.  The content is split into chunks that are
// deflated concurrently, each as its own gzip member; a multi-member file is
// one valid gzip stream to gunzip and zlib readers.  Returns false on an I/O
// or zlib error.
struct GzPart {
  const void* data;
  size_t bytes;
};
bool gz_write(const std::string& path, const std::vector<GzPart>& parts, int threads);

// Reference array-file layout: int32 dims[3], then the values.
bool write_array_gz(const std::string& path, const int dims[3], const void* data, size_t bytes);

// Asynchronous array files (data_step_N, the checkpoint files).  submit()
// enqueues a device-to-host copy of the field into one of `slots` pinned
// snapshot buffers on the caller's stream (so later kernels on that stream
// cannot change the snapshot) and returns; a worker thread waits for the copy
// and writes `name`.gz, in submission order.  With every buffer busy,
// submit() waits for the oldest file to finish.  finish() waits for every
// file and reports the first failure.
class StepWriter {
 public:
  static constexpr int MAX_SLOTS = 4;
  StepWriter(size_t bytes, const int dims[3], int slots = 2);
  ~StepWriter();
  void submit(hipStream_t st, const void* device_field, const std::string& name);
  void finish();

 private:
  struct Slot {
    void* host = nullptr;
    hipEvent_t copied = nullptr;
    bool busy = false;
  };
  struct Job {
    int slot;
    std::string name;
  };
  void run();

  Slot* free_slot();
  size_t bytes_;
  int dims_[3];
  int nslots_;
  Slot slots_[MAX_SLOTS];
  std::deque<Job> jobs_;
  std::mutex mu_;
  std::condition_variable cv_work_, cv_free_;
  bool stop_ = false;
  int pending_ = 0;
  std::string error_;
  std::thread worker_;
};

}  // namespace lsm
