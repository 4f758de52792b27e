// lsm_xfer.h -- large host <-> device copies of whole fields.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace lsm {

// Copies between a caller's (pageable) host array and device memory, in
// stream order on `st`, returning once the data has arrived.  Copies of
// XFER_MIN bytes or more go through a few pinned staging buffers, with the
// host side of each chunk split over threads, so the DMA of one chunk
// overlaps the host copy of the previous one (MI355X, 556 MB field: about
// 50 GB/s each way, against 16 GB/s device-to-host and 27 GB/s
// host-to-device for a pageable hipMemcpyAsync).  Smaller copies are one
// hipMemcpyAsync followed by a stream synchronization.  Staged copies from
// several threads take turns.
constexpr size_t XFER_MIN = size_t(64) << 20;
hipError_t copy_to_host(void* h, const void* d, size_t bytes, hipStream_t st);
hipError_t copy_to_device(void* d, const void* h, size_t bytes, hipStream_t st);

}  // namespace lsm
