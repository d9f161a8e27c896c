// gpp_stage.h — host->device copies of caller-owned (pageable) arrays.
//
// A DMA engine cannot read pageable memory, so HIP copies such a source
// through its own pinned buffer one piece at a time, with one host thread
// (about 28 GB/s on the MI355X host; pinned sources reach 57 GB/s).  The
// Stager keeps a ring of pinned slots: a small pool of host threads copies
// the next piece into a free slot while the DMA engine drains the previous
// ones on a copy stream of its own, which measured 54 GB/s on the same host.
// Compute streams wait on the copy stream with fence(), so kernels can start
// on the first part of the inputs while the rest is still in flight.
#pragma once

#include <hip/hip_runtime.h>

#include <condition_variable>
#include <cstddef>
#include <mutex>
#include <thread>
#include <vector>

namespace gpp {

// Fork-join memcpy over a fixed set of helper threads (the calling thread
// takes one share).
class CopyPool {
 public:
  explicit CopyPool(int helpers);
  ~CopyPool();
  CopyPool(const CopyPool &) = delete;
  CopyPool &operator=(const CopyPool &) = delete;
  void copy(void *dst, const void *src, size_t bytes);

 private:
  void run(int id);
  void share(int part, char *dst, const char *src, size_t bytes) const;
  std::vector<std::thread> threads_;
  std::mutex mu_;
  std::condition_variable go_, done_;
  unsigned long long gen_ = 0;
  int pending_ = 0;
  bool stop_ = false;
  char *dst_ = nullptr;
  const char *src_ = nullptr;
  size_t bytes_ = 0;
};

class Stager {
 public:
  static constexpr int kRing = 4;                      // pinned slots
  static constexpr size_t kSlotBytes = size_t(32) << 20;

  Stager() = default;
  ~Stager();
  Stager(const Stager &) = delete;
  Stager &operator=(const Stager &) = delete;

  // Allocates the ring and the copy stream on the current device and starts
  // the helper threads.  Idempotent.
  hipError_t init();
  // Enqueues dst (device) <- src (host, any memory) on the copy stream.
  // Returns when every piece has been copied into a pinned slot, so the
  // caller may reuse src; the device side completes in copy-stream order.
  hipError_t copy(void *dst, const void *src, size_t bytes);
  // Makes `s` wait for every copy enqueued so far.
  hipError_t fence(hipStream_t s);

 private:
  hipStream_t cs_ = nullptr;
  char *ring_ = nullptr;
  hipEvent_t slot_done_[kRing] = {};
  bool slot_busy_[kRing] = {};
  hipEvent_t arrived_ = nullptr;
  int next_ = 0;
  CopyPool *pool_ = nullptr;
};

}  // namespace gpp
