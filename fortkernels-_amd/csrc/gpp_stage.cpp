// gpp_stage.cpp — pinned-ring host->device staging (see gpp_stage.h).
#include "gpp_stage.h"

#include <algorithm>
#include <cstring>

namespace gpp {

// ---- CopyPool

CopyPool::CopyPool(int helpers) {
  for (int i = 0; i < helpers; ++i)
    threads_.emplace_back([this, i] { run(i + 1); });
}

CopyPool::~CopyPool() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  go_.notify_all();
  for (auto &t : threads_) t.join();
}

// Part `part` of (threads + 1) equal shares, 4 KiB aligned.
void CopyPool::share(int part, char *dst, const char *src,
                     size_t bytes) const {
  const size_t parts = threads_.size() + 1;
  size_t per = (bytes + parts - 1) / parts;
  per = (per + 4095) / 4096 * 4096;
  const size_t b = std::min(bytes, per * part);
  const size_t e = std::min(bytes, b + per);
  if (e > b) std::memcpy(dst + b, src + b, e - b);
}

void CopyPool::run(int id) {
  unsigned long long seen = 0;
  for (;;) {
    char *dst;
    const char *src;
    size_t bytes;
    {
      std::unique_lock<std::mutex> lk(mu_);
      go_.wait(lk, [&] { return stop_ || gen_ != seen; });
      if (stop_) return;
      seen = gen_;
      dst = dst_;
      src = src_;
      bytes = bytes_;
    }
    share(id, dst, src, bytes);
    {
      std::lock_guard<std::mutex> lk(mu_);
      if (--pending_ == 0) done_.notify_one();
    }
  }
}

void CopyPool::copy(void *dst, const void *src, size_t bytes) {
  if (threads_.empty() || bytes < (size_t(1) << 20)) {
    std::memcpy(dst, src, bytes);
    return;
  }
  {
    std::lock_guard<std::mutex> lk(mu_);
    dst_ = (char *)dst;
    src_ = (const char *)src;
    bytes_ = bytes;
    pending_ = (int)threads_.size();
    ++gen_;
  }
  go_.notify_all();
  share(0, (char *)dst, (const char *)src, bytes);
  std::unique_lock<std::mutex> lk(mu_);
  done_.wait(lk, [&] { return pending_ == 0; });
}

// ---- Stager

Stager::~Stager() {
  if (cs_) (void)hipStreamSynchronize(cs_);
  delete pool_;
  for (auto &e : slot_done_)
    if (e) (void)hipEventDestroy(e);
  if (arrived_) (void)hipEventDestroy(arrived_);
  if (ring_) (void)hipHostFree(ring_);
  if (cs_) (void)hipStreamDestroy(cs_);
}

hipError_t Stager::init() {
  if (pool_) return hipSuccess;
  hipError_t e = hipSuccess;
  if (!cs_ && (e = hipStreamCreateWithFlags(&cs_, hipStreamNonBlocking)))
    return e;
  if (!ring_ && (e = hipHostMalloc((void **)&ring_, kRing * kSlotBytes,
                                   hipHostMallocDefault)))
    return e;
  for (auto &ev : slot_done_)
    if (!ev && (e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)))
      return e;
  if (!arrived_ &&
      (e = hipEventCreateWithFlags(&arrived_, hipEventDisableTiming)))
    return e;
  // 8 threads in all saturated the DMA engine on the MI355X host (54 GB/s
  // from 4 to 16 threads); fewer if the host has fewer cores.
  const unsigned hw = std::thread::hardware_concurrency();
  const int threads = (int)std::min(8u, hw > 1 ? hw : 1u);
  pool_ = new CopyPool(threads - 1);
  return hipSuccess;
}

hipError_t Stager::copy(void *dst, const void *src, size_t bytes) {
  hipError_t e;
  for (size_t off = 0; off < bytes; off += kSlotBytes) {
    const size_t n = std::min(kSlotBytes, bytes - off);
    const int slot = next_;
    next_ = (next_ + 1) % kRing;
    char *pin = ring_ + (size_t)slot * kSlotBytes;
    if (slot_busy_[slot] && (e = hipEventSynchronize(slot_done_[slot])))
      return e;
    pool_->copy(pin, (const char *)src + off, n);
    if ((e = hipMemcpyAsync((char *)dst + off, pin, n, hipMemcpyHostToDevice,
                            cs_)))
      return e;
    if ((e = hipEventRecord(slot_done_[slot], cs_))) return e;
    slot_busy_[slot] = true;
  }
  return hipSuccess;
}

hipError_t Stager::fence(hipStream_t s) {
  hipError_t e;
  if ((e = hipEventRecord(arrived_, cs_))) return e;
  return hipStreamWaitEvent(s, arrived_, 0);
}

}  // namespace gpp
