// host_pool.h — a fixed set of host worker threads for the FF host
// planning and assembly (ff_api.cpp).  Header-only so that the CPU tests can
// build it under ThreadSanitizer without the GPU runtime.  Not public ABI.
#pragma once

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace gpp {

// Up to `workers` host threads (created once, joined on destruction) that
// run fn(i) for i in [0, n) together with the calling thread; one run at a
// time (the caller serialises runs).  fn must not throw.
class HostPool {
 public:
  explicit HostPool(int workers = 7) {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int nw = std::max(0, std::min(workers, (int)hw - 1));
    for (int t = 0; t < nw; ++t) {
      try {
        th_.emplace_back([this] { work(); });
      } catch (...) {
        break;
      }
    }
  }
  ~HostPool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      quit_ = true;
    }
    go_.notify_all();
    for (auto &t : th_) t.join();
  }
  int threads() const { return (int)th_.size() + 1; }
  void run(int n, const std::function<void(int)> &fn) {
    {
      std::lock_guard<std::mutex> lk(mu_);
      fn_ = &fn;
      n_ = n;
      next_.store(0);
      ++gen_;
    }
    go_.notify_all();
    for (int i; (i = next_.fetch_add(1)) < n;) fn(i);
    std::unique_lock<std::mutex> lk(mu_);
    done_.wait(lk, [this] { return busy_ == 0; });
    fn_ = nullptr;
  }

 private:
  void work() {
    unsigned seen = 0;
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
      go_.wait(lk, [&] { return quit_ || gen_ != seen; });
      if (quit_) return;
      seen = gen_;
      const std::function<void(int)> *fn = fn_;
      const int n = n_;
      if (!fn) continue;
      ++busy_;
      lk.unlock();
      for (int i; (i = next_.fetch_add(1)) < n;) (*fn)(i);
      lk.lock();
      if (--busy_ == 0) done_.notify_all();
    }
  }
  std::vector<std::thread> th_;
  std::mutex mu_;
  std::condition_variable go_, done_;
  const std::function<void(int)> *fn_ = nullptr;
  int n_ = 0;
  std::atomic<int> next_{0};
  int busy_ = 0;
  unsigned gen_ = 0;
  bool quit_ = false;
};

// fn(i) for i in [0, n): on the pool when the call's work (in the caller's
// units) is at least twice `grain`, else on the calling thread alone.
template <typename Fn>
inline void parallel_ranges(HostPool *pool, int n, long long work,
                            long long grain, Fn &&fn) {
  if (!pool || pool->threads() < 2 || n < 2 || work < 2 * grain) {
    for (int i = 0; i < n; ++i) fn(i);
    return;
  }
  std::function<void(int)> f;
  try {
    f = fn;  // may allocate
  } catch (...) {
    for (int i = 0; i < n; ++i) fn(i);
    return;
  }
  pool->run(n, f);
}

}  // namespace gpp
