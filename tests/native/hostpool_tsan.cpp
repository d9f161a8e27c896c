// The host worker pool of the FF planning and assembly (fortkernels-_amd/
// csrc/host_pool.h) under ThreadSanitizer (and AddressSanitizer): runs of
// many sizes, each index executed exactly once and its result visible to
// the caller after the run; runs back to back on one pool (the workers'
// generations), pools of 0-7 workers, pools built and destroyed while idle
// and right after a run.  Exit status 0 and no sanitizer report is the pass
// condition.
#include <atomic>
#include <cstdio>
#include <vector>

#include "host_pool.h"

int main() {
  for (int workers : {0, 1, 3, 7}) {
    gpp::HostPool pool(workers);
    for (int rep = 0; rep < 50; ++rep)
      for (int n : {0, 1, 2, 7, 64, 1000}) {
        std::vector<int> hits(n, 0);           // plain ints: one writer each
        std::vector<std::atomic<int>> seen(n);  // counts, to catch doubles
        for (auto &x : seen) x.store(0);
        gpp::parallel_ranges(&pool, n, 1LL << 40, 1, [&](int i) {
          hits[i] = i + 1;
          seen[i].fetch_add(1);
        });
        for (int i = 0; i < n; ++i)
          if (hits[i] != i + 1 || seen[i].load() != 1) {
            std::fprintf(stderr, "index %d of %d: hits %d seen %d (workers %d)\n",
                         i, n, hits[i], seen[i].load(), workers);
            return 1;
          }
      }
  }
  for (int i = 0; i < 20; ++i) {
    gpp::HostPool idle(5);  // created and destroyed without work
  }
  {
    gpp::HostPool pool(3);
    long long sum = 0;
    std::vector<long long> part(16, 0);
    gpp::parallel_ranges(&pool, 16, 1LL << 40, 1,
                         [&](int i) { part[i] = (long long)i * i; });
    for (long long v : part) sum += v;
    if (sum != 1240) {
      std::fprintf(stderr, "sum %lld\n", sum);
      return 1;
    }
  }  // destroyed right after a run
  std::printf("host pool sanitizer run ok\n");
  return 0;
}
