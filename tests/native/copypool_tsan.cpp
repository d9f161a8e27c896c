// The host copy pool of the pinned-ring stager (fortkernels-_amd/csrc/
// gpp_stage.cpp) under ThreadSanitizer (and AddressSanitizer): fork-join
// copies of many sizes (below and above the 1 MiB threshold, unaligned
// tails) from one thread, then pools built and destroyed while idle and
// right after a copy.  Every destination byte is checked.  Exit status 0 and
// no sanitizer report is the pass condition.  Only the CopyPool part of the
// stager runs: it is the part with threads; the rest needs a GPU.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "gpp_stage.h"

int main() {
  std::vector<unsigned char> src(size_t(9) << 20), dst(src.size());
  for (size_t i = 0; i < src.size(); ++i) src[i] = (unsigned char)(i * 131 + 7);
  const size_t sizes[] = {0, 1, 4095, 4096, 1 << 20, (1 << 20) + 3,
                          (3 << 20) + 12345, size_t(9) << 20};
  for (int helpers : {0, 1, 3, 7}) {
    gpp::CopyPool pool(helpers);
    for (int rep = 0; rep < 4; ++rep)
      for (size_t n : sizes) {
        std::memset(dst.data(), 0xAB, dst.size());
        pool.copy(dst.data(), src.data(), n);
        if (std::memcmp(dst.data(), src.data(), n) != 0 ||
            (n < dst.size() && dst[n] != 0xAB)) {
          std::fprintf(stderr, "copy mismatch: helpers %d bytes %zu\n", helpers, n);
          return 1;
        }
      }
  }
  for (int i = 0; i < 20; ++i) {
    gpp::CopyPool idle(5);  // created and destroyed without work
  }
  std::printf("copy pool sanitizer run ok\n");
  return 0;
}
