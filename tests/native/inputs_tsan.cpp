// The host input generator (fortkernels-_amd/csrc/gpp_inputs.cpp), whose
// fills run on up to 16 std::threads, under ThreadSanitizer: configs 1-3 at
// sizes that take the threaded path, each generated twice and compared
// bit for bit (the generator is counter-based, so the thread split must not
// change a value).  Exit status 0 and no sanitizer report is the pass
// condition.
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/gpp_sigma.h"

int main() {
  const int nb = 24, nc = 8192, npg = 8, ng = nc / npg;
  for (int config : {1, 2, 3, 3 | GPP_FILL_VCOUL_LINEAR}) {
    std::vector<gpp_c128> a[2][4];
    std::vector<double> v[2], w[2];
    std::vector<int> inv[2], ind[2];
    for (int r = 0; r < 2; ++r) {
      a[r][0].resize((size_t)nb * nc);
      a[r][1].resize((size_t)nb * nc);
      a[r][2].resize((size_t)ng * nc);
      a[r][3].resize((size_t)ng * nc);
      v[r].resize(nc);
      w[r].resize((size_t)nb * 3);
      inv[r].resize(ng);
      ind[r].resize(nc);
      if (gpp_fill_inputs(config, 99, nb, nc, ng, npg, a[r][0].data(),
                          a[r][1].data(), a[r][2].data(), a[r][3].data(),
                          v[r].data(), inv[r].data(), ind[r].data(),
                          w[r].data()) != GPP_OK) {
        std::fprintf(stderr, "fill failed: config %d\n", config);
        return 1;
      }
    }
    for (int j = 0; j < 4; ++j)
      if (std::memcmp(a[0][j].data(), a[1][j].data(),
                      a[0][j].size() * sizeof(gpp_c128)) != 0) {
        std::fprintf(stderr, "fill differs: config %d array %d\n", config, j);
        return 1;
      }
  }
  std::printf("input generator sanitizer run ok\n");
  return 0;
}
