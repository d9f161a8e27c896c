// gpp_main.cpp — the gppKer mini-app driver on top of the C ABI.
//
//   gpp <number_bands> <number_valence_bands> <number_plane_waves>
//       <nodes_per_mpi_group> [--precision=fp64|mixed|factored]
//       [--semantics=gppker|strict|bgw] [--config=1|2|3] [--seed=N]
//
// Same positional arguments, usage text, banners and result lines as
// gppKer.cpp:151-201,237-239,376-381; the band loop (:270-361) is replaced
// by one gpp_sigma_ex() call.  Inputs default to the file's constant fills
// (config 1, :241-264).  asxtemp, which the reference accumulates but never
// prints, is printed after achtemp.  Without --precision the sum is taken
// band sum first (factored) when all bands share the frequency grid (the
// file's fills, config 1 and 2), as the drop-in gpp_sigma() does, and per
// term (fp64) for band-resolved frequencies (config 3).
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/gpp_sigma.h"

static std::complex<double> cx(const gpp_c128 &z) { return {z.re, z.im}; }

int main(int argc, char **argv) {
  if (argc < 5) {
    std::cout << "The correct form of input is : " << std::endl;
    std::cout << " ./a.out <number_bands> <number_valence_bands> "
                 "<number_plane_waves> <nodes_per_mpi_group> "
              << std::endl;
    return 0;
  }
  const int number_bands = std::atoi(argv[1]);
  const int nvband = std::atoi(argv[2]);
  const int ncouls = std::atoi(argv[3]);
  const int nodes_per_group = std::atoi(argv[4]);

  gpp_options opt;
  gpp_default_options(&opt);
  bool precision_given = false;
  int config = 1;
  unsigned long long seed = 12345;
  for (int i = 5; i < argc; ++i) {
    const std::string a = argv[i];
    auto val = [&](const char *k) -> const char * {
      const size_t n = std::strlen(k);
      return a.compare(0, n, k) == 0 ? a.c_str() + n : nullptr;
    };
    if (const char *v = val("--precision=")) {
      precision_given = true;
      opt.precision = std::strcmp(v, "mixed") == 0      ? GPP_PREC_MIXED
                      : std::strcmp(v, "factored") == 0 ? GPP_PREC_FP64_FACTORED
                                                        : GPP_PREC_FP64;
    } else if (const char *v = val("--semantics=")) {
      opt.semantics = std::strcmp(v, "bgw") == 0      ? GPP_SEM_BGW
                      : std::strcmp(v, "strict") == 0 ? GPP_SEM_GPPKER_STRICT
                                                      : GPP_SEM_GPPKER;
    } else if (const char *v = val("--config=")) {
      config = std::atoi(v);
    } else if (const char *v = val("--seed=")) {
      seed = std::strtoull(v, nullptr, 10);
    } else {
      std::cerr << "unknown option " << a << std::endl;
      return 2;
    }
  }
  if (nodes_per_group <= 0 || ncouls <= 0 || number_bands < 0) {
    std::cerr << "error: need number_plane_waves > 0 and "
                 "nodes_per_mpi_group > 0"
              << std::endl;
    return 2;
  }
  if (!precision_given)
    opt.precision = config == 3 ? GPP_PREC_FP64 : GPP_PREC_FP64_FACTORED;
  const int npes = 1;
  const int ngpown = ncouls / (nodes_per_group * npes);  // gppKer.cpp:163
  const int nstart = 0, nend = 3;
  const double to1 = 1e-6, gamma = 0.5, sexcut = 4.0;
  const double limitone = 1.0 / (to1 * 4.0), limittwo = std::pow(0.5, 2);

  std::cout << "Number of OpenMP Threads = 1" << std::endl;
  std::cout << "number_bands = " << number_bands << "\t nvband = " << nvband
            << "\t ncouls = " << ncouls
            << "\t nodes_per_group  = " << nodes_per_group
            << "\t ngpown = " << ngpown << "\t nend = " << nend
            << "\t nstart = " << nstart << "\t gamma = " << gamma
            << "\t sexcut = " << sexcut << "\t limitone = " << limitone
            << "\t limittwo = " << limittwo << std::endl;

  const size_t nbnc = (size_t)number_bands * ncouls;
  const size_t ngnc = (size_t)ngpown * ncouls;
  std::vector<gpp_c128> aqsm(nbnc), aqsn(nbnc), ieps(ngnc), wt(ngnc);
  std::vector<double> vcoul(ncouls);
  std::vector<int> inv(ngpown > 0 ? ngpown : 1), indinv(ncouls);
  std::vector<double> wx(config == 3 ? (size_t)number_bands * 3 + 3 : 3);

  std::cout << "Size of wtilde_array = "
            << (ncouls * (double)ngpown * 2.0 * 8) / std::pow(1024, 2)
            << " Mbytes" << std::endl;
  std::cout << "Size of aqsntemp = "
            << (ncouls * (double)number_bands * 2.0 * 8) / std::pow(1024, 2)
            << " Mbytes" << std::endl;
  std::cout << "Size of I_eps_array array = "
            << (ncouls * (double)ngpown * 2.0 * 8) / std::pow(1024, 2)
            << " Mbytes" << std::endl;

  int rc = gpp_fill_inputs(config, seed, number_bands, ncouls, ngpown,
                           nodes_per_group, aqsm.data(), aqsn.data(),
                           ieps.data(), wt.data(), vcoul.data(), inv.data(),
                           indinv.data(), wx.data());
  if (rc) {
    std::cerr << "error: gpp_fill_inputs failed (" << rc << ")" << std::endl;
    return 3;
  }

  gpp_handle h = nullptr;
  if ((rc = gpp_create(&h, -1)) != GPP_OK) {
    std::cerr << "error: gpp_create failed (" << rc << ")" << std::endl;
    return 3;
  }
  gpp_problem p = {number_bands, nvband,     ncouls,      ngpown,
                   aqsm.data(),  aqsn.data(), ieps.data(), wt.data(),
                   vcoul.data(), inv.data(),  indinv.data(), wx.data(),
                   config == 3 ? 1 : 0};
  gpp_result res;
  std::memset(&res, 0, sizeof res);
  const auto t0 = std::chrono::high_resolution_clock::now();
  rc = gpp_sigma_ex(h, &p, &opt, &res);
  const std::chrono::duration<double> el =
      std::chrono::high_resolution_clock::now() - t0;
  if (rc) {
    std::cerr << "error: gpp_sigma_ex failed (" << rc
              << "): " << gpp_last_error(h) << std::endl;
    gpp_destroy(h);
    return 4;
  }
  std::cout << "achstemp = " << cx(res.achstemp) << std::endl;
  for (int iw = nstart; iw < nend; ++iw)
    std::cout << "achtemp[" << iw << "] = " << std::setprecision(15)
              << cx(res.achtemp[iw]) << std::endl;
  for (int iw = nstart; iw < nend; ++iw)
    std::cout << "asxtemp[" << iw << "] = " << std::setprecision(15)
              << cx(res.asxtemp[iw]) << std::endl;
  std::cout << "********** Chrono Time Taken **********= " << el.count()
            << " secs" << std::endl;
  std::cout << "Kernel time (non-occupied, occupied) = " << res.unocc_kernel_ms
            << " ms, " << res.occ_kernel_ms << " ms" << std::endl;
  gpp_destroy(h);
  return 0;
}
