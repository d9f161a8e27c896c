// gpp_main.cpp — the gppKer mini-app driver on top of the C ABI.
//
//   gpp <number_bands> <number_valence_bands> <number_plane_waves>
//       <nodes_per_mpi_group> [--precision=fp64|mixed|factored]
//       [--semantics=gppker|strict|bgw] [--config=1|2|3] [--seed=N]
//       [--vcoul=file|linear] [--tune] [--batch=N]
//
// --vcoul=linear fills vcoul[i] = i / ncouls instead of the configuration's
// fill.  --tune runs the tile autotuner for the shape first (and prints the
// tile it picked).  --batch=N evaluates N copies of the problem (seeds
// seed .. seed+N-1) in one gpp_sigma_batch call and prints each result.
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
  bool vcoul_linear = false, tune = false;
  int batch = 0;
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
    } else if (const char *v = val("--vcoul=")) {
      if (std::strcmp(v, "linear") == 0)
        vcoul_linear = true;
      else if (std::strcmp(v, "file") != 0) {
        std::cerr << "--vcoul takes file or linear" << std::endl;
        return 2;
      }
    } else if (a == "--tune") {
      tune = true;
    } else if (const char *v = val("--batch=")) {
      batch = std::atoi(v);
      if (batch < 1) {
        std::cerr << "--batch takes a positive count" << std::endl;
        return 2;
      }
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

  const int fill = config | (vcoul_linear ? GPP_FILL_VCOUL_LINEAR : 0);
  int rc = gpp_fill_inputs(fill, seed, number_bands, ncouls, ngpown,
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
  auto fail = [&](const char *what, int code) {
    std::cerr << "error: " << what << " failed (" << code
              << "): " << gpp_last_error(h) << std::endl;
    gpp_destroy(h);
    return 4;
  };
  if (tune) {
    int k = 0, w = 0;
    if ((rc = gpp_autotune(h, &p, &opt, &k, &w))) return fail("gpp_autotune", rc);
    std::cout << "tuned tile: " << k << " G-vectors x " << w
              << " waves per block" << std::endl;
  }
  auto print = [&](const gpp_result &res, const char *tag) {
    std::cout << tag << "achstemp = " << cx(res.achstemp) << std::endl;
    for (int iw = nstart; iw < nend; ++iw)
      std::cout << tag << "achtemp[" << iw << "] = " << std::setprecision(15)
                << cx(res.achtemp[iw]) << std::endl;
    for (int iw = nstart; iw < nend; ++iw)
      std::cout << tag << "asxtemp[" << iw << "] = " << std::setprecision(15)
                << cx(res.asxtemp[iw]) << std::endl;
  };

  if (batch > 0) {
    // N problems of this shape, seeds seed .. seed+N-1 (config 1 ignores
    // the seed, so its copies are identical).
    std::vector<std::vector<gpp_c128>> bc(4 * (size_t)batch);
    std::vector<std::vector<double>> bv((size_t)batch), bw((size_t)batch);
    std::vector<std::vector<int>> bi((size_t)batch), bd((size_t)batch);
    std::vector<gpp_problem> probs((size_t)batch, p);
    for (int b = 0; b < batch; ++b) {
      for (int j = 0; j < 4; ++j) bc[4 * b + j].resize(j < 2 ? nbnc : ngnc);
      bv[b].resize(ncouls);
      bw[b].resize(wx.size());
      bi[b].resize(inv.size());
      bd[b].resize(ncouls);
      rc = gpp_fill_inputs(fill, seed + b, number_bands, ncouls, ngpown,
                           nodes_per_group, bc[4 * b].data(),
                           bc[4 * b + 1].data(), bc[4 * b + 2].data(),
                           bc[4 * b + 3].data(), bv[b].data(), bi[b].data(),
                           bd[b].data(), bw[b].data());
      if (rc) return fail("gpp_fill_inputs", rc);
      probs[b].aqsmtemp = bc[4 * b].data();
      probs[b].aqsntemp = bc[4 * b + 1].data();
      probs[b].I_eps_array = bc[4 * b + 2].data();
      probs[b].wtilde_array = bc[4 * b + 3].data();
      probs[b].vcoul = bv[b].data();
      probs[b].inv_igp_index = bi[b].data();
      probs[b].indinv = bd[b].data();
      probs[b].wx_array = bw[b].data();
    }
    std::vector<gpp_result> res((size_t)batch);
    std::memset(res.data(), 0, res.size() * sizeof(gpp_result));
    const auto t0 = std::chrono::high_resolution_clock::now();
    rc = gpp_sigma_batch(h, probs.data(), batch, &opt, res.data());
    const std::chrono::duration<double> el =
        std::chrono::high_resolution_clock::now() - t0;
    if (rc) return fail("gpp_sigma_batch", rc);
    for (int b = 0; b < batch; ++b) {
      const std::string tag = "[" + std::to_string(b) + "] ";
      print(res[b], tag.c_str());
    }
    std::cout << "********** Chrono Time Taken **********= " << el.count()
              << " secs" << std::endl;
    gpp_destroy(h);
    return 0;
  }

  gpp_result res;
  std::memset(&res, 0, sizeof res);
  const auto t0 = std::chrono::high_resolution_clock::now();
  rc = gpp_sigma_ex(h, &p, &opt, &res);
  const std::chrono::duration<double> el =
      std::chrono::high_resolution_clock::now() - t0;
  if (rc) return fail("gpp_sigma_ex", rc);
  print(res, "");
  std::cout << "********** Chrono Time Taken **********= " << el.count()
            << " secs" << std::endl;
  std::cout << "Kernel time (non-occupied, occupied) = " << res.unocc_kernel_ms
            << " ms, " << res.occ_kernel_ms << " ms" << std::endl;
  gpp_destroy(h);
  return 0;
}
