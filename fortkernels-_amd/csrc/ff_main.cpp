// ff_main.cpp — command-line driver for the full-frequency kernels, same
// positional arguments, input fills and result lines as the FF mini-app
// (FFKer_cori/ffKernel.cpp:305-505):
//
//   ff number_bands nvband ncouls ngpown nFreq nfreqeval
//      [--semantics cpp|f90] [--config 1|2] [--seed S] [--host]
//
// Inputs are generated in device memory (the mini-app's own fills by
// default; its example runs need tens of GB of I_eps) unless --host is
// given.  As in ffKernel.cpp's main, which hands asxDtemp to
// achDtemp_Kernel (:495), the printed asxDtemp includes achDtemp for the
// cpp semantics.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ff_sigma.h"

namespace {

void usage() {
  std::printf("Incorrect Parameters!!! The correct form is \n");
  std::printf("./a.out number_bands nvband ncouls ngpown nFreq nfreqeval \n");
}

void print_c(const char *name, gpp_c128 z) {
  std::printf("%s = ( %f, %f) \n", name, z.re, z.im);
}

template <class T>
T *dalloc(size_t n) {
  void *p = nullptr;
  if (hipMalloc(&p, n * sizeof(T) + 16) != hipSuccess) return nullptr;
  return (T *)p;
}

}  // namespace

int main(int argc, char **argv) {
  std::vector<std::string> pos;
  int sem = FF_SEM_CPP, config = 1;
  unsigned long long seed = 2024;
  bool host = false;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    if (a == "--semantics" && i + 1 < argc) {
      const std::string v = argv[++i];
      sem = v == "f90" ? FF_SEM_F90 : FF_SEM_CPP;
    } else if (a == "--config" && i + 1 < argc) {
      config = std::atoi(argv[++i]);
    } else if (a == "--seed" && i + 1 < argc) {
      seed = std::strtoull(argv[++i], nullptr, 10);
    } else if (a == "--host") {
      host = true;
    } else {
      pos.push_back(a);
    }
  }
  if (pos.size() != 6) {
    usage();
    return 0;
  }
  const auto t_start = std::chrono::high_resolution_clock::now();
  const int NB = std::atoi(pos[0].c_str()), NV = std::atoi(pos[1].c_str());
  const int NC = std::atoi(pos[2].c_str()), NG = std::atoi(pos[3].c_str());
  const int nF = std::atoi(pos[4].c_str()), NE = std::atoi(pos[5].c_str());
  if (NG > NC) {
    std::printf("Incorrect Parameters!!! ngpown cannot be greater than "
                "ncouls. The correct form is \n");
    std::printf("./a.out number_bands nvband ncouls ngpown nFreq nfreqeval \n");
    return 0;
  }
  std::printf(" number_bands = %d\n nvband = %d\n ncouls = %d\n ngpown = %d\n"
              " nFreq = %d\n nfreqeval = %d\n semantics = %s\n",
              NB, NV, NC, NG, nF, NE, sem == FF_SEM_F90 ? "f90" : "cpp");
  const size_t nbnc = (size_t)NB * NC, eps = (size_t)nF * NG * NC;
  std::printf("Memory Used = %g GB\n",
              (double)(2 * nbnc + 2 * eps) * 16.0 / (1024.0 * 1024 * 1024));

  ff_problem p{};
  p.number_bands = NB;
  p.nvband = NV;
  p.ncouls = NC;
  p.ngpown = NG;
  p.nFreq = nF;
  p.nfreqeval = NE;
  double scal[4];
  std::vector<gpp_c128> h_aqsm, h_aqsn, h_er, h_ea, h_brd, h_sd;
  std::vector<double> h_vc, h_ek, h_gr, h_pr;
  std::vector<int> h_inv, h_ind;
  std::vector<void *> dev;
  int rc;
  if (host) {
    h_aqsm.resize(nbnc);
    h_aqsn.resize(nbnc);
    h_er.resize(eps);
    h_ea.resize(eps);
    h_vc.resize(NC);
    h_inv.resize(NG);
    h_ind.resize(NC);
    h_ek.resize(NB);
    h_gr.resize(nF);
    h_pr.resize(nF);
    h_brd.resize(nF);
    h_sd.resize((size_t)NB * nF);
    rc = ff_fill_inputs(config, sem, seed, NB, NC, NG, nF, h_aqsm.data(),
                        h_aqsn.data(), h_er.data(), h_ea.data(), h_vc.data(),
                        h_inv.data(), h_ind.data(), h_ek.data(), h_gr.data(),
                        h_pr.data(), h_brd.data(), h_sd.data(), scal);
    p.aqsmtemp = h_aqsm.data();
    p.aqsntemp = h_aqsn.data();
    p.I_epsR_array = h_er.data();
    p.I_epsA_array = h_ea.data();
    p.vcoul = h_vc.data();
    p.inv_igp_index = h_inv.data();
    p.indinv = h_ind.data();
    p.ekq = h_ek.data();
    p.dFreqGrid = h_gr.data();
    p.pref = h_pr.data();
    p.dFreqBrd = h_brd.data();
    p.schDt_matrix = h_sd.data();
  } else {
    gpp_c128 *aqsm = dalloc<gpp_c128>(nbnc), *aqsn = dalloc<gpp_c128>(nbnc);
    gpp_c128 *er = dalloc<gpp_c128>(eps), *ea = dalloc<gpp_c128>(eps);
    double *vc = dalloc<double>(NC), *ek = dalloc<double>(NB);
    double *gr = dalloc<double>(nF), *pr = dalloc<double>(nF);
    int *inv = dalloc<int>(NG), *ind = dalloc<int>(NC);
    gpp_c128 *brd = dalloc<gpp_c128>(nF), *sd = dalloc<gpp_c128>((size_t)NB * nF);
    dev = {aqsm, aqsn, er, ea, vc, ek, gr, pr, inv, ind, brd, sd};
    for (void *q : dev)
      if (!q) {
        std::fprintf(stderr, "device allocation failed\n");
        return 1;
      }
    rc = ff_fill_inputs_device(config, sem, seed, NB, NC, NG, nF, aqsm, aqsn,
                               er, ea, vc, inv, ind, ek, gr, pr, brd, sd, scal,
                               nullptr);
    p.aqsmtemp = aqsm;
    p.aqsntemp = aqsn;
    p.I_epsR_array = er;
    p.I_epsA_array = ea;
    p.vcoul = vc;
    p.inv_igp_index = inv;
    p.indinv = ind;
    p.ekq = ek;
    p.dFreqGrid = gr;
    p.pref = pr;
    p.dFreqBrd = brd;
    p.schDt_matrix = sd;
  }
  if (rc) {
    std::fprintf(stderr, "input generation failed: %d\n", rc);
    return 1;
  }
  p.freqevalmin = scal[0];
  p.freqevalstep = scal[1];
  p.pref_zb = scal[2];
  p.occ = scal[3];
  const std::chrono::duration<double> pre =
      std::chrono::high_resolution_clock::now() - t_start;
  std::printf("********** PreLoop **********= %g secs\n", pre.count());
  std::printf("starting Kernels\n");

  gpp_handle h = nullptr;
  if ((rc = gpp_create(&h, -1))) {
    std::fprintf(stderr, "gpp_create failed: %d\n", rc);
    return 1;
  }
  std::vector<gpp_c128> asx(NE), ach(NE), cor(NE), corb(NE), a2(NE);
  ff_options o;
  ff_default_options(&o);
  o.semantics = sem;
  o.memory = host ? GPP_MEM_HOST : GPP_MEM_DEVICE;
  ff_result r{};
  r.asxDtemp = asx.data();
  r.achDtemp = ach.data();
  r.achDtemp_cor = cor.data();
  r.achDtemp_corb = corb.data();
  r.ach2Dtemp = a2.data();
  const auto t0 = std::chrono::high_resolution_clock::now();
  rc = ff_sigma_ex(h, &p, &o, &r);
  const std::chrono::duration<double> dt =
      std::chrono::high_resolution_clock::now() - t0;
  if (rc) {
    std::fprintf(stderr, "ff_sigma_ex failed: %d %s\n", rc, gpp_last_error(h));
    return 1;
  }
  std::printf("********** slice sums (I_eps) kernel time  **********= %g secs\n",
              r.slice_kernel_ms * 1e-3);
  std::printf("********** frequency-grid kernel time **********= %g secs\n",
              r.freq_kernel_ms * 1e-3);
  print_c("achsDtemp", r.achsDtemp);
  gpp_c128 a0 = NE > 0 ? asx[0] : gpp_c128{0, 0};
  if (sem == FF_SEM_CPP && NE > 0) {
    a0.re += ach[0].re;
    a0.im += ach[0].im;
  }
  print_c("asxDtemp", a0);
  print_c("achDtemp_cor", NE > 0 ? cor[0] : gpp_c128{0, 0});
  std::printf("********** Total Kernel Time **********= %g secs\n", dt.count());
  gpp_destroy(h);
  for (void *q : dev) (void)hipFree(q);
  return 0;
}
