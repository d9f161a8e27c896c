// ref_kat.cpp — known-answer harness around the reference's own occupied-band
// solver (TEST INFRASTRUCTURE ONLY).
//
// Linked against an object compiled, unmodified, from one of the reference
// sources (main renamed on the compiler command line, see Makefile):
//   REF_BGWK : BGW-Kernels/gppKer.cpp, flagOCC_solver at :130 (I_eps read at
//              [my_igp][ig], :155)
//   REF_XL   : gppKerOpenMPFor_xl.cpp, which is gppKer.cpp without the ITT
//              calls; flagOCC_solver at :125 (I_eps read at
//              my_igp*ngpown+ig, gppKer.cpp:139)
//
// Input (binary, little endian, from stdin): int32 NB, NC, NG; then
// aqsm[NB][NC], aqsn[NB][NC], I_eps[NG][NC], wtilde[NG][NC] as complex128;
// wx[NB][3] as float64; igp[NG] as int32.
// Output (text, stdout): one line per (n1, my_igp, iw), n1 fastest last:
//   n1 my_igp iw ssxt.re ssxt.im scht.re scht.im
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <vector>

typedef std::complex<double> cd;

#if defined(REF_BGWK)
void flagOCC_solver(double wxt, cd *wtilde_array_tmp, int my_igp, int n1,
                    cd *aqsmtemp_arr, cd *aqsntemp_arr, cd *I_eps_array_tmp,
                    cd &ssxt, cd &scht, int igmax, int ncouls, int igp,
                    int number_bands, int ngpown);
#elif defined(REF_XL)
void flagOCC_solver(double wxt, cd *wtilde_array, int my_igp, int n1,
                    cd *aqsmtemp, cd *aqsntemp, cd *I_eps_array, cd &ssxt,
                    cd &scht, int ncouls, int igp, int number_bands,
                    int ngpown);
#else
#error "define REF_BGWK or REF_XL"
#endif

template <class T>
static void rd(std::vector<T> &v, size_t n) {
  v.resize(n);
  if (n && std::fread(v.data(), sizeof(T), n, stdin) != n) {
    std::fprintf(stderr, "short input\n");
    std::exit(2);
  }
}

int main() {
  int hdr[3];
  if (std::fread(hdr, sizeof(int), 3, stdin) != 3) return 2;
  const int NB = hdr[0], NC = hdr[1], NG = hdr[2];
  std::vector<cd> aqsm, aqsn, ieps, wt;
  std::vector<double> wx;
  std::vector<int> igp;
  rd(aqsm, (size_t)NB * NC);
  rd(aqsn, (size_t)NB * NC);
  rd(ieps, (size_t)NG * NC);
  rd(wt, (size_t)NG * NC);
  rd(wx, (size_t)NB * 3);
  rd(igp, (size_t)NG);
  for (int n1 = 0; n1 < NB; ++n1)
    for (int my_igp = 0; my_igp < NG; ++my_igp)
      for (int iw = 0; iw < 3; ++iw) {
        cd ssxt(0.0, 0.0), scht(0.0, 0.0);
#if defined(REF_BGWK)
        flagOCC_solver(wx[n1 * 3 + iw], wt.data(), my_igp, n1, aqsm.data(),
                       aqsn.data(), ieps.data(), ssxt, scht, NC, NC,
                       igp[my_igp], NB, NG);
#else
        flagOCC_solver(wx[n1 * 3 + iw], wt.data(), my_igp, n1, aqsm.data(),
                       aqsn.data(), ieps.data(), ssxt, scht, NC, igp[my_igp],
                       NB, NG);
#endif
        std::printf("%d %d %d %.17g %.17g %.17g %.17g\n", n1, my_igp, iw,
                    ssxt.real(), ssxt.imag(), scht.real(), scht.imag());
      }
  return 0;
}
