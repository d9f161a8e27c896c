// ff_ref_kat.cpp — known-answer harness around the reference's own
// full-frequency kernels (TEST INFRASTRUCTURE ONLY).
//
// Linked against an object compiled, unmodified, from
// FFKer_cori/ffKernel.cpp with its main renamed on the compiler command line
// (see Makefile).  Calls, in the order the reference's main does
// (ffKernel.cpp:482-505):
//   achsDtemp_Kernel :107, asxDtemp_Kernel :133, achDtemp_Kernel :155,
//   achDtemp_cor_Kernel :216
// with its own output arrays (achDtemp_Kernel gets a separate achDtemp
// array rather than main's asxDtemp).  Run it with OMP_NUM_THREADS=1:
// achDtemp_Kernel updates shared arrays from a parallel loop.
//
// Input (binary, little endian, stdin): int32 NB, NV, NC, NG, nFreq,
// nfreqeval; float64 freqevalmin, freqevalstep, pref_zb, occ; then
// aqsm[NB][NC], aqsn[NB][NC], I_epsR[nFreq][NG][NC], I_epsA[nFreq][NG][NC]
// (complex128); vcoul[NC] (float64); inv_igp_index[NG], indinv[NC] (int32);
// ekq[NB], dFreqGrid[nFreq], pref[nFreq] (float64); dFreqBrd[nFreq],
// schDt_matrix[NB][nFreq] (complex128).
// Output (text): "achsDtemp re im", then one line per iw:
//   iw asxDtemp.re .im achDtemp.re .im achDtemp_cor.re .im
#include <cstdio>
#include <cstdlib>
#include <vector>

// Layout-compatible stand-in for the reference's two-double class (only
// pointers and references to it cross the call).
struct GPUComplex {
  double re, im;
};

void achsDtemp_Kernel(int number_bands, int ngpown, int ncouls,
                      int *inv_igp_index, int *indinv, GPUComplex *aqsntemp,
                      GPUComplex *aqsmtemp, GPUComplex *I_epsR_array,
                      double *vcoul, GPUComplex &achsDtemp);
void asxDtemp_Kernel(int nvband, int nfreqeval, int ncouls, int ngpown,
                     int numThreads, int nFreq, double freqevalmin,
                     double freqevalstep, double occ, double *ekq,
                     double *dFreqGrid, int *inv_igp_index, int *indinv,
                     GPUComplex *aqsmtemp, GPUComplex *aqsntemp, double *vcoul,
                     GPUComplex *I_epsR_array, GPUComplex *ssxDittt,
                     GPUComplex *asxDtemp);
void achDtemp_Kernel(int number_bands, int nvband, int nfreqeval, int ncouls,
                     int ngpown, int nFreq, double freqevalmin,
                     double freqevalstep, double *ekq, double pref_zb,
                     double *pref, double *dFreqGrid, GPUComplex *dFreqBrd,
                     GPUComplex *schDt_matrix, GPUComplex *schDi,
                     GPUComplex *schDi_cor, GPUComplex *sch2Di,
                     GPUComplex *achDtemp);
void achDtemp_cor_Kernel(int number_bands, int nvband, int nfreqeval,
                         int ncouls, int ngpown, int nFreq, double freqevalmin,
                         double freqevalstep, double *ekq, double *dFreqGrid,
                         int *inv_igp_index, int *indinv, GPUComplex *aqsmtemp,
                         GPUComplex *aqsntemp, double *vcoul,
                         GPUComplex *I_epsR_array, GPUComplex *I_epsA_array,
                         GPUComplex *schDi_cor, GPUComplex *schDi_corb,
                         GPUComplex *sch2Di, GPUComplex *ach2Dtemp,
                         GPUComplex *achDtemp_cor, GPUComplex *achDtemp_corb);

template <class T>
static void rd(std::vector<T> &v, size_t n) {
  v.assign(n, T());
  if (n && std::fread(v.data(), sizeof(T), n, stdin) != n) {
    std::fprintf(stderr, "short input\n");
    std::exit(2);
  }
}

int main() {
  int h[6];
  double d[4];
  if (std::fread(h, sizeof(int), 6, stdin) != 6) return 2;
  if (std::fread(d, sizeof(double), 4, stdin) != 4) return 2;
  const int NB = h[0], NV = h[1], NC = h[2], NG = h[3], NF = h[4], NE = h[5];
  std::vector<GPUComplex> aqsm, aqsn, er, ea, brd, schdt;
  std::vector<double> vcoul, ekq, grid, pref;
  std::vector<int> inv, indinv;
  rd(aqsm, (size_t)NB * NC);
  rd(aqsn, (size_t)NB * NC);
  rd(er, (size_t)NF * NG * NC);
  rd(ea, (size_t)NF * NG * NC);
  rd(vcoul, (size_t)NC);
  rd(inv, (size_t)NG);
  rd(indinv, (size_t)NC);
  rd(ekq, (size_t)NB);
  rd(grid, (size_t)NF);
  rd(pref, (size_t)NF);
  rd(brd, (size_t)NF);
  rd(schdt, (size_t)NB * NF);

  GPUComplex achs = {0, 0};
  std::vector<GPUComplex> asx(NE), ach(NE), cor(NE), corb(NE), ach2(NE);
  std::vector<GPUComplex> schDi(NE), schDi_cor(NE), sch2Di(NE), ssx(64);
  achsDtemp_Kernel(NB, NG, NC, inv.data(), indinv.data(), aqsn.data(),
                   aqsm.data(), er.data(), vcoul.data(), achs);
  asxDtemp_Kernel(NV, NE, NC, NG, 1, NF, d[0], d[1], d[3], ekq.data(),
                  grid.data(), inv.data(), indinv.data(), aqsm.data(),
                  aqsn.data(), vcoul.data(), er.data(), ssx.data(),
                  asx.data());
  achDtemp_Kernel(NB, NV, NE, NC, NG, NF, d[0], d[1], ekq.data(), d[2],
                  pref.data(), grid.data(), brd.data(), schdt.data(),
                  schDi.data(), schDi_cor.data(), sch2Di.data(), ach.data());
  std::vector<GPUComplex> cor_w(NE), corb_w(NE), s2_w(NE);
  achDtemp_cor_Kernel(NB, NV, NE, NC, NG, NF, d[0], d[1], ekq.data(),
                      grid.data(), inv.data(), indinv.data(), aqsm.data(),
                      aqsn.data(), vcoul.data(), er.data(), ea.data(),
                      cor_w.data(), corb_w.data(), s2_w.data(), ach2.data(),
                      cor.data(), corb.data());
  std::printf("achsDtemp %.17g %.17g\n", achs.re, achs.im);
  for (int iw = 0; iw < NE; ++iw)
    std::printf("%d %.17g %.17g %.17g %.17g %.17g %.17g\n", iw, asx[iw].re,
                asx[iw].im, ach[iw].re, ach[iw].im, cor[iw].re, cor[iw].im);
  return 0;
}
