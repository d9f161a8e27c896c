/* ff_oracle.h — CPU restatement of the full-frequency (FF) self-energy
 * mini-app of FFKer_cori (TEST INFRASTRUCTURE ONLY: used by tests/, never
 * by the product library).
 *
 * Two semantics, as the reference's variants disagree:
 *   FF_ORC_CPP : FFKer_cori/ffKernel.cpp as written (:80-300)
 *   FF_ORC_F90 : FFKer_cori/ffKernel.f90 (:199-467)
 * Arrays use the C++ layout: aqsm/aqsn [NB][NC], I_epsR/I_epsA
 * [nFreq][NG][NC], schDt_matrix [NB][nFreq]; complex = {re, im} doubles.
 */
#ifndef FF_ORACLE_H
#define FF_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double re, im; } ffo_c;

#define FF_ORC_CPP 0
#define FF_ORC_F90 1

typedef struct {
  int number_bands, nvband, ncouls, ngpown, nFreq, nfreqeval;
  const ffo_c *aqsmtemp, *aqsntemp, *I_epsR_array, *I_epsA_array;
  const double *vcoul;
  const int *inv_igp_index, *indinv;
  const double *ekq, *dFreqGrid, *pref;
  const ffo_c *dFreqBrd, *schDt_matrix;
  double freqevalmin, freqevalstep, pref_zb, occ;
} ffo_problem;

/* Outputs: achsDtemp (scalar) and five [nfreqeval] arrays (any may be
 * NULL).  Returns 0, or -1 on a bad shape. */
int ff_oracle(int sem, const ffo_problem *p, ffo_c *achsDtemp,
              ffo_c *asxDtemp, ffo_c *achDtemp, ffo_c *achDtemp_cor,
              ffo_c *achDtemp_corb, ffo_c *ach2Dtemp);

/* The frequency-grid search and interpolation weights of the two variants
 * (exposed for unit tests).  loop 1 = asx, 2 = cor; returns ifreq (0-based
 * lower slice) or -1 when the variant computes nothing for this case. */
int ff_oracle_fact(int sem, int loop, int flag_occ, double wx, int nFreq,
                   const double *dFreqGrid, double *fact1, double *fact2);

#ifdef __cplusplus
}
#endif
#endif
