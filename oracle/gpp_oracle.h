/*
 * gpp_oracle.h — CPU FP64 oracle of the GPP sigma sum (test infrastructure;
 * only tests/, the smoke check and the benchmark's CPU-baseline leg use it).
 */
#ifndef GPP_ORACLE_H
#define GPP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define GPP_ORACLE_SEM_GPPKER 0
#define GPP_ORACLE_SEM_GPPKER_STRICT 1
#define GPP_ORACLE_SEM_BGW 2

#define GPP_ORACLE_ALL_BANDS_UNOCC 1
#define GPP_ORACLE_SKIP_UNOCC 2

/* Same fields and layout as gpp_problem in include/gpp_sigma.h. */
typedef struct {
  int number_bands, nvband, ncouls, ngpown;
  const double *aqsmtemp;     /* [NB][NC] interleaved re,im */
  const double *aqsntemp;     /* [NB][NC] */
  const double *I_eps_array;  /* [NG][NC] */
  const double *wtilde_array; /* [NG][NC] */
  const double *vcoul;        /* [NC] */
  const int *inv_igp_index;   /* [NG] */
  const int *indinv;          /* [NC] */
  const double *wx_array;     /* [3] or [NB][3] */
  int wx_per_band;
} gpp_oracle_problem;

/* out[0..5]: achtemp re/im per iw; out[6..11]: asxtemp re/im per iw.
 * achstemp (2 doubles) and acht_n1_loc (2*NB doubles) may be NULL.
 * Returns 0, or a negative value for bad arguments. */
int gpp_oracle_sigma(const gpp_oracle_problem *p, int semantics, int flags,
                     double *out, double *achstemp, double *acht_n1_loc);

int gpp_oracle_resolve_igp(int my_igp, int ncouls, const int *inv_igp_index,
                           const int *indinv);

int gpp_oracle_occ_sums(const gpp_oracle_problem *p, int semantics, int n1,
                        int my_igp, int igp, double wxt, double *ssxt,
                        double *scht);

#ifdef __cplusplus
}
#endif
#endif
