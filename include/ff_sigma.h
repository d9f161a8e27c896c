/*
 * ff_sigma.h — C ABI of the MI355X (gfx950) full-frequency (FF) self-energy
 * kernels of the FortKernels FF mini-app (FFKer_cori).
 *
 * For bands n1, evaluation frequencies iw and the frequency grid
 * dFreqGrid[nFreq], the mini-app forms
 *   achsDtemp  = -1/2 sum_{n1,G',G} aqsn[n1][G] conj(aqsm[n1][G']) I_epsR[f0][G'][G] vcoul[G']
 *   asxDtemp   occupied-band screened exchange, I_eps interpolated at wx
 *   achDtemp, achDtemp_corb, ach2Dtemp, achDtemp_cor
 *              Coulomb-hole terms: a sum over the frequency grid of
 *              schDt_matrix and, for achDtemp_cor/ach2Dtemp, I_epsR - I_epsA
 *              interpolated at wx
 * (FFKer_cori/ffKernel.cpp:80-300, ffKernel.f90:199-467).  The two sources
 * disagree in several places, so the semantics is an option:
 *   FF_SEM_CPP : ffKernel.cpp as written (the same wx for every iw in the
 *                asx and cor loops, I_epsR slice 1 in achsDtemp, I_epsR only
 *                in asxDtemp, compute_fact's grid search)
 *   FF_SEM_F90 : ffKernel.f90 (wx = freqevalmin - ekq + (iw-1)*freqevalstep,
 *                I_epsA below zero frequency, slice 0 in achsDtemp, the
 *                full frequency-grid terms)
 *
 * Layout as ffKernel.cpp (row-major, ig fastest): aqsm/aqsn [NB][NC],
 * I_epsR/I_epsA [nFreq][NG][NC], schDt_matrix [NB][nFreq]; complex numbers
 * are gpp_c128 {re, im}.  Conventions (return codes, handles, ownership) as
 * in gpp_sigma.h.  The host planning and assembly of an FF call run on up
 * to 7 worker threads of the handle besides the caller's, started by the
 * handle's first FF call and joined by gpp_destroy.
 */
#ifndef FF_SIGMA_H
#define FF_SIGMA_H

#include "gpp_sigma.h"

#ifdef __cplusplus
extern "C" {
#endif

#define FF_SEM_CPP 0
#define FF_SEM_F90 1

/* ff_options.parts: which outputs to compute (0 = all). */
#define FF_PART_ACHS  1   /* achsDtemp                                  */
#define FF_PART_ASX   2   /* asxDtemp                                   */
#define FF_PART_ACH   4   /* achDtemp (+ achDtemp_corb, the grid terms of
                             ach2Dtemp/achDtemp_cor for FF_SEM_F90)     */
#define FF_PART_COR   8   /* the I_eps terms of achDtemp_cor, ach2Dtemp */

typedef struct {
  int number_bands;               /* NB                                    */
  int nvband;                     /* occupied bands, n1 < nvband           */
  int ncouls;                     /* NC                                    */
  int ngpown;                     /* NG                                    */
  int nFreq;                      /* frequency grid points (>= 2)          */
  int nfreqeval;                  /* evaluation frequencies                */
  const gpp_c128 *aqsmtemp;       /* [NB][NC]                              */
  const gpp_c128 *aqsntemp;       /* [NB][NC]                              */
  const gpp_c128 *I_epsR_array;   /* [nFreq][NG][NC]                       */
  const gpp_c128 *I_epsA_array;   /* [nFreq][NG][NC]                       */
  const double   *vcoul;          /* [NC]                                  */
  const int      *inv_igp_index;  /* [NG]                                  */
  const int      *indinv;         /* [NC]                                  */
  const double   *ekq;            /* [NB]                                  */
  const double   *dFreqGrid;      /* [nFreq]                               */
  const double   *pref;           /* [nFreq]                               */
  const gpp_c128 *dFreqBrd;       /* [nFreq]                               */
  const gpp_c128 *schDt_matrix;   /* [NB][nFreq]                           */
  double freqevalmin;
  double freqevalstep;
  double pref_zb;
  double occ;
} ff_problem;

typedef struct {
  int semantics;                  /* FF_SEM_*                              */
  int memory;                     /* GPP_MEM_* for every array of the
                                     problem                               */
  void *stream;                   /* hipStream_t for GPP_MEM_DEVICE work;
                                     NULL = the handle's own stream        */
  int parts;                      /* FF_PART_* mask, 0 = all               */
} ff_options;

typedef struct {
  gpp_c128 achsDtemp;
  gpp_c128 *asxDtemp;             /* [nfreqeval] host buffers supplied by  */
  gpp_c128 *achDtemp;             /* the caller (NULL to skip); they are   */
  gpp_c128 *achDtemp_cor;         /* overwritten                           */
  gpp_c128 *achDtemp_corb;
  gpp_c128 *ach2Dtemp;
  double slice_kernel_ms;         /* I_eps slice sums (HIP events)         */
  double freq_kernel_ms;          /* frequency-grid terms                  */
  double total_ms;                /* whole device section                  */
  double slice_bytes;             /* I_eps bytes the slice kernel read     */
  long long slices;               /* distinct (array, frequency) slices    */
  long long slice_terms;          /* (n1, slice, G', G) products formed    */
  long long freq_terms;           /* (n1, ifreq, iw) grid terms            */
  double host_plan_ms;            /* host: grid searches and pass packing  */
  double host_assemble_ms;        /* host: interpolations of the sums      */
} ff_result;

GPP_API void ff_default_options(ff_options *o);

/* The general entry point. */
GPP_API int ff_sigma_ex(gpp_handle h, const ff_problem *p, const ff_options *o,
                ff_result *r);

/*
 * Drop-ins for the kernels of FFKer_cori/ffKernel.cpp (FF_SEM_CPP, host
 * pointers, a library-owned handle on the current device), same argument
 * order with gpp_c128 for GPUComplex.  Like the reference they add into
 * their output arrays (asxDtemp, achDtemp, ach2Dtemp, achDtemp_cor,
 * achDtemp_corb) and set achsDtemp.  Arguments that are the reference's
 * scratch (numThreads, ssxDittt, schDi, schDi_cor, schDi_corb, sch2Di) are
 * accepted and not used.
 */
GPP_API int ff_achsDtemp_Kernel(int number_bands, int ngpown, int ncouls,
                        const int *inv_igp_index, const int *indinv,
                        const gpp_c128 *aqsntemp, const gpp_c128 *aqsmtemp,
                        const gpp_c128 *I_epsR_array, const double *vcoul,
                        gpp_c128 *achsDtemp);                      /* :107 */

GPP_API int ff_asxDtemp_Kernel(int nvband, int nfreqeval, int ncouls, int ngpown,
                       int numThreads, int nFreq, double freqevalmin,
                       double freqevalstep, double occ, const double *ekq,
                       const double *dFreqGrid, const int *inv_igp_index,
                       const int *indinv, const gpp_c128 *aqsmtemp,
                       const gpp_c128 *aqsntemp, const double *vcoul,
                       const gpp_c128 *I_epsR_array, gpp_c128 *ssxDittt,
                       gpp_c128 *asxDtemp);                        /* :133 */

GPP_API int ff_achDtemp_Kernel(int number_bands, int nvband, int nfreqeval,
                       int ncouls, int ngpown, int nFreq, double freqevalmin,
                       double freqevalstep, const double *ekq, double pref_zb,
                       const double *pref, const double *dFreqGrid,
                       const gpp_c128 *dFreqBrd, const gpp_c128 *schDt_matrix,
                       gpp_c128 *schDi, gpp_c128 *schDi_cor, gpp_c128 *sch2Di,
                       gpp_c128 *achDtemp);                        /* :155 */

GPP_API int ff_achDtemp_cor_Kernel(int number_bands, int nvband, int nfreqeval,
                           int ncouls, int ngpown, int nFreq,
                           double freqevalmin, double freqevalstep,
                           const double *ekq, const double *dFreqGrid,
                           const int *inv_igp_index, const int *indinv,
                           const gpp_c128 *aqsmtemp, const gpp_c128 *aqsntemp,
                           const double *vcoul, const gpp_c128 *I_epsR_array,
                           const gpp_c128 *I_epsA_array, gpp_c128 *schDi_cor,
                           gpp_c128 *schDi_corb, gpp_c128 *sch2Di,
                           gpp_c128 *ach2Dtemp, gpp_c128 *achDtemp_cor,
                           gpp_c128 *achDtemp_corb);               /* :216 */

/* Synthetic inputs in host memory (caller-allocated, layout above).
 * config 1 = the mini-app's own fills (ffKernel.cpp:385-432 for FF_SEM_CPP;
 * ffKernel.f90:87-188 for FF_SEM_F90, whose dFreqBrd differs); config 2 =
 * seeded random values (aqsm/aqsn ~ N(0,1)/sqrt(NC), I_eps smooth in
 * frequency, sorted ekq in [-10, 10), schDt_matrix ~ N(0,1), strided
 * inv_igp_index).  Scalars (freqevalmin, freqevalstep, pref_zb, occ) are
 * written to *scal[4]. */
GPP_API int ff_fill_inputs(int config, int semantics, unsigned long long seed,
                   int number_bands, int ncouls, int ngpown, int nFreq,
                   gpp_c128 *aqsm, gpp_c128 *aqsn, gpp_c128 *I_epsR,
                   gpp_c128 *I_epsA, double *vcoul, int *inv_igp_index,
                   int *indinv, double *ekq, double *dFreqGrid, double *pref,
                   gpp_c128 *dFreqBrd, gpp_c128 *schDt_matrix, double *scal);

/* The same values generated in device memory on `stream` (hipStream_t,
 * NULL = default); the I_eps arrays are the bulk of an FF problem. */
GPP_API int ff_fill_inputs_device(int config, int semantics, unsigned long long seed,
                          int number_bands, int ncouls, int ngpown, int nFreq,
                          gpp_c128 *aqsm, gpp_c128 *aqsn, gpp_c128 *I_epsR,
                          gpp_c128 *I_epsA, double *vcoul, int *inv_igp_index,
                          int *indinv, double *ekq, double *dFreqGrid,
                          double *pref, gpp_c128 *dFreqBrd,
                          gpp_c128 *schDt_matrix, double *scal, void *stream);

#ifdef __cplusplus
}
#endif
#endif
