/*
 * gpp_sigma.h — C ABI of the MI355X (gfx950) general-plasmon-pole (GPP)
 * self-energy sigma sum.
 *
 * The library replaces the band loop of the gppKer mini-app
 * (gppKer.cpp:270-361): for every band n1, G' column my_igp, frequency iw
 * and G-vector ig it evaluates
 *
 *   wdiff = wx[n1][iw] - wtilde[my_igp][ig]
 *   delw  = wtilde[my_igp][ig] * conj(wdiff) / |wdiff|^2
 *   sch   = conj(aqsm[n1][igp]) * aqsn[n1][ig] * delw * I_eps[my_igp][ig]
 *
 * (plus the occupied-band screened-exchange branch for n1 < nvband,
 * gppKer.cpp:16-62,122-146) and returns achtemp[3] and asxtemp[3].
 *
 * Every array argument uses the reference layout: row-major, ig fastest,
 * complex numbers as interleaved {double re, im} (gppKer.cpp:215-221,
 * 244-252).  gpp_c128 is layout-compatible with std::complex<double>, the
 * reference's GPUComplex (Complex.h:18-22) and HIP's double2.
 *
 * Conventions (all entry points):
 *   - return 0 (GPP_OK) on success, a negative GPP_E* code on failure;
 *     nothing throws or exits across the ABI;
 *   - the caller owns every input and output buffer;
 *   - results are deterministic for a given shape, precision and tile
 *     configuration (fixed-order two-pass reduction, no atomics);
 *   - one gpp_handle per host thread; calls on one handle are serialised.
 */
#ifndef GPP_SIGMA_H
#define GPP_SIGMA_H

/* Exported entry points (the library hides every other symbol). */
#ifndef GPP_API
#define GPP_API __attribute__((visibility("default")))
#endif

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double re, im; } gpp_c128;

/* Return codes. */
#define GPP_OK           0
#define GPP_EINVAL      -1   /* bad shape, null pointer, bad option        */
#define GPP_ENOMEM      -2   /* device allocation failed                   */
#define GPP_EDEVICE     -3   /* HIP runtime error (launch, copy, device)   */
#define GPP_ENONFINITE  -4   /* a result is NaN/Inf and the caller asked
                                for GPP_FLAG_CHECK_FINITE                  */
#define GPP_EUNSUPPORTED -5  /* option combination not implemented         */
#define GPP_ESCRATCH    -6   /* gpp_sigma_noalloc: the problem needs more
                                scratch than GPP_NOALLOC_SCRATCH_BYTES      */

/* Arithmetic of the non-occupied (n1 >= nvband) sum. */
#define GPP_PREC_FP64    0   /* complex128 throughout                      */
#define GPP_PREC_MIXED   1   /* FP32 resolvent and products summed in FP32
                                over each lane's run of tile_ig G-vectors,
                                FP64 band and cross-tile accumulation; a run
                                holding a term with |wx - wtilde|^2 below
                                resonance_threshold (or, bgw, a term within
                                4e-6 relative of the mask's threshold, or a
                                non-finite wtilde or frequency) is
                                evaluated in FP64 instead, through a per-
                                wave work list                            */
#define GPP_PREC_FP64_FACTORED 2 /* FP64 with the band sum taken first:
                                P[my_igp][ig] = sum_n1 conj(aqsm[n1][igp]) *
                                aqsn[n1][ig] on the FP64 matrix cores
                                (three real products per complex one; the
                                error is bounded normwise, as for four),
                                then sum P * delw * I_eps; the occupied bands'
                                products are likewise summed before their
                                solver; acht_n1_loc by a second product
                                over the G-vectors.  Needs wx shared by
                                all bands (wx_per_band = 0); batches run
                                their problems one after another; no
                                tile tuning                               */

/* Which variant of the reference physics to follow. */
#define GPP_SEM_GPPKER         0 /* gppKer.cpp: unmasked non-occupied sch;
                                    occupied branch of ssxt_scht_solver with
                                    I_eps indexed [my_igp][ig];
                                    asxtemp += ssx*occ*vcoul               */
#define GPP_SEM_GPPKER_STRICT  1 /* as GPPKER but the occupied-branch sch and
                                    cutoff read I_eps[my_igp*ngpown+ig], the
                                    index gppKer.cpp:139 uses              */
#define GPP_SEM_BGW            2 /* gppkernel.deslippe.09.30.2014.f90: masked
                                    non-occupied sch, cden = wx^2 - wtilde^2,
                                    cutoff for wx < 0, asxtemp -= ...      */

/* Where the problem arrays live. */
#define GPP_MEM_HOST     0   /* host pointers (pageable is fine); copied
                                through a pinned ring by host threads, and
                                I_eps/wtilde in chunks of G' rows whose
                                kernel launches overlap the next chunk's
                                copy; same bits as device inputs         */
#define GPP_MEM_DEVICE   1   /* device pointers on the handle's device     */

/* gpp_options.flags */
#define GPP_FLAG_ALL_BANDS_UNOCC  1  /* treat every band as non-occupied
                                        (achtemp_kernel semantics,
                                        gppKer_gpuComplex.cpp:148-206)     */
#define GPP_FLAG_SKIP_UNOCC       2  /* evaluate only n1 < nvband
                                        (tillNvband_kernel semantics,
                                        gppKer_gpuComplex.cpp:127-145)     */
#define GPP_FLAG_CHECK_FINITE     4  /* return GPP_ENONFINITE on NaN/Inf   */
#define GPP_FLAG_NO_ACHSTEMP      8  /* skip the static remainder achstemp */
#define GPP_FLAG_AUTOTUNE        16  /* tune the tile shape for this shape
                                        class first if it is not cached    */
#define GPP_FLAG_MIXED_KAHAN     32  /* GPP_PREC_MIXED: Kahan-compensated
                                        FP32 run sums (4 more FP32
                                        operations per term)               */

/* One sigma evaluation.  Shapes and arrays as in gppKer.cpp:157-264. */
typedef struct {
  int number_bands;               /* NB                                    */
  int nvband;                     /* occupied bands, n1 < nvband           */
  int ncouls;                     /* NC, G-vectors                         */
  int ngpown;                     /* NG, G' columns owned (ncouls /
                                     nodes_per_group, gppKer.cpp:163)      */
  const gpp_c128 *aqsmtemp;       /* [NB][NC]                              */
  const gpp_c128 *aqsntemp;       /* [NB][NC]                              */
  const gpp_c128 *I_eps_array;    /* [NG][NC]                              */
  const gpp_c128 *wtilde_array;   /* [NG][NC]                              */
  const double   *vcoul;          /* [NC]                                  */
  const int      *inv_igp_index;  /* [NG]  my_igp -> indigp               */
  const int      *indinv;         /* [NC]  indigp -> igp; indigp == NC maps
                                     to igp = NC-1 (gppKer.cpp:287-290)    */
  const double   *wx_array;       /* [3] or [NB][3] (wx_per_band)         */
  int wx_per_band;                /* 0: one wx[3] for all bands;
                                     1: band-resolved wx[n1][3]            */
} gpp_problem;

typedef struct {
  int precision;                  /* GPP_PREC_*                            */
  int semantics;                  /* GPP_SEM_*                             */
  int memory;                     /* GPP_MEM_*                             */
  int flags;                      /* GPP_FLAG_*                            */
  void *stream;                   /* hipStream_t for GPP_MEM_DEVICE work;
                                     NULL = the handle's own stream        */
  double resonance_threshold;     /* mixed: a lane run with a term of
                                     |wdiff|^2 below this is evaluated in
                                     FP64; <= 0 selects the default 1e-6
                                     (FP32 underflow and exact resonances);
                                     larger values trade speed for
                                     accuracy where near-resonant terms
                                     dominate the sum (4e-3 met 1e-8 on
                                     the band-resolved configurations
                                     measured)                            */
  int tile_ig;                    /* G-vectors per lane run (0 = tuned)    */
  int waves_per_block;            /* 0 = tuned                             */
} gpp_options;

typedef struct {
  gpp_c128 achtemp[3];
  gpp_c128 asxtemp[3];
  gpp_c128 achstemp;              /* static remainder (gppKer.cpp:64-120)  */
  gpp_c128 *acht_n1_loc;          /* optional [NB] host buffer supplied by
                                     the caller, NULL to skip
                                     (gppKer.cpp:357,370-373)             */
  double unocc_kernel_ms;         /* n1 >= nvband kernel, HIP events on the
                                     stream it ran on (host inputs: its
                                     chunked launches and the copies they
                                     wait for)                             */
  double occ_kernel_ms;           /* n1 < nvband kernel                    */
  double total_ms;                /* whole device section incl. reduction  */
  double terms;                   /* NB * NG * NC * 3 evaluated            */
  long long fp64_fallback_terms;  /* mixed: terms redone in FP64 (0 if not
                                     counted)                              */
} gpp_result;

typedef struct gpp_handle_s *gpp_handle;

/* Handle lifetime.  device < 0 selects the current HIP device. */
GPP_API int gpp_create(gpp_handle *h, int device);
GPP_API int gpp_destroy(gpp_handle h);
/* Error text of the handle's last failed call; gpp_last_error(NULL) gives
 * that of the calling thread's last failed call without a handle (the
 * drop-in entry points, gpp_sigma_noalloc).  Never NULL. */
GPP_API const char *gpp_last_error(gpp_handle h);
GPP_API void gpp_default_options(gpp_options *o);

/* The general entry point. */
GPP_API int gpp_sigma_ex(gpp_handle h, const gpp_problem *p, const gpp_options *o,
                 gpp_result *r);

/* Many independent evaluations of uneven size (FP64 or mixed) in one
 * launch per band class: the blocks of all problems form one work list,
 * sorted largest-first, run by a single grid; the sequence is captured in a
 * hipGraph that is replayed while the batch (pointers, shapes, options) is
 * unchanged (a repeated device-memory call replays it without planning
 * again; the arrays' contents may change between calls).  All problems use
 * the same options and memory space; band-range flags and acht_n1_loc are
 * not available here.  GPP_PREC_FP64_FACTORED runs the problems one after
 * another. */
GPP_API int gpp_sigma_batch(gpp_handle h, const gpp_problem *p, int n,
                    const gpp_options *o, gpp_result *r);

/* Tile autotuner: times the non-occupied kernel on a slice of at most 32
 * unoccupied bands of `p` for every candidate (G-vectors per lane run x
 * waves per block), keeps the fastest for the problem's shape class
 * (precision, semantics, log2 ncouls, log2 ngpown) in the handle, appends it
 * to the file named by $GPP_TUNE_CACHE (read back by gpp_create), and
 * returns it.  Later calls with tile_ig = waves_per_block = 0 use it. */
GPP_API int gpp_autotune(gpp_handle h, const gpp_problem *p, const gpp_options *o,
                 int *tile_ig, int *waves_per_block);

/*
 * Drop-in for the band loop of gppKer.cpp:270-361.  Host pointers,
 * FP64, GPP_SEM_GPPKER, a library-owned handle on the current device.
 * wx_array holds 3 values shared by all bands, as gppKer.cpp:276-280
 * computes them; |wx| < 1e-6 is clamped to 1e-6 as there.  Since the
 * frequencies are shared, the sum is evaluated band sum first
 * (GPP_PREC_FP64_FACTORED: the same terms in FP64, another summation
 * order).  achtemp and asxtemp receive 3 values each.
 */
GPP_API int gpp_sigma(int number_bands, int nvband, int ncouls, int ngpown,
              const gpp_c128 *aqsmtemp, const gpp_c128 *aqsntemp,
              const gpp_c128 *I_eps_array, const gpp_c128 *wtilde_array,
              const double *vcoul, const int *inv_igp_index,
              const int *indinv, const double *wx_array,
              gpp_c128 *achtemp, gpp_c128 *asxtemp);

/*
 * Drop-ins for the reference's operator split (gppKer_gpuComplex.cpp).
 * gpp_achtemp_kernel replaces achtemp_kernel (:148): every band taken as
 * non-occupied (band sum first, its wx_array being shared by all bands),
 * results added into the six out-references.
 * gpp_tillNvband_kernel replaces tillNvband_kernel (:127): bands
 * n1 < nvband, asxtemp[iw] += ssxt * occ * vcoul[igp].
 * Both take host pointers and run FP64 on the current device.
 */
GPP_API int gpp_achtemp_kernel(int ncouls, int ngpown, const double *wx_array,
                       const gpp_c128 *wtilde_array, const gpp_c128 *aqsmtemp,
                       const gpp_c128 *aqsntemp, const gpp_c128 *I_eps_array,
                       const double *vcoul, const int *inv_igp_index,
                       const int *indinv, int number_bands,
                       double *achtemp_re0, double *achtemp_re1,
                       double *achtemp_re2, double *achtemp_im0,
                       double *achtemp_im1, double *achtemp_im2);

GPP_API int gpp_tillNvband_kernel(const int *inv_igp_index, const int *indinv,
                          gpp_c128 *asxtemp, const double *wx_array,
                          const gpp_c128 *wtilde_array,
                          const gpp_c128 *aqsmtemp, const gpp_c128 *aqsntemp,
                          const gpp_c128 *I_eps_array, int ncouls, int nvband,
                          int ngpown, const double *vcoul, int nstart,
                          int nend, double occ, int number_bands);

/* Synthetic inputs (host memory, caller-allocated, reference layout).
 * config: 1 = gppKer.cpp constant fills; 2 = canonical seeded random;
 * 3 = band-resolved energies with near-resonant wtilde (needs wx[NB][3]).
 * nodes_per_group gives the inv_igp_index stride of configs 2 and 3.
 * config | GPP_FILL_VCOUL_LINEAR replaces vcoul by vcoul[i] = i / ncouls
 * (a Coulomb factor proportional to i, in place of the file's 1.0 or the
 * canonical 1/(0.01 + (i/NC)^(2/3))).
 * Deterministic for (config, seed, shape) on every host. */
#define GPP_FILL_VCOUL_LINEAR 256
GPP_API int gpp_fill_inputs(int config, unsigned long long seed, int number_bands,
                    int ncouls, int ngpown, int nodes_per_group,
                    gpp_c128 *aqsmtemp, gpp_c128 *aqsntemp,
                    gpp_c128 *I_eps_array, gpp_c128 *wtilde_array,
                    double *vcoul, int *inv_igp_index, int *indinv,
                    double *wx_array);

/* The same inputs generated in device memory (all pointers are device
 * pointers on the current device; stream is a hipStream_t or NULL).  Returns
 * after the stream has finished.  I_eps, wtilde, vcoul, the index tables and
 * wx are bit-identical to gpp_fill_inputs; aqsm/aqsn (normal variates) use
 * the device's log/sin/cos and agree with the host's to a few ulp.  For
 * shapes too large to stage from the host on every run. */
GPP_API int gpp_fill_inputs_device(int config, unsigned long long seed,
                           int number_bands, int ncouls, int ngpown,
                           int nodes_per_group, gpp_c128 *aqsmtemp,
                           gpp_c128 *aqsntemp, gpp_c128 *I_eps_array,
                           gpp_c128 *wtilde_array, double *vcoul,
                           int *inv_igp_index, int *indinv, double *wx_array,
                           void *stream);

/*
 * Allocation-free evaluation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory during the call (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * It is gpp_sigma_ex for device-resident inputs (o->memory must be
 * GPP_MEM_DEVICE) with FP64 or GPP_PREC_FP64_FACTORED arithmetic: the same
 * kernels and record layout, so the same bits for the same tile (tile_ig /
 * waves_per_block 0 select K=8 x 4 waves, the untuned default).  Its scratch
 * is a library-owned device array of GPP_NOALLOC_SCRATCH_BYTES per GPU,
 * held until the process ends; a problem that needs more (about
 * 16*number_bands*ngpown bytes plus 96 per block record; the canonical
 * 512/2/32768/20 shape takes 19 MB) is refused with GPP_ESCRATCH before
 * anything is launched.  The call enqueues its kernels and one
 * device-to-device copy on o->stream (NULL: the null stream) and returns
 * without waiting; out is a device buffer of 7 gpp_c128: achtemp[3],
 * asxtemp[3], achstemp.  Every call on one GPU shares that GPU's scratch:
 * calls that overlap on one GPU (two streams, two host threads) race and
 * must be serialised by the caller.  Not supported: mixed precision,
 * acht_n1_loc, GPP_FLAG_AUTOTUNE and GPP_FLAG_CHECK_FINITE (EUNSUPPORTED).
 * Errors: gpp_last_error(NULL).
 */
#define GPP_NOALLOC_SCRATCH_BYTES (64ull << 20)
GPP_API int gpp_sigma_noalloc(const gpp_problem *p, const gpp_options *o,
                              gpp_c128 *out);

/* Version string of the library build. */
GPP_API const char *gpp_version(void);

#ifdef __cplusplus
}
#endif
#endif /* GPP_SIGMA_H */
