/*
 * gpp_oracle.c — CPU FP64 restatement of the gppKer GPP sigma sum.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the GPU library is
 * compared against.  Only tests/, the smoke check and the CPU leg of the
 * benchmark (baseline timing, result checks) load it; the product library
 * never links or calls it.
 *
 * What it restates (reference files under rgayatri23/FortKernels-):
 *   gpp_oracle_sigma     the band loop of gppKer.cpp:270-361 with the
 *                        solver ssxt_scht_solver (:16-62) and
 *                        flagOCC_solver (:122-146); the static remainder
 *                        reduce_achstemp (:64-120); the per-band log
 *                        acht_n1_loc (:357,370-373).  Semantics variants:
 *                          0 gppker  gppKer.cpp, occupied-branch I_eps read
 *                                    at [my_igp][ig] (as
 *                                    BGW-Kernels/gppKer.cpp:155)
 *                          1 strict  gppKer.cpp incl. its
 *                                    I_eps[my_igp*ngpown+ig] read (:139)
 *                          2 bgw     gppkernel.deslippe.09.30.2014.f90:
 *                                    masked non-occupied sum (:364-366),
 *                                    cden = wx^2 - wtilde^2 (:303), cutoff
 *                                    for wx < 0 (:321), asxtemp -= (:393),
 *                                    full static remainder (:195-229).
 *   Every term is evaluated in the reference's operation order with
 *   complex products written out as std::complex / GPUComplex_product
 *   define them (Complex.h:171-178).  Sums over ig use Neumaier
 *   compensation; the cross-(band, my_igp) sums use double-double, in a
 *   fixed order, so the result does not depend on the thread count.
 *
 * Deliberate differences from gppKer.cpp, all memory or race fixes:
 *   - a my_igp row whose igp is outside [0, ncouls) is skipped (the guard at
 *     :293 admits igp == ncouls and then reads past the row);
 *   - indigp == ncouls maps to igp = ncouls-1 without reading indinv[ncouls]
 *     (:288-290 read it first);
 *   - asxtemp[iw] accumulates ssx_array[iw]*occ*vcoul[igp] for each row (the
 *     reference indexes a shared slot, :352, and never prints it);
 *   - reduce_achstemp's accumulator schstemp is per row (the reference
 *     shares it between threads, :67,97).
 *
 * Parity pins (tests/): the reference binary's printed achtemp for its
 * constant fills; the reference's own flagOCC_solver on random inputs
 * (both index variants); closed forms; invariants.
 */
#include "gpp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  double re, im;
} c128;

static inline c128 cmk(double re, double im) {
  c128 z;
  z.re = re;
  z.im = im;
  return z;
}
static inline c128 cmul(c128 a, c128 b) {
  return cmk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static inline c128 cconj(c128 a) { return cmk(a.re, -a.im); }
static inline c128 cadd(c128 a, c128 b) { return cmk(a.re + b.re, a.im + b.im); }
static inline c128 cscale(c128 a, double s) { return cmk(a.re * s, a.im * s); }
/* real(z * conj(z)) */
static inline double cnorm2(c128 a) { return a.re * a.re + a.im * a.im; }
static inline double cabs_(c128 a) { return sqrt(a.re * a.re + a.im * a.im); }

/* Neumaier-compensated complex sum. */
typedef struct {
  double sr, cr, si, ci;
} ksum;
static inline void kadd1(double *s, double *c, double x) {
  const double t = *s + x;
  if (fabs(*s) >= fabs(x))
    *c += (*s - t) + x;
  else
    *c += (x - t) + *s;
  *s = t;
}
static inline void kadd(ksum *k, c128 z) {
  kadd1(&k->sr, &k->cr, z.re);
  kadd1(&k->si, &k->ci, z.im);
}
static inline c128 kval(const ksum *k) { return cmk(k->sr + k->cr, k->si + k->ci); }

/* Double-double accumulator (Dekker/Knuth two-sum). */
typedef struct {
  double hi, lo;
} dd;
static inline void dd_add(dd *a, double x) {
  const double s = a->hi + x;
  const double bp = s - a->hi;
  const double err = (a->hi - (s - bp)) + (x - bp);
  const double lo = a->lo + err;
  const double hi = s + lo;
  a->lo = lo - (hi - s);
  a->hi = hi;
}

static const double kTo1 = 1e-6;
static const double kSexcut = 4.0;
static const double kLimitOne = 1.0 / (1e-6 * 4.0);
static const double kLimitTwo = 0.25; /* pow(0.5, 2) */
static const double kOcc = 1.0;

/* gppKer.cpp:287-293 with the fixes listed above; -1 = skip the row. */
int gpp_oracle_resolve_igp(int my_igp, int ncouls, const int *inv_igp_index,
                           const int *indinv) {
  const int indigp = inv_igp_index[my_igp];
  int igp;
  if (indigp == ncouls)
    igp = ncouls - 1;
  else if (indigp < 0 || indigp > ncouls)
    return -1;
  else
    igp = indinv[indigp];
  if (igp < 0 || igp >= ncouls) return -1;
  return igp;
}

/* ssxt_scht_solver, gppKer.cpp:16-62 (sem 0/1) or the occupied branch of
 * the f90 :290-329 (sem 2).  ieps_x is the I_eps value used for sch and the
 * cutoff; Omega2 was formed from the correctly indexed I_eps. */
static void occ_solver(int sem, double wxt, c128 wtilde, c128 wtilde2,
                       c128 Omega2, c128 matngmatmgp, c128 ieps_x, c128 *ssxa,
                       c128 *scha) {
  const c128 zero = cmk(0.0, 0.0);
  c128 sch, ssx;
  const c128 wdiff = cmk(wxt - wtilde.re, -wtilde.im);
  c128 cden = wdiff;
  double rden = 1.0 / cnorm2(cden);
  const c128 delw = cscale(cmul(wtilde, cconj(cden)), rden);
  const double delwr = cnorm2(delw);
  const double wdiffr = cnorm2(wdiff);
  if (wdiffr > kLimitTwo && delwr < kLimitOne) {
    sch = cmul(delw, ieps_x);
    if (sem == GPP_ORACLE_SEM_BGW)
      cden = cmk(wxt * wxt - wtilde2.re, -wtilde2.im);
    else
      cden = cmk(wxt * wxt, 0.0);
    rden = cnorm2(cden);
    rden = 1.0 / rden;
    ssx = cscale(cmul(Omega2, cconj(cden)), rden);
  } else if (delwr > kTo1) {
    sch = zero;
    cden = cmul(cscale(wtilde2, 4.0), cmk(delw.re + 0.5, delw.im));
    rden = cnorm2(cden);
    rden = 1.0 / rden;
    ssx = cmul(cscale(cmul(cmk(-Omega2.re, -Omega2.im), cconj(cden)), rden),
               delw);
  } else {
    sch = zero;
    ssx = zero;
  }
  if (sem == GPP_ORACLE_SEM_BGW) {
    const double ssxcutoff = kSexcut * cabs_(ieps_x);
    if (cabs_(ssx) > ssxcutoff && wxt < 0.0) ssx = zero;
  }
  /* gppKer.cpp:58 tests abs(wxt) < 0, which never holds. */
  *ssxa = cmul(matngmatmgp, ssx);
  *scha = cmul(matngmatmgp, sch);
}

/* One (n1, my_igp) row: adds sch_array[iw] and ssx_array[iw] (gppKer.cpp
 * :295-348) to row[0..5] (sch re/im) and row[6..11] (ssx re/im). */
static void sigma_row(const gpp_oracle_problem *p, int sem, int flags, int n1,
                      int my_igp, int igp, double *row) {
  const int NC = p->ncouls, NG = p->ngpown;
  const int NV = p->nvband < p->number_bands ? p->nvband : p->number_bands;
  const int occ = !(flags & GPP_ORACLE_ALL_BANDS_UNOCC) && n1 < NV;
  const c128 *aqsm = (const c128 *)p->aqsmtemp;
  const c128 *aqsn = (const c128 *)p->aqsntemp;
  const c128 *ieps = (const c128 *)p->I_eps_array;
  const c128 *wt = (const c128 *)p->wtilde_array;
  const double *wx = p->wx_array + (p->wx_per_band ? 3 * n1 : 0);
  memset(row, 0, 12 * sizeof(double));
  if (occ) {
    const c128 mygpvar1 = cconj(aqsm[(size_t)n1 * NC + igp]);
    for (int iw = 0; iw < 3; ++iw) {
      double wxt = wx[iw];
      if (fabs(wxt) < kTo1) wxt = kTo1; /* gppKer.cpp:279 */
      ksum ss = {0, 0, 0, 0}, sc = {0, 0, 0, 0};
      for (int ig = 0; ig < NC; ++ig) {
        const c128 wtilde = wt[(size_t)my_igp * NC + ig];
        const c128 wtilde2 = cmul(wtilde, wtilde);
        const c128 Omega2 = cmul(wtilde2, ieps[(size_t)my_igp * NC + ig]);
        const c128 mat = cmul(aqsn[(size_t)n1 * NC + ig], mygpvar1);
        const c128 ieps_x = (sem == GPP_ORACLE_SEM_GPPKER_STRICT)
                                ? ieps[(size_t)my_igp * NG + ig]
                                : ieps[(size_t)my_igp * NC + ig];
        c128 a, b;
        occ_solver(sem, wxt, wtilde, wtilde2, Omega2, mat, ieps_x, &a, &b);
        kadd(&ss, a);
        kadd(&sc, b);
      }
      const c128 ssxt = kval(&ss), scht = kval(&sc);
      row[6 + 2 * iw] = ssxt.re;
      row[6 + 2 * iw + 1] = ssxt.im;
      row[2 * iw] = 0.5 * scht.re;
      row[2 * iw + 1] = 0.5 * scht.im;
    }
  } else {
    if (flags & GPP_ORACLE_SKIP_UNOCC) return;
    const c128 mygpvar1 = cconj(aqsm[(size_t)n1 * NC + igp]);
    for (int iw = 0; iw < 3; ++iw) {
      double wxt = wx[iw];
      if (fabs(wxt) < kTo1) wxt = kTo1;
      ksum sc = {0, 0, 0, 0};
      for (int ig = 0; ig < NC; ++ig) {
        const c128 wtilde = wt[(size_t)my_igp * NC + ig];
        const c128 wdiff = cmk(wxt - wtilde.re, -wtilde.im);
        const c128 cden = wdiff;
        double rden = cnorm2(cden);
        rden = 1.0 / rden;
        const c128 delw = cscale(cmul(wtilde, cconj(cden)), rden);
        const c128 scha =
            cmul(cmul(cmul(mygpvar1, aqsn[(size_t)n1 * NC + ig]), delw),
                 ieps[(size_t)my_igp * NC + ig]);
        if (sem == GPP_ORACLE_SEM_BGW) {
          const double delwr = cnorm2(delw);
          const double wdiffr = cnorm2(wdiff);
          if (!(wdiffr > kLimitTwo && delwr < kLimitOne)) continue;
        }
        kadd(&sc, scha);
      }
      const c128 scht = kval(&sc);
      row[2 * iw] = 0.5 * scht.re;
      row[2 * iw + 1] = 0.5 * scht.im;
    }
  }
}

/* reduce_achstemp contribution of one (n1, my_igp) row. */
static c128 achstemp_row(const gpp_oracle_problem *p, int sem, int n1,
                         int my_igp, int igp) {
  const int NC = p->ncouls;
  const c128 *aqsm = (const c128 *)p->aqsmtemp;
  const c128 *aqsn = (const c128 *)p->aqsntemp;
  const c128 *ieps = (const c128 *)p->I_eps_array;
  const c128 mygpvar1 = cconj(aqsm[(size_t)n1 * NC + igp]);
  c128 schstemp = cmk(0.0, 0.0);
  if (sem == GPP_ORACLE_SEM_BGW) {
    /* f90:214-224: schstemp -= aqsn(ig) * I_eps(ig,my_igp) * mygpvar1 */
    ksum k = {0, 0, 0, 0};
    for (int ig = 0; ig < NC; ++ig) {
      const c128 t = cmul(cmul(aqsn[(size_t)n1 * NC + ig],
                               ieps[(size_t)my_igp * NC + ig]),
                          mygpvar1);
      kadd(&k, cmk(-t.re, -t.im));
    }
    schstemp = kval(&k);
  } else {
    /* gppKer.cpp:93-97 */
    const c128 schs = ieps[(size_t)my_igp * NC + igp];
    const c128 mat = cmul(aqsn[(size_t)n1 * NC + igp], mygpvar1);
    if (cabs_(schs) > kTo1) schstemp = cadd(schstemp, cmul(mat, schs));
  }
  return cscale(cscale(schstemp, p->vcoul[igp]), 0.5);
}

int gpp_oracle_sigma(const gpp_oracle_problem *p, int sem, int flags,
                     double *out, double *achstemp, double *acht_n1_loc) {
  if (!p || !out) return -1;
  const int NB = p->number_bands, NC = p->ncouls, NG = p->ngpown;
  if (NB < 0 || NC <= 0 || NG < 0 || p->nvband < 0) return -1;
  if (sem < 0 || sem > 2) return -1;
  if (sem == GPP_ORACLE_SEM_GPPKER_STRICT && NG > NC) return -1;

  int *igp_of = (int *)malloc((size_t)(NG > 0 ? NG : 1) * sizeof(int));
  double *rows = (double *)malloc((size_t)(NG > 0 ? NG : 1) * 14 * sizeof(double));
  if (!igp_of || !rows) {
    free(igp_of);
    free(rows);
    return -2;
  }
  for (int i = 0; i < NG; ++i)
    igp_of[i] = gpp_oracle_resolve_igp(i, NC, p->inv_igp_index, p->indinv);

  dd acc[14];
  memset(acc, 0, sizeof acc);
  if (acht_n1_loc) memset(acht_n1_loc, 0, (size_t)2 * NB * sizeof(double));
  const double sx_sign = (sem == GPP_ORACLE_SEM_BGW) ? -1.0 : 1.0;

  for (int n1 = 0; n1 < NB; ++n1) {
#pragma omp parallel for schedule(dynamic)
    for (int my_igp = 0; my_igp < NG; ++my_igp) {
      double *row = rows + (size_t)my_igp * 14;
      const int igp = igp_of[my_igp];
      if (igp < 0) {
        memset(row, 0, 14 * sizeof(double));
        continue;
      }
      sigma_row(p, sem, flags, n1, my_igp, igp, row);
      const double vc = p->vcoul[igp];
      for (int i = 0; i < 6; ++i) row[i] *= vc;               /* :355 */
      for (int i = 6; i < 12; ++i) row[i] *= sx_sign * kOcc * vc; /* :352 */
      if (achstemp) {
        const c128 a = achstemp_row(p, sem, n1, my_igp, igp);
        row[12] = a.re;
        row[13] = a.im;
      } else {
        row[12] = row[13] = 0.0;
      }
    }
    /* Fixed-order accumulation over my_igp. */
    dd band_re, band_im;
    band_re.hi = band_re.lo = band_im.hi = band_im.lo = 0.0;
    for (int my_igp = 0; my_igp < NG; ++my_igp) {
      const double *row = rows + (size_t)my_igp * 14;
      for (int i = 0; i < 14; ++i) dd_add(&acc[i], row[i]);
      dd_add(&band_re, row[4]); /* sch_array[2]*vcoul, :357 */
      dd_add(&band_im, row[5]);
    }
    if (acht_n1_loc) {
      acht_n1_loc[2 * n1] = band_re.hi + band_re.lo;
      acht_n1_loc[2 * n1 + 1] = band_im.hi + band_im.lo;
    }
  }
  for (int i = 0; i < 12; ++i) out[i] = acc[i].hi + acc[i].lo;
  if (achstemp) {
    achstemp[0] = acc[12].hi + acc[12].lo;
    achstemp[1] = acc[13].hi + acc[13].lo;
  }
  free(igp_of);
  free(rows);
  return 0;
}

/* Per-row entry point used by the known-answer tests: the occupied-solver
 * sums ssxt/scht of gppKer.cpp:303-310 for one (n1, my_igp, wxt), before the
 * 0.5 and vcoul weights. */
int gpp_oracle_occ_sums(const gpp_oracle_problem *p, int sem, int n1,
                        int my_igp, int igp, double wxt, double *ssxt,
                        double *scht) {
  if (!p || !ssxt || !scht) return -1;
  const int NC = p->ncouls, NG = p->ngpown;
  const c128 *aqsm = (const c128 *)p->aqsmtemp;
  const c128 *aqsn = (const c128 *)p->aqsntemp;
  const c128 *ieps = (const c128 *)p->I_eps_array;
  const c128 *wt = (const c128 *)p->wtilde_array;
  const c128 mygpvar1 = cconj(aqsm[(size_t)n1 * NC + igp]);
  ksum ss = {0, 0, 0, 0}, sc = {0, 0, 0, 0};
  for (int ig = 0; ig < NC; ++ig) {
    const c128 wtilde = wt[(size_t)my_igp * NC + ig];
    const c128 wtilde2 = cmul(wtilde, wtilde);
    const c128 Omega2 = cmul(wtilde2, ieps[(size_t)my_igp * NC + ig]);
    const c128 mat = cmul(aqsn[(size_t)n1 * NC + ig], mygpvar1);
    const c128 ieps_x = (sem == GPP_ORACLE_SEM_GPPKER_STRICT)
                            ? ieps[(size_t)my_igp * NG + ig]
                            : ieps[(size_t)my_igp * NC + ig];
    c128 a, b;
    occ_solver(sem, wxt, wtilde, wtilde2, Omega2, mat, ieps_x, &a, &b);
    kadd(&ss, a);
    kadd(&sc, b);
  }
  const c128 x = kval(&ss), y = kval(&sc);
  ssxt[0] = x.re;
  ssxt[1] = x.im;
  scht[0] = y.re;
  scht[1] = y.im;
  return 0;
}
