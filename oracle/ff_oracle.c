/* ff_oracle.c — CPU restatement of the FFKer_cori full-frequency kernels.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/ load it (the benchmark checks the
 * FF path against closed forms); the product library never does.  Plain loops in the reference's order of
 * evaluation (per-element interpolation, then the sums), compensated
 * (Neumaier) accumulation over ig and my_igp, no FMA contraction
 * (-ffp-contract=off), OpenMP over my_igp.
 *
 * FF_ORC_CPP follows FFKer_cori/ffKernel.cpp:
 *   achsDtemp_Kernel :107  (I_epsR slice 1)
 *   asxDtemp_Kernel  :133  (wx = freqevalmin - ekq[n1] + freqevalstep, the
 *                           same for every iw; I_epsR only; compute_fact
 *                           loop 1 with flag_occ = 0)
 *   achDtemp_Kernel  :155  (its output term schDi, :200-206, summed over
 *                           n1 and ifreq without the reference's data race
 *                           on schDi/achDtemp; the sch2Di/schDi_cor updates
 *                           at :186-196 go to arrays that achDtemp_cor_Kernel
 *                           resets before reading, so they are not outputs)
 *   achDtemp_cor_Kernel :216 with schDttt_corKernel1/2 :250,279
 *   compute_fact     :28   (strict grid search; quirks kept)
 * ach2Dtemp here is the sum of schDttt over the bands where
 * schDttt_corKernel1 runs (the reference adds a stale sch2Di[iw] for the
 * other bands, :245).
 *
 * FF_ORC_F90 follows FFKer_cori/ffKernel.f90:
 *   achsDtemp :205-219 (I_epsR frequency 1, i.e. slice 0)
 *   asxDtemp  :225-313 (wx = freqevalmin - ekq + (iw-1)*freqevalstep;
 *                       I_epsR for wx >= 0, I_epsA below)
 *   frequency loop :335-384 (schDi, schDi_corb, sch2Di, schDi_cor)
 *   cor sums  :391-461
 * Rows whose igp falls outside [0, ncouls) are skipped in both (the F90
 * cycles them, :209; the C++ reads them unchecked).
 */
#include "ff_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  double re, im, cre, cim;
} acc_t; /* Neumaier-compensated complex sum */

static void acc_add(acc_t *a, double re, double im) {
  double t = a->re + re;
  a->cre += (fabs(a->re) >= fabs(re)) ? ((a->re - t) + re) : ((re - t) + a->re);
  a->re = t;
  t = a->im + im;
  a->cim += (fabs(a->im) >= fabs(im)) ? ((a->im - t) + im) : ((im - t) + a->im);
  a->im = t;
}
static ffo_c acc_get(const acc_t *a) {
  ffo_c r = {a->re + a->cre, a->im + a->cim};
  return r;
}
static ffo_c cmul(ffo_c a, ffo_c b) {
  ffo_c r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
  return r;
}
static ffo_c cconj(ffo_c a) {
  ffo_c r = {a.re, -a.im};
  return r;
}
static ffo_c cscale(ffo_c a, double s) {
  ffo_c r = {a.re * s, a.im * s};
  return r;
}
static ffo_c cadd(ffo_c a, ffo_c b) {
  ffo_c r = {a.re + b.re, a.im + b.im};
  return r;
}
static ffo_c csub(ffo_c a, ffo_c b) {
  ffo_c r = {a.re - b.re, a.im - b.im};
  return r;
}
/* a / b as Complex.h:208 forms it: a*conj(b) / (b*conj(b)).re */
static ffo_c cdiv(ffo_c a, ffo_c b) {
  ffo_c n = cmul(a, cconj(b));
  double d = cmul(b, cconj(b)).re;
  ffo_c r = {n.re / d, n.im / d};
  return r;
}

static int resolve_igp(const ffo_problem *p, int my_igp) {
  int indigp = p->inv_igp_index[my_igp];
  if (indigp < 0 || indigp >= p->ncouls) return -1;
  int igp = p->indinv[indigp];
  return (igp < 0 || igp >= p->ncouls) ? -1 : igp;
}

/* Grid search of the C++ compute_fact (ffKernel.cpp:28-78): the last ijk in
 * [0, nFreq-2] with grid[ijk] < x < grid[ijk+1], else 0. */
static int search_strict(double x, int nFreq, const double *g) {
  int f = 0;
  for (int k = 0; k < nFreq - 1; ++k)
    if (x > g[k] && x < g[k + 1]) f = k;
  return f;
}
/* Grid search of the F90 (ffKernel.f90:231-245): the last ijk with
 * grid(ijk) <= x < grid(ijk+1), 0-based, else nFreq-2. */
static int search_f90(double x, int nFreq, const double *g) {
  int f = -1;
  for (int k = 0; k < nFreq - 1; ++k)
    if (x >= g[k] && x < g[k + 1]) f = k;
  return f < 0 ? nFreq - 2 : f;
}

int ff_oracle_fact(int sem, int loop, int flag_occ, double wx, int nFreq,
                   const double *g, double *fact1, double *fact2) {
  int f;
  double d;
  *fact1 = *fact2 = 0.0;
  if (sem == FF_ORC_CPP) {
    if (loop == 1) {
      if (wx > 0.0) {
        f = search_strict(wx, nFreq, g);
        if (f == 0) f = nFreq - 2;
        d = g[f + 1] - g[f];
        *fact1 = (g[f + 1] - wx) / d;
        *fact2 = (wx - g[f]) / d;
      } else {
        f = search_strict(-wx, nFreq, g);
        if (f == 0) f = nFreq - 2;
        d = g[f + 1] - g[f];
        *fact1 = (g[f + 1] + wx) / d;
        *fact2 = (-g[f] - wx) / d;
      }
      return f;
    }
    if (wx > 0.0) {
      if (flag_occ) return -1; /* achDtemp_cor_Kernel calls nothing */
      f = search_strict(wx, nFreq, g); /* no fallback: :55 tests -1 */
      d = g[f + 1] - g[f];
      *fact1 = -0.5 * (g[f + 1] - wx) / d;
      *fact2 = -0.5 * (wx - g[f]) / d;
      return f;
    }
    if (!flag_occ) return -1;
    wx = -wx;
    f = search_strict(wx, nFreq, g);
    if (f == 0) f = nFreq - 2;
    d = g[f + 1] - g[f];
    *fact1 = (g[f + 1] - wx) / d;
    *fact2 = (wx - g[f]) / d;
    return f;
  }
  /* F90 */
  if (loop == 1) {
    if (!flag_occ) return -1;
    if (wx >= 0.0) {
      f = search_f90(wx, nFreq, g);
      d = g[f + 1] - g[f];
      *fact1 = (g[f + 1] - wx) / d;
      *fact2 = (wx - g[f]) / d;
    } else {
      f = search_f90(-wx, nFreq, g);
      d = g[f + 1] - g[f];
      *fact1 = (g[f + 1] + wx) / d;
      *fact2 = (-g[f] - wx) / d;
    }
    return f;
  }
  if (wx >= 0.0) {
    f = search_f90(wx, nFreq, g);
    d = g[f + 1] - g[f];
    *fact1 = -0.5 * (g[f + 1] - wx) / d;
    *fact2 = -0.5 * (wx - g[f]) / d;
    return f;
  }
  if (!flag_occ) return -1;
  wx = -wx;
  f = search_f90(wx, nFreq, g);
  d = g[f + 1] - g[f];
  *fact1 = (g[f + 1] - wx) / d;
  *fact2 = (wx - g[f]) / d;
  return f;
}

/* sum over valid my_igp, ig of
 *   aqsn[n1][ig] * conj(aqsm[n1][igp]) * X[my_igp][ig] * vcoul[igp]
 * with X = s*(Y1[f]*fact1 + Y1[f+1]*fact2) (Y2 == NULL) or
 *      X = s*((Y1[f]-Y2[f])*fact1 + (Y1[f+1]-Y2[f+1])*fact2).
 * The product is formed as the reference does: cpp_order groups
 * aqsn * (conj(aqsm) * X) (ffKernel.cpp:98); otherwise (aqsn * conj(aqsm))
 * * X (ffKernel.cpp:122,267; ffKernel.f90:269). */
static ffo_c slice_sum(const ffo_problem *p, int n1, const ffo_c *Y1,
                       const ffo_c *Y2, int f, double fact1, double fact2,
                       double s, int cpp_order) {
  const int NC = p->ncouls, NG = p->ngpown;
  const size_t sl = (size_t)NG * NC;
  acc_t tot = {0, 0, 0, 0};
#pragma omp parallel
  {
    acc_t mine = {0, 0, 0, 0};
#pragma omp for schedule(static)
    for (int my_igp = 0; my_igp < NG; ++my_igp) {
      const int igp = resolve_igp(p, my_igp);
      if (igp < 0) continue;
      const ffo_c am = cconj(p->aqsmtemp[(size_t)n1 * NC + igp]);
      acc_t row = {0, 0, 0, 0};
      for (int ig = 0; ig < NC; ++ig) {
        const size_t o = (size_t)my_igp * NC + ig;
        ffo_c y1 = Y1[(size_t)f * sl + o], y2 = Y1[(size_t)(f + 1) * sl + o];
        if (Y2) {
          y1 = csub(y1, Y2[(size_t)f * sl + o]);
          y2 = csub(y2, Y2[(size_t)(f + 1) * sl + o]);
        }
        ffo_c x = cadd(cscale(y1, fact1), cscale(y2, fact2));
        if (s != 1.0) x = cscale(x, s);
        const ffo_c an = p->aqsntemp[(size_t)n1 * NC + ig];
        const ffo_c t = cpp_order ? cmul(an, cmul(am, x)) : cmul(cmul(an, am), x);
        acc_add(&row, t.re, t.im);
      }
      const ffo_c r = cscale(acc_get(&row), p->vcoul[igp]);
      acc_add(&mine, r.re, r.im);
    }
#pragma omp critical
    acc_add(&tot, mine.re + mine.cre, mine.im + mine.cim);
  }
  return acc_get(&tot);
}

/* achsDtemp: sum over n1, valid my_igp of
 *   -sum_ig aqsn*conj(aqsm)*I_epsR[slice] * vcoul * 0.5 */
static ffo_c achs(const ffo_problem *p, int slice) {
  const int NC = p->ncouls, NG = p->ngpown;
  const size_t so = (size_t)slice * NG * NC;
  acc_t tot = {0, 0, 0, 0};
  for (int n1 = 0; n1 < p->number_bands; ++n1) {
#pragma omp parallel
    {
      acc_t mine = {0, 0, 0, 0};
#pragma omp for schedule(static)
      for (int my_igp = 0; my_igp < NG; ++my_igp) {
        const int igp = resolve_igp(p, my_igp);
        if (igp < 0) continue;
        const ffo_c am = cconj(p->aqsmtemp[(size_t)n1 * NC + igp]);
        acc_t row = {0, 0, 0, 0};
        for (int ig = 0; ig < NC; ++ig) {
          const ffo_c t = cmul(cmul(p->aqsntemp[(size_t)n1 * NC + ig], am),
                               p->I_epsR_array[so + (size_t)my_igp * NC + ig]);
          acc_add(&row, -t.re, -t.im);
        }
        const ffo_c r = cscale(acc_get(&row), p->vcoul[igp] * 0.5);
        acc_add(&mine, r.re, r.im);
      }
#pragma omp critical
      acc_add(&tot, mine.re + mine.cre, mine.im + mine.cim);
    }
  }
  return acc_get(&tot);
}

static double clampd(double x) {
  if (x < 1e-4) x = 1e-4;
  if (x > 1e4) x = 1e4;
  return x;
}

int ff_oracle(int sem, const ffo_problem *p, ffo_c *achsDtemp,
              ffo_c *asxDtemp, ffo_c *achDtemp, ffo_c *achDtemp_cor,
              ffo_c *achDtemp_corb, ffo_c *ach2Dtemp) {
  const int NB = p->number_bands, NF = p->nFreq, NE = p->nfreqeval;
  const int NV = p->nvband < NB ? p->nvband : NB;
  if (NB < 0 || NV < 0 || p->ncouls <= 0 || p->ngpown < 0 || NF < 2 ||
      NE < 0)
    return -1;
  const double *g = p->dFreqGrid;
  acc_t *asx = calloc(NE, sizeof(acc_t)), *ach = calloc(NE, sizeof(acc_t));
  acc_t *cor = calloc(NE, sizeof(acc_t)), *corb = calloc(NE, sizeof(acc_t));
  acc_t *a2 = calloc(NE, sizeof(acc_t));
  const ffo_c i_pz = {0.0, p->pref_zb};

  if (achsDtemp) *achsDtemp = achs(p, sem == FF_ORC_CPP ? 1 : 0);

  /* asxDtemp */
  for (int n1 = 0; n1 < NV; ++n1)
    for (int iw = 0; iw < NE; ++iw) {
      const double wx = sem == FF_ORC_CPP
                            ? p->freqevalmin - p->ekq[n1] + p->freqevalstep
                            : p->freqevalmin - p->ekq[n1] + iw * p->freqevalstep;
      double f1, f2;
      const int f = ff_oracle_fact(sem, 1, sem == FF_ORC_CPP ? 0 : 1, wx, NF,
                                   g, &f1, &f2);
      if (f < 0) continue;
      const ffo_c *Y = (sem == FF_ORC_F90 && wx < 0.0) ? p->I_epsA_array
                                                        : p->I_epsR_array;
      const ffo_c s = cscale(slice_sum(p, n1, Y, NULL, f, f1, f2, 1.0,
                                       sem == FF_ORC_CPP),
                             p->occ);
      acc_add(&asx[iw], s.re, s.im);
    }

  /* frequency-grid terms (no large arrays) */
  for (int n1 = 0; n1 < NB; ++n1) {
    const int occ_band = n1 < NV;
    for (int f = 0; f < NF; ++f) {
      const ffo_c schDt = p->schDt_matrix[(size_t)n1 * NF + f];
      const double zb = g[f];
      const ffo_c coh = {zb - p->dFreqBrd[f].re, -p->dFreqBrd[f].im};
      const ffo_c ipref = {0.0, p->pref[f]};
      if (sem == FF_ORC_CPP) {
        /* ffKernel.cpp:200-206: schDi = -(i pref * schDt) / (wx - coh) */
        for (int iw = 0; iw < NE; ++iw) {
          const double wx =
              p->freqevalmin - p->ekq[n1] + (iw - 1) * p->freqevalstep;
          const ffo_c den = {wx - coh.re, -coh.im};
          const ffo_c t = cdiv(cmul(ipref, schDt), den);
          acc_add(&ach[iw], -t.re, -t.im);
        }
        continue;
      }
      /* ffKernel.f90:361-409 */
      const ffo_c cor_c = {(occ_band ? -zb : zb) - p->dFreqBrd[f].re,
                           -p->dFreqBrd[f].im};
      if (f != NF - 1)
        for (int iw = 0; iw < NE; ++iw) {
          const double wx = p->freqevalmin - p->ekq[n1] + iw * p->freqevalstep;
          const ffo_c num = cmul(ipref, schDt);
          const ffo_c d1 = {wx - coh.re, -coh.im};
          const ffo_c d2 = {wx - cor_c.re, -cor_c.im};
          const ffo_c t1 = cdiv(num, d1), t2 = cdiv(num, d2);
          acc_add(&ach[iw], -t1.re, -t1.im);
          acc_add(&corb[iw], -t2.re, -t2.im);
        }
      if (f != 0) {
        const double right = zb, left = g[f - 1];
        const ffo_c sr = schDt, sl = p->schDt_matrix[(size_t)n1 * NF + f - 1];
        const ffo_c avg = cscale(cadd(sr, sl), 0.5);
        const ffo_c lin = csub(sr, sl);
        const ffo_c lin2 = cscale(lin, 1.0 / (right - left));
        for (int iw = 0; iw < NE; ++iw) {
          const double wv = p->freqevalmin - p->ekq[n1] + iw * p->freqevalstep;
          double it = -log(clampd(fabs((wv - right) / (wv - left))));
          const ffo_c t = cscale(cmul(i_pz, avg), it);
          acc_add(&a2[iw], -t.re, -t.im);
          ffo_c lin3;
          if (occ_band) {
            it = log(clampd(fabs((wv + right) / (wv + left))));
            const double a =
                -p->freqevalmin - p->ekq[n1] + iw * p->freqevalstep - left;
            lin3 = cscale(cadd(sl, cscale(lin2, a)), it);
          } else {
            lin3 = cscale(cadd(sl, cscale(lin2, wv - left)), it);
          }
          lin3 = cadd(lin3, lin);
          const ffo_c u = cmul(i_pz, lin3);
          acc_add(&cor[iw], -u.re, -u.im);
        }
      }
    }
  }

  /* cor sums */
  for (int n1 = 0; n1 < NB; ++n1) {
    const int occ_band = n1 < NV;
    for (int iw = 0; iw < NE; ++iw) {
      const double wx = sem == FF_ORC_CPP
                            ? p->freqevalmin - p->ekq[n1] + p->freqevalstep
                            : p->freqevalmin - p->ekq[n1] + iw * p->freqevalstep;
      double f1, f2;
      const int f = ff_oracle_fact(sem, 2, occ_band, wx, NF, g, &f1, &f2);
      if (f < 0) continue;
      const int positive = sem == FF_ORC_CPP ? (wx > 0.0) : (wx >= 0.0);
      const ffo_c s = slice_sum(p, n1, p->I_epsR_array, p->I_epsA_array, f,
                                f1, f2, positive ? 1.0 : -0.5, 0);
      if (positive) {
        acc_add(&a2[iw], s.re, s.im);
        if (!occ_band) acc_add(&cor[iw], s.re, s.im);
      } else {
        acc_add(&cor[iw], s.re, s.im);
      }
    }
  }

  for (int iw = 0; iw < NE; ++iw) {
    if (asxDtemp) asxDtemp[iw] = acc_get(&asx[iw]);
    if (achDtemp) achDtemp[iw] = acc_get(&ach[iw]);
    if (achDtemp_cor) achDtemp_cor[iw] = acc_get(&cor[iw]);
    if (achDtemp_corb) achDtemp_corb[iw] = acc_get(&corb[iw]);
    if (ach2Dtemp) ach2Dtemp[iw] = acc_get(&a2[iw]);
  }
  free(asx);
  free(ach);
  free(cor);
  free(corb);
  free(a2);
  return 0;
}
