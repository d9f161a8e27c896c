/* The CPU oracle under AddressSanitizer / UndefinedBehaviorSanitizer (and
 * ThreadSanitizer): gpp_oracle_sigma on small inputs with the edge cases the
 * tests use (skipped G' rows, indigp == ncouls, an exact resonance, nvband 0
 * and NB, band-resolved frequencies, all semantics and band-range flags,
 * the per-band log), and ff_oracle in both semantics.  Exit status 0 and no
 * sanitizer report is the pass condition. */
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../oracle/ff_oracle.h"
#include "../../oracle/gpp_oracle.h"

static double uni(unsigned *s) {
  *s = *s * 1664525u + 1013904223u;
  return (*s >> 8) * (1.0 / 16777216.0);
}

int main(void) {
  const int shapes[][4] = {{6, 2, 48, 4}, {5, 0, 40, 8}, {3, 3, 30, 3}, {9, 4, 64, 5}};
  unsigned seed = 12345;
  for (size_t si = 0; si < sizeof shapes / sizeof shapes[0]; ++si) {
    const int nb = shapes[si][0], nv = shapes[si][1], nc = shapes[si][2];
    const int ng = nc / shapes[si][3];
    double *aqsm = malloc(sizeof(double) * 2 * nb * nc);
    double *aqsn = malloc(sizeof(double) * 2 * nb * nc);
    double *ieps = malloc(sizeof(double) * 2 * ng * nc);
    double *wt = malloc(sizeof(double) * 2 * ng * nc);
    double *vc = malloc(sizeof(double) * nc);
    int *inv = malloc(sizeof(int) * ng);
    int *ind = malloc(sizeof(int) * nc);
    double *wx = malloc(sizeof(double) * 3 * nb);
    for (int i = 0; i < 2 * nb * nc; ++i) {
      aqsm[i] = uni(&seed) - 0.5;
      aqsn[i] = uni(&seed) - 0.5;
    }
    for (int i = 0; i < ng * nc; ++i) {
      ieps[2 * i] = uni(&seed);
      ieps[2 * i + 1] = 0.2 * uni(&seed) - 0.1;
      wt[2 * i] = 0.5 + 2.5 * uni(&seed);
      wt[2 * i + 1] = 0.1 * uni(&seed) - 0.05;
    }
    for (int i = 0; i < nc; ++i) {
      vc[i] = 1.0 / (0.01 + pow((double)i / nc, 2.0 / 3.0));
      ind[i] = i;
    }
    for (int k = 0; k < ng; ++k) inv[k] = k * shapes[si][3];
    inv[0] = nc;              /* indigp == ncouls -> igp = ncouls - 1 */
    if (ng > 2) inv[1] = -3;  /* skipped row */
    if (ng > 3) ind[inv[3]] = nc + 5; /* igp out of range -> skipped */
    wt[2 * (1 * nc + 5)] = 4.0;       /* exact resonance at wx = 4 */
    wt[2 * (1 * nc + 5) + 1] = 0.0;
    for (int n = 0; n < nb; ++n)
      for (int iw = 0; iw < 3; ++iw) wx[3 * n + iw] = 3.0 + iw - 0.4 * n;
    for (int per_band = 0; per_band < 2; ++per_band) {
      gpp_oracle_problem p = {nb, nv, nc, ng, aqsm, aqsn, ieps, wt, vc, inv,
                              ind, wx, per_band};
      for (int sem = 0; sem < 3; ++sem)
        for (int flags = 0; flags < 3; ++flags) {
          double out[12], ach[2], *n1 = malloc(sizeof(double) * 2 * nb);
          if (gpp_oracle_sigma(&p, sem, flags, out, ach, n1) != 0) {
            fprintf(stderr, "oracle failed: shape %zu sem %d\n", si, sem);
            return 1;
          }
          free(n1);
        }
    }
    free(aqsm); free(aqsn); free(ieps); free(wt); free(vc); free(inv);
    free(ind); free(wx);
  }
  /* full-frequency oracle, both semantics */
  {
    const int nb = 5, nv = 2, nc = 40, ng = 8, nf = 7, ne = 9;
    ffo_c *aqsm = malloc(sizeof(ffo_c) * nb * nc), *aqsn = malloc(sizeof(ffo_c) * nb * nc);
    ffo_c *er = malloc(sizeof(ffo_c) * nf * ng * nc), *ea = malloc(sizeof(ffo_c) * nf * ng * nc);
    ffo_c *brd = malloc(sizeof(ffo_c) * nf), *sch = malloc(sizeof(ffo_c) * nb * nf);
    double *vc = malloc(sizeof(double) * nc), *ekq = malloc(sizeof(double) * nb);
    double *grid = malloc(sizeof(double) * nf), *pref = malloc(sizeof(double) * nf);
    int *inv = malloc(sizeof(int) * ng), *ind = malloc(sizeof(int) * nc);
    for (int i = 0; i < nb * nc; ++i) {
      aqsm[i].re = uni(&seed) - 0.5; aqsm[i].im = uni(&seed) - 0.5;
      aqsn[i].re = uni(&seed) - 0.5; aqsn[i].im = uni(&seed) - 0.5;
    }
    for (int i = 0; i < nf * ng * nc; ++i) {
      er[i].re = uni(&seed); er[i].im = uni(&seed) - 0.5;
      ea[i].re = uni(&seed); ea[i].im = uni(&seed) - 0.5;
    }
    for (int i = 0; i < nc; ++i) { vc[i] = 1.0 + i; ind[i] = i; }
    for (int k = 0; k < ng; ++k) inv[k] = 5 * k;
    for (int n = 0; n < nb; ++n) ekq[n] = -6.0 + 3.0 * n;
    for (int f = 0; f < nf; ++f) {
      grid[f] = 2.0 * f; pref[f] = 2.0;
      brd[f].re = 0.0; brd[f].im = -0.1;
    }
    for (int i = 0; i < nb * nf; ++i) { sch[i].re = uni(&seed); sch[i].im = 0.0; }
    ffo_problem p = {nb, nv, nc, ng, nf, ne, aqsm, aqsn, er, ea, vc, inv, ind,
                     ekq, grid, pref, brd, sch, -3.0, 1.5, 0.5, 1.0};
    for (int sem = 0; sem < 2; ++sem) {
      ffo_c achs, asx[9], ach[9], cor[9], corb[9], ach2[9];
      if (ff_oracle(sem, &p, &achs, asx, ach, cor, corb, ach2) != 0) {
        fprintf(stderr, "ff oracle failed: sem %d\n", sem);
        return 1;
      }
    }
    free(aqsm); free(aqsn); free(er); free(ea); free(brd); free(sch);
    free(vc); free(ekq); free(grid); free(pref); free(inv); free(ind);
  }
  printf("oracle sanitizer run ok\n");
  return 0;
}
