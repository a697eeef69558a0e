/*
 * Eigenvalue extraction from a plain (unreorthogonalised) Lanczos tridiagonal.
 *
 * The reference Lanczos has no reorthogonalisation (src/Lanczos_EigenValue.c
 * :153-286), so once a Ritz value converges its copies ("ghosts") reappear and
 * not-yet-converged copies appear as spurious values.  To report the lowest
 * n distinct eigenvalues this file applies the Cullum-Willoughby test
 * (J. Cullum, R. Willoughby, Lanczos Algorithms for Large Symmetric
 * Eigenvalue Computations, 1985): an eigenvalue of T_m that is simple in T_m
 * and is also an eigenvalue of T_m with its first row and column removed is
 * spurious.  The remaining values are merged when they agree to `tol`.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "hphi_amd.h"

#define CW_EXTRA 6

/* lowest `want` good distinct eigenvalues of T(alpha[0..m), beta[0..m-1)); returns count */
int hphi_lanczos_good_eigenvalues(const double *alpha, const double *beta, int m, int want, double *out) {
  if (m <= 0 || want <= 0) return 0;
  if (m == 1) {
    out[0] = alpha[0];
    return 1;
  }
  double scale = 0.0;
  for (int i = 0; i < m; i++) {
    double r = fabs(alpha[i]) + (i < m - 1 ? fabs(beta[i]) : 0.0) + (i > 0 ? fabs(beta[i - 1]) : 0.0);
    if (r > scale) scale = r;
  }
  const double tol = 1e-11 * (scale > 1.0 ? scale : 1.0);
  /* Without reorthogonalisation every converged eigenvalue keeps gaining ghost
   * copies, so the lowest want + CW_EXTRA eigenvalues of T_m fill up with copies
   * as m grows: the window doubles until it holds `want` good values or all of T_m. */
  int L = want + CW_EXTRA, got = 0;
  if (L > m) L = m;
  double *ev = (double *)malloc(sizeof(double) * (size_t)m);
  double *evh = (double *)malloc(sizeof(double) * (size_t)m);
  for (;;) {
    int n = hphi_tridiag_lowest(alpha, beta, m, L, ev);
    int Lh = L < m - 1 ? L : m - 1;
    int nh = hphi_tridiag_lowest(alpha + 1, beta + 1, m - 1, Lh, evh);
    got = 0;
    for (int i = 0; i < n && got < want; i++) {
      int mult = 0;
      for (int j = 0; j < n; j++)
        if (fabs(ev[j] - ev[i]) <= tol) mult++;
      int spurious = 0;
      if (mult == 1 && i < n - 1) {  /* the highest computed value may have an uncounted twin */
        for (int j = 0; j < nh; j++)
          if (fabs(evh[j] - ev[i]) <= tol) spurious = 1;
      }
      if (spurious) continue;
      if (got > 0 && fabs(ev[i] - out[got - 1]) <= tol) continue; /* ghost copy of a converged value */
      out[got++] = ev[i];
    }
    if (got >= want || L >= m) break;
    L = 2 * L < m ? 2 * L : m;
  }
  free(ev);
  free(evh);
  return got;
}

/* convergence of the lowest `want` good eigenvalues between two checks */
int hphi_lanczos_converged(const double *prev, int nprev, const double *cur, int ncur, int want, double tol) {
  if (ncur < want || nprev < want) return 0;
  for (int i = 0; i < want; i++) {
    double s = fabs(cur[i]) > 1.0 ? fabs(cur[i]) : 1.0;
    if (fabs(cur[i] - prev[i]) > tol * s) return 0;
  }
  return 1;
}
