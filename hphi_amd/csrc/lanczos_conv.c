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

int hphi_set_error(const char *fmt, ...);

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

/* Eigenvector of the symmetric tridiagonal (a[0..n), b[0..n-1)) for the
 * eigenvalue lam: inverse iteration, two solves of (T - lam) x = y from y = 1,
 * with an LU factorisation by Gaussian elimination with partial pivoting (row
 * i+1 may swap with the running row i).  Normalised, first non-zero entry > 0. */
void hphi_tridiag_eigvec(const double *a, const double *b, int n, double lam, double *x) {
  if (n == 1) {
    x[0] = 1.0;
    return;
  }
  double *U = (double *)calloc(3 * (size_t)n, sizeof(double)), *L = (double *)calloc(n, sizeof(double));
  int *sw = (int *)calloc(n, sizeof(int));
  double *y = (double *)malloc(sizeof(double) * n);
  double nrm_t = 0.0;
  for (int i = 0; i < n; i++) nrm_t = fmax(nrm_t, fabs(a[i]) + (i ? fabs(b[i - 1]) : 0) + (i < n - 1 ? fabs(b[i]) : 0));
  const double tiny = 1e-15 * (nrm_t > 0 ? nrm_t : 1.0);
  /* running row (columns i, i+1, i+2) */
  double c = a[0] - lam, d = b[0], f = 0.0;
  for (int i = 0; i < n - 1; i++) {
    const double r0 = b[i], r1 = a[i + 1] - lam, r2 = (i + 2 < n) ? b[i + 1] : 0.0;
    if (fabs(r0) > fabs(c)) { /* row i+1 is the pivot */
      const double m = c / r0;
      U[3 * i] = r0, U[3 * i + 1] = r1, U[3 * i + 2] = r2;
      L[i] = m, sw[i] = 1;
      c = d - m * r1, d = f - m * r2, f = 0.0;
    } else {
      if (fabs(c) < tiny) c = tiny;
      const double m = r0 / c;
      U[3 * i] = c, U[3 * i + 1] = d, U[3 * i + 2] = f;
      L[i] = m, sw[i] = 0;
      c = r1 - m * d, d = r2 - m * f, f = 0.0;
    }
  }
  if (fabs(c) < tiny) c = tiny;
  U[3 * (n - 1)] = c;
  for (int i = 0; i < n; i++) y[i] = 1.0;
  for (int it = 0; it < 2; it++) {
    /* forward: the row operations of the factorisation on y */
    double run = y[0];
    for (int i = 0; i < n - 1; i++) {
      if (sw[i]) {
        const double piv = y[i + 1];
        y[i] = piv;
        run = run - L[i] * piv;
      } else {
        y[i] = run;
        run = y[i + 1] - L[i] * run;
      }
    }
    y[n - 1] = run;
    /* back substitution */
    for (int i = n - 1; i >= 0; i--) {
      double t = y[i];
      if (i + 1 < n) t -= U[3 * i + 1] * x[i + 1];
      if (i + 2 < n) t -= U[3 * i + 2] * x[i + 2];
      x[i] = t / U[3 * i];
    }
    double s = 0.0;
    for (int i = 0; i < n; i++) s += x[i] * x[i];
    s = 1.0 / sqrt(s);
    for (int i = 0; i < n; i++) y[i] = x[i] *= s;
  }
  int k0 = 0;
  while (k0 < n - 1 && x[k0] == 0.0) k0++;
  if (x[k0] < 0)
    for (int i = 0; i < n; i++) x[i] = -x[i];
  free(U);
  free(L);
  free(sw);
  free(y);
}

/* Residual bound of a Ritz value of the Lanczos tridiagonal: for the eigenvalue
 * theta of T_m (alpha[0..m), beta[0..m-1)) with unit eigenvector y, the Ritz
 * vector z = V_m y has |H z - theta z| = beta_m |y_m| (and, without
 * reorthogonalisation, up to rounding of the order of |H| times machine
 * precision: C. C. Paige, J. Inst. Math. Appl. 18 (1976) 373-381).  y comes
 * from inverse iteration at theta. */
int hphi_lanczos_ritz_residual(const double *alpha, const double *beta, int m, double beta_m, double theta,
                               double *res) {
  if (!alpha || !res || m < 1 || (m > 1 && !beta)) return hphi_set_error("hphi_lanczos_ritz_residual: bad argument");
  double *y = (double *)malloc(sizeof(double) * (size_t)m);
  if (!y) return hphi_set_error("hphi_lanczos_ritz_residual: out of memory");
  hphi_tridiag_eigvec(alpha, beta, m, theta, y);
  *res = fabs(beta_m) * fabs(y[m - 1]);
  free(y);
  return 0;
}

/* the Ritz residual test of the distinct mode: every one of the lowest `want`
 * values has beta_m |y_m| <= 0.1 sqrt(tol) max(1, |theta|) (with tol = 1e-12:
 * 1e-7, so an eigenvalue error of about residual^2 / gap, 1e-14 / gap, is left) */
int hphi_lanczos_residuals_small(const double *alpha, const double *beta, int m, double beta_m, const double *cur,
                                 int want, double tol) {
  const double rt = 0.1 * sqrt(tol);
  for (int i = 0; i < want; i++) {
    double r = 0.0;
    if (hphi_lanczos_ritz_residual(alpha, beta, m, beta_m, cur[i], &r)) return 0;
    if (r > rt * (fabs(cur[i]) > 1.0 ? fabs(cur[i]) : 1.0)) return 0;
  }
  return 1;
}
