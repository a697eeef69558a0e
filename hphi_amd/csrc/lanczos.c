/*
 * Device-resident Lanczos eigenvalue driver (host C over the hphi_amd C ABI).
 *
 * Recurrence of the reference (src/Lanczos_EigenValue.c:50-310):
 *   v1 = random, normalised; v0 = 0; v0 += H v1; alpha_1 = Re prdct;
 *   beta_1 = |v0 - alpha_1 v1|; then for stp = 2, 3, ...:
 *     v1 <- (v0 - alpha v1)/beta,  v0 <- -beta v1_old          (:154-160)
 *     v0 += H v1;  alpha = Re prdct;  beta = |v0 - alpha v1|   (:163-177)
 *     eigenvalues of the stp x stp tridiagonal (alpha, beta)   (:190-246)
 *     stop when |(E_target - E_target_prev)/E_target| < eps or beta < 1e-14 (:265)
 * No reorthogonalisation, as in the reference.  The tridiagonal eigenvalues
 * come from Sturm-sequence bisection (this file) instead of LAPACK dsyev.
 * Vectors stay in HBM; only alpha, beta cross PCIe.
 */
#define _POSIX_C_SOURCE 200809L
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "hphi_amd.h"

int hphi_set_error(const char *fmt, ...);

static double wall(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* number of eigenvalues of the symmetric tridiagonal (a[0..n), b[0..n-1)) below x */
static int sturm_count(const double *a, const double *b, int n, double x) {
  int c = 0;
  double q = a[0] - x;
  const double tiny = 1e-300;
  if (q < 0) c++;
  for (int i = 1; i < n; i++) {
    if (fabs(q) < tiny) q = -tiny;
    q = (a[i] - x) - b[i - 1] * b[i - 1] / q;
    if (q < 0) c++;
  }
  return c;
}

/* k lowest eigenvalues (ascending) of the tridiagonal by bisection */
int hphi_tridiag_lowest(const double *a, const double *b, int n, int k, double *ev) {
  if (n <= 0 || k <= 0) return 0;
  if (k > n) k = n;
  double lo = a[0], hi = a[0];
  for (int i = 0; i < n; i++) {
    double r = (i > 0 ? fabs(b[i - 1]) : 0.0) + (i < n - 1 ? fabs(b[i]) : 0.0);
    if (a[i] - r < lo) lo = a[i] - r;
    if (a[i] + r > hi) hi = a[i] + r;
  }
  double span = hi - lo;
  lo -= 1e-12 * span + 1e-300;
  hi += 1e-12 * span + 1e-300;
  for (int m = 0; m < k; m++) {
    double l = (m > 0) ? ev[m - 1] - 1e-15 * (fabs(ev[m - 1]) + span) : lo, h = hi;
    if (l < lo) l = lo;
    for (int it = 0; it < 200; it++) {
      double mid = 0.5 * (l + h);
      if (mid <= l || mid >= h) break;
      if (sturm_count(a, b, n, mid) > m) h = mid;
      else l = mid;
    }
    ev[m] = 0.5 * (l + h);
  }
  return k;
}

int hphi_lanczos(hphi_engine *e, const hphi_lanczos_params *p, hphi_lanczos_result *r) {
  if (!e || !p || !r) return hphi_set_error("hphi_lanczos: NULL argument");
  memset(r, 0, sizeof *r);
  uint64_t dim;
  if (hphi_engine_info(e, &dim, NULL, NULL, NULL, NULL, NULL)) return -1;
  int maxit = p->max_iter > 0 ? p->max_iter : 2000;
  if ((uint64_t)maxit > dim) maxit = (int)dim; /* Lanczos_EigenValue.c:126-133 */
  int target = p->target;
  if (target < p->n_eigen) target = p->n_eigen; /* readdef.c:969 */
  int nwant = p->n_eigen > 16 ? 16 : (p->n_eigen < 1 ? 1 : p->n_eigen);
  double eps = p->eps > 0 ? p->eps : 1e-14;
  int every = p->check_every > 0 ? p->check_every : 1;
  double t0 = wall(), thv = 0.0;

  double *v0 = NULL, *v1 = NULL;
  size_t bytes = (size_t)dim * 16;
  if (hphi_dev_alloc(bytes, (void **)&v0) || hphi_dev_alloc(bytes, (void **)&v1)) {
    if (v0) hphi_dev_free(v0);
    return -1;
  }
  double *alpha = (double *)calloc(maxit + 2, sizeof(double));
  double *beta = (double *)calloc(maxit + 2, sizeof(double));
  double *ev = (double *)calloc(maxit + 2, sizeof(double));
  int rc = -1, stp = 1, nprev = 0;
  double pr[2], nrm, b2, ebefor = 0.0, prevv[16];

  /* initial vector (SetInitialVector): random, normalised */
  if (hphi_fill_random(v1, dim, 1, p->seed) || hphi_vec_norm2(e, v1, 1, &nrm)) goto done;
  {
    double ab[4] = {0.0, 0.0, 1.0 / sqrt(nrm), 0.0}; /* v1 <- 0*v1 + (1/|v1|) v1 */
    if (hphi_vec_axpby(e, v1, v1, 1, ab)) goto done;
  }
  if (hphi_dev_zero(v0, bytes)) goto done;
  {
    double t = wall();
    if (hphi_mltply_dev(e, v0, v1, pr)) goto done;
    thv += wall() - t;
  }
  alpha[1] = pr[0];
  if (hphi_vec_lanczos_beta(e, v0, v1, alpha[1], &b2)) goto done;
  beta[1] = sqrt(b2);
  if (dim == 1) {
    r->eigenvalues[0] = alpha[1];
    r->n_found = 1;
    r->iterations = 1;
    r->converged = 1;
    rc = 0;
    goto done;
  }
  for (stp = 2; stp <= maxit; stp++) {
    if (hphi_vec_lanczos_update(e, v0, v1, alpha[stp - 1], beta[stp - 1])) goto done;
    double t = wall();
    if (hphi_mltply_dev(e, v0, v1, pr)) goto done;
    thv += wall() - t;
    alpha[stp] = pr[0];
    if (hphi_vec_lanczos_beta(e, v0, v1, alpha[stp], &b2)) goto done;
    beta[stp] = sqrt(b2);
    if (stp % every != 0 && stp != maxit && beta[stp] >= 1e-14) continue;
    if (p->mode == 1) {
      double cur[16];
      int nc = hphi_lanczos_good_eigenvalues(alpha + 1, beta + 1, stp, nwant, cur);
      r->iterations = stp;
      r->n_found = nc;
      for (int i = 0; i < nc; i++) r->eigenvalues[i] = cur[i];
      if (hphi_lanczos_converged(prevv, nprev, cur, nc, nwant, eps) || fabs(beta[stp]) < 1e-14 ||
          (uint64_t)stp >= dim) {
        r->converged = (nc >= nwant || (uint64_t)stp >= dim || fabs(beta[stp]) < 1e-14);
        break;
      }
      memcpy(prevv, cur, sizeof cur);
      nprev = nc;
      continue;
    }
    int want = (target + 1 > nwant) ? target + 1 : nwant;
    if (want > stp) want = stp;
    hphi_tridiag_lowest(alpha + 1, beta + 1, stp, want, ev);
    for (int i = 0; i < nwant && i < want; i++) r->eigenvalues[i] = ev[i];
    r->n_found = nwant < want ? nwant : want;
    r->iterations = stp;
    if (stp > target) {
      double et = ev[target];
      if (fabs((et - ebefor) / et) < eps || fabs(beta[stp]) < 1e-14) {
        r->converged = 1;
        break;
      }
      ebefor = et;
    }
  }
  rc = 0;
done:
  r->seconds = wall() - t0;
  r->hv_seconds = thv;
  hphi_dev_free(v0);
  hphi_dev_free(v1);
  free(alpha);
  free(beta);
  free(ev);
  return rc;
}
