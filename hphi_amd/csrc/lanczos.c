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
void hphi_tridiag_eigvec(const double *a, const double *b, int n, double lam, double *x);

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

/* State of a run continued from, or handed to, the reference's restart files
 * (Lanczos_EigenValue.c:78-100): n steps done, alpha_1..alpha_n and beta_1..beta_n
 * (0-based arrays), and the host vectors v0, v1 after step n's H·v. */
typedef struct {
  int n;
  double *alpha, *beta;
  double *v0, *v1;
  double norm;
} lz_state;

/* The eigenvalue pass.  On success *alpha_out / *beta_out (1-based, length
 * max_iter + 2) hold the recurrence coefficients of the r->iterations steps;
 * the caller frees them (NULL pointers: freed here).  With in->n > 0 the run
 * continues from that state for up to max_iter further steps; out (if not
 * NULL) receives the final state (arrays allocated here, freed by the caller). */
static int lanczos_core(hphi_engine *e, const hphi_lanczos_params *p, hphi_lanczos_result *r, double **alpha_out,
                        double **beta_out, const lz_state *in, lz_state *out) {
  memset(r, 0, sizeof *r);
  uint64_t dim;
  if (hphi_engine_info(e, &dim, NULL, NULL, NULL, NULL, NULL)) return -1;
  const int n0 = (in && in->n > 0) ? in->n : 0;
  int maxit = (p->max_iter > 0 ? p->max_iter : 2000) + n0; /* Lanczos_EigenValue.c:93 */
  if ((uint64_t)maxit > dim) maxit = (int)dim; /* Lanczos_EigenValue.c:126-133 */
  if (n0 && n0 >= maxit) return hphi_set_error("restart: %d steps done, nothing left to run", n0);
  int target = p->target;
  if (target < p->n_eigen) target = p->n_eigen; /* readdef.c:969 */
  int nwant = p->n_eigen > 16 ? 16 : (p->n_eigen < 1 ? 1 : p->n_eigen);
  double eps = p->eps > 0 ? p->eps : 1e-14;
  int every = p->check_every > 0 ? p->check_every : 1;
  double t0 = wall(), thv = 0.0;

  /* The recurrence of Lanczos_EigenValue.c:100-177 with the normalisation
   * deferred: w_m = q_m / s_m is kept unscaled (s_m = 1/beta_{m-1}), so a step
   * is one H·v that also forms r_m = s_m H w_m - beta_{m-1} s_{m-1} w_{m-1}
   * (hphi_mltply_dev_scaled), and one residual pass w_{m+1} = r_m -
   * alpha_m s_m w_m with |w_{m+1}| = beta_m (hphi_vec_lanczos_resid), instead
   * of the reference's H·v, beta pass and update pass.  alpha_m = s_m^2 <w_m|H w_m>.
   * Buffers: ra holds r_m, then w_{m+1}; wb holds w_m. */
  double *ra = NULL, *wb = NULL;
  size_t bytes = (size_t)dim * 16;
  if (hphi_dev_alloc(bytes, (void **)&ra) || hphi_dev_alloc(bytes, (void **)&wb)) {
    if (ra) hphi_dev_free(ra);
    return -1;
  }
  double *alpha = (double *)calloc(maxit + 2, sizeof(double));
  double *beta = (double *)calloc(maxit + 2, sizeof(double));
  double *ev = (double *)calloc(maxit + 2, sizeof(double));
  int rc = -1, stp = 1, nprev = 0, last = 0, stp0 = 2;
  double pr[2], nrm, b2, ebefor = 0.0, prevv[16], norm0 = 0.0, s_prev = 1.0, s_cur = 1.0;

  if (n0) { /* continue: the reference's state of step n0, v0 = r_n0 and v1 = q_n0 */
    for (int i = 0; i < n0; i++) {
      alpha[i + 1] = in->alpha[i];
      beta[i + 1] = in->beta[i];
    }
    if (hphi_memcpy_h2d(ra, in->v0, bytes) || hphi_memcpy_h2d(wb, in->v1, bytes)) goto done;
    norm0 = in->norm;
    /* w_{n0+1} = r - alpha q (s = 1); beta_n0 is the file's, as in the reference */
    if (hphi_vec_lanczos_resid(e, ra, wb, NULL, alpha[n0], 0.0, &b2)) goto done;
    s_cur = 1.0;
    stp0 = n0 + 1;
    last = n0;
    goto loop;
  }
  /* initial vector (SetInitialVector): random; q_1 = s_1 w_1 */
  if (hphi_fill_random(wb, dim, 1, p->seed) || hphi_vec_norm2(e, wb, 1, &nrm)) goto done;
  norm0 = sqrt(nrm);
  s_cur = 1.0 / sqrt(nrm);
  {
    double t = wall();
    if (hphi_mltply_dev_scaled(e, ra, wb, 0.0, s_cur, 1, pr)) goto done; /* r_1 = H q_1 */
    thv += wall() - t;
  }
  alpha[1] = s_cur * s_cur * pr[0];
  if (hphi_vec_lanczos_resid(e, ra, wb, NULL, alpha[1] * s_cur, 0.0, &b2)) goto done;
  beta[1] = sqrt(b2);
  last = 1;
  if (dim == 1) {
    r->eigenvalues[0] = alpha[1];
    r->n_found = 1;
    r->iterations = 1;
    r->converged = 1;
    rc = 0;
    goto done;
  }
loop:
  for (stp = stp0; stp <= maxit; stp++) {
    s_prev = s_cur;
    s_cur = 1.0 / beta[stp - 1];
    {
      double t = wall();
      /* wb <- -beta_{m-1} s_{m-1} w_{m-1} + s_m H w_m = r_m; then ra, wb swap roles */
      if (hphi_mltply_dev_scaled(e, wb, ra, -beta[stp - 1] * s_prev, s_cur, 0, pr)) goto done;
      thv += wall() - t;
      double *tmp = ra;
      ra = wb;
      wb = tmp;
    }
    alpha[stp] = s_cur * s_cur * pr[0];
    if (hphi_vec_lanczos_resid(e, ra, wb, NULL, alpha[stp] * s_cur, 0.0, &b2)) goto done;
    beta[stp] = sqrt(b2);
    last = stp;
    if (stp % every != 0 && stp != maxit && beta[stp] >= 1e-14) continue;
    if (p->mode == 1) {
      double cur[16];
      int nc = hphi_lanczos_good_eigenvalues(alpha + 1, beta + 1, stp, nwant, cur);
      r->iterations = stp;
      r->n_found = nc;
      for (int i = 0; i < nc; i++) r->eigenvalues[i] = cur[i];
      /* converged: the values stopped moving and their Ritz residuals are small
       * (a stalled value can pass the first test while still off) */
      if ((hphi_lanczos_converged(prevv, nprev, cur, nc, nwant, eps) &&
           hphi_lanczos_residuals_small(alpha + 1, beta + 1, stp, beta[stp], cur, nwant, eps)) ||
          fabs(beta[stp]) < 1e-14 || (uint64_t)stp >= dim) {
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
  if (out) { /* the reference's state of the last completed step: v0 = r, v1 = q */
    out->n = last;
    out->norm = norm0;
    out->alpha = (double *)malloc(sizeof(double) * (last + 1));
    out->beta = (double *)malloc(sizeof(double) * (last + 1));
    out->v0 = (double *)malloc(bytes);
    out->v1 = (double *)malloc(bytes);
    if (!out->alpha || !out->beta || !out->v0 || !out->v1) {
      rc = hphi_set_error("restart: host allocation failed");
    } else {
      for (int i = 0; i < last; i++) {
        out->alpha[i] = alpha[i + 1];
        out->beta[i] = beta[i + 1];
      }
      /* r_last = w_{last+1} + alpha_last s_last w_last, q_last = s_last w_last */
      double ar[4] = {alpha[last] * s_cur, 0.0, 1.0, 0.0}, aq[4] = {0.0, 0.0, s_cur, 0.0};
      if (hphi_vec_axpby(e, wb, ra, 1, ar) || hphi_vec_axpby(e, wb, wb, 1, aq) ||
          hphi_memcpy_d2h(out->v0, ra, bytes) || hphi_memcpy_d2h(out->v1, wb, bytes))
        rc = -1;
    }
  }
done:
  r->seconds = wall() - t0;
  r->hv_seconds = thv;
  hphi_dev_free(ra);
  hphi_dev_free(wb);
  if (rc == 0 && alpha_out && beta_out) {
    *alpha_out = alpha;
    *beta_out = beta;
  } else {
    free(alpha);
    free(beta);
  }
  free(ev);
  return rc;
}

int hphi_lanczos(hphi_engine *e, const hphi_lanczos_params *p, hphi_lanczos_result *r) {
  if (!e || !p || !r) return hphi_set_error("hphi_lanczos: NULL argument");
  return lanczos_core(e, p, r, NULL, NULL, NULL, NULL);
}

static void lz_free(lz_state *s) {
  free(s->alpha);
  free(s->beta);
  free(s->v0);
  free(s->v1);
  memset(s, 0, sizeof *s);
}

int hphi_lanczos_restart(hphi_engine *e, const hphi_lanczos_params *p, const char *tm_in, const char *vec_in,
                         const char *tm_out, const char *vec_out, hphi_lanczos_result *r) {
  if (!e || !p || !r) return hphi_set_error("hphi_lanczos_restart: NULL argument");
  if ((tm_in == NULL) != (vec_in == NULL) || (tm_out == NULL) != (vec_out == NULL))
    return hphi_set_error("hphi_lanczos_restart: the components and vector files come in pairs");
  uint64_t dim;
  if (hphi_engine_info(e, &dim, NULL, NULL, NULL, NULL, NULL)) return -1;
  lz_state in = {0}, out = {0};
  int rc = -1;
  if (tm_in) { /* ReadInitialVector + ReadTMComponents (Lanczos_EigenValue.c:78-91) */
    const int cap = 1 << 20;
    uint64_t nvec = 0;
    in.alpha = (double *)malloc(sizeof(double) * cap);
    in.beta = (double *)malloc(sizeof(double) * cap);
    in.v0 = (double *)malloc((size_t)dim * 16);
    in.v1 = (double *)malloc((size_t)dim * 16);
    if (!in.alpha || !in.beta || !in.v0 || !in.v1) {
      hphi_set_error("restart: host allocation failed");
      goto end;
    }
    if (hphi_read_tmcomponents(tm_in, cap, &in.n, &in.norm, in.alpha, in.beta) ||
        hphi_read_restart_vecs(vec_in, dim, in.v0, in.v1, &nvec))
      goto end;
    if (nvec != (uint64_t)in.n) {
      hphi_set_error("restart: %s holds step %llu but %s has %d components", vec_in, (unsigned long long)nvec,
                     tm_in, in.n);
      goto end;
    }
  }
  if (lanczos_core(e, p, r, NULL, NULL, &in, tm_out ? &out : NULL)) goto end;
  if (tm_out) { /* OutputTMComponents + OutputLanczosVector (Lanczos_EigenValue.c:282-287,456-489) */
    if (hphi_write_tmcomponents(tm_out, out.n, out.norm, out.alpha, out.beta) ||
        hphi_write_restart_vecs(vec_out, (uint64_t)out.n, dim, out.v0, out.v1))
      goto end;
  }
  rc = 0;
end:
  lz_free(&in);
  lz_free(&out);
  return rc;
}


int hphi_lanczos_eigenvector(hphi_engine *e, const hphi_lanczos_params *p, int k, double *vout,
                             hphi_lanczos_result *r) {
  if (!e || !p || !r || !vout) return hphi_set_error("hphi_lanczos_eigenvector: NULL argument");
  if (k < 0 || k >= 16) return hphi_set_error("hphi_lanczos_eigenvector: k must be in 0..15");
  double *alpha = NULL, *beta = NULL;
  if (lanczos_core(e, p, r, &alpha, &beta, NULL, NULL)) return -1;
  uint64_t dim;
  if (hphi_engine_info(e, &dim, NULL, NULL, NULL, NULL, NULL)) return -1;
  const int m = r->iterations > 0 ? r->iterations : 1;
  if (k >= m) {
    free(alpha);
    free(beta);
    return hphi_set_error("hphi_lanczos_eigenvector: k = %d but only %d Lanczos steps", k, m);
  }
  double *ev = (double *)calloc(k + 1, sizeof(double)), *c = (double *)calloc(m, sizeof(double));
  hphi_tridiag_lowest(alpha + 1, beta + 1, m, k + 1, ev);
  hphi_tridiag_eigvec(alpha + 1, beta + 1, m, ev[k], c);
  r->eigenvalues[0] = ev[k];
  double t0 = wall(), thv = 0.0, pr[2], nrm, b2;
  double *ra = NULL, *wb = NULL;
  const size_t bytes = (size_t)dim * 16;
  int rc = -1;
  if (hphi_dev_alloc(bytes, (void **)&ra) || hphi_dev_alloc(bytes, (void **)&wb)) goto done;
  /* The eigenvalue pass's operations again, in the same order on the same start
   * vector (Lanczos_EigenVector.c:65-190), so every w_s is bitwise the same; the
   * residual pass of step s+1 also adds c_s q_s = c_s s_s w_s to v. */
  if (hphi_fill_random(wb, dim, 1, p->seed) || hphi_vec_norm2(e, wb, 1, &nrm)) goto done;
  if (hphi_dev_zero(vout, bytes)) goto done;
  {
    double s_cur = 1.0 / sqrt(nrm), s_prev;
    if (m == 1) {
      double ab[4] = {c[0] * s_cur, 0.0, 0.0, 0.0}; /* v = c_1 q_1 */
      if (hphi_vec_axpby(e, wb, vout, 1, ab)) goto done;
    } else {
      double t = wall();
      if (hphi_mltply_dev_scaled(e, ra, wb, 0.0, s_cur, 1, pr)) goto done;
      thv += wall() - t;
    }
    for (int st = 2; st <= m; st++) {
      if (hphi_vec_lanczos_resid(e, ra, wb, vout, alpha[st - 1] * s_cur, c[st - 2] * s_cur, &b2)) goto done;
      s_prev = s_cur;
      s_cur = 1.0 / beta[st - 1];
      if (st < m) {
        double t = wall();
        if (hphi_mltply_dev_scaled(e, wb, ra, -beta[st - 1] * s_prev, s_cur, 0, pr)) goto done;
        thv += wall() - t;
        double *tmp = ra;
        ra = wb;
        wb = tmp;
      } else {
        double ab[4] = {c[m - 1] * s_cur, 0.0, 1.0, 0.0}; /* v += c_m q_m, q_m = s_m w_m (in ra) */
        if (hphi_vec_axpby(e, ra, vout, 1, ab)) goto done;
      }
    }
  }
  if (hphi_vec_norm2(e, vout, 1, &nrm)) goto done;
  {
    double ab[4] = {1.0 / sqrt(nrm), 0.0, 0.0, 0.0};
    if (hphi_vec_axpby(e, vout, vout, 1, ab)) goto done;
  }
  rc = 0;
done:
  r->seconds += wall() - t0;
  r->hv_seconds += thv;
  hphi_dev_free(ra);
  hphi_dev_free(wb);
  free(alpha);
  free(beta);
  free(ev);
  free(c);
  return rc;
}
