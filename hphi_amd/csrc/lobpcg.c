/*
 * Block LOBPCG eigenvalue driver (host C over the hphi_amd C ABI).
 *
 * The algorithm of the reference (src/CalcByLOBPCG.c:328-632; A. V. Knyazev,
 * SIAM J. Sci. Comput. 23, 517 (2001); S. Yamada et al., JSCES 20060027):
 *   x = initial block (k columns, normalised), X = H x, p = P = 0,
 *   mu_c = <x_c|X_c>;  for stp = 1, 2, ...:
 *     w = X - mu x; stop when max_c |w_c| < eps (eps scaled by max|mu| > 1);
 *     for stp > 1: optional point-Jacobi preconditioner, w <- w/|w|;
 *     W = H w;  Hsub = {w,x,p}^+ {W,X,P},  O = {w,x,p}^+ {w,x,p}  (3k x 3k);
 *     Loewdin: O = U o U^+, drop o <= 1e-10, T = U o^-1/2, solve T^+ Hsub T;
 *     mu_c <- (<x_c|X_c> + mu_sub_c)/2;
 *     x <- {w,x,p} C_c, X <- {W,X,P} C_c, p <- {w,p} C_c, P <- {W,P} C_c;
 *     x, X scaled by 1/|x|, p, P by 1/|p|.
 * The whole block goes through one block H·v launch per step (the reference
 * calls mltply once per column, :371,470).  Normalisations are carried as
 * per-column scale factors: a stored column times its scale is the
 * reference's vector, and the factors enter the Gram matrices, the residual
 * and the recombination coefficients instead of extra passes over the vectors.
 * The small dense problems use this file's complex Jacobi eigensolver in
 * place of LAPACK zheevd.
 */
#define _POSIX_C_SOURCE 200809L
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "hphi_amd.h"

int hphi_set_error(const char *fmt, ...);

typedef struct hphi_lb_ws hphi_lb_ws;
int hphi_lb_ws_create(hphi_engine *e, int k, hphi_lb_ws **out);
void hphi_lb_ws_destroy(hphi_lb_ws *w);
int hphi_lb_resid(hphi_lb_ws *ws, uint64_t dim, const double *X, const double *x, double *w, const double *mu,
                  const double *s, double *norms2);
int hphi_lb_precond(hphi_lb_ws *ws, uint64_t dim, double *w, const double *diag, const double *shift, double eps,
                    double *norms2);
int hphi_lb_load(hphi_lb_ws *ws, uint64_t dim, const double *src, double *x);
int hphi_lb_store(hphi_lb_ws *ws, uint64_t dim, const double *x, const double *sx, double *dst);
int hphi_lb_gram(hphi_lb_ws *ws, uint64_t dim, const double *w, const double *x, const double *p, const double *W,
                 const double *X, const double *P, double *gram);
int hphi_lb_update(hphi_lb_ws *ws, uint64_t dim, const double *w, double *x, double *p, const double *W, double *X,
                   double *P, const double *coef, double *norms2);
int hphi_lb_hv(hphi_engine *e, int k, double *W, const double *w);
const double *hphi_lb_diag(hphi_engine *e);
int hphi_lb_nsite(const hphi_engine *e);

static double wall(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

typedef double complex zc;

/* Eigen-decomposition of an n x n Hermitian matrix (column-major, interleaved
 * complex) by cyclic Jacobi rotations: w ascending, V's columns the
 * orthonormal eigenvectors.  Each rotation first removes the phase of the
 * pivot (diag(1, e^{-i phi})), then applies the real 2x2 Jacobi rotation. */
int hphi_herm_eig(int n, const double *A_in, double *w, double *V_out) {
  if (n <= 0 || !A_in || !w) return hphi_set_error("hphi_herm_eig: bad argument");
  zc *A = (zc *)malloc(sizeof(zc) * n * n), *V = (zc *)malloc(sizeof(zc) * n * n);
  if (!A || !V) {
    free(A);
    free(V);
    return hphi_set_error("hphi_herm_eig: out of memory");
  }
  memcpy(A, A_in, sizeof(zc) * n * n);
  for (int i = 0; i < n * n; i++) V[i] = 0.0;
  for (int i = 0; i < n; i++) V[i + i * n] = 1.0;
  double scale = 0.0;
  for (int i = 0; i < n * n; i++) scale += creal(A[i] * conj(A[i]));
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0;
    for (int q = 0; q < n; q++)
      for (int p = 0; p < q; p++) off += creal(A[p + q * n] * conj(A[p + q * n]));
    if (off <= 1e-32 * scale || off == 0.0) break;
    for (int q = 1; q < n; q++)
      for (int p = 0; p < q; p++) {
        const double g = cabs(A[p + q * n]);
        if (g == 0.0) continue;
        const zc ph = A[p + q * n] / g; /* e^{i phi} */
        const double tau = (creal(A[q + q * n]) - creal(A[p + p * n])) / (2.0 * g);
        const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
        const double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
        /* U: U_pp = c, U_pq = s, U_qp = -s conj(ph), U_qq = c conj(ph) */
        const zc upp = c, upq = s, uqp = -s * conj(ph), uqq = c * conj(ph);
        for (int k = 0; k < n; k++) { /* A <- A U */
          const zc akp = A[k + p * n], akq = A[k + q * n];
          A[k + p * n] = akp * upp + akq * uqp;
          A[k + q * n] = akp * upq + akq * uqq;
        }
        for (int k = 0; k < n; k++) { /* A <- U^+ A */
          const zc apk = A[p + k * n], aqk = A[q + k * n];
          A[p + k * n] = conj(upp) * apk + conj(uqp) * aqk;
          A[q + k * n] = conj(upq) * apk + conj(uqq) * aqk;
        }
        A[p + q * n] = 0.0;
        A[q + p * n] = 0.0;
        A[p + p * n] = creal(A[p + p * n]);
        A[q + q * n] = creal(A[q + q * n]);
        for (int k = 0; k < n; k++) { /* V <- V U */
          const zc vkp = V[k + p * n], vkq = V[k + q * n];
          V[k + p * n] = vkp * upp + vkq * uqp;
          V[k + q * n] = vkp * upq + vkq * uqq;
        }
      }
  }
  /* ascending order (selection sort; n <= 48) */
  int *ord = (int *)malloc(sizeof(int) * n);
  for (int i = 0; i < n; i++) ord[i] = i;
  for (int i = 0; i < n; i++) {
    int m = i;
    for (int j = i + 1; j < n; j++)
      if (creal(A[ord[j] + ord[j] * n]) < creal(A[ord[m] + ord[m] * n])) m = j;
    int t = ord[i];
    ord[i] = ord[m];
    ord[m] = t;
  }
  for (int i = 0; i < n; i++) w[i] = creal(A[ord[i] + ord[i] * n]);
  if (V_out) {
    zc *Vo = (zc *)V_out;
    for (int i = 0; i < n; i++)
      for (int k = 0; k < n; k++) Vo[k + i * n] = V[k + ord[i] * n];
  }
  free(ord);
  free(A);
  free(V);
  return 0;
}

/* Subspace problem with Loewdin orthogonalisation (CalcByLOBPCG.c:43-123):
 * hsub (n x n, column-major) is replaced by the coefficient matrix whose
 * first columns are the generalised eigenvectors; eig gets the eigenvalues;
 * returns the kept dimension. */
static int diag_ovrp(int n, zc *hsub, zc *ovlp, double *eig) {
  zc *U = (zc *)malloc(sizeof(zc) * n * n), *T = (zc *)calloc((size_t)n * n, sizeof(zc));
  zc *M = (zc *)calloc((size_t)n * n, sizeof(zc)), *Hp = (zc *)calloc((size_t)n * n, sizeof(zc));
  zc *Y = (zc *)malloc(sizeof(zc) * n * n);
  double *o = (double *)malloc(sizeof(double) * n);
  int n2 = 0;
  if (hphi_herm_eig(n, (const double *)ovlp, o, (double *)U)) goto out;
  for (int i = 0; i < n; i++)
    if (o[i] > 1.0e-10) {
      for (int j = 0; j < n; j++) T[j + n * n2] = U[j + n * i] / sqrt(o[i]);
      n2++;
    }
  /* M = hsub T (n x n2); Hp = T^+ M (n2 x n2) */
  for (int c = 0; c < n2; c++)
    for (int r = 0; r < n; r++) {
      zc s = 0.0;
      for (int q = 0; q < n; q++) s += hsub[r + q * n] * T[q + c * n];
      M[r + c * n] = s;
    }
  for (int c = 0; c < n2; c++)
    for (int r = 0; r < n2; r++) {
      zc s = 0.0;
      for (int q = 0; q < n; q++) s += conj(T[q + r * n]) * M[q + c * n];
      Hp[r + c * n2] = s;
    }
  if (n2 > 0 && hphi_herm_eig(n2, (const double *)Hp, eig, (double *)Y)) {
    n2 = 0;
    goto out;
  }
  /* hsub <- T Y (n x n2), remaining columns zero */
  for (int c = 0; c < n; c++)
    for (int r = 0; r < n; r++) {
      zc s = 0.0;
      if (c < n2)
        for (int q = 0; q < n2; q++) s += T[r + q * n] * Y[q + c * n2];
      hsub[r + c * n] = s;
    }
out:
  free(U);
  free(T);
  free(M);
  free(Hp);
  free(Y);
  free(o);
  return n2;
}

/* adaptive shift of the preconditioner (CalcByLOBPCG.c:130-147) */
static double calc_preshift(double eig, double res, double eps) {
  if (res >= 1.0) return 0.0;
  const double k = trunc(log10(fabs(eig)));
  const double i = (eps > res) ? ceil(log10(eps)) : ceil(log10(res));
  return trunc(eig / pow(10.0, k + i)) * pow(10.0, k + i);
}

static int lobpcg_run(hphi_engine *e, const hphi_lobpcg_params *prm, const double *X0, double *Xout, double *history,
                      hphi_lobpcg_result *r);

int hphi_lobpcg(hphi_engine *e, const hphi_lobpcg_params *prm, double *history, hphi_lobpcg_result *r) {
  return lobpcg_run(e, prm, NULL, NULL, history, r);
}

int hphi_lobpcg_ex(hphi_engine *e, const hphi_lobpcg_params *prm, const double *X0, double *Xout, double *history,
                   hphi_lobpcg_result *r) {
  return lobpcg_run(e, prm, X0, Xout, history, r);
}

/* X0 (may be NULL): device, k start vectors of dim complex128 one after
 * another (a restart's tmpvec_set<c> vectors, Initialize_wave :172-213), in
 * place of the synthetic vectors of seed + c; Xout (may be NULL): the final
 * normalised block, the same layout (wxp[1] of :716-724). */
static int lobpcg_run(hphi_engine *e, const hphi_lobpcg_params *prm, const double *X0, double *Xout, double *history,
                      hphi_lobpcg_result *r) {
  if (!e || !prm || !r) return hphi_set_error("hphi_lobpcg: NULL argument");
  memset(r, 0, sizeof *r);
  const int k = prm->k;
  if (k != 1 && k != 2 && k != 4 && k != 8 && k != 16)
    return hphi_set_error("hphi_lobpcg: block size %d not in {1,2,4,8,16}", k);
  uint64_t dim;
  if (hphi_engine_info(e, &dim, NULL, NULL, NULL, NULL, NULL)) return -1;
  if ((uint64_t)(3 * k) > dim) return hphi_set_error("hphi_lobpcg: 3k = %d exceeds the dimension %llu", 3 * k,
                                                     (unsigned long long)dim);
  const int maxit = prm->max_iter > 0 ? prm->max_iter : 500;
  const int nsub = 3 * k;
  const double eig_pos_shift = prm->large_value * hphi_lb_nsite(e);
  double t0 = wall(), thv = 0.0;
  int rc = -1;

  /* the reference's tables are needed before the first H·v */
  if (hphi_sz(e, NULL, NULL, NULL) || hphi_diagonalcalc(e, NULL)) return -1;
  hphi_lb_ws *ws = NULL;
  if (hphi_lb_ws_create(e, k, &ws)) return -1;
  double *blk[6] = {NULL, NULL, NULL, NULL, NULL, NULL}; /* w x p W X P */
  const size_t bytes = (size_t)dim * k * 16;
  zc *hsub = (zc *)calloc((size_t)nsub * nsub, sizeof(zc)), *ovlp = (zc *)calloc((size_t)nsub * nsub, sizeof(zc));
  double *gram = (double *)malloc(sizeof(double) * 2 * nsub * 2 * nsub);
  zc *coef = (zc *)malloc(sizeof(zc) * nsub * k);
  double *eigsub = (double *)calloc(nsub, sizeof(double));
  double mu[16], sx[16], sp[16], sw[16], nrm[32], shift[16];
  for (int i = 0; i < 6; i++)
    if (hphi_dev_alloc(bytes, (void **)&blk[i])) goto done;
  double *w = blk[0], *x = blk[1], *p = blk[2], *W = blk[3], *X = blk[4], *P = blk[5];

  /* initial block: the given vectors, or column c from seed + c; normalised
   * (Initialize_wave, :166-318) */
  if (X0 ? hphi_lb_load(ws, dim, X0, x) : hphi_fill_random(x, dim, k, prm->seed)) goto done;
  if (hphi_vec_norm2(e, x, k, nrm)) goto done;
  for (int c = 0; c < k; c++) sx[c] = 1.0 / sqrt(nrm[c]), sp[c] = 1.0, sw[c] = 1.0;
  if (hphi_dev_zero(p, bytes) || hphi_dev_zero(P, bytes)) goto done;
  {
    double t = wall();
    if (hphi_lb_hv(e, k, X, x) || hphi_engine_sync(e)) goto done;
    thv += wall() - t;
  }
  {
    double d[32];
    if (hphi_vec_dot(e, x, X, k, d)) goto done;
    for (int c = 0; c < k; c++) mu[c] = sx[c] * sx[c] * d[2 * c];
  }

  int stp, conv = 0;
  double dmax = 0.0;
  for (stp = 1; stp <= maxit; stp++) {
    double amax = 0.0;
    for (int c = 0; c < k; c++)
      if (fabs(mu[c]) > amax) amax = fabs(mu[c]);
    double eps = prm->eps;
    if (amax > 1.0) eps *= amax;
    /* residual w = X - mu x of the normalised columns */
    if (hphi_lb_resid(ws, dim, X, x, w, mu, sx, nrm)) goto done;
    dmax = 0.0;
    for (int c = 0; c < k; c++) {
      const double dn = sqrt(nrm[c]);
      r->residuals[c] = dn;
      if (dn > dmax) dmax = dn;
      shift[c] = calc_preshift(mu[c] + eig_pos_shift, dn, eps) - eig_pos_shift;
      sw[c] = 1.0;
    }
    if (history) {
      history[(size_t)(stp - 1) * (k + 1)] = dmax;
      for (int c = 0; c < k; c++) history[(size_t)(stp - 1) * (k + 1) + 1 + c] = mu[c];
    }
    r->iterations = stp;
    if (dmax < eps) {
      conv = 1;
      break;
    }
    if (stp != 1) {
      if (prm->precg) {
        if (hphi_lb_precond(ws, dim, w, hphi_lb_diag(e), shift, eps, nrm)) goto done;
      }
      for (int c = 0; c < k; c++) sw[c] = 1.0 / sqrt(nrm[c]);
    }
    {
      double t = wall(); /* the stream is idle here: the last reduction synchronised it */
      if (hphi_lb_hv(e, k, W, w) || hphi_engine_sync(e)) goto done;
      thv += wall() - t;
    }
    /* Gram matrices of the stored vectors, then the column scales */
    if (hphi_lb_gram(ws, dim, w, x, p, W, X, P, gram)) goto done;
    {
      double sc[48];
      for (int c = 0; c < k; c++) sc[c] = sw[c], sc[k + c] = sx[c], sc[2 * k + c] = sp[c];
      for (int rr = 0; rr < nsub; rr++)
        for (int cc = 0; cc < nsub; cc++) {
          const double *go = gram + 2 * ((size_t)rr * 2 * nsub + cc);
          const double *gh = gram + 2 * ((size_t)rr * 2 * nsub + nsub + cc);
          const double f = sc[rr] * sc[cc];
          ovlp[rr + cc * nsub] = f * (go[0] + I * go[1]);
          hsub[rr + cc * nsub] = f * (gh[0] + I * gh[1]);
        }
    }
    for (int c = 0; c < k; c++) mu[c] = creal(hsub[(k + c) + (k + c) * nsub]);
    const int n2 = diag_ovrp(nsub, hsub, ovlp, eigsub);
    for (int c = 0; c < k; c++)
      if (c < n2) mu[c] = 0.5 * (mu[c] + eigsub[c]);
    /* recombination coefficients for the stored rows: C[r][c] * scale_r */
    {
      double sc[48];
      for (int c = 0; c < k; c++) sc[c] = sw[c], sc[k + c] = sx[c], sc[2 * k + c] = sp[c];
      for (int rr = 0; rr < nsub; rr++)
        for (int c = 0; c < k; c++) coef[rr * k + c] = sc[rr] * hsub[rr + c * nsub];
    }
    if (hphi_lb_update(ws, dim, w, x, p, W, X, P, (const double *)coef, nrm)) goto done;
    for (int c = 0; c < k; c++) {
      sx[c] = nrm[c] > 0.0 ? 1.0 / sqrt(nrm[c]) : 1.0;
      sp[c] = nrm[k + c] > 0.0 ? 1.0 / sqrt(nrm[k + c]) : 1.0;
    }
  }
  /* Rayleigh quotients of the final normalised block */
  {
    double d[32];
    if (hphi_vec_dot(e, x, X, k, d)) goto done;
    for (int c = 0; c < k; c++) r->eigenvalues[c] = sx[c] * sx[c] * d[2 * c];
  }
  r->converged = conv;
  r->max_residual = dmax;
  if (Xout && hphi_lb_store(ws, dim, x, sx, Xout)) goto done;
  rc = 0;
done:
  for (int i = 0; i < 6; i++)
    if (blk[i]) hphi_dev_free(blk[i]);
  hphi_lb_ws_destroy(ws);
  free(hsub);
  free(ovlp);
  free(gram);
  free(coef);
  free(eigsub);
  r->seconds = wall() - t0;
  r->hv_seconds = thv;
  return rc;
}
