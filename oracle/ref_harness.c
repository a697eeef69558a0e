/*
 * Reference harness — TEST INFRASTRUCTURE ONLY.
 *
 * A main() of this project's own, linked against the reference's hot-path
 * translation units compiled in place from the reference tree by
 * oracle/Makefile (nothing of the reference is copied here).  It runs the
 * reference's own start-up sequence (src/HPhiMain.c:222-288: ReadDefFileNInt,
 * setmem_def, ReadDefFileIdxPara, SetConvergenceFactor, HPhiTrans, check,
 * setmem_large, sz, diagonalcalc) on an expert-mode input and then calls the
 * reference mltply (src/mltply.c:56).
 *
 * usage: ref_harness <def-dir> <mode> <out-dir> [seed] [reps]
 *   mode tables : write list_1.bin list_2_1.bin list_2_2.bin list_Diagonal.bin
 *                 (raw little-endian arrays exactly as the reference holds them)
 *                 and meta.txt
 *   mode hv     : v1 = synthetic vector(seed) (same SplitMix64 generator as the
 *                 oracle), v0 = 0, one mltply; write v0.bin (dim complex128,
 *                 0-based), prdct.txt
 *   mode sums   : as hv but write only checksums (sums.txt) — for large dims
 *   mode sums_seeds: one mltply per seed s = seed .. seed+reps-1 of the synthetic
 *                 vector of s; sums_<s>.txt and prdct_<s>.txt
 *   mode bench  : reps timed mltply calls; print one JSON line
 *   mode green  : v0 = the synthetic vector(seed) normalised; expec_energy_flct
 *                 (src/expec_energy_flct.c:40-160: moments on v0, then v1 = v0,
 *                 v0 = H v1) -> phys.txt; expec_cisajs and expec_cisajscktaltdc on
 *                 v1 write output/zvo_cisajs.dat and output/zvo_cisajscktalt.dat
 *   mode tpq    : the reference's CalcByTPQ (src/CalcByTPQ.c:52-294) as HPhiMain.c:315-321
 *                 calls it (input written with CalcType 1); it writes output/SS_rand*.dat,
 *                 Norm_rand*.dat, Flct_rand*.dat
 *   mode ctpq   : the same for CalcByCanonicalTPQ (src/CalcByCanonicalTPQ.c:49-300,
 *                 HPhiMain.c:323-329; CalcType 5)
 *   mode lanczos: the recurrence of src/Lanczos_EigenValue.c:102-177 (v1 = the
 *                 synthetic vector(seed) normalised, v0 = 0; mltply; alpha = Re prdct;
 *                 beta = |v0 - alpha v1|; then v1 <- (v0 - alpha v1)/beta,
 *                 v0 <- -beta v1_old, mltply, ...) for [reps] steps, driven by the
 *                 reference mltply; ab.txt gets "step alpha beta" per step, flushed
 *                 as it goes (the tridiagonal eigenvalues are taken from these
 *                 numbers by the caller: Lanczos_EigenValue's DSEVvalue needs
 *                 LAPACK, which is absent)
 *                 Both TPQ modes first write each sample's start vector as the
 *                 reference makes it in this process (MakeIniVec(rand_i), which
 *                 depends on the thread count): ini_rand<i>.bin = global_1st_norm
 *                 (double) then v1[1..dim] (complex128, normalised)
 * The input directory is entered with chdir, as the reference opens its def
 * files by relative name and writes ./output/.
 *
 * Built with -DHPHI_DROPIN (oracle/Makefile target `dropin`, _ref/ref_dropin)
 * the same program is the reference with this project's drop-in binding
 * (integration/mltply_hphi_amd.c) linked in place of src/mltply.c:
 * hphi_amd_setup() replaces sz() and diagonalcalc(), and every mltply call of
 * the reference's own callers (CalcByTPQ, CalcByCanonicalTPQ, expec_*, the
 * modes above) runs on the GPU through libhphi_amd.  One more mode there:
 *   mode desc   : host only, right after the reference's reader: the
 *                 binding's descriptor of X->Def against hphi_read_def's of the
 *                 same namelist, field by field; prints "desc ok" or the first
 *                 difference (exit 1)
 */
#include <errno.h>
#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include "CalcByCanonicalTPQ.h"
#include "CalcByTPQ.h"
#include "MakeIniVec.h"
#include "Common.h"
#include "HPhiTrans.h"
#include "check.h"
#include "diagonalcalc.h"
#include "expec_cisajs.h"
#include "expec_cisajscktaltdc.h"
#include "expec_energy_flct.h"
#include "global.h"
#include "mltply.h"
#include "readdef.h"
#include "sz.h"
#include "wrapperMPI.h"
#include "xsetmem.h"

#ifdef HPHI_DROPIN
#include "hphi_amd.h"
struct hphi_amd_desc_buf {
  int *ex, *pl, *tr, *ci, *cv, *hd, *ia, *iad, *ph;
  double *trt, *iav;
};
int hphi_amd_fill_desc(const struct DefineList *D, hphi_model_desc *d, struct hphi_amd_desc_buf *b);
void hphi_amd_desc_buf_free(struct hphi_amd_desc_buf *b);
int hphi_amd_setup(struct BindStruct *X);

static int cmp_i(const char *what, int n, int w, const int *a, const int *b) {
  for (int i = 0; i < n * w; i++)
    if (a[i] != b[i]) {
      printf("desc differs: %s[%d][%d] %d vs %d\n", what, i / w, i % w, a[i], b[i]);
      return 1;
    }
  return 0;
}
static int cmp_d(const char *what, int n, int w, const double *a, const double *b) {
  for (int i = 0; i < n * w; i++)
    if (a[i] != b[i]) {
      printf("desc differs: %s[%d][%d] %.17g vs %.17g\n", what, i / w, i % w, a[i], b[i]);
      return 1;
    }
  return 0;
}
/* the binding's descriptor of the reference's X->Def against hphi_read_def's */
static int desc_check(struct DefineList *D) {
  hphi_model_desc a, *bp = NULL;
  struct hphi_amd_desc_buf buf;
  if (hphi_amd_fill_desc(D, &a, &buf)) return 4;
  if (hphi_read_def("namelist.def", &bp, NULL)) {
    printf("hphi_read_def: %s\n", hphi_last_error());
    return 1;
  }
  const hphi_model_desc *b = bp;
  int bad = 0;
#define N(f)                                                                       \
  if (!bad && a.f != b->f) {                                                       \
    printf("desc differs: %s %d vs %d\n", #f, a.f, b->f);                          \
    bad = 1;                                                                       \
  }
  N(model) N(nsite) N(nup) N(ndown) N(n_exchange) N(n_pairlift) N(n_transfer) N(n_chemi) N(n_intra) N(n_inter)
  N(n_hund) N(n_interall) N(n_interall_diag) N(n_pairhop)
#undef N
  bad = bad || cmp_i("exchange_ij", a.n_exchange, 2, a.exchange_ij, b->exchange_ij) ||
        cmp_d("exchange_J", a.n_exchange, 1, a.exchange_J, b->exchange_J) ||
        cmp_i("pairlift_ij", a.n_pairlift, 2, a.pairlift_ij, b->pairlift_ij) ||
        cmp_d("pairlift_J", a.n_pairlift, 1, a.pairlift_J, b->pairlift_J) ||
        cmp_i("transfer_idx", a.n_transfer, 4, a.transfer_idx, b->transfer_idx) ||
        cmp_d("transfer_t", a.n_transfer, 2, a.transfer_t, b->transfer_t) ||
        cmp_i("chemi_site", a.n_chemi, 1, a.chemi_site, b->chemi_site) ||
        cmp_i("chemi_spin", a.n_chemi, 1, a.chemi_spin, b->chemi_spin) ||
        cmp_d("chemi_mu", a.n_chemi, 1, a.chemi_mu, b->chemi_mu) ||
        cmp_i("intra_site", a.n_intra, 1, a.intra_site, b->intra_site) ||
        cmp_d("intra_U", a.n_intra, 1, a.intra_U, b->intra_U) ||
        cmp_i("inter_ij", a.n_inter, 2, a.inter_ij, b->inter_ij) ||
        cmp_d("inter_V", a.n_inter, 1, a.inter_V, b->inter_V) ||
        cmp_i("hund_ij", a.n_hund, 2, a.hund_ij, b->hund_ij) ||
        cmp_d("hund_J", a.n_hund, 1, a.hund_J, b->hund_J) ||
        cmp_i("interall_idx", a.n_interall, 8, a.interall_idx, b->interall_idx) ||
        cmp_d("interall_V", a.n_interall, 2, a.interall_V, b->interall_V) ||
        cmp_i("interall_diag_idx", a.n_interall_diag, 4, a.interall_diag_idx, b->interall_diag_idx) ||
        cmp_d("interall_diag_V", a.n_interall_diag, 1, a.interall_diag_V, b->interall_diag_V) ||
        cmp_i("pairhop_ij", a.n_pairhop, 2, a.pairhop_ij, b->pairhop_ij) ||
        cmp_d("pairhop_J", a.n_pairhop, 1, a.pairhop_J, b->pairhop_J);
  hphi_model_desc_free(bp);
  hphi_amd_desc_buf_free(&buf);
  if (!bad) printf("desc ok\n");
  return bad;
}
#endif

static uint64_t sm64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
static double unit_rng(uint64_t seed, uint64_t c) {
  return (double)(sm64((seed * 0x9E3779B97F4A7C15ULL) ^ c) >> 11) * (1.0 / 9007199254740992.0);
}

static int dump(const char *dir, const char *name, const void *p, size_t bytes) {
  char path[4096];
  snprintf(path, sizeof path, "%s/%s", dir, name);
  FILE *f = fopen(path, "wb");
  if (!f) return -1;
  size_t w = fwrite(p, 1, bytes, f);
  fclose(f);
  return w == bytes ? 0 : -1;
}

static double now(void) { return omp_get_wtime(); }

/* checksums of v0[1..dim]: S1 = sum v0, S2 = sum |v0|^2, S3 = sum w_j v0[j] with
 * w_j = (j mod 97) + 1, and 64 sampled entries at j = 1 + (k * 2654435761) mod dim */
static int write_sums(const char *path, const double complex *v0, unsigned long dim) {
  double s1r = 0, s1i = 0, s2 = 0, s3r = 0, s3i = 0;
#pragma omp parallel for reduction(+ : s1r, s1i, s2, s3r, s3i)
  for (unsigned long j = 1; j <= dim; j++) {
    double a = creal(v0[j]), b = cimag(v0[j]);
    double w = (double)((j % 97) + 1);
    s1r += a; s1i += b; s2 += a * a + b * b; s3r += w * a; s3i += w * b;
  }
  FILE *f = fopen(path, "w");
  if (!f) return -1;
  fprintf(f, "S1 %.17g %.17g\nS2 %.17g\nS3 %.17g %.17g\n", s1r, s1i, s2, s3r, s3i);
  for (unsigned long k = 0; k < 64; k++) {
    unsigned long j = 1 + (unsigned long)((k * 2654435761ULL) % dim);
    fprintf(f, "E %lu %.17g %.17g\n", j, creal(v0[j]), cimag(v0[j]));
  }
  fclose(f);
  return 0;
}

int main(int argc, char *argv[]) {
  if (argc < 4) {
    fprintf(stderr, "usage: %s <def-dir> tables|hv|sums|bench <out-dir> [seed] [reps]\n", argv[0]);
    return 2;
  }
  char outdir[4096];
  if (argv[3][0] == '/') snprintf(outdir, sizeof outdir, "%s", argv[3]);
  else {
    char cwd[3000];
    if (!getcwd(cwd, sizeof cwd)) return 2;
    snprintf(outdir, sizeof outdir, "%s/%s", cwd, argv[3]);
  }
  uint64_t seed = argc > 4 ? strtoull(argv[4], NULL, 10) : 0;
  int reps = argc > 5 ? atoi(argv[5]) : 3;
  mkdir(outdir, 0777);
  if (chdir(argv[1]) != 0) {
    perror("chdir");
    return 2;
  }
  mkdir("output", 0777);

  /* start-up sequence of src/HPhiMain.c:222-288 */
  InitializeMPI(argc, argv);
  InitTimer();
  char nl[] = "namelist.def";
  setmem_HEAD(&X.Bind);
  if (ReadDefFileNInt(nl, &(X.Bind.Def), &(X.Bind.Boost)) != 0) return 3;
  setmem_def(&X.Bind, &X.Bind.Boost);
  if (ReadDefFileIdxPara(&(X.Bind.Def), &(X.Bind.Boost)) != 0) return 3;
  SetConvergenceFactor(&(X.Bind.Def));
  if (HPhiTrans(&(X.Bind)) != 0) return 3;
  const char *mode = argv[2];
#ifdef HPHI_DROPIN
  if (strcmp(mode, "desc") == 0) return desc_check(&X.Bind.Def);
#endif
  if (check(&(X.Bind)) == MPIFALSE) return 3;
  if (setmem_large(&X.Bind) != 0) return 3;
  double t0 = now();
#ifdef HPHI_DROPIN
  /* the binding's set-up in place of sz() and diagonalcalc() (HPhiMain.c:278,288) */
  if (hphi_amd_setup(&(X.Bind)) != 0) return 3;
  double t_sz = now() - t0, t_diag = 0.0;
#else
  if (sz(&(X.Bind), list_1, list_2_1, list_2_2) != 0) return 3;
  double t_sz = now() - t0;
  t0 = now();
  diagonalcalc(&(X.Bind));
  double t_diag = now() - t0;
#endif

  unsigned long dim = X.Bind.Check.idim_max;
  int model = X.Bind.Def.iCalcModel;
  int canon = (model == Spin || model == Hubbard);

  if (strcmp(mode, "tables") == 0) {
    FILE *f;
    char path[4200];
    snprintf(path, sizeof path, "%s/meta.txt", outdir);
    f = fopen(path, "w");
    fprintf(f, "dim %lu\nsdim %lu\nsize21 %lu\nsize22 %lu\nmodel %d\n", dim, X.Bind.Check.sdim,
            (unsigned long)X.Bind.Large.SizeOflist_2_1, (unsigned long)X.Bind.Large.SizeOflist_2_2, model);
    fclose(f);
    if (canon) {
      if (dump(outdir, "list_1.bin", list_1, (dim + 1) * sizeof(long unsigned int))) return 4;
      if (dump(outdir, "list_2_1.bin", list_2_1, X.Bind.Large.SizeOflist_2_1 * sizeof(long unsigned int))) return 4;
      if (dump(outdir, "list_2_2.bin", list_2_2, X.Bind.Large.SizeOflist_2_2 * sizeof(long unsigned int))) return 4;
    }
    if (dump(outdir, "list_Diagonal.bin", list_Diagonal, (dim + 1) * sizeof(double))) return 4;
    return 0;
  }

  if (strcmp(mode, "tpq") == 0 || strcmp(mode, "ctpq") == 0) {
    const int canonical = strcmp(mode, "ctpq") == 0;
    if (X.Bind.Def.iCalcType != (canonical ? cTPQ : TPQCalc)) return 7;
    for (int i = 0; i < NumAve; i++) {
      if (MakeIniVec(i, &(X.Bind)) != 0) return 7;
      char name[64];
      snprintf(name, sizeof name, "ini_rand%d.bin", i);
      char path[4200];
      snprintf(path, sizeof path, "%s/%s", outdir, name);
      FILE *f = fopen(path, "wb");
      if (!f) return 4;
      size_t w = fwrite(&global_1st_norm, sizeof(double), 1, f);
      w += fwrite(v1 + 1, sizeof(double complex), dim, f);
      fclose(f);
      if (w != dim + 1) return 4;
    }
    if (canonical) return CalcByCanonicalTPQ(NumAve, X.Bind.Def.Param.ExpecInterval, &X) == TRUE ? 0 : 7;
    return CalcByTPQ(NumAve, X.Bind.Def.Param.ExpecInterval, &X) == TRUE ? 0 : 7;
  }

  if (strcmp(mode, "sums_seeds") == 0) {
    /* one mltply per seed s = seed .. seed+reps-1 on the synthetic vector of s:
     * sums_<s>.txt as mode sums, prdct_<s>.txt = Re Im of prdct */
    for (int r = 0; r < reps; r++) {
      const uint64_t s = seed + (uint64_t)r;
#pragma omp parallel for
      for (unsigned long j = 1; j <= dim; j++) {
        v1[j] = (2.0 * unit_rng(s, 2 * (j - 1)) - 1.0) + I * (2.0 * unit_rng(s, 2 * (j - 1) + 1) - 1.0);
        v0[j] = 0.0;
      }
      if (mltply(&(X.Bind), v0, v1) != 0) return 5;
      char path[4200];
      snprintf(path, sizeof path, "%s/sums_%lu.txt", outdir, (unsigned long)s);
      if (write_sums(path, v0, dim)) return 4;
      snprintf(path, sizeof path, "%s/prdct_%lu.txt", outdir, (unsigned long)s);
      FILE *f = fopen(path, "w");
      if (!f) return 4;
      fprintf(f, "%.17g %.17g\n", creal(X.Bind.Large.prdct), cimag(X.Bind.Large.prdct));
      fclose(f);
    }
    return 0;
  }

  /* synthetic input, element j (1-based) from counter 2(j-1), 2(j-1)+1 */
#pragma omp parallel for
  for (unsigned long j = 1; j <= dim; j++) {
    v1[j] = (2.0 * unit_rng(seed, 2 * (j - 1)) - 1.0) + I * (2.0 * unit_rng(seed, 2 * (j - 1) + 1) - 1.0);
    v0[j] = 0.0;
  }

  if (strcmp(mode, "green") == 0) {
    double n2 = 0.0;
    for (unsigned long j = 1; j <= dim; j++) n2 += creal(conj(v1[j]) * v1[j]);
    for (unsigned long j = 1; j <= dim; j++) {
      v0[j] = v1[j] / sqrt(n2);
      v1[j] = 0.0;
    }
    X.Bind.Def.St = 0;
    if (expec_energy_flct(&(X.Bind)) != 0) return 6;
    char path[4200];
    snprintf(path, sizeof path, "%s/phys.txt", outdir);
    FILE *f = fopen(path, "w");
    fprintf(f, "energy %.17g\nvar %.17g\ndoublon %.17g\ndoublon2 %.17g\nnum %.17g\nnum2 %.17g\nSz %.17g\n"
               "Sz2 %.17g\n",
            X.Bind.Phys.energy, X.Bind.Phys.var, X.Bind.Phys.doublon, X.Bind.Phys.doublon2,
            X.Bind.Phys.num, X.Bind.Phys.num2, X.Bind.Phys.Sz, X.Bind.Phys.Sz2);
    fclose(f);
    if (expec_cisajs(&(X.Bind), v1) != 0) return 6;
    if (expec_cisajscktaltdc(&(X.Bind), v1) != 0) return 6;
    return 0;
  }

  if (strcmp(mode, "lanczos") == 0) {
    double n2 = 0.0;
#pragma omp parallel for reduction(+ : n2)
    for (unsigned long j = 1; j <= dim; j++) n2 += creal(conj(v1[j]) * v1[j]);
    const double inv = 1.0 / sqrt(n2);
#pragma omp parallel for
    for (unsigned long j = 1; j <= dim; j++) v1[j] *= inv;
    char path[4200];
    snprintf(path, sizeof path, "%s/ab.txt", outdir);
    FILE *f = fopen(path, "w");
    if (!f) return 4;
    double alpha1 = 0.0, beta1 = 0.0;
    for (int stp = 1; stp <= reps; stp++) {
      if (stp > 1) {
#pragma omp parallel for
        for (unsigned long i = 1; i <= dim; i++) {
          double complex temp1 = v1[i];
          double complex temp2 = (v0[i] - alpha1 * v1[i]) / beta1;
          v0[i] = -beta1 * temp1;
          v1[i] = temp2;
        }
      }
      t0 = now();
      if (mltply(&(X.Bind), v0, v1) != 0) return 5;
      const double dt = now() - t0;
      alpha1 = creal(X.Bind.Large.prdct);
      double complex cbeta1 = 0.0;
#pragma omp parallel for reduction(+ : cbeta1)
      for (unsigned long i = 1; i <= dim; i++) cbeta1 += conj(v0[i] - alpha1 * v1[i]) * (v0[i] - alpha1 * v1[i]);
      beta1 = sqrt(creal(cbeta1));
      fprintf(f, "%d %.17g %.17g %.3f\n", stp, alpha1, beta1, dt);
      fflush(f);
    }
    fclose(f);
    return 0;
  }

  if (strcmp(mode, "bench") == 0) {
    double best = 1e300, tot = 0.0;
    for (int r = 0; r < reps; r++) {
#pragma omp parallel for
      for (unsigned long j = 1; j <= dim; j++) v0[j] = 0.0;
      t0 = now();
      mltply(&(X.Bind), v0, v1);
      double dt = now() - t0;
      tot += dt;
      if (dt < best) best = dt;
    }
    printf("{\"dim\": %lu, \"threads\": %d, \"reps\": %d, \"hv_s_best\": %.6f, \"hv_s_mean\": %.6f, "
           "\"sz_s\": %.6f, \"diag_s\": %.6f}\n",
           dim, nthreads, reps, best, tot / reps, t_sz, t_diag);
    return 0;
  }

  t0 = now();
  if (mltply(&(X.Bind), v0, v1) != 0) return 5;
  double t_hv = now() - t0;
  double complex pr = X.Bind.Large.prdct;

  char path[4200];
  FILE *f;
  if (strcmp(mode, "hv") == 0) {
    if (dump(outdir, "v0.bin", v0 + 1, dim * sizeof(double complex))) return 4;
  } else if (strcmp(mode, "sums") == 0) {
    snprintf(path, sizeof path, "%s/sums.txt", outdir);
    if (write_sums(path, v0, dim)) return 4;
  } else {
    fprintf(stderr, "unknown mode %s\n", mode);
    return 2;
  }
  snprintf(path, sizeof path, "%s/prdct.txt", outdir);
  f = fopen(path, "w");
  fprintf(f, "%.17g %.17g\n%lu\n%.6f %.6f %.6f\n", creal(pr), cimag(pr), dim, t_sz, t_diag, t_hv);
  fclose(f);
  return 0;
}
