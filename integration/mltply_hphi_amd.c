/*
 * mltply_hphi_amd.c — the reference-side binding of libhphi_amd.
 *
 * Compiled inside the HPhi source tree in place of src/mltply.c, against the
 * reference's own headers (src/include).  It keeps the reference's symbol
 *
 *   int mltply(struct BindStruct *X, double complex *tmp_v0, double complex *tmp_v1)
 *
 * (src/include/mltply.h:22, src/mltply.c:56) and adds one set-up call,
 * hphi_amd_setup(X), that takes the place of the sz() and diagonalcalc()
 * calls of the start-up sequence (src/HPhiMain.c:278,288): it builds the
 * GPU engine from the term tables the reference's reader left in X->Def and
 * fills the reference's global list_1 / list_2_1 / list_2_2 / list_Diagonal
 * bit-exactly, so that every caller that reads them (expec_*, the output
 * writers) works unchanged.  CalcByLanczos, CalcByTPQ, CalcByCanonicalTPQ
 * and CalcByLOBPCG then call mltply unchanged; host vectors are staged through
 * the device on each call (hphi_mltply).
 *
 * Build (in the reference tree, after the library is built):
 *   cc -c -I<HPhi>/src/include -I<repo>/include mltply_hphi_amd.c
 *   link the HPhi objects without mltply.o, plus
 *   -L<repo>/hphi_amd -lhphi_amd -Wl,-rpath,<repo>/hphi_amd
 */
#include <complex.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "Common.h"
#include "global.h"
#include "mltply.h"
#include "mltplyCommon.h"

#include "hphi_amd.h"

static hphi_engine *g_eng = NULL;

/* Host copies of the term tables in the descriptor's flat layout; they live
 * as long as the descriptor that points into them. */
struct hphi_amd_desc_buf {
  int *ex, *pl, *tr, *ci, *cv, *hd, *ia, *iad, *ph;
  double *trt, *iav;
};

void hphi_amd_desc_buf_free(struct hphi_amd_desc_buf *b) {
  free(b->ex); free(b->pl); free(b->tr); free(b->ci); free(b->cv); free(b->hd);
  free(b->ia); free(b->iad); free(b->ph); free(b->trt); free(b->iav);
  memset(b, 0, sizeof *b);
}

static void *alloc0(size_t n, size_t sz) { return calloc(n ? n : 1, sz); }

/* DefineList (src/include/struct.h:43-310), after ReadDefFileIdxPara and
 * HPhiTrans, -> hphi_model_desc.  Every field of the descriptor is set; the
 * tables are copied into b.  Returns 0, or -1 when a copy cannot be made. */
int hphi_amd_fill_desc(const struct DefineList *D, hphi_model_desc *d, struct hphi_amd_desc_buf *b) {
  memset(d, 0, sizeof *d);
  memset(b, 0, sizeof *b);
  d->model = D->iCalcModel;
  d->nsite = (int)D->Nsite;
  d->nup = (int)D->Nup;
  d->ndown = (int)D->Ndown;

  const int nex = (int)D->NExchangeCoupling, npl = (int)D->NPairLiftCoupling, ntr = (int)D->EDNTransfer;
  const int nci = (int)D->NCoulombIntra, ncv = (int)D->NCoulombInter, nhd = (int)D->NHundCoupling;
  const int nia = (int)D->NInterAll_OffDiagonal, niad = (int)D->NInterAll_Diagonal, nph = (int)D->NPairHopping;
  b->ex = alloc0(2 * (size_t)nex, sizeof(int));
  b->pl = alloc0(2 * (size_t)npl, sizeof(int));
  b->tr = alloc0(4 * (size_t)ntr, sizeof(int));
  b->trt = alloc0(2 * (size_t)ntr, sizeof(double));
  b->ci = alloc0((size_t)nci, sizeof(int));
  b->cv = alloc0(2 * (size_t)ncv, sizeof(int));
  b->hd = alloc0(2 * (size_t)nhd, sizeof(int));
  b->ia = alloc0(8 * (size_t)nia, sizeof(int));
  b->iav = alloc0(2 * (size_t)nia, sizeof(double));
  b->iad = alloc0(4 * (size_t)niad, sizeof(int));
  b->ph = alloc0(2 * (size_t)nph, sizeof(int));
  if (!b->ex || !b->pl || !b->tr || !b->trt || !b->ci || !b->cv || !b->hd || !b->ia || !b->iav || !b->iad ||
      !b->ph) {
    hphi_amd_desc_buf_free(b);
    return -1;
  }
  for (int i = 0; i < nex; i++) {
    b->ex[2 * i] = D->ExchangeCoupling[i][0];
    b->ex[2 * i + 1] = D->ExchangeCoupling[i][1];
  }
  d->n_exchange = nex; d->exchange_ij = b->ex; d->exchange_J = D->ParaExchangeCoupling;
  for (int i = 0; i < npl; i++) {
    b->pl[2 * i] = D->PairLiftCoupling[i][0];
    b->pl[2 * i + 1] = D->PairLiftCoupling[i][1];
  }
  d->n_pairlift = npl; d->pairlift_ij = b->pl; d->pairlift_J = D->ParaPairLiftCoupling;
  for (int i = 0; i < ntr; i++) {
    for (int k = 0; k < 4; k++) b->tr[4 * i + k] = D->EDGeneralTransfer[i][k];
    b->trt[2 * i] = creal(D->EDParaGeneralTransfer[i]);
    b->trt[2 * i + 1] = cimag(D->EDParaGeneralTransfer[i]);
  }
  d->n_transfer = ntr; d->transfer_idx = b->tr; d->transfer_t = b->trt;
  d->n_chemi = (int)D->EDNChemi;
  d->chemi_site = D->EDChemi; d->chemi_spin = D->EDSpinChemi; d->chemi_mu = D->EDParaChemi;
  for (int i = 0; i < nci; i++) b->ci[i] = D->CoulombIntra[i][0];
  d->n_intra = nci; d->intra_site = b->ci; d->intra_U = D->ParaCoulombIntra;
  for (int i = 0; i < ncv; i++) {
    b->cv[2 * i] = D->CoulombInter[i][0];
    b->cv[2 * i + 1] = D->CoulombInter[i][1];
  }
  d->n_inter = ncv; d->inter_ij = b->cv; d->inter_V = D->ParaCoulombInter;
  for (int i = 0; i < nhd; i++) {
    b->hd[2 * i] = D->HundCoupling[i][0];
    b->hd[2 * i + 1] = D->HundCoupling[i][1];
  }
  d->n_hund = nhd; d->hund_ij = b->hd; d->hund_J = D->ParaHundCoupling;
  /* two-body terms as GetDiagonalInterAll left them (readdef.c:2415-2530) */
  for (int i = 0; i < nia; i++) {
    for (int k = 0; k < 8; k++) b->ia[8 * i + k] = D->InterAll_OffDiagonal[i][k];
    b->iav[2 * i] = creal(D->ParaInterAll_OffDiagonal[i]);
    b->iav[2 * i + 1] = cimag(D->ParaInterAll_OffDiagonal[i]);
  }
  d->n_interall = nia; d->interall_idx = b->ia; d->interall_V = b->iav;
  for (int i = 0; i < niad; i++)
    for (int k = 0; k < 4; k++) b->iad[4 * i + k] = D->InterAll_Diagonal[i][k];
  d->n_interall_diag = niad; d->interall_diag_idx = b->iad; d->interall_diag_V = D->ParaInterAll_Diagonal;
  for (int i = 0; i < nph; i++) {
    b->ph[2 * i] = D->PairHopping[i][0];
    b->ph[2 * i + 1] = D->PairHopping[i][1];
  }
  d->n_pairhop = nph; d->pairhop_ij = b->ph; d->pairhop_J = D->ParaPairHopping;
  return 0;
}

/* In place of sz() and diagonalcalc() (src/HPhiMain.c:278,288), after
 * setmem_large(): builds the engine, then list_1 / list_2_1 / list_2_2
 * (sz.c:62) and list_Diagonal (diagonalcalc.c:84) on the device and copies
 * them into the reference's globals. */
int hphi_amd_setup(struct BindStruct *X) {
  if (X->Def.iFlgGeneralSpin) {
    fprintf(stderr, "hphi_amd: general spin (S > 1/2) is not supported\n");
    return -1;
  }
  hphi_model_desc d;
  struct hphi_amd_desc_buf b;
  if (hphi_amd_fill_desc(&X->Def, &d, &b)) {
    fprintf(stderr, "hphi_amd: host allocation failed\n");
    return -1;
  }
  if (g_eng) hphi_engine_destroy(g_eng);
  g_eng = NULL;
  int rc = hphi_engine_create(&d, &g_eng);
  hphi_amd_desc_buf_free(&b); /* the engine keeps its own compiled tables */
  if (rc) {
    fprintf(stderr, "hphi_amd: %s\n", hphi_last_error());
    return -1;
  }
  uint64_t dim = 0, s21 = 0, s22 = 0;
  hphi_engine_info(g_eng, &dim, &s21, &s22, NULL, NULL, NULL);
  if (dim != (uint64_t)X->Check.idim_max) {
    fprintf(stderr, "hphi_amd: sector dimension %llu, the reference's check() gave %lu\n",
            (unsigned long long)dim, X->Check.idim_max);
    return -1;
  }
  const int canon = X->Def.iCalcModel == Spin || X->Def.iCalcModel == Hubbard;
  if (canon && (s21 > (uint64_t)X->Large.SizeOflist_2_1 || s22 > (uint64_t)X->Large.SizeOflist_2_2)) {
    fprintf(stderr, "hphi_amd: list_2_1/list_2_2 sizes exceed setmem_large's\n");
    return -1;
  }
  if (hphi_sz(g_eng, canon ? (uint64_t *)list_1 : NULL, canon ? (uint64_t *)list_2_1 : NULL,
              canon ? (uint64_t *)list_2_2 : NULL) ||
      hphi_diagonalcalc(g_eng, list_Diagonal)) {
    fprintf(stderr, "hphi_amd: %s\n", hphi_last_error());
    return -1;
  }
  return 0;
}

void hphi_amd_teardown(void) {
  if (g_eng) hphi_engine_destroy(g_eng);
  g_eng = NULL;
}

/* src/mltply.c:56 replacement: tmp_v0[1..i_max] += H tmp_v1[1..i_max], the same
 * side effects on X->Large (:90-94, :129), 0 or -1 as there. */
int mltply(struct BindStruct *X, double complex *tmp_v0, double complex *tmp_v1) {
  X->Large.prdct = 0.0;
  if (!g_eng) {
    fprintf(stderr, "hphi_amd: mltply before hphi_amd_setup\n");
    return -1;
  }
  if (X->Def.iCalcType == TimeEvolution) {
    fprintf(stderr, "hphi_amd: time evolution is not supported\n");
    return -1;
  }
  uint64_t dim = 0, irght = 0, ilft = 0, ihfbit = 0;
  hphi_engine_info(g_eng, &dim, NULL, NULL, &irght, &ilft, &ihfbit);
  if (dim != (uint64_t)X->Check.idim_max) {
    fprintf(stderr, "hphi_amd: X->Check.idim_max changed since hphi_amd_setup\n");
    return -1;
  }
  X->Large.i_max = X->Check.idim_max;
  X->Large.irght = dim ? irght : 0;
  X->Large.ilft = dim ? ilft : 0;
  X->Large.ihfbit = dim ? ihfbit : 0;
  X->Large.mode = M_MLTPLY;
  if (dim == 0) return 0;
  double pr[2];
  if (hphi_mltply(g_eng, (double *)tmp_v0, (const double *)tmp_v1, pr)) {
    fprintf(stderr, "hphi_amd: %s\n", hphi_last_error());
    return -1;
  }
  X->Large.prdct = pr[0] + I * pr[1];
  return 0;
}
