/*
 * CPU oracle for the H·v engine — TEST INFRASTRUCTURE ONLY.
 *
 * A plain C + OpenMP restatement of the reference's hot path, used by tests,
 * by the smoke check and as the CPU baseline of the benchmark.  Nothing in
 * the product links or calls it.  Every function cites the reference
 * file:line it follows (paths relative to the reference tree).
 *
 * Pinning of each function (what checks it, beyond retyping):
 *   or_sz            golden tables from the compiled reference
 *                    (oracle/_ref), brute-force enumeration, closed-form
 *                    ranks/prefixes in the tests.
 *   or_diagonalcalc  golden list_Diagonal from oracle/_ref (bit-exact), and
 *                    brute-force dense diagonal in numpy.
 *   or_mltply        golden H·v and prdct from oracle/_ref, dense numpy H
 *                    built from the operator definitions, known-answer
 *                    energies of the reference tests and tutorial.
 */
#ifndef HPHI_ORACLE_H
#define HPHI_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Model description: term tables in the order the reference holds them after
 * ReadDefFileIdxPara (src/readdef.c:1006-2470).  Sites are 0-based.
 * model: 0 Hubbard, 1 Spin, 4 SpinGC (calcmod.def CalcModel). */
typedef struct {
  int model, nsite, nup, ndown;
  int n_ex;    const int *ex_ij;     const double *ex_J;      /* Exchange [n][2]          */
  int n_pl;    const int *pl_ij;     const double *pl_J;      /* PairLift [n][2] (SpinGC) */
  int n_tr;    const int *tr_idx;    const double *tr_t;      /* EDGeneralTransfer [n][4], t as [n][2] */
  int n_ch;    const int *ch_site;   const int *ch_spin;  const double *ch_mu; /* EDChemi */
  int n_intra; const int *intra_site; const double *intra_U;  /* CoulombIntra             */
  int n_inter; const int *inter_ij;  const double *inter_V;   /* CoulombInter (+Ising)    */
  int n_hund;  const int *hund_ij;   const double *hund_J;    /* Hund (+Ising)            */
} or_model;

/* Sizes: dim (idim_max), list_2_1 and list_2_2 lengths, split masks. */
int or_dims(const or_model *m, uint64_t *dim, uint64_t *size21, uint64_t *size22,
            uint64_t *irght, uint64_t *ilft, uint64_t *ihfbit);

/* Canonical basis tables, arrays shaped like the reference's
 * (list_1[dim+1], 1-indexed; list_2_1[size21]; list_2_2[size22]; zero-filled by caller). */
int or_sz(const or_model *m, uint64_t *list_1, uint64_t *list_2_1, uint64_t *list_2_2);

/* list_Diagonal[1..dim]; list_1 may be NULL for SpinGC. */
int or_diagonalcalc(const or_model *m, const uint64_t *list_1, uint64_t dim, double *list_Diagonal);

/* v0[1..dim] += H v1[1..dim]; prdct (re, im) = <v1|H|v1>.  Per-term scatter
 * sweeps over source states jlo..jhi (1-based, inclusive); the full product is
 * jlo = 1, jhi = dim.  Vectors are interleaved complex128, 1-indexed. */
int or_mltply(const or_model *m, const uint64_t *list_1, const uint64_t *list_2_1,
              const uint64_t *list_2_2, const double *list_Diagonal, uint64_t dim,
              double *v0, const double *v1, double *prdct, uint64_t jlo, uint64_t jhi);

/* Off-diagonal (term, source) pairs for sources jlo..jhi (Spin model). */
uint64_t or_count_nnz(const or_model *m, const uint64_t *list_1, uint64_t jlo, uint64_t jhi);

/* Synthetic input vector: v[1+j] = (2u(2j)-1) + i(2u(2j+1)-1), u from SplitMix64. */
void or_fill_vector(uint64_t seed, uint64_t dim, double *v);

int or_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
