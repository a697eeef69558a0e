/*
 * CPU oracle for the H·v engine — TEST INFRASTRUCTURE ONLY (see hphi_oracle.h).
 *
 * Restates, in plain C + OpenMP, the reference algorithm of:
 *   basis build      src/sz.c:276-399,636-734,993-1094,1495-1546, src/check.c:91-413,
 *                    src/xsetmem.c:284-318, src/bitcalc.c:45-62,473-480
 *   diagonal         src/diagonalcalc.c:84-187,260-1100
 *   H·v              src/mltply.c:56-132, src/mltplySpin.c:192-293,500-703,954-991,
 *                    src/mltplySpinCore.c:35-56,203-227,273-382,
 *                    src/mltplyHubbard.c:150-209,582-602, src/mltplyHubbardCore.c:34-69,305-347,
 *                    src/bitcalc.c:166-223,338-352
 * It keeps the reference's structure (one OpenMP sweep over all states per
 * term, scatter into v0, two-level index lookup, 1-indexed arrays) so that it
 * also serves as the CPU baseline of the benchmark.
 */
#include "hphi_oracle.h"

#include <complex.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_HUBBARD 0
#define OR_SPIN 1
#define OR_SPINGC 4

int or_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* Binomial coefficient, 0 outside 0<=k<=n (src/sz.c:848-880). */
static uint64_t binom(int n, int k) {
  if (k < 0 || n < 0 || k > n) return 0;
  uint64_t r = 1;
  for (int i = 1; i <= k; i++) r = r * (uint64_t)(n - k + i) / (uint64_t)i;
  return r;
}

/* next larger integer with the same popcount (src/bitcalc.c:473-480) */
static uint64_t snoob(uint64_t x) {
  uint64_t smallest = x & (~x + 1);
  uint64_t ripple = x + smallest;
  uint64_t ones = x ^ ripple;
  ones = (ones >> 2) / smallest;
  return ripple | ones;
}

static int popc(uint64_t x) { return __builtin_popcountll(x); }

/* src/bitcalc.c:45-62 (GetSplitBit) with n = Nsite or 2*Nsite (bitcalc.c:78-108) */
static void split_bits(int n, uint64_t *irght, uint64_t *ilft, uint64_t *ihfbit) {
  *ihfbit = 1ULL << ((n + 1) / 2);
  *irght = *ihfbit - 1;
  *ilft = (((uint64_t)1 << n) - 1) ^ *irght;
}

int or_dims(const or_model *m, uint64_t *dim, uint64_t *size21, uint64_t *size22,
            uint64_t *irght, uint64_t *ilft, uint64_t *ihfbit) {
  int N = m->nsite;
  uint64_t sdim = 1;
  switch (m->model) {
    case OR_SPIN: /* check.c:177, check.c:310-316, xsetmem.c:294-299 */
      *dim = binom(N, m->nup);
      sdim = 1ULL << (N / 2);
      *size21 = (N % 2 == 1) ? sdim * 2 + 2 : sdim + 2;
      *size22 = sdim + 2;
      split_bits(N, irght, ilft, ihfbit);
      return 0;
    case OR_HUBBARD: /* check.c:114-118, check.c:301-307 */
      *dim = binom(N, m->nup) * binom(N, m->ndown);
      sdim = 1ULL << N;
      *size21 = sdim + 2;
      *size22 = sdim + 2;
      split_bits(2 * N, irght, ilft, ihfbit);
      return 0;
    case OR_SPINGC: /* check.c:99-110: 2^N, no tables */
      *dim = 1ULL << N;
      *size21 = *size22 = 0;
      split_bits(N, irght, ilft, ihfbit);
      return 0;
  }
  return -1;
}

/* ---------------------------------------------------------------- sz ---- */

/* Spin canonical, CalcHS=1 path: sz.c:684-703 then omp_sz_spin_hacker
 * (sz.c:1495-1546). */
static int sz_spin(const or_model *m, uint64_t *list_1, uint64_t *list_2_1, uint64_t *list_2_2) {
  int N = m->nsite, Ne = m->nup;
  uint64_t irght, ilft, ihfbit;
  split_bits(N, &irght, &ilft, &ihfbit);
  uint64_t sdim = 1ULL << (N / 2);
  int all_up = (N + 1) / 2;
  uint64_t *list_jb = (uint64_t *)calloc(sdim + 2, sizeof(uint64_t));
  uint64_t jb = 0;
  for (uint64_t ib = 0; ib < sdim; ib++) { /* serial block prefix, sz.c:686-698 */
    list_jb[ib] = jb;
    int num_up = popc(ib * ihfbit);
    jb += binom(all_up, Ne - num_up);
  }
  uint64_t icnt = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : icnt)
  for (uint64_t ib = 0; ib < sdim; ib++) {
    uint64_t jbb = list_jb[ib], ja = 1;
    int num_up = popc(ib * ihfbit);
    if (num_up <= Ne && (Ne - num_up) <= N - 1) {
      uint64_t ia = (1ULL << (Ne - num_up)) - 1;
      if (ia < ihfbit) {
        list_1[ja + jbb] = ia + ib * ihfbit;
        list_2_1[ia] = ja + 1;
        list_2_2[ib] = jbb + 1;
        ja++;
        if (ia != 0) {
          ia = snoob(ia);
          while (ia < ihfbit) {
            list_1[ja + jbb] = ia + ib * ihfbit;
            list_2_1[ia] = ja + 1;
            list_2_2[ib] = jbb + 1;
            ja++;
            ia = snoob(ia);
          }
        }
      }
    }
    icnt += ja - 1;
  }
  free(list_jb);
  return (icnt == binom(N, Ne)) ? 0 : -1; /* sz.c:810-820 */
}

/* Hubbard canonical, CalcHS=1 path: sz.c:316-394 then omp_sz_hacker
 * (sz.c:993-1094).  Up spin of site i is bit 2i, down spin bit 2i+1. */
static void count_updown(uint64_t x, int nsite, int *nu, int *nd) {
  uint64_t even = 0x5555555555555555ULL, odd = 0xAAAAAAAAAAAAAAAAULL;
  uint64_t mask = (2 * nsite >= 64) ? ~0ULL : ((1ULL << (2 * nsite)) - 1);
  *nu = popc(x & even & mask);
  *nd = popc(x & odd & mask);
}

static int sz_hubbard(const or_model *m, uint64_t *list_1, uint64_t *list_2_1, uint64_t *list_2_2) {
  int N = m->nsite, Nup = m->nup, Ndn = m->ndown;
  uint64_t irght, ilft, ihfbit;
  split_bits(2 * N, &irght, &ilft, &ihfbit);
  uint64_t sdim = 1ULL << N;
  int tmp_res = N % 2, all_up = (N + tmp_res) / 2, all_down = (N - tmp_res) / 2;
  uint64_t *list_jb = (uint64_t *)calloc(sdim + 2, sizeof(uint64_t));
  uint64_t jb = 0;
  for (uint64_t ib = 0; ib < sdim; ib++) {
    list_jb[ib] = jb;
    int nu, nd;
    count_updown(ib * ihfbit, N, &nu, &nd);
    jb += binom(all_up, Nup - nu) * binom(all_down, Ndn - nd);
  }
  uint64_t icnt = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : icnt)
  for (uint64_t ib = 0; ib < sdim; ib++) {
    uint64_t jbb = list_jb[ib], ja = 1;
    int nu0, nd0;
    count_updown(ib * ihfbit, N, &nu0, &nd0);
    if (nu0 <= Nup && nd0 <= Ndn) {
      uint64_t ia = (1ULL << (Nup + Ndn - nu0 - nd0)) - 1;
      if (ia < sdim) {
        for (;;) {
          int nu, nd;
          count_updown(ia, N, &nu, &nd);
          if (nu0 + nu == Nup && nd0 + nd == Ndn) {
            list_1[ja + jbb] = ia + ib * ihfbit;
            list_2_1[ia] = ja + 1;
            list_2_2[ib] = jbb + 1;
            ja++;
          }
          if (ia == 0) break;
          ia = snoob(ia);
          if (ia >= sdim) break;
        }
      }
    }
    icnt += ja - 1;
  }
  free(list_jb);
  return (icnt == binom(N, Nup) * binom(N, Ndn)) ? 0 : -1;
}

int or_sz(const or_model *m, uint64_t *list_1, uint64_t *list_2_1, uint64_t *list_2_2) {
  if (m->model == OR_SPIN) return sz_spin(m, list_1, list_2_1, list_2_2);
  if (m->model == OR_HUBBARD) return sz_hubbard(m, list_1, list_2_1, list_2_2);
  if (m->model == OR_SPINGC) return 0;
  return -1;
}

/* ---------------------------------------------------------- diagonal ---- */

/* diagonalcalc.c:84-187: one pass per term, in the order CoulombIntra, Chemi,
 * CoulombInter, Hund (InterAll diagonal is not part of this oracle). */
int or_diagonalcalc(const or_model *m, const uint64_t *list_1, uint64_t dim, double *D) {
  int gc = (m->model == OR_SPINGC);
#define STATE(j) (gc ? (uint64_t)((j)-1) : list_1[(j)])
#pragma omp parallel for
  for (uint64_t j = 1; j <= dim; j++) D[j] = 0.0;

  for (int t = 0; t < m->n_intra; t++) { /* SetDiagonalCoulombIntra, :260-355 */
    if (m->model != OR_HUBBARD) break;
    int s = m->intra_site[t];
    double V = m->intra_U[t];
    uint64_t is = (1ULL << (2 * s)) | (1ULL << (2 * s + 1));
#pragma omp parallel for
    for (uint64_t j = 1; j <= dim; j++)
      if ((list_1[j] & is) == is) D[j] += V;
  }
  for (int t = 0; t < m->n_ch; t++) { /* SetDiagonalChemi, :373-530 */
    int s = m->ch_site[t], spin = m->ch_spin[t];
    double V = -m->ch_mu[t];
    if (m->model == OR_HUBBARD) {
      uint64_t is1 = 1ULL << (2 * s + spin);
#pragma omp parallel for
      for (uint64_t j = 1; j <= dim; j++) {
        uint64_t num1 = (list_1[j] & is1) / is1;
        D[j] += num1 * V;
      }
    } else {
      uint64_t is1 = 1ULL << s;
#pragma omp parallel for
      for (uint64_t j = 1; j <= dim; j++) {
        uint64_t b = ((STATE(j) & is1) / is1) ^ (uint64_t)(1 - spin);
        D[j] += V * b;
      }
    }
  }
  for (int t = 0; t < m->n_inter; t++) { /* SetDiagonalCoulombInter, :543-770 */
    int s1 = m->inter_ij[2 * t], s2 = m->inter_ij[2 * t + 1];
    double V = m->inter_V[t];
    if (m->model == OR_HUBBARD) {
      uint64_t u1 = 1ULL << (2 * s1), d1 = 1ULL << (2 * s1 + 1);
      uint64_t u2 = 1ULL << (2 * s2), d2 = 1ULL << (2 * s2 + 1);
#pragma omp parallel for
      for (uint64_t j = 1; j <= dim; j++) {
        uint64_t x = list_1[j];
        uint64_t n1 = (x & u1) / u1 + (x & d1) / d1;
        uint64_t n2 = (x & u2) / u2 + (x & d2) / d2;
        D[j] += n1 * n2 * V;
      }
    } else {
#pragma omp parallel for
      for (uint64_t j = 1; j <= dim; j++) D[j] += V;
    }
  }
  for (int t = 0; t < m->n_hund; t++) { /* SetDiagonalHund, :788-1100 */
    int s1 = m->hund_ij[2 * t], s2 = m->hund_ij[2 * t + 1];
    double V = -m->hund_J[t];
    if (m->model == OR_HUBBARD) {
      uint64_t u1 = 1ULL << (2 * s1), d1 = 1ULL << (2 * s1 + 1);
      uint64_t u2 = 1ULL << (2 * s2), d2 = 1ULL << (2 * s2 + 1);
#pragma omp parallel for
      for (uint64_t j = 1; j <= dim; j++) {
        uint64_t x = list_1[j];
        uint64_t a = (x & u1) / u1, b = (x & d1) / d1, c = (x & u2) / u2, d = (x & d2) / d2;
        D[j] += V * (a * c + b * d);
      }
    } else {
      uint64_t is = (1ULL << s1) | (1ULL << s2);
#pragma omp parallel for
      for (uint64_t j = 1; j <= dim; j++) {
        uint64_t ib = STATE(j) & is;
        if (ib == 0 || ib == is) D[j] += V;
      }
    }
  }
#undef STATE
  return 0;
}

/* ----------------------------------------------------------- mltply ---- */

/* GetOffComp, bitcalc.c:195-223: 1-based position of state x, 0 if absent. */
static inline uint64_t off_comp(const uint64_t *l21, const uint64_t *l22, uint64_t x,
                                uint64_t irght, uint64_t ilft, uint64_t ihfbit) {
  uint64_t ia = x & irght, ib = (x & ilft) / ihfbit;
  if (l21[ia] * l22[ib] == 0) return 0;
  return l21[ia] - 1 + l22[ib] - 1;
}

/* parity sign, bitcalc.c:338-352 */
static inline int sgn_bit(uint64_t b) { return 1 - 2 * (popc(b) & 1); }

int or_mltply(const or_model *m, const uint64_t *list_1, const uint64_t *l21, const uint64_t *l22,
              const double *D, uint64_t dim, double *v0d, const double *v1d, double *prdct,
              uint64_t jlo, uint64_t jhi) {
  double complex *v0 = (double complex *)v0d;
  const double complex *v1 = (const double complex *)v1d;
  uint64_t irght, ilft, ihfbit, s21, s22, dd;
  if (or_dims(m, &dd, &s21, &s22, &irght, &ilft, &ihfbit) != 0 || dd != dim) return -1;
  if (jlo < 1) jlo = 1;
  if (jhi > dim) jhi = dim;
  double pr_re = 0.0, pr_im = 0.0;

  /* diagonal, mltply.c:96-102 */
  {
    double complex dam = 0.0;
    double dre = 0.0, dim_ = 0.0;
#pragma omp parallel for reduction(+ : dre, dim_) private(dam)
    for (uint64_t j = jlo; j <= jhi; j++) {
      v0[j] += D[j] * v1[j];
      dam = D[j] * conj(v1[j]) * v1[j];
      dre += creal(dam);
      dim_ += cimag(dam);
    }
    pr_re += dre;
    pr_im += dim_;
  }

  if (m->model == OR_SPIN) {
    /* Exchange loop, mltplySpin.c:254-288 -> exchange_spin (:954-969)
     * -> exchange_spin_element (mltplySpinCore.c:273-307) */
    for (int t = 0; t < m->n_ex; t++) {
      uint64_t is = (1ULL << m->ex_ij[2 * t]) + (1ULL << m->ex_ij[2 * t + 1]);
      double J = m->ex_J[t];
      double dre = 0.0, dim_ = 0.0;
#pragma omp parallel for reduction(+ : dre, dim_)
      for (uint64_t j = jlo; j <= jhi; j++) {
        uint64_t ibit = list_1[j] & is;
        if (ibit == 0 || ibit == is) continue;
        uint64_t off = off_comp(l21, l22, list_1[j] ^ is, irght, ilft, ihfbit);
        double complex dmv = J * v1[j];
        v0[off] += dmv;
        double complex d = dmv * conj(v1[off]);
        dre += creal(d);
        dim_ += cimag(d);
      }
      pr_re += dre;
      pr_im += dim_;
    }
  } else if (m->model == OR_SPINGC) {
    /* transverse Trans pairs, mltplySpin.c:527-578 -> child_SpinGC_CisAit
     * (mltplySpinCore.c:203-227) */
    for (int t = 0; t < m->n_tr; t++) {
      int s1 = m->tr_idx[4 * t], sig1 = m->tr_idx[4 * t + 1];
      int sig2 = m->tr_idx[4 * t + 3];
      if (sig1 == sig2) continue; /* "Transverse_OffDiagonal component is illegal" */
      double complex tt = -(m->tr_t[2 * t] + I * m->tr_t[2 * t + 1]);
      uint64_t is1 = 1ULL << s1;
      double dre = 0.0, dim_ = 0.0;
#pragma omp parallel for reduction(+ : dre, dim_)
      for (uint64_t j = jlo; j <= jhi; j++) {
        uint64_t x = j - 1, b = x & is1, off;
        if (b == 0 && sig2 == 0) off = x + is1;
        else if (b != 0 && sig2 == 1) off = x - is1;
        else continue;
        v0[off + 1] += v1[j] * tt;
        double complex d = tt * conj(v1[off + 1]) * v1[j];
        dre += creal(d);
        dim_ += cimag(d);
      }
      pr_re += dre;
      pr_im += dim_;
    }
    /* GC exchange, mltplySpin.c:624-658 -> GC_exchange_spin_element (mltplySpinCore.c:314-345) */
    for (int t = 0; t < m->n_ex; t++) {
      uint64_t is = (1ULL << m->ex_ij[2 * t]) + (1ULL << m->ex_ij[2 * t + 1]);
      double J = m->ex_J[t];
      double dre = 0.0, dim_ = 0.0;
#pragma omp parallel for reduction(+ : dre, dim_)
      for (uint64_t j = jlo; j <= jhi; j++) {
        uint64_t x = j - 1, ibit = x & is;
        if (ibit == 0 || ibit == is) continue;
        uint64_t off = x ^ is;
        double complex dmv = J * v1[j];
        v0[off + 1] += dmv;
        double complex d = dmv * conj(v1[off + 1]);
        dre += creal(d);
        dim_ += cimag(d);
      }
      pr_re += dre;
      pr_im += dim_;
    }
    /* GC pair lift, mltplySpin.c:664-698 -> GC_pairlift_spin_element (mltplySpinCore.c:352-382) */
    for (int t = 0; t < m->n_pl; t++) {
      uint64_t is = (1ULL << m->pl_ij[2 * t]) + (1ULL << m->pl_ij[2 * t + 1]);
      double J = m->pl_J[t];
      double dre = 0.0, dim_ = 0.0;
#pragma omp parallel for reduction(+ : dre, dim_)
      for (uint64_t j = jlo; j <= jhi; j++) {
        uint64_t x = j - 1, ibit = x & is;
        if (!(ibit == 0 || ibit == is)) continue;
        uint64_t off = x ^ is;
        double complex dmv = J * v1[j];
        v0[off + 1] += dmv;
        double complex d = dmv * conj(v1[off + 1]);
        dre += creal(d);
        dim_ += cimag(d);
      }
      pr_re += dre;
      pr_im += dim_;
    }
  } else if (m->model == OR_HUBBARD) {
    /* Transfer loop, mltplyHubbard.c:174-209: general_hopp_GetInfo
     * (mltplyHubbardCore.c:34-69), general_hopp (mltplyHubbard.c:582-602),
     * CisAjt (mltplyHubbardCore.c:305-347). */
    for (int t = 0; t < m->n_tr; t++) {
      int i1 = m->tr_idx[4 * t] + 1, sg1 = m->tr_idx[4 * t + 1];
      int i2 = m->tr_idx[4 * t + 2] + 1, sg2 = m->tr_idx[4 * t + 3];
      double complex trans = -(m->tr_t[2 * t] + I * m->tr_t[2 * t + 1]);
      uint64_t is1 = 1ULL << (2 * i1 - 2 + sg1), is2 = 1ULL << (2 * i2 - 2 + sg2), A;
      if (i1 > i2) A = (1ULL << (2 * i1 - 2 + sg1)) - (1ULL << (2 * i2 - 1 + sg2));
      else if (i1 < i2) A = (1ULL << (2 * i2 - 2 + sg2)) - (1ULL << (2 * i1 - 1 + sg1));
      else if (sg1 > sg2) A = (1ULL << (2 * i1 - 2 + sg1)) - (1ULL << (2 * i2 - 1 + sg2));
      else A = (1ULL << (2 * i2 - 2 + sg2)) - (1ULL << (2 * i1 - 1 + sg1));
      uint64_t isA = is1 + is2;
      double dre = 0.0, dim_ = 0.0;
#pragma omp parallel for reduction(+ : dre, dim_)
      for (uint64_t j = jlo; j <= jhi; j++) {
        uint64_t x = list_1[j];
        if ((x & is1) != 0 || (x & is2) == 0) continue;
        int sgn = sgn_bit(x & A);
        uint64_t off = off_comp(l21, l22, x ^ isA, irght, ilft, ihfbit);
        if (off == 0) continue;
        double complex dmv = sgn * v1[j];
        v0[off] += trans * dmv;
        double complex d = dmv * conj(v1[off]) * trans;
        dre += creal(d);
        dim_ += cimag(d);
      }
      pr_re += dre;
      pr_im += dim_;
    }
  } else {
    return -1;
  }
  prdct[0] = pr_re;
  prdct[1] = pr_im;
  return 0;
}

/* Number of off-diagonal (term, source) pairs or_mltply processes for sources
 * jlo..jhi: the NNZ that the algorithmic-byte count of a sweep uses. */
uint64_t or_count_nnz(const or_model *m, const uint64_t *list_1, uint64_t jlo, uint64_t jhi) {
  uint64_t n = 0;
  if (m->model == OR_SPIN) {
    for (int t = 0; t < m->n_ex; t++) {
      uint64_t is = (1ULL << m->ex_ij[2 * t]) + (1ULL << m->ex_ij[2 * t + 1]), c = 0;
#pragma omp parallel for reduction(+ : c)
      for (uint64_t j = jlo; j <= jhi; j++) {
        uint64_t b = list_1[j] & is;
        c += (b != 0 && b != is);
      }
      n += c;
    }
    return n;
  }
  return 0; /* other models: not needed by the benchmark */
}

/* ------------------------------------------------------ synthetic input -- */

static inline uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

static inline double unit(uint64_t seed, uint64_t c) {
  uint64_t h = splitmix64((seed * 0x9E3779B97F4A7C15ULL) ^ c);
  return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

void or_fill_vector(uint64_t seed, uint64_t dim, double *v) {
#pragma omp parallel for
  for (uint64_t j = 0; j < dim; j++) {
    v[2 * (j + 1)] = 2.0 * unit(seed, 2 * j) - 1.0;
    v[2 * (j + 1) + 1] = 2.0 * unit(seed, 2 * j + 1) - 1.0;
  }
}
