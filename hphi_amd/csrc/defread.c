/*
 * Expert-mode input reader for the hphi_amd engine (host C).
 *
 * Reads namelist.def and the files it names, with the reference's formats and
 * post-processing rules (paths relative to the HPhi source tree):
 *   namelist.def    "Keyword file" pairs, '#' comments       src/readdef.c:398-460
 *   calcmod.def     "Keyword int" lines, '#'/'=' comments     src/readdef.c:190-221,233-397
 *   modpara.def     8 header lines, then "Keyword value"     src/readdef.c:535-650
 *   term files      2nd line "Name count", 5 header lines    src/readdef.c:655-760,1006-1043
 *   Trans           duplicate (i,s,j,t) entries summed        src/readdef.c:1130-1150
 *                   Hermitian pairing into EDGeneralTransfer,
 *                   diagonal entries to EDChemi                src/readdef.c:2136-2234
 *   Ising J         Hund -J/2 and CoulombInter -J/4, appended after the
 *                   explicit Hund / CoulombInter entries      src/readdef.c:1300-1327,2788-2792
 *   sector          Spin: Nup = (NLocSpn + 2Sz)/2; Hubbard: Nup = (NCond + 2Sz)/2
 *                                                             src/readdef.c:843-905
 * Supported subset: CalcModel 0 (Hubbard), 1 (Spin, S=1/2), 3 (HubbardGC),
 * 4 (SpinGC, S=1/2);
 * term files Trans, CoulombIntra, CoulombInter, Hund, Exchange, Ising, PairLift,
 * InterAll (every model), PairHop (Hubbard, HubbardGC).  Other keywords (Boost, ...)
 * are rejected with a message.
 */
#include <complex.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "hphi_amd.h"

int hphi_set_error(const char *fmt, ...);

#define MAXLINE 1024

enum { KW_CALCMOD, KW_MODPARA, KW_LOCSPIN, KW_TRANS, KW_INTRA, KW_INTER, KW_HUND, KW_EXCHANGE,
       KW_ISING, KW_PAIRLIFT, KW_PAIRHOP, KW_INTERALL, KW_N };
static const char *kw_names[KW_N] = {"CalcMod", "ModPara", "LocSpin", "Trans", "CoulombIntra",
                                     "CoulombInter", "Hund", "Exchange", "Ising", "PairLift",
                                     "PairHop", "InterAll"};
/* keywords of the reference reader (readdef.c:46-71) that this engine does not implement */
static const char *kw_unsupported[] = {"OneBodyG", "TwoBodyG", "Boost",
                                       "SingleExcitation", "PairExcitation", "SpectrumVec", "Laser",
                                       "TEOneBody", "TETwoBody", "ThreeBodyG", "FourBodyG", "SixBodyG"};

static int ci_eq(const char *a, const char *b) {
  for (; *a && *b; a++, b++)
    if ((*a | 32) != (*b | 32)) return 0;
  return *a == *b;
}

typedef struct {
  int n, cap, w;    /* w = ints per entry */
  int *idx;
  double *val;      /* 2 per entry (re, im) */
} tlist;

static void tl_push(tlist *t, const int *ix, double re, double im) {
  if (t->n == t->cap) {
    t->cap = t->cap ? 2 * t->cap : 64;
    t->idx = (int *)realloc(t->idx, sizeof(int) * t->w * t->cap);
    t->val = (double *)realloc(t->val, sizeof(double) * 2 * t->cap);
  }
  memcpy(t->idx + t->w * t->n, ix, sizeof(int) * t->w);
  t->val[2 * t->n] = re;
  t->val[2 * t->n + 1] = im;
  t->n++;
}
static void tl_free(tlist *t) {
  free(t->idx);
  free(t->val);
}

static void join_path(char *out, size_t n, const char *dir, const char *file) {
  if (file[0] == '/' || dir[0] == 0) snprintf(out, n, "%s", file);
  else snprintf(out, n, "%s/%s", dir, file);
}

/* term file: header line, "Name count", 3 more header lines, then entries */
static int read_terms(const char *path, int nint, int nval, tlist *out, int *count_decl) {
  FILE *f = fopen(path, "r");
  char line[MAXLINE], name[256];
  if (!f) return hphi_set_error("cannot open %s", path);
  int count = 0;
  if (!fgets(line, sizeof line, f) || !fgets(line, sizeof line, f) || sscanf(line, "%255s %d", name, &count) != 2) {
    fclose(f);
    return hphi_set_error("%s: bad header", path);
  }
  for (int i = 0; i < 3; i++)
    if (!fgets(line, sizeof line, f)) {
      fclose(f);
      return hphi_set_error("%s: bad header", path);
    }
  out->w = nint;
  int got = 0;
  while (fgets(line, sizeof line, f)) {
    if (line[0] == '\n' || line[0] == 0) continue;
    int ix[8] = {0};
    double v[2] = {0.0, 0.0};
    int r;
    if (nint == 1) r = sscanf(line, "%d %lf %lf", &ix[0], &v[0], &v[1]);
    else if (nint == 2) r = sscanf(line, "%d %d %lf %lf", &ix[0], &ix[1], &v[0], &v[1]);
    else if (nint == 4) r = sscanf(line, "%d %d %d %d %lf %lf", &ix[0], &ix[1], &ix[2], &ix[3], &v[0], &v[1]);
    else r = sscanf(line, "%d %d %d %d %d %d %d %d %lf %lf", &ix[0], &ix[1], &ix[2], &ix[3], &ix[4], &ix[5],
                    &ix[6], &ix[7], &v[0], &v[1]);
    if (r < nint + nval) {
      fclose(f);
      return hphi_set_error("%s: bad entry line '%s'", path, line);
    }
    if (got == count) {  /* more entries than declared: readdef.c ReadDefFileError */
      fclose(f);
      return hphi_set_error("%s: more entries than the declared %d", path, count);
    }
    tl_push(out, ix, v[0], nval > 1 ? v[1] : 0.0);
    got++;
  }
  fclose(f);
  if (count_decl) *count_decl = count;
  return 0;
}

static int read_locspn(const char *path, int nsite, int *nlocspn, int *kind) {
  FILE *f = fopen(path, "r");
  char line[MAXLINE], name[256];
  if (!f) return hphi_set_error("cannot open %s", path);
  if (!fgets(line, sizeof line, f) || !fgets(line, sizeof line, f) || sscanf(line, "%255s %d", name, nlocspn) != 2) {
    fclose(f);
    return hphi_set_error("%s: bad header", path);
  }
  for (int i = 0; i < 3; i++)
    if (!fgets(line, sizeof line, f)) {
      fclose(f);
      return hphi_set_error("%s: bad header", path);
    }
  int n = 0, k0 = -1;
  while (fgets(line, sizeof line, f)) {
    int i, s;
    if (sscanf(line, "%d %d", &i, &s) != 2) continue;
    if (i < 0 || i >= nsite) {
      fclose(f);
      return hphi_set_error("%s: site %d out of range", path, i);
    }
    if (k0 < 0) k0 = s;
    if (s != k0) {
      fclose(f);
      return hphi_set_error("%s: mixed local spins (Kondo) are not supported", path);
    }
    n++;
  }
  fclose(f);
  *kind = k0;
  return 0;
}

/* owned storage behind a descriptor returned by hphi_read_def */
typedef struct {
  hphi_model_desc d;
  void *blocks[32];
  int nblocks;
} desc_store;

static void *keep(desc_store *s, void *p) {
  s->blocks[s->nblocks++] = p;
  return p;
}

void hphi_model_desc_free(hphi_model_desc *d) {
  if (!d) return;
  desc_store *s = (desc_store *)d; /* d is the first member */
  for (int i = 0; i < s->nblocks; i++) free(s->blocks[i]);
  free(s);
}

int hphi_engine_create_from_def(const char *namelist_path, hphi_engine **out, hphi_modpara *mp_out) {
  hphi_model_desc *d = NULL;
  if (!out) return hphi_set_error("NULL argument");
  *out = NULL;
  if (hphi_read_def(namelist_path, &d, mp_out)) return -1;
  int rc = hphi_engine_create(d, out);
  hphi_model_desc_free(d);
  return rc;
}

int hphi_read_def(const char *namelist_path, hphi_model_desc **dout, hphi_modpara *mp_out) {
  char dir[2048] = "", files[KW_N][2048], line[MAXLINE];
  int i;
  if (!namelist_path || !dout) return hphi_set_error("NULL argument");
  *dout = NULL;
  const char *slash = strrchr(namelist_path, '/');
  if (slash) {
    size_t n = (size_t)(slash - namelist_path);
    if (n >= sizeof dir) n = sizeof dir - 1;
    memcpy(dir, namelist_path, n);
    dir[n] = 0;
  }
  for (i = 0; i < KW_N; i++) files[i][0] = 0;
  FILE *f = fopen(namelist_path, "r");
  if (!f) return hphi_set_error("cannot open %s", namelist_path);
  while (fgets(line, sizeof line, f)) {
    char kw[256] = "", fn[2048] = "";
    int r = sscanf(line, "%255s %2047s", kw, fn);
    if (r <= 0 || kw[0] == '#') continue;
    if (r == 1) {
      fclose(f);
      return hphi_set_error("%s: keyword %s without a file", namelist_path, kw);
    }
    int found = -1;
    for (i = 0; i < KW_N; i++)
      if (ci_eq(kw, kw_names[i])) found = i;
    if (found < 0) {
      for (i = 0; i < (int)(sizeof kw_unsupported / sizeof kw_unsupported[0]); i++)
        if (ci_eq(kw, kw_unsupported[i])) {
          fclose(f);
          return hphi_set_error("%s: keyword %s is not supported by this engine", namelist_path, kw);
        }
      fclose(f);
      return hphi_set_error("%s: unknown keyword %s", namelist_path, kw);
    }
    if (files[found][0]) {
      fclose(f);
      return hphi_set_error("%s: keyword %s given twice", namelist_path, kw);
    }
    join_path(files[found], sizeof files[found], dir, fn);
  }
  fclose(f);
  for (i = 0; i < 3; i++)
    if (!files[i][0]) return hphi_set_error("%s: %s file is required", namelist_path, kw_names[i]);

  /* calcmod.def (ReadcalcmodFile, readdef.c:233-380): every keyword of the
   * reference is accepted and an unknown one is an error, as there; the
   * driver flags (ReStart, InitialVecType, CalcEigenVec, OutputEigenVec,
   * InputEigenVec) go to the caller in hphi_modpara; the inputs that select
   * paths this library does not have (InputHam / OutputHam, CalcSpec,
   * OutputExcitedVec) are rejected by name. */
  int calc_type = 0, calc_model = 0, restart = 0, ini_type = 0, calc_vec = 0, out_vec = 0, in_vec = 0;
  f = fopen(files[KW_CALCMOD], "r");
  if (!f) return hphi_set_error("cannot open %s", files[KW_CALCMOD]);
  while (fgets(line, sizeof line, f)) {
    char kw[256];
    int v;
    if (line[0] == '\n' || line[0] == '#' || line[0] == '=') continue;
    if (sscanf(line, "%255s %d", kw, &v) != 2) continue;
    if (ci_eq(kw, "CalcType")) calc_type = v;
    else if (ci_eq(kw, "CalcModel")) calc_model = v;
    else if (ci_eq(kw, "ReStart")) restart = v;
    else if (ci_eq(kw, "InitialVecType")) ini_type = v;
    else if (ci_eq(kw, "CalcEigenVec")) calc_vec = v;
    else if (ci_eq(kw, "OutputEigenVec") || ci_eq(kw, "OEV")) out_vec = v;
    else if (ci_eq(kw, "InputEigenVec") || ci_eq(kw, "IEV")) in_vec = v;
    else if (ci_eq(kw, "OutputMode") || ci_eq(kw, "FlgFiniteTemperature") || ci_eq(kw, "NGPU") ||
             ci_eq(kw, "ScaLAPACK")) {
      /* no effect on this library's results */
    } else if (ci_eq(kw, "InputHam") || ci_eq(kw, "OutputHam") || ci_eq(kw, "CalcSpec") ||
               ci_eq(kw, "CalcSpectrum") || ci_eq(kw, "OutputExcitedVec") || ci_eq(kw, "OutputExVec")) {
      if (v != 0) {
        fclose(f);
        return hphi_set_error("%s: %s %d is not supported by this library", files[KW_CALCMOD], kw, v);
      }
    } else {
      fclose(f);
      return hphi_set_error("%s: unknown keyword %s", files[KW_CALCMOD], kw); /* readdef.c:317-320 */
    }
  }
  fclose(f);
  if (restart < 0 || restart > 3) return hphi_set_error("%s: ReStart must be 0..3", files[KW_CALCMOD]);
  if (ini_type < -1 || ini_type > 1) return hphi_set_error("%s: InitialVecType must be -1..1", files[KW_CALCMOD]);
  if (calc_vec < -1 || calc_vec > 1) return hphi_set_error("%s: CalcEigenVec must be -1..1", files[KW_CALCMOD]);

  /* modpara.def */
  hphi_modpara mp;
  memset(&mp, 0, sizeof mp);
  mp.calc_type = calc_type;
  mp.lanczos_max = 2000;
  mp.initial_iv = -1;
  mp.exct = 1;
  mp.lanczos_eps = 14;
  mp.lanczos_target = 2;
  mp.num_ave = 1;
  mp.expec_interval = 1;
  mp.restart = restart;
  mp.initial_vec_type = ini_type;
  mp.calc_eigenvec = calc_vec;
  mp.output_eigenvec = out_vec;
  mp.input_eigenvec = in_vec;
  int nsite = 0, two_sz = 0, have_sz = 0, ncond = -1, nup = -1, ndown = -1;
  f = fopen(files[KW_MODPARA], "r");
  if (!f) return hphi_set_error("cannot open %s", files[KW_MODPARA]);
  for (i = 0; i < 8; i++)
    if (!fgets(line, sizeof line, f)) {
      fclose(f);
      return hphi_set_error("%s: short header", files[KW_MODPARA]);
    }
  while (fgets(line, sizeof line, f)) {
    char kw[256];
    double v = 0.0;
    if (line[0] == '\n') continue;
    if (sscanf(line, "%255s %lf", kw, &v) != 2) continue;
    if (ci_eq(kw, "Nsite")) nsite = (int)v;
    else if (ci_eq(kw, "2Sz")) { two_sz = (int)v; have_sz = 1; }
    else if (ci_eq(kw, "Ncond")) ncond = (int)v;
    else if (ci_eq(kw, "Nup")) nup = (int)v;
    else if (ci_eq(kw, "Ndown")) ndown = (int)v;
    else if (ci_eq(kw, "Lanczos_max")) mp.lanczos_max = (int)v;
    else if (ci_eq(kw, "initial_iv")) mp.initial_iv = (int)v;
    else if (ci_eq(kw, "exct")) mp.exct = (int)v;
    else if (ci_eq(kw, "LanczosEps")) mp.lanczos_eps = (int)v;
    else if (ci_eq(kw, "LanczosTarget")) mp.lanczos_target = (int)v;
    else if (ci_eq(kw, "LargeValue")) mp.large_value = v;
    else if (ci_eq(kw, "NumAve")) mp.num_ave = (int)v;
    else if (ci_eq(kw, "ExpecInterval")) mp.expec_interval = (int)v;
    else if (ci_eq(kw, "ExpandCoef")) mp.expand_coef = (int)v;
  }
  fclose(f);
  if (nsite <= 0) return hphi_set_error("%s: Nsite must be positive", files[KW_MODPARA]);
  if (mp.lanczos_target < mp.exct) mp.lanczos_target = mp.exct; /* readdef.c:969 */

  int nloc = 0, lkind = 0;
  if (read_locspn(files[KW_LOCSPIN], nsite, &nloc, &lkind)) return -1;

  hphi_model_desc d;
  memset(&d, 0, sizeof d);
  d.model = calc_model;
  d.nsite = nsite;
  if (calc_model == HPHI_MODEL_SPIN) {
    if (lkind != 1) return hphi_set_error("only S=1/2 spins are supported (LocSpin 1)");
    if (have_sz) {
      d.nup = (nloc + two_sz) / 2;
      d.ndown = (nloc - two_sz) / 2;
    } else if (nup >= 0 && ndown >= 0) {
      d.nup = nup;
      d.ndown = ndown;
    } else {
      return hphi_set_error("2Sz is not defined");
    }
    if (d.nup + d.ndown != nsite) return hphi_set_error(" 2Sz is incorrect."); /* check.c:165 */
  } else if (calc_model == HPHI_MODEL_HUBBARD) {
    if (lkind != 0) return hphi_set_error("Kondo sites are not supported");
    if (ncond >= 0 && have_sz) {
      d.nup = (nloc + ncond + two_sz) / 2;
      d.ndown = (nloc + ncond - two_sz) / 2;
    } else if (nup >= 0 && ndown >= 0) {
      d.nup = nup;
      d.ndown = ndown;
    } else {
      return hphi_set_error("canonical Hubbard needs Ncond and 2Sz (HubbardNConserved is not supported)");
    }
  } else if (calc_model == HPHI_MODEL_SPINGC) {
    if (lkind != 1) return hphi_set_error("only S=1/2 spins are supported (LocSpin 1)");
  } else if (calc_model == HPHI_MODEL_HUBBARDGC) {
    if (lkind != 0) return hphi_set_error("Kondo sites are not supported");
  } else {
    return hphi_set_error("CalcModel %d is not supported (0 Hubbard, 1 Spin, 3 HubbardGC, 4 SpinGC)", calc_model);
  }
  const int hubbard_model = calc_model == HPHI_MODEL_HUBBARD || calc_model == HPHI_MODEL_HUBBARDGC;

  tlist tr = {0}, intra = {0}, inter = {0}, hund = {0}, ex = {0}, ising = {0}, pl = {0};
  tlist edt = {0}, chemi = {0}, ph = {0}, ia = {0}, iaoff = {0}, iadiag = {0}, ph2 = {0};
  int rc = 0;
  if (files[KW_TRANS][0] && read_terms(files[KW_TRANS], 4, 2, &tr, NULL)) rc = -1;
  if (!rc && files[KW_INTRA][0] && read_terms(files[KW_INTRA], 1, 1, &intra, NULL)) rc = -1;
  if (!rc && files[KW_INTER][0] && read_terms(files[KW_INTER], 2, 1, &inter, NULL)) rc = -1;
  if (!rc && files[KW_HUND][0] && read_terms(files[KW_HUND], 2, 1, &hund, NULL)) rc = -1;
  if (!rc && files[KW_EXCHANGE][0] && read_terms(files[KW_EXCHANGE], 2, 1, &ex, NULL)) rc = -1;
  if (!rc && files[KW_ISING][0] && read_terms(files[KW_ISING], 2, 1, &ising, NULL)) rc = -1;
  if (!rc && files[KW_PAIRLIFT][0] && read_terms(files[KW_PAIRLIFT], 2, 1, &pl, NULL)) rc = -1;
  if (!rc && files[KW_PAIRHOP][0] && read_terms(files[KW_PAIRHOP], 2, 1, &ph, NULL)) rc = -1;
  if (!rc && files[KW_INTERALL][0] && read_terms(files[KW_INTERALL], 8, 2, &ia, NULL)) rc = -1;
  if (!rc && ph.n && !hubbard_model) rc = hphi_set_error("PairHop needs a Hubbard model");
  const int spin_model = !hubbard_model;

  if (!rc) {
    /* merge duplicates, keeping first-occurrence order (readdef.c:1130-1150) */
    tlist m = {0};
    m.w = 4;
    for (int a = 0; a < tr.n && calc_model == HPHI_MODEL_SPIN; a++)
      if (tr.idx[4 * a] != tr.idx[4 * a + 2]) { /* readdef.c:1106-1110 */
        rc = hphi_set_error("%s: a Spin transfer must act on one site (%d %d)", files[KW_TRANS], tr.idx[4 * a],
                            tr.idx[4 * a + 2]);
        break;
      }
    for (int a = 0; a < tr.n && !rc; a++) {
      int hit = -1;
      for (int b = 0; b < m.n && hit < 0; b++)
        if (!memcmp(m.idx + 4 * b, tr.idx + 4 * a, 4 * sizeof(int))) hit = b;
      if (hit >= 0) {
        m.val[2 * hit] += tr.val[2 * a];
        m.val[2 * hit + 1] += tr.val[2 * a + 1];
      } else {
        tl_push(&m, tr.idx + 4 * a, tr.val[2 * a], tr.val[2 * a + 1]);
      }
    }
    /* Hermitian pairing (readdef.c:2136-2234) */
    edt.w = 4;
    chemi.w = 2;
    for (int a = 0; a < m.n && !rc; a++) {
      const int *A = m.idx + 4 * a;
      int found = 0;
      for (int b = 0; b < m.n; b++) {
        const int *B = m.idx + 4 * b;
        if (A[0] == B[2] && A[2] == B[0] && A[1] == B[3] && A[3] == B[1]) {
          double dre = m.val[2 * a] - m.val[2 * b], dim = m.val[2 * a + 1] + m.val[2 * b + 1];
          if (sqrt(dre * dre + dim * dim) > 1e-12) {
            rc = hphi_set_error("non-Hermitian transfer (%d %d %d %d)", A[0], A[1], A[2], A[3]);
            break;
          }
          if (!found) {
            if (a <= b) {
              if (A[0] != A[2] || A[1] != A[3]) {
                tl_push(&edt, A, m.val[2 * a], m.val[2 * a + 1]);
                tl_push(&edt, B, m.val[2 * b], m.val[2 * b + 1]);
              } else {
                int cs[2] = {A[0], A[1]};
                tl_push(&chemi, cs, m.val[2 * a], 0.0);
              }
            }
            found = 1;
          }
        }
      }
      if (!found && !rc) rc = hphi_set_error("transfer (%d %d %d %d) has no Hermitian partner", A[0], A[1], A[2], A[3]);
    }
    tl_free(&m);
  }
  if (!rc) {
    /* PairHop: each entry and its reverse (readdef.c:1257-1267) */
    ph2.w = 2;
    for (int a = 0; a < ph.n; a++) {
      int rev[2] = {ph.idx[2 * a + 1], ph.idx[2 * a]};
      tl_push(&ph2, ph.idx + 2 * a, ph.val[2 * a], 0.0);
      tl_push(&ph2, rev, ph.val[2 * a], 0.0);
    }
    /* InterAll: duplicates summed in first-occurrence order (InputInterAllInfo,
     * readdef.c:3054-3090), then split (GetDiagonalInterAll, readdef.c:2415-2530) */
    tlist m = {0};
    m.w = 8;
    for (int a = 0; a < ia.n && !rc; a++) {
      /* CheckInterAllCondition (readdef.c:2996-3033): sites in range; spin
       * entries are (i s1 i s2 k s3 k s4) */
      const int *q = ia.idx + 8 * a;
      for (int u = 0; u < 8; u += 2)
        if (q[u] < 0 || q[u] >= nsite) rc = hphi_set_error("InterAll site index out of range");
      if (!rc && spin_model && (q[0] != q[2] || q[4] != q[6]))
        rc = hphi_set_error("spin InterAll (%d %d %d %d %d %d %d %d) needs i == j and k == l", q[0], q[1], q[2], q[3],
                            q[4], q[5], q[6], q[7]);
      if (rc) break;
      int hit = -1;
      for (int b = 0; b < m.n && hit < 0; b++)
        if (!memcmp(m.idx + 8 * b, ia.idx + 8 * a, 8 * sizeof(int))) hit = b;
      if (hit >= 0) {
        m.val[2 * hit] += ia.val[2 * a];
        m.val[2 * hit + 1] += ia.val[2 * a + 1];
      } else {
        tl_push(&m, ia.idx + 8 * a, ia.val[2 * a], ia.val[2 * a + 1]);
      }
    }
    iaoff.w = 8;
    iadiag.w = 4;
    chemi.w = 2;
    for (int a = 0; a < m.n && !rc; a++) {
      const int *q = m.idx + 8 * a; /* i s1 j s2 k s3 l s4 */
      const double vr = m.val[2 * a], vi = m.val[2 * a + 1];
      if (q[0] == q[2] && q[4] == q[6] && q[1] == q[3] && q[5] == q[7]) {
        int d4[4] = {q[0], q[1], q[4], q[5]};
        tl_push(&iadiag, d4, vr, 0.0);
      } else if (q[0] == q[6] && q[2] == q[4] && q[1] == q[7] && q[3] == q[5]) {
        int d4[4] = {q[0], q[1], q[2], q[3]}, cs[2] = {q[0], q[1]};
        tl_push(&iadiag, d4, -vr, 0.0);
        tl_push(&chemi, cs, -vr, 0.0);
      } else if (spin_model || (q[1] == q[3] && q[5] == q[7])) {
        tl_push(&iaoff, q, vr, vi);
      } else if (q[1] == q[7] && q[3] == q[5]) {
        int r8[8] = {q[0], q[1], q[6], q[1], q[4], q[3], q[2], q[3]};
        tl_push(&iaoff, r8, -vr, -vi);
      } else if (calc_model == HPHI_MODEL_HUBBARDGC) {
        tl_push(&iaoff, q, vr, vi); /* HubbardGC keeps Sz-changing entries */
      } else {
        rc = hphi_set_error("InterAll (%d %d %d %d %d %d %d %d) does not conserve Sz", q[0], q[1], q[2], q[3], q[4],
                            q[5], q[6], q[7]);
      }
    }
    tl_free(&m);
    /* Hermitian pairing: entry a followed by its reverse with the conjugate
     * value (CheckInterAllHermite, readdef.c:2257-2360) */
    tlist pr = {0};
    pr.w = 8;
    for (int a = 0; a < iaoff.n && !rc; a++) {
      const int *A = iaoff.idx + 8 * a;
      int hit = -1;
      for (int b = 0; b < iaoff.n && hit < 0; b++) {
        const int *B = iaoff.idx + 8 * b;
        if (A[0] == B[6] && A[1] == B[7] && A[2] == B[4] && A[3] == B[5] && A[4] == B[2] && A[5] == B[3] &&
            A[6] == B[0] && A[7] == B[1]) {
          const double dre = iaoff.val[2 * a] - iaoff.val[2 * b], dim = iaoff.val[2 * a + 1] + iaoff.val[2 * b + 1];
          if (sqrt(dre * dre + dim * dim) < 1e-12) hit = b;
        }
      }
      if (hit >= 0) {
        if (a <= hit) {
          tl_push(&pr, A, iaoff.val[2 * a], iaoff.val[2 * a + 1]);
          tl_push(&pr, iaoff.idx + 8 * hit, iaoff.val[2 * hit], iaoff.val[2 * hit + 1]);
        }
        continue;
      }
      /* spin: the partner may list the two site operators the other way
       * round, (i s2 i s1 k s4 k s3); the pair is then entry a and its
       * reverse, with the partner's value (readdef.c:2321-2352) */
      if (spin_model && A[0] != A[4] && A[1] != A[5] && A[2] != A[6] && A[3] != A[7]) {
        for (int b = 0; b < iaoff.n && hit < 0; b++) {
          const int *B = iaoff.idx + 8 * b;
          if (A[0] == B[2] && A[2] == B[0] && A[4] == B[6] && A[6] == B[4] && A[1] == B[3] && A[3] == B[1] &&
              A[5] == B[7] && A[7] == B[5]) {
            const double dre = iaoff.val[2 * a] - iaoff.val[2 * b], dim = iaoff.val[2 * a + 1] + iaoff.val[2 * b + 1];
            if (sqrt(dre * dre + dim * dim) < 1e-12) hit = b;
          }
        }
        if (hit >= 0) {
          if (a <= hit) {
            const int rev[8] = {A[6], A[7], A[4], A[5], A[2], A[3], A[0], A[1]};
            tl_push(&pr, A, iaoff.val[2 * a], iaoff.val[2 * a + 1]);
            tl_push(&pr, rev, iaoff.val[2 * hit], iaoff.val[2 * hit + 1]);
          }
          continue;
        }
      }
      rc = hphi_set_error("InterAll (%d %d %d %d %d %d %d %d) has no Hermitian partner", A[0], A[1], A[2], A[3], A[4],
                          A[5], A[6], A[7]);
    }
    if (!rc && pr.n != iaoff.n) rc = hphi_set_error("InterAll entries do not pair into Hermitian conjugates");
    tl_free(&iaoff);
    iaoff = pr;
  }
  if (!rc) {
    /* Ising -> Hund (-J/2) and CoulombInter (-J/4), appended (readdef.c:1320-1327) */
    hund.w = 2;
    inter.w = 2;
    for (int a = 0; a < ising.n; a++) tl_push(&hund, ising.idx + 2 * a, -ising.val[2 * a] / 2.0, 0.0);
    for (int a = 0; a < ising.n; a++) tl_push(&inter, ising.idx + 2 * a, -ising.val[2 * a] / 4.0, 0.0);
  }
  /* flatten into an owned descriptor */
  desc_store *st = NULL;
  int *ch_site = NULL, *ch_spin = NULL;
  double *ch_mu = NULL, *v_intra = NULL, *v_inter = NULL, *v_hund = NULL, *v_ex = NULL, *v_pl = NULL;
  double *v_ph = NULL, *v_iad = NULL;
  int *s_intra = NULL;
  if (!rc) {
    ch_site = (int *)calloc(chemi.n + 1, sizeof(int));
    ch_spin = (int *)calloc(chemi.n + 1, sizeof(int));
    ch_mu = (double *)calloc(chemi.n + 1, sizeof(double));
    for (int a = 0; a < chemi.n; a++) {
      ch_site[a] = chemi.idx[2 * a];
      ch_spin[a] = chemi.idx[2 * a + 1];
      ch_mu[a] = chemi.val[2 * a];
    }
    s_intra = (int *)calloc(intra.n + 1, sizeof(int));
    v_intra = (double *)calloc(intra.n + 1, sizeof(double));
    for (int a = 0; a < intra.n; a++) {
      s_intra[a] = intra.idx[a];
      v_intra[a] = intra.val[2 * a];
    }
#define REALS(dst, src)                                                    \
  dst = (double *)calloc((src).n + 1, sizeof(double));                     \
  for (int a = 0; a < (src).n; a++) dst[a] = (src).val[2 * a];
    REALS(v_inter, inter);
    REALS(v_hund, hund);
    REALS(v_ex, ex);
    REALS(v_pl, pl);
    REALS(v_ph, ph2);
    REALS(v_iad, iadiag);
#undef REALS
    double *v_iao = (double *)calloc(2 * iaoff.n + 2, sizeof(double));
    for (int a = 0; a < 2 * iaoff.n; a++) v_iao[a] = iaoff.val[a];
    double *tt = (double *)calloc(2 * edt.n + 2, sizeof(double));
    for (int a = 0; a < 2 * edt.n; a++) tt[a] = edt.val[a];
    d.n_exchange = ex.n; d.exchange_ij = ex.idx; d.exchange_J = v_ex;
    d.n_pairlift = pl.n; d.pairlift_ij = pl.idx; d.pairlift_J = v_pl;
    d.n_transfer = edt.n; d.transfer_idx = edt.idx; d.transfer_t = tt;
    d.n_chemi = chemi.n; d.chemi_site = ch_site; d.chemi_spin = ch_spin; d.chemi_mu = ch_mu;
    d.n_intra = intra.n; d.intra_site = s_intra; d.intra_U = v_intra;
    d.n_inter = inter.n; d.inter_ij = inter.idx; d.inter_V = v_inter;
    d.n_hund = hund.n; d.hund_ij = hund.idx; d.hund_J = v_hund;
    d.n_interall = iaoff.n; d.interall_idx = iaoff.idx; d.interall_V = v_iao;
    d.n_interall_diag = iadiag.n; d.interall_diag_idx = iadiag.idx; d.interall_diag_V = v_iad;
    d.n_pairhop = ph2.n; d.pairhop_ij = ph2.idx; d.pairhop_J = v_ph;
    st = (desc_store *)calloc(1, sizeof(desc_store));
    st->d = d;
    void *owned[] = {tt, ch_site, ch_spin, ch_mu, s_intra, v_intra, v_inter, v_hund, v_ex, v_pl, v_ph, v_iad,
                     v_iao, ex.idx, pl.idx, edt.idx, inter.idx, hund.idx, iaoff.idx, iadiag.idx, ph2.idx};
    for (int a = 0; a < (int)(sizeof owned / sizeof owned[0]); a++) keep(st, owned[a]);
    ex.idx = pl.idx = edt.idx = inter.idx = hund.idx = iaoff.idx = iadiag.idx = ph2.idx = NULL;
    *dout = &st->d;
  } else {
    free(ch_site); free(ch_spin); free(ch_mu); free(s_intra); free(v_intra);
    free(v_inter); free(v_hund); free(v_ex); free(v_pl); free(v_ph); free(v_iad);
  }
  tl_free(&tr); tl_free(&intra); tl_free(&inter); tl_free(&hund); tl_free(&ex);
  tl_free(&ising); tl_free(&pl); tl_free(&edt); tl_free(&chemi);
  tl_free(&ph); tl_free(&ia); tl_free(&iaoff); tl_free(&iadiag); tl_free(&ph2);
  if (!rc && mp_out) *mp_out = mp;
  return rc;
}
