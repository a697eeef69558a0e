/*
 * The reference's output and restart files, written and read from host
 * memory in its byte layouts and printf formats:
 *
 *   eigenvector / TPQ vector   <prefix>_eigenvec_<k>_rank_<r>.dat (CalcByLanczos.c:160-198,262-272),
 *                              tmpvec_set<i>_rank_<r>.dat (CalcByTPQ.c:88-104,280-287):
 *                              int step, unsigned long dim, complex128[dim + 1] (element 0 unused)
 *   Lanczos restart vectors    <prefix>_recalcvec_rank_<r>.dat (Lanczos_EigenValue.c:428-489):
 *                              unsigned long step, unsigned long dim, v0[dim + 1], v1[dim + 1]
 *   tridiagonal components     <prefix>_TMComponents.dat (Lanczos_EigenValue.c:624-720)
 *   energy                     <prefix>_energy.dat (CalcByLanczos.c:238-252)
 *   Green's functions          <prefix>_cisajs.dat, <prefix>_cisajscktalt.dat
 *                              (expec_cisajs.c, expec_cisajscktaltdc.c fprintf lines)
 *   TPQ                        SS_rand<i>.dat, Norm_rand<i>.dat, Flct_rand<i>.dat (CalcByTPQ.c:150-256,
 *                              headers LogMessage.c:143-145)
 *
 * Vectors here are 0-based complex128 host arrays of dim elements; the files
 * carry the reference's 1-based arrays, whose element 0 is written as 0.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "hphi_amd.h"

int hphi_set_error(const char *fmt, ...);

static FILE *open_or_fail(const char *path, const char *mode) {
  FILE *f = fopen(path, mode);
  if (!f) hphi_set_error("cannot open %s", path);
  return f;
}

static int write_vec(FILE *f, uint64_t dim, const double *v) {
  const double zero[2] = {0.0, 0.0};
  if (fwrite(zero, sizeof zero, 1, f) != 1) return -1;
  return fwrite(v, 2 * sizeof(double), dim, f) == dim ? 0 : -1;
}

static int read_vec(FILE *f, uint64_t dim, double *v) {
  double first[2];
  if (fread(first, sizeof first, 1, f) != 1) return -1;
  return fread(v, 2 * sizeof(double), dim, f) == dim ? 0 : -1;
}

int hphi_write_vec_file(const char *path, int step, uint64_t dim, const double *v) {
  if (!path || (dim && !v)) return hphi_set_error("hphi_write_vec_file: NULL argument");
  FILE *f = open_or_fail(path, "wb");
  if (!f) return -1;
  const unsigned long d = (unsigned long)dim;
  int rc = (fwrite(&step, sizeof step, 1, f) == 1 && fwrite(&d, sizeof d, 1, f) == 1 && !write_vec(f, dim, v))
               ? 0
               : hphi_set_error("write failed: %s", path);
  fclose(f);
  return rc;
}

int hphi_read_vec_file(const char *path, uint64_t dim, double *v, int *step) {
  if (!path || (dim && !v)) return hphi_set_error("hphi_read_vec_file: NULL argument");
  FILE *f = open_or_fail(path, "rb");
  if (!f) return -1;
  int s = 0, rc = 0;
  unsigned long d = 0;
  if (fread(&s, sizeof s, 1, f) != 1 || fread(&d, sizeof d, 1, f) != 1) {
    rc = hphi_set_error("%s: truncated header", path);
  } else if (d != dim) { /* CalcByLanczos.c:194-198 */
    rc = hphi_set_error("%s: dimension %lu differs from %llu (a file of Inputvector is incorrect)", path, d,
                        (unsigned long long)dim);
  } else if (read_vec(f, dim, v)) {
    rc = hphi_set_error("%s: truncated vector", path);
  }
  fclose(f);
  if (!rc && step) *step = s;
  return rc;
}

int hphi_write_restart_vecs(const char *path, uint64_t step, uint64_t dim, const double *v0, const double *v1) {
  if (!path || (dim && (!v0 || !v1))) return hphi_set_error("hphi_write_restart_vecs: NULL argument");
  FILE *f = open_or_fail(path, "wb");
  if (!f) return -1;
  const unsigned long s = (unsigned long)step, d = (unsigned long)dim;
  int rc = (fwrite(&s, sizeof s, 1, f) == 1 && fwrite(&d, sizeof d, 1, f) == 1 && !write_vec(f, dim, v0) &&
            !write_vec(f, dim, v1))
               ? 0
               : hphi_set_error("write failed: %s", path);
  fclose(f);
  return rc;
}

int hphi_read_restart_vecs(const char *path, uint64_t dim, double *v0, double *v1, uint64_t *step) {
  if (!path || (dim && (!v0 || !v1))) return hphi_set_error("hphi_read_restart_vecs: NULL argument");
  FILE *f = open_or_fail(path, "rb");
  if (!f) return -1;
  unsigned long s = 0, d = 0;
  int rc = 0;
  if (fread(&s, sizeof s, 1, f) != 1 || fread(&d, sizeof d, 1, f) != 1) {
    rc = hphi_set_error("%s: truncated header", path);
  } else if (d != dim) { /* Lanczos_EigenValue.c:442-446 */
    rc = hphi_set_error("%s: dimension %lu differs from %llu (a size of Inputvector is incorrect)", path, d,
                        (unsigned long long)dim);
  } else if (read_vec(f, dim, v0) || read_vec(f, dim, v1)) {
    rc = hphi_set_error("%s: truncated vectors", path);
  }
  fclose(f);
  if (!rc && step) *step = s;
  return rc;
}

int hphi_write_tmcomponents(const char *path, int steps, double norm, const double *alpha, const double *beta) {
  if (!path || (steps > 0 && (!alpha || !beta))) return hphi_set_error("hphi_write_tmcomponents: NULL argument");
  FILE *f = open_or_fail(path, "w");
  if (!f) return -1;
  fprintf(f, "%d \n", steps);
  fprintf(f, "%.10lf \n", norm);
  for (int i = 0; i < steps; i++) fprintf(f, "%.10lf %.10lf \n", alpha[i], beta[i]);
  return fclose(f) == 0 ? 0 : hphi_set_error("write failed: %s", path);
}

int hphi_read_tmcomponents(const char *path, int cap, int *steps, double *norm, double *alpha, double *beta) {
  if (!path || !steps || !norm) return hphi_set_error("hphi_read_tmcomponents: NULL argument");
  FILE *f = open_or_fail(path, "r");
  if (!f) return -1;
  char line[256];
  long n = 0;
  int rc = 0, i = 0;
  if (!fgets(line, sizeof line, f) || sscanf(line, "%ld", &n) != 1 || !fgets(line, sizeof line, f) ||
      sscanf(line, "%lf", norm) != 1) {
    rc = hphi_set_error("%s: bad header", path);
  } else {
    /* ReadTMComponents (Lanczos_EigenValue.c:677-681): every further line is
     * one (alpha, beta) pair */
    while (fgets(line, sizeof line, f)) {
      double a, b;
      if (sscanf(line, "%lf %lf", &a, &b) != 2) continue;
      if (i >= cap) {
        rc = hphi_set_error("%s: more than %d components", path, cap);
        break;
      }
      if (alpha) alpha[i] = a;
      if (beta) beta[i] = b;
      i++;
    }
    if (!rc && i != n) rc = hphi_set_error("%s: %d components, header says %ld", path, i, n);
  }
  fclose(f);
  if (!rc) *steps = i;
  return rc;
}

int hphi_write_energy(const char *path, const hphi_phys *p) {
  if (!path || !p) return hphi_set_error("hphi_write_energy: NULL argument");
  FILE *f = open_or_fail(path, "w");
  if (!f) return -1;
  fprintf(f, "Energy  %.16lf \n", p->energy);
  fprintf(f, "Doublon  %.16lf \n", p->doublon);
  fprintf(f, "Sz  %.16lf \n", p->sz);
  return fclose(f) == 0 ? 0 : hphi_set_error("write failed: %s", path);
}

int hphi_write_energy_states(const char *path, int n, const hphi_phys *p) {
  if (!path || (n > 0 && !p)) return hphi_set_error("hphi_write_energy_states: NULL argument");
  FILE *f = open_or_fail(path, "w");
  if (!f) return -1;
  for (int i = 0; i < n; i++) {
    fprintf(f, "State %ld\n", (long)i);
    fprintf(f, "  Energy  %.16lf \n", p[i].energy);
    fprintf(f, "  Doublon  %.16lf \n", p[i].doublon);
    fprintf(f, "  Sz  %.16lf \n", p[i].sz);
    fprintf(f, "\n");
  }
  return fclose(f) == 0 ? 0 : hphi_set_error("write failed: %s", path);
}

int hphi_write_cisajs(const char *path, int n, const int *idx, const double *val) {
  if (!path || (n > 0 && (!idx || !val))) return hphi_set_error("hphi_write_cisajs: NULL argument");
  FILE *f = open_or_fail(path, "w");
  if (!f) return -1;
  for (int t = 0; t < n; t++) {
    const int *q = idx + 4 * t;
    fprintf(f, " %4ld %4ld %4ld %4ld %.10lf %.10lf\n", (long)q[0], (long)q[1], (long)q[2], (long)q[3], val[2 * t],
            val[2 * t + 1]);
  }
  return fclose(f) == 0 ? 0 : hphi_set_error("write failed: %s", path);
}

int hphi_write_cisajscktalt(const char *path, int model, int n, const int *idx, const double *val) {
  if (!path || (n > 0 && (!idx || !val))) return hphi_set_error("hphi_write_cisajscktalt: NULL argument");
  FILE *f = open_or_fail(path, "w");
  if (!f) return -1;
  for (int t = 0; t < n; t++) {
    const int *q = idx + 8 * t;
    /* line ends: HubbardGC and computed canonical-Hubbard lines "\n"
     * (expec_cisajscktaltdc.c:495,616); the Sz-skipped canonical-Hubbard
     * lines (:538) and every spin line (:682,:767,:1448,:1519) " \n" */
    int trail = 1;
    if (model == HPHI_MODEL_HUBBARDGC) trail = 0;
    if (model == HPHI_MODEL_HUBBARD) trail = (q[1] + q[5] != q[3] + q[7]);
    fprintf(f, " %4ld %4ld %4ld %4ld %4ld %4ld %4ld %4ld %.10lf %.10lf%s", (long)q[0], (long)q[1], (long)q[2],
            (long)q[3], (long)q[4], (long)q[5], (long)q[6], (long)q[7], val[2 * t], val[2 * t + 1],
            trail ? " \n" : "\n");
  }
  return fclose(f) == 0 ? 0 : hphi_set_error("write failed: %s", path);
}

/* flct_from: first step with a Flct_rand line (microcanonical 1, CalcByTPQ.c:150-170;
 * canonical 0, CalcByCanonicalTPQ.c:193-200) */
static int write_tpq_files(const char *dir, int rand_i, int steps, const double *rows, int row_stride, int flct_from) {
  if (!dir || (steps > 0 && !rows)) return hphi_set_error("hphi_write_tpq_files: NULL argument");
  char p_ss[4096], p_norm[4096], p_flct[4096];
  snprintf(p_ss, sizeof p_ss, "%s/SS_rand%d.dat", dir, rand_i);
  snprintf(p_norm, sizeof p_norm, "%s/Norm_rand%d.dat", dir, rand_i);
  snprintf(p_flct, sizeof p_flct, "%s/Flct_rand%d.dat", dir, rand_i);
  FILE *fs = open_or_fail(p_ss, "w");
  if (!fs) return -1;
  FILE *fn = open_or_fail(p_norm, "w");
  if (!fn) {
    fclose(fs);
    return -1;
  }
  FILE *ff = open_or_fail(p_flct, "w");
  if (!ff) {
    fclose(fs);
    fclose(fn);
    return -1;
  }
  fprintf(fs, "%s", " # inv_tmp, energy, phys_var, phys_doublon, phys_num, step_i\n");
  fprintf(fn, "%s", " # inv_temp, global_norm, global_1st_norm, step_i \n");
  fprintf(ff, "%s", " # inv_temp, N, N^2, D, D^2, Sz, Sz^2, step_i \n");
  const double norm1 = steps > 0 ? rows[3] : 0.0;
  for (int s = 0; s < steps; s++) {
    const double *o = rows + (size_t)s * row_stride;
    fprintf(fs, "%.16lf  %.16lf %.16lf %.16lf %.16lf %d\n", o[0], o[1], o[2], o[4], o[6], s);
    fprintf(fn, "%.16lf %.16lf %.16lf %d\n", o[0], s == 0 ? norm1 : o[3], norm1, s);
    if (s >= flct_from)
      fprintf(ff, "%.16lf %.16lf %.16lf %.16lf %.16lf %.16lf %.16lf %d\n", o[0], o[6], o[7], o[4], o[5], o[8], o[9],
              s);
  }
  const int e1 = fclose(fs), e2 = fclose(fn), e3 = fclose(ff);
  return (e1 || e2 || e3) ? hphi_set_error("write failed in %s", dir) : 0;
}

int hphi_write_tpq_files(const char *dir, int rand_i, int steps, const double *rows, int row_stride) {
  return write_tpq_files(dir, rand_i, steps, rows, row_stride, 1);
}

int hphi_write_ctpq_files(const char *dir, int rand_i, int steps, const double *rows, int row_stride) {
  return write_tpq_files(dir, rand_i, steps, rows, row_stride, 0);
}

int hphi_append_tpq_files(const char *dir, int rand_i, int step0, int steps, const double *rows, int row_stride,
                          double norm1) {
  if (!dir || (steps > step0 && !rows)) return hphi_set_error("hphi_append_tpq_files: NULL argument");
  char p_ss[4096], p_norm[4096], p_flct[4096];
  snprintf(p_ss, sizeof p_ss, "%s/SS_rand%d.dat", dir, rand_i);
  snprintf(p_norm, sizeof p_norm, "%s/Norm_rand%d.dat", dir, rand_i);
  snprintf(p_flct, sizeof p_flct, "%s/Flct_rand%d.dat", dir, rand_i);
  FILE *fs = open_or_fail(p_ss, "a");
  if (!fs) return -1;
  FILE *fn = open_or_fail(p_norm, "a");
  if (!fn) {
    fclose(fs);
    return -1;
  }
  FILE *ff = open_or_fail(p_flct, "a");
  if (!ff) {
    fclose(fs);
    fclose(fn);
    return -1;
  }
  for (int s = step0; s < steps; s++) {
    const double *o = rows + (size_t)(s - step0) * row_stride;
    fprintf(fs, "%.16lf  %.16lf %.16lf %.16lf %.16lf %d\n", o[0], o[1], o[2], o[4], o[6], s);
    fprintf(fn, "%.16lf %.16lf %.16lf %d\n", o[0], o[3], norm1, s);
    fprintf(ff, "%.16lf %.16lf %.16lf %.16lf %.16lf %.16lf %.16lf %d\n", o[0], o[6], o[7], o[4], o[5], o[8], o[9], s);
  }
  const int e1 = fclose(fs), e2 = fclose(fn), e3 = fclose(ff);
  return (e1 || e2 || e3) ? hphi_set_error("write failed in %s", dir) : 0;
}

/* SpinGC three-, four- and six-body Green's function files
 * (expec_cisajscktaltdc.c:1028-1085,1243-1273,1398-1420): every index of an
 * evaluated entry, with a double space before the 13th and the 17th, or the
 * first eight indices and zeros for an entry the reference does not evaluate */
int hphi_write_nbody_green(const char *path, int nbody, int n, const int *idx, const double *val, const int *ok) {
  if (!path || (n > 0 && (!idx || !val || !ok))) return hphi_set_error("hphi_write_nbody_green: NULL argument");
  if (nbody != 3 && nbody != 4 && nbody != 6) return hphi_set_error("hphi_write_nbody_green: nbody must be 3, 4 or 6");
  FILE *f = open_or_fail(path, "w");
  if (!f) return -1;
  const int w = 4 * nbody;
  for (int t = 0; t < n; t++) {
    const int *q = idx + (size_t)w * t;
    if (!ok[t]) {
      for (int u = 0; u < 8; u++) fprintf(f, " %4ld", (long)q[u]);
      fprintf(f, " %.10lf %.10lf \n", 0.0, 0.0);
      continue;
    }
    for (int u = 0; u < w; u++) fprintf(f, (u == 12 || u == 16) ? "  %4ld" : " %4ld", (long)q[u]);
    fprintf(f, " %.10lf %.10lf \n", val[2 * t], val[2 * t + 1]);
  }
  return fclose(f) == 0 ? 0 : hphi_set_error("write failed: %s", path);
}
