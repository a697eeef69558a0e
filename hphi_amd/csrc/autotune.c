/*
 * Launch-shape autotuner for the H·v kernels (host C over the hphi_amd C ABI).
 *
 * For the single-vector canonical kernel it sweeps workgroup size, workgroups
 * per CU, list_2_1 in LDS or L2, and the XCD-contiguous tile mapping; for the
 * block kernel (k > 1) it sweeps columns per lane, workgroup size, workgroups
 * per CU and the XCD mapping; the per-column kernel is always a candidate.  Each candidate is timed with HIP events over `reps` H·v calls on
 * the caller's vectors (v0 accumulates: the caller re-initialises it), and
 * the fastest is left set on the engine.  A one-line report per candidate is
 * written to `report` when it is non-NULL.
 */
#include <stdio.h>
#include <string.h>

#include "hphi_amd.h"

int hphi_set_error(const char *fmt, ...);

typedef struct {
  int n;
  const char *knob[6];
  int val[6];
} cand_t;

static void add(cand_t *c, const char *knob, int val) {
  c->knob[c->n] = knob;
  c->val[c->n] = val;
  c->n++;
}

static int apply(hphi_engine *e, const cand_t *c) {
  for (int i = 0; i < c->n; i++)
    if (hphi_engine_tune(e, c->knob[i], c->val[i])) return -1;
  return 0;
}

int hphi_autotune(hphi_engine *e, int k, double *v0, const double *v1, int reps, char *report,
                  size_t report_len) {
  if (!e || !v0 || !v1) return hphi_set_error("hphi_autotune: NULL argument");
  if (reps < 1) reps = 2;
  cand_t cands[96];
  int n = 0, c2 = 0, blk = 0, lds_ok = 0, crank_ok = 0;
  memset(cands, 0, sizeof cands);
  if (k == 1 && hphi_engine_get_tune(e, "use_c2", &c2) == 0 && c2) {
    /* is list_2_1 allowed in LDS for this model? */
    int saved = 0;
    hphi_engine_get_tune(e, "lds21", &saved);
    lds_ok = hphi_engine_tune(e, "lds21", 1) == 0;
    hphi_engine_tune(e, "lds21", saved);
    hphi_engine_get_tune(e, "c2_crank", &saved);
    crank_ok = hphi_engine_tune(e, "c2_crank", 1) == 0;
    hphi_engine_tune(e, "c2_crank", saved);
    for (int xcd = 1; xcd >= 0; xcd--) {
      int tpbs[3] = {256, 512, 1024};
      if (lds_ok) {
        for (int a = 1; a < 3; a++) {
          cand_t *c = &cands[n++];
          add(c, "use_c2", 1), add(c, "xcd_map", xcd), add(c, "tpb", tpbs[a]);
          add(c, "wgs_per_cu", 1), add(c, "lds21", 1), add(c, "c2_crank", 0);
        }
      }
      for (int a = 0; a < 3; a++)
        for (int w = 2; w <= 8; w *= 2)
          if (tpbs[a] * w <= 4096) {
            cand_t *c = &cands[n++];
            add(c, "use_c2", 1), add(c, "xcd_map", xcd), add(c, "tpb", tpbs[a]);
            add(c, "wgs_per_cu", w), add(c, "lds21", 0), add(c, "c2_crank", 0);
          }
      if (crank_ok) { /* spin rank tables: large grids of small workgroups */
        const int ws[4] = {14, 21, 40, 48};
        for (int a = 0; a < 4; a++) {
          cand_t *c = &cands[n++];
          add(c, "use_c2", 1), add(c, "xcd_map", xcd), add(c, "tpb", 256);
          add(c, "wgs_per_cu", ws[a]), add(c, "c2_crank", 1);
        }
      }
    }
  }
  if (k > 1 && hphi_engine_get_tune(e, "use_blk", &blk) == 0 && blk) {
    /* block kernel: columns per lane x workgroup size x workgroups per CU x XCD map */
    static char knob[16];
    snprintf(knob, sizeof knob, "blk_cpl%d", k);
    int cpls[3] = {1, 4, k}, ncpl = k > 4 ? 3 : 2;
    if (k <= 4) cpls[1] = k;
    for (int xcd = 1; xcd >= 0; xcd--)
      for (int a = 0; a < ncpl; a++) {
        /* wide lanes hold 2*cpl accumulators: keep them on 256-thread groups */
        for (int w = 2; w <= 8; w *= 2) {
          cand_t *c = &cands[n++];
          add(c, "use_blk", 1), add(c, "xcd_map", xcd), add(c, knob, cpls[a]);
          add(c, "blk_tpb", 256), add(c, "blk_wgs_per_cu", w);
        }
        if (cpls[a] <= 4) {
          cand_t *c = &cands[n++];
          add(c, "use_blk", 1), add(c, "xcd_map", xcd), add(c, knob, cpls[a]);
          add(c, "blk_tpb", 1024), add(c, "blk_wgs_per_cu", 1);
        }
      }
  }
  for (int xcd = 1; xcd >= 0; xcd--)
    for (int g = 2; g <= 16; g *= 2) {
      cand_t *c = &cands[n++];
      add(c, "use_c2", 0), add(c, "use_blk", 0), add(c, "xcd_map", xcd), add(c, "grid_per_cu", g);
    }

  double best = 1e300;
  int bi = -1;
  size_t used = 0;
  if (report && report_len) report[0] = 0;
  for (int i = 0; i < n; i++) {
    double ms = 0.0;
    if (apply(e, &cands[i])) continue;
    if (hphi_time_mltply_dev(e, k, v0, v1, 1, reps, &ms)) return -1;
    if (report && used + 160 < report_len) {
      for (int q = 0; q < cands[i].n; q++)
        used += (size_t)snprintf(report + used, report_len - used, "%s=%d ", cands[i].knob[q], cands[i].val[q]);
      used += (size_t)snprintf(report + used, report_len - used, ": %.3f ms\n", ms);
    }
    if (ms < best) {
      best = ms;
      bi = i;
    }
  }
  if (bi < 0) return hphi_set_error("hphi_autotune: no candidate ran");
  /* leave every kernel enabled, then the winner's settings */
  hphi_engine_tune(e, "use_c2", 1);
  hphi_engine_tune(e, "use_blk", 1);
  return apply(e, &cands[bi]);
}
