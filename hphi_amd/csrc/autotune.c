/*
 * Launch-shape autotuner for the H·v kernels (host C over the hphi_amd C ABI).
 *
 * For the single-vector canonical kernel it sweeps workgroup size, workgroups
 * per CU, list_2_1 in LDS or L2, and the XCD-contiguous tile mapping; for the
 * SpinGC and block (k > 1) kernel it sweeps workgroups per CU and the XCD
 * mapping.  Each candidate is timed with HIP events over `reps` H·v calls on
 * the caller's vectors (v0 accumulates: the caller re-initialises it), and
 * the fastest is left set on the engine.  A one-line report per candidate is
 * written to `report` when it is non-NULL.
 */
#include <stdio.h>
#include <string.h>

#include "hphi_amd.h"

int hphi_set_error(const char *fmt, ...);

typedef struct {
  int use_c2, tpb, wpc, lds, xcd, gpc;
} cand_t;

static int apply(hphi_engine *e, const cand_t *c) {
  if (hphi_engine_tune(e, "use_c2", c->use_c2)) return -1;
  if (hphi_engine_tune(e, "xcd_map", c->xcd)) return -1;
  if (c->use_c2) {
    if (hphi_engine_tune(e, "tpb", c->tpb)) return -1;
    if (hphi_engine_tune(e, "wgs_per_cu", c->wpc)) return -1;
    if (hphi_engine_tune(e, "lds21", c->lds)) return -1;
  } else {
    if (hphi_engine_tune(e, "grid_per_cu", c->gpc)) return -1;
  }
  return 0;
}

int hphi_autotune(hphi_engine *e, int k, double *v0, const double *v1, int reps, char *report,
                  size_t report_len) {
  if (!e || !v0 || !v1) return hphi_set_error("hphi_autotune: NULL argument");
  if (reps < 1) reps = 2;
  cand_t cands[64];
  int n = 0, c2 = 0, lds_ok = 0;
  if (k == 1 && hphi_engine_get_tune(e, "use_c2", &c2) == 0 && c2) {
    /* is list_2_1 allowed in LDS for this model? */
    int saved = 0;
    hphi_engine_get_tune(e, "lds21", &saved);
    lds_ok = hphi_engine_tune(e, "lds21", 1) == 0;
    hphi_engine_tune(e, "lds21", saved);
    for (int xcd = 1; xcd >= 0; xcd--) {
      if (lds_ok) {
        cands[n++] = (cand_t){1, 1024, 1, 1, xcd, 8};
        cands[n++] = (cand_t){1, 512, 1, 1, xcd, 8};
      }
      int tpbs[3] = {256, 512, 1024};
      for (int a = 0; a < 3; a++)
        for (int w = 2; w <= 8; w *= 2)
          if (tpbs[a] * w <= 4096) cands[n++] = (cand_t){1, tpbs[a], w, 0, xcd, 8};
    }
  }
  for (int xcd = 1; xcd >= 0; xcd--)
    for (int g = 2; g <= 16; g *= 2) cands[n++] = (cand_t){0, 256, 1, 0, xcd, g};

  double best = 1e300;
  int bi = -1;
  size_t used = 0;
  if (report && report_len) report[0] = 0;
  for (int i = 0; i < n; i++) {
    double ms = 0.0;
    if (apply(e, &cands[i])) continue;
    if (hphi_time_mltply_dev(e, k, v0, v1, 1, reps, &ms)) return -1;
    if (report && used + 96 < report_len)
      used += (size_t)snprintf(report + used, report_len - used,
                               "use_c2=%d tpb=%d wgs_per_cu=%d lds21=%d xcd_map=%d grid_per_cu=%d : %.3f ms\n",
                               cands[i].use_c2, cands[i].tpb, cands[i].wpc, cands[i].lds, cands[i].xcd,
                               cands[i].gpc, ms);
    if (ms < best) {
      best = ms;
      bi = i;
    }
  }
  if (bi < 0) return hphi_set_error("hphi_autotune: no candidate ran");
  return apply(e, &cands[bi]);
}
