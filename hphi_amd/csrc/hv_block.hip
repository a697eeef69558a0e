// Block H·v over K vectors in one launch: V0 (+)= H V1 for the K columns of
// a state-major block (column c of state j at element j*K + c), with
// prdct[c] = <V1_c | H V1_c> for every column.
//
// The bits, the predicate, the Jordan-Wigner sign and the target index of a
// state are decoded once for all K columns: a group of LPS = K/CPL lanes
// owns one state, every lane of the group computes the same indices (LPS-fold,
// not K-fold, redundancy) and gathers its own CPL consecutive columns, so one
// term costs one index decode and K*16 contiguous bytes.  CPL trades index
// work (small LPS) against lanes per memory instruction touching one row
// (large LPS); the autotuner picks it per model and K.
//
// Canonical models (Spin, Hubbard; states of at most 32 bits) keep list_2_1
// in LDS as uint16 and take the wave-uniform path for high-half terms when
// the wave's states share one high block, as the single-vector kernel does.
// Grand-canonical spin states are their own index (target = x ^ flip).
// Same arithmetic as the per-term scatter of the reference
// (mltplySpinCore.c:273-345, mltplyHubbardCore.c:305-347), in gather form.
#include <hip/hip_runtime.h>

#include "engine_internal.h"

#define HIPCHK(x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) return hphi_set_error("%s failed: %s", #x, hipGetErrorString(e_)); \
  } while (0)

__device__ __forceinline__ bool bk_pred(int kind, uint32_t tmask, uint32_t tval, uint32_t x) {
  const uint32_t ym = x & tmask;
  const int pc = __popc(ym);
  return (kind == P_EQ) ? (ym == tval) : ((kind == P_DIFF) ? (pc == 1) : (pc != 1));
}

__device__ __forceinline__ double bk_sign(uint32_t sgn, uint32_t x) {
  return (sgn != 0u && (__popc(x & sgn) & 1)) ? -1.0 : 1.0;
}

// acc += (cr + i ci) * g for CPL columns
template <int CPL>
__device__ __forceinline__ void bk_mac(double *ar, double *ai, double cr, double ci, const double2 *g, bool realc) {
  if (realc) {
#pragma unroll
    for (int c = 0; c < CPL; c++) {
      ar[c] += cr * g[c].x;
      ai[c] += cr * g[c].y;
    }
  } else {
#pragma unroll
    for (int c = 0; c < CPL; c++) {
      ar[c] += cr * g[c].x - ci * g[c].y;
      ai[c] += cr * g[c].y + ci * g[c].x;
    }
  }
}

template <int CPL>
__device__ __forceinline__ void bk_load(const double2 *__restrict__ p, bool act, double2 *g) {
#pragma unroll
  for (int c = 0; c < CPL; c++) g[c] = make_double2(0.0, 0.0);
  if (act) {
#pragma unroll
    for (int c = 0; c < CPL; c++) g[c] = p[c];
  }
}

// rank lookup: the uint16 list_2_1 staged in LDS, or (CRANK, canonical Spin)
// the byte tables of the combinatorial rank
template <int CRANK>
__device__ __forceinline__ int64_t blk_l21(const uint16_t *s21, uint32_t ia) {
  return CRANK == 2 ? hrank_l21(s21, ia) : CRANK == 1 ? crank_l21(s21, ia) : (int64_t)s21[ia];
}

// CRANK: 0 the uint16 list_2_1 in LDS, 1 spin rank tables, 2 Hubbard rank tables
template <int TPB, int K, int CPL, bool CANON, int CRANK>
__global__ __launch_bounds__(TPB) void k_hv_blk(const HvArgs A, const int realc) {
  constexpr int LPS = K / CPL;                // lanes per state
  constexpr int SPT = TPB / LPS;              // states per tile
  constexpr int U = CPL >= 8 ? 1 : 8 / CPL;   // terms per batch (8 or 16 gathers in flight)
  static_assert(K % CPL == 0 && LPS <= 64, "bad block shape");
  extern __shared__ uint16_t s21[];
  const uint32_t tid = threadIdx.x;
  if (CANON && CRANK) {
    const uint32_t n = CRANK == 2 ? (uint32_t)HRANK_N : (uint32_t)CRANK_N;
    for (uint32_t i = tid; i < n; i += TPB) s21[i] = A.crank[i];
    __syncthreads();
  } else if (CANON) {
    for (uint32_t i = tid; i < (uint32_t)A.size21; i += TPB) s21[i] = (uint16_t)A.l21[i];
    __syncthreads();
  }
  const bool rc = realc != 0;
  const double2 *__restrict__ v1 = (const double2 *)A.v1;
  double2 *__restrict__ v0 = (double2 *)A.v0;
  const uint32_t irght = (uint32_t)A.irght;
  const int nlow = A.nlow;
  const uint64_t dim = A.dim;
  const CTerm tl = (CTerm)A.terms;
  const CTerm th = tl + A.n_low;
  const CTerm tm = th + A.n_high;
  const int sub = (int)(tid % LPS);
  const uint32_t sl = tid / LPS;
  const int c0 = sub * CPL;
  double pre[CPL], pim[CPL];
#pragma unroll
  for (int c = 0; c < CPL; c++) pre[c] = pim[c] = 0.0;

  const uint64_t ntiles = (dim + SPT - 1) / SPT;
  uint64_t t_begin, t_end, t_step;
  if (A.xcd_map) {
    const uint32_t xcd = blockIdx.x & 7, slot = blockIdx.x >> 3, nslot = gridDim.x >> 3;
    t_begin = (ntiles * xcd) / 8 + slot;
    t_end = (ntiles * (xcd + 1)) / 8;
    t_step = nslot;
  } else {
    t_begin = blockIdx.x;
    t_end = ntiles;
    t_step = gridDim.x;
  }
  for (uint64_t tile = t_begin; tile < t_end; tile += t_step) {
    const uint64_t j = tile * SPT + sl;
    const bool valid = j < dim;
    const uint64_t jj = valid ? j : dim - 1;
    const uint32_t x = CANON ? (uint32_t)A.list1[jj] : (uint32_t)jj;
    const double d = A.diag[jj];
    double ar[CPL], ai[CPL];
    {
      const double2 *own = v1 + jj * K + c0;
#pragma unroll
      for (int c = 0; c < CPL; c++) {
        const double2 a = own[c];
        ar[c] = d * a.x;
        ai[c] = d * a.y;
      }
    }
    if (CANON) {
      const uint32_t ia = x & irght, ib = x >> nlow;
      const int64_t b21 = blk_l21<CRANK>(s21, ia);
      // ---- low-half terms: same block, rank lookup in LDS
      for (int t = 0; t < A.e_low; t += U) {
        uint64_t idx[U];
        bool act[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
          const GTerm T = load_term(tl + t + u);
          const uint32_t y = x ^ (uint32_t)T.flip;
          const int64_t l = blk_l21<CRANK>(s21, y & irght);
          idx[u] = (uint64_t)((int64_t)j + l - b21);
          act[u] = valid & bk_pred(T.kind, (uint32_t)T.tmask, (uint32_t)T.tval, x) & (l != 0) & (idx[u] < dim);
        }
        double2 g[U][CPL];
#pragma unroll
        for (int u = 0; u < U; u++) bk_load<CPL>(v1 + idx[u] * K + c0, act[u], g[u]);
#pragma unroll
        for (int u = 0; u < U; u++) {
          const GTerm T = load_term(tl + t + u);
          const double s = bk_sign((uint32_t)T.sgn, x);
          bk_mac<CPL>(ar, ai, s * T.cre, s * T.cim, g[u], rc);
        }
      }
      // ---- high-half and mixed terms
      const uint32_t ib0 = __builtin_amdgcn_readfirstlane(ib);
      const bool uni = (__ballot(ib != ib0) == 0ull);
      if (uni) {
        const uint32_t xh = ib0 << nlow;
        const int64_t l22_0 = (int64_t)A.l22[ib0];
        for (int t = 0; t < A.e_high; t += U) {
          uint64_t idx[U];
          bool act[U];
#pragma unroll
          for (int u = 0; u < U; u++) {
            const GTerm T = load_term(th + t + u);
            bool a = bk_pred(T.kind, (uint32_t)T.tmask, (uint32_t)T.tval, xh);
            int64_t l = 0;
            if (a) l = (int64_t)A.l22[(xh ^ (uint32_t)T.flip) >> nlow];
            idx[u] = (uint64_t)((int64_t)j + l - l22_0);
            act[u] = a & valid & (l != 0) & (idx[u] < dim);
          }
          double2 g[U][CPL];
#pragma unroll
          for (int u = 0; u < U; u++) bk_load<CPL>(v1 + idx[u] * K + c0, act[u], g[u]);
#pragma unroll
          for (int u = 0; u < U; u++) {
            const GTerm T = load_term(th + t + u);
            const double s = bk_sign((uint32_t)T.sgn, xh);
            bk_mac<CPL>(ar, ai, s * T.cre, s * T.cim, g[u], rc);
          }
        }
        for (int t = 0; t < A.e_mixed; t += U) {
          uint64_t idx[U];
          bool act[U];
#pragma unroll
          for (int u = 0; u < U; u++) {
            const GTerm T = load_term(tm + t + u);
            const uint32_t y = x ^ (uint32_t)T.flip;
            const int64_t lb = (int64_t)A.l22[(xh ^ (uint32_t)T.flip) >> nlow];  // uniform
            const int64_t la = blk_l21<CRANK>(s21, y & irght);
            idx[u] = (uint64_t)(la + lb - 3);
            act[u] = valid & bk_pred(T.kind, (uint32_t)T.tmask, (uint32_t)T.tval, x) & (la != 0) & (lb != 0) &
                     (idx[u] < dim);
          }
          double2 g[U][CPL];
#pragma unroll
          for (int u = 0; u < U; u++) bk_load<CPL>(v1 + idx[u] * K + c0, act[u], g[u]);
#pragma unroll
          for (int u = 0; u < U; u++) {
            const GTerm T = load_term(tm + t + u);
            const double s = bk_sign((uint32_t)T.sgn, x);
            bk_mac<CPL>(ar, ai, s * T.cre, s * T.cim, g[u], rc);
          }
        }
      } else {
        // lanes span several blocks: per-lane list_2_2 lookups
        for (int seg = 0; seg < 2; seg++) {  // the high group, then the mixed group
          const CTerm tg = seg ? tm : th;
          const int ne = seg ? A.e_mixed : A.e_high;
          for (int t = 0; t < ne; t += U) {
            uint64_t idx[U];
            bool act[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
              const GTerm T = load_term(tg + t + u);
              const uint32_t y = x ^ (uint32_t)T.flip;
              bool a = valid & bk_pred(T.kind, (uint32_t)T.tmask, (uint32_t)T.tval, x);
              int64_t la = b21, lb = 0;
              if (a) {
                lb = (int64_t)A.l22[y >> nlow];
                if ((y & irght) != ia) la = blk_l21<CRANK>(s21, y & irght);
              }
              idx[u] = (uint64_t)(la + lb - 3);
              act[u] = a & (la != 0) & (lb != 0) & (idx[u] < dim);
            }
            double2 g[U][CPL];
#pragma unroll
            for (int u = 0; u < U; u++) bk_load<CPL>(v1 + idx[u] * K + c0, act[u], g[u]);
#pragma unroll
            for (int u = 0; u < U; u++) {
              const GTerm T = load_term(tg + t + u);
              const double s = bk_sign((uint32_t)T.sgn, x);
              bk_mac<CPL>(ar, ai, s * T.cre, s * T.cim, g[u], rc);
            }
          }
        }
      }
    } else {
      // grand canonical: the state is its own index
      for (int t = 0; t < A.e_low; t += U) {
        uint64_t idx[U];
        bool act[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
          const GTerm T = load_term(tl + t + u);
          idx[u] = (uint64_t)(x ^ (uint32_t)T.flip);
          act[u] = valid & bk_pred(T.kind, (uint32_t)T.tmask, (uint32_t)T.tval, x) & (idx[u] < dim);
        }
        double2 g[U][CPL];
#pragma unroll
        for (int u = 0; u < U; u++) bk_load<CPL>(v1 + idx[u] * K + c0, act[u], g[u]);
#pragma unroll
        for (int u = 0; u < U; u++) {
          const GTerm T = load_term(tl + t + u);
          const double s = bk_sign((uint32_t)T.sgn, x);
          bk_mac<CPL>(ar, ai, s * T.cre, s * T.cim, g[u], rc);
        }
      }
    }
    if (valid) {
      double2 *out = v0 + j * K + c0;
      const double2 *own = v1 + j * K + c0;
#pragma unroll
      for (int c = 0; c < CPL; c++) {
        double2 o = make_double2(A.c1 * ar[c], A.c1 * ai[c]);
        if (!A.overwrite) {
          const double2 p = out[c];
          o.x += A.c0 * p.x;
          o.y += A.c0 * p.y;
        }
        out[c] = o;
        const double2 a = own[c];
        pre[c] += a.x * ar[c] + a.y * ai[c];
        pim[c] += a.x * ai[c] - a.y * ar[c];
      }
    }
  }

  // prdct per column: lanes with equal (lane % LPS) hold the same columns
#pragma unroll
  for (int c = 0; c < CPL; c++) {
#pragma unroll
    for (int off = 32; off >= LPS; off >>= 1) {
      pre[c] += __shfl_xor(pre[c], off, 64);
      pim[c] += __shfl_xor(pim[c], off, 64);
    }
  }
  __shared__ double2 red[TPB / 64][K];
  const int lane = tid & 63, wave = tid >> 6;
  if (lane < LPS) {
#pragma unroll
    for (int c = 0; c < CPL; c++) red[wave][lane * CPL + c] = make_double2(pre[c], pim[c]);
  }
  __syncthreads();
  if (tid < (uint32_t)K) {
    double2 s = red[0][tid];
    for (int w = 1; w < TPB / 64; w++) {
      s.x += red[w][tid].x;
      s.y += red[w][tid].y;
    }
    A.partial[(uint64_t)blockIdx.x * K + tid] = s;
  }
}

// ---------------------------------------------------------------- host ----

template <int TPB, int K, int CPL, bool CANON, int CRANK>
static int bk_launch(hphi_engine *e, const HvArgs &A, int g) {
  const size_t lds = CANON ? (CRANK == 2 ? 2 * (size_t)HRANK_N : CRANK == 1 ? 2 * (size_t)CRANK_N : 2 * (size_t)e->size21) : 0;
  if (lds > 65536)
    (void)hipFuncSetAttribute((const void *)k_hv_blk<TPB, K, CPL, CANON, CRANK>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL((k_hv_blk<TPB, K, CPL, CANON, CRANK>), dim3(g), dim3(TPB), lds, e->stream, A, e->c2_realc);
  HIPCHK(hipGetLastError());
  return 0;
}

template <int K, int CPL>
static int bk_launch_tpb(hphi_engine *e, const HvArgs &A, int g, int tpb, bool canon) {
  if (!canon) return bk_launch<256, K, CPL, false, 0>(e, A, g);
  if (e->c2_crank) {
    if (!A.crank) return hphi_set_error("rank tables are not built");
    if (e->model == HPHI_MODEL_HUBBARD) return bk_launch<256, K, CPL, true, 2>(e, A, g);
    return bk_launch<256, K, CPL, true, 1>(e, A, g);
  }
  if (tpb == 1024) return bk_launch<1024, K, CPL, true, 0>(e, A, g);
  return bk_launch<256, K, CPL, true, 0>(e, A, g);
}

// columns per lane actually instantiated for each K
int hphi_blk_cpl_ok(int k, int cpl) {
  if (cpl == 1 || cpl == k) return 1;
  return cpl == 4 && k > 4;
}

int hphi_blk_ok(const hphi_engine *e) {
  if (e->nbits > 32) return 0;
  if (hphi_model_gc(e->model)) return 1;
  return e->c2_lds21 || e->c2_crank;
}

// default launch shape for the block kernel
void hphi_blk_shape(const hphi_engine *e, int k, int *cpl, int *tpb, int *wgs_per_cu) {
  int c = 0;
  switch (k) {
    case 2: c = e->bk_cpl[0]; break;
    case 4: c = e->bk_cpl[1]; break;
    case 8: c = e->bk_cpl[2]; break;
    case 16: c = e->bk_cpl[3]; break;
  }
  const bool canon = !hphi_model_gc(e->model);
  // grand canonical, k = 8 and 16: one column per lane and 4 workgroups per CU,
  // XCD-contiguous (13x2 triangular SpinGC: k = 16 137 ms per block H·v vs 156-170
  // ms for 4 or 16 columns per lane and 8-32 workgroups per CU, mTPQ step 143.8 ms
  // vs 162.6 ms; k = 8 62.3 ms vs 69-100 ms)
  const bool gc_wide = !canon && k >= 8;
  if (c <= 0 || !hphi_blk_cpl_ok(k, c)) c = gc_wide ? 1 : (k >= 8 ? 4 : k);
  *cpl = c;
  int t = e->bk_tpb;
  if (t != 256 && t != 1024) t = (canon && !e->c2_crank && 2 * e->size21 > 32768) ? 1024 : 256;
  if (!canon || e->c2_crank) t = 256;  // the rank tables leave room for many 256-thread workgroups
  *tpb = t;
  int w = e->bk_wgs_per_cu;
  if (w <= 0) w = (t == 1024) ? 1 : (canon && e->c2_crank ? 48 : (gc_wide ? 4 : 8));  // 8x4 J1-J2, k=4: 106 ms/column
  *wgs_per_cu = w;
}

// workgroups of the block kernel's launch for k columns (each writes k partials)
int hphi_blk_grid(const hphi_engine *e, int k) {
  int cpl, tpb, wpc;
  hphi_blk_shape(e, k, &cpl, &tpb, &wpc);
  const uint64_t spt = (uint64_t)tpb / (uint64_t)(k / cpl);
  const uint64_t ntiles = (e->dim + spt - 1) / spt;
  uint64_t g = std::min<uint64_t>(ntiles, (uint64_t)wpc * std::max(e->num_cu, 1));
  g = std::max<uint64_t>(g, 8);
  g = (g + 7) / 8 * 8;
  if (g > HV_MAX_GRID) g = HV_MAX_GRID;
  return (int)g;
}

int hphi_hv_launch_blk(hphi_engine *e, const HvArgs &A, int k, int *grid) {
  int cpl, tpb, wpc;
  hphi_blk_shape(e, k, &cpl, &tpb, &wpc);
  const bool canon = !hphi_model_gc(e->model);
  const int g = hphi_blk_grid(e, k);
  *grid = g;
  // each group's loop stops at its last batch holding a compiled term (the
  // groups are padded to HV_TERM_PAD, a multiple of every batch size)
  const int u = cpl >= 8 ? 1 : 8 / cpl;  // terms per batch of k_hv_blk
  HvArgs B = A;
  if (!canon && k >= 8 && !e->xcd_tuned) B.xcd_map = 1;  // with the grand-canonical k >= 8 shape above
  B.e_low = std::min(A.n_low, (e->r_low + u - 1) / u * u);
  B.e_high = std::min(A.n_high, (e->r_high + u - 1) / u * u);
  B.e_mixed = std::min(A.n_mixed, (e->r_mixed + u - 1) / u * u);
#define CASE(KK, CC) \
  if (k == KK && cpl == CC) return bk_launch_tpb<KK, CC>(e, B, g, tpb, canon);
  CASE(2, 1)
  CASE(2, 2)
  CASE(4, 1)
  CASE(4, 4)
  CASE(8, 1)
  CASE(8, 4)
  CASE(8, 8)
  CASE(16, 1)
  CASE(16, 4)
  CASE(16, 16)
#undef CASE
  return hphi_set_error("block H·v: no kernel for k=%d, %d columns per lane", k, cpl);
}
