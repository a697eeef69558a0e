// Device code of the gather-form H·v shared by the engine's kernels and the
// multi-job serving kernels (included by engine.hip and serve.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "engine_internal.h"

template <int K>
__device__ __forceinline__ void wg_reduce_store(double pre, double pim, double2 *partial) {
  // lanes with equal (lane % K) hold the same vector column
#pragma unroll
  for (int off = 32; off >= K; off >>= 1) {
    pre += __shfl_xor(pre, off, 64);
    pim += __shfl_xor(pim, off, 64);
  }
  __shared__ double2 red[HV_TPB / 64][K];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane < K) red[wave][lane] = make_double2(pre, pim);
  __syncthreads();
  if (threadIdx.x < K) {
    double2 s = red[0][threadIdx.x];
#pragma unroll
    for (int w = 1; w < HV_TPB / 64; w++) {
      s.x += red[w][threadIdx.x].x;
      s.y += red[w][threadIdx.x].y;
    }
    partial[(uint64_t)blockIdx.x * K + threadIdx.x] = s;
  }
}

__device__ __forceinline__ bool term_active(const GTerm &T, uint64_t x) {
  uint64_t ym = x & T.tmask;
  int pc = __popcll(ym);
  return (T.kind == P_EQ) ? (ym == T.tval) : ((T.kind == P_DIFF) ? (pc == 1) : (pc != 1));
}

// (H v1)_j of column `vec` for state j < A.dim: the diagonal plus every
// off-diagonal term in gather form.  a = v1[j] of that column.
template <int K, bool CANON>
__device__ __forceinline__ void hv_row(const HvArgs &A, uint64_t j, int vec, double2 &a, double &ar, double &ai) {
  const double2 *__restrict__ v1 = (const double2 *)A.v1;
  {
    const uint64_t x = CANON ? A.list1[j] : j;
    a = v1[j * K + vec];
    const double d = A.diag[j];
    ar = d * a.x;
    ai = d * a.y;
    if (CANON) {
      const uint64_t ia0 = x & A.irght, ib0 = x >> A.nlow;
      const int64_t b21 = (int64_t)A.l21[ia0], b22 = (int64_t)A.l22[ib0];
      // low group: only the low half changes, same block
      for (int t = 0; t < A.e_low; t += HV_UNROLL) {
        double2 g[HV_UNROLL];
        double cr[HV_UNROLL], ci[HV_UNROLL];
#pragma unroll
        for (int u = 0; u < HV_UNROLL; u++) {
          const GTerm T = load_term((CTerm)A.terms + t + u);
          const uint64_t y = x ^ T.flip;
          const int64_t l = (int64_t)A.l21[y & A.irght];
          bool act = term_active(T, x) && (l != 0);
          const double s = (__popcll(x & T.sgn) & 1) ? -1.0 : 1.0;
          uint64_t idx = (uint64_t)((int64_t)j + l - b21);
          act = act && idx < A.dim;
          idx = act ? idx : j;
          g[u] = v1[idx * K + vec];
          cr[u] = act ? s * T.cre : 0.0;
          ci[u] = act ? s * T.cim : 0.0;
        }
#pragma unroll
        for (int u = 0; u < HV_UNROLL; u++) {
          ar += cr[u] * g[u].x - ci[u] * g[u].y;
          ai += cr[u] * g[u].y + ci[u] * g[u].x;
        }
      }
      // high group: only the high half changes, same rank in another block
      const CTerm th = (CTerm)A.terms + A.n_low;
      for (int t = 0; t < A.e_high; t += HV_UNROLL) {
        double2 g[HV_UNROLL];
        double cr[HV_UNROLL], ci[HV_UNROLL];
#pragma unroll
        for (int u = 0; u < HV_UNROLL; u++) {
          const GTerm T = load_term(th + t + u);
          const uint64_t y = x ^ T.flip;
          const int64_t l = (int64_t)A.l22[y >> A.nlow];
          bool act = term_active(T, x) && (l != 0);
          const double s = (__popcll(x & T.sgn) & 1) ? -1.0 : 1.0;
          uint64_t idx = (uint64_t)((int64_t)j + l - b22);
          act = act && idx < A.dim;
          idx = act ? idx : j;
          g[u] = v1[idx * K + vec];
          cr[u] = act ? s * T.cre : 0.0;
          ci[u] = act ? s * T.cim : 0.0;
        }
#pragma unroll
        for (int u = 0; u < HV_UNROLL; u++) {
          ar += cr[u] * g[u].x - ci[u] * g[u].y;
          ai += cr[u] * g[u].y + ci[u] * g[u].x;
        }
      }
      // mixed group: both halves change, two lookups
      const CTerm tm = th + A.n_high;
      for (int t = 0; t < A.e_mixed; t += HV_UNROLL) {
        double2 g[HV_UNROLL];
        double cr[HV_UNROLL], ci[HV_UNROLL];
#pragma unroll
        for (int u = 0; u < HV_UNROLL; u++) {
          const GTerm T = load_term(tm + t + u);
          const uint64_t y = x ^ T.flip;
          const int64_t la = (int64_t)A.l21[y & A.irght];
          const int64_t lb = (int64_t)A.l22[y >> A.nlow];
          bool act = term_active(T, x) && (la != 0) && (lb != 0);
          const double s = (__popcll(x & T.sgn) & 1) ? -1.0 : 1.0;
          uint64_t idx = (uint64_t)(la + lb - 3);
          act = act && idx < A.dim;
          idx = act ? idx : j;
          g[u] = v1[idx * K + vec];
          cr[u] = act ? s * T.cre : 0.0;
          ci[u] = act ? s * T.cim : 0.0;
        }
#pragma unroll
        for (int u = 0; u < HV_UNROLL; u++) {
          ar += cr[u] * g[u].x - ci[u] * g[u].y;
          ai += cr[u] * g[u].y + ci[u] * g[u].x;
        }
      }
    } else {
      for (int t = 0; t < A.e_low; t += HV_UNROLL) {
        double2 g[HV_UNROLL];
        double cr[HV_UNROLL], ci[HV_UNROLL];
#pragma unroll
        for (int u = 0; u < HV_UNROLL; u++) {
          const GTerm T = load_term((CTerm)A.terms + t + u);
          bool act = term_active(T, x);
          const double s = (__popcll(x & T.sgn) & 1) ? -1.0 : 1.0;
          uint64_t idx = x ^ T.flip;
          act = act && idx < A.dim;
          idx = act ? idx : j;
          g[u] = v1[idx * K + vec];
          cr[u] = act ? s * T.cre : 0.0;
          ci[u] = act ? s * T.cim : 0.0;
        }
#pragma unroll
        for (int u = 0; u < HV_UNROLL; u++) {
          ar += cr[u] * g[u].x - ci[u] * g[u].y;
          ai += cr[u] * g[u].y + ci[u] * g[u].x;
        }
      }
    }
  }
}

// one tile of HV_TPB/K states x K vectors; adds this lane's share of prdct to (pre, pim)
template <int K, bool CANON>
__device__ __forceinline__ void hv_tile(const HvArgs &A, uint64_t tile, double &pre, double &pim) {
  constexpr int SPT = HV_TPB / K;  // states per tile
  const int vec = threadIdx.x & (K - 1);
  const int sl = threadIdx.x / K;
  double2 *__restrict__ v0 = (double2 *)A.v0;
  {
    const uint64_t j = tile * SPT + sl;
    if (j >= A.dim) return;
    double2 a;
    double ar, ai;
    hv_row<K, CANON>(A, j, vec, a, ar, ai);
    double2 o = make_double2(A.c1 * ar, A.c1 * ai);
    if (!A.overwrite) {
      const double2 p = v0[j * K + vec];
      o.x += A.c0 * p.x;
      o.y += A.c0 * p.y;
    }
    v0[j * K + vec] = o;
    // conj(a) * acc
    pre += a.x * ar + a.y * ai;
    pim += a.x * ai - a.y * ar;
  }
}

