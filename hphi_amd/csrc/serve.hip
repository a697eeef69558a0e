// Continuous-batching Lanczos service: many small diagonalisation jobs share
// every kernel launch.
//
// Each job is a canonical Hamiltonian (Spin or Hubbard) whose lowest n
// eigenvalues are wanted.  Jobs whose basis and term structure agree (same
// lattice and sector; couplings may differ) share a slab: one copy of the
// basis tables, and SL_K interleaved columns (element j*SL_K + c) for the
// Lanczos vectors, the diagonal and the couplings.  One lane per basis state
// decodes the state's bits, predicates, Jordan-Wigner signs and target
// indices once and applies them to every running column of its slab, so up
// to SL_K jobs pay for one index decode and read one 64-byte row per term.
//
// A tick
//   1. admits waiting jobs: first, free columns of running slabs go to arrived
//      jobs of the same structure among the first ADMIT_WINDOW waiting (slab
//      steps cost about the same however many columns run); then jobs are
//      admitted in arrival order (the largest first among jobs that arrive
//      together), into a free column of a slab of the same
//      structure, or into a new slab while its HBM footprint (tables, SL_K
//      columns of diagonal and two Lanczos vectors, alpha/beta histories) fits
//      the arena; a new column gets its couplings, its diagonal and its start
//      vector on the device;
//   2. advances every running job by S steps (S = the convergence-check
//      interval, at most SL_SMAX; fewer when a job reaches its dimension or
//      the step limit) of the reference's plain Lanczos recurrence
//      (src/Lanczos_EigenValue.c:153-177): per step four packed launches over
//      all slabs (residual + norm, beta, H·v, alpha), then one launch of the
//      batched eigenvalue check, with no host round trip between the steps
//      (a workgroup takes one tile of SV_TPB states, XCD-contiguous);
//   3. copies the (alpha, beta) of every step and the check results of all
//      jobs to the host in one transfer each, tests convergence of the lowest
//      n distinct Ritz values (Cullum-Willoughby filtered) and retires
//      converged jobs, freeing a slab's arena memory when its last job leaves.
//
// The recurrence is rearranged so that every vector is touched by two passes
// per step.  A slab keeps two arrays, R and Q.  At step m of a column, R
// holds r = v0 after the previous H·v and Q holds w_{m-1}, the previous
// Lanczos vector before normalisation (v_{m-1} = s_{m-1} w_{m-1}):
//   pass 1: w_m = r - alpha_{m-1} s_{m-1} w_{m-1}, written over R, and |w_m|^2;
//   beta_{m-1} = |w_m|, s_m = 1/beta_{m-1};
//   pass 2: Q <- -beta_{m-1} s_{m-1} Q + s_m (H w_m)  (= -beta v_{m-1} + H v_m)
//           and alpha_m = s_m^2 <w_m|H w_m>;
// then R and Q swap roles for the whole slab (every column steps every tick;
// a new column is started in the slab's current R).  This is the
// reference's v1 <- (v0 - alpha v1)/beta, v0 <- -beta v1 + H v1 with the
// division by beta folded into the H·v.
//
// Tiles map to slabs through a prefix table.  A job's reductions run over its
// slab's tiles in a fixed order and its column's arithmetic does not depend
// on the other columns, so its result is bit-identical alone and in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <time.h>
#include <vector>

#include "engine_internal.h"
#include "hphi_amd.h"

#define HIPCHK(x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) return hphi_set_error("%s failed: %s", #x, hipGetErrorString(e_)); \
  } while (0)

#define SV_TPB 256  // one tile = SV_TPB states of one slab
#define SL_K 4      // columns (jobs) per slab (8 with one lane per state: 2-3 waves per SIMD, slower per column)
#define SL_U 2      // terms per gather batch (SL_U * SL_K gathers in flight)
#define SL_SMAX 16  // Lanczos steps per tick at most
#define SL_TPW 1    // tiles per workgroup (a chunk): 1, 2, 4, 8 measured, 1 fastest at every job size

struct SlabDev {
  uint64_t dim;
  const uint64_t *list1;
  const uint4 *l21p;  // list_2_1 as uint16, zero-padded to a multiple of 8 entries
  const uint64_t *l22;
  const GTerm *terms;   // term structure of the slab (couplings in coef)
  const double2 *coef;  // [n_terms][SL_K]
  const double *diag;   // [dim][SL_K]
  double2 *R, *Q;       // [dim][SL_K]
  uint64_t irght, size21;
  int n_low, n_high, n_mixed, nlow;
  int mask;   // columns with a running job
  int realc;  // columns whose couplings are all real
  int slot[SL_K];
};
struct ColDev {  // one running job in one step of a tick
  int slab, col, slot;
  int step, check;  // Lanczos step computed (1-based); run the eigenvalue check after it
  int job;          // index of the job in the tick (rows of the outputs)
  double *ahist, *bhist;
};
typedef const SlabDev __attribute__((address_space(4))) *CSlab;
typedef const ColDev __attribute__((address_space(4))) *CCol;
typedef const uint64_t __attribute__((address_space(4))) *CU64;
typedef const double __attribute__((address_space(4))) *CD2;  // couplings as (re, im) doubles

// job state per slot (device): [0] alpha_{m-1}, [1] s_{m-1}, [2] beta_{m-1}, [3] s_m
#define JS_STRIDE 4

__device__ __forceinline__ int find_slab(CU64 prefix, int n, uint64_t tile) {
  int lo = 0, hi = n;  // prefix[lo] <= tile < prefix[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (prefix[mid] <= tile) lo = mid;
    else hi = mid;
  }
  return lo;
}

// XCD-contiguous chunk order.  Workgroups are dealt to the 8 XCDs round-robin
// by blockIdx, so workgroup b of a grid of G (a multiple of 8) takes logical
// chunk (b % 8) G/8 + b / 8: each XCD sweeps one contiguous range of chunks,
// and the gathers of neighbouring states meet in that XCD's own L2.  Logical
// chunks at or past the launch's total (the rounding) have no work.
__device__ __forceinline__ uint64_t sl_chunk() {
  return (uint64_t)(blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
}

// pass 1: w = r - alpha_{m-1} s_{m-1} q over R per running column, |w|^2 per
// chunk of SL_TPW tiles.  Lane (state, column) = (base + t / SL_K, t % SL_K), so
// every load and store of a wave covers one contiguous range of rows.
__global__ __launch_bounds__(SV_TPB) void k_sl_resid(const SlabDev *slabs, const uint64_t *prefix, int ns,
                                                     const double *js, double *partial) {
  const uint64_t chunk = sl_chunk();
  if (chunk >= ((CU64)prefix)[ns]) return;
  const int sb = find_slab((CU64)prefix, ns, chunk);
  CSlab S = (CSlab)slabs + sb;
  const int c = threadIdx.x % SL_K;
  const bool run = (S->mask >> c) & 1;
  const int sl = S->slot[c];
  const double f = js[sl * JS_STRIDE + 0] * js[sl * JS_STRIDE + 1];
  const uint64_t e0 = (chunk - ((CU64)prefix)[sb]) * SL_TPW * SV_TPB * SL_K + threadIdx.x;
  const uint64_t n = S->dim * SL_K, cend = e0 - threadIdx.x + (uint64_t)SL_TPW * SV_TPB * SL_K;
  const uint64_t eend = n < cend ? n : cend;
  double2 *__restrict__ R = S->R;
  const double2 *__restrict__ Q = S->Q;
  double acc = 0.0;
  if (run) {
    for (uint64_t e = e0; e < eend; e += SV_TPB) {
      double2 w = R[e];
      if (f != 0.0) {
        const double2 q = Q[e];
        w = make_double2(w.x - f * q.x, w.y - f * q.y);
        R[e] = w;
      }
      acc += w.x * w.x + w.y * w.y;
    }
  }
  // lanes with equal t % SL_K hold one column
  for (int off = 32; off >= SL_K; off >>= 1) acc += __shfl_xor(acc, off, 64);
  __shared__ double red[SV_TPB / 64][SL_K];
  if ((threadIdx.x & 63) < SL_K) red[threadIdx.x >> 6][threadIdx.x & 63] = acc;
  __syncthreads();
  if (threadIdx.x < SL_K) {
    double t = 0.0;
    for (int w = 0; w < SV_TPB / 64; w++) t += red[w][threadIdx.x];
    partial[chunk * SL_K + threadIdx.x] = t;
  }
}

__device__ __forceinline__ bool sl_pred(int kind, uint32_t tmask, uint32_t tval, uint32_t x) {
  const uint32_t ym = x & tmask;
  const int pc = __popc(ym);
  return (kind == P_EQ) ? (ym == tval) : ((kind == P_DIFF) ? (pc == 1) : (pc != 1));
}

__device__ __forceinline__ double sl_sign(uint32_t sgn, uint32_t x) {
  return (sgn != 0u && (__popc(x & sgn) & 1)) ? -1.0 : 1.0;
}

// acc_c += coef[t][c] * s * g_c for every column of the slab.  Columns without
// a running job are computed too (their results are never stored), so the
// term loop carries no per-column tests.  REAL: the couplings of every
// running column are real.  The complex form applies the imaginary part in
// separate fused steps after the real one, so a real column gets the same
// values on either path: a fused step with a zero factor adds a zero.
template <bool REAL>
__device__ __forceinline__ void sl_mac(double *ar, double *ai, CD2 cf, double s, const double2 *g) {
#pragma unroll
  for (int c = 0; c < SL_K; c++) {
    const double cr = s * cf[2 * c];
    ar[c] = __builtin_fma(cr, g[c].x, ar[c]);
    ai[c] = __builtin_fma(cr, g[c].y, ai[c]);
    if (!REAL) {
      const double ci = s * cf[2 * c + 1];
      ar[c] = __builtin_fma(-ci, g[c].y, ar[c]);
      ai[c] = __builtin_fma(ci, g[c].x, ai[c]);
    }
  }
}

__device__ __forceinline__ void sl_load(const double2 *__restrict__ row, bool act, double2 *g) {
#pragma unroll
  for (int c = 0; c < SL_K; c++) g[c] = make_double2(0.0, 0.0);
  if (act) {
#pragma unroll
    for (int c = 0; c < SL_K; c++) g[c] = row[c];
  }
}

// (H w)_c of one state j of slab S for the running columns, then
// Q_c <- cc_c Q_c + s_c (H w)_c and <w_c|H w_c> accumulated into (pre, pim).
// Gather form of the reference's per-term scatter (mltplySpinCore.c:273-307,
// mltplyHubbardCore.c:305-347); list_2_1 is read from the LDS copy s21.
template <bool REAL>
__device__ __forceinline__ void sl_hv_state(CSlab S, uint64_t j, const uint16_t *s21, const double *js, double *pre,
                                            double *pim) {
  const uint64_t dim = S->dim;
  const bool valid = j < dim;
  const uint64_t jj = valid ? j : dim - 1;
  const int mask = S->mask;
  const double2 *__restrict__ R = S->R;
  const uint64_t *__restrict__ l22 = S->l22;
  const uint32_t irght = (uint32_t)S->irght;
  const int nlow = S->nlow;
  const CTerm tl = (CTerm)S->terms;
  const CTerm th = tl + S->n_low;
  const CTerm tm = th + S->n_high;
  const CD2 cl = (CD2)S->coef;
  const CD2 ch = cl + (uint64_t)S->n_low * SL_K * 2;
  const CD2 cm = ch + (uint64_t)S->n_high * SL_K * 2;
  const uint32_t x = (uint32_t)S->list1[jj];
  double ar[SL_K], ai[SL_K];
  double2 own[SL_K];
#pragma unroll
  for (int c = 0; c < SL_K; c++) {
    own[c] = R[jj * SL_K + c];
    const double d = S->diag[jj * SL_K + c];
    ar[c] = d * own[c].x;
    ai[c] = d * own[c].y;
  }
  const uint32_t ia = x & irght, ib = x >> nlow;
  const int64_t b21 = (int64_t)s21[ia];
  // ---- low-half terms: same block
  for (int t = 0; t < S->n_low; t += SL_U) {
    uint64_t idx[SL_U];
    bool act[SL_U];
#pragma unroll
    for (int u = 0; u < SL_U; u++) {
      const GTerm T = load_term(tl + t + u);
      const uint32_t y = x ^ (uint32_t)T.flip;
      const int64_t l = (int64_t)s21[y & irght];
      idx[u] = (uint64_t)((int64_t)j + l - b21);
      act[u] = valid & sl_pred(T.kind, (uint32_t)T.tmask, (uint32_t)T.tval, x) & (l != 0) & (idx[u] < dim);
    }
    double2 g[SL_U][SL_K];
#pragma unroll
    for (int u = 0; u < SL_U; u++) sl_load(R + idx[u] * SL_K, act[u], g[u]);
#pragma unroll
    for (int u = 0; u < SL_U; u++) {
      const GTerm T = load_term(tl + t + u);
      sl_mac<REAL>(ar, ai, cl + (uint64_t)(t + u) * SL_K * 2, sl_sign((uint32_t)T.sgn, x), g[u]);
    }
  }
  // ---- high-half and mixed terms
  const uint32_t ib0 = __builtin_amdgcn_readfirstlane(ib);
  const bool uni = (__ballot(ib != ib0) == 0ull);
  if (uni) {  // the wave's states share one high block: high-term work is scalar
    const uint32_t xh = ib0 << nlow;
    const int64_t l22_0 = (int64_t)l22[ib0];
    for (int t = 0; t < S->n_high; t += SL_U) {
      uint64_t idx[SL_U];
      bool act[SL_U];
#pragma unroll
      for (int u = 0; u < SL_U; u++) {
        const GTerm T = load_term(th + t + u);
        const bool a = sl_pred(T.kind, (uint32_t)T.tmask, (uint32_t)T.tval, xh);
        int64_t l = 0;
        if (a) l = (int64_t)l22[(xh ^ (uint32_t)T.flip) >> nlow];
        idx[u] = (uint64_t)((int64_t)j + l - l22_0);
        act[u] = a & valid & (l != 0) & (idx[u] < dim);
      }
      double2 g[SL_U][SL_K];
#pragma unroll
      for (int u = 0; u < SL_U; u++) sl_load(R + idx[u] * SL_K, act[u], g[u]);
#pragma unroll
      for (int u = 0; u < SL_U; u++) {
        const GTerm T = load_term(th + t + u);
        sl_mac<REAL>(ar, ai, ch + (uint64_t)(t + u) * SL_K * 2, sl_sign((uint32_t)T.sgn, xh), g[u]);
      }
    }
    for (int t = 0; t < S->n_mixed; t += SL_U) {
      uint64_t idx[SL_U];
      bool act[SL_U];
#pragma unroll
      for (int u = 0; u < SL_U; u++) {
        const GTerm T = load_term(tm + t + u);
        const uint32_t y = x ^ (uint32_t)T.flip;
        const int64_t lb = (int64_t)l22[(xh ^ (uint32_t)T.flip) >> nlow];
        const int64_t la = (int64_t)s21[y & irght];
        idx[u] = (uint64_t)(la + lb - 3);
        act[u] = valid & sl_pred(T.kind, (uint32_t)T.tmask, (uint32_t)T.tval, x) & (la != 0) & (lb != 0) &
                 (idx[u] < dim);
      }
      double2 g[SL_U][SL_K];
#pragma unroll
      for (int u = 0; u < SL_U; u++) sl_load(R + idx[u] * SL_K, act[u], g[u]);
#pragma unroll
      for (int u = 0; u < SL_U; u++) {
        const GTerm T = load_term(tm + t + u);
        sl_mac<REAL>(ar, ai, cm + (uint64_t)(t + u) * SL_K * 2, sl_sign((uint32_t)T.sgn, x), g[u]);
      }
    }
  } else {  // lanes span several blocks: per-lane list_2_2 lookups
    for (int t = 0; t < S->n_high + S->n_mixed; t += SL_U) {
      uint64_t idx[SL_U];
      bool act[SL_U];
#pragma unroll
      for (int u = 0; u < SL_U; u++) {
        const GTerm T = load_term(th + t + u);
        const uint32_t y = x ^ (uint32_t)T.flip;
        bool a = valid & sl_pred(T.kind, (uint32_t)T.tmask, (uint32_t)T.tval, x);
        int64_t la = b21, lb = 0;
        if (a) {
          lb = (int64_t)l22[y >> nlow];
          if ((y & irght) != ia) la = (int64_t)s21[y & irght];
        }
        idx[u] = (uint64_t)(la + lb - 3);
        act[u] = a & (la != 0) & (lb != 0) & (idx[u] < dim);
      }
      double2 g[SL_U][SL_K];
#pragma unroll
      for (int u = 0; u < SL_U; u++) sl_load(R + idx[u] * SL_K, act[u], g[u]);
#pragma unroll
      for (int u = 0; u < SL_U; u++) {
        const GTerm T = load_term(th + t + u);
        sl_mac<REAL>(ar, ai, ch + (uint64_t)(t + u) * SL_K * 2, sl_sign((uint32_t)T.sgn, x), g[u]);
      }
    }
  }
#pragma unroll
  for (int c = 0; c < SL_K; c++) {
    if (!((mask >> c) & 1) || !valid) continue;
    const int sl = S->slot[c];
    const double sc = js[sl * JS_STRIDE + 3], cc = -js[sl * JS_STRIDE + 2] * js[sl * JS_STRIDE + 1];
    double2 *q = S->Q + j * SL_K + c;
    const double2 p = *q;
    *q = make_double2(cc * p.x + sc * ar[c], cc * p.y + sc * ai[c]);
    pre[c] += own[c].x * ar[c] + own[c].y * ai[c];
    pim[c] += own[c].x * ai[c] - own[c].y * ar[c];
  }
}

// pass 2 over one chunk of SL_TPW tiles of one slab: Q_c <- cc_c Q_c + s_c (H w)_c
// for the running columns; <w_c|H w_c> per chunk
__global__ __launch_bounds__(SV_TPB) void k_sl_hv(const SlabDev *slabs, const uint64_t *prefix, int ns,
                                                  const double *js, double2 *partial) {
  extern __shared__ uint4 s21v[];
  const uint16_t *s21 = (const uint16_t *)s21v;
  const uint64_t chunk = sl_chunk();
  if (chunk >= ((CU64)prefix)[ns]) return;
  const int sb = find_slab((CU64)prefix, ns, chunk);
  CSlab S = (CSlab)slabs + sb;
  const uint32_t n8 = (uint32_t)((S->size21 + 7) / 8);  // 16-byte rows of the packed table
  for (uint32_t i = threadIdx.x; i < n8; i += SV_TPB) s21v[i] = S->l21p[i];
  __syncthreads();
  double pre[SL_K], pim[SL_K];
#pragma unroll
  for (int c = 0; c < SL_K; c++) pre[c] = pim[c] = 0.0;
  const uint64_t t0 = (chunk - ((CU64)prefix)[sb]) * SL_TPW;
  const int full = (1 << SL_K) - 1;
  const bool real = ((S->realc | ~S->mask) & full) == full;  // every running column has real couplings
  const uint64_t dim = S->dim;
  for (int t = 0; t < SL_TPW; t++) {
    const uint64_t j = (t0 + t) * SV_TPB + threadIdx.x;
    if ((t0 + t) * SV_TPB >= dim) break;  // uniform: the chunk's last tiles may lie past the slab
    if (real) sl_hv_state<true>(S, j, s21, js, pre, pim);
    else sl_hv_state<false>(S, j, s21, js, pre, pim);
  }
  // per-column prdct partials: [blk][SL_K] double2
  __shared__ double2 red[SV_TPB / 64][SL_K];
#pragma unroll
  for (int c = 0; c < SL_K; c++) {
    for (int off = 32; off > 0; off >>= 1) {
      pre[c] += __shfl_xor(pre[c], off, 64);
      pim[c] += __shfl_xor(pim[c], off, 64);
    }
  }
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int c = 0; c < SL_K; c++) red[threadIdx.x >> 6][c] = make_double2(pre[c], pim[c]);
  __syncthreads();
  if (threadIdx.x < SL_K) {
    double2 s = red[0][threadIdx.x];
    for (int w = 1; w < SV_TPB / 64; w++) {
      s.x += red[w][threadIdx.x].x;
      s.y += red[w][threadIdx.x].y;
    }
    partial[chunk * SL_K + threadIdx.x] = s;
  }
}

// per job (one workgroup each): fixed-order sum over the slab's tiles of one column
__device__ double col_sum(const double *part, int stride, uint64_t t0, uint64_t t1) {
  __shared__ double s[256];
  double a = 0.0;
  for (uint64_t t = t0 + threadIdx.x; t < t1; t += 256) a += part[t * stride];
  s[threadIdx.x] = a;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) s[threadIdx.x] += s[threadIdx.x + st];
    __syncthreads();
  }
  const double r = s[0];
  __syncthreads();
  return r;
}

// beta_{m-1} = |w_m| and s_m = 1/beta_{m-1}; a beta below 1e-14 ends the
// recurrence (the Krylov space is exhausted) and s_m = 0 keeps pass 2 finite
__global__ __launch_bounds__(256) void k_sl_norm(const ColDev *cols, const uint64_t *prefix, const double *partial,
                                                 double *js) {
  CCol J = (CCol)cols + blockIdx.x;
  const double b2 = col_sum(partial + J->col, SL_K, ((CU64)prefix)[J->slab], ((CU64)prefix)[J->slab + 1]);
  if (threadIdx.x == 0) {
    const double b = sqrt(b2);
    double *st = js + J->slot * JS_STRIDE;
    st[2] = b;
    st[3] = (b > 0.0 && (J->step == 1 || b >= 1e-14)) ? 1.0 / b : 0.0;
    if (J->step >= 2) J->bhist[J->step - 1] = b;
  }
}

// alpha_m = s_m^2 Re <w|H w>; out = (alpha_m, beta_{m-1}) per job
__global__ __launch_bounds__(256) void k_sl_alpha(const ColDev *cols, const uint64_t *prefix, const double2 *partial,
                                                  double *js, double *out) {
  CCol J = (CCol)cols + blockIdx.x;
  const double a = col_sum((const double *)partial + 2 * J->col, 2 * SL_K, ((CU64)prefix)[J->slab],
                           ((CU64)prefix)[J->slab + 1]);
  if (threadIdx.x == 0) {
    double *st = js + J->slot * JS_STRIDE;
    const double sc = st[3];
    const double al = a * sc * sc;
    st[0] = al;
    st[1] = sc;
    J->ahist[J->step] = al;
    out[2 * J->job] = al;
    out[2 * J->job + 1] = st[2];
  }
}

// Lowest good eigenvalues of each checking job's tridiagonal T_m (alpha[1..m],
// beta[1..m-1]): thread i < L bisects eigenvalue i of T_m, thread 32+i
// eigenvalue i of T_m without its first row and column; thread 0 then applies
// the Cullum-Willoughby test and merges ghost copies (the rules of
// hphi_lanczos_good_eigenvalues).  out: [16 values, count] per job.
#define CK_EXTRA 6
__device__ int ck_sturm(const double *a, const double *b, int n, double x) {
  int c = 0;
  double q = a[0] - x;
  if (q < 0) c++;
  for (int i = 1; i < n; i++) {
    if (fabs(q) < 1e-300) q = -1e-300;
    q = (a[i] - x) - b[i - 1] * b[i - 1] / q;
    if (q < 0) c++;
  }
  return c;
}

#define CK_LMAX 256  // largest window of T_m eigenvalues examined
__global__ __launch_bounds__(64) void k_sl_check(const ColDev *cols, int want, double *out) {
  CCol J = (CCol)cols + blockIdx.x;
  if (!J->check) return;
  const int m = J->step;
  const double *a = J->ahist + 1, *b = J->bhist + 1;
  __shared__ double ev[CK_LMAX], evh[CK_LMAX];
  __shared__ double lo_s, hi_s, scale_s;
  __shared__ int next_s;
  const int cap = m < CK_LMAX ? m : CK_LMAX;
  if (threadIdx.x == 0) {
    double lo = a[0], hi = a[0], sc = 0.0;
    for (int i = 0; i < m; i++) {
      const double r = (i > 0 ? fabs(b[i - 1]) : 0.0) + (i < m - 1 ? fabs(b[i]) : 0.0);
      lo = fmin(lo, a[i] - r);
      hi = fmax(hi, a[i] + r);
      sc = fmax(sc, fabs(a[i]) + r);
    }
    const double span = hi - lo;
    lo_s = lo - 1e-12 * span - 1e-300;
    hi_s = hi + 1e-12 * span + 1e-300;
    scale_s = sc;
  }
  __syncthreads();
  const int t = threadIdx.x;
  // Without reorthogonalisation converged eigenvalues keep gaining ghost copies,
  // which fill a fixed window of T_m's lowest eigenvalues as m grows: the window
  // starts at want + CK_EXTRA and doubles until it holds `want` good values or
  // reaches min(m, CK_LMAX).  Each round bisects only the new indices.
  int L0 = 0, L = want + CK_EXTRA < cap ? want + CK_EXTRA : cap;
  for (;;) {
    const int Lh = L < m - 1 ? L : m - 1, Lh0 = L0 < m - 1 ? L0 : m - 1;
    const bool hat = t >= 32;
    const int lane = t & 31;
    const double *aa = hat ? a + 1 : a, *bb = hat ? b + 1 : b;
    const int n = hat ? m - 1 : m;
    for (int idx = (hat ? Lh0 : L0) + lane; idx < (hat ? Lh : L); idx += 32) {
      double l = lo_s, h = hi_s;
      for (int it = 0; it < 200; it++) {
        const double mid = 0.5 * (l + h);
        if (mid <= l || mid >= h) break;
        if (ck_sturm(aa, bb, n, mid) > idx) h = mid;
        else l = mid;
      }
      if (hat) evh[idx] = 0.5 * (l + h);
      else ev[idx] = 0.5 * (l + h);
    }
    __syncthreads();
    if (t == 0) {
      double *o = out + (uint64_t)J->job * 17;
      const double tol = 1e-11 * (scale_s > 1.0 ? scale_s : 1.0);
      int got = 0;
      if (m == 1) {
        o[0] = a[0];
        got = 1;
      }
      for (int i = 0; i < L && got < want && m > 1; i++) {
        int mult = 0;
        for (int j = 0; j < L; j++)
          if (fabs(ev[j] - ev[i]) <= tol) mult++;
        bool spurious = false;
        if (mult == 1 && i < L - 1)
          for (int j = 0; j < Lh; j++)
            if (fabs(evh[j] - ev[i]) <= tol) spurious = true;
        if (spurious) continue;
        if (got > 0 && fabs(ev[i] - o[got - 1]) <= tol) continue;
        o[got++] = ev[i];
      }
      o[16] = (double)got;
      next_s = (got >= want || L >= cap) ? 0 : (2 * L < cap ? 2 * L : cap);
    }
    __syncthreads();
    const int nx = next_s;
    if (nx == 0) break;
    L0 = L;
    L = nx;
    __syncthreads();  // next_s is read by every thread before thread 0 writes it again
  }
}

// start vector of a new column: R_c = SplitMix64 uniform complex (normalised
// by the first pass), Q_c = 0, slot state zero
__global__ void k_sl_start(double2 *R, double2 *Q, int col, uint64_t dim, uint64_t seed, double *js, int slot) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < JS_STRIDE) js[slot * JS_STRIDE + i] = 0.0;
  if (i >= dim) return;
  auto mix = [](uint64_t x) {
    uint64_t y = x + 0x9E3779B97F4A7C15ull;
    y = (y ^ (y >> 30)) * 0xBF58476D1CE4E5B9ull;
    y = (y ^ (y >> 27)) * 0x94D049BB133111EBull;
    return y ^ (y >> 31);
  };
  const uint64_t s0 = seed * 0x9E3779B97F4A7C15ull;
  const double u1 = (double)(mix(s0 ^ (2 * i)) >> 11) * (1.0 / 9007199254740992.0);
  const double u2 = (double)(mix(s0 ^ (2 * i + 1)) >> 11) * (1.0 / 9007199254740992.0);
  R[i * SL_K + col] = make_double2(2.0 * u1 - 1.0, 2.0 * u2 - 1.0);
  Q[i * SL_K + col] = make_double2(0.0, 0.0);
}

// ==================================================================== host
namespace {

struct Arena {
  char *base = nullptr;
  size_t size = 0;
  std::map<size_t, size_t> fre;  // offset -> length
  void init(char *b, size_t n) {
    base = b;
    size = n;
    fre.clear();
    fre[0] = n;
  }
  bool alloc(size_t n, size_t *off) {
    n = (n + 255) & ~(size_t)255;
    for (auto it = fre.begin(); it != fre.end(); ++it) {
      if (it->second >= n) {
        *off = it->first;
        const size_t rest = it->second - n, o = it->first;
        fre.erase(it);
        if (rest) fre[o + n] = rest;
        return true;
      }
    }
    return false;
  }
  void release(size_t off, size_t n) {
    n = (n + 255) & ~(size_t)255;
    auto it = fre.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != fre.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      fre.erase(nx);
    }
    if (it != fre.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        fre.erase(it);
      }
    }
  }
};

size_t al256(size_t n) { return (n + 255) & ~(size_t)255; }

struct Slab;

struct Job {
  int id = 0;
  uint64_t seed = 0;
  double t_arrival = 0, t_admit = 0, t_done = 0;
  hphi_engine h;  // host-side preparation only (no device members used)
  uint64_t shash = 0;  // structure_hash(h)
  Slab *slab = nullptr;
  int col = -1, slot = -1;
  int realc = 1;
  std::vector<double> alpha, beta;
  int step = 0, nprev = 0, tick_steps = 0;
  double prev[16];
  hphi_job_result res{};
};

// The engine pads each term group (low, high, mixed) to the plain kernels'
// batch; the slab kernel reads SL_U terms per batch, so its groups keep only
// the padding that batch needs.  A padding term never matches (P_EQ with an
// empty mask and a nonzero value); no compiled term has that form.
void slab_repad(hphi_engine &h) {
  const int n[3] = {h.n_low, h.n_high, h.n_mixed};
  std::vector<GTerm> out;
  int m[3], at = 0;
  for (int g = 0; g < 3; g++) {
    const size_t start = out.size();
    for (int t = at; t < at + n[g]; t++) {
      const GTerm &x = h.h_terms_dev[t];
      if (!(x.kind == P_EQ && x.tmask == 0 && x.tval != 0)) out.push_back(x);
    }
    GTerm z{};
    z.kind = P_EQ;
    z.tmask = 0;
    z.tval = 1;
    while ((out.size() - start) % SL_U) out.push_back(z);
    m[g] = (int)(out.size() - start);
    at += n[g];
  }
  h.h_terms_dev.swap(out);
  h.n_low = m[0];
  h.n_high = m[1];
  h.n_mixed = m[2];
}

// FNV-1a over what same_structure compares: equal structures hash equally
uint64_t structure_hash(const hphi_engine &h) {
  uint64_t x = 1469598103934665603ull;
  auto mix = [&x](uint64_t v) {
    for (int b = 0; b < 8; b++) {
      x ^= (v >> (8 * b)) & 0xffu;
      x *= 1099511628211ull;
    }
  };
  mix((uint64_t)h.model);
  mix((uint64_t)h.nsite);
  mix((uint64_t)h.nup);
  mix((uint64_t)h.ndown);
  mix(h.dim);
  mix((uint64_t)h.n_low);
  mix((uint64_t)h.n_high);
  mix((uint64_t)h.n_mixed);
  mix((uint64_t)h.h_terms_dev.size());
  mix((uint64_t)h.h_dterms.size());
  for (const GTerm &t : h.h_terms_dev) {
    mix(t.flip);
    mix(t.tmask);
    mix(t.tval);
    mix(t.sgn);
    mix((uint64_t)t.kind);
  }
  return x;
}

// Jobs share a slab when basis and term structure agree (couplings may differ).
bool same_structure(const hphi_engine &a, const hphi_engine &b) {
  if (a.model != b.model || a.nsite != b.nsite || a.nup != b.nup || a.ndown != b.ndown || a.dim != b.dim ||
      a.n_low != b.n_low || a.n_high != b.n_high || a.n_mixed != b.n_mixed ||
      a.h_terms_dev.size() != b.h_terms_dev.size() || a.h_dterms.size() != b.h_dterms.size())
    return false;
  for (size_t t = 0; t < a.h_terms_dev.size(); t++) {
    const GTerm &x = a.h_terms_dev[t], &y = b.h_terms_dev[t];
    if (x.flip != y.flip || x.tmask != y.tmask || x.tval != y.tval || x.sgn != y.sgn || x.kind != y.kind)
      return false;
  }
  return true;
}

struct Slab {
  hphi_engine proto;  // structure: the first job's preparation (host side)
  uint64_t shash = 0;  // structure_hash(proto)
  size_t off = 0, bytes = 0;
  uint64_t *list1 = nullptr, *l22 = nullptr, *jb = nullptr;
  uint32_t *l21 = nullptr;
  uint16_t *l21p = nullptr;  // list_2_1 packed for the LDS copy of every workgroup
  GTerm *terms = nullptr;
  double2 *coef = nullptr;
  double *diag = nullptr;
  double2 *R = nullptr, *Q = nullptr;
  DTerm *dterms = nullptr;  // [SL_K][nd]
  double *hist = nullptr;   // [SL_K][2][max_iter + 2]
  std::vector<double2> coef_host;
  uint64_t nnz = 0;  // off-diagonal elements of one column's H (hphi_term_nnz)
  Job *col[SL_K] = {};
  int used() const {
    int n = 0;
    for (int c = 0; c < SL_K; c++) n += col[c] != nullptr;
    return n;
  }
};

size_t slab_footprint(const hphi_engine &h, int max_iter) {
  const size_t nt = std::max<size_t>(h.h_terms_dev.size(), 1), nd = std::max<size_t>(h.h_dterms.size(), 1);
  return al256(8 * h.dim) + al256(4 * h.size21) + al256(16 * ((h.size21 + 7) / 8)) + al256(8 * h.size22) + al256(8 * (h.sdim + 1)) +
         al256(sizeof(GTerm) * nt) + al256(16 * SL_K * nt) + al256(8 * SL_K * h.dim) + 2 * al256(16 * SL_K * h.dim) +
         al256(sizeof(DTerm) * SL_K * nd) + al256(8 * SL_K * 2 * (size_t)(max_iter + 2));
}

double mono_now() {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

}  // namespace

struct hphi_server {
  hphi_server_params p{};
  hipStream_t stream = nullptr;
  Arena arena;
  char *d_arena = nullptr;
  double *d_js = nullptr, *d_out = nullptr, *h_out = nullptr, *d_part = nullptr;
  double *d_ev = nullptr, *h_ev = nullptr;  // per-job check results [17]
  SlabDev *d_slabs = nullptr;
  ColDev *d_cols = nullptr;
  uint64_t *d_prefix = nullptr;
  char *h_stage = nullptr;  // pinned staging for uploads
  size_t stage_size = 0, stage_used = 0;
  std::deque<Job *> waiting;
  std::vector<Job *> active;
  std::vector<Slab *> slabs;
  std::vector<int> free_slots;
  std::vector<hphi_job_result> done;
  int next_id = 0;
  double t0 = 0;
  size_t max_tiles = 0;
  double st_admit = 0, st_gpu = 0, st_check = 0;  // seconds spent per phase
  uint64_t st_ticks = 0, st_job_steps = 0, st_slab_steps = 0;
  // slab H·v launches: HIP-event time, algorithmic bytes, launches
  hipEvent_t ev_hv[2 * SL_SMAX] = {};
  double st_hv_s = 0, st_hv_bytes = 0;
  uint64_t st_hv_launches = 0;
  uint64_t st_moves = 0;  // columns moved between slabs (compact)
};

extern "C" double hphi_server_now(const hphi_server *s) { return s ? mono_now() - s->t0 : 0.0; }

extern "C" int hphi_server_create(const hphi_server_params *p, hphi_server **out) {
  if (!p || !out) return hphi_set_error("hphi_server_create: NULL argument");
  *out = nullptr;
  if (hphi_ensure_binom_table()) return -1;
  hphi_server *s = new hphi_server();
  s->p = *p;
  if (s->p.max_jobs <= 0) s->p.max_jobs = 64;
  if (s->p.max_jobs > 4096) s->p.max_jobs = 4096;
  if (s->p.n_eigen <= 0) s->p.n_eigen = 4;
  if (s->p.n_eigen > 16) s->p.n_eigen = 16;
  if (s->p.eps <= 0) s->p.eps = 1e-12;
  if (s->p.max_iter <= 0) s->p.max_iter = 2000;
  if (s->p.check_every <= 0) s->p.check_every = 5;
  const size_t budget = s->p.hbm_budget ? s->p.hbm_budget : ((size_t)16 << 30);
  s->max_tiles = (size_t)1 << 22;  // 4M tiles x 256 states = 1.07e9 states in flight
  int rc = 0;
  const int mj = s->p.max_jobs;
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&s->d_arena, budget) != hipSuccess ||
      hipMalloc(&s->d_js, sizeof(double) * JS_STRIDE * mj) != hipSuccess ||
      hipMalloc(&s->d_out, sizeof(double) * 2 * mj * SL_SMAX) != hipSuccess ||
      hipMalloc(&s->d_slabs, sizeof(SlabDev) * mj * SL_SMAX) != hipSuccess ||
      hipMalloc(&s->d_cols, sizeof(ColDev) * mj * SL_SMAX) != hipSuccess ||
      hipMalloc(&s->d_prefix, sizeof(uint64_t) * (mj + 1) * SL_SMAX) != hipSuccess ||
      hipMalloc(&s->d_part, sizeof(double2) * SL_K * s->max_tiles) != hipSuccess ||
      hipHostMalloc(&s->h_out, sizeof(double) * 2 * mj * SL_SMAX, hipHostMallocDefault) != hipSuccess ||
      hipMalloc(&s->d_ev, sizeof(double) * 17 * mj) != hipSuccess ||
      hipHostMalloc(&s->h_ev, sizeof(double) * 17 * mj, hipHostMallocDefault) != hipSuccess)
    rc = hphi_set_error("hphi_server_create: device allocation failed");
  s->stage_size = (size_t)16 << 20;
  if (!rc && hipHostMalloc(&s->h_stage, s->stage_size, hipHostMallocDefault) != hipSuccess)
    rc = hphi_set_error("hphi_server_create: pinned staging allocation failed");
  for (int i = 0; i < 2 * SL_SMAX && !rc; i++)
    if (hipEventCreate(&s->ev_hv[i]) != hipSuccess) rc = hphi_set_error("hphi_server_create: hipEventCreate failed");
  if (rc) {
    hphi_server_destroy(s);
    return -1;
  }
  s->arena.init(s->d_arena, budget);
  for (int i = mj - 1; i >= 0; i--) s->free_slots.push_back(i);
  s->t0 = mono_now();
  *out = s;
  return 0;
}

extern "C" void hphi_server_destroy(hphi_server *s) {
  if (!s) return;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (Job *j : s->active) delete j;
  for (Job *j : s->waiting) delete j;
  for (Slab *b : s->slabs) delete b;
  (void)hipFree(s->d_arena);
  (void)hipFree(s->d_js);
  (void)hipFree(s->d_out);
  (void)hipFree(s->d_slabs);
  (void)hipFree(s->d_cols);
  (void)hipFree(s->d_prefix);
  (void)hipFree(s->d_part);
  if (s->h_out) (void)hipHostFree(s->h_out);
  (void)hipFree(s->d_ev);
  if (s->h_ev) (void)hipHostFree(s->h_ev);
  if (s->h_stage) (void)hipHostFree(s->h_stage);
  for (int i = 0; i < 2 * SL_SMAX; i++)
    if (s->ev_hv[i]) (void)hipEventDestroy(s->ev_hv[i]);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

extern "C" int hphi_server_submit(hphi_server *s, const hphi_model_desc *d, uint64_t seed, double t_arrival,
                                  int *id) {
  if (!s || !d) return hphi_set_error("hphi_server_submit: NULL argument");
  if (d->model != HPHI_MODEL_SPIN && d->model != HPHI_MODEL_HUBBARD)
    return hphi_set_error("the service runs canonical Spin and Hubbard models");
  Job *j = new Job();
  if (hphi_engine_prepare(d, &j->h)) {
    delete j;
    return -1;
  }
  if (j->h.nbits > 32) {
    delete j;
    return hphi_set_error("the service runs states of at most 32 bits");
  }
  slab_repad(j->h);
  j->shash = structure_hash(j->h);
  for (const GTerm &t : j->h.h_terms_dev)
    if (t.cim != 0.0) j->realc = 0;
  const size_t fp = slab_footprint(j->h, s->p.max_iter);
  if (fp > s->arena.size) {
    delete j;
    return hphi_set_error("job footprint %zu exceeds the HBM budget", fp);
  }
  j->id = s->next_id++;
  j->seed = seed;
  j->t_arrival = t_arrival;
  // the queue is ordered by arrival; among jobs that arrive together the
  // largest goes first, so the last jobs to run are small ones (1,000-request
  // stream back-to-back: 171.2 -> 172.6 solves/s)
  auto before = [](const Job *a, const Job *b) {
    return a->t_arrival < b->t_arrival || (a->t_arrival == b->t_arrival && a->h.dim > b->h.dim);
  };
  s->waiting.insert(std::upper_bound(s->waiting.begin(), s->waiting.end(), j, before), j);
  if (id) *id = j->id;
  return 0;
}

// copy into pinned staging (valid until the tick's synchronisation)
static void *stage(hphi_server *s, const void *src, size_t n) {
  if (s->stage_used + n > s->stage_size) return nullptr;
  char *p = s->h_stage + s->stage_used;
  memcpy(p, src, n);
  s->stage_used += al256(n);
  return p;
}

// list_2_1 as uint16 (rank + 2 < 2^16 for <= 16 low bits), zero-padded to n entries
__global__ void k_pack21(const uint32_t *l21, uint64_t size21, uint64_t n, uint16_t *out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = i < size21 ? (uint16_t)l21[i] : (uint16_t)0;
}

// a new slab with the job's structure: tables on the device, terms uploaded;
// returns 1 when the staging buffer is full (retry next tick)
static int slab_create(hphi_server *s, const Job *j, size_t off, size_t bytes, Slab **out) {
  const hphi_engine &h = j->h;
  const size_t nt = std::max<size_t>(h.h_terms_dev.size(), 1), nd = std::max<size_t>(h.h_dterms.size(), 1);
  const size_t tb = sizeof(GTerm) * h.h_terms_dev.size();
  if (s->stage_used + al256(tb) > s->stage_size) return 1;
  Slab *b = new Slab();
  b->proto = h;
  b->nnz = hphi_term_nnz(&h);
  b->shash = j->shash;
  b->off = off;
  b->bytes = bytes;
  char *p = s->d_arena + off;
  b->list1 = (uint64_t *)p, p += al256(8 * h.dim);
  b->l21 = (uint32_t *)p, p += al256(4 * h.size21);
  b->l21p = (uint16_t *)p, p += al256(16 * ((h.size21 + 7) / 8));
  b->l22 = (uint64_t *)p, p += al256(8 * h.size22);
  b->jb = (uint64_t *)p, p += al256(8 * (h.sdim + 1));
  b->terms = (GTerm *)p, p += al256(sizeof(GTerm) * nt);
  b->coef = (double2 *)p, p += al256(16 * SL_K * nt);
  b->diag = (double *)p, p += al256(8 * SL_K * h.dim);
  b->R = (double2 *)p, p += al256(16 * SL_K * h.dim);
  b->Q = (double2 *)p, p += al256(16 * SL_K * h.dim);
  b->dterms = (DTerm *)p, p += al256(sizeof(DTerm) * SL_K * nd);
  b->hist = (double *)p;
  b->coef_host.assign(SL_K * nt, make_double2(0.0, 0.0));
  int rc = 0;
  if (tb) {
    void *st = stage(s, h.h_terms_dev.data(), tb);
    if (hipMemcpyAsync(b->terms, st, tb, hipMemcpyHostToDevice, s->stream) != hipSuccess)
      rc = hphi_set_error("slab term upload failed");
  }
  const SzParams P = hphi_sz_params(&h);
  if (!rc && hphi_launch_tables(P, h.sdim, h.size21, b->jb, b->l21, b->l22, b->list1, s->stream)) rc = -1;
  if (!rc) {
    const uint64_t n = (h.size21 + 7) / 8 * 8;
    hipLaunchKernelGGL(k_pack21, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s->stream, b->l21, h.size21, n,
                       b->l21p);
    if (hipGetLastError() != hipSuccess) rc = hphi_set_error("slab table packing failed");
  }
  if (rc) {
    delete b;
    return -1;
  }
  *out = b;
  return 0;
}

// the job takes column c of slab b: couplings, diagonal, start vector;
// returns 1 when the staging buffer is full (retry next tick)
static int column_start(hphi_server *s, Slab *b, int c, Job *j) {
  const hphi_engine &h = j->h;
  const size_t nt = h.h_terms_dev.size(), nd = h.h_dterms.size();
  const size_t cb = sizeof(double2) * b->coef_host.size(), db = sizeof(DTerm) * nd;
  if (s->stage_used + al256(cb) + al256(db) > s->stage_size) return 1;
  for (size_t t = 0; t < nt; t++)
    b->coef_host[t * SL_K + c] = make_double2(h.h_terms_dev[t].cre, h.h_terms_dev[t].cim);
  HIPCHK(hipMemcpyAsync(b->coef, stage(s, b->coef_host.data(), cb), cb, hipMemcpyHostToDevice, s->stream));
  DTerm *dt = b->dterms + (size_t)c * std::max<size_t>(nd, 1);
  if (db) HIPCHK(hipMemcpyAsync(dt, stage(s, h.h_dterms.data(), db), db, hipMemcpyHostToDevice, s->stream));
  if (hphi_launch_diag(h.dim, b->list1, 0, dt, (int)nd, b->diag, s->stream, SL_K, c)) return -1;
  const unsigned g = (unsigned)((h.dim + 255) / 256);
  hipLaunchKernelGGL(k_sl_start, dim3(g), dim3(256), 0, s->stream, b->R, b->Q, c, h.dim, j->seed, s->d_js, j->slot);
  HIPCHK(hipGetLastError());
  b->col[c] = j;
  j->slab = b;
  j->col = c;
  j->step = 0;
  j->alpha.assign(1, 0.0);
  j->beta.assign(1, 0.0);
  return 0;
}

// moves column cs of one slab into column cd of another slab of the same
// structure: the Lanczos vectors R and Q, the diagonal, the couplings and the
// alpha/beta history (nh doubles per column)
__global__ void k_sl_move(const double2 *R0, const double2 *Q0, const double *d0, const double2 *c0, const double *h0,
                          double2 *R1, double2 *Q1, double *d1, double2 *c1, double *h1, int cs, int cd,
                          uint64_t dim, uint64_t nt, uint64_t nh) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < dim) {
    R1[i * SL_K + cd] = R0[i * SL_K + cs];
    Q1[i * SL_K + cd] = Q0[i * SL_K + cs];
    d1[i * SL_K + cd] = d0[i * SL_K + cs];
  }
  if (i < nt) c1[i * SL_K + cd] = c0[i * SL_K + cs];
  if (i < nh) h1[(uint64_t)cd * nh + i] = h0[(uint64_t)cs * nh + i];
}

static void retire(hphi_server *s, Job *j, double now, int converged, const double *ev, int n) {
  j->res.id = j->id;
  j->res.iterations = j->step;
  j->res.converged = converged;
  j->res.n_found = n;
  for (int i = 0; i < n && i < 16; i++) j->res.eigenvalues[i] = ev[i];
  j->res.t_arrival = j->t_arrival;
  j->res.t_admit = j->t_admit;
  j->res.t_done = now;
  j->res.dim = j->h.dim;
  s->done.push_back(j->res);
  s->free_slots.push_back(j->slot);
  j->slab->col[j->col] = nullptr;
}

// admission in arrival order: a free column of a slab of the same structure,
// else a new slab while the arena has room
// The job takes column c of running slab b (slot and bookkeeping); 1: retry
// next tick (staging full)
static int admit_into(hphi_server *s, Job *j, Slab *b, int c, double now) {
  j->slot = s->free_slots.back();
  const int rc = column_start(s, b, c, j);
  if (rc) return rc;
  s->free_slots.pop_back();
  j->t_admit = now;
  s->active.push_back(j);
  return 0;
}

// Free columns of running slabs go first to arrived jobs of their structure
// among the first ADMIT_WINDOW waiting jobs, ahead of queue order: a slab step
// costs about the same however many of its columns run, so full slabs advance
// more jobs per launch.  The rest is admitted in arrival order (admit).
#define ADMIT_WINDOW 256
static int admit_fill(hphi_server *s, double now) {
  for (Slab *b : s->slabs) {
    while (b->used() < SL_K && !s->free_slots.empty()) {
      int c = -1;
      for (int q = SL_K - 1; q >= 0; q--)
        if (!b->col[q]) c = q;
      auto it = s->waiting.begin();
      int seen = 0;
      for (; it != s->waiting.end() && seen < ADMIT_WINDOW; ++it, ++seen) {
        const Job *w = *it;
        if (w->t_arrival <= now && w->shash == b->shash && same_structure(b->proto, w->h)) break;
      }
      if (it == s->waiting.end() || seen == ADMIT_WINDOW) break;
      Job *j = *it;
      const int rc = admit_into(s, j, b, c, now);
      if (rc == 1) return 0;
      if (rc) return -1;
      s->waiting.erase(it);
    }
  }
  return 0;
}

static int admit(hphi_server *s, double now) {
  if (admit_fill(s, now)) return -1;
  while (!s->waiting.empty() && !s->free_slots.empty()) {
    Job *j = s->waiting.front();
    if (j->t_arrival > now) break;
    Slab *b = nullptr;
    int c = -1;
    for (Slab *x : s->slabs) {
      if (x->used() == SL_K || x->shash != j->shash || !same_structure(x->proto, j->h)) continue;
      b = x;
      for (int q = SL_K - 1; q >= 0; q--)
        if (!x->col[q]) c = q;
      break;
    }
    bool fresh = false;
    if (!b) {
      const size_t bytes = slab_footprint(j->h, s->p.max_iter);
      size_t off;
      if (!s->arena.alloc(bytes, &off)) break;
      const int rc = slab_create(s, j, off, bytes, &b);
      if (rc) {
        s->arena.release(off, bytes);
        if (rc == 1) break;
        return -1;
      }
      s->slabs.push_back(b);
      c = 0;
      fresh = true;
    }
    j->slot = s->free_slots.back();
    const int rc = column_start(s, b, c, j);
    if (rc) {
      if (fresh) {  // nothing was launched on the new slab's columns: release it
        s->slabs.pop_back();
        s->arena.release(b->off, b->bytes);
        delete b;
      }
      if (rc == 1) break;
      return -1;
    }
    s->free_slots.pop_back();
    s->waiting.pop_front();
    j->t_admit = now;
    s->active.push_back(j);
  }
  return 0;
}

// Slabs of one structure whose running columns fit in fewer slabs are packed:
// columns move from the emptiest slabs into the fullest ones with free
// columns, so a structure whose queue has drained runs ceil(columns / SL_K)
// slab steps per step instead of one per slab.  Structures with arrived
// waiting jobs are left alone (admission fills their free columns).  A moved
// column keeps its arithmetic (same tiles, same per-column reductions, and a
// real column computes the same values on the complex path), so its results do
// not change.
static int compact(hphi_server *s, double now) {
  std::vector<uint64_t> busy;  // structures with arrived waiting jobs in the admission window
  int seen = 0;
  for (const Job *w : s->waiting) {
    if (seen++ == ADMIT_WINDOW) break;
    if (w->t_arrival <= now) busy.push_back(w->shash);
  }
  std::map<uint64_t, std::vector<Slab *>> groups;
  for (Slab *b : s->slabs) {
    const int u = b->used();
    if (u > 0 && u < SL_K && std::find(busy.begin(), busy.end(), b->shash) == busy.end())
      groups[b->shash].push_back(b);
  }
  const uint64_t nh = 2 * ((uint64_t)s->p.max_iter + 2);
  for (auto &g : groups) {
    std::vector<Slab *> &L = g.second;
    int cols = 0;
    for (const Slab *b : L) cols += b->used();
    if ((size_t)((cols + SL_K - 1) / SL_K) >= L.size()) continue;  // no slab to free
    std::sort(L.begin(), L.end(), [](const Slab *a, const Slab *b) { return a->used() > b->used(); });
    size_t i = 0, k = L.size() - 1;
    while (i < k) {
      Slab *A = L[i], *B = L[k];
      if (A->used() == SL_K) { i++; continue; }
      if (B->used() == 0 || !same_structure(A->proto, B->proto)) { k--; continue; }
      int cs = 0, cd = 0;
      while (!B->col[cs]) cs++;
      while (A->col[cd]) cd++;
      const hphi_engine &h = A->proto;
      const uint64_t nt = h.h_terms_dev.size();
      const uint64_t n = std::max<uint64_t>(std::max<uint64_t>(h.dim, nt), nh);
      hipLaunchKernelGGL(k_sl_move, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s->stream, B->R, B->Q,
                         B->diag, B->coef, B->hist, A->R, A->Q, A->diag, A->coef, A->hist, cs, cd, h.dim, nt, nh);
      HIPCHK(hipGetLastError());
      for (uint64_t t = 0; t < nt; t++) A->coef_host[t * SL_K + cd] = B->coef_host[t * SL_K + cs];
      Job *j = B->col[cs];
      B->col[cs] = nullptr;
      A->col[cd] = j;
      j->slab = A;
      j->col = cd;
      s->st_moves++;
    }
  }
  return 0;
}

extern "C" int hphi_server_tick(hphi_server *s, int *n_active, int *n_waiting) {
  if (!s) return hphi_set_error("hphi_server_tick: NULL argument");
  double now = hphi_server_now(s);
  const double ta = now;
  s->stage_used = 0;
  if (admit(s, now)) return -1;
  s->st_admit += hphi_server_now(s) - ta;
  if (!s->active.empty()) {
    const double tg = hphi_server_now(s);
    // 2. S Lanczos steps for every running job, packed launches over the slabs
    const int S = std::min(s->p.check_every, SL_SMAX);
    const int n = (int)s->active.size();
    std::vector<int> nst(n);  // steps of each job this tick
    for (int i = 0; i < n; i++) {
      Job *j = s->active[i];
      const uint64_t lim = std::min<uint64_t>(j->h.dim, (uint64_t)s->p.max_iter);
      nst[i] = (int)std::min<uint64_t>((uint64_t)S, lim - (uint64_t)j->step);
      j->tick_steps = nst[i];
    }
    std::vector<int> s_chk(S, 0);  // jobs whose last step of the tick is step st
    for (int i = 0; i < n; i++) s_chk[nst[i] - 1]++;
    std::vector<SlabDev> sd;
    std::vector<ColDev> cd;
    std::vector<uint64_t> prefix;
    std::vector<int> s_ns(S), s_nc(S), s_sd0(S), s_cd0(S), s_pf0(S);
    // algorithmic bytes of each step's slab H·v: per slab 8 dim (list_1), per
    // running column 56 dim (own row, diagonal, Q read and write) + 16 per
    // off-diagonal element (the gathered row entry)
    std::vector<double> s_bytes(S, 0.0);
    std::map<const Slab *, int> index;
    const size_t hl = (size_t)s->p.max_iter + 2;
    uint64_t tiles = 0, max21 = 0;
    for (int st = 0; st < S; st++) {
      s_sd0[st] = (int)sd.size();
      s_cd0[st] = (int)cd.size();
      s_pf0[st] = (int)prefix.size();
      prefix.push_back(0);
      index.clear();
      for (Slab *b : s->slabs) {
        const hphi_engine &h = b->proto;
        SlabDev d{};
        for (int c = 0; c < SL_K; c++) {
          const Job *j = b->col[c];
          d.slot[c] = j ? j->slot : 0;
          if (j && j->tick_steps > st) {
            d.mask |= 1 << c;
            d.realc |= j->realc << c;
          }
        }
        if (!d.mask) continue;
        d.dim = h.dim;
        d.list1 = b->list1;
        d.l21p = (const uint4 *)b->l21p;
        d.l22 = b->l22;
        d.terms = b->terms;
        d.coef = b->coef;
        d.diag = b->diag;
        d.R = b->R;
        d.Q = b->Q;
        d.irght = h.irght;
        d.size21 = h.size21;
        d.n_low = h.n_low;
        d.n_high = h.n_high;
        d.n_mixed = h.n_mixed;
        d.nlow = h.nlow;
        index[b] = (int)sd.size() - s_sd0[st];
        s_bytes[st] += 8.0 * (double)h.dim +
                       (double)__builtin_popcount(d.mask) * (56.0 * (double)h.dim + 16.0 * (double)b->nnz);
        sd.push_back(d);
        prefix.push_back(prefix.back() + (h.dim + (uint64_t)SL_TPW * SV_TPB - 1) / ((uint64_t)SL_TPW * SV_TPB));
        max21 = std::max<uint64_t>(max21, h.size21);
        std::swap(b->R, b->Q);  // roles after this step
      }
      for (int i = 0; i < n; i++) {
        if (nst[i] <= st) continue;
        Job *j = s->active[i];
        double *hb = j->slab->hist + (size_t)j->col * 2 * hl;
        cd.push_back(ColDev{index[j->slab], j->col, j->slot, j->step + st + 1, st == nst[i] - 1, i, hb, hb + hl});
      }
      s_ns[st] = (int)sd.size() - s_sd0[st];
      s_nc[st] = (int)cd.size() - s_cd0[st];
      tiles = std::max<uint64_t>(tiles, prefix.back());
    }
    if (tiles > s->max_tiles) return hphi_set_error("too many tiles in flight");
    const size_t lds = 16 * (size_t)((max21 + 7) / 8);  // list_2_1 as uint16 (rank + 2 < 2^16 for <= 16 low bits)
    if (lds > 65536)
      (void)hipFuncSetAttribute((const void *)k_sl_hv, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    HIPCHK(hipMemcpyAsync(s->d_slabs, sd.data(), sizeof(SlabDev) * sd.size(), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(s->d_cols, cd.data(), sizeof(ColDev) * cd.size(), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(s->d_prefix, prefix.data(), sizeof(uint64_t) * prefix.size(), hipMemcpyHostToDevice,
                          s->stream));
    for (int st = 0; st < S; st++) {
      const int ns = s_ns[st], nc = s_nc[st];
      if (!nc) continue;
      const SlabDev *dsl = s->d_slabs + s_sd0[st];
      const ColDev *dcl = s->d_cols + s_cd0[st];
      const uint64_t *dpf = s->d_prefix + s_pf0[st];
      const unsigned T = (unsigned)((prefix[s_pf0[st] + ns] + 7) / 8 * 8);  // grid of sl_chunk()
      double *out = s->d_out + (size_t)st * 2 * n;
      hipLaunchKernelGGL(k_sl_resid, dim3(T), dim3(SV_TPB), 0, s->stream, dsl, dpf, ns, (const double *)s->d_js,
                         s->d_part);
      hipLaunchKernelGGL(k_sl_norm, dim3(nc), dim3(256), 0, s->stream, dcl, dpf, (const double *)s->d_part,
                         s->d_js);
      HIPCHK(hipEventRecord(s->ev_hv[2 * st], s->stream));
      hipLaunchKernelGGL(k_sl_hv, dim3(T), dim3(SV_TPB), lds, s->stream, dsl, dpf, ns, (const double *)s->d_js,
                         (double2 *)s->d_part);
      HIPCHK(hipEventRecord(s->ev_hv[2 * st + 1], s->stream));
      hipLaunchKernelGGL(k_sl_alpha, dim3(nc), dim3(256), 0, s->stream, dcl, dpf, (const double2 *)s->d_part,
                         s->d_js, out);
      if (s_chk[st]) hipLaunchKernelGGL(k_sl_check, dim3(nc), dim3(64), 0, s->stream, dcl, s->p.n_eigen, s->d_ev);
      HIPCHK(hipGetLastError());
      s->st_slab_steps += ns;
    }
    HIPCHK(hipMemcpyAsync(s->h_out, s->d_out, sizeof(double) * 2 * n * S, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipMemcpyAsync(s->h_ev, s->d_ev, sizeof(double) * 17 * n, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    now = hphi_server_now(s);
    s->st_gpu += now - tg;
    for (int st = 0; st < S; st++) {
      if (!s_nc[st]) continue;
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, s->ev_hv[2 * st], s->ev_hv[2 * st + 1]));
      s->st_hv_s += 1e-3 * ms;
      s->st_hv_bytes += s_bytes[st];
      s->st_hv_launches++;
    }
    s->st_ticks++;
    // 3. convergence and retirement
    std::vector<Job *> keep;
    keep.reserve(n);
    for (int i = 0; i < n; i++) {
      Job *j = s->active[i];
      s->st_job_steps += nst[i];
      bool broken = false;
      for (int st = 0; st < nst[i]; st++) {
        const double *o = s->h_out + ((size_t)st * n + i) * 2;
        const int m = j->step + 1;
        j->alpha.push_back(o[0]);                  // alpha[m]
        if (m >= 2) j->beta.push_back(o[1]);      // beta[m-1]
        j->step = m;
        // beta[m-1] ~ 0: the recurrence ended at step m-1 (alpha[m] and later are void)
        if (m >= 2 && j->beta[m - 1] < 1e-14) {
          broken = true;
          j->step = m - 1;
          break;
        }
      }
      const int m = j->step;
      const bool exhausted = broken || (uint64_t)m >= j->h.dim;
      double cur[16];
      int nc;
      if (broken) {
        nc = hphi_lanczos_good_eigenvalues(j->alpha.data() + 1, j->beta.data() + 1, m, s->p.n_eigen, cur);
      } else {  // the check after the job's last step of the tick ran on the device
        const double *o = s->h_ev + 17 * (size_t)i;
        nc = (int)o[16];
        for (int q = 0; q < nc; q++) cur[q] = o[q];
        // the device window stops at CK_LMAX values of T_m; with more ghost copies
        // than that below the wanted values the host filter widens it up to m
        if (nc < (int)std::min<uint64_t>((uint64_t)s->p.n_eigen, j->h.dim) && m > CK_LMAX)
          nc = hphi_lanczos_good_eigenvalues(j->alpha.data() + 1, j->beta.data() + 1, m, s->p.n_eigen, cur);
      }
      const int want = (int)std::min<uint64_t>((uint64_t)s->p.n_eigen, j->h.dim);
      // converged: the values stopped moving and their Ritz residuals are small
      // (T_{m-1} with beta_{m-1}: beta_m comes with the next step)
      bool conv = (exhausted && nc >= 1);
      if (!conv && hphi_lanczos_converged(j->prev, j->nprev, cur, nc, want, s->p.eps))
        conv = m < 2 || hphi_lanczos_residuals_small(j->alpha.data() + 1, j->beta.data() + 1, m - 1, j->beta[m - 1],
                                                     cur, want, s->p.eps);
      if (conv || m >= s->p.max_iter) {
        retire(s, j, now, conv ? 1 : 0, cur, nc);
        delete j;
        continue;
      }
      memcpy(j->prev, cur, sizeof cur);
      j->nprev = nc;
      keep.push_back(j);
    }
    s->active.swap(keep);
    if (compact(s, now)) return -1;
    // slabs whose last job left give their memory back
    std::vector<Slab *> live;
    for (Slab *b : s->slabs) {
      if (b->used()) {
        live.push_back(b);
      } else {
        s->arena.release(b->off, b->bytes);
        delete b;
      }
    }
    s->slabs.swap(live);
    s->st_check += hphi_server_now(s) - now;
  }
  if (n_active) *n_active = (int)s->active.size();
  if (n_waiting) *n_waiting = (int)s->waiting.size();
  return 0;
}

extern "C" int hphi_server_stats(const hphi_server *s, double *out) {
  if (!s || !out) return hphi_set_error("hphi_server_stats: NULL argument");
  out[0] = s->st_admit;
  out[1] = s->st_gpu;
  out[2] = s->st_check;
  out[3] = (double)s->st_ticks;
  out[4] = (double)s->st_job_steps;
  out[5] = (double)s->st_slab_steps;
  out[6] = s->st_hv_s;
  out[7] = s->st_hv_bytes;
  out[8] = (double)s->st_hv_launches;
  out[9] = (double)s->st_moves;
  return 0;
}

extern "C" int hphi_server_results(hphi_server *s, hphi_job_result *out, int max, int *n) {
  if (!s || !n) return hphi_set_error("hphi_server_results: NULL argument");
  const int k = (int)std::min<size_t>(s->done.size(), (size_t)std::max(max, 0));
  for (int i = 0; i < k; i++) out[i] = s->done[i];
  s->done.erase(s->done.begin(), s->done.begin() + k);
  *n = k;
  return 0;
}
