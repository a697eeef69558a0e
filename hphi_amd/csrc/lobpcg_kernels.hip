// Streaming and reduction kernels of the block LOBPCG driver (lobpcg.c).
//
// The k vectors of each LOBPCG block (w, x, p and their images W, X, P) are
// stored state-major, element j*k + c, so the block H·v applies to them
// directly.  Normalisations of the reference (src/CalcByLOBPCG.c:430-455,
// 508-556) are carried as per-column scale factors by the host loop (lobpcg.c) instead of
// separate passes over the vectors: a stored column times its scale is the
// reference's vector.  Every reduction writes per-workgroup partials that a
// fixed-order sum finishes, so results do not depend on scheduling.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "engine_internal.h"
#include "hphi_amd.h"

#define HIPCHK(x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) return hphi_set_error("%s failed: %s", #x, hipGetErrorString(e_)); \
  } while (0)

#define LB_TPB 256
#define LB_GRID 1024
#define LB_GRAM_GRID 4096  // workgroups of the Gram pass: its load and compute phases overlap across more groups (cfg4 block 8: 2 % per iteration over 1024)

struct LbCols {
  double a[16], b[16];
};

// per-column sums of lanes with equal (lane % K), then over the waves
template <int K>
__device__ __forceinline__ void lb_colsum_store(double v, double *partial) {
  for (int off = 32; off >= K; off >>= 1) v += __shfl_xor(v, off, 64);
  __shared__ double red[LB_TPB / 64][K];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane < K) red[wave][lane] = v;
  __syncthreads();
  if (threadIdx.x < K) {
    double s = 0.0;
    for (int w = 0; w < LB_TPB / 64; w++) s += red[w][threadIdx.x];
    partial[(uint64_t)blockIdx.x * K + threadIdx.x] = s;
  }
}

// w = s_c (X - mu_c x), |w_c|^2   (CalcByLOBPCG.c:430-433)
template <int K>
__global__ __launch_bounds__(LB_TPB) void k_lb_resid(const double2 *X, const double2 *x, double2 *w, uint64_t dim,
                                                     LbCols cf, double *partial) {
  const int c = threadIdx.x % K;
  const double mu = cf.a[c], s = cf.b[c];
  double acc = 0.0;
  const uint64_t n = dim * K;
  for (uint64_t i = (uint64_t)blockIdx.x * LB_TPB + threadIdx.x; i < n; i += (uint64_t)gridDim.x * LB_TPB) {
    const double2 hx = X[i], xx = x[i];
    const double2 r = make_double2(s * (hx.x - mu * xx.x), s * (hx.y - mu * xx.y));
    w[i] = r;
    acc += r.x * r.x + r.y * r.y;
  }
  lb_colsum_store<K>(acc, partial);
}

// point-Jacobi preconditioner w /= (D_j - shift_c) where |D_j - shift_c| > eps, |w_c|^2
// (CalcByLOBPCG.c:440-447)
template <int K>
__global__ __launch_bounds__(LB_TPB) void k_lb_precond(double2 *w, const double *diag, uint64_t dim, LbCols cf,
                                                       double eps, double *partial) {
  const int c = threadIdx.x % K;
  const double shift = cf.a[c];
  double acc = 0.0;
  const uint64_t n = dim * K;
  for (uint64_t i = (uint64_t)blockIdx.x * LB_TPB + threadIdx.x; i < n; i += (uint64_t)gridDim.x * LB_TPB) {
    double2 v = w[i];
    const double pc = diag[i / K] - shift;
    if (fabs(pc) > eps) {
      v.x /= pc;
      v.y /= pc;
      w[i] = v;
    }
    acc += v.x * v.x + v.y * v.y;
  }
  lb_colsum_store<K>(acc, partial);
}

// Gram blocks of the stored vectors: G[r][c] = <b_r | y_c>, r over the 3K
// columns of (w, x, p), c over the 6K columns of (w, x, p, W, X, P)
// (CalcByLOBPCG.c:479-490, all pairs in one pass).  A workgroup stages a tile
// of states in LDS; each thread owns a 3 x 6 block of outputs and a subset of
// the tile's states.  Output: partial[blk][r][c] complex.
template <int K>
__global__ __launch_bounds__(LB_TPB) void k_lb_gram(const double2 *w, const double2 *x, const double2 *p,
                                                    const double2 *W, const double2 *X, const double2 *P,
                                                    uint64_t dim, double2 *partial) {
  constexpr int R = 3 * K, C = 6 * K;
  constexpr int MB = K * K;                    // 3x6 micro-blocks
  constexpr int G = LB_TPB / MB;               // thread groups
  constexpr int S = K == 16 ? 16 : (G > 32 ? G : 32);  // states per tile
  static_assert(LB_TPB % MB == 0 && S % G == 0, "bad Gram shape");
  __shared__ double2 tile[S][C];
  const int tid = threadIdx.x;
  const int mb = tid % MB, g = tid / MB;
  const int r0 = 3 * (mb / K), c0 = 6 * (mb % K);
  double2 acc[3][6];
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 6; b++) acc[a][b] = make_double2(0.0, 0.0);
  const double2 *src[6] = {w, x, p, W, X, P};
  const uint64_t ntiles = (dim + S - 1) / S;
  for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const uint64_t j0 = t * S;
    __syncthreads();
    for (int e = tid; e < S * C; e += LB_TPB) {
      const int s = e / C, col = e % C, blk = col / K, cc = col % K;
      const uint64_t j = j0 + s;
      tile[s][col] = j < dim ? src[blk][j * K + cc] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    for (int s = g; s < S; s += G) {
      double2 br[3], yc[6];
#pragma unroll
      for (int a = 0; a < 3; a++) br[a] = tile[s][r0 + a];
#pragma unroll
      for (int b = 0; b < 6; b++) yc[b] = tile[s][c0 + b];
#pragma unroll
      for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 6; b++) {  // conj(br) * yc
          acc[a][b].x += br[a].x * yc[b].x + br[a].y * yc[b].y;
          acc[a][b].y += br[a].x * yc[b].y - br[a].y * yc[b].x;
        }
    }
  }
  double2 *out = partial + (uint64_t)blockIdx.x * R * C;
  if (G == 1) {  // one group: every output has one owner
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = 0; b < 6; b++) out[(r0 + a) * C + (c0 + b)] = acc[a][b];
    return;
  }
  // sum the G groups in a fixed order through LDS (the tile, R*C <= S*C), then store
  static_assert(G == 1 || R * C <= S * C, "Gram staging does not fit the tile");
  __syncthreads();
  double2 *red = &tile[0][0];
  for (int q = 0; q < G; q++) {
    if (g == q) {
#pragma unroll
      for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 6; b++) {
          const int idx = (r0 + a) * C + (c0 + b);
          if (q == 0) red[idx] = acc[a][b];
          else {
            red[idx].x += acc[a][b].x;
            red[idx].y += acc[a][b].y;
          }
        }
    }
    __syncthreads();
  }
  for (int e = tid; e < R * C; e += LB_TPB) out[e] = red[e];
}

// x' = sum_r b_r cx[r][c], X' = sum_r B_r cx[r][c]  (all 3K rows);
// p' = sum_{r in w,p} b_r cx[r][c], P' likewise   (CalcByLOBPCG.c:508-556);
// in place; |x'_c|^2 and |p'_c|^2 per workgroup.  K lanes own one state, lane
// c writes column c after the whole wave has read the state's rows.
template <int K>
__global__ __launch_bounds__(LB_TPB) void k_lb_update(const double2 *w, double2 *x, double2 *p, const double2 *W,
                                                      double2 *X, double2 *P, uint64_t dim, const double2 *coef,
                                                      double *partial) {
  constexpr int R = 3 * K;
  constexpr int SPB = LB_TPB / K;  // states per tile: one (state, column) element per thread
  __shared__ double2 cf[R * K];
  __shared__ double2 rows[6][LB_TPB];  // the tile's w, x, p, W, X, P rows, one coalesced load each
  for (int e = threadIdx.x; e < R * K; e += LB_TPB) cf[e] = coef[e];
  const int c = threadIdx.x % K, sl = threadIdx.x / K;
  const uint64_t n = dim * K, ntiles = (dim + SPB - 1) / SPB;
  double nx = 0.0, np = 0.0;
  for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const uint64_t e = t * LB_TPB + threadIdx.x;  // = j * K + c
    const bool ok = e < n;
    const double2 z = make_double2(0.0, 0.0);
    rows[0][threadIdx.x] = ok ? w[e] : z;
    rows[1][threadIdx.x] = ok ? x[e] : z;
    rows[2][threadIdx.x] = ok ? p[e] : z;
    rows[3][threadIdx.x] = ok ? W[e] : z;
    rows[4][threadIdx.x] = ok ? X[e] : z;
    rows[5][threadIdx.x] = ok ? P[e] : z;
    __syncthreads();
    if (ok) {
      double2 xa, pa = z, Xa, Pa = z;
      const int row = sl * K;
#pragma unroll 4
      for (int r = 0; r < K; r++) {  // w rows: into x' and p'
        const double2 k0 = cf[r * K + c], bv = rows[0][row + r], hv = rows[3][row + r];
        const double2 tb = make_double2(k0.x * bv.x - k0.y * bv.y, k0.x * bv.y + k0.y * bv.x);
        const double2 th = make_double2(k0.x * hv.x - k0.y * hv.y, k0.x * hv.y + k0.y * hv.x);
        pa.x += tb.x, pa.y += tb.y, Pa.x += th.x, Pa.y += th.y;
      }
#pragma unroll 4
      for (int r = 0; r < K; r++) {  // p rows: into x' and p'
        const double2 k0 = cf[(2 * K + r) * K + c], bv = rows[2][row + r], hv = rows[5][row + r];
        pa.x += k0.x * bv.x - k0.y * bv.y, pa.y += k0.x * bv.y + k0.y * bv.x;
        Pa.x += k0.x * hv.x - k0.y * hv.y, Pa.y += k0.x * hv.y + k0.y * hv.x;
      }
      xa = pa;
      Xa = Pa;
#pragma unroll 4
      for (int r = 0; r < K; r++) {  // x rows: into x' only
        const double2 k0 = cf[(K + r) * K + c], bv = rows[1][row + r], hv = rows[4][row + r];
        xa.x += k0.x * bv.x - k0.y * bv.y, xa.y += k0.x * bv.y + k0.y * bv.x;
        Xa.x += k0.x * hv.x - k0.y * hv.y, Xa.y += k0.x * hv.y + k0.y * hv.x;
      }
      x[e] = xa;
      X[e] = Xa;
      p[e] = pa;
      P[e] = Pa;
      nx += xa.x * xa.x + xa.y * xa.y;
      np += pa.x * pa.x + pa.y * pa.y;
    }
    __syncthreads();  // the next tile overwrites rows
  }
  // two column sums: reuse the per-column reduction twice
  __shared__ double red2[LB_TPB / 64][2][K];
  for (int off = 32; off >= K; off >>= 1) {
    nx += __shfl_xor(nx, off, 64);
    np += __shfl_xor(np, off, 64);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane < K) {
    red2[wave][0][lane] = nx;
    red2[wave][1][lane] = np;
  }
  __syncthreads();
  if (threadIdx.x < 2 * K) {
    const int q = threadIdx.x / K, cc = threadIdx.x % K;
    double s = 0.0;
    for (int wv = 0; wv < LB_TPB / 64; wv++) s += red2[wv][q][cc];
    partial[(uint64_t)blockIdx.x * 2 * K + threadIdx.x] = s;
  }
}

// out[c] = sum_b partial[b*m + c]: one workgroup per c, fixed order
__global__ __launch_bounds__(256) void k_lb_sum(const double *partial, int nblk, int m, double *out) {
  __shared__ double sh[256];
  const int c = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < nblk; b += 256) s += partial[(uint64_t)b * m + c];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = sh[0];
}

// ---------------------------------------------------------------- host ----
// Scratch for one LOBPCG run: partial sums and small results (device), a
// pinned host mirror of the results.
struct hphi_lb_ws {
  double *part = nullptr, *res = nullptr, *h_res = nullptr;
  double2 *coef = nullptr;
  hipStream_t stream = nullptr;
  int k = 0;
};

static int lb_grid(uint64_t n) {
  uint64_t g = (n + LB_TPB - 1) / LB_TPB;
  if (g > LB_GRID) g = LB_GRID;
  if (g < 1) g = 1;
  return (int)g;
}

extern "C" int hphi_lb_ws_create(hphi_engine *e, int k, hphi_lb_ws **out) {
  if (k != 1 && k != 2 && k != 4 && k != 8 && k != 16) return hphi_set_error("LOBPCG block size %d not in {1,2,4,8,16}", k);
  hphi_lb_ws *w = new hphi_lb_ws();
  w->k = k;
  w->stream = e->stream;
  const size_t gram = (size_t)(LB_GRAM_GRID > LB_GRID ? LB_GRAM_GRID : LB_GRID) * 18 * k * k * 2;  // doubles
  if (hipMalloc(&w->part, sizeof(double) * gram) != hipSuccess ||
      hipMalloc(&w->res, sizeof(double) * 18 * k * k * 2) != hipSuccess ||
      hipMalloc(&w->coef, sizeof(double2) * 3 * k * k) != hipSuccess ||
      hipHostMalloc(&w->h_res, sizeof(double) * 18 * k * k * 2, hipHostMallocDefault) != hipSuccess) {
    (void)hipFree(w->part);
    (void)hipFree(w->res);
    (void)hipFree(w->coef);
    delete w;
    return hphi_set_error("LOBPCG workspace allocation failed");
  }
  *out = w;
  return 0;
}

extern "C" void hphi_lb_ws_destroy(hphi_lb_ws *w) {
  if (!w) return;
  (void)hipStreamSynchronize(w->stream);
  (void)hipFree(w->part);
  (void)hipFree(w->res);
  (void)hipFree(w->coef);
  if (w->h_res) (void)hipHostFree(w->h_res);
  delete w;
}

static int lb_finish(hphi_lb_ws *w, int g, int m, double *host) {
  hipLaunchKernelGGL(k_lb_sum, dim3(m), dim3(256), 0, w->stream, (const double *)w->part, g, m, w->res);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(w->h_res, w->res, sizeof(double) * m, hipMemcpyDeviceToHost, w->stream));
  HIPCHK(hipStreamSynchronize(w->stream));
  memcpy(host, w->h_res, sizeof(double) * m);
  return 0;
}

#define LB_DISPATCH(KK, CALL)  \
  switch (KK) {                \
    case 1: CALL(1); break;    \
    case 2: CALL(2); break;    \
    case 4: CALL(4); break;    \
    case 8: CALL(8); break;    \
    case 16: CALL(16); break;  \
  }

// w = s_c (X - mu_c x); norms2[c] = |w_c|^2
extern "C" int hphi_lb_resid(hphi_lb_ws *ws, uint64_t dim, const double *X, const double *x, double *w,
                             const double *mu, const double *s, double *norms2) {
  LbCols cf;
  for (int c = 0; c < 16; c++) cf.a[c] = c < ws->k ? mu[c] : 0.0, cf.b[c] = c < ws->k ? s[c] : 0.0;
  const int g = lb_grid(dim * ws->k);
#define CALL(K)                                                                                     \
  hipLaunchKernelGGL(k_lb_resid<K>, dim3(g), dim3(LB_TPB), 0, ws->stream, (const double2 *)X,       \
                     (const double2 *)x, (double2 *)w, dim, cf, ws->part)
  LB_DISPATCH(ws->k, CALL)
#undef CALL
  HIPCHK(hipGetLastError());
  return lb_finish(ws, g, ws->k, norms2);
}

extern "C" int hphi_lb_precond(hphi_lb_ws *ws, uint64_t dim, double *w, const double *diag, const double *shift,
                               double eps, double *norms2) {
  LbCols cf;
  for (int c = 0; c < 16; c++) cf.a[c] = c < ws->k ? shift[c] : 0.0, cf.b[c] = 0.0;
  const int g = lb_grid(dim * ws->k);
#define CALL(K)                                                                                          \
  hipLaunchKernelGGL(k_lb_precond<K>, dim3(g), dim3(LB_TPB), 0, ws->stream, (double2 *)w, diag, dim, cf, \
                     eps, ws->part)
  LB_DISPATCH(ws->k, CALL)
#undef CALL
  HIPCHK(hipGetLastError());
  return lb_finish(ws, g, ws->k, norms2);
}

// gram[(r * 6k + c) * 2 + {re, im}] = <b_r | y_c> of the stored vectors
extern "C" int hphi_lb_gram(hphi_lb_ws *ws, uint64_t dim, const double *w, const double *x, const double *p,
                            const double *W, const double *X, const double *P, double *gram) {
  const int k = ws->k;
  const int g = (int)std::min<uint64_t>(LB_GRAM_GRID, std::max<uint64_t>(1, (dim + 15) / 16));
#define CALL(K)                                                                                              \
  hipLaunchKernelGGL(k_lb_gram<K>, dim3(g), dim3(LB_TPB), 0, ws->stream, (const double2 *)w,                 \
                     (const double2 *)x, (const double2 *)p, (const double2 *)W, (const double2 *)X,         \
                     (const double2 *)P, dim, (double2 *)ws->part)
  LB_DISPATCH(k, CALL)
#undef CALL
  HIPCHK(hipGetLastError());
  return lb_finish(ws, g, 18 * k * k * 2, gram);
}

// coef: (3k x k) complex, row-major [r][c] as (re, im); norms2 = (|x'_c|^2, |p'_c|^2)
extern "C" int hphi_lb_update(hphi_lb_ws *ws, uint64_t dim, const double *w, double *x, double *p, const double *W,
                              double *X, double *P, const double *coef, double *norms2) {
  const int k = ws->k;
  HIPCHK(hipMemcpyAsync(ws->coef, coef, sizeof(double2) * 3 * k * k, hipMemcpyHostToDevice, ws->stream));
  const int g = lb_grid(dim * k);
#define CALL(K)                                                                                              \
  hipLaunchKernelGGL(k_lb_update<K>, dim3(g), dim3(LB_TPB), 0, ws->stream, (const double2 *)w, (double2 *)x, \
                     (double2 *)p, (const double2 *)W, (double2 *)X, (double2 *)P, dim,                      \
                     (const double2 *)ws->coef, ws->part)
  LB_DISPATCH(k, CALL)
#undef CALL
  HIPCHK(hipGetLastError());
  return lb_finish(ws, g, 2 * k, norms2);
}

// W = H w (overwrite) for a block of k columns, no prdct read-back
extern "C" int hphi_lb_hv(hphi_engine *e, int k, double *W, const double *w) {
  return hphi_hv_launch_ex(e, k, W, w, 1);
}

extern "C" const double *hphi_lb_diag(hphi_engine *e) { return e->d_diag; }

extern "C" int hphi_lb_nsite(const hphi_engine *e) { return e->nsite; }
