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

// ---- Gram blocks on the FP64 matrix cores (k = 8) --------------------------
// Per state the 24 complex columns of (w, x, p) are 48 reals r (block b's 8
// columns are reals 16b .. 16b+15, interleaved re/im as stored) and the 24 of
// (W, X, P) are 48 reals y.  Over all states the real products R2 = r^T r and
// R1 = r^T y hold both complex Gram matrices:
//   <b_a|c_b> = (R[2a][2b] + R[2a+1][2b+1]) + i (R[2a][2b+1] - R[2a+1][2b]).
// Only the 16 x 16 real tiles (I, J) with I <= J are formed (6 per matrix):
// they hold every complex entry a <= b, the triangle zheevd('U') reads
// (CalcByLOBPCG.c:51,69,105); the other triangle is its conjugate.
// One v_mfma_f64_16x16x4f64 per tile and 4 states: lane l loads real column
// l % 16 of state 4q + l / 16 from each block (64 lanes read 512 contiguous
// bytes), which is operand A[i = l % 16][kk = l / 16] of r and B[kk][j] of y;
// D[i][j] sits in lane 16 (i % 4) + j, register i / 4 (checked on the device).
#define LBM_TPB 256
#define LBM_GRID 768       // 3 workgroups per CU = 3 waves per SIMD (152 VGPRs); each wave owns a contiguous range of states
#define LBM_TILES 12
#define LBM_PART (LBM_TILES * 256)
#define LBM_SPLIT 16       // first reduction pass: partial groups per output
typedef double lbm_d4 __attribute__((ext_vector_type(4)));

// the 12 tile products of one group of 4 states (operands as loaded)
__device__ __forceinline__ void lbm_tiles(lbm_d4 *acc, const double *v) {
  const double a0 = v[0], a1 = v[1], a2 = v[2], b0 = v[3], b1 = v[4], b2 = v[5];
  acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, a0, acc[0], 0, 0, 0);    // R2 (0,0)
  acc[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, a1, acc[1], 0, 0, 0);    // R2 (0,1)
  acc[2] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, a2, acc[2], 0, 0, 0);    // R2 (0,2)
  acc[3] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, a1, acc[3], 0, 0, 0);    // R2 (1,1)
  acc[4] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, a2, acc[4], 0, 0, 0);    // R2 (1,2)
  acc[5] = __builtin_amdgcn_mfma_f64_16x16x4f64(a2, a2, acc[5], 0, 0, 0);    // R2 (2,2)
  acc[6] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[6], 0, 0, 0);    // R1 (0,0)
  acc[7] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[7], 0, 0, 0);    // R1 (0,1)
  acc[8] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b2, acc[8], 0, 0, 0);    // R1 (0,2)
  acc[9] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[9], 0, 0, 0);    // R1 (1,1)
  acc[10] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b2, acc[10], 0, 0, 0);  // R1 (1,2)
  acc[11] = __builtin_amdgcn_mfma_f64_16x16x4f64(a2, b2, acc[11], 0, 0, 0);  // R1 (2,2)
}

__global__ __launch_bounds__(LBM_TPB) __attribute__((amdgpu_waves_per_eu(3, 3))) void k_lb_gram_mfma8(const double *w, const double *x, const double *p,
                                                           const double *W, const double *X, const double *P,
                                                           uint64_t dim, double *partial) {
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int col = lane & 15, sub = lane >> 4;
  const double *src[6] = {w, x, p, W, X, P};
  lbm_d4 acc[LBM_TILES];
#pragma unroll
  for (int t = 0; t < LBM_TILES; t++) acc[t] = lbm_d4{0.0, 0.0, 0.0, 0.0};
  // whole groups of 4 states, a contiguous range per wave; loads run two
  // groups ahead of their products through three register sets used in
  // rotation (no copies, so no wait on a load in flight).  Load indices are
  // clamped into the range: every load is in bounds, with no branch around it.
  const uint64_t nq = dim / 4;
  const uint64_t nw = (uint64_t)gridDim.x * (LBM_TPB / 64);
  const uint64_t gw = (uint64_t)blockIdx.x * (LBM_TPB / 64) + wave;
  const uint64_t q0 = nq * gw / nw, q1 = nq * (gw + 1) / nw;
  if (q1 > q0) {
    // wave-uniform bases at the wave's first group and 32-bit lane offsets (a
    // wave's range is a few MB), so a load is a scalar base plus one VGPR
    const double *bw[6];
#pragma unroll
    for (int b = 0; b < 6; b++) bw[b] = src[b] + 64 * q0;
    const uint32_t lo = (uint32_t)(sub * 16 + col), last = (uint32_t)(64 * (q1 - 1 - q0)) + lo;
    double ra[6], rb[6], rc[6];
    auto ld = [&](double *v, uint32_t e) {
      e = e < last ? e : last;
#pragma unroll
      for (int b = 0; b < 6; b++) v[b] = bw[b][e];
    };
    ld(ra, lo);
    ld(rb, lo + 64);
    uint64_t q = q0;
    uint32_t e = lo;
    for (; q + 3 <= q1; q += 3, e += 192) {
      ld(rc, e + 128);
      lbm_tiles(acc, ra);
      ld(ra, e + 192);
      lbm_tiles(acc, rb);
      ld(rb, e + 256);
      lbm_tiles(acc, rc);
    }
    if (q < q1) lbm_tiles(acc, ra);
    if (q + 1 < q1) lbm_tiles(acc, rb);
  }
  // the last dim % 4 states: one group, zero-padded, on the grid's first wave
  if (dim % 4 && blockIdx.x == 0 && wave == 0) {
    const uint64_t j = 4 * nq + sub;
    double v[6];
#pragma unroll
    for (int b = 0; b < 6; b++) v[b] = j < dim ? src[b][j * 16 + col] : 0.0;
    lbm_tiles(acc, v);
  }
  // the waves' tiles summed in a fixed order through LDS: red[t][i][j]
  __shared__ double red[LBM_PART];
  for (int wv = 0; wv < LBM_TPB / 64; wv++) {
    if (wave == wv) {
#pragma unroll
      for (int t = 0; t < LBM_TILES; t++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int idx = t * 256 + (sub + 4 * r) * 16 + col;
          red[idx] = wv == 0 ? acc[t][r] : red[idx] + acc[t][r];
        }
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < LBM_PART; i += LBM_TPB) partial[(uint64_t)blockIdx.x * LBM_PART + i] = red[i];
}

// first pass: out[s][i] = sum of partial[b][i] over the blocks b of group s
// (LBM_SPLIT contiguous groups), fixed order
__global__ __launch_bounds__(256) void k_lb_gram_sum1(const double *partial, int nblk, double *out) {
  const int i = blockIdx.x * 256 + threadIdx.x, s = blockIdx.y;
  if (i >= LBM_PART) return;
  const int b0 = (int)((int64_t)nblk * s / LBM_SPLIT), b1 = (int)((int64_t)nblk * (s + 1) / LBM_SPLIT);
  double a = 0.0;
  for (int b = b0; b < b1; b++) a += partial[(uint64_t)b * LBM_PART + i];
  out[s * LBM_PART + i] = a;
}

// second pass and the complex Gram matrices: gram[(r * 48 + c) * 2 + {re, im}]
// = <b_r|y_c>, r < 24 over (w, x, p), c < 48 over (w, x, p, W, X, P)
__device__ __forceinline__ double lbm_real(const double *sums, int t, int i, int j) {
  double a = 0.0;
  for (int s = 0; s < LBM_SPLIT; s++) a += sums[s * LBM_PART + t * 256 + i * 16 + j];
  return a;
}

__global__ __launch_bounds__(256) void k_lb_gram_fin8(const double *sums, double *gram) {
  const int id = blockIdx.x * 256 + threadIdx.x;
  if (id >= 24 * 48) return;
  const int r = id / 48, c = id % 48;
  const bool h = c >= 24;
  const int cb = h ? c - 24 : c;
  const bool mirror = r > cb;  // the entry is conj(<b_cb|y_r>)
  const int a = mirror ? cb : r, b = mirror ? r : cb;
  const int I = a / 8, J = b / 8;  // I <= J
  const int t = (I == 0 ? J : (I == 1 ? 2 + J : 5)) + (h ? 6 : 0);  // (0,0) (0,1) (0,2) (1,1) (1,2) (2,2)
  const int i0 = 2 * (a % 8), j0 = 2 * (b % 8);
  const double rr = lbm_real(sums, t, i0, j0), ii = lbm_real(sums, t, i0 + 1, j0 + 1);
  const double ri = lbm_real(sums, t, i0, j0 + 1), ir = lbm_real(sums, t, i0 + 1, j0);
  const double re = rr + ii, im = ri - ir;
  gram[2 * id] = re;
  gram[2 * id + 1] = mirror ? -im : im;
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
  int mfma = 1;  // k = 8 Gram on the matrix cores (0: the scalar kernel)
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
  w->mfma = e->lb_mfma;
  size_t gram = (size_t)(LB_GRAM_GRID > LB_GRID ? LB_GRAM_GRID : LB_GRID) * 18 * k * k * 2;  // doubles
  gram = std::max(gram, (size_t)(LBM_GRID + LBM_SPLIT) * LBM_PART);  // the matrix-core Gram's partials
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
  if (k == 8 && ws->mfma) {
    double *sums = ws->part + (size_t)LBM_GRID * LBM_PART;
    hipLaunchKernelGGL(k_lb_gram_mfma8, dim3(LBM_GRID), dim3(LBM_TPB), 0, ws->stream, w, x, p, W, X, P, dim,
                       ws->part);
    hipLaunchKernelGGL(k_lb_gram_sum1, dim3((LBM_PART + 255) / 256, LBM_SPLIT), dim3(256), 0, ws->stream,
                       (const double *)ws->part, LBM_GRID, sums);
    hipLaunchKernelGGL(k_lb_gram_fin8, dim3((24 * 48 + 255) / 256), dim3(256), 0, ws->stream,
                       (const double *)sums, ws->res);
    HIPCHK(hipGetLastError());
    const size_t m = 18 * k * k * 2;
    HIPCHK(hipMemcpyAsync(ws->h_res, ws->res, sizeof(double) * m, hipMemcpyDeviceToHost, ws->stream));
    HIPCHK(hipStreamSynchronize(ws->stream));
    memcpy(gram, ws->h_res, sizeof(double) * m);
    return 0;
  }
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

// x[j][c] = src[c*dim + j] (k vectors one after another -> state-major block)
__global__ __launch_bounds__(256) void k_lb_load(const double2 *src, double2 *x, uint64_t dim, int k) {
  const uint64_t n = dim * (uint64_t)k;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    const uint64_t c = i / dim, j = i % dim;
    x[j * k + c] = src[i];
  }
}

// dst[c*dim + j] = sx_c x[j][c]: the normalised columns, one vector after another
__global__ __launch_bounds__(256) void k_lb_store(const double2 *x, double2 *dst, uint64_t dim, int k, LbCols cf) {
  const uint64_t n = dim * (uint64_t)k;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    const uint64_t c = i / dim, j = i % dim;
    const double2 a = x[j * k + c];
    const double f = cf.a[c];
    dst[i] = make_double2(f * a.x, f * a.y);
  }
}

extern "C" int hphi_lb_load(hphi_lb_ws *ws, uint64_t dim, const double *src, double *x) {
  hipLaunchKernelGGL(k_lb_load, dim3(lb_grid(dim * ws->k)), dim3(256), 0, ws->stream, (const double2 *)src,
                     (double2 *)x, dim, ws->k);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(ws->stream));
  return 0;
}

extern "C" int hphi_lb_store(hphi_lb_ws *ws, uint64_t dim, const double *x, const double *sx, double *dst) {
  LbCols cf;
  for (int c = 0; c < 16; c++) cf.a[c] = c < ws->k ? sx[c] : 0.0, cf.b[c] = 0.0;
  hipLaunchKernelGGL(k_lb_store, dim3(lb_grid(dim * ws->k)), dim3(256), 0, ws->stream, (const double2 *)x,
                     (double2 *)dst, dim, ws->k, cf);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(ws->stream));
  return 0;
}

// W = H w (overwrite) for a block of k columns, no prdct read-back
extern "C" int hphi_lb_hv(hphi_engine *e, int k, double *W, const double *w) {
  return hphi_hv_launch_ex(e, k, W, w, 1);
}

extern "C" const double *hphi_lb_diag(hphi_engine *e) { return e->d_diag; }

extern "C" int hphi_lb_nsite(const hphi_engine *e) { return e->nsite; }
