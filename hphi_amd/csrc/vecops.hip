// Streaming vector kernels used by the host C drivers (Lanczos, TPQ, LOBPCG).
// All reductions are two-stage with a fixed grid and a fixed summation order,
// so a result depends only on the data, never on scheduling.
#include <hip/hip_runtime.h>

#include <cstring>

#include "engine_internal.h"
#include "hphi_amd.h"

#define HIPCHK(x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) return hphi_set_error("%s failed: %s", #x, hipGetErrorString(e_)); \
  } while (0)

#define VO_TPB 256
#define VO_GRID 2048

__device__ __forceinline__ double block_sum(double v, double *sh) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < VO_TPB / 64; w++) s += sh[w];
  __syncthreads();
  return s;
}

// beta^2 = sum |v0 - alpha v1|^2   (Lanczos_EigenValue.c:170-177)
__global__ __launch_bounds__(VO_TPB) void k_lanczos_beta(const double2 *v0, const double2 *v1, uint64_t n,
                                                         double alpha, double *partial) {
  __shared__ double sh[VO_TPB / 64];
  double s = 0.0;
  for (uint64_t i = (uint64_t)blockIdx.x * VO_TPB + threadIdx.x; i < n; i += (uint64_t)gridDim.x * VO_TPB) {
    double2 a = v0[i], b = v1[i];
    double r = a.x - alpha * b.x, m = a.y - alpha * b.y;
    s += r * r + m * m;
  }
  s = block_sum(s, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// Lanczos residual with the normalisation deferred: r <- r - f w (f = alpha_m
// s_m, w the unnormalised Lanczos vector of step m), |r|^2 per block; with vg
// also vg += cg w (the eigenvector sum).  One pass of 48 bytes per state.
__global__ __launch_bounds__(VO_TPB) void k_lz_resid(double2 *r, const double2 *w, double2 *vg, uint64_t n, double f,
                                                     double cg, double *partial) {
  __shared__ double sh[VO_TPB / 64];
  double s = 0.0;
  for (uint64_t i = (uint64_t)blockIdx.x * VO_TPB + threadIdx.x; i < n; i += (uint64_t)gridDim.x * VO_TPB) {
    const double2 a = r[i], b = w[i];
    const double2 t = make_double2(a.x - f * b.x, a.y - f * b.y);
    r[i] = t;
    s += t.x * t.x + t.y * t.y;
    if (vg) {
      const double2 g = vg[i];
      vg[i] = make_double2(g.x + cg * b.x, g.y + cg * b.y);
    }
  }
  s = block_sum(s, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// temp1 = v1; temp2 = (v0 - alpha v1)/beta; v0 = -beta temp1; v1 = temp2
// (Lanczos_EigenValue.c:154-160)
__global__ __launch_bounds__(VO_TPB) void k_lanczos_update(double2 *v0, double2 *v1, uint64_t n, double alpha,
                                                           double beta) {
  for (uint64_t i = (uint64_t)blockIdx.x * VO_TPB + threadIdx.x; i < n; i += (uint64_t)gridDim.x * VO_TPB) {
    double2 a = v0[i], b = v1[i];
    double2 t2 = make_double2((a.x - alpha * b.x) / beta, (a.y - alpha * b.y) / beta);
    v0[i] = make_double2(-beta * b.x, -beta * b.y);
    v1[i] = t2;
  }
}

// per-column squared norms of a [n][k] block: partial[blk*k + c]
__global__ __launch_bounds__(VO_TPB) void k_norm2(const double2 *v, uint64_t n, int k, double *partial) {
  __shared__ double sh[VO_TPB / 64];
  for (int c = 0; c < k; c++) {
    double s = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * VO_TPB + threadIdx.x; i < n; i += (uint64_t)gridDim.x * VO_TPB) {
      double2 a = v[i * k + c];
      s += a.x * a.x + a.y * a.y;
    }
    s = block_sum(s, sh);
    if (threadIdx.x == 0) partial[(uint64_t)blockIdx.x * k + c] = s;
  }
}

// per-column dot <x_c|y_c> of [n][k] blocks
__global__ __launch_bounds__(VO_TPB) void k_dot(const double2 *x, const double2 *y, uint64_t n, int k,
                                                double *partial) {
  __shared__ double sh[VO_TPB / 64];
  for (int c = 0; c < k; c++) {
    double sr = 0.0, si = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * VO_TPB + threadIdx.x; i < n; i += (uint64_t)gridDim.x * VO_TPB) {
      double2 a = x[i * k + c], b = y[i * k + c];
      sr += a.x * b.x + a.y * b.y;
      si += a.x * b.y - a.y * b.x;
    }
    sr = block_sum(sr, sh);
    si = block_sum(si, sh);
    if (threadIdx.x == 0) {
      partial[2 * ((uint64_t)blockIdx.x * k + c)] = sr;
      partial[2 * ((uint64_t)blockIdx.x * k + c) + 1] = si;
    }
  }
}

// y_c = a_c x_c + b_c y_c  (complex a, b per column)
__global__ __launch_bounds__(VO_TPB) void k_axpby(const double2 *x, double2 *y, uint64_t n, int k,
                                                  const double2 *ab) {
  for (uint64_t i = (uint64_t)blockIdx.x * VO_TPB + threadIdx.x; i < n * k; i += (uint64_t)gridDim.x * VO_TPB) {
    int c = (int)(i % k);
    double2 a = ab[2 * c], b = ab[2 * c + 1], xv = x[i], yv = y[i];
    y[i] = make_double2(a.x * xv.x - a.y * xv.y + b.x * yv.x - b.y * yv.y,
                        a.x * xv.y + a.y * xv.x + b.x * yv.y + b.y * yv.x);
  }
}

// out[c] = sum_b partial[b*m + c]: one workgroup per c, strided per-thread
// sums then a tree, the same order on every call
__global__ __launch_bounds__(VO_TPB) void k_sum_partials(const double *partial, int nblk, int m, double *out) {
  __shared__ double sh[VO_TPB];
  const int c = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < nblk; b += VO_TPB) s += partial[(uint64_t)b * m + c];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int st = VO_TPB / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = sh[0];
}

static int vo_grid(uint64_t n) {
  uint64_t g = (n + VO_TPB - 1) / VO_TPB;
  if (g > VO_GRID) g = VO_GRID;
  if (g < 1) g = 1;
  return (int)g;
}

// scratch: e->d_partial holds HV_MAX_GRID*16 double2 = enough for VO_GRID*32 doubles
static int finish_sum(hphi_engine *e, int g, int m, double *host_out) {
  double *part = (double *)e->d_partial;
  double *out = (double *)e->d_prdct;
  hipLaunchKernelGGL(k_sum_partials, dim3(m), dim3(VO_TPB), 0, e->stream, part, g, m, out);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(e->h_prdct, out, sizeof(double) * m, hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  memcpy(host_out, e->h_prdct, sizeof(double) * m);
  return 0;
}

extern "C" int hphi_vec_lanczos_resid(hphi_engine *e, double *r, const double *w, double *vg, double f, double cg,
                                      double *r2) {
  if (!e || !r || !w || !r2) return hphi_set_error("hphi_vec_lanczos_resid: NULL argument");
  int g = vo_grid(e->dim);
  hipLaunchKernelGGL(k_lz_resid, dim3(g), dim3(VO_TPB), 0, e->stream, (double2 *)r, (const double2 *)w,
                     (double2 *)vg, e->dim, f, cg, (double *)e->d_partial);
  return finish_sum(e, g, 1, r2);
}

extern "C" int hphi_vec_lanczos_beta(hphi_engine *e, const double *v0, const double *v1, double alpha,
                                     double *beta2) {
  if (!e || !v0 || !v1 || !beta2) return hphi_set_error("hphi_vec_lanczos_beta: NULL argument");
  int g = vo_grid(e->dim);
  hipLaunchKernelGGL(k_lanczos_beta, dim3(g), dim3(VO_TPB), 0, e->stream, (const double2 *)v0,
                     (const double2 *)v1, e->dim, alpha, (double *)e->d_partial);
  return finish_sum(e, g, 1, beta2);
}

// the update above plus vg += c * v1_new (eigenvector accumulation)
__global__ __launch_bounds__(VO_TPB) void k_lanczos_update_acc(double2 *v0, double2 *v1, double2 *vg, uint64_t n,
                                                               double alpha, double beta, double c) {
  for (uint64_t i = (uint64_t)blockIdx.x * VO_TPB + threadIdx.x; i < n; i += (uint64_t)gridDim.x * VO_TPB) {
    double2 a = v0[i], b = v1[i], g = vg[i];
    double2 t2 = make_double2((a.x - alpha * b.x) / beta, (a.y - alpha * b.y) / beta);
    v0[i] = make_double2(-beta * b.x, -beta * b.y);
    v1[i] = t2;
    vg[i] = make_double2(g.x + c * t2.x, g.y + c * t2.y);
  }
}
extern "C" int hphi_vec_lanczos_update_acc(hphi_engine *e, double *v0, double *v1, double *vg, double alpha,
                                           double beta, double c) {
  if (!e || !v0 || !v1 || !vg) return hphi_set_error("hphi_vec_lanczos_update_acc: NULL argument");
  hipLaunchKernelGGL(k_lanczos_update_acc, dim3(vo_grid(e->dim)), dim3(VO_TPB), 0, e->stream, (double2 *)v0,
                     (double2 *)v1, (double2 *)vg, e->dim, alpha, beta, c);
  HIPCHK(hipGetLastError());
  // the copy / zero / fill helpers run on the null stream, which is not ordered
  // against the engine's non-blocking stream: finish before returning
  HIPCHK(hipStreamSynchronize(e->stream));
  return 0;
}
extern "C" int hphi_vec_lanczos_update(hphi_engine *e, double *v0, double *v1, double alpha, double beta) {
  if (!e || !v0 || !v1) return hphi_set_error("hphi_vec_lanczos_update: NULL argument");
  hipLaunchKernelGGL(k_lanczos_update, dim3(vo_grid(e->dim)), dim3(VO_TPB), 0, e->stream, (double2 *)v0,
                     (double2 *)v1, e->dim, alpha, beta);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(e->stream));
  return 0;
}

extern "C" int hphi_vec_norm2(hphi_engine *e, const double *v, int k, double *out) {
  if (!e || !v || !out || k < 1 || k > 16) return hphi_set_error("hphi_vec_norm2: bad argument");
  int g = vo_grid(e->dim);
  hipLaunchKernelGGL(k_norm2, dim3(g), dim3(VO_TPB), 0, e->stream, (const double2 *)v, e->dim, k,
                     (double *)e->d_partial);
  return finish_sum(e, g, k, out);
}

extern "C" int hphi_vec_dot(hphi_engine *e, const double *x, const double *y, int k, double *out) {
  if (!e || !x || !y || !out || k < 1 || k > 16) return hphi_set_error("hphi_vec_dot: bad argument");
  int g = vo_grid(e->dim);
  hipLaunchKernelGGL(k_dot, dim3(g), dim3(VO_TPB), 0, e->stream, (const double2 *)x, (const double2 *)y,
                     e->dim, k, (double *)e->d_partial);
  return finish_sum(e, g, 2 * k, out);
}

extern "C" int hphi_vec_axpby(hphi_engine *e, const double *x, double *y, int k, const double *ab) {
  if (!e || !x || !y || !ab || k < 1 || k > 16) return hphi_set_error("hphi_vec_axpby: bad argument");
  double2 *d_ab = (double2 *)e->d_coef;
  HIPCHK(hipMemcpyAsync(d_ab, ab, sizeof(double2) * 2 * k, hipMemcpyHostToDevice, e->stream));
  hipLaunchKernelGGL(k_axpby, dim3(vo_grid(e->dim * k)), dim3(VO_TPB), 0, e->stream, (const double2 *)x,
                     (double2 *)y, e->dim, k, (const double2 *)d_ab);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(e->stream));
  return 0;
}

extern "C" int hphi_engine_sync(hphi_engine *e) {
  if (!e) return hphi_set_error("NULL engine");
  HIPCHK(hipStreamSynchronize(e->stream));
  return 0;
}
