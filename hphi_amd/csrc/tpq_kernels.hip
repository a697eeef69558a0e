// Microcanonical TPQ on the device for k samples at once.
//
// Restates the step of the reference (src/CalcByTPQ.c:216-272):
//   Multiply         v0 = l v1 - v0/Ns, v0 /= |v0|          (src/Multiply.c:42-66)
//   expec_energy_flct v1 <- v0, v0 = H v1, E = <v1|v0>, E2 = <v0|v0>
//                    (+ Sz, Sz^2 for SpinGC)               (src/expec_energy_flct.c:126-157,429-500)
//   inv_temp = 2 step/Ns / (l - E/Ns)                      (src/CalcByTPQ.c:235)
// with the normalisation carried as a scalar: the vector kept in HBM is
// w_s = (l w_{s-1} - H w_{s-1}/Ns)/|w_{s-1}|, whose norm is the reference's
// global_norm, and E, E2 are divided by |w_s|^2.  One step is an overwriting
// H·v (prdct fused) plus one streaming kernel (k_tpq_fused) plus a one-block
// finalisation that writes the step's output row and the next scale on the
// device, so a step needs no host synchronisation and is captured once in a
// hipGraph.  Layout: column c of sample vectors at element j*sj + c*sc
// (state-major block: sj = k, sc = 1; vector-major: sj = 1, sc = dim).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "engine_internal.h"
#include "hphi_amd.h"

#define HIPCHK(x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) return hphi_set_error("%s failed: %s", #x, hipGetErrorString(e_)); \
  } while (0)

#define TQ_TPB 256
#define TQ_GRID 1024
#define TQ_NRED 8  // per column: |Hw|^2, D, D^2, N, N^2, S, S^2, |w_next|^2
#define TQ_NOUT 12  // output doubles per (step, sample)

__device__ __forceinline__ uint64_t tq_splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double tq_unit(uint64_t seed, uint64_t c) {
  return (double)(tq_splitmix64((seed * 0x9E3779B97F4A7C15ull) ^ c) >> 11) * (1.0 / 9007199254740992.0);
}

// column c from seed0 + c; kind 0: (2u-1) + i(2u-1) (MakeIniVec.c:69, InitialVecType 0),
// kind -1: complex Gaussian by Box-Muller (MakeIniVec.c:81-91, InitialVecType -1)
__global__ void k_tpq_init(double2 *v, uint64_t dim, int k, uint64_t sj, uint64_t sc, uint64_t seed0, int kind) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= dim * (uint64_t)k) return;
  uint64_t j = i / k;
  int c = (int)(i % k);
  uint64_t s = seed0 + (uint64_t)c;
  double u1 = tq_unit(s, 2 * j), u2 = tq_unit(s, 2 * j + 1);
  double2 z;
  if (kind == -1) {
    double r = sqrt(-2.0 * log(1.0 - u1));
    double sn, cs;
    sincospi(2.0 * u2, &sn, &cs);
    z = make_double2(r * cs, r * sn);
  } else {
    z = make_double2(2.0 * u1 - 1.0, 2.0 * u2 - 1.0);
  }
  v[j * sj + c * sc] = z;
}

// squared norms of the columns (initial |w_0|^2)
__global__ __launch_bounds__(TQ_TPB) void k_tpq_colnorm(const double2 *v, uint64_t dim, int k, uint64_t sj,
                                                        uint64_t sc, double *partial) {
  const int c = threadIdx.x % k, jl = threadIdx.x / k, spb = TQ_TPB / k;
  double s = 0.0;
  for (uint64_t j = (uint64_t)blockIdx.x * spb + jl; j < dim; j += (uint64_t)gridDim.x * spb) {
    double2 a = v[j * sj + c * sc];
    s += a.x * a.x + a.y * a.y;
  }
  for (int off = 32; off >= k; off >>= 1) s += __shfl_xor(s, off, 64);
  __shared__ double red[TQ_TPB / 64][16];
  if ((threadIdx.x & 63) < k) red[threadIdx.x >> 6][threadIdx.x & 63] = s;
  __syncthreads();
  if (threadIdx.x < k) {
    double t = 0.0;
    for (int w = 0; w < TQ_TPB / 64; w++) t += red[w][threadIdx.x];
    partial[(uint64_t)blockIdx.x * k + threadIdx.x] = t;
  }
}


// step s: var += |h|^2, the moments of w (expec_energy_flct: obs_kind 1 SpinGC
// S = 2 popcount - Nsite; 2 Hubbard models D doublons, N electrons,
// S = N_up - N_down), w_next = (l w - h/Ns) * inv_norm, |w_next|^2
__global__ __launch_bounds__(TQ_TPB) void k_tpq_fused(double2 *v1, const double2 *v0, uint64_t dim, int k,
                                                      uint64_t sj, uint64_t sc, const double *state, double lv,
                                                      double invNs, int obs_kind, int nsite,
                                                      const uint64_t *list1, double *partial) {
  const int c = threadIdx.x % k, jl = threadIdx.x / k, spb = TQ_TPB / k;
  const double sc_c = state[16 + c];  // 1/|w_s|
  double r[TQ_NRED] = {};
  for (uint64_t j = (uint64_t)blockIdx.x * spb + jl; j < dim; j += (uint64_t)gridDim.x * spb) {
    const uint64_t idx = j * sj + c * sc;
    const double2 w = v1[idx], h = v0[idx];
    r[0] += h.x * h.x + h.y * h.y;
    if (obs_kind) {
      const double ww = w.x * w.x + w.y * w.y;
      const uint64_t x = list1 ? list1[j] : j;
      const uint64_t EVEN = 0x5555555555555555ull;
      double S;
      if (obs_kind == 1) {
        S = (double)(2 * (int)__popcll(x) - nsite);
      } else {
        const double D = (double)(int)__popcll(x & (x >> 1) & EVEN);
        const int up = (int)__popcll(x & EVEN), dn = (int)__popcll(x & (EVEN << 1));
        const double N = (double)(up + dn);
        S = (double)(up - dn);
        r[1] += ww * D;
        r[2] += ww * D * D;
        r[3] += ww * N;
        r[4] += ww * N * N;
      }
      r[5] += ww * S;
      r[6] += ww * S * S;
    }
    const double2 wn = make_double2((lv * w.x - h.x * invNs) * sc_c, (lv * w.y - h.y * invNs) * sc_c);
    v1[idx] = wn;
    r[7] += wn.x * wn.x + wn.y * wn.y;
  }
  __shared__ double red[TQ_TPB / 64][16][TQ_NRED];
#pragma unroll
  for (int q = 0; q < TQ_NRED; q++) {
    double s = r[q];
    for (int off = 32; off >= k; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) < k) red[threadIdx.x >> 6][threadIdx.x & 63][q] = s;
  }
  __syncthreads();
  if (threadIdx.x < k * TQ_NRED) {
    const int cc = threadIdx.x / TQ_NRED, q = threadIdx.x % TQ_NRED;
    double t = 0.0;
    for (int w = 0; w < TQ_TPB / 64; w++) t += red[w][cc][q];
    partial[((uint64_t)blockIdx.x * k + cc) * TQ_NRED + q] = t;
  }
}

// state layout (doubles): [0..16) |w_s|^2, [16..32) 1/|w_s|, [32] step counter
__global__ void k_tpq_start(const double *partial, int nblk, int k, double *state) {
  const int c = threadIdx.x;
  if (c < k) {
    double n = 0.0;
    for (int b = 0; b < nblk; b++) n += partial[(uint64_t)b * k + c];
    state[c] = n;
    state[16 + c] = 1.0 / sqrt(n);
  }
  if (c == 0) state[32] = 0.0;
}

// one block: output row of step s, next scale
__global__ void k_tpq_finalize(const double *partial, int nblk, int k, const double2 *prdct, double *state,
                               double lv, double Ns, int obs_kind, int nsite, double sz_const, double *out,
                               int max_steps, int row0) {
  const int c = threadIdx.x;
  const int s = (int)state[32];
  __syncthreads();
  if (c < k) {
    double r[TQ_NRED] = {};
    for (int b = 0; b < nblk; b++)
      for (int q = 0; q < TQ_NRED; q++) r[q] += partial[((uint64_t)b * k + c) * TQ_NRED + q];
    const double n = state[c];
    const double E = prdct[c].x / n;
    const double var = r[0] / n;
    if (s >= row0 && s < max_steps) {
      // row of SS_rand / Norm_rand / Flct_rand (CalcByTPQ.c:150-210,235-256) with
      // the per-model moments of expec_energy_flct (expec_energy_flct.c:186-500)
      double *o = out + ((uint64_t)(s - row0) * k + c) * TQ_NOUT;
      o[0] = (s == 0) ? 0.0 : (2.0 * s / Ns) / (lv - E / Ns);
      o[1] = E;
      o[2] = var;
      o[3] = sqrt(n);
      if (obs_kind == 2) {
        o[4] = r[1] / n;
        o[5] = r[2] / n;
        o[6] = r[3] / n;
        o[7] = r[4] / n;
      } else {
        o[4] = o[5] = 0.0;
        o[6] = (double)nsite;
        o[7] = (double)nsite * nsite;
      }
      o[8] = obs_kind ? 0.5 * r[5] / n : sz_const;
      o[9] = obs_kind ? 0.25 * r[6] / n : sz_const * sz_const;
      o[10] = prdct[c].y / n;
      o[11] = 0.0;
    }
    state[c] = r[7];
    state[16 + c] = 1.0 / sqrt(r[7]);
  }
  __syncthreads();
  if (c == 0) state[32] = (double)(s + 1);
}

// max_j |list_Diagonal[j]| (for l = 3 max|D| / Ns)
__global__ __launch_bounds__(256) void k_maxabs(const double *d, uint64_t n, double *partial) {
  double m = 0.0;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
    m = fmax(m, fabs(d[i]));
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  __shared__ double sh[4];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}

__global__ void k_copy_prdct(const double2 *src, double2 *dst, int c) { dst[c] = src[0]; }

// ------------------------------------------------------------------ host --
static int tq_grid(uint64_t dim, int k) {
  uint64_t spb = TQ_TPB / k;
  uint64_t g = (dim + spb - 1) / spb;
  if (g > TQ_GRID) g = TQ_GRID;
  return (int)(g < 1 ? 1 : g);
}

extern "C" int hphi_diag_maxabs(hphi_engine *e, double *out) {
  if (!e || !out) return hphi_set_error("hphi_diag_maxabs: NULL argument");
  if (hphi_build_diag(e)) return -1;
  int g = (int)std::min<uint64_t>((e->dim + 255) / 256, 1024);
  hipLaunchKernelGGL(k_maxabs, dim3(g), dim3(256), 0, e->stream, e->d_diag, e->dim, (double *)e->d_partial);
  std::vector<double> h(g);
  HIPCHK(hipMemcpyAsync(h.data(), e->d_partial, sizeof(double) * g, hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  double m = 0.0;
  for (double x : h) m = std::max(m, x);
  *out = m;
  return 0;
}

// LargeValue when the caller gives none: the reference has no default rule
// (modpara LargeValue, global.c); here l = max(3 max|diag|, B) / Nsite with
// B = max|diag| + sum of |coupling| over the compiled off-diagonal terms, a
// Gershgorin bound of |E| (each gather term contributes at most one element
// per row), so l - H/Nsite stays positive definite also for Hamiltonians
// without a diagonal part (pure hopping, XY, transverse field).
static int default_large_value(hphi_engine *e, double *lv) {
  double m = 0.0;
  if (hphi_diag_maxabs(e, &m)) return -1;
  double off = 0.0;
  for (const GTerm &t : e->h_terms_dev) off += std::hypot(t.cre, t.cim);
  const double b = std::max(3.0 * m, m + off);
  if (!(b > 0.0))
    return hphi_set_error("LargeValue not given and the Hamiltonian is zero: no bound of its spectrum");
  *lv = b / (double)e->nsite;
  return 0;
}

struct TpqPlan {
  double *V0, *V1, *state, *part, *out;
  int k, layout, g, grid_hv;
  uint64_t sj, sc;
  double lv, Ns, sz_const;
  int obs_kind, steps;
  int row0;  // first step whose row is kept (out holds steps row0 .. steps-1)
};

// enqueue one step on e->stream (capturable: kernels only)
static int tpq_enqueue_step(hphi_engine *e, const TpqPlan &P) {
  const uint64_t dim = e->dim;
  if (P.layout == 1) {
    if (hphi_hv_launch_ex(e, P.k, P.V0, P.V1, 1)) return -1;
  } else {
    // vector-major: one single-vector H·v per column; their prdct go to columns of a
    // per-column buffer through the shared reduction kernel
    for (int c = 0; c < P.k; c++) {
      if (hphi_hv_launch_ex(e, 1, P.V0 + 2 * (uint64_t)c * dim, P.V1 + 2 * (uint64_t)c * dim, 1)) return -1;
      hipLaunchKernelGGL(k_copy_prdct, dim3(1), dim3(1), 0, e->stream, (const double2 *)e->d_prdct,
                         (double2 *)e->d_coef, c);
    }
  }
  hipLaunchKernelGGL(k_tpq_fused, dim3(P.g), dim3(TQ_TPB), 0, e->stream, (double2 *)P.V1, (const double2 *)P.V0,
                     dim, P.k, P.sj, P.sc, (const double *)P.state, P.lv, 1.0 / P.Ns, P.obs_kind, e->nsite,
                     hphi_model_gc(e->model) ? (const uint64_t *)nullptr : (const uint64_t *)e->d_list1,
                     P.part);
  hipLaunchKernelGGL(k_tpq_finalize, dim3(1), dim3(64), 0, e->stream, (const double *)P.part, P.g, P.k,
                     (const double2 *)(P.layout == 1 ? e->d_prdct : e->d_coef), P.state, P.lv, P.Ns, P.obs_kind,
                     e->nsite, P.sz_const, P.out, P.steps, P.row0);
  HIPCHK(hipGetLastError());
  return 0;
}

// w_c <- W0 column c (vector-major input, element c*dim + j) in the run's layout
__global__ void k_tpq_load(double2 *v, const double2 *W0, uint64_t dim, int k, uint64_t sj, uint64_t sc) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= dim * (uint64_t)k) return;
  const uint64_t c = i / dim, j = i % dim;
  v[j * sj + c * sc] = W0[i];
}

// W_end column c = w_c / |w_c|, the normalised vector of the last step
__global__ void k_tpq_store(double2 *W_end, const double2 *v, uint64_t dim, int k, uint64_t sj, uint64_t sc,
                            const double *state) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= dim * (uint64_t)k) return;
  const uint64_t c = i / dim, j = i % dim;
  const double2 a = v[j * sj + c * sc];
  const double f = state[16 + c];
  W_end[i] = make_double2(a.x * f, a.y * f);
}

// the step counter of a resumed run (the first enqueued step is step0 - 1)
__global__ void k_tpq_set_step(double *state, double s) { state[32] = s; }

static int tpq_run(hphi_engine *e, const hphi_tpq_params *p, const double *W0, int step0, double *out_host,
                   hphi_tpq_result *r, double *W_end);

extern "C" int hphi_tpq(hphi_engine *e, const hphi_tpq_params *p, double *out_host, hphi_tpq_result *r) {
  return tpq_run(e, p, nullptr, 0, out_host, r, nullptr);
}

extern "C" int hphi_tpq_from(hphi_engine *e, const hphi_tpq_params *p, const double *W0, int step0, double *out_host,
                             hphi_tpq_result *r, double *W_end) {
  if (!W0) return hphi_set_error("hphi_tpq_from: NULL start vectors");
  if (step0 < 0 || step0 >= (p ? p->steps : 0)) return hphi_set_error("hphi_tpq_from: step0 must be in [0, steps)");
  return tpq_run(e, p, W0, step0, out_host, r, W_end);
}

// W0 == nullptr: random start vectors (seed); else column c of W0 starts
// sample c: with step0 = 0 as w_0 (its norm is global_1st_norm), with step0 > 0
// as the normalised vector of step step0 - 1 of an earlier run, whose row is
// recomputed and dropped.  Rows of steps step0 .. steps-1 go to out_host.
static int tpq_run(hphi_engine *e, const hphi_tpq_params *p, const double *W0, int step0, double *out_host,
                   hphi_tpq_result *r, double *W_end) {
  if (!e || !p || !out_host) return hphi_set_error("hphi_tpq: NULL argument");
  const int k = p->k;
  if (k != 1 && k != 2 && k != 4 && k != 8 && k != 16) return hphi_set_error("hphi_tpq: k must be 1, 2, 4, 8 or 16");
  if (p->steps < 1) return hphi_set_error("hphi_tpq: steps must be positive");
  if (hphi_build_tables(e) || hphi_build_diag(e)) return -1;
  const uint64_t dim = e->dim;
  TpqPlan P{};
  P.k = k;
  P.layout = (p->layout == 1) ? 1 : 0;
  P.sj = P.layout == 1 ? (uint64_t)k : 1;
  P.sc = P.layout == 1 ? 1 : dim;
  P.Ns = (double)e->nsite;
  P.steps = p->steps;
  if (p->large_value > 0.0) {
    P.lv = p->large_value;
  } else if (default_large_value(e, &P.lv)) {
    return -1;
  }
  P.obs_kind = e->model == HPHI_MODEL_SPINGC ? 1 : hphi_model_hubbard(e->model) ? 2 : 0;
  P.sz_const = 0.5 * (double)(e->nup - e->ndown);  // canonical Spin (expec_energy_flct.c:104-118)
  P.g = tq_grid(dim, k);
  size_t vbytes = sizeof(double2) * dim * k;
  int rc = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ev0, ev1;
  (void)hipEventCreate(&ev0);
  (void)hipEventCreate(&ev1);
  do {
    // every buffer is freed below, also when a later allocation fails
    if (hipMalloc(&P.V0, vbytes) != hipSuccess || hipMalloc(&P.V1, vbytes) != hipSuccess ||
        hipMalloc(&P.state, sizeof(double) * 64) != hipSuccess ||
        hipMalloc(&P.part, sizeof(double) * (size_t)TQ_GRID * k * TQ_NRED) != hipSuccess ||
        hipMalloc(&P.out, sizeof(double) * (size_t)(p->steps - step0) * k * TQ_NOUT) != hipSuccess) {
      (void)hipGetLastError();
      rc = hphi_set_error("hphi_tpq: device allocation failed");
      break;
    }
    uint64_t n = dim * (uint64_t)k;
    const unsigned gn = (unsigned)((n + 255) / 256);
    if (W0)
      hipLaunchKernelGGL(k_tpq_load, dim3(gn), dim3(256), 0, e->stream, (double2 *)P.V1, (const double2 *)W0, dim, k,
                         P.sj, P.sc);
    else
      hipLaunchKernelGGL(k_tpq_init, dim3(gn), dim3(256), 0, e->stream, (double2 *)P.V1, dim, k, P.sj, P.sc, p->seed,
                         p->init_kind);
    hipLaunchKernelGGL(k_tpq_colnorm, dim3(P.g), dim3(TQ_TPB), 0, e->stream, (const double2 *)P.V1, dim, k, P.sj,
                       P.sc, P.part);
    hipLaunchKernelGGL(k_tpq_start, dim3(1), dim3(64), 0, e->stream, (const double *)P.part, P.g, k, P.state);
    if (step0 > 0) {  // resumed: the given vector is step step0 - 1's, its row is not kept
      hipLaunchKernelGGL(k_tpq_set_step, dim3(1), dim3(1), 0, e->stream, P.state, (double)(step0 - 1));
      P.row0 = step0;
    }
    if (hipGetLastError() != hipSuccess) { rc = hphi_set_error("hphi_tpq: init launch failed"); break; }
    if (hipEventRecord(ev0, e->stream) != hipSuccess) { rc = hphi_set_error("event record failed"); break; }
    // W_end receives the vectors of the last step (steps - 1), normalised, as
    // the reference writes them for a restart (CalcByTPQ.c:272-287): they are
    // taken right before that step, whose fused pass moves on to the next vector
    const int s_first = step0 > 0 ? step0 - 1 : 0, last = p->steps - 1;
    auto store = [&]() {
      hipLaunchKernelGGL(k_tpq_store, dim3(gn), dim3(256), 0, e->stream, (double2 *)W_end, (const double2 *)P.V1, dim,
                         k, P.sj, P.sc, (const double *)P.state);
    };
    if (W_end && s_first == last) store();
    if ((rc = tpq_enqueue_step(e, P))) break;  // the first step eagerly (also loads the code objects)
    int s = s_first + 1;
    if (p->use_graph && s <= last) {
      if (hipStreamBeginCapture(e->stream, hipStreamCaptureModeRelaxed) != hipSuccess) {
        rc = hphi_set_error("hipStreamBeginCapture failed");
        break;
      }
      int erc = tpq_enqueue_step(e, P);
      hipError_t ce = hipStreamEndCapture(e->stream, &graph);
      if (erc || ce != hipSuccess) { rc = hphi_set_error("graph capture of the TPQ step failed"); break; }
      if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        rc = hphi_set_error("hipGraphInstantiate failed");
        break;
      }
      for (; s <= last && !rc; s++) {
        if (W_end && s == last) store();
        if (hipGraphLaunch(exec, e->stream) != hipSuccess) rc = hphi_set_error("hipGraphLaunch failed");
      }
    } else {
      for (; s <= last && !rc; s++) {
        if (W_end && s == last) store();
        rc = tpq_enqueue_step(e, P);
      }
    }
    if (rc) break;
    if (hipGetLastError() != hipSuccess) { rc = hphi_set_error("hphi_tpq: launch failed"); break; }
    if (hipEventRecord(ev1, e->stream) != hipSuccess) { rc = hphi_set_error("event record failed"); break; }
    if (hipStreamSynchronize(e->stream) != hipSuccess) { rc = hphi_set_error("hphi_tpq: stream sync failed"); break; }
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ev0, ev1);
    if (hipMemcpy(out_host, P.out, sizeof(double) * (size_t)(p->steps - step0) * k * TQ_NOUT,
                  hipMemcpyDeviceToHost) != hipSuccess) {
      rc = hphi_set_error("hphi_tpq: copy of results failed");
      break;
    }
    if (r) {
      r->seconds = ms * 1e-3;
      r->large_value = P.lv;
      r->steps = p->steps;
    }
  } while (0);
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  (void)hipFree(P.V0);
  (void)hipFree(P.V1);
  (void)hipFree(P.state);
  (void)hipFree(P.part);
  (void)hipFree(P.out);
  return rc;
}

// ---------------------------------------------------------------------------
// Canonical TPQ (CalcByCanonicalTPQ, src/CalcByCanonicalTPQ.c:52-300): each
// step applies exp(-dtau H / 2) to the normalised psi by its Taylor series to
// order ExpandCoef (MultiplyForCanonicalTPQ, src/Multiply.c:152-205), reusing H psi
// of the previous measurement, normalises, and measures (expec_energy_flct):
// ExpandCoef block H·v per step.  Vectors are state-major [dim][k].
// ---------------------------------------------------------------------------

// A = P + t Hp
__global__ __launch_bounds__(256) void k_ct_acc0(double2 *A, const double2 *P, const double2 *Hp, uint64_t n, double t) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    const double2 p = P[i], h = Hp[i];
    A[i] = make_double2(p.x + t * h.x, p.y + t * h.y);
  }
}

// A += t C
__global__ __launch_bounds__(256) void k_ct_acc(double2 *A, const double2 *Cv, uint64_t n, double t) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    const double2 a = A[i], c = Cv[i];
    A[i] = make_double2(a.x + t * c.x, a.y + t * c.y);
  }
}

// norms of the k columns from k_tpq_colnorm partials: state[c] = |A_c|,
// state[16 + c] = 1 / |A_c|
__global__ void k_ct_norms(const double *partial, int nblk, int k, double *state) {
  const int c = threadIdx.x;
  if (c < k) {
    double n = 0.0;
    for (int b = 0; b < nblk; b++) n += partial[(uint64_t)b * k + c];
    state[c] = sqrt(n);
    state[16 + c] = 1.0 / sqrt(n);
  }
}

// P[j][c] = A[j][c] / |A_c|
__global__ __launch_bounds__(256) void k_ct_scale(double2 *P, const double2 *A, uint64_t dim, int k,
                                                  const double *state) {
  const uint64_t n = dim * (uint64_t)k;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    const double s = state[16 + (int)(i % (uint64_t)k)];
    const double2 a = A[i];
    P[i] = make_double2(a.x * s, a.y * s);
  }
}

// per column: |h|^2 and the moments of |p|^2 (kinds as k_tpq_fused)
__global__ __launch_bounds__(TQ_TPB) void k_ct_moments(const double2 *P, const double2 *Hp, uint64_t dim, int k,
                                                       int obs_kind, int nsite, const uint64_t *list1,
                                                       double *partial) {
  const int c = threadIdx.x % k, jl = threadIdx.x / k, spb = TQ_TPB / k;
  double r[TQ_NRED] = {};
  const uint64_t EVEN = 0x5555555555555555ull;
  for (uint64_t j = (uint64_t)blockIdx.x * spb + jl; j < dim; j += (uint64_t)gridDim.x * spb) {
    const double2 w = P[j * k + c], h = Hp[j * k + c];
    r[0] += h.x * h.x + h.y * h.y;
    if (obs_kind) {
      const double ww = w.x * w.x + w.y * w.y;
      const uint64_t x = list1 ? list1[j] : j;
      double S;
      if (obs_kind == 1) {
        S = (double)(2 * (int)__popcll(x) - nsite);
      } else {
        const double D = (double)(int)__popcll(x & (x >> 1) & EVEN);
        const int up = (int)__popcll(x & EVEN), dn = (int)__popcll(x & (EVEN << 1));
        const double N = (double)(up + dn);
        S = (double)(up - dn);
        r[1] += ww * D;
        r[2] += ww * D * D;
        r[3] += ww * N;
        r[4] += ww * N * N;
      }
      r[5] += ww * S;
      r[6] += ww * S * S;
    }
  }
  __shared__ double red[TQ_TPB / 64][16][TQ_NRED];
#pragma unroll
  for (int q = 0; q < TQ_NRED; q++) {
    double s = r[q];
    for (int off = 32; off >= k; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) < k) red[threadIdx.x >> 6][threadIdx.x & 63][q] = s;
  }
  __syncthreads();
  if (threadIdx.x < k * TQ_NRED) {
    const int cc = threadIdx.x / TQ_NRED, q = threadIdx.x % TQ_NRED;
    double t = 0.0;
    for (int w = 0; w < TQ_TPB / 64; w++) t += red[w][cc][q];
    partial[((uint64_t)blockIdx.x * k + cc) * TQ_NRED + q] = t;
  }
}

// W_end column c (element c*dim + j) = P[j][c]
__global__ void k_ct_store(double2 *W_end, const double2 *P, uint64_t dim, int k) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= dim * (uint64_t)k) return;
  const uint64_t c = i / dim, j = i % dim;
  W_end[i] = P[j * (uint64_t)k + c];
}

// row s of sample c (the mTPQ row layout): inv_temp (s additions of dtau), E, <H^2>, the
// norm before normalisation (step 0: the start vector's), the moments
__global__ void k_ct_finalize(const double *partial, int nblk, int k, const double2 *prdct, const double *state,
                              int s, double dtau, int obs_kind, int nsite, double sz_const, double *out, int row0) {
  const int c = threadIdx.x;
  if (c >= k || s < row0) return;
  double r[TQ_NRED] = {};
  for (int b = 0; b < nblk; b++)
    for (int q = 0; q < TQ_NRED; q++) r[q] += partial[((uint64_t)b * k + c) * TQ_NRED + q];
  double *o = out + ((uint64_t)(s - row0) * k + c) * TQ_NOUT;
  double it = 0.0;  // inv_temp += delta_tau once per step (CalcByCanonicalTPQ.c:227)
  for (int i = 0; i < s; i++) it += dtau;
  o[0] = it;
  o[1] = prdct[c].x;
  o[2] = r[0];
  o[3] = state[c];
  if (obs_kind == 2) {
    o[4] = r[1];
    o[5] = r[2];
    o[6] = r[3];
    o[7] = r[4];
  } else {
    o[4] = o[5] = 0.0;
    o[6] = (double)nsite;
    o[7] = (double)nsite * nsite;
  }
  o[8] = obs_kind ? 0.5 * r[5] : sz_const;
  o[9] = obs_kind ? 0.25 * r[6] : sz_const * sz_const;
  o[10] = prdct[c].y;
  o[11] = 0.0;
}

static int ctpq_run(hphi_engine *e, const hphi_ctpq_params *p, const double *W0, int step0, double *out_host,
                    hphi_tpq_result *r, double *W_end);

extern "C" int hphi_ctpq(hphi_engine *e, const hphi_ctpq_params *p, double *out_host, hphi_tpq_result *r) {
  return ctpq_run(e, p, nullptr, 0, out_host, r, nullptr);
}

extern "C" int hphi_ctpq_from(hphi_engine *e, const hphi_ctpq_params *p, const double *W0, int step0,
                              double *out_host, hphi_tpq_result *r, double *W_end) {
  if (!W0) return hphi_set_error("hphi_ctpq_from: NULL start vectors");
  if (step0 < 0 || step0 >= (p ? p->steps : 0)) return hphi_set_error("hphi_ctpq_from: step0 must be in [0, steps)");
  return ctpq_run(e, p, W0, step0, out_host, r, W_end);
}

// W0 as in tpq_run: the start vectors (step0 = 0), or the normalised vectors of
// step step0 - 1 of an earlier run, measured again without a row; rows of steps
// step0 .. steps-1 go to out_host; W_end receives the last step's vectors.
static int ctpq_run(hphi_engine *e, const hphi_ctpq_params *p, const double *W0, int step0, double *out_host,
                    hphi_tpq_result *r, double *W_end) {
  if (!e || !p || !out_host) return hphi_set_error("hphi_ctpq: NULL argument");
  const int k = p->k;
  if (k != 1 && k != 2 && k != 4 && k != 8 && k != 16) return hphi_set_error("hphi_ctpq: k must be 1, 2, 4, 8 or 16");
  if (p->steps < 1) return hphi_set_error("hphi_ctpq: steps must be positive");
  const int nc = p->expand_coef > 0 ? p->expand_coef : 10;  // CalcByCanonicalTPQ.c:71-76
  if (hphi_build_tables(e) || hphi_build_diag(e)) return -1;
  const uint64_t dim = e->dim, n = dim * (uint64_t)k;
  double lv = p->large_value;
  if (lv <= 0.0 && default_large_value(e, &lv)) return -1;
  const double dtau = 1.0 / lv;  // CalcByCanonicalTPQ.c:143
  const int obs_kind = e->model == HPHI_MODEL_SPINGC ? 1 : hphi_model_hubbard(e->model) ? 2 : 0;
  const double sz_const = 0.5 * (double)(e->nup - e->ndown);
  const uint64_t *list1 = hphi_model_gc(e->model) ? (const uint64_t *)nullptr : (const uint64_t *)e->d_list1;
  const int g = tq_grid(dim, k);
  const unsigned gs = (unsigned)std::min<uint64_t>((n + 255) / 256, 65536);
  const size_t vbytes = sizeof(double2) * n;
  double2 *P = nullptr, *Hp = nullptr, *A = nullptr, *Cv = nullptr;
  double *state = nullptr, *part = nullptr, *out = nullptr;
  int rc = 0;
  hipEvent_t ev0, ev1;
  (void)hipEventCreate(&ev0);
  (void)hipEventCreate(&ev1);
  do {
    if (hipMalloc(&P, vbytes) != hipSuccess || hipMalloc(&Hp, vbytes) != hipSuccess ||
        hipMalloc(&A, vbytes) != hipSuccess || hipMalloc(&Cv, vbytes) != hipSuccess ||
        hipMalloc(&state, sizeof(double) * 64) != hipSuccess ||
        hipMalloc(&part, sizeof(double) * (size_t)TQ_GRID * k * TQ_NRED) != hipSuccess ||
        hipMalloc(&out, sizeof(double) * (size_t)(p->steps - step0) * k * TQ_NOUT) != hipSuccess) {
      (void)hipGetLastError();
      rc = hphi_set_error("hphi_ctpq: device allocation failed");
      break;
    }
    if (hipEventRecord(ev0, e->stream) != hipSuccess) { rc = hphi_set_error("event record failed"); break; }
    // step 0: start vector (MakeIniVec), normalised, measured (CalcByCanonicalTPQ.c:140-175)
    if (W0)
      hipLaunchKernelGGL(k_tpq_load, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, e->stream, A,
                         (const double2 *)W0, dim, k, (uint64_t)k, (uint64_t)1);
    else
      hipLaunchKernelGGL(k_tpq_init, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, e->stream, A, dim, k,
                         (uint64_t)k, (uint64_t)1, p->seed, p->init_kind);
    const int s_first = step0 > 0 ? step0 - 1 : 0;  // a resumed run re-measures the given vector first
    for (int s = s_first; s < p->steps && !rc; s++) {
      if (s > s_first) {  // A = sum_{m <= nc} (-dtau/2)^m / m! H^m P, with H P = Hp (MultiplyForCanonicalTPQ)
        double t = -0.5 * dtau;
        hipLaunchKernelGGL(k_ct_acc0, dim3(gs), dim3(256), 0, e->stream, A, (const double2 *)P,
                           (const double2 *)Hp, n, t);
        double2 *cur = Hp, *nxt = Cv;
        for (int m = 2; m <= nc && !rc; m++) {
          t *= (-0.5 * dtau) / (double)m;
          if (hphi_hv_launch_ex(e, k, (double *)nxt, (const double *)cur, 1)) { rc = -1; break; }
          hipLaunchKernelGGL(k_ct_acc, dim3(gs), dim3(256), 0, e->stream, A, (const double2 *)nxt, n, t);
          std::swap(cur, nxt);
        }
        if (rc) break;
      }
      // normalise A into P (global_norm; step 0: global_1st_norm)
      hipLaunchKernelGGL(k_tpq_colnorm, dim3(g), dim3(TQ_TPB), 0, e->stream, (const double2 *)A, dim, k, (uint64_t)k,
                         (uint64_t)1, part);
      hipLaunchKernelGGL(k_ct_norms, dim3(1), dim3(64), 0, e->stream, (const double *)part, g, k, state);
      hipLaunchKernelGGL(k_ct_scale, dim3(gs), dim3(256), 0, e->stream, P, (const double2 *)A, dim, k,
                         (const double *)state);
      // measure: Hp = H P (its prdct is E), <H^2> and the moments (expec_energy_flct)
      if (hphi_hv_launch_ex(e, k, (double *)Hp, (const double *)P, 1)) { rc = -1; break; }
      hipLaunchKernelGGL(k_ct_moments, dim3(g), dim3(TQ_TPB), 0, e->stream, (const double2 *)P, (const double2 *)Hp,
                         dim, k, obs_kind, e->nsite, list1, part);
      hipLaunchKernelGGL(k_ct_finalize, dim3(1), dim3(64), 0, e->stream, (const double *)part, g, k,
                         (const double2 *)e->d_prdct, (const double *)state, s, dtau, obs_kind, e->nsite, sz_const,
                         out, step0);
      if (hipGetLastError() != hipSuccess) { rc = hphi_set_error("hphi_ctpq: launch failed"); break; }
    }
    if (rc) break;
    if (hipEventRecord(ev1, e->stream) != hipSuccess || hipStreamSynchronize(e->stream) != hipSuccess) {
      rc = hphi_set_error("hphi_ctpq: stream sync failed");
      break;
    }
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ev0, ev1);
    if (W_end) {  // P holds the last step's normalised vectors, state-major
      hipLaunchKernelGGL(k_ct_store, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, e->stream, (double2 *)W_end,
                         (const double2 *)P, dim, k);
      if (hipGetLastError() != hipSuccess || hipStreamSynchronize(e->stream) != hipSuccess) {
        rc = hphi_set_error("hphi_ctpq: copy of the final vectors failed");
        break;
      }
    }
    if (hipMemcpy(out_host, out, sizeof(double) * (size_t)(p->steps - step0) * k * TQ_NOUT,
                  hipMemcpyDeviceToHost) != hipSuccess) {
      rc = hphi_set_error("hphi_ctpq: copy of results failed");
      break;
    }
    if (r) {
      r->seconds = ms * 1e-3;
      r->large_value = lv;
      r->steps = p->steps;
    }
  } while (0);
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  (void)hipFree(P);
  (void)hipFree(Hp);
  (void)hipFree(A);
  (void)hipFree(Cv);
  (void)hipFree(state);
  (void)hipFree(part);
  (void)hipFree(out);
  return rc;
}
