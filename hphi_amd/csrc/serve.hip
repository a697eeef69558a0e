// Continuous-batching Lanczos service: many small diagonalisation jobs share
// every kernel launch.
//
// Each job is a canonical Hamiltonian (Spin or Hubbard) whose lowest n
// eigenvalues are wanted.  Jobs arrive at any time; a tick
//   1. admits waiting jobs in arrival order while their HBM footprint
//      (tables + diagonal + two Lanczos vectors) fits the arena and a slot is
//      free, and builds their tables, diagonal and start vector on the device;
//   2. advances every running job by one step of the reference's plain
//      Lanczos recurrence (src/Lanczos_EigenValue.c:153-177) with four packed
//      launches for all jobs together (residual + norm, beta, H·v, alpha)
//      and one launch of the batched eigenvalue check;
//   3. copies the new (alpha, beta) and the check results of all jobs to the
//      host in one transfer each, tests convergence of the lowest n distinct
//      Ritz values (Cullum-Willoughby filtered) and retires converged jobs,
//      freeing their arena memory for the next admissions.
//
// The recurrence is rearranged so that every vector is touched by two passes
// per step instead of four.  A job keeps two arrays, R and Q.  At step m,
// R holds r = v0 after the previous H·v and Q holds w_{m-1}, the previous
// Lanczos vector before normalisation (v_{m-1} = s_{m-1} w_{m-1}).
//   pass 1: w_m = r - alpha_{m-1} s_{m-1} w_{m-1}, written over R, and |w_m|^2;
//   beta_{m-1} = |w_m|, s_m = 1/beta_{m-1};
//   pass 2: Q <- -beta_{m-1} s_{m-1} Q + s_m (H w_m)  (= -beta v_{m-1} + H v_m)
//           and alpha_m = s_m^2 <w_m|H w_m>;
// then R and Q swap roles.  This is the reference's v1 <- (v0 - alpha v1)/beta,
// v0 <- -beta v1 + H v1 with the division by beta folded into the H·v.
// Tiles map to jobs through a prefix table; a job's reductions run over its
// own tiles in a fixed order, so its result is bit-identical whether it runs
// alone or among other jobs.
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
#include "hv_device.h"

#define HIPCHK(x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) return hphi_set_error("%s failed: %s", #x, hipGetErrorString(e_)); \
  } while (0)

#define SV_TPB HV_TPB  // one tile = HV_TPB states

struct JobDev {
  HvArgs hv;              // hv.v1 = R (holds w_m after pass 1), hv.v0 = Q
  int slot, pad;
  int step, check;        // Lanczos step computed this tick (1-based); run the eigenvalue check
  double *ahist, *bhist;  // alpha[1..max_iter], beta[1..max_iter] on the device
};
typedef const JobDev __attribute__((address_space(4))) *CJob;
typedef const uint64_t __attribute__((address_space(4))) *CU64;

// job state per slot (device): [0] alpha_{m-1}, [1] s_{m-1}, [2] beta_{m-1}, [3] s_m
#define JS_STRIDE 4

__device__ __forceinline__ int find_job(CU64 prefix, int njobs, uint64_t tile) {
  int lo = 0, hi = njobs;  // prefix[lo] <= tile < prefix[hi]
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (prefix[mid] <= tile) lo = mid;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ HvArgs load_args(CJob J) {
  HvArgs A;
  A.dim = J->hv.dim;
  A.list1 = J->hv.list1;
  A.l21 = J->hv.l21;
  A.l22 = J->hv.l22;
  A.diag = J->hv.diag;
  A.terms = J->hv.terms;
  A.n_low = J->hv.n_low;
  A.n_high = J->hv.n_high;
  A.n_mixed = J->hv.n_mixed;
  A.nlow = J->hv.nlow;
  A.irght = J->hv.irght;
  A.v0 = J->hv.v0;
  A.v1 = J->hv.v1;
  A.partial = nullptr;
  A.xcd_map = 0;
  A.size21 = J->hv.size21;
  A.overwrite = 0;
  return A;
}

__device__ __forceinline__ void tile_sum_store(double s, double *partial) {
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  __shared__ double red[SV_TPB / 64];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < SV_TPB / 64; w++) t += red[w];
    partial[blockIdx.x] = t;
  }
}

// pass 1: w = r - alpha_{m-1} s_{m-1} q over R, |w|^2 per tile
// (the residual of Lanczos_EigenValue.c:154-160,170-177)
__global__ __launch_bounds__(SV_TPB) void k_mj_resid(const JobDev *jobs, const uint64_t *prefix, int njobs,
                                                     const double *js, double *partial) {
  const int jb = find_job((CU64)prefix, njobs, blockIdx.x);
  CJob J = (CJob)jobs + jb;
  const uint64_t j = (blockIdx.x - ((CU64)prefix)[jb]) * SV_TPB + threadIdx.x;
  double s = 0.0;
  if (j < J->hv.dim) {
    const double f = js[J->slot * JS_STRIDE + 0] * js[J->slot * JS_STRIDE + 1];
    double2 *R = (double2 *)J->hv.v1;
    double2 w = R[j];
    if (f != 0.0) {
      const double2 q = ((const double2 *)J->hv.v0)[j];
      w = make_double2(w.x - f * q.x, w.y - f * q.y);
      R[j] = w;
    }
    s = w.x * w.x + w.y * w.y;
  }
  tile_sum_store(s, partial);
}

// pass 2: Q <- c Q + s_m (H w) for one tile of one job, <w|H w> per tile
__global__ __launch_bounds__(SV_TPB) void k_mj_hv(const JobDev *jobs, const uint64_t *prefix, int njobs,
                                                  const double *js, double2 *partial) {
  const int jb = find_job((CU64)prefix, njobs, blockIdx.x);
  CJob J = (CJob)jobs + jb;
  const HvArgs A = load_args(J);
  const double *st = js + J->slot * JS_STRIDE;
  const double sc = st[3], c = -st[2] * st[1];
  const uint64_t j = (blockIdx.x - ((CU64)prefix)[jb]) * SV_TPB + threadIdx.x;
  double pre = 0.0, pim = 0.0;
  if (j < A.dim) {
    double2 a;
    double ar, ai;
    hv_row<1, true>(A, j, 0, a, ar, ai);
    double2 *Q = (double2 *)A.v0;
    const double2 p = Q[j];
    Q[j] = make_double2(c * p.x + sc * ar, c * p.y + sc * ai);
    pre = a.x * ar + a.y * ai;
    pim = a.x * ai - a.y * ar;
  }
  wg_reduce_store<1>(pre, pim, partial);
}

// per job (one workgroup each): fixed-order sum over the job's tiles
__device__ double job_sum(const double *part, int stride, int comp, uint64_t t0, uint64_t t1) {
  __shared__ double s[256];
  double a = 0.0;
  for (uint64_t t = t0 + threadIdx.x; t < t1; t += 256) a += part[t * stride + comp];
  s[threadIdx.x] = a;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) s[threadIdx.x] += s[threadIdx.x + st];
    __syncthreads();
  }
  double r = s[0];
  __syncthreads();
  return r;
}

// beta_{m-1} = |w_m| and s_m = 1/beta_{m-1}; a beta below 1e-14 ends the
// recurrence (the Krylov space is exhausted) and s_m = 0 keeps pass 2 finite
__global__ __launch_bounds__(256) void k_mj_norm(const JobDev *jobs, const uint64_t *prefix, const double *partial,
                                                 double *js) {
  CJob J = (CJob)jobs + blockIdx.x;
  const double b2 = job_sum(partial, 1, 0, ((CU64)prefix)[blockIdx.x], ((CU64)prefix)[blockIdx.x + 1]);
  if (threadIdx.x == 0) {
    const double b = sqrt(b2);
    double *st = js + J->slot * JS_STRIDE;
    st[2] = b;
    st[3] = (b > 0.0 && (J->step == 1 || b >= 1e-14)) ? 1.0 / b : 0.0;
    if (J->step >= 2) J->bhist[J->step - 1] = b;
  }
}

// alpha_m = s_m^2 Re <w|H w>; out = (alpha_m, beta_{m-1}) per job
__global__ __launch_bounds__(256) void k_mj_alpha(const JobDev *jobs, const uint64_t *prefix, const double2 *partial,
                                                  double *js, double *out) {
  CJob J = (CJob)jobs + blockIdx.x;
  const double a = job_sum((const double *)partial, 2, 0, ((CU64)prefix)[blockIdx.x], ((CU64)prefix)[blockIdx.x + 1]);
  if (threadIdx.x == 0) {
    double *st = js + J->slot * JS_STRIDE;
    const double sc = st[3];
    const double al = a * sc * sc;
    st[0] = al;
    st[1] = sc;
    J->ahist[J->step] = al;
    out[2 * blockIdx.x] = al;
    out[2 * blockIdx.x + 1] = st[2];
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

__global__ __launch_bounds__(64) void k_mj_check(const JobDev *jobs, int want, double *out) {
  CJob J = (CJob)jobs + blockIdx.x;
  if (!J->check) return;
  const int m = J->step;
  const double *a = J->ahist + 1, *b = J->bhist + 1;
  __shared__ double ev[32], evh[32];
  __shared__ double lo_s, hi_s, scale_s;
  int L = want + CK_EXTRA;
  if (L > m) L = m;
  if (L > 32) L = 32;
  int Lh = L < m - 1 ? L : m - 1;
  if (threadIdx.x == 0) {
    double lo = a[0], hi = a[0], sc = 0.0;
    for (int i = 0; i < m; i++) {
      double r = (i > 0 ? fabs(b[i - 1]) : 0.0) + (i < m - 1 ? fabs(b[i]) : 0.0);
      lo = fmin(lo, a[i] - r);
      hi = fmax(hi, a[i] + r);
      sc = fmax(sc, fabs(a[i]) + r);
    }
    double span = hi - lo;
    lo_s = lo - 1e-12 * span - 1e-300;
    hi_s = hi + 1e-12 * span + 1e-300;
    scale_s = sc;
  }
  __syncthreads();
  const int t = threadIdx.x;
  const bool main = t < L, hat = (t >= 32) && (t - 32 < Lh);
  if (main || hat) {
    const int idx = main ? t : t - 32;
    const double *aa = main ? a : a + 1, *bb = main ? b : b + 1;
    const int n = main ? m : m - 1;
    double l = lo_s, h = hi_s;
    for (int it = 0; it < 200; it++) {
      double mid = 0.5 * (l + h);
      if (mid <= l || mid >= h) break;
      if (ck_sturm(aa, bb, n, mid) > idx) h = mid;
      else l = mid;
    }
    if (main) ev[idx] = 0.5 * (l + h);
    else evh[idx] = 0.5 * (l + h);
  }
  __syncthreads();
  if (t == 0) {
    double *o = out + (uint64_t)blockIdx.x * 17;
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
  }
}

// start vector of one job: R = SplitMix64 uniform complex (normalised by the
// first pass), Q = 0, slot state zero
__global__ void k_sv_fill(double2 *v1, double2 *v0, uint64_t dim, uint64_t seed, double *js, int slot) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < JS_STRIDE) js[slot * JS_STRIDE + i] = 0.0;
  if (i >= dim) return;
  uint64_t z = (seed * 0x9E3779B97F4A7C15ull) ^ (2 * i);
  auto mix = [](uint64_t x) {
    uint64_t y = x + 0x9E3779B97F4A7C15ull;
    y = (y ^ (y >> 30)) * 0xBF58476D1CE4E5B9ull;
    y = (y ^ (y >> 27)) * 0x94D049BB133111EBull;
    return y ^ (y >> 31);
  };
  const double u1 = (double)(mix(z) >> 11) * (1.0 / 9007199254740992.0);
  const double u2 = (double)(mix((seed * 0x9E3779B97F4A7C15ull) ^ (2 * i + 1)) >> 11) * (1.0 / 9007199254740992.0);
  v1[i] = make_double2(2.0 * u1 - 1.0, 2.0 * u2 - 1.0);
  v0[i] = make_double2(0.0, 0.0);
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
        size_t rest = it->second - n, o = it->first;
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

struct Job {
  int id = 0;
  uint64_t seed = 0;
  double t_arrival = 0, t_admit = 0, t_done = 0;
  hphi_engine h;  // host-side preparation only (no device members used)
  size_t off = 0, bytes = 0;
  int slot = -1;
  JobDev dev{};
  std::vector<double> alpha, beta;
  int step = 0, nprev = 0;
  double prev[16];
  hphi_job_result res{};
};

size_t job_footprint(const hphi_engine &h, int max_iter) {
  auto al = [](size_t n) { return (n + 255) & ~(size_t)255; };
  return al(8 * h.dim) + al(4 * h.size21) + al(8 * h.size22) + al(8 * (h.sdim + 1)) + al(8 * h.dim) +
         2 * al(16 * h.dim) + al(sizeof(GTerm) * std::max<size_t>(h.h_terms_dev.size(), 1)) +
         al(sizeof(DTerm) * std::max<size_t>(h.h_dterms.size(), 1)) + 2 * al(8 * (size_t)(max_iter + 2));
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
  JobDev *d_jobs = nullptr;
  uint64_t *d_prefix = nullptr;
  char *h_stage = nullptr;  // pinned staging for term uploads
  size_t stage_size = 0, stage_used = 0;
  std::deque<Job *> waiting;
  std::vector<Job *> active;
  std::vector<int> free_slots;
  std::vector<hphi_job_result> done;
  int next_id = 0;
  double t0 = 0;
  size_t max_tiles = 0;
  double st_admit = 0, st_gpu = 0, st_check = 0;  // seconds spent per phase
  uint64_t st_ticks = 0, st_job_steps = 0;
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
  size_t budget = s->p.hbm_budget ? s->p.hbm_budget : ((size_t)16 << 30);
  s->max_tiles = (size_t)1 << 22;  // 4M tiles x 256 states = 1.07e9 states in flight
  int rc = 0;
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&s->d_arena, budget) != hipSuccess ||
      hipMalloc(&s->d_js, sizeof(double) * JS_STRIDE * s->p.max_jobs) != hipSuccess ||
      hipMalloc(&s->d_out, sizeof(double) * 2 * s->p.max_jobs) != hipSuccess ||
      hipMalloc(&s->d_jobs, sizeof(JobDev) * s->p.max_jobs) != hipSuccess ||
      hipMalloc(&s->d_prefix, sizeof(uint64_t) * (s->p.max_jobs + 1)) != hipSuccess ||
      hipMalloc(&s->d_part, sizeof(double2) * s->max_tiles) != hipSuccess ||
      hipHostMalloc(&s->h_out, sizeof(double) * 2 * s->p.max_jobs, hipHostMallocDefault) != hipSuccess ||
      hipMalloc(&s->d_ev, sizeof(double) * 17 * s->p.max_jobs) != hipSuccess ||
      hipHostMalloc(&s->h_ev, sizeof(double) * 17 * s->p.max_jobs, hipHostMallocDefault) != hipSuccess)
    rc = hphi_set_error("hphi_server_create: device allocation failed");
  s->stage_size = (size_t)8 << 20;
  if (!rc && hipHostMalloc(&s->h_stage, s->stage_size, hipHostMallocDefault) != hipSuccess)
    rc = hphi_set_error("hphi_server_create: pinned staging allocation failed");
  if (rc) {
    hphi_server_destroy(s);
    return -1;
  }
  s->arena.init(s->d_arena, budget);
  for (int i = s->p.max_jobs - 1; i >= 0; i--) s->free_slots.push_back(i);
  s->t0 = mono_now();
  *out = s;
  return 0;
}

extern "C" void hphi_server_destroy(hphi_server *s) {
  if (!s) return;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (Job *j : s->active) delete j;
  for (Job *j : s->waiting) delete j;
  (void)hipFree(s->d_arena);
  (void)hipFree(s->d_js);
  (void)hipFree(s->d_out);
  (void)hipFree(s->d_jobs);
  (void)hipFree(s->d_prefix);
  (void)hipFree(s->d_part);
  if (s->h_out) (void)hipHostFree(s->h_out);
  (void)hipFree(s->d_ev);
  if (s->h_ev) (void)hipHostFree(s->h_ev);
  if (s->h_stage) (void)hipHostFree(s->h_stage);
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
  j->bytes = job_footprint(j->h, s->p.max_iter);
  if (j->bytes > s->arena.size) {
    delete j;
    return hphi_set_error("job footprint %zu exceeds the HBM budget", j->bytes);
  }
  j->id = s->next_id++;
  j->seed = seed;
  j->t_arrival = t_arrival;
  s->waiting.push_back(j);
  if (id) *id = j->id;
  return 0;
}

static int admit(hphi_server *s, Job *j) {
  hphi_engine &h = j->h;
  char *b = s->d_arena + j->off;
  auto al = [](size_t n) { return (n + 255) & ~(size_t)255; };
  uint64_t *list1 = (uint64_t *)b;
  b += al(8 * h.dim);
  uint32_t *l21 = (uint32_t *)b;
  b += al(4 * h.size21);
  uint64_t *l22 = (uint64_t *)b;
  b += al(8 * h.size22);
  uint64_t *jb = (uint64_t *)b;
  b += al(8 * (h.sdim + 1));
  double *diag = (double *)b;
  b += al(8 * h.dim);
  double *v0 = (double *)b;
  b += al(16 * h.dim);
  double *v1 = (double *)b;
  b += al(16 * h.dim);
  GTerm *terms = (GTerm *)b;
  b += al(sizeof(GTerm) * std::max<size_t>(h.h_terms_dev.size(), 1));
  DTerm *dterms = (DTerm *)b;
  b += al(sizeof(DTerm) * std::max<size_t>(h.h_dterms.size(), 1));
  double *ahist = (double *)b;
  b += al(8 * (size_t)(s->p.max_iter + 2));
  double *bhist = (double *)b;
  // term tables through pinned staging (asynchronous copies)
  size_t tb = sizeof(GTerm) * h.h_terms_dev.size(), db = sizeof(DTerm) * h.h_dterms.size();
  if (s->stage_used + tb + db > s->stage_size) return 1;  // staging full: admit next tick
  char *st = s->h_stage + s->stage_used;
  if (tb) memcpy(st, h.h_terms_dev.data(), tb);
  if (db) memcpy(st + tb, h.h_dterms.data(), db);
  s->stage_used += (tb + db + 255) & ~(size_t)255;
  if (tb) HIPCHK(hipMemcpyAsync(terms, st, tb, hipMemcpyHostToDevice, s->stream));
  if (db) HIPCHK(hipMemcpyAsync(dterms, st + tb, db, hipMemcpyHostToDevice, s->stream));
  SzParams P = hphi_sz_params(&h);
  if (hphi_launch_tables(P, h.sdim, h.size21, jb, l21, l22, list1, s->stream)) return -1;
  if (hphi_launch_diag(h.dim, list1, 0, dterms, (int)h.h_dterms.size(), diag, s->stream)) return -1;
  unsigned g = (unsigned)((h.dim + 255) / 256);
  hipLaunchKernelGGL(k_sv_fill, dim3(g), dim3(256), 0, s->stream, (double2 *)v1, (double2 *)v0, h.dim, j->seed,
                     s->d_js, j->slot);
  HIPCHK(hipGetLastError());
  HvArgs &A = j->dev.hv;
  memset(&A, 0, sizeof A);
  A.dim = h.dim;
  A.list1 = list1;
  A.l21 = l21;
  A.l22 = l22;
  A.diag = diag;
  A.terms = terms;
  A.n_low = h.n_low;
  A.n_high = h.n_high;
  A.n_mixed = h.n_mixed;
  A.nlow = h.nlow;
  A.irght = h.irght;
  A.v0 = v0;
  A.v1 = v1;
  A.size21 = h.size21;
  j->dev.slot = j->slot;
  j->dev.ahist = ahist;
  j->dev.bhist = bhist;
  j->step = 0;
  j->alpha.assign(1, 0.0);
  j->beta.assign(1, 0.0);
  return 0;
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
  s->arena.release(j->off, j->bytes);
  s->free_slots.push_back(j->slot);
}

extern "C" int hphi_server_tick(hphi_server *s, int *n_active, int *n_waiting) {
  if (!s) return hphi_set_error("hphi_server_tick: NULL argument");
  double now = hphi_server_now(s);
  const double ta = now;
  s->stage_used = 0;
  // 1. admission in arrival order
  while (!s->waiting.empty() && !s->free_slots.empty()) {
    Job *j = s->waiting.front();
    if (j->t_arrival > now) break;
    size_t off;
    if (!s->arena.alloc(j->bytes, &off)) break;
    j->off = off;
    j->slot = s->free_slots.back();
    int rc = admit(s, j);
    if (rc == 1) {
      s->arena.release(off, j->bytes);
      break;
    }
    if (rc) return -1;
    s->free_slots.pop_back();
    s->waiting.pop_front();
    j->t_admit = now;
    s->active.push_back(j);
  }
  s->st_admit += hphi_server_now(s) - ta;
  if (!s->active.empty()) {
    const double tg = hphi_server_now(s);
    // 2. one Lanczos step for every running job, packed launches
    const int n = (int)s->active.size();
    std::vector<JobDev> jd(n);
    std::vector<uint64_t> prefix(n + 1, 0);
    for (int i = 0; i < n; i++) {
      Job *j = s->active[i];
      const int m = j->step + 1;  // step computed this tick
      j->dev.step = m;
      j->dev.check = (m % s->p.check_every == 0) || (uint64_t)m >= j->h.dim || m >= s->p.max_iter;
      jd[i] = j->dev;
      prefix[i + 1] = prefix[i] + (j->h.dim + SV_TPB - 1) / SV_TPB;
    }
    const uint64_t T = prefix[n];
    if (T > s->max_tiles) return hphi_set_error("too many tiles in flight");
    HIPCHK(hipMemcpyAsync(s->d_jobs, jd.data(), sizeof(JobDev) * n, hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(s->d_prefix, prefix.data(), sizeof(uint64_t) * (n + 1), hipMemcpyHostToDevice, s->stream));
    hipLaunchKernelGGL(k_mj_resid, dim3((unsigned)T), dim3(SV_TPB), 0, s->stream, s->d_jobs, s->d_prefix, n,
                       (const double *)s->d_js, s->d_part);
    hipLaunchKernelGGL(k_mj_norm, dim3(n), dim3(256), 0, s->stream, s->d_jobs, s->d_prefix,
                       (const double *)s->d_part, s->d_js);
    hipLaunchKernelGGL(k_mj_hv, dim3((unsigned)T), dim3(SV_TPB), 0, s->stream, s->d_jobs, s->d_prefix, n,
                       (const double *)s->d_js, (double2 *)s->d_part);
    hipLaunchKernelGGL(k_mj_alpha, dim3(n), dim3(256), 0, s->stream, s->d_jobs, s->d_prefix,
                       (const double2 *)s->d_part, s->d_js, s->d_out);
    hipLaunchKernelGGL(k_mj_check, dim3(n), dim3(64), 0, s->stream, s->d_jobs, s->p.n_eigen, s->d_ev);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(s->h_out, s->d_out, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipMemcpyAsync(s->h_ev, s->d_ev, sizeof(double) * 17 * n, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    now = hphi_server_now(s);
    s->st_gpu += now - tg;
    s->st_ticks++;
    s->st_job_steps += n;
    // 3. convergence and retirement
    std::vector<Job *> keep;
    keep.reserve(n);
    for (int i = 0; i < n; i++) {
      Job *j = s->active[i];
      j->step++;
      const int m = j->step;
      j->alpha.push_back(s->h_out[2 * i]);                  // alpha[m]
      if (m >= 2) j->beta.push_back(s->h_out[2 * i + 1]);  // beta[m-1]
      double *r = (double *)j->dev.hv.v1;                   // R and Q swap roles
      j->dev.hv.v1 = j->dev.hv.v0;
      j->dev.hv.v0 = r;
      // beta[m-1] ~ 0: the recurrence ended at step m-1 (alpha[m] is void)
      const bool broken = m >= 2 && j->beta[m - 1] < 1e-14;
      const bool exhausted = broken || (uint64_t)m >= j->h.dim;
      if (j->dev.check || exhausted) {
        double cur[16];
        int nc;
        if (broken) {
          j->step = m - 1;
          nc = hphi_lanczos_good_eigenvalues(j->alpha.data() + 1, j->beta.data() + 1, m - 1, s->p.n_eigen, cur);
        } else if (j->dev.check) {  // computed on the device this tick
          const double *o = s->h_ev + 17 * (size_t)i;
          nc = (int)o[16];
          for (int q = 0; q < nc; q++) cur[q] = o[q];
        } else {
          nc = hphi_lanczos_good_eigenvalues(j->alpha.data() + 1, j->beta.data() + 1, m, s->p.n_eigen, cur);
        }
        int want = (int)std::min<uint64_t>((uint64_t)s->p.n_eigen, j->h.dim);
        bool conv = hphi_lanczos_converged(j->prev, j->nprev, cur, nc, want, s->p.eps) ||
                    (exhausted && nc >= 1);
        if (conv || j->step >= s->p.max_iter) {
          retire(s, j, now, conv ? 1 : 0, cur, nc);
          delete j;
          continue;
        }
        memcpy(j->prev, cur, sizeof cur);
        j->nprev = nc;
      }
      keep.push_back(j);
    }
    s->active.swap(keep);
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
  return 0;
}

extern "C" int hphi_server_results(hphi_server *s, hphi_job_result *out, int max, int *n) {
  if (!s || !n) return hphi_set_error("hphi_server_results: NULL argument");
  int k = (int)std::min<size_t>(s->done.size(), (size_t)std::max(max, 0));
  for (int i = 0; i < k; i++) out[i] = s->done[i];
  s->done.erase(s->done.begin(), s->done.begin() + k);
  *n = k;
  return 0;
}
