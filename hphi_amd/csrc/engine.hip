// hphi_amd engine: device basis tables, diagonal and gather-form H·v for
// CDNA4 (gfx950).  C ABI declared in include/hphi_amd.h.
//
// Layout in HBM (per engine):
//   list1  uint64[dim]      canonical states in ascending order (sz.c list_1[1..dim])
//   l21    uint32[size21]   rank-in-class + 2, 0 = not in sector (sz.c list_2_1)
//   l22    uint64[size22]   states before the block + 1, 0 = empty block (sz.c list_2_2)
//   diag   double[dim]      list_Diagonal[1..dim]
//   terms  GTerm[...]       off-diagonal terms, grouped low/high/mixed
// Vectors are 0-based complex128 (double2); block vectors are state-major
// [dim][k] so that one gather moves k*16 contiguous bytes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <map>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>

#include "hphi_amd.h"
#include "engine_internal.h"

// ------------------------------------------------------------------ errors --
static thread_local char g_err[1024] = "";

extern "C" int hphi_set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return -1;
}

extern "C" const char *hphi_last_error(void) { return g_err; }

#define HIPCHK(x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) return hphi_set_error("%s failed: %s", #x, hipGetErrorString(e_)); \
  } while (0)

// ------------------------------------------------------------ binomials ----
__constant__ uint64_t c_binom[65][65];

static uint64_t h_binom(int n, int k) {
  if (k < 0 || n < 0 || k > n) return 0;
  uint64_t r = 1;
  for (int i = 1; i <= k; i++) r = r * (uint64_t)(n - k + i) / (uint64_t)i;
  return r;
}

__device__ __forceinline__ uint64_t dbinom(int n, int k) {
  return (k < 0 || n < 0 || k > n) ? 0ull : c_binom[n][k];
}

// Off-diagonal elements of H in a canonical sector: for every compiled term,
// the states x whose bits on M = tmask | flip form a pattern p that passes the
// term's predicate and whose source x ^ flip keeps the particle numbers, times
// the ways to fill the other bits (C(N - |M|, n - |p|) for Spin; one binomial
// per species for Hubbard, up orbitals at even bits).  0 for grand-canonical
// models (the caller counts those another way).
uint64_t hphi_term_nnz(const hphi_engine *e) {
  if (hphi_model_gc(e->model)) return 0;
  const bool hub = hphi_model_hubbard(e->model);
  const uint64_t up_bits = hub ? 0x5555555555555555ull : ~0ull;
  const int nb = e->nbits;
  const uint64_t all = nb >= 64 ? ~0ull : ((1ull << nb) - 1);
  uint64_t total = 0;
  for (const GTerm &t : e->h_terms_dev) {
    const uint64_t M = (t.tmask | t.flip) & all;
    int pos[64], k = 0;
    for (int b = 0; b < nb; b++)
      if ((M >> b) & 1) pos[k++] = b;
    if (k > 20) continue;  // no compiled term touches that many bits
    const int ku = __builtin_popcountll(M & up_bits), kd = k - ku;
    for (uint64_t m = 0; m < (1ull << k); m++) {
      uint64_t x = 0;
      for (int q = 0; q < k; q++)
        if ((m >> q) & 1) x |= 1ull << pos[q];
      const uint64_t ym = x & t.tmask;
      const int pc = __builtin_popcountll(ym);
      const bool act = (t.kind == P_EQ) ? (ym == t.tval) : ((t.kind == P_DIFF) ? (pc == 1) : (pc != 1));
      if (!act) continue;
      const uint64_t y = x ^ t.flip;
      if (hub) {
        const int xu = __builtin_popcountll(x & up_bits), xd = __builtin_popcountll(x & ~up_bits);
        if (xu != __builtin_popcountll(y & up_bits) || xd != __builtin_popcountll(y & ~up_bits)) continue;
        total += h_binom(e->nsite - ku, e->nup - xu) * h_binom(e->nsite - kd, e->ndown - xd);
      } else {
        const int xs = __builtin_popcountll(x);
        if (xs != __builtin_popcountll(y)) continue;
        total += h_binom(nb - k, e->nup - xs);
      }
    }
  }
  return total;
}

int hphi_ensure_binom_table() {
  static bool done = false;
  if (done) return 0;
  static uint64_t t[65][65];
  for (int n = 0; n <= 64; n++)
    for (int k = 0; k <= 64; k++) t[n][k] = h_binom(n, k);
  HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_binom), t, sizeof t));
  done = true;
  return 0;
}

// ------------------------------------------------------- basis tables ----
// Closed forms restating sz.c (see engine_internal.h for the parameters):
//   block prefix  jb[ib] = sum_{ib'<ib} cnt(ib'), evaluated per set bit of ib
//                 with Vandermonde's identity (no scan)
//   list_2_2[ib]  = jb[ib] + 1 for a non-empty block, else 0   (sz.c:1525,1069)
//   list_2_1[ia]  = rank of ia among the same class + 2, 0 when no block
//                 pairs with its class                           (sz.c:1524,1068)
//   list_1        unranked per block in ascending order          (sz.c:1523,1067)
// Spin: class = popcount.  Hubbard: class = (up, down) counts, up = even
// absolute bit (mltplyHubbardCore.c:45-46).

__device__ __forceinline__ void hub_counts(uint64_t x, int shift, int *nu, int *nd) {
  // counts of set bits at even / odd ABSOLUTE positions, x holding bits from `shift`
  const uint64_t even = 0x5555555555555555ull, odd = 0xAAAAAAAAAAAAAAAAull;
  uint64_t e = (shift & 1) ? odd : even, o = (shift & 1) ? even : odd;
  *nu = (int)__popcll(x & e);
  *nd = (int)__popcll(x & o);
}

// number of even (up) absolute positions among relative positions [0, p) of a
// half starting at absolute bit `shift`
__device__ __forceinline__ int n_up_below(int p, int shift) {
  return (shift & 1) ? p / 2 : (p + 1) / 2;
}

__global__ void k_block_prefix(SzParams P, uint64_t *jb, uint64_t *l22) {
  uint64_t ib = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (ib > P.sdim) return;
  uint64_t acc = 0;
  if (P.hubbard) {
    int su = 0, sd = 0;
    for (int p = P.nhigh; p >= 0; p--) {
      if (!((ib >> p) & 1)) continue;
      int eu = n_up_below(p, P.nlow), ed = p - eu;
      acc += dbinom(eu + P.nlu, P.nup - su) * dbinom(ed + P.nld, P.ndn - sd);
      if (((p + P.nlow) & 1) == 0) su++; else sd++;
    }
  } else {
    int s = 0;
    for (int p = P.nhigh; p >= 0; p--) {
      if (!((ib >> p) & 1)) continue;
      acc += dbinom(p + P.nlow, P.nup - s);
      s++;
    }
  }
  jb[ib] = acc;
  if (ib < P.sdim) {
    uint64_t cnt;
    if (P.hubbard) {
      int nu, nd;
      hub_counts(ib, P.nlow, &nu, &nd);
      cnt = dbinom(P.nlu, P.nup - nu) * dbinom(P.nld, P.ndn - nd);
    } else {
      cnt = dbinom(P.nlow, P.nup - (int)__popcll(ib));
    }
    l22[ib] = cnt ? acc + 1 : 0;
  } else {
    l22[ib] = 0;
    l22[ib + 1] = 0;
  }
}

__global__ void k_list21(SzParams P, uint32_t *l21) {
  uint64_t ia = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (ia >= P.size21) return;
  uint32_t val = 0;
  if (ia < P.ihfbit) {
    if (P.hubbard) {
      int a, b;
      hub_counts(ia, 0, &a, &b);
      if (P.nup - a >= 0 && P.nup - a <= P.nhu && P.ndn - b >= 0 && P.ndn - b <= P.nhd) {
        uint64_t rank = 0;
        int su = 0, sd = 0;
        for (int p = P.nlow - 1; p >= 0; p--) {
          if (!((ia >> p) & 1)) continue;
          int eu = (p + 1) / 2, ed = p / 2;
          rank += dbinom(eu, a - su) * dbinom(ed, b - sd);
          if ((p & 1) == 0) su++; else sd++;
        }
        val = (uint32_t)(rank + 2);
      }
    } else {
      int k = (int)__popcll(ia);
      if (P.nup - k >= 0 && P.nup - k <= P.nhigh) {
        uint64_t rank = 0;
        int s = 0;
        for (int p = P.nlow - 1; p >= 0; p--) {
          if (!((ia >> p) & 1)) continue;
          rank += dbinom(p, k - s);
          s++;
        }
        val = (uint32_t)(rank + 2);
      }
    }
  }
  l21[ia] = val;
}

// one workgroup per high block; lanes unrank consecutive ranks (coalesced writes)
__global__ __launch_bounds__(256) void k_list1(SzParams P, const uint64_t *jb, uint64_t *list1) {
  for (uint64_t ib = blockIdx.x; ib < P.sdim; ib += gridDim.x) {
    uint64_t base = jb[ib], end = jb[ib + 1];
    if (end > P.dim) end = P.dim;  // never write past list1, even for inconsistent prefixes
    if (end <= base) continue;
    uint64_t cnt = end - base;
    int a, b;
    if (P.hubbard) {
      int nu, nd;
      hub_counts(ib, P.nlow, &nu, &nd);
      a = P.nup - nu;
      b = P.ndn - nd;
    } else {
      a = P.nup - (int)__popcll(ib);
      b = 0;
    }
    for (uint64_t r = threadIdx.x; r < cnt; r += blockDim.x) {
      uint64_t rem = r, ia = 0;
      int ra = a, rb = b;
      if (P.hubbard) {
        for (int p = P.nlow - 1; p >= 0; p--) {
          uint64_t c = dbinom((p + 1) / 2, ra) * dbinom(p / 2, rb);
          if (rem >= c) {
            ia |= 1ull << p;
            rem -= c;
            if ((p & 1) == 0) ra--; else rb--;
          }
        }
      } else {
        for (int p = P.nlow - 1; p >= 0 && ra > 0; p--) {
          uint64_t c = dbinom(p, ra);
          if (rem >= c) {
            ia |= 1ull << p;
            rem -= c;
            ra--;
          }
        }
      }
      list1[base + r] = (ib << P.nlow) | ia;
    }
  }
}

// --------------------------------------------------------------- diagonal --
// One lane per state, terms applied in the reference's order
// (diagonalcalc.c:105-180: CoulombIntra, Chemi, CoulombInter, Hund) with the
// same expressions, so the result is bit-identical.
__global__ void k_diag(uint64_t dim, const uint64_t *__restrict__ list1, int gc,
                       const DTerm *__restrict__ terms, int nterms, double *__restrict__ diag, int stride,
                       int col) {
  uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= dim) return;
  uint64_t x = gc ? j : list1[j];
  double D = 0.0;
  for (int t = 0; t < nterms; t++) {
    const DTerm T = terms[t];
    switch (T.kind) {
      case D_INTRA:
        if ((x & T.m1) == T.m1) D += T.V;
        break;
      case D_CHEMI_H: {
        uint64_t n = (x & T.m1) / T.m1;
        D += (double)n * T.V;
      } break;
      case D_CHEMI_S: {
        uint64_t b = ((x & T.m1) / T.m1) ^ T.m2;
        D += T.V * (double)b;
      } break;
      case D_INTER_H: {
        uint64_t n1 = (x & T.m1) / T.m1 + (x & T.m2) / T.m2;
        uint64_t n2 = (x & T.m3) / T.m3 + (x & T.m4) / T.m4;
        D += (double)(n1 * n2) * T.V;
      } break;
      case D_INTER_S:
        D += T.V;
        break;
      case D_HUND_H: {
        uint64_t a = (x & T.m1) / T.m1, b = (x & T.m2) / T.m2;
        uint64_t c = (x & T.m3) / T.m3, d = (x & T.m4) / T.m4;
        D += T.V * (double)(a * c + b * d);
      } break;
      case D_HUND_S: {
        uint64_t ibit = x & T.m1;
        if (ibit == 0 || ibit == T.m1) D += T.V;
      } break;
      case D_NN: {  // SetDiagonalInterAll, Hubbard (diagonalcalc.c:1113-1435)
        uint64_t n1 = (x & T.m1) / T.m1, n2 = (x & T.m2) / T.m2;
        D += (double)(n1 * n2) * T.V;
      } break;
      case D_SS: {  // SetDiagonalInterAll, spin: [bit_i == s1][bit_k == s3] V
        uint64_t n1 = ((x & T.m1) / T.m1) ^ T.m3, n2 = ((x & T.m2) / T.m2) ^ T.m4;
        D += (double)(n1 * n2) * T.V;
      } break;
    }
  }
  diag[j * stride + col] = D;
}

// ------------------------------------------------------------------- H·v --
// Gather form: lane (state j, vector v) accumulates
//   acc = diag[j] v1[j] + sum_t c_t * sgn_t(x) * v1[idx(x ^ flip_t)]
// over the terms whose predicate holds on x = list1[j] (or j for SpinGC).
// This is the transpose of the reference's per-term scatter
// v0[idx(s ^ flip)] += c v1[s]: every (term, source) pair of the reference
// appears once as a (term, target) pair here, with the same coefficient and
// the same Jordan-Wigner sign (the sign mask never overlaps the flipped bits).
// prdct = sum_j conj(v1[j]) * acc_j equals the reference's sum of
// c v1[s] conj(v1[target]) over the same pairs (mltply.c:100, mltplySpinCore.c:304).
//
// Loads are made unconditional (inactive lanes re-read their own element and
// weight it by zero) so the unrolled term loop keeps many gathers in flight.

#include "hv_device.h"

template <int K, bool CANON>
__global__ __launch_bounds__(HV_TPB) void k_hv(const HvArgs A) {
  constexpr int SPT = HV_TPB / K;
  double pre = 0.0, pim = 0.0;
  const uint64_t ntiles = (A.dim + SPT - 1) / SPT;
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
  for (uint64_t tile = t_begin; tile < t_end; tile += t_step) hv_tile<K, CANON>(A, tile, pre, pim);
  wg_reduce_store<K>(pre, pim, A.partial);
}

// deterministic final reduction: out[v] = sum_w partial[w*K + v] in fixed order
__global__ __launch_bounds__(256) void k_reduce_partials(const double2 *partial, int nblk, int K,
                                                         double2 *out) {
  __shared__ double2 s[256];
  for (int v = 0; v < K; v++) {
    double re = 0.0, im = 0.0;
    for (int w = threadIdx.x; w < nblk; w += 256) {
      re += partial[(uint64_t)w * K + v].x;
      im += partial[(uint64_t)w * K + v].y;
    }
    s[threadIdx.x] = make_double2(re, im);
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
      if (threadIdx.x < st) {
        s[threadIdx.x].x += s[threadIdx.x + st].x;
        s[threadIdx.x].y += s[threadIdx.x + st].y;
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) out[v] = s[0];
    __syncthreads();
  }
}

// ------------------------------------------------------------ vectors ----
__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double unit_rng(uint64_t seed, uint64_t c) {
  return (double)(splitmix64((seed * 0x9E3779B97F4A7C15ull) ^ c) >> 11) * (1.0 / 9007199254740992.0);
}

__global__ void k_fill_random(double2 *v, uint64_t dim, int k, uint64_t seed) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= dim * (uint64_t)k) return;
  uint64_t j = i / k;
  int c = (int)(i % k);
  uint64_t s = seed + (uint64_t)c;
  v[i] = make_double2(2.0 * unit_rng(s, 2 * j) - 1.0, 2.0 * unit_rng(s, 2 * j + 1) - 1.0);
}

__global__ void k_zero(double2 *v, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = make_double2(0.0, 0.0);
}

// ====================================================================== host
static int grid_for(uint64_t n, int tpb) { return (int)((n + tpb - 1) / tpb); }

int hphi_engine_alloc_scratch(hphi_engine *e) {
  if (e->d_partial) return 0;
  HIPCHK(hipMalloc(&e->d_partial, sizeof(double2) * (size_t)HV_MAX_GRID * 16));
  HIPCHK(hipMalloc(&e->d_prdct, sizeof(double2) * 16));
  HIPCHK(hipMalloc(&e->d_coef, sizeof(double2) * 64));
  HIPCHK(hipHostMalloc(&e->h_prdct, sizeof(double2) * 16, hipHostMallocDefault));
  e->bytes += sizeof(double2) * ((size_t)HV_MAX_GRID * 16 + 16);
  return 0;
}

static void classify_and_upload_terms(hphi_engine *e) {
  // Identical terms (same flip, predicate and sign mask) act on the same pairs
  // of states: they become one term with the summed coupling, in order of first
  // appearance (the 13x2 triangular torus lists each a2 bond twice).  Terms
  // whose couplings cancel exactly are dropped.
  std::vector<GTerm> merged;
  std::map<std::tuple<uint64_t, uint64_t, uint64_t, uint64_t, int>, size_t> first;
  for (const GTerm &t : e->h_terms) {
    const auto key = std::make_tuple(t.flip, t.tmask, t.tval, t.sgn, t.kind);
    const auto it = first.find(key);
    if (it == first.end()) {
      first.emplace(key, merged.size());
      merged.push_back(t);
    } else {
      merged[it->second].cre += t.cre;
      merged[it->second].cim += t.cim;
    }
  }
  // A hop and its Hermitian partner (c+_a c_b and c+_b c_a, same coupling):
  // both gather from x ^ (a|b) with the same between-bits sign, one when x
  // holds a and not b, the other when x holds b and not a, i.e. together
  // when exactly one of the two bits is set.  The pair becomes one P_DIFF
  // term: half the predicates and rank lookups for the same gathers.
  {
    std::map<std::tuple<uint64_t, uint64_t, double, double>, size_t> open;  // (flip, sgn, c) -> index
    std::vector<char> drop(merged.size(), 0);
    for (size_t i = 0; i < merged.size(); i++) {
      GTerm &t = merged[i];
      if (t.kind != P_EQ || t.tmask != t.flip || __builtin_popcountll(t.flip) != 2 ||
          __builtin_popcountll(t.tval) != 1 || (t.tval & ~t.flip))
        continue;
      const auto key = std::make_tuple(t.flip, t.sgn, t.cre, t.cim);
      const auto it = open.find(key);
      if (it != open.end() && merged[it->second].tval == (t.flip ^ t.tval)) {
        GTerm &p = merged[it->second];
        p.kind = P_DIFF;
        p.tval = 0;
        drop[i] = 1;
        open.erase(it);
      } else if (it == open.end()) {
        open.emplace(key, i);
      }
    }
    std::vector<GTerm> kept;
    for (size_t i = 0; i < merged.size(); i++)
      if (!drop[i]) kept.push_back(merged[i]);
    merged.swap(kept);
  }
  // groups: low (flip within the low half), high (within the high half), mixed
  std::vector<GTerm> low, high, mixed;
  for (const GTerm &t : merged) {
    if (t.cre == 0.0 && t.cim == 0.0) continue;
    if (hphi_model_gc(e->model)) {
      low.push_back(t);
      continue;
    }
    // the high group's wave-uniform path reads predicate and sign from the high
    // half only: terms that test or count low bits go to the general mixed group
    bool lo = (t.flip & ~e->irght) == 0;
    bool hi = (t.flip & e->irght) == 0 && (t.tmask & e->irght) == 0 && (t.sgn & e->irght) == 0;
    if (lo) low.push_back(t);
    else if (hi) high.push_back(t);
    else mixed.push_back(t);
  }
  // A mixed term gathers from the block whose high half is x's with the
  // term's high flip bits changed; terms that share those bits read the same
  // target block, and together they fill its cache lines (on the 8x4 J1-J2
  // sector one term alone touches 1.62 bytes of lines per gathered byte, the
  // three that share a high site 1.05 together).  Consecutive, they reuse
  // those lines in L2 instead of re-reading them from HBM some tens of
  // microseconds later: 655 -> 494 GB fetched per H·v there.
  std::stable_sort(mixed.begin(), mixed.end(), [&](const GTerm &a, const GTerm &b) {
    return (a.flip & ~e->irght) < (b.flip & ~e->irght);
  });
  // The same for the low group (and every grand-canonical term): a 64-state
  // wave gathers from the region its target bits above the lowest six select,
  // so terms that agree there read the same lines back to back.
  std::stable_sort(low.begin(), low.end(), [](const GTerm &a, const GTerm &b) { return (a.flip >> 6) < (b.flip >> 6); });
  auto pad = [](std::vector<GTerm> &v) {
    GTerm z{};
    z.kind = P_EQ;
    z.tmask = 0;
    z.tval = 1;  // never active
    while (v.size() % HV_TERM_PAD) v.push_back(z);
  };
  e->r_low = (int)low.size();
  e->r_high = (int)high.size();
  e->r_mixed = (int)mixed.size();
  pad(low);
  pad(high);
  pad(mixed);
  e->n_low = (int)low.size();
  e->n_high = (int)high.size();
  e->n_mixed = (int)mixed.size();
  e->h_terms_dev.clear();
  e->h_terms_dev.insert(e->h_terms_dev.end(), low.begin(), low.end());
  e->h_terms_dev.insert(e->h_terms_dev.end(), high.begin(), high.end());
  e->h_terms_dev.insert(e->h_terms_dev.end(), mixed.begin(), mixed.end());
}

// Host-side preparation: sector sizes, split masks, gather terms (grouped
// low/high/mixed) and diagonal terms.  No device work.
// bits strictly between orbitals a and b
uint64_t hphi_between_mask(int a, int b) {
  const int lo = std::min(a, b), hi = std::max(a, b);
  return (hi - lo > 1) ? (((1ull << hi) - 1) & ~((1ull << (lo + 1)) - 1)) : 0ull;
}

// one hop c+_a c_b on state x as the reference's child functions apply it
// (child_GC_CisAjt, mltplyHubbardCore.c:431-456: a empty, b occupied, sign
// from the occupied bits strictly between; a == b is the occupation n_a,
// child_CisAis :288-298); sign parity restricted to `inside`
static bool apply_hop(uint64_t &x, int a, int b, uint64_t inside, int &parity) {
  if (a == b) return (x >> a) & 1;
  if (((x >> a) & 1) || !((x >> b) & 1)) return false;
  parity ^= __builtin_popcountll(x & hphi_between_mask(a, b) & inside) & 1;
  x ^= (1ull << a) | (1ull << b);
  return true;
}

// V c+_{a1} c_{b1} c+_{a2} c_{b2} in gather form: the reference applies the
// hop (a2 <- b2) to the source and then (a1 <- b1) to the result
// (CisAjtCkuAlv_element, CisAisCjtAku_element, CisAjtCkuAku_element,
// mltplyHubbardCore.c:718-830).  For every occupation pattern p of the involved
// orbitals with a non-zero result q, the target state x (pattern q) gathers
// from x ^ (p ^ q); the sign splits into a constant (the involved bits) and the
// parity of the target's other bits under the XOR of the two hops' between-masks
// (those bits do not change).
void hphi_compile_two_hops(int a1, int b1, int a2, int b2, double vre, double vim, std::vector<GTerm> &out) {
  const uint64_t B = (1ull << a1) | (1ull << b1) | (1ull << a2) | (1ull << b2);
  int orb[4], no = 0;
  for (int k = 0; k < 64; k++)
    if ((B >> k) & 1) orb[no++] = k;
  const uint64_t m1 = a1 != b1 ? hphi_between_mask(a1, b1) : 0, m2 = a2 != b2 ? hphi_between_mask(a2, b2) : 0;
  for (int p = 0; p < (1 << no); p++) {
    uint64_t s = 0;
    for (int k = 0; k < no; k++)
      if ((p >> k) & 1) s |= 1ull << orb[k];
    uint64_t x = s;
    int parity = 0;
    if (!apply_hop(x, a2, b2, B, parity) || !apply_hop(x, a1, b1, B, parity)) continue;
    GTerm g{};
    g.kind = P_EQ;
    g.tmask = B;
    g.tval = x;
    g.flip = s ^ x;
    g.sgn = (m1 ^ m2) & ~B;
    g.cre = parity ? -vre : vre;
    g.cim = parity ? -vim : vim;
    out.push_back(g);
  }
}

// V (|s1><s2|)_A (|s3><s4|)_B for spin-1/2 sites (bit value = sigma), applied
// to the source as the reference does: site B first, then site A
// (GC_general_int_spin and its four element functions,
// mltplySpinCore.c:427-583; canonical general_int_spin, mltplySpin.c:1020-1051,
// checks bit_A == s2, bit_B == s4 and flips both, which is the same operator for
// the Sz-conserving entries the canonical model accepts).  For every pattern p
// of the involved sites with a non-zero result q, the target (pattern q)
// gathers from target ^ (p ^ q); no sign.
void hphi_compile_spin_pair(int A, int s1, int s2, int B, int s3, int s4, double vre, double vim,
                            std::vector<GTerm> &out) {
  const uint64_t mA = 1ull << A, mB = 1ull << B, M = mA | mB;
  for (uint64_t p = 0; p < 4; p++) {
    const uint64_t src = ((p & 1) ? mA : 0) | ((p & 2) ? mB : 0);
    if ((p & 2) && A == B) continue;  // one site: patterns 0 and 1 only
    uint64_t x = src;
    if ((uint64_t)((x & mB) != 0) != (uint64_t)s4) continue;
    x = s3 ? (x | mB) : (x & ~mB);
    if ((uint64_t)((x & mA) != 0) != (uint64_t)s2) continue;
    x = s1 ? (x | mA) : (x & ~mA);
    GTerm g{};
    g.kind = P_EQ;
    g.tmask = M;
    g.tval = x;
    g.flip = src ^ x;
    g.cre = vre;
    g.cim = vim;
    out.push_back(g);
  }
}

int hphi_engine_prepare(const hphi_model_desc *d, hphi_engine *e) {
  e->model = d->model;
  e->nsite = d->nsite;
  e->nup = d->nup;
  e->ndown = d->ndown;
  const int N = d->nsite;
  if (N < 2) {
    return hphi_set_error("Nsite must be at least 2");
  }
  switch (d->model) {
    case HPHI_MODEL_SPIN:  // check.c:177, check.c:310-316, bitcalc.c:45-62
      if (N > 64 || d->nup < 0 || d->nup > N) { return hphi_set_error("bad Spin sector"); }
      e->nbits = N;
      e->dim = h_binom(N, d->nup);
      e->sdim = 1ull << (N / 2);
      e->size21 = (N % 2) ? e->sdim * 2 + 2 : e->sdim + 2;
      e->size22 = e->sdim + 2;
      break;
    case HPHI_MODEL_HUBBARD:  // check.c:114-118, check.c:301-307
      if (N > 32 || d->nup < 0 || d->nup > N || d->ndown < 0 || d->ndown > N) {
        return hphi_set_error("bad Hubbard sector");
      }
      e->nbits = 2 * N;
      e->dim = h_binom(N, d->nup) * h_binom(N, d->ndown);
      e->sdim = 1ull << N;
      e->size21 = e->sdim + 2;
      e->size22 = e->sdim + 2;
      break;
    case HPHI_MODEL_SPINGC:
      if (N > 40) { return hphi_set_error("SpinGC Nsite too large"); }
      e->nbits = N;
      e->dim = 1ull << N;
      e->sdim = 0;
      e->size21 = e->size22 = 0;
      break;
    case HPHI_MODEL_HUBBARDGC:  // check.c:120-124: idim_max = 4^Nsite
      if (N > 20) { return hphi_set_error("HubbardGC Nsite too large"); }
      e->nbits = 2 * N;
      e->dim = 1ull << (2 * N);
      e->sdim = 0;
      e->size21 = e->size22 = 0;
      break;
    default:
      return hphi_set_error("unsupported CalcModel %d (supported: 0 Hubbard, 1 Spin, 3 HubbardGC, 4 SpinGC)",
                            d->model);
  }
  e->nlow = (e->nbits + 1) / 2;
  e->ihfbit = 1ull << e->nlow;
  e->irght = e->ihfbit - 1;
  e->ilft = ((e->nbits >= 64) ? ~0ull : ((1ull << e->nbits) - 1)) ^ e->irght;
  if (e->dim == 0) { return hphi_set_error("empty sector"); }

  // ---- off-diagonal terms (gather form) ----
  const int gc = hphi_model_gc(d->model);
  const bool hubbard = hphi_model_hubbard(d->model);
  for (int t = 0; t < d->n_exchange; t++) {  // mltplySpinCore.c:273-345
    int i = d->exchange_ij[2 * t], j = d->exchange_ij[2 * t + 1];
    if (i < 0 || j < 0 || i >= N || j >= N || i == j) { return hphi_set_error("bad Exchange sites"); }
    if (hubbard) {
      // exchange_element (mltplyHubbardCore.c:472-525): coefficient -J, the
      // up/down pattern of sites i, j swapped, no sign; gather form: the
      // target has the swapped pattern
      const uint64_t iu = 1ull << (2 * i), id = 1ull << (2 * i + 1), ju = 1ull << (2 * j), jd = 1ull << (2 * j + 1);
      const uint64_t pat[2] = {iu | jd, id | ju};
      for (int q = 0; q < 2; q++) {
        GTerm g{};
        g.flip = iu | id | ju | jd;
        g.tmask = g.flip;
        g.tval = pat[q];
        g.kind = P_EQ;
        g.cre = -d->exchange_J[t];
        e->h_terms.push_back(g);
      }
      continue;
    }
    GTerm g{};
    g.flip = (1ull << i) | (1ull << j);
    g.tmask = g.flip;
    g.kind = P_DIFF;
    g.cre = d->exchange_J[t];
    e->h_terms.push_back(g);
  }
  for (int t = 0; t < d->n_pairlift; t++) {  // mltplySpinCore.c:352-382
    if (!gc) { return hphi_set_error("PairLift is active only in SpinGC"); }
    int i = d->pairlift_ij[2 * t], j = d->pairlift_ij[2 * t + 1];
    if (i < 0 || j < 0 || i >= N || j >= N || i == j) { return hphi_set_error("bad PairLift sites"); }
    GTerm g{};
    g.flip = (1ull << i) | (1ull << j);
    g.tmask = g.flip;
    g.kind = P_SAME;
    g.cre = d->pairlift_J[t];
    e->h_terms.push_back(g);
  }
  for (int t = 0; t < d->n_transfer; t++) {
    int i1 = d->transfer_idx[4 * t], s1 = d->transfer_idx[4 * t + 1];
    int i2 = d->transfer_idx[4 * t + 2], s2 = d->transfer_idx[4 * t + 3];
    double tre = d->transfer_t[2 * t], tim = d->transfer_t[2 * t + 1];
    if (i1 < 0 || i2 < 0 || i1 >= N || i2 >= N || s1 < 0 || s1 > 1 || s2 < 0 || s2 > 1) {
      return hphi_set_error("bad transfer indices");
    }
    GTerm g{};
    if (hubbard) {
      // CisAjt (mltplyHubbardCore.c:305-347): c+_a c_b with a = 2 i1 + s1,
      // b = 2 i2 + s2; target has bit a set and bit b clear; sign from the
      // bits strictly between a and b (general_hopp_GetInfo, :34-69)
      int a = 2 * i1 + s1, b = 2 * i2 + s2;
      if (a == b) continue;  // diagonal: EDChemi, not a transfer
      if (s1 != s2 && !gc) {  // HubbardGC takes spin-flip transfers (GC_CisAjt, mltplyHubbardCore.c:354-392)
        return hphi_set_error("spin-flip transfer (%d,%d,%d,%d) breaks the canonical Hubbard sector", i1, s1, i2, s2);
      }
      int lo = std::min(a, b), hi = std::max(a, b);
      g.flip = (1ull << a) | (1ull << b);
      g.tmask = g.flip;
      g.tval = 1ull << a;
      g.kind = P_EQ;
      g.sgn = (hi - lo > 1) ? (((1ull << hi) - 1) & ~((1ull << (lo + 1)) - 1)) : 0ull;
      g.cre = -tre;
      g.cim = -tim;
    } else if (gc) {
      // child_SpinGC_CisAit (mltplySpinCore.c:203-227): source has bit == s2,
      // target bit == s1; coefficient -t (mltplySpin.c:547)
      if (i1 != i2 || s1 == s2) continue;  // the reference ignores these (mltplySpin.c:553-556)
      g.flip = 1ull << i1;
      g.tmask = g.flip;
      g.tval = s1 ? g.flip : 0ull;
      g.kind = P_EQ;
      g.cre = -tre;
      g.cim = -tim;
    } else {
      continue;  // canonical spin ignores off-diagonal transfers (mltplySpin.c:192-293)
    }
    e->h_terms.push_back(g);
  }
  if (d->n_pairhop && !hubbard) return hphi_set_error("PairHop needs a Hubbard model");
  for (int t = 0; t < d->n_pairhop; t++) {
    // pairhopp_element (mltplyHubbardCore.c:533-572): a doublon moves from
    // site j to the empty site i, coefficient J, no sign
    int i = d->pairhop_ij[2 * t], j = d->pairhop_ij[2 * t + 1];
    if (i < 0 || j < 0 || i >= N || j >= N || i == j) return hphi_set_error("bad PairHop sites");
    GTerm g{};
    const uint64_t di = 3ull << (2 * i), dj = 3ull << (2 * j);
    g.flip = di | dj;
    g.tmask = g.flip;
    g.tval = di;
    g.kind = P_EQ;
    g.cre = d->pairhop_J[t];
    e->h_terms.push_back(g);
  }
  for (int t = 0; t < d->n_interall; t++) {
    const int *q = d->interall_idx + 8 * t;
    for (int u = 0; u < 8; u += 2)
      if (q[u] < 0 || q[u] >= N || q[u + 1] < 0 || q[u + 1] > 1) return hphi_set_error("bad InterAll indices");
    if (!hubbard) {
      // spin entries are (i s1 i s2 k s3 k s4) (CheckFormatForSpinInt, readdef.c:2637-2650)
      if (q[0] != q[2] || q[4] != q[6]) return hphi_set_error("spin InterAll needs i == j and k == l");
      // canonical: only two opposite flips on distinct sites keep the sector;
      // the reference's canonical kernel would index outside it otherwise
      if (!gc && (q[0] == q[4] || q[1] == q[3] || q[5] == q[7] || q[3] == q[7]))
        return hphi_set_error("InterAll (%d %d %d %d %d %d %d %d) does not conserve Sz in the canonical Spin model",
                              q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7]);
      hphi_compile_spin_pair(q[0], q[1], q[3], q[4], q[5], q[7], d->interall_V[2 * t], d->interall_V[2 * t + 1],
                        e->h_terms);
      continue;
    }
    hphi_compile_two_hops(2 * q[0] + q[1], 2 * q[2] + q[3], 2 * q[4] + q[5], 2 * q[6] + q[7], d->interall_V[2 * t],
                     d->interall_V[2 * t + 1], e->h_terms);
  }
  classify_and_upload_terms(e);

  // ---- diagonal terms, in diagonalcalc order ----
  for (int t = 0; t < d->n_intra; t++) {
    if (!hubbard) break;
    int s = d->intra_site[t];
    DTerm x{};
    x.kind = D_INTRA;
    x.m1 = (1ull << (2 * s)) | (1ull << (2 * s + 1));
    x.V = d->intra_U[t];
    e->h_dterms.push_back(x);
  }
  for (int t = 0; t < d->n_chemi; t++) {
    int s = d->chemi_site[t], sp = d->chemi_spin[t];
    DTerm x{};
    x.V = -d->chemi_mu[t];
    if (hubbard) {
      x.kind = D_CHEMI_H;
      x.m1 = 1ull << (2 * s + sp);
    } else {
      x.kind = D_CHEMI_S;
      x.m1 = 1ull << s;
      x.m2 = (uint64_t)(1 - sp);
    }
    e->h_dterms.push_back(x);
  }
  for (int t = 0; t < d->n_inter; t++) {
    int s1 = d->inter_ij[2 * t], s2 = d->inter_ij[2 * t + 1];
    DTerm x{};
    x.V = d->inter_V[t];
    if (hubbard) {
      x.kind = D_INTER_H;
      x.m1 = 1ull << (2 * s1);
      x.m2 = 1ull << (2 * s1 + 1);
      x.m3 = 1ull << (2 * s2);
      x.m4 = 1ull << (2 * s2 + 1);
    } else {
      x.kind = D_INTER_S;
    }
    e->h_dterms.push_back(x);
  }
  for (int t = 0; t < d->n_hund; t++) {
    int s1 = d->hund_ij[2 * t], s2 = d->hund_ij[2 * t + 1];
    DTerm x{};
    x.V = -d->hund_J[t];
    if (hubbard) {
      x.kind = D_HUND_H;
      x.m1 = 1ull << (2 * s1);
      x.m2 = 1ull << (2 * s1 + 1);
      x.m3 = 1ull << (2 * s2);
      x.m4 = 1ull << (2 * s2 + 1);
    } else {
      x.kind = D_HUND_S;
      x.m1 = (1ull << s1) | (1ull << s2);
    }
    e->h_dterms.push_back(x);
  }
  for (int t = 0; t < d->n_interall_diag; t++) {  // diagonalcalc.c:168-181
    const int *q = d->interall_diag_idx + 4 * t;
    if (q[0] < 0 || q[0] >= N || q[2] < 0 || q[2] >= N || q[1] < 0 || q[1] > 1 || q[3] < 0 || q[3] > 1)
      return hphi_set_error("bad InterAll diagonal indices");
    DTerm x{};
    if (hubbard) {
      x.kind = D_NN;
      x.m1 = 1ull << (2 * q[0] + q[1]);
      x.m2 = 1ull << (2 * q[2] + q[3]);
    } else {  // child_Spin_CisAis / child_SpinGC_CisAis products (diagonalcalc.c:1380-1420)
      x.kind = D_SS;
      x.m1 = 1ull << q[0];
      x.m2 = 1ull << q[2];
      x.m3 = (uint64_t)(1 - q[1]);
      x.m4 = (uint64_t)(1 - q[3]);
    }
    x.V = d->interall_diag_V[t];
    e->h_dterms.push_back(x);
  }

  return 0;
}

extern "C" int hphi_model_nnz(const hphi_model_desc *d, uint64_t *dim, uint64_t *nnz) {
  if (!d) return hphi_set_error("hphi_model_nnz: NULL argument");
  hphi_engine e;
  if (hphi_engine_prepare(d, &e)) return -1;
  if (hphi_model_gc(e.model)) return hphi_set_error("hphi_model_nnz: canonical Spin and Hubbard models only");
  if (dim) *dim = e.dim;
  if (nnz) *nnz = hphi_term_nnz(&e);
  return 0;
}

extern "C" int hphi_engine_create(const hphi_model_desc *d, hphi_engine **out) {
  if (!d || !out) return hphi_set_error("hphi_engine_create: NULL argument");
  *out = nullptr;
  if (hphi_ensure_binom_table()) return -1;
  hphi_engine *e = new hphi_engine();
  if (hphi_engine_prepare(d, e)) {
    delete e;
    return -1;
  }
  if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
    delete e;
    return hphi_set_error("hipStreamCreate failed");
  }
  if (hipGetDevice(&e->device) != hipSuccess) e->device = 0;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, e->device) == hipSuccess) e->num_cu = prop.multiProcessorCount;
  size_t tb = std::max<size_t>(e->h_terms_dev.size(), 1) * sizeof(GTerm);
  size_t db = std::max<size_t>(e->h_dterms.size(), 1) * sizeof(DTerm);
  if (hipMalloc(&e->d_terms, tb) != hipSuccess || hipMalloc(&e->d_dterms, db) != hipSuccess) {
    hphi_engine_destroy(e);
    return hphi_set_error("hipMalloc(terms) failed");
  }
  if (!e->h_terms_dev.empty())
    (void)hipMemcpy(e->d_terms, e->h_terms_dev.data(), e->h_terms_dev.size() * sizeof(GTerm), hipMemcpyHostToDevice);
  if (!e->h_dterms.empty())
    (void)hipMemcpy(e->d_dterms, e->h_dterms.data(), e->h_dterms.size() * sizeof(DTerm), hipMemcpyHostToDevice);
  e->bytes += tb + db;
  hphi_canon_config(e);
  if (e->c2_crank_ok) {
    const bool hub = e->model == HPHI_MODEL_HUBBARD;
    const size_t n = hub ? HRANK_N : CRANK_N;
    std::vector<uint16_t> t(n);
    if (hub) hphi_hrank_tables(t.data());
    else hphi_crank_tables(t.data());
    if (hipMalloc(&e->d_crank, sizeof(uint16_t) * n) != hipSuccess ||
        hipMemcpy(e->d_crank, t.data(), sizeof(uint16_t) * n, hipMemcpyHostToDevice) != hipSuccess) {
      hphi_engine_destroy(e);
      return hphi_set_error("hipMalloc(rank tables) failed");
    }
    e->bytes += sizeof(uint16_t) * n;
  }
  if (e->c2_crank_ok) {  // rank tables (5 / 13 KB of LDS): many resident workgroups
    e->c2_crank = 1;
    e->c2_tpb = 256;
    e->c2_wgs_per_cu = 48;  // 8x4 J1-J2: 107.3 ms vs 138.9 ms with the LDS table, grids of 40-64 workgroups
                            // per CU within 1 %; 4x4 Hubbard: 20.1 ms vs 27.3 ms
    if (e->model == HPHI_MODEL_HUBBARD) e->c2_wgs_per_cu = 160;  // 4x4 Hubbard: 14.42 ms against 14.98 at 48
    e->c2_unroll = 8;
    e->xcd_map = e->model == HPHI_MODEL_SPIN ? 1 : 0;
  } else if (e->c2_lds21) {  // list_2_1 in LDS: one 1024-thread workgroup per CU
    e->c2_tpb = 1024;
    e->c2_wgs_per_cu = 1;
    e->c2_unroll = 16;  // 16 gathers in flight per lane: 27.3 ms vs 28.7 ms (4x4 Hubbard)
    e->xcd_map = 0;
  } else {
    e->c2_tpb = 512;
    e->c2_wgs_per_cu = 8;
  }
  if (hphi_engine_alloc_scratch(e) || hphi_noalloc_setup(e->device)) {
    hphi_engine_destroy(e);
    return -1;
  }
  *out = e;
  return 0;
}

extern "C" void hphi_engine_destroy(hphi_engine *e) {
  if (!e) return;
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  (void)hipFree(e->d_list1);
  (void)hipFree(e->d_l21);
  (void)hipFree(e->d_l22);
  (void)hipFree(e->d_crank);
  (void)hipFree(e->d_hb);
  (void)hipFree(e->d_blkpos);
  (void)hipFree(e->d_v1t);
  (void)hipFree(e->d_tofs);
  (void)hipFree(e->d_jb);
  (void)hipFree(e->d_diag);
  (void)hipFree(e->d_terms);
  (void)hipFree(e->d_dterms);
  (void)hipFree(e->d_partial);
  (void)hipFree(e->d_prdct);
  (void)hipFree(e->d_coef);
  (void)hipFree(e->d_stage0);
  (void)hipFree(e->d_stage1);
  for (int t = 0; t < e->n_host; t++) {
    if (e->h_stream[t]) (void)hipStreamSynchronize(e->h_stream[t]);
    for (int b = 0; b < 2; b++) {
      if (e->h_bounce[t][b]) (void)hipHostFree(e->h_bounce[t][b]);
      if (e->h_ev[t][b]) (void)hipEventDestroy(e->h_ev[t][b]);
    }
    if (e->h_stream[t]) (void)hipStreamDestroy(e->h_stream[t]);
  }
  if (e->h_prdct) (void)hipHostFree(e->h_prdct);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
}

extern "C" int hphi_engine_info(const hphi_engine *e, uint64_t *dim, uint64_t *s21, uint64_t *s22,
                                uint64_t *irght, uint64_t *ilft, uint64_t *ihfbit) {
  if (!e) return hphi_set_error("NULL engine");
  if (dim) *dim = e->dim;
  if (s21) *s21 = e->size21;
  if (s22) *s22 = e->size22;
  if (irght) *irght = e->irght;
  if (ilft) *ilft = e->ilft;
  if (ihfbit) *ihfbit = e->ihfbit;
  return 0;
}

extern "C" uint64_t hphi_engine_device_bytes(const hphi_engine *e) { return e ? e->bytes : 0; }

extern "C" int hphi_engine_tune(hphi_engine *e, const char *knob, int value) {
  if (!e || !knob) return hphi_set_error("hphi_engine_tune: NULL argument");
  std::string k(knob);
  if (k == "tpb") {
    if (value != 256 && value != 512 && value != 1024) return hphi_set_error("tpb must be 256, 512 or 1024");
    e->c2_tpb = value;
  } else if (k == "wgs_per_cu") {
    if (value < 1 || value > 256) return hphi_set_error("wgs_per_cu out of range");  // grids cap at HV_MAX_GRID
    e->c2_wgs_per_cu = value;
  } else if (k == "lds21") {
    if (value) {
      int saved = e->c2_lds21;
      hphi_canon_config(e);
      if (!e->c2_lds21) {
        e->c2_lds21 = saved;
        return hphi_set_error("list_2_1 does not fit LDS for this model");
      }
    } else {
      e->c2_lds21 = 0;
    }
  } else if (k == "c2_crank") {
    if (value && !e->c2_crank_ok)
      return hphi_set_error("rank tables need a canonical Spin or Hubbard model with nlow <= 16");
    e->c2_crank = value ? 1 : 0;
  } else if (k == "c2_unroll") {
    if (value != 8 && value != 16) return hphi_set_error("c2_unroll must be 8 or 16");
    e->c2_unroll = value;
  } else if (k == "hi_split") {
    if (value < -1 || value > 1) return hphi_set_error("hi_split must be -1, 0 or 1");
    e->hi_split = value;
  } else if (k == "hi_p1_wgs_per_cu") {
    if (value < 1 || value > 256) return hphi_set_error("hi_p1_wgs_per_cu out of range");
    e->hi_p1_wgs_per_cu = value;
  } else if (k == "hi_fuse_copy") {
    e->hi_fuse_copy = value ? 1 : 0;
  } else if (k == "hi_wgs_per_cu") {
    if (value < 1 || value > 64) return hphi_set_error("hi_wgs_per_cu out of range");
    e->hi_wgs_per_cu = value;
  } else if (k == "xcd_map") {
    e->xcd_map = value ? 1 : 0;
    e->xcd_tuned = 1;
  } else if (k == "use_c2") {
    e->use_c2 = value ? 1 : 0;
  } else if (k == "grid_per_cu") {
    if (value < 1 || value > 64) return hphi_set_error("grid_per_cu out of range");
    e->grid_per_cu = value;
  } else if (k == "use_blk") {
    e->use_blk = value ? 1 : 0;
  } else if (k == "lb_mfma") {
    e->lb_mfma = value ? 1 : 0;
  } else if (k == "blk_tpb") {
    if (value != 0 && value != 256 && value != 1024) return hphi_set_error("blk_tpb must be 0, 256 or 1024");
    e->bk_tpb = value;
  } else if (k == "blk_wgs_per_cu") {
    if (value < 0 || value > 64) return hphi_set_error("blk_wgs_per_cu out of range");
    e->bk_wgs_per_cu = value;
  } else if (k.rfind("blk_cpl", 0) == 0) {
    int kk = atoi(k.c_str() + 7), slot = kk == 2 ? 0 : kk == 4 ? 1 : kk == 8 ? 2 : kk == 16 ? 3 : -1;
    if (slot < 0) return hphi_set_error("unknown tuning knob %s", knob);
    if (value != 0 && !hphi_blk_cpl_ok(kk, value))
      return hphi_set_error("%s: %d columns per lane not available", knob, value);
    e->bk_cpl[slot] = value;
  } else {
    return hphi_set_error("unknown tuning knob %s", knob);
  }
  return 0;
}

extern "C" int hphi_engine_get_tune(const hphi_engine *e, const char *knob, int *value) {
  if (!e || !knob || !value) return hphi_set_error("hphi_engine_get_tune: NULL argument");
  std::string k(knob);
  if (k == "tpb") *value = e->c2_tpb;
  else if (k == "wgs_per_cu") *value = e->c2_wgs_per_cu;
  else if (k == "lds21") *value = e->c2_lds21;
  else if (k == "c2_crank") *value = e->c2_crank;
  else if (k == "c2_unroll") *value = e->c2_unroll;
  else if (k == "xcd_map") *value = e->xcd_map;
  else if (k == "hi_split") *value = hphi_hi_wanted(e);
  else if (k == "hi_wgs_per_cu") *value = e->hi_wgs_per_cu;
  else if (k == "hi_fuse_copy") *value = e->hi_fuse_copy;
  else if (k == "hi_p1_wgs_per_cu") *value = e->hi_p1_wgs_per_cu;
  else if (k == "use_c2") *value = e->use_c2 && e->c2_ok;
  else if (k == "grid_per_cu") *value = e->grid_per_cu;
  else if (k == "use_blk") *value = e->use_blk && hphi_blk_ok(e);
  else if (k == "lb_mfma") *value = e->lb_mfma;
  else if (k == "blk_tpb" || k == "blk_wgs_per_cu" || k.rfind("blk_cpl", 0) == 0) {
    int kk = k.rfind("blk_cpl", 0) == 0 ? atoi(k.c_str() + 7) : 16;
    if (kk != 2 && kk != 4 && kk != 8 && kk != 16) return hphi_set_error("unknown tuning knob %s", knob);
    int cpl, tpb, wpc;
    hphi_blk_shape(e, kk, &cpl, &tpb, &wpc);
    *value = k == "blk_tpb" ? tpb : k == "blk_wgs_per_cu" ? wpc : cpl;
  } else return hphi_set_error("unknown tuning knob %s", knob);
  return 0;
}

SzParams hphi_sz_params(const hphi_engine *e) {
  SzParams P{};
  P.hubbard = e->model == HPHI_MODEL_HUBBARD;
  P.nlow = e->nlow;
  P.nhigh = e->nbits - e->nlow;
  P.nup = e->nup;
  P.ndn = e->ndown;
  P.sdim = e->sdim;
  P.ihfbit = e->ihfbit;
  P.size21 = e->size21;
  P.dim = e->dim;
  // Hubbard: up positions are even absolute bits; the low half is bits [0, nlow)
  P.nlu = (e->nlow + 1) / 2;
  P.nld = e->nlow / 2;
  P.nhu = e->nsite - P.nlu;
  P.nhd = e->nsite - P.nld;
  return P;
}

// enqueue the basis-table kernels (block prefix + list_2_2, list_2_1, list_1) on `st`
int hphi_launch_tables(const SzParams &P, uint64_t sdim, uint64_t size21, uint64_t *jb, uint32_t *l21,
                       uint64_t *l22, uint64_t *list1, hipStream_t st) {
  hipLaunchKernelGGL(k_block_prefix, dim3(grid_for(sdim + 1, 256)), dim3(256), 0, st, P, jb, l22);
  hipLaunchKernelGGL(k_list21, dim3(grid_for(size21, 256)), dim3(256), 0, st, P, l21);
  int g = (int)std::min<uint64_t>(sdim, 1u << 20);
  hipLaunchKernelGGL(k_list1, dim3(g), dim3(256), 0, st, P, (const uint64_t *)jb, list1);
  HIPCHK(hipGetLastError());
  return 0;
}

int hphi_launch_diag(uint64_t dim, const uint64_t *list1, int gc, const DTerm *dterms, int nd, double *diag,
                     hipStream_t st, int stride, int col) {
  hipLaunchKernelGGL(k_diag, dim3(grid_for(dim, 256)), dim3(256), 0, st, dim, list1, gc, dterms, nd, diag, stride,
                     col);
  HIPCHK(hipGetLastError());
  return 0;
}

int hphi_build_tables(hphi_engine *e) {
  if (e->tables_ready) return 0;
  if (hphi_model_gc(e->model)) {
    e->tables_ready = true;
    return 0;
  }
  SzParams P = hphi_sz_params(e);
  HIPCHK(hipMalloc(&e->d_list1, sizeof(uint64_t) * e->dim));
  HIPCHK(hipMalloc(&e->d_l21, sizeof(uint32_t) * e->size21));
  HIPCHK(hipMalloc(&e->d_l22, sizeof(uint64_t) * e->size22));
  HIPCHK(hipMalloc(&e->d_jb, sizeof(uint64_t) * (e->sdim + 1)));
  e->bytes += sizeof(uint64_t) * (e->dim + e->size22 + e->sdim + 1) + sizeof(uint32_t) * e->size21;
  if (hphi_launch_tables(P, e->sdim, e->size21, e->d_jb, e->d_l21, e->d_l22, e->d_list1, e->stream)) return -1;
  uint64_t total = 0;
  HIPCHK(hipMemcpyAsync(&total, e->d_jb + e->sdim, sizeof total, hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  if (total != e->dim)  // sz.c:810-820
    return hphi_set_error("basis count %llu differs from idim_max %llu", (unsigned long long)total,
                          (unsigned long long)e->dim);
  e->tables_ready = true;
  // plan of the high-half pass now, outside any H·v launch (it copies list_2_2 to the host)
  if (hphi_hi_split_on(e) < 0) return -1;
  return 0;
}

extern "C" int hphi_sz(hphi_engine *e, uint64_t *list_1, uint64_t *list_2_1, uint64_t *list_2_2) {
  if (!e) return hphi_set_error("NULL engine");
  if (hphi_build_tables(e)) return -1;
  if (hphi_model_gc(e->model)) return 0;
  if (list_1) {
    list_1[0] = 0;
    HIPCHK(hipMemcpy(list_1 + 1, e->d_list1, sizeof(uint64_t) * e->dim, hipMemcpyDeviceToHost));
  }
  if (list_2_1) {
    std::vector<uint32_t> t(e->size21);
    HIPCHK(hipMemcpy(t.data(), e->d_l21, sizeof(uint32_t) * e->size21, hipMemcpyDeviceToHost));
    for (uint64_t i = 0; i < e->size21; i++) list_2_1[i] = t[i];
  }
  if (list_2_2)
    HIPCHK(hipMemcpy(list_2_2, e->d_l22, sizeof(uint64_t) * e->size22, hipMemcpyDeviceToHost));
  return 0;
}

int hphi_build_diag(hphi_engine *e) {
  if (e->diag_ready) return 0;
  if (hphi_build_tables(e)) return -1;
  HIPCHK(hipMalloc(&e->d_diag, sizeof(double) * e->dim));
  e->bytes += sizeof(double) * e->dim;
  if (hphi_launch_diag(e->dim, e->d_list1, hphi_model_gc(e->model) ? 1 : 0, e->d_dterms,
                       (int)e->h_dterms.size(), e->d_diag, e->stream))
    return -1;
  HIPCHK(hipStreamSynchronize(e->stream));
  e->diag_ready = true;
  return 0;
}

extern "C" int hphi_diagonalcalc(hphi_engine *e, double *list_Diagonal) {
  if (!e) return hphi_set_error("NULL engine");
  if (hphi_build_diag(e)) return -1;
  if (list_Diagonal) {
    list_Diagonal[0] = 0.0;
    HIPCHK(hipMemcpy(list_Diagonal + 1, e->d_diag, sizeof(double) * e->dim, hipMemcpyDeviceToHost));
  }
  return 0;
}

HvArgs hphi_hv_args(hphi_engine *e, int k, double *v0, const double *v1) {
  HvArgs A{};
  A.dim = e->dim;
  A.list1 = e->d_list1;
  A.l21 = e->d_l21;
  A.l22 = e->d_l22;
  A.diag = e->d_diag;
  A.terms = e->d_terms;
  A.n_low = e->n_low;
  A.n_high = e->n_high;
  A.n_mixed = e->n_mixed;
  A.c0 = 1.0;
  A.c1 = 1.0;
  A.e_low = e->n_low;
  A.e_high = e->n_high;
  A.e_mixed = e->n_mixed;
  A.irght = e->irght;
  A.nlow = e->nlow;
  A.v0 = v0;
  A.v1 = v1;
  A.partial = (double2 *)e->d_partial;
  A.xcd_map = e->xcd_map;
  A.crank = e->d_crank;
  A.size21 = e->size21;
  (void)k;
  return A;
}

int hphi_hv_grid(const hphi_engine *e, int k) {
  uint64_t spt = HV_TPB / k;
  uint64_t ntiles = (e->dim + spt - 1) / spt;
  uint64_t g = std::min<uint64_t>(ntiles, (uint64_t)e->grid_per_cu * std::max(e->num_cu, 1));
  g = std::max<uint64_t>(g, 8);
  g = (g + 7) / 8 * 8;
  if (g > HV_MAX_GRID) g = HV_MAX_GRID;
  return (int)g;
}

// launches H·v and the final reduction on e->stream; result in e->d_prdct[0..k)
static int canon_grid(const hphi_engine *e, int wgs_per_cu = 0) {
  uint64_t ntiles = (e->dim + e->c2_tpb - 1) / e->c2_tpb;
  const uint64_t w = (uint64_t)(wgs_per_cu > 0 ? wgs_per_cu : e->c2_wgs_per_cu);
  uint64_t g = std::min<uint64_t>(ntiles, w * std::max(e->num_cu, 1));
  g = std::max<uint64_t>(g, 8);
  g = (g + 7) / 8 * 8;
  if (g > HV_MAX_GRID) g = HV_MAX_GRID;
  return (int)g;
}

int hphi_hv_launch(hphi_engine *e, int k, double *v0, const double *v1) {
  return hphi_hv_launch_ex(e, k, v0, v1, 0);
}

static int hv_launch_args(hphi_engine *e, int k, HvArgs &A, double2 *prdct_out);

int hphi_hv_launch_ex(hphi_engine *e, int k, double *v0, const double *v1, int overwrite) {
  HvArgs A = hphi_hv_args(e, k, v0, v1);
  A.overwrite = overwrite;
  return hv_launch_args(e, k, A, (double2 *)e->d_prdct);
}

// which kernel a launch for k columns uses, and its workgroups (each writes k
// partial sums of prdct)
enum { HV_PATH_CANON, HV_PATH_BLK, HV_PATH_GENERIC };
static int hv_path(const hphi_engine *e, int k, int *grid) {
  if (k == 1 && e->c2_ok && e->use_c2) {
    *grid = canon_grid(e);
    return HV_PATH_CANON;
  }
  if (k > 1 && e->use_blk && hphi_blk_ok(e)) {
    *grid = hphi_blk_grid(e, k);
    return HV_PATH_BLK;
  }
  *grid = hphi_hv_grid(e, k);
  return HV_PATH_GENERIC;
}

// launch of prepared arguments: the canonical single-vector kernel, the block
// kernel, or the generic kernel; then the partials in A.partial are reduced
// into prdct_out[0..k)
static int hv_launch_args(hphi_engine *e, int k, HvArgs &A, double2 *prdct_out) {
  int g = 0;
  const int path = hv_path(e, k, &g);
  if (path == HV_PATH_CANON) {
    // canonical Spin: the high-half terms in a second pass over a tile-major
    // copy of v1 (only with the engine's own partial-sum buffer, which holds
    // both grids)
    const int split = A.partial == (double2 *)e->d_partial ? hphi_hi_split_on(e) : 0;
    if (split < 0) return -1;
    if (split && e->hi_fuse_copy) {  // the first pass also writes the tile-major copy of v1
      A.v1t = e->d_v1t;
      A.tofs = e->d_tofs;
    }
    if (split) g = canon_grid(e, e->hi_p1_wgs_per_cu);  // the first pass without the high terms
    if (hphi_hv_launch_canon(e, A, g, split)) return -1;
    if (split) {
      int g2 = 0;
      if (hphi_hv_launch_high(e, A, A.partial + g, &g2)) return -1;
      g += g2;
    }
  } else if (path == HV_PATH_BLK) {
    int gb = 0;
    if (hphi_hv_launch_blk(e, A, k, &gb)) return -1;
    g = gb;
  } else {
    bool canon = !hphi_model_gc(e->model);
    static_assert(HV_TERM_PAD % HV_UNROLL == 0, "a batch must not cross a term group");
    A.e_low = std::min(A.n_low, (e->r_low + HV_UNROLL - 1) / HV_UNROLL * HV_UNROLL);
    A.e_high = std::min(A.n_high, (e->r_high + HV_UNROLL - 1) / HV_UNROLL * HV_UNROLL);
    A.e_mixed = std::min(A.n_mixed, (e->r_mixed + HV_UNROLL - 1) / HV_UNROLL * HV_UNROLL);
#define LAUNCH(KK)                                                                                   \
  if (canon) hipLaunchKernelGGL((k_hv<KK, true>), dim3(g), dim3(HV_TPB), 0, e->stream, A);          \
  else hipLaunchKernelGGL((k_hv<KK, false>), dim3(g), dim3(HV_TPB), 0, e->stream, A);
    switch (k) {
      case 1: LAUNCH(1); break;
      case 2: LAUNCH(2); break;
      case 4: LAUNCH(4); break;
      case 8: LAUNCH(8); break;
      case 16: LAUNCH(16); break;
      default: return hphi_set_error("block size k=%d not in {1,2,4,8,16}", k);
    }
#undef LAUNCH
  }
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, e->stream, (const double2 *)A.partial, g, k,
                     prdct_out);
  HIPCHK(hipGetLastError());
  return 0;
}

static int ensure_ready(hphi_engine *e) {
  if (hphi_build_tables(e)) return -1;
  if (hphi_build_diag(e)) return -1;
  return 0;
}

static int fetch_prdct_from(hphi_engine *e, const double2 *d_src, int k, double *prdct) {
  HIPCHK(hipMemcpyAsync(e->h_prdct, d_src, sizeof(double2) * k, hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  if (prdct) memcpy(prdct, e->h_prdct, sizeof(double2) * k);
  return 0;
}

int hphi_fetch_prdct(hphi_engine *e, int k, double *prdct) {
  return fetch_prdct_from(e, (const double2 *)e->d_prdct, k, prdct);
}

extern "C" int hphi_mltply_dev(hphi_engine *e, double *v0, const double *v1, double *prdct) {
  if (!e || !v0 || !v1) return hphi_set_error("hphi_mltply_dev: NULL argument");
  if (ensure_ready(e)) return -1;
  if (hphi_hv_launch(e, 1, v0, v1)) return -1;
  return hphi_fetch_prdct(e, 1, prdct);
}

extern "C" int hphi_mltply_dev_scaled(hphi_engine *e, double *v0, const double *v1, double c0, double c1,
                                      int overwrite, double *prdct) {
  if (!e || !v0 || !v1) return hphi_set_error("hphi_mltply_dev_scaled: NULL argument");
  if (ensure_ready(e)) return -1;
  HvArgs A = hphi_hv_args(e, 1, v0, v1);
  A.overwrite = overwrite ? 1 : 0;
  A.c0 = c0;
  A.c1 = c1;
  if (hv_launch_args(e, 1, A, (double2 *)e->d_prdct)) return -1;
  return hphi_fetch_prdct(e, 1, prdct);
}

extern "C" int hphi_mltply_block_dev(hphi_engine *e, int k, double *V0, const double *V1, double *prdct) {
  if (!e || !V0 || !V1) return hphi_set_error("hphi_mltply_block_dev: NULL argument");
  if (ensure_ready(e)) return -1;
  if (hphi_hv_launch(e, k, V0, V1)) return -1;
  return hphi_fetch_prdct(e, k, prdct);
}

// ------------------------------------------------- allocation-free H·v ----
// hphi_mltply_block_dev_noalloc keeps the per-workgroup partial sums of prdct
// and their reduction in this fixed array rather than in the engine's
// buffers.  Every device holds its own copy; host code never passes the
// array's name to a kernel (on the host it is only a stand-in), but the
// device address hipGetSymbolAddress returned for that device, recorded
// once per device when an engine is created there.
#define HV_NOALLOC_PARTIALS (1 << 18)  // k x workgroups of one launch (4 MiB)
static __device__ __align__(256) double2 scratch[HV_NOALLOC_PARTIALS + 16];
#define HV_MAX_DEVICES 64
static double2 *g_scratch_dev[HV_MAX_DEVICES];
static std::mutex g_scratch_mu;

int hphi_noalloc_setup(int device) {
  if (device < 0 || device >= HV_MAX_DEVICES) return hphi_set_error("device %d out of range", device);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_dev[device]) return 0;
  void *ptr = nullptr;
  hipError_t err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (err != hipSuccess || !ptr)
    return hphi_set_error("hipGetSymbolAddress(scratch) failed on device %d: %s", device, hipGetErrorString(err));
  g_scratch_dev[device] = (double2 *)ptr;
  return 0;
}

extern "C" int hphi_mltply_block_dev_noalloc(hphi_engine *e, int k, double *V0, const double *V1, double *prdct) {
  if (!e || !V0 || !V1) return hphi_set_error("hphi_mltply_block_dev_noalloc: NULL argument");
  if (k != 1 && k != 2 && k != 4 && k != 8 && k != 16)
    return hphi_set_error("hphi_mltply_block_dev_noalloc: block size k=%d not in {1,2,4,8,16}", k);
  if (!e->tables_ready || !e->diag_ready)
    return hphi_set_error("hphi_mltply_block_dev_noalloc: basis tables and diagonal are not built; call hphi_sz "
                          "and hphi_diagonalcalc first (this entry point allocates nothing)");
  int cur = -1;
  HIPCHK(hipGetDevice(&cur));
  if (cur != e->device)
    return hphi_set_error("hphi_mltply_block_dev_noalloc: the engine lives on device %d, device %d is current",
                          e->device, cur);
  double2 *scr = g_scratch_dev[e->device];
  if (!scr) return hphi_set_error("hphi_mltply_block_dev_noalloc: no scratch address for device %d", e->device);
  int g = 0;
  (void)hv_path(e, k, &g);
  if ((uint64_t)g * (uint64_t)k > HV_NOALLOC_PARTIALS)
    return hphi_set_error("hphi_mltply_block_dev_noalloc: %d workgroups x k=%d partial sums exceed the fixed "
                          "scratch of %d; use hphi_mltply_block_dev or a smaller launch shape",
                          g, k, HV_NOALLOC_PARTIALS);
  HvArgs A = hphi_hv_args(e, k, V0, V1);
  A.partial = scr;  // every workgroup of the launch writes its slots, nothing needs clearing
  if (hv_launch_args(e, k, A, scr + HV_NOALLOC_PARTIALS)) return -1;
  return fetch_prdct_from(e, scr + HV_NOALLOC_PARTIALS, k, prdct);
}

static bool is_device_ptr(const void *p) {
  hipPointerAttribute_t at;
  hipError_t err = hipPointerGetAttributes(&at, p);
  if (err != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return at.type == hipMemoryTypeDevice;
}

// Host-pointer calls move v1 in and H v1 out through pinned bounce buffers:
// n_host worker threads each own a contiguous share of the vector, a stream
// and two HOST_CHUNK buffers, so CPU copies (pageable <-> pinned) of one chunk
// overlap the DMA of the other and the threads' DMAs run side by side.  v0
// never crosses to the device: the kernel writes H v1 (overwrite) and the
// threads add it to the host v0 as it comes back, which is the kernel's own
// v0 + H v1 rounding (one addition).
#define HOST_CHUNK ((size_t)16 << 20)
#define HOST_THREADS 16

static int host_setup(hphi_engine *e) {
  if (e->n_host) return 0;
  unsigned hw = std::thread::hardware_concurrency();
  int n = (int)std::min<unsigned>(HOST_THREADS, std::max(1u, hw));
  for (int t = 0; t < n; t++) {
    if (hipStreamCreateWithFlags(&e->h_stream[t], hipStreamNonBlocking) != hipSuccess) break;
    bool ok = true;
    for (int b = 0; b < 2 && ok; b++)
      ok = hipHostMalloc((void **)&e->h_bounce[t][b], HOST_CHUNK, hipHostMallocDefault) == hipSuccess &&
           hipEventCreateWithFlags(&e->h_ev[t][b], hipEventDisableTiming) == hipSuccess;
    e->n_host = t + 1;  // destroy frees whatever of this slot exists
    if (!ok) return hphi_set_error("pinned bounce buffers for host-pointer calls: allocation failed");
  }
  if (!e->n_host) return hphi_set_error("streams for host-pointer calls: creation failed");
  return 0;
}

// thread t's share [lo, hi) of n bytes, in whole complex128 elements
static void share(size_t n, int t, int nt, size_t *lo, size_t *hi) {
  const size_t ne = n / 16;
  *lo = ne * (size_t)t / (size_t)nt * 16;
  *hi = ne * (size_t)(t + 1) / (size_t)nt * 16;
}

extern "C" int hphi_mltply(hphi_engine *e, double *v0, const double *v1, double *prdct) {
  if (!e || !v0 || !v1) return hphi_set_error("hphi_mltply: NULL argument");
  if (ensure_ready(e)) return -1;
  bool d0 = is_device_ptr(v0), d1 = is_device_ptr(v1);
  if (d0 && d1) return hphi_mltply_dev(e, v0 + 2, v1 + 2, prdct);
  if (d0 != d1) return hphi_set_error("hphi_mltply: v0 and v1 must both be host or both be device memory");
  // compatibility path: the reference's 1-indexed host arrays
  const size_t bytes = sizeof(double2) * e->dim;
  if (!e->d_stage0) {
    HIPCHK(hipMalloc(&e->d_stage0, bytes));
    HIPCHK(hipMalloc(&e->d_stage1, bytes));
    e->bytes += 2 * bytes;
  }
  if (host_setup(e)) return -1;
  const int nt = e->n_host;
  char *dv1 = (char *)e->d_stage1, *dout = (char *)e->d_stage0;
  const char *hv1 = (const char *)(v1 + 2);
  double *hv0 = v0 + 2;
  std::vector<int> err(nt, 0);
  // 1. v1 in
  {
    std::vector<std::thread> th;
    for (int t = 0; t < nt; t++)
      th.emplace_back([&, t]() {
        size_t lo, hi;
        share(bytes, t, nt, &lo, &hi);
        int b = 0;
        for (size_t o = lo; o < hi; o += HOST_CHUNK, b ^= 1) {
          const size_t n = std::min(HOST_CHUNK, hi - o);
          if (hipEventSynchronize(e->h_ev[t][b]) != hipSuccess) { err[t] = 1; return; }
          memcpy(e->h_bounce[t][b], hv1 + o, n);
          if (hipMemcpyAsync(dv1 + o, e->h_bounce[t][b], n, hipMemcpyHostToDevice, e->h_stream[t]) != hipSuccess ||
              hipEventRecord(e->h_ev[t][b], e->h_stream[t]) != hipSuccess) {
            err[t] = 1;
            return;
          }
        }
      });
    for (auto &x : th) x.join();
  }
  for (int t = 0; t < nt; t++) {
    if (err[t]) return hphi_set_error("hphi_mltply: host-to-device copy failed");
    for (int b = 0; b < 2; b++) HIPCHK(hipStreamWaitEvent(e->stream, e->h_ev[t][b], 0));
  }
  // 2. H v1 on the device (overwrite)
  if (hphi_hv_launch_ex(e, 1, e->d_stage0, e->d_stage1, 1)) return -1;
  hipEvent_t done;
  HIPCHK(hipEventCreateWithFlags(&done, hipEventDisableTiming));
  HIPCHK(hipEventRecord(done, e->stream));
  // 3. H v1 out, added to v0 on the host; each thread keeps one chunk in flight
  {
    std::vector<std::thread> th;
    for (int t = 0; t < nt; t++)
      th.emplace_back([&, t]() {
        size_t lo, hi;
        share(bytes, t, nt, &lo, &hi);
        if (hipStreamWaitEvent(e->h_stream[t], done, 0) != hipSuccess) { err[t] = 1; return; }
        auto issue = [&](size_t o, int b) {
          const size_t n = std::min(HOST_CHUNK, hi - o);
          return hipMemcpyAsync(e->h_bounce[t][b], dout + o, n, hipMemcpyDeviceToHost, e->h_stream[t]) == hipSuccess &&
                 hipEventRecord(e->h_ev[t][b], e->h_stream[t]) == hipSuccess;
        };
        int b = 0;
        if (lo < hi && !issue(lo, 0)) { err[t] = 1; return; }
        for (size_t o = lo; o < hi; o += HOST_CHUNK, b ^= 1) {
          if (o + HOST_CHUNK < hi && !issue(o + HOST_CHUNK, b ^ 1)) { err[t] = 1; return; }
          if (hipEventSynchronize(e->h_ev[t][b]) != hipSuccess) { err[t] = 1; return; }
          const size_t n = std::min(HOST_CHUNK, hi - o) / sizeof(double);
          const double *src = (const double *)e->h_bounce[t][b];
          double *dst = hv0 + o / sizeof(double);
          for (size_t i = 0; i < n; i++) dst[i] += src[i];
        }
      });
    for (auto &x : th) x.join();
  }
  (void)hipEventDestroy(done);
  for (int t = 0; t < nt; t++)
    if (err[t]) return hphi_set_error("hphi_mltply: device-to-host copy failed");
  return hphi_fetch_prdct(e, 1, prdct);
}

// ----------------------------------------------------------- helpers ----
extern "C" int hphi_dev_alloc(size_t bytes, void **ptr) {
  if (!ptr) return hphi_set_error("NULL ptr");
  HIPCHK(hipMalloc(ptr, bytes ? bytes : 1));
  return 0;
}
extern "C" int hphi_dev_free(void *ptr) {
  HIPCHK(hipFree(ptr));
  return 0;
}
extern "C" int hphi_memcpy_h2d(void *dst, const void *src, size_t bytes) {
  HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return 0;
}
extern "C" int hphi_memcpy_d2h(void *dst, const void *src, size_t bytes) {
  HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return 0;
}
extern "C" int hphi_dev_zero(void *ptr, size_t bytes) {
  uint64_t n = bytes / sizeof(double2);
  if (n) hipLaunchKernelGGL(k_zero, dim3(grid_for(n, 256)), dim3(256), 0, 0, (double2 *)ptr, n);
  HIPCHK(hipGetLastError());
  HIPCHK(hipDeviceSynchronize());
  return 0;
}
extern "C" int hphi_synchronize(void) {
  HIPCHK(hipDeviceSynchronize());
  return 0;
}
extern "C" int hphi_fill_random(double *v, uint64_t dim, int k, uint64_t seed) {
  if (!v || k < 1) return hphi_set_error("hphi_fill_random: bad argument");
  uint64_t n = dim * (uint64_t)k;
  hipLaunchKernelGGL(k_fill_random, dim3(grid_for(n, 256)), dim3(256), 0, 0, (double2 *)v, dim, k, seed);
  HIPCHK(hipGetLastError());
  HIPCHK(hipDeviceSynchronize());
  return 0;
}

extern "C" int hphi_time_mltply_dev(hphi_engine *e, int k, double *v0, const double *v1, int warmup,
                                    int reps, double *avg_ms) {
  if (!e || !v0 || !v1 || reps < 1) return hphi_set_error("hphi_time_mltply_dev: bad argument");
  if (ensure_ready(e)) return -1;
  for (int i = 0; i < warmup; i++)
    if (hphi_hv_launch(e, k, v0, v1)) return -1;
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  HIPCHK(hipEventRecord(a, e->stream));
  for (int i = 0; i < reps; i++)
    if (hphi_hv_launch(e, k, v0, v1)) return -1;
  HIPCHK(hipEventRecord(b, e->stream));
  HIPCHK(hipEventSynchronize(b));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  *avg_ms = ms / reps;
  return 0;
}
