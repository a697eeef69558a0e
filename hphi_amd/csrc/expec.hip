// Expectation values of a device vector v (normalised by the caller):
//
//   hphi_expec_energy_flct   <- expec_energy_flct (src/expec_energy_flct.c:40-160,186-660):
//                               <v|H|v>, <Hv|Hv> and the doublon / electron-number / Sz moments
//   hphi_expec_cisajs        <- expec_cisajs (src/expec_cisajs.c:69-650): <v|c+_{i s1} c_{j s2}|v>
//   hphi_expec_cisajscktalt  <- expec_cisajscktaltdc (src/expec_cisajscktaltdc.c:87-1647):
//                               <v|c+_{i s1} c_{j s2} c+_{k s3} c_{l s4}|v>
//
// Every Green's function entry is compiled on the host into gather terms (the
// same forms the Hamiltonian uses), each tagged with its output slot.  One
// kernel pass evaluates EX_CW slots for all states: lane j holds target state
// x = list_1[j] (or j), and for each active term adds
// conj(v[j]) * c * sign * v[idx(x ^ flip)], i.e. <v|O|v> = sum_x conj(v_x) (O v)_x.
// Reductions run in a fixed order (waves, then blocks), so results are
// reproducible run to run.
#include <cmath>
#include <cstring>
#include <utility>
#include <vector>

#include "engine_internal.h"

#define EX_TPB 256
#define EX_GRID 512
#define EX_CW 8  // output slots per pass

namespace {

struct CorrTerm {
  uint64_t flip, tmask, tval, sgn;
  double cre, cim;
};

__device__ __forceinline__ double wave_sum(double s) {
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  return s;
}

// per state: w = |v_j|^2; kind 1 (SpinGC): S = 2 popcount(x) - N;
// kind 2 (Hubbard models): D doublons, N electrons, S = N_up - N_down
__global__ __launch_bounds__(EX_TPB) void k_moments(const double2 *v, const uint64_t *list1, uint64_t dim, int kind,
                                                     int nsite, double *partial) {
  double r[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const uint64_t EVEN = 0x5555555555555555ull;
  for (uint64_t j = (uint64_t)blockIdx.x * EX_TPB + threadIdx.x; j < dim; j += (uint64_t)gridDim.x * EX_TPB) {
    const double2 a = v[j];
    const double w = a.x * a.x + a.y * a.y;
    const uint64_t x = list1 ? list1[j] : j;
    if (kind == 1) {
      const double S = (double)(2 * (int)__popcll(x) - nsite);
      r[4] += w * S;
      r[5] += w * S * S;
    } else {
      const double D = (double)(int)__popcll(x & (x >> 1) & EVEN);
      const int up = (int)__popcll(x & EVEN), dn = (int)__popcll(x & (EVEN << 1));
      const double N = (double)(up + dn), S = (double)(up - dn);
      r[0] += w * D;
      r[1] += w * D * D;
      r[2] += w * N;
      r[3] += w * N * N;
      r[4] += w * S;
      r[5] += w * S * S;
    }
  }
  __shared__ double sh[EX_TPB / 64][6];
#pragma unroll
  for (int q = 0; q < 6; q++) {
    const double s = wave_sum(r[q]);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6][q] = s;
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    double s = 0.0;
    for (int w = 0; w < EX_TPB / 64; w++) s += sh[w][threadIdx.x];
    partial[(uint64_t)blockIdx.x * 6 + threadIdx.x] = s;
  }
}

// block (bx, c): slots c*EX_CW .. c*EX_CW+EX_CW-1; the terms of slot s are
// terms[toff[s] .. toff[s+1])
__global__ __launch_bounds__(EX_TPB) void k_corr(const double2 *v, const uint64_t *list1, const uint32_t *l21,
                                                  const uint64_t *l22, uint64_t dim, uint64_t irght, int nlow,
                                                  const CorrTerm *terms, const int *toff, int nslot,
                                                  double2 *partial) {
  const int s0 = blockIdx.y * EX_CW;
  double ar[EX_CW], ai[EX_CW];
#pragma unroll
  for (int q = 0; q < EX_CW; q++) ar[q] = ai[q] = 0.0;
  for (uint64_t j = (uint64_t)blockIdx.x * EX_TPB + threadIdx.x; j < dim; j += (uint64_t)gridDim.x * EX_TPB) {
    const uint64_t x = list1 ? list1[j] : j;
    const double2 vj = v[j];
#pragma unroll
    for (int q = 0; q < EX_CW; q++) {
      if (s0 + q >= nslot) break;
      for (int t = toff[s0 + q]; t < toff[s0 + q + 1]; t++) {
        const CorrTerm T = terms[t];
        if ((x & T.tmask) != T.tval) continue;
        const uint64_t y = x ^ T.flip;
        uint64_t idx = y;
        if (list1) {  // canonical: rank tables, 0 = not in the sector
          const uint64_t la = l21[y & irght], lb = l22[y >> nlow];
          if (la == 0 || lb == 0) continue;
          idx = la + lb - 3;
        }
        if (idx >= dim) continue;
        const double2 u = v[idx];
        // c * sign * u
        double cr = T.cre, ci = T.cim;
        if (__popcll(x & T.sgn) & 1) {
          cr = -cr;
          ci = -ci;
        }
        const double wr = cr * u.x - ci * u.y, wi = cr * u.y + ci * u.x;
        // conj(vj) * w
        ar[q] += vj.x * wr + vj.y * wi;
        ai[q] += vj.x * wi - vj.y * wr;
      }
    }
  }
  __shared__ double sh[EX_TPB / 64][EX_CW][2];
#pragma unroll
  for (int q = 0; q < EX_CW; q++) {
    const double sr = wave_sum(ar[q]), si = wave_sum(ai[q]);
    if ((threadIdx.x & 63) == 0) {
      sh[threadIdx.x >> 6][q][0] = sr;
      sh[threadIdx.x >> 6][q][1] = si;
    }
  }
  __syncthreads();
  if (threadIdx.x < EX_CW) {
    double sr = 0.0, si = 0.0;
    for (int w = 0; w < EX_TPB / 64; w++) {
      sr += sh[w][threadIdx.x][0];
      si += sh[w][threadIdx.x][1];
    }
    partial[((uint64_t)blockIdx.y * gridDim.x + blockIdx.x) * EX_CW + threadIdx.x] = make_double2(sr, si);
  }
}

// slot s: sum over the blocks of its chunk in block order
__global__ void k_corr_final(const double2 *partial, int nblk, int nslot, double2 *out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nslot) return;
  const int c = s / EX_CW, q = s % EX_CW;
  double sr = 0.0, si = 0.0;
  for (int b = 0; b < nblk; b++) {
    const double2 p = partial[((uint64_t)c * nblk + b) * EX_CW + q];
    sr += p.x;
    si += p.y;
  }
  out[s] = make_double2(sr, si);
}

int ex_grid(uint64_t dim) {
  uint64_t g = (dim + EX_TPB - 1) / EX_TPB;
  return (int)(g < EX_GRID ? (g ? g : 1) : EX_GRID);
}

// density-like projector: target (and source) has (x & mask) == val
void push_proj(std::vector<GTerm> &out, uint64_t mask, uint64_t val) {
  GTerm g{};
  g.kind = P_EQ;
  g.tmask = mask;
  g.tval = val;
  g.cre = 1.0;
  out.push_back(g);
}

// evaluate the compiled slots on v; out[2*s], out[2*s+1] = re, im
int run_corr(hphi_engine *e, const double *v, const std::vector<std::vector<GTerm>> &slots, double *out) {
  const int nslot = (int)slots.size();
  if (nslot == 0) return 0;
  if (hphi_build_tables(e)) return -1;
  std::vector<CorrTerm> h;
  std::vector<int> toff(1, 0);
  for (const auto &sl : slots) {
    for (const GTerm &g : sl) h.push_back(CorrTerm{g.flip, g.tmask, g.tval, g.sgn, g.cre, g.cim});
    toff.push_back((int)h.size());
  }
  const bool gc = hphi_model_gc(e->model);
  const int g = ex_grid(e->dim), nchunk = (nslot + EX_CW - 1) / EX_CW;
  CorrTerm *d_terms = nullptr;
  int *d_toff = nullptr;
  double2 *d_part = nullptr, *d_out = nullptr;
  int rc = -1;
  if (hphi_dev_alloc(sizeof(CorrTerm) * (h.size() + 1), (void **)&d_terms) ||
      hphi_dev_alloc(sizeof(int) * toff.size(), (void **)&d_toff) ||
      hphi_dev_alloc(sizeof(double2) * (size_t)nchunk * g * EX_CW, (void **)&d_part) ||
      hphi_dev_alloc(sizeof(double2) * nslot, (void **)&d_out))
    goto done;
  if (!h.empty() && hipMemcpyAsync(d_terms, h.data(), sizeof(CorrTerm) * h.size(), hipMemcpyHostToDevice,
                                   e->stream) != hipSuccess) {
    hphi_set_error("hphi_expec: term upload failed");
    goto done;
  }
  if (hipMemcpyAsync(d_toff, toff.data(), sizeof(int) * toff.size(), hipMemcpyHostToDevice, e->stream) !=
      hipSuccess) {
    hphi_set_error("hphi_expec: offset upload failed");
    goto done;
  }
  for (int c0 = 0; c0 < nchunk; c0 += 65535) {  // grid.y limit
    const int nc = nchunk - c0 < 65535 ? nchunk - c0 : 65535;
    const int ns = (c0 + nc) * EX_CW < nslot ? (c0 + nc) * EX_CW : nslot;
    // chunks c0.. run with blockIdx.y offset folded into the slot base
    hipLaunchKernelGGL(k_corr, dim3(g, nc), dim3(EX_TPB), 0, e->stream, (const double2 *)v,
                       gc ? (const uint64_t *)nullptr : e->d_list1, e->d_l21, e->d_l22, e->dim, e->irght, e->nlow,
                       d_terms, d_toff + (size_t)c0 * EX_CW, ns - c0 * EX_CW, d_part + (size_t)c0 * g * EX_CW);
    if (hipGetLastError() != hipSuccess) {
      hphi_set_error("hphi_expec: k_corr launch failed");
      goto done;
    }
  }
  hipLaunchKernelGGL(k_corr_final, dim3((nslot + 255) / 256), dim3(256), 0, e->stream, (const double2 *)d_part, g,
                     nslot, d_out);
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(out, d_out, sizeof(double2) * nslot, hipMemcpyDeviceToHost, e->stream) != hipSuccess ||
      hipStreamSynchronize(e->stream) != hipSuccess) {
    hphi_set_error("hphi_expec: correlator pass failed");
    goto done;
  }
  rc = 0;
done:
  hphi_dev_free(d_terms);
  hphi_dev_free(d_toff);
  hphi_dev_free(d_part);
  hphi_dev_free(d_out);
  return rc;
}

}  // namespace

extern "C" int hphi_expec_energy_flct(hphi_engine *e, const double *v, hphi_phys *out) {
  if (!e || !v || !out) return hphi_set_error("hphi_expec_energy_flct: NULL argument");
  memset(out, 0, sizeof *out);
  if (hphi_build_tables(e)) return -1;
  // moments (expec_energy_flct_HalfSpinGC / _Hubbard / _HubbardGC)
  const int kind = e->model == HPHI_MODEL_SPINGC ? 1 : hphi_model_hubbard(e->model) ? 2 : 0;
  if (kind) {
    const int g = ex_grid(e->dim);
    double *d_part = nullptr;
    std::vector<double> part((size_t)g * 6);
    if (hphi_dev_alloc(sizeof(double) * part.size(), (void **)&d_part)) return -1;
    hipLaunchKernelGGL(k_moments, dim3(g), dim3(EX_TPB), 0, e->stream, (const double2 *)v,
                       hphi_model_gc(e->model) ? (const uint64_t *)nullptr : e->d_list1, e->dim, kind, e->nsite,
                       d_part);
    const bool ok = hipGetLastError() == hipSuccess &&
                    hipMemcpyAsync(part.data(), d_part, sizeof(double) * part.size(), hipMemcpyDeviceToHost,
                                   e->stream) == hipSuccess &&
                    hipStreamSynchronize(e->stream) == hipSuccess;
    hphi_dev_free(d_part);
    if (!ok) return hphi_set_error("hphi_expec_energy_flct: moment pass failed");
    double m[6] = {0, 0, 0, 0, 0, 0};
    for (int b = 0; b < g; b++)
      for (int q = 0; q < 6; q++) m[q] += part[(size_t)b * 6 + q];
    if (kind == 1) {  // expec_energy_flct.c:484-491
      out->num = e->nsite;
      out->num2 = (double)e->nsite * e->nsite;
    } else {  // :290-297, :414-421
      out->doublon = m[0];
      out->doublon2 = m[1];
      out->num = m[2];
      out->num2 = m[3];
    }
    out->sz = 0.5 * m[4];
    out->sz2 = 0.25 * m[5];
  } else {  // canonical Spin: constants (expec_energy_flct.c:104-118)
    const double sz = 0.5 * (double)(e->nup - e->ndown);
    out->num = e->nsite;
    out->num2 = (double)e->nsite * e->nsite;
    out->sz = sz;
    out->sz2 = sz * sz;
  }
  // energy <v|H|v> (the H·v's prdct) and variance <Hv|Hv> (:141-158)
  double *hv = nullptr, pr[2], n2;
  if (hphi_dev_alloc(16 * e->dim, (void **)&hv)) return -1;
  int rc = -1;
  if (!hphi_dev_zero(hv, 16 * e->dim) && !hphi_mltply_dev(e, hv, v, pr) && !hphi_vec_norm2(e, hv, 1, &n2)) {
    out->energy = pr[0];
    out->var = n2;
    rc = 0;
  }
  hphi_dev_free(hv);
  return rc;
}

extern "C" int hphi_expec_cisajs(hphi_engine *e, const double *v, int n, const int *idx, double *out) {
  if (!e || !v || (n > 0 && (!idx || !out))) return hphi_set_error("hphi_expec_cisajs: NULL argument");
  const int N = e->nsite;
  const bool hub = hphi_model_hubbard(e->model), gc = hphi_model_gc(e->model);
  std::vector<std::vector<GTerm>> slots(n);
  for (int t = 0; t < n; t++) {
    const int i = idx[4 * t], s1 = idx[4 * t + 1], j = idx[4 * t + 2], s2 = idx[4 * t + 3];
    if (i < 0 || j < 0 || i >= N || j >= N || s1 < 0 || s1 > 1 || s2 < 0 || s2 > 1)
      return hphi_set_error("OneBodyG entry %d (%d %d %d %d) out of range", t, i, s1, j, s2);
    std::vector<GTerm> &sl = slots[t];
    if (hub) {
      if (!gc && s1 != s2) continue;  // Sz-conserving sector (expec_cisajs.c:276-282)
      const int a = 2 * i + s1, b = 2 * j + s2;
      if (a == b) {  // n_a (:330-338)
        push_proj(sl, 1ull << a, 1ull << a);
      } else {  // c+_a c_b with its Jordan-Wigner sign (general_hopp / GC_general_hopp)
        GTerm g{};
        g.kind = P_EQ;
        g.flip = (1ull << a) | (1ull << b);
        g.tmask = g.flip;
        g.tval = 1ull << a;
        g.sgn = hphi_between_mask(a, b);
        g.cre = 1.0;
        sl.push_back(g);
      }
    } else {
      if (i != j) continue;  // no hopping between localized spins (:406-408, :571-574)
      if (s1 == s2) {        // [bit_i == s1] (:389-403, :546-552)
        push_proj(sl, 1ull << i, s1 ? 1ull << i : 0ull);
      } else if (gc) {  // transverse: target bit s1, source bit s2 (:553-566)
        GTerm g{};
        g.kind = P_EQ;
        g.flip = 1ull << i;
        g.tmask = g.flip;
        g.tval = s1 ? g.flip : 0ull;
        g.cre = 1.0;
        sl.push_back(g);
      }  // canonical Spin: s1 != s2 changes Sz, 0 (:409-412)
    }
  }
  return run_corr(e, v, slots, out);
}

extern "C" int hphi_expec_cisajscktalt(hphi_engine *e, const double *v, int n, const int *idx, double *out) {
  if (!e || !v || (n > 0 && (!idx || !out))) return hphi_set_error("hphi_expec_cisajscktalt: NULL argument");
  const int N = e->nsite;
  const bool hub = hphi_model_hubbard(e->model), gc = hphi_model_gc(e->model);
  std::vector<std::vector<GTerm>> slots(n);
  for (int t = 0; t < n; t++) {
    const int *q = idx + 8 * t;
    for (int u = 0; u < 8; u += 2)
      if (q[u] < 0 || q[u] >= N || q[u + 1] < 0 || q[u + 1] > 1)
        return hphi_set_error("TwoBodyG entry %d out of range", t);
    std::vector<GTerm> &sl = slots[t];
    if (hub) {
      if (!gc && q[1] + q[5] != q[3] + q[7]) continue;  // expec_cisajscktaltdc.c:535-541
      hphi_compile_two_hops(2 * q[0] + q[1], 2 * q[2] + q[3], 2 * q[4] + q[5], 2 * q[6] + q[7], 1.0, 0.0, sl);
      continue;
    }
    // spin entries are (i s1 i s2 k s3 k s4) (readdef.c:1497-1503); the pair
    // with the smaller site comes first (Rearray_Interactions, :274-381)
    if (q[0] != q[2] || q[4] != q[6])
      return hphi_set_error("spin TwoBodyG entry %d needs i == j and k == l", t);
    int A = q[0], s1 = q[1], s2 = q[3], B = q[4], s3 = q[5], s4 = q[7];
    if (A > B) {
      std::swap(A, B);
      std::swap(s1, s3);
      std::swap(s2, s4);
    }
    if (!gc) {  // canonical branches (:736-772): diagonal, one-site, exchange; others 0
      const bool keep = (s1 == s2 && s3 == s4) || (A == B && s1 == s4 && s3 == s2) || (s1 == s4 && s2 == s3);
      if (!keep) continue;
    }
    hphi_compile_spin_pair(A, s1, s2, B, s3, s4, 1.0, 0.0, sl);
  }
  return run_corr(e, v, slots, out);
}

// SpinGC three-, four- and six-body Green's functions
// (expec_Threebody_SpinGCHalf / expec_Fourbody_SpinGCHalf /
// expec_Sixbody_SpinGCHalf, src/expec_cisajscktaltdc.c:901-1420).  An entry
// holds nbody pairs (site, sigma)(site, sigma); pair p is O_p = |s><s'| on its
// site when both sites agree.  The value is <v| O_1 O_2 O_3 ... O_nbody |v>:
//  * pairs 3 .. nbody are applied to v right to left (mltplyHalfSpinGC_mini,
//    src/mltplySpin.c:416-490); one on two different sites leaves zero;
//  * the first two follow Rearray_Interactions (:274-381): (i,i)(k,k) as given,
//    the cross form (i,k)(k,i) as -|s1><s4|_i |s3><s2|_k; any other first two
//    pairs are not evaluated (ok = 0; the reference prints the first eight
//    indices and zeros).
// The product of on-site projectors and flips is one gather term: per site the
// composed operator |a><b| (|a><b| after |c><d| is delta(b, c) |a><d|), the
// target x has bits a, the source is x with bits b.
extern "C" int hphi_expec_spin_nbody(hphi_engine *e, const double *v, int nbody, int n, const int *idx, double *out,
                                     int *ok) {
  if (!e || !v || (n > 0 && (!idx || !out || !ok))) return hphi_set_error("hphi_expec_spin_nbody: NULL argument");
  if (nbody != 3 && nbody != 4 && nbody != 6) return hphi_set_error("hphi_expec_spin_nbody: nbody must be 3, 4 or 6");
  if (e->model != HPHI_MODEL_SPINGC)
    return hphi_set_error("hphi_expec_spin_nbody: the reference defines these for SpinGC (spin 1/2) only");
  const int N = e->nsite;
  std::vector<std::vector<GTerm>> slots(n);
  for (int t = 0; t < n; t++) {
    const int *q = idx + 4 * nbody * t;
    for (int u = 0; u < 4 * nbody; u += 2)
      if (q[u] < 0 || q[u] >= N || q[u + 1] < 0 || q[u + 1] > 1)
        return hphi_set_error("%d-body Green's function entry %d out of range", nbody, t);
    ok[t] = 1;
    // the first two pairs after Rearray_Interactions: (site, s, s') x 2 and the sign
    int A, a1, a2, B, b1, b2;
    double sign = 1.0;
    if (q[0] == q[2] && q[4] == q[6]) {
      A = q[0], a1 = q[1], a2 = q[3], B = q[4], b1 = q[5], b2 = q[7];
    } else if (q[0] == q[6] && q[4] == q[2]) {
      A = q[0], a1 = q[1], a2 = q[7], B = q[4], b1 = q[5], b2 = q[3];
      sign = -1.0;
    } else {
      ok[t] = 0;
      continue;
    }
    if (A > B) {  // the pair with the smaller site first
      std::swap(A, B);
      std::swap(a1, b1);
      std::swap(a2, b2);
    }
    // compose, in the order the operators act on the ket
    int ca[64], cb[64];
    for (int s = 0; s < N; s++) ca[s] = cb[s] = -1;
    bool zero = false;
    auto apply = [&](int site, int s_to, int s_from) {
      if (ca[site] < 0) {
        ca[site] = s_to;
        cb[site] = s_from;
      } else if (s_from == ca[site]) {
        ca[site] = s_to;
      } else {
        zero = true;
      }
    };
    for (int p = nbody - 1; p >= 2 && !zero; p--) {
      const int *r = q + 4 * p;
      if (r[0] != r[2]) zero = true;  // mltplyHalfSpinGC_mini adds nothing for two sites
      else apply(r[0], r[1], r[3]);
    }
    if (!zero) apply(B, b1, b2);
    if (!zero) apply(A, a1, a2);
    if (zero) continue;
    GTerm g{};
    g.kind = P_EQ;
    for (int s = 0; s < N; s++) {
      if (ca[s] < 0) continue;
      const uint64_t bit = 1ull << s;
      g.tmask |= bit;
      if (ca[s]) g.tval |= bit;
      if (ca[s] != cb[s]) g.flip |= bit;
    }
    g.cre = sign;
    slots[t].push_back(g);
  }
  return run_corr(e, v, slots, out);
}
