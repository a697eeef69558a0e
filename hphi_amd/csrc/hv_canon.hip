// Gather-form H·v for canonical models (Spin, Hubbard) with states of at
// most 32 bits, one lane per basis state — the hot kernel of the plain path.
//
// Same arithmetic as k_hv in engine.hip (transpose of the reference's
// per-term scatter, mltplySpinCore.c:273-307, mltplyHubbardCore.c:305-347),
// organised for the CDNA4 memory pipeline:
//  * list_2_1 lives in LDS as uint16 when it fits (rank-in-class + 2 <= 65535):
//    the per-lane rank lookups of low-half and mixed terms never touch L1/L2;
//  * a wave whose 64 states share one high block (the common case: blocks
//    hold up to C(16,8) = 12870 states) takes a wave-uniform path for the
//    high-half terms: predicate, list_2_2 lookups and sign are scalar, the
//    gather is one contiguous 1 KB read, and inactive terms issue nothing;
//  * gathers are masked (only active lanes present an address to L1);
//  * state words are 32-bit; real couplings skip the imaginary products.
#include <hip/hip_runtime.h>

#include <map>

#include "engine_internal.h"

#define HIPCHK(x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) return hphi_set_error("%s failed: %s", #x, hipGetErrorString(e_)); \
  } while (0)

// rank lookup of list_2_1 (rank-in-class + 2): RK 0 the global table, 1 the
// table staged in LDS as uint16, 2 (canonical Spin) the combinatorial rank
// from two byte tables in LDS: the k-th lowest set bit at position p of ia
// contributes C(p, k), split into the low byte and the high byte given the
// low byte's popcount; 3 (canonical Hubbard) the same with one binomial per
// spin species, the high byte's table indexed by the low byte's up and down
// counts
template <int RK>
__device__ __forceinline__ int64_t r21(const HvArgs &A, const uint16_t *s21, uint32_t ia) {
  if (RK == 3) return hrank_l21(s21, ia);
  if (RK == 2) return crank_l21(s21, ia);
  if (RK == 1) return (int64_t)s21[ia];
  return (int64_t)A.l21[ia];
}

__device__ __forceinline__ bool pred32(int kind, uint32_t tmask, uint32_t tval, uint32_t x) {
  const uint32_t ym = x & tmask;
  const int pc = __popc(ym);
  return (kind == P_EQ) ? (ym == tval) : ((kind == P_DIFF) ? (pc == 1) : (pc != 1));
}

template <bool REALC>
__device__ __forceinline__ void cmac(double &ar, double &ai, double cr, double ci, double2 g) {
  if (REALC) {
    ar += cr * g.x;
    ai += cr * g.y;
  } else {
    ar += cr * g.x - ci * g.y;
    ai += cr * g.y + ci * g.x;
  }
}

template <int TPB, bool REALC, int RK, int UNR>
__global__ __launch_bounds__(TPB) void k_hv_canon(const HvArgs A) {
  extern __shared__ uint16_t s21[];
  const uint32_t tid = threadIdx.x;
  if (RK == 1) {
    for (uint32_t i = tid; i < (uint32_t)A.size21; i += TPB) s21[i] = (uint16_t)A.l21[i];
    __syncthreads();
  } else if (RK >= 2) {
    const uint32_t n = RK == 3 ? (uint32_t)HRANK_N : (uint32_t)CRANK_N;
    for (uint32_t i = tid; i < n; i += TPB) s21[i] = A.crank[i];
    __syncthreads();
  }
  const double2 *__restrict__ v1 = (const double2 *)A.v1;
  double2 *__restrict__ v0 = (double2 *)A.v0;
  const uint32_t irght = (uint32_t)A.irght;
  const int nlow = A.nlow;
  const uint64_t dim = A.dim;
  const CTerm tl = (CTerm)A.terms;
  const CTerm th = tl + A.n_low;
  const CTerm tm = th + A.n_high;
  double pre = 0.0, pim = 0.0;

  const uint64_t ntiles = (dim + TPB - 1) / TPB;
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
    const uint64_t j = tile * TPB + tid;
    const bool valid = j < dim;
    const uint64_t jj = valid ? j : dim - 1;
    const uint32_t x = (uint32_t)A.list1[jj];
    const uint32_t ia = x & irght, ib = x >> nlow;
    const double2 a = v1[jj];
    const double d = A.diag[jj];
    double ar = d * a.x, ai = d * a.y;
    const int64_t b21 = r21<RK>(A, s21, ia);
    if (A.v1t && valid) {  // tile-major copy for the high-half pass (block ib, rank b21 - 2)
      const ulonglong2 ts = A.tofs[ib];
      const uint64_t r = (uint64_t)(b21 - 2);
      A.v1t[ts.x + (r / HI_RT) * ts.y + r % HI_RT] = a;
    }

    // ---- low-half terms: target in the same block, per-lane rank lookup
    for (int t = 0; t < A.e_low; t += UNR) {
      double2 g[UNR];
      int64_t l[UNR];
      bool act[UNR];
      // phase 1: all rank lookups of the batch (LDS, or L2 issued back to back)
#pragma unroll
      for (int u = 0; u < UNR; u++) {
        const GTerm T = load_term(tl + t + u);
        const uint32_t y = x ^ (uint32_t)T.flip;
        act[u] = valid & pred32(T.kind, (uint32_t)T.tmask, (uint32_t)T.tval, x);
        l[u] = r21<RK>(A, s21, y & irght);
      }
      // phase 2: all gathers of the batch, masked to the active lanes
#pragma unroll
      for (int u = 0; u < UNR; u++) {
        const uint64_t idx = (uint64_t)((int64_t)j + l[u] - b21);
        const bool a2 = act[u] & (l[u] != 0) & (idx < dim);
        g[u] = make_double2(0.0, 0.0);
        if (a2) g[u] = v1[idx];
      }
#pragma unroll
      for (int u = 0; u < UNR; u++) {
        const GTerm T = load_term(tl + t + u);
        double2 gg = g[u];
        if ((uint32_t)T.sgn != 0u && (__popc(x & (uint32_t)T.sgn) & 1)) {
          gg.x = -gg.x;
          gg.y = -gg.y;
        }
        cmac<REALC>(ar, ai, T.cre, T.cim, gg);
      }
    }

    // ---- high-half and mixed terms
    const uint32_t ib0 = __builtin_amdgcn_readfirstlane(ib);
    const bool uni = (__ballot(ib != ib0) == 0ull);
    if (uni) {
      // wave-uniform block: predicate, lookups and sign of high terms are scalar
      const uint32_t xh = ib0 << nlow;
      const int64_t l22_0 = (int64_t)A.l22[ib0];
      for (int t = 0; t < A.e_high; t += UNR) {
        double2 g[UNR];
        bool au[UNR];
#pragma unroll
        for (int u = 0; u < UNR; u++) {
          const GTerm T = load_term(th + t + u);
          au[u] = pred32(T.kind, (uint32_t)T.tmask, (uint32_t)T.tval, xh);
          g[u] = make_double2(0.0, 0.0);
          if (au[u]) {
            const int64_t l = (int64_t)A.l22[(xh ^ (uint32_t)T.flip) >> nlow];
            au[u] = (l != 0);
            if (au[u] && valid) g[u] = v1[(uint64_t)((int64_t)j + l - l22_0)];
          }
        }
#pragma unroll
        for (int u = 0; u < UNR; u++) {
          const GTerm T = load_term(th + t + u);
          if (au[u]) {
            double s = ((uint32_t)T.sgn != 0u && (__popc(xh & (uint32_t)T.sgn) & 1)) ? -1.0 : 1.0;
            cmac<REALC>(ar, ai, s * T.cre, s * T.cim, g[u]);
          }
        }
      }
      for (int t = 0; t < A.e_mixed; t += UNR) {
        double2 g[UNR];
        int64_t la[UNR], lb[UNR];
        bool act[UNR];
#pragma unroll
        for (int u = 0; u < UNR; u++) {
          const GTerm T = load_term(tm + t + u);
          const uint32_t y = x ^ (uint32_t)T.flip;
          act[u] = valid & pred32(T.kind, (uint32_t)T.tmask, (uint32_t)T.tval, x);
          lb[u] = (int64_t)A.l22[(xh ^ (uint32_t)T.flip) >> nlow];  // uniform
          la[u] = r21<RK>(A, s21, y & irght);
        }
#pragma unroll
        for (int u = 0; u < UNR; u++) {
          const uint64_t idx = (uint64_t)(la[u] + lb[u] - 3);
          const bool a2 = act[u] & (la[u] != 0) & (lb[u] != 0) & (idx < dim);
          g[u] = make_double2(0.0, 0.0);
          if (a2) g[u] = v1[idx];
        }
#pragma unroll
        for (int u = 0; u < UNR; u++) {
          const GTerm T = load_term(tm + t + u);
          double2 gg = g[u];
          if ((uint32_t)T.sgn != 0u && (__popc(x & (uint32_t)T.sgn) & 1)) {
            gg.x = -gg.x;
            gg.y = -gg.y;
          }
          cmac<REALC>(ar, ai, T.cre, T.cim, gg);
        }
      }
    } else {
      // lanes span several blocks: per-lane list_2_2 lookups
      const int64_t b22 = (int64_t)A.l22[ib];
      for (int seg = 0; seg < 2; seg++) {  // the high group, then the mixed group
        const CTerm tg = seg ? tm : th;
        const int ne = seg ? A.e_mixed : A.e_high;
        for (int t = 0; t < ne; t += UNR) {
          double2 g[UNR];
#pragma unroll
          for (int u = 0; u < UNR; u++) {
            const GTerm T = load_term(tg + t + u);
            const uint32_t y = x ^ (uint32_t)T.flip;
            bool act = valid & pred32(T.kind, (uint32_t)T.tmask, (uint32_t)T.tval, x);
            int64_t la = b21, lb = 0;
            if (act) {
              lb = (int64_t)A.l22[y >> nlow];
              if ((y & irght) != ia) la = r21<RK>(A, s21, y & irght);
            }
            const uint64_t idx = (uint64_t)(la + lb - 3);
            act = act & (la != 0) & (lb != 0) & (idx < dim);
            g[u] = make_double2(0.0, 0.0);
            if (act) g[u] = v1[idx];
          }
#pragma unroll
          for (int u = 0; u < UNR; u++) {
            const GTerm T = load_term(tg + t + u);
            double2 gg = g[u];
            if ((uint32_t)T.sgn != 0u && (__popc(x & (uint32_t)T.sgn) & 1)) {
              gg.x = -gg.x;
              gg.y = -gg.y;
            }
            cmac<REALC>(ar, ai, T.cre, T.cim, gg);
          }
        }
      }
      (void)b22;
    }

    if (valid) {
      double2 o = make_double2(A.c1 * ar, A.c1 * ai);
      if (!A.overwrite) {
        const double2 p = v0[j];
        o.x += A.c0 * p.x;
        o.y += A.c0 * p.y;
      }
      v0[j] = o;
      pre += a.x * ar + a.y * ai;
      pim += a.x * ai - a.y * ar;
    }
  }

  // workgroup reduction of prdct (fixed order)
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    pre += __shfl_xor(pre, off, 64);
    pim += __shfl_xor(pim, off, 64);
  }
  __shared__ double2 red[TPB / 64];
  if ((tid & 63) == 0) red[tid >> 6] = make_double2(pre, pim);
  __syncthreads();
  if (tid == 0) {
    double2 s = red[0];
    for (int w = 1; w < TPB / 64; w++) {
      s.x += red[w].x;
      s.y += red[w].y;
    }
    A.partial[blockIdx.x] = s;
  }
}

// ---- canonical Spin (and Hubbard): the high-half terms over a tile-major copy of v1
// A state of a canonical sector is (block ib = its high-half bits, rank r of
// its low-half bits among the low-half values of the same class), at index
// list_2_2[ib] - 1 + r.  A term that moves one spin (or one electron) between
// two high-half bits maps (ib, r) to (ib ^ m, r): a block of the same class at
// the same rank.  k_hv_canon visits all ranks of a block before the next, so
// each high term streams a whole other block and every block's rows come from
// HBM again for each block that gathers from them.  Here the high group runs
// in its own pass ordered by (class, tile of HI_RT ranks, group of HI_BPG
// blocks), over a copy of v1 in tile-major order (tile R of the class's block
// at position p at base[c] + (R nblk[c] + p) HI_RT; k_hv_canon writes it as
// it reads each state's own entry, or k_hi_copy does): one rank tile of a
// class is then one contiguous slab (HI_RT x 16 B per block: 6.6 MB for the
// largest class of the 8x4 sector), each XCD takes a contiguous range of
// tiles, and every high term of the tile gathers whole lines from the slab
// while it is cached.  On the 8x4 J1-J2 sector: 101.6 ms in one pass; 89.8 ms
// as k_hv_canon without the high group (62 ms), the copy (4.8 ms) and this
// pass (23 ms); 90.0 ms with the copy fused and the first pass on a larger
// grid, on a box where the single pass takes 107.4 ms.

__device__ __forceinline__ bool hi_unit(const HiPlan &P, uint64_t u, int &c, uint32_t &R, uint32_t &grp) {
  while (u >= P.wgp[c + 1]) c++;  // a workgroup's units only increase
  const uint64_t loc = u - P.wgp[c];
  R = (uint32_t)(loc / P.ngrp[c]);
  grp = (uint32_t)(loc % P.ngrp[c]);
  return true;
}

// v1 into tile-major order (the workgroups and units of k_hv_high)
__global__ __launch_bounds__(256) void k_hi_copy(const double2 *__restrict__ v1, const uint64_t *__restrict__ l22,
                                                 const HiPlan P, const uint32_t *__restrict__ hb,
                                                 double2 *__restrict__ v1t) {
  const uint32_t tid = threadIdx.x;
  const uint64_t U = P.wgp[P.ncls];
  const uint32_t xcd = blockIdx.x & 7, nslot = gridDim.x >> 3;
  const uint64_t u1 = U * (xcd + 1) / 8;
  int c = 0;
  for (uint64_t u = U * xcd / 8 + (blockIdx.x >> 3); u < u1; u += nslot) {
    uint32_t R, grp;
    hi_unit(P, u, c, R, grp);
    const uint32_t p = grp * HI_BPG + tid / HI_RT, rl = tid % HI_RT, r = R * HI_RT + rl;
    if (p < P.nblk[c] && r < P.size[c]) {
      const uint32_t ib = hb[P.start[c] + p];
      v1t[P.base[c] + ((uint64_t)R * P.nblk[c] + p) * HI_RT + rl] = v1[l22[ib] - 1 + r];
    }
  }
}

template <bool REALC, int UNR>
__global__ __launch_bounds__(256) void k_hv_high(const HvArgs A, const HiPlan P, const uint32_t *__restrict__ hb,
                                                 const uint32_t *__restrict__ blkpos,
                                                 const double2 *__restrict__ v1t, double2 *__restrict__ partial) {
  const uint32_t tid = threadIdx.x;
  double2 *__restrict__ v0 = (double2 *)A.v0;
  const int nlow = A.nlow;
  const CTerm th = (CTerm)A.terms + A.n_low;
  const uint64_t U = P.wgp[P.ncls];
  const uint32_t xcd = blockIdx.x & 7, nslot = gridDim.x >> 3;
  const uint64_t u1 = U * (xcd + 1) / 8;
  double pre = 0.0, pim = 0.0;
  int c = 0;
  for (uint64_t u = U * xcd / 8 + (blockIdx.x >> 3); u < u1; u += nslot) {
    uint32_t R, grp;
    hi_unit(P, u, c, R, grp);
    const uint32_t nb = P.nblk[c];
    const uint32_t p = grp * HI_BPG + tid / HI_RT, rl = tid % HI_RT, r = R * HI_RT + rl;
    const bool valid = p < nb && r < P.size[c];
    const uint32_t ib = hb[P.start[c] + (p < nb ? p : 0u)];
    const uint32_t xh = ib << nlow;
    const double2 *__restrict__ tile = v1t + P.base[c] + (uint64_t)R * nb * HI_RT + rl;  // this tile's slab
    double ar = 0.0, ai = 0.0;
    for (int t = 0; t < A.e_high; t += UNR) {
      uint32_t q[UNR];
#pragma unroll
      for (int u2 = 0; u2 < UNR; u2++) {
        const GTerm T = load_term(th + t + u2);
        const bool act = valid & pred32(T.kind, (uint32_t)T.tmask, (uint32_t)T.tval, xh);
        q[u2] = act ? blkpos[(xh ^ (uint32_t)T.flip) >> nlow] : ~0u;
      }
      double2 g[UNR];
#pragma unroll
      for (int u2 = 0; u2 < UNR; u2++) {
        g[u2] = make_double2(0.0, 0.0);
        if (q[u2] < nb) g[u2] = tile[(uint64_t)q[u2] * HI_RT];
      }
#pragma unroll
      for (int u2 = 0; u2 < UNR; u2++) {
        const GTerm T = load_term(th + t + u2);
        double2 gg = g[u2];
        if ((uint32_t)T.sgn != 0u && (__popc(xh & (uint32_t)T.sgn) & 1)) {
          gg.x = -gg.x;
          gg.y = -gg.y;
        }
        cmac<REALC>(ar, ai, T.cre, T.cim, gg);
      }
    }
    if (valid) {  // after k_hv_canon in stream order: accumulate onto its output
      const uint64_t j = A.l22[ib] - 1 + r;
      const double2 a = tile[(uint64_t)p * HI_RT];
      const double2 o = v0[j];
      v0[j] = make_double2(o.x + A.c1 * ar, o.y + A.c1 * ai);
      pre += a.x * ar + a.y * ai;
      pim += a.x * ai - a.y * ar;
    }
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    pre += __shfl_xor(pre, off, 64);
    pim += __shfl_xor(pim, off, 64);
  }
  __shared__ double2 red[4];
  if ((tid & 63) == 0) red[tid >> 6] = make_double2(pre, pim);
  __syncthreads();
  if (tid == 0) {
    double2 s2 = red[0];
    for (int w = 1; w < 4; w++) {
      s2.x += red[w].x;
      s2.y += red[w].y;
    }
    partial[blockIdx.x] = s2;
  }
}

#define HI_UNR 8  // high terms per batch of the high-half pass (16: 95-101 ms against 93.5 on one box)

// sectors of at least this many states take the high-half pass by default
// (J1-J2 7x4, 40M states: 5.19 ms either way; 6x4, 2.7M: 0.29 against 0.29 ms)
#define HI_AUTO_DIM ((uint64_t)1 << 27)

// Hubbard orbitals 2i + sigma: up electrons at even bit positions
#define HI_EVEN 0x5555555555555555ull

int hphi_hi_wanted(const hphi_engine *e) {
  const bool hub = e->model == HPHI_MODEL_HUBBARD;
  if (e->hi_split == 0 || e->hi_ok == 0 || !(e->model == HPHI_MODEL_SPIN || hub) || !e->c2_ok || !e->use_c2)
    return 0;
  const int nhigh = e->nbits - e->nlow;
  if (e->r_high == 0 || e->nlow > 16 || nhigh < 1 || nhigh > 16) return 0;
  // every high term moves one particle (spin) between two high-half bits of
  // one species, so the block changes and its class and the rank do not:
  // an exchange or hop pair (P_DIFF) or one hop (P_EQ: exactly one of the two
  // bits set); padding terms have an empty mask
  for (int t = e->n_low; t < e->n_low + e->n_high; t++) {
    const GTerm &x = e->h_terms_dev[t];
    if (x.kind == P_EQ && x.tmask == 0) continue;
    const bool two = x.tmask == x.flip && __builtin_popcountll(x.flip) == 2;
    const bool kind_ok = x.kind == P_DIFF || (x.kind == P_EQ && __builtin_popcountll(x.tval) == 1);
    const bool species = !hub || __builtin_popcountll(x.flip & HI_EVEN) != 1;
    if (!two || !kind_ok || !species || (x.flip & e->irght) || (x.sgn & e->irght)) return 0;
  }
  // by default Spin only: on the 4x4 Hubbard sector the two passes took 14.9 ms against 14.0
  return e->hi_split == 1 || (!hub && e->dim >= HI_AUTO_DIM);
}

// The plan: the non-empty blocks from list_2_2 grouped by class, checked to be
// laid out as the pass assumes (consecutive in block order, every block of a
// class the same size); a sector that does not match, or a device without
// room for the tile-major copy, keeps the single pass.
int hphi_hi_split_on(hphi_engine *e) {
  if (!hphi_hi_wanted(e)) return 0;
  if (e->d_v1t) return 1;
  const int nhigh = e->nbits - e->nlow;
  const uint32_t nb = 1u << nhigh;
  if (e->size22 < nb) {
    e->hi_ok = 0;
    return 0;
  }
  std::vector<uint64_t> l22(nb);
  if (hipMemcpyAsync(l22.data(), e->d_l22, sizeof(uint64_t) * nb, hipMemcpyDeviceToHost, e->stream) != hipSuccess ||
      hipStreamSynchronize(e->stream) != hipSuccess)
    return hphi_set_error("high-half pass: reading list_2_2 failed");
  const bool hub = e->model == HPHI_MODEL_HUBBARD;
  const uint64_t even = HI_EVEN >> (e->nlow & 1);  // up positions of the high half's own bits
  auto key = [&](uint32_t ib) -> uint32_t {
    return hub ? (uint32_t)(__builtin_popcountll(ib & even) * 64 + __builtin_popcountll(ib & ~even & (nb - 1)))
               : (uint32_t)__builtin_popcount(ib);
  };
  // block sizes from consecutive offsets (non-empty blocks are consecutive in block order)
  std::vector<uint32_t> ne;
  for (uint32_t ib = 0; ib < nb; ib++)
    if (l22[ib]) ne.push_back(ib);
  if (ne.empty() || l22[ne[0]] != 1) {
    e->hi_ok = 0;
    return 0;
  }
  std::map<uint32_t, uint64_t> csize;  // class key -> block size
  for (size_t i = 0; i < ne.size(); i++) {
    const uint64_t end = i + 1 < ne.size() ? l22[ne[i + 1]] : e->dim + 1;
    if (end <= l22[ne[i]]) {
      e->hi_ok = 0;
      return 0;
    }
    const uint64_t sz = end - l22[ne[i]];
    auto it = csize.emplace(key(ne[i]), sz).first;
    if (it->second != sz || sz > 0xffffffffull) {
      e->hi_ok = 0;
      return 0;
    }
  }
  if (csize.size() > HI_MAXC) {
    e->hi_ok = 0;
    return 0;
  }
  HiPlan P{};
  std::vector<uint32_t> hb, pos(nb, ~0u);
  std::vector<ulonglong2> tofs(nb, make_ulonglong2(0, 0));
  P.ncls = (int)csize.size();
  uint64_t w = 0, base = 0;
  int c = 0;
  for (const auto &kv : csize) {
    P.start[c] = (uint32_t)hb.size();
    for (uint32_t ib : ne)
      if (key(ib) == kv.first) {
        pos[ib] = (uint32_t)hb.size() - P.start[c];
        hb.push_back(ib);
      }
    P.nblk[c] = (uint32_t)hb.size() - P.start[c];
    P.size[c] = (uint32_t)kv.second;
    for (uint32_t q = 0; q < P.nblk[c]; q++)
      tofs[hb[P.start[c] + q]] = make_ulonglong2(base + (uint64_t)q * HI_RT, (uint64_t)P.nblk[c] * HI_RT);
    P.ngrp[c] = (P.nblk[c] + HI_BPG - 1) / HI_BPG;
    P.wgp[c] = w;
    P.base[c] = base;
    const uint64_t nr = (P.size[c] + HI_RT - 1) / HI_RT;
    w += nr * P.ngrp[c];
    base += nr * P.nblk[c] * HI_RT;
    c++;
  }
  P.wgp[P.ncls] = w;
  P.base[P.ncls] = base;
  if (hipMalloc(&e->d_hb, sizeof(uint32_t) * hb.size()) != hipSuccess ||
      hipMalloc(&e->d_blkpos, sizeof(uint32_t) * nb) != hipSuccess ||
      hipMalloc(&e->d_v1t, sizeof(double2) * base) != hipSuccess ||
      hipMalloc(&e->d_tofs, sizeof(ulonglong2) * nb) != hipSuccess ||
      hipMemcpy(e->d_tofs, tofs.data(), sizeof(ulonglong2) * nb, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(e->d_hb, hb.data(), sizeof(uint32_t) * hb.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(e->d_blkpos, pos.data(), sizeof(uint32_t) * nb, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipFree(e->d_hb);
    (void)hipFree(e->d_blkpos);
    (void)hipFree(e->d_v1t);
    (void)hipFree(e->d_tofs);
    e->d_hb = nullptr;
    e->d_blkpos = nullptr;
    e->d_v1t = nullptr;
    e->d_tofs = nullptr;
    e->hi_ok = 0;  // no room for the tile-major copy: the single pass serves this engine
    return 0;
  }
  e->bytes += sizeof(uint32_t) * (hb.size() + nb) + sizeof(double2) * base + sizeof(ulonglong2) * nb;
  e->hi_plan = P;
  e->hi_ok = 1;
  return 1;
}

int hphi_hv_launch_high(hphi_engine *e, const HvArgs &A, double2 *partial, int *grid) {
  uint64_t g = (uint64_t)e->hi_wgs_per_cu * std::max(e->num_cu, 1);
  g = std::max<uint64_t>((g + 7) / 8 * 8, 8);
  if (g > HV_MAX_GRID) g = HV_MAX_GRID;
  if (!A.v1t)  // the first pass did not write the tile-major copy
    hipLaunchKernelGGL(k_hi_copy, dim3((unsigned)g), dim3(256), 0, e->stream, (const double2 *)A.v1, A.l22,
                       e->hi_plan, (const uint32_t *)e->d_hb, e->d_v1t);
  HvArgs B = A;
  B.e_high = std::min(A.n_high, (e->r_high + HI_UNR - 1) / HI_UNR * HI_UNR);
  static_assert(HV_TERM_PAD % HI_UNR == 0, "a batch must not cross a term group");
  if (e->c2_realc)
    hipLaunchKernelGGL((k_hv_high<true, HI_UNR>), dim3((unsigned)g), dim3(256), 0, e->stream, B, e->hi_plan,
                       (const uint32_t *)e->d_hb, (const uint32_t *)e->d_blkpos, (const double2 *)e->d_v1t, partial);
  else
    hipLaunchKernelGGL((k_hv_high<false, HI_UNR>), dim3((unsigned)g), dim3(256), 0, e->stream, B, e->hi_plan,
                       (const uint32_t *)e->d_hb, (const uint32_t *)e->d_blkpos, (const double2 *)e->d_v1t, partial);
  HIPCHK(hipGetLastError());
  *grid = (int)g;
  return 0;
}

// Which terms are real and whether the list_2_1 table fits LDS as uint16.
int hphi_canon_config(hphi_engine *e) {
  e->c2_realc = 1;
  for (const GTerm &t : e->h_terms_dev)
    if (t.cim != 0.0) e->c2_realc = 0;
  int maxlds = 0;
  if (hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, 0) != hipSuccess) maxlds = 65536;
  e->c2_max_lds = maxlds;
  size_t need = 2 * (size_t)e->size21;
  // rank + 2 must fit uint16: the largest class has C(nlow, nlow/2) members (spin)
  bool small_rank = e->nlow <= 16;
  e->c2_lds21 = small_rank && need + 1024 <= (size_t)maxlds;
  e->c2_ok = e->nbits <= 32 && !hphi_model_gc(e->model);
  e->c2_crank_ok = e->c2_ok && (e->model == HPHI_MODEL_SPIN || e->model == HPHI_MODEL_HUBBARD) && e->nlow <= 16;
  return 0;
}

// t[b] (b < 256): sum over the k-th set bit of b (k = 1, 2, ...) at position p
// of C(p, k); t[256 + 9 h + c]: the same for the bits of h at positions 8 + q
// with k counted from c + 1 (c = set bits below position 8).  For a canonical
// spin sector list_2_1[ia] - 2 is the rank of ia among the values of its
// popcount in ascending order, which is this sum (combinatorial number system).
void hphi_crank_tables(uint16_t *t) {
  uint32_t C[17][17] = {};
  for (int n = 0; n <= 16; n++) {
    C[n][0] = 1;
    for (int k = 1; k <= n; k++) C[n][k] = C[n - 1][k - 1] + (k <= n - 1 ? C[n - 1][k] : 0);
  }
  for (int b = 0; b < 256; b++) {
    uint32_t r = 0;
    int k = 0;
    for (int p = 0; p < 8; p++)
      if ((b >> p) & 1) r += C[p][++k];
    t[b] = (uint16_t)r;
  }
  for (int h = 0; h < 256; h++)
    for (int c = 0; c <= 8; c++) {
      uint32_t r = 0;
      int k = c;
      for (int q = 0; q < 8; q++)
        if ((h >> q) & 1) r += C[8 + q][++k];
      t[256 + 9 * h + c] = (uint16_t)r;
    }
}

// Hubbard: list_2_1[ia] - 2 = sum over the set bits p of ia of
// C(ceil(p/2), up bits at positions <= p) C(floor(p/2), down bits at positions
// <= p) (the class rank the sz kernel builds, C = 0 outside 0..n).  t[b]: the
// bits of the low byte; t[256 + 25 h + 5 u + d]: the bits of the high byte h
// given u up and d down bits in the low byte.
void hphi_hrank_tables(uint16_t *t) {
  auto C = [](int n, int k) -> uint32_t {
    if (k < 0 || k > n) return 0;
    uint32_t r = 1;
    for (int i = 1; i <= k; i++) r = r * (uint32_t)(n - k + i) / (uint32_t)i;
    return r;
  };
  auto part = [&](uint32_t bits, int p0, int u0, int d0) -> uint32_t {
    // bits of one byte at positions p0..p0+7, counts below p0 given
    uint32_t r = 0;
    int u = u0, d = d0;
    for (int q = 0; q < 8; q++) {
      if (!((bits >> q) & 1)) continue;
      const int p = p0 + q;
      if ((p & 1) == 0) u++; else d++;
      r += C((p + 1) / 2, u) * C(p / 2, d);
    }
    return r;
  };
  for (int b = 0; b < 256; b++) t[b] = (uint16_t)part((uint32_t)b, 0, 0, 0);
  for (int h = 0; h < 256; h++)
    for (int u = 0; u <= 4; u++)
      for (int d = 0; d <= 4; d++) t[256 + 25 * h + 5 * u + d] = (uint16_t)part((uint32_t)h, 8, u, d);
}

template <int TPB, bool REALC, int RK, int UNR>
static int launch_one(hphi_engine *e, const HvArgs &A, int g, int skip_high) {
  const size_t lds = RK == 1 ? 2 * (size_t)e->size21 : RK == 2 ? 2 * (size_t)CRANK_N : RK == 3 ? 2 * (size_t)HRANK_N : 0;
  if (lds > 65536)
    (void)hipFuncSetAttribute((const void *)k_hv_canon<TPB, REALC, RK, UNR>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  // each group's loop stops at its last batch holding a compiled term (the
  // groups are padded to HV_TERM_PAD, a multiple of UNR)
  static_assert(HV_TERM_PAD % UNR == 0, "a batch must not cross a term group");
  HvArgs B = A;
  B.e_low = std::min(A.n_low, (e->r_low + UNR - 1) / UNR * UNR);
  B.e_high = skip_high ? 0 : std::min(A.n_high, (e->r_high + UNR - 1) / UNR * UNR);
  B.e_mixed = std::min(A.n_mixed, (e->r_mixed + UNR - 1) / UNR * UNR);
  hipLaunchKernelGGL((k_hv_canon<TPB, REALC, RK, UNR>), dim3(g), dim3(TPB), lds, e->stream, B);
  HIPCHK(hipGetLastError());
  return 0;
}

template <int TPB, bool REALC>
static int launch_rk(hphi_engine *e, const HvArgs &A, int g, int sk) {
  const int rk = e->c2_crank ? (e->model == HPHI_MODEL_HUBBARD ? 3 : 2) : (e->c2_lds21 ? 1 : 0);
  if (e->c2_unroll == 16) {
    if (rk == 3) return launch_one<TPB, REALC, 3, 16>(e, A, g, sk);
    if (rk == 2) return launch_one<TPB, REALC, 2, 16>(e, A, g, sk);
    if (rk == 1) return launch_one<TPB, REALC, 1, 16>(e, A, g, sk);
    return launch_one<TPB, REALC, 0, 16>(e, A, g, sk);
  }
  if (rk == 3) return launch_one<TPB, REALC, 3, 8>(e, A, g, sk);
  if (rk == 2) return launch_one<TPB, REALC, 2, 8>(e, A, g, sk);
  if (rk == 1) return launch_one<TPB, REALC, 1, 8>(e, A, g, sk);
  return launch_one<TPB, REALC, 0, 8>(e, A, g, sk);
}

int hphi_hv_launch_canon(hphi_engine *e, const HvArgs &A, int g, int skip_high) {
  const int tpb = e->c2_tpb;
  if (e->c2_crank && !A.crank) return hphi_set_error("rank tables are not built");
#define SEL(TPB)                                                                   \
  if (tpb == TPB) return e->c2_realc ? launch_rk<TPB, true>(e, A, g, skip_high) : launch_rk<TPB, false>(e, A, g, skip_high);
  SEL(256)
  SEL(512)
  SEL(1024)
#undef SEL
  return hphi_set_error("unsupported workgroup size %d", tpb);
}
