// Internal types of the hphi_amd engine (device term tables, kernel argument
// blocks, engine state).  Not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "hphi_amd.h"

#define HV_TPB 256        // threads per workgroup of the H·v kernels
#define HV_UNROLL 8       // terms per unrolled gather batch of the generic kernels
#define HV_TERM_PAD 16    // term groups are padded with never-active terms to a multiple of
                          // this, so every batch size in use (8, 16) reads whole groups
#define HV_MAX_GRID 65536 // upper bound of H·v workgroups (sizes the partial-sum buffer)

// predicate kinds of a gather term, evaluated on the TARGET state x
// grand-canonical models index the vector by the state's bits (no tables)
static inline bool hphi_model_gc(int m) { return m == HPHI_MODEL_SPINGC || m == HPHI_MODEL_HUBBARDGC; }
// Hubbard models: site i's orbitals are bits 2i (up) and 2i+1 (down)
static inline bool hphi_model_hubbard(int m) { return m == HPHI_MODEL_HUBBARD || m == HPHI_MODEL_HUBBARDGC; }

enum { P_EQ = 0,    // (x & tmask) == tval            (hopping, transverse field)
       P_DIFF = 1,  // popcount(x & tmask) == 1       (exchange: the two spins differ)
       P_SAME = 2   // popcount(x & tmask) != 1       (pair lift: the two spins agree)
};

// one off-diagonal term in gather form: source = x ^ flip, weight
// (cre + i cim) * (-1)^popcount(x & sgn) when the predicate holds
struct GTerm {
  uint64_t flip, tmask, tval, sgn;
  double cre, cim;
  int kind, pad;
};

// Term tables are read through the constant address space: the loads are
// known invariant, so uniform accesses become scalar (SMEM) loads instead of
// per-lane vector loads that would queue behind the gathers.
typedef const GTerm __attribute__((address_space(4))) *CTerm;

__device__ __forceinline__ GTerm load_term(CTerm p) {
  GTerm t;
  t.flip = p->flip;
  t.tmask = p->tmask;
  t.tval = p->tval;
  t.sgn = p->sgn;
  t.cre = p->cre;
  t.cim = p->cim;
  t.kind = p->kind;
  t.pad = 0;
  return t;
}

// host-side term compilers shared by the Hamiltonian and the Green's functions:
// bits strictly between orbitals a and b; V c+_{a1} c_{b1} c+_{a2} c_{b2}
// (fermions, Jordan-Wigner sign); V (|s1><s2|)_A (|s3><s4|)_B (spin-1/2)
uint64_t hphi_between_mask(int a, int b);
void hphi_compile_two_hops(int a1, int b1, int a2, int b2, double vre, double vim, std::vector<GTerm> &out);
void hphi_compile_spin_pair(int A, int s1, int s2, int B, int s3, int s4, double vre, double vim,
                            std::vector<GTerm> &out);

// diagonal term kinds (diagonalcalc.c set functions, per model)
enum { D_INTRA = 0, D_CHEMI_H, D_CHEMI_S, D_INTER_H, D_INTER_S, D_HUND_H, D_HUND_S, D_NN, D_SS };

struct DTerm {
  int kind, pad;
  uint64_t m1, m2, m3, m4;
  double V;
};

struct SzParams {
  int hubbard, nlow, nhigh, nup, ndn;
  int nlu, nld, nhu, nhd;  // Hubbard: up/down positions in the low / high half
  uint64_t sdim, ihfbit, size21, dim;
};

struct HvArgs {
  uint64_t dim;
  const uint64_t *list1;
  const uint32_t *l21;
  const uint64_t *l22;
  const double *diag;
  const GTerm *terms;
  int n_low, n_high, n_mixed;  // group sizes (padded): the groups' offsets
  int e_low, e_high, e_mixed;  // loop extents: terms up to the launch's batch (<= the group sizes)
  int nlow;
  uint64_t irght;
  double *v0;
  const double *v1;
  double2 *partial;
  int xcd_map;
  uint64_t size21;
  int overwrite;  // 1: v0 = c1 H v1 (v0 not read); 0: v0 = c0 v0 + c1 H v1 (reference: c0 = c1 = 1)
  double c0, c1;  // scale factors of the output (prdct stays <v1|H v1>)
  // canonical Spin: list_2_1[ia] = rank of ia among the values of its popcount
  // + 2, from two byte tables (hphi_crank_tables)
  const uint16_t *crank;
  // k_hv_canon also writes the tile-major copy of v1 for the high-half pass
  // when v1t is set: state (block ib, rank r) at tofs[ib].x + (r / HI_RT) tofs[ib].y + r % HI_RT
  double2 *v1t;
  const ulonglong2 *tofs;
};

// byte tables of the combinatorial rank (spin, nlow <= 16): rank(ia) =
// t[ia & 255] + t[256 + 9 * ((ia >> 8) & 255) + popcount(ia & 255)]
#define CRANK_N (256 + 256 * 9)

// list_2_1[ia] (rank + 2) of a canonical spin sector from the byte tables t
// staged in LDS (hphi_crank_tables)
__device__ __forceinline__ int64_t crank_l21(const uint16_t *t, uint32_t ia) {
  return (int64_t)t[ia & 255u] + (int64_t)t[256u + 9u * ((ia >> 8) & 255u) + (uint32_t)__popc(ia & 255u)] + 2;
}

// canonical Hubbard (up orbitals even bits, down odd): the rank of ia among the
// values with its up and down counts, t[ia & 255] + t[256 + 25 h + 5 u + d]
// with u, d the up / down counts of the low byte (hphi_hrank_tables)
#define HRANK_N (256 + 256 * 25)
__device__ __forceinline__ int64_t hrank_l21(const uint16_t *t, uint32_t ia) {
  return (int64_t)t[ia & 255u] +
         (int64_t)t[256u + 25u * ((ia >> 8) & 255u) + 5u * (uint32_t)__popc(ia & 0x55u) + (uint32_t)__popc(ia & 0xAAu)] +
         2;
}

// Canonical Spin and Hubbard, high-half terms by rank tiles (hv_canon.hip):
// for each class c (blocks whose high half holds the same number of up
// spins, or of up and of down electrons) its blocks (high-half values,
// ascending), the block size, the block groups of HI_BPG blocks, the
// workgroups before the class, and where the class's rank tiles start in the
// tile-major copy of v1 (tile R of block position p at base[c] + (R nblk[c] + p) HI_RT)
#define HI_MAXC 96
#define HI_RT 32              // ranks per tile (8x4 J1-J2: 89.8 ms at 32, 90.5 at 16, 92.8 at 8; 64 no better)
#define HI_BPG (256 / HI_RT)  // blocks per workgroup
struct HiPlan {
  int ncls;
  uint32_t start[HI_MAXC], nblk[HI_MAXC], size[HI_MAXC], ngrp[HI_MAXC];
  uint64_t wgp[HI_MAXC + 1], base[HI_MAXC + 1];
};

struct hphi_engine {
  int model = 0, nsite = 0, nup = 0, ndown = 0, nbits = 0, nlow = 0;
  uint64_t dim = 0, sdim = 0, size21 = 0, size22 = 0, irght = 0, ilft = 0, ihfbit = 0;
  hipStream_t stream = nullptr;
  int num_cu = 256;
  int grid_per_cu = 8;  // H·v workgroups per CU (tunable)
  int xcd_map = 1;      // contiguous tile ranges per XCD (tunable)
  int xcd_tuned = 0;    // xcd_map was set through hphi_engine_tune
  bool tables_ready = false, diag_ready = false;
  uint64_t *d_list1 = nullptr, *d_l22 = nullptr, *d_jb = nullptr;
  uint32_t *d_l21 = nullptr;
  double *d_diag = nullptr;
  GTerm *d_terms = nullptr;
  DTerm *d_dterms = nullptr;
  std::vector<GTerm> h_terms, h_terms_dev;
  std::vector<DTerm> h_dterms;
  int n_low = 0, n_high = 0, n_mixed = 0;
  int r_low = 0, r_high = 0, r_mixed = 0;  // compiled terms per group, before padding
  void *d_partial = nullptr, *d_prdct = nullptr, *h_prdct = nullptr, *d_coef = nullptr;
  double *d_stage0 = nullptr, *d_stage1 = nullptr;
  // host-pointer calls: per worker thread a stream and two pinned bounce
  // buffers of HOST_CHUNK bytes, with the events that free them
  int n_host = 0;
  hipStream_t h_stream[16] = {};
  char *h_bounce[16][2] = {};
  hipEvent_t h_ev[16][2] = {};
  uint64_t bytes = 0;
  // canonical single-vector kernel (hv_canon.hip) configuration
  int c2_ok = 0, c2_realc = 1, c2_lds21 = 0, c2_max_lds = 0;
  int c2_tpb = 1024, c2_wgs_per_cu = 1, use_c2 = 1;
  int c2_crank = 0, c2_crank_ok = 0, c2_unroll = 8;
  uint16_t *d_crank = nullptr;
  // block kernel (hv_block.hip): columns per lane for k = 2, 4, 8, 16 (0: default),
  // workgroup size (0: by table size), workgroups per CU (0: by workgroup size)
  int bk_cpl[4] = {0, 0, 0, 0};
  int bk_tpb = 0, bk_wgs_per_cu = 0, use_blk = 1;
  int device = 0;  // the device current when the engine was created
  int lb_mfma = 1;  // LOBPCG block-8 Gram on the FP64 matrix cores
  // canonical Spin and Hubbard: high-half terms in their own pass over a tile-major copy of v1
  int hi_split = -1;      // -1: by sector size, 0: off, 1: on when the sector allows it
  int hi_ok = -1;         // -1: not checked yet, 0: the sector does not allow it
  int hi_wgs_per_cu = 6;  // workgroups per CU of that pass (8x4 J1-J2: 6 best of 4-16)
  int hi_fuse_copy = 1;   // the first pass writes the tile-major copy (0: a copy kernel does)
  int hi_p1_wgs_per_cu = 256;  // workgroups per CU of that first pass (8x4 J1-J2: 90.4 ms at 256, 93.6 at 48)
  uint32_t *d_hb = nullptr;      // blocks of each class (hi_plan)
  uint32_t *d_blkpos = nullptr;  // position of every block within its class (~0u: empty block)
  double2 *d_v1t = nullptr;      // tile-major copy of v1, hi_plan.base[ncls] elements
  ulonglong2 *d_tofs = nullptr;  // per block: its first tile-major element, and the stride between its tiles
  HiPlan hi_plan{};
};

extern "C" int hphi_set_error(const char *fmt, ...);
int hphi_build_tables(hphi_engine *e);
int hphi_build_diag(hphi_engine *e);
int hphi_hv_launch(hphi_engine *e, int k, double *v0, const double *v1);
int hphi_hv_launch_ex(hphi_engine *e, int k, double *v0, const double *v1, int overwrite);
int hphi_fetch_prdct(hphi_engine *e, int k, double *prdct);
int hphi_hv_grid(const hphi_engine *e, int k);
int hphi_canon_config(hphi_engine *e);
void hphi_crank_tables(uint16_t *t);
void hphi_hrank_tables(uint16_t *t);
int hphi_engine_prepare(const hphi_model_desc *d, hphi_engine *e);
int hphi_ensure_binom_table();
// off-diagonal elements of H in a canonical sector (0 for grand-canonical models)
uint64_t hphi_term_nnz(const hphi_engine *e);
SzParams hphi_sz_params(const hphi_engine *e);
int hphi_launch_tables(const SzParams &P, uint64_t sdim, uint64_t size21, uint64_t *jb, uint32_t *l21, uint64_t *l22, uint64_t *list1, hipStream_t st);
// diag[j * stride + col] for j < dim
int hphi_launch_diag(uint64_t dim, const uint64_t *list1, int gc, const DTerm *dterms, int nd, double *diag,
                     hipStream_t st, int stride = 1, int col = 0);
int hphi_hv_launch_canon(hphi_engine *e, const HvArgs &A, int g, int skip_high = 0);
// canonical Spin: is the high-half pass usable and wanted for this engine (host check only)
int hphi_hi_wanted(const hphi_engine *e);
// the same, building its plan and buffers on first use (-1: error)
int hphi_hi_split_on(hphi_engine *e);
// launches the high-half pass: accumulates into A.v0, partials at partial[0..*grid)
int hphi_hv_launch_high(hphi_engine *e, const HvArgs &A, double2 *partial, int *grid);
int hphi_blk_ok(const hphi_engine *e);
int hphi_blk_cpl_ok(int k, int cpl);
void hphi_blk_shape(const hphi_engine *e, int k, int *cpl, int *tpb, int *wgs_per_cu);
int hphi_blk_grid(const hphi_engine *e, int k);
// records the device address of the allocation-free entry point's scratch array
int hphi_noalloc_setup(int device);
int hphi_hv_launch_blk(hphi_engine *e, const HvArgs &A, int k, int *grid);
