/*
 * hphi_amd — MI355X-native matrix-free H·v engine for HPhi, C ABI.
 *
 * Plain C types only (pointers, sizes, doubles); complex128 data is passed as
 * interleaved (re, im) doubles, layout-compatible with C99 `double complex`.
 * Every entry point returns 0 on success and -1 on failure; the message is
 * available from hphi_last_error().  No entry point aborts the process.
 *
 * Reference interfaces replaced (paths relative to the HPhi source tree):
 *   hphi_sz            <- int sz(struct BindStruct*, long unsigned int *list_1,
 *                          long unsigned int *list_2_1, long unsigned int *list_2_2)
 *                          src/include/sz.h:20, src/sz.c:62
 *   hphi_diagonalcalc  <- int diagonalcalc(struct BindStruct*)
 *                          src/include/diagonalcalc.h:22, src/diagonalcalc.c:84
 *   hphi_mltply        <- int mltply(struct BindStruct*, double complex *tmp_v0,
 *                          double complex *tmp_v1)   src/include/mltply.h:22, src/mltply.c:56
 *   hphi_engine_create_from_def
 *                      <- ReadDefFileNInt + ReadDefFileIdxPara (src/readdef.c:463,1006)
 *                          followed by check() (src/check.c:51) for the sizes
 *   hphi_lanczos       <- CalcByLanczos -> Lanczos_EigenValue (src/CalcByLanczos.c:55,
 *                          src/Lanczos_EigenValue.c:50), eigenvalues only
 *   hphi_tpq           <- CalcByTPQ (src/CalcByTPQ.c:52) step loop with
 *                          Multiply (src/Multiply.c:42) and expec_energy_flct
 *                          (src/expec_energy_flct.c:34), k samples at once
 *   hphi_lobpcg        <- CalcByLOBPCG (src/CalcByLOBPCG.c:328), eigenvalues only
 */
#ifndef HPHI_AMD_H
#define HPHI_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* calcmod.def CalcModel values (src/include/DefCommon.h:22-37) */
#define HPHI_MODEL_HUBBARD 0
#define HPHI_MODEL_SPIN 1
#define HPHI_MODEL_HUBBARDGC 3
#define HPHI_MODEL_SPINGC 4

typedef struct hphi_engine hphi_engine;

/* Term tables in the order the reference holds them after ReadDefFileIdxPara
 * (struct DefineList, src/include/struct.h:43-310).  Sites are 0-based.
 *   exchange  ExchangeCoupling[n][2], ParaExchangeCoupling[n]
 *   pairlift  PairLiftCoupling[n][2], ParaPairLiftCoupling[n]       (SpinGC)
 *   transfer  EDGeneralTransfer[n][4] (i, s1, j, s2), EDParaGeneralTransfer[n]
 *             as (re, im); term is -t c+_{i s1} c_{j s2}
 *   chemi     EDChemi[n], EDSpinChemi[n], EDParaChemi[n]
 *   intra     CoulombIntra[n], ParaCoulombIntra[n]
 *   inter     CoulombInter[n][2], ParaCoulombInter[n]   (Ising part included)
 *   hund      HundCoupling[n][2], ParaHundCoupling[n]   (Ising part included)
 * nup/ndown: Spin — nup = number of 1-bits (X->Def.Ne); Hubbard — electrons
 * per spin (X->Def.Nup, X->Def.Ndown); SpinGC — ignored. */
typedef struct {
  int model, nsite, nup, ndown;
  int n_exchange;  const int *exchange_ij;  const double *exchange_J;
  int n_pairlift;  const int *pairlift_ij;  const double *pairlift_J;
  int n_transfer;  const int *transfer_idx; const double *transfer_t;
  int n_chemi;     const int *chemi_site;   const int *chemi_spin; const double *chemi_mu;
  int n_intra;     const int *intra_site;   const double *intra_U;
  int n_inter;     const int *inter_ij;     const double *inter_V;
  int n_hund;      const int *hund_ij;      const double *hund_J;
  /* Hubbard two-body terms as the reference's reader leaves them
   * (InputInterAllInfo readdef.c:3054, GetDiagonalInterAll :2415-2530):
   *   interall       InterAll_OffDiagonal[n][8] (i, s1, j, s2, k, s3, l, s4) for
   *                  V c+_{i s1} c_{j s2} c+_{k s3} c_{l s4}, V as (re, im)
   *   interall_diag  InterAll_Diagonal[n][4] (i, s1, k, s3) for V n_{i s1} n_{k s3}
   *   pairhop        PairHopping[n][2] (both directions, readdef.c:1257-1267), J
   * For Hubbard, exchange_ij/exchange_J hold the Hubbard ExchangeCoupling
   * (mltplyHubbardCore.c:472-525). */
  int n_interall;      const int *interall_idx;      const double *interall_V;
  int n_interall_diag; const int *interall_diag_idx; const double *interall_diag_V;
  int n_pairhop;       const int *pairhop_ij;        const double *pairhop_J;
} hphi_model_desc;

/* Modelling parameters read from modpara.def (src/readdef.c:555-650) and the
 * driver flags of calcmod.def (src/readdef.c:233-320): restart = ReStart
 * (0 none, 1 out, 2 in+out, 3 in), initial_vec_type = InitialVecType (0
 * uniform complex, 1 real, -1 Gaussian), calc_eigenvec, output_eigenvec,
 * input_eigenvec as there; expand_coef = ExpandCoef (0: not given). */
typedef struct {
  int calc_type, lanczos_max, initial_iv, exct, lanczos_eps, lanczos_target, num_ave,
      expec_interval;
  double large_value;
  int restart, initial_vec_type, calc_eigenvec, output_eigenvec, input_eigenvec, expand_coef;
} hphi_modpara;

const char *hphi_last_error(void);

/* ---- engine ------------------------------------------------------------ */
int hphi_engine_create(const hphi_model_desc *desc, hphi_engine **out);
/* Reads an expert-mode namelist.def and the files it names (subset: CalcMod,
 * ModPara, LocSpin, Trans, CoulombIntra, CoulombInter, Hund, Exchange, Ising,
 * PairLift); modpara may be NULL. */
int hphi_engine_create_from_def(const char *namelist_path, hphi_engine **out, hphi_modpara *modpara);
void hphi_engine_destroy(hphi_engine *e);
/* Host-only parse of an expert-mode input into a descriptor (no device work);
 * release it with hphi_model_desc_free. */
int hphi_read_def(const char *namelist_path, hphi_model_desc **desc, hphi_modpara *modpara);
void hphi_model_desc_free(hphi_model_desc *desc);
/* dim = idim_max; size21/size22 = lengths of list_2_1/list_2_2 (xsetmem.c:284-318);
 * irght/ilft/ihfbit = split masks (bitcalc.c:45-62). */
int hphi_engine_info(const hphi_engine *e, uint64_t *dim, uint64_t *size21, uint64_t *size22,
                     uint64_t *irght, uint64_t *ilft, uint64_t *ihfbit);
/* Launch-shape knobs of the H·v kernels (set by the autotuner or by hand):
 *   "tpb" 256|512|1024, "wgs_per_cu" 1..256, "lds21" 0|1 (list_2_1 staged in
 *   LDS), "xcd_map" 0|1 (contiguous state ranges per XCD), "use_c2" 0|1
 *   (32-bit canonical kernel), "grid_per_cu" 1..64 (generic kernel);
 *   two-pass H·v of canonical sectors (high-half terms over a tile-major copy
 *   of v1, one extra vector of device memory): "hi_split" -1|0|1 (-1: Spin
 *   sectors of at least 2^27 states), "hi_wgs_per_cu" 1..64 (high pass),
 *   "hi_p1_wgs_per_cu" 1..256 (first pass), "hi_fuse_copy" 0|1;
 *   block kernel (k > 1): "use_blk" 0|1, "blk_cpl2|4|8|16" columns per lane
 *   (1, 4 or k; 0 = default), "blk_tpb" 0|256|1024, "blk_wgs_per_cu" 0..64;
 *   LOBPCG: "lb_mfma" 0|1 (block-8 Gram matrices on the FP64 matrix cores). */
int hphi_engine_tune(hphi_engine *e, const char *knob, int value);
int hphi_engine_get_tune(const hphi_engine *e, const char *knob, int *value);
/* Times every candidate launch shape on the caller's device vectors (k = block
 * size; v0 accumulates, re-initialise it afterwards) and keeps the fastest;
 * one line per candidate goes to `report` when non-NULL. */
int hphi_autotune(hphi_engine *e, int k, double *v0, const double *v1, int reps, char *report,
                  size_t report_len);
/* Host only (no device work): dim of the sector and the number of nonzero
 * off-diagonal elements of H in it, counted per compiled term by
 * combinatorics (canonical Spin and Hubbard models); the algorithmic bytes of
 * one H·v are 16 per off-diagonal element plus the per-state streams. */
int hphi_model_nnz(const hphi_model_desc *desc, uint64_t *dim, uint64_t *nnz);
/* HBM bytes the engine holds (tables, diagonal, scratch). */
uint64_t hphi_engine_device_bytes(const hphi_engine *e);

/* ---- basis, diagonal, H·v (drop-in for sz / diagonalcalc / mltply) ----- */
/* Builds list_1 / list_2_1 / list_2_2 on the device (bit-exact with sz()).
 * Host arrays, when non-NULL, receive the reference-shaped copies:
 * list_1[0..dim] (1-indexed, [0] = 0), list_2_1[0..size21-1], list_2_2[0..size22-1]. */
int hphi_sz(hphi_engine *e, uint64_t *list_1, uint64_t *list_2_1, uint64_t *list_2_2);
/* Precomputes list_Diagonal on the device (bit-exact with diagonalcalc());
 * a non-NULL host array receives list_Diagonal[0..dim] ([0] = 0). */
int hphi_diagonalcalc(hphi_engine *e, double *list_Diagonal);
/* v0[1..dim] += H v1[1..dim], prdct[2] = <v1|H|v1>.  v0/v1 are the reference's
 * 1-indexed arrays of dim+1 complex128 and may be host or device memory
 * (host memory is staged through the device: PCIe-bound).  Builds tables and
 * diagonal on first use. */
int hphi_mltply(hphi_engine *e, double *v0, const double *v1, double *prdct);
/* Device-resident form: v0/v1 are 0-indexed device arrays of dim complex128. */
int hphi_mltply_dev(hphi_engine *e, double *v0, const double *v1, double *prdct);
/* Scaled device form: v0 = c0 v0 + c1 H v1 (v0 = c1 H v1 with overwrite, v0
 * not read), prdct = <v1|H|v1>.  No counterpart in the reference: the device
 * Lanczos folds its normalisation 1/beta and -beta into this launch. */
int hphi_mltply_dev_scaled(hphi_engine *e, double *v0, const double *v1, double c0, double c1, int overwrite,
                           double *prdct);
/* Block form: k vectors stored state-major, V[(j*k + v)*2 + {0,1}], device
 * memory; prdct[2k].  k in {1, 2, 4, 8, 16}. */
int hphi_mltply_block_dev(hphi_engine *e, int k, double *V0, const double *V1, double *prdct);
/* Allocation-free form of hphi_mltply_block_dev (same kernels, same results,
 * bit for bit).  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  It calls no hipMalloc/hipFree:
 * the prdct partial sums live in a fixed per-device array of 2^18 + 16
 * complex128 (4 MiB, held until the process ends), so a launch of g
 * workgroups needs g*k <= 2^18 (every default and tuned shape on <= 256 CUs
 * with <= 64 workgroups per CU); larger launches are rejected with an error.
 * hphi_sz and hphi_diagonalcalc must have run on the engine (no lazy build),
 * and the engine's device must be current.  Calls that overlap on one device
 * share that array and race. */
int hphi_mltply_block_dev_noalloc(hphi_engine *e, int k, double *V0, const double *V1, double *prdct);

/* ---- device memory and vector helpers for host drivers ----------------- */
int hphi_dev_alloc(size_t bytes, void **ptr);
int hphi_dev_free(void *ptr);
int hphi_memcpy_h2d(void *dst, const void *src, size_t bytes);
int hphi_memcpy_d2h(void *dst, const void *src, size_t bytes);
int hphi_dev_zero(void *ptr, size_t bytes);
int hphi_synchronize(void);
/* Synthetic vector: v[j*k+v] = (2u(2c)-1) + i(2u(2c+1)-1), c = j (k = 1) with
 * u from SplitMix64 of seed (seed+v for column v of a block).  Device memory. */
int hphi_fill_random(double *v, uint64_t dim, int k, uint64_t seed);

/* Streaming kernels on the engine's stream for the host drivers; vectors are
 * device arrays of dim (x k) complex128.  Results are copied to host memory.
 *   lanczos_beta    beta2 = sum |v0 - alpha v1|^2      (Lanczos_EigenValue.c:170-177)
 *   lanczos_update  v1 <- (v0 - alpha v1)/beta, v0 <- -beta v1_old  (:154-160)
 *   norm2 / dot     per-column squared norms / <x_c|y_c> of [dim][k] blocks
 *   axpby           y_c <- a_c x_c + b_c y_c, ab = {a_0, b_0, a_1, b_1, ...} complex */
int hphi_vec_lanczos_beta(hphi_engine *e, const double *v0, const double *v1, double alpha, double *beta2);
/* Residual of the recurrence with the normalisation deferred (the device
 * Lanczos): r <- r - f w, r2 = |r|^2; with vg non-NULL also vg += cg w.  With
 * r = H q_m - beta_{m-1} q_{m-1}, w = q_m / s_m and f = alpha_m s_m this is
 * (Lanczos_EigenValue.c:154-177) with q_{m+1} = r / beta_m left unscaled. */
int hphi_vec_lanczos_resid(hphi_engine *e, double *r, const double *w, double *vg, double f, double cg, double *r2);
int hphi_vec_lanczos_update(hphi_engine *e, double *v0, double *v1, double alpha, double beta);
/* the update, and vg += c * v1_new (Lanczos_EigenVector.c:170-185) */
int hphi_vec_lanczos_update_acc(hphi_engine *e, double *v0, double *v1, double *vg, double alpha, double beta,
                                double c);
int hphi_vec_norm2(hphi_engine *e, const double *v, int k, double *out);
int hphi_vec_dot(hphi_engine *e, const double *x, const double *y, int k, double *out);
int hphi_vec_axpby(hphi_engine *e, const double *x, double *y, int k, const double *ab);
int hphi_engine_sync(hphi_engine *e);

/* Times `reps` device H·v calls on the engine's stream with HIP events
 * (after `warmup` untimed calls); avg_ms = mean per call. */
int hphi_time_mltply_dev(hphi_engine *e, int k, double *v0, const double *v1, int warmup,
                         int reps, double *avg_ms);

/* ---- solvers (host C drivers, device-resident vectors) ----------------- */
typedef struct {
  int max_iter;      /* Lanczos_max */
  int n_eigen;       /* number of lowest distinct eigenvalues wanted (exct) */
  double eps;        /* relative convergence of the target eigenvalue (10^-LanczosEps) */
  int target;        /* LanczosTarget (0-based index of the converged value tested) */
  uint64_t seed;     /* initial vector seed */
  int check_every;   /* tridiagonal solve interval (1 = every step, as the reference) */
  int mode;          /* 0: the reference's test on E[target]; 1: the lowest n_eigen distinct
                        eigenvalues (Cullum-Willoughby filtered) change by < eps relative and
                        their Ritz residuals are <= 0.1 sqrt(eps) max(1, |E|) */
} hphi_lanczos_params;

typedef struct {
  double eigenvalues[16];
  int n_found, iterations, converged;
  double seconds, hv_seconds;
} hphi_lanczos_result;

int hphi_lanczos(hphi_engine *e, const hphi_lanczos_params *p, hphi_lanczos_result *r);
/* Lanczos with the reference's restart files (ReStart, Lanczos_EigenValue.c:78-100,
 * 282-287, 428-489): with tm_in / vec_in (a TMComponents file and a recalcvec file
 * of the same step n) the recurrence continues from step n for up to max_iter
 * further steps; with tm_out / vec_out the components and vectors of the last
 * step are written on return.  Either pair may be NULL.  The continued run uses
 * the alpha, beta read from the file (10 decimals), as the reference does. */
int hphi_lanczos_restart(hphi_engine *e, const hphi_lanczos_params *p, const char *tm_in, const char *vec_in,
                         const char *tm_out, const char *vec_out, hphi_lanczos_result *r);
/* Lanczos eigenvector (Lanczos_EigenVector, src/Lanczos_EigenVector.c:46-208):
 * runs hphi_lanczos with p, takes the eigenvector c of the k-th lowest
 * eigenvalue (k = 0, 1, ...) of the final m x m tridiagonal matrix (vec12,
 * src/vec12.c:31-80), re-runs the recurrence from the same start vector with the
 * stored alpha/beta and accumulates v = sum_m c_m q_m into v (device array of
 * dim complex128), normalised to 1.  The sign of c is fixed by c_1 > 0 (the
 * reference takes LAPACK's).  r->eigenvalues[0] is the k-th Ritz value. */
int hphi_lanczos_eigenvector(hphi_engine *e, const hphi_lanczos_params *p, int k, double *v,
                             hphi_lanczos_result *r);

/* ---- expectation values of a normalised device vector v (dim complex128) ---- */
/* expec_energy_flct (src/expec_energy_flct.c:40-660): energy <v|H|v>, var
 * <Hv|Hv>; doublon, doublon2, num, num2 (Hubbard models: measured; spin:
 * 0, 0, Nsite, Nsite^2); sz, sz2 = <Sz>, <Sz^2> with Sz = (N_up - N_down)/2
 * (canonical Spin: the sector's Sz and its square, as the reference). */
typedef struct {
  double energy, var, doublon, doublon2, num, num2, sz, sz2;
} hphi_phys;
int hphi_expec_energy_flct(hphi_engine *e, const double *v, hphi_phys *out);
/* One-body Green's functions <v|c+_{i s1} c_{j s2}|v> (expec_cisajs,
 * src/expec_cisajs.c:69-650) for idx[n][4] = (i, s1, j, s2), 0-based sites;
 * out[2n] = (re, im).  Spin models: c+_{i s1} c_{i s2} = |s1><s2| on site i (bit
 * value = s); entries the reference sets to 0 are 0 (spin i != j; Sz-changing
 * entries in canonical sectors). */
int hphi_expec_cisajs(hphi_engine *e, const double *v, int n, const int *idx, double *out);
/* Two-body Green's functions <v|c+_{i s1} c_{j s2} c+_{k s3} c_{l s4}|v>
 * (expec_cisajscktaltdc, src/expec_cisajscktaltdc.c:87-1647) for idx[n][8];
 * spin entries need i == j and k == l. */
int hphi_expec_cisajscktalt(hphi_engine *e, const double *v, int n, const int *idx, double *out);
/* SpinGC three-, four- and six-body Green's functions (ThreeBodyG, FourBodyG,
 * SixBodyG; expec_Threebody/Fourbody/Sixbody_SpinGCHalf,
 * src/expec_cisajscktaltdc.c:901-1420) for idx[n][4*nbody] = nbody pairs
 * (i, s1, j, s2): <v| O_1 ... O_nbody |v> with O = |s1><s2| on site i when
 * i == j (a later pair on two sites gives 0); the first two pairs follow
 * Rearray_Interactions (the cross form (i,k)(k,i) counts with sign -1, other
 * forms are not evaluated: ok[t] = 0).  nbody in {3, 4, 6}. */
int hphi_expec_spin_nbody(hphi_engine *e, const double *v, int nbody, int n, const int *idx, double *out, int *ok);

/* ---- the reference's output and restart files (host memory; no device work) ----
 * Vectors are 0-based complex128 host arrays of dim elements; files hold the
 * reference's 1-based arrays (element 0 written as 0, skipped on reading).
 *   vec file      int step, unsigned long dim, complex128[dim+1]: the eigenvector
 *                 <head>_eigenvec_<k>_rank_<r>.dat (CalcByLanczos.c:160-198,262-272)
 *                 and the TPQ vector tmpvec_set<i>_rank_<r>.dat (CalcByTPQ.c:88-104,280-287)
 *   restart vecs  unsigned long step, unsigned long dim, v0[dim+1], v1[dim+1]:
 *                 <head>_recalcvec_rank_<r>.dat (Lanczos_EigenValue.c:428-489)
 *   TMComponents  "%d \n" steps, "%.10lf \n" norm, "%.10lf %.10lf \n" alpha_i beta_i
 *                 (OutputTMComponents / ReadTMComponents, Lanczos_EigenValue.c:624-720)
 *   energy        Energy / Doublon / Sz lines (CalcByLanczos.c:238-252)
 *   cisajs        " %4ld x4 %.10lf %.10lf\n" per entry (expec_cisajs.c)
 *   cisajscktalt  " %4ld x8 %.10lf %.10lf" with the reference's per-model line ends
 *                 (expec_cisajscktaltdc.c)
 *   TPQ files     SS_rand<i>.dat, Norm_rand<i>.dat, Flct_rand<i>.dat in dir from hphi_tpq
 *                 rows (12 doubles, rows[s*row_stride + q]) (CalcByTPQ.c:150-256)
 * Readers reject a file whose dimension differs from dim. */
int hphi_write_vec_file(const char *path, int step, uint64_t dim, const double *v);
int hphi_read_vec_file(const char *path, uint64_t dim, double *v, int *step);
int hphi_write_restart_vecs(const char *path, uint64_t step, uint64_t dim, const double *v0, const double *v1);
int hphi_read_restart_vecs(const char *path, uint64_t dim, double *v0, double *v1, uint64_t *step);
int hphi_write_tmcomponents(const char *path, int steps, double norm, const double *alpha, const double *beta);
int hphi_read_tmcomponents(const char *path, int cap, int *steps, double *norm, double *alpha, double *beta);
int hphi_write_energy(const char *path, const hphi_phys *p);
/* The energy file of the block solvers, one block per state: "State %ld",
 * "  Energy  %.16lf ", "  Doublon  %.16lf ", "  Sz  %.16lf ", blank line
 * (CalcByLOBPCG.c:744-760, <head>_energy.dat) */
int hphi_write_energy_states(const char *path, int n, const hphi_phys *p);
int hphi_write_cisajs(const char *path, int n, const int *idx, const double *val);
int hphi_write_cisajscktalt(const char *path, int model, int n, const int *idx, const double *val);
/* <head>_ThreeBody.dat / _FourBody.dat / _SixBody.dat lines for the entries
 * of hphi_expec_spin_nbody (the reference's spacing; ok[t] = 0: eight indices
 * and zeros) */
int hphi_write_nbody_green(const char *path, int nbody, int n, const int *idx, const double *val, const int *ok);
int hphi_write_tpq_files(const char *dir, int rand_i, int steps, const double *rows, int row_stride);
/* The same files for canonical TPQ rows (hphi_ctpq), which also write a
 * Flct_rand line at step 0 (CalcByCanonicalTPQ.c:160-200). */
int hphi_write_ctpq_files(const char *dir, int rand_i, int steps, const double *rows, int row_stride);
/* A resumed run's rows (steps step0 .. steps-1, rows[(s - step0)*row_stride + q])
 * appended to the SS_rand, Norm_rand and Flct_rand files of sample rand_i, as
 * the reference appends them under ReStart (no header); norm1 is the
 * global_1st_norm column (the reference prints its process value there). */
int hphi_append_tpq_files(const char *dir, int rand_i, int step0, int steps, const double *rows, int row_stride,
                          double norm1);
/* Lowest `want` distinct eigenvalues of the plain-Lanczos tridiagonal
 * (alpha[0..m), beta[0..m-1)) after Cullum-Willoughby removal of spurious
 * values and merging of ghost copies; returns how many were found. */
int hphi_lanczos_good_eigenvalues(const double *alpha, const double *beta, int m, int want, double *out);
int hphi_lanczos_converged(const double *prev, int nprev, const double *cur, int ncur, int want, double tol);
/* Ritz residual beta_m |y_m| of the eigenvalue theta of T_m (alpha[0..m),
 * beta[0..m-1); y its unit eigenvector, by inverse iteration): |H z - theta z|
 * for the Ritz vector z (host only) */
int hphi_lanczos_ritz_residual(const double *alpha, const double *beta, int m, double beta_m, double theta,
                               double *res);
/* 1 when every one of the `want` values cur[] has a Ritz residual <= 0.1
 * sqrt(tol) max(1, |theta|): the second half of the distinct mode's
 * convergence test (the first: hphi_lanczos_converged) */
int hphi_lanczos_residuals_small(const double *alpha, const double *beta, int m, double beta_m, const double *cur,
                                 int want, double tol);
/* Microcanonical TPQ (CalcByTPQ, src/CalcByTPQ.c:52-294) for k samples at once.
 * Sample c starts from a random vector of seed + c (init_kind 0: uniform
 * complex, -1: complex Gaussian); l = large_value, or, when large_value <= 0,
 * max(3 max|list_Diagonal|, max|list_Diagonal| + sum of the off-diagonal
 * |couplings|) / Nsite: the larger of 3 max|diag| and a Gershgorin bound of
 * |E|, so that l - H/Nsite is positive definite (the reference has no default
 * and reads LargeValue from modpara.def).  layout 1 keeps the k vectors state-major and uses
 * the block H·v; layout 0 keeps them vector-major with one H·v per sample.
 * use_graph captures one step in a hipGraph and replays it.
 * out[(s*k + c)*12 + q], s = 0..steps-1 (the reference's step_i), q:
 *   0 inv_temp, 1 energy, 2 <H^2>, 3 norm (global_norm; step 0: global_1st_norm),
 *   4 doublon, 5 doublon^2, 6 num, 7 num^2, 8 Sz, 9 Sz^2 (the values of
 *   expec_energy_flct per model: measured for Hubbard models and SpinGC Sz;
 *   Nsite, Nsite^2 and the sector's Sz for spin models), 10 Im <H> (rounding
 *   check), 11 reserved (0).  Rows map to SS_rand / Norm_rand / Flct_rand
 *   (hphi_write_tpq_files). */
typedef struct {
  int k, steps, init_kind, layout, use_graph;
  double large_value;
  uint64_t seed;
} hphi_tpq_params;
typedef struct {
  double seconds, large_value;
  int steps;
} hphi_tpq_result;
int hphi_tpq(hphi_engine *e, const hphi_tpq_params *p, double *out, hphi_tpq_result *r);
/* Canonical TPQ (CalcByCanonicalTPQ, src/CalcByCanonicalTPQ.c:52-300) for k
 * samples at once, state-major with the block H·v.  Each step s >= 1 applies
 * exp(-dtau H/2) to the normalised psi by its Taylor series to order
 * expand_coef (MultiplyForCanonicalTPQ, src/Multiply.c:152-205; 0 means the
 * reference's default 10), normalises and measures; dtau = 1/l with l as in
 * hphi_tpq.  Rows as hphi_tpq with q0 = inv_temp = s dtau and q3 = the norm
 * before normalisation (global_norm; step 0: global_1st_norm).  Costs
 * expand_coef block H·v per step. */
typedef struct {
  int k, steps, init_kind, expand_coef;
  double large_value;
  uint64_t seed;
} hphi_ctpq_params;
int hphi_ctpq(hphi_engine *e, const hphi_ctpq_params *p, double *out, hphi_tpq_result *r);
/* The TPQ runs from caller-supplied vectors and the reference's restart
 * (ReStart, CalcByTPQ.c:86-117,272-287; CalcByCanonicalTPQ.c:88-117,282-297).
 * W0: device array of k vectors of dim complex128, vector c at [c*dim ..).
 *   step0 = 0: the start vectors, unnormalised as MakeIniVec draws them (their
 *     norm is step 0's global_1st_norm; the reference's ini vector times that
 *     norm); rows of steps 0 .. steps-1 as hphi_tpq / hphi_ctpq.
 *   step0 > 0: resume: the normalised vectors of step step0 - 1 of an earlier
 *     run, as a restart file tmpvec_set<i> holds them with step field step0;
 *     they are measured again (no row), and out holds the rows of steps
 *     step0 .. steps-1 only (out[((s - step0)*k + c)*12 + q]).  The reference
 *     leaves inv_temp of a resumed canonical run uninitialised; here it
 *     continues as s * dtau.
 * W_end (may be NULL): device array as W0 receiving the normalised vectors of
 * step steps - 1, which a restart file stores with step field `steps`
 * (hphi_write_vec_file). */
int hphi_tpq_from(hphi_engine *e, const hphi_tpq_params *p, const double *W0, int step0, double *out,
                  hphi_tpq_result *r, double *W_end);
int hphi_ctpq_from(hphi_engine *e, const hphi_ctpq_params *p, const double *W0, int step0, double *out,
                   hphi_tpq_result *r, double *W_end);
/* max_j |list_Diagonal[j]| */
int hphi_diag_maxabs(hphi_engine *e, double *out);

/* Block LOBPCG (CalcByLOBPCG.c:328-632) for the k lowest eigenpairs: the
 * reference's residual, optional point-Jacobi preconditioner with adaptive
 * shift, 3k x 3k Gram matrices, Loewdin-orthogonalised subspace problem,
 * mu <- (mu + mu_sub)/2 and recombination, with one block H·v per step.
 * Column c of the initial block is the synthetic vector of seed + c.
 * eigenvalues = Rayleigh quotients of the final block; history (may be NULL,
 * max_iter rows of 1 + k doubles): (max residual, mu_0 .. mu_{k-1}) at the
 * start of each step, the reference's step log (:395-404). */
typedef struct {
  int k;              /* block size = eigenpairs (k_exct): 1, 2, 4, 8 or 16 */
  int max_iter;       /* Lanczos_max */
  double eps;         /* residual threshold: 10^(-LanczosEps/2) in the reference (:384) */
  int precg;          /* PreCG */
  double large_value; /* LargeValue: the preconditioner's positive shift is LargeValue*Nsite (:339) */
  uint64_t seed;
} hphi_lobpcg_params;
typedef struct {
  double eigenvalues[16], residuals[16];
  double max_residual;
  int iterations, converged;
  double seconds, hv_seconds;
} hphi_lobpcg_result;
int hphi_lobpcg(hphi_engine *e, const hphi_lobpcg_params *p, double *history, hphi_lobpcg_result *r);
/* hphi_lobpcg with the reference's start and result blocks (CalcByLOBPCG.c:
 * Initialize_wave :153-282, Output_restart :283-319, L_vec :716-724).  X0
 * (may be NULL): device array of k start vectors of dim complex128, vector c
 * at [c*dim ..) (e.g. the restart files tmpvec_set<c>_rank_0.dat read with
 * hphi_read_vec_file), normalised here; Xout (may be NULL): the final
 * eigenvector block, normalised, the same layout: write it with
 * hphi_write_vec_file for a restart or an eigenvector file, and measure each
 * column with hphi_expec_* (phys, src/phys.c:48-140). */
int hphi_lobpcg_ex(hphi_engine *e, const hphi_lobpcg_params *p, const double *X0, double *Xout, double *history,
                   hphi_lobpcg_result *r);
/* Eigen-decomposition of an n x n Hermitian matrix (column-major, interleaved
 * complex): w ascending, V (may be NULL) the eigenvectors as columns; the
 * subspace solver of hphi_lobpcg (replaces LAPACK zheevd, :60,100). */
int hphi_herm_eig(int n, const double *A, double *w, double *V);

/* ---- continuous-batching service ----------------------------------------
 * No counterpart in the reference (which solves one model per process).
 * Jobs (canonical Spin / Hubbard models) are submitted with an arrival time
 * on the server clock (hphi_server_now); each hphi_server_tick admits arrived
 * jobs in order while their HBM footprint fits the budget and a slot is
 * free, advances every running job by one plain-Lanczos step in packed
 * launches, and retires jobs whose lowest n_eigen distinct eigenvalues
 * (Cullum-Willoughby filtered, checked every check_every steps) changed by
 * less than eps relative.  A job's numbers do not depend on the other jobs. */
typedef struct hphi_server hphi_server;
typedef struct {
  uint64_t hbm_budget; /* arena bytes (0: 16 GiB) */
  int max_jobs, n_eigen;
  double eps;
  int max_iter, check_every;
} hphi_server_params;
typedef struct {
  int id, n_found, iterations, converged;
  uint64_t dim;
  double eigenvalues[16];
  double t_arrival, t_admit, t_done;
} hphi_job_result;
int hphi_server_create(const hphi_server_params *p, hphi_server **out);
void hphi_server_destroy(hphi_server *s);
int hphi_server_submit(hphi_server *s, const hphi_model_desc *d, uint64_t seed, double t_arrival, int *id);
int hphi_server_tick(hphi_server *s, int *n_active, int *n_waiting);
int hphi_server_results(hphi_server *s, hphi_job_result *out, int max, int *n);
double hphi_server_now(const hphi_server *s);
/* out[0..9]: seconds in admission, in launched work (including the wait), in
 * convergence checks; number of non-empty ticks; job-steps advanced;
 * slab-steps advanced (a slab holds up to 4 jobs of one structure); the slab
 * H·v launches: seconds between HIP events around them, their algorithmic
 * bytes (per slab 8 dim, per running column 56 dim + 16 per off-diagonal
 * element), their number; jobs moved between slabs of one structure to pack
 * their running columns into fewer slabs */
int hphi_server_stats(const hphi_server *s, double *out);

/* k lowest eigenvalues (ascending) of the symmetric tridiagonal matrix with
 * diagonal a[0..n) and off-diagonal b[0..n-1), by Sturm bisection (host only;
 * replaces DSEVvalue -> dsyev_, src/matrixlapack.c:93-150). */
int hphi_tridiag_lowest(const double *a, const double *b, int n, int k, double *ev);

#ifdef __cplusplus
}
#endif
#endif
