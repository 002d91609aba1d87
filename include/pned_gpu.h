/*
 * pned_gpu.h -- C ABI of the MI355X H.v + Krylov library (libpned_gpu.so).
 *
 * One context = one operator (cluster + bond matrices) on one GPU, with its own HIP
 * stream.  Vectors live in device memory and are referred to by opaque handles; the
 * caller only sees host scalars.  Every call returns 0 on success or a negative
 * PNED_E_* code; pned_last_error() gives the message (the reference prints a
 * message and stops instead, e.g. #(E07) at S_one_half/Hybrid_TPS_Lanczos.f90:126-130;
 * the thin Fortran layer restores that behaviour).
 *
 * Array conventions (identical to the reference's data_mem tables, S_one_half/
 * Hybrid_TPS_params.f90:313-329):
 *   ipair       int32[2*ibond], 1-based site numbers, reference bond order
 *   ibond_type  int32[ibond], 1..9
 *   xlocal_ham  double[2*qq*qq*9]  complex (re,im) in Fortran order of
 *               xlocal_Ham(0:qq-1, 0:qq-1, 9), qq = q_spin^2
 *   llocal_ham  int32[qq*qq*9]      Fortran order of llocal_Ham (nonzero = skipped)
 * The H.v applies v0(out) += xlocal_Ham(in,out,type) * v1(in) over every bond, the
 * formula of H_dot_v (S_one_half/Hybrid_TPS_parallel.f90:415-440), with the bond's
 * local index q*digit(max site) + digit(min site).
 * Complex vectors are (re,im) pairs: complex128 = 2 doubles, complex64 = 2 floats.
 * A vector with nvec > 1 holds nvec columns interleaved state-major: [dim][nvec].
 */
#ifndef PNED_GPU_H
#define PNED_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pned_ctx pned_ctx;
typedef struct pned_vec pned_vec;

enum {
    PNED_OK = 0,
    PNED_E_ARG = -1,        /* bad argument */
    PNED_E_NOMEM = -2,      /* device allocation failed */
    PNED_E_HIP = -3,        /* HIP runtime error */
    PNED_E_UNSUPPORTED = -4,/* lattice/model outside what the kernels handle */
    PNED_E_BREAKDOWN = -5,  /* Lanczos beta below 0.5e-30 (reference #(E07)) */
    PNED_E_NOCONV = -6      /* iteration limit reached without convergence */
};

enum { PNED_C128 = 0, PNED_C64 = 1 };

const char *pned_last_error(void);
int pned_version(void);
/* number of visible HIP devices (0 when no GPU is usable) */
int pned_device_count(void);

/* ---- context: replaces the data_mem tables + set_parallel_bits + make_cpu_list2 ----
 * set_xlocal_Ham / set_bond_info stay with the caller; their outputs come in here.
 * (S_one_half/Hybrid_TPS_main.f90:57-65; S_arbitrary/Hybrid_TPS9_main.f90)            */
int pned_ctx_create(int n_sites, int q_spin, int ibond, const int32_t *ipair,
                    const int32_t *ibond_type, const double *xlocal_ham,
                    const int32_t *llocal_ham, int device, pned_ctx **out);
/* destroying a context detaches its remaining vectors: they keep their device memory until
 * pned_vec_destroy and are refused by every call that takes a context */
int pned_ctx_destroy(pned_ctx *ctx);
/* batch context (no reference counterpart; for serving same-size requests together):
 * nop operators on the same cluster, one per column of the context's vectors (nop a power
 * of two <= 64).  xlocal_ham/llocal_ham hold nop consecutive tables.  The operators share
 * one term structure; a later operator must fit it (pned_ctx_set_column_matrices). */
int pned_ctx_create_batch(int n_sites, int q_spin, int ibond, const int32_t *ipair,
                          const int32_t *ibond_type, int nop, const double *xlocal_ham,
                          const int32_t *llocal_ham, int device, pned_ctx **out);
int pned_ctx_set_column_matrices(pned_ctx *ctx, int col, const double *xlocal_ham,
                                 const int32_t *llocal_ham);
/* basis dimension q^N (ldim at np=1, S_one_half/Hybrid_TPS_MPIlib.f90:74-84) */
int64_t pned_ctx_dim(const pned_ctx *ctx);
/* replace the bond matrices in place (same sparsity structure not required),
 * e.g. change_to_TPS_hamiltonian (S_one_half/Hybrid_TPS_params.f90:1600-1623) */
int pned_ctx_set_matrices(pned_ctx *ctx, const double *xlocal_ham, const int32_t *llocal_ham);
/* tile/pass tuning knobs of the S=1/2 H.v (no reference counterpart): tile bits <= 14,
 * forced low bits, threads per block 256, 512 or 1024 (1024: term-major kernel only);
 * 0 = default */
int pned_ctx_tune(pned_ctx *ctx, int tile_bits, int low_bits, int threads);
/* autotuner (no reference counterpart): times one fresh H.v v1 -> v0 for every
 * (tile bits, forced low bits, threads per block) candidate that fits LDS and keeps the
 * fastest for this context; v0 is overwritten.  Outputs may be NULL. */
int pned_ctx_autotune(pned_ctx *ctx, const pned_vec *v1, pned_vec *v0, int reps, int *best_tile_bits,
                      int *best_low_bits, int *best_threads, double *best_ms);
/* spin-S kernel choice: 1 = tiled passes (default, q = 3..5), 0 = per-state gather */
int pned_ctx_spin_kernel(pned_ctx *ctx, int tiled);
/* HIP graph replay in the Krylov loops (no reference counterpart): 1 (default) = pned_lnc1
 * and pned_dsflnc1z on one-column S=1/2 full-space contexts replay their steps between the
 * host's convergence checks from captured graphs of 5 steps; 0 = every step launched
 * and read back one by one.  Both give the same alpha/beta bit for bit. */
int pned_ctx_graphs(pned_ctx *ctx, int enable);
/* S^z sector basis (parity_type = 0, mk_list12: S_one_half/Hybrid_TPS_parallel.f90:452-578):
 * restrict the context to the states with n_down set bits, indexed as the reference's
 * list1/list3 tables do; n_down < 0 returns to the full 2^N space.  The context's dim
 * becomes C(N, n_down): create vectors after this call.  S=1/2 only; fails with
 * PNED_E_UNSUPPORTED when the operator does not conserve S^z.  H.v, BLAS-1 and the Krylov
 * drivers work in the sector; S^z(q) and mTPQ observables need the full basis. */
int pned_ctx_sz_sector(pned_ctx *ctx, int n_down);
/* S=1/2 pass kernel choice (no reference counterpart): 1 = term-major (default; per-term
 * selector and partner address resolved once per tile), 0 = state-major */
int pned_ctx_s12_kernel(pned_ctx *ctx, int kind);
/* diagnostic: number of tile passes one H.v makes, and terms per pass (max 64) */
int pned_ctx_plan_info(const pned_ctx *ctx, int *npasses, int *terms_per_pass, int max);
/* diagnostic, needs no GPU: the S=1/2 term decomposition and tile passes the planner
 * builds for these tables.  Outputs (caller-sized):
 *   pass_info[npass*4]   = {K, first_term, nterm, has_diag}
 *   pass_bits[npass*40]  = tile bit positions (memory index), K used
 *   terms[nterm*6]       = {tmask, selmask, c0r, c0i, c1r, c1i} (as doubles)
 *   term_gmask[nterm]    = flipped state bits
 *   diag[ngroup*5 + 2]   = {kind, shift, mask, wr, wi}..., c0r, c0i
 * Returns the number of passes in *npass, terms in *nterm, groups in *ngroup. */
int pned_s12_plan_export(int n_sites, int ibond, const int32_t *ipair, const int32_t *ibond_type,
                         const double *xlocal_ham, const int32_t *llocal_ham, int colbits,
                         int tile_bits, int low_bits, int *npass, int *nterm, int *ngroup,
                         int32_t *pass_info, int32_t *pass_bits, double *terms,
                         uint32_t *term_gmask, double *diag, int max_pass, int max_term,
                         const int32_t *site_bit /* memory layout (see below), NULL = identity */);
/* diagnostic, needs no GPU: the spin-S (q = 3..5) tile passes the planner builds for these
 * tables.  tile_sites <= 0: the largest tile (q^K <= 8192 states); low_sites 0: the kernel's
 * choice (3 forced low sites when that costs no pass, else 2).  Outputs (caller-sized):
 *   pass_sites[npass*32] = tile sites (0-based, lowest first), -1 padded
 *   bond_pass[ibond]     = pass applying each bond
 *   cls[npass*64*(3+2*q*q)] = per pass and class: {bond, delta_min_site, delta_max_site,
 *                          coefficient re/im by output local state} (as doubles, unused rows 0)
 *   ncls[npass]          = off-diagonal classes per pass
 * Returns the pass count in *npass and the forced low sites in *low_used. */
int pned_spin_plan_export(int n_sites, int q, int ibond, const int32_t *ipair, const int32_t *ibond_type,
                          const double *xlocal_ham, const int32_t *llocal_ham, int tile_sites, int low_sites,
                          int *npass, int *low_used, int32_t *pass_sites, int32_t *bond_pass, double *cls,
                          int32_t *ncls, int max_pass);
/* ---- memory layout of full-space S=1/2 vectors (no reference counterpart) ----
 * Device memory holds basis state s at memory index swap(s), where swap exchanges up to 4
 * disjoint pairs of bits: site k (0-based) lives in memory bit site_bit[k], and
 * site_bit[site_bit[k]] == k.  A new single-operator S=1/2 context picks the layout that
 * puts the three sites into the lowest memory bits (the bits every tile pass keeps, for
 * contiguous 128-byte runs) that need the fewest tile passes; identity when no choice
 * saves a pass.  From 30 sites on, the swaps left move sites into memory bits 3..7 when
 * the tile passes then hold more of those bits (longer runs per 4 KB of memory), with the
 * same pass count.  Every host boundary -- upload/download, vector files, the Gaussian start
 * vector (counter = basis index), S^z(q), S^+-(q), mTPQ observables, check1 -- maps
 * through it, so callers see the reference's basis order; only pned_vec_data exposes the
 * layout.  pned_ctx_sz_sector returns a context to the identity layout. */
int pned_ctx_layout(const pned_ctx *ctx, int32_t *site_bit /* n_sites */);
/* set the layout explicitly (NULL = identity); only before the context has vectors */
int pned_ctx_set_layout(pned_ctx *ctx, const int32_t *site_bit);
/* diagnostic, needs no GPU: the layout pned_ctx_create would choose for these tables */
int pned_s12_layout_choose(int n_sites, int ibond, const int32_t *ipair, const int32_t *ibond_type,
                           const double *xlocal_ham, const int32_t *llocal_ham, int32_t *site_bit);
/* synchronize the context stream */
int pned_ctx_sync(pned_ctx *ctx);
/* the hipStream_t the context launches on (for event timing by the caller) */
void *pned_ctx_stream(pned_ctx *ctx);

/* ---- vectors: replace the working_area arrays wwk/rr1/xvec (Hybrid_TPS_params.f90:272-311) */
int pned_vec_create(pned_ctx *ctx, int nvec, int dtype, pned_vec **out);
int pned_vec_destroy(pned_vec *v);
int pned_vec_info(const pned_vec *v, int64_t *dim, int *nvec, int *dtype);
/* copy count states (all columns) starting at state 'offset' to/from host memory */
int pned_vec_upload(pned_ctx *ctx, pned_vec *v, const void *host, int64_t offset, int64_t count);
int pned_vec_download(pned_ctx *ctx, const pned_vec *v, void *host, int64_t offset, int64_t count);
int pned_vec_zero(pned_ctx *ctx, pned_vec *v);
int pned_vec_copy(pned_ctx *ctx, const pned_vec *src, pned_vec *dst);
/* device pointer of the data (for zero-copy use by the caller's own kernels; states are in
 * the context's memory layout, pned_ctx_layout) */
void *pned_vec_data(pned_vec *v);

/* ---- parallelized_diag (S_one_half/Hybrid_TPS_parallel.f90:6-8) ---------------------- */

/* parallelized_set_initial_vec_p (Hybrid_TPS_parallel.f90:17-102), replaced: i.i.d. N(0,1)
 * real/imag parts from Philox-4x32-10 (key = seed, counter = (index, column)), each column
 * normalised in FP64. */
int pned_set_initial_vec(pned_ctx *ctx, pned_vec *v, uint64_t seed);

/* p_mltply_v3 (Hybrid_TPS_parallel.f90:190-263): v0 += (H - ibond*eperbd) v1 and
 * prdct[c] = <v1_c|(H - ibond*eperbd)|v1_c> per column (2 doubles each).
 * eperbd is one value for all columns (or an nvec array, see pned_hv_multi). */
int pned_p_mltply_v3(pned_ctx *ctx, const pned_vec *v1, pned_vec *v0, double eperbd, double *prdct);
/* same, with v0 treated as zero on input (fresh output: 2 instead of 3 vector streams) */
int pned_hv_fresh(pned_ctx *ctx, const pned_vec *v1, pned_vec *v0, double eperbd, double *prdct);

/* Allocation-free p_mltply_v3.  No counterpart in the reference: for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same v0 += (H - ibond*eperbd) v1 as pned_p_mltply_v3 on an S=1/2
 * full-space context, with the same kernels; prdct_dev (device memory, 2 doubles per column)
 * receives <v1_c|(H - ibond*eperbd)|v1_c>.  It allocates nothing and does not synchronise:
 * the pass tables must exist for this vector shape (one earlier pned_p_mltply_v3 with such
 * vectors builds them; otherwise PNED_E_ARG), and the per-block partial sums go to a static
 * device array of PNED_NOALLOC_SCRATCH_BYTES per GPU (passes x 4096 x 2 x nvec doubles must
 * fit: nvec <= 8 at up to 8 passes; larger requests fail with PNED_E_ARG).  Every call on
 * one GPU shares that array: calls that overlap on one GPU (several streams or threads)
 * race. */
#define PNED_NOALLOC_SCRATCH_BYTES (4u << 20)
int pned_p_mltply_v3_noalloc(pned_ctx *ctx, const pned_vec *v1, pned_vec *v0, double eperbd,
                             double *prdct_dev);

/* calc_vec_inprdct (Hybrid_TPS_parallel.f90:104-126): out[c] = sum conj(x_c) y_c, FP64 */
int pned_calc_vec_inprdct(pned_ctx *ctx, const pned_vec *x, const pned_vec *y, double *out);
/* calc_beta (Hybrid_TPS_parallel.f90:161-188): beta2[c] = ||v0_c - alpha_c v1_c||^2 */
int pned_calc_beta(pned_ctx *ctx, const double *alpha, const pned_vec *v1, const pned_vec *v0,
                   double *beta2);
/* calc_alpha_beta (Hybrid_TPS_parallel.f90:145-159): p_mltply_v3 then calc_beta */
int pned_calc_alpha_beta(pned_ctx *ctx, pned_vec *v1, pned_vec *v0, double *alpha, double *beta);
/* calc_alpha_beta2 (Hybrid_TPS_parallel.f90:128-142): p_mltply_v3 then ||v0|| */
int pned_calc_alpha_beta2(pned_ctx *ctx, pned_vec *v1, pned_vec *v0, double *alpha, double *beta);

/* BLAS-1 helpers used by the Krylov loops (per column scalars, complex = 2 doubles):
 * y = a*x + b*y */
int pned_axpby(pned_ctx *ctx, const double *a, const pned_vec *x, const double *b, pned_vec *y);
int pned_scale(pned_ctx *ctx, pned_vec *x, const double *a);

/* ---- Krylov drivers ---------------------------------------------------------------------- */

/* Lanczos three-term update of lnc1z / dsflnc1z (Hybrid_TPS_Lanczos.f90:111-116):
 *   t = v1; v1 = (v0 - alpha v1)/beta; v0 = -beta t      (per column)                    */
int pned_lanczos_update(pned_ctx *ctx, pned_vec *v1, pned_vec *v0, const double *alpha,
                        const double *beta);

/* parallelized_lnc1 + parallelized_lnc1z (Hybrid_TPS_Lanczos.f90:7-163) on column 0.
 * v1 holds the start vector (restart semantics: the caller fills it, e.g. with
 * pned_set_initial_vec), v0 is work.  alpha/beta need max_itr entries, ene nvec.
 * fixed_steps > 0 runs exactly that many steps (no convergence exit).  *itr gets the
 * step count.  Convergence: 8th (nvec-th) Ritz value changes by < 1e-13 relative,
 * checked every 5 steps after 4*nvec (Hybrid_TPS_Lanczos.f90:133-150).                    */
int pned_lnc1(pned_ctx *ctx, pned_vec *v1, pned_vec *v0, int max_itr, int nvec, int fixed_steps,
              double *alpha, double *beta, double *ene, int *itr);

/* Resumable Lanczos (no separate reference entry; the restart branch of lnc1z,
 * Hybrid_TPS_Lanczos.f90:87-93): the lnc1z loop from step `first` with the reference's
 * explicit update, so that the loop state at any step is the reference's vectors.
 * first <= 2 starts from the normalised v1; first > 2 continues from the loop state of the
 * top of step `first` (v1, v0 and alpha/beta[0..first-2], e.g. after
 * pned_interval_dump_read).  stop >= max(first,2) returns at the top of step `stop` with
 * *itr = stop and the loop state in v1, v0 (what interval_dump writes); stop = 0 runs to
 * convergence, max_itr or fixed_steps as pned_lnc1 does.  A run split into spans gives the
 * same alpha/beta bit for bit as one span. */
int pned_lnc1_span(pned_ctx *ctx, pned_vec *v1, pned_vec *v0, int first, int stop, int max_itr, int nvec,
                   int fixed_steps, double *alpha, double *beta, double *ene, int *itr);

/* Lanczos with full reorthogonalisation (no reference counterpart; the reference's unused
 * orthg, Hybrid_TPS_basic.f90:219-303, is the closest): basis[0] holds the normalised start
 * vector, basis[1..] receive the Lanczos vectors (two passes of classical Gram-Schmidt
 * against all earlier ones per step), w is work; all of one shape, complex128 or -- the
 * precision form -- complex64, with every dot, norm, alpha and beta in FP64.  fixed_steps
 * > 0 runs that many steps; else the lnc1z convergence rule up to nbasis - 1 steps. */
int pned_lnc1_reorth(pned_ctx *ctx, pned_vec *const *basis, int nbasis, pned_vec *w, int nvec, int fixed_steps,
                     double *alpha, double *beta, double *ene, int *itr);
/* the DSF form of it: s_maxitr - 1 steps from the normalised excitation in basis[0];
 * salpha[s_maxitr], sbeta[s_maxitr] with sbeta[0] = cnorm (as pned_dsflnc1z) */
int pned_dsflnc1z_reorth(pned_ctx *ctx, pned_vec *const *basis, int nbasis, pned_vec *w, double cnorm, int s_maxitr,
                         double *salpha, double *sbeta);

/* Serving (no reference counterpart): one Lanczos step of every slot (column) of a batch.
 * mode[c]: 0 idle (column zeroed, alpha = beta = 0), 1 first step (v1 holds the start
 * vector, v0 is cleared; the lnc1z start, Hybrid_TPS_Lanczos.f90:78-85), 2 continue (the
 * update of Hybrid_TPS_Lanczos.f90:111-116 with alpha_prev[c], beta_prev[c]).  Then one
 * batched H.v and per-column alpha[c] = <v1|H|v1>, beta[c] = ||v0 - alpha v1||.  Every
 * column's arithmetic is independent of the others (fixed-order per-column reductions).
 * One column of a plain context (S=1/2 full space, tiled spin-S) steps as pned_lnc1 does,
 * with the normalisations deferred: between calls v1 and v0 hold scaled vectors in roles
 * the context tracks (only alpha and beta are meaningful; mode 1 restarts from v1, mode 0
 * does nothing). */
int pned_lanczos_batch_step(pned_ctx *ctx, pned_vec *v1, pned_vec *v0, const int *mode,
                            const double *alpha_prev, const double *beta_prev, double *alpha,
                            double *beta);
/* nsteps calls of pned_lanczos_batch_step in one (no reference counterpart; for servers
 * that check convergence every few steps): the first step takes mode, alpha_prev and
 * beta_prev as given, each later one continues every stepped column (mode 2) from the
 * previous step's alpha and beta, and idles (mode 0) a column whose beta fell below 0.5e-30
 * or is not a number.  alpha and beta are nsteps x nvec (step-major); idle entries are 0.
 * Every value equals what the single steps would give. */
int pned_lanczos_batch_steps(pned_ctx *ctx, pned_vec *v1, pned_vec *v0, const int *mode,
                             const double *alpha_prev, const double *beta_prev, int nsteps,
                             double *alpha, double *beta);
/* column 'col' of v := normalised Gaussian start vector of 'seed' (the values of column 0
 * of pned_set_initial_vec for that seed, whatever the column) */
int pned_set_initial_column(pned_ctx *ctx, pned_vec *v, int col, uint64_t seed);

/* bisec (Hybrid_TPS_basic.f90:25-107) and vec12 (112-213); host FP64.  coef is ndim x nvec
 * column-major. */
int pned_bisec(const double *alpha, const double *beta, int ndim, double *ene, int ne, double eps);
int pned_vec12(const double *alpha, const double *beta, int ndim, const double *ene, int nvec,
               double *coef);
/* The same two solves (same arguments, errors and results, bit for bit: one source for both
 * sides) in FP64 on the context's GPU: a wavefront per eigenvalue evaluating six levels of
 * the bisection tree per round, a lane per cluster of eigenvectors; host arrays in and out.  pned_lnc1 and its span/reorth forms take their
 * convergence checks' eigenvalues from pned_ctx_bisec. */
int pned_ctx_bisec(pned_ctx *ctx, const double *alpha, const double *beta, int ndim, double *ene, int ne,
                   double eps);
int pned_ctx_vec12(pned_ctx *ctx, const double *alpha, const double *beta, int ndim, const double *ene, int nvec,
                   double *coef);

/* Ritz vector: x = sum_k coef[k] v_k regenerated by re-running the Lanczos recursion
 * from the start vector in v1 (3 resident vectors: v1, v0, x).  Uses the alpha/beta of the
 * first run.  */
int pned_ritz_vector(pned_ctx *ctx, pned_vec *v1, pned_vec *v0, pned_vec *x, int nsteps,
                     const double *alpha, const double *beta, const double *coef);

/* parallelized_inv1z + parallelized_cg1 (Hybrid_TPS_cg.f90:18-197): inverse iteration for
 * the eigenvector of eigenvalue eig.  b holds the start vector; x the result (normalised).
 * work must hold 3 vectors' handles (p, r, y).  reference_saved_loop = 1 reproduces the
 * SAVEd 'loop' flag of cg1 (Hybrid_TPS_cg.f90:118).  *iters = inverse iterations taken. */
int pned_inv1(pned_ctx *ctx, double eig, pned_vec *x, pned_vec *b, pned_vec *p, pned_vec *r,
              pned_vec *y, int max_inv, int max_cg, int reference_saved_loop, int *iters);

/* COCG for the complex-shifted system (H - e0 + i*eta) x = b (no reference counterpart:
 * the reference's CG shift is real, Hybrid_TPS_cg.f90:135).  Stops at ||r|| < tol ||b||. */
int pned_cocg(pned_ctx *ctx, double e0, double eta, const pned_vec *b, pned_vec *x, pned_vec *r,
              pned_vec *p, pned_vec *y, int max_it, double tol, int *iters, double *resid);

/* check1 (Hybrid_TPS_check.f90:8-54): expect = <x|H|x>, ratios (Hx)_j/x_j at
 * j = min(ldim/3,13) : ldim : max(1,ldim/20) (1-based), up to max_ratios; *nratios set. */
int pned_check1(pned_ctx *ctx, const pned_vec *x, pned_vec *work, double *expect, double *ratios,
                int max_ratios, int *nratios);

/* residual ||(H - e) x|| / ||x|| using 'work' as scratch vector */
int pned_residual(pned_ctx *ctx, const pned_vec *x, double e, pned_vec *work, double *res);

/* ---- mTPQ (Hybrid_TPS_TPS.f90:45-409) -------------------------------------------------- */
/* Observables of eval_phys_val: site positions (n x 2), the static-structure-factor
 * momenta (nsq x 2).  The operator of the context must already be the shifted one
 * (l - h_b, change_to_TPS_hamiltonian) and lmax = ibond*l_b. */
int pned_tps_set_observables(pned_ctx *ctx, const double *site_vec, int nsq, const double *sq_q);
/* Runs nsteps steps of parallelized_TPSz on every column of v1 (the normalised start
 * vectors); v0, vg, vt are work vectors with the same shape.  alpha/beta are
 * [nsteps][nvec], coef is [nsteps][nvec][7+nsq] (row of eval_phys_val per column). */
int pned_tps(pned_ctx *ctx, pned_vec *v1, pned_vec *v0, pned_vec *vg, pned_vec *vt, double lmax,
             int nsteps, double *alpha, double *beta, double *coef);

/* pned_tps from step `first` (the TPSz restart branch, Hybrid_TPS_TPS.f90:352-358):
 * first > 1 continues from the loop state of the top of step `first` (v0 = (L-H) v1 of the
 * step before, v1, vg; rows 0..first-2 of alpha/beta/coef filled); stop >= first returns at
 * the top of step `stop`.  Split runs give the rows of one span bit for bit. */
int pned_tps_span(pned_ctx *ctx, pned_vec *v1, pned_vec *v0, pned_vec *vg, pned_vec *vt, double lmax, int first,
                  int stop, int nsteps, double *alpha, double *beta, double *coef);

/* Same outputs as pned_tps with the vector work fused (rotation instead of copies, the
 * 1/beta normalisation folded into the next H.v, one reduction pass per step).  v1 holds
 * the normalised start columns; va, vb are work vectors; on return all three hold work
 * data.  At most 8 structure-factor entries. */
int pned_tps_fused(pned_ctx *ctx, pned_vec *v1, pned_vec *va, pned_vec *vb, double lmax, int nsteps,
                   double *alpha, double *beta, double *coef);
/* pned_tps_fused (the same outputs, bit for bit) with the steps replayed from a HIP graph
 * captured once: graph_steps steps per replay (a multiple of 3, at most 3000; 0 = 30), the
 * 1/beta scale of each step formed on the device, one host read-back per replay.  Same
 * counterpart (parallelized_TPSz, Hybrid_TPS_TPS.f90:45-409) and requirements; the call
 * allocates its step history (graph_steps x (2 + 26) x nvec doubles) and frees it on return. */
int pned_tps_fused_graph(pned_ctx *ctx, pned_vec *v1, pned_vec *va, pned_vec *vb, double lmax, int nsteps,
                         int graph_steps, double *alpha, double *beta, double *coef);

/* ---- DSF (Hybrid_TPS_DSF.f90) ------------------------------------------------------------ */
/* intalv_zz (Hybrid_TPS_DSF.f90:18-87): v = normalised sum_j w(d_j) e^{i q.r_j} x with
 * w = (d - 1/2) for S=1/2 (bit value d) or (d - S) for spin S; cnorm = ||.||^2 / N. */
int pned_intalv_zz(pned_ctx *ctx, const pned_vec *x, pned_vec *v, const double q[2],
                   const double *site_vec, double *cnorm);
/* intalv_xy (S_one_half/Hybrid_TPS_DSF.f90:89-193, S_arbitrary/Hybrid_TPS9_DSF.f90:91-286):
 * v = normalised sum_j e^{i q.r_j} S^{ipm}_j x, cnorm = ||.||^2 / N; ipm = +1 (the
 * reference's S+: digit j + 1) or -1.  Full space (parity_type != 0): x and v of one
 * cluster and spin (x may belong to another context); spin S > 1/2 applies the ladder
 * factor sqrt((S -/+ m)(S +/- m + 1)) of the source digit, the S=1/2 tree sets or clears
 * bit j with factor 1.  S=1/2 sector basis (parity_type = 0): x in a context with k set
 * bits, ctx and v in one with k + ipm. */
int pned_intalv_xy(pned_ctx *ctx, const pned_vec *x, pned_vec *v, const double q[2],
                   const double *site_vec, int ipm, double *cnorm);
/* parallelized_dsflnc1z (Hybrid_TPS_DSF.f90:195-261): s_maxitr-1 Lanczos steps from the
 * normalised excitation in v1; salpha[s_maxitr], sbeta[s_maxitr+1] (sbeta[0] = cnorm). */
int pned_dsflnc1z(pned_ctx *ctx, pned_vec *v1, pned_vec *v0, double cnorm, int s_maxitr,
                  double *salpha, double *sbeta);
/* dsflnc1z from step `first` with the reference's explicit update (restart after
 * interval_dump_read, Hybrid_TPS_DSF.f90:213-225, and the COMP=zr/xr extension of
 * calc_extended_DSF, :515-553, which reads SalphaSbeta_q* and vec_q* and continues from
 * first = indx + 1): first > 2 continues from v1, v0 of the top of step `first` with
 * salpha[0..first-2], sbeta[0..first-1]; runs to step s_maxitr - 1 or returns at the top of
 * step `stop` >= max(first, 2); *icf = last step taken.  Split runs equal one span bit for bit. */
int pned_dsflnc1z_span(pned_ctx *ctx, pned_vec *v1, pned_vec *v0, double cnorm, int first, int stop, int s_maxitr,
                       double *salpha, double *sbeta, int *icf);
/* exec_renbun (Hybrid_TPS_DSF.f90:263-334) on the GPU, one thread per omega, followed by
 * nomalization_DSF (337-351) when sq > 0.  w/dsf/conv have i_range entries.              */
int pned_renbun(pned_ctx *ctx, int i_range, int s_maxitr, double split, double ene, double eta,
                const double *salpha, const double *sbeta, double sq, double *w, double *dsf,
                int *conv);

/* ---- post-processing (host FP64) ---- */
/* cTPQ evaluation of mTPQ runs: the reference's evaluation program
 * (S_one_half/Hybrid_TPS_eval_mpi.f90:236-446, one process).  Inputs in the reference's
 * Fortran order: beta(i,seed) at beta[i + num_data1*seed] (alpha likewise, unused),
 * coef(i,m,seed) at coef[i + num_data1*(m + num_data2*seed)] (0-based; columns 4 and 5
 * are read).  Temperatures log-spaced on [t_min, t_max] with num_div intervals;
 * out[4*(num_div+1)] = per temperature {T, C = dE/dT, variance term, dC/dT}. */
int pned_ctpq_eval(int num_data1, int num_data2, int seed_num, const double *alpha, const double *beta,
                   const double *coef, double t_min, double t_max, int num_div, double *out);

/* ---- result and restart files (Fortran unformatted sequential records) ---- */
/* Records carry 4-byte length markers; records above the subrecord limit (default
 * 2^31-9 bytes) are split into subrecords with negated markers as gfortran documents.
 * Diagnostic (no reference counterpart): set the limit (0 = default) to exercise the
 * split at small sizes. */
int pned_io_subrecord_limit(uint64_t bytes);
/* dump_alpha_beta_coef / read_alpha_beta_coef (S_one_half/Hybrid_TPS_TPS.f90:507-583):
 * num_data1, num_data2, alpha(num_data1), beta(num_data1), coef(num_data1,num_data2) in
 * Fortran order.  read: with alpha = beta = coef = NULL returns the sizes only. */
int pned_dump_alpha_beta_coef(const char *path, int num_data1, int num_data2, const double *alpha,
                              const double *beta, const double *coef);
int pned_read_alpha_beta_coef(const char *path, int *num_data1, int *num_data2, double *alpha, double *beta,
                              double *coef, int max_data1, int max_data2);
/* DSF continued-fraction coefficients (S_one_half/Hybrid_TPS_DSF.f90:353-443 dump_alpha_beta
 * / read_alpha_beta2): sq (complex: sq[0] real, sq[1] imaginary part), salpha(n_alpha),
 * sbeta(n_beta) */
int pned_dump_dsf_alpha_beta(const char *path, const double *sq, int n_alpha, const double *salpha,
                             int n_beta, const double *sbeta);
int pned_read_dsf_alpha_beta(const char *path, double *sq, int n_alpha, double *salpha, int n_beta,
                             double *sbeta);
/* vector files (vec_q*, vec_me*: Hybrid_TPS_DSF.f90:374-379, Hybrid_TPS_TPS.f90:530-537):
 * an integer record, then one record per complex128 vector, streamed from / to device
 * memory through a pinned staging buffer */
int pned_dump_vectors(const char *path, int indx, int nvec, const pned_vec *const *vecs);
int pned_read_vectors(const char *path, int *indx, int nvec, pned_vec *const *vecs);
/* readwrite_vec_data (S_one_half/Hybrid_TPS_output.f90:12-36): e, then the vector */
int pned_basis_write(const char *path, double e, const pned_vec *x);
int pned_basis_read(const char *path, double *e, pned_vec *x);
/* interval_dump / interval_dump_read (S_one_half/Hybrid_TPS_output.f90:50-91, 147-201): the
 * checkpoint vecdump_me*: data_size, vec_size, then one record per given real array (the
 * reference's optional sa, sb, sc, se in that order, real_counts doubles each) and per
 * given vector (its optional vec0 .. vec4, streamed from / to device memory). */
int pned_interval_dump(const char *path, int32_t data_size, int32_t vec_size, int nreal,
                       const double *const *reals, const int64_t *real_counts, int nvec,
                       const pned_vec *const *vecs);
int pned_interval_dump_read(const char *path, int32_t *data_size, int32_t *vec_size, int nreal,
                            double *const *reals, const int64_t *real_counts, int nvec,
                            pned_vec *const *vecs);
/* the checkpoint's .info file (interval_dump / check_interval_dump, Hybrid_TPS_output.f90:
 * 93-146, 160-176): DUMP_NAME (character(128), blank padded), np, position(npos) (logical:
 * written as true_value -- 1 for gfortran/flang, -1 for Intel Fortran's default -- read as
 * non-zero), max_itr, dsf_comp (character(2)), s_maxitr, n, ibond, nvec, total_sz,
 * itr_counter(12).  read: dump_name gets the trimmed name (>= 129 bytes), dsf_comp 3 bytes. */
int pned_dump_info_write(const char *path, const char *dump_name, int32_t np, int npos, const int32_t *position,
                         int32_t true_value, int32_t max_itr, const char *dsf_comp, int32_t s_maxitr,
                         int32_t n, int32_t ibond, int32_t nvec, int32_t total_sz, const int32_t *itr_counter);
int pned_dump_info_read(const char *path, char *dump_name, int32_t *np, int npos, int32_t *position,
                        int32_t *max_itr, char *dsf_comp, int32_t *s_maxitr, int32_t *n, int32_t *ibond,
                        int32_t *nvec, int32_t *total_sz, int32_t *itr_counter);
/* write_dsf_res (Hybrid_TPS_output.f90:243-262): one momentum's block of dsf_res_<comp>.dat*
 * text -- '#' line with sq (F25.18), ll rows (2F18.10, 2F25.16, I5) of q, w, dsf, conv, four
 * list-directed blank lines; append = 1 adds to the file (the reference keeps it open). */
int pned_write_dsf_res(const char *path, int append, double sq, const double q[2], int ll, const double *w,
                       const double *dsf, const int32_t *conv);

#ifdef __cplusplus
}
#endif
#endif /* PNED_GPU_H */
