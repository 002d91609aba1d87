// internal.h -- shared declarations of libpned_gpu (not part of the ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/pned_gpu.h"

namespace pned {

// ------------------------------------------------------------------------------------
// error plumbing: every ABI entry returns a status; the message is kept per thread
// ------------------------------------------------------------------------------------
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define PNED_HIP(call)                                                                    \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            return ::pned::fail(PNED_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define PNED_TRY(expr)                                                                    \
    do {                                                                                  \
        int rc_ = (expr);                                                                 \
        if (rc_ != PNED_OK) return rc_;                                                   \
    } while (0)

// ------------------------------------------------------------------------------------
// S=1/2 H.v plan: the operator decomposed into a real-or-complex diagonal and flip terms,
// grouped into tile passes.  Memory index m = (state << colbits) | column.
// ------------------------------------------------------------------------------------
struct Term {            // one off-diagonal term: v0(i) += c(sel) * v1(i ^ gmask)
    uint32_t gmask;      // flipped state bits
    uint32_t selmask;    // coefficient selector = parity of (state & selmask)
    double c0r, c0i;     // coefficient when selector = 0
    double c1r, c1i;     // coefficient when selector = 1
};

struct DiagGroup {       // D(i) += w * popcount(f(i))
    int kind;            // 0: f = i & mask ; 1: f = (i ^ (i >> shift)) & mask
    int shift;
    uint32_t mask;
    double wr, wi;
};

// device-side term of one pass (tile-local partner mask precomputed)
struct DevTerm {
    uint32_t tmask;      // partner offset inside the tile (tile-local bit mask)
    uint32_t selmask;
    uint32_t is_real;    // coefficients purely real (fast path)
    uint32_t pad;
    double c0r, c0i, c1r, c1i;
};

// A term as the term-major pass kernel reads it, in the coordinates of one pass's tile
// (tile slot j = thread + TPB * register): the flip mask and the selector split into
// thread bits (lo), register bits (hi) and, for the selector, outer memory bits.
struct TmTerm {
    uint32_t mlo, mhi;   // flip mask: thread bits, register bits (>> log2 TPB)
    uint32_t slo, shi;   // selector mask: thread bits, register bits
    uint64_t sout;       // selector mask on the outer memory-index bits
    uint32_t flags;      // bit 0: complex coefficients; bit 1: selector has register bits
    uint32_t term;       // index into the pass-ordered DevTerm table (batch coefficients)
    double c0r, c0i, c1r, c1i;
};

struct PassPlan {
    int K;                   // tile bits (over memory-index bits)
    std::vector<int> tb;     // tile bit positions (memory index), ascending
    std::vector<int> ob;     // outer bit positions (memory index), ascending
    int first_term, nterm;   // range into the concatenated DevTerm table
    bool diag;               // this pass adds the diagonal (and the v0 input)
};

struct S12Plan {
    int n = 0, colbits = 0;
    std::vector<Term> terms;
    std::vector<DiagGroup> diag;
    double c0r = 0, c0i = 0;     // constant part of the diagonal
    std::vector<PassPlan> passes;
    std::vector<DevTerm> devterms;
    // batch contexts: one operator per column (same cluster, different couplings)
    int nop = 1;
    std::vector<double> colcoef;  // [term][nop][4] (c0r,c0i,c1r,c1i), plan.terms order
    std::vector<double> coldiag;  // [nop][ngroups+1][2]: group weights, then the constant
    std::vector<double> devcoef;  // [devterm][nop][4], pass order
};

// batch decomposition: nop operators over one cluster -> common term structure and
// diagonal groups (the common refinement of every operator's weight classes)
int s12_decompose_batch(int n, int ibond, const int32_t *ipair, const int32_t *btype, int nop,
                        const double *const *xl, const int32_t *const *ll, S12Plan &plan,
                        const int8_t *pos = nullptr);
// replace the coefficients of one column; the operator must fit the existing structure
int s12_set_column(S12Plan &plan, int ibond, const int32_t *ipair, const int32_t *btype, int col,
                   const double *xl, const int32_t *ll, const int8_t *pos = nullptr);

// build the S=1/2 decomposition of the reference tables (no passes yet); pos: memory bit
// of each site (null = identity layout)
int s12_decompose(int n, int ibond, const int32_t *ipair, const int32_t *btype,
                  const double *xl_fortran, const int32_t *ll_fortran, S12Plan &plan,
                  const int8_t *pos = nullptr);
// group terms into tile passes for a given memory layout
int s12_make_passes(S12Plan &plan, int colbits, int tile_bits, int low_bits);

// Memory layout of a full-space S=1/2 context: the memory index of basis state s is
// swap(s), which exchanges nswap disjoint pairs of bits (an involution, so the same map
// takes memory indices back to basis states).  Site k (0-based) sits at memory bit pos[k].
// Identity unless the planner finds fewer tile passes with other sites in the lowest bits.
constexpr int LAYOUT_MAX_SWAPS = 4;
struct Layout {
    int nswap = 0;
    int a[LAYOUT_MAX_SWAPS] = {0, 0, 0, 0};
    int b[LAYOUT_MAX_SWAPS] = {0, 0, 0, 0};
    int8_t pos[32];
    Layout() { for (int k = 0; k < 32; ++k) pos[k] = (int8_t)k; }
    bool identity() const { return nswap == 0; }
    bool same(const Layout &o) const
    {
        for (int k = 0; k < 32; ++k)
            if (pos[k] != o.pos[k]) return false;
        return true;
    }
};
// the swaps as a kernel argument (bit positions within the state part of the index)
struct SwapArgs {
    int n;
    uint32_t a[LAYOUT_MAX_SWAPS], b[LAYOUT_MAX_SWAPS];
};
inline SwapArgs swap_args(const Layout &l)
{
    SwapArgs s{};
    s.n = l.nswap;
    for (int k = 0; k < l.nswap; ++k) {
        s.a[k] = (uint32_t)l.a[k];
        s.b[k] = (uint32_t)l.b[k];
    }
    return s;
}
__host__ __device__ inline uint64_t swap_bits(const SwapArgs &s, uint64_t x)
{
    for (int k = 0; k < s.n; ++k) {
        const uint64_t t = ((x >> s.a[k]) ^ (x >> s.b[k])) & 1ull;
        x ^= (t << s.a[k]) | (t << s.b[k]);
    }
    return x;
}
// builds the layout that puts the given sites (at most 4, distinct, 0-based) into memory
// bits 0, 1, ... in that order of bits; sites already there stay
int layout_from_low_sites(int n, const int *sites, int nsites, Layout &out);
// layout from a site -> memory bit table (must be an involution with <= 4 swapped pairs)
int layout_from_table(int n, const int32_t *site_bit, Layout &out);
// chooses the layout of a new single-operator S=1/2 context from its decomposition (in
// the identity layout): the sites in the forced low tile bits that give the fewest tile
// passes at 2^11..2^13-state tiles; identity when nothing is gained
Layout choose_layout(const S12Plan &identity_plan);

struct DiagDev {             // flattened diagonal for the kernel (constant per context)
    int ngroups;
    int kind[48];
    int shift[48];
    uint32_t mask[48];
    double wr[48], wi[48];
    double c0r, c0i;
    // The same groups as a lane program for the term-major kernel, read with readlane so the
    // per-state evaluation waits on no memory.  Lane e < nprog of prog_a: shift (bits 0-5),
    // pair parity (bit 6), last group of its weight class (bit 7), class (bits 8-10); of
    // prog_b: the mask.  Groups are ordered by class.  Lanes 48 + c and 56 + c of prog_a hold
    // the low and high word of class c's real weight, of prog_b of its imaginary weight.
    // nprog = 0: not available (more than 48 groups or 8 weight classes).
    int nprog;
    uint32_t prog_a[64], prog_b[64];
    // nrest > 0 (the term-major kernel's diagonal pass): the program split by the pass's
    // register bits -- lanes 0..nprog-1 keep, per class, the bit positions that involve a
    // register bit (every class ends there, with an empty mask where it has none), lanes
    // nprog..nprog+nrest-1 the positions that do not, which are the same for all register
    // slots of a thread and are counted once per tile (at most 4 classes: 16-bit counts)
    int nrest;
};

// ------------------------------------------------------------------------------------
// context and vectors
// ------------------------------------------------------------------------------------
struct GpuTables {           // spin-S tables in device memory
    int32_t *pairs = nullptr;    // 2*ibond (0-based min,max)
    int32_t *btype = nullptr;    // ibond (0-based)
    double *xl = nullptr;        // [9][qq][qq] complex, [t][in][out]
    int32_t *mk = nullptr;       // [9][qq][qq]
};

}  // namespace pned

struct pned_vec {
    pned_ctx *ctx = nullptr;
    int64_t dim = 0;
    int nvec = 1;
    int dtype = PNED_C128;
    void *data = nullptr;        // device
    size_t bytes = 0;
};

// a vector whose context was destroyed keeps its memory but takes part in no operation
inline bool vec_detached(const pned_vec *v) { return v != nullptr && v->ctx == nullptr; }

struct pned_ctx {
    int device = 0;
    std::vector<pned_vec *> vecs;    // live vectors; detached (ctx = nullptr) when the context goes
    int n = 0, q = 2, ibond = 0;
    int64_t dim = 0;
    hipStream_t stream = nullptr;
    std::vector<int32_t> ipair, btype;
    std::vector<double> xl;      // Fortran order copy
    std::vector<int32_t> ll;
    std::vector<int32_t> site_pair;   // 0-based (min,max) per bond
    // S=1/2
    pned::S12Plan s12;
    pned::DiagDev diag_host;
    // per-(colbits, dtype) pass layouts are rebuilt on demand
    int tune_tile_bits = 0, tune_low_bits = 0, tune_threads = 0;
    int plan_colbits = -1, plan_dtype = -1;
    pned::DevTerm *d_terms = nullptr;
    size_t d_terms_cap = 0;
    pned::DiagDev *d_diag = nullptr;
    pned::DiagDev *d_diag_tm = nullptr;     // diagonal program split by the diag pass's register bits
    void *pd_dev = nullptr;          // per-pass address tables of the current plan
    pned::TmTerm *d_tm = nullptr;    // per-pass tile-coordinate terms (term-major kernel)
    std::vector<int> tm_ncat;        // [pass][4] terms per term-major category
    int s12_kernel = 1;              // 1: term-major pass kernel, 0: state-major pass kernel
    // one-column lanczos_batch_step with deferred normalisation: which handle holds x
    // (v1/sx) and the two factors the next step applies
    int lz_swapped = 0;
    double lz_sx = 1.0, lz_sy = 0.0;
    // memory layout of the full-space S=1/2 vectors (identity for every other context)
    pned::Layout lay;
    double *d_stage = nullptr;       // device staging of layout-mapped uploads/downloads
    size_t stage_bytes = 0;
    // S^z sector basis (parity_type = 0); sector_k < 0: full 2^N space
    int sector_k = -1;
    int sector_nterm = 0;
    uint32_t *d_list1 = nullptr;     // index -> state
    uint32_t *d_srank = nullptr;     // rank of a right half among its popcount class
    uint32_t *d_soffset = nullptr;   // first index of a left half's block
    void *d_sterms = nullptr;        // two-bit flip terms
    double *d_colcoef = nullptr;     // batch contexts: per-column term coefficients
    double *d_coldiag = nullptr;     // batch contexts: per-column diagonal weights
    int pd_threads = -1;
    int pd_gap = -1;                 // column-group bits the address tables were built for
    int plan_req_colbits = -1;       // column bits of the vectors the current plan was chosen for
    // spin-S
    pned::GpuTables spin;
    int spin_tiled = 1;              // tiled kernel (q = 3..5); 0 = per-state gather kernel
    int num_cu = 0;                  // compute units of the device (read on first use)
    int use_graphs = 1;              // Krylov loops replay captured HIP graphs of their steps
    void *spin_pd = nullptr;         // tiled-pass tables
    int spin_K = -1, spin_np = 0;
    std::vector<int> spin_nb, spin_tsz, spin_nlo, spin_ncls, spin_cplx;
    // scratch for reductions: per-block partials and final results (pinned host mirror)
    double *d_partial = nullptr;
    size_t partial_cap = 0;          // in doubles
    double *d_result = nullptr;
    double *h_result = nullptr;      // pinned
    size_t result_cap = 0;           // in doubles
    double *d_scalars = nullptr;     // small per-call scalar/argument uploads
    size_t scalars_cap = 0;          // in doubles
    // mTPQ observables
    std::vector<double> site_vec;
    std::vector<double> sq_q;
    int nsq = 0;
};

namespace pned {
// the H.v operand checks: both vectors of this context, same shape, the context's
// dimension, and two different buffers (every entry that launches an H.v calls it first)
int check_hv_pair(const pned_ctx *ctx, const pned_vec *v1, const pned_vec *v0);
// scratch management
int ensure_partial(pned_ctx *ctx, size_t ndoubles);
int ensure_result(pned_ctx *ctx, size_t ndoubles);
int upload_scalars(pned_ctx *ctx, const double *h, int n, double **dptr);
int ensure_scalars(pned_ctx *ctx, int n, double **dptr);
// v0 += H v1; pr = <v1|H v1> (2 per column), b2 = |v0 - pr v1|^2: one host readback
int hv_alpha_beta2(pned_ctx *ctx, pned_vec *v1, pned_vec *v0, double *pr, double *b2);
// the beta^2 half of it: alpha taken from ctx->d_result where the H.v left it
int beta2_after_hv(pned_ctx *ctx, const pned_vec *v1, const pned_vec *v0, double *pr, double *b2);
// One Lanczos step with deferred normalisation (S=1/2 full space): y = sy y + H(sx x), then
// y <- y - Re(alpha) sx x with alpha = sx <x|H|sx x>, and b2 = |y_before - alpha sx x|^2;
// pr = <x|H|sx x> (2 per column).  fresh: y is taken as zero (sy ignored).
int lanczos_lazy_step(pned_ctx *ctx, pned_vec *x, pned_vec *y, const double *sx, const double *sy, bool fresh,
                      double *pr, double *b2);
// COCG: x += a p, r -= a y; out4 = <r|r> (re, im), r^T r (re, im) of the new r
int cocg_update(pned_ctx *ctx, const pned_vec *p, const pned_vec *y, pned_vec *x, pned_vec *r, const double *a,
                double *out4);
// after such steps: v1 <- sx x, v0 <- y + a sx x, (x, y) = (v1, v0) or, swapped, (v0, v1)
int lanczos_lazy_restore(pned_ctx *ctx, pned_vec *v1, pned_vec *v0, bool swapped, double sx, double a);
// lanczos_lazy_step's enqueue-only form (one column, S=1/2 full space), which Lanczos
// loops capture into graphs: st = [sx, sy, prd re, prd im, b2 of the step before] on the
// device; the step sets sx = 1/sqrt(b2), sy = -sqrt(b2) sx (fresh: 1, 0), runs the step and
// writes [prd re, prd im, b2] to slot and b2 to st[4].  No allocation or host wait once
// lanczos_lazy_setup (which also checks the pair) and one un-captured step of the same
// shape have run.
int lanczos_lazy_setup(pned_ctx *ctx, const pned_vec *x, const pned_vec *y);
int lanczos_lazy_enqueue(pned_ctx *ctx, pned_vec *x, pned_vec *y, bool fresh, double *st, double *slot);
int ensure_s12_plan(pned_ctx *ctx, int colbits, int dtype);
// the plan for 2^colbits interleaved columns: with column groups (tiles holding run_bits
// column bits, the higher ones fixed per launch) when that takes fewer passes than all
// columns in the tile (single-operator contexts); *cg = the plan's column bits.  The choice
// is kept until the plan is rebuilt.
int ensure_s12_plan_cols(pned_ctx *ctx, int colbits, int run_bits, int dtype, int *cg);
// dh's lane program split by register bits (slot_bits: state bits held in register slots,
// in the diagonal's state coordinates); false when it does not fit (more than 4 classes or
// 48 lanes), then out is unchanged
bool split_diag_slots(const DiagDev &dh, uint32_t slot_bits, DiagDev &out);
// returns the context to the identity layout (re-decomposes its operator); only while no
// vector of the context holds data in the old layout matters to the caller
int reset_layout(pned_ctx *ctx);
// vectors of two contexts can exchange data only in one memory layout
int check_same_layout(const pned_vec *x, const pned_vec *y);
// layout-mapped copies of count states (all columns) starting at basis state offset
int vec_to_host(const pned_vec *v, void *host, int64_t offset, int64_t count);
int vec_from_host(pned_vec *v, const void *host, int64_t offset, int64_t count);
// S^z sector basis: H.v and table release
int launch_hv_sector(pned_ctx *ctx, const pned_vec *v1, pned_vec *v0, double eperbd, bool fresh,
                     double *prdct_host);
void free_sector(pned_ctx *ctx);
// intalv_xy, S=1/2 sectors: v (sector k + ipm of ctx) = sum_j e^{i q.r_j} (x with bit j
// set (ipm = +1) or cleared (-1)); x in sector k of x->ctx
int intalv_xy_apply(pned_ctx *ctx, const pned_vec *x, pned_vec *v, const double *qv, const double *site_vec,
                    int ipm);
// intalv_xy over the full q^N space: v = sum_j e^{i q.r_j} S^{ipm}_j x (ladder factors)
int intalv_pm_apply(pned_ctx *ctx, const pned_vec *x, pned_vec *v, const double *qv, const double *site_vec, int ipm);
int refresh_sector_terms(pned_ctx *ctx);   // after the bond matrices change
// untuned tile defaults: true for vectors above 2^25 elements (dim x columns)
bool default_large(const pned_ctx *ctx, int colbits);
int upload_batch_coefficients(pned_ctx *ctx);
// sums the rows of ctx->d_partial [nrows][ncol2] into ctx->d_result[0..ncol2): 256 threads
// per column, fixed strides and a fixed tree (the same partials give the same sums)
int reduce_rows_result(pned_ctx *ctx, int nrows, int ncol2);

// kernels (launchers), all asynchronous on ctx->stream unless stated
// d_colscale: optional device array of per-column factors s_c: v0 += (H - shift)(s_c v1_c);
// prdct is then <v1|H|s v1> (one factor s)
// d_yscale: optional per-column factors y_c on the v0 input: v0 = y_c v0 + (H - shift)(s_c v1)
// lz_fused (deferred-normalisation Lanczos step, one complex128 column): when not null and
// the last pass runs on the term-major kernel, that pass also forms v0 - Re(alpha) s v1;
// then d_result holds <v1|H s v1> (2 doubles) followed by [beta^2, 0] and *lz_fused = true
// resolves the device address of the allocation-free H.v's static scratch on one device
// (once per device; called at context creation with that device current)
int noalloc_scratch_init(int device);
int launch_hv_s12(pned_ctx *ctx, const pned_vec *v1, pned_vec *v0, double eperbd, bool fresh,
                  double *prdct_host, const double *d_colscale = nullptr, const double *d_yscale = nullptr,
                  bool *lz_fused = nullptr);
int launch_hv_spin(pned_ctx *ctx, const pned_vec *v1, pned_vec *v0, double eperbd, bool fresh,
                   double *prdct_host);
// the tiled spin-S kernel (q = 3..5) with v0 = ys v0 + xs (H - shift) v1 (ys on the first
// pass's input; <v1|.> carries the factor xs)
int launch_hv_spin_tiled(pned_ctx *ctx, const pned_vec *v1, pned_vec *v0, double eperbd, bool fresh,
                         double *prdct_host, double xs, double ys);
// the same with v0 = ys v0 + xs (H - shift + z) v1, z = zr + i zi, and with dotu the
// returned sum is v1^T (.) instead of v1^H (.) (shifted complex-symmetric solves)
int launch_hv_spin_tiled_shift(pned_ctx *ctx, const pned_vec *v1, pned_vec *v0, double eperbd, bool fresh,
                               double *prdct_host, double xs, double ys, double zr, double zi, bool dotu);

// reductions return host values (synchronizing)
int dot(pned_ctx *ctx, const pned_vec *x, const pned_vec *y, double *out /*2*nvec*/);
int dotu(pned_ctx *ctx, const pned_vec *x, const pned_vec *y, double *out /*2*nvec, no conjugate*/);
int beta2(pned_ctx *ctx, const double *alpha /*2*nvec complex*/, const pned_vec *v1,
          const pned_vec *v0, double *out /*nvec*/);
int axpby(pned_ctx *ctx, const double *a, const pned_vec *x, const double *b, pned_vec *y);
int scale(pned_ctx *ctx, pned_vec *x, const double *a);
int lanczos_update(pned_ctx *ctx, pned_vec *v1, pned_vec *v0, const double *alpha,
                   const double *beta);
int tps_shift_update(pned_ctx *ctx, pned_vec *v1, pned_vec *v0, pned_vec *vg, const double *xn);
int lanczos_slots(pned_ctx *ctx, pned_vec *v1, pned_vec *v0, const double *alpha, const double *beta,
                  const int *mode);
int gaussian_column(pned_ctx *ctx, pned_vec *v, int col, uint64_t seed);
int gaussian_fill(pned_ctx *ctx, pned_vec *v, uint64_t seed);
int sz_q_apply(pned_ctx *ctx, const pned_vec *x, pned_vec *v, const double *q, const double *site_vec);
int tps_observables(pned_ctx *ctx, const pned_vec *vg, const pned_vec *v1, double *out);
int tps_fused_reduce(pned_ctx *ctx, const pned_vec *up, const pned_vec *u, const pned_vec *w, double *out);
// fused mTPQ reduction outputs per column (10 + 2 x at most 8 structure-factor entries)
constexpr int TPS_RED_WIDTH = 26;
// the captured form of the fused mTPQ step: a device copy of the observable tables
// (*dobs, freed by the caller with hipFree), then per step, enqueued only (no allocation, no
// host round trip): su = 1/sqrt(red_prev[c][0]) (1 when red_prev is null), w = (L-H)(su u),
// slot[0..2nc) = the H.v's per-column product (what pned_tps_fused reads back),
// slot[2nc..) = the reduction of (up, u, w) ([column][TPS_RED_WIDTH])
int tps_graph_setup(pned_ctx *ctx, const pned_vec *u, void **dobs, int *nsq);
int tps_graph_step(pned_ctx *ctx, const pned_vec *up, const pned_vec *u, pned_vec *w, const void *dobs, int nsq,
                   const double *red_prev, double *su_dev, double *slot);
int renbun_gpu(pned_ctx *ctx, int i_range, int s_maxitr, double split, double ene, double eta,
               const double *salpha, const double *sbeta, double *w, double *dsf, int *conv);

// tridiagonal T (alpha, beta of length ndim, ndim - 1): b2 = beta^2 and the bisection's
// start bracket (Gershgorin bounds widened by round-off), stop width and pivot floor
struct TridiagBounds {
    double lo, hi, width, pivmin;
};
int tridiag_bounds(const double *alpha, const double *beta, int ndim, double eps, std::vector<double> &b2,
                   TridiagBounds &tb);
// inverse iteration's cluster starts (c0[k]: first eigenvalue of k's cluster, eigenvalues
// closer than 1e-10 of the spectrum's scale) and pivot floor
int tridiag_clusters(const double *alpha, const double *beta, int n, const double *E, int nvec,
                     std::vector<int> &c0, double &tiny);

inline size_t elem_bytes(int dtype) { return dtype == PNED_C128 ? 16 : 8; }
}  // namespace pned
