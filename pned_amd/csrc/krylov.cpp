// krylov.cpp -- Krylov drivers over the device H.v: Lanczos, inverse iteration (CG),
// COCG, mTPQ, DSF.  Vectors stay on the GPU; only FP64 scalars (alpha, beta, dots) come
// back to the host, where the tridiagonal algebra runs.
#include <algorithm>
#include <cmath>
#include <limits>
#include <complex>
#include <cstring>

#include "internal.h"
#include "tridiag.h"

using namespace pned;

namespace {

double cabs2(const double *z) { return std::sqrt(z[0] * z[0] + z[1] * z[1]); }

// The Lanczos recursion of lnc1z and dsflnc1z (Hybrid_TPS_Lanczos.f90:111-120,
// Hybrid_TPS_DSF.f90:228-250): v1 <- (v0 - alpha v1)/beta, v0 <- -beta v1, then
// v0 += H v1, alpha = Re <v1|H v1>, beta = |v0 - <v1|H v1> v1|.  In the S=1/2 full space
// (and with the tiled spin-S kernel) the divisions by beta are left to the next H.v: x
// holds v1/sx, the H.v input y is read times sy, and the two buffers swap roles each step,
// so a step streams the vectors 3 times besides the H.v instead of 6.  Other bases run the
// reference's update as written.
// finish() leaves v1 and v0 as the reference's loop leaves them (the last Lanczos vector
// and H v1 - beta v_prev), so a dump for restart holds the same vectors.
// With graphs on (one S=1/2 full-space column), the steps are enqueued in their device form:
// each step forms its sx, sy from the previous step's beta^2 on the device and leaves
// [prd, beta^2] in a history slot, so advance() replays 5 steps at a time from a captured
// graph (one per buffer role) and reads the 5 slots back once.  The host repeats the same
// sx, sy arithmetic to form alpha = sx prd and beta = sqrt(beta^2), so both forms give the
// same alpha/beta bit for bit.
struct LanczosSteps {
    static constexpr int GRAPH_STEPS = 5;
    pned_ctx *ctx;
    pned_vec *v1, *v0;
    bool lazy, dev;
    pned_vec *X, *Y;
    double sx = 1.0, sy = 0.0;
    double *st = nullptr;     // device: [sx, sy, prd re, prd im, beta^2], then GRAPH_STEPS slots of 3
    double *host = nullptr;   // pinned read-back of the slots
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    // explicit: the reference's update on the true vectors at every step (no deferred
    // normalisation), so the loop state can be checkpointed and resumed bit for bit
    LanczosSteps(pned_ctx *c, pned_vec *a, pned_vec *b, bool explicit_update = false)
        : ctx(c), v1(a), v0(b),
          lazy(!explicit_update && a->nvec == 1 &&
               ((c->q == 2 && c->sector_k < 0) ||
                (c->q >= 3 && c->q <= 5 && c->spin_tiled && a->dtype == PNED_C128))),
          dev(lazy && c->q == 2 && c->use_graphs),
          X(a), Y(b)
    {
    }
    LanczosSteps(const LanczosSteps &) = delete;
    LanczosSteps &operator=(const LanczosSteps &) = delete;
    ~LanczosSteps()
    {
        for (int k = 0; k < 2; ++k) {
            if (exec[k]) (void)hipGraphExecDestroy(exec[k]);
            if (graph[k]) (void)hipGraphDestroy(graph[k]);
        }
        if (host) (void)hipHostFree(host);
        if (st) (void)hipFree(st);
    }
    int lazy_ab(bool fresh, double &a1, double &b1)
    {
        double pr[2], b2;
        PNED_TRY(lanczos_lazy_step(ctx, X, Y, &sx, &sy, fresh, pr, &b2));
        a1 = sx * pr[0];
        b1 = std::sqrt(b2);
        return PNED_OK;
    }
    // the host's copy of the device's role swap and scales for the next step
    void roll(double b1)
    {
        const double nsx = 1.0 / b1, nsy = -b1 * sx;   // v1 is Y / b1, v0 is X times -b1 sx
        std::swap(X, Y);
        sx = nsx;
        sy = nsy;
    }
    // alpha, beta of the n steps whose slots were just filled (roll() before each but the
    // first step of a run); b1 carries the beta before them
    int harvest(int n, bool fresh, double &b1, double *a_out, double *b_out)
    {
        PNED_HIP(hipMemcpyAsync(host, st + 5, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, ctx->stream));
        PNED_HIP(hipStreamSynchronize(ctx->stream));
        for (int j = 0; j < n; ++j) {
            if (!(fresh && j == 0)) roll(b1);
            a_out[j] = sx * host[3 * j];
            b1 = std::sqrt(host[3 * j + 2]);
            b_out[j] = b1;
        }
        return PNED_OK;
    }
    // v0 = H v1 from v1 (normalised) and alpha, beta
    int first(double &a1, double &b1)
    {
        if (dev) {
            PNED_TRY(lanczos_lazy_setup(ctx, X, Y));
            PNED_HIP(hipMalloc((void **)&st, sizeof(double) * (5 + 3 * GRAPH_STEPS)));
            PNED_HIP(hipHostMalloc((void **)&host, sizeof(double) * 3 * GRAPH_STEPS, hipHostMallocDefault));
            PNED_TRY(lanczos_lazy_enqueue(ctx, X, Y, true, st, st + 5));
            return harvest(1, true, b1, &a1, &b1);
        }
        if (lazy) return lazy_ab(true, a1, b1);
        PNED_TRY(pned_vec_zero(ctx, v0));
        return pned_calc_alpha_beta(ctx, v1, v0, &a1, &b1);
    }
    // the update with the previous (a1, b1), then the next alpha, beta into them
    int next(double &a1, double &b1)
    {
        if (dev) {
            // enqueued with the roles after the swap that harvest's roll() mirrors
            PNED_TRY(lanczos_lazy_enqueue(ctx, Y, X, false, st, st + 5));
            return harvest(1, false, b1, &a1, &b1);
        }
        if (lazy) {
            roll(b1);
            return lazy_ab(false, a1, b1);
        }
        PNED_TRY(pned_lanczos_update(ctx, v1, v0, &a1, &b1));
        return pned_calc_alpha_beta(ctx, v1, v0, &a1, &b1);
    }
    int capture(int role)
    {
        // the buffers the captured kernels address must stay where the first step left them
        const void *keep[4] = {ctx->d_partial, ctx->d_result, ctx->pd_dev, ctx->d_tm};
        PNED_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
        pned_vec *x = X, *y = Y;
        int rc = PNED_OK;
        for (int j = 0; j < GRAPH_STEPS && rc == PNED_OK; ++j) {
            std::swap(x, y);
            rc = lanczos_lazy_enqueue(ctx, x, y, false, st, st + 5 + 3 * j);
        }
        const hipError_t ec = hipStreamEndCapture(ctx->stream, &graph[role]);
        PNED_TRY(rc);
        if (ec != hipSuccess) return fail(PNED_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
        const void *now[4] = {ctx->d_partial, ctx->d_result, ctx->pd_dev, ctx->d_tm};
        if (std::memcmp(keep, now, sizeof(keep)) != 0)
            return fail(PNED_E_HIP, "Lanczos graph: a work buffer moved while the steps were captured");
        PNED_HIP(hipGraphInstantiate(&exec[role], graph[role], nullptr, nullptr, 0));
        return PNED_OK;
    }
    // up to n more steps (next() n times): alpha, beta of each into a_out, b_out; a1, b1
    // in/out as in next(); *ran = the steps taken.  Whole groups of GRAPH_STEPS replay a captured
    // graph in the device form; one by one, with stop_at_breakdown a step with beta <
    // 0.5e-30 ends the run early.
    int advance(int n, double &a1, double &b1, double *a_out, double *b_out, int *ran, bool stop_at_breakdown)
    {
        int k = 0;
        *ran = 0;
        while (k < n) {
            if (stop_at_breakdown && k > 0 && b_out[k - 1] < 0.5e-30) break;
            if (dev && n - k >= GRAPH_STEPS) {
                const int role = X != v1;
                if (!exec[role]) PNED_TRY(capture(role));
                PNED_HIP(hipGraphLaunch(exec[role], ctx->stream));
                PNED_TRY(harvest(GRAPH_STEPS, false, b1, a_out + k, b_out + k));
                k += GRAPH_STEPS;
                a1 = a_out[k - 1];
            } else {
                PNED_TRY(next(a1, b1));
                a_out[k] = a1;
                b_out[k] = b1;
                ++k;
            }
            *ran = k;
        }
        return PNED_OK;
    }
    // a1: the last alpha
    int finish(double a1)
    {
        if (!lazy) return PNED_OK;
        return lanczos_lazy_restore(ctx, v1, v0, X != v1, sx, a1);
    }
};

}  // namespace

namespace pned {

int tridiag_bounds(const double *alpha, const double *beta, int ndim, double eps, std::vector<double> &b2,
                   TridiagBounds &tb)
{
    if (!alpha || !beta || ndim < 1) return fail(PNED_E_ARG, "null argument");
    if (ndim > 8000) return fail(PNED_E_ARG, "#(E04) ndim given to bisec exceeds 8000");
    b2.assign(ndim > 1 ? ndim - 1 : 1, 0.0);
    double lo = alpha[0], hi = alpha[0], b2max = 0.0;
    for (int i = 0; i < ndim; ++i) {
        const double r = (i > 0 ? std::fabs(beta[i - 1]) : 0.0) + (i < ndim - 1 ? std::fabs(beta[i]) : 0.0);
        lo = std::min(lo, alpha[i] - r);   // Gershgorin discs
        hi = std::max(hi, alpha[i] + r);
        if (i < ndim - 1) {
            b2[i] = beta[i] * beta[i];
            b2max = std::max(b2max, b2[i]);
        }
    }
    const double scale = std::max(std::fabs(lo), std::fabs(hi));
    tb.pivmin = std::max(std::numeric_limits<double>::min(), b2max * std::numeric_limits<double>::min());
    tb.width = std::max(scale * (eps > 0.0 ? eps : 0.0), 0.0);
    tb.lo = lo - (2.0 * std::numeric_limits<double>::epsilon() * scale + tb.pivmin);
    tb.hi = hi + (2.0 * std::numeric_limits<double>::epsilon() * scale + tb.pivmin);
    return PNED_OK;
}

int tridiag_clusters(const double *alpha, const double *beta, int n, const double *E, int nvec,
                     std::vector<int> &c0, double &tiny)
{
    double scale = 0.0;
    for (int i = 0; i < n; ++i)
        scale = std::max(scale, std::fabs(alpha[i]) + (i < n - 1 ? std::fabs(beta[i]) : 0.0) +
                                    (i > 0 ? std::fabs(beta[i - 1]) : 0.0));
    if (scale == 0.0) scale = 1.0;
    tiny = std::numeric_limits<double>::epsilon() * scale;
    c0.assign(nvec, 0);
    for (int k = 1; k < nvec; ++k) c0[k] = (std::fabs(E[k] - E[k - 1]) > 1e-10 * scale) ? k : c0[k - 1];
    return PNED_OK;
}

}  // namespace pned

extern "C" {

// Lowest eigenvalues and eigenvectors of the Lanczos tridiagonal T (diagonal alpha[0..n),
// off-diagonal beta[0..n-1)).  The entry points keep the names, arguments and errors of the
// reference's bisec / vec12 (S_one_half/Hybrid_TPS_basic.f90:25-213); the arithmetic is this
// library's own: Sturm counts from the pivots of the LDL^T factorisation of T - x I, each
// eigenvalue bisected down to adjacent doubles, eigenvectors by inverse iteration through a
// pivoted LU of T - E I.  The stop step of lnc1z (a 1e-13 relative change test) depends on
// rounding at that level, so it can differ from the reference's by a check or two.

// bisec (the reference's interface): the ne lowest eigenvalues, ascending; eps is the
// reference's relative width at which bisection may stop (below round-off, the bisection
// runs until the bracket holds two adjacent doubles).  Each eigenvalue is bisected from the
// Gershgorin bounds on its own, so the eigenvalues are independent lanes on the device; a
// bisection started elsewhere ends in the same adjacent-double bracket, or, where the stop
// width eps*scale is reached first (eigenvalues near zero), within that width of it
int pned_bisec(const double *alpha, const double *beta, int ndim, double *E, int ne, double eps)
{
    if (!alpha || !beta || !E || ndim < 1) return fail(PNED_E_ARG, "null argument");
    if (ne > ndim || ne < 0) return fail(PNED_E_ARG, "#(E05) ne given to bisec out of range");
    std::vector<double> b2;
    TridiagBounds tb;
    PNED_TRY(tridiag_bounds(alpha, beta, ndim, eps, b2, tb));
    for (int k = 1; k <= ne; ++k) E[k - 1] = bisect_kth(alpha, b2.data(), ndim, k, tb.lo, tb.hi, tb.width, tb.pivmin);
    return PNED_OK;
}

// vec12 (the reference's interface): eigenvectors of the eigenvalues E[0..nvec) in v (nvec
// columns of length ndim), unit norm.  Inverse iteration, three solves from a fixed start;
// vectors of eigenvalues closer than 1e-10 (relative to the spectrum's scale) to an earlier
// one are kept orthogonal to that cluster after every solve.
int pned_vec12(const double *alpha, const double *beta, int ndim, const double *E, int nvec, double *v)
{
    if (!alpha || !beta || !E || !v || ndim < 1 || nvec < 1) return fail(PNED_E_ARG, "bad argument");
    const int n = ndim;
    std::vector<int> c0;
    double tiny;
    PNED_TRY(tridiag_clusters(alpha, beta, n, E, nvec, c0, tiny));
    std::vector<double> dl(n), d(n), du(n), du2(n);
    std::vector<char> sw(n);
    for (int k = 0; k < nvec; ++k)
        inverse_iteration(alpha, beta, n, E[k], k, c0[k], tiny, v, dl.data(), d.data(), du.data(), du2.data(),
                          sw.data());
    return PNED_OK;
}

// parallelized_lnc1z (S_one_half/Hybrid_TPS_Lanczos.f90:48-163)
int pned_lnc1(pned_ctx *ctx, pned_vec *v1, pned_vec *v0, int max_itr, int nvec, int fixed_steps,
              double *alpha, double *beta, double *ene, int *itr)
{
    if (!ctx || !v1 || !v0 || !alpha || !beta || !ene || !itr) return fail(PNED_E_ARG, "null argument");
    if (nvec < 0) return fail(PNED_E_ARG, "#(E06) Wrong value given to nvec in lnc1");
    if (v1->nvec != 1) return fail(PNED_E_ARG, "lnc1 runs on one column");
    const int last = fixed_steps > 0 ? fixed_steps : max_itr;
    double a1, b1;
    LanczosSteps run(ctx, v1, v0);
    PNED_TRY(run.first(a1, b1));
    alpha[0] = a1;
    beta[0] = b1;
    double ebefor = 0.0, eps = 1e-25;
    const int ne = std::max(nvec, 1);
    std::vector<double> ab, bb;
    for (int i = 2; i <= last;) {
        // the steps up to the next one after which the host acts: the 4*nvec-th (first
        // bisection), then every 5th (the convergence test), or the last
        int e = last;
        if (fixed_steps <= 0) e = std::min(e, i <= 4 * ne ? 4 * ne : (i + 4) / 5 * 5);
        const int n = e - i + 1;
        ab.resize(n);
        bb.resize(n);
        int ran = 0;
        PNED_TRY(run.advance(n, a1, b1, ab.data(), bb.data(), &ran, true));
        for (int j = 0; j < ran; ++j) {
            alpha[i + j - 1] = ab[j];
            beta[i + j - 1] = bb[j];
            if (bb[j] < 0.5e-30) {
                *itr = i + j;
                PNED_TRY(run.finish(a1));
                return fail(PNED_E_BREAKDOWN, "#(E07) Tridiagonalization unsuccessful in lnc1: beta too small at i=" +
                                                  std::to_string(i + j));
            }
        }
        i = e + 1;
        if (fixed_steps > 0) continue;
        if (e > 4 * ne && e % 5 == 0) {
            PNED_TRY(pned_ctx_bisec(ctx, alpha, beta, e, ene, ne, eps));
            if (std::fabs((ebefor - ene[ne - 1]) / ene[ne - 1]) < 1e-13) {
                *itr = e;
                return run.finish(a1);
            }
            ebefor = ene[ne - 1];
        }
        if (e == 4 * ne) {
            eps = 1e-25;
            PNED_TRY(pned_ctx_bisec(ctx, alpha, beta, 4 * ne, ene, ne, eps));
            ebefor = ene[ne - 1];
        }
    }
    PNED_TRY(run.finish(a1));
    if (fixed_steps > 0) {
        PNED_TRY(pned_ctx_bisec(ctx, alpha, beta, last, ene, std::min(ne, last), 1e-25));
        *itr = last;
        return PNED_OK;
    }
    *itr = last;
    return fail(PNED_E_NOCONV, "#(W07) lnc1 did not converge within " + std::to_string(last) + " steps");
}

// parallelized_lnc1z (S_one_half/Hybrid_TPS_Lanczos.f90:48-163) from step `first`, with
// the reference's explicit update: first <= 2 starts from the normalised vector in v1
// (alpha(1), beta(1) by calc_alpha_beta, :78-85); first > 2 continues from the loop state a
// checkpoint holds -- v1, v0 of the top of step `first` and alpha/beta(1..first-1) -- as
// the restart branch does after interval_dump_read (:87-93).  The convergence reference
// ebefor is recomputed from the last check before `first`, where the uninterrupted loop
// set it (the reference leaves it undefined after a restart).  stop > first: return at the
// top of step `stop` (where the reference's time limit dumps) with *itr = stop and the loop
// state in v1, v0.  The same entry run in one span or split at any step gives the same
// alpha/beta bit for bit.
int pned_lnc1_span(pned_ctx *ctx, pned_vec *v1, pned_vec *v0, int first, int stop, int max_itr, int nvec,
                   int fixed_steps, double *alpha, double *beta, double *ene, int *itr)
{
    if (!ctx || !v1 || !v0 || !alpha || !beta || !ene || !itr) return fail(PNED_E_ARG, "null argument");
    if (nvec < 0) return fail(PNED_E_ARG, "#(E06) Wrong value given to nvec in lnc1");
    if (v1->nvec != 1) return fail(PNED_E_ARG, "lnc1 runs on one column");
    PNED_TRY(check_hv_pair(ctx, v1, v0));
    const int last = fixed_steps > 0 ? fixed_steps : max_itr;
    if (first < 1 || first > last || (stop > 0 && stop < std::max(first, 2)))
        return fail(PNED_E_ARG, "lnc1 span: need 1 <= first <= last and stop >= max(first, 2) (or 0)");
    double a1, b1;
    LanczosSteps run(ctx, v1, v0, true);
    int start = first;
    if (first <= 2) {
        PNED_TRY(run.first(a1, b1));
        alpha[0] = a1;
        beta[0] = b1;
        start = 2;
    } else {
        a1 = alpha[first - 2];
        b1 = beta[first - 2];
    }
    double ebefor = 0.0;
    const double eps = 1e-25;
    const int ne = std::max(nvec, 1);
    if (fixed_steps <= 0) {
        int prev = -1;   // the last step before `start` that set ebefor
        for (int i = start - 1; i >= 2 && prev < 0; --i)
            if ((i > 4 * ne && i % 5 == 0) || i == 4 * ne) prev = i;
        if (prev > 0) {
            PNED_TRY(pned_ctx_bisec(ctx, alpha, beta, prev, ene, ne, eps));
            ebefor = ene[ne - 1];
        }
    }
    for (int i = start; i <= last; ++i) {
        if (i == stop) {
            *itr = stop;
            return PNED_OK;
        }
        PNED_TRY(run.next(a1, b1));
        alpha[i - 1] = a1;
        beta[i - 1] = b1;
        if (b1 < 0.5e-30) {
            *itr = i;
            return fail(PNED_E_BREAKDOWN, "#(E07) Tridiagonalization unsuccessful in lnc1: beta too small at i=" +
                                              std::to_string(i));
        }
        if (fixed_steps > 0) continue;
        if (i > 4 * ne && i % 5 == 0) {
            PNED_TRY(pned_ctx_bisec(ctx, alpha, beta, i, ene, ne, eps));
            if (std::fabs((ebefor - ene[ne - 1]) / ene[ne - 1]) < 1e-13) {
                *itr = i;
                return PNED_OK;
            }
            ebefor = ene[ne - 1];
        }
        if (i == 4 * ne) {
            PNED_TRY(pned_ctx_bisec(ctx, alpha, beta, 4 * ne, ene, ne, eps));
            ebefor = ene[ne - 1];
        }
    }
    *itr = last;
    if (fixed_steps > 0) return pned_ctx_bisec(ctx, alpha, beta, last, ene, std::min(ne, last), eps);
    return fail(PNED_E_NOCONV, "#(W07) lnc1 did not converge within " + std::to_string(last) + " steps");
}

// Lanczos with full reorthogonalisation (no reference counterpart; the reference's orthg,
// S_one_half/Hybrid_TPS_basic.f90:219-303, is never called).  basis[0] holds the normalised
// start vector q_1; step i forms r = H q_i - alpha_i q_i - beta_{i-1} q_{i-1} in w, removes
// its components along q_1..q_i by two passes of classical Gram-Schmidt (all dots of a pass
// first, FP64, then r -= sum_j c_j q_j), and stores q_{i+1} = r / beta_i, beta_i = ||r||,
// in basis[i].  Every dot, norm, alpha and beta is FP64 whatever the vectors' storage, so
// with complex64 vectors only the stored components and the H.v products round to single
// precision.  fixed > 0: exactly that many steps; else the lnc1z rule (nvec-th Ritz value
// unchanged to 1e-13 relative between checks every 5 steps after 4 nvec), at most
// nbasis - 1 steps.  alpha[i-1] = alpha_i, beta[i-1] = beta_i.
static int lanczos_reorth(pned_ctx *ctx, pned_vec *const *basis, int nbasis, pned_vec *w, int fixed, int nvec,
                          double *alpha, double *beta, double *ene, int *itr)
{
    if (!ctx || !basis || !w || !alpha || !beta || !ene || !itr || nbasis < 2) return fail(PNED_E_ARG, "bad argument");
    for (int j = 0; j < nbasis; ++j) {
        if (!basis[j]) return fail(PNED_E_ARG, "null basis vector");
        PNED_TRY(check_hv_pair(ctx, basis[j], w));
        if (basis[j]->nvec != 1) return fail(PNED_E_ARG, "reorthogonalised Lanczos runs on one column");
    }
    const int last = fixed > 0 ? fixed : nbasis - 1;
    if (last > nbasis - 1) return fail(PNED_E_ARG, "basis holds fewer than steps + 1 vectors");
    const int ne = std::max(nvec, 1);
    const double one[2] = {1.0, 0.0};
    double ebefor = 0.0;
    std::vector<double> c(2 * (size_t)(last + 1));
    for (int i = 1; i <= last; ++i) {
        pned_vec *qi = basis[i - 1];
        double pr[2];
        PNED_TRY(pned_hv_fresh(ctx, qi, w, 0.0, pr));
        const double ai = pr[0];
        const double ma[2] = {-ai, 0.0};
        PNED_TRY(pned_axpby(ctx, ma, qi, one, w));
        if (i > 1) {
            const double mb[2] = {-beta[i - 2], 0.0};
            PNED_TRY(pned_axpby(ctx, mb, basis[i - 2], one, w));
        }
        for (int pass = 0; pass < 2; ++pass) {
            for (int j = 0; j < i; ++j) PNED_TRY(pned_calc_vec_inprdct(ctx, basis[j], w, &c[2 * j]));
            for (int j = 0; j < i; ++j) {
                const double mc[2] = {-c[2 * j], -c[2 * j + 1]};
                PNED_TRY(pned_axpby(ctx, mc, basis[j], one, w));
            }
        }
        double nn[2];
        PNED_TRY(pned_calc_vec_inprdct(ctx, w, w, nn));
        const double bi = std::sqrt(nn[0]);
        alpha[i - 1] = ai;
        beta[i - 1] = bi;
        if (bi < 0.5e-30) {
            *itr = i;
            return fail(PNED_E_BREAKDOWN, "#(E07) Tridiagonalization unsuccessful: beta too small at i=" +
                                              std::to_string(i));
        }
        PNED_TRY(pned_vec_copy(ctx, w, basis[i]));
        const double sc[2] = {1.0 / bi, 0.0};
        PNED_TRY(pned_scale(ctx, basis[i], sc));
        if (fixed > 0) continue;
        if (i >= 4 * ne && i % 5 == 0) {
            PNED_TRY(pned_ctx_bisec(ctx, alpha, beta, i, ene, ne, 1e-25));
            if (i > 4 * ne && std::fabs((ebefor - ene[ne - 1]) / ene[ne - 1]) < 1e-13) {
                *itr = i;
                return PNED_OK;
            }
            ebefor = ene[ne - 1];
        }
    }
    *itr = last;
    if (fixed > 0) return pned_ctx_bisec(ctx, alpha, beta, last, ene, std::min(ne, last), 1e-25);
    return fail(PNED_E_NOCONV, "reorthogonalised Lanczos did not converge within " + std::to_string(last) + " steps");
}

int pned_lnc1_reorth(pned_ctx *ctx, pned_vec *const *basis, int nbasis, pned_vec *w, int nvec, int fixed_steps,
                     double *alpha, double *beta, double *ene, int *itr)
{
    return lanczos_reorth(ctx, basis, nbasis, w, fixed_steps, nvec, alpha, beta, ene, itr);
}

int pned_dsflnc1z_reorth(pned_ctx *ctx, pned_vec *const *basis, int nbasis, pned_vec *w, double cnorm, int s_maxitr,
                         double *salpha, double *sbeta)
{
    if (!salpha || !sbeta || s_maxitr < 2) return fail(PNED_E_ARG, "bad argument");
    std::vector<double> b((size_t)s_maxitr), e(1);
    int itr = 0;
    PNED_TRY(lanczos_reorth(ctx, basis, nbasis, w, s_maxitr - 1, 1, salpha, b.data(), e.data(), &itr));
    sbeta[0] = cnorm;
    for (int i = 1; i <= s_maxitr - 1; ++i) sbeta[i] = b[(size_t)i - 1];
    return PNED_OK;
}

// Ritz vector of the Lanczos run: x = sum_k coef[k] q_k, regenerating q_k with the stored
// alpha/beta (same recursion as lnc1z, so the same basis up to rounding)
int pned_ritz_vector(pned_ctx *ctx, pned_vec *v1, pned_vec *v0, pned_vec *x, int nsteps, const double *alpha,
                     const double *beta, const double *coef)
{
    if (!ctx || !v1 || !v0 || !x || !alpha || !beta || !coef) return fail(PNED_E_ARG, "null argument");
    PNED_TRY(pned_vec_zero(ctx, v0));
    PNED_TRY(pned_vec_zero(ctx, x));
    const double zero[2] = {0.0, 0.0}, one[2] = {1.0, 0.0};
    double c[2] = {coef[0], 0.0};
    PNED_TRY(pned_axpby(ctx, c, v1, one, x));
    double pr[2];
    PNED_TRY(pned_p_mltply_v3(ctx, v1, v0, 0.0, pr));
    for (int i = 2; i <= nsteps; ++i) {
        double a = alpha[i - 2], b = beta[i - 2];
        PNED_TRY(pned_lanczos_update(ctx, v1, v0, &a, &b));
        c[0] = coef[i - 1];
        PNED_TRY(pned_axpby(ctx, c, v1, one, x));
        if (i < nsteps) PNED_TRY(pned_p_mltply_v3(ctx, v1, v0, 0.0, pr));
    }
    (void)zero;
    double nn[2];
    PNED_TRY(pned_calc_vec_inprdct(ctx, x, x, nn));
    const double s[2] = {1.0 / std::sqrt(nn[0]), 0.0};
    return pned_scale(ctx, x, s);
}

// parallelized_cg1 (S_one_half/Hybrid_TPS_cg.f90:104-197)
static int cg1(pned_ctx *ctx, double eig, pned_vec *x, pned_vec *b, pned_vec *p, pned_vec *r, pned_vec *y,
               int max_cg, int *saved_loop)
{
    double bc[2];
    PNED_TRY(pned_calc_vec_inprdct(ctx, b, b, bc));
    PNED_TRY(pned_vec_copy(ctx, b, r));
    PNED_TRY(pned_vec_copy(ctx, b, p));
    PNED_TRY(pned_vec_zero(ctx, x));
    const double eperbd = eig / (double)ctx->ibond;
    const double one[2] = {1.0, 0.0};
    for (int itr = 1; itr <= max_cg; ++itr) {
        double cdmm[2], crp[2], cyp[2], rc[2], rc2[2];
        PNED_TRY(pned_hv_fresh(ctx, p, y, eperbd, cdmm));
        PNED_TRY(pned_calc_vec_inprdct(ctx, p, r, crp));
        PNED_TRY(pned_calc_vec_inprdct(ctx, p, y, cyp));
        const std::complex<double> ca = std::complex<double>(crp[0], crp[1]) / std::complex<double>(cyp[0], cyp[1]);
        PNED_TRY(pned_calc_vec_inprdct(ctx, r, r, rc));
        const double pa[2] = {ca.real(), ca.imag()}, ma[2] = {-ca.real(), -ca.imag()};
        PNED_TRY(pned_axpby(ctx, pa, p, one, x));
        PNED_TRY(pned_axpby(ctx, ma, y, one, r));
        PNED_TRY(pned_calc_vec_inprdct(ctx, r, r, rc2));
        const double bt[2] = {rc2[0] / rc[0], 0.0};
        PNED_TRY(pned_axpby(ctx, one, r, bt, p));
        if (itr % 5 != 0) continue;
        if (std::sqrt(cabs2(rc2)) < 1.0e-9 * std::sqrt(cabs2(bc))) *saved_loop = 1;
        if (*saved_loop) return PNED_OK;
    }
    return PNED_OK;
}

// parallelized_inv1z (S_one_half/Hybrid_TPS_cg.f90:18-100)
int pned_inv1(pned_ctx *ctx, double eig, pned_vec *x, pned_vec *b, pned_vec *p, pned_vec *r, pned_vec *y,
              int max_inv, int max_cg, int reference_saved_loop, int *iters)
{
    if (!ctx || !x || !b || !p || !r || !y || !iters) return fail(PNED_E_ARG, "null argument");
    int saved = 0;
    for (int itr = 1; itr <= max_inv; ++itr) {
        int local = reference_saved_loop ? saved : 0;
        PNED_TRY(cg1(ctx, eig, x, b, p, r, y, max_cg, &local));
        if (reference_saved_loop) saved = local;
        double xn[2], cxb[2];
        PNED_TRY(pned_calc_vec_inprdct(ctx, x, x, xn));
        const double s[2] = {1.0 / std::sqrt(xn[0]), 0.0};
        PNED_TRY(pned_scale(ctx, x, s));
        PNED_TRY(pned_calc_vec_inprdct(ctx, x, b, cxb));
        if (std::fabs(cabs2(cxb) - 1.0) < 1.0e-12) {
            *iters = itr;
            return PNED_OK;
        }
        PNED_TRY(pned_vec_copy(ctx, x, b));
    }
    *iters = max_inv;
    return fail(PNED_E_NOCONV, "#(W11) inv1 did not converge");
}

// COCG (van der Vorst & Melissen, IEEE Trans. Magn. 26 (1990) 706) for the complex
// symmetric system (H - e0 + i eta) x = b.  It needs A^T = A, i.e. a real symmetric H,
// which every reference model has (real bond matrices); complex bond matrices (a field
// with an S^y part, Gamma terms) are rejected.
int pned_cocg(pned_ctx *ctx, double e0, double eta, const pned_vec *b, pned_vec *x, pned_vec *r, pned_vec *p,
              pned_vec *y, int max_it, double tol, int *iters, double *resid)
{
    if (!ctx || !b || !x || !r || !p || !y || !iters || !resid) return fail(PNED_E_ARG, "null argument");
    if (b->nvec != 1) return fail(PNED_E_ARG, "cocg runs on one column");
    for (size_t k = 1; k < ctx->xl.size(); k += 2)
        if (ctx->xl[k] != 0.0)
            return fail(PNED_E_UNSUPPORTED, "cocg needs real bond matrices (complex-symmetric H - z)");
    const double one[2] = {1.0, 0.0};
    const double ieta[2] = {0.0, eta};
    const double eperbd = e0 / (double)ctx->ibond;
    double bb[2];
    PNED_TRY(pned_calc_vec_inprdct(ctx, b, b, bb));
    const double bnorm = std::sqrt(bb[0]);
    PNED_TRY(pned_vec_zero(ctx, x));
    PNED_TRY(pned_vec_copy(ctx, b, r));
    PNED_TRY(pned_vec_copy(ctx, b, p));
    double rho[2];
    PNED_TRY(dotu(ctx, r, r, rho));
    // tiled spin-S kernel: the i eta shift and p^T y ride on the H.v itself
    // (other shapes take the generic path, whose calls report what is wrong)
    const bool fused = ctx->q >= 3 && ctx->q <= 5 && ctx->spin_tiled && ctx->sector_k < 0 && p->ctx == ctx &&
                       y->ctx == ctx && p != y && p->nvec == 1 && y->nvec == 1 && p->dtype == PNED_C128 &&
                       y->dtype == PNED_C128 && p->dim == ctx->dim && y->dim == ctx->dim;
    for (int it = 1; it <= max_it; ++it) {
        double pr[2], mu[2], rho2[2], rr[2];
        if (fused) {
            // y = (H - e0 + i eta) p, mu = p^T y
            PNED_TRY(launch_hv_spin_tiled_shift(ctx, p, y, eperbd, true, mu, 1.0, 1.0, 0.0, eta, true));
        } else {
            PNED_TRY(pned_hv_fresh(ctx, p, y, eperbd, pr));      // y = (H - e0) p
            PNED_TRY(pned_axpby(ctx, ieta, p, one, y));          // y += i eta p
            PNED_TRY(dotu(ctx, p, y, mu));
        }
        const std::complex<double> al = std::complex<double>(rho[0], rho[1]) / std::complex<double>(mu[0], mu[1]);
        const double pa[2] = {al.real(), al.imag()};
        double sums[4];
        PNED_TRY(cocg_update(ctx, p, y, x, r, pa, sums));   // x += al p, r -= al y, <r|r>, r^T r
        rr[0] = sums[0];
        rr[1] = sums[1];
        *resid = std::sqrt(rr[0]) / bnorm;
        *iters = it;
        if (*resid < tol) return PNED_OK;
        rho2[0] = sums[2];
        rho2[1] = sums[3];
        const std::complex<double> be = std::complex<double>(rho2[0], rho2[1]) / std::complex<double>(rho[0], rho[1]);
        const double bt[2] = {be.real(), be.imag()};
        PNED_TRY(pned_axpby(ctx, one, r, bt, p));             // p = r + beta p
        rho[0] = rho2[0];
        rho[1] = rho2[1];
    }
    return fail(PNED_E_NOCONV, "cocg did not reach the tolerance");
}

// One Lanczos step for every slot of a batch (serving): per-column update by mode, one
// batched H.v, per-column alpha and beta.  Idle columns return alpha = beta = 0.
int pned_lanczos_batch_step(pned_ctx *ctx, pned_vec *v1, pned_vec *v0, const int *mode, const double *alpha_prev,
                            const double *beta_prev, double *alpha, double *beta)
{
    if (!ctx || !v1 || !v0 || !mode || !alpha_prev || !beta_prev || !alpha || !beta)
        return fail(PNED_E_ARG, "null argument");
    const int nc = v1->nvec;
    // one column of a plain context: the deferred-normalisation step (as lnc1 runs it),
    // the buffer roles kept in the context between calls; mode 1 restarts from v1
    const bool lazy = nc == 1 && ctx->s12.nop <= 1 &&
                      ((ctx->q == 2 && ctx->sector_k < 0) ||
                       (ctx->q >= 3 && ctx->q <= 5 && ctx->spin_tiled && v1->dtype == PNED_C128));
    if (lazy) {
        if (mode[0] == 0) {
            alpha[0] = beta[0] = 0.0;
            return PNED_OK;
        }
        // the new buffer roles and factors are kept only once the step has run
        int swapped = 0;
        double sx = 1.0, sy = 0.0;
        if (mode[0] != 1) {
            if (!(beta_prev[0] > 0.0)) return fail(PNED_E_ARG, "stepping a column with beta <= 0");
            swapped = ctx->lz_swapped ^ 1;
            sx = 1.0 / beta_prev[0];
            sy = -beta_prev[0] * ctx->lz_sx;
        }
        pned_vec *X = swapped ? v0 : v1, *Y = swapped ? v1 : v0;
        double pr[2], b2;
        PNED_TRY(lanczos_lazy_step(ctx, X, Y, &sx, &sy, mode[0] == 1, pr, &b2));
        ctx->lz_swapped = swapped;
        ctx->lz_sx = sx;
        ctx->lz_sy = sy;
        alpha[0] = sx * pr[0];
        beta[0] = std::sqrt(b2);
        return PNED_OK;
    }
    PNED_TRY(lanczos_slots(ctx, v1, v0, alpha_prev, beta_prev, mode));
    std::vector<double> pr(2 * nc), b2(nc);
    PNED_TRY(hv_alpha_beta2(ctx, v1, v0, pr.data(), b2.data()));
    for (int c = 0; c < nc; ++c) {
        alpha[c] = mode[c] ? pr[2 * c] : 0.0;
        beta[c] = mode[c] ? std::sqrt(b2[c]) : 0.0;
    }
    return PNED_OK;
}

int pned_lanczos_batch_steps(pned_ctx *ctx, pned_vec *v1, pned_vec *v0, const int *mode, const double *alpha_prev,
                             const double *beta_prev, int nsteps, double *alpha, double *beta)
{
    if (!ctx || !v1 || !v0 || !mode || !alpha_prev || !beta_prev || !alpha || !beta)
        return fail(PNED_E_ARG, "null argument");
    if (nsteps < 1) return fail(PNED_E_ARG, "nsteps must be positive");
    const int nc = v1->nvec;
    std::vector<int> m(mode, mode + nc);
    std::vector<double> ap(alpha_prev, alpha_prev + nc), bp(beta_prev, beta_prev + nc);
    for (int s = 0; s < nsteps; ++s) {
        double *a = alpha + (size_t)s * nc, *b = beta + (size_t)s * nc;
        PNED_TRY(pned_lanczos_batch_step(ctx, v1, v0, m.data(), ap.data(), bp.data(), a, b));
        for (int c = 0; c < nc; ++c) {
            if (m[c] == 0) continue;
            ap[c] = a[c];
            bp[c] = b[c];
            m[c] = (b[c] >= 0.5e-30) ? 2 : 0;
        }
    }
    return PNED_OK;
}

int pned_set_initial_column(pned_ctx *ctx, pned_vec *v, int col, uint64_t seed)
{
    if (!ctx || !v) return fail(PNED_E_ARG, "null argument");
    return gaussian_column(ctx, v, col, seed);
}

int pned_check1(pned_ctx *ctx, const pned_vec *x, pned_vec *work, double *expect, double *ratios, int max_ratios,
                int *nratios)
{
    if (!ctx || !x || !work || !expect || !nratios) return fail(PNED_E_ARG, "null argument");
    double cd[2];
    PNED_TRY(pned_calc_vec_inprdct(ctx, x, x, cd));
    if (cabs2(cd) < 1e-30) return fail(PNED_E_ARG, "#(W09) Null vector given to check1");
    double pr[2];
    PNED_TRY(pned_hv_fresh(ctx, x, work, 0.0, pr));
    PNED_TRY(pned_calc_vec_inprdct(ctx, work, x, expect));
    const int64_t ld = ctx->dim;
    const int64_t j0 = std::min<int64_t>(ld / 3, 13), st = std::max<int64_t>(1, ld / 20);
    int k = 0;
    for (int64_t j = j0; j <= ld && k < max_ratios && ratios; j += st) {
        if (j < 1) continue;
        double hx[2], xv[2];
        PNED_TRY(pned_vec_download(ctx, work, hx, j - 1, 1));
        PNED_TRY(pned_vec_download(ctx, x, xv, j - 1, 1));
        const std::complex<double> q = std::complex<double>(hx[0], hx[1]) / std::complex<double>(xv[0], xv[1]);
        ratios[2 * k] = q.real();
        ratios[2 * k + 1] = q.imag();
        ++k;
    }
    *nratios = k;
    return PNED_OK;
}

int pned_residual(pned_ctx *ctx, const pned_vec *x, double e, pned_vec *work, double *res)
{
    if (!ctx || !x || !work || !res) return fail(PNED_E_ARG, "null argument");
    double pr[2], nn[2], xx[2];
    PNED_TRY(pned_hv_fresh(ctx, x, work, e / (double)ctx->ibond, pr));
    PNED_TRY(pned_calc_vec_inprdct(ctx, work, work, nn));
    PNED_TRY(pned_calc_vec_inprdct(ctx, x, x, xx));
    *res = std::sqrt(nn[0] / xx[0]);
    return PNED_OK;
}

int pned_tps_set_observables(pned_ctx *ctx, const double *site_vec, int nsq, const double *sq_q)
{
    if (!ctx || !site_vec || nsq < 0 || (nsq > 0 && !sq_q)) return fail(PNED_E_ARG, "bad argument");
    ctx->site_vec.assign(site_vec, site_vec + 2 * ctx->n);
    ctx->sq_q.assign(sq_q, sq_q + 2 * nsq);
    ctx->nsq = nsq;
    return PNED_OK;
}

// parallelized_TPSz (S_one_half/Hybrid_TPS_TPS.f90:325-409) with eval_phys_val (83-111)
int pned_tps(pned_ctx *ctx, pned_vec *v1, pned_vec *v0, pned_vec *vg, pned_vec *vt, double lmax, int nsteps,
             double *alpha, double *beta, double *coef)
{
    return pned_tps_span(ctx, v1, v0, vg, vt, lmax, 1, 0, nsteps, alpha, beta, coef);
}

// The same from step `first`: first <= 1 starts from the normalised columns of v1; first > 1
// continues from the loop state of the top of step `first` (the TPSz restart branch after
// interval_dump_read, :352-358): v0 = (L-H) v1 of step first-1, v1, vg, and rows
// 1..first-1 of alpha/beta/coef.  stop > first returns at the top of step `stop` (the
// time-limit dump point) with that loop state.  One span or several: the same rows.
int pned_tps_span(pned_ctx *ctx, pned_vec *v1, pned_vec *v0, pned_vec *vg, pned_vec *vt, double lmax, int first,
                  int stop, int nsteps, double *alpha, double *beta, double *coef)
{
    if (!ctx || !v1 || !v0 || !vg || !vt || !alpha || !beta || !coef) return fail(PNED_E_ARG, "null argument");
    if (ctx->site_vec.empty()) return fail(PNED_E_ARG, "call pned_tps_set_observables first");
    if (first < 1 || first > nsteps || (stop > 0 && stop < first))
        return fail(PNED_E_ARG, "tps span: need 1 <= first <= nsteps and stop >= first (or 0)");
    const int nc = v1->nvec, nsq = ctx->nsq, ncoef = 7 + nsq;
    const int width = 6 + 2 * nsq;
    std::vector<double> pr(2 * nc), n0(2 * nc), gv0(2 * nc), obs((size_t)width * nc), xn(nc);
    (void)vt;
    if (first <= 1) {
        PNED_TRY(pned_vec_zero(ctx, v0));
        PNED_TRY(pned_vec_zero(ctx, vg));
    } else {
        for (int c = 0; c < nc; ++c) xn[c] = 1.0 / beta[(size_t)(first - 2) * nc + c];
    }
    for (int i = std::max(first, 1); i <= nsteps; ++i) {
        if (i == stop) return PNED_OK;
        if (i > 1) {
            PNED_TRY(tps_shift_update(ctx, v1, v0, vg, xn.data()));
        }
        PNED_TRY(pned_hv_fresh(ctx, v1, v0, 0.0, pr.data()));
        PNED_TRY(pned_calc_vec_inprdct(ctx, v0, v0, n0.data()));
        PNED_TRY(pned_calc_vec_inprdct(ctx, vg, v0, gv0.data()));
        PNED_TRY(tps_observables(ctx, vg, v1, obs.data()));
        for (int c = 0; c < nc; ++c) {
            const double a1 = pr[2 * c], b1 = std::sqrt(n0[2 * c]);
            alpha[(size_t)(i - 1) * nc + c] = lmax - a1;
            beta[(size_t)(i - 1) * nc + c] = b1;
            xn[c] = 1.0 / b1;
            double *row = coef + ((size_t)(i - 1) * nc + c) * ncoef;
            const double *o = obs.data() + (size_t)c * width;
            row[0] = (double)(2 * i) / a1;
            row[1] = lmax;
            row[2] = b1;
            row[3] = lmax - a1;
            // calc_internal_energy: Re(l <vg|v1> - <vg|v0>)
            row[4] = lmax * o[0] - gv0[2 * c];
            row[5] = o[2];
            row[6] = o[4];
            for (int k = 0; k < nsq; ++k) row[7 + k] = o[6 + 2 * k];
        }
    }
    return PNED_OK;
}

// One column's alpha, beta and eval_phys_val row of fused mTPQ step i (1-based) from the
// step's H.v product pr (2 doubles: <u|(L-H)(su u)>) and reduction o; su is the step's input
// scale (1/beta of the step before, 1 at the first step) and sp the one before it.
static void tps_fused_row(double lmax, int i, int nsq, double su, double sp, const double *pr, const double *o,
                          double *alpha, double *beta, double *row)
{
    const double a1 = su * pr[0];                   // <v1|(L-H)|v1>
    const double b1 = std::sqrt(o[0]);              // ||v0||
    *alpha = lmax - a1;
    *beta = b1;
    const double gv = sp * su, vv = su * su;
    row[0] = (double)(2 * i) / a1;
    row[1] = lmax;
    row[2] = b1;
    row[3] = lmax - a1;
    row[4] = lmax * gv * o[4] - sp * o[2];          // Re(l <vg|v1> - <vg|v0>)
    row[5] = vv * o[6];
    row[6] = gv * o[8];
    for (int k = 0; k < nsq; ++k) {
        const bool left_g = (k % 4) % 2 == 1;
        row[7 + k] = (left_g ? gv : vv) * o[10 + 2 * k];
    }
}

static int tps_fused_check(pned_ctx *ctx, pned_vec *v1, pned_vec *va, pned_vec *vb)
{
    if (ctx->site_vec.empty()) return fail(PNED_E_ARG, "call pned_tps_set_observables first");
    if (ctx->q != 2) return fail(PNED_E_UNSUPPORTED, "fused mTPQ is S=1/2");
    PNED_TRY(check_hv_pair(ctx, v1, va));
    PNED_TRY(check_hv_pair(ctx, v1, vb));
    return check_hv_pair(ctx, va, vb);
}

// The same recursion and coef rows as pned_tps with the vector work fused: vectors are
// rotated instead of copied (vg <- v1 <- v0 by handle), the 1/beta normalisation rides
// on the next H.v as a per-column input scale, and every reduction of a step is one pass
// over (vg, v1, v0).  Per step: the H.v passes + 3 vector reads, instead of the
// reference's update (2 reads, 3 writes), two dots and the observable passes.
int pned_tps_fused(pned_ctx *ctx, pned_vec *v1, pned_vec *va, pned_vec *vb, double lmax, int nsteps, double *alpha,
                   double *beta, double *coef)
{
    if (!ctx || !v1 || !va || !vb || !alpha || !beta || !coef) return fail(PNED_E_ARG, "null argument");
    PNED_TRY(tps_fused_check(ctx, v1, va, vb));
    const int nc = v1->nvec, nsq = ctx->nsq, ncoef = 7 + nsq, width = 10 + 2 * nsq;
    pned_vec *u = v1, *up = va, *w = vb;
    PNED_TRY(pned_vec_zero(ctx, up));
    std::vector<double> su(nc, 1.0), sp(nc, 0.0), pr(2 * nc), red((size_t)width * nc);
    double *dscale = nullptr;
    for (int i = 1; i <= nsteps; ++i) {
        PNED_TRY(upload_scalars(ctx, su.data(), nc, &dscale));
        PNED_TRY(launch_hv_s12(ctx, u, w, 0.0, true, pr.data(), dscale));   // w = (L-H)(su u) = v0
        PNED_TRY(tps_fused_reduce(ctx, up, u, w, red.data()));
        for (int c = 0; c < nc; ++c) {
            const size_t at = (size_t)(i - 1) * nc + c;
            tps_fused_row(lmax, i, nsq, su[c], sp[c], pr.data() + 2 * c, red.data() + (size_t)c * width, alpha + at,
                          beta + at, coef + at * ncoef);
            sp[c] = su[c];
            su[c] = 1.0 / beta[at];
        }
        pned_vec *t = up;   // vg <- v1, v1 <- v0, v0 <- free
        up = u;
        u = w;
        w = t;
    }
    PNED_HIP(hipStreamSynchronize(ctx->stream));
    return PNED_OK;
}

namespace {
// device and pinned buffers and the graph of one pned_tps_fused_graph call
struct TpsGraphRun {
    void *dobs = nullptr;
    double *hist = nullptr, *su = nullptr, *host = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~TpsGraphRun()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (host) (void)hipHostFree(host);
        if (hist) (void)hipFree(hist);
        if (su) (void)hipFree(su);
        if (dobs) (void)hipFree(dobs);
    }
};
}  // namespace

// pned_tps_fused with the step sequence captured once into a HIP graph of graph_steps
// steps (a multiple of 3, so the rotated vector roles repeat; 0 = 30) and replayed: each
// step's 1/beta scale is formed on the device from the step before, each step's H.v
// product and reduction land in a per-step history slot, and the host reads the history
// once per replay.  The kernels, their inputs and their order are those of pned_tps_fused,
// so alpha, beta and coef are the same bit for bit.  Step 1 and the steps after the last
// whole replay run un-captured through the same enqueue path.
int pned_tps_fused_graph(pned_ctx *ctx, pned_vec *v1, pned_vec *va, pned_vec *vb, double lmax, int nsteps,
                         int graph_steps, double *alpha, double *beta, double *coef)
{
    if (!ctx || !v1 || !va || !vb || !alpha || !beta || !coef) return fail(PNED_E_ARG, "null argument");
    if (graph_steps == 0) graph_steps = 30;
    if (graph_steps < 3 || graph_steps % 3 != 0 || graph_steps > 3000)
        return fail(PNED_E_ARG, "graph_steps is a positive multiple of 3 (at most 3000), or 0 for 30");
    if (nsteps < 1) return fail(PNED_E_ARG, "nsteps must be positive");
    PNED_TRY(tps_fused_check(ctx, v1, va, vb));
    PNED_HIP(hipSetDevice(ctx->device));
    const int nc = v1->nvec, ncoef = 7 + ctx->nsq, G = graph_steps;
    const size_t slot_d = 2 * (size_t)nc + (size_t)TPS_RED_WIDTH * nc;
    TpsGraphRun run;
    int nsq = 0;
    PNED_TRY(tps_graph_setup(ctx, v1, &run.dobs, &nsq));
    PNED_HIP(hipMalloc((void **)&run.hist, sizeof(double) * slot_d * G));
    PNED_HIP(hipMalloc((void **)&run.su, sizeof(double) * nc));
    PNED_HIP(hipHostMalloc((void **)&run.host, sizeof(double) * slot_d * G, hipHostMallocDefault));
    pned_vec *u = v1, *up = va, *w = vb;
    PNED_TRY(pned_vec_zero(ctx, up));
    std::vector<double> su(nc, 1.0), sp(nc, 0.0);
    // step i (1-based) writes slot (i - 2) mod G and scales by the slot before it
    auto slot_of = [G](int i) { return ((i - 2) % G + G) % G; };
    auto enqueue = [&](int i) -> int {
        const int sl = slot_of(i);
        const double *prev = i == 1 ? nullptr : run.hist + (size_t)slot_of(i - 1) * slot_d + 2 * (size_t)nc;
        PNED_TRY(tps_graph_step(ctx, up, u, w, run.dobs, nsq, prev, run.su, run.hist + (size_t)sl * slot_d));
        pned_vec *t = up;   // vg <- v1, v1 <- v0, v0 <- free
        up = u;
        u = w;
        w = t;
        return PNED_OK;
    };
    // rows of steps first..first+count-1 (their slots contiguous) from the device history
    auto harvest = [&](int first, int count) -> int {
        const int s0 = slot_of(first);
        PNED_HIP(hipMemcpyAsync(run.host, run.hist + (size_t)s0 * slot_d, sizeof(double) * slot_d * count,
                                hipMemcpyDeviceToHost, ctx->stream));
        PNED_HIP(hipStreamSynchronize(ctx->stream));
        for (int j = 0; j < count; ++j) {
            const int i = first + j;
            const double *h = run.host + (size_t)j * slot_d;
            for (int c = 0; c < nc; ++c) {
                const size_t at = (size_t)(i - 1) * nc + c;
                tps_fused_row(lmax, i, ctx->nsq, su[c], sp[c], h + 2 * c, h + 2 * (size_t)nc + (size_t)c * TPS_RED_WIDTH,
                              alpha + at, beta + at, coef + at * ncoef);
                sp[c] = su[c];
                su[c] = 1.0 / beta[at];
            }
        }
        return PNED_OK;
    };
    // step 1 also builds the H.v plan, its tables and the partial buffers: nothing in the
    // captured steps allocates or waits on the host
    PNED_TRY(enqueue(1));
    PNED_TRY(harvest(1, 1));
    const int nrep = (nsteps - 1) / G;
    int i = 2;
    if (nrep > 0) {
        // the buffers the captured kernels address must stay where step 1 left them
        const void *keep[4] = {ctx->d_partial, ctx->d_result, ctx->pd_dev, ctx->d_tm};
        PNED_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
        int rc = PNED_OK;
        for (int j = 0; j < G && rc == PNED_OK; ++j) rc = enqueue(i + j);
        const hipError_t ec = hipStreamEndCapture(ctx->stream, &run.graph);
        PNED_TRY(rc);
        if (ec != hipSuccess) return fail(PNED_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
        const void *now[4] = {ctx->d_partial, ctx->d_result, ctx->pd_dev, ctx->d_tm};
        if (std::memcmp(keep, now, sizeof(keep)) != 0)
            return fail(PNED_E_HIP, "mTPQ graph: a work buffer moved while the steps were captured");
        PNED_HIP(hipGraphInstantiate(&run.exec, run.graph, nullptr, nullptr, 0));
        for (int r = 0; r < nrep; ++r, i += G) {
            PNED_HIP(hipGraphLaunch(run.exec, ctx->stream));
            PNED_TRY(harvest(i, G));
        }
    }
    const int rest = nsteps - i + 1;
    for (int j = 0; j < rest; ++j) PNED_TRY(enqueue(i + j));
    if (rest > 0) PNED_TRY(harvest(i, rest));
    return PNED_OK;
}

int pned_intalv_zz(pned_ctx *ctx, const pned_vec *x, pned_vec *v, const double q[2], const double *site_vec,
                   double *cnorm)
{
    if (!ctx || !x || !v || !q || !site_vec || !cnorm) return fail(PNED_E_ARG, "null argument");
    PNED_TRY(sz_q_apply(ctx, x, v, q, site_vec));
    double nn[2];
    PNED_TRY(pned_calc_vec_inprdct(ctx, v, v, nn));
    double cn = nn[0];
    if (cabs2(nn) < 1e-25) cn = 1e-25;   // 'vec norm is too small' (Hybrid_TPS_DSF.f90:72-76)
    const double s[2] = {1.0 / std::sqrt(cn), 0.0};
    PNED_TRY(pned_scale(ctx, v, s));
    *cnorm = cn / (double)ctx->n;
    return PNED_OK;
}

int pned_intalv_xy(pned_ctx *ctx, const pned_vec *x, pned_vec *v, const double q[2], const double *site_vec,
                   int ipm, double *cnorm)
{
    if (!ctx || !x || !v || !q || !site_vec || !cnorm || !x->ctx) return fail(PNED_E_ARG, "null argument");
    if (ipm != 1 && ipm != -1) return fail(PNED_E_ARG, "ipm is +1 (S+) or -1 (S-)");
    if (v->ctx != ctx) return fail(PNED_E_ARG, "v must belong to the target context");
    if (ctx->sector_k >= 0 || x->ctx->sector_k >= 0)
        PNED_TRY(intalv_xy_apply(ctx, x, v, q, site_vec, ipm));
    else
        PNED_TRY(intalv_pm_apply(ctx, x, v, q, site_vec, ipm));
    double nn[2];
    PNED_TRY(pned_calc_vec_inprdct(ctx, v, v, nn));
    double cn = nn[0];
    if (cabs2(nn) < 1e-25) cn = 1e-25;   // 'vec norm is too small' (Hybrid_TPS_DSF.f90:174-179)
    const double s[2] = {1.0 / std::sqrt(cn), 0.0};
    PNED_TRY(pned_scale(ctx, v, s));
    *cnorm = cn / (double)ctx->n;
    return PNED_OK;
}

// parallelized_dsflnc1z (S_one_half/Hybrid_TPS_DSF.f90:195-261)
int pned_dsflnc1z(pned_ctx *ctx, pned_vec *v1, pned_vec *v0, double cnorm, int s_maxitr, double *salpha,
                  double *sbeta)
{
    if (!ctx || !v1 || !v0 || !salpha || !sbeta || s_maxitr < 2) return fail(PNED_E_ARG, "bad argument");
    sbeta[0] = cnorm;
    double a1, b1;
    LanczosSteps run(ctx, v1, v0);
    PNED_TRY(run.first(a1, b1));
    salpha[0] = a1;
    sbeta[1] = b1;
    // steps 2 .. s_maxitr-1, with no test inside the loop (as the reference's)
    const int n = s_maxitr - 2;
    if (n > 0) {
        std::vector<double> ab(n), bb(n);
        int ran = 0;
        PNED_TRY(run.advance(n, a1, b1, ab.data(), bb.data(), &ran, false));
        for (int j = 0; j < n; ++j) {
            salpha[j + 1] = ab[j];
            sbeta[j + 2] = bb[j];
        }
    }
    return run.finish(a1);
}

// parallelized_dsflnc1z (S_one_half/Hybrid_TPS_DSF.f90:195-261) from step `first` with
// the reference's explicit update: first <= 2 starts from the normalised excitation in v1
// (sbeta(1) = cnorm); first > 2 continues from the loop state of the top of step `first`
// -- v1, v0 and salpha(1..first-1), sbeta(1..first) -- the interval_dump_read restart
// (itr_counter(11) > 2) and the calc_extended_DSF continuation (itr_counter(11) = indx + 1
// after read_alpha_beta2, :515-553) both use.  Runs to step s_maxitr - 1; stop > first
// returns at the top of step `stop` (the time-limit dump point).  *icf = the last step
// taken.  One span or several give the same salpha/sbeta bit for bit.
int pned_dsflnc1z_span(pned_ctx *ctx, pned_vec *v1, pned_vec *v0, double cnorm, int first, int stop, int s_maxitr,
                       double *salpha, double *sbeta, int *icf)
{
    if (!ctx || !v1 || !v0 || !salpha || !sbeta || !icf || s_maxitr < 2) return fail(PNED_E_ARG, "bad argument");
    PNED_TRY(check_hv_pair(ctx, v1, v0));
    if (first < 1 || first > s_maxitr || (stop > 0 && stop < std::max(first, 2)))
        return fail(PNED_E_ARG, "dsflnc1z span: need 1 <= first <= s_maxitr and stop >= max(first, 2) (or 0)");
    double a1, b1;
    LanczosSteps run(ctx, v1, v0, true);
    int start = first;
    if (first <= 2) {
        sbeta[0] = cnorm;
        PNED_TRY(run.first(a1, b1));
        salpha[0] = a1;
        sbeta[1] = b1;
        start = 2;
    } else {
        a1 = salpha[first - 2];
        b1 = sbeta[first - 1];
    }
    int i = start;
    for (; i <= s_maxitr - 1; ++i) {
        if (i == stop) break;
        PNED_TRY(run.next(a1, b1));
        salpha[i - 1] = a1;
        sbeta[i] = b1;
    }
    *icf = i - 1;
    return PNED_OK;
}

int pned_renbun(pned_ctx *ctx, int i_range, int s_maxitr, double split, double ene, double eta,
                const double *salpha, const double *sbeta, double sq, double *w, double *dsf, int *conv)
{
    if (!ctx || !salpha || !sbeta || !w || !dsf || !conv || i_range < 1) return fail(PNED_E_ARG, "bad argument");
    PNED_TRY(renbun_gpu(ctx, i_range, s_maxitr, split, ene, eta, salpha, sbeta, w, dsf, conv));
    if (sq > 0.0) {
        // nomalization_DSF (Hybrid_TPS_DSF.f90:337-351)
        double sum = 0.0;
        for (int i = 1; i < i_range; ++i) sum += 0.5 * split * (dsf[i] + dsf[i - 1]);
        for (int i = 0; i < i_range; ++i) dsf[i] = sq / sum * dsf[i];
    }
    return PNED_OK;
}

}  // extern "C"
