// tridiag.hip -- the Lanczos tridiagonal's lowest eigenpairs on the context's GPU
// (pned_ctx_bisec, pned_ctx_vec12): the arithmetic of the host entry points pned_bisec /
// pned_vec12 (tridiag.h, the same code on both sides): a wavefront per eigenvalue that
// walks the bisection six tree levels per round, and one thread per cluster of
// eigenvectors (inverse iteration is a sequential recurrence over the rows).
#include <algorithm>
#include <cstring>

#include "internal.h"
#include "tridiag.h"

namespace pned {

// One wavefront per eigenvalue (block k-1 finds the k-th): bisect_kth's steps, six at a
// time.  Each round every lane takes one node of the next six levels of the bisection tree
// (heap order, 63 nodes), forms its bracket by the same halvings the sequential loop would
// make on the way there, and evaluates that node's stop test and Sturm count; the wave then
// follows the decisions from the root.  The midpoints, stop tests and the 200-step cap are
// those of bisect_kth, so the result is the same double, after ~11 dependent Sturm counts
// instead of ~64.
__global__ __launch_bounds__(64) void k_bisect_tree(const double *__restrict__ alpha, const double *__restrict__ b2,
                                                    int n, TridiagBounds tb, double *__restrict__ E)
{
#pragma clang fp contract(off)
    const int k = blockIdx.x + 1, lane = threadIdx.x;
    const int m = lane + 1, depth = 31 - __clz(m);   // 1-based heap number: path bits below the top one
    double a = tb.lo, b = tb.hi;
    int it = 0;
    for (;;) {
        double na = a, nb = b;
        bool reach = true;
        for (int d = depth - 1; d >= 0 && reach; --d) {   // root -> this node (1 = right: a = c)
            const double c = 0.5 * (na + nb);
            if (c <= na || c >= nb || nb - na <= tb.width) reach = false;
            else if ((m >> d) & 1) na = c;
            else nb = c;
        }
        int left = 0, stop = 0;
        if (lane < 63 && reach) {
            const double c = 0.5 * (na + nb);
            if (c <= na || c >= nb || nb - na <= tb.width) stop = 1;
            else left = sturm_count(alpha, b2, n, c, tb.pivmin) >= k;
        }
        const uint64_t leftm = __ballot(left), stopm = __ballot(stop);
        int node = 0;
        bool done = false;
        for (int d = 0; d < 6; ++d) {
            if (it >= 200 || ((stopm >> node) & 1)) {
                done = true;
                break;
            }
            const double c = 0.5 * (a + b);
            ++it;
            if ((leftm >> node) & 1) {
                b = c;
                node = 2 * node + 1;
            } else {
                a = c;
                node = 2 * node + 2;
            }
        }
        if (done) break;
    }
    if (lane == 0) E[k - 1] = 0.5 * (a + b);
}

// thread j: the eigenvectors cl[j] .. cl[j+1]-1 of one cluster, in order; work: 4n doubles
// and n bytes per thread
__global__ void k_inverse_iteration(const double *__restrict__ alpha, const double *__restrict__ beta, int n,
                                    const double *__restrict__ E, const int *__restrict__ cl, int ncl, double tiny,
                                    double *__restrict__ v, double *__restrict__ work, char *__restrict__ swork)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ncl) return;
    double *w = work + (size_t)j * 4 * n;
    for (int k = cl[j]; k < cl[j + 1]; ++k)
        inverse_iteration(alpha, beta, n, E[k], k, cl[j], tiny, v, w, w + n, w + 2 * (size_t)n, w + 3 * (size_t)n,
                          swork + (size_t)j * n);
}

}  // namespace pned

using namespace pned;

extern "C" int pned_ctx_bisec(pned_ctx *ctx, const double *alpha, const double *beta, int ndim, double *E, int ne,
                              double eps)
{
    if (!ctx || !alpha || !beta || !E || ndim < 1) return fail(PNED_E_ARG, "null argument");
    if (ne > ndim || ne < 0) return fail(PNED_E_ARG, "#(E05) ne given to bisec out of range");
    std::vector<double> b2;
    TridiagBounds tb;
    PNED_TRY(tridiag_bounds(alpha, beta, ndim, eps, b2, tb));
    if (ne == 0) return PNED_OK;
    PNED_HIP(hipSetDevice(ctx->device));
    // device: alpha (ndim), b2 (ndim - 1, at least 1), E (ne)
    const size_t nb2 = b2.size();
    double *d;
    PNED_TRY(ensure_scalars(ctx, (int)(ndim + nb2 + ne), &d));
    PNED_HIP(hipMemcpyAsync(d, alpha, sizeof(double) * ndim, hipMemcpyHostToDevice, ctx->stream));
    PNED_HIP(hipMemcpyAsync(d + ndim, b2.data(), sizeof(double) * nb2, hipMemcpyHostToDevice, ctx->stream));
    double *dE = d + ndim + nb2;
    hipLaunchKernelGGL(k_bisect_tree, dim3(ne), dim3(64), 0, ctx->stream, d, d + ndim, ndim, tb, dE);
    PNED_HIP(hipGetLastError());
    PNED_HIP(hipMemcpyAsync(E, dE, sizeof(double) * ne, hipMemcpyDeviceToHost, ctx->stream));
    PNED_HIP(hipStreamSynchronize(ctx->stream));
    return PNED_OK;
}

extern "C" int pned_ctx_vec12(pned_ctx *ctx, const double *alpha, const double *beta, int ndim, const double *E,
                              int nvec, double *v)
{
    if (!ctx || !alpha || !beta || !E || !v || ndim < 1 || nvec < 1) return fail(PNED_E_ARG, "bad argument");
    const int n = ndim;
    std::vector<int> c0;
    double tiny;
    PNED_TRY(tridiag_clusters(alpha, beta, n, E, nvec, c0, tiny));
    std::vector<int> cl;   // cluster starts, then nvec
    for (int k = 0; k < nvec; ++k)
        if (c0[k] == k) cl.push_back(k);
    const int ncl = (int)cl.size();
    cl.push_back(nvec);
    PNED_HIP(hipSetDevice(ctx->device));
    // device: alpha (n), beta (n), E (nvec), v (nvec n), work (ncl 4n), cluster starts and
    // the pivot flags (ncl + 1 ints, ncl n bytes, rounded up to doubles)
    const size_t nbeta = (size_t)n, nwork = (size_t)ncl * 4 * n;
    const size_t nint = ((size_t)(ncl + 1) * sizeof(int) + 7) / 8, nsw = ((size_t)ncl * n + 7) / 8;
    const size_t total = (size_t)n + nbeta + nvec + (size_t)nvec * n + nwork + nint + nsw;
    if (total > (size_t)1 << 30) return fail(PNED_E_NOMEM, "vec12: more than 8 GiB of device work space");
    double *d;
    PNED_TRY(ensure_scalars(ctx, (int)total, &d));
    double *dal = d, *dbe = dal + n, *dE = dbe + nbeta, *dv = dE + nvec, *dw = dv + (size_t)nvec * n;
    int *dcl = reinterpret_cast<int *>(dw + nwork);
    char *dsw = reinterpret_cast<char *>(dw + nwork + nint);
    std::vector<double> bpad(nbeta, 0.0);
    std::copy(beta, beta + (n - 1), bpad.begin());
    PNED_HIP(hipMemcpyAsync(dal, alpha, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
    PNED_HIP(hipMemcpyAsync(dbe, bpad.data(), sizeof(double) * nbeta, hipMemcpyHostToDevice, ctx->stream));
    PNED_HIP(hipMemcpyAsync(dE, E, sizeof(double) * nvec, hipMemcpyHostToDevice, ctx->stream));
    PNED_HIP(hipMemcpyAsync(dcl, cl.data(), sizeof(int) * (ncl + 1), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_inverse_iteration, dim3((ncl + 63) / 64), dim3(64), 0, ctx->stream, dal, dbe, n, dE, dcl,
                       ncl, tiny, dv, dw, dsw);
    PNED_HIP(hipGetLastError());
    PNED_HIP(hipMemcpyAsync(v, dv, sizeof(double) * nvec * n, hipMemcpyDeviceToHost, ctx->stream));
    PNED_HIP(hipStreamSynchronize(ctx->stream));
    return PNED_OK;
}
