// tridiag.h -- lowest eigenpairs of a symmetric tridiagonal T (diagonal alpha[0..n),
// off-diagonal beta[0..n-1)), shared by the host entry points (pned_bisec, pned_vec12) and
// the device kernels behind pned_ctx_bisec / pned_ctx_vec12.  Both sides run this same code
// with contraction off and IEEE-rounded division and square root, so they give the same
// doubles.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pned {

// number of eigenvalues of T below x (LDL^T pivots; a vanishing pivot is moved to -pivmin)
__host__ __device__ inline int sturm_count(const double *alpha, const double *b2, int n, double x, double pivmin)
{
#pragma clang fp contract(off)
    int cnt = 0;
    double d = alpha[0] - x;
    if (fabs(d) < pivmin) d = -pivmin;
    cnt += d < 0.0;
    for (int i = 1; i < n; ++i) {
        d = (alpha[i] - x) - b2[i - 1] / d;
        if (fabs(d) < pivmin) d = -pivmin;
        cnt += d < 0.0;
    }
    return cnt;
}

// the k-th lowest eigenvalue (1-based): bisection of (lo, hi] (fewer than k eigenvalues
// below the left end, at least k below the right end) until the bracket is narrower than
// width or holds two adjacent doubles; the bracket's midpoint
__host__ __device__ inline double bisect_kth(const double *alpha, const double *b2, int n, int k, double lo, double hi,
                                             double width, double pivmin)
{
#pragma clang fp contract(off)
    double a = lo, b = hi;
    for (int it = 0; it < 200; ++it) {
        const double c = 0.5 * (a + b);
        if (c <= a || c >= b || b - a <= width) break;
        if (sturm_count(alpha, b2, n, c, pivmin) >= k) b = c;
        else a = c;
    }
    return 0.5 * (a + b);
}

// start vector of inverse iteration: entries in [0.5, 1.5) from an integer hash of (i, k),
// exact on both sides
__host__ __device__ inline double inviter_start(int i, int k)
{
    uint64_t z = (uint64_t)(uint32_t)i * 0x9E3779B97F4A7C15ull + (uint64_t)(uint32_t)k * 0xD1B54A32D192ED03ull +
                 0x2545F4914F6CDD1Dull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return 0.5 + (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

// eigenvector k of eigenvalue E into v + k*n (unit norm): LU with partial pivoting of
// T - E I (rows i, i+1 exchanged where sw[i]), three inverse-iteration solves from
// inviter_start, each followed by two Gram-Schmidt passes against vectors c0..k-1 (the
// earlier members of its cluster, already in v); work arrays of n entries each
__host__ __device__ inline void inverse_iteration(const double *alpha, const double *beta, int n, double E, int k,
                                                  int c0, double tiny, double *v, double *dl, double *d, double *du,
                                                  double *du2, char *sw)
{
#pragma clang fp contract(off)
    for (int i = 0; i < n; ++i) {
        d[i] = alpha[i] - E;
        dl[i] = (i < n - 1) ? beta[i] : 0.0;
        du[i] = (i < n - 1) ? beta[i] : 0.0;
        du2[i] = 0.0;
    }
    for (int i = 0; i < n - 1; ++i) {
        if (fabs(d[i]) >= fabs(dl[i])) {
            sw[i] = 0;
            if (fabs(d[i]) < tiny) d[i] = (d[i] < 0.0) ? -tiny : tiny;
            const double f = dl[i] / d[i];
            dl[i] = f;
            d[i + 1] -= f * du[i];
        } else {
            sw[i] = 1;
            const double f = d[i] / dl[i];
            d[i] = dl[i];
            dl[i] = f;
            const double t = du[i];
            du[i] = d[i + 1];
            d[i + 1] = t - f * d[i + 1];
            if (i < n - 2) {
                du2[i] = du[i + 1];
                du[i + 1] = -f * du[i + 1];
            }
        }
    }
    if (fabs(d[n - 1]) < tiny) d[n - 1] = (d[n - 1] < 0.0) ? -tiny : tiny;
    double *x = v + (size_t)k * n;
    for (int i = 0; i < n; ++i) x[i] = inviter_start(i, k);
    for (int sweep = 0; sweep < 3; ++sweep) {
        for (int i = 0; i < n - 1; ++i) {   // forward: L
            if (sw[i]) {
                const double t = x[i];
                x[i] = x[i + 1];
                x[i + 1] = t;
            }
            x[i + 1] -= dl[i] * x[i];
        }
        for (int i = n - 1; i >= 0; --i) {   // backward: U (two super-diagonals)
            double s = x[i];
            if (i + 1 < n) s -= du[i] * x[i + 1];
            if (i + 2 < n) s -= du2[i] * x[i + 2];
            x[i] = s / d[i];
        }
        for (int pass = 0; pass < 2; ++pass)   // orthogonal to the cluster so far
            for (int j = c0; j < k; ++j) {
                const double *y = v + (size_t)j * n;
                double pr = 0.0;
                for (int i = 0; i < n; ++i) pr += y[i] * x[i];
                for (int i = 0; i < n; ++i) x[i] -= pr * y[i];
            }
        double nrm = 0.0;
        for (int i = 0; i < n; ++i) nrm += x[i] * x[i];
        nrm = sqrt(nrm);
        if (nrm > 0.0)
            for (int i = 0; i < n; ++i) x[i] /= nrm;
    }
}

}  // namespace pned
