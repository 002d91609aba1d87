// vec_ops.hip -- BLAS-1, start vectors, excitations, observables, continued fractions.
//
// All reductions are two-stage and order-fixed (fixed grid, fixed per-thread strides,
// fixed tree), so a result depends only on the vector contents: the same column gives
// the same bits alone or batched with others.  Sums are FP64 for both storage types.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "internal.h"

namespace pned {

namespace {

constexpr int RED_TPB = 256;
constexpr int RED_GRID = 1024;

template <typename Real> struct V2;
template <> struct V2<double> { using T = double2; };
template <> struct V2<float> { using T = float2; };

__device__ __forceinline__ void ld(const double2 *p, size_t i, double &re, double &im)
{
    const double2 v = p[i];
    re = v.x;
    im = v.y;
}
__device__ __forceinline__ void ld(const float2 *p, size_t i, double &re, double &im)
{
    const float2 v = p[i];
    re = v.x;
    im = v.y;
}
__device__ __forceinline__ void st(double2 *p, size_t i, double re, double im) { p[i] = make_double2(re, im); }
__device__ __forceinline__ void st(float2 *p, size_t i, double re, double im)
{
    p[i] = make_float2((float)re, (float)im);
}

// per-column tree: after it, red[0..NV*ncol) holds the block sums of thread groups with
// equal (tid % ncol).  NV values per thread.
template <int NV>
__device__ void block_reduce_cols(double *red, double (&v)[NV], int ncol, double *out_block)
{
    const int tid = threadIdx.x;
#pragma unroll
    for (int k = 0; k < NV; ++k) red[NV * tid + k] = v[k];
    __syncthreads();
    for (int s = RED_TPB / 2; s >= ncol && s > 0; s >>= 1) {
        if (tid < s) {
#pragma unroll
            for (int k = 0; k < NV; ++k) red[NV * tid + k] += red[NV * (tid + s) + k];
        }
        __syncthreads();
    }
    if (tid < ncol) {
#pragma unroll
        for (int k = 0; k < NV; ++k) out_block[NV * tid + k] = red[NV * tid + k];
    }
}

// final stage: out[c] = sum over blocks of partial[block][c]; one block per output,
// fixed strides and tree (deterministic)
__global__ __launch_bounds__(256) void reduce_rows(const double *__restrict__ partial, int nrows, int width,
                                                   double *__restrict__ out)
{
    __shared__ double red[256];
    const int c = blockIdx.x, tid = threadIdx.x;
    double s = 0.0;
    for (int b = tid; b < nrows; b += 256) s += partial[(size_t)b * width + c];
    red[tid] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (tid < k) red[tid] += red[tid + k];
        __syncthreads();
    }
    if (tid == 0) out[c] = red[0];
}

// dot: sum conj(x) y per column -> 2 values per column
template <typename Real>
__global__ __launch_bounds__(RED_TPB) void k_dot(const typename V2<Real>::T *__restrict__ x,
                                                 const typename V2<Real>::T *__restrict__ y,
                                                 size_t nelem, int ncol, double *__restrict__ partial)
{
    __shared__ double red[2 * RED_TPB];
    const size_t stride = (size_t)RED_TPB * gridDim.x;
    double acc[2] = {0.0, 0.0};
    for (size_t e = (size_t)blockIdx.x * RED_TPB + threadIdx.x; e < nelem; e += stride) {
        double xr, xi, yr, yi;
        ld(x, e, xr, xi);
        ld(y, e, yr, yi);
        acc[0] = fma(xr, yr, fma(xi, yi, acc[0]));
        acc[1] = fma(xr, yi, fma(-xi, yr, acc[1]));
    }
    block_reduce_cols<2>(red, acc, ncol, partial + (size_t)blockIdx.x * 2 * ncol);
}

// bilinear (unconjugated) product: sum x y per column -> 2 values per column
template <typename Real>
__global__ __launch_bounds__(RED_TPB) void k_dotu(const typename V2<Real>::T *__restrict__ x,
                                                  const typename V2<Real>::T *__restrict__ y,
                                                  size_t nelem, int ncol, double *__restrict__ partial)
{
    __shared__ double red[2 * RED_TPB];
    const size_t stride = (size_t)RED_TPB * gridDim.x;
    double acc[2] = {0.0, 0.0};
    for (size_t e = (size_t)blockIdx.x * RED_TPB + threadIdx.x; e < nelem; e += stride) {
        double xr, xi, yr, yi;
        ld(x, e, xr, xi);
        ld(y, e, yr, yi);
        acc[0] = fma(xr, yr, fma(-xi, yi, acc[0]));
        acc[1] = fma(xr, yi, fma(xi, yr, acc[1]));
    }
    block_reduce_cols<2>(red, acc, ncol, partial + (size_t)blockIdx.x * 2 * ncol);
}

// beta2: sum |v0 - alpha_c v1|^2 per column (alpha complex per column, in constant args)
template <typename Real>
__global__ __launch_bounds__(RED_TPB) void k_beta2(const typename V2<Real>::T *__restrict__ v1,
                                                   const typename V2<Real>::T *__restrict__ v0,
                                                   const double *__restrict__ alpha, size_t nelem,
                                                   int ncol, double *__restrict__ partial)
{
    __shared__ double red[RED_TPB];
    const size_t stride = (size_t)RED_TPB * gridDim.x;
    const size_t e0 = (size_t)blockIdx.x * RED_TPB + threadIdx.x;
    const int col = (int)(e0 % (size_t)ncol);
    const double ar = alpha[2 * col], ai = alpha[2 * col + 1];
    double acc[1] = {0.0};
    for (size_t e = e0; e < nelem; e += stride) {
        double xr, xi, yr, yi;
        ld(v1, e, xr, xi);
        ld(v0, e, yr, yi);
        const double wr = yr - (ar * xr - ai * xi);
        const double wi = yi - (ar * xi + ai * xr);
        acc[0] = fma(wr, wr, fma(wi, wi, acc[0]));
    }
    block_reduce_cols<1>(red, acc, ncol, partial + (size_t)blockIdx.x * ncol);
}

// Lanczos step with the normalisations left to the next H.v (column c): x holds v1 / s_c,
// y holds H v1 - beta v_prev; with alpha_c = s_c prd_c (prd = <x|H|s x> from the H.v),
//   y <- y - Re(alpha_c) s_c x          (the next v1, times beta)
//   acc += |y - alpha_c s_c x|^2        (beta^2 as calc_beta forms it, complex alpha)
// one read of x and y and one write of y.  sc = [s (ncol)], prd = [re, im] per column.
template <typename Real>
__global__ __launch_bounds__(RED_TPB) void k_lanczos_lazy(const typename V2<Real>::T *__restrict__ x,
                                                          typename V2<Real>::T *__restrict__ y,
                                                          const double *__restrict__ sc, const double *__restrict__ prd,
                                                          size_t nelem, int ncol, double *__restrict__ partial)
{
    __shared__ double red[RED_TPB];
    const size_t stride = (size_t)RED_TPB * gridDim.x;
    const size_t e0 = (size_t)blockIdx.x * RED_TPB + threadIdx.x;
    const int col = (int)(e0 % (size_t)ncol);
    const double s = sc[col];
    const double ar = s * prd[2 * col], ai = s * prd[2 * col + 1];
    double acc[1] = {0.0};
    for (size_t e = e0; e < nelem; e += stride) {
        double xr, xi, yr, yi;
        ld(x, e, xr, xi);
        ld(y, e, yr, yi);
        xr *= s;
        xi *= s;
        const double wr = yr - (ar * xr - ai * xi);
        const double wi = yi - (ar * xi + ai * xr);
        acc[0] = fma(wr, wr, fma(wi, wi, acc[0]));
        st(y, e, yr - ar * xr, yi - ar * xi);
    }
    block_reduce_cols<1>(red, acc, ncol, partial + (size_t)blockIdx.x * ncol);
}

// back to the reference's vectors after deferred-normalisation steps: with (x, y) the
// buffers holding v1/sx and v0 - a v1 (swapped: x is p0, y is p1), p1 <- sx x and
// p0 <- y + a sx x.  Each element is read before it is written by the same thread.
template <typename Real>
__global__ void k_lanczos_restore(typename V2<Real>::T *__restrict__ p1, typename V2<Real>::T *__restrict__ p0,
                                  int swapped, double sx, double a, size_t nelem)
{
    const size_t stride = (size_t)blockDim.x * gridDim.x;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nelem; e += stride) {
        double xr, xi, yr, yi;
        if (swapped) {
            ld(p0, e, xr, xi);
            ld(p1, e, yr, yi);
        } else {
            ld(p1, e, xr, xi);
            ld(p0, e, yr, yi);
        }
        const double tr = sx * xr, ti = sx * xi;
        st(p1, e, tr, ti);
        st(p0, e, yr + a * tr, yi + a * ti);
    }
}

// one COCG update (one column): x += a p, r -= a y, and of the new r the sums
// <r|r> (re, im) and r^T r (re, im): one pass over p, y, x, r instead of two axpy passes
// and two reductions
template <typename Real>
__global__ __launch_bounds__(RED_TPB) void k_cocg_update(const typename V2<Real>::T *__restrict__ p,
                                                         const typename V2<Real>::T *__restrict__ y,
                                                         typename V2<Real>::T *__restrict__ x,
                                                         typename V2<Real>::T *__restrict__ r, double ar, double ai,
                                                         size_t nelem, double *__restrict__ partial)
{
    __shared__ double red[4 * RED_TPB];
    const size_t stride = (size_t)RED_TPB * gridDim.x;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (size_t e = (size_t)blockIdx.x * RED_TPB + threadIdx.x; e < nelem; e += stride) {
        double pr, pi, yr, yi, xr, xi, rr, ri;
        ld(p, e, pr, pi);
        ld(y, e, yr, yi);
        ld(x, e, xr, xi);
        ld(r, e, rr, ri);
        st(x, e, ar * pr - ai * pi + xr, ar * pi + ai * pr + xi);
        rr = rr - (ar * yr - ai * yi);
        ri = ri - (ar * yi + ai * yr);
        st(r, e, rr, ri);
        acc[0] = fma(rr, rr, fma(ri, ri, acc[0]));
        acc[1] = fma(rr, ri, fma(-ri, rr, acc[1]));
        acc[2] = fma(rr, rr, fma(-ri, ri, acc[2]));
        acc[3] = fma(rr, ri, fma(ri, rr, acc[3]));
    }
    block_reduce_cols<4>(red, acc, 1, partial + (size_t)blockIdx.x * 4);
}

// y = a_c x + b_c y  (complex scalars per column)
template <typename Real>
__global__ void k_axpby(const typename V2<Real>::T *__restrict__ x, typename V2<Real>::T *__restrict__ y,
                        const double *__restrict__ a, const double *__restrict__ b, size_t nelem, int ncol)
{
    const size_t stride = (size_t)blockDim.x * gridDim.x;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nelem; e += stride) {
        const int c = (int)(e % (size_t)ncol);
        double xr, xi, yr, yi;
        ld(x, e, xr, xi);
        ld(y, e, yr, yi);
        const double ar = a[2 * c], ai = a[2 * c + 1], br = b[2 * c], bi = b[2 * c + 1];
        st(y, e, ar * xr - ai * xi + br * yr - bi * yi, ar * xi + ai * xr + br * yi + bi * yr);
    }
}

template <typename Real>
__global__ void k_scale(typename V2<Real>::T *__restrict__ x, const double *__restrict__ a, size_t nelem,
                        int ncol)
{
    const size_t stride = (size_t)blockDim.x * gridDim.x;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nelem; e += stride) {
        const int c = (int)(e % (size_t)ncol);
        double xr, xi;
        ld(x, e, xr, xi);
        const double ar = a[2 * c], ai = a[2 * c + 1];
        st(x, e, ar * xr - ai * xi, ar * xi + ai * xr);
    }
}

// Lanczos three-term update (Hybrid_TPS_Lanczos.f90:111-116), alpha real per column
template <typename Real>
__global__ void k_lanczos_update(typename V2<Real>::T *__restrict__ v1, typename V2<Real>::T *__restrict__ v0,
                                 const double *__restrict__ ab, size_t nelem, int ncol)
{
    const size_t stride = (size_t)blockDim.x * gridDim.x;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nelem; e += stride) {
        const int c = (int)(e % (size_t)ncol);
        const double al = ab[2 * c], be = ab[2 * c + 1];
        double xr, xi, yr, yi;
        ld(v1, e, xr, xi);
        ld(v0, e, yr, yi);
        st(v0, e, -be * xr, -be * xi);
        st(v1, e, (yr - al * xr) / be, (yi - al * xi) / be);
    }
}

__device__ __forceinline__ void philox(uint32_t &c0, uint32_t &c1, uint32_t &c2, uint32_t &c3, uint32_t k0,
                                       uint32_t k1);

// per-column Lanczos update for slot batches: mode 0 idle (v1 = v0 = 0), 1 start (v0 = 0,
// v1 holds the start vector), 2 step (the update above with that column's alpha, beta)
template <typename Real>
__global__ void k_lanczos_slots(typename V2<Real>::T *__restrict__ v1, typename V2<Real>::T *__restrict__ v0,
                                const double *__restrict__ ab, const int *__restrict__ mode, size_t nelem,
                                int ncol)
{
    const size_t stride = (size_t)blockDim.x * gridDim.x;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nelem; e += stride) {
        const int c = (int)(e % (size_t)ncol);
        const int m = mode[c];
        if (m == 1) {
            st(v0, e, 0.0, 0.0);
        } else if (m == 2) {
            const double al = ab[2 * c], be = ab[2 * c + 1];
            double xr, xi, yr, yi;
            ld(v1, e, xr, xi);
            ld(v0, e, yr, yi);
            st(v0, e, -be * xr, -be * xi);
            st(v1, e, (yr - al * xr) / be, (yi - al * xi) / be);
        } else {
            st(v1, e, 0.0, 0.0);
            st(v0, e, 0.0, 0.0);
        }
    }
}

// Gaussian start vector of one column: values depend on (seed, state) only, not on
// the column a request occupies; returns the column's norm^2 partials
template <typename Real>
__global__ __launch_bounds__(RED_TPB) void k_gauss_col(typename V2<Real>::T *__restrict__ v, uint64_t seed,
                                                       size_t dim, int ncol, int col, SwapArgs sw,
                                                       double *__restrict__ partial)
{
    __shared__ double red[RED_TPB];
    const size_t stride = (size_t)RED_TPB * gridDim.x;
    double acc[1] = {0.0};
    for (size_t i = (size_t)blockIdx.x * RED_TPB + threadIdx.x; i < dim; i += stride) {
        const uint64_t bs = swap_bits(sw, i);   // basis state of memory state i
        uint32_t c0 = (uint32_t)bs, c1 = (uint32_t)(bs >> 32), c2 = 0u, c3 = 0u;
        philox(c0, c1, c2, c3, (uint32_t)seed, (uint32_t)(seed >> 32));
        const uint64_t aa = ((uint64_t)c0 << 21) ^ (c1 >> 11);
        const uint64_t bb = ((uint64_t)c2 << 21) ^ (c3 >> 11);
        const double u1 = ((double)(aa & ((1ull << 53) - 1)) + 1.0) * 0x1.0p-53;
        const double u2 = (double)(bb & ((1ull << 53) - 1)) * 0x1.0p-53;
        const double r = sqrt(-2.0 * log(u1));
        double sn, cs;
        sincos(6.283185307179586476925286766559 * u2, &sn, &cs);
        const size_t e = i * (size_t)ncol + col;
        st(v, e, r * cs, r * sn);
        double qr, qi;
        ld(v, e, qr, qi);
        acc[0] = fma(qr, qr, fma(qi, qi, acc[0]));
    }
    block_reduce_cols<1>(red, acc, 1, partial + blockIdx.x);
}

template <typename Real>
__global__ void k_scale_col(typename V2<Real>::T *__restrict__ v, double s, size_t dim, int ncol, int col)
{
    const size_t stride = (size_t)blockDim.x * gridDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += stride) {
        const size_t e = i * (size_t)ncol + col;
        double xr, xi;
        ld(v, e, xr, xi);
        st(v, e, s * xr, s * xi);
    }
}

// mTPQ update (Hybrid_TPS_TPS.f90:380-386): vg = v1; v1 = xnorm*v0; v0 = 0
template <typename Real>
__global__ void k_tps_update(typename V2<Real>::T *__restrict__ v1, typename V2<Real>::T *__restrict__ v0,
                             typename V2<Real>::T *__restrict__ vg, const double *__restrict__ xn,
                             size_t nelem, int ncol)
{
    const size_t stride = (size_t)blockDim.x * gridDim.x;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nelem; e += stride) {
        const int c = (int)(e % (size_t)ncol);
        double xr, xi, yr, yi;
        ld(v1, e, xr, xi);
        ld(v0, e, yr, yi);
        st(vg, e, xr, xi);
        st(v1, e, xn[c] * yr, xn[c] * yi);
        st(v0, e, 0.0, 0.0);
    }
}

// ---- Philox-4x32-10 (Salmon et al., SC'11) + Box-Muller --------------------------------
__device__ __forceinline__ void philox(uint32_t &c0, uint32_t &c1, uint32_t &c2, uint32_t &c3, uint32_t k0,
                                       uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
        const uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

template <typename Real>
__global__ __launch_bounds__(RED_TPB) void k_gauss(typename V2<Real>::T *__restrict__ v, uint64_t seed,
                                                   size_t nelem, int ncol, SwapArgs sw, double *__restrict__ partial)
{
    __shared__ double red[RED_TPB];
    const size_t stride = (size_t)RED_TPB * gridDim.x;
    double acc[1] = {0.0};
    for (size_t e = (size_t)blockIdx.x * RED_TPB + threadIdx.x; e < nelem; e += stride) {
        const uint64_t i = swap_bits(sw, e / (size_t)ncol);   // basis state
        const uint32_t col = (uint32_t)(e % (size_t)ncol);
        uint32_t c0 = (uint32_t)i, c1 = (uint32_t)(i >> 32), c2 = col, c3 = 0u;
        philox(c0, c1, c2, c3, (uint32_t)seed, (uint32_t)(seed >> 32));
        const uint64_t a = ((uint64_t)c0 << 21) ^ (c1 >> 11);
        const uint64_t b = ((uint64_t)c2 << 21) ^ (c3 >> 11);
        const double u1 = ((double)(a & ((1ull << 53) - 1)) + 1.0) * 0x1.0p-53;
        const double u2 = (double)(b & ((1ull << 53) - 1)) * 0x1.0p-53;
        const double r = sqrt(-2.0 * log(u1));
        const double th = 6.283185307179586476925286766559 * u2;
        double s, c;
        sincos(th, &s, &c);
        const double re = r * c, im = r * s;
        st(v, e, re, im);
        double qr, qi;
        ld(v, e, qr, qi);   // norm of the stored (possibly rounded) value
        acc[0] = fma(qr, qr, fma(qi, qi, acc[0]));
    }
    block_reduce_cols<1>(red, acc, ncol, partial + (size_t)blockIdx.x * ncol);
}

// ---- excitation S^z(q) x (intalv_zz, Hybrid_TPS_DSF.f90:18-87) ---------------------------
struct SzArgs {
    int n, q;
    double phr[32], phi[32];  // e^{i q.r_j}
    double wlow;              // weight offset: 0.5 (S=1/2) or S
    int8_t pos[32];           // S=1/2: memory bit of site j
};

template <typename Real>
__global__ void k_sz_q(const typename V2<Real>::T *__restrict__ x, typename V2<Real>::T *__restrict__ v,
                       SzArgs a, size_t dim)
{
    const size_t stride = (size_t)blockDim.x * gridDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += stride) {
        double cr = 0.0, ci = 0.0;
        uint64_t s = i;
        for (int j = 0; j < a.n; ++j) {
            int d;
            if (a.q == 2) {
                d = (int)((i >> a.pos[j]) & 1u);
            } else {
                d = (int)(s % (uint64_t)a.q);
                s /= (uint64_t)a.q;
            }
            const double w = (double)d - a.wlow;
            cr = fma(w, a.phr[j], cr);
            ci = fma(w, a.phi[j], ci);
        }
        double xr, xi;
        ld(x, i, xr, xi);
        st(v, i, cr * xr - ci * xi, cr * xi + ci * xr);
    }
}

// ---- excitation S^{+/-}(q) x over the full q^N space (intalv_xy at parity_type != 0:
// S_arbitrary/Hybrid_TPS9_DSF.f90:91-286, S_one_half/Hybrid_TPS_DSF.f90:89-193) ----------------
// One thread per target element t gathers, over the sites j in ascending order (the order
// the reference's site-outer scatter adds them to one element), fac(d_s) e^{i q.r_j} x(s)
// where s has digit j moved back by ipm and d_s is that source digit.
struct PmArgs {
    int n, q, ipm;
    double phr[32], phi[32];
    double fac[16];        // ladder factor by source digit (1 for the S=1/2 tree)
    uint64_t pw[32];       // q^j
    SwapArgs sw_v, sw_x;   // S=1/2: memory layouts of v and x (basis state <-> memory index)
};

__global__ __launch_bounds__(256) void k_intalv_pm(const double2 *__restrict__ x, double2 *__restrict__ v, PmArgs a,
                                                   uint64_t dim)
{
    const uint64_t stride = (uint64_t)blockDim.x * gridDim.x;
    for (uint64_t l = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; l < dim; l += stride) {
        double ar = 0.0, ai = 0.0;
        uint32_t rem = (uint32_t)l;   // q >= 3: dim < 2^32 (checked on the host)
        const uint64_t bt = (a.q == 2) ? swap_bits(a.sw_v, l) : l;   // basis state of the target
        for (int j = 0; j < a.n; ++j) {
            int dt;
            if (a.q == 2) {
                dt = (int)((bt >> j) & 1u);
            } else {
                dt = (int)(rem % (uint32_t)a.q);
                rem /= (uint32_t)a.q;
            }
            const int ds = dt - a.ipm;
            if (ds < 0 || ds >= a.q) continue;
            const uint64_t bsrc = a.ipm > 0 ? bt - a.pw[j] : bt + a.pw[j];
            const double2 y = x[(a.q == 2) ? swap_bits(a.sw_x, bsrc) : bsrc];
            const double rr = a.phr[j] * y.x - a.phi[j] * y.y, ri = a.phi[j] * y.x + a.phr[j] * y.y;
            ar += a.fac[ds] * rr;
            ai += a.fac[ds] * ri;
        }
        v[l] = make_double2(ar, ai);
    }
}

// ---- mTPQ observables of eval_phys_val (Hybrid_TPS_TPS.f90:83-322), fused ------------------
// Per column: out[0..1] = <vg|v1>, out[2..3] = <v1|Wzz|v1>, out[4..5] = <vg|Wzz|v1>,
// then for each requested structure-factor entry k: out[6+2k..] = <a_k|W_k|v1>.
constexpr int OBS_MAX_SQ = 32;
struct ObsArgs {
    int n;
    uint32_t zz_mask1[64], zz_mask2[64];  // z-type bond bit masks (off-by-one bits kept)
    int nzz;
    int nsq;                              // number of structure-factor entries computed
    int sq_mode[OBS_MAX_SQ];              // 0: zz weight, 1: pp weight
    int sq_left[OBS_MAX_SQ];              // 0: <v1|, 1: <vg|
    int sq_q[OBS_MAX_SQ];                 // momentum slot
    double phr[8][32], phi[8][32];        // e^{i q.r_j} per momentum slot (entry k uses slot k / 4)
    int8_t pos[32];                       // memory bit of site j
    // Entry k's weight of a state, W_k(s) = sum_j w_k(s_j, s_0) e^{i q.r_j}, with w_k(b, b0)
    // by mode (zz or pp) of the two bits: S=1/2 tree (b - 1/2)(b0 - 1/2) and [b != b0]/4;
    // spin-S tree (b - S)(b0 - S) and the 2x2 corner of sx + sy
    // (S_arbitrary/Hybrid_TPS9_TPS.f90:175-177, 224-243).  As w_k depends on site j only
    // through its bit, W_k(s) = A_k(s_0) + D_k(s_0) Q(s) with Q(s) the sum of the phases of the
    // sites whose bit is set: sa[k][b0] = w_k(0, b0) sum_j e^{i q.r_j}, sd[k][b0] =
    // w_k(1, b0) - w_k(0, b0), each (re, im).
    double sa[OBS_MAX_SQ][2][2], sd[OBS_MAX_SQ][2][2];
};

// W_k(s) of the entries k < nsq (at most MAXK) of one state, handed to emit(k, re, im)
template <int MAXK, typename F>
__device__ __forceinline__ void sq_weights(const ObsArgs &a, uint32_t s, int nsq, F &&emit)
{
    constexpr int NSLOT = (MAXK + 3) / 4;
    const int nslot = (nsq + 3) / 4;
    double qr[NSLOT], qi[NSLOT];
#pragma unroll
    for (int m = 0; m < NSLOT; ++m) {
        qr[m] = 0.0;
        qi[m] = 0.0;
    }
    for (int j = 0; j < a.n; ++j) {
        const double sj = (double)((s >> a.pos[j]) & 1u);
#pragma unroll
        for (int m = 0; m < NSLOT; ++m) {
            if (m >= nslot) break;
            qr[m] = fma(sj, a.phr[m][j], qr[m]);
            qi[m] = fma(sj, a.phi[m][j], qi[m]);
        }
    }
    const bool b0 = (s >> a.pos[0]) & 1u;
#pragma unroll
    for (int k = 0; k < MAXK; ++k) {
        if (k >= nsq) break;
        const double ar = b0 ? a.sa[k][1][0] : a.sa[k][0][0], ai = b0 ? a.sa[k][1][1] : a.sa[k][0][1];
        const double dr = b0 ? a.sd[k][1][0] : a.sd[k][0][0], di = b0 ? a.sd[k][1][1] : a.sd[k][0][1];
        emit(k, fma(dr, qr[k / 4], fma(-di, qi[k / 4], ar)), fma(dr, qi[k / 4], fma(di, qr[k / 4], ai)));
    }
}

template <typename Real, int NOUT>
__global__ __launch_bounds__(RED_TPB) void k_tps_obs(const typename V2<Real>::T *__restrict__ vg,
                                                     const typename V2<Real>::T *__restrict__ v1,
                                                     const ObsArgs *__restrict__ ap, size_t nelem,
                                                     int ncol, double *__restrict__ partial)
{
    extern __shared__ double red[];
    const ObsArgs &a = *ap;
    const size_t stride = (size_t)RED_TPB * gridDim.x;
    double acc[NOUT];
#pragma unroll
    for (int k = 0; k < NOUT; ++k) acc[k] = 0.0;
    const int nout = 6 + 2 * a.nsq;
    for (size_t e = (size_t)blockIdx.x * RED_TPB + threadIdx.x; e < nelem; e += stride) {
        const uint32_t s = (uint32_t)(e / (size_t)ncol);
        double gr, gi, xr, xi;
        ld(vg, e, gr, gi);
        ld(v1, e, xr, xi);
        acc[0] = fma(gr, xr, fma(gi, xi, acc[0]));
        acc[1] = fma(gr, xi, fma(-gi, xr, acc[1]));
        double wzz = 0.0;
        for (int b = 0; b < a.nzz; ++b) {
            const double s3 = (s & a.zz_mask1[b]) ? 0.5 : -0.5;
            const double s4 = (s & a.zz_mask2[b]) ? 0.5 : -0.5;
            wzz = fma(s3, s4, wzz);
        }
        const double n1 = xr * xr + xi * xi;
        acc[2] = fma(wzz, n1, acc[2]);
        acc[4] = fma(wzz, gr * xr + gi * xi, acc[4]);
        acc[5] = fma(wzz, gr * xi - gi * xr, acc[5]);
        constexpr int MAXK = (NOUT - 6) / 2;
        sq_weights<MAXK>(a, s, a.nsq < MAXK ? a.nsq : MAXK, [&](int k, double cr, double ci) {
            // W x
            const double wr = cr * xr - ci * xi, wi = cr * xi + ci * xr;
            double lr = xr, li = xi;
            if (a.sq_left[k]) { lr = gr; li = gi; }
            acc[6 + 2 * k] = fma(lr, wr, fma(li, wi, acc[6 + 2 * k]));
            acc[6 + 2 * k + 1] = fma(lr, wi, fma(-li, wr, acc[6 + 2 * k + 1]));
        });
    }
    (void)nout;
    // per-column block reduction of NOUT values
    const int tid = threadIdx.x;
#pragma unroll
    for (int k = 0; k < NOUT; ++k) red[NOUT * tid + k] = acc[k];
    __syncthreads();
    for (int s2 = RED_TPB / 2; s2 >= ncol && s2 > 0; s2 >>= 1) {
        if (tid < s2) {
#pragma unroll
            for (int k = 0; k < NOUT; ++k) red[NOUT * tid + k] += red[NOUT * (tid + s2) + k];
        }
        __syncthreads();
    }
    if (tid < ncol) {
#pragma unroll
        for (int k = 0; k < NOUT; ++k)
            partial[(size_t)blockIdx.x * NOUT * ncol + NOUT * tid + k] = red[NOUT * tid + k];
    }
}

// Fused mTPQ step reductions over (uprev, u, w), per column:
//  [0] |w|^2  [2..3] <uprev|w>  [4..5] <uprev|u>  [6] <u|Wzz|u>  [8..9] <uprev|Wzz|u>
//  [10+2k..] <a_k|W_k|u> (a_k = u or uprev) for the stored structure-factor entries
// Fused mTPQ reduction.  The observables' state weights (z-bond sum, structure-factor
// sums W_k(s) over sites) depend on the state only, and the ncol interleaved columns of a
// state share them: each chunk of CS states first computes them once per state into LDS
// (phase A, spread over the block's threads), then every element of the chunk reads its
// state's weights (phase B).  A thread's elements all lie in one column, so the block
// reduction per column is the fixed tree of reduce_rows.
constexpr int TPS_CS_MAX = 256;   // states per chunk

template <typename Real, int NOUT>
__global__ __launch_bounds__(RED_TPB) void k_tps_fused(const typename V2<Real>::T *__restrict__ up,
                                                       const typename V2<Real>::T *__restrict__ u,
                                                       const typename V2<Real>::T *__restrict__ w,
                                                       const ObsArgs *__restrict__ ap, size_t nstates, int ncol,
                                                       int cs, double *__restrict__ partial)
{
    extern __shared__ double red[];
    double *wt = red + NOUT * RED_TPB;   // [cs][1 + 2 nsq]: wzz, then (Re, Im) W_k
    const ObsArgs &a = *ap;
    const int tid = threadIdx.x;
    const int nsq = (10 + 2 * a.nsq <= NOUT) ? a.nsq : (NOUT - 10) / 2;
    const int nw = 1 + 2 * nsq;
    const int lc = __ffs(ncol) - 1;   // log2 ncol
    double acc[NOUT];
#pragma unroll
    for (int k = 0; k < NOUT; ++k) acc[k] = 0.0;
    const size_t nchunk = (nstates + cs - 1) / cs;
    for (size_t ch = blockIdx.x; ch < nchunk; ch += gridDim.x) {
        const size_t s0c = ch * (size_t)cs;
        const int ns = (int)((nstates - s0c < (size_t)cs) ? (nstates - s0c) : (size_t)cs);
        // phase A: per-state weights, one state per thread
        for (int sl = tid; sl < ns; sl += RED_TPB) {
            const uint32_t s = (uint32_t)(s0c + sl);
            double wzz = 0.0;
            for (int b = 0; b < a.nzz; ++b) {
                const double s3 = (s & a.zz_mask1[b]) ? 0.5 : -0.5;
                const double s4 = (s & a.zz_mask2[b]) ? 0.5 : -0.5;
                wzz = fma(s3, s4, wzz);
            }
            wt[sl * nw] = wzz;
            sq_weights<(NOUT - 10) / 2>(a, s, nsq, [&](int k, double cr, double ci) {
                wt[sl * nw + 1 + 2 * k] = cr;
                wt[sl * nw + 2 + 2 * k] = ci;
            });
        }
        __syncthreads();
        // phase B: the chunk's elements
        const int nel = ns << lc;
        for (int le = tid; le < nel; le += RED_TPB) {
            const size_t e = (s0c << lc) + (size_t)le;
            const double *ws = wt + (le >> lc) * nw;
            double pr, pi, xr, xi, wr, wi;
            ld(up, e, pr, pi);
            ld(u, e, xr, xi);
            ld(w, e, wr, wi);
            acc[0] = fma(wr, wr, fma(wi, wi, acc[0]));
            acc[2] = fma(pr, wr, fma(pi, wi, acc[2]));
            acc[3] = fma(pr, wi, fma(-pi, wr, acc[3]));
            acc[4] = fma(pr, xr, fma(pi, xi, acc[4]));
            acc[5] = fma(pr, xi, fma(-pi, xr, acc[5]));
            const double wzz = ws[0];
            acc[6] = fma(wzz, xr * xr + xi * xi, acc[6]);
            acc[8] = fma(wzz, pr * xr + pi * xi, acc[8]);
            acc[9] = fma(wzz, pr * xi - pi * xr, acc[9]);
#pragma unroll
            for (int k = 0; 10 + 2 * k + 1 < NOUT; ++k) {
                if (k >= nsq) break;
                const double cr = ws[1 + 2 * k], ci = ws[2 + 2 * k];
                const double yr = cr * xr - ci * xi, yi = cr * xi + ci * xr;
                double lr = xr, li = xi;
                if (a.sq_left[k]) { lr = pr; li = pi; }
                acc[10 + 2 * k] = fma(lr, yr, fma(li, yi, acc[10 + 2 * k]));
                acc[10 + 2 * k + 1] = fma(lr, yi, fma(-li, yr, acc[10 + 2 * k + 1]));
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < NOUT; ++k) red[NOUT * tid + k] = acc[k];
    __syncthreads();
    for (int s2 = RED_TPB / 2; s2 >= ncol && s2 > 0; s2 >>= 1) {
        if (tid < s2) {
#pragma unroll
            for (int k = 0; k < NOUT; ++k) red[NOUT * tid + k] += red[NOUT * (tid + s2) + k];
        }
        __syncthreads();
    }
    if (tid < ncol) {
#pragma unroll
        for (int k = 0; k < NOUT; ++k)
            partial[(size_t)blockIdx.x * NOUT * ncol + NOUT * tid + k] = red[NOUT * tid + k];
    }
}

// ---- continued fraction (exec_renbun, Hybrid_TPS_DSF.f90:263-334): one thread per omega ----
__global__ void k_renbun(int i_range, int s_maxitr, double split, double ene, double eps,
                         const double *__restrict__ salpha, const double *__restrict__ sbeta,
                         double *__restrict__ w, double *__restrict__ dsf, int *__restrict__ conv)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= i_range) return;
    const double omega = (double)j * split;
    w[j] = omega;
    int i_conv = 0;
    double xr = 1.0e-10, xi = 1.0e-10, cr = xr, ci = xi, dr = 0.0, di = 0.0;
    for (int m1 = 1; m1 <= s_maxitr - 1; ++m1) {
        const double f1 = omega + ene - salpha[m1 - 1];
        const double f2 = -1.0 * (sbeta[m1 - 1] * sbeta[m1 - 1]);
        const double xr1 = xr, xi1 = xi;
        dr = dr * f2 + f1;
        di = di * f2 + eps;
        double d = dr * dr + di * di;
        if (d == 0.0) { dr = 0.5e-28; di = 0.5e-28; }
        double c = cr * cr + ci * ci;
        cr = cr * f2 / c + f1;
        ci = -ci * f2 / c + eps;
        c = cr * cr + ci * ci;
        if (c == 0.0) { cr = 0.5e-28; ci = 0.5e-28; }
        d = dr * dr + di * di;
        dr = dr / d;
        di = -di / d;
        const double dlr = cr * dr - ci * di, dli = ci * dr + cr * di;
        xr = xr1 * dlr - xi1 * dli;
        xi = xr1 * dli + xi1 * dlr;
        const double dl = (dlr - 1.0) * (dlr - 1.0) + dli * dli;
        if (dl < 0.5e-28) { i_conv = m1; break; }
    }
    conv[j] = i_conv;
    dsf[j] = xi;
}

inline int grid_for(size_t n, int tpb, int cap = 8192)
{
    size_t g = (n + tpb - 1) / tpb;
    if (g > (size_t)cap) g = cap;
    if (g < 1) g = 1;
    return (int)g;
}

int check_pow2_cols(int ncol)
{
    if (ncol < 1 || (ncol & (ncol - 1)) || ncol > RED_TPB)
        return fail(PNED_E_UNSUPPORTED, "column count must be a power of two <= 256");
    return PNED_OK;
}

// run the final stage and copy 'width' doubles to host (synchronizing)
int finish_reduce(pned_ctx *ctx, int nrows, int width, double *out)
{
    PNED_TRY(ensure_result(ctx, (size_t)width));
    hipLaunchKernelGGL(reduce_rows, dim3(width), dim3(256), 0, ctx->stream, ctx->d_partial, nrows, width,
                       ctx->d_result);
    PNED_HIP(hipGetLastError());
    PNED_HIP(hipMemcpyAsync(ctx->h_result, ctx->d_result, sizeof(double) * width, hipMemcpyDeviceToHost,
                            ctx->stream));
    PNED_HIP(hipStreamSynchronize(ctx->stream));
    for (int c = 0; c < width; ++c) out[c] = ctx->h_result[c];
    return PNED_OK;
}

int same_shape(const pned_vec *a, const pned_vec *b)
{
    if (vec_detached(a) || vec_detached(b)) return fail(PNED_E_ARG, "vector of a destroyed context");
    if (!a || !b || a->dim != b->dim || a->nvec != b->nvec || a->dtype != b->dtype)
        return fail(PNED_E_ARG, "vector shapes or types differ");
    return check_same_layout(a, b);
}

// layout-mapped staging copy: memory state swap(first + i) <-> stage row i, every row
// 'words' 8-byte words (all columns of one state); to_dev = 1 writes the vector
__global__ void k_layout_copy(uint64_t *__restrict__ vec, uint64_t *__restrict__ stage, SwapArgs sw, uint64_t first,
                              uint64_t count, int words, int to_dev)
{
    const uint64_t total = count * (uint64_t)words;
    const uint64_t stride = (uint64_t)blockDim.x * gridDim.x;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
        const uint64_t i = e / (uint64_t)words, w = e % (uint64_t)words;
        const uint64_t m = swap_bits(sw, first + i) * (uint64_t)words + w;
        if (to_dev) vec[m] = stage[e];
        else stage[e] = vec[m];
    }
}

constexpr size_t LAYOUT_STAGE_BYTES = (size_t)64 << 20;

int ensure_stage(pned_ctx *ctx)
{
    if (ctx->d_stage) return PNED_OK;
    if (hipMalloc((void **)&ctx->d_stage, LAYOUT_STAGE_BYTES) != hipSuccess) {
        (void)hipGetLastError();
        ctx->d_stage = nullptr;
        return fail(PNED_E_NOMEM, "device staging buffer for a layout-mapped copy");
    }
    ctx->stage_bytes = LAYOUT_STAGE_BYTES;
    return PNED_OK;
}

// count states from basis state offset, through the staging buffer in chunks
int layout_copy(pned_vec *v, void *host, int64_t offset, int64_t count, bool to_dev)
{
    pned_ctx *ctx = v->ctx;
    const size_t row = elem_bytes(v->dtype) * (size_t)v->nvec;
    if (ctx->lay.identity()) {
        char *dev = (char *)v->data + row * (size_t)offset;
        if (to_dev) PNED_HIP(hipMemcpyAsync(dev, host, row * (size_t)count, hipMemcpyHostToDevice, ctx->stream));
        else PNED_HIP(hipMemcpyAsync(host, dev, row * (size_t)count, hipMemcpyDeviceToHost, ctx->stream));
        PNED_HIP(hipStreamSynchronize(ctx->stream));
        return PNED_OK;
    }
    PNED_TRY(ensure_stage(ctx));
    const SwapArgs sw = swap_args(ctx->lay);
    const int64_t per = (int64_t)(ctx->stage_bytes / row);
    for (int64_t done = 0; done < count;) {
        const int64_t c = std::min<int64_t>(per, count - done);
        char *h = (char *)host + row * (size_t)done;
        const int grid = grid_for((size_t)c * row / 8, 256);
        if (to_dev)
            PNED_HIP(hipMemcpyAsync(ctx->d_stage, h, row * (size_t)c, hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(k_layout_copy, dim3(grid), dim3(256), 0, ctx->stream, (uint64_t *)v->data,
                           (uint64_t *)ctx->d_stage, sw, (uint64_t)(offset + done), (uint64_t)c, (int)(row / 8),
                           to_dev ? 1 : 0);
        PNED_HIP(hipGetLastError());
        if (!to_dev)
            PNED_HIP(hipMemcpyAsync(h, ctx->d_stage, row * (size_t)c, hipMemcpyDeviceToHost, ctx->stream));
        PNED_HIP(hipStreamSynchronize(ctx->stream));
        done += c;
    }
    return PNED_OK;
}

}  // namespace

int vec_to_host(const pned_vec *v, void *host, int64_t offset, int64_t count)
{
    PNED_HIP(hipSetDevice(v->ctx->device));
    return layout_copy(const_cast<pned_vec *>(v), host, offset, count, false);
}

int vec_from_host(pned_vec *v, const void *host, int64_t offset, int64_t count)
{
    PNED_HIP(hipSetDevice(v->ctx->device));
    return layout_copy(v, const_cast<void *>(host), offset, count, true);
}

int beta2_after_hv(pned_ctx *ctx, const pned_vec *v1, const pned_vec *v0, double *pr, double *b2)
{
    PNED_TRY(same_shape(v1, v0));
    PNED_TRY(check_pow2_cols(v1->nvec));
    const int nc = v1->nvec;
    const size_t nelem = (size_t)v1->dim * nc;
    const int grid = RED_GRID;
    PNED_TRY(ensure_partial(ctx, (size_t)grid * nc));
    PNED_TRY(ensure_result(ctx, 3 * (size_t)nc));
    double *da;
    PNED_TRY(ensure_scalars(ctx, 2 * nc, &da));
    PNED_HIP(hipMemcpyAsync(da, ctx->d_result, sizeof(double) * 2 * nc, hipMemcpyDeviceToDevice, ctx->stream));
    if (v1->dtype == PNED_C128)
        hipLaunchKernelGGL(k_beta2<double>, dim3(grid), dim3(RED_TPB), 0, ctx->stream, (const double2 *)v1->data,
                           (const double2 *)v0->data, (const double *)da, nelem, nc, ctx->d_partial);
    else
        hipLaunchKernelGGL(k_beta2<float>, dim3(grid), dim3(RED_TPB), 0, ctx->stream, (const float2 *)v1->data,
                           (const float2 *)v0->data, (const double *)da, nelem, nc, ctx->d_partial);
    PNED_HIP(hipGetLastError());
    hipLaunchKernelGGL(reduce_rows, dim3(nc), dim3(256), 0, ctx->stream, ctx->d_partial, grid, nc, ctx->d_result);
    PNED_HIP(hipGetLastError());
    PNED_HIP(hipMemcpyAsync(ctx->h_result, ctx->d_result, sizeof(double) * nc, hipMemcpyDeviceToHost, ctx->stream));
    PNED_HIP(hipMemcpyAsync(ctx->h_result + nc, da, sizeof(double) * 2 * nc, hipMemcpyDeviceToHost, ctx->stream));
    PNED_HIP(hipStreamSynchronize(ctx->stream));
    for (int c = 0; c < nc; ++c) b2[c] = ctx->h_result[c];
    for (int c = 0; c < 2 * nc; ++c) pr[c] = ctx->h_result[nc + c];
    return PNED_OK;
}

int lanczos_lazy_step(pned_ctx *ctx, pned_vec *x, pned_vec *y, const double *sx, const double *sy, bool fresh,
                      double *pr, double *b2)
{
    PNED_TRY(check_hv_pair(ctx, x, y));
    PNED_TRY(check_pow2_cols(x->nvec));
    const bool spin = ctx->q >= 3 && ctx->q <= 5 && ctx->spin_tiled;
    if (!((ctx->q == 2 && ctx->sector_k < 0) || (spin && x->nvec == 1 && x->dtype == PNED_C128)))
        return fail(PNED_E_UNSUPPORTED, "lazy Lanczos step: S=1/2 full space or the tiled spin-S kernel");
    const int nc = x->nvec;
    const size_t nelem = (size_t)x->dim * nc;
    const int grid = RED_GRID;
    PNED_TRY(ensure_result(ctx, 3 * (size_t)nc + 8));
    double *d;
    PNED_TRY(ensure_scalars(ctx, 4 * nc, &d));
    std::vector<double> up(2 * nc);
    for (int c = 0; c < nc; ++c) {
        up[c] = sx[c];
        up[nc + c] = sy[c];
    }
    PNED_HIP(hipMemcpyAsync(d, up.data(), sizeof(double) * 2 * nc, hipMemcpyHostToDevice, ctx->stream));
    // y = sy y + H (sx x); the H.v's <x|H|sx x> lands in d_result
    bool fused = false;
    if (spin)
        PNED_TRY(launch_hv_spin_tiled(ctx, x, y, 0.0, fresh, nullptr, sx[0], fresh ? 1.0 : sy[0]));
    else
        PNED_TRY(launch_hv_s12(ctx, x, y, 0.0, fresh, nullptr, d, fresh ? nullptr : d + nc, &fused));
    if (fused) {   // the last pass formed y - Re(alpha) sx x and beta^2 already
        PNED_HIP(hipMemcpyAsync(ctx->h_result, ctx->d_result, sizeof(double) * 4 * nc, hipMemcpyDeviceToHost,
                                ctx->stream));
        PNED_HIP(hipStreamSynchronize(ctx->stream));
        for (int c = 0; c < 2 * nc; ++c) pr[c] = ctx->h_result[c];
        for (int c = 0; c < nc; ++c) b2[c] = ctx->h_result[2 * nc + 2 * c];
        return PNED_OK;
    }
    PNED_TRY(ensure_partial(ctx, (size_t)grid * nc));
    double *dprd = d + 2 * nc;
    PNED_HIP(hipMemcpyAsync(dprd, ctx->d_result, sizeof(double) * 2 * nc, hipMemcpyDeviceToDevice, ctx->stream));
    if (x->dtype == PNED_C128)
        hipLaunchKernelGGL(k_lanczos_lazy<double>, dim3(grid), dim3(RED_TPB), 0, ctx->stream, (const double2 *)x->data,
                           (double2 *)y->data, (const double *)d, (const double *)dprd, nelem, nc, ctx->d_partial);
    else
        hipLaunchKernelGGL(k_lanczos_lazy<float>, dim3(grid), dim3(RED_TPB), 0, ctx->stream, (const float2 *)x->data,
                           (float2 *)y->data, (const double *)d, (const double *)dprd, nelem, nc, ctx->d_partial);
    PNED_HIP(hipGetLastError());
    hipLaunchKernelGGL(reduce_rows, dim3(nc), dim3(256), 0, ctx->stream, ctx->d_partial, grid, nc, ctx->d_result);
    PNED_HIP(hipGetLastError());
    PNED_HIP(hipMemcpyAsync(ctx->h_result, ctx->d_result, sizeof(double) * nc, hipMemcpyDeviceToHost, ctx->stream));
    PNED_HIP(hipMemcpyAsync(ctx->h_result + nc, dprd, sizeof(double) * 2 * nc, hipMemcpyDeviceToHost, ctx->stream));
    PNED_HIP(hipStreamSynchronize(ctx->stream));
    for (int c = 0; c < nc; ++c) b2[c] = ctx->h_result[c];
    for (int c = 0; c < 2 * nc; ++c) pr[c] = ctx->h_result[nc + c];
    return PNED_OK;
}

__global__ void k_copy_f64(const double *__restrict__ src, double *__restrict__ dst, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// the captured form's scalars: st = [sx, sy, prd re, prd im, b2 of the step before]
__global__ void k_lz_scale(double *__restrict__ st, int fresh)
{
    if (threadIdx.x != 0) return;
    if (fresh) {
        st[0] = 1.0;
        st[1] = 0.0;
    } else {
        const double b = sqrt(st[4]), sxo = st[0];   // v1 is y / b, v0 is x times -b sx
        st[0] = 1.0 / b;
        st[1] = -b * sxo;
    }
}

// slot = [prd re, prd im, b2], st[4] = b2; fused: all three from the H.v's result
// ([prd re, prd im, b2]), else prd from st[2..4) and b2 from the beta kernel's result
__global__ void k_lz_store(const double *__restrict__ result, double *__restrict__ st, double *__restrict__ slot,
                           int fused)
{
    if (threadIdx.x != 0) return;
    const double pr = fused ? result[0] : st[2], pi = fused ? result[1] : st[3];
    const double b2 = fused ? result[2] : result[0];
    slot[0] = pr;
    slot[1] = pi;
    slot[2] = b2;
    st[4] = b2;
}

int lanczos_lazy_setup(pned_ctx *ctx, const pned_vec *x, const pned_vec *y)
{
    PNED_TRY(check_hv_pair(ctx, x, y));
    PNED_TRY(ensure_result(ctx, 3 * (size_t)x->nvec + 8));
    return ensure_partial(ctx, (size_t)RED_GRID * x->nvec);
}

int lanczos_lazy_enqueue(pned_ctx *ctx, pned_vec *x, pned_vec *y, bool fresh, double *st, double *slot)
{
    PNED_TRY(check_hv_pair(ctx, x, y));
    if (x->nvec != 1 || ctx->q != 2 || ctx->sector_k >= 0)
        return fail(PNED_E_UNSUPPORTED, "captured Lanczos step: one column of an S=1/2 full-space context");
    hipLaunchKernelGGL(k_lz_scale, dim3(1), dim3(64), 0, ctx->stream, st, fresh ? 1 : 0);
    PNED_HIP(hipGetLastError());
    bool fused = false;
    PNED_TRY(launch_hv_s12(ctx, x, y, 0.0, fresh, nullptr, st, fresh ? nullptr : st + 1, &fused));
    if (!fused) {
        const size_t nelem = (size_t)x->dim;
        hipLaunchKernelGGL(k_copy_f64, dim3(1), dim3(64), 0, ctx->stream, ctx->d_result, st + 2, 2);
        PNED_HIP(hipGetLastError());
        if (x->dtype == PNED_C128)
            hipLaunchKernelGGL(k_lanczos_lazy<double>, dim3(RED_GRID), dim3(RED_TPB), 0, ctx->stream,
                               (const double2 *)x->data, (double2 *)y->data, (const double *)st,
                               (const double *)(st + 2), nelem, 1, ctx->d_partial);
        else
            hipLaunchKernelGGL(k_lanczos_lazy<float>, dim3(RED_GRID), dim3(RED_TPB), 0, ctx->stream,
                               (const float2 *)x->data, (float2 *)y->data, (const double *)st,
                               (const double *)(st + 2), nelem, 1, ctx->d_partial);
        PNED_HIP(hipGetLastError());
        hipLaunchKernelGGL(reduce_rows, dim3(1), dim3(256), 0, ctx->stream, ctx->d_partial, RED_GRID, 1,
                           ctx->d_result);
        PNED_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_lz_store, dim3(1), dim3(64), 0, ctx->stream, ctx->d_result, st, slot, fused ? 1 : 0);
    PNED_HIP(hipGetLastError());
    return PNED_OK;
}

int cocg_update(pned_ctx *ctx, const pned_vec *p, const pned_vec *y, pned_vec *x, pned_vec *r, const double *a,
                double *out4)
{
    PNED_TRY(same_shape(p, y));
    PNED_TRY(same_shape(p, x));
    PNED_TRY(same_shape(p, r));
    if (p->nvec != 1) return fail(PNED_E_UNSUPPORTED, "cocg update: one column");
    const size_t nelem = (size_t)p->dim;
    const int grid = RED_GRID;
    PNED_TRY(ensure_partial(ctx, (size_t)grid * 4));
    if (p->dtype == PNED_C128)
        hipLaunchKernelGGL(k_cocg_update<double>, dim3(grid), dim3(RED_TPB), 0, ctx->stream, (const double2 *)p->data,
                           (const double2 *)y->data, (double2 *)x->data, (double2 *)r->data, a[0], a[1], nelem,
                           ctx->d_partial);
    else
        hipLaunchKernelGGL(k_cocg_update<float>, dim3(grid), dim3(RED_TPB), 0, ctx->stream, (const float2 *)p->data,
                           (const float2 *)y->data, (float2 *)x->data, (float2 *)r->data, a[0], a[1], nelem,
                           ctx->d_partial);
    PNED_HIP(hipGetLastError());
    return finish_reduce(ctx, grid, 4, out4);
}

int lanczos_lazy_restore(pned_ctx *ctx, pned_vec *v1, pned_vec *v0, bool swapped, double sx, double a)
{
    PNED_TRY(same_shape(v1, v0));
    const size_t nelem = (size_t)v1->dim * v1->nvec;
    const int grid = grid_for(nelem, 256);
    if (v1->dtype == PNED_C128)
        hipLaunchKernelGGL(k_lanczos_restore<double>, dim3(grid), dim3(256), 0, ctx->stream, (double2 *)v1->data,
                           (double2 *)v0->data, swapped ? 1 : 0, sx, a, nelem);
    else
        hipLaunchKernelGGL(k_lanczos_restore<float>, dim3(grid), dim3(256), 0, ctx->stream, (float2 *)v1->data,
                           (float2 *)v0->data, swapped ? 1 : 0, sx, a, nelem);
    PNED_HIP(hipGetLastError());
    return PNED_OK;
}

int intalv_pm_apply(pned_ctx *ctx, const pned_vec *x, pned_vec *v, const double *qv, const double *site_vec, int ipm)
{
    if (ctx->sector_k >= 0 || x->ctx->sector_k >= 0)
        return fail(PNED_E_ARG, "S^+/-(q): both vectors in the full space, or both in sector contexts");
    if (x->ctx->n != ctx->n || x->ctx->q != ctx->q || x->dim != v->dim)
        return fail(PNED_E_ARG, "S^+/-(q): x and v of the same cluster and spin");
    if (x->nvec != 1 || v->nvec != 1 || x->dtype != PNED_C128 || v->dtype != PNED_C128)
        return fail(PNED_E_UNSUPPORTED, "S^+/-(q) takes one complex128 column");
    if (ctx->n > 32 || ctx->q > 16 || (ctx->q > 2 && (uint64_t)ctx->dim > 0xFFFFFFFFull))
        return fail(PNED_E_UNSUPPORTED, "S^+/-(q): at most 32 sites and 2^32 states for spin S > 1/2");
    PmArgs a{};
    a.n = ctx->n;
    a.q = ctx->q;
    a.ipm = ipm;
    a.sw_v = swap_args(ctx->lay);
    a.sw_x = swap_args(x->ctx->lay);
    const double ss = 0.5 * (ctx->q - 1);
    for (int d = 0; d < ctx->q; ++d) {
        const double sm = (double)d - ss;
        a.fac[d] = (ctx->q == 2) ? 1.0 : (ipm > 0 ? std::sqrt((ss - sm) * (ss + sm + 1.0))
                                                 : std::sqrt((ss + sm) * (ss - sm + 1.0)));
    }
    uint64_t p = 1;
    for (int j = 0; j < ctx->n; ++j, p *= (uint64_t)ctx->q) {
        const double qn = qv[0] * site_vec[2 * j] + qv[1] * site_vec[2 * j + 1];
        a.phr[j] = std::cos(qn);
        a.phi[j] = std::sin(qn);
        a.pw[j] = p;
    }
    if (x->ctx != ctx) PNED_HIP(hipStreamSynchronize(x->ctx->stream));   // x written on its own stream
    const int grid = (int)std::min<uint64_t>(((uint64_t)v->dim + 255) / 256, 8192);
    hipLaunchKernelGGL(k_intalv_pm, dim3(grid), dim3(256), 0, ctx->stream, (const double2 *)x->data,
                       (double2 *)v->data, a, (uint64_t)v->dim);
    PNED_HIP(hipGetLastError());
    return PNED_OK;
}

int dot(pned_ctx *ctx, const pned_vec *x, const pned_vec *y, double *out)
{
    PNED_TRY(same_shape(x, y));
    PNED_TRY(check_pow2_cols(x->nvec));
    const size_t nelem = (size_t)x->dim * x->nvec;
    const int grid = RED_GRID;
    PNED_TRY(ensure_partial(ctx, (size_t)grid * 2 * x->nvec));
    if (x->dtype == PNED_C128)
        hipLaunchKernelGGL(k_dot<double>, dim3(grid), dim3(RED_TPB), 0, ctx->stream, (const double2 *)x->data,
                           (const double2 *)y->data, nelem, x->nvec, ctx->d_partial);
    else
        hipLaunchKernelGGL(k_dot<float>, dim3(grid), dim3(RED_TPB), 0, ctx->stream, (const float2 *)x->data,
                           (const float2 *)y->data, nelem, x->nvec, ctx->d_partial);
    PNED_HIP(hipGetLastError());
    return finish_reduce(ctx, grid, 2 * x->nvec, out);
}

int dotu(pned_ctx *ctx, const pned_vec *x, const pned_vec *y, double *out)
{
    PNED_TRY(same_shape(x, y));
    PNED_TRY(check_pow2_cols(x->nvec));
    const size_t nelem = (size_t)x->dim * x->nvec;
    const int grid = RED_GRID;
    PNED_TRY(ensure_partial(ctx, (size_t)grid * 2 * x->nvec));
    if (x->dtype == PNED_C128)
        hipLaunchKernelGGL(k_dotu<double>, dim3(grid), dim3(RED_TPB), 0, ctx->stream, (const double2 *)x->data,
                           (const double2 *)y->data, nelem, x->nvec, ctx->d_partial);
    else
        hipLaunchKernelGGL(k_dotu<float>, dim3(grid), dim3(RED_TPB), 0, ctx->stream, (const float2 *)x->data,
                           (const float2 *)y->data, nelem, x->nvec, ctx->d_partial);
    PNED_HIP(hipGetLastError());
    return finish_reduce(ctx, grid, 2 * x->nvec, out);
}

int beta2(pned_ctx *ctx, const double *alpha, const pned_vec *v1, const pned_vec *v0, double *out)
{
    PNED_TRY(same_shape(v1, v0));
    PNED_TRY(check_pow2_cols(v1->nvec));
    const size_t nelem = (size_t)v1->dim * v1->nvec;
    const int grid = RED_GRID;
    PNED_TRY(ensure_partial(ctx, (size_t)grid * v1->nvec));
    double *da;
    PNED_TRY(upload_scalars(ctx, alpha, 2 * v1->nvec, &da));
    if (v1->dtype == PNED_C128)
        hipLaunchKernelGGL(k_beta2<double>, dim3(grid), dim3(RED_TPB), 0, ctx->stream, (const double2 *)v1->data,
                           (const double2 *)v0->data, (const double *)da, nelem, v1->nvec, ctx->d_partial);
    else
        hipLaunchKernelGGL(k_beta2<float>, dim3(grid), dim3(RED_TPB), 0, ctx->stream, (const float2 *)v1->data,
                           (const float2 *)v0->data, (const double *)da, nelem, v1->nvec, ctx->d_partial);
    PNED_HIP(hipGetLastError());
    return finish_reduce(ctx, grid, v1->nvec, out);
}

int axpby(pned_ctx *ctx, const double *a, const pned_vec *x, const double *b, pned_vec *y)
{
    PNED_TRY(same_shape(x, y));
    const size_t nelem = (size_t)x->dim * x->nvec;
    std::vector<double> ab(4 * x->nvec);
    for (int c = 0; c < 2 * x->nvec; ++c) { ab[c] = a[c]; ab[2 * x->nvec + c] = b[c]; }
    double *d;
    PNED_TRY(upload_scalars(ctx, ab.data(), 4 * x->nvec, &d));
    const int grid = grid_for(nelem, 256);
    if (x->dtype == PNED_C128)
        hipLaunchKernelGGL(k_axpby<double>, dim3(grid), dim3(256), 0, ctx->stream, (const double2 *)x->data,
                           (double2 *)y->data, (const double *)d, (const double *)(d + 2 * x->nvec), nelem, x->nvec);
    else
        hipLaunchKernelGGL(k_axpby<float>, dim3(grid), dim3(256), 0, ctx->stream, (const float2 *)x->data,
                           (float2 *)y->data, (const double *)d, (const double *)(d + 2 * x->nvec), nelem, x->nvec);
    PNED_HIP(hipGetLastError());
    return PNED_OK;
}

int scale(pned_ctx *ctx, pned_vec *x, const double *a)
{
    const size_t nelem = (size_t)x->dim * x->nvec;
    double *d;
    PNED_TRY(upload_scalars(ctx, a, 2 * x->nvec, &d));
    const int grid = grid_for(nelem, 256);
    if (x->dtype == PNED_C128)
        hipLaunchKernelGGL(k_scale<double>, dim3(grid), dim3(256), 0, ctx->stream, (double2 *)x->data,
                           (const double *)d, nelem, x->nvec);
    else
        hipLaunchKernelGGL(k_scale<float>, dim3(grid), dim3(256), 0, ctx->stream, (float2 *)x->data,
                           (const double *)d, nelem, x->nvec);
    PNED_HIP(hipGetLastError());
    return PNED_OK;
}

int lanczos_update(pned_ctx *ctx, pned_vec *v1, pned_vec *v0, const double *alpha, const double *beta)
{
    PNED_TRY(same_shape(v1, v0));
    const size_t nelem = (size_t)v1->dim * v1->nvec;
    std::vector<double> ab(2 * v1->nvec);
    for (int c = 0; c < v1->nvec; ++c) { ab[2 * c] = alpha[c]; ab[2 * c + 1] = beta[c]; }
    double *d;
    PNED_TRY(upload_scalars(ctx, ab.data(), 2 * v1->nvec, &d));
    const int grid = grid_for(nelem, 256);
    if (v1->dtype == PNED_C128)
        hipLaunchKernelGGL(k_lanczos_update<double>, dim3(grid), dim3(256), 0, ctx->stream, (double2 *)v1->data,
                           (double2 *)v0->data, (const double *)d, nelem, v1->nvec);
    else
        hipLaunchKernelGGL(k_lanczos_update<float>, dim3(grid), dim3(256), 0, ctx->stream, (float2 *)v1->data,
                           (float2 *)v0->data, (const double *)d, nelem, v1->nvec);
    PNED_HIP(hipGetLastError());
    return PNED_OK;
}

int lanczos_slots(pned_ctx *ctx, pned_vec *v1, pned_vec *v0, const double *alpha, const double *beta,
                  const int *mode)
{
    PNED_TRY(same_shape(v1, v0));
    const int nc = v1->nvec;
    const size_t nelem = (size_t)v1->dim * nc;
    // scalars: [alpha,beta]*nc then the modes as doubles (converted on upload)
    std::vector<double> buf(2 * nc + nc);
    for (int c = 0; c < nc; ++c) {
        buf[2 * c] = alpha[c];
        buf[2 * c + 1] = beta[c];
        if (mode[c] == 2 && !(beta[c] > 0.0)) return fail(PNED_E_ARG, "stepping a column with beta <= 0");
    }
    std::vector<int> modes(mode, mode + nc);
    std::memcpy(buf.data() + 2 * nc, modes.data(), sizeof(int) * nc);
    double *d;
    PNED_TRY(upload_scalars(ctx, buf.data(), (int)buf.size(), &d));
    const int *dm = reinterpret_cast<const int *>(d + 2 * nc);
    const int grid = grid_for(nelem, 256);
    if (v1->dtype == PNED_C128)
        hipLaunchKernelGGL(k_lanczos_slots<double>, dim3(grid), dim3(256), 0, ctx->stream, (double2 *)v1->data,
                           (double2 *)v0->data, (const double *)d, dm, nelem, nc);
    else
        hipLaunchKernelGGL(k_lanczos_slots<float>, dim3(grid), dim3(256), 0, ctx->stream, (float2 *)v1->data,
                           (float2 *)v0->data, (const double *)d, dm, nelem, nc);
    PNED_HIP(hipGetLastError());
    return PNED_OK;
}

int gaussian_column(pned_ctx *ctx, pned_vec *v, int col, uint64_t seed)
{
    if (vec_detached(v)) return fail(PNED_E_ARG, "vector of a destroyed context");
    if (col < 0 || col >= v->nvec) return fail(PNED_E_ARG, "column outside the vector");
    const int grid = RED_GRID;
    PNED_TRY(ensure_partial(ctx, (size_t)grid));
    if (v->dtype == PNED_C128)
        hipLaunchKernelGGL(k_gauss_col<double>, dim3(grid), dim3(RED_TPB), 0, ctx->stream, (double2 *)v->data, seed,
                           (size_t)v->dim, v->nvec, col, swap_args(v->ctx->lay), ctx->d_partial);
    else
        hipLaunchKernelGGL(k_gauss_col<float>, dim3(grid), dim3(RED_TPB), 0, ctx->stream, (float2 *)v->data, seed,
                           (size_t)v->dim, v->nvec, col, swap_args(v->ctx->lay), ctx->d_partial);
    PNED_HIP(hipGetLastError());
    double nrm;
    PNED_TRY(finish_reduce(ctx, grid, 1, &nrm));
    const double sc = 1.0 / std::sqrt(nrm);
    const int g2 = grid_for((size_t)v->dim, 256);
    if (v->dtype == PNED_C128)
        hipLaunchKernelGGL(k_scale_col<double>, dim3(g2), dim3(256), 0, ctx->stream, (double2 *)v->data, sc,
                           (size_t)v->dim, v->nvec, col);
    else
        hipLaunchKernelGGL(k_scale_col<float>, dim3(g2), dim3(256), 0, ctx->stream, (float2 *)v->data, sc,
                           (size_t)v->dim, v->nvec, col);
    PNED_HIP(hipGetLastError());
    return PNED_OK;
}

int tps_shift_update(pned_ctx *ctx, pned_vec *v1, pned_vec *v0, pned_vec *vg, const double *xn)
{
    PNED_TRY(same_shape(v1, v0));
    PNED_TRY(same_shape(v1, vg));
    const size_t nelem = (size_t)v1->dim * v1->nvec;
    double *d;
    PNED_TRY(upload_scalars(ctx, xn, v1->nvec, &d));
    const int grid = grid_for(nelem, 256);
    if (v1->dtype == PNED_C128)
        hipLaunchKernelGGL(k_tps_update<double>, dim3(grid), dim3(256), 0, ctx->stream, (double2 *)v1->data,
                           (double2 *)v0->data, (double2 *)vg->data, (const double *)d, nelem, v1->nvec);
    else
        hipLaunchKernelGGL(k_tps_update<float>, dim3(grid), dim3(256), 0, ctx->stream, (float2 *)v1->data,
                           (float2 *)v0->data, (float2 *)vg->data, (const double *)d, nelem, v1->nvec);
    PNED_HIP(hipGetLastError());
    return PNED_OK;
}

int gaussian_fill(pned_ctx *ctx, pned_vec *v, uint64_t seed)
{
    PNED_TRY(check_pow2_cols(v->nvec));
    const size_t nelem = (size_t)v->dim * v->nvec;
    const int grid = RED_GRID;
    PNED_TRY(ensure_partial(ctx, (size_t)grid * v->nvec));
    if (v->dtype == PNED_C128)
        hipLaunchKernelGGL(k_gauss<double>, dim3(grid), dim3(RED_TPB), 0, ctx->stream, (double2 *)v->data, seed,
                           nelem, v->nvec, swap_args(v->ctx->lay), ctx->d_partial);
    else
        hipLaunchKernelGGL(k_gauss<float>, dim3(grid), dim3(RED_TPB), 0, ctx->stream, (float2 *)v->data, seed,
                           nelem, v->nvec, swap_args(v->ctx->lay), ctx->d_partial);
    PNED_HIP(hipGetLastError());
    std::vector<double> nrm(v->nvec);
    PNED_TRY(finish_reduce(ctx, grid, v->nvec, nrm.data()));
    std::vector<double> sc(2 * v->nvec);
    for (int c = 0; c < v->nvec; ++c) { sc[2 * c] = 1.0 / std::sqrt(nrm[c]); sc[2 * c + 1] = 0.0; }
    return scale(ctx, v, sc.data());
}

int sz_q_apply(pned_ctx *ctx, const pned_vec *x, pned_vec *v, const double *qv, const double *site_vec)
{
    if (ctx->sector_k >= 0) return fail(PNED_E_UNSUPPORTED, "observables need the full 2^N basis (sector basis active)");
    PNED_TRY(same_shape(x, v));
    if (x->nvec != 1) return fail(PNED_E_UNSUPPORTED, "S^z(q) excitation takes one column");
    if (ctx->n > 32) return fail(PNED_E_UNSUPPORTED, "at most 32 sites");
    PNED_TRY(check_same_layout(x, v));
    SzArgs a{};
    a.n = ctx->n;
    a.q = ctx->q;
    a.wlow = 0.5 * (ctx->q - 1);
    for (int j = 0; j < 32; ++j) a.pos[j] = ctx->lay.pos[j];
    for (int j = 0; j < ctx->n; ++j) {
        const double qn = qv[0] * site_vec[2 * j] + qv[1] * site_vec[2 * j + 1];
        a.phr[j] = std::cos(qn);
        a.phi[j] = std::sin(qn);
    }
    const int grid = grid_for((size_t)x->dim, 256);
    if (x->dtype == PNED_C128)
        hipLaunchKernelGGL(k_sz_q<double>, dim3(grid), dim3(256), 0, ctx->stream, (const double2 *)x->data,
                           (double2 *)v->data, a, (size_t)x->dim);
    else
        hipLaunchKernelGGL(k_sz_q<float>, dim3(grid), dim3(256), 0, ctx->stream, (const float2 *)x->data,
                           (float2 *)v->data, a, (size_t)x->dim);
    PNED_HIP(hipGetLastError());
    return PNED_OK;
}

// The reference computes the observables from the binary bits of the state index in both
// trees (iand(ishft(istate,-j),1)): in the spin-S tree these are bits of the base-q index,
// not its digits (S_arbitrary/Hybrid_TPS9_TPS.f90:175-177, 238-243, 305-308), kept for parity.
static int build_obs_args(pned_ctx *ctx, ObsArgs &a)
{
    if (ctx->q > 2 && (uint64_t)ctx->dim > 0xFFFFFFFFull)
        return fail(PNED_E_UNSUPPORTED, "mTPQ observables: at most 2^32 states");
    a = ObsArgs{};
    a.n = ctx->n;
    a.nzz = 0;
    for (int j = 0; j < 32; ++j) a.pos[j] = ctx->lay.pos[j];
    // w[mode][b][b0] (re, im): the two-bit weights of the structure-factor modes
    double wtab[2][2][2][2];
    const double ss = 0.5 * (ctx->q - 1);
    for (int b = 0; b < 2; ++b)
        for (int b0 = 0; b0 < 2; ++b0) {
            wtab[0][b][b0][0] = ((double)b - ss) * ((double)b0 - ss);
            wtab[0][b][b0][1] = 0.0;
            wtab[1][b][b0][0] = wtab[1][b][b0][1] = 0.0;
        }
    if (ctx->q == 2) {
        wtab[1][0][1][0] = wtab[1][1][0][0] = 0.25;   // my_xor, Hybrid_TPS_TPS.f90:214-215
    } else {
        // my_xor = sx + sy with sx = (su + sd)/2, sy = -i (su - sd)/2, su(0,1) and sd(1,0)
        // the ladder factors of m = -S (Hybrid_TPS9_params.f90:1479-1495)
        const double su01 = std::sqrt((ss + ss) * (ss - ss + 1.0)), sd10 = std::sqrt((ss + 1.0 - ss) * (ss + ss));
        wtab[1][0][1][0] = 0.5 * su01;
        wtab[1][0][1][1] = -0.5 * su01;
        wtab[1][1][0][0] = 0.5 * sd10;
        wtab[1][1][0][1] = 0.5 * sd10;
    }
    for (int b = 0; b < ctx->ibond; ++b) {
        if (ctx->btype[b] != 3) continue;
        // H/TPS.f90:290-316 shifts by the 1-based site number: bit 'ipair'; past the last
        // site it reads the rank id (0 at one process) -> mask 0 -> bit value 0
        // (basis bit p sits at memory bit pos[p])
        const int p1 = ctx->ipair[2 * b], p2 = ctx->ipair[2 * b + 1];
        a.zz_mask1[a.nzz] = p1 < ctx->n ? (1u << ctx->lay.pos[p1]) : 0u;
        a.zz_mask2[a.nzz] = p2 < ctx->n ? (1u << ctx->lay.pos[p2]) : 0u;
        if (++a.nzz >= 64) break;
    }
    // entries stored by the reference: sq(1..nsq), sq(4k+1..4k+4) = (zz,v1),(zz,vg),(pp,v1),(pp,vg)
    const int nsq = ctx->nsq;
    a.nsq = nsq;
    if (nsq > OBS_MAX_SQ) return fail(PNED_E_UNSUPPORTED, "too many structure-factor momenta");
    int nslots = 0;
    for (int k = 0; k < nsq; ++k) {
        const int qi = k / 4, m = k % 4;
        a.sq_mode[k] = m / 2;
        a.sq_left[k] = m % 2;
        a.sq_q[k] = qi;
        if (qi + 1 > nslots) nslots = qi + 1;
    }
    if (nslots > 8) return fail(PNED_E_UNSUPPORTED, "too many momentum slots");
    double p0[8][2] = {};
    for (int s = 0; s < nslots; ++s)
        for (int j = 0; j < ctx->n; ++j) {
            const double qn = ctx->sq_q[2 * s] * ctx->site_vec[2 * j] + ctx->sq_q[2 * s + 1] * ctx->site_vec[2 * j + 1];
            a.phr[s][j] = std::cos(qn);
            a.phi[s][j] = std::sin(qn);
            p0[s][0] += a.phr[s][j];
            p0[s][1] += a.phi[s][j];
        }
    for (int k = 0; k < nsq; ++k) {
        const int md = a.sq_mode[k], q = a.sq_q[k];
        for (int b0 = 0; b0 < 2; ++b0) {
            const double *w0 = wtab[md][0][b0], *w1 = wtab[md][1][b0];
            a.sa[k][b0][0] = w0[0] * p0[q][0] - w0[1] * p0[q][1];
            a.sa[k][b0][1] = w0[0] * p0[q][1] + w0[1] * p0[q][0];
            a.sd[k][b0][0] = w1[0] - w0[0];
            a.sd[k][b0][1] = w1[1] - w0[1];
        }
    }
    return PNED_OK;
}

// out layout per column: [vg.v1 (2)] [v1 Wzz v1 (2)] [vg Wzz v1 (2)] [sq entries (2 each)]
int tps_observables(pned_ctx *ctx, const pned_vec *vg, const pned_vec *v1, double *out)
{
    if (ctx->sector_k >= 0) return fail(PNED_E_UNSUPPORTED, "observables need the full 2^N basis (sector basis active)");
    PNED_TRY(same_shape(vg, v1));
    PNED_TRY(check_pow2_cols(v1->nvec));
    ObsArgs a;
    PNED_TRY(build_obs_args(ctx, a));
    const int nsq = a.nsq;
    constexpr int NOUT = 6 + 2 * OBS_MAX_SQ;   // fixed instance; unused slots stay zero
    const int width = (6 + 2 * nsq);
    const size_t nelem = (size_t)v1->dim * v1->nvec;
    const int grid = RED_GRID / 4;
    PNED_TRY(ensure_partial(ctx, (size_t)grid * NOUT * v1->nvec));
    // arguments too large for the kernel parameter block: copy to device scratch
    double *dargs;
    const int nd = (int)((sizeof(ObsArgs) + sizeof(double) - 1) / sizeof(double));
    std::vector<double> hbuf(nd);
    std::memcpy(hbuf.data(), &a, sizeof(ObsArgs));
    PNED_TRY(upload_scalars(ctx, hbuf.data(), nd, &dargs));
    const size_t lds = sizeof(double) * NOUT * RED_TPB;
    if (v1->dtype == PNED_C128) {
        auto k = k_tps_obs<double, NOUT>;
        PNED_HIP(hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(k, dim3(grid), dim3(RED_TPB), lds, ctx->stream, (const double2 *)vg->data,
                           (const double2 *)v1->data, (const ObsArgs *)dargs, nelem, v1->nvec, ctx->d_partial);
    } else {
        auto k = k_tps_obs<float, NOUT>;
        PNED_HIP(hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(k, dim3(grid), dim3(RED_TPB), lds, ctx->stream, (const float2 *)vg->data,
                           (const float2 *)v1->data, (const ObsArgs *)dargs, nelem, v1->nvec, ctx->d_partial);
    }
    PNED_HIP(hipGetLastError());
    std::vector<double> all((size_t)NOUT * v1->nvec);
    PNED_TRY(finish_reduce(ctx, grid, NOUT * v1->nvec, all.data()));
    for (int c = 0; c < v1->nvec; ++c)
        for (int k = 0; k < width; ++k) out[(size_t)c * width + k] = all[(size_t)c * NOUT + k];
    return PNED_OK;
}

// out per column: the 10 + 2*nsq values of k_tps_fused
// launch shape of the fused mTPQ reduction for u's columns: chunk of cs states, grid,
// dynamic LDS
struct TpsReduceShape {
    int cs, grid;
    size_t lds;
};

static int tps_reduce_shape(const pned_vec *u, int nsq, TpsReduceShape &s)
{
    if (10 + 2 * nsq > TPS_RED_WIDTH)
        return fail(PNED_E_UNSUPPORTED, "fused mTPQ reduction: at most 8 structure-factor entries");
    if (u->nvec > RED_TPB) return fail(PNED_E_UNSUPPORTED, "fused mTPQ reduction: at most 256 columns");
    // chunks of TPS_CS_MAX states where that keeps up to 64 elements per thread: the per-state
    // weights (phase A) then keep every thread busy and a chunk's loads outweigh its two
    // barriers (N=24, 64 columns: 256-state chunks 23.5 / 39.2 ms per fused step, complex64 /
    // complex128, against 26.7 / 42.4 with 64-state chunks)
    s.cs = std::max(1, std::min(TPS_CS_MAX, 64 * RED_TPB / u->nvec));
    const size_t nchunk = ((size_t)u->dim + s.cs - 1) / s.cs;
    s.grid = (int)std::min<size_t>(nchunk, (size_t)RED_GRID);
    s.lds = sizeof(double) * ((size_t)TPS_RED_WIDTH * RED_TPB + (size_t)s.cs * (1 + 2 * nsq));
    return PNED_OK;
}

// k_tps_fused over (up, u, w) into ctx->d_partial, then the final reduction into out_dev
// ([column][TPS_RED_WIDTH]); dargs is a device copy of the context's ObsArgs.  Enqueued
// only: no allocation, no host round trip (the partial buffer must already hold the grid)
static int tps_fused_reduce_launch(pned_ctx *ctx, const pned_vec *up, const pned_vec *u, const pned_vec *w,
                                   const ObsArgs *dargs, int nsq, double *out_dev)
{
    TpsReduceShape s;
    PNED_TRY(tps_reduce_shape(u, nsq, s));
    const size_t nstates = (size_t)u->dim;
    if (u->dtype == PNED_C128) {
        auto k = k_tps_fused<double, TPS_RED_WIDTH>;
        PNED_HIP(hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)s.lds));
        hipLaunchKernelGGL(k, dim3(s.grid), dim3(RED_TPB), s.lds, ctx->stream, (const double2 *)up->data,
                           (const double2 *)u->data, (const double2 *)w->data, dargs, nstates, u->nvec, s.cs,
                           ctx->d_partial);
    } else {
        auto k = k_tps_fused<float, TPS_RED_WIDTH>;
        PNED_HIP(hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)s.lds));
        hipLaunchKernelGGL(k, dim3(s.grid), dim3(RED_TPB), s.lds, ctx->stream, (const float2 *)up->data,
                           (const float2 *)u->data, (const float2 *)w->data, dargs, nstates, u->nvec, s.cs,
                           ctx->d_partial);
    }
    PNED_HIP(hipGetLastError());
    hipLaunchKernelGGL(reduce_rows, dim3(TPS_RED_WIDTH * u->nvec), dim3(256), 0, ctx->stream, ctx->d_partial, s.grid,
                       TPS_RED_WIDTH * u->nvec, out_dev);
    PNED_HIP(hipGetLastError());
    return PNED_OK;
}

int tps_fused_reduce(pned_ctx *ctx, const pned_vec *up, const pned_vec *u, const pned_vec *w, double *out)
{
    if (ctx->sector_k >= 0) return fail(PNED_E_UNSUPPORTED, "observables need the full 2^N basis (sector basis active)");
    PNED_TRY(same_shape(up, u));
    PNED_TRY(same_shape(u, w));
    PNED_TRY(check_pow2_cols(u->nvec));
    ObsArgs a;
    PNED_TRY(build_obs_args(ctx, a));
    const int nsq = a.nsq, width = 10 + 2 * nsq;
    TpsReduceShape s;
    PNED_TRY(tps_reduce_shape(u, nsq, s));
    PNED_TRY(ensure_partial(ctx, (size_t)s.grid * TPS_RED_WIDTH * u->nvec));
    PNED_TRY(ensure_result(ctx, (size_t)TPS_RED_WIDTH * u->nvec));
    double *dargs;
    const int nd = (int)((sizeof(ObsArgs) + sizeof(double) - 1) / sizeof(double));
    std::vector<double> hbuf(nd);
    std::memcpy(hbuf.data(), &a, sizeof(ObsArgs));
    PNED_TRY(upload_scalars(ctx, hbuf.data(), nd, &dargs));
    PNED_TRY(tps_fused_reduce_launch(ctx, up, u, w, (const ObsArgs *)dargs, nsq, ctx->d_result));
    const size_t all = (size_t)TPS_RED_WIDTH * u->nvec;
    PNED_HIP(hipMemcpyAsync(ctx->h_result, ctx->d_result, sizeof(double) * all, hipMemcpyDeviceToHost, ctx->stream));
    PNED_HIP(hipStreamSynchronize(ctx->stream));
    for (int c = 0; c < u->nvec; ++c)
        for (int k = 0; k < width; ++k) out[(size_t)c * width + k] = ctx->h_result[(size_t)c * TPS_RED_WIDTH + k];
    return PNED_OK;
}

// su[c] = 1 / sqrt(red[c][0]) (1/||v0|| of the previous step), or 1 without a previous step
__global__ void k_tps_scale(const double *__restrict__ red, int nc, double *__restrict__ su)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < nc) su[c] = red ? 1.0 / sqrt(red[(size_t)c * TPS_RED_WIDTH]) : 1.0;
}

int tps_graph_setup(pned_ctx *ctx, const pned_vec *u, void **dobs, int *nsq)
{
    if (ctx->sector_k >= 0) return fail(PNED_E_UNSUPPORTED, "observables need the full 2^N basis (sector basis active)");
    PNED_TRY(check_pow2_cols(u->nvec));
    ObsArgs a;
    PNED_TRY(build_obs_args(ctx, a));
    TpsReduceShape s;
    PNED_TRY(tps_reduce_shape(u, a.nsq, s));
    PNED_TRY(ensure_partial(ctx, (size_t)s.grid * TPS_RED_WIDTH * u->nvec));
    PNED_HIP(hipMalloc(dobs, sizeof(ObsArgs)));
    const hipError_t e = hipMemcpy(*dobs, &a, sizeof(ObsArgs), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(*dobs);
        *dobs = nullptr;
        return fail(PNED_E_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
    }
    *nsq = a.nsq;
    return PNED_OK;
}

int tps_graph_step(pned_ctx *ctx, const pned_vec *up, const pned_vec *u, pned_vec *w, const void *dobs, int nsq,
                   const double *red_prev, double *su_dev, double *slot)
{
    PNED_TRY(check_hv_pair(ctx, u, w));
    PNED_TRY(check_hv_pair(ctx, up, u));
    const int nc = u->nvec;
    hipLaunchKernelGGL(k_tps_scale, dim3((nc + 63) / 64), dim3(64), 0, ctx->stream, red_prev, nc, su_dev);
    PNED_HIP(hipGetLastError());
    PNED_TRY(launch_hv_s12(ctx, u, w, 0.0, true, nullptr, su_dev));   // w = (L-H)(su u)
    hipLaunchKernelGGL(k_copy_f64, dim3((2 * nc + 63) / 64), dim3(64), 0, ctx->stream, ctx->d_result, slot, 2 * nc);
    PNED_HIP(hipGetLastError());
    return tps_fused_reduce_launch(ctx, up, u, w, (const ObsArgs *)dobs, nsq, slot + 2 * nc);
}

int renbun_gpu(pned_ctx *ctx, int i_range, int s_maxitr, double split, double ene, double eta,
               const double *salpha, const double *sbeta, double *w, double *dsf, int *conv)
{
    const size_t na = (size_t)s_maxitr, nb = (size_t)s_maxitr + 1;
    const size_t need = na + nb + 2 * (size_t)i_range + (size_t)i_range;
    double *dbuf = nullptr;
    PNED_HIP(hipMallocAsync((void **)&dbuf, sizeof(double) * need, ctx->stream));
    double *da = dbuf, *db = dbuf + na, *dw = db + nb, *dd = dw + i_range;
    int *dc = reinterpret_cast<int *>(dd + i_range);
    PNED_HIP(hipMemcpyAsync(da, salpha, sizeof(double) * na, hipMemcpyHostToDevice, ctx->stream));
    PNED_HIP(hipMemcpyAsync(db, sbeta, sizeof(double) * nb, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_renbun, dim3((i_range + 63) / 64), dim3(64), 0, ctx->stream, i_range, s_maxitr, split,
                       ene, eta, (const double *)da, (const double *)db, dw, dd, dc);
    PNED_HIP(hipGetLastError());
    PNED_HIP(hipMemcpyAsync(w, dw, sizeof(double) * i_range, hipMemcpyDeviceToHost, ctx->stream));
    PNED_HIP(hipMemcpyAsync(dsf, dd, sizeof(double) * i_range, hipMemcpyDeviceToHost, ctx->stream));
    PNED_HIP(hipMemcpyAsync(conv, dc, sizeof(int) * i_range, hipMemcpyDeviceToHost, ctx->stream));
    PNED_HIP(hipStreamSynchronize(ctx->stream));
    PNED_HIP(hipFreeAsync(dbuf, ctx->stream));
    PNED_HIP(hipStreamSynchronize(ctx->stream));
    return PNED_OK;
}

}  // namespace pned
