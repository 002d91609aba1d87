// hv_spin.hip -- spin-S (q = 2S+1 digits) matrix-free H.v for gfx950.
//
// p_mltply_v3 / H_dot_v of the spin-S tree (S_arbitrary/Hybrid_TPS9_parallel.f90:184-356):
//     v0(l) += sum_bonds sum_in xlocal_Ham(in, istat(l), t) * v1(l with the bond's digits = in)
// with istat = q*digit(max site) + digit(min site) and the diagonal shifted by -eperbd per
// bond.  One thread owns one output state and sums all bonds in registers (a gather: no
// write conflicts, no atomics); v0 is read and written once.  Per (type, output local
// state) the nonzero input states are precomputed on the host, so the inner loop visits
// only the elements the reference does not skip (llocal_Ham).
#include "internal.h"

namespace pned {

struct SpinArgs {
    int n, q, ibond, qq;
    int colbits;          // unused (one column), kept for symmetry
    double shift;         // -eperbd, added on each bond's diagonal
    int fresh;
    uint64_t pw[32];      // q^k
};

// device tables: nz[t][istat] = count, list[t][istat][k] = in-state, coefficients
struct SpinTables {
    int32_t *pairs;       // 2*ibond
    int32_t *btype;       // ibond (0-based)
    int32_t *nzcnt;       // [9][qq]
    int32_t *nzin;        // [9][qq][qq]
    double *nzc;          // [9][qq][qq][2]
    int32_t *diag_kept;   // [9][qq]: diagonal element present (shift applies)
};

template <int Q>
__global__ __launch_bounds__(256) void hv_spin_kernel(const double2 *__restrict__ v1, double2 *__restrict__ v0,
                                                      SpinTables tb, SpinArgs a, size_t dim,
                                                      double *__restrict__ partial)
{
    __shared__ double red[512];
    const int q = (Q > 0) ? Q : a.q;
    const int qq = q * q;
    double pr_re = 0.0, pr_im = 0.0;
    const size_t stride = (size_t)blockDim.x * gridDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += stride) {
        // digits of the output state, 4 bits each
        uint64_t dlo = 0, dhi = 0;
        {
            uint64_t s = i;
            for (int k = 0; k < a.n; ++k) {
                const uint64_t d = s % (uint64_t)q;
                s /= (uint64_t)q;
                if (k < 16) dlo |= d << (4 * k);
                else dhi |= d << (4 * (k - 16));
            }
        }
        double ar = 0.0, ai = 0.0;
        for (int b = 0; b < a.ibond; ++b) {
            const int s1 = tb.pairs[2 * b], s2 = tb.pairs[2 * b + 1];
            const int t = tb.btype[b];
            const int d1 = (int)(((s1 < 16) ? (dlo >> (4 * s1)) : (dhi >> (4 * (s1 - 16)))) & 15u);
            const int d2 = (int)(((s2 < 16) ? (dlo >> (4 * s2)) : (dhi >> (4 * (s2 - 16)))) & 15u);
            const int istat = q * d2 + d1;
            const int base = t * qq + istat;
            const int cnt = tb.nzcnt[base];
            const uint64_t p1 = a.pw[s1], p2 = a.pw[s2];
            for (int k = 0; k < cnt; ++k) {
                const int ml = tb.nzin[base * qq + k];
                double cr = tb.nzc[2 * (base * qq + k)], ci = tb.nzc[2 * (base * qq + k) + 1];
                if (ml == istat) cr += a.shift;
                const int l1 = ml % q, l2 = ml / q;
                const int64_t off = (int64_t)(l1 - d1) * (int64_t)p1 + (int64_t)(l2 - d2) * (int64_t)p2;
                const double2 p = v1[(size_t)((int64_t)i + off)];
                ar = fma(cr, p.x, fma(-ci, p.y, ar));
                ai = fma(cr, p.y, fma(ci, p.x, ai));
            }
        }
        const double2 x = v1[i];
        pr_re += x.x * ar + x.y * ai;
        pr_im += x.x * ai - x.y * ar;
        double2 o;
        if (a.fresh) {
            o = make_double2(ar, ai);
        } else {
            const double2 old = v0[i];
            o = make_double2(old.x + ar, old.y + ai);
        }
        v0[i] = o;
    }
    const int tid = threadIdx.x;
    red[2 * tid] = pr_re;
    red[2 * tid + 1] = pr_im;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) {
            red[2 * tid] += red[2 * (tid + s)];
            red[2 * tid + 1] += red[2 * (tid + s) + 1];
        }
        __syncthreads();
    }
    if (tid == 0) {
        partial[2 * blockIdx.x] = red[0];
        partial[2 * blockIdx.x + 1] = red[1];
    }
}

__global__ void reduce_two(const double *__restrict__ partial, int nrows, double *__restrict__ out)
{
    if (threadIdx.x >= 2) return;
    double s = 0.0;
    for (int b = 0; b < nrows; ++b) s += partial[2 * b + threadIdx.x];
    out[threadIdx.x] = s;
}

// Builds (once per matrix set) the nonzero tables.  Stored in ctx->spin.
static int build_spin_tables(pned_ctx *ctx, SpinTables &out)
{
    const int q = ctx->q, qq = q * q;
    if (!ctx->spin.pairs) {
        std::vector<int32_t> pairs(2 * ctx->ibond), bt(ctx->ibond);
        for (int b = 0; b < ctx->ibond; ++b) {
            pairs[2 * b] = ctx->site_pair[2 * b];
            pairs[2 * b + 1] = ctx->site_pair[2 * b + 1];
            bt[b] = ctx->btype[b] - 1;
        }
        PNED_HIP(hipMalloc((void **)&ctx->spin.pairs, sizeof(int32_t) * pairs.size()));
        PNED_HIP(hipMalloc((void **)&ctx->spin.btype, sizeof(int32_t) * bt.size()));
        PNED_HIP(hipMemcpy(ctx->spin.pairs, pairs.data(), sizeof(int32_t) * pairs.size(), hipMemcpyHostToDevice));
        PNED_HIP(hipMemcpy(ctx->spin.btype, bt.data(), sizeof(int32_t) * bt.size(), hipMemcpyHostToDevice));
    }
    if (!ctx->spin.xl) {
        // [9][qq] counts, [9][qq][qq] in-states, [9][qq][qq] complex coefficients, [9][qq] diag flag
        const size_t nc = 9 * (size_t)qq, nl = 9 * (size_t)qq * qq;
        std::vector<int32_t> cnt(nc, 0), lst(nl, 0), dk(nc, 0);
        std::vector<double> cf(2 * nl, 0.0);
        for (int t = 0; t < 9; ++t)
            for (int out = 0; out < qq; ++out) {
                int c = 0;
                for (int in = 0; in < qq; ++in) {
                    // Fortran order: xlocal_Ham(in, out, t) at in + qq*out + qq*qq*t
                    const size_t f = (size_t)in + (size_t)qq * out + (size_t)qq * qq * t;
                    if (ctx->ll[f]) continue;
                    const size_t slot = ((size_t)t * qq + out) * qq + c;
                    lst[slot] = in;
                    cf[2 * slot] = ctx->xl[2 * f];
                    cf[2 * slot + 1] = ctx->xl[2 * f + 1];
                    if (in == out) dk[(size_t)t * qq + out] = 1;
                    ++c;
                }
                cnt[(size_t)t * qq + out] = c;
            }
        // one allocation: ints then doubles (8-byte aligned)
        const size_t ints = nc + nl + nc;
        const size_t ib = ((ints * sizeof(int32_t)) + 15) & ~(size_t)15;
        char *buf;
        PNED_HIP(hipMalloc((void **)&buf, ib + sizeof(double) * cf.size()));
        std::vector<int32_t> allints;
        allints.reserve(ints);
        allints.insert(allints.end(), cnt.begin(), cnt.end());
        allints.insert(allints.end(), lst.begin(), lst.end());
        allints.insert(allints.end(), dk.begin(), dk.end());
        PNED_HIP(hipMemcpy(buf, allints.data(), sizeof(int32_t) * ints, hipMemcpyHostToDevice));
        PNED_HIP(hipMemcpy(buf + ib, cf.data(), sizeof(double) * cf.size(), hipMemcpyHostToDevice));
        ctx->spin.mk = reinterpret_cast<int32_t *>(buf);
        ctx->spin.xl = reinterpret_cast<double *>(buf + ib);
    }
    out.pairs = ctx->spin.pairs;
    out.btype = ctx->spin.btype;
    out.nzcnt = ctx->spin.mk;
    out.nzin = ctx->spin.mk + 9 * qq;
    out.diag_kept = ctx->spin.mk + 9 * qq + 9 * qq * qq;
    out.nzc = ctx->spin.xl;
    return PNED_OK;
}

int launch_hv_spin(pned_ctx *ctx, const pned_vec *v1, pned_vec *v0, double eperbd, bool fresh,
                   double *prdct_host)
{
    if (v1->nvec != 1 || v1->dtype != PNED_C128)
        return fail(PNED_E_UNSUPPORTED, "spin-S H.v takes one complex128 column");
    if (ctx->q > 15 || ctx->n > 32) return fail(PNED_E_UNSUPPORTED, "spin-S H.v: q <= 15, N <= 32");
    SpinTables tb;
    PNED_TRY(build_spin_tables(ctx, tb));
    SpinArgs a{};
    a.n = ctx->n;
    a.q = ctx->q;
    a.ibond = ctx->ibond;
    a.qq = ctx->q * ctx->q;
    a.shift = -eperbd;
    a.fresh = fresh ? 1 : 0;
    uint64_t p = 1;
    for (int k = 0; k < ctx->n; ++k) { a.pw[k] = p; p *= (uint64_t)ctx->q; }
    const size_t dim = (size_t)ctx->dim;
    int grid = (int)((dim + 255) / 256);
    if (grid > 4096) grid = 4096;
    PNED_TRY(ensure_partial(ctx, 2 * (size_t)grid));
    PNED_TRY(ensure_result(ctx, 8));
    if (ctx->q == 3)
        hipLaunchKernelGGL(hv_spin_kernel<3>, dim3(grid), dim3(256), 0, ctx->stream, (const double2 *)v1->data,
                           (double2 *)v0->data, tb, a, dim, ctx->d_partial);
    else if (ctx->q == 2)
        hipLaunchKernelGGL(hv_spin_kernel<2>, dim3(grid), dim3(256), 0, ctx->stream, (const double2 *)v1->data,
                           (double2 *)v0->data, tb, a, dim, ctx->d_partial);
    else
        hipLaunchKernelGGL(hv_spin_kernel<0>, dim3(grid), dim3(256), 0, ctx->stream, (const double2 *)v1->data,
                           (double2 *)v0->data, tb, a, dim, ctx->d_partial);
    PNED_HIP(hipGetLastError());
    hipLaunchKernelGGL(reduce_two, dim3(1), dim3(64), 0, ctx->stream, ctx->d_partial, grid, ctx->d_result);
    PNED_HIP(hipGetLastError());
    if (prdct_host) {
        PNED_HIP(hipMemcpyAsync(ctx->h_result, ctx->d_result, 2 * sizeof(double), hipMemcpyDeviceToHost,
                                ctx->stream));
        PNED_HIP(hipStreamSynchronize(ctx->stream));
        prdct_host[0] = ctx->h_result[0];
        prdct_host[1] = ctx->h_result[1];
    }
    return PNED_OK;
}

// drop cached device tables (matrices changed)
void reset_spin_tables(pned_ctx *ctx)
{
    if (ctx->spin.mk) (void)hipFree(ctx->spin.mk);
    ctx->spin.mk = nullptr;
    ctx->spin.xl = nullptr;
}

}  // namespace pned
