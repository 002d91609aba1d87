// hv_sector.hip -- S=1/2 H.v in a fixed-S^z sector basis (the reference's parity_type = 0).
//
// Basis (mk_list12, S_one_half/Hybrid_TPS_parallel.f90:452-578, at one process): the
// states with k set bits in increasing order.  A state's index comes from its two halves
// (right half: the low ihf = (N+1)/2 bits, left half: the rest):
//     index(s) = rank(s & right) + offset(s >> ihf)
// where rank is the position of the right half among right halves of the same popcount
// and offset counts the sector states of every smaller left half -- the reference's
// list3(1,.) - 1 and list3(2,.).  list1 (index -> state) is kept in device memory.
//
// H.v reuses the S=1/2 term decomposition: the diagonal groups act on the state, and a
// two-bit flip term keeps the state in the sector exactly when the two bits differ; the
// context checks at set-up that every other term vanishes (the operator conserves S^z).
// One thread per output index gathers its partners through the two half tables.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "internal.h"

namespace pned {

struct SectorTerm {
    uint32_t mask;       // the two flipped bits
    uint32_t selmask;
    double c0r, c0i, c1r, c1i;
};

__device__ __forceinline__ double sector_diag(const DiagDev &d, uint32_t s, double &im)
{
    double re = d.c0r;
    im = d.c0i;
    for (int g = 0; g < d.ngroups; ++g) {
        const uint32_t f = (d.kind[g] == 0) ? (s & d.mask[g]) : ((s ^ (s >> d.shift[g])) & d.mask[g]);
        const double c = (double)__popc(f);
        re = fma(c, d.wr[g], re);
        im = fma(c, d.wi[g], im);
    }
    return re;
}

__global__ __launch_bounds__(256) void hv_sector_kernel(const double2 *__restrict__ v1, double2 *__restrict__ v0,
                                                        const uint32_t *__restrict__ list1,
                                                        const uint32_t *__restrict__ rank,
                                                        const uint32_t *__restrict__ offset,
                                                        const SectorTerm *__restrict__ terms, int nterm,
                                                        const DiagDev *__restrict__ diagp, int ihf, double shift,
                                                        int fresh, uint64_t dim, double *__restrict__ partial)
{
    __shared__ double red[512];
    const uint32_t right = (1u << ihf) - 1u;
    const DiagDev &dg = *diagp;
    double pr_re = 0.0, pr_im = 0.0;
    const uint64_t stride = (uint64_t)blockDim.x * gridDim.x;
    for (uint64_t l = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; l < dim; l += stride) {
        const uint32_t s = list1[l];
        const double2 x = v1[l];
        double di;
        const double dr = sector_diag(dg, s, di) + shift;
        double ar = dr * x.x - di * x.y, ai = dr * x.y + di * x.x;
        for (int t = 0; t < nterm; ++t) {
            const SectorTerm T = terms[t];
            if (__popc(s & T.mask) != 1) continue;   // partner outside the sector: coefficient 0
            const uint32_t p = s ^ T.mask;
            const uint64_t idx = (uint64_t)rank[p & right] + (uint64_t)offset[p >> ihf];
            const double2 y = v1[idx];
            const bool sel = __popc(s & T.selmask) & 1;
            const double cr = sel ? T.c1r : T.c0r, ci = sel ? T.c1i : T.c0i;
            ar = fma(cr, y.x, fma(-ci, y.y, ar));
            ai = fma(cr, y.y, fma(ci, y.x, ai));
        }
        pr_re += x.x * ar + x.y * ai;
        pr_im += x.x * ai - x.y * ar;
        double2 o;
        if (fresh) {
            o = make_double2(ar, ai);
        } else {
            const double2 old = v0[l];
            o = make_double2(old.x + ar, old.y + ai);
        }
        v0[l] = o;
    }
    const int tid = threadIdx.x;
    red[2 * tid] = pr_re;
    red[2 * tid + 1] = pr_im;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (tid < st) {
            red[2 * tid] += red[2 * (tid + st)];
            red[2 * tid + 1] += red[2 * (tid + st) + 1];
        }
        __syncthreads();
    }
    if (tid == 0) {
        partial[2 * blockIdx.x] = red[0];
        partial[2 * blockIdx.x + 1] = red[1];
    }
}

// intalv_xy (S_one_half/Hybrid_TPS_DSF.f90:89-193) as a gather: one thread per target
// index l (sector k + ipm) sums e^{i q.r_j} x(source index of t with bit j flipped) over
// the bits j of its state t that ipm has just set (ipm = +1) or cleared (-1), in ascending
// j -- per element, the order in which the reference's site-outer scatter adds them.
struct XyArgs {
    int n, ipm;
    double phr[32], phi[32];
};

__global__ __launch_bounds__(256) void k_intalv_xy(const double2 *__restrict__ x, double2 *__restrict__ v,
                                                   const uint32_t *__restrict__ list1_dst,
                                                   const uint32_t *__restrict__ rank,
                                                   const uint32_t *__restrict__ offset_src, XyArgs a, int ihf,
                                                   uint64_t dim)
{
    const uint32_t right = (1u << ihf) - 1u;
    const uint64_t stride = (uint64_t)blockDim.x * gridDim.x;
    for (uint64_t l = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; l < dim; l += stride) {
        const uint32_t t = list1_dst[l];
        double ar = 0.0, ai = 0.0;
        for (int j = 0; j < a.n; ++j) {
            if ((int)((t >> j) & 1u) != (a.ipm > 0 ? 1 : 0)) continue;
            const uint32_t s = t ^ (1u << j);
            const double2 y = x[(uint64_t)rank[s & right] + (uint64_t)offset_src[s >> ihf]];
            ar += a.phr[j] * y.x - a.phi[j] * y.y;
            ai += a.phi[j] * y.x + a.phr[j] * y.y;
        }
        v[l] = make_double2(ar, ai);
    }
}

int intalv_xy_apply(pned_ctx *ctx, const pned_vec *x, pned_vec *v, const double *qv, const double *site_vec,
                    int ipm)
{
    const pned_ctx *src = x->ctx;
    if (v->ctx != ctx) return fail(PNED_E_ARG, "v must belong to the target context");
    if (ctx->q != 2 || !src || src->q != 2 || src->n != ctx->n)
        return fail(PNED_E_ARG, "S^+(q): two S=1/2 contexts of the same cluster");
    if (ctx->sector_k < 0 || src->sector_k < 0 || ctx->sector_k != src->sector_k + ipm)
        return fail(PNED_E_ARG, "S^+/-(q): x in the sector with k set bits, v in the one with k + ipm");
    if (src->device != ctx->device) return fail(PNED_E_ARG, "S^+(q): both contexts on one device");
    if (x->nvec != 1 || v->nvec != 1 || x->dtype != PNED_C128 || v->dtype != PNED_C128)
        return fail(PNED_E_UNSUPPORTED, "S^+(q) takes one complex128 column");
    XyArgs a{};
    a.n = ctx->n;
    a.ipm = ipm;
    for (int j = 0; j < ctx->n; ++j) {
        const double qn = qv[0] * site_vec[2 * j] + qv[1] * site_vec[2 * j + 1];
        a.phr[j] = std::cos(qn);
        a.phi[j] = std::sin(qn);
    }
    // x was written on the source context's stream
    PNED_HIP(hipStreamSynchronize(src->stream));
    const uint64_t dim = (uint64_t)ctx->dim;
    const int grid = (int)std::min<uint64_t>((dim + 255) / 256, 8192);
    hipLaunchKernelGGL(k_intalv_xy, dim3(grid), dim3(256), 0, ctx->stream, (const double2 *)x->data,
                       (double2 *)v->data, ctx->d_list1, src->d_srank, src->d_soffset, a, (ctx->n + 1) / 2, dim);
    PNED_HIP(hipGetLastError());
    return PNED_OK;
}


void free_sector(pned_ctx *ctx)
{
    if (ctx->d_list1) (void)hipFree(ctx->d_list1);
    if (ctx->d_srank) (void)hipFree(ctx->d_srank);
    if (ctx->d_soffset) (void)hipFree(ctx->d_soffset);
    if (ctx->d_sterms) (void)hipFree(ctx->d_sterms);
    ctx->d_list1 = ctx->d_srank = ctx->d_soffset = nullptr;
    ctx->d_sterms = nullptr;
    ctx->sector_k = -1;
    ctx->sector_nterm = 0;
}

// the two-bit flip terms of the decomposition, or an error if the operator moves states
// out of the sector (a one-bit flip, or a two-bit flip of aligned spins)
static int sector_terms(const S12Plan &pl, std::vector<SectorTerm> &out)
{
    out.clear();
    for (const Term &t : pl.terms) {
        const bool zero0 = t.c0r == 0.0 && t.c0i == 0.0, zero1 = t.c1r == 0.0 && t.c1i == 0.0;
        if (zero0 && zero1) continue;
        const int nb = __builtin_popcount(t.gmask);
        // aligned spins (the only states a two-bit flip takes out of the sector) select
        // c0 when the selector is exactly the two flipped bits
        if (nb != 2 || t.selmask != t.gmask || !zero0)
            return fail(PNED_E_UNSUPPORTED, "the operator does not conserve S^z (no sector basis)");
        out.push_back(SectorTerm{t.gmask, t.selmask, t.c0r, t.c0i, t.c1r, t.c1i});
    }
    return PNED_OK;
}

int refresh_sector_terms(pned_ctx *ctx)
{
    std::vector<SectorTerm> st;
    PNED_TRY(sector_terms(ctx->s12, st));
    PNED_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->d_sterms) PNED_HIP(hipFree(ctx->d_sterms));
    ctx->d_sterms = nullptr;
    PNED_HIP(hipMalloc(&ctx->d_sterms, sizeof(SectorTerm) * std::max<size_t>(st.size(), 1)));
    if (!st.empty())
        PNED_HIP(hipMemcpy(ctx->d_sterms, st.data(), sizeof(SectorTerm) * st.size(), hipMemcpyHostToDevice));
    ctx->sector_nterm = (int)st.size();
    return PNED_OK;
}

int launch_hv_sector(pned_ctx *ctx, const pned_vec *v1, pned_vec *v0, double eperbd, bool fresh, double *prdct_host)
{
    if (v1->nvec != 1 || v1->dtype != PNED_C128)
        return fail(PNED_E_UNSUPPORTED, "sector H.v takes one complex128 column");
    // the diagonal table is shared with the full-space path
    PNED_TRY(ensure_s12_plan(ctx, 0, PNED_C128));
    const uint64_t dim = (uint64_t)ctx->dim;
    int grid = (int)std::min<uint64_t>((dim + 255) / 256, 4096);
    PNED_TRY(ensure_partial(ctx, 2 * (size_t)grid));
    PNED_TRY(ensure_result(ctx, 8));
    hipLaunchKernelGGL(hv_sector_kernel, dim3(grid), dim3(256), 0, ctx->stream, (const double2 *)v1->data,
                       (double2 *)v0->data, ctx->d_list1, ctx->d_srank, ctx->d_soffset,
                       (const SectorTerm *)ctx->d_sterms, ctx->sector_nterm, ctx->d_diag, (ctx->n + 1) / 2,
                       -(double)ctx->ibond * eperbd, fresh ? 1 : 0, dim, ctx->d_partial);
    PNED_HIP(hipGetLastError());
    PNED_TRY(reduce_rows_result(ctx, grid, 2));
    if (prdct_host) {
        PNED_HIP(hipMemcpyAsync(ctx->h_result, ctx->d_result, 2 * sizeof(double), hipMemcpyDeviceToHost,
                                ctx->stream));
        PNED_HIP(hipStreamSynchronize(ctx->stream));
        prdct_host[0] = ctx->h_result[0];
        prdct_host[1] = ctx->h_result[1];
    }
    return PNED_OK;
}

}  // namespace pned

using namespace pned;

extern "C" int pned_ctx_sz_sector(pned_ctx *ctx, int n_down)
{
    if (!ctx) return fail(PNED_E_ARG, "null context");
    if (ctx->q != 2) return fail(PNED_E_UNSUPPORTED, "sector basis: S=1/2 only");
    PNED_HIP(hipSetDevice(ctx->device));
    PNED_HIP(hipStreamSynchronize(ctx->stream));
    const int n = ctx->n;
    if (n_down < 0) {   // back to the full 2^N space
        free_sector(ctx);
        ctx->dim = (int64_t)1 << n;
        return PNED_OK;
    }
    if (n_down > n) return fail(PNED_E_ARG, "#(E02A)# Incorrect ldim or n given to sz");
    if (n > 32) return fail(PNED_E_UNSUPPORTED, "sector basis: N <= 32");
    PNED_TRY(reset_layout(ctx));   // the sector tables index basis states directly
    std::vector<SectorTerm> st;
    PNED_TRY(sector_terms(ctx->s12, st));
    const int ihf = (n + 1) / 2, ihf0 = n / 2;
    // binomials and the two half tables
    std::vector<std::vector<uint64_t>> C(33, std::vector<uint64_t>(33, 0));
    for (int a = 0; a <= 32; ++a) {
        C[a][0] = 1;
        for (int b = 1; b <= a; ++b) C[a][b] = C[a - 1][b - 1] + (b <= a - 1 ? C[a - 1][b] : 0);
    }
    std::vector<uint32_t> rank((size_t)1 << ihf), offset((size_t)1 << ihf0);
    {
        std::vector<uint32_t> seen(ihf + 1, 0);
        for (uint64_t r = 0; r < ((uint64_t)1 << ihf); ++r) rank[r] = seen[__builtin_popcountll(r)]++;
    }
    uint64_t total = 0;
    for (uint64_t l = 0; l < ((uint64_t)1 << ihf0); ++l) {
        offset[l] = (uint32_t)total;
        const int need = n_down - __builtin_popcountll(l);
        if (need >= 0 && need <= ihf) total += C[ihf][need];
    }
    if (total < 3) return fail(PNED_E_ARG, "#(E02A)# Incorrect ldim or n given to sz");
    if (total > 0xFFFFFFFFull) return fail(PNED_E_UNSUPPORTED, "sector larger than 2^32 states");
    std::vector<uint32_t> list1(total);
    {
        uint64_t i = 0;
        for (uint64_t l = 0; l < ((uint64_t)1 << ihf0); ++l)
            for (uint64_t r = 0; r < ((uint64_t)1 << ihf); ++r)
                if (__builtin_popcountll(l) + __builtin_popcountll(r) == n_down)
                    list1[i++] = (uint32_t)((l << ihf) | r);
    }
    free_sector(ctx);
    PNED_HIP(hipMalloc((void **)&ctx->d_list1, sizeof(uint32_t) * list1.size()));
    PNED_HIP(hipMalloc((void **)&ctx->d_srank, sizeof(uint32_t) * rank.size()));
    PNED_HIP(hipMalloc((void **)&ctx->d_soffset, sizeof(uint32_t) * offset.size()));
    PNED_HIP(hipMalloc(&ctx->d_sterms, sizeof(SectorTerm) * std::max<size_t>(st.size(), 1)));
    PNED_HIP(hipMemcpy(ctx->d_list1, list1.data(), sizeof(uint32_t) * list1.size(), hipMemcpyHostToDevice));
    PNED_HIP(hipMemcpy(ctx->d_srank, rank.data(), sizeof(uint32_t) * rank.size(), hipMemcpyHostToDevice));
    PNED_HIP(hipMemcpy(ctx->d_soffset, offset.data(), sizeof(uint32_t) * offset.size(), hipMemcpyHostToDevice));
    if (!st.empty())
        PNED_HIP(hipMemcpy(ctx->d_sterms, st.data(), sizeof(SectorTerm) * st.size(), hipMemcpyHostToDevice));
    ctx->sector_k = n_down;
    ctx->sector_nterm = (int)st.size();
    ctx->dim = (int64_t)total;
    return PNED_OK;
}
