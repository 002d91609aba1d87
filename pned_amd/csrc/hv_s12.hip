// hv_s12.hip -- S=1/2 matrix-free H.v for gfx950 (CDNA4).
//
// Computes the reference's p_mltply_v3 (S_one_half/Hybrid_TPS_parallel.f90:190-263):
//     v0 += (H - ibond*eperbd) v1,   prdct = <v1|(H - ibond*eperbd)|v1>
// with H given by the bond matrices of H_dot_v (same file, 305-442).  Instead of the
// reference's 4*ibond full passes over the vector, the host planner (context.cpp)
// rewrites H as a diagonal D(i) plus flip terms  c_t(sel(i)) * v1(i ^ mask_t)  and groups
// the terms into a few "tile passes".  A pass picks K memory-index bits (the tile bits);
// every term of the pass flips only tile bits, so a workgroup that stages one tile
// (2^K consecutive-in-tile states, strided in memory over the non-tile bits) in LDS
// can apply all of the pass's terms without touching HBM again.  Each pass reads v1
// once and read-modify-writes v0 once.
//
// Memory index m = (state << colbits) | column: batched vectors ([dim][nvec] interleaved)
// add never-flipped column bits at the bottom, which the planner keeps in the tile so
// the 16-byte lanes of a wave form contiguous runs.
#include "internal.h"

namespace pned {

struct PassArgs {
    uint64_t tile_count;   // number of tiles = 2^(outer bits)
    int K;                 // tile bits
    int nouter;
    int colbits;
    int first_term, nterm;
    int diag;              // add the diagonal (first pass)
    int fresh;             // first pass: v0 input is zero
    int ncol;              // columns (1 << colbits)
    double shift;          // -ibond * eperbd, added to the diagonal
    uint8_t tb[40];        // tile bit -> memory bit
    uint8_t ob[40];        // outer bit -> memory bit
};

template <typename R> struct C2;
template <> struct C2<double> { using T = double2; };
template <> struct C2<float> { using T = float2; };

__device__ __forceinline__ double diag_value_re(const DiagDev &d, uint32_t s, double &im)
{
    double re = d.c0r;
    im = d.c0i;
    for (int g = 0; g < d.ngroups; ++g) {
        uint32_t f = (d.kind[g] == 0) ? (s & d.mask[g]) : ((s ^ (s >> d.shift[g])) & d.mask[g]);
        const double c = (double)__popc(f);
        re = fma(c, d.wr[g], re);
        im = fma(c, d.wi[g], im);
    }
    return re;
}

// Direct HBM -> LDS load of one wave's 64 consecutive tile slots (gfx950
// global_load_lds_dwordx4 / _dwordx2): lane l writes slot (lds_wave_base + l).
__device__ __forceinline__ void load_to_lds(const double2 *g, double2 *lds_wave_base)
{
    __builtin_amdgcn_global_load_lds((const void *)g, (__attribute__((address_space(3))) void *)lds_wave_base, 16, 0, 0);
}
// no 8-byte form of the LDS-direct load exists: complex64 goes through a register
__device__ __forceinline__ void load_to_lds(const float2 *g, float2 *lds_wave_base)
{
    lds_wave_base[threadIdx.x & 63] = *g;
}

// TPB threads, R states per thread: tile of K = log2(TPB*R) bits.  Tile slot
// j = tid + TPB*r holds the state with memory index base | offLo(tid) | offHi(r).
// Per tile: (1) every wave streams its R slices of v1 straight into LDS; (2) states are
// processed CH at a time, the v0 values of the next chunk being fetched while the current
// one computes; (3) outputs are written in place.
template <typename Real, int TPB, int R>
__global__ __launch_bounds__(TPB) void hv_s12_pass(const typename C2<Real>::T *__restrict__ v1,
                                                   typename C2<Real>::T *__restrict__ v0,
                                                   const DevTerm *__restrict__ terms,
                                                   const DiagDev *__restrict__ diagp, PassArgs a,
                                                   double *__restrict__ partial)
{
    using V = typename C2<Real>::T;
    extern __shared__ unsigned char smem_raw[];
    V *tile = reinterpret_cast<V *>(smem_raw);
    constexpr int LT = __builtin_ctz(TPB);
    constexpr int LR = __builtin_ctz(R);
    constexpr int CH = (R < 4) ? R : 4;      // states per compute chunk
    const int tid = threadIdx.x;
    const int wave_base = tid & ~63;

    uint64_t offLo = 0;
#pragma unroll
    for (int k = 0; k < LT; ++k)
        if ((tid >> k) & 1) offLo |= 1ull << a.tb[k];
    uint64_t offHi[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        uint64_t o = 0;
#pragma unroll
        for (int k = 0; k < LR; ++k)
            if ((r >> k) & 1) o |= 1ull << a.tb[LT + k];
        offHi[r] = o;
    }
    const bool rd0 = !(a.diag && a.fresh);

    double pr_re = 0.0, pr_im = 0.0;
    for (uint64_t tix = blockIdx.x; tix < a.tile_count; tix += gridDim.x) {
        uint64_t base = 0;
        for (int k = 0; k < a.nouter; ++k)
            if ((tix >> k) & 1) base |= 1ull << a.ob[k];
        const uint64_t tb0 = base | offLo;
#pragma unroll
        for (int r = 0; r < R; ++r) load_to_lds(v1 + (tb0 | offHi[r]), tile + wave_base + TPB * r);
        V y[CH];
        if (rd0) {
#pragma unroll
            for (int c = 0; c < CH; ++c) y[c] = v0[tb0 | offHi[c]];
        }
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();

#pragma unroll
        for (int r0 = 0; r0 < R; r0 += CH) {
            uint32_t st[CH];
            double ar[CH], ai[CH];
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                st[c] = (uint32_t)((tb0 | offHi[r0 + c]) >> a.colbits);
                ar[c] = 0.0;
                ai[c] = 0.0;
            }
            if (a.diag) {
                const DiagDev &dg = *diagp;
#pragma unroll
                for (int c = 0; c < CH; ++c) {
                    double di;
                    const double dr = diag_value_re(dg, st[c], di) + a.shift;
                    const V x = tile[tid + TPB * (r0 + c)];
                    ar[c] = dr * (double)x.x - di * (double)x.y;
                    ai[c] = dr * (double)x.y + di * (double)x.x;
                }
            }
            for (int t = 0; t < a.nterm; ++t) {
                const DevTerm T = terms[a.first_term + t];
                if (T.is_real) {
                    const double d = T.c1r - T.c0r;
#pragma unroll
                    for (int c = 0; c < CH; ++c) {
                        const V p = tile[(tid + TPB * (r0 + c)) ^ T.tmask];
                        const double sel = (double)(__popc(st[c] & T.selmask) & 1);
                        const double cf = fma(sel, d, T.c0r);
                        ar[c] = fma(cf, (double)p.x, ar[c]);
                        ai[c] = fma(cf, (double)p.y, ai[c]);
                    }
                } else {
                    const double dr = T.c1r - T.c0r, di = T.c1i - T.c0i;
#pragma unroll
                    for (int c = 0; c < CH; ++c) {
                        const V p = tile[(tid + TPB * (r0 + c)) ^ T.tmask];
                        const double sel = (double)(__popc(st[c] & T.selmask) & 1);
                        const double cr = fma(sel, dr, T.c0r), ci = fma(sel, di, T.c0i);
                        ar[c] = fma(cr, (double)p.x, ar[c]);
                        ar[c] = fma(-ci, (double)p.y, ar[c]);
                        ai[c] = fma(cr, (double)p.y, ai[c]);
                        ai[c] = fma(ci, (double)p.x, ai[c]);
                    }
                }
            }
            // fetch the next chunk's v0 before writing this one
            V yn[CH];
            if (rd0 && r0 + CH < R) {
#pragma unroll
                for (int c = 0; c < CH; ++c) yn[c] = v0[tb0 | offHi[r0 + CH + c]];
            }
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                const V x = tile[tid + TPB * (r0 + c)];
                pr_re += (double)x.x * ar[c] + (double)x.y * ai[c];
                pr_im += (double)x.x * ai[c] - (double)x.y * ar[c];
                V o;
                if (rd0) {
                    o.x = (Real)((double)y[c].x + ar[c]);
                    o.y = (Real)((double)y[c].y + ai[c]);
                } else {
                    o.x = (Real)ar[c];
                    o.y = (Real)ai[c];
                }
                v0[tb0 | offHi[r0 + c]] = o;
            }
            if (rd0 && r0 + CH < R) {
#pragma unroll
                for (int c = 0; c < CH; ++c) y[c] = yn[c];
            }
        }
        __syncthreads();
    }

    // deterministic block reduction per column: lanes with equal (tid & (ncol-1)) belong to
    // the same column (the planner keeps the column bits as the lowest tile bits)
    double *red = reinterpret_cast<double *>(smem_raw);
    red[2 * tid] = pr_re;
    red[2 * tid + 1] = pr_im;
    __syncthreads();
    for (int s = TPB / 2; s >= a.ncol && s > 0; s >>= 1) {
        if (tid < s) {
            red[2 * tid] += red[2 * (tid + s)];
            red[2 * tid + 1] += red[2 * (tid + s) + 1];
        }
        __syncthreads();
    }
    if (tid < a.ncol) {
        partial[(size_t)blockIdx.x * 2 * a.ncol + 2 * tid] = red[2 * tid];
        partial[(size_t)blockIdx.x * 2 * a.ncol + 2 * tid + 1] = red[2 * tid + 1];
    }
}

// out[c] = sum over rows of partial[row][c]: one block per column, fixed strides and a
// fixed tree, so the result is the same for the same partials
__global__ __launch_bounds__(256) void reduce_partials(const double *__restrict__ partial, int nrows, int ncol2,
                                                       double *__restrict__ out, int accumulate)
{
    __shared__ double red[256];
    const int c = blockIdx.x, tid = threadIdx.x;
    double s = 0.0;
    for (int b = tid; b < nrows; b += 256) s += partial[(size_t)b * ncol2 + c];
    red[tid] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (tid < k) red[tid] += red[tid + k];
        __syncthreads();
    }
    if (tid == 0) out[c] = accumulate ? out[c] + red[0] : red[0];
}

template <typename Real, int TPB, int R>
static int launch_pass(pned_ctx *ctx, const void *v1, void *v0, const DiagDev *ddiag,
                       const PassArgs &a, double *partial, int grid)
{
    using V = typename C2<Real>::T;
    size_t lds = sizeof(V) * (size_t)TPB * R;
    size_t need_red = sizeof(double) * 2 * TPB;
    if (lds < need_red) lds = need_red;
    auto kern = hv_s12_pass<Real, TPB, R>;
    PNED_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(kern, dim3(grid), dim3(TPB), lds, ctx->stream, (const V *)v1, (V *)v0,
                       ctx->d_terms, ddiag, a, partial);
    PNED_HIP(hipGetLastError());
    return PNED_OK;
}

template <typename Real>
static int dispatch_pass(pned_ctx *ctx, int TPB, int R, const void *v1, void *v0,
                         const DiagDev *ddiag, const PassArgs &a, double *partial, int grid)
{
#define PNED_CASE(T_, R_) \
    if (TPB == T_ && R == R_) return launch_pass<Real, T_, R_>(ctx, v1, v0, ddiag, a, partial, grid);
    PNED_CASE(256, 32) PNED_CASE(256, 16) PNED_CASE(256, 8) PNED_CASE(256, 4) PNED_CASE(256, 2)
    PNED_CASE(256, 1) PNED_CASE(128, 1) PNED_CASE(64, 1) PNED_CASE(32, 1) PNED_CASE(16, 1)
    PNED_CASE(512, 32) PNED_CASE(512, 16) PNED_CASE(512, 8) PNED_CASE(512, 4) PNED_CASE(512, 2)
#undef PNED_CASE
    return fail(PNED_E_UNSUPPORTED, "no H.v kernel instance for this tile shape");
}

int launch_hv_s12(pned_ctx *ctx, const pned_vec *v1, pned_vec *v0, double eperbd, bool fresh,
                  double *prdct_host)
{
    const int nvec = v1->nvec;
    int colbits = 0;
    while ((1 << colbits) < nvec) ++colbits;
    if ((1 << colbits) != nvec)
        return fail(PNED_E_UNSUPPORTED, "batched H.v needs a power-of-two column count");
    PNED_TRY(ensure_s12_plan(ctx, colbits, v1->dtype));
    const S12Plan &pl = ctx->s12;
    const int npass = (int)pl.passes.size();
    const int grid_cap = 4096;
    // partial layout: [pass][block][2*ncol]
    size_t need = (size_t)npass * grid_cap * 2 * nvec + 64;
    PNED_TRY(ensure_partial(ctx, need));
    PNED_TRY(ensure_result(ctx, 2 * (size_t)nvec + 8));
    int total_rows = 0;
    for (int p = 0; p < npass; ++p) {
        const PassPlan &pp = pl.passes[p];
        PassArgs a{};
        a.K = pp.K;
        a.nouter = (int)pp.ob.size();
        a.tile_count = 1ull << a.nouter;
        a.colbits = colbits;
        a.first_term = pp.first_term;
        a.nterm = pp.nterm;
        a.diag = pp.diag ? 1 : 0;
        a.fresh = fresh ? 1 : 0;
        a.ncol = nvec;
        a.shift = -(double)ctx->ibond * eperbd;
        for (int k = 0; k < pp.K; ++k) a.tb[k] = (uint8_t)pp.tb[k];
        for (int k = 0; k < a.nouter; ++k) a.ob[k] = (uint8_t)pp.ob[k];
        // threads per block: tunable (256 or 512); R = tile / threads states per thread
        int TPB = (ctx->tune_threads == 256 || ctx->tune_threads == 512) ? ctx->tune_threads : 512;
        while (TPB > 16 && TPB > (1 << pp.K)) TPB >>= 1;
        const int R = (1 << pp.K) / TPB;
        if (R > 32) return fail(PNED_E_UNSUPPORTED, "tile too large for the thread count");
        int grid = (int)((a.tile_count < (uint64_t)grid_cap) ? a.tile_count : (uint64_t)grid_cap);
        double *part = ctx->d_partial + (size_t)total_rows * 2 * nvec;
        int rc;
        if (v1->dtype == PNED_C128)
            rc = dispatch_pass<double>(ctx, TPB, R, v1->data, v0->data, ctx->d_diag, a, part, grid);
        else
            rc = dispatch_pass<float>(ctx, TPB, R, v1->data, v0->data, ctx->d_diag, a, part, grid);
        if (rc != PNED_OK) return rc;
        total_rows += grid;
    }
    const int ncol2 = 2 * nvec;
    hipLaunchKernelGGL(reduce_partials, dim3(ncol2), dim3(256), 0, ctx->stream, ctx->d_partial, total_rows, ncol2,
                       ctx->d_result, 0);
    PNED_HIP(hipGetLastError());
    if (prdct_host) {
        PNED_HIP(hipMemcpyAsync(ctx->h_result, ctx->d_result, sizeof(double) * ncol2,
                                hipMemcpyDeviceToHost, ctx->stream));
        PNED_HIP(hipStreamSynchronize(ctx->stream));
        for (int c = 0; c < ncol2; ++c) prdct_host[c] = ctx->h_result[c];
    }
    return PNED_OK;
}

}  // namespace pned
