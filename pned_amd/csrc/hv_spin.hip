// hv_spin.hip -- spin-S (q = 2S+1 digits) matrix-free H.v for gfx950.
//
// p_mltply_v3 / H_dot_v of the spin-S tree (S_arbitrary/Hybrid_TPS9_parallel.f90:184-356):
//     v0(l) += sum_bonds sum_in xlocal_Ham(in, istat(l), t) * v1(l with the bond's digits = in)
// with istat = q*digit(max site) + digit(min site) and the diagonal shifted by -eperbd per
// bond.  One thread owns one output state and sums all bonds in registers (a gather: no
// write conflicts, no atomics); v0 is read and written once.  Per (type, output local
// state) the nonzero input states are precomputed on the host, so the inner loop visits
// only the elements the reference does not skip (llocal_Ham).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <utility>

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

int launch_hv_spin_tiled(pned_ctx *ctx, const pned_vec *v1, pned_vec *v0, double eperbd, bool fresh,
                         double *prdct_host, double xs, double ys);

int launch_hv_spin(pned_ctx *ctx, const pned_vec *v1, pned_vec *v0, double eperbd, bool fresh,
                   double *prdct_host)
{
    if (v1->nvec != 1 || v1->dtype != PNED_C128)
        return fail(PNED_E_UNSUPPORTED, "spin-S H.v takes one complex128 column");
    if (ctx->spin_tiled && ctx->q >= 3 && ctx->q <= 5)
        return launch_hv_spin_tiled(ctx, v1, v0, eperbd, fresh, prdct_host, 1.0, 1.0);
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

// ------------------------------------------------------------------------------------
// Tiled spin-S H.v.  A pass owns K tile sites (the lowest L sites always among them, so a
// wave's lanes read contiguous runs of q^L states); every bond with both sites in the tile
// is applied in that pass from the tile staged in LDS.  The tile is the q^K states sharing
// the digits of the other sites; tile-local index j = sum digit_k q^k over the tile sites in
// order.  Each pass reads v1 once and read-modify-writes v0 once (pass 0 may start from
// zero).
//
// A bond's elements that the reference's llocal_Ham does not skip are grouped by their
// partner shift (input minus output digit on each of the two sites).  Within such a class
// the partner's tile offset is the same for every state; only the coefficient depends on
// the state's two digits: a q^2-entry table, zero where the element is skipped or the
// partner digit would leave [0, q).  A thread owns R tile slots j = tid + TPB r for the whole
// launch, so each slot's digits are fixed: the diagonals of all the pass's bonds are summed
// into one coefficient per slot at launch start, and per tile the slots visit the
// off-diagonal classes under uniform (scalar) loop control, one coefficient and one partner
// read from LDS each, S slots at a time for independent reads.  The next tile's v1 is
// prefetched into registers while the current one is computed.  Tiles are handed out so that
// the blocks of one XCD (blockIdx mod 8) work on neighbouring tiles at the same time: the
// runs of neighbouring tiles share cache lines at their ends.
// ------------------------------------------------------------------------------------
constexpr int SP_MAX_TILE = 8192;     // states per tile (128 KB of complex128)
constexpr int SP_MAX_BOND = 64;       // bonds per pass: one record per lane of a wave
constexpr int SP_BASE_TAB = 4096;
constexpr int SP_MAX_QQ = 25;         // q <= 5
constexpr int SP_MAX_CLS = 64;        // off-diagonal (bond, partner shift) classes per pass: one per lane
constexpr int SP_MAX_COEF = SP_MAX_CLS * SP_MAX_QQ;

struct SpinPassDev {
    uint32_t off[SP_MAX_TILE];        // memory offset of tile-local state j (dim <= 2^32)
    uint64_t blo[SP_BASE_TAB];        // tile base, low outer digits
    uint64_t bhi[SP_BASE_TAB];        // tile base, high outer digits
    // bond record: digit-word shift (4 * tile position) of the min site (bits 0-4) and of the
    // max site (5-9), first off-diagonal class (10-15), one past the last (16-22)
    int32_t bond_rec[SP_MAX_BOND];
    int32_t cls_delta[SP_MAX_CLS];    // (partner - j) in bytes of a complex128 tile entry
    int32_t diag_kept[SP_MAX_BOND * SP_MAX_QQ];   // diagonal element present: the shift applies
    double diag[2 * SP_MAX_BOND * SP_MAX_QQ];     // bond diagonal by output local state
    double cls_c[2 * SP_MAX_COEF];    // class coefficient by output local state (0 where none)
};

struct SpinPassArgs {
    uint64_t tile_count;
    int tsz;                          // q^K
    int K, nbond, ncls, diag_first, fresh;
    int nlo_mod;                      // tile_count split: tix % nlo_mod, tix / nlo_mod
    int xcd;                          // 1: neighbouring tiles on one XCD (grid a multiple of 8)
    double shift;                     // -eperbd per bond diagonal
    double xs;                        // v0 += xs (H v1); <v1|.> then carries one factor xs
    double ys;                        // first pass: factor on the v0 input
    double zr, zi;                    // first pass: v0 += (zr + i zi) v1 (complex diagonal shift)
    int dotu;                         // 1: accumulate sum v1 * (H v1) without the conjugate
};

// Slots are visited in sweeps of S; a sweep is skipped by the waves that hold no state of the
// tile in its first slot, and within a sweep every slot is computed (a slot beyond the tile
// reads clamped, in-bounds entries and is not stored), so the class loop has no branches.
// S is chosen so that only the last sweep can be partial.
template <int Q, int TPB, int R, int S, bool CPLX>
__global__ __launch_bounds__(TPB) void hv_spin_pass(const double2 *__restrict__ v1, double2 *__restrict__ v0,
                                                    const SpinPassDev *__restrict__ pd, SpinPassArgs a,
                                                    double *__restrict__ partial)
{
    constexpr int QQ = Q * Q;
    constexpr int CW = CPLX ? 16 : 8;     // bytes per staged coefficient
    extern __shared__ unsigned char spin_smem[];
    double2 *tile = reinterpret_cast<double2 *>(spin_smem);
    unsigned char *cfb = spin_smem + sizeof(double2) * (size_t)a.tsz;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    {
        const int ncoef = a.ncls * QQ;
        double *cf = reinterpret_cast<double *>(cfb);
        for (int i = tid; i < ncoef; i += TPB) {
            if (CPLX) {
                cf[2 * i] = pd->cls_c[2 * i];
                cf[2 * i + 1] = pd->cls_c[2 * i + 1];
            } else {
                cf[i] = pd->cls_c[2 * i];
            }
        }
    }
    // the pass's bond and class records, lane b / c of every wave
    const int brec = (lane < a.nbond) ? pd->bond_rec[lane] : 0;
    const int crec = (lane < a.ncls) ? pd->cls_delta[lane] : 0;
    const bool rd0 = !(a.diag_first && a.fresh);
    const double ysc = a.diag_first ? a.ys : 1.0;
    const uint32_t tmax = (uint32_t)(a.tsz - 1) * 16u;
    // first thread of this wave: slot r holds a state of the tile for some lane of the wave
    // iff wfirst + TPB r < tsz (a uniform test)
    const int wfirst = __builtin_amdgcn_readfirstlane(tid & ~63);
    // per slot for the launch: digit word (4 bits per tile digit), memory offset, diagonal
    uint32_t dw[R];
    uint32_t offj[R];
    double gr[R], gi[CPLX ? R : 1];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int j = tid + TPB * r;
        uint32_t w = 0;
        int x = j;
        for (int k = 0; k < a.K; ++k) {
            w |= (uint32_t)(x % Q) << (4 * k);
            x /= Q;
        }
        dw[r] = w;
        offj[r] = (j < a.tsz) ? pd->off[j] : 0;
        double sr = 0.0, si = (CPLX && a.diag_first) ? a.zi : 0.0;
        if (j < a.tsz) {
            for (int b = 0; b < a.nbond; ++b) {
                const int rec = __builtin_amdgcn_readlane(brec, b);
                const int o = b * QQ + Q * (int)((w >> ((rec >> 5) & 31)) & 15u) + (int)((w >> (rec & 31)) & 15u);
                sr += pd->diag[2 * o] + (pd->diag_kept[o] ? a.shift : 0.0);
                si += pd->diag[2 * o + 1];
            }
            if (a.diag_first) sr += a.zr;
        }
        gr[r] = sr;
        if (CPLX) gi[r] = si;
    }
    // pass 0 of a real pass: the imaginary part of the complex diagonal shift
    const double zi = (a.diag_first && !CPLX) ? a.zi : 0.0;
    double pr_re = 0.0, pr_im = 0.0;
    double xr[R], xi[R];
    auto tile_base = [&](uint64_t tix) {
        return pd->blo[tix % (uint64_t)a.nlo_mod] + pd->bhi[tix / (uint64_t)a.nlo_mod];
    };
    auto fetch_x = [&](uint64_t base) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (tid + TPB * r < a.tsz) {
                const double2 t = v1[base + offj[r]];
                xr[r] = t.x;
                xi[r] = t.y;
            }
        }
    };
    const uint64_t G = gridDim.x;
    uint64_t tix = a.xcd ? (uint64_t)(blockIdx.x & 7u) * (G >> 3) + (blockIdx.x >> 3) : (uint64_t)blockIdx.x;
    if (tix < a.tile_count) fetch_x(tile_base(tix));
    __syncthreads();   // staged coefficients visible
    for (; tix < a.tile_count; tix += G) {
        const uint64_t base = tile_base(tix);
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (tid + TPB * r < a.tsz) tile[tid + TPB * r] = make_double2(xr[r], xi[r]);
        __syncthreads();
        const uint64_t nx = tix + G;
        // v0 of the first sweep before the next tile's v1 (in-order vector memory counters)
        double yr[S], yi[S];
        auto fetch_y = [&](int r0) {
#pragma unroll
            for (int u = 0; u < S; ++u) {
                const int r = r0 + u;
                yr[u] = 0.0;
                yi[u] = 0.0;
                if (r < R && rd0 && tid + TPB * r < a.tsz) {
                    const double2 t = v0[base + offj[r]];
                    yr[u] = t.x;
                    yi[u] = t.y;
                }
            }
        };
        fetch_y(0);
        if (nx < a.tile_count) fetch_x(tile_base(nx));
#pragma unroll
        for (int r0 = 0; r0 < R; r0 += S) {
            if (r0 > 0) fetch_y(r0);
            if (wfirst + TPB * r0 >= a.tsz) continue;   // uniform: no state of this sweep in the wave
            double ar[S], ai[S];
            uint32_t jb[S];   // byte offset of the slot's state in the tile (clamped)
#pragma unroll
            for (int u = 0; u < S; ++u) {
                const int r = (r0 + u < R) ? r0 + u : R - 1;   // (slots past R are never used)
                jb[u] = min((uint32_t)(tid + TPB * r) * 16u, tmax);
                const double2 x = *reinterpret_cast<const double2 *>(spin_smem + jb[u]);
                const double g_i = CPLX ? gi[r] : zi;
                ar[u] = fma(gr[r], x.x, -g_i * x.y);
                ai[u] = fma(gr[r], x.y, g_i * x.x);
            }
#pragma unroll 1
            for (int b = 0; b < a.nbond; ++b) {
                const int rec = __builtin_amdgcn_readlane(brec, b);
                const int sa = rec & 31, sb = (rec >> 5) & 31;
                const int c0 = (rec >> 10) & 63, c1 = (rec >> 16) & 127;
                uint32_t ob[S];   // byte offset of the slot's output local state in a class table
#pragma unroll
                for (int u = 0; u < S; ++u) {
                    const int r = (r0 + u < R) ? r0 + u : R - 1;
                    ob[u] = (uint32_t)(Q * (int)((dw[r] >> sb) & 15u) + (int)((dw[r] >> sa) & 15u)) * CW;
                }
#pragma unroll 1
                for (int c = c0; c < c1; ++c) {
                    const uint32_t dl = (uint32_t)__builtin_amdgcn_readlane(crec, c);
                    const unsigned char *cb = cfb + (size_t)c * QQ * CW;
                    // all of the sweep's reads first, so they are in flight together
                    double2 p[S], cc[S];
#pragma unroll
                    for (int u = 0; u < S; ++u) {
                        if (r0 + u >= R) break;   // resolved at compile time
                        p[u] = *reinterpret_cast<const double2 *>(spin_smem + min(jb[u] + dl, tmax));
                        if (CPLX) cc[u] = *reinterpret_cast<const double2 *>(cb + ob[u]);
                        else cc[u].x = *reinterpret_cast<const double *>(cb + ob[u]);
                    }
#pragma unroll
                    for (int u = 0; u < S; ++u) {
                        if (r0 + u >= R) break;
                        if (CPLX) {
                            ar[u] = fma(cc[u].x, p[u].x, fma(-cc[u].y, p[u].y, ar[u]));
                            ai[u] = fma(cc[u].x, p[u].y, fma(cc[u].y, p[u].x, ai[u]));
                        } else {
                            ar[u] = fma(cc[u].x, p[u].x, ar[u]);
                            ai[u] = fma(cc[u].x, p[u].y, ai[u]);
                        }
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < S; ++u) {
                const int r = r0 + u;
                if (r >= R || tid + TPB * r >= a.tsz) continue;
                const double2 x = tile[tid + TPB * r];
                const double hr = ar[u] * a.xs, hi = ai[u] * a.xs;
                if (a.dotu) {
                    pr_re += x.x * hr - x.y * hi;
                    pr_im += x.x * hi + x.y * hr;
                } else {
                    pr_re += x.x * hr + x.y * hi;
                    pr_im += x.x * hi - x.y * hr;
                }
                v0[base + offj[r]] = make_double2(fma(ysc, yr[u], hr), fma(ysc, yi[u], hi));
            }
        }
        __syncthreads();
    }
    double *red = reinterpret_cast<double *>(spin_smem);
    red[2 * tid] = pr_re;
    red[2 * tid + 1] = pr_im;
    __syncthreads();
    for (int st = TPB / 2; st > 0; st >>= 1) {
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

// ---- planner (host only) ------------------------------------------------------------

namespace {

// greedy cover of the bonds by tile-site sets (lowest L sites forced): returns one site set
// per pass
std::vector<std::vector<int>> spin_greedy_cover(int n, const std::vector<int32_t> &sp, int ibond, int K, int L)
{
    std::vector<char> done(ibond, 0);
    int remaining = ibond;
    std::vector<std::vector<int>> out;
    while (remaining > 0) {
        std::vector<char> in(n, 0);
        int cnt = 0;
        for (int k = 0; k < L && k < n; ++k) { in[k] = 1; ++cnt; }
        while (cnt < K && cnt < n) {
            int best = -1;
            long bs = -1;
            for (int st = 0; st < n; ++st) {
                if (in[st]) continue;
                long newly = 0, touch = 0;
                for (int b = 0; b < ibond; ++b) {
                    if (done[b]) continue;
                    const int x = sp[2 * b], y = sp[2 * b + 1];
                    if (x != st && y != st) continue;
                    ++touch;
                    const int other = (x == st) ? y : x;
                    if (in[other]) ++newly;
                }
                const long score = newly * (1L << 30) + touch * (1L << 10) + (n - st);
                if (score > bs) { bs = score; best = st; }
            }
            if (best < 0) break;
            in[best] = 1;
            ++cnt;
        }
        bool progress = false;
        for (int b = 0; b < ibond; ++b)
            if (!done[b] && in[sp[2 * b]] && in[sp[2 * b + 1]]) { done[b] = 1; --remaining; progress = true; }
        if (!progress) return {};
        std::vector<int> sites;
        for (int st = 0; st < n; ++st)
            if (in[st]) sites.push_back(st);
        out.push_back(sites);
    }
    return out;
}

uint64_t sp_mix(uint64_t &s)
{
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// P tile-site sets (each holding sites 0..L-1 and K-L others) covering every bond, by
// annealing from random sets: a move exchanges one free site of one set for a site outside
// it; the cost is the number of bonds no set covers (Metropolis at a falling temperature).
// Returns the sets, or nothing when the search ends without a cover.
std::vector<std::vector<int>> spin_anneal_cover(int n, const std::vector<int32_t> &sp, int ibond, int K, int L,
                                                int P, uint64_t seed, int moves)
{
    if (K >= n || K - L < 1) return {};
    std::vector<std::vector<char>> in(P, std::vector<char>(n, 0));
    for (int p = 0; p < P; ++p) {
        for (int k = 0; k < L; ++k) in[p][k] = 1;
        std::vector<int> fr;
        for (int st = L; st < n; ++st) fr.push_back(st);
        for (int k = 0; k < K - L; ++k) {
            const int i = k + (int)(sp_mix(seed) % (uint64_t)(fr.size() - k));
            std::swap(fr[k], fr[i]);
            in[p][fr[k]] = 1;
        }
    }
    std::vector<int> cov(ibond, 0);
    for (int p = 0; p < P; ++p)
        for (int b = 0; b < ibond; ++b) cov[b] += in[p][sp[2 * b]] && in[p][sp[2 * b + 1]];
    int unc = 0;
    for (int b = 0; b < ibond; ++b) unc += cov[b] == 0;
    double T = 1.0;
    for (int it = 0; it < moves && unc > 0; ++it) {
        const int p = (int)(sp_mix(seed) % (uint64_t)P);
        int so, si;
        do { so = L + (int)(sp_mix(seed) % (uint64_t)(n - L)); } while (!in[p][so]);
        do { si = L + (int)(sp_mix(seed) % (uint64_t)(n - L)); } while (in[p][si]);
        // change of the uncovered count when site so leaves set p and si joins it
        int d = 0;
        for (int b = 0; b < ibond; ++b) {
            const int x = sp[2 * b], y = sp[2 * b + 1];
            const bool before = in[p][x] && in[p][y];
            const bool after = (x != so && y != so) && (in[p][x] || x == si) && (in[p][y] || y == si);
            if (before == after) continue;
            if (before && cov[b] == 1) ++d;
            if (after && cov[b] == 0) --d;
        }
        const double u = (double)(sp_mix(seed) >> 11) * (1.0 / 9007199254740992.0);
        if (d <= 0 || u < std::exp(-d / T)) {
            for (int b = 0; b < ibond; ++b) {
                const int x = sp[2 * b], y = sp[2 * b + 1];
                const bool before = in[p][x] && in[p][y];
                const bool after = (x != so && y != so) && (in[p][x] || x == si) && (in[p][y] || y == si);
                cov[b] += (int)after - (int)before;
            }
            in[p][so] = 0;
            in[p][si] = 1;
            unc += d;
        }
        T = std::max(0.05, T * 0.9999);
    }
    if (unc > 0) return {};
    std::vector<std::vector<int>> out(P);
    for (int p = 0; p < P; ++p)
        for (int st = 0; st < n; ++st)
            if (in[p][st]) out[p].push_back(st);
    return out;
}

constexpr uint64_t SPIN_COVER_SEED = 0x5EED0003ull;
constexpr int SPIN_COVER_TRIES = 4;
constexpr int SPIN_COVER_MOVES = 60000;

// the fewest passes found: the greedy cover, then annealed covers with one set fewer until
// a search fails
std::vector<std::vector<int>> spin_cover(int n, const std::vector<int32_t> &sp, int ibond, int K, int L)
{
    auto best = spin_greedy_cover(n, sp, ibond, K, L);
    if (best.empty()) return best;
    uint64_t seed = SPIN_COVER_SEED ^ ((uint64_t)K << 8) ^ (uint64_t)L;
    for (int P = (int)best.size() - 1; P >= 1; --P) {
        std::vector<std::vector<int>> got;
        for (int t = 0; t < SPIN_COVER_TRIES && got.empty(); ++t)
            got = spin_anneal_cover(n, sp, ibond, K, L, P, sp_mix(seed), SPIN_COVER_MOVES);
        if (got.empty()) break;
        best.swap(got);
    }
    return best;
}

// Spreads the bonds over the passes whose tiles hold both of their sites: bonds with the
// fewest covering passes are placed first, each on the pass with the least work so far
// (w[b]: the bond's cost in the kernel), so that no pass carries much more than the others.
// Passes left without bonds are dropped.  Returns per pass its sites and bonds.
std::vector<std::pair<std::vector<int>, std::vector<int>>> spin_balance(const std::vector<std::vector<int>> &cov,
                                                                        const std::vector<int32_t> &sp, int n,
                                                                        int ibond, const std::vector<int> &w)
{
    const int np = (int)cov.size();
    std::vector<std::vector<char>> has(np, std::vector<char>(n, 0));
    for (int p = 0; p < np; ++p)
        for (int st : cov[p]) has[p][st] = 1;
    std::vector<std::vector<int>> choice(ibond);
    for (int b = 0; b < ibond; ++b)
        for (int p = 0; p < np; ++p)
            if (has[p][sp[2 * b]] && has[p][sp[2 * b + 1]]) choice[b].push_back(p);
    std::vector<int> order(ibond);
    for (int b = 0; b < ibond; ++b) order[b] = b;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return choice[a].size() < choice[b].size(); });
    std::vector<std::vector<int>> bonds(np);
    std::vector<long> load(np, 0);
    for (int b : order) {
        int best = choice[b][0];
        for (int p : choice[b])
            if (load[p] < load[best]) best = p;
        bonds[best].push_back(b);
        load[best] += w[b];
    }
    std::vector<std::pair<std::vector<int>, std::vector<int>>> out;
    for (int p = 0; p < np; ++p) {
        if (bonds[p].empty()) continue;
        std::sort(bonds[p].begin(), bonds[p].end());
        out.emplace_back(cov[p], bonds[p]);
    }
    return out;
}

// Splits a pass whose bonds hold more off-diagonal classes or bonds than one wave holds
// records for (64 each) into passes over the same tile sites.
std::vector<std::pair<std::vector<int>, std::vector<int>>> spin_split(
    const std::vector<std::pair<std::vector<int>, std::vector<int>>> &passes, const std::vector<int32_t> &btype,
    const std::vector<std::vector<std::pair<int, int>>> &tcls)
{
    std::vector<std::pair<std::vector<int>, std::vector<int>>> out;
    for (const auto &p : passes) {
        std::vector<int> cur;
        int nc = 0;
        for (int b : p.second) {
            const int k = (int)tcls[btype[b] - 1].size();
            if (!cur.empty() && (nc + k > SP_MAX_CLS || (int)cur.size() == SP_MAX_BOND)) {
                out.emplace_back(p.first, cur);
                cur.clear();
                nc = 0;
            }
            cur.push_back(b);
            nc += k;
        }
        if (!cur.empty()) out.emplace_back(p.first, cur);
    }
    return out;
}

// off-diagonal partner shifts of a bond type: (in - out) digit of the min and max site
// among the elements that are not skipped
std::vector<std::pair<int, int>> spin_classes(int q, const std::vector<int32_t> &ll, int t)
{
    const int qq = q * q;
    std::vector<std::pair<int, int>> cls;
    for (int out = 0; out < qq; ++out)
        for (int in = 0; in < qq; ++in) {
            if (in == out) continue;
            if (ll[(size_t)in + (size_t)qq * out + (size_t)qq * qq * t]) continue;
            const std::pair<int, int> k(in % q - out % q, in / q - out / q);
            if (std::find(cls.begin(), cls.end(), k) == cls.end()) cls.push_back(k);
        }
    std::sort(cls.begin(), cls.end());
    return cls;
}

// LDS bytes of one pass: tile, then the class coefficients
size_t spin_lds_bytes(int tsz, int ncls, int qq, bool cplx, int tpb)
{
    const size_t t = sizeof(double2) * (size_t)tsz + (size_t)ncls * qq * (cplx ? 16 : 8);
    return std::max(t, sizeof(double) * 2 * (size_t)tpb);
}

struct SpinPlan {
    int K = 0, L = 0;
    std::vector<SpinPassDev> dev;
    std::vector<std::vector<int>> sites, bonds;
    std::vector<int> tsz, nlo, ncls, cplx;
    std::vector<std::vector<std::array<int, 3>>> cls_key;   // per pass and class: bond, digit shifts
};

// Builds the passes for tile sites K (at most, fewer when the tables do not fit) and L forced
// low sites (0: 3 if that costs no pass over 2, else 2).
int spin_make_plan(int n, int q, int ibond, const std::vector<int32_t> &sp, const std::vector<int32_t> &btype,
                   const std::vector<double> &xl, const std::vector<int32_t> &ll, int Kmax, int Lreq, SpinPlan &plan)
{
    const int qq = q * q;
    if (q < 3 || q > 5) return fail(PNED_E_UNSUPPORTED, "tiled spin-S kernel: q = 3..5");
    std::vector<std::vector<std::pair<int, int>>> tcls(9);
    for (int t = 0; t < 9; ++t) tcls[t] = spin_classes(q, ll, t);
    std::vector<int> w(ibond);
    for (int b = 0; b < ibond; ++b) w[b] = 1 + 4 * (int)tcls[btype[b] - 1].size();
    std::vector<uint64_t> pw(n);
    {
        uint64_t x = 1;
        for (int k = 0; k < n; ++k) { pw[k] = x; x *= (uint64_t)q; }
    }
    for (int K = Kmax; K >= 1; --K) {
        std::vector<std::vector<int>> cov;
        int L = 0;
        if (Lreq > 0) {
            L = std::min(Lreq, K);
            cov = spin_cover(n, sp, ibond, K, L);
        } else {
            auto c2 = spin_cover(n, sp, ibond, K, std::min(2, K));
            auto c3 = (K >= 3) ? spin_cover(n, sp, ibond, K, 3) : std::vector<std::vector<int>>{};
            if (!c3.empty() && (c2.empty() || c3.size() <= c2.size())) { cov = c3; L = 3; }
            else { cov = c2; L = std::min(2, K); }
        }
        // tiles too small to hold every bond next to the forced sites: force fewer
        while (cov.empty() && L > 0) cov = spin_cover(n, sp, ibond, K, --L);
        if (cov.empty()) return fail(PNED_E_UNSUPPORTED, "spin-S tile planner made no progress");
        auto passes = spin_split(spin_balance(cov, sp, n, ibond, w), btype, tcls);
        const int np = (int)passes.size();
        plan = SpinPlan{};
        plan.K = K;
        plan.L = L;
        plan.dev.assign(np, SpinPassDev{});
        bool fits = true;
        for (int p = 0; p < np && fits; ++p) {
            const std::vector<int> &sites = passes[p].first, &bonds = passes[p].second;
            SpinPassDev &d = plan.dev[p];
            if ((int)bonds.size() > SP_MAX_BOND) return fail(PNED_E_UNSUPPORTED, "too many bonds in a spin-S pass");
            const int Kp = (int)sites.size();
            int ts = 1;
            for (int k = 0; k < Kp; ++k) ts *= q;
            for (int j = 0; j < ts; ++j) {
                uint64_t o = 0;
                int y = j;
                for (int k = 0; k < Kp; ++k) { o += (uint64_t)(y % q) * pw[sites[k]]; y /= q; }
                d.off[j] = (uint32_t)o;
            }
            std::vector<int> outer;
            for (int st = 0; st < n; ++st)
                if (std::find(sites.begin(), sites.end(), st) == sites.end()) outer.push_back(st);
            const int no = (int)outer.size();
            int nl = 0;
            int64_t lm = 1;
            while (nl < no && lm * q <= SP_BASE_TAB) { lm *= q; ++nl; }
            int64_t hm = 1;
            for (int k = nl; k < no; ++k) hm *= q;
            if (hm > SP_BASE_TAB) return fail(PNED_E_UNSUPPORTED, "too many outer sites for the spin-S tables");
            for (int64_t t2 = 0; t2 < lm; ++t2) {
                uint64_t o = 0;
                int64_t y = t2;
                for (int k = 0; k < nl; ++k) { o += (uint64_t)(y % q) * pw[outer[k]]; y /= q; }
                d.blo[t2] = o;
            }
            for (int64_t t2 = 0; t2 < hm; ++t2) {
                uint64_t o = 0;
                int64_t y = t2;
                for (int k = nl; k < no; ++k) { o += (uint64_t)(y % q) * pw[outer[k]]; y /= q; }
                d.bhi[t2] = o;
            }
            int nc = 0;
            bool cplx = false;
            std::vector<std::array<int, 3>> keys;
            for (size_t bi = 0; bi < bonds.size(); ++bi) {
                const int bond = bonds[bi];
                const int sa = sp[2 * bond], sb = sp[2 * bond + 1];
                const int pa = (int)(std::find(sites.begin(), sites.end(), sa) - sites.begin());
                const int pbp = (int)(std::find(sites.begin(), sites.end(), sb) - sites.begin());
                int wa = 1, wb = 1;
                for (int k = 0; k < pa; ++k) wa *= q;
                for (int k = 0; k < pbp; ++k) wb *= q;
                const int t = btype[bond] - 1;
                const int nct = (int)tcls[t].size();
                if (nc + nct > SP_MAX_CLS) return fail(PNED_E_UNSUPPORTED, "spin-S pass tables too large");
                d.bond_rec[bi] = (4 * pa) | ((4 * pbp) << 5) | (nc << 10) | ((nc + nct) << 16);
                for (int out = 0; out < qq; ++out) {
                    const size_t f = (size_t)out + (size_t)qq * out + (size_t)qq * qq * t;
                    const size_t e = bi * qq + out;
                    d.diag_kept[e] = ll[f] ? 0 : 1;
                    d.diag[2 * e] = ll[f] ? 0.0 : xl[2 * f];
                    d.diag[2 * e + 1] = ll[f] ? 0.0 : xl[2 * f + 1];
                    cplx = cplx || d.diag[2 * e + 1] != 0.0;   // the real kernel drops it
                }
                for (const auto &k : tcls[t]) {
                    keys.push_back({bond, k.first, k.second});
                    d.cls_delta[nc] = 16 * (k.first * wa + k.second * wb);
                    for (int out = 0; out < qq; ++out) {
                        const int d1 = out % q + k.first, d2 = out / q + k.second;
                        double cr = 0.0, ci = 0.0;
                        if (d1 >= 0 && d1 < q && d2 >= 0 && d2 < q) {
                            const int in = d1 + q * d2;
                            const size_t f = (size_t)in + (size_t)qq * out + (size_t)qq * qq * t;
                            if (!ll[f]) {
                                cr = xl[2 * f];
                                ci = xl[2 * f + 1];
                            }
                        }
                        d.cls_c[2 * ((size_t)nc * qq + out)] = cr;
                        d.cls_c[2 * ((size_t)nc * qq + out) + 1] = ci;
                        cplx = cplx || ci != 0.0;
                    }
                    ++nc;
                }
            }
            plan.sites.push_back(sites);
            plan.bonds.push_back(bonds);
            plan.tsz.push_back(ts);
            plan.nlo.push_back((int)lm);
            plan.ncls.push_back(nc);
            plan.cplx.push_back(cplx ? 1 : 0);
            plan.cls_key.push_back(keys);
            fits = spin_lds_bytes(ts, nc, qq, cplx, 1024) <= 160 * 1024;
        }
        if (fits) return PNED_OK;
    }
    return fail(PNED_E_UNSUPPORTED, "spin-S tables do not fit LDS");
}

int spin_kmax(int n, int q)
{
    int K = 0;
    int64_t p = 1;
    while (K < n && p * q <= SP_MAX_TILE) { p *= q; ++K; }
    return K;
}

}  // namespace

extern "C" int pned_spin_plan_export(int n_sites, int q, int ibond, const int32_t *ipair, const int32_t *ibond_type,
                                     const double *xlocal_ham, const int32_t *llocal_ham, int tile_sites,
                                     int low_sites, int *npass, int *low_used, int32_t *pass_sites,
                                     int32_t *bond_pass, double *cls, int32_t *ncls, int max_pass)
{
    if (!ipair || !ibond_type || !xlocal_ham || !llocal_ham || !npass || !low_used || !pass_sites || !bond_pass ||
        !cls || !ncls)
        return fail(PNED_E_ARG, "null argument");
    if (q < 3 || q > 5 || n_sites < 2 || n_sites > 32 || ibond < 1) return fail(PNED_E_ARG, "q = 3..5, 2..32 sites");
    const int qq = q * q;
    std::vector<int32_t> sp(2 * ibond), bt(ibond_type, ibond_type + ibond);
    for (int b = 0; b < ibond; ++b) {
        const int a = ipair[2 * b] - 1, c = ipair[2 * b + 1] - 1;
        if (a < 0 || a >= n_sites || c < 0 || c >= n_sites) return fail(PNED_E_ARG, "ipair out of range");
        sp[2 * b] = std::min(a, c);
        sp[2 * b + 1] = std::max(a, c);
    }
    std::vector<double> xl(xlocal_ham, xlocal_ham + 2 * (size_t)qq * qq * 9);
    std::vector<int32_t> ll(llocal_ham, llocal_ham + (size_t)qq * qq * 9);
    int K = spin_kmax(n_sites, q);
    if (tile_sites > 0 && tile_sites < K) K = tile_sites;
    SpinPlan plan;
    PNED_TRY(spin_make_plan(n_sites, q, ibond, sp, bt, xl, ll, K, low_sites, plan));
    const int np = (int)plan.dev.size();
    *npass = np;
    *low_used = plan.L;
    if (np > max_pass) return fail(PNED_E_ARG, "output arrays too small");
    const int row = 3 + 2 * qq;
    for (int p = 0; p < np; ++p) {
        for (int k = 0; k < 32; ++k) pass_sites[p * 32 + k] = k < (int)plan.sites[p].size() ? plan.sites[p][k] : -1;
        ncls[p] = plan.ncls[p];
        const SpinPassDev &d = plan.dev[p];
        for (size_t bi = 0; bi < plan.bonds[p].size(); ++bi) {
            const int bond = plan.bonds[p][bi];
            bond_pass[bond] = p;
        }
        for (int c = 0; c < plan.ncls[p]; ++c) {
            double *o = cls + ((size_t)p * SP_MAX_CLS + c) * row;
            o[0] = plan.cls_key[p][c][0];
            o[1] = plan.cls_key[p][c][1];
            o[2] = plan.cls_key[p][c][2];
            for (int e = 0; e < 2 * qq; ++e) o[3 + e] = d.cls_c[2 * (size_t)c * qq + e];
        }
    }
    return PNED_OK;
}

// Launches one pass on a grid of resident blocks (compute units x the blocks of this
// instance that fit on one at once, at most grid_cap and the tile count), so every block
// lives for the whole pass and walks its tiles with the next tile's loads in flight; the
// XCD-local tile order needs the grid to be a multiple of 8.  Spin-1 N=18: 256 blocks (one
// per CU) 19.38-19.40 ms per H.v against 20.27 with 2048, 20.06 with 1024, 19.87-20.03 with
// 512 and 24.4 with 384 (a second, partial round of blocks).  Returns the grid used.
template <int Q, int TPB, int R>
static int launch_spin_pass(pned_ctx *ctx, int grid_cap, size_t lds, bool cplx, const pned_vec *v1, pned_vec *v0,
                            const SpinPassDev *pd, SpinPassArgs a, double *part)
{
    // slots per sweep: only the last sweep may be partial (the instances' tiles fill all but
    // the last slot), and 7 slots of 1024 threads keep 2 (real) or 1 (complex coefficients)
    // so that 128 VGPRs hold every slot's state
    constexpr int S = (R == 7) ? 2 : (R == 4 && Q == 4) ? 4 : (R > 1) ? R - 1 : 1;
    constexpr int SC = (R == 7) ? 1 : S;
    auto k = cplx ? hv_spin_pass<Q, TPB, R, SC, true> : hv_spin_pass<Q, TPB, R, S, false>;
    (void)hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (ctx->num_cu <= 0 &&
        hipDeviceGetAttribute(&ctx->num_cu, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess)
        ctx->num_cu = 0;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k, TPB, lds) != hipSuccess) per_cu = 0;
    uint64_t want = (ctx->num_cu > 0 && per_cu > 0) ? (uint64_t)ctx->num_cu * per_cu : (uint64_t)grid_cap;
    want = std::min<uint64_t>(std::min<uint64_t>(want, (uint64_t)grid_cap), a.tile_count);
    if (want >= 8) want &= ~(uint64_t)7;
    const int grid = (int)std::max<uint64_t>(want, 1);
    a.xcd = (grid % 8 == 0) ? 1 : 0;
    hipLaunchKernelGGL(k, dim3(grid), dim3(TPB), lds, ctx->stream, (const double2 *)v1->data, (double2 *)v0->data,
                       pd, a, part);
    return grid;
}

int launch_hv_spin_tiled(pned_ctx *ctx, const pned_vec *v1, pned_vec *v0, double eperbd, bool fresh,
                         double *prdct_host, double xs, double ys)
{
    return launch_hv_spin_tiled_shift(ctx, v1, v0, eperbd, fresh, prdct_host, xs, ys, 0.0, 0.0, false);
}

int launch_hv_spin_tiled_shift(pned_ctx *ctx, const pned_vec *v1, pned_vec *v0, double eperbd, bool fresh,
                               double *prdct_host, double xs, double ys, double zr, double zi, bool dotu)
{
    const int q = ctx->q, qq = q * q;
    if (ctx->spin_pd == nullptr) {
        int K = spin_kmax(ctx->n, q);
        if (ctx->tune_tile_bits > 0 && ctx->tune_tile_bits < K) K = ctx->tune_tile_bits;
        SpinPlan plan;
        PNED_TRY(spin_make_plan(ctx->n, q, ctx->ibond, ctx->site_pair, ctx->btype, ctx->xl, ctx->ll, K,
                                ctx->tune_low_bits, plan));
        const int np = (int)plan.dev.size();
        PNED_HIP(hipStreamSynchronize(ctx->stream));
        PNED_HIP(hipMalloc(&ctx->spin_pd, sizeof(SpinPassDev) * np));
        PNED_HIP(hipMemcpy(ctx->spin_pd, plan.dev.data(), sizeof(SpinPassDev) * np, hipMemcpyHostToDevice));
        ctx->spin_K = plan.K;
        ctx->spin_np = np;
        ctx->spin_nb.clear();
        for (const auto &b : plan.bonds) ctx->spin_nb.push_back((int)b.size());
        ctx->spin_tsz = plan.tsz;
        ctx->spin_nlo = plan.nlo;
        ctx->spin_ncls = plan.ncls;
        ctx->spin_cplx = plan.cplx;
    }
    const int np = ctx->spin_np;
    const int grid_cap = 2048;
    PNED_TRY(ensure_partial(ctx, 2 * (size_t)grid_cap * np + 16));
    PNED_TRY(ensure_result(ctx, 8));
    int rows = 0;
    for (int p = 0; p < np; ++p) {
        SpinPassArgs a{};
        a.tsz = ctx->spin_tsz[p];
        a.tile_count = (uint64_t)(ctx->dim / a.tsz);
        a.K = 0;
        for (int t = 1; t < a.tsz; t *= q) ++a.K;
        a.nbond = ctx->spin_nb[p];
        a.ncls = ctx->spin_ncls[p];
        a.diag_first = (p == 0);
        a.fresh = fresh ? 1 : 0;
        a.nlo_mod = ctx->spin_nlo[p];
        a.shift = -eperbd;
        a.xs = xs;
        a.ys = ys;
        a.zr = zr;
        a.zi = zi;
        a.dotu = dotu ? 1 : 0;
        int grid = 0;
        const SpinPassDev *pd = reinterpret_cast<const SpinPassDev *>(ctx->spin_pd) + p;
        double *part = ctx->d_partial + 2 * (size_t)rows;
        // 1024 threads for tiles above 2048 states, else 256; R = ceil(tsz / threads) slots
        const int tpb = (a.tsz > 2048) ? 1024 : 256;
        const int R = (a.tsz + tpb - 1) / tpb;
        const bool cplx = ctx->spin_cplx[p] != 0;
        const size_t lds = spin_lds_bytes(a.tsz, a.ncls, qq, cplx, tpb);
#define PNED_SPIN_CASE(QV, TV, RV)                                                                                   \
        if (q == QV && tpb == TV && R == RV) {                                                                      \
            grid = launch_spin_pass<QV, TV, RV>(ctx, grid_cap, lds, cplx, v1, v0, pd, a, part);                     \
        } else
        // every tile size q^K (q^K <= 8192) maps to one of these
        PNED_SPIN_CASE(3, 1024, 7) PNED_SPIN_CASE(3, 1024, 3) PNED_SPIN_CASE(3, 256, 3) PNED_SPIN_CASE(3, 256, 1)
        PNED_SPIN_CASE(4, 1024, 4) PNED_SPIN_CASE(4, 256, 4) PNED_SPIN_CASE(4, 256, 1)
        PNED_SPIN_CASE(5, 1024, 4) PNED_SPIN_CASE(5, 256, 3) PNED_SPIN_CASE(5, 256, 1)
        { return fail(PNED_E_UNSUPPORTED, "no tiled spin-S kernel instance for this q / tile"); }
#undef PNED_SPIN_CASE
        PNED_HIP(hipGetLastError());
        rows += grid;
    }
    PNED_TRY(reduce_rows_result(ctx, rows, 2));
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

void free_spin_plan(pned_ctx *ctx)
{
    if (ctx->spin_pd) (void)hipFree(ctx->spin_pd);
    ctx->spin_pd = nullptr;
    ctx->spin_K = -1;
}

}  // namespace pned
