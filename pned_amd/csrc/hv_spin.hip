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
#include <cstring>

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
// Tiled spin-S H.v.  A pass owns K tile sites (the lowest L digits always among them, so a
// wave's lanes read contiguous runs); every bond with both sites in the tile is applied in
// that pass from the tile staged in LDS.  The tile is the q^K states sharing the digits
// of the other sites; tile-local index j = sum digit_k q^k over the tile sites in order.
// Each pass reads v1 once and read-modify-writes v0 once (pass 0 may start from zero).
//
// Per pass the host lists, for every (bond, output local state), the nonzero elements as
// (partner offset in tile units, coefficient); the kernel stages that list in LDS once per
// block.  A thread owns R tile slots j = tid + TPB r: their digit words and memory offsets
// are computed once per launch and stay in registers, the next tile's v1 is prefetched
// into registers while the current one is computed, and the slots are accumulated in
// sweeps of S so only S accumulators and v0 values are live.
// ------------------------------------------------------------------------------------
constexpr int SP_MAX_TILE = 8192;     // states per tile (128 KB of complex128)
constexpr int SP_MAX_BOND = 64;
constexpr int SP_BASE_TAB = 4096;
constexpr int SP_MAX_QQ = 25;         // q <= 5
constexpr int SP_MAX_ENT = 4096;      // nonzero entries per pass

struct SpinPassDev {
    uint64_t off[SP_MAX_TILE];        // memory offset of tile-local state j
    uint64_t blo[SP_BASE_TAB];        // tile base, low outer digits
    uint64_t bhi[SP_BASE_TAB];        // tile base, high outer digits
    int32_t bond_pa[SP_MAX_BOND];     // tile position of the min site
    int32_t bond_pb[SP_MAX_BOND];     // tile position of the max site
    int32_t row_start[SP_MAX_BOND * SP_MAX_QQ + 1];   // entries of (bond, istat) row
    int32_t ent_delta[SP_MAX_ENT];    // partner - j in tile units (0 = diagonal)
    double ent_c[2 * SP_MAX_ENT];     // coefficient (re, im)
};

struct SpinPassArgs {
    uint64_t tile_count;
    int tsz;                          // q^K
    int K, nbond, diag_first, fresh;
    int nent;                         // entries of this pass
    int nlo_mod;                      // tile_count split: tix % nlo_mod, tix / nlo_mod
    double shift;                     // -eperbd per bond diagonal
    double xs;                        // v0 += xs (H v1); <v1|.> then carries one factor xs
    double ys;                        // first pass: factor on the v0 input
    double zr, zi;                    // first pass: v0 += (zr + i zi) v1 (complex diagonal shift)
    int dotu;                         // 1: accumulate sum v1 * (H v1) without the conjugate
};

template <int Q, int TPB, int R, int S>
__global__ __launch_bounds__(TPB) void hv_spin_pass(const double2 *__restrict__ v1, double2 *__restrict__ v0,
                                                    const SpinPassDev *__restrict__ pd, SpinPassArgs a,
                                                    double *__restrict__ partial)
{
    constexpr int QQ = Q * Q;
    extern __shared__ unsigned char spin_smem[];
    double2 *tile = reinterpret_cast<double2 *>(spin_smem);
    // staged row table: starts, deltas, coefficients
    const int nrow = a.nbond * QQ;
    int32_t *rs = reinterpret_cast<int32_t *>(spin_smem + sizeof(double2) * (size_t)a.tsz);
    int32_t *dl = rs + ((nrow + 1 + 3) & ~3);
    double2 *cf = reinterpret_cast<double2 *>(dl + ((a.nent + 3) & ~3));
    const int tid = threadIdx.x;
    for (int i = tid; i <= nrow; i += TPB) rs[i] = pd->row_start[i];
    for (int i = tid; i < a.nent; i += TPB) {
        dl[i] = pd->ent_delta[i];
        cf[i] = make_double2(pd->ent_c[2 * i], pd->ent_c[2 * i + 1]);
    }
    const bool rd0 = !(a.diag_first && a.fresh);
    const double ysc = a.diag_first ? a.ys : 1.0;
    // per-slot digit words (4 bits per tile digit) and memory offsets, fixed for the launch
    uint32_t dw[R];
    uint64_t offj[R];
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
    }
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
    uint64_t tix = blockIdx.x;
    if (tix < a.tile_count) fetch_x(tile_base(tix));
    __syncthreads();   // staged table visible
    for (; tix < a.tile_count; tix += gridDim.x) {
        const uint64_t base = tile_base(tix);
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (tid + TPB * r < a.tsz) tile[tid + TPB * r] = make_double2(xr[r], xi[r]);
        __syncthreads();
        const uint64_t nx = tix + gridDim.x;
        // v0 of the first sweep before the next tile's v1 (in-order vector memory counters)
        double yr[S], yi[S];
        auto fetch_y = [&](int r0) {
#pragma unroll
            for (int u = 0; u < S; ++u) {
                const int r = r0 + u;
                yr[u] = 0.0;
                yi[u] = 0.0;
                if (rd0 && tid + TPB * r < a.tsz) {
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
#pragma unroll
            for (int u = 0; u < S; ++u) {
                const int r = r0 + u;
                const int j = tid + TPB * r;
                if (j >= a.tsz) continue;
                double ar = 0.0, ai = 0.0;
                for (int b = 0; b < a.nbond; ++b) {
                    const int d1 = (int)((dw[r] >> (4 * pd->bond_pa[b])) & 15u);
                    const int d2 = (int)((dw[r] >> (4 * pd->bond_pb[b])) & 15u);
                    const int row = b * QQ + Q * d2 + d1;
                    const int e1 = rs[row + 1];
                    for (int e = rs[row]; e < e1; ++e) {
                        const int d = dl[e];
                        const double2 c = cf[e];
                        const double cr = (d == 0) ? c.x + a.shift : c.x;
                        const double2 p = tile[j + d];
                        ar = fma(cr, p.x, fma(-c.y, p.y, ar));
                        ai = fma(cr, p.y, fma(c.y, p.x, ai));
                    }
                }
                const double2 x = tile[j];
                if (a.diag_first) {
                    ar = fma(a.zr, x.x, fma(-a.zi, x.y, ar));
                    ai = fma(a.zr, x.y, fma(a.zi, x.x, ai));
                }
                ar *= a.xs;
                ai *= a.xs;
                if (a.dotu) {
                    pr_re += x.x * ar - x.y * ai;
                    pr_im += x.x * ai + x.y * ar;
                } else {
                    pr_re += x.x * ar + x.y * ai;
                    pr_im += x.x * ai - x.y * ar;
                }
                v0[base + offj[r]] = make_double2(fma(ysc, yr[u], ar), fma(ysc, yi[u], ai));
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

// greedy cover of the bonds by tile-site sets (lowest L sites forced)
static std::vector<std::vector<int>> spin_cover(int n, const std::vector<int32_t> &sp, int ibond, int K, int L)
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
        std::vector<int> bonds;
        for (int b = 0; b < ibond; ++b)
            if (!done[b] && in[sp[2 * b]] && in[sp[2 * b + 1]]) { done[b] = 1; --remaining; bonds.push_back(b); }
        if (bonds.empty()) return {};
        std::vector<int> sites;
        for (int st = 0; st < n; ++st)
            if (in[st]) sites.push_back(st);
        out.push_back(sites);
        out.push_back(bonds);
    }
    return out;
}

// Spreads the bonds over the passes whose tiles hold both of their sites: bonds with the
// fewest covering passes are placed first, each on the pass with the fewest nonzero
// elements so far (w[b]: elements of bond b's matrix that are not skipped), so that no pass
// carries many more (bond, state) elements than the others. Passes left without bonds are
// dropped.
static void spin_balance(std::vector<std::vector<int>> &cov, const std::vector<int32_t> &sp, int n,
                         const std::vector<int> &w)
{
    const int np = (int)cov.size() / 2;
    std::vector<std::vector<char>> has(np, std::vector<char>(n, 0));
    std::vector<int> all;
    for (int p = 0; p < np; ++p) {
        for (int st : cov[2 * p]) has[p][st] = 1;
        for (int b : cov[2 * p + 1]) all.push_back(b);
    }
    std::vector<std::vector<int>> choice(all.size());
    for (size_t i = 0; i < all.size(); ++i)
        for (int p = 0; p < np; ++p)
            if (has[p][sp[2 * all[i]]] && has[p][sp[2 * all[i] + 1]]) choice[i].push_back(p);
    std::vector<size_t> order(all.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t a, size_t b) { return choice[a].size() < choice[b].size(); });
    std::vector<std::vector<int>> bonds(np);
    std::vector<long> load(np, 0);
    for (size_t i : order) {
        int best = choice[i][0];
        for (int p : choice[i])
            if (load[p] < load[best]) best = p;
        bonds[best].push_back(all[i]);
        load[best] += w[all[i]];
    }
    std::vector<std::vector<int>> out;
    for (int p = 0; p < np; ++p) {
        if (bonds[p].empty()) continue;
        std::sort(bonds[p].begin(), bonds[p].end());
        out.push_back(cov[2 * p]);
        out.push_back(bonds[p]);
    }
    cov.swap(out);
}

// LDS bytes of one pass: tile, then the row table (starts, deltas, coefficients)
static size_t spin_lds_bytes(int tsz, int nrow, int nent, int tpb)
{
    const size_t t = sizeof(double2) * (size_t)tsz + sizeof(int32_t) * (size_t)((nrow + 1 + 3) & ~3) +
                     sizeof(int32_t) * (size_t)((nent + 3) & ~3) + sizeof(double2) * (size_t)nent;
    return std::max(t, sizeof(double) * 2 * (size_t)tpb);
}

template <int Q, int TPB, int R>
static void launch_spin_pass(pned_ctx *ctx, int grid, size_t lds, const pned_vec *v1, pned_vec *v0,
                             const SpinPassDev *pd, const SpinPassArgs &a, double *part)
{
    constexpr int S = (R < 4) ? R : 4;
    auto k = hv_spin_pass<Q, TPB, R, S>;
    (void)hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(k, dim3(grid), dim3(TPB), lds, ctx->stream, (const double2 *)v1->data, (double2 *)v0->data,
                       pd, a, part);
}

int launch_hv_spin_tiled(pned_ctx *ctx, const pned_vec *v1, pned_vec *v0, double eperbd, bool fresh,
                         double *prdct_host, double xs, double ys)
{
    return launch_hv_spin_tiled_shift(ctx, v1, v0, eperbd, fresh, prdct_host, xs, ys, 0.0, 0.0, false);
}

int launch_hv_spin_tiled_shift(pned_ctx *ctx, const pned_vec *v1, pned_vec *v0, double eperbd, bool fresh,
                               double *prdct_host, double xs, double ys, double zr, double zi, bool dotu)
{
    const int q = ctx->q, n = ctx->n, qq = q * q;
    const int L0 = (ctx->tune_low_bits > 0) ? ctx->tune_low_bits : 2;
    if (ctx->spin_pd == nullptr) {
        // tile sites: the largest K with q^K <= 8192 whose passes fit LDS with their tables
        int K = 0;
        {
            int64_t p = 1;
            while (K < n && p * q <= SP_MAX_TILE) { p *= q; ++K; }
        }
        if (ctx->tune_tile_bits > 0 && ctx->tune_tile_bits < K) K = ctx->tune_tile_bits;
        std::vector<SpinPassDev> host;
        std::vector<int> nb, tsz, nlo, nent;
        for (; K >= 1; --K) {
            const int L = std::min(L0, K);
            auto cov = spin_cover(n, ctx->site_pair, ctx->ibond, K, L);
            if (cov.empty()) return fail(PNED_E_UNSUPPORTED, "spin-S tile planner made no progress");
            {
                std::vector<int> w(ctx->ibond, 0);
                for (int b = 0; b < ctx->ibond; ++b) {
                    const size_t t = (size_t)(ctx->btype[b] - 1);
                    for (int e = 0; e < qq * qq; ++e) w[b] += ctx->ll[(size_t)e + (size_t)qq * qq * t] ? 0 : 1;
                }
                spin_balance(cov, ctx->site_pair, n, w);
            }
            const int np = (int)cov.size() / 2;
            host.assign(np, SpinPassDev{});
            nb.assign(np, 0);
            tsz.assign(np, 0);
            nlo.assign(np, 0);
            nent.assign(np, 0);
            bool fits = true;
            std::vector<uint64_t> pw(n);
            {
                uint64_t x = 1;
                for (int k = 0; k < n; ++k) { pw[k] = x; x *= (uint64_t)q; }
            }
            for (int p = 0; p < np && fits; ++p) {
                const std::vector<int> &sites = cov[2 * p], &bonds = cov[2 * p + 1];
                SpinPassDev &d = host[p];
                if ((int)bonds.size() > SP_MAX_BOND) return fail(PNED_E_UNSUPPORTED, "too many bonds in a spin-S pass");
                const int Kp = (int)sites.size();
                int ts = 1;
                for (int k = 0; k < Kp; ++k) ts *= q;
                for (int j = 0; j < ts; ++j) {
                    uint64_t o = 0;
                    int y = j;
                    for (int k = 0; k < Kp; ++k) { o += (uint64_t)(y % q) * pw[sites[k]]; y /= q; }
                    d.off[j] = o;
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
                // (bond, output local state) rows of nonzero elements, as the reference's
                // H_dot_v visits them (llocal_Ham skips), partners as tile offsets
                int ne = 0;
                for (size_t bi = 0; bi < bonds.size(); ++bi) {
                    const int bond = bonds[bi];
                    const int sa = ctx->site_pair[2 * bond], sb = ctx->site_pair[2 * bond + 1];
                    const int pa = (int)(std::find(sites.begin(), sites.end(), sa) - sites.begin());
                    const int pbp = (int)(std::find(sites.begin(), sites.end(), sb) - sites.begin());
                    int wa = 1, wb = 1;
                    for (int k = 0; k < pa; ++k) wa *= q;
                    for (int k = 0; k < pbp; ++k) wb *= q;
                    d.bond_pa[bi] = pa;
                    d.bond_pb[bi] = pbp;
                    const int t = ctx->btype[bond] - 1;
                    for (int out = 0; out < qq; ++out) {
                        d.row_start[bi * qq + out] = ne;
                        const int d1 = out % q, d2 = out / q;
                        for (int in = 0; in < qq; ++in) {
                            const size_t f = (size_t)in + (size_t)qq * out + (size_t)qq * qq * t;
                            if (ctx->ll[f]) continue;
                            if (ne >= SP_MAX_ENT) return fail(PNED_E_UNSUPPORTED, "spin-S pass table too large");
                            d.ent_delta[ne] = (in % q - d1) * wa + (in / q - d2) * wb;
                            d.ent_c[2 * ne] = ctx->xl[2 * f];
                            d.ent_c[2 * ne + 1] = ctx->xl[2 * f + 1];
                            ++ne;
                        }
                    }
                }
                d.row_start[bonds.size() * qq] = ne;
                nb[p] = (int)bonds.size();
                tsz[p] = ts;
                nlo[p] = (int)lm;
                nent[p] = ne;
                fits = spin_lds_bytes(ts, (int)bonds.size() * qq, ne, 1024) <= 160 * 1024;
            }
            if (fits) break;
        }
        if (K < 1) return fail(PNED_E_UNSUPPORTED, "spin-S tables do not fit LDS");
        const int np = (int)host.size();
        PNED_HIP(hipStreamSynchronize(ctx->stream));
        PNED_HIP(hipMalloc(&ctx->spin_pd, sizeof(SpinPassDev) * np));
        PNED_HIP(hipMemcpy(ctx->spin_pd, host.data(), sizeof(SpinPassDev) * np, hipMemcpyHostToDevice));
        ctx->spin_K = K;
        ctx->spin_np = np;
        ctx->spin_nb = nb;
        ctx->spin_tsz = tsz;
        ctx->spin_nlo = nlo;
        ctx->spin_nent = nent;
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
        a.nent = ctx->spin_nent[p];
        a.diag_first = (p == 0);
        a.fresh = fresh ? 1 : 0;
        a.nlo_mod = ctx->spin_nlo[p];
        a.shift = -eperbd;
        a.xs = xs;
        a.ys = ys;
        a.zr = zr;
        a.zi = zi;
        a.dotu = dotu ? 1 : 0;
        const int grid = (int)std::min<uint64_t>(a.tile_count, (uint64_t)grid_cap);
        const SpinPassDev *pd = reinterpret_cast<const SpinPassDev *>(ctx->spin_pd) + p;
        double *part = ctx->d_partial + 2 * (size_t)rows;
        // 1024 threads for tiles above 2048 states, else 256; R slots per thread (power of 2)
        const int tpb = (a.tsz > 2048) ? 1024 : 256;
        int R = 1;
        while (R * tpb < a.tsz) R *= 2;
        const size_t lds = spin_lds_bytes(a.tsz, a.nbond * qq, a.nent, tpb);
#define PNED_SPIN_CASE(QV, TV, RV)                                                                                   \
        if (q == QV && tpb == TV && R == RV) {                                                                      \
            launch_spin_pass<QV, TV, RV>(ctx, grid, lds, v1, v0, pd, a, part);                                      \
        } else
        PNED_SPIN_CASE(3, 1024, 8) PNED_SPIN_CASE(3, 1024, 4) PNED_SPIN_CASE(3, 256, 8) PNED_SPIN_CASE(3, 256, 4)
        PNED_SPIN_CASE(3, 256, 2) PNED_SPIN_CASE(3, 256, 1)
        PNED_SPIN_CASE(4, 1024, 8) PNED_SPIN_CASE(4, 1024, 4) PNED_SPIN_CASE(4, 256, 8) PNED_SPIN_CASE(4, 256, 4)
        PNED_SPIN_CASE(4, 256, 2) PNED_SPIN_CASE(4, 256, 1)
        PNED_SPIN_CASE(5, 1024, 8) PNED_SPIN_CASE(5, 1024, 4) PNED_SPIN_CASE(5, 256, 8) PNED_SPIN_CASE(5, 256, 4)
        PNED_SPIN_CASE(5, 256, 2) PNED_SPIN_CASE(5, 256, 1)
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
