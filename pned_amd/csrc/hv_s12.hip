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
#include <cstring>
#include <mutex>
#include <utility>

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
    int nlo;               // outer bits resolved by the first base table
    int nop;               // >1: one operator per column (batch context)
    double shift;          // -ibond * eperbd, added to the diagonal
    const double *colcoef; // [devterm][nop][4] when nop > 1
    const double *coldiag; // [nop][ngroups+1][2] when nop > 1
    const double *colscale;// per-column factor on v1 (H (s v1)); null = 1
    const double *yscale;  // diagonal pass: per-column factor on the v0 input (v0 = y v0 + ...); null = 1
    int ncat[4];           // term-major kernel: terms per category (see build_tm_terms)
    int ngroups;           // diagonal groups (batch contexts stage their weights)
    // term-major kernel, last pass of a deferred-normalisation Lanczos step:
    //   1  only the <v1|H_pass v1> partials (no v0 traffic)
    //   2  v0 := w - Re(alpha) s v1 with w = v0 + H_pass (s v1), alpha = s * lz_prd[c],
    //      partials [|w - alpha s v1|^2, 0]
    int lz_mode;
    const double *lz_prd;
};

// Per-pass address tables in device memory (read with scalar loads):
//   lo[t]  memory-index offset of thread t's slot (tile bits 0..LT-1)
//   hi[r]  offset of register slot r (tile bits LT..K-1)
//   blo/bhi  tile base = blo[tix & (2^nlo - 1)] | bhi[tix >> nlo]  (outer bits)
constexpr int PD_LO = 1024, PD_HI = 32, PD_B = 2048;
struct PassDev {
    uint64_t lo[PD_LO];
    uint64_t hi[PD_HI];
    uint64_t blo[PD_B];
    uint64_t bhi[PD_B];
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

template <int... Is, typename F>
__device__ __forceinline__ void static_for_impl(std::integer_sequence<int, Is...>, F &&f)
{
    (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F &&f)
{
    static_for_impl(std::make_integer_sequence<int, N>{}, f);
}

// a wave-uniform double held in scalar registers
__device__ __forceinline__ double uniform(double v)
{
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)u);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(u >> 32));
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// TPB threads, R states per thread: tile of K = log2(TPB*R) bits.  Tile slot
// j = tid + TPB*r holds the state with memory index base | offLo(tid) | offHi(r).
// Software pipeline over the tiles a block owns: while tile t is computed from LDS, the
// v1 values of tile t+1 are in flight into registers (written to LDS after the barrier
// that ends tile t), and the v0 values being accumulated arrive one chunk of CH states
// ahead of their use, across tile boundaries.  HBM latency is hidden inside one
// workgroup, which matters when a 128 KB tile leaves room for one workgroup per CU.
template <typename Real, int TPB, int R>
__global__ __launch_bounds__(TPB) void hv_s12_pass(const typename C2<Real>::T *__restrict__ v1,
                                                   typename C2<Real>::T *__restrict__ v0,
                                                   const DevTerm *__restrict__ terms,
                                                   const DiagDev *__restrict__ diagp,
                                                   const PassDev *__restrict__ pd, PassArgs a,
                                                   double *__restrict__ partial)
{
    using V = typename C2<Real>::T;
    extern __shared__ unsigned char smem_raw[];
    V *tile = reinterpret_cast<V *>(smem_raw);
    const unsigned char *tbytes = smem_raw;  // partner slot j ^ m sits at byte (16 j) ^ (16 m)
    constexpr int CH = (R < 4) ? R : 4;      // states per compute chunk
    constexpr int NCH = R / CH;
    const int tid = threadIdx.x;

    const uint64_t offLo = pd->lo[tid];
    auto offHi = [&](int r) { return pd->hi[r]; };
    const bool rd0 = !(a.diag && a.fresh);
    const double ys = (a.diag && a.yscale) ? a.yscale[tid & (a.ncol - 1)] : 1.0;
    const uint64_t lomask = (1ull << a.nlo) - 1;
    auto tile_base = [&](uint64_t tix) { return pd->blo[tix & lomask] | pd->bhi[tix >> a.nlo] | offLo; };

    double pr_re = 0.0, pr_im = 0.0;
    // prefetch registers kept as scalar components (a V[] array is not promoted to
    // registers for the 16-byte type)
    Real xr_[R], xi_[R];
    Real ycr[CH], yci[CH];
    auto fetch_x = [&](uint64_t b) {
        static_for<R>([&](auto I) {
            const V t = v1[b | offHi(I)];
            xr_[I] = t.x;
            xi_[I] = t.y;
        });
    };
    uint64_t tix = blockIdx.x;
    if (tix < a.tile_count) {
        const uint64_t tb0 = tile_base(tix);
        fetch_x(tb0);
        if (rd0) static_for<CH>([&](auto I) { const V t = v0[tb0 | offHi(I)]; ycr[I] = t.x; yci[I] = t.y; });
    }
    for (; tix < a.tile_count; tix += gridDim.x) {
        const uint64_t tb0 = tile_base(tix);
        static_for<R>([&](auto I) {
            V t;
            t.x = xr_[I];
            t.y = xi_[I];
            tile[tid + TPB * I] = t;
        });
        __syncthreads();
        const uint64_t nx = tix + gridDim.x;
        const bool has_next = nx < a.tile_count;
        const uint64_t nb = has_next ? tile_base(nx) : 0;
        if (has_next) fetch_x(nb);

        static_for<NCH>([&](auto CI) {
            constexpr int r0 = CI * CH;
            // v0 of the next chunk (or of the next tile's first chunk)
            Real ynr[CH], yni[CH];
            if (rd0) {
                if constexpr (r0 + CH < R) {
                    static_for<CH>([&](auto I) {
                        const V t = v0[tb0 | offHi(r0 + CH + I)];
                        ynr[I] = t.x;
                        yni[I] = t.y;
                    });
                } else {
                    if (has_next)
                        static_for<CH>([&](auto I) {
                            const V t = v0[nb | offHi(I)];
                            ynr[I] = t.x;
                            yni[I] = t.y;
                        });
                }
            }
            uint32_t st[CH], jb[CH];
            double ar[CH], ai[CH];
            static_for<CH>([&](auto c) {
                jb[c] = (uint32_t)((tid + TPB * (r0 + c)) * sizeof(V));   // LDS byte offset of the slot
                st[c] = (uint32_t)((tb0 | offHi(r0 + c)) >> a.colbits);
                ar[c] = 0.0;
                ai[c] = 0.0;
            });
            if (a.diag) {
                const DiagDev &dg = *diagp;
                static_for<CH>([&](auto c) {
                    double di, dr;
                    if (a.nop > 1) {
                        // per-column weights; the thread's column is tid & (nop-1)
                        const double *w = a.coldiag + (size_t)(tid & (a.nop - 1)) * (dg.ngroups + 1) * 2;
                        dr = w[2 * dg.ngroups];
                        di = w[2 * dg.ngroups + 1];
                        for (int g = 0; g < dg.ngroups; ++g) {
                            const uint32_t f = (dg.kind[g] == 0) ? (st[c] & dg.mask[g])
                                                                 : ((st[c] ^ (st[c] >> dg.shift[g])) & dg.mask[g]);
                            const double cnt = (double)__popc(f);
                            dr = fma(cnt, w[2 * g], dr);
                            di = fma(cnt, w[2 * g + 1], di);
                        }
                        dr += a.shift;
                    } else {
                        dr = diag_value_re(dg, st[c], di) + a.shift;
                    }
                    const V x = tile[tid + TPB * (r0 + c)];
                    ar[c] = dr * (double)x.x - di * (double)x.y;
                    ai[c] = dr * (double)x.y + di * (double)x.x;
                });
            }
            if (a.nop > 1) {
                const double *cbase = a.colcoef + (size_t)(tid & (a.nop - 1)) * 4;
                for (int t = 0; t < a.nterm; ++t) {
                    const DevTerm T = terms[a.first_term + t];
                    const double *cc = cbase + (size_t)(a.first_term + t) * a.nop * 4;
                    const double c0r = cc[0], c0i = cc[1];
                    const double dr = cc[2] - c0r, di = cc[3] - c0i;
                    const uint32_t tmb = T.tmask * (uint32_t)sizeof(V);
                    static_for<CH>([&](auto c) {
                        const V p = *reinterpret_cast<const V *>(tbytes + (jb[c] ^ tmb));
                        const double sel = (double)(__popc(st[c] & T.selmask) & 1);
                        const double cr = fma(sel, dr, c0r), ci = fma(sel, di, c0i);
                        ar[c] = fma(cr, (double)p.x, ar[c]);
                        ar[c] = fma(-ci, (double)p.y, ar[c]);
                        ai[c] = fma(cr, (double)p.y, ai[c]);
                        ai[c] = fma(ci, (double)p.x, ai[c]);
                    });
                }
            } else
            for (int t = 0; t < a.nterm; ++t) {
                const DevTerm T = terms[a.first_term + t];
                if (T.is_real) {
                    const double d = T.c1r - T.c0r;
                    const uint32_t tmb = T.tmask * (uint32_t)sizeof(V);
                    static_for<CH>([&](auto c) {
                        const V p = *reinterpret_cast<const V *>(tbytes + (jb[c] ^ tmb));
                        const double sel = (double)(__popc(st[c] & T.selmask) & 1);
                        const double cf = fma(sel, d, T.c0r);
                        ar[c] = fma(cf, (double)p.x, ar[c]);
                        ai[c] = fma(cf, (double)p.y, ai[c]);
                    });
                } else {
                    const double dr = T.c1r - T.c0r, di = T.c1i - T.c0i;
                    const uint32_t tmb = T.tmask * (uint32_t)sizeof(V);
                    static_for<CH>([&](auto c) {
                        const V p = *reinterpret_cast<const V *>(tbytes + (jb[c] ^ tmb));
                        const double sel = (double)(__popc(st[c] & T.selmask) & 1);
                        const double cr = fma(sel, dr, T.c0r), ci = fma(sel, di, T.c0i);
                        ar[c] = fma(cr, (double)p.x, ar[c]);
                        ar[c] = fma(-ci, (double)p.y, ar[c]);
                        ai[c] = fma(cr, (double)p.y, ai[c]);
                        ai[c] = fma(ci, (double)p.x, ai[c]);
                    });
                }
            }
            if (a.colscale) {
                const double sc = a.colscale[tid & (a.ncol - 1)];
                static_for<CH>([&](auto c) {
                    ar[c] *= sc;
                    ai[c] *= sc;
                });
            }
            static_for<CH>([&](auto c) {
                const V x = tile[tid + TPB * (r0 + c)];
                pr_re += (double)x.x * ar[c] + (double)x.y * ai[c];
                pr_im += (double)x.x * ai[c] - (double)x.y * ar[c];
                V o;
                if (rd0) {
                    o.x = (Real)fma(ys, (double)ycr[c], ar[c]);
                    o.y = (Real)fma(ys, (double)yci[c], ai[c]);
                } else {
                    o.x = (Real)ar[c];
                    o.y = (Real)ai[c];
                }
                v0[tb0 | offHi(r0 + c)] = o;
            });
            if (rd0) static_for<CH>([&](auto I) { ycr[I] = ynr[I]; yci[I] = yni[I]; });
        });
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

// Term-major pass kernel.  Same tiles, tables and results as hv_s12_pass, but the loop
// runs over terms outside and over register slots of a thread inside, with one
// accumulator per slot.  For a term with flip mask m and selector mask s (tile slot
// j = tid + TPB r):
//   partner slot  j ^ m = (tid ^ m_lo) + TPB (r ^ m_hi)
//   selector      parity(tid & s_lo) ^ parity(outer & s_out) ^ parity(r & s_hi)
// The first two parts are fixed for the (thread, tile): they pick the two coefficients
// once per term per sweep; the register part is uniform across the wave.  A state-term
// then costs one address add, one LDS read and 2 (real) or 4 (complex) FMAs, plus a
// 64-bit select for terms whose selector has register bits.  The host sorts each pass's
// terms into these four categories so every category is one straight-line loop (the
// accumulators never pass through divergent merges).
//
// The R slots of a tile are accumulated in R/S sweeps of S slots, so only S accumulators
// and S v0 values are live next to the R prefetched v1 values of the next tile; a sweep's
// v0 values are in flight while its terms are applied.  Accumulates in Real (FP64 for complex128, FP32 for complex64
// storage); the <x|Hx> partials are FP64.
template <typename Real, int TPB, int R, int S>
__global__ __launch_bounds__(TPB) void hv_s12_tm(const typename C2<Real>::T *__restrict__ v1,
                                                 typename C2<Real>::T *__restrict__ v0,
                                                 const TmTerm *__restrict__ terms,
                                                 const DiagDev *__restrict__ diagp,
                                                 const PassDev *__restrict__ pd, PassArgs a,
                                                 double *__restrict__ partial)
{
    static_assert(R % S == 0, "sweeps must tile the register slots");
    using V = typename C2<Real>::T;
    using A = Real;
    constexpr int NS = R / S;
    extern __shared__ unsigned char smem_raw[];
    V *tile = reinterpret_cast<V *>(smem_raw);
    const int tid = threadIdx.x;
    const uint64_t offLo = pd->lo[tid];
    const bool rd0 = !(a.diag && a.fresh) && a.lz_mode != 1;
    const uint64_t lomask = (1ull << a.nlo) - 1;
    auto outer_base = [&](uint64_t tix) { return pd->blo[tix & lomask] | pd->bhi[tix >> a.nlo]; };
    constexpr uint32_t RSTRIDE = (uint32_t)(TPB * sizeof(V));   // bytes between register slots
    const uint32_t lane_b = (uint32_t)tid * (uint32_t)sizeof(V);

    // batch contexts: this pass's per-column coefficients [term][column][4] and, in the
    // diagonal pass, the per-column diagonal weights [column][groups+1][2], after the tile
    double *ccol = reinterpret_cast<double *>(smem_raw + sizeof(V) * TPB * R);
    double *cdiag = ccol + (a.nop > 1 ? (size_t)a.nterm * a.nop * 4 : 0);
    if (a.nop > 1) {
        const int n = a.nterm * a.nop * 4;
        const double *src = a.colcoef + (size_t)a.first_term * a.nop * 4;
        for (int i = tid; i < n; i += TPB) ccol[i] = src[i];
        if (a.diag)
            for (int i = tid; i < a.nop * (a.ngroups + 1) * 2; i += TPB) cdiag[i] = a.coldiag[i];
    }
    const A colsc = a.colscale ? (A)a.colscale[tid & (a.ncol - 1)] : (A)1;
    const A ys = (a.diag && a.yscale) ? (A)a.yscale[tid & (a.ncol - 1)] : (A)1;
    double lar = 0.0, lai = 0.0;
    if (a.lz_mode == 2) {
        lar = (double)colsc * a.lz_prd[2 * (tid & (a.ncol - 1))];
        lai = (double)colsc * a.lz_prd[2 * (tid & (a.ncol - 1)) + 1];
    }

    // the diagonal as a lane program (see DiagDev): group e in lane e, class weights in
    // lanes 48..63; one operator only (batch contexts weight per column)
    const int dprog = (a.diag && a.nop <= 1) ? diagp->nprog : 0;
    const int drest = dprog ? diagp->nrest : 0;   // split program: lanes dprog.. counted per tile
    const uint32_t dpa = dprog ? diagp->prog_a[tid & 63] : 0u;
    const uint32_t dpb = dprog ? diagp->prog_b[tid & 63] : 0u;

    double pr_re = 0.0, pr_im = 0.0;
    Real xr[R], xi[R];
    auto fetch_x = [&](uint64_t b) {
        static_for<R>([&](auto I) {
            const V t = v1[b | pd->hi[I]];
            xr[I] = t.x;
            xi[I] = t.y;
        });
    };
    auto fetch_y = [&](uint64_t b, int r0, Real *yr, Real *yi) {
        static_for<S>([&](auto I) {
            const V t = v0[b | pd->hi[r0 + I]];
            yr[I] = t.x;
            yi[I] = t.y;
        });
    };
    // each block takes a contiguous range of tiles: consecutive tiles of a block differ in the
    // lowest outer bits, so its runs follow each other in memory (measured at N=32: a few per
    // cent on the passes whose tiles hold few low memory bits, against tiles strided by the grid)
    const uint64_t per = (a.tile_count + gridDim.x - 1) / gridDim.x;
    uint64_t tix = (uint64_t)blockIdx.x * per;
    const uint64_t tend = (tix + per < a.tile_count) ? tix + per : a.tile_count;
    if (tix < tend) fetch_x(outer_base(tix) | offLo);
    for (; tix < tend; ++tix) {
        const uint64_t ob = outer_base(tix);
        const uint64_t tb0 = ob | offLo;
        static_for<R>([&](auto I) {
            V t;
            t.x = xr[I];
            t.y = xi[I];
            tile[tid + TPB * I] = t;
        });
        __syncthreads();
        // v0 of the first sweep goes out before the next tile's v1, so waiting for it does
        // not wait for the prefetch (vector memory counters retire in order)
        Real ycr[S], yci[S];
        if (rd0) fetch_y(tb0, 0, ycr, yci);
        const uint64_t nx = tix + 1;
        if (nx < tend) fetch_x(outer_base(nx) | offLo);
        // the diagonal's per-class counts over the bit positions that no register bit takes
        // part in: the same for every slot of this thread and tile (16 bits per class)
        uint64_t rcnt = 0;
        if (drest) {
            const uint32_t s0 = (uint32_t)(tb0 >> a.colbits);
            uint32_t acc = 0;
            for (int e = dprog; e < dprog + drest; ++e) {
                const uint32_t pa = (uint32_t)__builtin_amdgcn_readlane((int)dpa, e);
                const uint32_t m = (uint32_t)__builtin_amdgcn_readlane((int)dpb, e);
                acc += (uint32_t)__popc(((pa & 64u) ? (s0 ^ (s0 >> (pa & 63u))) : s0) & m);
                if (pa & 128u) {
                    rcnt |= (uint64_t)acc << (16u * ((pa >> 8) & 7u));
                    acc = 0u;
                }
            }
        }

        static_for<NS>([&](auto HI) {
            constexpr int r0 = HI * S;
            if constexpr (r0 > 0) {
                if (rd0) fetch_y(tb0, r0, ycr, yci);
            }
            A ar[S], ai[S];
            if (dprog) {
                // all S slots at once, group-outer: per group two readlanes, per slot a shift,
                // an xor and an and with a popcount; per weight class one convert and FMA pair
                // a slot's state is sb | hs[I] (disjoint bits), and x -> x ^ (x >> sh) is linear
                // over xor, so a group's bits of slot I are its bits of sb (per thread, formed
                // once per group) xor those of hs[I] (uniform, on the scalar unit)
                uint32_t cnt[S], hs[S];
                double dr[S], di[S];
                const double d0r = diagp->c0r + a.shift, d0i = diagp->c0i;
                const uint32_t sb = (uint32_t)(tb0 >> a.colbits);
                static_for<S>([&](auto I) {
                    hs[I] = (uint32_t)(pd->hi[r0 + I] >> a.colbits);
                    cnt[I] = 0u;
                    dr[I] = d0r;
                    di[I] = d0i;
                });
                for (int e = 0; e < dprog; ++e) {
                    const uint32_t pa = (uint32_t)__builtin_amdgcn_readlane((int)dpa, e);
                    const uint32_t m = (uint32_t)__builtin_amdgcn_readlane((int)dpb, e);
                    const uint32_t sh = pa & 63u;
                    if (pa & 64u) {
                        const uint32_t f = (sb ^ (sb >> sh)) & m;
                        static_for<S>([&](auto I) {
                            cnt[I] += (uint32_t)__popc(f ^ ((hs[I] ^ (hs[I] >> sh)) & m));
                        });
                    } else {
                        const uint32_t f = sb & m;
                        static_for<S>([&](auto I) { cnt[I] += (uint32_t)__popc(f ^ (hs[I] & m)); });
                    }
                    if (pa & 128u) {
                        const int c = (int)((pa >> 8) & 7u);
                        const uint32_t rc = (uint32_t)(rcnt >> (16 * c)) & 0xFFFFu;
                        const uint64_t wrb = (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)dpa, 48 + c) |
                                             ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)dpa, 56 + c) << 32);
                        const uint64_t wib = (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)dpb, 48 + c) |
                                             ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)dpb, 56 + c) << 32);
                        const double wr = __longlong_as_double((long long)wrb);
                        const double wi = __longlong_as_double((long long)wib);
                        static_for<S>([&](auto I) {
                            dr[I] = fma((double)(cnt[I] + rc), wr, dr[I]);
                            di[I] = fma((double)(cnt[I] + rc), wi, di[I]);
                            cnt[I] = 0u;
                        });
                    }
                }
                static_for<S>([&](auto I) {
                    const V x = tile[tid + TPB * (r0 + I)];
                    ar[I] = (A)(dr[I] * (double)x.x - di[I] * (double)x.y);
                    ai[I] = (A)(dr[I] * (double)x.y + di[I] * (double)x.x);
                });
            } else if (a.diag) {
                const DiagDev &dg = *diagp;
                static_for<S>([&](auto I) {
                    const uint32_t st = (uint32_t)((tb0 | pd->hi[r0 + I]) >> a.colbits);
                    double dr, di;
                    if (a.nop > 1) {
                        const double *w = cdiag + (size_t)(tid & (a.nop - 1)) * (dg.ngroups + 1) * 2;
                        dr = w[2 * dg.ngroups];
                        di = w[2 * dg.ngroups + 1];
                        for (int g = 0; g < dg.ngroups; ++g) {
                            const uint32_t f = (dg.kind[g] == 0) ? (st & dg.mask[g])
                                                                 : ((st ^ (st >> dg.shift[g])) & dg.mask[g]);
                            const double cnt = (double)__popc(f);
                            dr = fma(cnt, w[2 * g], dr);
                            di = fma(cnt, w[2 * g + 1], di);
                        }
                        dr += a.shift;
                    } else {
                        dr = diag_value_re(dg, st, di) + a.shift;
                    }
                    const V x = tile[tid + TPB * (r0 + I)];
                    ar[I] = (A)(dr * (double)x.x - di * (double)x.y);
                    ai[I] = (A)(dr * (double)x.y + di * (double)x.x);
                });
            } else {
                static_for<S>([&](auto I) {
                    ar[I] = (A)0;
                    ai[I] = (A)0;
                });
            }

            // terms come sorted by category: [real | real, register selector | complex |
            // complex, register selector]; one straight-line loop per category
            int t = a.first_term;
            auto prelude = [&](const TmTerm &T, A &car, A &cai, A &cbr, A &cbi) {
                const uint32_t par = (uint32_t)(__popc((uint32_t)tid & T.slo) + __popcll(ob & T.sout)) & 1u;
                A c0r, c0i, c1r, c1i;
                if (a.nop > 1) {
                    // per-column coefficients, staged in LDS for this pass
                    const double *cc =
                        ccol + ((size_t)(T.term - a.first_term) * a.nop + (size_t)(tid & (a.nop - 1))) * 4;
                    c0r = (A)cc[0];
                    c0i = (A)cc[1];
                    c1r = (A)cc[2];
                    c1i = (A)cc[3];
                } else {
                    // uniform: kept in scalar registers (a per-lane select of two loaded
                    // values would otherwise become a per-lane load)
                    c0r = (A)uniform(T.c0r);
                    c0i = (A)uniform(T.c0i);
                    c1r = (A)uniform(T.c1r);
                    c1i = (A)uniform(T.c1i);
                }
                car = par ? c1r : c0r;   // coefficient where the register part of the selector is even
                cai = par ? c1i : c0i;
                cbr = par ? c0r : c1r;   // ... and where it is odd
                cbi = par ? c0i : c1i;
                return smem_raw + (lane_b ^ (T.mlo * (uint32_t)sizeof(V)));
            };
            constexpr int G = (S < 4) ? S : 4;   // partner reads issued ahead of their FMAs
            auto run_cat = [&](int count, auto CXc, auto SELc) {
                constexpr bool CX = decltype(CXc)::value;
                constexpr bool SEL = decltype(SELc)::value;
                for (const int e = t + count; t < e; ++t) {
                    const TmTerm T = terms[t];
                    A car, cai, cbr, cbi;
                    const unsigned char *lb = prelude(T, car, cai, cbr, cbi);
                    static_for<S / G>([&](auto GI) {
                        Real qr[G], qi[G];
                        static_for<G>([&](auto k) {
                            constexpr int r = r0 + GI * G + k;
                            const V p = *reinterpret_cast<const V *>(lb + ((uint32_t)r ^ T.mhi) * RSTRIDE);
                            qr[k] = p.x;
                            qi[k] = p.y;
                        });
                        static_for<G>([&](auto k) {
                            constexpr int r = r0 + GI * G + k;
                            constexpr int u = GI * G + k;
                            A cr = car, ci = cai;
                            if constexpr (SEL) {
                                const bool f = __popc((uint32_t)r & T.shi) & 1;
                                cr = f ? cbr : car;
                                if constexpr (CX) ci = f ? cbi : cai;
                            }
                            if constexpr (CX) {
                                ar[u] = fma(cr, (A)qr[k], ar[u]);
                                ar[u] = fma(-ci, (A)qi[k], ar[u]);
                                ai[u] = fma(cr, (A)qi[k], ai[u]);
                                ai[u] = fma(ci, (A)qr[k], ai[u]);
                            } else {
                                ar[u] = fma(cr, (A)qr[k], ar[u]);
                                ai[u] = fma(cr, (A)qi[k], ai[u]);
                            }
                        });
                    });
                }
            };
            run_cat(a.ncat[0], std::false_type{}, std::false_type{});
            run_cat(a.ncat[1], std::false_type{}, std::true_type{});
            run_cat(a.ncat[2], std::true_type{}, std::false_type{});
            run_cat(a.ncat[3], std::true_type{}, std::true_type{});

            if (a.colscale) {
                static_for<S>([&](auto I) {
                    ar[I] *= colsc;
                    ai[I] *= colsc;
                });
            }
            static_for<S>([&](auto I) {
                const V x = tile[tid + TPB * (r0 + I)];
                if (a.lz_mode == 2) {
                    const double wr0 = (double)fma(ys, (A)ycr[I], ar[I]);
                    const double wi0 = (double)fma(ys, (A)yci[I], ai[I]);
                    const double xr_ = (double)x.x * (double)colsc, xi_ = (double)x.y * (double)colsc;
                    const double wr = wr0 - (lar * xr_ - lai * xi_);
                    const double wi = wi0 - (lar * xi_ + lai * xr_);
                    pr_re = fma(wr, wr, fma(wi, wi, pr_re));
                    V o;
                    o.x = (Real)(wr0 - lar * xr_);
                    o.y = (Real)(wi0 - lar * xi_);
                    v0[tb0 | pd->hi[r0 + I]] = o;
                } else {
                    const double hr = (double)ar[I], hi = (double)ai[I];
                    pr_re += (double)x.x * hr + (double)x.y * hi;
                    pr_im += (double)x.x * hi - (double)x.y * hr;
                    if (a.lz_mode != 1) {
                        V o;
                        if (rd0) {
                            o.x = (Real)fma(ys, (A)ycr[I], ar[I]);
                            o.y = (Real)fma(ys, (A)yci[I], ai[I]);
                        } else {
                            o.x = (Real)ar[I];
                            o.y = (Real)ai[I];
                        }
                        v0[tb0 | pd->hi[r0 + I]] = o;
                    }
                }
            });
        });
        __syncthreads();
    }

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

int reduce_rows_result(pned_ctx *ctx, int nrows, int ncol2)
{
    hipLaunchKernelGGL(reduce_partials, dim3(ncol2), dim3(256), 0, ctx->stream, ctx->d_partial, nrows, ncol2,
                       ctx->d_result, 0);
    PNED_HIP(hipGetLastError());
    return PNED_OK;
}

template <typename Real, int TPB, int R>
static int launch_pass(pned_ctx *ctx, const void *v1, void *v0, const DiagDev *ddiag, const PassDev *pd,
                       const PassArgs &a, double *partial, int grid)
{
    using V = typename C2<Real>::T;
    size_t lds = sizeof(V) * (size_t)TPB * R;
    size_t need_red = sizeof(double) * 2 * TPB;
    if (lds < need_red) lds = need_red;
    auto kern = hv_s12_pass<Real, TPB, R>;
    PNED_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(kern, dim3(grid), dim3(TPB), lds, ctx->stream, (const V *)v1, (V *)v0,
                       ctx->d_terms, ddiag, pd, a, partial);
    PNED_HIP(hipGetLastError());
    return PNED_OK;
}

template <typename Real>
static int dispatch_pass(pned_ctx *ctx, int TPB, int R, const void *v1, void *v0,
                         const DiagDev *ddiag, const PassDev *pd, const PassArgs &a, double *partial, int grid)
{
#define PNED_CASE(T_, R_) \
    if (TPB == T_ && R == R_) return launch_pass<Real, T_, R_>(ctx, v1, v0, ddiag, pd, a, partial, grid);
    PNED_CASE(256, 32) PNED_CASE(256, 16) PNED_CASE(256, 8) PNED_CASE(256, 4) PNED_CASE(256, 2)
    PNED_CASE(256, 1) PNED_CASE(128, 1) PNED_CASE(64, 1) PNED_CASE(32, 1) PNED_CASE(16, 1)
    PNED_CASE(512, 32) PNED_CASE(512, 16) PNED_CASE(512, 8) PNED_CASE(512, 4) PNED_CASE(512, 2)
#undef PNED_CASE
    return fail(PNED_E_UNSUPPORTED, "no H.v kernel instance for this tile shape");
}

// LDS the term-major kernel stages per-column coefficients in (batch contexts)
static size_t tm_coef_bytes(const PassArgs &a)
{
    if (a.nop <= 1) return 0;
    return sizeof(double) * ((size_t)4 * a.nterm * a.nop + (a.diag ? (size_t)2 * (a.ngroups + 1) * a.nop : 0));
}
constexpr size_t LDS_LIMIT = 160 * 1024;
// a batched tile of a single-operator context holds the columns of one 256-byte run (32
// complex64 or 16 complex128 values per state); more columns run in groups
constexpr size_t COL_GROUP_BYTES = 256;

template <typename Real, int TPB, int R, int S>
static int launch_tm(pned_ctx *ctx, const void *v1, void *v0, const TmTerm *tm, const DiagDev *ddiag,
                     const PassDev *pd, const PassArgs &a, double *partial, int grid)
{
    using V = typename C2<Real>::T;
    size_t lds = sizeof(V) * (size_t)TPB * R + tm_coef_bytes(a);
    size_t need_red = sizeof(double) * 2 * TPB;
    if (lds < need_red) lds = need_red;
    auto kern = hv_s12_tm<Real, TPB, R, S>;
    PNED_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(kern, dim3(grid), dim3(TPB), lds, ctx->stream, (const V *)v1, (V *)v0, tm, ddiag, pd, a,
                       partial);
    PNED_HIP(hipGetLastError());
    return PNED_OK;
}

// term-major instances: the prefetched v1 of R slots plus the accumulators and v0 of one
// sweep of S slots stay in registers; S is chosen so that nothing spills at the block's
// occupancy (1024 threads leave 128 registers per lane, 512 threads 256)
template <typename Real>
static bool has_tm(int TPB, int R)
{
    if (TPB != 256 && TPB != 512 && TPB != 1024) return false;
    if (sizeof(Real) == 8) return R == 1 || R == 2 || R == 4 || R == 8 || (R == 16 && TPB <= 512);
    return R == 1 || R == 2 || R == 4 || R == 8 || R == 16 || (R == 32 && TPB <= 512);
}

template <typename Real>
static int dispatch_tm(pned_ctx *ctx, int TPB, int R, const void *v1, void *v0, const TmTerm *tm,
                       const DiagDev *ddiag, const PassDev *pd, const PassArgs &a, double *partial, int grid)
{
#define PNED_CASE(T_, R_, S_) \
    if (TPB == T_ && R == R_) return launch_tm<Real, T_, R_, S_>(ctx, v1, v0, tm, ddiag, pd, a, partial, grid);
    if constexpr (sizeof(Real) == 8) {
        PNED_CASE(256, 16, 8) PNED_CASE(256, 8, 8) PNED_CASE(256, 4, 4) PNED_CASE(256, 2, 2) PNED_CASE(256, 1, 1)
        PNED_CASE(512, 16, 8) PNED_CASE(512, 8, 8) PNED_CASE(512, 4, 4) PNED_CASE(512, 2, 2) PNED_CASE(512, 1, 1)
        PNED_CASE(1024, 8, 4) PNED_CASE(1024, 4, 4) PNED_CASE(1024, 2, 2) PNED_CASE(1024, 1, 1)
    } else {
        PNED_CASE(256, 32, 16) PNED_CASE(256, 16, 16) PNED_CASE(256, 8, 8) PNED_CASE(256, 4, 4)
        PNED_CASE(256, 2, 2) PNED_CASE(256, 1, 1)
        PNED_CASE(512, 32, 16) PNED_CASE(512, 16, 16) PNED_CASE(512, 8, 8) PNED_CASE(512, 4, 4)
        PNED_CASE(512, 2, 2) PNED_CASE(512, 1, 1)
        PNED_CASE(1024, 16, 8) PNED_CASE(1024, 8, 8) PNED_CASE(1024, 4, 4) PNED_CASE(1024, 2, 2)
        PNED_CASE(1024, 1, 1)
    }
#undef PNED_CASE
    return fail(PNED_E_UNSUPPORTED, "no term-major H.v kernel instance for this tile shape");
}

// Tile-coordinate term table of one pass (see TmTerm), sorted into the term-major
// kernel's categories: 0 real, 1 real with register selector bits, 2 complex, 3 complex
// with register selector bits (batch contexts: every term complex).  ncat gets the counts.
// colbits: the plan's column bits; outer selector bits at or above them sit gap bits higher
// in memory (column groups)
static void build_tm_terms(const S12Plan &pl, const PassPlan &pp, int TPB, int colbits, int gap,
                           std::vector<TmTerm> &out, int ncat[4])
{
    const int LT = __builtin_ctz(TPB);
    uint64_t tilemask = 0;
    for (int b : pp.tb) tilemask |= 1ull << b;
    std::vector<TmTerm> cat[4];
    for (int i = 0; i < pp.nterm; ++i) {
        const DevTerm &d = pl.devterms[pp.first_term + i];
        TmTerm t{};
        t.mlo = d.tmask & (uint32_t)(TPB - 1);
        t.mhi = d.tmask >> LT;
        const uint64_t selm = (uint64_t)d.selmask << colbits;
        uint32_t stile = 0;
        for (int k = 0; k < pp.K; ++k)
            if (selm >> pp.tb[k] & 1) stile |= 1u << k;
        t.slo = stile & (uint32_t)(TPB - 1);
        t.shi = stile >> LT;
        const uint64_t so = selm & ~tilemask;
        t.sout = (so & ((1ull << colbits) - 1)) | ((so >> colbits) << (colbits + gap));
        const bool cx = !d.is_real || pl.nop > 1;
        t.flags = (cx ? 1u : 0u) | ((t.shi != 0) ? 2u : 0u);
        t.term = (uint32_t)(pp.first_term + i);
        t.c0r = d.c0r;
        t.c0i = d.c0i;
        t.c1r = d.c1r;
        t.c1i = d.c1i;
        cat[(cx ? 2 : 0) + (t.shi != 0 ? 1 : 0)].push_back(t);
    }
    for (int c = 0; c < 4; ++c) {
        ncat[c] = (int)cat[c].size();
        out.insert(out.end(), cat[c].begin(), cat[c].end());
    }
}

// ------------------------------------------------------------------------------------
// Scratch of the allocation-free H.v (pned_p_mltply_v3_noalloc): the per-block partial
// sums of <v1|H|v1> of every pass.  One array per GPU for the life of the process; its
// device address is looked up once per device when the first context on that device is
// created, and host code only ever passes that address (the array's name is a host-side
// stand-in).  Calls on one GPU share it, so overlapping calls on one GPU race.
// ------------------------------------------------------------------------------------
constexpr size_t NOALLOC_SCRATCH_DOUBLES = PNED_NOALLOC_SCRATCH_BYTES / sizeof(double);
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];
constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[NOALLOC_MAX_DEVICES];

int noalloc_scratch_init(int device)
{
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return fail(PNED_E_ARG, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return PNED_OK;
    int cur = 0;
    PNED_HIP(hipGetDevice(&cur));
    if (cur != device) PNED_HIP(hipSetDevice(device));
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) PNED_HIP(hipSetDevice(cur));
    if (e != hipSuccess || p == nullptr)
        return fail(PNED_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<double *>(p);
    return PNED_OK;
}

static double *noalloc_scratch(int device)
{
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_ptr[device];
}

// noalloc: the caller's plan, pass tables and buffers must already exist; partials go to
// part_ext (cap doubles) and the sums of <v1|H|v1> to result_ext (device memory).  Nothing
// is allocated and the stream is not synchronised.
static int hv_s12_run(pned_ctx *ctx, const pned_vec *v1, pned_vec *v0, double eperbd, bool fresh,
                      double *prdct_host, const double *d_colscale, const double *d_yscale, bool *lz_fused,
                      bool noalloc, double *part_ext, size_t part_cap, double *result_ext);

int launch_hv_s12(pned_ctx *ctx, const pned_vec *v1, pned_vec *v0, double eperbd, bool fresh,
                  double *prdct_host, const double *d_colscale, const double *d_yscale, bool *lz_fused)
{
    return hv_s12_run(ctx, v1, v0, eperbd, fresh, prdct_host, d_colscale, d_yscale, lz_fused, false, nullptr, 0,
                      nullptr);
}

static int hv_s12_run(pned_ctx *ctx, const pned_vec *v1, pned_vec *v0, double eperbd, bool fresh,
                      double *prdct_host, const double *d_colscale, const double *d_yscale, bool *lz_fused,
                      bool noalloc, double *part_ext, size_t part_cap, double *result_ext)
{
    if (lz_fused) *lz_fused = false;
    const int nvec = v1->nvec;
    int colbits = 0;
    while ((1 << colbits) < nvec) ++colbits;
    if ((1 << colbits) != nvec)
        return fail(PNED_E_UNSUPPORTED, "batched H.v needs a power-of-two column count");
    // A single-operator batch wider than one 256-byte run of columns runs in column groups: the
    // plan's tiles hold that run's column bits, the higher column bits are fixed per launch
    // (one launch per group and pass, base pointers moved by the group), so a tile spends its
    // other bits on sites and a pass covers more terms (N=24 with 64 columns: 4 -> 3 passes).
    // Plan positions at or above cg move up by gap bits in memory (expand).
    // Groups are taken only where they save a pass (ensure_s12_plan_cols).
    const int run_bits = __builtin_ctzll(COL_GROUP_BYTES / (v1->dtype == PNED_C128 ? 16 : 8));
    int cg = colbits;
    if (noalloc) {
        if (ctx->plan_req_colbits != colbits || ctx->plan_colbits < 0 || ctx->plan_dtype != v1->dtype ||
            ctx->pd_dev == nullptr || ctx->pd_gap != colbits - ctx->plan_colbits)
            return fail(PNED_E_ARG, "no H.v plan for this vector shape yet: call pned_p_mltply_v3 once with "
                                    "vectors of this shape first (that call allocates the plan's tables)");
        cg = ctx->plan_colbits;
    } else {
        PNED_TRY(ensure_s12_plan_cols(ctx, colbits, run_bits, v1->dtype, &cg));
    }
    const int gap = colbits - cg, ngrp = 1 << gap, ncol_t = 1 << cg;
    auto expand = [cg, gap](int b) { return b < cg ? b : b + gap; };
    const S12Plan &pl = ctx->s12;
    if (pl.nop > 1 && pl.nop != nvec)
        return fail(PNED_E_ARG, "a batch context's vectors need one column per operator");
    const int npass = (int)pl.passes.size();
    const int grid_cap = 4096;
    // partial layout: [pass][block][2*ncol] (plus one pass for a fused Lanczos step)
    if (noalloc) {
        const size_t need = (size_t)npass * grid_cap * 2 * nvec;
        if (need > part_cap)
            return fail(PNED_E_ARG, "allocation-free H.v: " + std::to_string(npass) + " passes x " +
                                        std::to_string(nvec) + " columns need " + std::to_string(need * 8) +
                                        " bytes of partial sums, more than the " +
                                        std::to_string(part_cap * 8) + "-byte static scratch");
    } else {
        size_t need = (size_t)(npass + 1) * grid_cap * 2 * nvec + 64;
        PNED_TRY(ensure_partial(ctx, need));
        PNED_TRY(ensure_result(ctx, 4 * (size_t)nvec + 8));
    }
    double *const part_base = noalloc ? part_ext : ctx->d_partial;
    // per-pass address tables (rebuilt when the plan or the thread count changes)
    const int TPBpref = (ctx->tune_threads == 256 || ctx->tune_threads == 512 || ctx->tune_threads == 1024)
                            ? ctx->tune_threads
                            : (ctx->s12_kernel == 1 && default_large(ctx, colbits) ? 1024
                                                                                    : (ctx->s12_kernel == 1 ? 256 : 512));
    std::vector<int> tpb(npass), rr(npass);
    for (int p = 0; p < npass; ++p) {
        int T = TPBpref;
        while (T > 16 && T > (1 << pl.passes[p].K)) T >>= 1;
        tpb[p] = T;
        rr[p] = (1 << pl.passes[p].K) / T;
        if (rr[p] > PD_HI) return fail(PNED_E_UNSUPPORTED, "tile too large for the thread count");
        if ((int)pl.passes[p].ob.size() > 22) return fail(PNED_E_UNSUPPORTED, "too many outer bits");
    }
    if (noalloc && ctx->pd_threads != TPBpref)
        return fail(PNED_E_ARG, "no H.v plan for this thread count yet: call pned_p_mltply_v3 once after "
                                "changing the tuning (that call rebuilds the plan's tables)");
    if (ctx->pd_threads != TPBpref || ctx->pd_gap != gap || ctx->pd_dev == nullptr) {
        std::vector<PassDev> host(npass);
        for (int p = 0; p < npass; ++p) {
            const PassPlan &pp = pl.passes[p];
            PassDev &d = host[p];
            std::memset(&d, 0, sizeof(PassDev));
            const int LT = __builtin_ctz(tpb[p]);
            for (int t = 0; t < tpb[p]; ++t) {
                uint64_t o = 0;
                for (int k = 0; k < LT; ++k)
                    if ((t >> k) & 1) o |= 1ull << expand(pp.tb[k]);
                d.lo[t] = o;
            }
            for (int r = 0; r < rr[p]; ++r) {
                uint64_t o = 0;
                for (int k = 0; LT + k < pp.K; ++k)
                    if ((r >> k) & 1) o |= 1ull << expand(pp.tb[LT + k]);
                d.hi[r] = o;
            }
            const int no = (int)pp.ob.size(), nlo = no < 11 ? no : 11, nhi = no - nlo;
            for (int x = 0; x < (1 << nlo); ++x) {
                uint64_t o = 0;
                for (int k = 0; k < nlo; ++k)
                    if ((x >> k) & 1) o |= 1ull << expand(pp.ob[k]);
                d.blo[x] = o;
            }
            for (int x = 0; x < (1 << nhi); ++x) {
                uint64_t o = 0;
                for (int k = 0; k < nhi; ++k)
                    if ((x >> k) & 1) o |= 1ull << expand(pp.ob[nlo + k]);
                d.bhi[x] = o;
            }
        }
        std::vector<TmTerm> tmh;
        ctx->tm_ncat.assign(4 * (size_t)npass, 0);
        for (int p = 0; p < npass; ++p)
            build_tm_terms(pl, pl.passes[p], tpb[p], cg, gap, tmh, &ctx->tm_ncat[4 * p]);
        if (tmh.empty()) tmh.push_back(TmTerm{});
        PNED_HIP(hipStreamSynchronize(ctx->stream));
        if (ctx->pd_dev) PNED_HIP(hipFree(ctx->pd_dev));
        ctx->pd_dev = nullptr;
        if (ctx->d_tm) PNED_HIP(hipFree(ctx->d_tm));
        ctx->d_tm = nullptr;
        PNED_HIP(hipMalloc(&ctx->pd_dev, sizeof(PassDev) * npass));
        PNED_HIP(hipMemcpy(ctx->pd_dev, host.data(), sizeof(PassDev) * npass, hipMemcpyHostToDevice));
        PNED_HIP(hipMalloc((void **)&ctx->d_tm, sizeof(TmTerm) * tmh.size()));
        PNED_HIP(hipMemcpy(ctx->d_tm, tmh.data(), sizeof(TmTerm) * tmh.size(), hipMemcpyHostToDevice));
        // the diagonal pass's lane program split by that pass's register bits
        if (ctx->d_diag_tm) PNED_HIP(hipFree(ctx->d_diag_tm));
        ctx->d_diag_tm = nullptr;
        for (int p = 0; p < npass && pl.nop <= 1 && ctx->diag_host.nprog > 0; ++p) {
            const PassPlan &pp = pl.passes[p];
            if (!pp.diag) continue;
            uint32_t slot_bits = 0;
            for (int k = __builtin_ctz(tpb[p]); k < pp.K; ++k) slot_bits |= 1u << (pp.tb[k] - cg);
            DiagDev dt;
            if (split_diag_slots(ctx->diag_host, slot_bits, dt)) {
                PNED_HIP(hipMalloc((void **)&ctx->d_diag_tm, sizeof(DiagDev)));
                PNED_HIP(hipMemcpy(ctx->d_diag_tm, &dt, sizeof(DiagDev), hipMemcpyHostToDevice));
            }
        }
        ctx->pd_threads = TPBpref;
        ctx->pd_gap = gap;
    }
    // blocks per pass: at least 8 contiguous tiles each (so a block's next-tile loads overlap
    // its terms for most of its tiles) while that leaves 256 or more, at most grid_cap.  At
    // N=24 (2048 tiles of 2^13) 256 blocks ran 0.524 ms per H.v against 0.532 with 1024 and
    // 0.554 with 2048; at N=32 4096 blocks of 128 tiles beat 1024 and 256 by about 0.5 ms
    auto pass_grid = [&](int p) {
        const uint64_t tc = 1ull << pl.passes[p].ob.size();
        const uint64_t want = std::max<uint64_t>(256, tc / 8);
        return (int)std::min<uint64_t>(std::min<uint64_t>(tc, want), (uint64_t)grid_cap);
    };
    // partial rows: per column group [pass][block] of 2 ncol_t values, the groups one after another
    int rows_grp = 0;
    for (int p = 0; p < npass; ++p) rows_grp += pass_grid(p);
    const bool c128 = v1->dtype == PNED_C128;
    const size_t esz = c128 ? 16 : 8;
    int total_rows = 0;
    for (int p = 0; p < npass; ++p) {
        const PassPlan &pp = pl.passes[p];
        PassArgs a{};
        a.K = pp.K;
        a.nouter = (int)pp.ob.size();
        a.tile_count = 1ull << a.nouter;
        a.nlo = a.nouter < 11 ? a.nouter : 11;
        a.colbits = colbits;
        a.first_term = pp.first_term;
        a.nterm = pp.nterm;
        a.diag = pp.diag ? 1 : 0;
        a.fresh = fresh ? 1 : 0;
        a.ncol = ncol_t;
        a.shift = -(double)ctx->ibond * eperbd;
        a.nop = pl.nop;
        a.colcoef = ctx->d_colcoef;
        a.coldiag = ctx->d_coldiag;
        a.colscale = d_colscale;
        a.yscale = d_yscale;
        for (int c = 0; c < 4; ++c) a.ncat[c] = ctx->tm_ncat[4 * p + c];
        a.ngroups = ctx->diag_host.ngroups;
        const int TPB = tpb[p], R = rr[p];
        const int grid = pass_grid(p);
        const PassDev *pdp = reinterpret_cast<const PassDev *>(ctx->pd_dev) + p;
        const size_t tile_bytes = esz * (size_t)TPB * R;
        const bool tm = ctx->s12_kernel == 1 && (c128 ? has_tm<double>(TPB, R) : has_tm<float>(TPB, R)) &&
                        tile_bytes + tm_coef_bytes(a) <= LDS_LIMIT;
        const DiagDev *dgp = (pp.diag && ctx->d_diag_tm) ? ctx->d_diag_tm : ctx->d_diag;
        if (!noalloc && lz_fused && p == npass - 1 && p > 0 && tm && c128 && nvec == 1 && pl.nop <= 1) {
            // deferred-normalisation Lanczos step: this pass's <v1|H v1> share first (no v0
            // traffic), alpha from all passes on the device, then the pass itself forms
            // v0 - Re(alpha) s v1 and the partials of beta^2
            double *part = part_base + (size_t)total_rows * 2 * nvec;
            a.lz_mode = 1;
            PNED_TRY(dispatch_tm<double>(ctx, TPB, R, v1->data, v0->data, ctx->d_tm, dgp, pdp, a, part, grid));
            total_rows += grid;
            hipLaunchKernelGGL(reduce_partials, dim3(2 * nvec), dim3(256), 0, ctx->stream, ctx->d_partial, total_rows,
                               2 * nvec, ctx->d_result, 0);
            PNED_HIP(hipGetLastError());
            a.lz_mode = 2;
            a.lz_prd = ctx->d_result;
            part = ctx->d_partial + (size_t)total_rows * 2 * nvec;
            PNED_TRY(dispatch_tm<double>(ctx, TPB, R, v1->data, v0->data, ctx->d_tm, dgp, pdp, a, part, grid));
            hipLaunchKernelGGL(reduce_partials, dim3(2 * nvec), dim3(256), 0, ctx->stream, part, grid, 2 * nvec,
                               ctx->d_result + 2 * nvec, 0);
            PNED_HIP(hipGetLastError());
            *lz_fused = true;
            return PNED_OK;
        }
        for (int g = 0; g < ngrp; ++g) {
            // column group g: columns g*ncol_t .. (g+1)*ncol_t-1 of every state
            const size_t eo = (size_t)g * ncol_t * esz;
            const void *x = static_cast<const char *>(v1->data) + eo;
            void *y = static_cast<char *>(v0->data) + eo;
            a.colscale = d_colscale ? d_colscale + (size_t)g * ncol_t : nullptr;
            a.yscale = d_yscale ? d_yscale + (size_t)g * ncol_t : nullptr;
            double *part = part_base + ((size_t)g * rows_grp + total_rows) * 2 * ncol_t;
            int rc;
            if (tm) {
                // the term table is in pass order, so the kernel indexes it by first_term as well
                if (c128)
                    rc = dispatch_tm<double>(ctx, TPB, R, x, y, ctx->d_tm, dgp, pdp, a, part, grid);
                else
                    rc = dispatch_tm<float>(ctx, TPB, R, x, y, ctx->d_tm, dgp, pdp, a, part, grid);
            } else if (c128) {
                rc = dispatch_pass<double>(ctx, TPB, R, x, y, ctx->d_diag, pdp, a, part, grid);
            } else {
                rc = dispatch_pass<float>(ctx, TPB, R, x, y, ctx->d_diag, pdp, a, part, grid);
            }
            if (rc != PNED_OK) return rc;
        }
        total_rows += grid;
    }
    const int ncol2 = 2 * nvec;
    double *const out = noalloc ? result_ext : ctx->d_result;
    for (int g = 0; g < ngrp; ++g) {
        hipLaunchKernelGGL(reduce_partials, dim3(2 * ncol_t), dim3(256), 0, ctx->stream,
                           part_base + (size_t)g * rows_grp * 2 * ncol_t, rows_grp, 2 * ncol_t,
                           out + (size_t)g * 2 * ncol_t, 0);
        PNED_HIP(hipGetLastError());
    }
    if (noalloc) return PNED_OK;
    if (prdct_host) {
        PNED_HIP(hipMemcpyAsync(ctx->h_result, ctx->d_result, sizeof(double) * ncol2,
                                hipMemcpyDeviceToHost, ctx->stream));
        PNED_HIP(hipStreamSynchronize(ctx->stream));
        for (int c = 0; c < ncol2; ++c) prdct_host[c] = ctx->h_result[c];
    }
    return PNED_OK;
}

}  // namespace pned

extern "C" int pned_p_mltply_v3_noalloc(pned_ctx *ctx, const pned_vec *v1, pned_vec *v0, double eperbd,
                                        double *prdct_dev)
{
    using namespace pned;
    PNED_TRY(check_hv_pair(ctx, v1, v0));
    if (!prdct_dev) return fail(PNED_E_ARG, "null prdct_dev");
    if (ctx->q != 2 || ctx->sector_k >= 0)
        return fail(PNED_E_UNSUPPORTED, "the allocation-free H.v covers S=1/2 full-space contexts only");
    double *part = noalloc_scratch(ctx->device);
    if (!part) return fail(PNED_E_HIP, "allocation-free H.v: no scratch address for this device");
    PNED_HIP(hipSetDevice(ctx->device));
    return hv_s12_run(ctx, v1, v0, eperbd, false, nullptr, nullptr, nullptr, nullptr, true, part,
                      NOALLOC_SCRATCH_DOUBLES, prdct_dev);
}
