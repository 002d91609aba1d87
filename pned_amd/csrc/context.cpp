// context.cpp -- contexts, vectors, the S=1/2 term planner and the hot-path ABI entries.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <set>
#include <tuple>

#include "internal.h"

namespace pned {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }
int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

void reset_spin_tables(pned_ctx *ctx);
void free_spin_plan(pned_ctx *ctx);

// ------------------------------------------------------------------------------------
// S=1/2 decomposition.  For every bond (sites a < b, local index t = 2*s_b + s_a) the
// reference applies v0(t) += h(in, t) v1(in).  Each flip pattern f = in ^ t (1: flip a,
// 2: flip b, 3: both) gives a coefficient function g_f(t) of the OUTPUT local state,
// which is expanded in the Walsh basis {1, sigma_a, sigma_b, sigma_a sigma_b}
// (sigma = 1 - 2 s).  Every component becomes a term whose selector is the parity of the
// output-state bits it depends on.  The diagonal (f = 0) is expanded the same way into a
// constant, per-site and per-bond-parity popcount groups.
// ------------------------------------------------------------------------------------
namespace {
struct Cx {
    double r = 0, i = 0;
};
inline Cx operator+(Cx a, Cx b) { return {a.r + b.r, a.i + b.i}; }
inline Cx operator-(Cx a, Cx b) { return {a.r - b.r, a.i - b.i}; }
inline Cx operator*(double s, Cx a) { return {s * a.r, s * a.i}; }
inline bool nz(Cx a) { return a.r * a.r + a.i * a.i > 0.5e-28; }

void walsh(const Cx g[4], Cx &e0, Cx &e1, Cx &e2, Cx &e12)
{
    // g index t = 2*s2 + s1 ; sigma = +1 for s = 0
    e0 = 0.25 * (g[0] + g[1] + g[2] + g[3]);
    e1 = 0.25 * (g[0] - g[1] + g[2] - g[3]);
    e2 = 0.25 * (g[0] + g[1] - g[2] - g[3]);
    e12 = 0.25 * (g[0] - g[1] - g[2] + g[3]);
}
}  // namespace

struct RawDecomp {
    std::map<std::pair<uint32_t, uint32_t>, Term> terms;
    std::vector<Cx> site_w;
    std::map<std::pair<int, int>, Cx> pair_w;   // (a, b) -> weight of parity(a,b)
    Cx c0;
};

// pos: memory bit of each site (null: bit k = site k).  The bond's local index keeps the
// reference's site order (a < b by site number); only the bits the terms act on move.
static int decompose_raw(int n, int ibond, const int32_t *ipair, const int32_t *btype, const double *xl,
                         const int32_t *ll, RawDecomp &rd, const int8_t *pos)
{
    rd = RawDecomp();
    rd.site_w.assign(n, Cx());
    auto X = [&](int t, int in, int out) -> Cx {
        const size_t f = (size_t)in + 4 * (size_t)out + 16 * (size_t)t;
        if (ll[f]) return Cx();
        return Cx{xl[2 * f], xl[2 * f + 1]};
    };
    auto add_term = [&](uint32_t gmask, uint32_t selmask, Cx cs0, Cx cs1) {
        Term &T = rd.terms[std::make_pair(gmask, selmask)];
        T.gmask = gmask;
        T.selmask = selmask;
        T.c0r += cs0.r;
        T.c0i += cs0.i;
        T.c1r += cs1.r;
        T.c1i += cs1.i;
    };
    for (int b = 0; b < ibond; ++b) {
        int s_a = ipair[2 * b] - 1, s_b = ipair[2 * b + 1] - 1;
        if (s_a < 0 || s_b < 0 || s_a >= n || s_b >= n || s_a == s_b)
            return fail(PNED_E_ARG, "#(E03) Incorrect data in ipair at bond " + std::to_string(b + 1));
        if (s_a > s_b) std::swap(s_a, s_b);
        const int t = btype[b] - 1;
        if (t < 0 || t > 8) return fail(PNED_E_ARG, "bond type outside 1..9");
        for (int l = 0; l < 4; ++l)
            if (ll[(size_t)l + 4 * l + 16 * t])
                return fail(PNED_E_UNSUPPORTED, "a masked bond diagonal (llocal_Ham(k,k)) is not supported");
        Cx g[4];
        for (int l = 0; l < 4; ++l) g[l] = X(t, l, l);
        Cx e0, e1, e2, e12;
        walsh(g, e0, e1, e2, e12);
        // d = e0 + e1 sa + e2 sb + e12 sa sb, sigma = 1 - 2s, sigma_a sigma_b = 1 - 2 par
        const int m_a = pos ? pos[s_a] : s_a, m_b = pos ? pos[s_b] : s_b;   // memory bits
        rd.c0 = rd.c0 + e0 + e1 + e2 + e12;
        rd.site_w[m_a] = rd.site_w[m_a] - 2.0 * e1;
        rd.site_w[m_b] = rd.site_w[m_b] - 2.0 * e2;
        const std::pair<int, int> pk{std::min(m_a, m_b), std::max(m_a, m_b)};
        rd.pair_w[pk] = rd.pair_w[pk] - 2.0 * e12;
        const uint32_t ba = 1u << m_a, bb = 1u << m_b;
        for (int f = 1; f <= 3; ++f) {
            for (int l = 0; l < 4; ++l) g[l] = X(t, l ^ f, l);
            walsh(g, e0, e1, e2, e12);
            const uint32_t gm = (f == 1) ? ba : (f == 2) ? bb : (ba | bb);
            Cx base = e0;
            bool used = false;
            if (nz(e12)) { add_term(gm, ba | bb, base + e12, base - e12); base = Cx(); used = true; }
            if (nz(e1)) { add_term(gm, ba, base + e1, base - e1); base = Cx(); used = true; }
            if (nz(e2)) { add_term(gm, bb, base + e2, base - e2); base = Cx(); used = true; }
            if (!used && nz(base)) add_term(gm, 0u, base, base);
        }
    }
    return PNED_OK;
}

static bool term_nz(const Term &T) { return nz(Cx{T.c0r, T.c0i}) || nz(Cx{T.c1r, T.c1i}); }

int s12_decompose(int n, int ibond, const int32_t *ipair, const int32_t *btype, const double *xl,
                  const int32_t *ll, S12Plan &plan, const int8_t *pos)
{
    const double *xls[1] = {xl};
    const int32_t *lls[1] = {ll};
    return s12_decompose_batch(n, ibond, ipair, btype, 1, xls, lls, plan, pos);
}

int s12_decompose_batch(int n, int ibond, const int32_t *ipair, const int32_t *btype, int nop,
                        const double *const *xl, const int32_t *const *ll, S12Plan &plan, const int8_t *pos)
{
    plan = S12Plan();
    plan.n = n;
    plan.nop = nop;
    std::vector<RawDecomp> rd(nop);
    for (int o = 0; o < nop; ++o) PNED_TRY(decompose_raw(n, ibond, ipair, btype, xl[o], ll[o], rd[o], pos));
    // term structure: union of the nonzero (mask, selector) keys
    std::map<std::pair<uint32_t, uint32_t>, int> key_index;
    for (int o = 0; o < nop; ++o)
        for (auto &kv : rd[o].terms)
            if (term_nz(kv.second)) key_index[kv.first] = 0;
    for (auto &kv : key_index) {
        kv.second = (int)plan.terms.size();
        Term t{};
        t.gmask = kv.first.first;
        t.selmask = kv.first.second;
        auto it = rd[0].terms.find(kv.first);
        if (it != rd[0].terms.end()) t = it->second;
        plan.terms.push_back(t);
    }
    plan.colcoef.assign(plan.terms.size() * nop * 4, 0.0);
    for (int o = 0; o < nop; ++o)
        for (auto &kv : rd[o].terms) {
            auto it = key_index.find(kv.first);
            if (it == key_index.end()) continue;   // zero in every operator
            double *c = &plan.colcoef[((size_t)it->second * nop + o) * 4];
            c[0] = kv.second.c0r;
            c[1] = kv.second.c0i;
            c[2] = kv.second.c1r;
            c[3] = kv.second.c1i;
        }
    // diagonal groups: sites (pairs per distance) with the same weight in every operator
    std::map<std::vector<double>, uint32_t> sg;
    for (int k = 0; k < n; ++k) {
        std::vector<double> key;
        bool any = false;
        for (int o = 0; o < nop; ++o) {
            key.push_back(rd[o].site_w[k].r);
            key.push_back(rd[o].site_w[k].i);
            any = any || nz(rd[o].site_w[k]);
        }
        if (any) sg[key] |= 1u << k;
    }
    std::map<std::vector<double>, uint32_t> pg;   // key: distance, weights...
    for (auto &kv : rd[0].pair_w) {
        std::vector<double> key{(double)(kv.first.second - kv.first.first)};
        bool any = false;
        for (int o = 0; o < nop; ++o) {
            const Cx w = rd[o].pair_w[kv.first];
            key.push_back(w.r);
            key.push_back(w.i);
            any = any || nz(w);
        }
        if (any) pg[key] |= 1u << kv.first.first;
    }
    for (auto &kv : sg) plan.diag.push_back(DiagGroup{0, 0, kv.second, kv.first[0], kv.first[1]});
    for (auto &kv : pg) plan.diag.push_back(DiagGroup{1, (int)kv.first[0], kv.second, kv.first[1], kv.first[2]});
    if (plan.diag.size() > 48) return fail(PNED_E_UNSUPPORTED, "diagonal has more than 48 coefficient groups");
    plan.c0r = rd[0].c0.r;
    plan.c0i = rd[0].c0.i;
    const size_t ng = plan.diag.size();
    plan.coldiag.assign((size_t)nop * (ng + 1) * 2, 0.0);
    size_t g = 0;
    for (auto &kv : sg) {
        for (int o = 0; o < nop; ++o) {
            plan.coldiag[((size_t)o * (ng + 1) + g) * 2] = kv.first[2 * o];
            plan.coldiag[((size_t)o * (ng + 1) + g) * 2 + 1] = kv.first[2 * o + 1];
        }
        ++g;
    }
    for (auto &kv : pg) {
        for (int o = 0; o < nop; ++o) {
            plan.coldiag[((size_t)o * (ng + 1) + g) * 2] = kv.first[1 + 2 * o];
            plan.coldiag[((size_t)o * (ng + 1) + g) * 2 + 1] = kv.first[2 + 2 * o];
        }
        ++g;
    }
    for (int o = 0; o < nop; ++o) {
        plan.coldiag[((size_t)o * (ng + 1) + ng) * 2] = rd[o].c0.r;
        plan.coldiag[((size_t)o * (ng + 1) + ng) * 2 + 1] = rd[o].c0.i;
    }
    return PNED_OK;
}

int s12_set_column(S12Plan &plan, int ibond, const int32_t *ipair, const int32_t *btype, int col,
                   const double *xl, const int32_t *ll, const int8_t *pos)
{
    if (col < 0 || col >= plan.nop) return fail(PNED_E_ARG, "column outside the batch");
    RawDecomp rd;
    PNED_TRY(decompose_raw(plan.n, ibond, ipair, btype, xl, ll, rd, pos));
    std::map<std::pair<uint32_t, uint32_t>, int> idx;
    for (size_t t = 0; t < plan.terms.size(); ++t) idx[{plan.terms[t].gmask, plan.terms[t].selmask}] = (int)t;
    std::vector<double> cc(plan.terms.size() * 4, 0.0);
    for (auto &kv : rd.terms) {
        if (!term_nz(kv.second)) continue;
        auto it = idx.find(kv.first);
        if (it == idx.end()) return fail(PNED_E_UNSUPPORTED, "operator has a term outside the batch structure");
        double *c = &cc[(size_t)it->second * 4];
        c[0] = kv.second.c0r;
        c[1] = kv.second.c0i;
        c[2] = kv.second.c1r;
        c[3] = kv.second.c1i;
    }
    // diagonal: every group must carry one weight for this operator.  Weights that are
    // equal in exact arithmetic can differ in the last bits (a site's weight is summed over
    // its bonds in bond order, and the x/y/z bond contributions round differently), so a
    // group accepts sites within a few ulps of the largest weight and uses its first site's.
    const size_t ng = plan.diag.size();
    std::vector<double> dw((ng + 1) * 2, 0.0);
    double wscale = 0.0;
    for (int k = 0; k < plan.n; ++k) wscale = std::max(wscale, std::hypot(rd.site_w[k].r, rd.site_w[k].i));
    for (auto &kv : rd.pair_w) wscale = std::max(wscale, std::hypot(kv.second.r, kv.second.i));
    const double wtol = 64.0 * std::numeric_limits<double>::epsilon() * wscale;
    std::vector<char> seen_site(plan.n, 0);
    std::map<std::pair<int, int>, char> seen_pair;
    for (size_t g = 0; g < ng; ++g) {
        const DiagGroup &G = plan.diag[g];
        bool first = true;
        Cx w;
        for (int k = 0; k < plan.n; ++k) {
            if (!(G.mask >> k & 1)) continue;
            Cx x;
            if (G.kind == 0) {
                x = rd.site_w[k];
                seen_site[k] = 1;
            } else {
                auto it = rd.pair_w.find({k, k + G.shift});
                if (it != rd.pair_w.end()) x = it->second;
                seen_pair[{k, k + G.shift}] = 1;
            }
            if (first) { w = x; first = false; }
            else if (std::hypot(x.r - w.r, x.i - w.i) > wtol)
                return fail(PNED_E_UNSUPPORTED, "diagonal weights do not fit the batch groups");
        }
        dw[2 * g] = w.r;
        dw[2 * g + 1] = w.i;
    }
    for (int k = 0; k < plan.n; ++k)
        if (!seen_site[k] && nz(rd.site_w[k])) return fail(PNED_E_UNSUPPORTED, "site weight outside the batch groups");
    for (auto &kv : rd.pair_w)
        if (!seen_pair.count(kv.first) && nz(kv.second))
            return fail(PNED_E_UNSUPPORTED, "bond weight outside the batch groups");
    dw[2 * ng] = rd.c0.r;
    dw[2 * ng + 1] = rd.c0.i;
    for (size_t t = 0; t < plan.terms.size(); ++t)
        for (int k = 0; k < 4; ++k) plan.colcoef[((size_t)t * plan.nop + col) * 4 + k] = cc[t * 4 + k];
    for (size_t k = 0; k < dw.size(); ++k) plan.coldiag[(size_t)col * (ng + 1) * 2 + k] = dw[k];
    // keep devcoef (pass order) in step
    if (!plan.devcoef.empty()) {
        for (size_t d = 0; d < plan.devterms.size(); ++d) {
            const int t = plan.devterms[d].pad;   // index into plan.terms
            for (int k = 0; k < 4; ++k)
                plan.devcoef[((size_t)d * plan.nop + col) * 4 + k] = cc[(size_t)t * 4 + k];
        }
    }
    return PNED_OK;
}

// Tile-pass cover.  Memory bits: [colbits column bits | n state bits].  Every term must
// lie inside one pass's tile bits; each pass costs one full stream of v1 and v0, so the
// planner minimises the number of passes.  Randomised greedy with restarts (fixed seed,
// so the plan is reproducible): a pass starts from the forced low bits plus the bits of one
// uncovered term and grows by the bit that completes the most uncovered terms (ties:
// terms touching it, then a random key).
namespace {
struct Lcg {
    uint64_t s;
    uint32_t next() { s = s * 6364136223846793005ull + 1442695040888963407ull; return (uint32_t)(s >> 33); }
};
}  // namespace

static std::vector<uint64_t> cover_once(const std::vector<uint64_t> &mm, int colbits, int nb, int K, int L,
                                        Lcg &rng, bool randomised)
{
    std::vector<char> done(mm.size(), 0);
    size_t remaining = mm.size();
    std::vector<uint64_t> tiles;
    const uint64_t low = (L >= 64) ? ~0ull : ((1ull << L) - 1);
    bool first = true;
    while (first || remaining > 0) {
        uint64_t T = low;
        int cnt = L;
        if (remaining > 0 && randomised) {
            // seed with one uncovered term
            size_t pick = rng.next() % remaining, k = 0;
            for (size_t i = 0; i < mm.size(); ++i)
                if (!done[i] && k++ == pick) {
                    const uint64_t add = mm[i] & ~T;
                    if (__builtin_popcountll(add) + cnt <= K) {
                        T |= add;
                        cnt += __builtin_popcountll(add);
                    }
                    break;
                }
        }
        while (cnt < K) {
            int best = -1;
            uint64_t best_score = 0;
            for (int bit = colbits; bit < nb; ++bit) {
                if (T >> bit & 1) continue;
                const uint64_t T2 = T | (1ull << bit);
                uint64_t newly = 0, touch = 0;
                for (size_t i = 0; i < mm.size(); ++i) {
                    if (done[i] || !(mm[i] >> bit & 1)) continue;
                    ++touch;
                    if ((mm[i] & ~T2) == 0) ++newly;
                }
                const uint64_t key = randomised ? (rng.next() & 0xFFFFF) : (uint64_t)(nb - bit);
                const uint64_t score = (newly << 44) | (touch << 24) | key;
                if (best < 0 || score > best_score) { best_score = score; best = bit; }
            }
            if (best < 0) break;
            T |= 1ull << best;
            ++cnt;
        }
        size_t got = 0;
        for (size_t i = 0; i < mm.size(); ++i)
            if (!done[i] && (mm[i] & ~T) == 0) { done[i] = 1; --remaining; ++got; }
        if (got == 0 && !first) return {};   // stuck
        tiles.push_back(T);
        first = false;
        if (tiles.size() > 64) return {};
    }
    return tiles;
}

// Local search for a cover with P tiles: every tile is the forced low bits plus K-L free
// bits; a move swaps one free bit of one tile for a bit outside it, accepted by the
// Metropolis rule on the number of flip masks no tile contains.  Starts from `init`
// (P tiles), fixed seed, so the result is reproducible.
static bool anneal_cover(const std::vector<uint64_t> &masks, int nb, int L, int K, std::vector<uint64_t> tiles,
                         Lcg &rng, int iters, std::vector<uint64_t> &out)
{
    const int P = (int)tiles.size();
    const int m = K - L;
    if (P == 0 || m <= 0 || m >= nb - L) return false;
    const size_t nm = masks.size();
    std::vector<int> cov(nm, 0);
    int unc = 0;
    for (size_t i = 0; i < nm; ++i) {
        for (int p = 0; p < P; ++p) cov[i] += ((masks[i] & ~tiles[p]) == 0);
        unc += (cov[i] == 0);
    }
    if (unc == 0) { out = tiles; return true; }
    std::vector<int> in_bits, out_bits;
    for (int it = 0; it < iters; ++it) {
        const double temp = 1.5 * (1.0 - (double)it / iters) + 0.02;
        const int p = (int)(rng.next() % (uint32_t)P);
        const uint64_t T = tiles[p];
        in_bits.clear();
        out_bits.clear();
        for (int b = L; b < nb; ++b) (T >> b & 1 ? in_bits : out_bits).push_back(b);
        const int a = in_bits[rng.next() % in_bits.size()];
        const int c = out_bits[rng.next() % out_bits.size()];
        const uint64_t T2 = (T & ~(1ull << a)) | (1ull << c);
        int delta = 0;
        for (size_t i = 0; i < nm; ++i) {
            const int was = (masks[i] & ~T) == 0, now = (masks[i] & ~T2) == 0;
            if (was == now) continue;
            const int c2 = cov[i] - was + now;
            delta += (c2 == 0) - (cov[i] == 0);
        }
        const double u = (double)(rng.next() & 0xFFFFFF) / 16777216.0;
        if (delta <= 0 || u < std::exp(-(double)delta / temp)) {
            for (size_t i = 0; i < nm; ++i) {
                const int was = (masks[i] & ~T) == 0, now = (masks[i] & ~T2) == 0;
                cov[i] += now - was;
            }
            tiles[p] = T2;
            unc += delta;
            if (unc == 0) { out = tiles; return true; }
        }
    }
    return false;
}

// The same local search from random tiles (the forced low bits plus K-L random free
// bits), `tries` seeds derived from `seed`: finds covers the greedy start misses.  Fixed
// seeds make the outcome a function of the mask set alone.
static bool random_cover(const std::vector<uint64_t> &um, int nb, int L, int K, int P, uint64_t seed, int tries,
                         int iters, std::vector<uint64_t> &out)
{
    if (P < 1 || nb <= K || K <= L) return false;
    for (int t = 0; t < tries; ++t) {
        Lcg rng{seed * 0x9E3779B97F4A7C15ull + 0xD1B54A32D192ED03ull * (uint64_t)(t + 1)};
        std::vector<uint64_t> init((size_t)P);
        for (int p = 0; p < P; ++p) {
            uint64_t T = (1ull << L) - 1;
            int cnt = L;
            while (cnt < K) {
                const int b = L + (int)(rng.next() % (uint32_t)(nb - L));
                if (!(T >> b & 1)) { T |= 1ull << b; ++cnt; }
            }
            init[(size_t)p] = T;
        }
        if (anneal_cover(um, nb, L, K, init, rng, iters, out)) return true;
    }
    return false;
}

// Memory locality of a cover: a tile that holds at least two of the memory-index bits
// colbits+3 .. colbits+7 reads at least 512 contiguous-ish bytes of every 4 KB it touches;
// tiles without such bits read one 128-byte run per 4 KB and stream slower (measured at
// N=32: 46 ms per pass against 39-40 ms).  Score = sum over tiles of min(2, such bits).
static int locality_score(const std::vector<uint64_t> &tiles, int colbits)
{
    const uint64_t low = 0x1Full << (colbits + 3);
    int s = 0;
    for (uint64_t t : tiles) s += std::min(2, __builtin_popcountll(t & low));
    return s;
}

// Anneals a complete cover towards a higher locality score: moves exchange one free bit of
// one tile; the cost is 10 per uncovered mask minus the score (so the walk may pass through
// incomplete covers), and the best complete cover seen is kept.
static void refine_locality(const std::vector<uint64_t> &masks, int nb, int L, int colbits,
                            std::vector<uint64_t> &tiles, Lcg &rng, int iters)
{
    const int P = (int)tiles.size();
    const size_t nm = masks.size();
    std::vector<int> cov(nm, 0);
    for (size_t i = 0; i < nm; ++i)
        for (int p = 0; p < P; ++p) cov[i] += ((masks[i] & ~tiles[p]) == 0);
    int unc = 0;
    for (size_t i = 0; i < nm; ++i) unc += cov[i] == 0;
    int score = locality_score(tiles, colbits), best = unc == 0 ? score : -1;
    std::vector<uint64_t> best_tiles = tiles;
    std::vector<int> in_bits, out_bits;
    for (int it = 0; it < iters && best < 2 * P; ++it) {
        const double temp = std::max(0.05, 2.0 * std::pow(0.9997, (double)it));
        const int p = (int)(rng.next() % (uint32_t)P);
        const uint64_t T = tiles[p];
        in_bits.clear();
        out_bits.clear();
        for (int b = L; b < nb; ++b) (T >> b & 1 ? in_bits : out_bits).push_back(b);
        if (in_bits.empty() || out_bits.empty()) break;
        const int a = in_bits[rng.next() % in_bits.size()];
        const int c = out_bits[rng.next() % out_bits.size()];
        const uint64_t T2 = (T & ~(1ull << a)) | (1ull << c);
        int du = 0;
        for (size_t i = 0; i < nm; ++i) {
            const int was = (masks[i] & ~T) == 0, now = (masks[i] & ~T2) == 0;
            if (was == now) continue;
            du += ((cov[i] - was + now) == 0) - (cov[i] == 0);
        }
        tiles[p] = T2;
        const int sc = locality_score(tiles, colbits);
        const double dcost = 10.0 * du - (sc - score);
        const double u = (double)(rng.next() & 0xFFFFFF) / 16777216.0;
        if (dcost <= 0 || u < std::exp(-dcost / temp)) {
            for (size_t i = 0; i < nm; ++i) cov[i] += ((masks[i] & ~T2) == 0) - ((masks[i] & ~T) == 0);
            unc += du;
            score = sc;
            if (unc == 0 && score > best) { best = score; best_tiles = tiles; }
        } else {
            tiles[p] = T;
        }
    }
    tiles = best_tiles;
}

static std::vector<uint64_t> unique_masks(const S12Plan &plan, int colbits)
{
    std::vector<uint64_t> um;
    for (const Term &t : plan.terms) um.push_back((uint64_t)t.gmask << colbits);
    std::sort(um.begin(), um.end());
    um.erase(std::unique(um.begin(), um.end()), um.end());
    return um;
}

constexpr uint64_t RANDOM_COVER_SEED = 0x5EEDull;
constexpr int RANDOM_COVER_TRIES = 2, RANDOM_COVER_ITERS = 30000;
constexpr uint64_t LOCALITY_SEED = 0x10CA1ull;
constexpr int LOCALITY_TRIES = 8, LOCALITY_ITERS = 40000;

int s12_make_passes(S12Plan &plan, int colbits, int tile_bits, int low_bits)
{
    plan.passes.clear();
    plan.devterms.clear();
    plan.colbits = colbits;
    const int nb = colbits + plan.n;
    const int K = std::min(tile_bits, nb);
    const int L = std::min(std::max(low_bits, colbits), K);
    std::vector<uint64_t> mmask(plan.terms.size());
    for (size_t i = 0; i < plan.terms.size(); ++i) mmask[i] = (uint64_t)plan.terms[i].gmask << colbits;
    Lcg rng{0x9E3779B97F4A7C15ull};
    std::vector<uint64_t> best = cover_once(mmask, colbits, nb, K, L, rng, false);
    const int restarts = (nb > K) ? 400 : 0;
    for (int r = 0; r < restarts; ++r) {
        std::vector<uint64_t> t = cover_once(mmask, colbits, nb, K, L, rng, true);
        if (!t.empty() && (best.empty() || t.size() < best.size())) best = t;
    }
    if (best.empty()) return fail(PNED_E_UNSUPPORTED, "tile planner made no progress");
    // refine: try one tile fewer (start from the greedy tiles minus the one that covers
    // the fewest masks on its own) until the local search fails
    if (nb > K && best.size() > 1) {
        std::vector<uint64_t> um(mmask);
        std::sort(um.begin(), um.end());
        um.erase(std::unique(um.begin(), um.end()), um.end());
        const uint64_t low = (1ull << L) - 1;
        for (;;) {
            size_t drop = 0, fewest = SIZE_MAX;
            for (size_t p = 0; p < best.size(); ++p) {
                size_t own = 0;
                for (uint64_t mk : um) {
                    if ((mk & ~best[p]) != 0) continue;
                    bool other = false;
                    for (size_t q = 0; q < best.size() && !other; ++q) other = (q != p) && (mk & ~best[q]) == 0;
                    own += !other;
                }
                if (own < fewest) { fewest = own; drop = p; }
            }
            std::vector<uint64_t> init;
            for (size_t p = 0; p < best.size(); ++p)
                if (p != drop) init.push_back(best[p] | low);
            std::vector<uint64_t> found;
            bool ok = false;
            for (int attempt = 0; attempt < 3 && !ok; ++attempt)
                ok = anneal_cover(um, nb, L, K, init, rng, 60000, found);
            if (!ok) break;
            best = found;
            if (best.size() <= 1) break;
        }
        // then from random tiles, one tile below the best cover so far
        while (best.size() > 1) {
            std::vector<uint64_t> found;
            if (!random_cover(um, nb, L, K, (int)best.size() - 1, RANDOM_COVER_SEED, RANDOM_COVER_TRIES,
                              RANDOM_COVER_ITERS, found))
                break;
            best = found;
        }
        // with the pass count fixed, give the tiles low memory bits (a few restarts)
        std::vector<uint64_t> loc = best;
        int loc_score = locality_score(best, colbits);
        for (int t = 0; t < LOCALITY_TRIES && loc_score < 2 * (int)best.size(); ++t) {
            std::vector<uint64_t> trial = best;
            Lcg lr{LOCALITY_SEED + 0x9E3779B97F4A7C15ull * (uint64_t)(t + 1)};
            refine_locality(um, nb, L, colbits, trial, lr, LOCALITY_ITERS);
            const int sc = locality_score(trial, colbits);
            if (sc > loc_score) { loc_score = sc; loc = trial; }
        }
        best = loc;
    }
    // Each term goes to the first tile that contains it; the diagonal goes to the pass with
    // the least term work (1 per real term, 2 per complex term: twice the FMAs), which then
    // runs first (it also takes the v0 input).  Measured at N=32 with 4 passes (MI355X):
    // 206.3 ms per H.v, against 209-210 ms with the diagonal on the first tile's pass and
    // 211.3 ms when multiply covered terms are also spread over the passes by work.
    const size_t P = best.size(), nt = plan.terms.size();
    std::vector<int> best_owner(nt, -1);
    std::vector<double> load(P, 0.0);
    for (size_t i = 0; i < nt; ++i)
        for (size_t p = 0; p < P; ++p)
            if ((mmask[i] & ~best[p]) == 0) {
                best_owner[i] = (int)p;
                const Term &t = plan.terms[i];
                load[p] += (plan.nop > 1 || t.c0i != 0.0 || t.c1i != 0.0) ? 2.0 : 1.0;
                break;
            }
    size_t diag_pass = 0;
    for (size_t p = 1; p < P; ++p)
        if (load[p] < load[diag_pass]) diag_pass = p;
    std::vector<size_t> pass_order{diag_pass};
    for (size_t p = 0; p < P; ++p)
        if (p != diag_pass) pass_order.push_back(p);
    for (size_t p : pass_order) {
        const uint64_t T = best[p];
        PassPlan pp;
        pp.K = __builtin_popcountll(T);
        for (int bit = 0; bit < nb; ++bit) {
            if (T >> bit & 1) pp.tb.push_back(bit);
            else pp.ob.push_back(bit);
        }
        pp.diag = (p == diag_pass);
        pp.first_term = (int)plan.devterms.size();
        for (size_t i = 0; i < plan.terms.size(); ++i) {
            if (best_owner[i] != (int)p) continue;
            const Term &t = plan.terms[i];
            DevTerm d{};
            uint32_t tm = 0;
            for (int k = 0; k < pp.K; ++k)
                if (mmask[i] >> pp.tb[k] & 1) tm |= 1u << k;
            d.tmask = tm;
            d.selmask = t.selmask;
            d.pad = (uint32_t)i;   // index into plan.terms (batch coefficients)
            d.is_real = (t.c0i == 0.0 && t.c1i == 0.0) ? 1u : 0u;
            d.c0r = t.c0r;
            d.c0i = t.c0i;
            d.c1r = t.c1r;
            d.c1i = t.c1i;
            plan.devterms.push_back(d);
        }
        pp.nterm = (int)plan.devterms.size() - pp.first_term;
        if (pp.nterm == 0 && !pp.diag) continue;   // every term already placed
        plan.passes.push_back(pp);
    }
    plan.devcoef.clear();
    if (plan.nop > 1) {
        plan.devcoef.resize(plan.devterms.size() * plan.nop * 4);
        for (size_t d = 0; d < plan.devterms.size(); ++d) {
            const size_t t = plan.devterms[d].pad;
            for (int o = 0; o < plan.nop; ++o)
                for (int k = 0; k < 4; ++k)
                    plan.devcoef[(d * plan.nop + o) * 4 + k] = plan.colcoef[(t * plan.nop + o) * 4 + k];
        }
    }
    return PNED_OK;
}

int ensure_partial(pned_ctx *ctx, size_t nd)
{
    if (nd <= ctx->partial_cap) return PNED_OK;
    if (ctx->d_partial) PNED_HIP(hipFree(ctx->d_partial));
    ctx->d_partial = nullptr;
    size_t cap = std::max(nd, (size_t)1 << 16);
    PNED_HIP(hipMalloc((void **)&ctx->d_partial, sizeof(double) * cap));
    ctx->partial_cap = cap;
    return PNED_OK;
}

static int hv_dispatch(pned_ctx *ctx, const pned_vec *v1, pned_vec *v0, double eperbd, bool fresh, double *prdct);

// v0 += H v1, then alpha = <v1|H v1> (the H.v's fused reduction, complex per column) and
// beta^2 = |v0 - alpha v1|^2 with one readback: alpha stays on the device for the beta
// kernel (the values it reads are the ones a host round trip would hand it)
int hv_alpha_beta2(pned_ctx *ctx, pned_vec *v1, pned_vec *v0, double *pr, double *b2)
{
    PNED_TRY(ensure_result(ctx, 3 * (size_t)v1->nvec + 8));   // no reallocation between the two
    PNED_TRY(hv_dispatch(ctx, v1, v0, 0.0, false, nullptr));
    return beta2_after_hv(ctx, v1, v0, pr, b2);
}

int ensure_result(pned_ctx *ctx, size_t nd)
{
    if (nd <= ctx->result_cap) return PNED_OK;
    PNED_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->d_result) PNED_HIP(hipFree(ctx->d_result));
    if (ctx->h_result) PNED_HIP(hipHostFree(ctx->h_result));
    ctx->d_result = nullptr;
    ctx->h_result = nullptr;
    size_t cap = std::max(nd, (size_t)8192);
    PNED_HIP(hipMalloc((void **)&ctx->d_result, sizeof(double) * cap));
    PNED_HIP(hipHostMalloc((void **)&ctx->h_result, sizeof(double) * cap, hipHostMallocDefault));
    ctx->result_cap = cap;
    return PNED_OK;
}

// small argument arrays for kernels: uploaded in stream order into a dedicated buffer
// (each kernel that reads them is enqueued before the next upload overwrites them)
int ensure_scalars(pned_ctx *ctx, int n, double **dptr)
{
    if ((size_t)n > ctx->scalars_cap) {
        PNED_HIP(hipStreamSynchronize(ctx->stream));
        if (ctx->d_scalars) PNED_HIP(hipFree(ctx->d_scalars));
        ctx->d_scalars = nullptr;
        const size_t cap = std::max<size_t>((size_t)n, 8192);
        PNED_HIP(hipMalloc((void **)&ctx->d_scalars, sizeof(double) * cap));
        ctx->scalars_cap = cap;
    }
    *dptr = ctx->d_scalars;
    return PNED_OK;
}

int upload_scalars(pned_ctx *ctx, const double *h, int n, double **dptr)
{
    PNED_TRY(ensure_scalars(ctx, n, dptr));
    PNED_HIP(hipMemcpyAsync(ctx->d_scalars, h, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
    return PNED_OK;
}

int layout_from_low_sites(int n, const int *sites, int nsites, Layout &out)
{
    out = Layout();
    if (nsites < 0 || nsites > LAYOUT_MAX_SWAPS || n < nsites || n > 32) return fail(PNED_E_ARG, "bad layout sites");
    std::vector<char> in_f(n, 0);
    for (int k = 0; k < nsites; ++k) {
        if (sites[k] < 0 || sites[k] >= n || in_f[sites[k]]) return fail(PNED_E_ARG, "bad layout sites");
        in_f[sites[k]] = 1;
    }
    // sites outside the low bits trade places with low bits that hold no chosen site
    int free_low = 0;
    for (int k = 0; k < nsites; ++k) {
        const int s = sites[k];
        if (s < nsites) continue;
        while (in_f[free_low]) ++free_low;
        out.a[out.nswap] = free_low;
        out.b[out.nswap] = s;
        ++out.nswap;
        out.pos[free_low] = (int8_t)s;
        out.pos[s] = (int8_t)free_low;
        ++free_low;
    }
    return PNED_OK;
}

int layout_from_table(int n, const int32_t *site_bit, Layout &out)
{
    out = Layout();
    if (n < 1 || n > 32 || !site_bit) return fail(PNED_E_ARG, "bad layout table");
    for (int k = 0; k < n; ++k) {
        const int b = site_bit[k];
        if (b < 0 || b >= n || site_bit[b] != k)
            return fail(PNED_E_ARG, "layout table must exchange pairs of sites (site_bit[site_bit[k]] == k)");
        if (b > k) {
            if (out.nswap == LAYOUT_MAX_SWAPS) return fail(PNED_E_ARG, "layout table swaps more than 4 pairs");
            out.a[out.nswap] = k;
            out.b[out.nswap] = b;
            ++out.nswap;
        }
        out.pos[k] = (int8_t)b;
    }
    return PNED_OK;
}

// The tile plan of a decomposition seen through a layout (flip masks mapped to memory bits):
// the tile size the untuned plan of a large vector takes (K = 12, or 13 when that saves a
// pass) and the plan's covers.  Returns false when the planner fails.
static bool plan_through(const S12Plan &ip, const Layout &lay, int &K, std::vector<uint64_t> &tiles)
{
    const SwapArgs sw = swap_args(lay);
    S12Plan tmp[2];
    for (int k = 0; k < 2; ++k) {
        tmp[k].n = ip.n;
        tmp[k].terms = ip.terms;
        for (Term &t : tmp[k].terms) {
            t.gmask = (uint32_t)swap_bits(sw, t.gmask);
            t.selmask = (uint32_t)swap_bits(sw, t.selmask);
        }
        if (s12_make_passes(tmp[k], 0, 12 + k, 3) != PNED_OK) return false;
    }
    const int k = tmp[1].passes.size() < tmp[0].passes.size() ? 1 : 0;
    K = 12 + k;
    tiles.clear();
    for (const PassPlan &pp : tmp[k].passes) {
        uint64_t t = 0;
        for (int b : pp.tb) t |= 1ull << b;
        tiles.push_back(t);
    }
    return true;
}

// Spends the swaps a layout has left on locality: memory bits 3..7 hold sites whose
// exchange with them gives the plan's tiles the most of those bits (two per tile count; see
// locality_score).  Each step ranks the possible exchanges by the score of the current tiles
// relabelled, reruns the planner through the best LOCALITY_CANDIDATES of them and keeps the
// one whose own plan has the same tile size, no more passes and the highest score; the
// search stops when no exchange raises the score.  Clusters of 30 sites and more only:
// measured on one box, N=32 (4x4 torus, field; tiles with 3 / 0 / 1 / 3 such bits ->
// 2 / 2 / 4 / 1) 190.8-191.2 -> 185.9 ms per H.v, but N=28 (2x7) 9.52-9.55 -> 9.85-9.97 ms.
constexpr int LOCALITY_MIN_SITES = 30, LOCALITY_CANDIDATES = 6;
static Layout add_locality_swaps(const S12Plan &ip, const Layout &base)
{
    if (ip.nop != 1 || ip.n < LOCALITY_MIN_SITES || ip.n > 32 || base.nswap >= LAYOUT_MAX_SWAPS) return base;
    int K0;
    std::vector<uint64_t> tiles;
    if (!plan_through(ip, base, K0, tiles) || tiles.size() < 2) return base;
    const size_t P0 = tiles.size();
    int score = locality_score(tiles, 0);
    std::vector<char> used(ip.n, 0);
    for (int i = 0; i < base.nswap; ++i) used[base.a[i]] = used[base.b[i]] = 1;
    for (int b = 0; b < 3; ++b) used[b] = 1;   // the forced low bits stay
    Layout lay = base;
    auto relabel = [](std::vector<uint64_t> t, int l, int h) {
        for (uint64_t &x : t) {
            const uint64_t bl = x >> l & 1, bh = x >> h & 1;
            x = (x & ~((1ull << l) | (1ull << h))) | (bh << l) | (bl << h);
        }
        return t;
    };
    while (lay.nswap < LAYOUT_MAX_SWAPS && score < 2 * (int)tiles.size()) {
        std::vector<std::tuple<int, int, int>> cand;   // (-predicted score, l, h)
        for (int l = 3; l < 8 && l < ip.n; ++l) {
            if (used[l]) continue;
            for (int h = 8; h < ip.n; ++h)
                if (!used[h]) cand.emplace_back(-locality_score(relabel(tiles, l, h), 0), l, h);
        }
        std::sort(cand.begin(), cand.end());
        int best_sc = score;
        Layout best_lay;
        std::vector<uint64_t> best_tiles;
        for (size_t c = 0; c < cand.size() && c < (size_t)LOCALITY_CANDIDATES; ++c) {
            const int l = std::get<1>(cand[c]), h = std::get<2>(cand[c]);
            Layout t = lay;
            t.a[t.nswap] = l;
            t.b[t.nswap] = h;
            ++t.nswap;
            t.pos[l] = (int8_t)h;
            t.pos[h] = (int8_t)l;
            int K1;
            std::vector<uint64_t> t1;
            if (!plan_through(ip, t, K1, t1) || K1 != K0 || t1.size() > P0) continue;
            const int sc = locality_score(t1, 0);
            if (sc > best_sc) { best_sc = sc; best_lay = t; best_tiles = t1; }
        }
        if (best_sc <= score) break;
        lay = best_lay;
        tiles = best_tiles;
        score = best_sc;
        used[lay.a[lay.nswap - 1]] = used[lay.b[lay.nswap - 1]] = 1;
    }
    return lay;
}

// Candidate layouts put three sites into the forced low tile bits (L = 3).  For the tile
// sizes of large vectors, K = 12 and 13 (2^K complex128 states fill the LDS at 13), the
// identity layout's pass count P_K is the planner's own; a candidate gains one for every
// K at which the random-start local search (the planner's last stage, same seeds) covers
// the mapped flip masks with P_K - 1 tiles.  The best candidate must gain; ties keep the
// first found.  Seeded, so a cluster always gets the same layout.  Vectors below 2^18
// states keep the identity: their H.v is bound by launches, not by passes.
constexpr int LAYOUT_K0 = 12, LAYOUT_NK = 2, LAYOUT_CANDIDATES = 32, LAYOUT_MIN_SITES = 18;
Layout choose_layout(const S12Plan &ip)
{
    Layout best;
    const int n = ip.n, L = 3;
    if (ip.nop != 1 || n < LAYOUT_MIN_SITES || n > 32 || ip.terms.empty()) return best;
    const std::vector<uint64_t> um = unique_masks(ip, 0);
    int pnat[LAYOUT_NK];
    for (int ki = 0; ki < LAYOUT_NK; ++ki) {
        S12Plan tmp;
        tmp.n = ip.n;
        tmp.terms = ip.terms;
        if (s12_make_passes(tmp, 0, LAYOUT_K0 + ki, L) != PNED_OK) return best;
        pnat[ki] = (int)tmp.passes.size();
    }
    Lcg rng{0xC0FFEE123ull};
    int best_gain = 0;
    for (int c = 0; c < LAYOUT_CANDIDATES && best_gain < LAYOUT_NK; ++c) {
        int st[3];
        for (int k = 0; k < 3; ++k) {
            bool dup;
            do {
                st[k] = (int)(rng.next() % (uint32_t)n);
                dup = false;
                for (int j = 0; j < k; ++j) dup = dup || st[j] == st[k];
            } while (dup);
        }
        Layout lay;
        if (layout_from_low_sites(n, st, 3, lay) != PNED_OK || lay.identity()) continue;
        const SwapArgs sw = swap_args(lay);
        std::vector<uint64_t> mm;
        for (uint64_t m : um) mm.push_back(swap_bits(sw, m));
        std::sort(mm.begin(), mm.end());
        int gain = 0;
        for (int ki = 0; ki < LAYOUT_NK; ++ki) {
            std::vector<uint64_t> found;
            if (pnat[ki] > 1 && random_cover(mm, n, L, LAYOUT_K0 + ki, pnat[ki] - 1, RANDOM_COVER_SEED,
                                             RANDOM_COVER_TRIES, RANDOM_COVER_ITERS, found))
                ++gain;
        }
        if (gain > best_gain) {
            best_gain = gain;
            best = lay;
        }
    }
    return add_locality_swaps(ip, best);
}

int reset_layout(pned_ctx *ctx)
{
    if (ctx->lay.identity()) return PNED_OK;
    ctx->lay = Layout();
    PNED_TRY(s12_decompose(ctx->n, ctx->ibond, ctx->ipair.data(), ctx->btype.data(), ctx->xl.data(), ctx->ll.data(),
                           ctx->s12, nullptr));
    ctx->plan_colbits = -1;
    ctx->plan_dtype = -1;
    ctx->pd_threads = -1;
    return PNED_OK;
}

int check_same_layout(const pned_vec *x, const pned_vec *y)
{
    if (!x->ctx || !y->ctx || x->ctx == y->ctx) return PNED_OK;
    if (!x->ctx->lay.same(y->ctx->lay))
        return fail(PNED_E_ARG, "the two vectors' contexts lay their states out differently in memory");
    return PNED_OK;
}

bool default_large(const pned_ctx *ctx, int colbits)
{
    return (double)ctx->dim * (double)(1ull << colbits) > (double)(1ull << 25);
}

// The lane program of a diagonal (see DiagDev): groups sorted by weight class, so the
// kernel sums a class's popcounts as integers and converts them once.
static void build_diag_program(DiagDev &dh)
{
    dh.nprog = 0;
    std::memset(dh.prog_a, 0, sizeof(dh.prog_a));
    std::memset(dh.prog_b, 0, sizeof(dh.prog_b));
    if (dh.ngroups < 1 || dh.ngroups > 48) return;
    std::vector<std::pair<double, double>> cls;
    std::vector<int> cof(dh.ngroups);
    for (int g = 0; g < dh.ngroups; ++g) {
        const std::pair<double, double> w(dh.wr[g], dh.wi[g]);
        auto it = std::find(cls.begin(), cls.end(), w);
        cof[g] = (int)(it - cls.begin());
        if (it == cls.end()) cls.push_back(w);
        if (dh.shift[g] < 0 || dh.shift[g] > 63) return;
    }
    if (cls.size() > 8) return;
    int e = 0;
    for (int c = 0; c < (int)cls.size(); ++c) {
        int last = -1;
        for (int g = 0; g < dh.ngroups; ++g)
            if (cof[g] == c) {
                dh.prog_a[e] = (uint32_t)dh.shift[g] | (dh.kind[g] != 0 ? 64u : 0u) | ((uint32_t)c << 8);
                dh.prog_b[e] = dh.mask[g];
                last = e++;
            }
        dh.prog_a[last] |= 128u;
        uint64_t wr, wi;
        std::memcpy(&wr, &cls[c].first, 8);
        std::memcpy(&wi, &cls[c].second, 8);
        dh.prog_a[48 + c] = (uint32_t)wr;
        dh.prog_a[56 + c] = (uint32_t)(wr >> 32);
        dh.prog_b[48 + c] = (uint32_t)wi;
        dh.prog_b[56 + c] = (uint32_t)(wi >> 32);
    }
    dh.nprog = e;
}

bool split_diag_slots(const DiagDev &dh, uint32_t slot_bits, DiagDev &out)
{
    if (dh.nprog < 1) return false;
    int ncls = 0;
    for (int e = 0; e < dh.nprog; ++e) ncls = std::max(ncls, (int)((dh.prog_a[e] >> 8) & 7u) + 1);
    if (ncls > 4) return false;
    std::vector<uint32_t> sa, sb, ra, rb;
    for (int c = 0; c < ncls; ++c) {
        const size_t s0 = sa.size(), r0 = ra.size();
        for (int e = 0; e < dh.nprog; ++e) {
            const uint32_t pa = dh.prog_a[e] & ~128u;
            if ((int)((pa >> 8) & 7u) != c) continue;
            const uint32_t sh = pa & 63u, m = dh.prog_b[e];
            // positions b of the mask whose bit b, or bit b + sh of a pair, is a register bit
            const uint32_t touch = slot_bits | ((pa & 64u) ? (uint32_t)((uint64_t)slot_bits >> sh) : 0u);
            if (m & touch) {
                sa.push_back(pa);
                sb.push_back(m & touch);
            }
            if (m & ~touch) {
                ra.push_back(pa);
                rb.push_back(m & ~touch);
            }
        }
        if (sa.size() == s0) {   // the class still ends in the slot program
            sa.push_back((uint32_t)c << 8);
            sb.push_back(0u);
        }
        sa.back() |= 128u;
        if (ra.size() > r0) ra.back() |= 128u;
    }
    if (sa.size() + ra.size() > 48) return false;
    out = dh;
    std::memset(out.prog_a, 0, sizeof(uint32_t) * 48);
    std::memset(out.prog_b, 0, sizeof(uint32_t) * 48);
    for (size_t i = 0; i < sa.size(); ++i) {
        out.prog_a[i] = sa[i];
        out.prog_b[i] = sb[i];
    }
    for (size_t i = 0; i < ra.size(); ++i) {
        out.prog_a[sa.size() + i] = ra[i];
        out.prog_b[sa.size() + i] = rb[i];
    }
    out.nprog = (int)sa.size();
    out.nrest = (int)ra.size();
    return true;
}

int ensure_s12_plan(pned_ctx *ctx, int colbits, int dtype)
{
    if (ctx->plan_colbits == colbits && ctx->plan_dtype == dtype) return PNED_OK;
    ctx->plan_req_colbits = -1;   // set again by ensure_s12_plan_cols once it has chosen
    // the whole tile sits in LDS (160 KB on gfx950): 2^13 complex128 or 2^14 complex64.
    // Untuned default (the autotuner's usual choice): vectors of up to 2^25 elements two
    // bits below that limit (small tiles, 256 threads, more blocks in flight), larger
    // vectors one bit below (1024 threads)
    const int kmax = (dtype == PNED_C128) ? 13 : 14;
    int K = ctx->tune_tile_bits;
    const bool untuned = K <= 0;
    if (untuned) K = default_large(ctx, colbits) ? kmax - 1 : kmax - 2;
    if (K > kmax) K = kmax;
    int L = ctx->tune_low_bits > 0 ? ctx->tune_low_bits : 3;
    PNED_TRY(s12_make_passes(ctx->s12, colbits, K, L));
    if (untuned && K < kmax) {
        // one tile bit more when it saves a whole pass (a pass streams v1 and v0 once)
        S12Plan wider;
        wider.n = ctx->s12.n;
        wider.terms = ctx->s12.terms;
        PNED_TRY(s12_make_passes(wider, colbits, K + 1, L));
        if (wider.passes.size() < ctx->s12.passes.size()) PNED_TRY(s12_make_passes(ctx->s12, colbits, K + 1, L));
    }
    S12Plan &pl = ctx->s12;
    const size_t nt = std::max<size_t>(pl.devterms.size(), 1);
    PNED_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->d_terms) PNED_HIP(hipFree(ctx->d_terms));
    ctx->d_terms = nullptr;
    PNED_HIP(hipMalloc((void **)&ctx->d_terms, sizeof(DevTerm) * nt));
    ctx->d_terms_cap = nt;
    if (!pl.devterms.empty())
        PNED_HIP(hipMemcpy(ctx->d_terms, pl.devterms.data(), sizeof(DevTerm) * pl.devterms.size(),
                           hipMemcpyHostToDevice));
    DiagDev &dh = ctx->diag_host;
    std::memset(&dh, 0, sizeof(dh));
    dh.ngroups = (int)pl.diag.size();
    for (int g = 0; g < dh.ngroups; ++g) {
        dh.kind[g] = pl.diag[g].kind;
        dh.shift[g] = pl.diag[g].shift;
        dh.mask[g] = pl.diag[g].mask;
        dh.wr[g] = pl.diag[g].wr;
        dh.wi[g] = pl.diag[g].wi;
    }
    dh.c0r = pl.c0r;
    dh.c0i = pl.c0i;
    build_diag_program(dh);
    if (!ctx->d_diag) PNED_HIP(hipMalloc((void **)&ctx->d_diag, sizeof(DiagDev)));
    PNED_HIP(hipMemcpy(ctx->d_diag, &dh, sizeof(DiagDev), hipMemcpyHostToDevice));
    ctx->plan_colbits = colbits;
    ctx->plan_dtype = dtype;
    ctx->pd_threads = -1;    // address tables follow the new plan
    return upload_batch_coefficients(ctx);
}

int ensure_s12_plan_cols(pned_ctx *ctx, int colbits, int run_bits, int dtype, int *cg)
{
    if (ctx->plan_req_colbits == colbits && ctx->plan_colbits >= 0 && ctx->plan_dtype == dtype) {
        *cg = ctx->plan_colbits;
        return PNED_OK;
    }
    PNED_TRY(ensure_s12_plan(ctx, colbits, dtype));
    *cg = colbits;
    if (ctx->s12.nop <= 1 && colbits > run_bits) {
        const size_t whole = ctx->s12.passes.size();
        PNED_TRY(ensure_s12_plan(ctx, run_bits, dtype));
        if (ctx->s12.passes.size() < whole) *cg = run_bits;
        else PNED_TRY(ensure_s12_plan(ctx, colbits, dtype));   // the same plan again (deterministic)
    }
    ctx->plan_req_colbits = colbits;
    return PNED_OK;
}

int upload_batch_coefficients(pned_ctx *ctx)
{
    const S12Plan &pl = ctx->s12;
    if (pl.nop <= 1) return PNED_OK;
    PNED_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->d_colcoef) PNED_HIP(hipFree(ctx->d_colcoef));
    if (ctx->d_coldiag) PNED_HIP(hipFree(ctx->d_coldiag));
    ctx->d_colcoef = ctx->d_coldiag = nullptr;
    if (!pl.devcoef.empty()) {
        PNED_HIP(hipMalloc((void **)&ctx->d_colcoef, sizeof(double) * pl.devcoef.size()));
        PNED_HIP(hipMemcpy(ctx->d_colcoef, pl.devcoef.data(), sizeof(double) * pl.devcoef.size(),
                           hipMemcpyHostToDevice));
    }
    PNED_HIP(hipMalloc((void **)&ctx->d_coldiag, sizeof(double) * pl.coldiag.size()));
    PNED_HIP(hipMemcpy(ctx->d_coldiag, pl.coldiag.data(), sizeof(double) * pl.coldiag.size(), hipMemcpyHostToDevice));
    return PNED_OK;
}

int check_hv_pair(const pned_ctx *ctx, const pned_vec *v1, const pned_vec *v0)
{
    if (!ctx || !v1 || !v0) return fail(PNED_E_ARG, "null argument");
    if (v1->ctx != ctx || v0->ctx != ctx) return fail(PNED_E_ARG, "vector belongs to another context");
    if (v1->dim != v0->dim || v1->nvec != v0->nvec || v1->dtype != v0->dtype)
        return fail(PNED_E_ARG, "v1 and v0 differ in shape or type");
    if (v1 == v0) return fail(PNED_E_ARG, "v1 and v0 must be different vectors");
    if (v1->dim != ctx->dim) return fail(PNED_E_ARG, "vector dimension differs from the context's basis");
    return PNED_OK;
}

static int hv_dispatch(pned_ctx *ctx, const pned_vec *v1, pned_vec *v0, double eperbd, bool fresh, double *prdct)
{
    PNED_TRY(check_hv_pair(ctx, v1, v0));
    if (ctx->sector_k >= 0) return launch_hv_sector(ctx, v1, v0, eperbd, fresh, prdct);
    if (ctx->q == 2) return launch_hv_s12(ctx, v1, v0, eperbd, fresh, prdct);
    return launch_hv_spin(ctx, v1, v0, eperbd, fresh, prdct);
}

}  // namespace pned

using namespace pned;

extern "C" {

const char *pned_last_error(void) { return g_err.c_str(); }
int pned_version(void) { return 1; }

int pned_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

static int ctx_create(int n_sites, int q_spin, int ibond, const int32_t *ipair, const int32_t *ibond_type,
                      const double *xlocal_ham, const int32_t *llocal_ham, int device, bool pick_layout,
                      pned_ctx **out);

int pned_ctx_create(int n_sites, int q_spin, int ibond, const int32_t *ipair, const int32_t *ibond_type,
                    const double *xlocal_ham, const int32_t *llocal_ham, int device, pned_ctx **out)
{
    return ctx_create(n_sites, q_spin, ibond, ipair, ibond_type, xlocal_ham, llocal_ham, device, true, out);
}

static int ctx_create(int n_sites, int q_spin, int ibond, const int32_t *ipair, const int32_t *ibond_type,
                      const double *xlocal_ham, const int32_t *llocal_ham, int device, bool pick_layout,
                      pned_ctx **out)
{
    if (!out || !ipair || !ibond_type || !xlocal_ham || !llocal_ham) return fail(PNED_E_ARG, "null argument");
    *out = nullptr;
    if (n_sites < 2 || q_spin < 2 || ibond < 1) return fail(PNED_E_ARG, "bad sizes");
    if (q_spin == 2 && n_sites > 32) return fail(PNED_E_UNSUPPORTED, "S=1/2: at most 32 sites on one GPU");
    double dimd = std::pow((double)q_spin, (double)n_sites);
    if (dimd > 9.3e18) return fail(PNED_E_UNSUPPORTED, "dimension overflows 64 bits");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(PNED_E_HIP, "no HIP device available (libpned_gpu needs an MI355X)");
    if (device < 0 || device >= ndev) return fail(PNED_E_ARG, "device index out of range");
    PNED_HIP(hipSetDevice(device));
    PNED_TRY(noalloc_scratch_init(device));
    // datack (Hybrid_TPS_basic.f90:8-21)
    for (int k = 0; k < 2 * ibond; ++k)
        if (ipair[k] <= 0 || ipair[k] > n_sites)
            return fail(PNED_E_ARG, "#(E03) Incorrect data in ipair at " + std::to_string(k + 1));
    auto *ctx = new pned_ctx();
    ctx->device = device;
    ctx->n = n_sites;
    ctx->q = q_spin;
    ctx->ibond = ibond;
    int64_t d = 1;
    for (int k = 0; k < n_sites; ++k) d *= q_spin;
    ctx->dim = d;
    ctx->ipair.assign(ipair, ipair + 2 * ibond);
    ctx->btype.assign(ibond_type, ibond_type + ibond);
    const size_t qq = (size_t)q_spin * q_spin;
    ctx->xl.assign(xlocal_ham, xlocal_ham + 2 * qq * qq * 9);
    ctx->ll.assign(llocal_ham, llocal_ham + qq * qq * 9);
    // the tiled spin-S kernel keeps 32-bit tile offsets; larger spaces use the gather kernel
    if (d > (int64_t(1) << 32)) ctx->spin_tiled = 0;
    ctx->site_pair.resize(2 * ibond);
    for (int b = 0; b < ibond; ++b) {
        const int a = ipair[2 * b] - 1, c = ipair[2 * b + 1] - 1;
        ctx->site_pair[2 * b] = std::min(a, c);
        ctx->site_pair[2 * b + 1] = std::max(a, c);
    }
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return fail(PNED_E_HIP, "hipStreamCreate failed");
    }
    if (q_spin == 2) {
        int rc = s12_decompose(n_sites, ibond, ipair, ibond_type, xlocal_ham, llocal_ham, ctx->s12);
        if (rc == PNED_OK && pick_layout) {
            ctx->lay = choose_layout(ctx->s12);
            if (!ctx->lay.identity())
                rc = s12_decompose(n_sites, ibond, ipair, ibond_type, xlocal_ham, llocal_ham, ctx->s12, ctx->lay.pos);
        }
        if (rc != PNED_OK) {
            pned_ctx_destroy(ctx);
            return rc;
        }
    }
    *out = ctx;
    return PNED_OK;
}

int pned_ctx_create_batch(int n_sites, int q_spin, int ibond, const int32_t *ipair, const int32_t *ibond_type,
                          int nop, const double *xlocal_ham, const int32_t *llocal_ham, int device, pned_ctx **out)
{
    if (q_spin != 2) return fail(PNED_E_UNSUPPORTED, "batch contexts are S=1/2");
    if (nop < 1 || nop > 64 || (nop & (nop - 1))) return fail(PNED_E_ARG, "nop must be a power of two <= 64");
    PNED_TRY(ctx_create(n_sites, q_spin, ibond, ipair, ibond_type, xlocal_ham, llocal_ham, device, false, out));
    pned_ctx *ctx = *out;
    const size_t nx = 2 * 16 * 9, nl = 16 * 9;
    std::vector<const double *> xs(nop);
    std::vector<const int32_t *> ls(nop);
    for (int o = 0; o < nop; ++o) {
        xs[o] = xlocal_ham + o * nx;
        ls[o] = llocal_ham + o * nl;
    }
    int rc = s12_decompose_batch(n_sites, ibond, ipair, ibond_type, nop, xs.data(), ls.data(), ctx->s12);
    if (rc != PNED_OK) {
        pned_ctx_destroy(ctx);
        *out = nullptr;
        return rc;
    }
    ctx->plan_colbits = -1;
    return PNED_OK;
}

int pned_ctx_set_column_matrices(pned_ctx *ctx, int col, const double *xlocal_ham, const int32_t *llocal_ham)
{
    if (!ctx || !xlocal_ham || !llocal_ham) return fail(PNED_E_ARG, "null argument");
    if (ctx->q != 2 || ctx->s12.nop <= 1) return fail(PNED_E_ARG, "not a batch context");
    PNED_TRY(s12_set_column(ctx->s12, ctx->ibond, ctx->ipair.data(), ctx->btype.data(), col, xlocal_ham, llocal_ham));
    if (ctx->plan_colbits >= 0) PNED_TRY(upload_batch_coefficients(ctx));
    return PNED_OK;
}

int pned_ctx_destroy(pned_ctx *ctx)
{
    if (!ctx) return PNED_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (pned_vec *v : ctx->vecs) v->ctx = nullptr;   // they stay valid for pned_vec_destroy
    if (ctx->d_terms) (void)hipFree(ctx->d_terms);
    if (ctx->d_diag) (void)hipFree(ctx->d_diag);
    if (ctx->d_diag_tm) (void)hipFree(ctx->d_diag_tm);
    if (ctx->pd_dev) (void)hipFree(ctx->pd_dev);
    if (ctx->d_tm) (void)hipFree(ctx->d_tm);
    free_sector(ctx);
    if (ctx->d_colcoef) (void)hipFree(ctx->d_colcoef);
    if (ctx->d_coldiag) (void)hipFree(ctx->d_coldiag);
    if (ctx->d_partial) (void)hipFree(ctx->d_partial);
    if (ctx->d_result) (void)hipFree(ctx->d_result);
    if (ctx->h_result) (void)hipHostFree(ctx->h_result);
    if (ctx->d_scalars) (void)hipFree(ctx->d_scalars);
    if (ctx->d_stage) (void)hipFree(ctx->d_stage);
    if (ctx->spin.pairs) (void)hipFree(ctx->spin.pairs);
    if (ctx->spin.btype) (void)hipFree(ctx->spin.btype);
    reset_spin_tables(ctx);
    free_spin_plan(ctx);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return PNED_OK;
}

int64_t pned_ctx_dim(const pned_ctx *ctx) { return ctx ? ctx->dim : -1; }

// New bond matrices whose terms and diagonal groups fit the current single-operator plan
// (same rule as a batch column): replace the coefficients and diagonal weights and keep
// the tile passes, so a sequence of operators on one cluster is planned once.  Returns 1
// when the plan has to be rebuilt instead.
static int refit_plan(pned_ctx *ctx, const double *xl, const int32_t *ll)
{
    S12Plan &pl = ctx->s12;
    if (pl.nop != 1 || pl.passes.empty() || ctx->plan_colbits < 0 || pl.colcoef.empty()) return 1;
    S12Plan trial = pl;
    if (s12_set_column(trial, ctx->ibond, ctx->ipair.data(), ctx->btype.data(), 0, xl, ll, ctx->lay.pos) != PNED_OK)
        return 1;
    for (size_t t = 0; t < trial.terms.size(); ++t) {
        Term &T = trial.terms[t];
        T.c0r = trial.colcoef[t * 4 + 0];
        T.c0i = trial.colcoef[t * 4 + 1];
        T.c1r = trial.colcoef[t * 4 + 2];
        T.c1i = trial.colcoef[t * 4 + 3];
    }
    for (DevTerm &d : trial.devterms) {
        const Term &T = trial.terms[d.pad];
        d.c0r = T.c0r;
        d.c0i = T.c0i;
        d.c1r = T.c1r;
        d.c1i = T.c1i;
        d.is_real = (T.c0i == 0.0 && T.c1i == 0.0) ? 1u : 0u;
    }
    const size_t ng = trial.diag.size();
    for (size_t g = 0; g < ng; ++g) {
        trial.diag[g].wr = trial.coldiag[2 * g];
        trial.diag[g].wi = trial.coldiag[2 * g + 1];
    }
    trial.c0r = trial.coldiag[2 * ng];
    trial.c0i = trial.coldiag[2 * ng + 1];
    pl = std::move(trial);
    if (!pl.devterms.empty())
        PNED_HIP(hipMemcpy(ctx->d_terms, pl.devterms.data(), sizeof(DevTerm) * pl.devterms.size(),
                           hipMemcpyHostToDevice));
    DiagDev &dh = ctx->diag_host;
    for (size_t g = 0; g < ng; ++g) {
        dh.wr[g] = pl.diag[g].wr;
        dh.wi[g] = pl.diag[g].wi;
    }
    dh.c0r = pl.c0r;
    dh.c0i = pl.c0i;
    build_diag_program(dh);
    PNED_HIP(hipMemcpy(ctx->d_diag, &dh, sizeof(DiagDev), hipMemcpyHostToDevice));
    ctx->pd_threads = -1;   // the term-major table carries the coefficients
    return PNED_OK;
}

int pned_ctx_set_matrices(pned_ctx *ctx, const double *xlocal_ham, const int32_t *llocal_ham)
{
    if (!ctx || !xlocal_ham || !llocal_ham) return fail(PNED_E_ARG, "null argument");
    const size_t qq = (size_t)ctx->q * ctx->q;
    ctx->xl.assign(xlocal_ham, xlocal_ham + 2 * qq * qq * 9);
    ctx->ll.assign(llocal_ham, llocal_ham + qq * qq * 9);
    PNED_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->q == 2) {
        const int rc = refit_plan(ctx, xlocal_ham, llocal_ham);
        if (rc < 0) return rc;   // a HIP error
        if (rc > 0) {
            PNED_TRY(s12_decompose(ctx->n, ctx->ibond, ctx->ipair.data(), ctx->btype.data(), xlocal_ham, llocal_ham,
                                   ctx->s12, ctx->lay.pos));
            ctx->plan_colbits = -1;
            ctx->plan_dtype = -1;
        }
        if (ctx->sector_k >= 0) PNED_TRY(refresh_sector_terms(ctx));
    } else {
        reset_spin_tables(ctx);
        free_spin_plan(ctx);   // the tiled passes carry the coefficients
    }
    return PNED_OK;
}

int pned_ctx_spin_kernel(pned_ctx *ctx, int tiled)
{
    if (!ctx) return fail(PNED_E_ARG, "null context");
    if (tiled && ctx->dim > (int64_t(1) << 32))
        return fail(PNED_E_UNSUPPORTED, "tiled spin-S kernel: at most 2^32 states (32-bit tile offsets)");
    ctx->spin_tiled = tiled ? 1 : 0;
    return PNED_OK;
}

int pned_ctx_graphs(pned_ctx *ctx, int enable)
{
    if (!ctx) return fail(PNED_E_ARG, "null context");
    ctx->use_graphs = enable ? 1 : 0;
    return PNED_OK;
}

int pned_ctx_s12_kernel(pned_ctx *ctx, int kind)
{
    if (!ctx) return fail(PNED_E_ARG, "null context");
    if (kind != 0 && kind != 1) return fail(PNED_E_ARG, "kernel kind 0 or 1");
    ctx->s12_kernel = kind;
    return PNED_OK;
}

int pned_ctx_tune(pned_ctx *ctx, int tile_bits, int low_bits, int threads)
{
    if (!ctx) return fail(PNED_E_ARG, "null context");
    if (tile_bits < 0 || tile_bits > 14 || low_bits < 0 || low_bits > 8 ||
        (threads != 0 && threads != 256 && threads != 512 && threads != 1024))
        return fail(PNED_E_ARG, "tile_bits in 0..14, low_bits in 0..8, threads 0, 256, 512 or 1024");
    ctx->tune_tile_bits = tile_bits;
    ctx->tune_low_bits = low_bits;
    ctx->tune_threads = threads;
    ctx->plan_colbits = -1;
    free_spin_plan(ctx);
    return PNED_OK;
}

// Autotuner: times one fresh H.v v1 -> v0 for each candidate (tile bits, forced low bits,
// threads per block) that fits LDS, keeps the fastest (v0 is overwritten).
int pned_ctx_autotune(pned_ctx *ctx, const pned_vec *v1, pned_vec *v0, int reps, int *best_tile_bits,
                      int *best_low_bits, int *best_threads, double *best_ms)
{
    PNED_TRY(check_hv_pair(ctx, v1, v0));
    if (ctx->q != 2) {
        if (best_ms) *best_ms = 0.0;
        return PNED_OK;   // spin-S kernel has no tile knobs
    }
    if (reps < 1) reps = 2;
    const int kmax = (v1->dtype == PNED_C128) ? 13 : 14;
    hipEvent_t e0, e1;
    PNED_HIP(hipEventCreate(&e0));
    PNED_HIP(hipEventCreate(&e1));
    double best = 1e300;
    int bk = 0, bl = 0, bt = 0;
    for (int K = kmax - 2; K <= kmax; ++K)
        for (int L = 3; L <= 4; ++L)
            for (int T : {256, 512, 1024}) {
                if (T == 1024 && ctx->s12_kernel != 1) continue;   // term-major kernel only
                ctx->tune_tile_bits = K;
                ctx->tune_low_bits = L;
                ctx->tune_threads = T;
                ctx->plan_colbits = -1;
                int rc = launch_hv_s12(ctx, v1, v0, 0.0, true, nullptr);   // warm-up, builds the plan
                if (rc != PNED_OK) continue;
                PNED_HIP(hipEventRecord(e0, ctx->stream));
                for (int k = 0; k < reps; ++k) PNED_TRY(launch_hv_s12(ctx, v1, v0, 0.0, true, nullptr));
                PNED_HIP(hipEventRecord(e1, ctx->stream));
                PNED_HIP(hipEventSynchronize(e1));
                float ms = 0.f;
                PNED_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (ms / reps < best) { best = ms / reps; bk = K; bl = L; bt = T; }
            }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ctx->tune_tile_bits = bk;
    ctx->tune_low_bits = bl;
    ctx->tune_threads = bt;
    ctx->plan_colbits = -1;
    if (best_tile_bits) *best_tile_bits = bk;
    if (best_low_bits) *best_low_bits = bl;
    if (best_threads) *best_threads = bt;
    if (best_ms) *best_ms = best;
    return PNED_OK;
}

int pned_ctx_plan_info(const pned_ctx *ctx, int *npasses, int *terms_per_pass, int max)
{
    if (!ctx || !npasses) return fail(PNED_E_ARG, "null argument");
    if (ctx->q != 2) {
        if (ctx->spin_tiled && ctx->spin_pd) {   // tiled passes: bonds per pass
            *npasses = ctx->spin_np;
            for (int p = 0; p < ctx->spin_np && p < max && terms_per_pass; ++p) terms_per_pass[p] = ctx->spin_nb[p];
            return PNED_OK;
        }
        *npasses = 1;
        if (terms_per_pass && max > 0) terms_per_pass[0] = ctx->ibond;
        return PNED_OK;
    }
    *npasses = (int)ctx->s12.passes.size();
    for (int p = 0; p < *npasses && p < max && terms_per_pass; ++p) terms_per_pass[p] = ctx->s12.passes[p].nterm;
    return PNED_OK;
}

int pned_s12_plan_export(int n_sites, int ibond, const int32_t *ipair, const int32_t *ibond_type,
                         const double *xlocal_ham, const int32_t *llocal_ham, int colbits, int tile_bits,
                         int low_bits, int *npass, int *nterm, int *ngroup, int32_t *pass_info,
                         int32_t *pass_bits, double *terms, uint32_t *term_gmask, double *diag, int max_pass,
                         int max_term, const int32_t *site_bit)
{
    if (!ipair || !ibond_type || !xlocal_ham || !llocal_ham || !npass || !nterm || !ngroup)
        return fail(PNED_E_ARG, "null argument");
    Layout lay;
    if (site_bit) PNED_TRY(layout_from_table(n_sites, site_bit, lay));
    S12Plan pl;
    PNED_TRY(s12_decompose(n_sites, ibond, ipair, ibond_type, xlocal_ham, llocal_ham, pl, lay.pos));
    PNED_TRY(s12_make_passes(pl, colbits, tile_bits, low_bits));
    *npass = (int)pl.passes.size();
    *nterm = (int)pl.devterms.size();
    *ngroup = (int)pl.diag.size();
    if (*npass > max_pass || *nterm > max_term) return fail(PNED_E_ARG, "output arrays too small");
    // term gmask in devterm order: recover from tmask and the pass's tile bits
    for (int p = 0; p < *npass; ++p) {
        const PassPlan &pp = pl.passes[p];
        pass_info[4 * p] = pp.K;
        pass_info[4 * p + 1] = pp.first_term;
        pass_info[4 * p + 2] = pp.nterm;
        pass_info[4 * p + 3] = pp.diag ? 1 : 0;
        for (int k = 0; k < 40; ++k) pass_bits[40 * p + k] = k < pp.K ? pp.tb[k] : -1;
        for (int t = pp.first_term; t < pp.first_term + pp.nterm; ++t) {
            const DevTerm &d = pl.devterms[t];
            uint64_t m = 0;
            for (int k = 0; k < pp.K; ++k)
                if (d.tmask >> k & 1) m |= 1ull << pp.tb[k];
            term_gmask[t] = (uint32_t)(m >> colbits);
            terms[6 * t] = d.tmask;
            terms[6 * t + 1] = d.selmask;
            terms[6 * t + 2] = d.c0r;
            terms[6 * t + 3] = d.c0i;
            terms[6 * t + 4] = d.c1r;
            terms[6 * t + 5] = d.c1i;
        }
    }
    for (int g = 0; g < *ngroup; ++g) {
        diag[5 * g] = pl.diag[g].kind;
        diag[5 * g + 1] = pl.diag[g].shift;
        diag[5 * g + 2] = pl.diag[g].mask;
        diag[5 * g + 3] = pl.diag[g].wr;
        diag[5 * g + 4] = pl.diag[g].wi;
    }
    diag[5 * *ngroup] = pl.c0r;
    diag[5 * *ngroup + 1] = pl.c0i;
    return PNED_OK;
}

int pned_s12_layout_choose(int n_sites, int ibond, const int32_t *ipair, const int32_t *ibond_type,
                           const double *xlocal_ham, const int32_t *llocal_ham, int32_t *site_bit)
{
    if (!ipair || !ibond_type || !xlocal_ham || !llocal_ham || !site_bit) return fail(PNED_E_ARG, "null argument");
    if (n_sites < 2 || n_sites > 32) return fail(PNED_E_ARG, "bad site count");
    S12Plan pl;
    PNED_TRY(s12_decompose(n_sites, ibond, ipair, ibond_type, xlocal_ham, llocal_ham, pl));
    const Layout lay = choose_layout(pl);
    for (int k = 0; k < n_sites; ++k) site_bit[k] = lay.pos[k];
    return PNED_OK;
}

int pned_ctx_layout(const pned_ctx *ctx, int32_t *site_bit)
{
    if (!ctx || !site_bit) return fail(PNED_E_ARG, "null argument");
    for (int k = 0; k < ctx->n; ++k) site_bit[k] = ctx->lay.pos[k];
    return PNED_OK;
}

int pned_ctx_set_layout(pned_ctx *ctx, const int32_t *site_bit)
{
    if (!ctx) return fail(PNED_E_ARG, "null context");
    if (ctx->q != 2 || ctx->s12.nop != 1 || ctx->sector_k >= 0)
        return fail(PNED_E_UNSUPPORTED, "memory layouts apply to single-operator S=1/2 full-space contexts");
    if (!ctx->vecs.empty()) return fail(PNED_E_ARG, "set the layout before creating the context's vectors");
    Layout lay;
    if (site_bit) PNED_TRY(layout_from_table(ctx->n, site_bit, lay));
    PNED_HIP(hipStreamSynchronize(ctx->stream));
    PNED_TRY(reset_layout(ctx));
    if (lay.identity()) return PNED_OK;
    PNED_TRY(s12_decompose(ctx->n, ctx->ibond, ctx->ipair.data(), ctx->btype.data(), ctx->xl.data(), ctx->ll.data(),
                           ctx->s12, lay.pos));
    ctx->lay = lay;
    ctx->plan_colbits = -1;
    ctx->plan_dtype = -1;
    ctx->pd_threads = -1;
    return PNED_OK;
}

int pned_ctx_sync(pned_ctx *ctx)
{
    if (!ctx) return fail(PNED_E_ARG, "null context");
    PNED_HIP(hipStreamSynchronize(ctx->stream));
    return PNED_OK;
}

void *pned_ctx_stream(pned_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int pned_vec_create(pned_ctx *ctx, int nvec, int dtype, pned_vec **out)
{
    if (!ctx || !out) return fail(PNED_E_ARG, "null argument");
    *out = nullptr;
    if (nvec < 1 || (dtype != PNED_C128 && dtype != PNED_C64)) return fail(PNED_E_ARG, "bad nvec or dtype");
    auto *v = new pned_vec();
    v->ctx = ctx;
    v->dim = ctx->dim;
    v->nvec = nvec;
    v->dtype = dtype;
    v->bytes = (size_t)ctx->dim * nvec * elem_bytes(dtype);
    PNED_HIP(hipSetDevice(ctx->device));
    if (hipMalloc(&v->data, v->bytes) != hipSuccess) {
        (void)hipGetLastError();
        delete v;
        return fail(PNED_E_NOMEM, "device allocation of " + std::to_string(v->bytes) + " bytes failed");
    }
    ctx->vecs.push_back(v);
    *out = v;
    return PNED_OK;
}

int pned_vec_destroy(pned_vec *v)
{
    if (!v) return PNED_OK;
    if (v->ctx) {
        if (v->ctx->stream) (void)hipStreamSynchronize(v->ctx->stream);
        auto &vs = v->ctx->vecs;
        vs.erase(std::remove(vs.begin(), vs.end(), v), vs.end());
    }
    if (v->data) (void)hipFree(v->data);
    delete v;
    return PNED_OK;
}

int pned_vec_info(const pned_vec *v, int64_t *dim, int *nvec, int *dtype)
{
    if (!v) return fail(PNED_E_ARG, "null vector");
    if (dim) *dim = v->dim;
    if (nvec) *nvec = v->nvec;
    if (dtype) *dtype = v->dtype;
    return PNED_OK;
}

void *pned_vec_data(pned_vec *v) { return v ? v->data : nullptr; }

int pned_vec_upload(pned_ctx *ctx, pned_vec *v, const void *host, int64_t offset, int64_t count)
{
    if (!ctx || !v || !host) return fail(PNED_E_ARG, "null argument");
    if (vec_detached(v)) return fail(PNED_E_ARG, "vector of a destroyed context");
    if (offset < 0 || count < 0 || offset + count > v->dim) return fail(PNED_E_ARG, "range outside the vector");
    if (v->ctx != ctx) return fail(PNED_E_ARG, "vector belongs to another context");
    return vec_from_host(v, host, offset, count);
}

int pned_vec_download(pned_ctx *ctx, const pned_vec *v, void *host, int64_t offset, int64_t count)
{
    if (!ctx || !v || !host) return fail(PNED_E_ARG, "null argument");
    if (vec_detached(v)) return fail(PNED_E_ARG, "vector of a destroyed context");
    if (offset < 0 || count < 0 || offset + count > v->dim) return fail(PNED_E_ARG, "range outside the vector");
    if (v->ctx != ctx) return fail(PNED_E_ARG, "vector belongs to another context");
    return vec_to_host(v, host, offset, count);
}

int pned_vec_zero(pned_ctx *ctx, pned_vec *v)
{
    if (!ctx || !v) return fail(PNED_E_ARG, "null argument");
    if (vec_detached(v)) return fail(PNED_E_ARG, "vector of a destroyed context");
    PNED_HIP(hipMemsetAsync(v->data, 0, v->bytes, ctx->stream));
    return PNED_OK;
}

int pned_vec_copy(pned_ctx *ctx, const pned_vec *src, pned_vec *dst)
{
    if (!ctx || !src || !dst) return fail(PNED_E_ARG, "null argument");
    if (vec_detached(src) || vec_detached(dst)) return fail(PNED_E_ARG, "vector of a destroyed context");
    if (src->bytes != dst->bytes || src->dtype != dst->dtype) return fail(PNED_E_ARG, "shape mismatch");
    PNED_TRY(check_same_layout(src, dst));
    PNED_HIP(hipMemcpyAsync(dst->data, src->data, src->bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return PNED_OK;
}

int pned_set_initial_vec(pned_ctx *ctx, pned_vec *v, uint64_t seed)
{
    if (!ctx || !v) return fail(PNED_E_ARG, "null argument");
    if (vec_detached(v)) return fail(PNED_E_ARG, "vector of a destroyed context");
    return gaussian_fill(ctx, v, seed);
}

int pned_p_mltply_v3(pned_ctx *ctx, const pned_vec *v1, pned_vec *v0, double eperbd, double *prdct)
{
    return hv_dispatch(ctx, v1, v0, eperbd, false, prdct);
}

int pned_hv_fresh(pned_ctx *ctx, const pned_vec *v1, pned_vec *v0, double eperbd, double *prdct)
{
    return hv_dispatch(ctx, v1, v0, eperbd, true, prdct);
}

int pned_calc_vec_inprdct(pned_ctx *ctx, const pned_vec *x, const pned_vec *y, double *out)
{
    if (!ctx || !x || !y || !out) return fail(PNED_E_ARG, "null argument");
    return dot(ctx, x, y, out);
}

int pned_calc_beta(pned_ctx *ctx, const double *alpha, const pned_vec *v1, const pned_vec *v0, double *b2)
{
    if (!ctx || !alpha || !v1 || !v0 || !b2) return fail(PNED_E_ARG, "null argument");
    return beta2(ctx, alpha, v1, v0, b2);
}

int pned_calc_alpha_beta(pned_ctx *ctx, pned_vec *v1, pned_vec *v0, double *alpha, double *beta)
{
    if (!ctx || !v1 || !v0 || !alpha || !beta) return fail(PNED_E_ARG, "null argument");
    const int nc = v1->nvec;
    std::vector<double> pr(2 * nc), b2(nc);
    PNED_TRY(hv_alpha_beta2(ctx, v1, v0, pr.data(), b2.data()));
    for (int c = 0; c < nc; ++c) {
        alpha[c] = pr[2 * c];
        beta[c] = std::sqrt(b2[c]);
    }
    return PNED_OK;
}

int pned_calc_alpha_beta2(pned_ctx *ctx, pned_vec *v1, pned_vec *v0, double *alpha, double *beta)
{
    if (!ctx || !v1 || !v0 || !alpha || !beta) return fail(PNED_E_ARG, "null argument");
    const int nc = v1->nvec;
    std::vector<double> pr(2 * nc), nn(2 * nc);
    PNED_TRY(hv_dispatch(ctx, v1, v0, 0.0, false, pr.data()));
    PNED_TRY(dot(ctx, v0, v0, nn.data()));
    for (int c = 0; c < nc; ++c) {
        alpha[c] = pr[2 * c];
        beta[c] = std::sqrt(nn[2 * c]);
    }
    return PNED_OK;
}

int pned_axpby(pned_ctx *ctx, const double *a, const pned_vec *x, const double *b, pned_vec *y)
{
    if (!ctx || !a || !x || !b || !y) return fail(PNED_E_ARG, "null argument");
    return axpby(ctx, a, x, b, y);
}

int pned_scale(pned_ctx *ctx, pned_vec *x, const double *a)
{
    if (!ctx || !x || !a) return fail(PNED_E_ARG, "null argument");
    if (vec_detached(x)) return fail(PNED_E_ARG, "vector of a destroyed context");
    return scale(ctx, x, a);
}

int pned_lanczos_update(pned_ctx *ctx, pned_vec *v1, pned_vec *v0, const double *alpha, const double *beta)
{
    if (!ctx || !v1 || !v0 || !alpha || !beta) return fail(PNED_E_ARG, "null argument");
    return lanczos_update(ctx, v1, v0, alpha, beta);
}

}  // extern "C"
