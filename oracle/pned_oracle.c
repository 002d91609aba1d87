/*
 * pned_oracle.c -- CPU restatement of the reference's H.v and Krylov algorithms.
 *
 * TEST INFRASTRUCTURE ONLY.  The HIP path is compared against this file: only tests/,
 * the smoke entry and the bench's cpu_baseline leg may load it.  The product path never
 * calls it.
 *
 * Every function names the reference routine it follows (paths relative to the
 * reference root; H/ = S_one_half/Hybrid_TPS_*.f90, S/ = S_arbitrary/Hybrid_TPS9_*.f90).
 *
 * Conventions shared with the product ABI:
 *   xl[t*qq*qq + row*qq + col] == xlocal_Ham(row, col, t+1)   (complex double)
 *   mk[t*qq*qq + row*qq + col] == llocal_Ham(row, col, t+1)   (1 = element skipped)
 *   site_pair: 0-based (min site, max site) per bond; local index = q*d(max) + d(min)
 *
 * Pinning (what checks each function, beyond retyping it) is listed in the header of
 * tests/test_oracle.py.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

typedef double complex zc;

/* ------------------------------------------------------------------------------ */
/* basis helpers                                                                   */
/* ------------------------------------------------------------------------------ */
static int64_t ipow64(int q, int n) { int64_t r = 1; for (int i = 0; i < n; ++i) r *= q; return r; }

/* ------------------------------------------------------------------------------ */
/* H.v: p_mltply_v3 -> local_prdct -> H_dot_v at np=1                              */
/*   H/parallel.f90:190-263 (element loop), 265-302 (decode), 415-440 (case 3)     */
/*   S/parallel.f90:184-261, 298-356 (base-q digits)                                */
/* One pass over all output states per (bond, input local state) element, in bond  */
/* order, exactly the reference's pass structure.  v0 += sum (h - eperbd d)^T v1.  */
/* Rows [row_lo, row_hi) only (full H.v: 0..D); prdct gets <v1|(H - ibond eperbd)|v1> */
/* restricted to those rows.                                                        */
/* ------------------------------------------------------------------------------ */
void orc_hv_rows(int n, int q, int ibond, const int *site_pair, const int *btype,
                 const zc *xl, const unsigned char *mk, double eperbd,
                 const zc *v1, zc *v0, zc *prdct, int64_t row_lo, int64_t row_hi)
{
    const int qq = q * q;
    double pr_re = 0.0, pr_im = 0.0;
    (void)n;
    for (int b = 0; b < ibond; ++b) {
        const int s1 = site_pair[2 * b], s2 = site_pair[2 * b + 1];
        const int t = btype[b] - 1;
        const int64_t is1 = ipow64(q, s1), is2 = ipow64(q, s2);
        for (int ml = 0; ml < qq; ++ml) {
            const int l1 = ml % q, l2 = ml / q;
            #pragma omp parallel for schedule(static) reduction(+:pr_re,pr_im)
            for (int64_t l0 = row_lo; l0 < row_hi; ++l0) {
                const int d1 = (int)((l0 / is1) % q);
                const int d2 = (int)((l0 / is2) % q);
                const int istat = q * d2 + d1;
                if (mk[(size_t)t * qq * qq + (size_t)ml * qq + istat]) continue;
                const int64_t it = l0 + (int64_t)(l1 - d1) * is1 + (int64_t)(l2 - d2) * is2;
                zc f = xl[(size_t)t * qq * qq + (size_t)ml * qq + istat];
                if (ml == istat) f -= eperbd;
                const zc w = f * v1[it];
                v0[l0] += w;
                const zc p = conj(v1[l0]) * w;
                pr_re += creal(p);
                pr_im += cimag(p);
            }
        }
    }
    *prdct = pr_re + I * pr_im;
}

void orc_hv(int n, int q, int ibond, const int *site_pair, const int *btype,
            const zc *xl, const unsigned char *mk, double eperbd,
            const zc *v1, zc *v0, zc *prdct)
{
    orc_hv_rows(n, q, ibond, site_pair, btype, xl, mk, eperbd, v1, v0, prdct, 0, ipow64(q, n));
}

/* ------------------------------------------------------------------------------ */
/* S^z sector basis (parity_type = 0): mk_list12, H/parallel.f90:452-578, at np=1   */
/* (no rank bits: psz = 0) and dsf_comp != 'xx' (is = 0).  The b_bit state bits     */
/* split into a right half of ihf = (b_bit+1)/2 bits and a left half of             */
/* ihf0 = b_bit/2 bits; states with popcount == tot_sz are enumerated left-major,   */
/* right-minor.  list1[i] = state of 0-based index i; la[ia] = 1-based rank of ia   */
/* among right halves of its popcount (list3(1,ia)); lb[ib] = index offset of the   */
/* left half ib's block (list3(2,ib)).  Index (1-based) of state s =                */
/* la[s & right] + lb[s >> ihf].  Returns ldim; list1 may be NULL to count only.    */
/* ------------------------------------------------------------------------------ */
int64_t orc_mk_list12(int n, int tot_sz, uint32_t *list1, uint32_t *la, uint32_t *lb)
{
    const int ihf0 = n / 2, ihf = (n + 1) / 2;
    int64_t icnt = 0, ja = 0, jb = 0;
    int64_t ibpatn = 0;
    for (int64_t i2 = 0; i2 < ((int64_t)1 << ihf0); ++i2) {
        for (int64_t i1 = 0; i1 < ((int64_t)1 << ihf); ++i1) {
            const int isz = __builtin_popcountll((unsigned long long)i1) + __builtin_popcountll((unsigned long long)i2);
            if (isz != tot_sz) continue;
            ++icnt;
            if (i2 == ibpatn) {
                ++ja;
            } else {
                ibpatn = i2;
                ja = 1;
                jb = icnt - 1;
            }
            if (list1) list1[icnt - 1] = (uint32_t)((i2 << ihf) + i1);
            if (la) la[i1] = (uint32_t)ja;
            if (lb) lb[i2] = (uint32_t)jb;
        }
    }
    return icnt;
}

/* H_dot_v over the sector basis (H/parallel.f90:305-442 with list_s = list1 and    */
/* list_lib = list3): per (bond, input local state) element, every output index l0  */
/* decodes its state, forms the partner state and looks its index up in the two    */
/* half tables.  S=1/2 only.  Returns -1 if an element not skipped by llocal_Ham    */
/* leaves the sector (the operator does not conserve S^z), else 0.                  */
int orc_hv_sector(int n, int ibond, const int *site_pair, const int *btype, const zc *xl, const unsigned char *mk,
                  double eperbd, int tot_sz, int64_t ldim, const uint32_t *list1, const uint32_t *la,
                  const uint32_t *lb, const zc *v1, zc *v0, zc *prdct)
{
    const int ihf = (n + 1) / 2;
    const uint32_t right = (1u << ihf) - 1u;
    double pr_re = 0.0, pr_im = 0.0;
    int bad = 0;
    for (int b = 0; b < ibond; ++b) {
        const int s1 = site_pair[2 * b], s2 = site_pair[2 * b + 1];
        const int t = btype[b] - 1;
        for (int ml = 0; ml < 4; ++ml) {
            const int l1 = ml % 2, l2 = ml / 2;
            #pragma omp parallel for schedule(static) reduction(+:pr_re,pr_im) reduction(|:bad)
            for (int64_t l0 = 0; l0 < ldim; ++l0) {
                const uint32_t st = list1[l0];
                const int d1 = (int)((st >> s1) & 1u), d2 = (int)((st >> s2) & 1u);
                const int istat = 2 * d2 + d1;
                if (mk[(size_t)t * 16 + (size_t)ml * 4 + istat]) continue;
                const int64_t pst = (int64_t)st + (int64_t)(l1 - d1) * ((int64_t)1 << s1) +
                                    (int64_t)(l2 - d2) * ((int64_t)1 << s2);
                if (__builtin_popcountll((unsigned long long)pst) != tot_sz) { bad = 1; continue; }
                const int64_t it = (int64_t)la[(uint32_t)pst & right] + (int64_t)lb[(uint32_t)pst >> ihf] - 1;
                zc f = xl[(size_t)t * 16 + (size_t)ml * 4 + istat];
                if (ml == istat) f -= eperbd;
                const zc w = f * v1[it];
                v0[l0] += w;
                const zc p = conj(v1[l0]) * w;
                pr_re += creal(p);
                pr_im += cimag(p);
            }
        }
    }
    *prdct = pr_re + I * pr_im;
    return bad ? -1 : 0;
}

/* ------------------------------------------------------------------------------ */
/* cTPQ post-processing, H/eval_mpi.f90:236-446 (program main + calc_fac, eval_q,   */
/* lsort, calc_ave).  Arrays in the reference's Fortran order: alpha(i,seed) at      */
/* i + nd1*seed, beta likewise, coef(i,m,seed) at i + nd1*(m + nd2*seed), 0-based.  */
/* ------------------------------------------------------------------------------ */
static void ctpq_calc_fac(int nd1, double bt, const double *beta_s, double *fac1, double *fac2)
{
    double y = 0.0, maxfac = -1e200;
    for (int i = 0; i < nd1; ++i) { fac1[i] = 0.0; fac2[i] = 0.0; }
    for (int i = 1; i <= 2 * (nd1 - 1); ++i) {
        const int m = ((i + 1) >> 1) + 1;                 /* 1-based */
        const double z = log(bt / (double)(2 * i) * beta_s[m - 1]);
        y += z;
        fac1[m - 1] = y;
        if (y > maxfac) maxfac = y;
    }
    for (int i = 1; i <= nd1 - 1; ++i) {
        fac1[i - 1] = exp(fac1[i - 1] - maxfac);
        fac2[i - 1] = fac1[i - 1] * bt * beta_s[i] / (double)(2 * i - 1);
    }
    fac1[nd1 - 1] = exp(fac1[nd1 - 1] - maxfac);
    fac2[nd1 - 1] = 0.0;
}

/* lsort: the reference's exchange sort, largest first */
static void ctpq_lsort_desc(int n, double *v)
{
    for (int j = 0; j < n - 1; ++j)
        for (int i = j; i < n; ++i)
            if (v[i] > v[j]) { const double t = v[i]; v[i] = v[j]; v[j] = t; }
}

double orc_ctpq_sample(int nd1, double bt, const double *beta_s, const double *c4, const double *c5)
{
    double *fac1 = malloc(sizeof(double) * nd1), *fac2 = malloc(sizeof(double) * nd1);
    double *t1 = malloc(sizeof(double) * nd1), *t2 = malloc(sizeof(double) * nd1);
    ctpq_calc_fac(nd1, bt, beta_s, fac1, fac2);
    double z = 0.0;
    for (int i = 0; i < nd1 - 1; ++i) {
        t1[i] = fac1[i] * c4[i];
        t2[i] = fac2[i] * c5[i];
        z += fac1[i] + fac2[i];
    }
    ctpq_lsort_desc(nd1 - 1, t1);
    ctpq_lsort_desc(nd1 - 1, t2);
    double x = 0.0, y = 0.0;
    for (int i = 0; i < nd1 - 1; ++i) x += t1[i];
    for (int i = 0; i < nd1 - 1; ++i) y += t2[i];
    free(fac1); free(fac2); free(t1); free(t2);
    return (x + y) / z;
}

/* out[4*(num_div+1)]: per temperature T, C = dE/dT, Var(C) proxy, dC/dT (the program's columns) */
void orc_ctpq(int nd1, int nd2, int seed_num, const double *alpha, const double *beta, const double *coef,
              double t_min, double t_max, int num_div, double *out)
{
    (void)alpha;
    const double log_t_max = log(t_max), log_t_min = log(t_min);
    const double delta = (log_t_max - log_t_min) / (double)num_div;
    double t_delta = 1e-4;
    if (t_min < t_delta) t_delta = t_min * 0.9;
    double *tmp = malloc(sizeof(double) * seed_num);
    for (int i1 = 0; i1 <= num_div; ++i1) {
        const double t0 = exp(log_t_min + (double)i1 * delta);
        double dat[9];
        for (int i2 = 0; i2 <= 2; ++i2) {
            double bt = 1.0 / t0;
            if (i2 == 0) bt = 1.0 / (t0 - t_delta);
            if (i2 == 2) bt = 1.0 / (t0 + t_delta);
            for (int sd = 0; sd < seed_num; ++sd) {
                const double *bs = beta + (size_t)nd1 * sd;
                const double *c4 = coef + (size_t)nd1 * (3 + (size_t)nd2 * sd);
                const double *c5 = coef + (size_t)nd1 * (4 + (size_t)nd2 * sd);
                tmp[sd] = orc_ctpq_sample(nd1, bt, bs, c4, c5);
            }
            double ave = 0.0, var = 0.0;
            for (int sd = 0; sd < seed_num; ++sd) { ave += tmp[sd]; var += tmp[sd] * tmp[sd]; }
            const double r1 = ave / (double)seed_num;
            dat[i2 * 3 + 0] = bt;
            dat[i2 * 3 + 1] = r1;
            dat[i2 * 3 + 2] = var / (double)seed_num - r1 * r1;
        }
        const double d_e = 0.5 * (dat[7] - dat[1]);
        const double spec = d_e / t_delta;
        const double var_spec = 0.5 * (dat[2] + dat[8]);
        const double spec_d1 = (dat[4] - dat[1]) / t_delta;
        const double spec_d2 = (dat[7] - dat[4]) / t_delta;
        const double del_c = (spec_d2 - spec_d1) / t_delta;
        out[4 * i1 + 0] = 1.0 / dat[3];
        out[4 * i1 + 1] = spec;
        out[4 * i1 + 2] = var_spec;
        out[4 * i1 + 3] = del_c;
    }
    free(tmp);
}

/* calc_vec_inprdct  H/parallel.f90:104-126: sum conj(x) y */
zc orc_dot(int64_t d, const zc *x, const zc *y)
{
    double re = 0.0, im = 0.0;
    #pragma omp parallel for schedule(static) reduction(+:re,im)
    for (int64_t i = 0; i < d; ++i) { zc p = conj(x[i]) * y[i]; re += creal(p); im += cimag(p); }
    return re + I * im;
}

/* calc_beta  H/parallel.f90:161-188: sum |v0 - alpha v1|^2 */
double orc_beta2(int64_t d, zc alpha, const zc *v1, const zc *v0)
{
    double s = 0.0;
    #pragma omp parallel for schedule(static) reduction(+:s)
    for (int64_t i = 0; i < d; ++i) { zc w = v0[i] - alpha * v1[i]; s += creal(conj(w) * w); }
    return s;
}

/* ------------------------------------------------------------------------------ */
/* Start vector.  The reference draws a thread-dependent random phase vector       */
/* (H/parallel.f90:17-102).  The build replaces it by i.i.d. N(0,1) real and       */
/* imaginary parts from Philox-4x32-10 (Salmon et al., SC'11, "Parallel random     */
/* numbers: as easy as 1, 2, 3"), keyed by the 64-bit seed, counter =              */
/* (index lo, index hi, column, 0), followed by Box-Muller, normalised in FP64.    */
/* ------------------------------------------------------------------------------ */
void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Gaussian pair for basis index i, column col: u1 in (0,1], u2 in [0,1) from 53-bit words */
zc orc_gauss_elem(uint64_t seed, uint64_t i, uint32_t col)
{
    uint32_t ctr[4] = {(uint32_t)i, (uint32_t)(i >> 32), col, 0u};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    orc_philox4x32_10(ctr, key, o);
    const uint64_t a = ((uint64_t)o[0] << 21) ^ (o[1] >> 11);   /* 53 bits */
    const uint64_t b = ((uint64_t)o[2] << 21) ^ (o[3] >> 11);
    const double u1 = ((double)(a & ((1ull << 53) - 1)) + 1.0) * 0x1.0p-53;
    const double u2 = (double)(b & ((1ull << 53) - 1)) * 0x1.0p-53;
    const double r = sqrt(-2.0 * log(u1));
    const double th = 6.283185307179586476925286766559 * u2;
    return r * cos(th) + I * (r * sin(th));
}

/* unnormalised fill; returns ||v||^2 */
double orc_gauss_fill(int64_t d, uint64_t seed, uint32_t col, zc *v)
{
    double s = 0.0;
    #pragma omp parallel for schedule(static) reduction(+:s)
    for (int64_t i = 0; i < d; ++i) { v[i] = orc_gauss_elem(seed, (uint64_t)i, col); s += creal(conj(v[i]) * v[i]); }
    return s;
}

void orc_set_initial_vec(int64_t d, uint64_t seed, uint32_t col, zc *v)
{
    const double s = orc_gauss_fill(d, seed, col, v);
    const double r = 1.0 / sqrt(s);
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < d; ++i) v[i] *= r;
}

/* ------------------------------------------------------------------------------ */
/* Tridiagonal eigen-solvers: basic.f90 bisec (H/basic.f90:25-107) and vec12       */
/* (H/basic.f90:112-213).  1-based arrays of the reference become 0-based here.     */
/* ------------------------------------------------------------------------------ */
int orc_bisec(const double *alpha, const double *beta, int ndim, double *E, int ne, double eps)
{
    if (ndim > 8000 || ne > ndim || ne < 0) return -1;
    double *b2 = (double *)malloc(sizeof(double) * (size_t)ndim);
    double range = fabs(alpha[0]) + fabs(beta[0]);
    for (int k = 1; k < ndim - 1; ++k)
        range = fmax(range, fabs(beta[k - 1]) + fabs(alpha[k]) + fabs(beta[k]));
    range = fmax(range, fabs(beta[ndim - 2]) + fabs(alpha[ndim - 1]));
    range = -range;
    b2[0] = 0.0;
    for (int i = 1; i < ndim; ++i) b2[i] = beta[i - 1] * beta[i - 1];
    const double epsabs = fabs(range) * eps;
    for (int i = 0; i < ne; ++i) E[i] = -range;
    double b = range;
    for (int k = 1; k <= ne; ++k) {
        double a = E[k - 1];
        for (int j = 0; j < 100; ++j) {
            const double c = (a + b) / 2.0;
            if (fabs(a - b) < epsabs) continue;
            int numneg = 0, ipass = 0;
            double g = 1.0;
            for (int i = 0; i < ndim; ++i) {
                if (ipass == 0) g = c - alpha[i] - b2[i] / g;
                else if (ipass == 1) ipass = 2;
                else { g = c - alpha[i]; ipass = 0; }
                if (ipass == 0) {
                    if (g < 0.0) ++numneg;
                    if (fabs(g) < fabs(b2[i] * epsabs * eps)) ipass = 1;
                }
            }
            numneg = ndim - numneg;
            if (numneg < k) b = c;
            else {
                a = c;
                const int hi = numneg < ne ? numneg : ne;
                for (int i = k; i <= hi; ++i) E[i - 1] = c;
            }
        }
    }
    free(b2);
    return 0;
}

/* v is column-major (ndim x nvec); returns Ritz coefficient vectors of the nvec lowest */
void orc_vec12(const double *alpha, const double *beta, int ndim, const double *E, int nvec, double *v)
{
    double *di = malloc(sizeof(double) * ndim), *bl = malloc(sizeof(double) * ndim);
    double *bu = malloc(sizeof(double) * ndim), *bv = malloc(sizeof(double) * ndim);
    double *cm = malloc(sizeof(double) * (ndim + 1));
    int *lex = malloc(sizeof(int) * ndim);
    int km = 0;
    for (int k = 0; k < nvec; ++k) {
        double *vk = v + (size_t)k * ndim;
        for (int j = 0; j < ndim; ++j) { di[j] = E[k] - alpha[j]; bl[j] = -beta[j]; bu[j] = -beta[j]; }
        for (int j = 0; j < ndim - 1; ++j) {
            if (fabs(di[j]) > fabs(bl[j])) {
                lex[j] = 0;
                if (fabs(di[j]) < 1e-13) di[j] = 1e-13;
                cm[j + 1] = bl[j] / di[j];
                di[j + 1] = di[j + 1] - cm[j + 1] * bu[j];
                bv[j] = 0.0;
            } else {
                lex[j] = 1;
                cm[j + 1] = di[j] / bl[j];
                di[j] = bl[j];
                const double s = bu[j];
                bu[j] = di[j + 1];
                bv[j] = bu[j + 1];
                di[j + 1] = s - cm[j + 1] * bu[j];
                bu[j + 1] = -cm[j + 1] * bv[j];
            }
        }
        if (fabs(di[ndim - 1]) < 1e-13) di[ndim - 1] = 1e-13;
        for (int j = 0; j < ndim; ++j) vk[j] = 1.0 / ((double)(j + 1) * 5.0);
        if (k == 0) km = k;
        else if (fabs(E[k] - E[km]) > 1e-13) km = k;
        else {
            for (int i = km; i < k; ++i) {
                double prd = 0.0;
                for (int j = 0; j < ndim; ++j) prd += v[(size_t)i * ndim + j] * vk[j];
                for (int j = 0; j < ndim; ++j) vk[j] -= prd * v[(size_t)i * ndim + j];
            }
        }
        for (int l = 1; l <= k - km + 3; ++l) {
            if (l != 1 || k != km) {
                for (int j = 0; j < ndim - 1; ++j) {
                    if (lex[j] == 0) vk[j + 1] -= cm[j + 1] * vk[j];
                    else { const double s = vk[j]; vk[j] = vk[j + 1]; vk[j + 1] = s - cm[j + 1] * vk[j]; }
                }
            }
            for (int j = ndim - 1; j >= 0; --j) {
                double s = vk[j];
                if (j <= ndim - 2) s -= bu[j] * vk[j + 1];
                if (j <= ndim - 3) s -= bv[j] * vk[j + 2];
                vk[j] = s / di[j];
            }
            double dn = 0.0;
            for (int j = 0; j < ndim; ++j) dn += vk[j] * vk[j];
            if (dn > 1e-13) dn = 1.0 / sqrt(dn);
            for (int j = 0; j < ndim; ++j) vk[j] *= dn;
        }
    }
    free(di); free(bl); free(bu); free(bv); free(cm); free(lex);
}

/* ------------------------------------------------------------------------------ */
/* Lanczos: parallelized_lnc1z  H/Lanczos.f90:48-163 (start vector injected: v1).   */
/* fixed_steps > 0 disables the convergence exit and runs exactly that many steps.  */
/* Returns the number of steps taken; alpha/beta get 'itr' entries; ene gets nvec.  */
/* ------------------------------------------------------------------------------ */
typedef struct {
    int n, q, ibond;
    const int *site_pair, *btype;
    const zc *xl;
    const unsigned char *mk;
} orc_op;

static void hv_op(const orc_op *op, double eperbd, const zc *v1, zc *v0, zc *pr)
{
    orc_hv(op->n, op->q, op->ibond, op->site_pair, op->btype, op->xl, op->mk, eperbd, v1, v0, pr);
}

int orc_lanczos(int n, int q, int ibond, const int *site_pair, const int *btype,
                const zc *xl, const unsigned char *mk,
                zc *v1, zc *v0, int max_itr, int nvec, int fixed_steps,
                double *alpha, double *beta, double *ene)
{
    orc_op op = {n, q, ibond, site_pair, btype, xl, mk};
    const int64_t d = ipow64(q, n);
    memset(v0, 0, sizeof(zc) * (size_t)d);
    zc pr;
    hv_op(&op, 0.0, v1, v0, &pr);
    double alpha1 = creal(pr);
    double beta1 = sqrt(orc_beta2(d, alpha1, v1, v0));
    alpha[0] = alpha1;
    beta[0] = beta1;
    double ebefor = 0.0, eps = 1e-25;
    const int last = fixed_steps > 0 ? fixed_steps : max_itr;
    for (int i = 2; i <= last; ++i) {
        #pragma omp parallel for schedule(static)
        for (int64_t j = 0; j < d; ++j) {
            const zc t1 = v1[j];
            const zc t2 = (v0[j] - alpha1 * v1[j]) / beta1;
            v0[j] = -beta1 * t1;
            v1[j] = t2;
        }
        hv_op(&op, 0.0, v1, v0, &pr);
        alpha1 = creal(pr);
        beta1 = sqrt(orc_beta2(d, alpha1, v1, v0));
        alpha[i - 1] = alpha1;
        beta[i - 1] = beta1;
        if (beta1 < 0.5e-30) return -i;
        if (fixed_steps > 0) continue;
        if (i > 4 * nvec && i % 5 == 0) {
            orc_bisec(alpha, beta, i, ene, nvec, eps);
            if (fabs((ebefor - ene[nvec - 1]) / ene[nvec - 1]) < 1e-13) return i;
            ebefor = ene[nvec - 1];
        }
        if (i == 4 * nvec) {
            eps = 1e-25;
            orc_bisec(alpha, beta, 4 * nvec, ene, nvec, eps);
            ebefor = ene[nvec - 1];
        }
    }
    if (fixed_steps > 0) orc_bisec(alpha, beta, last, ene, nvec, 1e-25);
    return last;
}

/* ------------------------------------------------------------------------------ */
/* Inverse iteration by CG: parallelized_inv1z H/cg.f90:18-100 with parallelized_cg1 */
/* H/cg.f90:104-197.  b is the (injected) start vector.  'saved_loop' reproduces    */
/* the implicitly SAVEd 'loop' flag of cg1 (H/cg.f90:118): once one cg1 call has    */
/* converged, later calls leave at their first 5-step check.                        */
/* Returns the inverse-iteration count, or -1 if it did not converge.               */
/* ------------------------------------------------------------------------------ */
static int cg1(const orc_op *op, double eig, zc *x, const zc *b, zc *p, zc *r, zc *y,
               int max_cg, int *saved_loop)
{
    const int64_t d = ipow64(op->q, op->n);
    const zc bcn = orc_dot(d, b, b);
    for (int64_t i = 0; i < d; ++i) { r[i] = b[i]; p[i] = b[i]; x[i] = 0.0; }
    const double eperbd = eig / (double)op->ibond;
    for (int itr = 1; itr <= max_cg; ++itr) {
        memset(y, 0, sizeof(zc) * (size_t)d);
        zc cdmm;
        hv_op(op, eperbd, p, y, &cdmm);
        const zc crp = orc_dot(d, p, r);
        const zc cyp = orc_dot(d, p, y);
        const zc calpha = crp / cyp;
        const zc rc = orc_dot(d, r, r);
        #pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < d; ++i) { x[i] += calpha * p[i]; r[i] -= calpha * y[i]; }
        const zc rc2 = orc_dot(d, r, r);
        const double bt = creal(rc2) / creal(rc);
        #pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < d; ++i) p[i] = r[i] + bt * p[i];
        if (itr % 5 != 0) continue;
        if (sqrt(cabs(rc2)) < 1e-9 * sqrt(cabs(bcn))) *saved_loop = 1;
        if (*saved_loop) return itr;
    }
    return max_cg;
}

int orc_inverse_iteration(int n, int q, int ibond, const int *site_pair, const int *btype,
                          const zc *xl, const unsigned char *mk, double eig,
                          zc *x, zc *b, zc *p, zc *r, zc *y, int max_inv, int max_cg,
                          int reproduce_saved_loop)
{
    orc_op op = {n, q, ibond, site_pair, btype, xl, mk};
    const int64_t d = ipow64(q, n);
    int saved = 0;
    for (int itr = 1; itr <= max_inv; ++itr) {
        int local = reproduce_saved_loop ? saved : 0;
        cg1(&op, eig, x, b, p, r, y, max_cg, &local);
        if (reproduce_saved_loop) saved = local;
        const double xn = 1.0 / sqrt(creal(orc_dot(d, x, x)));
        for (int64_t i = 0; i < d; ++i) x[i] *= xn;
        const zc cxb = orc_dot(d, x, b);
        if (fabs(cabs(cxb) - 1.0) < 1e-12) return itr;
        memcpy(b, x, sizeof(zc) * (size_t)d);
    }
    return -1;
}

/* ------------------------------------------------------------------------------ */
/* Diagonal observables of mTPQ and DSF                                            */
/* ------------------------------------------------------------------------------ */
/* intalv_zz  H/DSF.f90:18-87: v = sum_j (s_j - 1/2) e^{i q.r_j} x (S=1/2, bit value s)
 * spin-S (S/DSF.f90 intalv_zz): weight (digit_j - S).  Not normalised here.         */
void orc_sz_q(int n, int q, const double *site_vec /* n x 2 */, const double qv[2],
              const zc *x, zc *v)
{
    const int64_t d = ipow64(q, n);
    const double ss = 0.5 * (q - 1);
    zc *ph = malloc(sizeof(zc) * n);
    for (int j = 0; j < n; ++j) {
        const double qn = qv[0] * site_vec[2 * j] + qv[1] * site_vec[2 * j + 1];
        ph[j] = cos(qn) + I * sin(qn);
    }
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < d; ++i) {
        zc acc = 0.0;
        int64_t s = i;
        for (int j = 0; j < n; ++j) {
            const int dj = (int)(s % q);
            s /= q;
            const double w = (q == 2) ? ((double)dj - 0.5) : ((double)dj - ss);
            acc += w * ph[j] * x[i];
        }
        v[i] = acc;
    }
    free(ph);
}

/* intalv_xy  H/DSF.f90:89-193 (S=1/2, one process): v = sum_j e^{i q.r_j} S^+_j x from
 * the sector with tot_sz set bits (x indexed by that sector's list1) to the sector with
 * tot_sz + 1 (v indexed by the target's la/lb tables).  Site-outer loop over the source
 * list, as the reference scatters; S^+_j sets bit j.  Not normalised here. */
void orc_intalv_xy(int n, const double *site_vec, const double qv[2], int64_t pdim, const uint32_t *listt,
                   const zc *x, int64_t ldim, const uint32_t *la, const uint32_t *lb, zc *v)
{
    const int ihf = (n + 1) / 2;
    const uint32_t right = (1u << ihf) - 1u;
    for (int64_t l = 0; l < ldim; ++l) v[l] = 0.0;
    for (int j = 0; j < n; ++j) {
        const double qn = qv[0] * site_vec[2 * j] + qv[1] * site_vec[2 * j + 1];
        const double c = cos(qn), s = sin(qn);
        for (int64_t i = 0; i < pdim; ++i) {
            const uint32_t st = listt[i];
            if ((st >> j) & 1u) continue;
            const uint32_t t = st + (1u << j);
            const int64_t it = (int64_t)la[t & right] + (int64_t)lb[t >> ihf] - 1;
            v[it] += (c * creal(x[i]) - s * cimag(x[i])) + I * (s * creal(x[i]) + c * cimag(x[i]));
        }
    }
}

/* intalv_xy in the full q^n space (parity_type != 0; S/DSF.f90:91-286 local_operation at
 * one process, H/DSF.f90:89-193 with listt = every state): site-outer scatter of
 * fac * e^{i q.r_j} x(i) onto the state whose digit j is moved by ipm.  q >= 3: the
 * spin-S ladder factor of the source digit d (m = d - S): sqrt((S-m)(S+m+1)) for ipm = +1,
 * sqrt((S+m)(S-m+1)) for ipm = -1.  q = 2 (the S=1/2 tree): bit j set (+1) or cleared
 * (-1), factor 1.  Not normalised here. */
void orc_intalv_pm(int n, int q, const double *site_vec, const double qv[2], int ipm, const zc *x, zc *v)
{
    const int64_t d = ipow64(q, n);
    const double ss = 0.5 * (q - 1);
    for (int64_t i = 0; i < d; ++i) v[i] = 0.0;
    int64_t base = 1;
    for (int j = 0; j < n; ++j, base *= q) {
        const double qn = qv[0] * site_vec[2 * j] + qv[1] * site_vec[2 * j + 1];
        const double c = cos(qn), s = sin(qn);
        for (int64_t i = 0; i < d; ++i) {
            const int is = (int)((i / base) % q);
            if (is + ipm < 0 || is + ipm > q - 1) continue;
            const double sm = (double)is - ss;
            const double fac = (q == 2) ? 1.0 : (ipm == 1 ? sqrt((ss - sm) * (ss + sm + 1.0))
                                                           : sqrt((ss + sm) * (ss - sm + 1.0)));
            const int64_t it = i + (int64_t)ipm * base;
            v[it] += fac * ((c * creal(x[i]) - s * cimag(x[i])) + I * (s * creal(x[i]) + c * cimag(x[i])));
        }
    }
}

/* calc_sq_zz / calc_sq_pp  H/TPS.f90:146-251 (S=1/2 tree), S/TPS.f90:152-247 (spin-S tree).
 * Both trees weight bit j of the state index (iand(ishft(istate,-j),1); in the spin-S tree
 * these are bits of the base-q index, not its digits) against bit 0:
 *   mode 0 (zz): (b_j - S)(b_0 - S) e^{iqr_j}, S = (q-1)/2;
 *   mode 1 (pp): my_xor(b_j, b_0) e^{iqr_j}, my_xor = [b != b0]/4 in the S=1/2 tree
 *   (H/TPS.f90:214-215) and sx + sy of the spin-S tree's ladder matrices
 *   (S/params.f90:1479-1495), whose (0,1) and (1,0) entries are (1 - i) su(0,1)/2 and
 *   (1 + i) sd(1,0)/2. */
static void my_xor_table(int q, zc t[2][2])
{
    t[0][0] = t[1][1] = 0.0;
    if (q == 2) {
        t[0][1] = t[1][0] = 0.25;
        return;
    }
    const double ss = 0.5 * (q - 1);
    const double su01 = sqrt((ss - (-ss)) * (ss + (-ss) + 1.0));       /* su(0,1), m = -S      */
    const double sd10 = sqrt((ss + (1.0 - ss)) * (ss - (1.0 - ss) + 1.0)); /* sd(1,0), m = 1 - S */
    t[0][1] = 0.5 * su01 - 0.5 * I * su01;
    t[1][0] = 0.5 * sd10 + 0.5 * I * sd10;
}

void orc_sq_apply(int n, int q, const double *site_vec, const double qv[2], int mode, const zc *x, zc *v)
{
    const int64_t d = ipow64(q, n);
    const double ss = 0.5 * (q - 1);
    zc tab[2][2];
    my_xor_table(q, tab);
    zc *ph = malloc(sizeof(zc) * n);
    for (int j = 0; j < n; ++j) {
        const double qn = qv[0] * site_vec[2 * j] + qv[1] * site_vec[2 * j + 1];
        ph[j] = cos(qn) + I * sin(qn);
    }
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < d; ++i) {
        zc acc = 0.0;
        const int s0 = (int)(i & 1);
        for (int j = 0; j < n; ++j) {
            const int sj = (int)((i >> j) & 1);
            const zc w = (mode == 0) ? ((double)sj - ss) * ((double)s0 - ss) : tab[sj][s0];
            acc += w * ph[j] * x[i];
        }
        v[i] = acc;
    }
    free(ph);
}

/* calc_zz_bond_corr  H/TPS.f90:270-321: for z-type bonds (type 3), weight
 * (b(is1)-1/2)(b(is2)-1/2) where the bit positions are the 1-BASED ipair entries
 * (reference off-by-one, kept); a position >= n reads bit (pos-n) of the rank id 0. */
void orc_zz_bond_apply(int n, int q, int ibond, const int *ipair, const int *btype, const zc *x, zc *v)
{
    const int64_t d = ipow64(q, n);   /* bits of the base-q index in the spin-S tree (S/TPS.f90:269-320) */
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < d; ++i) {
        double w = 0.0;
        for (int b = 0; b < ibond; ++b) {
            if (btype[b] != 3) continue;
            const int p1 = ipair[2 * b], p2 = ipair[2 * b + 1];
            const int s3 = p1 < n ? (int)((i >> p1) & 1) : 0;
            const int s4 = p2 < n ? (int)((i >> p2) & 1) : 0;
            w += ((double)s3 - 0.5) * ((double)s4 - 0.5);
        }
        v[i] = w * x[i];
    }
}

/* ------------------------------------------------------------------------------ */
/* mTPQ: parallelized_TPSz H/TPS.f90:325-409 with eval_phys_val H/TPS.f90:83-111.   */
/* The operator must already be the shifted (l - h_b) one (change_to_TPS_hamiltonian).
 * v1 holds the normalised start vector.  coef is row-major (steps x ncoef),        */
/* ncoef = 7 + sq_num.  Returns 0.                                                   */
/* ------------------------------------------------------------------------------ */
static void eval_phys(int n, int q, int ibond, const int *ipair, const int *btype, const double *site_vec,
                      int sq_num, const double *sq_q, int ict, double alp, double bet, double lmax,
                      const zc *vg, const zc *v1, const zc *v0, zc *vt, double *row)
{
    const int64_t d = ipow64(q, n);
    row[0] = (double)(2 * ict) / alp;
    row[1] = lmax;
    row[2] = bet;
    row[3] = lmax - alp;
    const zc val1 = orc_dot(d, vg, v0), val2 = orc_dot(d, vg, v1);
    row[4] = creal(lmax * val2 - val1);
    orc_zz_bond_apply(n, q, ibond, ipair, btype, v1, vt);
    row[5] = creal(orc_dot(d, v1, vt));
    row[6] = creal(orc_dot(d, vg, vt));
    double *sq = malloc(sizeof(double) * (4 * sq_num + 1));
    for (int i = 0; i < sq_num; ++i) {
        orc_sq_apply(n, q, site_vec, sq_q + 2 * i, 0, v1, vt);
        sq[4 * i + 0] = creal(orc_dot(d, v1, vt));
        sq[4 * i + 1] = creal(orc_dot(d, vg, vt));
        orc_sq_apply(n, q, site_vec, sq_q + 2 * i, 1, v1, vt);
        sq[4 * i + 2] = creal(orc_dot(d, v1, vt));
        sq[4 * i + 3] = creal(orc_dot(d, vg, vt));
    }
    /* the reference stores only sq(1..sq_num) (H/TPS.f90:106-109) */
    for (int i = 0; i < sq_num; ++i) row[7 + i] = sq[i];
    free(sq);
}

int orc_tps(int n, int q, int ibond, const int *site_pair, const int *btype, const int *ipair,
            const zc *xl, const unsigned char *mk, const double *site_vec,
            int sq_num, const double *sq_q, double lmax, int nsteps,
            zc *v1, zc *v0, zc *vg, zc *vt, double *alpha, double *beta, double *coef)
{
    orc_op op = {n, q, ibond, site_pair, btype, xl, mk};
    const int64_t d = ipow64(q, n);
    const int nc = 7 + sq_num;
    memset(v0, 0, sizeof(zc) * (size_t)d);
    memset(vg, 0, sizeof(zc) * (size_t)d);
    zc pr;
    hv_op(&op, 0.0, v1, v0, &pr);
    double a1 = creal(pr);
    double b1 = sqrt(creal(orc_dot(d, v0, v0)));
    alpha[0] = lmax - a1;
    beta[0] = b1;
    eval_phys(n, q, ibond, ipair, btype, site_vec, sq_num, sq_q, 1, a1, b1, lmax, vg, v1, v0, vt, coef);
    for (int i = 2; i <= nsteps; ++i) {
        const double xn = 1.0 / b1;
        #pragma omp parallel for schedule(static)
        for (int64_t j = 0; j < d; ++j) { vg[j] = v1[j]; v1[j] = xn * v0[j]; v0[j] = 0.0; }
        hv_op(&op, 0.0, v1, v0, &pr);
        a1 = creal(pr);
        b1 = sqrt(creal(orc_dot(d, v0, v0)));
        alpha[i - 1] = lmax - a1;
        beta[i - 1] = b1;
        eval_phys(n, q, ibond, ipair, btype, site_vec, sq_num, sq_q, i, a1, b1, lmax, vg, v1, v0, vt,
                  coef + (size_t)(i - 1) * nc);
    }
    return 0;
}

/* ------------------------------------------------------------------------------ */
/* DSF: parallelized_dsflnc1z H/DSF.f90:195-261 (v1 = normalised excitation,        */
/* sbeta(1) = cnorm), exec_renbun H/DSF.f90:263-334, nomalization_DSF 337-351.       */
/* ------------------------------------------------------------------------------ */
int orc_dsf_lanczos(int n, int q, int ibond, const int *site_pair, const int *btype,
                    const zc *xl, const unsigned char *mk, double cnorm, int s_maxitr,
                    zc *v1, zc *v0, double *salpha, double *sbeta)
{
    orc_op op = {n, q, ibond, site_pair, btype, xl, mk};
    const int64_t d = ipow64(q, n);
    memset(v0, 0, sizeof(zc) * (size_t)d);
    sbeta[0] = cnorm;
    zc pr;
    hv_op(&op, 0.0, v1, v0, &pr);
    double a1 = creal(pr), b1 = sqrt(orc_beta2(d, a1, v1, v0));
    salpha[0] = a1;
    sbeta[1] = b1;
    for (int i = 2; i <= s_maxitr - 1; ++i) {
        #pragma omp parallel for schedule(static)
        for (int64_t j = 0; j < d; ++j) {
            const zc t1 = v1[j];
            const zc t0 = (v0[j] - a1 * v1[j]) / b1;
            v0[j] = -b1 * t1;
            v1[j] = t0;
        }
        hv_op(&op, 0.0, v1, v0, &pr);
        a1 = creal(pr);
        b1 = sqrt(orc_beta2(d, a1, v1, v0));
        salpha[i - 1] = a1;
        sbeta[i] = b1;
    }
    return s_maxitr - 1;
}

void orc_renbun(int i_range, int s_maxitr, double split, double ene, double epsln,
                const double *salpha, const double *sbeta, double *w, int *conv, double *dsf)
{
    #pragma omp parallel for schedule(static, 1)
    for (int j = 1; j <= i_range; ++j) {
        const double omega = (double)(j - 1) * split;
        w[j - 1] = omega;
        int i_conv = 0;
        double xr = 1.0e-10, xi = 1.0e-10, cr = xr, ci = xi, dr = 0.0, di = 0.0;
        for (int m1 = 1; m1 <= s_maxitr - 1; ++m1) {
            const double f1 = omega + ene - salpha[m1 - 1];
            const double f2 = -1.0 * pow(sbeta[m1 - 1], 2.0);
            const double xr1 = xr, xi1 = xi;
            dr = dr * f2; dr = dr + f1;
            di = di * f2; di = di + epsln;
            double dd = dr * dr + di * di;
            if (dd == 0.0) { dr = 0.5e-28; di = 0.5e-28; }
            double c = cr * cr + ci * ci;
            cr = cr * f2 / c; cr = cr + f1;
            ci = -ci * f2 / c; ci = ci + epsln;
            c = cr * cr + ci * ci;
            if (c == 0.0) { cr = 0.5e-28; ci = 0.5e-28; }
            dd = dr * dr + di * di;
            dr = dr / dd; di = -di / dd;
            const double dlr = cr * dr - ci * di, dli = ci * dr + cr * di;
            xr = xr1 * dlr - xi1 * dli;
            xi = xr1 * dli + xi1 * dlr;
            const double dl = (dlr - 1.0) * (dlr - 1.0) + dli * dli;
            if (dl < 0.5e-28) { i_conv = m1; break; }
        }
        conv[j - 1] = i_conv;
        dsf[j - 1] = xi;
    }
}

void orc_normalize_dsf(int i_range, double split, double sq, double *dsf)
{
    double sum = 0.0;
    for (int i = 2; i <= i_range; ++i) sum += 0.5 * split * (dsf[i - 1] + dsf[i - 2]);
    for (int i = 0; i < i_range; ++i) dsf[i] = sq / sum * dsf[i];
}

/* check1 H/check.f90:8-54: <x|H|x> and (Hx)_j/x_j at j = min(ldim/3,13) : ldim : max(1,ldim/20) */
int orc_check1(int n, int q, int ibond, const int *site_pair, const int *btype,
               const zc *xl, const unsigned char *mk, const zc *x, zc *work,
               zc *expect, zc *ratios, int max_ratios)
{
    orc_op op = {n, q, ibond, site_pair, btype, xl, mk};
    const int64_t d = ipow64(q, n);
    memset(work, 0, sizeof(zc) * (size_t)d);
    zc pr;
    hv_op(&op, 0.0, x, work, &pr);
    *expect = orc_dot(d, work, x);
    int64_t j0 = d / 3 < 13 ? d / 3 : 13;
    int64_t st = d / 20 > 1 ? d / 20 : 1;
    int k = 0;
    for (int64_t j = j0; j <= d && k < max_ratios; j += st) ratios[k++] = work[j - 1] / x[j - 1];
    return k;
}

/* CPU baseline helper: fill rows of v1 with the Gaussian generator (no normalisation) */
void orc_fill_range(uint64_t seed, int64_t lo, int64_t hi, zc *v)
{
    #pragma omp parallel for schedule(static)
    for (int64_t i = lo; i < hi; ++i) v[i] = orc_gauss_elem(seed, (uint64_t)i, 0);
}

int orc_num_threads(void) { return omp_get_max_threads(); }
