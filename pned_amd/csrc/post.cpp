// post.cpp -- host-side post-processing of Krylov results (FP64, no device work).
//
// cTPQ: the reference's evaluation program for mTPQ runs (S_one_half/Hybrid_TPS_eval_mpi.f90,
// program main + calc_fac / eval_q / lsort / calc_ave, lines 236-446), one process.
#include <algorithm>
#include <cmath>
#include <functional>
#include <vector>

#include "internal.h"

namespace pned {

// calc_fac (eval_mpi.f90:236-283): log-domain weights of the even (fac1) and odd (fac2)
// terms for one sample; beta_s = beta(:, seed).  As in the reference, the k-th weight
// takes the norms beta(2..k) and the maximum log weight is divided out.
static void ctpq_fac(int nd1, double bt, const double *beta_s, std::vector<double> &fac1, std::vector<double> &fac2)
{
    fac1.assign(nd1, 0.0);
    fac2.assign(nd1, 0.0);
    double y = 0.0, maxfac = -1e200;
    for (int i = 1; i <= 2 * (nd1 - 1); ++i) {
        const int m = ((i + 1) >> 1) + 1;
        y += std::log(bt / (double)(2 * i) * beta_s[m - 1]);
        fac1[m - 1] = y;
        maxfac = std::max(maxfac, y);
    }
    for (int i = 1; i <= nd1 - 1; ++i) {
        fac1[i - 1] = std::exp(fac1[i - 1] - maxfac);
        fac2[i - 1] = fac1[i - 1] * bt * beta_s[i] / (double)(2 * i - 1);
    }
    fac1[nd1 - 1] = std::exp(fac1[nd1 - 1] - maxfac);
    fac2[nd1 - 1] = 0.0;
}

// eval_q (eval_mpi.f90:285-308): weighted energy of one sample; each weighted column is
// summed largest term first (the reference sorts before summing; any descending order
// gives the same sums, ties included)
static double ctpq_sample(int nd1, double bt, const double *beta_s, const double *c4, const double *c5)
{
    std::vector<double> fac1, fac2;
    ctpq_fac(nd1, bt, beta_s, fac1, fac2);
    std::vector<double> t1(nd1 - 1), t2(nd1 - 1);
    double z = 0.0;
    for (int i = 0; i < nd1 - 1; ++i) {
        t1[i] = fac1[i] * c4[i];
        t2[i] = fac2[i] * c5[i];
        z += fac1[i] + fac2[i];
    }
    std::sort(t1.begin(), t1.end(), std::greater<double>());
    std::sort(t2.begin(), t2.end(), std::greater<double>());
    double x = 0.0, y = 0.0;
    for (double v : t1) x += v;
    for (double v : t2) y += v;
    return (x + y) / z;
}

}  // namespace pned

using namespace pned;

extern "C" int pned_ctpq_eval(int num_data1, int num_data2, int seed_num, const double *alpha, const double *beta,
                              const double *coef, double t_min, double t_max, int num_div, double *out)
{
    (void)alpha;
    if (!beta || !coef || !out) return fail(PNED_E_ARG, "null argument");
    if (num_data1 < 2 || num_data2 < 5 || seed_num < 1 || num_div < 1 || !(t_min > 0.0) || !(t_max > t_min))
        return fail(PNED_E_ARG, "cTPQ: num_data1 >= 2, num_data2 >= 5, seeds >= 1, num_div >= 1, 0 < t_min < t_max");
    const double log_t_max = std::log(t_max), log_t_min = std::log(t_min);
    const double delta = (log_t_max - log_t_min) / (double)num_div;
    double t_delta = 1e-4;
    if (t_min < t_delta) t_delta = t_min * 0.9;
    std::vector<double> tmp(seed_num);
    for (int i1 = 0; i1 <= num_div; ++i1) {
        const double t0 = std::exp(log_t_min + (double)i1 * delta);
        double dat[9];
        for (int i2 = 0; i2 <= 2; ++i2) {
            const double bt = (i2 == 0) ? 1.0 / (t0 - t_delta) : (i2 == 1) ? 1.0 / t0 : 1.0 / (t0 + t_delta);
            for (int sd = 0; sd < seed_num; ++sd) {
                const double *bs = beta + (size_t)num_data1 * sd;
                const double *c4 = coef + (size_t)num_data1 * (3 + (size_t)num_data2 * sd);
                const double *c5 = coef + (size_t)num_data1 * (4 + (size_t)num_data2 * sd);
                tmp[sd] = ctpq_sample(num_data1, bt, bs, c4, c5);
            }
            // calc_ave (eval_mpi.f90:339-352)
            double ave = 0.0, var = 0.0;
            for (double v : tmp) {
                ave += v;
                var += v * v;
            }
            const double r1 = ave / (double)seed_num;
            dat[3 * i2] = bt;
            dat[3 * i2 + 1] = r1;
            dat[3 * i2 + 2] = var / (double)seed_num - r1 * r1;
        }
        const double spec = 0.5 * (dat[7] - dat[1]) / t_delta;
        const double spec_d1 = (dat[4] - dat[1]) / t_delta, spec_d2 = (dat[7] - dat[4]) / t_delta;
        out[4 * i1 + 0] = 1.0 / dat[3];
        out[4 * i1 + 1] = spec;
        out[4 * i1 + 2] = 0.5 * (dat[2] + dat[8]);
        out[4 * i1 + 3] = (spec_d2 - spec_d1) / t_delta;
    }
    return PNED_OK;
}
