// Host-only check of the batch planner: build the batch term structure from template
// operators, then fit every request operator into it (no GPU calls).
// Input file: int32 {n, ibond, ntemplates, nrequests}, ipair[2*ibond], ibond_type[ibond],
// then per operator xlocal_Ham (9*16 complex128) and llocal_Ham (9*16 int32).
#include <cstdio>
#include <vector>

#include "../../pned_amd/csrc/internal.h"

extern "C" const char *pned_last_error(void);

int main(int argc, char **argv)
{
    if (argc != 2) return 2;
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 2;
    int h[4];
    if (fread(h, 4, 4, f) != 4) return 2;
    const int n = h[0], ib = h[1], nt = h[2], nr = h[3];
    std::vector<int32_t> ip(2 * ib), bt(ib);
    if (fread(ip.data(), 4, 2 * ib, f) != (size_t)(2 * ib) || fread(bt.data(), 4, ib, f) != (size_t)ib) return 2;
    std::vector<std::vector<double>> xl(nt + nr, std::vector<double>(288));
    std::vector<std::vector<int32_t>> ll(nt + nr, std::vector<int32_t>(144));
    for (int o = 0; o < nt + nr; ++o)
        if (fread(xl[o].data(), 8, 288, f) != 288 || fread(ll[o].data(), 4, 144, f) != 144) return 2;
    fclose(f);
    std::vector<const double *> xp;
    std::vector<const int32_t *> lp;
    for (int o = 0; o < nt; ++o) {
        xp.push_back(xl[o].data());
        lp.push_back(ll[o].data());
    }
    pned::S12Plan plan;
    if (pned::s12_decompose_batch(n, ib, ip.data(), bt.data(), nt, xp.data(), lp.data(), plan)) {
        printf("decompose: %s\n", pned_last_error());
        return 1;
    }
    int bad = 0;
    for (int r = 0; r < nr; ++r)
        if (pned::s12_set_column(plan, ib, ip.data(), bt.data(), r % nt, xl[nt + r].data(), ll[nt + r].data())) {
            printf("request %d: %s\n", r, pned_last_error());
            ++bad;
        }
    printf("terms %zu groups %zu rejected %d of %d\n", plan.terms.size(), plan.diag.size(), bad, nr);
    return bad != 0;
}
