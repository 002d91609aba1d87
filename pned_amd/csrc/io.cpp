// io.cpp -- the reference's result and restart files (Fortran unformatted sequential).
//
// Every file the reference writes with form='unformatted', access='sequential' is a list
// of records; a record is its bytes between two 4-byte length markers.  Records longer
// than a subrecord limit (2^31 - 9 bytes, the largest a 4-byte marker describes) are
// split into subrecords as gfortran documents: the leading marker of a subrecord that is
// followed by another one is negated, the trailing marker of a subrecord that follows
// another one is negated.  The reader also accepts one record of up to 4 GiB with
// unsigned markers (what amdflang writes).
//
// Files (S_one_half/Hybrid_TPS_*.f90):
//   SalphaSbeta_me*seed*<comp>  TPS.f90:507-538   num_data1, num_data2, alpha, beta, coef
//   SalphaSbeta_q*me*<comp>     DSF.f90:353-372   sq, salpha, sbeta
//   vec_q*me*<comp>, vec_me*    DSF.f90:374-379   indx, v1, v2 [, v3]
//   basisme*                    output.f90:12-36  e, x
//   vecdump_me* (interval_dump) output.f90:147-201 data_size, vec_size, sa, sb, sc, se,
//                                                  vec0..vec4 (the ones a caller dumps)
//   vecdump_me*.info            output.f90:160-176 DUMP_NAME, np, position, max_itr,
//                                                  dsf_comp, s_maxitr, n, ibond, nvec,
//                                                  total_sz, itr_counter(1..12)
//   dsf_res_<comp>.dat*         output.f90:243-262 text: sq, then q, w, dsf, conv rows
// Device vectors are streamed through a pinned host buffer, never held whole on the host.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "internal.h"

namespace pned {
namespace {

size_t g_sub_limit = 2147483639;   // 2^31 - 9

struct File {
    FILE *f = nullptr;
    ~File()
    {
        if (f) fclose(f);
    }
};

int put_marker(FILE *f, int32_t m)
{
    return fwrite(&m, 4, 1, f) == 1 ? PNED_OK : fail(PNED_E_ARG, "write failed");
}

// writes one record of `total` bytes whose data come from fill(buf, offset, n)
template <typename Fill>
int write_record(FILE *f, size_t total, Fill fill)
{
    const size_t lim = g_sub_limit;
    size_t done = 0;
    std::vector<unsigned char> buf;
    const size_t nsub = total == 0 ? 1 : (total + lim - 1) / lim;
    for (size_t k = 0; k < nsub; ++k) {
        const size_t len = std::min(lim, total - done);
        const int32_t m = (int32_t)len;
        PNED_TRY(put_marker(f, k + 1 < nsub ? -m : m));
        PNED_TRY(fill(f, done, len));
        PNED_TRY(put_marker(f, k > 0 ? -m : m));
        done += len;
    }
    return PNED_OK;
}

int write_host(FILE *f, const void *p, size_t bytes)
{
    const unsigned char *c = (const unsigned char *)p;
    return write_record(f, bytes, [&](FILE *ff, size_t off, size_t n) -> int {
        return fwrite(c + off, 1, n, ff) == n ? PNED_OK : fail(PNED_E_ARG, "write failed");
    });
}

// reads one record of exactly `bytes`, handing the data to take(offset, ptr, n)
template <typename Take>
int read_record(FILE *f, size_t bytes, Take take)
{
    size_t got = 0;
    std::vector<unsigned char> buf(1 << 20);
    bool first = true;
    for (;;) {
        int32_t lead;
        if (fread(&lead, 4, 1, f) != 1) return fail(PNED_E_ARG, "unexpected end of file");
        size_t len;
        bool more;
        if (lead >= 0) {
            len = (size_t)lead;
            more = false;
        } else if (first && (uint32_t)lead == bytes) {   // one record with unsigned markers
            len = (size_t)(uint32_t)lead;
            more = false;
        } else {
            len = (size_t)(-(int64_t)lead);
            more = true;
        }
        if (got + len > bytes) return fail(PNED_E_ARG, "record longer than expected");
        size_t left = len;
        while (left > 0) {
            const size_t n = std::min(left, buf.size());
            if (fread(buf.data(), 1, n, f) != n) return fail(PNED_E_ARG, "unexpected end of file");
            PNED_TRY(take(got, buf.data(), n));
            got += n;
            left -= n;
        }
        int32_t trail;
        if (fread(&trail, 4, 1, f) != 1) return fail(PNED_E_ARG, "unexpected end of file");
        first = false;
        if (!more) break;
    }
    if (got != bytes) return fail(PNED_E_ARG, "record shorter than expected");
    return PNED_OK;
}

int read_host(FILE *f, void *p, size_t bytes)
{
    unsigned char *c = (unsigned char *)p;
    return read_record(f, bytes, [&](size_t off, const unsigned char *src, size_t n) -> int {
        std::memcpy(c + off, src, n);
        return PNED_OK;
    });
}

constexpr size_t STAGE = (size_t)64 << 20;

int write_vec(FILE *f, const pned_vec *v)
{
    if (vec_detached(v)) return fail(PNED_E_ARG, "vector of a destroyed context");
    if (v->dtype != PNED_C128 || v->nvec != 1)
        return fail(PNED_E_UNSUPPORTED, "vector files hold one complex128 column");
    PNED_HIP(hipSetDevice(v->ctx->device));
    PNED_HIP(hipStreamSynchronize(v->ctx->stream));
    void *pin = nullptr;
    PNED_HIP(hipHostMalloc(&pin, STAGE, hipHostMallocDefault));
    const unsigned char *dev = (const unsigned char *)v->data;
    const bool mapped = !v->ctx->lay.identity();   // file order = basis order
    const int rc = write_record(f, v->bytes, [&](FILE *ff, size_t off, size_t n) -> int {
        size_t done = 0;
        while (done < n) {
            if (!mapped) {
                const size_t m = std::min(STAGE, n - done);
                if (hipMemcpy(pin, dev + off + done, m, hipMemcpyDeviceToHost) != hipSuccess)
                    return fail(PNED_E_HIP, "device to host copy failed");
                if (fwrite(pin, 1, m, ff) != m) return fail(PNED_E_ARG, "write failed");
                done += m;
                continue;
            }
            // subrecords need not end on a state: fetch whole states around the bytes
            const size_t eb = 16, b0 = off + done, m = std::min(STAGE - 2 * eb, n - done);
            const size_t s0 = b0 / eb, s1 = (b0 + m + eb - 1) / eb;
            PNED_TRY(vec_to_host(v, pin, (int64_t)s0, (int64_t)(s1 - s0)));
            if (fwrite((const unsigned char *)pin + (b0 - s0 * eb), 1, m, ff) != m)
                return fail(PNED_E_ARG, "write failed");
            done += m;
        }
        return PNED_OK;
    });
    (void)hipHostFree(pin);
    return rc;
}

int read_vec(FILE *f, pned_vec *v)
{
    if (vec_detached(v)) return fail(PNED_E_ARG, "vector of a destroyed context");
    if (v->dtype != PNED_C128 || v->nvec != 1)
        return fail(PNED_E_UNSUPPORTED, "vector files hold one complex128 column");
    PNED_HIP(hipSetDevice(v->ctx->device));
    PNED_HIP(hipStreamSynchronize(v->ctx->stream));
    void *pin = nullptr;
    PNED_HIP(hipHostMalloc(&pin, STAGE, hipHostMallocDefault));
    unsigned char *dev = (unsigned char *)v->data;
    size_t fill = 0, base = 0;   // staged bytes, byte offset of the stage in the vector
    // the stage is flushed at multiples of STAGE and at the end, so it always holds whole
    // states; a layout-mapped vector takes them through the state map
    auto flush = [&]() -> int {
        if (v->ctx->lay.identity()) {
            if (hipMemcpy(dev + base, pin, fill, hipMemcpyHostToDevice) != hipSuccess)
                return fail(PNED_E_HIP, "host to device copy failed");
            return PNED_OK;
        }
        return vec_from_host(v, pin, (int64_t)(base / 16), (int64_t)(fill / 16));
    };
    int rc = read_record(f, v->bytes, [&](size_t off, const unsigned char *src, size_t n) -> int {
        size_t done = 0;
        while (done < n) {
            if (fill == 0) base = off + done;
            const size_t m = std::min(STAGE - fill, n - done);
            std::memcpy((unsigned char *)pin + fill, src + done, m);
            fill += m;
            done += m;
            if (fill == STAGE) {
                PNED_TRY(flush());
                fill = 0;
            }
        }
        return PNED_OK;
    });
    if (rc == PNED_OK && fill > 0) rc = flush();
    (void)hipHostFree(pin);
    return rc;
}

int open_file(File &ff, const char *path, const char *mode)
{
    if (!path) return fail(PNED_E_ARG, "null path");
    ff.f = fopen(path, mode);
    if (!ff.f) return fail(PNED_E_ARG, std::string("cannot open ") + path);
    return PNED_OK;
}

}  // namespace
}  // namespace pned

using namespace pned;

extern "C" {

int pned_io_subrecord_limit(uint64_t bytes)
{
    if (bytes != 0 && (bytes < 8 || bytes > 2147483639ull))
        return fail(PNED_E_ARG, "subrecord limit in 8 .. 2^31-9 (0 = default)");
    g_sub_limit = bytes ? (size_t)bytes : 2147483639;
    return PNED_OK;
}

int pned_dump_alpha_beta_coef(const char *path, int num_data1, int num_data2, const double *alpha,
                              const double *beta, const double *coef)
{
    if (num_data1 < 1 || num_data2 < 1 || !alpha || !beta || !coef) return fail(PNED_E_ARG, "bad argument");
    File ff;
    PNED_TRY(open_file(ff, path, "wb"));
    const int32_t n1 = num_data1, n2 = num_data2;
    PNED_TRY(write_host(ff.f, &n1, 4));
    PNED_TRY(write_host(ff.f, &n2, 4));
    PNED_TRY(write_host(ff.f, alpha, sizeof(double) * (size_t)num_data1));
    PNED_TRY(write_host(ff.f, beta, sizeof(double) * (size_t)num_data1));
    PNED_TRY(write_host(ff.f, coef, sizeof(double) * (size_t)num_data1 * num_data2));
    return PNED_OK;
}

int pned_read_alpha_beta_coef(const char *path, int *num_data1, int *num_data2, double *alpha, double *beta,
                              double *coef, int max_data1, int max_data2)
{
    if (!num_data1 || !num_data2) return fail(PNED_E_ARG, "null argument");
    File ff;
    PNED_TRY(open_file(ff, path, "rb"));
    int32_t n1, n2;
    PNED_TRY(read_host(ff.f, &n1, 4));
    PNED_TRY(read_host(ff.f, &n2, 4));
    *num_data1 = n1;
    *num_data2 = n2;
    if (!alpha && !beta && !coef) return PNED_OK;   // sizes only
    if (!alpha || !beta || !coef || n1 > max_data1 || n2 > max_data2 || n1 < 1 || n2 < 1)
        return fail(PNED_E_ARG, "arrays missing or smaller than the file's num_data1 x num_data2");
    PNED_TRY(read_host(ff.f, alpha, sizeof(double) * (size_t)n1));
    PNED_TRY(read_host(ff.f, beta, sizeof(double) * (size_t)n1));
    PNED_TRY(read_host(ff.f, coef, sizeof(double) * (size_t)n1 * n2));
    return PNED_OK;
}

int pned_dump_dsf_alpha_beta(const char *path, const double *sq, int n_alpha, const double *salpha,
                             int n_beta, const double *sbeta)
{
    if (!sq || n_alpha < 1 || n_beta < 1 || !salpha || !sbeta) return fail(PNED_E_ARG, "bad argument");
    File ff;
    PNED_TRY(open_file(ff, path, "wb"));
    PNED_TRY(write_host(ff.f, sq, 2 * sizeof(double)));   // complex(kind(0d0)) sq
    PNED_TRY(write_host(ff.f, salpha, sizeof(double) * (size_t)n_alpha));
    PNED_TRY(write_host(ff.f, sbeta, sizeof(double) * (size_t)n_beta));
    return PNED_OK;
}

int pned_read_dsf_alpha_beta(const char *path, double *sq, int n_alpha, double *salpha, int n_beta, double *sbeta)
{
    if (!sq || n_alpha < 1 || n_beta < 1 || !salpha || !sbeta) return fail(PNED_E_ARG, "bad argument");
    File ff;
    PNED_TRY(open_file(ff, path, "rb"));
    PNED_TRY(read_host(ff.f, sq, 2 * sizeof(double)));
    PNED_TRY(read_host(ff.f, salpha, sizeof(double) * (size_t)n_alpha));
    PNED_TRY(read_host(ff.f, sbeta, sizeof(double) * (size_t)n_beta));
    return PNED_OK;
}

int pned_dump_vectors(const char *path, int indx, int nvec, const pned_vec *const *vecs)
{
    if (nvec < 1 || !vecs) return fail(PNED_E_ARG, "bad argument");
    for (int k = 0; k < nvec; ++k)
        if (!vecs[k]) return fail(PNED_E_ARG, "null vector");
    File ff;
    PNED_TRY(open_file(ff, path, "wb"));
    const int32_t ix = indx;
    PNED_TRY(write_host(ff.f, &ix, 4));
    for (int k = 0; k < nvec; ++k) PNED_TRY(write_vec(ff.f, vecs[k]));
    return PNED_OK;
}

int pned_read_vectors(const char *path, int *indx, int nvec, pned_vec *const *vecs)
{
    if (!indx || nvec < 1 || !vecs) return fail(PNED_E_ARG, "bad argument");
    for (int k = 0; k < nvec; ++k)
        if (!vecs[k]) return fail(PNED_E_ARG, "null vector");
    File ff;
    PNED_TRY(open_file(ff, path, "rb"));
    int32_t ix;
    PNED_TRY(read_host(ff.f, &ix, 4));
    *indx = ix;
    for (int k = 0; k < nvec; ++k) PNED_TRY(read_vec(ff.f, vecs[k]));
    return PNED_OK;
}

int pned_basis_write(const char *path, double e, const pned_vec *x)
{
    if (!x) return fail(PNED_E_ARG, "null vector");
    File ff;
    PNED_TRY(open_file(ff, path, "wb"));
    PNED_TRY(write_host(ff.f, &e, sizeof(double)));
    return write_vec(ff.f, x);
}

int pned_basis_read(const char *path, double *e, pned_vec *x)
{
    if (!e || !x) return fail(PNED_E_ARG, "null argument");
    File ff;
    PNED_TRY(open_file(ff, path, "rb"));
    PNED_TRY(read_host(ff.f, e, sizeof(double)));
    return read_vec(ff.f, x);
}

int pned_interval_dump(const char *path, int32_t data_size, int32_t vec_size, int nreal, const double *const *reals,
                       const int64_t *real_counts, int nvec, const pned_vec *const *vecs)
{
    if (nreal < 0 || nvec < 0 || (nreal > 0 && (!reals || !real_counts)) || (nvec > 0 && !vecs))
        return fail(PNED_E_ARG, "bad argument");
    for (int k = 0; k < nreal; ++k)
        if (!reals[k] || real_counts[k] < 0) return fail(PNED_E_ARG, "bad real array");
    for (int k = 0; k < nvec; ++k)
        if (!vecs[k]) return fail(PNED_E_ARG, "null vector");
    File ff;
    PNED_TRY(open_file(ff, path, "wb"));
    PNED_TRY(write_host(ff.f, &data_size, 4));
    PNED_TRY(write_host(ff.f, &vec_size, 4));
    for (int k = 0; k < nreal; ++k) PNED_TRY(write_host(ff.f, reals[k], sizeof(double) * (size_t)real_counts[k]));
    for (int k = 0; k < nvec; ++k) PNED_TRY(write_vec(ff.f, vecs[k]));
    return PNED_OK;
}

int pned_interval_dump_read(const char *path, int32_t *data_size, int32_t *vec_size, int nreal, double *const *reals,
                            const int64_t *real_counts, int nvec, pned_vec *const *vecs)
{
    if (!data_size || !vec_size || nreal < 0 || nvec < 0 || (nreal > 0 && (!reals || !real_counts)) ||
        (nvec > 0 && !vecs))
        return fail(PNED_E_ARG, "bad argument");
    for (int k = 0; k < nreal; ++k)
        if (!reals[k] || real_counts[k] < 0) return fail(PNED_E_ARG, "bad real array");
    for (int k = 0; k < nvec; ++k)
        if (!vecs[k]) return fail(PNED_E_ARG, "null vector");
    File ff;
    PNED_TRY(open_file(ff, path, "rb"));
    PNED_TRY(read_host(ff.f, data_size, 4));
    PNED_TRY(read_host(ff.f, vec_size, 4));
    for (int k = 0; k < nreal; ++k) PNED_TRY(read_host(ff.f, reals[k], sizeof(double) * (size_t)real_counts[k]));
    for (int k = 0; k < nvec; ++k) PNED_TRY(read_vec(ff.f, vecs[k]));
    return PNED_OK;
}

int pned_dump_info_write(const char *path, const char *dump_name, int32_t np, int npos, const int32_t *position,
                         int32_t true_value, int32_t max_itr, const char *dsf_comp, int32_t s_maxitr, int32_t n,
                         int32_t ibond, int32_t nvec, int32_t total_sz, const int32_t *itr_counter)
{
    if (!dump_name || npos < 0 || (npos > 0 && !position) || !dsf_comp || !itr_counter)
        return fail(PNED_E_ARG, "bad argument");
    char name[128], comp[2];
    std::memset(name, ' ', sizeof(name));                 // character(len=128), blank padded
    std::memcpy(name, dump_name, std::min(std::strlen(dump_name), sizeof(name)));
    std::memset(comp, ' ', sizeof(comp));
    std::memcpy(comp, dsf_comp, std::min(std::strlen(dsf_comp), sizeof(comp)));
    std::vector<int32_t> pos((size_t)npos);
    for (int k = 0; k < npos; ++k) pos[k] = position[k] ? true_value : 0;   // logical (default kind)
    File ff;
    PNED_TRY(open_file(ff, path, "wb"));
    PNED_TRY(write_host(ff.f, name, sizeof(name)));
    PNED_TRY(write_host(ff.f, &np, 4));
    PNED_TRY(write_host(ff.f, pos.data(), 4 * (size_t)npos));
    PNED_TRY(write_host(ff.f, &max_itr, 4));
    PNED_TRY(write_host(ff.f, comp, sizeof(comp)));
    for (int32_t v : {s_maxitr, n, ibond, nvec, total_sz}) PNED_TRY(write_host(ff.f, &v, 4));
    for (int i = 0; i < 12; ++i) PNED_TRY(write_host(ff.f, &itr_counter[i], 4));
    return PNED_OK;
}

int pned_dump_info_read(const char *path, char *dump_name, int32_t *np, int npos, int32_t *position,
                        int32_t *max_itr, char *dsf_comp, int32_t *s_maxitr, int32_t *n, int32_t *ibond,
                        int32_t *nvec, int32_t *total_sz, int32_t *itr_counter)
{
    if (!dump_name || !np || npos < 0 || (npos > 0 && !position) || !max_itr || !dsf_comp || !s_maxitr || !n ||
        !ibond || !nvec || !total_sz || !itr_counter)
        return fail(PNED_E_ARG, "bad argument");
    char name[128], comp[2];
    std::vector<int32_t> pos((size_t)npos);
    File ff;
    PNED_TRY(open_file(ff, path, "rb"));
    PNED_TRY(read_host(ff.f, name, sizeof(name)));
    PNED_TRY(read_host(ff.f, np, 4));
    PNED_TRY(read_host(ff.f, pos.data(), 4 * (size_t)npos));
    PNED_TRY(read_host(ff.f, max_itr, 4));
    PNED_TRY(read_host(ff.f, comp, sizeof(comp)));
    for (int32_t *v : {s_maxitr, n, ibond, nvec, total_sz}) PNED_TRY(read_host(ff.f, v, 4));
    for (int i = 0; i < 12; ++i) PNED_TRY(read_host(ff.f, &itr_counter[i], 4));
    int len = (int)sizeof(name);
    while (len > 0 && name[len - 1] == ' ') --len;        // trim(DUMP_NAME)
    std::memcpy(dump_name, name, (size_t)len);
    dump_name[len] = '\0';
    dsf_comp[0] = comp[0];
    dsf_comp[1] = comp[1];
    dsf_comp[2] = '\0';
    for (int k = 0; k < npos; ++k) position[k] = pos[k] != 0;   // any non-zero value is .true.
    return PNED_OK;
}

// Fortran F<w>.<d> / I<w> edit: right-justified in w columns, asterisks when it does not fit
static void put_f(std::string &out, double v, int w, int d)
{
    char buf[512];
    const int len = std::snprintf(buf, sizeof(buf), "%*.*f", w, d, v);
    if (len < 0 || len > w)
        out.append((size_t)w, '*');
    else
        out.append(buf, (size_t)len);
}

static void put_i(std::string &out, long v, int w)
{
    char buf[64];
    const int len = std::snprintf(buf, sizeof(buf), "%*ld", w, v);
    if (len < 0 || len > w)
        out.append((size_t)w, '*');
    else
        out.append(buf, (size_t)len);
}

int pned_write_dsf_res(const char *path, int append, double sq, const double q[2], int ll, const double *w,
                       const double *dsf, const int32_t *conv)
{
    if (!q || ll < 0 || (ll > 0 && (!w || !dsf || !conv))) return fail(PNED_E_ARG, "bad argument");
    std::string out = "#        ";
    put_f(out, sq, 25, 18);
    out += '\n';
    for (int k = 0; k < ll; ++k) {
        put_f(out, q[0], 18, 10);
        put_f(out, q[1], 18, 10);
        put_f(out, w[k], 25, 16);
        put_f(out, dsf[k], 25, 16);
        put_i(out, conv[k], 5);
        out += '\n';
    }
    for (int k = 0; k < 4; ++k) out += " \n";           // write(j,*) '' four times
    File ff;
    PNED_TRY(open_file(ff, path, append ? "ab" : "wb"));
    if (fwrite(out.data(), 1, out.size(), ff.f) != out.size()) return fail(PNED_E_ARG, "write failed");
    return PNED_OK;
}

}  // extern "C"
