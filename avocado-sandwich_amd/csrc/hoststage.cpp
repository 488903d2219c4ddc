// hoststage.cpp -- the host stage of the svdvals path, kept on the host in FP64
// as the reference runs it:
//   bidiagonalize / block_bidiagonalize! / QR_row! / QR_col!
//       KAbasedSVD/src/bulgechasing.jl:1-42
//   diag_lapack (LAPACK.bdsdc!('U','N', ...))   KAbasedSVD/src/bulgechasing.jl:50-53
// The reference works on the dense n x n matrix; here the matrix lives in a
// compact band store wide enough for every bulge the block sweeps create (a
// pass from bandwidth bw to tbw fills at most bw+tbw-2 sub-diagonals and
// 2*bw+tbw-2 super-diagonals), so memory is O(n*b) instead of O(n^2).  Each
// block operation gathers its block into a dense buffer, factors it with
// Householder reflectors (keeping R, or L for the row variant, exactly like
// the reference keeps triu of the factors) and scatters it back.
#include <dlfcn.h>
#include <glob.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "avsvd.h"
#include "internal.hpp"

namespace {

struct Band {
    long n, kl, ku, ld;
    std::vector<double> a;
    bool overflow = false;
    Band(long n_, long kl_, long ku_) : n(n_), kl(kl_), ku(ku_), ld(kl_ + ku_ + 1), a((size_t)(kl_ + ku_ + 1) * n_, 0.0) {}
    bool in(long i, long j) const { return j - i <= ku && i - j <= kl; }
    double get(long i, long j) const { return in(i, j) ? a[(size_t)(ku + i - j) + (size_t)j * ld] : 0.0; }
    void set(long i, long j, double v)
    {
        if (in(i, j)) a[(size_t)(ku + i - j) + (size_t)j * ld] = v;
        else if (v != 0.0) overflow = true;
    }
};

inline double sgn1(double x) { return x >= 0.0 ? 1.0 : -1.0; }

// Householder QR of the m x nc column-major block B, keeping R (zeros below).
void qr_keep_r(long m, long nc, double *B, long ld)
{
    const long kmax = std::min(m, nc);
    for (long j = 0; j < kmax; ++j) {
        double *cj = B + j * ld;
        double s = 0.0;
        for (long r = j + 1; r < m; ++r) s += cj[r] * cj[r];
        const double alpha = cj[j];
        const double nrm = std::sqrt(s + alpha * alpha);
        if (nrm == 0.0) continue;
        const double beta = -sgn1(alpha) * nrm;
        const double tau = (beta - alpha) / beta;
        const double scal = 1.0 / (alpha - beta);
        for (long r = j + 1; r < m; ++r) cj[r] *= scal;
        for (long c = j + 1; c < nc; ++c) {
            double *cc = B + c * ld;
            double w = cc[j];
            for (long r = j + 1; r < m; ++r) w += cj[r] * cc[r];
            w *= tau;
            cc[j] -= w;
            for (long r = j + 1; r < m; ++r) cc[r] -= w * cj[r];
        }
        cj[j] = beta;
        for (long r = j + 1; r < m; ++r) cj[r] = 0.0;
    }
}

// QR_col!(A, s, l, gap) (1-based, bulgechasing.jl:9-13): rows s..l, columns
// s..l+gap replaced by R.
void qr_col(Band &A, std::vector<double> &buf, long s, long l, long gap)
{
    const long m = l - s + 1, nc = l + gap - s + 1;
    if (m <= 0 || nc <= 0) return;
    buf.assign((size_t)(m * nc), 0.0);
    const long r0 = s - 1, c0 = s - 1;
    for (long c = 0; c < nc; ++c)
        for (long r = 0; r < m; ++r) buf[r + c * m] = A.get(r0 + r, c0 + c);
    qr_keep_r(m, nc, buf.data(), m);
    for (long c = 0; c < nc; ++c)
        for (long r = 0; r < m; ++r) A.set(r0 + r, c0 + c, buf[r + c * m]);
}

// QR_row!(A, s, l, gap) (bulgechasing.jl:1-7): rows s..l, columns s+gap..l
// replaced by L of their LQ factorisation.
void qr_row(Band &A, std::vector<double> &buf, long s, long l, long gap)
{
    const long m = l - s + 1, nc = l - s - gap + 1;
    if (m <= 0 || nc <= 0) return;
    buf.assign((size_t)(m * nc), 0.0); // transposed block: nc x m
    const long r0 = s - 1, c0 = s - 1 + gap;
    for (long r = 0; r < m; ++r)
        for (long c = 0; c < nc; ++c) buf[c + r * nc] = A.get(r0 + r, c0 + c);
    qr_keep_r(nc, m, buf.data(), nc);
    for (long r = 0; r < m; ++r)
        for (long c = 0; c < nc; ++c) A.set(r0 + r, c0 + c, buf[c + r * nc]);
}

// block_bidiagonalize!(A, n, bandwidth, target_bandwidth), bulgechasing.jl:17-31
void block_bidiagonalize(Band &A, long n, long bw, long tbw)
{
    std::vector<double> buf;
    for (long j = 1; j <= n - tbw; j += tbw) {
        qr_row(A, buf, j, std::min(j + bw + tbw - 1, n), tbw);
        for (long i = j; i <= n; i += bw) {
            const long last = std::min(i + bw + tbw - 1, n);
            const long sc = std::min(bw + tbw - 1, std::max(n - i - bw - tbw + 1, 0L));
            qr_col(A, buf, i + tbw, last, sc);
            if (i + tbw > n - bw) break;
            qr_row(A, buf, i + tbw, last + sc, bw);
        }
    }
}

// ------------------------------------------------------------------ LAPACK
typedef void (*dbdsdc_t)(const char *, const char *, const int *, double *, double *, double *, const int *,
                         double *, const int *, double *, int *, double *, int *, int *, size_t, size_t);

std::once_flag g_lapack_once;
dbdsdc_t g_dbdsdc = nullptr;
std::string g_lapack_path;

void load_lapack()
{
    std::vector<std::string> cands;
    if (const char *e = getenv("AVSVD_LAPACK")) cands.push_back(e);
    const char *pats[] = {"/usr/local/lib/python3*/dist-packages/scipy.libs/libscipy_openblas*.so",
                          "/usr/lib/python3*/dist-packages/scipy.libs/libscipy_openblas*.so",
                          "/usr/local/lib/python3*/site-packages/scipy.libs/libscipy_openblas*.so"};
    for (const char *p : pats) {
        glob_t g;
        if (glob(p, 0, nullptr, &g) == 0)
            for (size_t i = 0; i < g.gl_pathc; ++i) cands.push_back(g.gl_pathv[i]);
        globfree(&g);
    }
    cands.push_back("liblapack.so.3");
    cands.push_back("libopenblas.so.0");
    for (auto &c : cands) {
        void *h = dlopen(c.c_str(), RTLD_NOW | RTLD_LOCAL);
        if (!h) continue;
        void *f = dlsym(h, "scipy_dbdsdc_");
        if (!f) f = dlsym(h, "dbdsdc_");
        if (f) {
            g_dbdsdc = (dbdsdc_t)f;
            g_lapack_path = c;
            return;
        }
        dlclose(h);
    }
}

} // namespace

extern "C" {

avs_status avs_bidiagonalize_band(int64_t n, int32_t bw, const double *band, int64_t ldband, double *d, double *e)
{
    using avs::fail;
    if (n <= 0) return fail(-1, "n must be positive");
    if (bw < 1) return fail(-2, "bandwidth must be >= 1");
    if (!band) return fail(-3, "null band");
    if (ldband < bw + 1) return fail(-4, "ldband < bw+1");
    if (!d) return fail(-5, "null d");
    if (n > 1 && !e) return fail(-6, "null e");
    // widest pass: bw -> round(bw/2) (or bw -> 1 when bw <= 16)
    const long tb0 = bw > 16 ? (long)std::nearbyint(bw / 2.0) : 1;
    Band A(n, bw + tb0, 2 * bw + tb0);
    for (long j = 0; j < n; ++j)
        for (long i = std::max(0L, j - bw); i <= j; ++i) A.set(i, j, band[(bw + i - j) + j * ldband]);
    long b = bw;
    while (b > 16) {
        const long half = (long)std::nearbyint(b / 2.0);
        block_bidiagonalize(A, n, half * 2, half);
        b = half;
    }
    block_bidiagonalize(A, n, b, 1);
    if (A.overflow) return fail(AVS_ERR_STATE, "bulge outside the band store (internal error)");
    for (long i = 0; i < n; ++i) d[i] = A.get(i, i);
    for (long i = 0; i + 1 < n; ++i) e[i] = A.get(i, i + 1);
    return AVS_OK;
}

avs_status avs_bdsdc_values(int64_t n, const double *d, const double *e, double *sigma)
{
    using avs::fail;
    if (n <= 0) return fail(-1, "n must be positive");
    if (!d) return fail(-2, "null d");
    if (n > 1 && !e) return fail(-3, "null e");
    if (!sigma) return fail(-4, "null sigma");
    std::call_once(g_lapack_once, load_lapack);
    if (!g_dbdsdc) return fail(AVS_ERR_LAPACK, "no LAPACK with dbdsdc found (set AVSVD_LAPACK)");
    const int nn = (int)n, one = 1;
    std::vector<double> ee(std::max<int64_t>(n, 1), 0.0), work(std::max<int64_t>(4 * n, 1));
    std::vector<int> iwork(std::max<int64_t>(8 * n, 1));
    memcpy(sigma, d, sizeof(double) * n);
    if (n > 1) memcpy(ee.data(), e, sizeof(double) * (n - 1));
    double dummy[1] = {0.0};
    int idummy[1] = {0};
    int info = 0;
    g_dbdsdc("U", "N", &nn, sigma, ee.data(), dummy, &one, dummy, &one, dummy, idummy, work.data(), iwork.data(),
             &info, 1, 1);
    if (info != 0) return fail(AVS_ERR_LAPACK, "dbdsdc info=%d", info);
    return AVS_OK;
}

} // extern "C"
