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
#include <atomic>
#include <climits>
#include <cmath>
#include <thread>
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

// Householder QR of a wide m x nc block stored ROW-major (row r at B + r*ld),
// keeping R (zeros below).  Same arithmetic, in the same order, as a
// column-major unblocked QR: w_c = B[j][c] + sum_{r>j} v_r B[r][c] accumulated
// over r ascending, then B[r][c] -= (tau w_c) v_r.  Row-major makes both inner
// loops run over contiguous c, so they vectorise without reassociation.
void qr_keep_r_rows(long m, long nc, double *B, long ld, double *w)
{
    const long kmax = std::min(m, nc);
    for (long j = 0; j < kmax; ++j) {
        double s = 0.0;
        for (long r = j + 1; r < m; ++r) s += B[r * ld + j] * B[r * ld + j];
        const double alpha = B[j * ld + j];
        const double nrm = std::sqrt(s + alpha * alpha);
        if (nrm == 0.0) continue;
        const double beta = -sgn1(alpha) * nrm;
        const double tau = (beta - alpha) / beta;
        const double scal = 1.0 / (alpha - beta);
        for (long r = j + 1; r < m; ++r) B[r * ld + j] *= scal;
        const long c0 = j + 1, nw = nc - c0;
        if (nw > 0) {
            const double *bj = B + j * ld + c0;
            for (long c = 0; c < nw; ++c) w[c] = bj[c];
            for (long r = j + 1; r < m; ++r) {
                const double v = B[r * ld + j];
                const double *br = B + r * ld + c0;
                for (long c = 0; c < nw; ++c) w[c] += v * br[c];
            }
            for (long c = 0; c < nw; ++c) w[c] *= tau;
            double *bjw = B + j * ld + c0;
            for (long c = 0; c < nw; ++c) bjw[c] -= w[c];
            for (long r = j + 1; r < m; ++r) {
                const double v = B[r * ld + j];
                double *br = B + r * ld + c0;
                for (long c = 0; c < nw; ++c) br[c] -= w[c] * v;
            }
        }
        B[j * ld + j] = beta;
        for (long r = j + 1; r < m; ++r) B[r * ld + j] = 0.0;
    }
}

struct Scratch {
    std::vector<double> buf, w;
};

// QR_col!(A, s, l, gap) (1-based, bulgechasing.jl:9-13): rows s..l, columns
// s..l+gap replaced by R.
void qr_col(Band &A, Scratch &S, long s, long l, long gap)
{
    const long m = l - s + 1, nc = l + gap - s + 1;
    if (m <= 0 || nc <= 0) return;
    S.buf.assign((size_t)(m * nc), 0.0);
    S.w.resize((size_t)nc);
    const long r0 = s - 1, c0 = s - 1;
    for (long r = 0; r < m; ++r)
        for (long c = 0; c < nc; ++c) S.buf[r * nc + c] = A.get(r0 + r, c0 + c);
    qr_keep_r_rows(m, nc, S.buf.data(), nc, S.w.data()); // m <= nc since gap >= 0
    for (long r = 0; r < m; ++r)
        for (long c = 0; c < nc; ++c) A.set(r0 + r, c0 + c, S.buf[r * nc + c]);
}

// QR_row!(A, s, l, gap) (bulgechasing.jl:1-7): rows s..l, columns s+gap..l
// replaced by L of their LQ factorisation (QR of the transposed block).
void qr_row(Band &A, Scratch &S, long s, long l, long gap)
{
    const long m = l - s + 1, nc = l - s - gap + 1;
    if (m <= 0 || nc <= 0) return;
    // the transposed block (nc x m) stored row-major is the block column-major:
    // row c of the transpose = column c of the block
    S.buf.assign((size_t)(m * nc), 0.0);
    S.w.resize((size_t)m);
    const long r0 = s - 1, c0 = s - 1 + gap;
    for (long c = 0; c < nc; ++c)
        for (long r = 0; r < m; ++r) S.buf[c * m + r] = A.get(r0 + r, c0 + c);
    qr_keep_r_rows(nc, m, S.buf.data(), m, S.w.data()); // nc <= m since gap >= 0
    for (long c = 0; c < nc; ++c)
        for (long r = 0; r < m; ++r) A.set(r0 + r, c0 + c, S.buf[c * m + r]);
}

// One block operation of block_bidiagonalize!; its block spans rows [s, l].
struct BOp {
    int kind; // 0 = QR_row, 1 = QR_col
    long s, l, gap;
};

// ops of sweep j of block_bidiagonalize!(A, n, bw, tbw), bulgechasing.jl:18-29
void sweep_ops(long n, long bw, long tbw, long j, std::vector<BOp> &ops)
{
    ops.clear();
    ops.push_back({0, j, std::min(j + bw + tbw - 1, n), tbw});
    for (long i = j; i <= n; i += bw) {
        const long last = std::min(i + bw + tbw - 1, n);
        const long sc = std::min(bw + tbw - 1, std::max(n - i - bw - tbw + 1, 0L));
        ops.push_back({1, i + tbw, last, sc});
        if (i + tbw > n - bw) break;
        ops.push_back({0, i + tbw, last + sc, bw});
    }
}

void run_op(Band &A, Scratch &S, const BOp &o)
{
    if (o.kind == 0) qr_row(A, S, o.s, o.l, o.gap);
    else qr_col(A, S, o.s, o.l, o.gap);
}

int host_threads()
{
    if (const char *e = getenv("AVSVD_HOST_THREADS")) {
        const int v = atoi(e);
        if (v > 0) return v;
    }
    const unsigned hc = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(16u, hc));
}

// block_bidiagonalize!(A, n, bandwidth, target_bandwidth), bulgechasing.jl:17-31.
// Sweep j+tbw trails sweep j: before an operation whose block spans rows
// [s, l], a sweep waits until the previous sweep's next operation starts below
// row l.  Rows of a sweep's operations only move down, so every pair of
// operations with overlapping blocks runs in the sequential order and the
// result is bit-identical to the sequential loop.
void block_bidiagonalize(Band &A, long n, long bw, long tbw, int nthreads)
{
    std::vector<long> starts;
    for (long j = 1; j <= n - tbw; j += tbw) starts.push_back(j);
    const long ns = (long)starts.size();
    if (ns == 0) return;
    if (nthreads <= 1 || ns < 4) {
        Scratch S;
        std::vector<BOp> ops;
        for (long j : starts) {
            sweep_ops(n, bw, tbw, j, ops);
            for (auto &o : ops) run_op(A, S, o);
        }
        return;
    }
    const long DONE = LONG_MAX;
    std::vector<std::atomic<long>> front(ns);
    for (auto &f : front) f.store(LONG_MIN, std::memory_order_relaxed); // not started
    auto worker = [&](int t) {
        Scratch S;
        std::vector<BOp> ops;
        for (long w = t; w < ns; w += nthreads) {
            sweep_ops(n, bw, tbw, starts[w], ops);
            for (size_t k = 0; k < ops.size(); ++k) {
                front[w].store(ops[k].s, std::memory_order_release);
                if (w > 0) {
                    const long need_below = ops[k].l; // previous sweep must be past row l
                    while (true) {
                        const long f = front[w - 1].load(std::memory_order_acquire);
                        if (f == DONE || f > need_below) break;
                        std::this_thread::yield();
                    }
                }
                run_op(A, S, ops[k]);
            }
            front[w].store(DONE, std::memory_order_release);
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nthreads; ++t) th.emplace_back(worker, t);
    worker(0);
    for (auto &x : th) x.join();
}

// ------------------------------------------------------------------ LAPACK
typedef void (*dbdsdc_t)(const char *, const char *, const int *, double *, double *, double *, const int *,
                         double *, const int *, double *, int *, double *, int *, int *, size_t, size_t);

// dgbbrd(vect, m, n, ncc, kl, ku, ab, ldab, d, e, q, ldq, pt, ldpt, c, ldc, work, info)
typedef void (*dgbbrd_t)(const char *, const int *, const int *, const int *, const int *, const int *, double *,
                         const int *, double *, double *, double *, const int *, double *, const int *, double *,
                         const int *, double *, int *, size_t);

std::once_flag g_lapack_once;
dbdsdc_t g_dbdsdc = nullptr;
dgbbrd_t g_dgbbrd = nullptr;
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
            void *g = dlsym(h, "scipy_dgbbrd_");
            if (!g) g = dlsym(h, "dgbbrd_");
            g_dgbbrd = (dgbbrd_t)g;
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
    return avs::bidiagonalize_band_threads(n, bw, band, ldband, d, e, host_threads());
}

avs_status avs_bidiagonalize_band_lapack(int64_t n, int32_t bw, const double *band, int64_t ldband, double *d,
                                         double *e)
{
    return avs::bidiagonalize_band_lapack(n, bw, band, ldband, d, e);
}

avs_status avs_bidiagonalize_band_chase(int64_t n, int32_t bw, const double *band, int64_t ldband, double *d,
                                        double *e)
{
    return avs::bidiagonalize_band_chase(n, bw, band, ldband, d, e, host_threads());
}

} // extern "C"

namespace avs {

avs_status bidiagonalize_band_threads(int64_t n, int32_t bw, const double *band, int64_t ldband, double *d, double *e,
                                      int nth)
{
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
        block_bidiagonalize(A, n, half * 2, half, nth);
        b = half;
    }
    block_bidiagonalize(A, n, b, 1, nth);
    if (A.overflow) return fail(AVS_ERR_STATE, "bulge outside the band store (internal error)");
    for (long i = 0; i < n; ++i) d[i] = A.get(i, i);
    for (long i = 0; i + 1 < n; ++i) e[i] = A.get(i, i + 1);
    return AVS_OK;
}

int host_thread_count() { return host_threads(); }

// Band -> bidiagonal on one thread with LAPACK dgbbrd (Givens rotations, the
// band chased in place); selected by AVSVD_STAGE2=dgbbrd.  Our band layout
// (band[(bw + i - j) + j*ldband] = A(i, j)) is LAPACK's band storage with
// kl = 0, ku = bw.
avs_status bidiagonalize_band_lapack(int64_t n, int32_t bw, const double *band, int64_t ldband, double *d, double *e)
{
    if (n <= 0) return fail(-1, "n must be positive");
    if (bw < 1) return fail(-2, "bandwidth must be >= 1");
    if (!band) return fail(-3, "null band");
    if (ldband < bw + 1) return fail(-4, "ldband < bw+1");
    if (!d) return fail(-5, "null d");
    if (n > 1 && !e) return fail(-6, "null e");
    if (n > INT_MAX / (bw + 1)) return fail(-1, "n too large for LAPACK");
    std::call_once(g_lapack_once, load_lapack);
    if (!g_dgbbrd) return fail(AVS_ERR_LAPACK, "no LAPACK with dgbbrd found (set AVSVD_LAPACK)");
    const int nn = (int)n, zero = 0, ku = bw, ldab = bw + 1, one = 1;
    std::vector<double> ab((size_t)ldab * n), ee(std::max<int64_t>(n, 1)), work(2 * (size_t)n);
    for (int64_t j = 0; j < n; ++j) memcpy(ab.data() + j * ldab, band + j * ldband, sizeof(double) * ldab);
    // the entries of the band store above row 0 are outside the matrix
    for (int64_t j = 0; j < std::min<int64_t>(n, bw); ++j)
        for (int64_t r = 0; r < bw - j; ++r) ab[r + j * ldab] = 0.0;
    double dummy[1] = {0.0};
    int info = 0;
    g_dgbbrd("N", &nn, &nn, &zero, &zero, &ku, ab.data(), &ldab, d, ee.data(), dummy, &one, dummy, &one, dummy, &one,
             work.data(), &info, 1);
    if (info != 0) return fail(AVS_ERR_LAPACK, "dgbbrd info=%d", info);
    if (n > 1) memcpy(e, ee.data(), sizeof(double) * (n - 1));
    return AVS_OK;
}

// The stage-2 reduction for a caller that gives each matrix `nthreads` host
// threads (the batched form: 1; the serving loop: 1, or 4 for n >= 4096):
// Householder bulge chasing (bandbrd.cpp), 1.5-2.8x faster than dgbbrd on
// one thread at n = 256..4096, bw = 32..64 and 4x faster than the
// bandwidth-halving sweeps on 8 threads at n = 8192.  AVSVD_STAGE2=dgbbrd or
// =halving selects LAPACK dgbbrd or the reference's halving sweeps instead.
avs_status bidiagonalize_band_one(int64_t n, int32_t bw, const double *band, int64_t ldband, double *d, double *e,
                                  int nthreads)
{
    const char *pol = getenv("AVSVD_STAGE2");
    if (pol && strcmp(pol, "halving") == 0) return bidiagonalize_band_threads(n, bw, band, ldband, d, e, nthreads);
    if (pol && strcmp(pol, "dgbbrd") == 0) {
        std::call_once(g_lapack_once, load_lapack);
        if (g_dgbbrd) return bidiagonalize_band_lapack(n, bw, band, ldband, d, e);
    }
    return bidiagonalize_band_chase(n, bw, band, ldband, d, e, nthreads);
}

// The stage-2 reduction for a caller that has the host threads to itself (the
// plain per-matrix svdvals): the bulge chasing on one thread below
// n*bw = 65536 (threads do not pay off there), pipelined over every host
// thread above.
avs_status bidiagonalize_band_auto(int64_t n, int32_t bw, const double *band, int64_t ldband, double *d, double *e)
{
    return bidiagonalize_band_one(n, bw, band, ldband, d, e, n * (int64_t)bw <= 65536 ? 1 : host_threads());
}

} // namespace avs

extern "C" {

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
