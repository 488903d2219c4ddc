// bandbrd.cpp -- upper band (bandwidth bw) -> upper bidiagonal on the host, in
// FP64, by Householder bulge chasing (one reflector per annihilated row or
// column segment, in the order of Lang's band-to-bidiagonal algorithm).
//
// The reference runs this stage as bandwidth-halving block sweeps
// (KAbasedSVD/src/bulgechasing.jl:33-42); LAPACK dgbbrd does it with Givens
// rotations.  Both are orthogonal reductions and leave the singular values
// unchanged; this one applies length-bw reflectors to contiguous columns of a
// compact band store, so every inner loop is a unit-stride AXPY or dot product
// of length <= 2bw.
//
// Sweep i makes row i bidiagonal:
//   T1  right reflector on columns i+1 .. i+bw annihilates A(i, i+2 .. i+bw);
//       applied to rows i+1 .. i+bw it fills the lower triangle of that block;
//   T2  left reflector on rows i+1 .. i+bw annihilates A(i+2 .. i+bw, i+1);
//       applied to columns i+1 .. i+2bw it fills rows i+1 .. i+bw up to
//       column i+2bw;
//   then, with r = i+1, while row r reaches beyond column r+bw:
//   T3  right reflector on columns r+bw .. r+2bw-1 annihilates row r beyond
//       its band; applied to rows r+1 .. r+2bw-1 it fills the lower triangle of
//       the block at (r+bw, r+bw);
//   T4  left reflector on rows r+bw .. r+2bw-1 annihilates that block's first
//       column below the diagonal; applied to columns r+bw .. r+3bw-1;
//   r += bw.
// The rest of each lower-triangular bulge is annihilated by the following
// sweeps (their T2 / T4 reach it), so the working matrix never has more than
// bw-1 subdiagonals or 2bw-1 superdiagonals (the store keeps bw and 2bw).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#include "avsvd.h"
#include "internal.hpp"

namespace {

// Householder vector (LAPACK dlarfg convention) of x[0..m-1]: x <- v with
// v[0] = 1, returns tau and sets beta; tau = 0 (H = I, beta = x[0]) when
// x[1..m-1] is zero.
double house(double *x, int m, double &beta)
{
    const double alpha = x[0];
    double scale = 0.0;
    for (int k = 1; k < m; ++k) scale = std::max(scale, std::fabs(x[k]));
    if (scale == 0.0) {
        beta = alpha;
        return 0.0;
    }
    // divisions, not reciprocals: the chased bulges decay, and a subnormal
    // scale or alpha - beta has no finite reciprocal
    scale = std::max(scale, std::fabs(alpha));
    double ss = 0.0;
    for (int k = 0; k < m; ++k) {
        const double t = x[k] / scale;
        ss += t * t;
    }
    const double nrm = scale * std::sqrt(ss);
    beta = alpha >= 0.0 ? -nrm : nrm;
    const double tau = (beta - alpha) / beta;
    const double den = alpha - beta;
    x[0] = 1.0;
    for (int k = 1; k < m; ++k) x[k] /= den;
    return tau;
}

struct Store {
    long n, ku, ld;
    std::vector<double> a;
    // element (r, c), r in [c - ku, c + kl]
    double *col(long c) { return a.data() + c * ld + ku - c; } // col(c)[r] = A(r, c)
};

// right reflector on columns c0 .. c0+m-1 applied to rows r0 .. r1
void apply_right(Store &S, long c0, int m, const double *v, double tau, long r0, long r1, double *w)
{
    if (r1 < r0) return;
    const long len = r1 - r0 + 1;
    std::memset(w, 0, sizeof(double) * len);
    for (int k = 0; k < m; ++k) {
        const double vk = v[k];
        const double *ck = S.col(c0 + k) + r0;
#pragma omp simd
        for (long r = 0; r < len; ++r) w[r] += ck[r] * vk;
    }
    for (int k = 0; k < m; ++k) {
        const double f = tau * v[k];
        double *ck = S.col(c0 + k) + r0;
#pragma omp simd
        for (long r = 0; r < len; ++r) ck[r] -= w[r] * f;
    }
}

// left reflector on rows r0 .. r0+m-1 applied to columns c0 .. c1
void apply_left(Store &S, long r0, int m, const double *v, double tau, long c0, long c1)
{
    for (long c = c0; c <= c1; ++c) {
        double *cc = S.col(c) + r0;
        double s = 0.0;
#pragma omp simd reduction(+ : s)
        for (int k = 0; k < m; ++k) s += v[k] * cc[k];
        const double f = tau * s;
#pragma omp simd
        for (int k = 0; k < m; ++k) cc[k] -= f * v[k];
    }
}

// Sweep i, step s: s = 0 is T1 + T2 (rows i .. i+bw, columns i+1 .. i+2bw);
// s >= 1 is T3 + T4 at r = i+1+(s-1)bw (rows r .. r+2bw-1, columns
// r+bw .. r+3bw-1).  Step s of sweep i+1 and step s+k of sweep i touch
// disjoint rows or disjoint columns once k >= 3, which is the lag the threaded
// chase keeps; every element then sees its updates in sweep order, so the
// result is the same for any thread count.
struct Chase {
    Store S;
    long b, last;

    long steps(long i) const
    {
        long s = 1;
        for (long r = i + 1; std::min(r + 2 * b - 1, last) > r + b; r += b) ++s;
        return s;
    }

    void step(long i, long s, double *v, double *w)
    {
        double beta;
        if (s == 0) {
            // T1: row i beyond the superdiagonal
            const long c0 = i + 1, c1 = std::min(i + b, last);
            if (c1 > c0) {
                const int m = (int)(c1 - c0 + 1);
                for (int k = 0; k < m; ++k) v[k] = S.col(c0 + k)[i];
                const double tau = house(v, m, beta);
                if (tau != 0.0) {
                    apply_right(S, c0, m, v, tau, i + 1, std::min(i + b, last), w);
                    S.col(c0)[i] = beta;
                    for (int k = 1; k < m; ++k) S.col(c0 + k)[i] = 0.0;
                }
            }
            // T2: column i+1 below the diagonal
            const long col = i + 1, r1 = std::min(i + b, last);
            if (r1 > col) {
                const int m = (int)(r1 - col + 1);
                std::memcpy(v, S.col(col) + col, sizeof(double) * m);
                const double tau = house(v, m, beta);
                if (tau != 0.0) {
                    apply_left(S, col, m, v, tau, col + 1, std::min(col + 2 * b - 1, last));
                    double *cc = S.col(col) + col;
                    cc[0] = beta;
                    for (int k = 1; k < m; ++k) cc[k] = 0.0;
                }
            }
            return;
        }
        const long r = i + 1 + (s - 1) * b;
        const long ca = r + b, cb = std::min(r + 2 * b - 1, last);
        // T3: row r beyond its band
        const int m = (int)(cb - ca + 1);
        for (int k = 0; k < m; ++k) v[k] = S.col(ca + k)[r];
        double tau = house(v, m, beta);
        if (tau != 0.0) {
            apply_right(S, ca, m, v, tau, r + 1, std::min(r + 2 * b - 1, last), w);
            S.col(ca)[r] = beta;
            for (int k = 1; k < m; ++k) S.col(ca + k)[r] = 0.0;
        }
        // T4: column ca below the diagonal
        const long r1 = std::min(ca + b - 1, last);
        if (r1 > ca) {
            const int mm = (int)(r1 - ca + 1);
            std::memcpy(v, S.col(ca) + ca, sizeof(double) * mm);
            tau = house(v, mm, beta);
            if (tau != 0.0) {
                apply_left(S, ca, mm, v, tau, ca + 1, std::min(ca + 2 * b - 1, last));
                double *cc = S.col(ca) + ca;
                cc[0] = beta;
                for (int k = 1; k < mm; ++k) cc[k] = 0.0;
            }
        }
    }
};

} // namespace

namespace avs {

avs_status bidiagonalize_band_chase(int64_t n, int32_t bw, const double *band, int64_t ldband, double *d, double *e,
                                    int nthreads)
{
    if (n <= 0) return fail(-1, "n must be positive");
    if (bw < 1) return fail(-2, "bandwidth must be >= 1");
    if (!band) return fail(-3, "null band");
    if (ldband < bw + 1) return fail(-4, "ldband < bw+1");
    if (!d) return fail(-5, "null d");
    if (n > 1 && !e) return fail(-6, "null e");
    const long b = bw;
    Chase C;
    C.b = b;
    C.last = n - 1;
    C.S.n = n;
    C.S.ku = 2 * b;
    C.S.ld = 3 * b + 1; // rows c-2b .. c+b of column c
    C.S.a.assign((size_t)C.S.ld * n, 0.0);
    for (long c = 0; c < n; ++c) {
        double *cc = C.S.col(c);
        for (long r = std::max(0L, c - b); r <= c; ++r) cc[r] = band[(b + r - c) + c * ldband];
    }
    const long nsweeps = n - 1;
    if (nthreads <= 1 || nsweeps < 2) {
        std::vector<double> v(2 * b + 1), w(2 * b + 1);
        for (long i = 0; i < nsweeps; ++i) {
            const long ns = C.steps(i);
            for (long s = 0; s < ns; ++s) C.step(i, s, v.data(), w.data());
        }
    } else {
        // sweep i runs on thread i % T; step s of sweep i waits until sweep
        // i-1 has finished step s+2 (or all its steps)
        const int T = (int)std::min<long>(nthreads, nsweeps);
        std::vector<long> nsteps(nsweeps);
        for (long i = 0; i < nsweeps; ++i) nsteps[i] = C.steps(i);
        std::unique_ptr<std::atomic<long>[]> done(new std::atomic<long>[nsweeps]);
        for (long i = 0; i < nsweeps; ++i) done[i].store(0, std::memory_order_relaxed);
        auto worker = [&](int t) {
            std::vector<double> v(2 * b + 1), w(2 * b + 1);
            for (long i = t; i < nsweeps; i += T) {
                for (long s = 0; s < nsteps[i]; ++s) {
                    if (i > 0) {
                        const long need = std::min(s + 3, nsteps[i - 1]);
                        int spins = 0;
                        while (done[i - 1].load(std::memory_order_acquire) < need)
                            if (++spins > 64) std::this_thread::yield();
                    }
                    C.step(i, s, v.data(), w.data());
                    done[i].store(s + 1, std::memory_order_release);
                }
            }
        };
        std::vector<std::thread> th;
        for (int t = 1; t < T; ++t) th.emplace_back(worker, t);
        worker(0);
        for (auto &x : th) x.join();
    }
    for (long c = 0; c < n; ++c) d[c] = C.S.col(c)[c];
    for (long c = 1; c < n; ++c) e[c - 1] = C.S.col(c)[c - 1];
    return AVS_OK;
}

} // namespace avs
