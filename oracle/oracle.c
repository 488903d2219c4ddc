/*
 * oracle.c -- CPU restatement of the reference's tiled SVD band-reduction path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this file; it is imported by tests/, by __graft_entry__.smoke() and by
 * bench.py's cpu_baseline leg, always as the checker (or the timed CPU
 * baseline), never as the product path.
 *
 * What it restates (paths relative to the reference repository root):
 *   or_geqrt   KAbasedSVD/src/QRkernels.jl:3-66     QR_unsafe_kernel!   (GEQRT)
 *   or_tsqrt   KAbasedSVD/src/QRkernels.jl:71-129   QR_unsafe_kernel2!  (TSQRT)
 *   or_unmqr   KAbasedSVD/src/QRkernels.jl:131-161  applyQorQt_unsafe_kernel!  (UNMQR)
 *   or_tsmqr   KAbasedSVD/src/QRkernels.jl:165-197  applyQorQt_unsafe_kernel2! (TSMQR)
 *   or_blockbidiag / or_tiled_qr
 *              KAbasedSVD/src/tiledalgos.jl:6-59 with the tile/tau mapping of
 *              KAbasedSVD/src/TiledMatrix.jl:85-113 (LQ = QR on adjoint views)
 *   or_bidiagonalize
 *              KAbasedSVD/src/bulgechasing.jl:1-42 (block bulge chasing)
 *   or_philox_normal
 *              the synthetic-input generator (Philox4x32-10, Salmon et al. SC'11,
 *              plus Box-Muller); the reference draws unseeded randn, this is the
 *              seeded stand-in used on both sides of every comparison.
 *
 * Arithmetic follows the reference loop order, one reflector at a time, in
 * FP64.  Householder convention (QRkernels.jl:23-31): nu = alpha + sign(alpha)*
 * ||[alpha;x]||, v = [1; x/nu], tau = 2*(nu/sqrt(||x||^2+nu^2))^2, R_jj =
 * -sign(alpha)*||[alpha;x]||.  One deliberate deviation: the reference's
 * sign(0)=0 turns an exactly-zero pivot into 0/0 = NaN; here sign(0)=+1 and an
 * all-zero column gives tau=0 (H = I), the LAPACK dlarfg convention.
 *
 * Pins (see tests/test_oracle.py): GEQRT/TSQRT factors and tau against LAPACK
 * dgeqrf/dtpqrt, UNMQR/TSMQR against dormqr/dtpmqrt, Blockbidiag against
 * sigma(A) from dgesdd plus band structure and Frobenius-norm invariants,
 * bidiagonalize against sigma of the band, and brute force at n <= 16.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define EL(p, ld, tr, r, c) \
    (*((tr) ? &(p)[(long)(c) + (long)(r) * (long)(ld)] : &(p)[(long)(r) + (long)(c) * (long)(ld)]))

static inline double sgn1(double x) { return x >= 0.0 ? 1.0 : -1.0; }

/* GEQRT, QRkernels.jl:3-66.  b-1 reflectors; tau[b-1] is left untouched like
 * the reference.  `tr` != 0 runs the kernel on the adjoint of the tile
 * (the LQ twin, TiledMatrix.jl:85). */
void or_geqrt(int b, double *A, long lda, int tr, double *tau)
{
    double *cache = (double *)malloc(sizeof(double) * (size_t)b);
    for (int it = 0; it < b - 1; ++it) {
        double s = 0.0;
        for (int j = it + 1; j < b; ++j) {
            double x = EL(A, lda, tr, j, it);
            s += x * x;
        }
        double alpha = EL(A, lda, tr, it, it);
        double nrm = sqrt(s + alpha * alpha);
        if (nrm == 0.0) { /* all-zero column: H = I */
            tau[it] = 0.0;
            continue;
        }
        double nu = alpha + sgn1(alpha) * nrm;
        double t2 = sqrt(s + nu * nu);
        double t = 2.0 * (nu / t2) * (nu / t2);
        tau[it] = t;
        /* cache[i] = nu*A[it,i] + sum_{j>it} A[j,it]*A[j,i]   (QRkernels.jl:36-42) */
        for (int i = it; i < b; ++i) {
            double acc = nu * EL(A, lda, tr, it, i);
            for (int j = it + 1; j < b; ++j) acc += EL(A, lda, tr, j, it) * EL(A, lda, tr, j, i);
            cache[i] = acc;
        }
        double taucorr = t / nu;
        /* row updates (QRkernels.jl:43-57) */
        for (int i = it; i < b; ++i) {
            double f = (i == it) ? taucorr : EL(A, lda, tr, i, it) * taucorr / nu;
            for (int j = it + 1; j < b; ++j) EL(A, lda, tr, i, j) -= f * cache[j];
            if (i > it) EL(A, lda, tr, i, it) = f / taucorr;
        }
        EL(A, lda, tr, it, it) -= cache[it] * taucorr;
    }
    free(cache);
}

/* TSQRT, QRkernels.jl:71-129: QR of [triu(A1); A2].  Only the upper triangle
 * of A1 is read and written; A2 is overwritten by the dense reflector block. */
void or_tsqrt(int b, double *A1, long ld1, double *A2, long ld2, int tr, double *tau)
{
    double *cache = (double *)malloc(sizeof(double) * (size_t)b);
    for (int it = 0; it < b; ++it) {
        double s = 0.0;
        for (int j = 0; j < b; ++j) {
            double x = EL(A2, ld2, tr, j, it);
            s += x * x;
        }
        double alpha = EL(A1, ld1, tr, it, it);
        double nrm = sqrt(s + alpha * alpha);
        if (nrm == 0.0) {
            tau[it] = 0.0;
            continue;
        }
        double nu = alpha + sgn1(alpha) * nrm;
        double t2 = sqrt(s + nu * nu);
        double t = 2.0 * (nu / t2) * (nu / t2);
        tau[it] = t;
        for (int i = it; i < b; ++i) {
            double acc = nu * EL(A1, ld1, tr, it, i);
            for (int j = 0; j < b; ++j) acc += EL(A2, ld2, tr, j, it) * EL(A2, ld2, tr, j, i);
            cache[i] = acc;
        }
        double taucorr = t / nu;
        for (int i = it; i < b; ++i) EL(A1, ld1, tr, it, i) -= cache[i] * taucorr;
        for (int i = 0; i < b; ++i) {
            double f = EL(A2, ld2, tr, i, it) / nu * taucorr;
            for (int j = it + 1; j < b; ++j) EL(A2, ld2, tr, i, j) -= f * cache[j];
            EL(A2, ld2, tr, i, it) = f / taucorr;
        }
    }
    free(cache);
}

/* UNMQR, QRkernels.jl:131-161.  Applies Q^T (applyt) or Q of a GEQRT tile to
 * the b x ncols block C.  vtr/ctr select adjoint access of V and C. */
void or_unmqr(int b, const double *V, long ldv, int vtr, const double *tau,
              double *C, long ldc, int ctr, long ncols, int applyt)
{
    for (long col = 0; col < ncols; ++col) {
        for (int s = 0; s < b - 1; ++s) {
            int k = applyt ? s : (b - 2 - s);
            double acc = EL(C, ldc, ctr, k, col);
            for (int j = k + 1; j < b; ++j) acc += EL(V, ldv, vtr, j, k) * EL(C, ldc, ctr, j, col);
            acc *= tau[k];
            for (int j = k + 1; j < b; ++j) EL(C, ldc, ctr, j, col) -= acc * EL(V, ldv, vtr, j, k);
            EL(C, ldc, ctr, k, col) -= acc;
        }
    }
}

/* TSMQR, QRkernels.jl:165-197.  Applies Q^T (applyt) or Q of a TSQRT pair to
 * [C1; C2], each b x ncols. */
void or_tsmqr(int b, double *C1, long ld1, double *C2, long ld2, int ctr,
              const double *V, long ldv, int vtr, const double *tau, long ncols, int applyt)
{
    for (long col = 0; col < ncols; ++col) {
        for (int s = 0; s < b; ++s) {
            int k = applyt ? s : (b - 1 - s);
            double acc = EL(C1, ld1, ctr, k, col);
            for (int j = 0; j < b; ++j) acc += EL(V, ldv, vtr, j, k) * EL(C2, ld2, ctr, j, col);
            acc *= tau[k];
            EL(C1, ld1, ctr, k, col) -= acc;
            for (int j = 0; j < b; ++j) EL(C2, ld2, ctr, j, col) -= acc * EL(V, ldv, vtr, j, k);
        }
    }
}

static void zero_strict_lower(int b, double *T, long ld, int tr)
{
    for (int c = 0; c < b; ++c)
        for (int r = c + 1; r < b; ++r) EL(T, ld, tr, r, c) = 0.0;
}

static void zero_tile(int b, double *T, long ld)
{
    for (int c = 0; c < b; ++c) memset(T + (long)c * ld, 0, sizeof(double) * (size_t)b);
}

#define TILE(A, n, b, i, j) ((A) + (long)(i) * (b) + (long)(j) * (b) * (long)(n))

/* tiledalgos.jl:17-34 (svd=1, Blockbidiag!) and :6-15 (svd=0, tiled_QR!), with
 * the helpers QRandmulQt1!/QRandmulQt2!/LQandmulQt1!/LQandmulQt2! (:37-59).
 * A is column-major n x n (lda = n), reduced in place; kend steps (1-based
 * count, like the reference's kend). */
static void or_sweeps(long n, int b, double *A, int kend, int svd)
{
    int nt = (int)(n / b);
    double *tau = (double *)calloc((size_t)b, sizeof(double));
    for (int k = 0; k < kend && k < nt; ++k) {
        /* QR sweep on tile column k */
        or_geqrt(b, TILE(A, n, b, k, k), n, 0, tau);
        if (k + 1 < nt)
            or_unmqr(b, TILE(A, n, b, k, k), n, 0, tau, TILE(A, n, b, k, k + 1), n, 0,
                     (long)(nt - k - 1) * b, 1);
        if (svd) zero_strict_lower(b, TILE(A, n, b, k, k), n, 0);
        for (int i = k + 1; i < nt; ++i) {
            or_tsqrt(b, TILE(A, n, b, k, k), n, TILE(A, n, b, i, k), n, 0, tau);
            if (k + 1 < nt)
                or_tsmqr(b, TILE(A, n, b, k, k + 1), n, TILE(A, n, b, i, k + 1), n, 0,
                         TILE(A, n, b, i, k), n, 0, tau, (long)(nt - k - 1) * b, 1);
            if (svd) zero_tile(b, TILE(A, n, b, i, k), n);
        }
        if (!svd) continue;
        if (k == nt - 1) break;
        /* LQ sweep on tile row k from column k+1 (adjoint views) */
        or_geqrt(b, TILE(A, n, b, k, k + 1), n, 1, tau);
        /* LQtapply1_par!: C = rows below tile k in tile column k+1, adjoint:
         * the adjoint of an (n-(k+1)b) x b block is b x (n-(k+1)b). */
        or_unmqr(b, TILE(A, n, b, k, k + 1), n, 1, tau, TILE(A, n, b, k + 1, k + 1), n, 1,
                 (long)(nt - k - 1) * b, 1);
        /* tril!: zero the strict upper triangle of tile (k,k+1) */
        zero_strict_lower(b, TILE(A, n, b, k, k + 1), n, 1);
        for (int j = k + 2; j < nt; ++j) {
            or_tsqrt(b, TILE(A, n, b, k, k + 1), n, TILE(A, n, b, k, j), n, 1, tau);
            or_tsmqr(b, TILE(A, n, b, k + 1, k + 1), n, TILE(A, n, b, k + 1, j), n, 1,
                     TILE(A, n, b, k, j), n, 1, tau, (long)(nt - k - 1) * b, 1);
            zero_tile(b, TILE(A, n, b, k, j), n);
        }
    }
    free(tau);
}

void or_blockbidiag(long n, int b, double *A, int kend) { or_sweeps(n, b, A, kend, 1); }
void or_tiled_qr(long n, int b, double *A, int kend) { or_sweeps(n, b, A, kend, 0); }

/* ---------------------------------------------------------------------------
 * Host stage: KAbasedSVD/src/bulgechasing.jl:1-42 on a dense column-major
 * n x n matrix.  qr! on a block is restated as an unblocked Householder QR
 * that keeps R and drops the reflectors (the reference keeps triu of the
 * factors, :4, :11).
 * ------------------------------------------------------------------------- */

/* In-place Householder QR of an m x nc block B (ld), leaving R in the upper
 * trapezoid and exact zeros below it. */
static void hh_qr_keepR(long m, long nc, double *B, long ld)
{
    long kmax = m < nc ? m : nc;
    for (long j = 0; j < kmax; ++j) {
        double s = 0.0;
        for (long r = j + 1; r < m; ++r) s += B[r + j * ld] * B[r + j * ld];
        double alpha = B[j + j * ld];
        double nrm = sqrt(s + alpha * alpha);
        if (nrm == 0.0) continue;
        double beta = -sgn1(alpha) * nrm;
        double t = (beta - alpha) / beta;
        double scal = 1.0 / (alpha - beta);
        for (long r = j + 1; r < m; ++r) B[r + j * ld] *= scal;
        for (long c = j + 1; c < nc; ++c) {
            double w = B[j + c * ld];
            for (long r = j + 1; r < m; ++r) w += B[r + j * ld] * B[r + c * ld];
            w *= t;
            B[j + c * ld] -= w;
            for (long r = j + 1; r < m; ++r) B[r + c * ld] -= w * B[r + j * ld];
        }
        B[j + j * ld] = beta;
        for (long r = j + 1; r < m; ++r) B[r + j * ld] = 0.0;
    }
}

/* QR_col!(A, s, l, gap), bulgechasing.jl:9-13 (1-based inclusive indices). */
static void qr_col(double *A, long n, long s, long l, long gap)
{
    long m = l - s + 1, nc = l + gap - s + 1;
    if (m <= 0 || nc <= 0) return;
    hh_qr_keepR(m, nc, A + (s - 1) + (s - 1) * n, n);
}

/* QR_row!(A, s, l, gap), bulgechasing.jl:1-7: the block rows s..l, columns
 * s+gap..l is replaced by L from its LQ factorisation. */
static void qr_row(double *A, long n, long s, long l, long gap)
{
    long m = l - s + 1, nc = l - s - gap + 1;
    if (m <= 0 || nc <= 0) return;
    double *blk = A + (s - 1) + (s - 1 + gap) * n;
    double *t = (double *)malloc(sizeof(double) * (size_t)(m * nc));
    for (long r = 0; r < m; ++r)
        for (long c = 0; c < nc; ++c) t[c + r * nc] = blk[r + c * n]; /* t = blk' (nc x m) */
    hh_qr_keepR(nc, m, t, nc);
    for (long r = 0; r < m; ++r)
        for (long c = 0; c < nc; ++c) blk[r + c * n] = t[c + r * nc];
    free(t);
}

static long lmin(long a, long b) { return a < b ? a : b; }
static long lmax(long a, long b) { return a > b ? a : b; }

/* block_bidiagonalize!(A, n, bandwidth, target_bandwidth), bulgechasing.jl:17-31 */
void or_block_bidiagonalize(double *A, long n, long bw, long tbw)
{
    for (long j = 1; j <= n - tbw; j += tbw) {
        qr_row(A, n, j, lmin(j + bw + tbw - 1, n), tbw);
        for (long i = j; i <= n; i += bw) {
            long last = lmin(i + bw + tbw - 1, n);
            long sc = lmin(bw + tbw - 1, lmax(n - i - bw - tbw + 1, 0));
            qr_col(A, n, i + tbw, last, sc);
            if (i + tbw > n - bw) break;
            qr_row(A, n, i + tbw, last + sc, bw);
        }
    }
}

/* bidiagonalize(A, bandwidth), bulgechasing.jl:33-42.  Julia's round() rounds
 * half to even; bandwidths here are powers of two so bw/2 is exact. */
void or_bidiagonalize(long n, long bw, double *A)
{
    while (bw > 16) {
        long half = (long)nearbyint((double)bw / 2.0);
        or_block_bidiagonalize(A, n, half * 2, half);
        bw = half;
    }
    or_block_bidiagonalize(A, n, bw, 1);
}

/* ---------------------------------------------------------------------------
 * Seeded synthetic inputs: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11)
 * and Box-Muller.  Counter = (pair index lo, pair index hi, 0, 0), key =
 * (seed lo, seed hi); one call yields two normals (elements 2p and 2p+1).
 * ------------------------------------------------------------------------- */
static void philox4x32_10(uint32_t ctr[4], uint32_t key[2])
{
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)M0 * ctr[0];
        uint64_t p1 = (uint64_t)M1 * ctr[2];
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ ctr[1] ^ k0, n2 = hi0 ^ ctr[3] ^ k1;
        ctr[0] = n0;
        ctr[1] = lo1;
        ctr[2] = n2;
        ctr[3] = lo0;
        k0 += W0;
        k1 += W1;
    }
}

static inline double u01(uint32_t a, uint32_t b)
{
    uint64_t x = (((uint64_t)a << 32) | b) >> 11; /* 53 random bits */
    return ((double)x + 0.5) * (1.0 / 9007199254740992.0); /* in (0,1) */
}

void or_philox_normal(uint64_t seed, long first, long count, double scale, double *out)
{
    for (long e = first; e < first + count; ++e) {
        uint64_t p = (uint64_t)e >> 1;
        uint32_t ctr[4] = {(uint32_t)p, (uint32_t)(p >> 32), 0u, 0u};
        uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
        philox4x32_10(ctr, key);
        double u1 = u01(ctr[0], ctr[1]), u2 = u01(ctr[2], ctr[3]);
        double rad = sqrt(-2.0 * log(u1));
        double th = 6.283185307179586476925286766559 * u2;
        out[e - first] = scale * ((e & 1) ? rad * sin(th) : rad * cos(th));
    }
}
