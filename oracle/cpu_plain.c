/*
 * cpu_plain.c -- the reference's CPU plain path for the band reduction,
 * restated in C.  TEST INFRASTRUCTURE / CPU BASELINE ONLY: imported through
 * oracle.py by tests/ and by bench.py's cpu_baseline leg; the shipped library
 * never links or calls it.
 *
 * Restates (paths relative to the reference repository root)
 * CUDAbasedSVD/SVD_GPU.jl:67-214, BandBidiagonal!(A, b, b, nt, nt, onGPU=false):
 *   for each pivot tile k:
 *     QR sweep (:173-187): the tiles of row k (columns k..nt) go into the top
 *       halves of nt-k+1 cached 2b x b blocks.  The first block's top tile is
 *       factored (Julia qr! = LAPACK dgeqrt, block size min(b, 36)) and Q' is
 *       applied to the other blocks' top tiles (lmul! = dgemqrt 'L','T');
 *       then for every row i > k the tiles of row i go into the bottom halves,
 *       the whole 2b x b first block is factored and Q' applied to the other
 *       2b x b blocks, the first block is made upper triangular (triu!), and
 *       the bottom halves go back to row i.  Finally the top halves go back
 *       to row k.
 *     LQ sweep (:190-211): the tiles of column k+1 (rows k..nt) go into
 *       2b x 2b cached blocks; the first block's tile is transposed, factored
 *       and Q applied from the right to the other tiles (rmul! = dgemqrt
 *       'R','N'); then for every column c > k+1 the tiles of column c go into
 *       the right halves, [L'; A(k,c)'] (2b x b) is factored and Q applied
 *       from the right to every [A(i,k+1), A(i,c)] (b x 2b), A(k,c) is zeroed,
 *       and the right halves go back to column c.  Finally the left halves go
 *       back to column k+1 (L = lower triangle, tril!).
 * The reference's per-block loop is sequential with a multithreaded BLAS (the
 * Threads.@spawn over the blocks is commented out, SVD_GPU.jl:127-128); here
 * the same loop over the cached blocks can run on `nthreads` threads with a
 * single-threaded BLAS per call, the parallel form the reference sketches.
 * The arithmetic of every block is unchanged by the thread count.
 *
 * LAPACK comes from a shared library the caller names (cp_init), resolved
 * with dlopen: dgeqrt_ and dgemqrt_ under a prefix (scipy's OpenBLAS exports
 * them as scipy_dgeqrt_ / scipy_dgemqrt_).
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <stdio.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

typedef void (*geqrt_t)(const int *m, const int *n, const int *nb, double *a, const int *lda, double *t,
                        const int *ldt, double *work, int *info);
typedef void (*gemqrt_t)(const char *side, const char *trans, const int *m, const int *n, const int *k,
                         const int *nb, const double *v, const int *ldv, const double *t, const int *ldt, double *c,
                         const int *ldc, double *work, int *info, size_t, size_t);
typedef void (*setthreads_t)(int);
typedef int (*getthreads_t)(void);

static geqrt_t p_geqrt;
static gemqrt_t p_gemqrt;
static setthreads_t p_setthreads;
static getthreads_t p_getthreads;

/* Resolves dgeqrt / dgemqrt (and the BLAS thread setter, if present) in the
 * LAPACK library at `path`, symbol names prefix + "dgeqrt_" etc.  0 on success. */
int cp_init(const char *path, const char *prefix)
{
    void *h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h) return 1;
    char name[128];
    snprintf(name, sizeof name, "%sdgeqrt_", prefix);
    p_geqrt = (geqrt_t)dlsym(h, name);
    snprintf(name, sizeof name, "%sdgemqrt_", prefix);
    p_gemqrt = (gemqrt_t)dlsym(h, name);
    snprintf(name, sizeof name, "%sopenblas_set_num_threads", prefix);
    p_setthreads = (setthreads_t)dlsym(h, name);
    if (!p_setthreads) p_setthreads = (setthreads_t)dlsym(h, "openblas_set_num_threads");
    snprintf(name, sizeof name, "%sopenblas_get_num_threads", prefix);
    p_getthreads = (getthreads_t)dlsym(h, name);
    return (p_geqrt && p_gemqrt) ? 0 : 2;
}

/* BLAS threads for the calls below (no-op if the library has no setter);
 * returns the previous count (0 if unknown). */
int cp_blas_threads(int n)
{
    const int prev = p_getthreads ? p_getthreads() : 0;
    if (p_setthreads) p_setthreads(n);
    return prev;
}

#define AT(A, n, r, c) ((A)[(long)(r) + (long)(c) * (long)(n)])

/* tile (ti, tj) of the n x n column-major A <-> block rows/cols offset in a
 * cached block with leading dimension ldb */
static void get_tile(const double *A, long n, int b, long ti, long tj, double *B, int ldb, int r0, int c0)
{
    for (int c = 0; c < b; ++c)
        memcpy(B + (r0 + (long)(c0 + c) * ldb), &AT(A, n, ti * b, tj * b + c), sizeof(double) * (size_t)b);
}
static void put_tile(double *A, long n, int b, long ti, long tj, const double *B, int ldb, int r0, int c0)
{
    for (int c = 0; c < b; ++c)
        memcpy(&AT(A, n, ti * b, tj * b + c), B + (r0 + (long)(c0 + c) * ldb), sizeof(double) * (size_t)b);
}

/* one parallel application: blocks 1..nb-1 of `blk` (each ld x cols, leading
 * dimension ld), the factor (V, T) of the first block */
typedef struct {
    double **blk;
    int first, last;     /* block range */
    char side, trans;
    int m, n, k, nb, ldv, ldt, ldc;
    const double *V, *T;
    int status;
} apply_job;

static void *apply_run(void *p)
{
    apply_job *j = (apply_job *)p;
    const int wsz = (j->side == 'L' ? j->n : j->m) * j->nb;
    double *work = (double *)malloc(sizeof(double) * (size_t)(wsz > 0 ? wsz : 1));
    for (int q = j->first; q < j->last; ++q) {
        int info = 0;
        p_gemqrt(&j->side, &j->trans, &j->m, &j->n, &j->k, &j->nb, j->V, &j->ldv, j->T, &j->ldt, j->blk[q], &j->ldc,
                 work, &info, 1, 1);
        if (info) j->status = info;
    }
    free(work);
    return NULL;
}

static int apply_all(double **blk, int nblk, char side, char trans, int m, int n, int k, int nb, const double *V,
                     int ldv, const double *T, int ldt, int ldc, int nthreads)
{
    const int count = nblk - 1; /* blocks 1 .. nblk-1 */
    if (count <= 0) return 0;
    int nth = nthreads < 1 ? 1 : nthreads;
    if (nth > count) nth = count;
    apply_job jobs[64];
    pthread_t th[64];
    if (nth > 64) nth = 64;
    for (int t = 0; t < nth; ++t) {
        jobs[t] = (apply_job){blk, 1 + (count * t) / nth, 1 + (count * (t + 1)) / nth, side, trans, m, n, k, nb,
                              ldv, ldt, ldc, V, T, 0};
    }
    for (int t = 1; t < nth; ++t) pthread_create(&th[t], NULL, apply_run, &jobs[t]);
    apply_run(&jobs[0]);
    int st = jobs[0].status;
    for (int t = 1; t < nth; ++t) {
        pthread_join(th[t], NULL);
        if (jobs[t].status) st = jobs[t].status;
    }
    return st;
}

/* BandBidiagonal!(A, b, b, nt, nt, false) on the n x n column-major A in
 * place (n a multiple of b).  Returns 0, or a LAPACK info / setup error. */
int cp_band_bidiagonal(long n, int b, double *A, int nthreads)
{
    if (!p_geqrt || !p_gemqrt) return -1;
    if (n % b) return -2;
    const int nt = (int)(n / b);
    const int nbq = b < 36 ? b : 36; /* Julia's qr! block size */
    const int b2 = 2 * b;
    double *T = (double *)malloc(sizeof(double) * (size_t)nbq * b);
    double *work = (double *)malloc(sizeof(double) * (size_t)nbq * b2);
    double **blk = (double **)malloc(sizeof(double *) * (size_t)(nt + 1));
    int st = 0, info = 0;
    for (int k = 0; k < nt && !st; ++k) {
        /* ---------------- QR sweep on tile column k */
        const int nq = nt - k; /* cached 2b x b blocks: columns k .. nt-1 */
        for (int q = 0; q < nq; ++q) {
            blk[q] = (double *)calloc((size_t)b2 * b, sizeof(double));
            get_tile(A, n, b, k, k + q, blk[q], b2, 0, 0);
        }
        /* QR_single_col!: top tile of the first block, Q' on the other tops */
        p_geqrt(&b, &b, &nbq, blk[0], &b2, T, &nbq, work, &info);
        if (info) st = info;
        st = st ? st : apply_all(blk, nq, 'L', 'T', b, b, b, nbq, blk[0], b2, T, nbq, b2, nthreads);
        for (int c = 0; c < b; ++c) /* triu! */
            for (int r = c + 1; r < b; ++r) blk[0][r + (long)c * b2] = 0.0;
        /* QR_double_col! for every row below */
        for (int i = k + 1; i < nt && !st; ++i) {
            for (int q = 0; q < nq; ++q) get_tile(A, n, b, i, k + q, blk[q], b2, b, 0);
            p_geqrt(&b2, &b, &nbq, blk[0], &b2, T, &nbq, work, &info);
            if (info) st = info;
            st = st ? st : apply_all(blk, nq, 'L', 'T', b2, b, b, nbq, blk[0], b2, T, nbq, b2, nthreads);
            for (int c = 0; c < b; ++c) /* triu! of the 2b x b block */
                for (int r = c + 1; r < b2; ++r) blk[0][r + (long)c * b2] = 0.0;
            for (int q = 0; q < nq; ++q) put_tile(A, n, b, i, k + q, blk[q], b2, b, 0);
        }
        for (int q = 0; q < nq; ++q) {
            put_tile(A, n, b, k, k + q, blk[q], b2, 0, 0);
            free(blk[q]);
        }
        if (k == nt - 1 || st) break;
        /* ---------------- LQ sweep on tile row k from column k+1 */
        const int nl = nt - k; /* cached 2b x 2b blocks: rows k .. nt-1 */
        for (int q = 0; q < nl; ++q) {
            blk[q] = (double *)calloc((size_t)b2 * b2, sizeof(double));
            get_tile(A, n, b, k + q, k + 1, blk[q], b2, 0, 0);
        }
        double *B0 = blk[0];
        /* QR_single_row!: transpose the first tile, factor, apply Q from the
         * right to the other blocks' first tiles, transpose back, tril! */
        for (int c = 0; c < b; ++c)
            for (int r = c + 1; r < b; ++r) {
                const double x = B0[r + (long)c * b2];
                B0[r + (long)c * b2] = B0[c + (long)r * b2];
                B0[c + (long)r * b2] = x;
            }
        p_geqrt(&b, &b, &nbq, B0, &b2, T, &nbq, work, &info);
        if (info) st = info;
        st = st ? st : apply_all(blk, nl, 'R', 'N', b, b, b, nbq, B0, b2, T, nbq, b2, nthreads);
        for (int c = 0; c < b; ++c)
            for (int r = c + 1; r < b; ++r) {
                const double x = B0[r + (long)c * b2];
                B0[r + (long)c * b2] = B0[c + (long)r * b2];
                B0[c + (long)r * b2] = x;
            }
        for (int c = 0; c < b; ++c) /* tril! */
            for (int r = 0; r < c; ++r) B0[r + (long)c * b2] = 0.0;
        /* QR_double_row! for every column to the right */
        for (int cc = k + 2; cc < nt && !st; ++cc) {
            for (int q = 0; q < nl; ++q) get_tile(A, n, b, k + q, cc, blk[q], b2, 0, b);
            /* first block: [L'; A(k,cc)'] in its left 2b x b */
            for (int c = 0; c < b; ++c)
                for (int r = c + 1; r < b; ++r) {
                    const double x = B0[r + (long)c * b2];
                    B0[r + (long)c * b2] = B0[c + (long)r * b2];
                    B0[c + (long)r * b2] = x;
                }
            for (int c = 0; c < b; ++c)
                for (int r = 0; r < b; ++r) B0[(b + r) + (long)c * b2] = B0[c + (long)(b + r) * b2];
            p_geqrt(&b2, &b, &nbq, B0, &b2, T, &nbq, work, &info);
            if (info) st = info;
            st = st ? st : apply_all(blk, nl, 'R', 'N', b, b2, b, nbq, B0, b2, T, nbq, b2, nthreads);
            for (int c = 0; c < b; ++c)
                for (int r = c + 1; r < b; ++r) {
                    const double x = B0[r + (long)c * b2];
                    B0[r + (long)c * b2] = B0[c + (long)r * b2];
                    B0[c + (long)r * b2] = x;
                }
            for (int c = 0; c < b; ++c) /* tril! of the left 2b x b (top tile) */
                for (int r = 0; r < c; ++r) B0[r + (long)c * b2] = 0.0;
            for (int c = 0; c < b; ++c) /* A(k, cc) = 0 */
                memset(B0 + (long)(b + c) * b2, 0, sizeof(double) * (size_t)b);
            for (int q = 0; q < nl; ++q) put_tile(A, n, b, k + q, cc, blk[q], b2, 0, b);
        }
        for (int q = 0; q < nl; ++q) {
            put_tile(A, n, b, k + q, k + 1, blk[q], b2, 0, 0);
            free(blk[q]);
        }
    }
    free(blk);
    free(work);
    free(T);
    return st;
}
