// chase.hip -- band -> upper bidiagonal on the device for the batched form:
// one wavefront per matrix runs the Householder bulge chasing of the host
// routine (bandbrd.cpp: sweep i = T1 + T2, then T3 + T4 steps down the band,
// Lang's order) in FP64 on a compact band store in HBM, so only the bidiagonal
// (d, e) crosses PCIe and the host runs dbdsdc alone.
//
// Each reflector has length <= b and touches at most 2b-1 rows or columns, so
// with b <= 32 one lane owns one row (right reflectors) or one column (left
// reflectors) of the update; the norm is a wavefront butterfly, and the
// reflector is shared through LDS.  There is no synchronisation between
// workgroups.  Every store access goes through a bounds check against the
// store's shape (rows c-2b .. c+b of column c, 0 <= r, c < N); an access
// outside it is skipped and flags the matrix, which the host reports.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace avs {
namespace {

constexpr int CHASE_THREADS = 64;

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = fmax(v, __shfl_xor(v, m, 64));
    return v;
}

struct BandStore {
    double *S;
    long ld, ku, b, last;
    int *err;

    __device__ bool inside(long r, long c) const
    {
        return r >= 0 && c >= 0 && r <= last && c <= last && r >= c - ku && r <= c + b;
    }
    __device__ long index(long r, long c) const { return c * ld + ku + r - c; }
    __device__ double *at(long r, long c) const
    {
        if (!inside(r, c)) {
            *err = 1;
            return nullptr;
        }
        return S + index(r, c);
    }
    __device__ double get(long r, long c) const
    {
        const double *q = at(r, c);
        return q ? *q : 0.0;
    }
    __device__ void set(long r, long c, double v) const
    {
        double *q = at(r, c);
        if (q) *q = v;
    }
};

// Householder reflector of the m values x held by lanes 0 .. m-1 (LAPACK
// dlarfg convention, as the host routine): returns tau (the same in every
// lane; 0 when x[1 .. m-1] is zero), sets beta, and leaves v in vsh[0 .. m-1]
// (v[0] = 1) behind a barrier.  Divisions, not reciprocals: the chased bulges
// decay, and a subnormal scale has no finite reciprocal.
__device__ double wave_house(double x, int m, int l, double &beta, double *vsh)
{
    const double alpha = __shfl(x, 0, 64);
    double sc = wave_max((l >= 1 && l < m) ? fabs(x) : 0.0);
    if (sc == 0.0) {
        beta = alpha;
        return 0.0;
    }
    sc = fmax(sc, fabs(alpha));
    const double t = l < m ? x / sc : 0.0;
    const double nrm = sc * sqrt(wave_sum(t * t));
    beta = alpha >= 0.0 ? -nrm : nrm;
    const double tau = (beta - alpha) / beta;
    const double den = alpha - beta;
    if (l < m) vsh[l] = l == 0 ? 1.0 : x / den;
    __syncthreads();
    return tau;
}

// The reflector updates load every element of a lane's row or column first
// (unrolled to the largest reflector length), so the loads are in flight
// together instead of one latency per element.  A lane's elements sit at
// base + k*stride (base >= 0); the in-store ones form an interval [klo, khi]
// of k and only those are used or written.  The loads for k outside it read
// other elements of the store allocation, which ends with MAXLEN*ld spare
// doubles so that base + (MAXLEN-1)*stride stays inside it.  A reflector
// element outside the store flags the matrix.
constexpr int MAXLEN = 32;

__device__ void apply_lane(const BandStore &A, long base, long stride, int klo, int khi, int m, double tau,
                           const double *vsh)
{
    if (klo > 0 || khi < m - 1) *A.err = 1;
    if (klo > khi) return;
    const double *src = A.S + base;
    double a[MAXLEN];
#pragma unroll
    for (int k = 0; k < MAXLEN; ++k) a[k] = src[k * stride];
    double w = 0.0;
#pragma unroll
    for (int k = 0; k < MAXLEN; ++k)
        if (k >= klo && k <= khi) w += a[k] * vsh[k];
    const double f = tau * w;
#pragma unroll
    for (int k = 0; k < MAXLEN; ++k)
        if (k >= klo && k <= khi) A.S[base + (long)k * stride] = a[k] - f * vsh[k];
}

// right reflector on columns c0 .. c0+m-1, applied to rows r0 .. r1 (lane = row);
// element (row, c0+k) is at index(row, c0) + k*(ld-1)
__device__ void apply_right(const BandStore &A, long c0, int m, double tau, long r0, long r1, int l,
                            const double *vsh)
{
    if (l > r1 - r0) return;
    const long row = r0 + l;
    const long klo = max(0L, row - c0 - A.b);
    const long khi = min(min((long)m - 1, A.last - c0), row - c0 + A.ku);
    apply_lane(A, A.index(row, c0), A.ld - 1, (int)klo, (int)khi, m, tau, vsh);
}

// left reflector on rows r0 .. r0+m-1, applied to columns c0 .. c1 (lane = column);
// element (r0+k, c) is at index(r0, c) + k
__device__ void apply_left(const BandStore &A, long r0, int m, double tau, long c0, long c1, int l,
                           const double *vsh)
{
    if (l > c1 - c0) return;
    const long c = c0 + l;
    const long klo = max(max(0L, c - A.ku - r0), -r0);
    const long khi = min(min((long)m - 1, c + A.b - r0), A.last - r0);
    apply_lane(A, A.index(r0, c), 1, (int)klo, (int)khi, m, tau, vsh);
}

// annihilate row `row` over columns c0 .. c0+m-1 from the right, applying the
// reflector to rows r0 .. r1
__device__ void right_phase(const BandStore &A, long row, long c0, int m, long r0, long r1, int l, double *vsh)
{
    double beta;
    const double x = l < m ? A.get(row, c0 + l) : 0.0;
    const double tau = wave_house(x, m, l, beta, vsh);
    if (tau == 0.0) return;
    apply_right(A, c0, m, tau, r0, r1, l, vsh);
    if (l < m) A.set(row, c0 + l, l == 0 ? beta : 0.0);
    __syncthreads();
}

// annihilate column `col` over rows col .. col+m-1 from the left, applying
// the reflector to columns col+1 .. c1
__device__ void left_phase(const BandStore &A, long col, int m, long c1, int l, double *vsh)
{
    double beta;
    const double x = l < m ? A.get(col + l, col) : 0.0;
    const double tau = wave_house(x, m, l, beta, vsh);
    if (tau == 0.0) return;
    apply_left(A, col, m, tau, col + 1, c1, l, vsh);
    if (l < m) A.set(col + l, col, l == 0 ? beta : 0.0);
    __syncthreads();
}

template <typename TS>
__global__ void __launch_bounds__(CHASE_THREADS) batch_chase_kernel(const TS *__restrict__ arena,
                                                                    const long *__restrict__ off,
                                                                    const int *__restrict__ nts,
                                                                    const int *__restrict__ order,
                                                                    const long *__restrict__ col0, int bw,
                                                                    double *store, double *de, int *err)
{
    __shared__ double vsh[CHASE_THREADS];
    const int m = order[blockIdx.x];
    const int l = threadIdx.x;
    const long b = bw, nt = nts[m], N = nt * b;
    BandStore A;
    A.ld = 3 * b + 1;
    A.ku = 2 * b;
    A.b = b;
    A.last = N - 1;
    A.S = store + col0[m] * A.ld;
    A.err = err + m;
    const TS *tiles = arena + off[m];
    // fill the store: the band of the reduced matrix from its tiles, zero elsewhere
    for (long idx = l; idx < A.ld * N; idx += CHASE_THREADS) {
        const long c = idx / A.ld, r = c - A.ku + idx % A.ld;
        double v = 0.0;
        if (r >= 0 && r >= c - b && r <= c)
            v = (double)tiles[((c / b) * nt + r / b) * b * b + (r % b) + (c % b) * b];
        A.S[idx] = v;
    }
    __syncthreads();
    const long last = A.last;
    for (long i = 0; i < last; ++i) {
        // T1: row i beyond the superdiagonal
        {
            const long c0 = i + 1, c1 = min(i + b, last);
            if (c1 > c0) right_phase(A, i, c0, (int)(c1 - c0 + 1), i + 1, min(i + b, last), l, vsh);
        }
        // T2: column i+1 below the diagonal
        {
            const long col = i + 1, r1 = min(i + b, last);
            if (r1 > col) left_phase(A, col, (int)(r1 - col + 1), min(col + 2 * b - 1, last), l, vsh);
        }
        // chase: T3 on row r beyond its band, T4 on column r+b below the diagonal
        for (long r = i + 1;; r += b) {
            const long ca = r + b, cb = min(r + 2 * b - 1, last);
            if (cb <= ca) break;
            right_phase(A, r, ca, (int)(cb - ca + 1), r + 1, min(r + 2 * b - 1, last), l, vsh);
            const long r1 = min(ca + b - 1, last);
            if (r1 > ca) left_phase(A, ca, (int)(r1 - ca + 1), min(ca + 2 * b - 1, last), l, vsh);
        }
    }
    // d at de[2 col0 .. + N), e at de[2 col0 + N .. + N - 1)
    double *d = de + 2 * col0[m];
    for (long c = l; c < N; c += CHASE_THREADS) {
        d[c] = A.get(c, c);
        if (c + 1 < N) d[N + c] = A.get(c, c + 1);
    }
}

// Singular values of the bidiagonals by bisection on their Golub-Kahan
// tridiagonal T (order 2N, zero diagonal, off-diagonals d_0, e_0, d_1, ...,
// d_{N-1}), whose eigenvalues are +-sigma.  The Sturm count of T - xI (the
// pivots q_0 = -x, q_k = -x - t_{k-1}^2 / q_{k-1}, each kept at least pivmin
// away from zero as in LAPACK dstebz) is the number of eigenvalues below x,
// so for x > 0 the number of singular values below x is count - N.  One lane
// per singular value: 64 halvings of [0, Gershgorin bound of T] put it within
// 2^-64 of the bound, below the rounding of the count itself.  Workgroup g
// handles singular values wfirst[g] .. +BIS_THREADS-1 (ascending) of matrix
// wmat[g]; t^2 is staged in LDS (2N - 1 <= BIS_LDS); the padding's N - n
// zero singular values (the smallest) are not written.
constexpr int BIS_THREADS = 256;
constexpr int BIS_LDS = 4096;

__global__ void __launch_bounds__(BIS_THREADS) batch_bisect_kernel(const double *__restrict__ de,
                                                                   const long *__restrict__ col0,
                                                                   const int *__restrict__ nts, int b,
                                                                   const int *__restrict__ wmat,
                                                                   const int *__restrict__ wfirst,
                                                                   const long *__restrict__ nreal,
                                                                   const long *__restrict__ soff,
                                                                   double *__restrict__ sigma)
{
    __shared__ double t2[BIS_LDS];
    __shared__ double red[2][BIS_THREADS];
    const int m = wmat[blockIdx.x], t = threadIdx.x;
    const long N = (long)nts[m] * b, L = 2 * N - 1;
    const double *d = de + 2 * col0[m], *e = d + N;
    // t_k^2 into LDS; row bound |t_{k-1}| + |t_k| and the largest t^2
    double bnd = 0.0, tmx = 0.0;
    for (long k = t; k <= L; k += BIS_THREADS) {
        const double tk = k < L ? fabs((k & 1) ? e[k >> 1] : d[k >> 1]) : 0.0;
        const double tp = k > 0 ? fabs(((k - 1) & 1) ? e[(k - 1) >> 1] : d[(k - 1) >> 1]) : 0.0;
        if (k < L) t2[k] = tk * tk;
        bnd = fmax(bnd, tk + tp);
        tmx = fmax(tmx, tk * tk);
    }
    red[0][t] = bnd;
    red[1][t] = tmx;
    __syncthreads();
    for (int s = BIS_THREADS / 2; s > 0; s >>= 1) {
        if (t < s) {
            red[0][t] = fmax(red[0][t], red[0][t + s]);
            red[1][t] = fmax(red[1][t], red[1][t + s]);
        }
        __syncthreads();
    }
    const double hi0 = red[0][0] * (1.0 + 1e-14) + 1e-300;
    const double pivmin = 2.2250738585072014e-308 * fmax(1.0, red[1][0]);
    const long j = wfirst[blockIdx.x] + t; // ascending index among the N singular values
    if (j >= N || j < N - nreal[m]) return;
    double lo = 0.0, hi = hi0;
    for (int it = 0; it < 64; ++it) {
        const double x = 0.5 * (lo + hi);
        double q = -x;
        if (fabs(q) < pivmin) q = -pivmin;
        long cnt = q < 0.0;
        for (long k = 0; k < L; ++k) {
            q = -x - t2[k] / q;
            if (fabs(q) < pivmin) q = -pivmin;
            cnt += q < 0.0;
        }
        if (cnt - N >= j + 1) hi = x;
        else lo = x;
    }
    sigma[soff[m] + (N - 1 - j)] = 0.5 * (lo + hi);
}

// One bidiagonal's singular values by the same bisection, t^2 read from HBM
// (every lane reads the same t2[k] at the same step); lane j of the grid finds
// the j-th smallest and the nkeep largest are written, descending.
__global__ void __launch_bounds__(BIS_THREADS) bisect_kernel(const double *__restrict__ t2, long N, long nkeep,
                                                             double hi0, double pivmin, double *__restrict__ sigma)
{
    const long j = (long)blockIdx.x * BIS_THREADS + threadIdx.x;
    if (j >= N || j < N - nkeep) return;
    const long L = 2 * N - 1;
    double lo = 0.0, hi = hi0;
    for (int it = 0; it < 64; ++it) {
        const double x = 0.5 * (lo + hi);
        double q = -x;
        if (fabs(q) < pivmin) q = -pivmin;
        long cnt = q < 0.0;
        for (long k = 0; k < L; ++k) {
            q = -x - t2[k] / q;
            if (fabs(q) < pivmin) q = -pivmin;
            cnt += q < 0.0;
        }
        if (cnt - N >= j + 1) hi = x;
        else lo = x;
    }
    sigma[N - 1 - j] = 0.5 * (lo + hi);
}

} // namespace

void launch_bisect(const double *t2, long N, long nkeep, double hi0, double pivmin, double *sigma, hipStream_t s)
{
    const long grid = (N + BIS_THREADS - 1) / BIS_THREADS;
    hipLaunchKernelGGL(bisect_kernel, dim3((unsigned)grid), dim3(BIS_THREADS), 0, s, t2, N, nkeep, hi0, pivmin,
                       sigma);
}

void launch_batch_bisect(const double *de, const long *col0, const int *nts, int b, const int *wmat,
                         const int *wfirst, int nwg, const long *nreal, const long *soff, double *sigma,
                         hipStream_t s)
{
    hipLaunchKernelGGL(batch_bisect_kernel, dim3((unsigned)nwg), dim3(BIS_THREADS), 0, s, de, col0, nts, b, wmat,
                       wfirst, nreal, soff, sigma);
}

template <typename TS>
void launch_batch_chase(const void *arena, const long *off, const int *nts, const int *order, const long *col0,
                        int count, int b, double *store, double *de, int *err, hipStream_t s)
{
    hipLaunchKernelGGL(batch_chase_kernel<TS>, dim3((unsigned)count), dim3(CHASE_THREADS), 0, s,
                       (const TS *)arena, off, nts, order, col0, b, store, de, err);
}

template void launch_batch_chase<double>(const void *, const long *, const int *, const int *, const long *, int,
                                         int, double *, double *, int *, hipStream_t);
template void launch_batch_chase<float>(const void *, const long *, const int *, const int *, const long *, int,
                                        int, double *, double *, int *, hipStream_t);

} // namespace avs
