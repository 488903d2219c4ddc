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

#include <cstdlib>
#include <cstring>

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

// Wavefront reductions on the VALU, with no LDS traffic: within a row of 16
// lanes by DPP (quad_perm xor 1, xor 2, then row_half_mirror and row_mirror,
// which pair lanes of equal partial values), across the four rows by
// v_permlane16_swap / v_permlane32_swap.  Every step adds a lane's value and
// its partner's, the same two values in every lane of a pair, so all 64 lanes
// end with the same bits.
template <int CTRL> __device__ __forceinline__ double dpp_d(double v)
{
    return __builtin_amdgcn_update_dpp(0.0, v, CTRL, 0xf, 0xf, true);
}
struct PairD { double x, y; };
template <int W> __device__ __forceinline__ PairD swap_rows_d(double v)
{
    const unsigned lo = (unsigned)__double2loint(v), hi = (unsigned)__double2hiint(v);
    if constexpr (W == 16) {
        const auto rl = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
        const auto rh = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
        return {__hiloint2double((int)rh[0], (int)rl[0]), __hiloint2double((int)rh[1], (int)rl[1])};
    } else {
        const auto rl = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
        const auto rh = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
        return {__hiloint2double((int)rh[0], (int)rl[0]), __hiloint2double((int)rh[1], (int)rl[1])};
    }
}
__device__ __forceinline__ double wave_sum_v(double v)
{
    v += dpp_d<0xB1>(v);  // quad_perm [1,0,3,2]
    v += dpp_d<0x4E>(v);  // quad_perm [2,3,0,1]
    v += dpp_d<0x141>(v); // row_half_mirror
    v += dpp_d<0x140>(v); // row_mirror
    const PairD a = swap_rows_d<16>(v);
    const PairD b = swap_rows_d<32>(a.x + a.y);
    return b.x + b.y;
}
__device__ __forceinline__ double wave_max_v(double v)
{
    v = fmax(v, dpp_d<0xB1>(v));
    v = fmax(v, dpp_d<0x4E>(v));
    v = fmax(v, dpp_d<0x141>(v));
    v = fmax(v, dpp_d<0x140>(v));
    const PairD a = swap_rows_d<16>(v);
    const PairD b = swap_rows_d<32>(fmax(a.x, a.y));
    return fmax(b.x, b.y);
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
// The per-wave helpers below synchronise the lanes of one wavefront: with a
// workgroup barrier in a one-wave workgroup (WG), or, when each wavefront of a
// larger workgroup runs its own step (!WG), with a wavefront-scope fence and
// wave barrier (no workgroup barrier may be skipped by some waves only).
template <bool WG> __device__ __forceinline__ void lanes_sync()
{
    if constexpr (WG) {
        __syncthreads();
    } else {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

template <bool WG = true>
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
    lanes_sync<WG>();
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
template <bool WG = true>
__device__ void right_phase(const BandStore &A, long row, long c0, int m, long r0, long r1, int l, double *vsh)
{
    double beta;
    const double x = l < m ? A.get(row, c0 + l) : 0.0;
    const double tau = wave_house<WG>(x, m, l, beta, vsh);
    if (tau == 0.0) return;
    apply_right(A, c0, m, tau, r0, r1, l, vsh);
    if (l < m) A.set(row, c0 + l, l == 0 ? beta : 0.0);
    lanes_sync<WG>();
}

// annihilate column `col` over rows col .. col+m-1 from the left, applying
// the reflector to columns col+1 .. c1
template <bool WG = true>
__device__ void left_phase(const BandStore &A, long col, int m, long c1, int l, double *vsh)
{
    double beta;
    const double x = l < m ? A.get(col + l, col) : 0.0;
    const double tau = wave_house<WG>(x, m, l, beta, vsh);
    if (tau == 0.0) return;
    apply_left(A, col, m, tau, col + 1, c1, l, vsh);
    if (l < m) A.set(col + l, col, l == 0 ? beta : 0.0);
    lanes_sync<WG>();
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

// The same chase with one workgroup of CHASE_WAVES wavefronts per matrix and
// the pipelined schedule of the single-matrix chase: step s of sweep i runs at
// time 3i + s (s = 0: T1 + T2; s >= 1: T3 + T4 at r = i+1+(s-1)b), the steps
// of one time touch disjoint rows or columns, and wavefront w takes the steps
// of sweeps ihi - w, ihi - w - CHASE_WAVES, ... of each time.  Each step is
// the one-wavefront code above with wavefront-local synchronisation; one
// workgroup barrier per time (the same count in every wavefront) orders the
// times.  Every element sees its updates in sweep order, so d and e are those
// of the one-wavefront kernel, bit for bit.
constexpr int CHASE_WAVES = 8;
template <typename TS>
__global__ void __launch_bounds__(CHASE_WAVES * 64) batch_chase_wg_kernel(const TS *__restrict__ arena,
                                                                        const long *__restrict__ off,
                                                                        const int *__restrict__ nts,
                                                                        const int *__restrict__ order,
                                                                        const long *__restrict__ col0, int bw,
                                                                        double *store, double *de, int *err)
{
    __shared__ double vsh_all[CHASE_WAVES * 64];
    constexpr int NT = CHASE_WAVES * 64;
    const int m = order[blockIdx.x];
    const int t0 = threadIdx.x, w = t0 >> 6, l = t0 & 63;
    double *vsh = vsh_all + 64 * w;
    const long b = bw, nt = nts[m], N = nt * b;
    BandStore A;
    A.ld = 3 * b + 1;
    A.ku = 2 * b;
    A.b = b;
    A.last = N - 1;
    A.S = store + col0[m] * A.ld;
    A.err = err + m;
    const TS *tiles = arena + off[m];
    for (long idx = t0; idx < A.ld * N; idx += NT) {
        const long c = idx / A.ld, r = c - A.ku + idx % A.ld;
        double v = 0.0;
        if (r >= 0 && r >= c - b && r <= c)
            v = (double)tiles[((c / b) * nt + r / b) * b * b + (r % b) + (c % b) * b];
        A.S[idx] = v;
    }
    __syncthreads();
    const long last = A.last;
    // steps of sweep i: T1 + T2, then T3 + T4 for r = i+1, i+1+b, ... with r + b < last
    auto nsteps = [&](long i) {
        const long x = last - i - 1 - b;
        return 1 + (x > 0 ? (x + b - 1) / b : 0);
    };
    // the last sweep (i = last - 1) has one step, at time 3(last - 1)
    const long tend = last >= 1 ? 3 * (last - 1) + 1 : 0;
    for (long t = 0; t < tend; ++t) {
        const long ihi = min(t / 3, last - 1);
        for (long k = w;; k += CHASE_WAVES) {
            const long i = ihi - k;
            if (i < 0) break;
            const long st = t - 3 * i;
            if (st >= nsteps(i)) break; // finished sweeps form a prefix
            if (st == 0) {
                const long c0 = i + 1, c1 = min(i + b, last);
                if (c1 > c0) right_phase<false>(A, i, c0, (int)(c1 - c0 + 1), i + 1, min(i + b, last), l, vsh);
                const long col = i + 1, r1 = min(i + b, last);
                if (r1 > col) left_phase<false>(A, col, (int)(r1 - col + 1), min(col + 2 * b - 1, last), l, vsh);
            } else {
                const long r = i + 1 + (st - 1) * b;
                const long ca = r + b, cb = min(r + 2 * b - 1, last);
                right_phase<false>(A, r, ca, (int)(cb - ca + 1), r + 1, min(r + 2 * b - 1, last), l, vsh);
                const long r1 = min(ca + b - 1, last);
                if (r1 > ca) left_phase<false>(A, ca, (int)(r1 - ca + 1), min(ca + 2 * b - 1, last), l, vsh);
            }
        }
        __syncthreads();
    }
    double *d = de + 2 * col0[m];
    for (long c = t0; c < N; c += NT) {
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

// One bidiagonal's singular values by the same Sturm counts, t^2 read from
// HBM (every lane reads the same t2[k] at the same step), sixteen lanes per
// singular value: in each of 16 rounds lanes 0..14 of the group count at the
// points lo + (k+1)/16 (hi - lo), the group's counts meet in one ballot, and
// all sixteen lanes keep the same sixteenth (16^-16 = 2^-64 of the
// Gershgorin bound, as 64 halvings, in a quarter of the sequential rounds and
// on sixteen times the lanes, so four wavefronts per SIMD at n = 16384 hide
// the latency of the dependent divisions).  Group j of the grid finds the
// j-th smallest; the nkeep largest are written, descending.
constexpr int BIS_LANES = 16;
__global__ void __launch_bounds__(BIS_THREADS) bisect_kernel(const double *__restrict__ t2, long N, long nkeep,
                                                             double hi0, double pivmin, double *__restrict__ sigma)
{
    const long g = (long)blockIdx.x * BIS_THREADS + threadIdx.x;
    const long j = g / BIS_LANES; // the same for the lanes of a group
    const int sub = (int)(g % BIS_LANES);
    // groups are 16-lane aligned, so a group returns or stays as a whole
    if (j >= N || j < N - nkeep) return;
    const int base = (int)(threadIdx.x & 63) & ~(BIS_LANES - 1);
    const long L = 2 * N - 1;
    double lo = 0.0, hi = hi0;
    for (int it = 0; it < 16; ++it) {
        const double w = hi - lo;
        const double x = lo + w * ((double)(min(sub, 14) + 1) * 0.0625);
        double q = -x;
        if (fabs(q) < pivmin) q = -pivmin;
        long cnt = q < 0.0;
        for (long k = 0; k < L; ++k) {
            q = -x - t2[k] / q;
            if (fabs(q) < pivmin) q = -pivmin;
            cnt += q < 0.0;
        }
        // point k lies at or above the j-th smallest value when more than j
        // singular values are below it; the first such point ends the interval
        const bool above = cnt - N >= j + 1;
        const unsigned long long bal = __ballot(above);
        const unsigned bits = (unsigned)(bal >> base) & 0x7fffu;
        const int k = bits ? __builtin_ctz(bits) : 15;
        const double nlo = k == 0 ? lo : lo + w * ((double)k * 0.0625);
        if (k < 15) hi = lo + w * ((double)(k + 1) * 0.0625);
        lo = nlo;
    }
    if (sub == 0) sigma[N - 1 - j] = 0.5 * (lo + hi);
}

// ---------------------------------------------------------------------------
// One large matrix: the same chase with the steps of many sweeps in flight.
// Step s of sweep i (s = 0: T1 + T2 on rows i .. i+b; s >= 1: T3 + T4 at
// r = i+1+(s-1)b) touches rows and columns that step s+3 of sweep i-1 and
// every later step of earlier sweeps leave alone, so with the time of (i, s)
// set to 3i + s, the steps of one time are disjoint and each element sees its
// updates in sweep order, as in the sequential loop.  One launch per time;
// workgroup k runs step tstep - 3i of sweep i = ihi - k.  A phase (one
// reflector and its update) costs one round trip to the store and two
// barriers: every thread first requests its part of the update (at most 2b-1
// rows for right reflectors, two threads per row; at most 2b-1 columns for
// left ones, a wavefront per column with lanes along it; all loads in flight
// together), while wave 0 loads the annihilated segment (two elements per
// lane, m <= 128) and forms the reflector with lane exchanges; a barrier
// publishes v and tau through LDS, and the threads update.  b <= 128.
// ---------------------------------------------------------------------------
constexpr int WCH_THREADS = 512;
constexpr int WCH_HALF = 64; // elements of a line per thread

// The reflector of a phase, on wave 0 (two elements per lane, m <= 128):
// RIGHT annihilates row p over columns c0 .. c0+m-1, LEFT column c0 over rows
// c0 .. c0+m-1.  Householder vector as wave_house (LAPACK dlarfg convention,
// divisions; tau = 0 and nothing written when the rest of the segment is
// zero): v to vsh, tau to sc2[0], and the segment becomes (beta, 0, ..., 0).
// No update line holds the segment, so it is written before the barrier.
// The reflector of the m values x held by the lanes of wave 0 (x0 for k =
// lane, x1 for k = lane + 64; zero past m): v to vsh, tau to sc2[0], and the
// segment's new values (beta, 0, ..., 0) through put(k, value).  ALWAYS: put
// is also called when tau = 0, with the unchanged values.
template <bool ALWAYS, typename Put>
__device__ __forceinline__ void wave0_house2(double x0, double x1, int m, double *vsh, double *sc2, Put put)
{
    const int k0 = threadIdx.x, k1 = k0 + 64;
    const double alpha = __shfl(x0, 0, 64);
    double sc = wave_max_v(fmax((k0 >= 1 && k0 < m) ? fabs(x0) : 0.0, k1 < m ? fabs(x1) : 0.0));
    double tau = 0.0;
    if (sc != 0.0) {
        sc = fmax(sc, fabs(alpha));
        const double u0 = x0 / sc, u1 = x1 / sc;
        const double nrm = sc * sqrt(wave_sum_v(u0 * u0 + u1 * u1));
        const double beta = alpha >= 0.0 ? -nrm : nrm;
        tau = (beta - alpha) / beta;
        const double den = alpha - beta;
        if (k0 < m) {
            vsh[k0] = k0 == 0 ? 1.0 : x0 / den;
            put(k0, k0 == 0 ? beta : 0.0);
        }
        if (k1 < m) {
            vsh[k1] = x1 / den;
            put(k1, 0.0);
        }
    } else if (ALWAYS) {
        if (k0 < m) put(k0, x0);
        if (k1 < m) put(k1, x1);
    }
    if (k0 == 0) sc2[0] = tau;
}

// The reflector of a phase, on wave 0 (two elements per lane, m <= 128):
// RIGHT annihilates row p over columns c0 .. c0+m-1, LEFT column c0 over rows
// c0 .. c0+m-1.  Householder vector as wave_house (LAPACK dlarfg convention,
// divisions; tau = 0 and nothing written when the rest of the segment is
// zero): v to vsh, tau to sc2[0], and the segment becomes (beta, 0, ..., 0).
// No update line holds the segment, so it is written before the barrier.
template <bool RIGHT>
__device__ __forceinline__ void wave0_reflector(const BandStore &A, long p, long c0, int m, double *vsh, double *sc2)
{
    const int t = threadIdx.x;
    if (t >= 64) return;
    const int k0 = t, k1 = t + 64;
    auto seg = [&](int k) { return RIGHT ? A.get(p, c0 + k) : A.get(c0 + k, c0); };
    const double x0 = k0 < m ? seg(k0) : 0.0, x1 = k1 < m ? seg(k1) : 0.0;
    wave0_house2<false>(x0, x1, m, vsh, sc2, [&](int k, double v) {
        if (RIGHT) A.set(p, c0 + k, v);
        else A.set(c0 + k, c0, v);
    });
}

// Right phase: the reflector of row p over columns c0 .. c0+m-1, applied to
// rows p+1 .. lend.  Two threads per row (element (row, c0+k) at
// index(row, c0) + k*(ld-1), consecutive rows in consecutive lane pairs), each
// holding up to 64 elements of its row's in-store interval [klo, khi]; all
// loads go out before the reflector is known.  The loads past a thread's last
// element read other elements of the store allocation (CHASE_SPARE_COLS spare
// columns at its end keep base + 127*(ld-1) inside it) and are not used.
__device__ __forceinline__ void wg_right_phase(const BandStore &A, long p, long c0, int m, long lend, double *vsh,
                                               double *sc2)
{
    const int t = threadIdx.x, line = t >> 1, h = t & 1;
    const long ln = p + 1 + line;
    const bool valid = ln <= lend;
    const long base = A.index(ln, c0), stride = A.ld - 1;
    const long klo = max(0L, ln - c0 - A.b);
    const long khi = min(min((long)m - 1, A.last - c0), ln - c0 + A.ku);
    if (valid && (klo > 0 || khi < m - 1)) *A.err = 1;
    const int kb = max((int)klo, WCH_HALF * h), ke = min((int)khi, WCH_HALF * h + WCH_HALF - 1);
    const bool mine = valid && kb <= ke;
    double a[WCH_HALF];
    if (mine) {
        const double *src = A.S + base;
#pragma unroll
        for (int u = 0; u < WCH_HALF; ++u) a[u] = src[(long)(kb + u) * stride];
    }
    wave0_reflector<true>(A, p, c0, m, vsh, sc2);
    __syncthreads();
    const double tau = sc2[0];
    if (tau != 0.0) {
        double w = 0.0;
        if (mine) {
#pragma unroll
            for (int u = 0; u < WCH_HALF; ++u) w = kb + u <= ke ? fma(a[u], vsh[min(kb + u, ke)], w) : w;
        }
        // the row's dot product, the same in both threads: lower half + upper half
        const double o = __shfl_xor(w, 1, 64);
        const double f = tau * (h == 0 ? w + o : o + w);
        if (mine) {
            double *dst = A.S + base;
#pragma unroll
            for (int u = 0; u < WCH_HALF; ++u)
                if (kb + u <= ke) dst[(long)(kb + u) * stride] = fma(-f, vsh[kb + u], a[u]);
        }
    }
    __syncthreads();
}

// Left phase: the reflector of column c0 over rows c0 .. c0+m-1, applied to
// columns c0+1 .. lend.  Wave w takes columns c0+1+w, c0+1+w+8, ... (at most
// 32, as 2b-1 <= 255); lane l holds elements k = l and l+64 of each, so the
// loads and stores of a column are contiguous (element (c0+k, c) at
// index(c0, c) + k) and its dot product is a wavefront sum.  All loads go out
// before the reflector is known; those outside a column's in-store interval
// [klo, khi] read other elements of the store allocation (index(c0, c) + 127
// stays inside it) and are not used.
__device__ __forceinline__ void wg_left_phase(const BandStore &A, long c0, int m, long lend, double *vsh,
                                              double *sc2)
{
    constexpr int NW = WCH_THREADS / 64, CPW = 32;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    double a[CPW][2];
#pragma unroll
    for (int j = 0; j < CPW; ++j) {
        const long c = c0 + 1 + wv + (long)NW * j;
        if (c <= lend) {
            const double *src = A.S + A.index(c0, c);
            a[j][0] = src[lane];
            a[j][1] = src[lane + 64];
        }
    }
    wave0_reflector<false>(A, c0, c0, m, vsh, sc2);
    __syncthreads();
    const double tau = sc2[0];
    if (tau != 0.0) {
        const double v0 = vsh[lane], v1 = vsh[lane + 64];
#pragma unroll
        for (int j = 0; j < CPW; ++j) {
            const long c = c0 + 1 + wv + (long)NW * j;
            if (c > lend) continue; // wave-uniform
            const long klo = max(max(0L, c - A.ku - c0), -c0);
            const long khi = min(min((long)m - 1, c + A.b - c0), A.last - c0);
            if (lane == 0 && (klo > 0 || khi < m - 1)) *A.err = 1;
            const bool in0 = lane >= klo && lane <= khi, in1 = lane + 64 >= klo && lane + 64 <= khi;
            double pr = in0 ? a[j][0] * v0 : 0.0;
            pr = in1 ? fma(a[j][1], v1, pr) : pr;
            const double f = tau * wave_sum(pr);
            double *dst = A.S + A.index(c0, c);
            if (in0) dst[lane] = fma(-f, v0, a[j][0]);
            if (in1) dst[lane + 64] = fma(-f, v1, a[j][1]);
        }
    }
    __syncthreads();
}

// T3 + T4 of one step (s >= 1) with the right phase's output for the left
// phase kept in LDS.  The right reflector (row r over columns ca .. ca+m-1,
// m = min(b, N - ca)) updates rows r+1 .. ca+m-1; the left reflector (column
// ca over rows ca .. ca+m-1) then reads and rewrites rows ca .. ca+m-1 of
// columns ca .. min(ca+2b-1, N-1).  Rows ca .. ca+m-1 of the right phase's
// columns therefore go to LDS (Lt[k*b + row - ca], column ca + k) instead of
// a store to HBM and a load back; the left phase writes them to HBM whether
// or not its reflector is the identity.  Columns ca+m .. of the left phase,
// which the right phase does not touch, come from HBM.  The arithmetic and its
// order are those of wg_right_phase + wg_left_phase.
// CPW: left-phase columns per wave (2b-1 <= 8*CPW); JF: the first j whose
// column lies beyond the right phase's when m = b (b = 4*CPW).
// timing hooks of wg_t34 (event ev); empty unless a profiling build defines them
#ifndef AVS_CHASE_MARK
#define AVS_CHASE_MARK0
#define AVS_CHASE_MARK(ev)
#endif
template <int CPW, int JF>
__device__ __forceinline__ void wg_t34(const BandStore &A, long r, long ca, int m, double *vsh, double *sc2,
                                       double *Lt)
{
    AVS_CHASE_MARK0;
    constexpr int NW = WCH_THREADS / 64;
    const long b = A.b, last = A.last;
    const int ldt = (int)b;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    // ---- right phase loads: two threads per row, rows r+1 .. lendR
    const long lendR = ca + m - 1; // = min(r + 2b - 1, last)
    const int line = t >> 1, h = t & 1;
    const long ln = r + 1 + line;
    const bool valid = ln <= lendR;
    const long base = A.index(ln, ca), stride = A.ld - 1;
    const long klo = max(0L, ln - ca - A.b);
    const long khi = min(min((long)m - 1, A.last - ca), ln - ca + A.ku);
    if (valid && (klo > 0 || khi < m - 1)) *A.err = 1;
    const int kb = max((int)klo, WCH_HALF * h), ke = min((int)khi, WCH_HALF * h + WCH_HALF - 1);
    const bool mine = valid && kb <= ke;
    double a[WCH_HALF];
    if (mine) {
        const double *src = A.S + base;
#pragma unroll
        for (int u = 0; u < WCH_HALF; ++u) a[u] = src[(long)(kb + u) * stride];
    }
    // ---- right reflector (row r's segment written to HBM)
    wave0_reflector<true>(A, r, ca, m, vsh, sc2);
    AVS_CHASE_MARK(1);
    __syncthreads();
    AVS_CHASE_MARK(2);
    {
        const double tau = sc2[0];
        double w = 0.0;
        if (tau != 0.0 && mine) {
#pragma unroll
            for (int u = 0; u < WCH_HALF; ++u) w = kb + u <= ke ? fma(a[u], vsh[min(kb + u, ke)], w) : w;
        }
        const double o = __shfl_xor(w, 1, 64);
        const double f = tau * (h == 0 ? w + o : o + w);
        if (mine) {
            if (tau != 0.0) {
#pragma unroll
                for (int u = 0; u < WCH_HALF; ++u)
                    if (kb + u <= ke) a[u] = fma(-f, vsh[kb + u], a[u]);
            }
            if (ln >= ca) {
                double *dst = Lt + (ln - ca);
#pragma unroll
                for (int u = 0; u < WCH_HALF; ++u)
                    if (kb + u <= ke) dst[(kb + u) * ldt] = a[u];
            } else if (tau != 0.0) {
                double *dst = A.S + base;
#pragma unroll
                for (int u = 0; u < WCH_HALF; ++u)
                    if (kb + u <= ke) dst[(long)(kb + u) * stride] = a[u];
            }
        }
    }
    AVS_CHASE_MARK(3);
    __syncthreads();
    AVS_CHASE_MARK(4);
    // ---- left phase: columns ca+1+wv+NW*j.  Those of the right phase (offset
    // < m) are read from LDS when used; the rest, untouched by it, come from
    // HBM: for j >= JF (all of them when m = b = 4*CPW) requested now, while wave 0
    // forms the reflector from column ca in LDS (written to HBM in any case)
    const long lendL = min(ca + 2 * b - 1, last);
    double ff[CPW - JF][2];
#pragma unroll
    for (int j = JF; j < CPW; ++j) {
        const long c = ca + 1 + wv + (long)NW * j;
        ff[j - JF][0] = ff[j - JF][1] = 0.0;
        if (c <= lendL && c - ca >= m) { // wave-uniform
            const double *src = A.S + A.index(ca, c);
            ff[j - JF][0] = src[lane];
            ff[j - JF][1] = src[lane + 64];
        }
    }
    if (t < 64) {
        const double x0 = lane < m ? Lt[lane] : 0.0, x1 = lane + 64 < m ? Lt[lane + 64] : 0.0;
        wave0_house2<true>(x0, x1, m, vsh, sc2, [&](int k, double v) { A.set(ca + k, ca, v); });
    }
    AVS_CHASE_MARK(5);
    __syncthreads();
    AVS_CHASE_MARK(6);
    const double tau = sc2[0];
    // column j's values: LDS (offset < m; Lt holds b*b + 128 doubles, so lane
    // + 64 of column m-1 stays inside it), the early HBM request, or HBM now
    auto colv = [&](int j, double &x0, double &x1) {
        const long c = ca + 1 + wv + (long)NW * j;
        const int off = (int)(c - ca);
        if (j >= JF && off >= m) {
            x0 = ff[j - JF][0];
            x1 = ff[j - JF][1];
        } else if (off < m) {
            x0 = Lt[off * ldt + lane];
            x1 = Lt[off * ldt + lane + 64];
        } else {
            const double *src = A.S + A.index(ca, c);
            x0 = src[lane];
            x1 = src[lane + 64];
        }
    };
    if (tau != 0.0) {
        const double v0 = vsh[lane], v1 = vsh[lane + 64];
        // the columns' dot products, 4 at a time with no branch between them,
        // so their reductions overlap
        constexpr int G = 4;
#pragma unroll
        for (int g = 0; g < CPW; g += G) {
            if (ca + 1 + wv + (long)NW * g > lendL) break; // wave-uniform
            double x[G][2], pr[G];
#pragma unroll
            for (int jj = 0; jj < G; ++jj) {
                const long c = ca + 1 + wv + (long)NW * (g + jj);
                const long klo2 = max(max(0L, c - A.ku - ca), -ca);
                const long khi2 = min(min((long)m - 1, c + A.b - ca), A.last - ca);
                const bool ok = c <= lendL;
                x[jj][0] = x[jj][1] = 0.0;
                if (ok) colv(g + jj, x[jj][0], x[jj][1]);
                const bool in0 = ok && lane >= klo2 && lane <= khi2;
                const bool in1 = ok && lane + 64 >= klo2 && lane + 64 <= khi2;
                const double p = in0 ? x[jj][0] * v0 : 0.0;
                pr[jj] = in1 ? fma(x[jj][1], v1, p) : p;
            }
#pragma unroll
            for (int jj = 0; jj < G; ++jj) pr[jj] = wave_sum_v(pr[jj]);
#pragma unroll
            for (int jj = 0; jj < G; ++jj) {
                const long c = ca + 1 + wv + (long)NW * (g + jj);
                if (c > lendL) continue; // wave-uniform
                const long klo2 = max(max(0L, c - A.ku - ca), -ca);
                const long khi2 = min(min((long)m - 1, c + A.b - ca), A.last - ca);
                if (lane == 0 && (klo2 > 0 || khi2 < m - 1)) *A.err = 1;
                const bool in0 = lane >= klo2 && lane <= khi2, in1 = lane + 64 >= klo2 && lane + 64 <= khi2;
                const double f = tau * pr[jj];
                double *dst = A.S + A.index(ca, c);
                if (in0) dst[lane] = fma(-f, v0, x[jj][0]);
                if (in1) dst[lane + 64] = fma(-f, v1, x[jj][1]);
            }
        }
    } else {
        // the identity: the right phase's rows still go to HBM
#pragma unroll
        for (int j = 0; j < CPW; ++j) {
            const long c = ca + 1 + wv + (long)NW * j;
            if (c > lendL || c - ca >= m) continue; // wave-uniform
            const int off = (int)(c - ca);
            double *dst = A.S + A.index(ca, c);
            if (lane < m) dst[lane] = Lt[off * ldt + lane];
            if (lane + 64 < m) dst[lane + 64] = Lt[off * ldt + lane + 64];
        }
    }
    AVS_CHASE_MARK(7);
}

__global__ void __launch_bounds__(WCH_THREADS) chase_step_kernel(double *store, long N, int bw, long tstep, long ihi,
                                                                 int *err, int lds)
{
    __shared__ double vsh[2 * WCH_HALF];
    __shared__ double sc2[2];
    extern __shared__ double Lt[]; // b*b + 128: the T3 -> T4 rows of a step
    const long b = bw;
    BandStore A;
    A.ld = 3 * b + 1;
    A.ku = 2 * b;
    A.b = b;
    A.last = N - 1;
    A.S = store;
    A.err = err;
    const long last = A.last;
    const long i = ihi - blockIdx.x, s = tstep - 3 * i;
    if (i < 0 || s < 0) {
        if (threadIdx.x == 0) *err = 1;
        return;
    }
    if (s == 0) {
        // T1: row i beyond the superdiagonal, applied to rows i+1 .. i+b;
        // T2: column i+1 below the diagonal, applied to columns i+2 .. i+2b
        const long c0 = i + 1, c1 = min(i + b, last);
        if (c1 > c0) wg_right_phase(A, i, c0, (int)(c1 - c0 + 1), min(i + b, last), vsh, sc2);
        const long col = i + 1, r1 = min(i + b, last);
        if (r1 > col) wg_left_phase(A, col, (int)(r1 - col + 1), min(col + 2 * b - 1, last), vsh, sc2);
        return;
    }
    // T3 on row r beyond its band, T4 on column r+b below the diagonal
    const long r = i + 1 + (s - 1) * b;
    const long ca = r + b, cb = min(r + 2 * b - 1, last);
    if (cb <= ca) {
        if (threadIdx.x == 0) *err = 1; // the host schedules only existing steps
        return;
    }
    if (lds) {
        const int m = (int)(cb - ca + 1);
        if (b <= 32) wg_t34<8, 3>(A, r, ca, m, vsh, sc2, Lt);
        else if (b <= 64) wg_t34<16, 7>(A, r, ca, m, vsh, sc2, Lt);
        else wg_t34<32, 15>(A, r, ca, m, vsh, sc2, Lt);
        return;
    }
    wg_right_phase(A, r, ca, (int)(cb - ca + 1), min(r + 2 * b - 1, last), vsh, sc2);
    const long r1 = min(ca + b - 1, last);
    if (r1 > ca) wg_left_phase(A, ca, (int)(r1 - ca + 1), min(ca + 2 * b - 1, last), vsh, sc2);
}

// de[c] = A(c, c), de[N + c] = A(c, c+1) from the store
__global__ void __launch_bounds__(256) chase_diag_kernel(const double *__restrict__ store, long N, int bw,
                                                         double *__restrict__ de)
{
    const long c = (long)blockIdx.x * 256 + threadIdx.x;
    if (c >= N) return;
    const long ld = 3L * bw + 1, ku = 2L * bw;
    de[c] = store[c * ld + ku];
    if (c + 1 < N) de[N + c] = store[(c + 1) * ld + ku - 1];
}

} // namespace

void launch_bisect(const double *t2, long N, long nkeep, double hi0, double pivmin, double *sigma, hipStream_t s)
{
    const long grid = (N * BIS_LANES + BIS_THREADS - 1) / BIS_THREADS;
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

void launch_chase_step(double *store, long N, int b, long tstep, long ilo, long ihi, int *err, hipStream_t s)
{
    // T3 -> T4 through LDS when b*b + 128 doubles fit beside the static LDS
    // (AVSVD_CHASE_LDS=0: through HBM; read once per process)
    static const long lds_max = [] {
        const char *e = getenv("AVSVD_CHASE_LDS");
        if (e && strcmp(e, "0") == 0) return 0L;
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
            return 0L;
        return (long)v - 4096;
    }();
    const size_t lds = ((size_t)b * b + 128) * sizeof(double);
    const int use = (long)lds <= lds_max ? 1 : 0;
    hipLaunchKernelGGL(chase_step_kernel, dim3((unsigned)(ihi - ilo + 1)), dim3(WCH_THREADS), use ? lds : 0, s,
                       store, N, b, tstep, ihi, err, use);
}

void launch_chase_diag(const double *store, long N, int b, double *de, hipStream_t s)
{
    hipLaunchKernelGGL(chase_diag_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, store, N, b, de);
}

template <typename TS>
void launch_batch_chase(const void *arena, const long *off, const int *nts, const int *order, const long *col0,
                        int count, int b, double *store, double *de, int *err, hipStream_t s)
{
    // AVSVD_BATCH_CHASE=wave: one wavefront per matrix, sweeps in sequence
    // (read once per process); default: a workgroup per matrix, pipelined
    static const bool one_wave = [] {
        const char *e = getenv("AVSVD_BATCH_CHASE");
        return e && strcmp(e, "wave") == 0;
    }();
    if (!one_wave) {
        hipLaunchKernelGGL(batch_chase_wg_kernel<TS>, dim3((unsigned)count), dim3(CHASE_WAVES * 64), 0, s,
                           (const TS *)arena, off, nts, order, col0, b, store, de, err);
        return;
    }
    hipLaunchKernelGGL(batch_chase_kernel<TS>, dim3((unsigned)count), dim3(CHASE_THREADS), 0, s,
                       (const TS *)arena, off, nts, order, col0, b, store, de, err);
}

template void launch_batch_chase<double>(const void *, const long *, const int *, const int *, const long *, int,
                                         int, double *, double *, int *, hipStream_t);
template void launch_batch_chase<float>(const void *, const long *, const int *, const int *, const long *, int,
                                        int, double *, double *, int *, hipStream_t);

} // namespace avs
