// device.hpp -- gfx950 device code of the tiled band-reduction engine.
//
// Tile operators (names after PLASMA; reference kernels in
// KAbasedSVD/src/QRkernels.jl):
//   panel_body<MODE_GE>  GEQRT  (QR_unsafe_kernel!,  QRkernels.jl:3-66)
//   panel_body<MODE_TS>  TSQRT  (QR_unsafe_kernel2!, QRkernels.jl:71-129)
//   panel_body<MODE_TT>  TSQRT whose bottom tile is upper triangular (tree merges)
//   update_body<MODE_GE> UNMQR  (applyQorQt_unsafe_kernel!,  QRkernels.jl:131-161)
//   update_body<MODE_TS> TSMQR  (applyQorQt_unsafe_kernel2!, QRkernels.jl:165-197)
//   update_body<MODE_TT> TSMQR with an upper-triangular reflector block
// `tr` runs an operator on the adjoint of its tiles: the LQ twins of the
// reference (TiledMatrix.jl:85-88), here a change of addressing inside one
// tile-major layout rather than a strided adjoint view.
//
// Numerics: the panel keeps the tile in registers and does all Householder
// arithmetic (norms, tau, dot products, rank-1 updates) in FP64 whatever the
// storage type.  The update applies the b reflectors of a panel as b/16
// compact-WY blocks Q_J = I - V_J T_JJ V_J^T (V_J = 16 reflectors), each as
// MFMA contractions; only the 16 x 16 diagonal blocks T_JJ of the triangular
// factor are needed, and the panel produces them from the in-block Gram
// entries it computes alongside the reflectors.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "task.hpp"

namespace avs {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));
typedef short s4_t __attribute__((ext_vector_type(4)));
typedef unsigned short u8_t __attribute__((ext_vector_type(8)));

template <typename T> struct AccT;
template <> struct AccT<double> { typedef d4_t type; };
template <> struct AccT<float> { typedef f4_t type; };

__device__ __forceinline__ d4_t mfma16(double a, double b, d4_t c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4_t mfma16(float a, float b, f4_t c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

constexpr int NTHREADS = 512; // 8 wave64 per workgroup
constexpr int NWAVES = NTHREADS / 64;

// element (r, c) of a tile with leading dimension ld, or of its adjoint
__device__ __forceinline__ long toff(int r, int c, long ld, bool tr)
{
    return tr ? (long)c + (long)r * ld : (long)r + (long)c * ld;
}

// LDS footprint (in doubles) of each body; the executor allocates the max.
template <typename TS, int B> struct Lds {
    static constexpr int G = NTHREADS / B;
    static constexpr int panel = B * (B + 1) / 2 + 2 * G + G * B + B + 4 + 16 * B + B + 16 * B;
    static constexpr int update = (B * (B + 1) * (int)sizeof(TS) + 16 * B * (int)sizeof(TS) + 7) / 8;
    static constexpr int tjj = 16 * B + B + 16 * B;
    // split-BF16 update (FP32 storage): Y and Y^T as bf16 hi/lo pairs + T_JJ in FP32
    static constexpr int SPLIT_S = B + 8;
    static constexpr int update_split = sizeof(TS) == 4 ? (4 * B * SPLIT_S * 2 + 16 * B * 4 + 7) / 8 : 0;
    static constexpr int upd = update > update_split ? update : update_split;
    static constexpr int maxv = panel > upd ? panel : upd; // bodies run by the persistent executor
};

// ---------------------------------------------------------------------------
// Panel: Householder QR of one tile (GE), of [triu(top); bot] (TS) or of
// [triu(top); triu(bot)] (TT).  Thread t owns column c = t % B and the RP
// consecutive rows g*RP.. of it (g = t / B).  Per reflector j: (1) column-j
// norm partials, (2) tau/beta/scale (FP64; an all-zero column gives tau = 0),
// scaled reflector broadcast through LDS, (3) dot products v_j . a_c for all c
// (c > j: the update coefficients; c < j in j's 16-block: Gram entries for
// T_JJ), (4) rank-1 update.  Three workgroup barriers per reflector.
// Outputs: reflectors in place (GE: strict lower part, unit diagonal
// implicit; TS: the whole bottom tile; TT: the upper triangle of bot), R in
// the upper triangle of top, tau[B] (FP64) and tjj[B*16] (FP64, row i holds
// T[i][16*(i/16) + 0..15], the diagonal block of the compact-WY factor).
// ---------------------------------------------------------------------------
template <typename TS, int B, int MODE>
__device__ void panel_body(TS *__restrict__ top, long ldt, TS *__restrict__ bot, long ldb, bool tr,
                           double *__restrict__ tau, double *__restrict__ tjj, double *smem)
{
    constexpr int G = NTHREADS / B;
    constexpr int RP = B / G;
    const int t = threadIdx.x;
    const int c = t % B, g = t / B;

    double *Rp = smem;                       // packed upper triangle of R (TS/TT)
    double *partN = Rp + B * (B + 1) / 2;    // G
    double *partD = partN + 2 * G;           // G*B
    double *ybuf = partD + G * B;            // B
    double *misc = ybuf + B;                 // 4
    double *gramL = misc + 4;                // 16*B: gramL[j*16 + (c&15)] = v_c . v_j
    double *tauL = gramL + 16 * B;           // B
    double *TJ = tauL + B;                   // 16*B

    auto rpk = [](int r, int cc) { return r * B - (r * (r - 1)) / 2 + (cc - r); };

    TS *src = (MODE == MODE_GE) ? top : bot;
    const long lds = (MODE == MODE_GE) ? ldt : ldb;

    // this thread's column c, rows g*RP + r at src_c + r*rs
    const long rs = tr ? lds : 1;
    TS *src_c = src + toff(g * RP, c, lds, tr);
    double a[RP];
#pragma unroll
    for (int r = 0; r < RP; ++r) {
        const int row = g * RP + r;
        a[r] = (MODE == MODE_TT && row > c) ? 0.0 : (double)src_c[(long)r * rs];
    }
    if (MODE != MODE_GE) {
        for (int idx = t; idx < B * B; idx += NTHREADS) {
            int r, cc;
            if (tr) { cc = idx % B; r = idx / B; } else { r = idx % B; cc = idx / B; }
            if (r <= cc) Rp[rpk(r, cc)] = (double)top[toff(r, cc, ldt, tr)];
        }
    }
    for (int idx = t; idx < 16 * B; idx += NTHREADS) gramL[idx] = 0.0;
    for (int idx = t; idx < B; idx += NTHREADS) tauL[idx] = 0.0;
    __syncthreads();

    const int jend = (MODE == MODE_GE) ? B - 1 : B;
    for (int j = 0; j < jend; ++j) {
        // (1) norm partials of column j
        if (c == j) {
            double ss = 0.0;
#pragma unroll
            for (int r = 0; r < RP; ++r) {
                const int row = g * RP + r;
                if (MODE != MODE_GE || row > j) ss += a[r] * a[r];
                if (MODE == MODE_GE && row == j) misc[0] = a[r];
            }
            partN[g] = ss;
        }
        __syncthreads();
        // (2) reflector
        double s = 0.0;
#pragma unroll
        for (int gg = 0; gg < G; ++gg) s += partN[gg];
        const double alpha = (MODE != MODE_GE) ? Rp[rpk(j, j)] : misc[0];
        const double nrm = sqrt(s + alpha * alpha);
        if (nrm == 0.0) { // H_j = I
            if (t == 0) tau[j] = 0.0;
            __syncthreads();
            continue;
        }
        const double beta = alpha >= 0.0 ? -nrm : nrm;
        const double taur = (beta - alpha) / beta;
        const double scal = 1.0 / (alpha - beta);
        if (c == j) {
#pragma unroll
            for (int r = 0; r < RP; ++r) {
                const int row = g * RP + r;
                double y;
                if (MODE != MODE_GE) {
                    y = a[r] * scal;
                    a[r] = y;
                } else if (row > j) {
                    y = a[r] * scal;
                    a[r] = y;
                } else if (row == j) {
                    y = 1.0;
                    a[r] = beta;
                } else {
                    y = 0.0;
                }
                ybuf[row] = y;
            }
        }
        if (t == 0) { tau[j] = taur; tauL[j] = taur; }
        __syncthreads();
        // (3) dot products
        double rjc = 0.0;
        if (c != j) {
            double d = 0.0;
#pragma unroll
            for (int r = 0; r < RP; ++r) d += ybuf[g * RP + r] * a[r];
            partD[g * B + c] = d;
            if (MODE != MODE_GE && c > j) rjc = Rp[rpk(j, c)];
        }
        __syncthreads();
        // (4) rank-1 update / Gram entries
        if (c > j) {
            double w = rjc;
#pragma unroll
            for (int gg = 0; gg < G; ++gg) w += partD[gg * B + c];
            const double tw = taur * w;
#pragma unroll
            for (int r = 0; r < RP; ++r) a[r] -= tw * ybuf[g * RP + r];
            if (MODE != MODE_GE && g == 0) Rp[rpk(j, c)] = rjc - tw;
        } else if (c < j && g == 0 && c >= (j & ~15)) {
            double d = 0.0;
#pragma unroll
            for (int gg = 0; gg < G; ++gg) d += partD[gg * B + c];
            gramL[j * 16 + (c & 15)] = d;
        }
        if (MODE != MODE_GE && c == j && g == 0) Rp[rpk(j, j)] = beta;
    }
    if (MODE == MODE_GE && t == 0) tau[B - 1] = 0.0;
    __syncthreads();

    // write back
#pragma unroll
    for (int r = 0; r < RP; ++r) {
        const int row = g * RP + r;
        if (MODE != MODE_TT || row <= c) src_c[(long)r * rs] = (TS)a[r];
    }
    if (MODE != MODE_GE) {
        for (int idx = t; idx < B * B; idx += NTHREADS) {
            int r, cc;
            if (tr) { cc = idx % B; r = idx / B; } else { r = idx % B; cc = idx / B; }
            if (r <= cc) top[toff(r, cc, ldt, tr)] = (TS)Rp[rpk(r, cc)];
        }
    }

    // diagonal blocks of the compact-WY factor: T[i][j] = -tau_j sum_{m=i}^{j-1} T[i][m] G[m][j]
    if (tjj) {
        for (int jl = 0; jl < 16; ++jl) {
            if (t < B) {
                const int J = t >> 4, il = t & 15;
                const int j = J * 16 + jl;
                double v;
                if (il == jl) {
                    v = tauL[j];
                } else if (il < jl) {
                    double acc = 0.0;
                    for (int m = il; m < jl; ++m) acc += TJ[(J * 16 + il) * 16 + m] * gramL[j * 16 + m];
                    v = -tauL[j] * acc;
                } else {
                    v = 0.0;
                }
                TJ[(J * 16 + il) * 16 + jl] = v;
            }
            __syncthreads();
        }
        for (int idx = t; idx < 16 * B; idx += NTHREADS) tjj[idx] = TJ[idx];
    }
}

// ---------------------------------------------------------------------------
// T_JJ from stored reflectors and tau (used by the tile operator API, whose
// callers pass V and tau only, like the reference kernels).
// ---------------------------------------------------------------------------
template <typename TS, int B, int MODE>
__device__ void tjj_body(const TS *__restrict__ V, long ldv, bool vtr, const double *__restrict__ tau,
                         double *__restrict__ tjj, double *smem)
{
    const int t = threadIdx.x;
    double *GL = smem;              // 16*B
    double *TL = GL + 16 * B;       // B (tau)
    double *TJ = TL + B;            // 16*B
    auto vget = [&](int k, int i) -> double {
        double v = (double)V[toff(k, i, ldv, vtr)];
        if (MODE == MODE_GE) v = (k < i) ? 0.0 : (k == i ? 1.0 : v);
        if (MODE == MODE_TT && k > i) v = 0.0;
        return v;
    };
    for (int idx = t; idx < B; idx += NTHREADS) TL[idx] = tau[idx];
    for (int idx = t; idx < 16 * B; idx += NTHREADS) {
        const int j = idx / 16, ml = idx % 16;
        const int m = (j & ~15) + ml;
        double d = 0.0;
        for (int k = 0; k < B; ++k) d += vget(k, m) * vget(k, j);
        GL[j * 16 + ml] = d;
    }
    __syncthreads();
    for (int jl = 0; jl < 16; ++jl) {
        if (t < B) {
            const int J = t >> 4, il = t & 15;
            const int j = J * 16 + jl;
            double v;
            if (il == jl) {
                v = TL[j];
            } else if (il < jl) {
                double acc = 0.0;
                for (int m = il; m < jl; ++m) acc += TJ[(J * 16 + il) * 16 + m] * GL[j * 16 + m];
                v = -TL[j] * acc;
            } else {
                v = 0.0;
            }
            TJ[(J * 16 + il) * 16 + jl] = v;
        }
        __syncthreads();
    }
    for (int idx = t; idx < 16 * B; idx += NTHREADS) tjj[idx] = TJ[idx];
}

// ---------------------------------------------------------------------------
// Update: apply the B reflectors of one panel to ntiles column tiles.
//   GE: C <- Q^T C,            Q = I - V T V^T, V unit lower (UNMQR)
//   TS/TT: [C1; C2] <- Q^T [C1; C2], Q = I - [I; Y] T [I; Y]^T (TSMQR)
// (Q instead of Q^T when !applyt: blocks in reverse order with T_JJ.)
// Each wave owns 16-column strips of C; per strip it keeps C2 (and C1) as
// B/16 MFMA accumulator blocks and, for every 16-reflector block J,
//   W  = C1[J] + Y_J^T C2      (B/16 x 4 MFMA, K order chosen so the
//                               accumulators feed the B operand directly)
//   W' = T_JJ^T W              (4 MFMA)
//   C1[J] -= W';  C2 -= Y_J W' (B/16 x 4 MFMA)
// Operand layout of the 16x16x4 MFMAs: A[i][k] from lane i+16k, B[k][j] from
// lane j+16k.  Accumulator element m of lane l holds column l%16 and row
// 4*(l/16)+m for f32 but row (l/16)+4*m for f64 (measured on gfx950); AccRow
// gives row = base*(l/16) + step*m.  The K index of a product only has to agree
// between the A and B operands, so every K loop below walks C rows in the
// accumulator's own order.
// C1 tile k sits at C1 + k*s1 (elements), C2 tile k at C2 + k*s2.
// ---------------------------------------------------------------------------
template <typename V4> __device__ __forceinline__ V4 vzero4()
{
    V4 v;
    v[0] = 0; v[1] = 0; v[2] = 0; v[3] = 0;
    return v;
}

template <typename TS> struct AccRow;
template <> struct AccRow<float> { static constexpr int base = 4, step = 1; };
template <> struct AccRow<double> { static constexpr int base = 1, step = 4; };

template <typename TS, int B, int MODE, bool APPLYT>
__device__ void update_body(const TS *__restrict__ V, long ldv, bool vtr, const double *__restrict__ tjj,
                            TS *__restrict__ C1, long ld1, long s1, TS *__restrict__ C2, long ld2, long s2,
                            bool ctr, int ntiles, double *smem)
{
    typedef typename AccT<TS>::type acc_t;
    constexpr int S = B + 1;
    constexpr int NB = B / 16;
    TS *YL = reinterpret_cast<TS *>(smem); // YL[k*S + i] = Y[k][i]
    TS *TL = YL + B * S;                   // TL[i*16 + jl] = T[i][16*(i/16)+jl]
    const int t = threadIdx.x;
    for (int idx = t; idx < B * B; idx += NTHREADS) {
        int k, i;
        if (vtr) { i = idx % B; k = idx / B; } else { k = idx % B; i = idx / B; }
        TS v = V[toff(k, i, ldv, vtr)];
        if (MODE == MODE_GE) v = (k < i) ? (TS)0 : (k == i ? (TS)1 : v);
        if (MODE == MODE_TT && k > i) v = (TS)0;
        YL[k * S + i] = v;
    }
    for (int idx = t; idx < 16 * B; idx += NTHREADS) TL[idx] = (TS)tjj[idx];
    __syncthreads();

    const int lane = t & 63, w = t >> 6, q = lane >> 4, il = lane & 15;
    constexpr int RB = AccRow<TS>::base, RS = AccRow<TS>::step;
    // logical element (row, col) of a C tile is at base + row*rs + col*cs
    const long rs1 = ctr ? ld1 : 1, cs1 = ctr ? 1 : ld1;
    const long rs2 = ctr ? ld2 : 1, cs2 = ctr ? 1 : ld2;
    const int nstrips = ntiles * NB;
    for (int sidx = w; sidx < nstrips; sidx += NWAVES) {
        const int tile = sidx / NB;
        const int col = (sidx % NB) * 16 + il;
        // this lane's column of each tile, starting at its first accumulator row
        TS *c1 = C1 + (long)tile * s1 + (long)col * cs1 + (long)(RB * q) * rs1;
        TS *c2 = C2 + (long)tile * s2 + (long)col * cs2 + (long)(RB * q) * rs2;
        acc_t acc2[NB];
#pragma unroll
        for (int tb = 0; tb < NB; ++tb) {
#pragma unroll
            for (int m = 0; m < 4; ++m) acc2[tb][m] = c2[(long)(16 * tb + RS * m) * rs2];
        }
        acc_t c1n;
        if (MODE != MODE_GE) {
            const int J0 = APPLYT ? 0 : NB - 1;
#pragma unroll
            for (int m = 0; m < 4; ++m) c1n[m] = c1[(long)(16 * J0 + RS * m) * rs1];
        }
#pragma unroll 1
        for (int jj = 0; jj < NB; ++jj) {
            const int J = APPLYT ? jj : (NB - 1 - jj);
            acc_t c1v, Wa;
            if (MODE != MODE_GE) {
                c1v = c1n;
                if (jj + 1 < NB) {
                    const int Jn = APPLYT ? jj + 1 : (NB - 2 - jj);
#pragma unroll
                    for (int m = 0; m < 4; ++m) c1n[m] = c1[(long)(16 * Jn + RS * m) * rs1];
                }
                Wa = c1v;
            } else {
                Wa = vzero4<acc_t>();
            }
            // W = C1[J] + Y_J^T C2   (K order: lane group q supplies rows 16tb+RB*q+RS*m)
#pragma unroll
            for (int tb = 0; tb < NB; ++tb) {
                if (MODE == MODE_GE && tb < J) continue;
                if (MODE == MODE_TT && tb > J) continue;
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const TS av = YL[(16 * tb + RB * q + RS * m) * S + 16 * J + il];
                    Wa = mfma16(av, acc2[tb][m], Wa);
                }
            }
            // W' = T_JJ^T W (Q^T) or T_JJ W (Q)
            acc_t Wp = vzero4<acc_t>();
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const TS av = APPLYT ? TL[(16 * J + RB * q + RS * m) * 16 + il] : TL[(16 * J + il) * 16 + RB * q + RS * m];
                Wp = mfma16(av, Wa[m], Wp);
            }
            if (MODE != MODE_GE) {
#pragma unroll
                for (int m = 0; m < 4; ++m) c1[(long)(16 * J + RS * m) * rs1] = c1v[m] - Wp[m];
            }
            // C2 -= Y_J W'
#pragma unroll
            for (int tb = 0; tb < NB; ++tb) {
                if (MODE == MODE_GE && tb < J) continue;
                if (MODE == MODE_TT && tb > J) continue;
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const TS av = -YL[(16 * tb + il) * S + 16 * J + RB * q + RS * m];
                    acc2[tb] = mfma16(av, Wp[m], acc2[tb]);
                }
            }
        }
#pragma unroll
        for (int tb = 0; tb < NB; ++tb) {
#pragma unroll
            for (int m = 0; m < 4; ++m) c2[(long)(16 * tb + RS * m) * rs2] = acc2[tb][m];
        }
    }
}


// ---------------------------------------------------------------------------
// Split-BF16 update for FP32 storage.  Every FP32 operand x of the two large
// products is split as x = hi + lo with hi = bf16(x) and lo = bf16(x - hi)
// (round to nearest even), and a*b is taken as hi*hi + hi*lo + lo*hi on the
// bf16 matrix cores with FP32 accumulation (the lo*lo term, ~2^-16 relative,
// is dropped).  Same blocking and K ordering as update_body:
//   W  = C1[J] + Y_J^T C2   v_mfma_f32_16x16x32_bf16, K chunk p covers C2 rows
//                           32p + 16(e/4) + 4q + e%4 for lane group q, e = 0..7
//   W' = T_JJ^T W           native FP32 (v_mfma_f32_16x16x4f32)
//   C2 -= Y_J W'            v_mfma_f32_16x16x16bf16_1k, K = 16 reflectors
// Y is staged as bf16 hi/lo pairs twice: row-major for the C2 product and
// transposed for the W product, so every A operand is two 8-byte LDS reads.
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned short f2bf(float x)
{
    unsigned u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }
__device__ __forceinline__ void split_bf(float x, unsigned short &h, unsigned short &l)
{
    h = f2bf(x);
    l = f2bf(x - bf2f(h));
}

template <int B, int MODE, bool APPLYT>
__device__ void update_body_split(const float *__restrict__ V, long ldv, bool vtr, const double *__restrict__ tjj,
                                  float *__restrict__ C1, long ld1, long s1, float *__restrict__ C2, long ld2, long s2,
                                  bool ctr, int ntiles, double *smem)
{
    typedef f4_t acc_t;
    constexpr int S = B + 8;
    constexpr int NB = B / 16;
    unsigned short *YTh = reinterpret_cast<unsigned short *>(smem); // YT[i*S + k] = Y[k][i]
    unsigned short *YTl = YTh + B * S;
    unsigned short *Yh = YTl + B * S; // Y[k*S + i] = Y[k][i]
    unsigned short *Yl = Yh + B * S;
    float *TL = reinterpret_cast<float *>(Yl + B * S);
    const int t = threadIdx.x;
    for (int idx = t; idx < B * B; idx += NTHREADS) {
        int k, i;
        if (vtr) { i = idx % B; k = idx / B; } else { k = idx % B; i = idx / B; }
        float v = V[toff(k, i, ldv, vtr)];
        if (MODE == MODE_GE) v = (k < i) ? 0.f : (k == i ? 1.f : v);
        if (MODE == MODE_TT && k > i) v = 0.f;
        unsigned short h, l;
        split_bf(v, h, l);
        YTh[i * S + k] = h;
        YTl[i * S + k] = l;
        Yh[k * S + i] = h;
        Yl[k * S + i] = l;
    }
    for (int idx = t; idx < 16 * B; idx += NTHREADS) TL[idx] = (float)tjj[idx];
    __syncthreads();

    const int lane = t & 63, w = t >> 6, q = lane >> 4, il = lane & 15;
    const long rs1 = ctr ? ld1 : 1, cs1 = ctr ? 1 : ld1;
    const long rs2 = ctr ? ld2 : 1, cs2 = ctr ? 1 : ld2;
    const int nstrips = ntiles * NB;
    for (int sidx = w; sidx < nstrips; sidx += NWAVES) {
        const int tile = sidx / NB;
        const int col = (sidx % NB) * 16 + il;
        float *c1 = C1 + (long)tile * s1 + (long)col * cs1 + (long)(4 * q) * rs1;
        float *c2 = C2 + (long)tile * s2 + (long)col * cs2 + (long)(4 * q) * rs2;
        acc_t acc2[NB];
#pragma unroll
        for (int tb = 0; tb < NB; ++tb) {
#pragma unroll
            for (int m = 0; m < 4; ++m) acc2[tb][m] = c2[(long)(16 * tb + m) * rs2];
        }
#pragma unroll 1
        for (int jj = 0; jj < NB; ++jj) {
            const int J = APPLYT ? jj : (NB - 1 - jj);
            acc_t c1v, Wa;
            if (MODE != MODE_GE) {
#pragma unroll
                for (int m = 0; m < 4; ++m) c1v[m] = c1[(long)(16 * J + m) * rs1];
                Wa = c1v;
            } else {
                Wa = vzero4<acc_t>();
            }
#pragma unroll
            for (int p = 0; p < NB / 2; ++p) {
                if (MODE == MODE_GE && 2 * p + 1 < J) continue;
                if (MODE == MODE_TT && 2 * p > J) continue;
                const unsigned short *th = YTh + (16 * J + il) * S + 32 * p + 4 * q;
                const unsigned short *tl = YTl + (16 * J + il) * S + 32 * p + 4 * q;
                u8_t ah, al, bh, bl;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    ah[e] = th[e];
                    ah[e + 4] = th[16 + e];
                    al[e] = tl[e];
                    al[e + 4] = tl[16 + e];
                    unsigned short h, l;
                    split_bf(acc2[2 * p][e], h, l);
                    bh[e] = h;
                    bl[e] = l;
                    split_bf(acc2[2 * p + 1][e], h, l);
                    bh[e + 4] = h;
                    bl[e + 4] = l;
                }
                const bf8_t Ah = __builtin_bit_cast(bf8_t, ah), Al = __builtin_bit_cast(bf8_t, al);
                const bf8_t Bh = __builtin_bit_cast(bf8_t, bh), Bl = __builtin_bit_cast(bf8_t, bl);
                Wa = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah, Bh, Wa, 0, 0, 0);
                Wa = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah, Bl, Wa, 0, 0, 0);
                Wa = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Al, Bh, Wa, 0, 0, 0);
            }
            acc_t Wp = vzero4<acc_t>();
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const float av = APPLYT ? TL[(16 * J + 4 * q + m) * 16 + il] : TL[(16 * J + il) * 16 + 4 * q + m];
                Wp = mfma16(av, Wa[m], Wp);
            }
            if (MODE != MODE_GE) {
#pragma unroll
                for (int m = 0; m < 4; ++m) c1[(long)(16 * J + m) * rs1] = c1v[m] - Wp[m];
            }
            s4_t wh, wl;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                unsigned short h, l;
                split_bf(-Wp[e], h, l);
                wh[e] = (short)h;
                wl[e] = (short)l;
            }
#pragma unroll
            for (int tb = 0; tb < NB; ++tb) {
                if (MODE == MODE_GE && tb < J) continue;
                if (MODE == MODE_TT && tb > J) continue;
                const unsigned short *yh = Yh + (16 * tb + il) * S + 16 * J + 4 * q;
                const unsigned short *yl = Yl + (16 * tb + il) * S + 16 * J + 4 * q;
                s4_t ah, al;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    ah[e] = (short)yh[e];
                    al[e] = (short)yl[e];
                }
                acc2[tb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, wh, acc2[tb], 0, 0, 0);
                acc2[tb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, wl, acc2[tb], 0, 0, 0);
                acc2[tb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, wh, acc2[tb], 0, 0, 0);
            }
        }
#pragma unroll
        for (int tb = 0; tb < NB; ++tb) {
#pragma unroll
            for (int m = 0; m < 4; ++m) c2[(long)(16 * tb + m) * rs2] = acc2[tb][m];
        }
    }
}

// update with the arithmetic selected at run time: split-BF16 (FP32 storage
// only) or the native MFMA of the storage type
template <typename TS, int B, int MODE>
__device__ __forceinline__ void update_sel(bool split, const TS *V, long ldv, bool vtr, const double *tjj, TS *C1,
                                           long ld1, long s1, TS *C2, long ld2, long s2, bool ctr, int ntiles,
                                           double *smem)
{
    if constexpr (sizeof(TS) == 4) {
        if (split) {
            update_body_split<B, MODE, true>((const float *)V, ldv, vtr, tjj, (float *)C1, ld1, s1, (float *)C2, ld2,
                                             s2, ctr, ntiles, smem);
            return;
        }
    }
    update_body<TS, B, MODE, true>(V, ldv, vtr, tjj, C1, ld1, s1, C2, ld2, s2, ctr, ntiles, smem);
}

} // namespace avs
