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

// Timing hook inside the panel loop (k = stage, j = reflector, pub = this lane
// publishes the next column); empty unless a profiling build defines it.
#ifndef AVS_PANEL_MARK
#define AVS_PANEL_MARK(k, j, pub)
#endif

// element (r, c) of a tile with leading dimension ld, or of its adjoint
__device__ __forceinline__ long toff(int r, int c, long ld, bool tr)
{
    return tr ? (long)c + (long)r * ld : (long)r + (long)c * ld;
}

// Panel thread geometry for NT threads: the NT/64 waves split the B columns
// evenly; the LPC lanes of one column (LPC consecutive lanes of its wave)
// hold RPL consecutive rows each, so every column-wise reduction is a DPP
// butterfly inside a row of lanes.
template <int B, int NT = NTHREADS> struct PanelGeo {
    static constexpr int CPW = B / (NT / 64); // columns per wave
    static constexpr int LPC = 64 / CPW;      // lanes per column
    static constexpr int RPL = B / LPC;       // rows per lane
    static constexpr int XS = B + 2 * LPC;    // padded broadcast vector (2 doubles after each lane's rows)
};
// threads of the standalone / chained panel launches
template <int B> struct PanelThreads { static constexpr int value = B >= 128 ? 1024 : NTHREADS; };

// LDS footprint (in doubles) of each body; the executor allocates the max.
template <typename TS, int B> struct Lds {
    // packed R, broadcast vectors and their norm partials (sized for the widest
    // geometry), Gram rows, tau, beta
    static constexpr int panel = B * (B + 1) / 2 + 2 * (B + 128) + 2 * 64 + 16 * B + 2 * B + 2 * B;
    static constexpr int update = (B * (B + 1) * (int)sizeof(TS) + 16 * B * (int)sizeof(TS) + 7) / 8;
    static constexpr int tjj = 16 * B + B + 16 * B;
    // split-BF16 update (FP32 storage): Y and Y^T as bf16 hi/lo pairs + T_JJ in FP32
    static constexpr int SPLIT_S = B + 8;
    static constexpr int update_split = sizeof(TS) == 4 ? (4 * B * SPLIT_S * 2 + 16 * B * 4 + 7) / 8 : 0;
    static constexpr int upd = update > update_split ? update : update_split;
    static constexpr int maxv = panel > upd ? panel : upd; // bodies run by the persistent executor
};

// Lane exchanges v -> v[lane ^ m] on the VALU (no LDS round trip):
//   ^4  : row_half_mirror (lane ^ 7 within 8) then quad_perm [3,2,1,0] (^ 3)
//   ^8  : row_ror:8 inside each row of 16 lanes
//   ^16 : v_permlane16_swap (odd rows of the first operand <-> even rows of the second)
//   ^32 : v_permlane32_swap (upper half of the first operand <-> lower half of the second)
// (every control used here reads a lane inside the wave's row, so no lane
// takes the bound_ctrl zero)
template <int CTRL> __device__ __forceinline__ double dpp_mov(double v)
{
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}
template <int M> __device__ __forceinline__ double lane_xor(double v)
{
    static_assert(M == 4 || M == 8 || M == 16 || M == 32, "lane_xor: 4, 8, 16 or 32");
    if constexpr (M == 4) {
        return dpp_mov<0x1B>(dpp_mov<0x141>(v));
    } else if constexpr (M == 8) {
        return dpp_mov<0x128>(v);
    } else {
        const unsigned lo = (unsigned)__double2loint(v), hi = (unsigned)__double2hiint(v);
        bool upper;
        unsigned nlo, nhi;
        if constexpr (M == 16) {
            const auto rl = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
            const auto rh = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
            upper = (__lane_id() >> 4) & 1; // odd row: partner is the row below
            nlo = upper ? rl[0] : rl[1];
            nhi = upper ? rh[0] : rh[1];
        } else {
            const auto rl = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
            const auto rh = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
            upper = __lane_id() >= 32;
            nlo = upper ? rl[0] : rl[1];
            nhi = upper ? rh[0] : rh[1];
        }
        return __hiloint2double((int)nhi, (int)nlo);
    }
}

// sum over the lanes of one panel column (lanes l, l ^ CPW, l ^ 2CPW, ...),
// identical in every one of them: each butterfly step adds the same two
// values in either order
template <int CPW> __device__ __forceinline__ double colsum(double v)
{
    if constexpr (CPW <= 4) v += lane_xor<4>(v);
    if constexpr (CPW <= 8) v += lane_xor<8>(v);
    if constexpr (CPW <= 16) v += lane_xor<16>(v);
    v += lane_xor<32>(v);
    return v;
}

// sum over a group of LPC consecutive lanes (LPC = 4, 8 or 16), identical in
// every lane of the group, with DPP moves only: quad_perm [1,0,3,2] (lane ^ 1),
// quad_perm [2,3,0,1] (lane ^ 2), then row_half_mirror (pairs the two quads
// of 8 lanes once each quad holds its sum) and row_mirror (the two halves of
// 16).  Each step adds the same two values in either order.
template <int LPC> __device__ __forceinline__ double grpsum(double v)
{
    static_assert(LPC == 4 || LPC == 8 || LPC == 16, "grpsum: 4, 8 or 16 lanes");
    v += dpp_mov<0xB1>(v);
    v += dpp_mov<0x4E>(v);
    if constexpr (LPC >= 8) v += dpp_mov<0x141>(v);
    if constexpr (LPC >= 16) v += dpp_mov<0x140>(v);
    return v;
}

// sqrt(x) and 1/sqrt(x) for x > 0 finite: the hardware reciprocal square
// root refined by two coupled Newton (Goldschmidt) steps; 1/x from the
// hardware reciprocal refined by two Newton steps.  Accurate to a few ulp and
// identical in every lane that runs them on the same input.
__device__ __forceinline__ void sqrt_rsqrt(double x, double &sq, double &rs)
{
    const double r0 = __builtin_amdgcn_rsq(x);
    double y = x * r0, h = 0.5 * r0;
    double e = fma(-y, h, 0.5);
    y = fma(y, e, y);
    h = fma(h, e, h);
    e = fma(-y, h, 0.5);
    y = fma(y, e, y);
    h = fma(h, e, h);
    sq = y;
    rs = 2.0 * h;
}
__device__ __forceinline__ double rcp_nr(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    r = fma(r, fma(-x, r, 1.0), r);
    return r;
}
// pairwise sum of N values (log2 N dependent adds)
template <int N> __device__ __forceinline__ double tree_sum(double (&v)[N])
{
#pragma unroll
    for (int h = 1; h < N; h <<= 1) {
#pragma unroll
        for (int u = 0; u + h < N; u += 2 * h) v[u] += v[u + h];
    }
    return v[0];
}

// a[s] for a wave-uniform s: a register-indexed move when the array is small
// enough to stay in registers, a select chain otherwise
template <int R> __device__ __forceinline__ double reg_get(const double (&a)[R], int s)
{
    if constexpr (R <= 16) {
        // the slot in a scalar register: VGPR-indexed moves instead of a
        // select over all R registers
        return a[__builtin_amdgcn_readfirstlane(s)];
    } else {
        double v = 0.0;
#pragma unroll
        for (int u = 0; u < R; ++u) v = (u == s) ? a[u] : v;
        return v;
    }
}
template <int R> __device__ __forceinline__ void reg_set(double (&a)[R], int s, double d)
{
    if constexpr (R <= 16) {
        a[__builtin_amdgcn_readfirstlane(s)] = d;
    } else {
#pragma unroll
        for (int u = 0; u < R; ++u) a[u] = (u == s) ? d : a[u];
    }
}
template <int R> __device__ __forceinline__ void reg_add(double (&a)[R], int s, double d)
{
    if constexpr (R <= 16) {
        a[s] += d;
    } else {
#pragma unroll
        for (int u = 0; u < R; ++u)
            if (u == s) a[u] += d;
    }
}

// ---------------------------------------------------------------------------
// Panel: Householder QR of one tile (GE), of [triu(top); bot] (TS) or of
// [triu(top); triu(bot)] (TT), computed in FP64 with one workgroup barrier
// per reflector.
// Reflector j: the lanes owning column j publish it in a double-buffered LDS
// vector x, already restricted to the rows the reflector touches (GE: rows
// >= j; TT: rows <= j by the triangle), with the partial sums of x^2 of each
// lane group.  After the barrier every column group forms S_c = x . a_c over
// its rows, sums it over its lanes, and derives
//   beta = -sgn(alpha) ||(alpha, x)||,  tau = (beta - alpha) / beta,
//   y = (x - beta e_j) / (alpha - beta)              (GE; y_j = 1)
//   y = x / (alpha - beta)                           (TS, TT; 1 in R's row j)
//   w_c = y . a_c = (S_c - beta a_c[j]) / (alpha - beta)   (GE)
//   w_c = R[j][c] + S_c / (alpha - beta)                    (TS, TT)
// then a_c -= tau w_c y (and R[j][c] -= tau w_c).  Columns c < j of j's
// 16-block hold finished reflectors; the same formula gives the Gram entries
// y_c . y_j of the compact-WY factor.  Waves whose columns all lie left of j's
// 16-block only take part in the barrier.  An all-zero column gives tau = 0.
// Outputs: reflectors in place (GE: strict lower part, unit diagonal
// implicit; TS: the whole bottom tile; TT: the upper triangle of bot), R in
// the upper triangle of top, tau[B] (FP64) and tjj[B*16] (FP64, row i holds
// T[i][16*(i/16) + 0..15], the diagonal block of the compact-WY factor).
// ---------------------------------------------------------------------------
template <typename TS, int B, int MODE, int NT = NTHREADS>
__device__ __forceinline__ void panel_body(TS *__restrict__ top, long ldt, TS *__restrict__ bot, long ldb, bool tr,
                                           double *__restrict__ tau, double *__restrict__ tjj, double *smem)
{
    typedef PanelGeo<B, NT> PG;
    constexpr int CPW = PG::CPW, LPC = PG::LPC, RPL = PG::RPL, XS = PG::XS;
    static_assert(XS <= B + 128 && LPC <= 64 && CPW >= 4, "panel geometry");
    // The thread index passes through an empty asm so that, where the body
    // sits in a loop over a chain of tiles, nothing derived from it is
    // hoisted out of the loop and held in registers across the whole body.
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    const int l = t & 63, w = t >> 6;
#ifdef AVS_PANEL_STRIDED
    const int c = w * CPW + (l % CPW), q = l / CPW, r0 = q * RPL;
#else
    const int c = w * CPW + l / LPC, q = l % LPC, r0 = q * RPL;
#endif

    double *Rp = smem;                    // packed upper triangle of R (TS/TT); diagonal read as alpha only
    double *xbuf = Rp + B * (B + 1) / 2;  // 2 x XS: published column
    double *nbuf = xbuf + 2 * XS;         // 2 x LPC: its partial sums of squares, per lane group
    double *gramL = nbuf + 2 * LPC;       // gramL[j*16 + (c&15)] = y_c . y_j
    double *tauL = gramL + 16 * B;        // B
    double *dgL = tauL + B;               // B: R diagonal (beta_j) as produced
    double *rowb = dgL + B;               // 2 x B (GE): row j of every column before reflector j
    auto rpk = [](int r, int cc) { return r * B - (r * (r - 1)) / 2 + (cc - r); };
    auto xi = [](int row) { return row + (row / RPL) * 2; };

    TS *src = (MODE == MODE_GE) ? top : bot;
    const long lds = (MODE == MODE_GE) ? ldt : ldb;
    const long rs = tr ? lds : 1;
    TS *src_c = src + toff(r0, c, lds, tr);
    double a[RPL];
#pragma unroll
    for (int s = 0; s < RPL; ++s) {
        const int row = r0 + s;
        a[s] = (MODE == MODE_TT && row > c) ? 0.0 : (double)src_c[(long)s * rs];
    }
    if (MODE != MODE_GE) {
        // R from this lane's rows of column c of the top tile, loaded like the
        // column itself: every load of the lane is in flight at once (the
        // whole column is read; only its upper part is kept)
        const long rt = tr ? ldt : 1;
        TS *top_c = top + toff(r0, c, ldt, tr);
        double rv[RPL];
#pragma unroll
        for (int s = 0; s < RPL; ++s) rv[s] = (double)top_c[(long)s * rt];
#pragma unroll
        for (int s = 0; s < RPL; ++s) {
            const int row = r0 + s;
            if (row <= c) Rp[rpk(row, c)] = rv[s];
            if (row == c) dgL[c] = rv[s];
        }
    }
    for (int idx = t; idx < 16 * B; idx += NT) gramL[idx] = 0.0;
    for (int idx = t; idx < B; idx += NT) tauL[idx] = 0.0;

    // lanes of column p publish it for reflector p
    auto publish = [&](int p) {
        double *xn = xbuf + (p & 1) * XS;
        double sq[RPL];
#pragma unroll
        for (int s = 0; s < RPL; ++s) {
            // GE: the rows above p are zero in registers by now (each is zeroed when it
            // becomes final R, which is kept in LDS)
            const double v = a[s];
            xn[xi(r0 + s)] = v;
            sq[s] = v * v;
        }
        nbuf[(p & 1) * LPC + q] = tree_sum(sq);
    };
    if (c == 0) publish(0);
    // GE: the lane holding row p of column c hands a_c[p] to reflector p
    auto pass_row = [&](int p) {
        if (q == p / RPL) rowb[(p & 1) * B + c] = reg_get(a, p % RPL);
    };
    if (MODE == MODE_GE) pass_row(0);

    const int jend = (MODE == MODE_GE) ? B - 1 : B;
    for (int jv = 0; jv < jend; ++jv) {
        __syncthreads(); // column j published, previous reflector's writes visible
        // the reflector index in a scalar register: every index derived from
        // it (packed R, broadcast buffers, Gram rows) is formed on the SALU
        const int j = __builtin_amdgcn_readfirstlane(jv);
        const int j16 = j & ~15;
        if ((w + 1) * CPW <= j16) continue; // every column of this wave is finished
#ifdef AVS_PANEL_PRIO
        // the wave that publishes the next pivot column issues first
        if (__builtin_amdgcn_readfirstlane(w) == __builtin_amdgcn_readfirstlane((j + 1) / CPW))
            __builtin_amdgcn_s_setprio(3);
        else __builtin_amdgcn_s_setprio(0);
#endif
        AVS_PANEL_MARK(0, j, c == j + 1 && q == 0);
        const double *xb = xbuf + (j & 1) * XS;
        const double *nb = nbuf + (j & 1) * LPC;
        // everything this reflector reads from LDS is requested up front
        const double alpha = (MODE == MODE_GE) ? xb[xi(j)] : Rp[rpk(j, j)];
        double rjc = 0.0; // R[j][c]; the packed index stays inside R for every c < B
        if (MODE != MODE_GE) {
            const double v = Rp[rpk(j, c)];
            rjc = c > j ? v : 0.0;
        }
        double nv[LPC];
#pragma unroll
        for (int u = 0; u < LPC; ++u) nv[u] = nb[u];
        // GE: a_c[j], the pivot-row entry of this column (R[j][c] once reflector j is applied)
        const double pj = (MODE == MODE_GE) ? rowb[(j & 1) * B + c] : 0.0;
        double xr[RPL], Sp[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s = 0; s < RPL; ++s) {
            xr[s] = xb[xi(r0 + s)];
            Sp[s & 3] = fma(xr[s], a[s], Sp[s & 3]);
        }
        // Householder scalars from ||x||^2 and alpha alone, overlapping the
        // reduction of S.  With rn = 1/nrm:  beta = -sgn(alpha) nrm,
        // tau = (beta - alpha)/beta = 1 + |alpha| rn,
        // scal = 1/(alpha - beta) = sgn(alpha) rn / tau,  tau scal^2 = rn^2 / tau.
        const double nn = tree_sum(nv);
        const double n2 = (MODE == MODE_GE) ? nn : fma(alpha, alpha, nn);
        AVS_PANEL_MARK(1, j, c == j + 1 && q == 0 && n2 == n2);
        double nrm = 0.0, rn = 0.0;
        if (n2 != 0.0) sqrt_rsqrt(n2, nrm, rn);
        const double sg = alpha >= 0.0 ? 1.0 : -1.0;
        const double beta = -sg * nrm;
        const double taur = fma(fabs(alpha), rn, 1.0);
        const double itau = rcp_nr(taur);
        const double srn = sg * rn;
        const double scal = srn * itau;
        const double c2 = rn * rn * itau;
        // S_c = x . a_c
#ifdef AVS_PANEL_STRIDED
        const double S = colsum<CPW>((Sp[0] + Sp[1]) + (Sp[2] + Sp[3]));
#else
        const double S = grpsum<LPC>((Sp[0] + Sp[1]) + (Sp[2] + Sp[3]));
#endif
        AVS_PANEL_MARK(2, j, c == j + 1 && q == 0 && S == S && c2 == c2);
        if (n2 != 0.0) {
            // GE: w_c = scal (S - beta a_c[j]);  TS/TT: w_c = R[j][c] + scal S
            // a_c -= tau w_c y  ==  a_c -= (tau w_c scal) x on every row but GE's row j
            const double sgp = (MODE == MODE_GE) ? fma(-beta, pj, S) : S;
            if (c > j) {
                const double tws = (MODE == MODE_GE) ? c2 * sgp : fma(c2, S, srn * rjc);
#pragma unroll
                for (int s = 0; s < RPL; ++s) a[s] = fma(-tws, xr[s], a[s]);
                // row j is final now: R[j][c] = a_c[j] - tau w_c (GE: y_j = 1), kept in LDS;
                // GE zeroes its register copy, so later publishes need no mask
                if (q == 0) Rp[rpk(j, c)] = (MODE == MODE_GE) ? pj - srn * sgp : rjc - fma(taur, rjc, srn * S);
                if (MODE == MODE_GE && q == j / RPL) reg_set(a, j % RPL, 0.0);
            } else if (c == j) {
                // the stored reflector (GE: rows <= j are written from LDS at the end)
#pragma unroll
                for (int s = 0; s < RPL; ++s) a[s] = xr[s] * scal;
                if (q == 0) {
                    tauL[j] = taur;
                    dgL[j] = beta;
                }
            } else if (c >= j16 && q == 0) {
                gramL[j * 16 + (c & 15)] = scal * sgp;
            }
        }
        if (MODE == MODE_GE && n2 == 0.0) { // H_j = I: row j stays as it is
            if (q == 0) {
                if (c > j) Rp[rpk(j, c)] = pj;
                else if (c == j) dgL[j] = pj;
            }
            if (c >= j && q == j / RPL) reg_set(a, j % RPL, 0.0);
        }
        AVS_PANEL_MARK(3, j, c == j + 1 && q == 0 && a[0] == a[0]);
        if (MODE == MODE_GE && j + 1 < jend) pass_row(j + 1);
        if (c == j + 1) publish(j + 1);
        AVS_PANEL_MARK(4, j, c == j + 1 && q == 0);
    }
    __syncthreads();

    // write back (tau from LDS: tauL holds 0 for every identity reflector and
    // for GE's last column).  The tile pointer goes through an empty asm so
    // the per-row addresses are formed again here instead of being kept live
    // in registers across the reflector loop.
    for (int idx = t; idx < B; idx += NT) tau[idx] = tauL[idx];
    TS *src_w = src;
    asm volatile("" : "+s"(src_w));
    // the lane's column and rows again, from a fresh thread index: nothing of
    // the prologue's addressing (the same expressions) is kept across the loop
    int te = threadIdx.x;
    asm volatile("" : "+v"(te));
#ifdef AVS_PANEL_STRIDED
    const int ce = (te >> 6) * CPW + ((te & 63) % CPW), r0e = ((te & 63) / CPW) * RPL;
#else
    const int ce = (te >> 6) * CPW + (te & 63) / LPC, r0e = ((te & 63) % LPC) * RPL;
#endif
    TS *dst_c = src_w + toff(r0e, ce, lds, tr);
    if (MODE == MODE_GE) {
        // every row of column c: R above the diagonal and the diagonal (up to
        // column B-2, which has no reflector) from LDS, the reflector below
#pragma unroll
        for (int s = 0; s < RPL; ++s) {
            const int row = r0e + s;
            if (row > ce || (row == ce && ce == B - 1)) dst_c[(long)s * rs] = (TS)a[s];
        }
#pragma unroll
        for (int s = 0; s < RPL; ++s) {
            const int row = r0e + s;
            if (row < ce || (row == ce && ce < B - 1)) dst_c[(long)s * rs] = (TS)(row < ce ? Rp[rpk(row, ce)] : dgL[ce]);
        }
    } else {
#pragma unroll
        for (int s = 0; s < RPL; ++s) {
            const int row = r0e + s;
            if (MODE != MODE_TT || row <= ce) dst_c[(long)s * rs] = (TS)a[s];
        }
        // R: rows <= c of column c of the top tile; its lower part stays as it was
        TS *top_w = top;
        asm volatile("" : "+s"(top_w));
        const long rt = tr ? ldt : 1;
        TS *top_c = top_w + toff(r0e, ce, ldt, tr);
#pragma unroll
        for (int s = 0; s < RPL; ++s) {
            const int row = r0e + s;
            if (row <= ce) top_c[(long)s * rt] = (TS)(row < ce ? Rp[rpk(row, ce)] : dgL[ce]);
        }
    }

    // diagonal blocks of the compact-WY factor, one row per lane:
    // T[i][j] = -tau_j sum_{m=i}^{j-1} T[i][m] G[m][j]  (i < j in one 16-block).
    // trow[m] = T[i][m] is exactly zero for m < i, so the sum runs over m < j.
    if (tjj && t < B) {
        const int J = t >> 4, il = t & 15;
        double trow[16];
#pragma unroll
        for (int jl = 0; jl < 16; ++jl) {
            const int j = J * 16 + jl;
            double acc = 0.0;
#pragma unroll
            for (int m = 0; m < jl; ++m) acc += trow[m] * gramL[j * 16 + m];
            const double tj = tauL[j];
            const double v = il == jl ? tj : (il < jl ? -tj * acc : 0.0);
            trow[jl] = v;
        }
#pragma unroll
        for (int jl = 0; jl < 16; ++jl) tjj[t * 16 + jl] = trow[jl];
    }
}

// ---------------------------------------------------------------------------
// T_JJ from stored reflectors and tau (used by the tile operator API, whose
// callers pass V and tau only, like the reference kernels).
// ---------------------------------------------------------------------------
template <typename TS, int B, int MODE>
__device__ __forceinline__ void tjj_body(const TS *__restrict__ V, long ldv, bool vtr, const double *__restrict__ tau,
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
__device__ __forceinline__ void update_body(const TS *__restrict__ V, long ldv, bool vtr, const double *__restrict__ tjj,
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

template <int B, int MODE, bool APPLYT, int NT = NTHREADS>
__device__ __forceinline__ void update_body_split(const float *__restrict__ V, long ldv, bool vtr, const double *__restrict__ tjj,
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
    for (int idx = t; idx < B * B; idx += NT) {
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
    for (int idx = t; idx < 16 * B; idx += NT) TL[idx] = (float)tjj[idx];
    __syncthreads();

    const int lane = t & 63, w = t >> 6, q = lane >> 4, il = lane & 15;
    const long rs1 = ctr ? ld1 : 1, cs1 = ctr ? 1 : ld1;
    const long rs2 = ctr ? ld2 : 1, cs2 = ctr ? 1 : ld2;
    const int nstrips = ntiles * NB;
    for (int sidx = w; sidx < nstrips; sidx += NT / 64) {
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

// ---------------------------------------------------------------------------
// Update, blocked-staging form (the stream-ordered executor and the operator
// API): the same compact-WY arithmetic as update_body, with
//  - only the 16 reflectors of block J in LDS (B x 16, double-buffered; the
//    next block's global loads are in flight while block J is computed), so a
//    workgroup needs ~2*B*17 + 16*B elements of LDS and two workgroups share a CU;
//  - each wave owning two 16-column strips and interleaving them, so every
//    A-operand read from LDS feeds two MFMAs and the two accumulation chains
//    are independent.
// Handles up to 2*NT/64 strips (ntiles * B/16 of them) per call.
// ---------------------------------------------------------------------------
constexpr int UPD_THREADS = 256;
// timing hook of update2_body (event ev at block jj); empty unless a probe
// build defines it
#ifndef AVS_UPD_MARK
#define AVS_UPD_MARK(ev, jj)
#endif
template <typename TS, int B> struct Lds2 {
    static constexpr int YS = 17; // row stride of a staged 16-reflector block
    static constexpr int upd = (2 * B * YS * (int)sizeof(TS) + 16 * B * (int)sizeof(TS) + 7) / 8;
};

// SPW = 1: a half-tile workgroup for launches too small to fill the GPU:
// wave w takes strip s0 + w alone (s0 = 0 or NB/2, ntiles = 1), so one column
// tile is spread over two workgroups and twice as many SIMDs.
template <typename TS, int B, bool APPLYT, int NT, bool CTR, int SPW = 2>
__device__ __forceinline__ void update2_body(int mode, const TS *__restrict__ V, long ldv, bool vtr,
                                             const double *__restrict__ tjj, TS *__restrict__ C1, long ld1, long s1,
                                             TS *__restrict__ C2, long ld2, long s2, int ntiles, double *smem,
                                             int s0 = 0)
{
    static_assert(SPW == 1 || SPW == 2, "update2 strips per wave");
    typedef typename AccT<TS>::type acc_t;
    constexpr int NB = B / 16, YS = Lds2<TS, B>::YS;
    // Rows of C (and reflectors of block J) are taken in "memory order": in
    // step m of a K loop, lane group q supplies row 4q + m, so each lane owns
    // 4 consecutive rows and the C accesses of the column-major orientation are
    // 32-byte runs.  The f64 accumulator puts hardware row q + 4m in element m;
    // the products stay exact because Y's rows, C's rows and the block's
    // reflector index are permuted consistently: hardware row d <-> memory row
    // perm(d) = 4*(d % 4) + d / 4 (the identity for the f32 layout, whose
    // hardware row already is 4q + m).
    auto perm = [](int d) { return sizeof(TS) == 8 ? 4 * (d & 3) + (d >> 2) : d; };
    AVS_UPD_MARK(0, 0);
    constexpr int PER = (B * 16) / NT; // staged block elements per thread
    static_assert((B * 16) % NT == 0 && PER >= 1, "update2 staging split");
    TS *YB = reinterpret_cast<TS *>(smem); // YB[buf*B*YS + k*YS + i] = Y[k][16J + i]
    TS *TL = YB + 2 * B * YS;              // TL[i*16 + jl] = T[i][16*(i/16)+jl]
    // thread index through an empty asm: nothing derived from it is hoisted
    // out of a caller's loop over a chain of operations
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    const int lane = t & 63, w = t >> 6, q = lane >> 4, il = lane & 15;

    // (k, i) of the u-th element this thread stages; coalesced along the
    // contiguous direction of V
    auto kidx = [&](int u, int &k, int &i) {
        const int idx = t + NT * u;
        if (vtr) { i = idx % 16; k = idx / 16; } else { k = idx % B; i = idx / B; }
    };
    auto yload = [&](int J, TS(&r)[PER]) {
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            int k, i;
            kidx(u, k, i);
            const int ic = 16 * J + i;
            TS v = V[toff(k, ic, ldv, vtr)];
            if (mode == MODE_GE) v = (k < ic) ? (TS)0 : (k == ic ? (TS)1 : v);
            if (mode == MODE_TT && k > ic) v = (TS)0;
            r[u] = v;
        }
    };
    auto ystore = [&](int buf, const TS(&r)[PER]) {
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            int k, i;
            kidx(u, k, i);
            YB[buf * B * YS + k * YS + i] = r[u];
        }
    };
    for (int idx = t; idx < 16 * B; idx += NT) TL[idx] = (TS)tjj[idx];
    {
        TS r[PER];
        yload(APPLYT ? 0 : NB - 1, r);
        ystore(0, r);
    }

    // logical element (row, col) of a C tile is at base + row*rs + col*cs
    const long rs1 = CTR ? ld1 : 1, cs1 = CTR ? 1 : ld1;
    const long rs2 = CTR ? ld2 : 1, cs2 = CTR ? 1 : ld2;
    const int pil = perm(il);
    const int nstrips = ntiles * NB;
    // wave-uniform; strips come in pairs (NB is even)
    const bool mine = SPW == 2 ? 2 * w < nstrips : w < NB / 2;
    TS *c1p[SPW], *c2p[SPW];
#pragma unroll
    for (int sp = 0; sp < SPW; ++sp) {
        const int strip = SPW == 2 ? 2 * w + sp : s0 + w;
        const int tile = strip / NB, col = (strip % NB) * 16 + il;
        c1p[sp] = C1 + (long)tile * s1 + (long)col * cs1 + (long)(4 * q) * rs1;
        c2p[sp] = C2 + (long)tile * s2 + (long)col * cs2 + (long)(4 * q) * rs2;
    }
    acc_t acc2[SPW][NB];
    if (mine) {
#pragma unroll
        for (int sp = 0; sp < SPW; ++sp) {
#pragma unroll
            for (int tb = 0; tb < NB; ++tb) {
#pragma unroll
                for (int m = 0; m < 4; ++m) acc2[sp][tb][m] = c2p[sp][(long)(16 * tb + m) * rs2];
            }
        }
    }
    __syncthreads();
    AVS_UPD_MARK(1, 0);

    // SPW = 1 has the registers to request C1[J] one block ahead (c1nx); SPW
    // = 2 requests it at the start of block J, before the W MFMAs
    acc_t c1nx[SPW];
    if (SPW == 1 && mine && mode != MODE_GE) {
        const int J0 = APPLYT ? 0 : NB - 1;
#pragma unroll
        for (int m = 0; m < 4; ++m) c1nx[0][m] = c1p[0][(long)(16 * J0 + m) * rs1];
    }

#pragma unroll 1
    for (int jj = 0; jj < NB; ++jj) {
        const int J = APPLYT ? jj : (NB - 1 - jj);
        const TS *YJ = YB + (jj & 1) * B * YS;
        TS yr[PER];
        if (jj + 1 < NB) yload(APPLYT ? jj + 1 : NB - 2 - jj, yr);
        if (mine) {
            // C1[J] is requested before the W MFMAs and added after them
            acc_t c1v[SPW], Wa[SPW];
#pragma unroll
            for (int sp = 0; sp < SPW; ++sp) {
                Wa[sp] = vzero4<acc_t>();
                if (mode != MODE_GE) {
                    if constexpr (SPW == 1) {
                        c1v[sp] = c1nx[sp];
                        if (jj + 1 < NB) {
                            const int Jn = APPLYT ? jj + 1 : NB - 2 - jj;
#pragma unroll
                            for (int m = 0; m < 4; ++m) c1nx[sp][m] = c1p[sp][(long)(16 * Jn + m) * rs1];
                        }
                    } else {
#pragma unroll
                        for (int m = 0; m < 4; ++m) c1v[sp][m] = c1p[sp][(long)(16 * J + m) * rs1];
                    }
                }
            }
            // W = Y_J^T C2 + C1[J]
#pragma unroll
            for (int tb = 0; tb < NB; ++tb) {
                if (mode == MODE_GE && tb < J) continue;
                if (mode == MODE_TT && tb > J) continue;
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const TS av = YJ[(16 * tb + 4 * q + m) * YS + pil];
#pragma unroll
                    for (int sp = 0; sp < SPW; ++sp) Wa[sp] = mfma16(av, acc2[sp][tb][m], Wa[sp]);
                }
            }
            AVS_UPD_MARK(2, jj);
            if (mode != MODE_GE) {
#pragma unroll
                for (int sp = 0; sp < SPW; ++sp) {
#pragma unroll
                    for (int m = 0; m < 4; ++m) Wa[sp][m] += c1v[sp][m];
                }
            }
            AVS_UPD_MARK(3, jj);
            // W' = T_JJ^T W (Q^T) or T_JJ W (Q)
            acc_t Wp[SPW];
#pragma unroll
            for (int sp = 0; sp < SPW; ++sp) Wp[sp] = vzero4<acc_t>();
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const TS av = APPLYT ? TL[(16 * J + 4 * q + m) * 16 + pil] : TL[(16 * J + pil) * 16 + 4 * q + m];
#pragma unroll
                for (int sp = 0; sp < SPW; ++sp) Wp[sp] = mfma16(av, Wa[sp][m], Wp[sp]);
            }
            if (mode != MODE_GE) {
#pragma unroll
                for (int sp = 0; sp < SPW; ++sp) {
#pragma unroll
                    for (int m = 0; m < 4; ++m) c1p[sp][(long)(16 * J + m) * rs1] = c1v[sp][m] - Wp[sp][m];
                }
            }
            AVS_UPD_MARK(4, jj);
            // C2 -= Y_J W'
#pragma unroll
            for (int tb = 0; tb < NB; ++tb) {
                if (mode == MODE_GE && tb < J) continue;
                if (mode == MODE_TT && tb > J) continue;
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const TS av = -YJ[(16 * tb + pil) * YS + 4 * q + m];
#pragma unroll
                    for (int sp = 0; sp < SPW; ++sp) acc2[sp][tb] = mfma16(av, Wp[sp][m], acc2[sp][tb]);
                }
            }
        }
        AVS_UPD_MARK(5, jj);
        if (jj + 1 < NB) ystore((jj + 1) & 1, yr);
        AVS_UPD_MARK(6, jj);
        __syncthreads();
        AVS_UPD_MARK(7, jj);
    }
    if (mine) {
#pragma unroll
        for (int sp = 0; sp < SPW; ++sp) {
#pragma unroll
            for (int tb = 0; tb < NB; ++tb) {
#pragma unroll
                for (int m = 0; m < 4; ++m) c2p[sp][(long)(16 * tb + m) * rs2] = acc2[sp][tb][m];
            }
        }
    }
}

// blocked-staging update with the arithmetic fixed at compile time: split-BF16
// (FP32 storage only) or the native MFMA of the storage type
template <typename TS, int B, int MODE, bool APPLYT, bool SPLIT, int SPW = 2>
__device__ __forceinline__ void update_fixed(const TS *V, long ldv, bool vtr, const double *tjj, TS *C1, long ld1,
                                             long s1, TS *C2, long ld2, long s2, bool ctr, int ntiles, double *smem,
                                             int s0 = 0)
{
    if constexpr (SPLIT && sizeof(TS) == 4) {
        update_body_split<B, MODE, APPLYT, UPD_THREADS>((const float *)V, ldv, vtr, tjj, (float *)C1, ld1, s1,
                                                         (float *)C2, ld2, s2, ctr, ntiles, smem);
    } else {
        if (ctr)
            update2_body<TS, B, APPLYT, UPD_THREADS, true, SPW>(MODE, V, ldv, vtr, tjj, C1, ld1, s1, C2, ld2, s2, ntiles,
                                                                smem, s0);
        else
            update2_body<TS, B, APPLYT, UPD_THREADS, false, SPW>(MODE, V, ldv, vtr, tjj, C1, ld1, s1, C2, ld2, s2,
                                                                 ntiles, smem, s0);
    }
}
// the same with the tile mode chosen at run time (one body for all modes) and
// the orientation fixed (V and C both in adjoint addressing when TR)
template <typename TS, int B, bool SPLIT, bool TR>
__device__ __forceinline__ void update_any(int mode, const TS *V, long ldv, const double *tjj, TS *C1, long ld1,
                                           long s1, TS *C2, long ld2, long s2, int ntiles, double *smem)
{
    const bool vtr = TR, ctr = TR;
    if constexpr (SPLIT && sizeof(TS) == 4) {
        if (mode == MODE_GE)
            update_fixed<TS, B, MODE_GE, true, true>(V, ldv, vtr, tjj, C1, ld1, s1, C2, ld2, s2, ctr, ntiles, smem);
        else if (mode == MODE_TS)
            update_fixed<TS, B, MODE_TS, true, true>(V, ldv, vtr, tjj, C1, ld1, s1, C2, ld2, s2, ctr, ntiles, smem);
        else
            update_fixed<TS, B, MODE_TT, true, true>(V, ldv, vtr, tjj, C1, ld1, s1, C2, ld2, s2, ctr, ntiles, smem);
    } else {
        update2_body<TS, B, true, UPD_THREADS, TR>(mode, V, ldv, vtr, tjj, C1, ld1, s1, C2, ld2, s2, ntiles, smem);
    }
}
// LDS (doubles) of update_fixed
template <typename TS, int B, bool SPLIT> struct LdsU {
    static constexpr int value = (SPLIT && sizeof(TS) == 4) ? Lds<TS, B>::update_split : Lds2<TS, B>::upd;
};
// column tiles one update workgroup handles (8 strips of 16 columns)
template <int B> struct UpdTiles { static constexpr int value = (2 * UPD_THREADS / 64) / (B / 16); };

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
