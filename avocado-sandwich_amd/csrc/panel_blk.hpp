// panel_blk.hpp -- blocked Householder panel for 128 x 128 tiles on gfx950.
//
// The same operators as panel_body (GEQRT, TSQRT, tree-merge TSQRT; reference
// kernels QR_unsafe_kernel! / QR_unsafe_kernel2!, KAbasedSVD/src/QRkernels.jl:
// 3-129) with the same outputs and Householder convention, organised for a
// short critical path instead of one workgroup barrier per reflector:
//
//  - 8 waves; wave w owns the 16-column strip w of the factored tile in the
//    FP64 MFMA accumulator layout (lane (c, q) = (lane & 15, lane >> 4) holds
//    memory rows 16 rb + 4 q + m, m = 0..3, of column 16 w + c for every row
//    block rb), so the strip is both the operand of the block updates and the
//    data of its own factorization;
//  - block J (reflectors 16 J .. 16 J + 15) is factored by wave J alone, with
//    no workgroup barrier: the pivot column reaches the 16 lanes of each row
//    group by one 64-bit DPP row_newbcast per register, dot products are
//    summed over the four row groups by two lane swaps, and the Householder
//    scalars come from ||x||^2 (the pivot column's own dot product) and alpha;
//  - the strips right of J then take block J as a compact-WY update on FP64
//    MFMA, W = C1[J] + Y_J^T A, W' = T_JJ^T W, C1[J] -= W', A -= Y_J W' (the
//    arithmetic of update2_body), while wave J forms T_JJ from the Gram
//    entries it collected; two barriers per block.
// Waves left of J are finished and only take part in the barriers; the wave
// that factors next runs at a raised wave priority.
//
// Reflector j of a TS panel touches row j of R and every row of the bottom
// tile; of a TT panel (bottom tile upper triangular) rows 0..j of the bottom
// tile; of a GE panel rows j..B-1 of its tile (unit pivot, implicit).
#pragma once
#include <utility>

#include "device.hpp"

namespace avs {

constexpr int PBLK_THREADS = 512;

// LDS of panel_blk_body (doubles): Y_J and T_JJ double-buffered (Y row
// stride 17), the block's Gram entries, tau
template <int B> struct LdsBlk {
    static constexpr int YS = 17;
    static constexpr int value = 2 * B * YS + 2 * 256 + 256 + B;
};

template <typename F, int... I>
__device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, I...>)
{
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F> __device__ __forceinline__ void static_for(F &&f)
{
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// lane L of each row of 16 lanes, to every lane of that row (every lane is
// written, so the old value is not needed)
template <int L> __device__ __forceinline__ double row_bcast(double v)
{
    return __builtin_amdgcn_update_dpp(0.0, v, 0x150 + L, 0xf, 0xf, true);
}
// Cross-row exchanges with v_permlane{16,32}_swap applied to two copies of
// one value: permlane16 gives X = rows (0,0,2,2) and Y = rows (1,1,3,3) of
// rows of 16 lanes, permlane32 gives X = rows (0,1,0,1) and Y = (2,3,2,3).
struct SwapXY { double x, y; };
template <int W> __device__ __forceinline__ SwapXY swap_rows(double v)
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
// sum over the four row groups (lanes c, c+16, c+32, c+48), the same in every
// lane (each step adds the same two values in the same order everywhere)
__device__ __forceinline__ double rowgroup_sum(double v)
{
    const SwapXY a = swap_rows<16>(v);
    const SwapXY b = swap_rows<32>(a.x + a.y);
    return b.x + b.y;
}
// the value of lane (c, QJ) to every lane (c, q)
template <int QJ> __device__ __forceinline__ double rowgroup_bcast(double v)
{
    const SwapXY a = swap_rows<32>(v);
    const SwapXY b = swap_rows<16>(QJ < 2 ? a.x : a.y);
    return (QJ & 1) ? b.y : b.x;
}
// The four registers of one row block through v_fmac_f64 with the first
// source read through DPP row_newbcast:L (the pivot lane of each row of 16):
//   dot4:  acc[m] += x[m][lane L of the row] * y[m]
//   axpy4: a[m]   += a[m][lane L of the row] * s[m]
// Written as asm because the compiler does not fold a 64-bit DPP move into
// its user.  A DPP source written by a VALU instruction needs 2 wait states
// before the read: each block starts with s_nop 1, so a copy the compiler
// places just before the block is covered; the sources are not written
// inside the block (the outputs are tied to their own inputs).  EXEC writes
// (5 wait states) and MFMA results (up to 18) are covered by the callers.
template <int L> __device__ __forceinline__ d4_t dot4(d4_t acc, d4_t x, d4_t y)
{
    double a0 = acc[0], a1 = acc[1], a2 = acc[2], a3 = acc[3];
    asm volatile("s_nop 1\n\t"
                 "v_fmac_f64_dpp %0, %4, %8 row_newbcast:%12 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
                 "v_fmac_f64_dpp %1, %5, %9 row_newbcast:%12 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
                 "v_fmac_f64_dpp %2, %6, %10 row_newbcast:%12 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
                 "v_fmac_f64_dpp %3, %7, %11 row_newbcast:%12 row_mask:0xf bank_mask:0xf bound_ctrl:1"
                 : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3)
                 : "v"(x[0]), "v"(x[1]), "v"(x[2]), "v"(x[3]), "v"(y[0]), "v"(y[1]), "v"(y[2]), "v"(y[3]), "n"(L));
    return d4_t{a0, a1, a2, a3};
}
// dot4 with y == x
template <int L> __device__ __forceinline__ d4_t dot4(d4_t acc, d4_t x)
{
    double a0 = acc[0], a1 = acc[1], a2 = acc[2], a3 = acc[3];
    asm volatile("s_nop 1\n\t"
                 "v_fmac_f64_dpp %0, %4, %4 row_newbcast:%8 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
                 "v_fmac_f64_dpp %1, %5, %5 row_newbcast:%8 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
                 "v_fmac_f64_dpp %2, %6, %6 row_newbcast:%8 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
                 "v_fmac_f64_dpp %3, %7, %7 row_newbcast:%8 row_mask:0xf bank_mask:0xf bound_ctrl:1"
                 : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3)
                 : "v"(x[0]), "v"(x[1]), "v"(x[2]), "v"(x[3]), "n"(L));
    return d4_t{a0, a1, a2, a3};
}
// p_i += x[lane L0 + i of the row] * t_i for i = 0..3: four terms of four
// independent sums, one fused DPP FMA each (the T_JJ recurrence)
template <int L0>
__device__ __forceinline__ void fmac_bc4(double &p0, double &p1, double &p2, double &p3, double x, double t0, double t1,
                                         double t2, double t3)
{
    asm volatile("s_nop 1\n\t"
                 "v_fmac_f64_dpp %0, %4, %5 row_newbcast:%9 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
                 "v_fmac_f64_dpp %1, %4, %6 row_newbcast:%10 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
                 "v_fmac_f64_dpp %2, %4, %7 row_newbcast:%11 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
                 "v_fmac_f64_dpp %3, %4, %8 row_newbcast:%12 row_mask:0xf bank_mask:0xf bound_ctrl:1"
                 : "+v"(p0), "+v"(p1), "+v"(p2), "+v"(p3)
                 : "v"(x), "v"(t0), "v"(t1), "v"(t2), "v"(t3), "n"(L0), "n"(L0 + 1), "n"(L0 + 2), "n"(L0 + 3));
}
// p += x[lane L of the row] * t
template <int L> __device__ __forceinline__ void fmac_bc1(double &p, double x, double t)
{
    asm volatile("s_nop 1\n\t"
                 "v_fmac_f64_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf bound_ctrl:1"
                 : "+v"(p)
                 : "v"(x), "v"(t), "n"(L));
}
// axpy4 with one s for the four registers
template <int L> __device__ __forceinline__ d4_t axpy4(d4_t a, double s)
{
    double a0 = a[0], a1 = a[1], a2 = a[2], a3 = a[3];
    asm volatile("s_nop 1\n\t"
                 "v_fmac_f64_dpp %0, %0, %4 row_newbcast:%5 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
                 "v_fmac_f64_dpp %1, %1, %4 row_newbcast:%5 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
                 "v_fmac_f64_dpp %2, %2, %4 row_newbcast:%5 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
                 "v_fmac_f64_dpp %3, %3, %4 row_newbcast:%5 row_mask:0xf bank_mask:0xf bound_ctrl:1"
                 : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3)
                 : "v"(s), "n"(L));
    return d4_t{a0, a1, a2, a3};
}
template <int L> __device__ __forceinline__ d4_t axpy4(d4_t a, d4_t s)
{
    double a0 = a[0], a1 = a[1], a2 = a[2], a3 = a[3];
    asm volatile("s_nop 1\n\t"
                 "v_fmac_f64_dpp %0, %0, %4 row_newbcast:%8 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
                 "v_fmac_f64_dpp %1, %1, %5 row_newbcast:%8 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
                 "v_fmac_f64_dpp %2, %2, %6 row_newbcast:%8 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
                 "v_fmac_f64_dpp %3, %3, %7 row_newbcast:%8 row_mask:0xf bank_mask:0xf bound_ctrl:1"
                 : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3)
                 : "v"(s[0]), "v"(s[1]), "v"(s[2]), "v"(s[3]), "n"(L));
    return d4_t{a0, a1, a2, a3};
}
__device__ __forceinline__ double readlane_d(double v, int lane)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

// Timing hook (event id, block J); empty unless a profiling build defines it.
#ifndef AVS_BLK_MARK
#define AVS_BLK_MARK(ev, J)
#endif

template <int I> using IC = std::integral_constant<int, I>;

template <typename TS, int MODE>
__device__ __forceinline__ void panel_blk_body(TS *__restrict__ top, long ldt, TS *__restrict__ bot, long ldb,
                                               bool tr, double *__restrict__ tau, double *__restrict__ tjj,
                                               double *smem)
{
    constexpr int B = 128, NB = B / 16, YS = LdsBlk<B>::YS;
    double *YB = smem;             // YB[buf*B*YS + k*YS + i] = Y_J[k][i]
    double *TB = YB + 2 * B * YS;  // TB[buf*256 + i*16 + k] = T_JJ[i][k], double-buffered like Y_J
    double *GL = TB + 512;         // GL[k*16 + i] = y_i . y_k (i < k)
    double *tauL = GL + 256;       // tau of every reflector

    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    const int lane = t & 63, c = lane & 15, q = lane >> 4;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
    const int col = 16 * w + c;
    auto perm = [](int d) { return 4 * (d & 3) + (d >> 2); };
    const int pil = perm(c);

    TS *src = (MODE == MODE_GE) ? top : bot;
    const long lds = (MODE == MODE_GE) ? ldt : ldb;
    const long rs = tr ? lds : 1, rt = tr ? ldt : 1; // row steps of the source and top tiles

    // the strip, in the accumulator layout
    d4_t a[NB];
    {
        const TS *p = src + toff(4 * q, col, lds, tr);
#pragma unroll
        for (int rb = 0; rb < NB; ++rb) {
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const int row = 16 * rb + 4 * q + m;
                a[rb][m] = (MODE == MODE_TT && row > col) ? 0.0 : (double)p[(long)(16 * rb + m) * rs];
            }
        }
    }
    // TS/TT: R's diagonal block of this strip (rows 16w .. 16w+15 of the top
    // tile) in the accumulator layout: lane (c, q) holds rows 4q .. 4q+3 of
    // column c, zero below the diagonal
    d4_t rd = {0.0, 0.0, 0.0, 0.0};
    if (MODE != MODE_GE) {
        const TS *p = top + toff(16 * w + 4 * q, col, ldt, tr);
#pragma unroll
        for (int m = 0; m < 4; ++m) rd[m] = (4 * q + m <= c) ? (double)p[(long)m * rt] : 0.0;
    }
    for (int idx = t; idx < B; idx += PBLK_THREADS) tauL[idx] = 0.0;
    AVS_BLK_MARK(0, 0);
    __syncthreads();
    AVS_BLK_MARK(1, 0);

    // one block step; JS = J at compile time (GE/TT, whose row-block ranges
    // depend on J: one straight-line instance per J) or -1 (TS: every row
    // block, one instance for all J)
    auto step = [&](auto jsc, const int J) {
        constexpr int JS = decltype(jsc)::value;
        // row blocks that reflectors of block J touch
        auto tb_on = [&](int tb) {
            if constexpr (MODE == MODE_TS) return true;
            else if constexpr (MODE == MODE_TT) return JS >= 0 ? tb <= JS : tb <= J;
            else return JS >= 0 ? tb >= JS : tb >= J;
        };
        if (w == J || w == J + 1) __builtin_amdgcn_s_setprio(2);
        else __builtin_amdgcn_s_setprio(0);
        if (w == J) {
            // the strip's last MFMA results are read through DPP below
            asm volatile("s_nop 7\n\ts_nop 7\n\ts_nop 7");
            AVS_BLK_MARK(9, J);
            // ---- factor strip J: reflectors j = 16J + jl.  GE/TT touch a
            // prefix or suffix of the row blocks that depends on J: one
            // straight-line instance per J (JS), selected once per strip; TS
            // touches every block and has one instance (JS = -1).
            auto factor = [&](auto) {
                // row blocks the reflectors touch (1) or not (0)
                auto blk = [](int rb) constexpr {
                    return MODE == MODE_TS ? 1 : MODE == MODE_TT ? (rb <= JS ? 1 : 0) : (rb >= JS ? 1 : 0);
                };
                static_for<16>([&](auto jc) {
                    constexpr int jl = decltype(jc)::value;
                    constexpr int qj = jl / 4, mj = jl % 4;
                    const int j = 16 * J + jl;
                    d4_t sp = {0.0, 0.0, 0.0, 0.0};
                    // row j of this lane's column (row group qj): TS/TT R's row j;
                    // GE the tile's row j, taken out of the columns c >= jl of the
                    // strip (so column jl is zero on rows <= j and x needs no row
                    // mask; the finished columns keep their reflector entry in
                    // row j; the R rows collect in rd and go back at the end)
                    double pv;
                    if constexpr (MODE == MODE_GE) {
                        pv = a[JS][mj];
                        if (q == qj && c >= jl) a[JS][mj] = 0.0;
                    } else {
                        pv = rd[mj];
                    }
                    // EXEC writes of the previous reflector: 5 wait states before a DPP
                    asm volatile("s_nop 4");
                    static_for<NB>([&](auto rbc) {
                        constexpr int rb = decltype(rbc)::value;
                        if constexpr (blk(rb) != 0) sp = dot4<jl>(sp, a[rb]);
                    });
                    const double S = rowgroup_sum((sp[0] + sp[1]) + (sp[2] + sp[3]));
                    // S_c = x . a_c in every lane of column c; S_jl = ||x||^2
                    const double nn = readlane_d(S, jl);
                    // row j of the factored columns (GE: the tile's row j, TS/TT: R's
                    // row j): alpha its entry in column jl, rowv the entry in this lane's column
                    const double alpha = readlane_d(pv, 16 * qj + jl);
                    const double rowv = rowgroup_bcast<qj>(pv);
                    const bool none = (MODE == MODE_GE && j == B - 1); // GE's last column has no reflector
                    const double n2 = none ? 0.0 : fma(alpha, alpha, nn);
                    const bool nz = n2 != 0.0; // H_j = I for an all-zero column
                    double nrm, rn;
                    sqrt_rsqrt(nz ? n2 : 1.0, nrm, rn);
                    const double sg = alpha >= 0.0 ? 1.0 : -1.0;
                    const double beta = -sg * nrm;
                    const double taur = fma(fabs(alpha), rn, 1.0);
                    const double itau = rcp_nr(taur);
                    const double scal = sg * rn * itau; // 1 / (alpha - beta)
                    // w_c = v_jl . a_c; for c < jl the Gram entry y_c . y_jl
                    const double wc = fma(S, scal, rowv);
                    const double tw = taur * wc;
                    // c > jl: a_c -= tau w_c y, R[j][c] -= tau w_c; c == jl: a_jl = x -> x scal
                    // (the stored reflector), R[j][j] = beta; c < jl: finished columns
                    // (as arithmetic on 0/1 lane factors, not as branches)
                    const double nzf = nz ? 1.0 : 0.0;
                    const double gt = c > jl ? nzf : 0.0, eq = c == jl ? nzf : 0.0;
                    const double tws = fma(gt, tw * scal, eq * (1.0 - scal));
                    const double newrow = (c == jl && nz) ? beta : fma(-gt, tw, rowv);
                    // a_c -= tws x, x broadcast again from the pivot column (its
                    // lanes read their own value before they update it)
                    const double ntws = -tws;
                    asm volatile("s_nop 4");
                    static_for<NB>([&](auto rbc) {
                        constexpr int rb = decltype(rbc)::value;
                        if constexpr (blk(rb) != 0) a[rb] = axpy4<jl>(a[rb], ntws);
                    });
                    if (q == qj && c >= jl) rd[mj] = newrow;
                    // one LDS store: the Gram entry (c < jl) or tau_j (c == jl)
                    if (q == 0 && c <= jl) {
                        double *dst = c < jl ? GL + jl * 16 + c : tauL + j;
                        *dst = !nz ? 0.0 : (c < jl ? wc : taur);
                    }
                });
            };
            factor(jsc);
            if constexpr (MODE == MODE_GE) {
                // the strip's diagonal block: R (rows <= column, in rd) and the
                // reflectors (rows > column, in the strip) have disjoint supports
#pragma unroll
                for (int m = 0; m < 4; ++m) a[JS][m] += rd[m];
                rd = d4_t{0.0, 0.0, 0.0, 0.0};
            }
            AVS_BLK_MARK(2, J);
            // ---- Y_J for the MFMA updates (reflector rows with the implicit
            // unit pivot and zeros written out)
            double *Y = YB + (J & 1) * B * YS;
#pragma unroll
            for (int rb = 0; rb < NB; ++rb) {
                if (tb_on(rb)) {
#pragma unroll
                    for (int m = 0; m < 4; ++m) {
                        const int r = 4 * q + m;
                        double v = a[rb][m];
                        if (MODE == MODE_GE && rb == J) v = r < c ? 0.0 : (r == c ? 1.0 : v);
                        Y[(16 * rb + r) * YS + c] = v;
                    }
                }
            }
        }
        // C1 rows of block J for this strip (TS/TT: R rows 16J .. 16J+15 of the top tile)
        d4_t c1v = {0.0, 0.0, 0.0, 0.0};
        if (MODE != MODE_GE && w > J) {
#pragma unroll
            for (int m = 0; m < 4; ++m) c1v[m] = (double)top[toff(16 * J + 4 * q, col, ldt, tr) + (long)m * rt];
        }
        AVS_BLK_MARK(3, J);
        __syncthreads(); // Y_J published
        AVS_BLK_MARK(4, J);
        const double *Y = YB + (J & 1) * B * YS;
        d4_t W = {0.0, 0.0, 0.0, 0.0};
        if (w > J) {
            // W = Y_J^T A (+ C1[J])
#pragma unroll
            for (int tb = 0; tb < NB; ++tb) {
                if (tb_on(tb)) {
#pragma unroll
                    for (int m = 0; m < 4; ++m) W = mfma16(Y[(16 * tb + 4 * q + m) * YS + pil], a[tb][m], W);
                }
            }
            if (MODE != MODE_GE) {
#pragma unroll
                for (int m = 0; m < 4; ++m) W[m] += c1v[m];
            }
        } else if (w == J) {
            // T_JJ while the other strips form W, row c in the lanes of column c:
            // T[c][k] = -tau_k sum_{m=c}^{k-1} T[c][m] G[m][k] (T[c][m] = 0 for m < c).
            // Each lane reads its own row of G once (gr[k] = G[c][k]); G[m][k]
            // then comes from lane m by DPP, so the recurrence waits on no LDS
            // (term m of row c's sum goes to partial sum m % 4; one fused DPP
            // FMA per term)
            double gr[16], trow[16];
#pragma unroll
            for (int k = 0; k < 16; ++k) gr[k] = GL[k * 16 + c];
            const double *tb = tauL + 16 * J;
            asm volatile("s_nop 4"); // EXEC written by the branch above: 5 wait states before a DPP read
            static_for<16>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                double p0 = 0.0, p1 = 0.0, p2 = 0.0, p3 = 0.0;
                static_for<k / 4>([&](auto gc) {
                    constexpr int g4 = 4 * decltype(gc)::value;
                    fmac_bc4<g4>(p0, p1, p2, p3, gr[k], trow[g4], trow[g4 + 1], trow[g4 + 2], trow[g4 + 3]);
                });
                static_for<k % 4>([&](auto rc) {
                    constexpr int mm = 4 * (k / 4) + decltype(rc)::value;
                    if constexpr ((mm & 3) == 0) fmac_bc1<mm>(p0, gr[k], trow[mm]);
                    else if constexpr ((mm & 3) == 1) fmac_bc1<mm>(p1, gr[k], trow[mm]);
                    else if constexpr ((mm & 3) == 2) fmac_bc1<mm>(p2, gr[k], trow[mm]);
                    else fmac_bc1<mm>(p3, gr[k], trow[mm]);
                });
                const double tk = tb[k];
                trow[k] = c == k ? tk : (c < k ? -tk * ((p0 + p1) + (p2 + p3)) : 0.0);
            });
            if (q == 0) {
                double *Tr = TB + (J & 1) * 256 + c * 16;
#pragma unroll
                for (int k = 0; k < 16; ++k) Tr[k] = trow[k];
                if (tjj) {
#pragma unroll
                    for (int k = 0; k < 16; ++k) tjj[(16 * J + c) * 16 + k] = trow[k];
                }
            }
        }
        AVS_BLK_MARK(5, J);
        __syncthreads(); // T_JJ published
        AVS_BLK_MARK(6, J);
        if (w > J) {
            // W' = T_JJ^T W
            const double *TL = TB + (J & 1) * 256;
            d4_t Wp = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int m = 0; m < 4; ++m) Wp = mfma16(TL[(4 * q + m) * 16 + pil], W[m], Wp);
            if (MODE != MODE_GE) {
#pragma unroll
                for (int m = 0; m < 4; ++m) top[toff(16 * J + 4 * q, col, ldt, tr) + (long)m * rt] = (TS)(c1v[m] - Wp[m]);
            }
            // A -= Y_J W'
#pragma unroll
            for (int tb = 0; tb < NB; ++tb) {
                if (tb_on(tb)) {
#pragma unroll
                    for (int m = 0; m < 4; ++m) a[tb] = mfma16(-Y[(16 * tb + pil) * YS + 4 * q + m], Wp[m], a[tb]);
                }
            }
        }
    };
    if constexpr (MODE == MODE_TS) {
#pragma unroll 1
        for (int J = 0; J < NB; ++J) step(IC<-1>{}, J);
    } else {
        static_for<NB>([&](auto jc) { step(jc, decltype(jc)::value); });
    }
    __builtin_amdgcn_s_setprio(0);
    AVS_BLK_MARK(7, 0);

    // ---- write back: the strip (GE: R and reflectors; TS: reflectors; TT:
    // reflectors in the upper triangle only) and R's diagonal block (TS/TT).
    // The lane's addresses are formed again from a fresh thread index and
    // tile pointers, so none of the prologue's is kept live across the loop.
    int te = threadIdx.x;
    asm volatile("" : "+v"(te));
    TS *srcw = src, *topw = top;
    asm volatile("" : "+s"(srcw), "+s"(topw));
    const int ce = te & 15, qe = (te & 63) >> 4, cole = 16 * (te >> 6) + ce;
    {
        TS *p = srcw + toff(4 * qe, cole, lds, tr);
#pragma unroll
        for (int rb = 0; rb < NB; ++rb) {
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const int row = 16 * rb + 4 * qe + m;
                if (MODE != MODE_TT || row <= cole) p[(long)(16 * rb + m) * rs] = (TS)a[rb][m];
            }
        }
    }
    if (MODE != MODE_GE) {
        TS *p = topw + toff(16 * (te >> 6) + 4 * qe, cole, ldt, tr);
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            if (4 * qe + m <= ce) p[(long)m * rt] = (TS)rd[m];
        }
    }
    AVS_BLK_MARK(8, 0);
    __syncthreads();
    for (int idx = t; idx < B; idx += PBLK_THREADS) tau[idx] = tauL[idx];
}

} // namespace avs
