// kernels.hip -- launchable kernels around the tile-operator bodies of
// device.hpp: standalone operator kernels (the host-enqueued plain path and the
// tile-operator ABI), layout conversion, band extraction and the persistent
// dependency-counting executor of the tile DAG.
#include <stdlib.h>
#include <string.h>

#include "device.hpp"
#include "kernels.hpp"
#include "panel_blk.hpp"

namespace avs {

// ----------------------------------------------------------------- layout
// column-major (lda) <-> tile-major: tile (i,j) at (j*nt+i)*b*b, column-major inside
template <typename TS>
__global__ void __launch_bounds__(256) pack_kernel(const TS *__restrict__ A, long lda, TS *__restrict__ tiles,
                                                   int nt, int b)
{
    const int i = blockIdx.x, j = blockIdx.y;
    TS *dst = tiles + ((long)j * nt + i) * b * b;
    const TS *src = A + (long)i * b + (long)j * b * lda;
    for (int idx = threadIdx.x; idx < b * b; idx += blockDim.x) {
        const int r = idx % b, c = idx / b;
        dst[idx] = src[r + (long)c * lda];
    }
}

// mode 0: raw tiles; mode 1: the reference's Blockbidiag! output (band of the
// first kend steps, exact zeros outside it, raw trailing block)
template <typename TS>
__global__ void __launch_bounds__(256) unpack_kernel(const TS *__restrict__ tiles, TS *__restrict__ A, long lda,
                                                     int nt, int b, int mode, int kend)
{
    const int i = blockIdx.x, j = blockIdx.y;
    const TS *src = tiles + ((long)j * nt + i) * b * b;
    TS *dst = A + (long)i * b + (long)j * b * lda;
    const bool banded = mode == 1 && (i < kend || j < kend);
    for (int idx = threadIdx.x; idx < b * b; idx += blockDim.x) {
        const int r = idx % b, c = idx / b;
        TS v = src[idx];
        if (banded) {
            if (j == i) v = (r <= c) ? v : (TS)0;
            else if (j == i + 1) v = (r >= c) ? v : (TS)0;
            else v = (TS)0;
        }
        dst[r + (long)c * lda] = v;
    }
}

// In place, the zeros the reference's Blockbidiag! writes (tiledalgos.jl:46,58)
// for its first kend steps: strict lower part of the diagonal tiles, strict
// upper part of the super-diagonal tiles, every other tile of those rows and
// columns; the trailing block is left as it is
template <typename TS>
__global__ void __launch_bounds__(256) zero_band_kernel(TS *__restrict__ tiles, int nt, int b, int kend)
{
    const int i = blockIdx.x, j = blockIdx.y;
    if (i >= kend && j >= kend) return;
    TS *t = tiles + ((long)j * nt + i) * b * b;
    for (int idx = threadIdx.x; idx < b * b; idx += blockDim.x) {
        const int r = idx % b, c = idx / b;
        const bool keep = (j == i) ? r <= c : (j == i + 1 ? r >= c : false);
        if (!keep) t[idx] = (TS)0;
    }
}

// band[(b + i - j) + j*ldband] = A(i,j), max(0,j-b) <= i <= j, as FP64
template <typename TS>
__global__ void __launch_bounds__(256) band_kernel(const TS *__restrict__ tiles, double *__restrict__ band,
                                                   long ldband, int nt, int b, long n)
{
    const long j = blockIdx.x;
    for (int d = threadIdx.x; d <= b; d += blockDim.x) {
        const long i = j - b + d;
        double v = 0.0;
        if (i >= 0) {
            const long I = i / b, J = j / b;
            const int r = (int)(i % b), c = (int)(j % b);
            v = (double)tiles[(J * nt + I) * b * b + r + (long)c * b];
        }
        band[d + j * ldband] = v;
    }
}

// ----------------------------------------------------- standalone operators
template <typename TS, int B, int MODE>
__global__ void __launch_bounds__(PanelThreads<B>::value) panel_kernel(TS *top, long ldt, TS *bot, long ldb, int tr,
                                                                       double *tau, double *tjj)
{
    __shared__ double smem[Lds<TS, B>::panel];
    panel_body<TS, B, MODE, PanelThreads<B>::value>(top, ldt, bot, ldb, tr != 0, tau, tjj, smem);
}

// the blocked panel of 128 x 128 tiles (panel_blk.hpp)
template <typename TS, int MODE>
__global__ void __launch_bounds__(PBLK_THREADS) panel_blk_kernel(TS *top, long ldt, TS *bot, long ldb, int tr,
                                                                 double *tau, double *tjj)
{
    __shared__ double smem[LdsBlk<128>::value];
    panel_blk_body<TS, MODE>(top, ldt, bot, ldb, tr != 0, tau, tjj, smem);
}

// b = 128 panels use the blocked body unless AVSVD_PANEL=rows selects the
// one-barrier-per-reflector body (read once per process)
static bool panel_blk_on()
{
    static const bool on = [] {
        const char *e = getenv("AVSVD_PANEL");
        return !(e && strcmp(e, "rows") == 0);
    }();
    return on;
}

template <typename TS, int B, int MODE>
__global__ void __launch_bounds__(NTHREADS) tjj_kernel(const TS *V, long ldv, int vtr, const double *tau,
                                                       double *tjj)
{
    __shared__ double smem[Lds<TS, B>::tjj];
    tjj_body<TS, B, MODE>(V, ldv, vtr != 0, tau, tjj, smem);
}

// one block per group of `per_block` column tiles (at most UpdTiles<B>);
// group g starts at tile g*per_block
template <typename TS, int B, int MODE, bool APPLYT, bool SPLIT>
__global__ void __launch_bounds__(UPD_THREADS, 2) update_kernel(const TS *V, long ldv, int vtr, const double *tjj,
                                                             TS *C1, long ld1, long s1, TS *C2, long ld2, long s2,
                                                             int ctr, int ntiles_total, int per_block)
{
    __shared__ double smem[LdsU<TS, B, SPLIT>::value];
    const int first = blockIdx.x * per_block;
    int cnt = ntiles_total - first;
    if (cnt > per_block) cnt = per_block;
    if (cnt <= 0) return;
    update_fixed<TS, B, MODE, APPLYT, SPLIT>(V, ldv, vtr != 0, tjj, C1 + (long)first * s1, ld1, s1,
                                             C2 + (long)first * s2, ld2, s2, ctr != 0, cnt, smem);
}

// ------------------------------------------------------ persistent executor
// relaxed agent-scope load for spinning (no cache invalidate per poll); the
// acquire fence after the wait orders the data reads that follow
__device__ __forceinline__ int ld_relaxed(const int *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename TS, int B>
__global__ void __launch_bounds__(NTHREADS, 1) persistent_kernel(TS *tiles, double *taus, double *tjjs,
                                                                  const Task *__restrict__ tasks, int ntasks,
                                                                  const int *__restrict__ dep_tile,
                                                                  const int *__restrict__ dep_val,
                                                                  const int *__restrict__ sig_tile, int *cnt,
                                                                  int *ticket, long long wait_ticks, int *status,
                                                                  int split)
{
    __shared__ double smem[Lds<TS, B>::maxv];
    __shared__ int sh_task;
    __shared__ int sh_fail;
    const int t = threadIdx.x;
    const long tsz = (long)B * B;
    for (;;) {
        // a failed wait anywhere ends the launch: no new ticket once status is set
        if (t == 0) {
            sh_task = ld_relaxed(status) != 0 ? ntasks : atomicAdd(ticket, 1);
            sh_fail = 0;
        }
        __syncthreads();
        const int tk = sh_task;
        if (tk >= ntasks) break;
        const Task T = tasks[tk];
        // wait for dependencies, bounded in wall-clock time so that a schedule
        // error ends the launch with a status instead of hanging it.  Thread 0
        // polls with relaxed loads, then one agent-scope acquire fence (cache
        // invalidate) orders the workgroup's reads after the producers' writes.
        if (t == 0) {
            const long long t0 = wall_clock64();
            for (int d = 0; d < T.ndep; ++d) {
                const int *p = cnt + dep_tile[T.dep0 + d];
                const int need = dep_val[T.dep0 + d];
                while (ld_relaxed(p) < need) {
                    __builtin_amdgcn_s_sleep(1);
                    // the bound covers the whole launch, not each wait: a
                    // workgroup whose wait ran out sets status, and every
                    // other wait gives up as soon as it sees that
                    if (wall_clock64() - t0 > wait_ticks || ld_relaxed(status) != 0) {
                        sh_fail = 1;
                        break;
                    }
                }
                if (sh_fail) break;
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        __syncthreads();
        if (sh_fail) {
            // the first failure keeps its task index
            if (t == 0) atomicCAS(status, 0, 1 + tk);
            break;
        }
        double *tau = taus + (long)T.fac * B;
        double *tjj = tjjs + (long)T.fac * 16 * B;
        const bool tr = T.trans != 0;
        if (T.kind == TK_PANEL) {
            TS *top = tiles + (long)T.vtile * tsz;
            TS *bot = tiles + (long)T.btile * tsz;
            if (T.mode == MODE_GE) panel_body<TS, B, MODE_GE>(top, B, top, B, tr, tau, tjj, smem);
            else if (T.mode == MODE_TS) panel_body<TS, B, MODE_TS>(top, B, bot, B, tr, tau, tjj, smem);
            else panel_body<TS, B, MODE_TT>(top, B, bot, B, tr, tau, tjj, smem);
        } else {
            const TS *V = tiles + (long)T.vtile * tsz;
            TS *C2 = tiles + (long)T.btile * tsz;
            TS *C1 = tiles + (long)T.c1tile * tsz;
            const long st = (long)T.cstep * tsz;
            if (T.mode == MODE_GE)
                update_sel<TS, B, MODE_GE>(split, V, B, tr, tjj, C2, B, st, C2, B, st, tr, T.ntiles, smem);
            else if (T.mode == MODE_TS)
                update_sel<TS, B, MODE_TS>(split, V, B, tr, tjj, C1, B, st, C2, B, st, tr, T.ntiles, smem);
            else
                update_sel<TS, B, MODE_TT>(split, V, B, tr, tjj, C1, B, st, C2, B, st, tr, T.ntiles, smem);
        }
        __builtin_amdgcn_s_waitcnt(0); // this thread's global stores are acknowledged
        __syncthreads();
        // every thread's stores are acknowledged before the barrier; thread 0's
        // agent-scope release fence (L2 write-back) then publishes them before
        // the counters move
        if (t == 0) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            for (int s = 0; s < T.nsig; ++s)
                __hip_atomic_fetch_add(cnt + sig_tile[T.sig0 + s], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// ------------------------------------------------ stream-ordered executor
// Panel phase: workgroup c runs the panel operations of chain c in order
// (a domain's GEQRT + TSQRT chain, or one tree merge).
template <typename TS, int B, int MODE>
__global__ void __launch_bounds__(PanelThreads<B>::value) chain_panel_kernel(TS *tiles, double *taus, double *tjjs,
                                                                             const ChainOp *__restrict__ ops,
                                                                             const int *__restrict__ chain_start,
                                                                             int chain0)
{
    constexpr int NT = PanelThreads<B>::value;
    __shared__ double smem[Lds<TS, B>::panel];
    const long tsz = (long)B * B;
    const int c = chain0 + blockIdx.x;
    for (int o = chain_start[c]; o < chain_start[c + 1]; ++o) {
        const ChainOp T = ops[o];
        TS *top = tiles + (long)T.vtile * tsz;
        TS *bot = tiles + (long)T.btile * tsz;
        double *tau = taus + (long)T.fac * B;
        double *tjj = tjjs + (long)T.fac * 16 * B;
        panel_body<TS, B, MODE, NT>(top, B, MODE == MODE_GE ? top : bot, B, T.trans != 0, tau, tjj, smem);
        __syncthreads();
    }
}

// Update phase: workgroup (chain c, column group g) applies every operation of
// chain c, in order, to the `per` column tiles of group g.  The chain's C1 row
// tile is rewritten by each operation and stays hot in L2.
// SPW = 1 (launches too small to fill the GPU, per = 1): two workgroups per
// column tile, each taking half of its 16-column strips, one strip per wave.
template <typename TS, int B, bool SPLIT, bool TR, int MODE, int SPW = 2>
__global__ void __launch_bounds__(UPD_THREADS, SPW == 1 ? 3 : 2) chain_update_kernel(TS *tiles, const double *tjjs,
                                                                       const ChainOp *__restrict__ ops,
                                                                       const int *__restrict__ chain_start, int chain0,
                                                                       int groups, int per)
{
    __shared__ double smem[LdsU<TS, B, SPLIT>::value];
    const long tsz = (long)B * B;
    constexpr int HALVES = SPW == 1 ? 2 : 1;
    const int c = chain0 + blockIdx.x / (groups * HALVES);
    const int g = (blockIdx.x / HALVES) % groups;
    const int s0 = (blockIdx.x % HALVES) * (B / 32);
    for (int o = chain_start[c]; o < chain_start[c + 1]; ++o) {
        const ChainOp T = ops[o];
        const int first = g * per;
        const int cnt = min(per, T.ncols - first);
        if (cnt <= 0) break;
        const TS *V = tiles + (long)T.vtile * tsz;
        const long st = (long)T.cstep * tsz;
        TS *C2 = tiles + (long)T.btile * tsz + (long)first * st;
        // GE (UNMQR) updates the one tile it is given: C1 = C2
        TS *C1 = MODE == MODE_GE ? C2 : tiles + (long)T.c1tile * tsz + (long)first * st;
        const double *tjj = tjjs + (long)T.fac * 16 * B;
        update_fixed<TS, B, MODE, true, SPLIT, SPW>(V, B, TR, tjj, C1, B, st, C2, B, st, TR, cnt, smem, s0);
        __syncthreads();
    }
}

template <typename TS, int MODE>
__global__ void __launch_bounds__(PBLK_THREADS) chain_panel_blk_kernel(TS *tiles, double *taus, double *tjjs,
                                                                       const ChainOp *__restrict__ ops,
                                                                       const int *__restrict__ chain_start, int chain0)
{
    constexpr int B = 128;
    __shared__ double smem[LdsBlk<B>::value];
    const long tsz = (long)B * B;
    const int c = chain0 + blockIdx.x;
    for (int o = chain_start[c]; o < chain_start[c + 1]; ++o) {
        const ChainOp T = ops[o];
        TS *top = tiles + (long)T.vtile * tsz;
        TS *bot = tiles + (long)T.btile * tsz;
        double *tau = taus + (long)T.fac * B;
        double *tjj = tjjs + (long)T.fac * 16 * B;
        panel_blk_body<TS, MODE>(top, B, MODE == MODE_GE ? top : bot, B, T.trans != 0, tau, tjj, smem);
        __syncthreads();
    }
}

template <typename TS, int B>
void launch_chain_panel(int nchains, int mode, void *tiles, double *taus, double *tjjs, const ChainOp *ops,
                        const int *chain_start, int chain0, hipStream_t s)
{
    if constexpr (B == 128) {
        if (panel_blk_on()) {
#define AVS_CHB(M)                                                                                                     \
    hipLaunchKernelGGL((chain_panel_blk_kernel<TS, M>), dim3(nchains), dim3(PBLK_THREADS), 0, s, (TS *)tiles, taus,    \
                       tjjs, ops, chain_start, chain0)
            if (mode == MODE_GE) AVS_CHB(MODE_GE);
            else if (mode == MODE_TS) AVS_CHB(MODE_TS);
            else AVS_CHB(MODE_TT);
#undef AVS_CHB
            return;
        }
    }
#define AVS_CHP(M)                                                                                                     \
    hipLaunchKernelGGL((chain_panel_kernel<TS, B, M>), dim3(nchains), dim3(PanelThreads<B>::value), 0, s, (TS *)tiles, \
                       taus, tjjs, ops, chain_start, chain0)
    if (mode == MODE_GE) AVS_CHP(MODE_GE);
    else if (mode == MODE_TS) AVS_CHP(MODE_TS);
    else AVS_CHP(MODE_TT);
#undef AVS_CHP
}
template <typename TS, int B>
void launch_chain_update(int nchains, int groups, int per, int mode, int trans, void *tiles, const double *tjjs,
                         const ChainOp *ops, const int *chain_start, int chain0, int split, hipStream_t s)
{
    // a launch of fewer workgroups than two per CU: half-tile workgroups
    // (b = 128, one column tile per workgroup, native arithmetic)
    if constexpr (B == 128) {
        // AVSVD_UPD_HALF=all: every such launch; =off: none (read once per process)
        static const int slots = [] {
            const char *e = getenv("AVSVD_UPD_HALF");
            if (e && strcmp(e, "all") == 0) return 1 << 30;
            if (e && strcmp(e, "off") == 0) return 0;
            int dev = 0, ncu = 0;
            if (hipGetDevice(&dev) != hipSuccess ||
                hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
                return 0;
            return 2 * ncu;
        }();
        if (per == 1 && !split && nchains * groups < slots) {
#define AVS_CHH(TR, M)                                                                                                 \
    hipLaunchKernelGGL((chain_update_kernel<TS, B, false, TR, M, 1>), dim3(2 * nchains * groups), dim3(UPD_THREADS), 0, \
                       s, (TS *)tiles, tjjs, ops, chain_start, chain0, groups, per)
#define AVS_CHH_M(TR)                                                                                                  \
    if (mode == MODE_GE) AVS_CHH(TR, MODE_GE);                                                                         \
    else if (mode == MODE_TS) AVS_CHH(TR, MODE_TS);                                                                    \
    else AVS_CHH(TR, MODE_TT)
            if (trans) { AVS_CHH_M(true); }
            else { AVS_CHH_M(false); }
#undef AVS_CHH_M
#undef AVS_CHH
            return;
        }
    }
#define AVS_CHU(SP, TR, M)                                                                                             \
    hipLaunchKernelGGL((chain_update_kernel<TS, B, SP, TR, M>), dim3(nchains * groups), dim3(UPD_THREADS), 0, s,      \
                       (TS *)tiles, tjjs, ops, chain_start, chain0, groups, per)
#define AVS_CHU_M(SP, TR)                                                                                              \
    if (mode == MODE_GE) AVS_CHU(SP, TR, MODE_GE);                                                                     \
    else if (mode == MODE_TS) AVS_CHU(SP, TR, MODE_TS);                                                                \
    else AVS_CHU(SP, TR, MODE_TT)
    if constexpr (sizeof(TS) == 4) {
        if (split) {
            if (trans) { AVS_CHU_M(true, true); }
            else { AVS_CHU_M(true, false); }
            return;
        }
    }
    if (trans) { AVS_CHU_M(false, true); }
    else { AVS_CHU_M(false, false); }
#undef AVS_CHU_M
#undef AVS_CHU
}

// ---------------------------------------------------------- batched form
// One grouped persistent launch for many independent matrices (possibly of
// different sizes, all with tile size B).  Each workgroup takes whole matrices
// from a queue (largest first) and runs the reference's flat-tree
// Blockbidiag! schedule on it (tiledalgos.jl:17-34) with the same tile
// operators as the single-matrix path; all reflector factors of the matrix
// being reduced live in a per-workgroup scratch slot.
// update of nc column tiles by a whole workgroup of NTHREADS, with the
// per-strip arithmetic of the single-matrix executors (so a matrix reduced in
// a batch gets the same bits as on its own)
template <typename TS, int B, int MODE>
__device__ __forceinline__ void batch_update(int split, const TS *V, bool vtr, const double *tjj, TS *C1, long s1,
                                             TS *C2, long s2, bool ctr, int nc, double *smem)
{
    if constexpr (sizeof(TS) == 4) {
        if (split) {
            update_body_split<B, MODE, true, NTHREADS>((const float *)V, B, vtr, tjj, (float *)C1, B, s1, (float *)C2,
                                                       B, s2, ctr, nc, smem);
            return;
        }
    }
    constexpr int CHUNK = (2 * NWAVES) / (B / 16); // column tiles per pass (16 strips)
    for (int c0 = 0; c0 < nc; c0 += CHUNK) {
        const int cnt = min(CHUNK, nc - c0);
        if (ctr)
            update2_body<TS, B, true, NTHREADS, true>(MODE, V, B, vtr, tjj, C1 + (long)c0 * s1, B, s1,
                                                      C2 + (long)c0 * s2, B, s2, cnt, smem);
        else
            update2_body<TS, B, true, NTHREADS, false>(MODE, V, B, vtr, tjj, C1 + (long)c0 * s1, B, s1,
                                                       C2 + (long)c0 * s2, B, s2, cnt, smem);
        __syncthreads();
    }
}

template <typename TS, int B>
__global__ void __launch_bounds__(NTHREADS, 1) batch_kernel(TS *arena, const long *__restrict__ off,
                                                             const int *__restrict__ nts,
                                                             const int *__restrict__ order, int count,
                                                             double *scratch, int *next, int split)
{
    __shared__ double smem[Lds<TS, B>::maxv];
    __shared__ int sh_m;
    const int t = threadIdx.x;
    const long tsz = (long)B * B;
    double *tau = scratch + (long)blockIdx.x * 17 * B;
    double *tjj = tau + B;
    for (;;) {
        if (t == 0) sh_m = atomicAdd(next, 1);
        __syncthreads();
        const int qi = sh_m;
        if (qi >= count) break;
        const int m = order[qi];
        const int nt = nts[m];
        TS *T = arena + off[m];
        auto tile = [&](int i, int j) { return T + ((long)j * nt + i) * tsz; };
        for (int k = 0; k < nt; ++k) {
            const int nc = nt - k - 1;
            // QR sweep on tile column k
            panel_body<TS, B, MODE_GE>(tile(k, k), B, tile(k, k), B, false, tau, tjj, smem);
            __syncthreads();
            if (nc > 0)
                batch_update<TS, B, MODE_GE>(split, tile(k, k), false, tjj, tile(k, k + 1), nt * tsz, tile(k, k + 1),
                                             nt * tsz, false, nc, smem);
            __syncthreads();
            for (int i = k + 1; i < nt; ++i) {
                panel_body<TS, B, MODE_TS>(tile(k, k), B, tile(i, k), B, false, tau, tjj, smem);
                __syncthreads();
                if (nc > 0)
                    batch_update<TS, B, MODE_TS>(split, tile(i, k), false, tjj, tile(k, k + 1), nt * tsz,
                                                 tile(i, k + 1), nt * tsz, false, nc, smem);
                __syncthreads();
            }
            if (k == nt - 1) break;
            // LQ sweep on tile row k from column k+1 (adjoint addressing)
            panel_body<TS, B, MODE_GE>(tile(k, k + 1), B, tile(k, k + 1), B, true, tau, tjj, smem);
            __syncthreads();
            batch_update<TS, B, MODE_GE>(split, tile(k, k + 1), true, tjj, tile(k + 1, k + 1), tsz, tile(k + 1, k + 1),
                                         tsz, true, nc, smem);
            __syncthreads();
            for (int j = k + 2; j < nt; ++j) {
                panel_body<TS, B, MODE_TS>(tile(k, k + 1), B, tile(k, j), B, true, tau, tjj, smem);
                __syncthreads();
                batch_update<TS, B, MODE_TS>(split, tile(k, j), true, tjj, tile(k + 1, k + 1), tsz, tile(k + 1, j),
                                             tsz, true, nc, smem);
                __syncthreads();
            }
        }
    }
}

// column-major padded staging (one N_m x N_m block per matrix, N_m = nt_m*B)
// -> tile-major arena; one block per tile of any matrix
template <typename TS>
__global__ void __launch_bounds__(256) batch_pack_kernel(const TS *__restrict__ stage, const long *__restrict__ soff,
                                                         TS *__restrict__ arena, const long *__restrict__ off,
                                                         const int *__restrict__ nts, const long *__restrict__ tile0,
                                                         int count, int b)
{
    const long g = blockIdx.x;
    int lo = 0, hi = count - 1;
    while (lo < hi) { // last matrix with tile0 <= g
        const int mid = (lo + hi + 1) / 2;
        if (tile0[mid] <= g) lo = mid; else hi = mid - 1;
    }
    const int m = lo, nt = nts[m];
    const long lt = g - tile0[m];
    const int i = (int)(lt % nt), j = (int)(lt / nt);
    const long N = (long)nt * b;
    const TS *src = stage + soff[m] + (long)i * b + (long)j * b * N;
    TS *dst = arena + off[m] + lt * b * b;
    for (int idx = threadIdx.x; idx < b * b; idx += blockDim.x) {
        const int r = idx % b, c = idx / b;
        dst[idx] = src[r + (long)c * N];
    }
}

// bands of all matrices: band m at boff[m], (b+1) x N_m, LAPACK upper-band layout
template <typename TS>
__global__ void __launch_bounds__(256) batch_band_kernel(const TS *__restrict__ arena, const long *__restrict__ off,
                                                         const int *__restrict__ nts, const long *__restrict__ col0,
                                                         double *__restrict__ bands, const long *__restrict__ boff,
                                                         int count, int b)
{
    const long g = blockIdx.x;
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (col0[mid] <= g) lo = mid; else hi = mid - 1;
    }
    const int m = lo, nt = nts[m];
    const long j = g - col0[m];
    const TS *tiles = arena + off[m];
    double *band = bands + boff[m];
    for (int d = threadIdx.x; d <= b; d += blockDim.x) {
        const long i = j - b + d;
        double v = 0.0;
        if (i >= 0) {
            const long I = i / b, J = j / b;
            v = (double)tiles[(J * nt + I) * b * b + (i % b) + (j % b) * (long)b];
        }
        band[d + j * (long)(b + 1)] = v;
    }
}

template <typename TS, int B>
void launch_batch(int workers, void *arena, const long *off, const int *nts, const int *order, int count,
                  double *scratch, int *next, int split, hipStream_t s)
{
    hipLaunchKernelGGL((batch_kernel<TS, B>), dim3(workers), dim3(NTHREADS), 0, s, (TS *)arena, off, nts, order, count,
                       scratch, next, split);
}
template <typename TS, int B> int batch_occupancy()
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_kernel<TS, B>, NTHREADS, 0) != hipSuccess) return 0;
    return nb;
}
template <typename TS>
void launch_batch_pack(const void *stage, const long *soff, void *arena, const long *off, const int *nts,
                       const long *tile0, long total_tiles, int count, int b, hipStream_t s)
{
    hipLaunchKernelGGL(batch_pack_kernel<TS>, dim3((unsigned)total_tiles), dim3(256), 0, s, (const TS *)stage, soff,
                       (TS *)arena, off, nts, tile0, count, b);
}
template <typename TS>
void launch_batch_band(const void *arena, const long *off, const int *nts, const long *col0, long total_cols,
                       double *bands, const long *boff, int count, int b, hipStream_t s)
{
    hipLaunchKernelGGL(batch_band_kernel<TS>, dim3((unsigned)total_cols), dim3(256), 0, s, (const TS *)arena, off, nts,
                       col0, bands, boff, count, b);
}

// ------------------------------------------------------------ debug probes
// One 16x16x4 MFMA per element type; lane l supplies a[l] and b[l] and the
// accumulator of every lane is written out so the host can read the lane layout.
__global__ void mfma_probe_kernel(const double *a, const double *b, double *out64, const float *af,
                                  const float *bf, float *out32, int lanes)
{
    const int l = threadIdx.x;
    if (lanes) { // lane-exchange check: v = a[l]; out64[4l + k] = v[l ^ (4 << k)]
        const double v = a[l];
        out64[l * 4 + 0] = lane_xor<4>(v);
        out64[l * 4 + 1] = lane_xor<8>(v);
        out64[l * 4 + 2] = lane_xor<16>(v);
        out64[l * 4 + 3] = lane_xor<32>(v);
        // out32[4l + k]: sum of v over the lanes congruent to l modulo 4 << k
        out32[l * 4 + 0] = (float)colsum<4>(v);
        out32[l * 4 + 1] = (float)colsum<8>(v);
        out32[l * 4 + 2] = (float)colsum<16>(v);
        out32[l * 4 + 3] = (float)colsum<32>(v);
        return;
    }
    d4_t acc = vzero4<d4_t>();
    acc = mfma16(a[l], b[l], acc);
    for (int m = 0; m < 4; ++m) out64[l * 4 + m] = acc[m];
    f4_t accf = vzero4<f4_t>();
    accf = mfma16(af[l], bf[l], accf);
    for (int m = 0; m < 4; ++m) out32[l * 4 + m] = accf[m];
}

// ------------------------------------------------------------ host launchers
template <typename TS> void launch_pack(const void *A, long lda, void *tiles, int nt, int b, hipStream_t s)
{
    hipLaunchKernelGGL(pack_kernel<TS>, dim3(nt, nt), dim3(256), 0, s, (const TS *)A, lda, (TS *)tiles, nt, b);
}
template <typename TS>
void launch_unpack(const void *tiles, void *A, long lda, int nt, int b, int mode, int kend, hipStream_t s)
{
    hipLaunchKernelGGL(unpack_kernel<TS>, dim3(nt, nt), dim3(256), 0, s, (const TS *)tiles, (TS *)A, lda, nt, b,
                       mode, kend);
}
template <typename TS> void launch_zero_band(void *tiles, int nt, int b, int kend, hipStream_t s)
{
    hipLaunchKernelGGL(zero_band_kernel<TS>, dim3(nt, nt), dim3(256), 0, s, (TS *)tiles, nt, b, kend);
}
template <typename TS> void launch_band(const void *tiles, double *band, long ldband, int nt, int b, hipStream_t s)
{
    hipLaunchKernelGGL(band_kernel<TS>, dim3((unsigned)((long)nt * b)), dim3(256), 0, s, (const TS *)tiles, band,
                       ldband, nt, b, (long)nt * b);
}

template <typename TS, int B>
void launch_panel(int mode, void *top, long ldt, void *bot, long ldb, int tr, double *tau, double *tjj, hipStream_t s)
{
    if constexpr (B == 128) {
        if (panel_blk_on()) {
            if (mode == MODE_GE)
                hipLaunchKernelGGL((panel_blk_kernel<TS, MODE_GE>), dim3(1), dim3(PBLK_THREADS), 0, s, (TS *)top, ldt,
                                   (TS *)top, ldt, tr, tau, tjj);
            else if (mode == MODE_TS)
                hipLaunchKernelGGL((panel_blk_kernel<TS, MODE_TS>), dim3(1), dim3(PBLK_THREADS), 0, s, (TS *)top, ldt,
                                   (TS *)bot, ldb, tr, tau, tjj);
            else
                hipLaunchKernelGGL((panel_blk_kernel<TS, MODE_TT>), dim3(1), dim3(PBLK_THREADS), 0, s, (TS *)top, ldt,
                                   (TS *)bot, ldb, tr, tau, tjj);
            return;
        }
    }
    if (mode == MODE_GE)
        hipLaunchKernelGGL((panel_kernel<TS, B, MODE_GE>), dim3(1), dim3(PanelThreads<B>::value), 0, s, (TS *)top, ldt, (TS *)top,
                           ldt, tr, tau, tjj);
    else if (mode == MODE_TS)
        hipLaunchKernelGGL((panel_kernel<TS, B, MODE_TS>), dim3(1), dim3(PanelThreads<B>::value), 0, s, (TS *)top, ldt, (TS *)bot,
                           ldb, tr, tau, tjj);
    else
        hipLaunchKernelGGL((panel_kernel<TS, B, MODE_TT>), dim3(1), dim3(PanelThreads<B>::value), 0, s, (TS *)top, ldt, (TS *)bot,
                           ldb, tr, tau, tjj);
}

template <typename TS, int B>
void launch_tjj(int mode, const void *V, long ldv, int vtr, const double *tau, double *tjj, hipStream_t s)
{
    if (mode == MODE_GE)
        hipLaunchKernelGGL((tjj_kernel<TS, B, MODE_GE>), dim3(1), dim3(NTHREADS), 0, s, (const TS *)V, ldv, vtr, tau,
                           tjj);
    else if (mode == MODE_TS)
        hipLaunchKernelGGL((tjj_kernel<TS, B, MODE_TS>), dim3(1), dim3(NTHREADS), 0, s, (const TS *)V, ldv, vtr, tau,
                           tjj);
    else
        hipLaunchKernelGGL((tjj_kernel<TS, B, MODE_TT>), dim3(1), dim3(NTHREADS), 0, s, (const TS *)V, ldv, vtr, tau,
                           tjj);
}

template <typename TS, int B, int MODE>
static void launch_update_m(bool applyt, const void *V, long ldv, int vtr, const double *tjj, void *C1, long ld1,
                            long s1, void *C2, long ld2, long s2, int ctr, int ntiles, int per_block, int split,
                            hipStream_t s)
{
    if (per_block > UpdTiles<B>::value || per_block < 1) per_block = UpdTiles<B>::value;
    const int nblk = (ntiles + per_block - 1) / per_block;
    if (nblk <= 0) return;
#define AVS_UPD(AT, SP)                                                                                                \
    hipLaunchKernelGGL((update_kernel<TS, B, MODE, AT, SP>), dim3(nblk), dim3(UPD_THREADS), 0, s, (const TS *)V, ldv,  \
                       vtr, tjj, (TS *)C1, ld1, s1, (TS *)C2, ld2, s2, ctr, ntiles, per_block)
    const bool sp = sizeof(TS) == 4 && split;
    if (applyt) {
        if (sp) AVS_UPD(true, true);
        else AVS_UPD(true, false);
    } else {
        if (sp) AVS_UPD(false, true);
        else AVS_UPD(false, false);
    }
#undef AVS_UPD
}

template <typename TS, int B>
void launch_update(int mode, bool applyt, const void *V, long ldv, int vtr, const double *tjj, void *C1, long ld1,
                   long s1, void *C2, long ld2, long s2, int ctr, int ntiles, int per_block, int split, hipStream_t s)
{
    if (mode == MODE_GE)
        launch_update_m<TS, B, MODE_GE>(applyt, V, ldv, vtr, tjj, C1, ld1, s1, C2, ld2, s2, ctr, ntiles, per_block,
                                        split, s);
    else if (mode == MODE_TS)
        launch_update_m<TS, B, MODE_TS>(applyt, V, ldv, vtr, tjj, C1, ld1, s1, C2, ld2, s2, ctr, ntiles, per_block,
                                        split, s);
    else
        launch_update_m<TS, B, MODE_TT>(applyt, V, ldv, vtr, tjj, C1, ld1, s1, C2, ld2, s2, ctr, ntiles, per_block,
                                        split, s);
}

template <typename TS, int B>
void launch_persistent(int workers, void *tiles, double *taus, double *tjjs, const Task *tasks, int ntasks,
                       const int *dep_tile, const int *dep_val, const int *sig_tile, int *cnt, int *ticket,
                       long long wait_ticks, int *status, int split, hipStream_t s)
{
    hipLaunchKernelGGL((persistent_kernel<TS, B>), dim3(workers), dim3(NTHREADS), 0, s, (TS *)tiles, taus, tjjs, tasks,
                       ntasks, dep_tile, dep_val, sig_tile, cnt, ticket, wait_ticks, status, split);
}

template <typename TS, int B> int persistent_occupancy()
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, persistent_kernel<TS, B>, NTHREADS, 0) != hipSuccess)
        return 0;
    return nb;
}

void launch_mfma_probe(const double *a, const double *b, double *o64, const float *af, const float *bf, float *o32,
                       int lanes, hipStream_t s)
{
    hipLaunchKernelGGL(mfma_probe_kernel, dim3(1), dim3(64), 0, s, a, b, o64, af, bf, o32, lanes);
}

#define AVS_INST_B(TS, B)                                                                                             \
    template void launch_panel<TS, B>(int, void *, long, void *, long, int, double *, double *, hipStream_t);         \
    template void launch_tjj<TS, B>(int, const void *, long, int, const double *, double *, hipStream_t);             \
    template void launch_update<TS, B>(int, bool, const void *, long, int, const double *, void *, long, long,        \
                                       void *, long, long, int, int, int, int, hipStream_t);                           \
    template void launch_persistent<TS, B>(int, void *, double *, double *, const Task *, int, const int *,           \
                                           const int *, const int *, int *, int *, long long, int *, int,             \
                                           hipStream_t);                                                               \
    template int persistent_occupancy<TS, B>();                                                                        \
    template void launch_batch<TS, B>(int, void *, const long *, const int *, const int *, int, double *, int *, int, \
                                      hipStream_t);                                                                    \
    template int batch_occupancy<TS, B>();                                                                             \
    template void launch_chain_panel<TS, B>(int, int, void *, double *, double *, const ChainOp *, const int *, int,   \
                                            hipStream_t);                                                              \
    template void launch_chain_update<TS, B>(int, int, int, int, int, void *, const double *, const ChainOp *,         \
                                             const int *, int, int, hipStream_t);

AVS_INST_B(double, 32)
AVS_INST_B(double, 64)
AVS_INST_B(double, 128)
AVS_INST_B(float, 32)
AVS_INST_B(float, 64)
AVS_INST_B(float, 128)

template void launch_pack<double>(const void *, long, void *, int, int, hipStream_t);
template void launch_pack<float>(const void *, long, void *, int, int, hipStream_t);
template void launch_unpack<double>(const void *, void *, long, int, int, int, int, hipStream_t);
template void launch_unpack<float>(const void *, void *, long, int, int, int, int, hipStream_t);
template void launch_band<double>(const void *, double *, long, int, int, hipStream_t);
template void launch_zero_band<double>(void *, int, int, int, hipStream_t);
template void launch_zero_band<float>(void *, int, int, int, hipStream_t);
template void launch_batch_pack<double>(const void *, const long *, void *, const long *, const int *, const long *,
                                        long, int, int, hipStream_t);
template void launch_batch_pack<float>(const void *, const long *, void *, const long *, const int *, const long *,
                                       long, int, int, hipStream_t);
template void launch_batch_band<double>(const void *, const long *, const int *, const long *, long, double *,
                                        const long *, int, int, hipStream_t);
template void launch_batch_band<float>(const void *, const long *, const int *, const long *, long, double *,
                                       const long *, int, int, hipStream_t);
template void launch_band<float>(const void *, double *, long, int, int, hipStream_t);

} // namespace avs
