// kernels.hpp -- host-side declarations of the kernel launchers (kernels.hip).
#pragma once
#include <hip/hip_runtime_api.h>
#include <stdint.h>

namespace avs {

struct Task;
struct ChainOp;

// T_JJ workspace of the allocation-free TSMQR: 16 * b doubles for b <= 128
constexpr int NOALLOC_SCRATCH_DOUBLES = 16 * 128;
// device address of `device`'s copy of that workspace (noalloc.hip); looked up
// on the first call for each device, with that device current
hipError_t noalloc_scratch(int device, double **ptr);

template <typename TS> void launch_pack(const void *A, long lda, void *tiles, int nt, int b, hipStream_t s);
template <typename TS>
void launch_unpack(const void *tiles, void *A, long lda, int nt, int b, int mode, int kend, hipStream_t s);
template <typename TS> void launch_band(const void *tiles, double *band, long ldband, int nt, int b, hipStream_t s);
template <typename TS> void launch_zero_band(void *tiles, int nt, int b, int kend, hipStream_t s);

template <typename TS, int B>
void launch_panel(int mode, void *top, long ldt, void *bot, long ldb, int tr, double *tau, double *tjj,
                  hipStream_t s);
template <typename TS, int B>
void launch_tjj(int mode, const void *V, long ldv, int vtr, const double *tau, double *tjj, hipStream_t s);
template <typename TS, int B>
void launch_update(int mode, bool applyt, const void *V, long ldv, int vtr, const double *tjj, void *C1, long ld1,
                   long s1, void *C2, long ld2, long s2, int ctr, int ntiles, int per_block, int split,
                   hipStream_t s);
template <typename TS, int B>
void launch_persistent(int workers, void *tiles, double *taus, double *tjjs, const Task *tasks, int ntasks,
                       const int *dep_tile, const int *dep_val, const int *sig_tile, int *cnt, int *ticket,
                       long long wait_ticks, int *status, int split, hipStream_t s);
template <typename TS, int B> int persistent_occupancy();

template <typename TS, int B>
void launch_batch(int workers, void *arena, const long *off, const int *nts, const int *order, int count,
                  double *scratch, int *next, int split, hipStream_t s);
template <typename TS, int B> int batch_occupancy();
template <typename TS>
void launch_batch_pack(const void *stage, const long *soff, void *arena, const long *off, const int *nts,
                       const long *tile0, long total_tiles, int count, int b, hipStream_t s);
template <typename TS>
void launch_batch_band(const void *arena, const long *off, const int *nts, const long *col0, long total_cols,
                       double *bands, const long *boff, int count, int b, hipStream_t s);

// band -> bidiagonal on the device, one wavefront per matrix (chase.hip); b <= 32
template <typename TS>
void launch_batch_chase(const void *arena, const long *off, const int *nts, const int *order, const long *col0,
                        int count, int b, double *store, double *de, int *err, hipStream_t s);

// band -> bidiagonal of one matrix in a band store (ld 3b+1, rows c-2b .. c+b
// of column c, chase.hip), b <= 128: the steps of sweeps ilo .. ihi at time
// tstep (step tstep - 3i of sweep i), one workgroup each.  The store's
// allocation holds N + CHASE_SPARE_COLS columns (the spare ones are read, never
// used or written).
constexpr long CHASE_SPARE_COLS = 128;
void launch_chase_step(double *store, long N, int b, long tstep, long ilo, long ihi, int *err, hipStream_t s);
// de[0 .. N) = diagonal, de[N .. 2N-1) = superdiagonal of the store
void launch_chase_diag(const double *store, long N, int b, double *de, hipStream_t s);

// singular values of the bidiagonals (de as launch_batch_chase writes it) by
// bisection, one lane per value; workgroup g: matrix wmat[g], ascending values
// wfirst[g] .. +255; 2N-1 <= 4096; sigma[soff[m] ..] gets the n_m largest
constexpr int BATCH_BISECT_THREADS = 256;
void launch_batch_bisect(const double *de, const long *col0, const int *nts, int b, const int *wmat,
                         const int *wfirst, int nwg, const long *nreal, const long *soff, double *sigma,
                         hipStream_t s);

// one bidiagonal: t2 = squared off-diagonals of its Golub-Kahan tridiagonal
// (2N - 1 values); sigma[0 .. nkeep) gets the nkeep largest, descending
void launch_bisect(const double *t2, long N, long nkeep, double hi0, double pivmin, double *sigma, hipStream_t s);

template <typename TS, int B>
void launch_chain_panel(int nchains, int mode, void *tiles, double *taus, double *tjjs, const ChainOp *ops,
                        const int *chain_start, int chain0, hipStream_t s);
template <typename TS, int B>
void launch_chain_update(int nchains, int groups, int per, int mode, int trans, void *tiles, const double *tjjs,
                         const ChainOp *ops, const int *chain_start, int chain0, int split, hipStream_t s);

void launch_mfma_probe(const double *a, const double *b, double *o64, const float *af, const float *bf, float *o32,
                       int lanes, hipStream_t s);

} // namespace avs
