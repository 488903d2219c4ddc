/*
 * avsvd.h -- C ABI of the MI355X tiled SVD band-reduction engine (libavsvd.so).
 *
 * Drop-in boundary for the reference's stage-1 path (paths relative to the
 * reference repository root):
 *   TiledMatrix(A, b, b; tiledim2d)        KAbasedSVD/src/TiledMatrix.jl:41-67
 *   tiled_QR!(A; kend)                     KAbasedSVD/src/tiledalgos.jl:6-15
 *   Blockbidiag!(A; kend)                  KAbasedSVD/src/tiledalgos.jl:17-34
 *   OOC_Bidiag!/OOC_SVD! (in-core branch)  KAbasedSVD/src/tiledalgos.jl:232-260
 *   bidiagonalize / diag_lapack / my_CU_svdval
 *                                          KAbasedSVD/src/bulgechasing.jl:33-53
 *   tile kernels GEQRT/TSQRT/UNMQR/TSMQR   KAbasedSVD/src/QRkernels.jl:3-197
 * Julia binds these with ccall (see avocado-sandwich_amd/julia/AvSVD.jl).
 *
 * Conventions
 *  - Every function returns an avs_status: 0 = success, -i = argument i is
 *    invalid, > 0 = runtime/device error (AVS_ERR_*).  No exceptions cross the
 *    ABI.  avs_last_error() returns a message for the calling thread.
 *  - Matrices are column-major with a leading dimension, like Julia arrays.
 *  - A handle is bound to one HIP stream (avs_opts.stream, 0 = a stream the
 *    handle creates).  Device work is asynchronous unless a function returns
 *    host data.  Calls on one handle are not thread-safe; distinct handles are.
 *  - Pointers named d* are device pointers, all others host pointers.
 */
#ifndef AVSVD_H
#define AVSVD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t avs_status;

enum {
    AVS_OK = 0,
    AVS_ERR_HIP = 1,        /* a HIP runtime call failed (message has the code) */
    AVS_ERR_ALLOC = 2,      /* device or host allocation failed */
    AVS_ERR_STATE = 3,      /* call out of order (e.g. no matrix set) */
    AVS_ERR_LAPACK = 4,     /* host LAPACK not found or returned info != 0 */
    AVS_ERR_UNSUPPORTED = 5,/* shape/tile/dtype combination not built */
    AVS_ERR_BUDGET = 6      /* request exceeds the configured memory budget */
};

typedef enum { AVS_F64 = 0, AVS_F32 = 1 } avs_dtype;

/* How the tile DAG is executed. */
typedef enum {
    AVS_SCHED_PERSISTENT = 0, /* one device-side dependency-counting persistent launch (disabled:
                                 returns AVS_ERR_UNSUPPORTED; its launch did not complete) */
    AVS_SCHED_HOST = 1,       /* host-enqueued launches in the reference order (plain path) */
    AVS_SCHED_STREAMS = 2     /* fused per-domain chain kernels, stream-ordered */
} avs_sched;

/* Reduction tree of each QR/LQ panel sweep.  FLAT is the reference's
 * (TSQRT chain on the diagonal R, tiledalgos.jl:19-32); HIER splits the
 * panel into domains of `domain` tiles, each a flat TS chain, whose R
 * factors are then merged by triangle-on-triangle steps.  Both compute an
 * orthogonal two-sided reduction with the same singular values. */
typedef enum { AVS_TREE_FLAT = 0, AVS_TREE_HIER = 1 } avs_tree;

/* Arithmetic of the FP32 trailing updates (FP64 handles ignore it). */
typedef enum {
    AVS_UPD_NATIVE = 0,     /* native FP32 MFMA, FP32 accumulate */
    AVS_UPD_SPLIT_BF16 = 1  /* 3-term split-BF16 MFMA products, FP32 accumulate */
} avs_update_mode;

typedef struct avs_opts {
    void *stream;          /* hipStream_t; 0 = create one */
    int32_t sched;         /* avs_sched */
    int32_t tree;          /* avs_tree */
    int32_t domain;        /* tiles per domain for AVS_TREE_HIER (0 = automatic per-sweep rule) */
    int32_t update_mode;   /* avs_update_mode */
    int32_t workers;       /* persistent workgroups (0 = one per CU) */
    int32_t wait_ms;       /* bound in ms on any one dependency wait of the persistent
                              executor (0 = 30000); exceeding it ends the launch and the
                              next sync returns AVS_ERR_STATE */
    int32_t phase_timing;  /* != 0: AVS_SCHED_STREAMS brackets every launch with timing
                              events for avs_phase_timing / avs_phase_trace */
    int32_t single_stream; /* != 0: AVS_SCHED_STREAMS runs the panels on the handle's
                              stream too (by default they run on a second stream of the
                              handle, ahead of the same sweep's updates) */
    int32_t reserve_cus;   /* K > 0: AVS_SCHED_STREAMS runs the updates on a stream whose
                              CU mask leaves K CUs (the same number on every XCD;
                              K <= CUs/8 - CUs/32) to the panels */
    int32_t graph;         /* != 0: AVS_SCHED_STREAMS captures the first run of a handle's
                              phase program as a HIP graph and replays it on every run
                              (not with phase_timing, single_stream or reserve_cus); by
                              default every launch is enqueued each run (faster on MI355X) */
    int32_t masked_modes;  /* with reserve_cus > 0: bit mask of the update modes (1 GEQRT's,
                              2 TSQRT's, 4 the tree merges') that run on the CU-masked
                              stream; the others run on the handle's stream; 0 = all */
    int32_t reserved[3];   /* must be 0 (avs_create returns -5 otherwise) */
} avs_opts;

typedef struct avs_handle avs_handle;

/* Default options. */
void avs_default_opts(avs_opts *opts);

/* TiledMatrix(A, b, b) for an n x n matrix: allocates the tile-major working
 * copy, tau/T factor storage and dependency counters.  Errors like the
 * reference's "Not supported" (TiledMatrix.jl:42-44): n % b != 0 -> -3,
 * unsupported b -> -4; nonzero opts->reserved -> -5. */
avs_status avs_create(avs_handle **h, int32_t dtype, int64_t n, int32_t b, const avs_opts *opts);
avs_status avs_destroy(avs_handle *h);

/* Load the matrix (element type = the handle's dtype), column-major, lda >= n.
 * The copy runs on the handle's stream, which does not wait for other
 * streams: a device matrix must be complete when the call is made (or be
 * produced on the stream given in avs_opts.stream). */
avs_status avs_set_matrix_host(avs_handle *h, const void *A, int64_t lda);
avs_status avs_set_matrix_device(avs_handle *h, const void *dA, int64_t lda);
/* Load an m x m matrix (m <= n) into the leading block and zero the rest:
 * the zero padding only adds n - m zero singular values.  For sizes that are
 * not multiples of b (the reference rejects them, TiledMatrix.jl:42-44). */
avs_status avs_set_matrix_host_sub(avs_handle *h, const void *A, int64_t m, int64_t lda);

/* Read back the reduced n x n matrix in the reference's in-place layout:
 * after Blockbidiag!, the band (upper-triangular diagonal tiles, lower-
 * triangular super-diagonal tiles) and exact zeros elsewhere; after
 * tiled_QR!, R in the upper triangle and the reflectors below it. */
avs_status avs_get_matrix_host(avs_handle *h, void *A, int64_t lda);
avs_status avs_get_matrix_device(avs_handle *h, void *dA, int64_t lda);

/* Blockbidiag!(A; kend): two-sided QR/LQ reduction to upper band form of
 * bandwidth b.  kend <= 0 means all n/b steps. */
avs_status avs_blockbidiag(avs_handle *h, int32_t kend);
/* tiled_QR!(A; kend). */
avs_status avs_tiled_qr(avs_handle *h, int32_t kend);
/* myormqr!(C, A, Tau) (CAQR/src/qr_kernels.jl:256-265): C <- Q^T C (applyt != 0) or
 * Q C, with Q the orthogonal factor of the handle's last avs_tiled_qr (its reflectors and
 * compact-WY factors stay in the handle until the next run), on the n x ncols column-major
 * device matrix dC (leading dimension ldc >= n, ncols a multiple of b, the handle's dtype),
 * enqueued on the handle's stream.  Errors: -2 null dC, -3 ldc < n, -4 ncols,
 * AVS_ERR_STATE when the handle holds no tiled_QR! factors. */
avs_status avs_apply_q(avs_handle *h, void *dC, int64_t ldc, int64_t ncols, int32_t applyt);

/* ---- out-of-core runs on a host matrix (OOC_alg!, KAbasedSVD/src/tiledalgos.jl:232-251) ----
 * avs_ooc_blockbidiag is OOC_Bidiag! (:254), avs_ooc_tiled_qr is OOC_QR! (:253): the
 * column-major n x n host matrix A (leading dimension lda) is reduced in place.  When
 * n > kswitch * b the first n/b - kswitch steps stream A through the GPU as the reference's
 * LargeTiledMatrix does (KAbasedSVD/src/TiledMatrix.jl:121-186, tiledalgos.jl:75-225): the
 * flat tree, one tile row (QR sweep) or tile column (LQ sweep) at a time against the
 * sweep's pivot row or column, so the device holds 3 (n/b - k) tiles in step k; the
 * trailing kswitch x kswitch tiles then run in core with `opts` (NULL = defaults).
 * Otherwise the whole matrix runs in core.  The result is what avs_get_matrix_host returns
 * after the same algorithm in core (the band with exact zeros outside it, or the raw R and
 * reflectors of tiled_QR!); with opts->tree = AVS_TREE_FLAT it is the same bits.
 * Errors: -1 dtype, -2 n not a positive multiple of b ("Not supported"), -3 b not built,
 * -4 null A, -5 lda < n, -6 kswitch < 0; device errors as avs_blockbidiag. */
avs_status avs_ooc_blockbidiag(int32_t dtype, int64_t n, int32_t b, void *A, int64_t lda, int32_t kswitch,
                               const avs_opts *opts);
avs_status avs_ooc_tiled_qr(int32_t dtype, int64_t n, int32_t b, void *A, int64_t lda, int32_t kswitch,
                            const avs_opts *opts);

/* Per-launch device time of the last avs_blockbidiag / avs_tiled_qr run under
 * AVS_SCHED_STREAMS, if that run recorded per-launch events (opts.phase_timing
 * != 0, or avs_set_phase_timing before it; AVS_ERR_STATE otherwise) (one event
 * pair per launch, on the stream the launch runs on; with the default second
 * panel stream, panel and update launches overlap in time, so the two sums can
 * exceed the run's wall time): summed milliseconds and launch counts of the panel and
 * update kernels, and the algorithmic flops of the update launches (4b^3 per
 * TS-operation column tile, 2b^3 per GE/TT-operation column tile).  Waits for
 * the run.  No counterpart in the reference (its benchmarks time whole calls). */
avs_status avs_phase_timing(avs_handle *h, double *panel_ms, int64_t *panel_launches, double *update_ms,
                            int64_t *update_launches, double *update_flops);
/* Turn the per-launch timing events of opts.phase_timing on (on != 0) or off
 * for the handle's next runs.  The events cost launch latency on the panel
 * chain (n=16384 FP64 b=128: 374 vs 356 ms per reduction), so a caller that
 * times several runs can record them in one run only.  No counterpart in the
 * reference. */
avs_status avs_set_phase_timing(avs_handle *h, int32_t on);
/* Per-launch timeline of the last stream-ordered run, if it recorded per-launch
 * events (as for avs_phase_timing):
 * for launch i, start and end in ms from the first launch's start event, and
 * kind*4 + mode (kind 0 panel, 1 update; mode 0 GE, 1 TS, 2 TT).  With null
 * arrays only *count is set.  No counterpart in the reference. */
avs_status avs_phase_trace(avs_handle *h, double *start_ms, double *end_ms, int32_t *kind_mode, int64_t cap,
                           int64_t *count);
/* Copy the band to the host in FP64 LAPACK upper-band layout:
 * band[(b + i - j) + j*ldband] = A(i, j) for max(0, j-b) <= i <= j,
 * ldband >= b+1.  Only (b+1) x n values cross PCIe, not n x n. */
avs_status avs_get_band(avs_handle *h, double *band, int64_t ldband);

/* Full svdvals (OOC_SVD! + diag_lapack, KAbasedSVD/src/tiledalgos.jl:256-260,
 * bulgechasing.jl:50-53): sigma[n], descending, FP64.
 *  1. Band reduction of the handle's matrix if it has not run to the end (after a
 *     partial avs_blockbidiag the remaining steps run on the band-zeroed matrix, as
 *     Blockbidiag! on the reference's array would).
 *  2. Band -> bidiagonal in FP64 by Householder bulge chasing: on the device
 *     (avs_bidiagonalize_band_device, no download of the band) for
 *     n >= AVSVD_DEVICE_CHASE_MIN_N (default 8192 for b > 64, else 16384) and
 *     2 <= b <= 128, otherwise on
 *     the host after avs_get_band (pipelined over the library's host threads).
 *     AVSVD_STAGE2=device | host force either side; AVSVD_STAGE2=halving runs the
 *     reference's block sweeps (bulgechasing.jl:33-42) and =dgbbrd LAPACK's chase.
 *  3. Singular values of the bidiagonal: Sturm-count bisection on the device for
 *     n >= 128, LAPACK dbdsdc('U','N') on the host below that or with AVSVD_BDSDC=host.
 * Errors: AVS_ERR_STATE if no matrix is set or the handle holds tiled_QR! factors
 * (set the matrix again first); AVS_ERR_LAPACK if the host LAPACK is needed and
 * missing. */
avs_status avs_svdvals(avs_handle *h, double *sigma);

avs_status avs_synchronize(avs_handle *h);

/* Device time of the last avs_blockbidiag/avs_tiled_qr, in milliseconds,
 * measured with HIP events on the handle's stream (waits for completion). */
avs_status avs_last_time_ms(avs_handle *h, double *ms);

const char *avs_last_error(void);
/* gfx architecture of HIP device `device` without target features (e.g.
 * "gfx950"), NUL-terminated in name[len] (the autotuner's table key).  No
 * reference counterpart.  Errors: -2 null name, -3 len < 2, AVS_ERR_HIP when
 * the device cannot be queried. */
avs_status avs_device_arch(int32_t device, char *name, int64_t len);

/* ---- host stage (kept on the host in FP64, as the reference runs it) ---- */

/* bidiagonalize(A, bw), KAbasedSVD/src/bulgechasing.jl:33-42, on an upper
 * band (layout as avs_get_band, bandwidth bw, ldband >= bw+1).  Outputs the
 * diagonal d[n] and super-diagonal e[n-1] of the bidiagonal result. */
avs_status avs_bidiagonalize_band(int64_t n, int32_t bw, const double *band, int64_t ldband,
                                  double *d, double *e);
/* The same reduction with LAPACK dgbbrd (Givens rotations) on one thread: no
 * counterpart in the reference.  The batched and served forms use it for each
 * matrix they give one host thread (AVSVD_STAGE2=halving selects the
 * reference's block sweeps instead). */
avs_status avs_bidiagonalize_band_lapack(int64_t n, int32_t bw, const double *band, int64_t ldband,
                                         double *d, double *e);
/* avs_bidiagonalize_band_chase: the same reduction by Householder bulge
 * chasing (one length-bw reflector per annihilated row or column segment,
 * Lang's order) on a compact band store with unit-stride inner loops, the
 * sweeps pipelined over the library's host threads (AVSVD_HOST_THREADS,
 * default min(16, cores)); the result does not depend on the thread count.
 * No reference counterpart (the reference halves the bandwidth with block
 * sweeps, KAbasedSVD/src/bulgechasing.jl:33-42); singular values are
 * unchanged, d and e may differ in sign from the other two routines.
 * Argument errors as avs_bidiagonalize_band. */
avs_status avs_bidiagonalize_band_chase(int64_t n, int32_t bw, const double *band, int64_t ldband, double *d,
                                        double *e);
/* avs_bidiagonalize_band_device: the same chase on the calling thread's
 * current GPU, bw <= 128.  The band goes to a band store in HBM, and the
 * steps of many sweeps run at once: step s of sweep i at time 3i + s, one
 * launch per time, one workgroup per step (the steps of one time touch
 * disjoint rows or columns, so every element sees its updates in sweep
 * order, as on the host); only d and e come back.  Same rounding class as the
 * host chase, not the same bits.  avs_svdvals runs this stage on the device
 * for n >= AVSVD_DEVICE_CHASE_MIN_N (default 8192 for b > 64, else 16384:
 * where it matches or beats the host chase on 16 threads) and 2 <= b <= 128,
 * or always with
 * AVSVD_STAGE2=device; AVSVD_STAGE2=host (or halving, dgbbrd) keeps it on the
 * host.  No reference counterpart.  Argument errors as
 * avs_bidiagonalize_band; -2 also for bw > 128. */
avs_status avs_bidiagonalize_band_device(int64_t n, int32_t bw, const double *band, int64_t ldband, double *d,
                                         double *e);
/* diag_lapack, KAbasedSVD/src/bulgechasing.jl:50-53: dbdsdc('U','N').
 * sigma[n] descending.  LAPACK is loaded at run time (AVSVD_LAPACK or the
 * OpenBLAS bundled with scipy). */
avs_status avs_bdsdc_values(int64_t n, const double *d, const double *e, double *sigma);

/* ---- tile operator API (mirrors KAbasedSVD/src/QRkernels.jl kernels) ----
 * One b x b tile per operand, element type dtype, device pointers with
 * leading dimensions.  tau is a device array of b doubles.  `trans` runs the
 * operator on the adjoint of every tile operand (the reference's LQ twins,
 * TiledMatrix.jl:85-88).  stream may be 0 (the null stream). */
/* QR_unsafe_kernel! (QRkernels.jl:3-66): in-place GEQRT, tau[0..b-2]. */
avs_status avs_geqrt(int32_t dtype, int32_t b, void *dA, int64_t lda, double *dtau,
                     int32_t trans, void *stream);
/* QR_unsafe_kernel2! (QRkernels.jl:71-129): TSQRT of [triu(A1); A2]. */
avs_status avs_tsqrt(int32_t dtype, int32_t b, void *dA1, int64_t lda1, void *dA2, int64_t lda2,
                     double *dtau, int32_t trans, void *stream);
/* applyQorQt_unsafe_kernel! (QRkernels.jl:131-161): C <- Q^T C (applyt) or
 * Q C for the GEQRT factors in dV; C is b x ncols (ncols % b == 0).
 * ctrans applies to C' (the reference's adjoint-view calls), vtrans reads V'. */
avs_status avs_unmqr(int32_t dtype, int32_t b, const void *dV, int64_t ldv, const double *dtau,
                     void *dC, int64_t ldc, int64_t ncols, int32_t applyt, int32_t vtrans,
                     int32_t ctrans, void *stream);
/* applyQorQt_unsafe_kernel2! (QRkernels.jl:165-197): [C1; C2] <- Q^T [C1; C2]
 * (applyt) or Q [C1; C2] for the TSQRT reflectors in dV. */
avs_status avs_tsmqr(int32_t dtype, int32_t b, void *dC1, int64_t ldc1, void *dC2, int64_t ldc2,
                     const void *dV, int64_t ldv, const double *dtau, int64_t ncols,
                     int32_t applyt, int32_t vtrans, int32_t ctrans, void *stream);
/* avs_tsmqr_noalloc: avs_tsmqr that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The T_JJ factor that avs_tsmqr builds in a buffer of its own goes to a
 * fixed per-GPU workspace of 16 * 128 doubles inside the library, which each
 * GPU holds until the process ends; its address is looked up on the first
 * call on each device.  b beyond that workspace (b > 128) is refused with -2.
 * Every call on one GPU shares the workspace: calls that overlap on one GPU
 * (on different streams, or from host threads without ordering between them)
 * race; calls ordered on one stream do not.  It does not synchronize: the
 * result is ready when `stream` has passed this call.  The same kernels and
 * results as avs_tsmqr; arguments and other errors as avs_tsmqr. */
avs_status avs_tsmqr_noalloc(int32_t dtype, int32_t b, void *dC1, int64_t ldc1, void *dC2, int64_t ldc2,
                             const void *dV, int64_t ldv, const double *dtau, int64_t ncols,
                             int32_t applyt, int32_t vtrans, int32_t ctrans, void *stream);

/* ---- seeded synthetic inputs (both sides of every comparison) ----
 * Philox4x32-10 + Box-Muller, element e of the column-major matrix from
 * counter e/2 and key = seed; entries N(0, scale^2). */
avs_status avs_gen_gaussian_host(uint64_t seed, int64_t n, double scale, int32_t dtype, void *A,
                                 int64_t lda);

/* ---- batched form ----
 * Many independent square matrices of mixed size n_i (each zero-padded to a
 * multiple of b; padding only adds zero singular values, which are dropped),
 * band-reduced in one grouped persistent launch: every workgroup takes whole
 * matrices from a queue, largest first, and runs the reference's Blockbidiag!
 * schedule on them.  The host stage then runs on a thread pool, one matrix
 * per thread.  No reference counterpart (the reference reduces one matrix per
 * call); it serves the batched svdvals workload. */
typedef struct avs_batch avs_batch;
avs_status avs_batch_create(avs_batch **batch, int32_t dtype, int32_t count, const int64_t *ns, int32_t b,
                            const avs_opts *opts);
avs_status avs_batch_destroy(avs_batch *batch);
/* Stage matrix idx (n_idx x n_idx, column-major, the batch dtype) for upload. */
avs_status avs_batch_set_matrix_host(avs_batch *batch, int32_t idx, const void *A, int64_t lda);
/* One grouped launch over all matrices (re-packs the staged inputs first). */
avs_status avs_batch_blockbidiag(avs_batch *batch);
avs_status avs_batch_last_time_ms(avs_batch *batch, double *ms);
/* Band of matrix idx, FP64 upper-band layout, ldband >= b+1, N_idx columns
 * (N_idx = n_idx rounded up to a multiple of b). */
avs_status avs_batch_get_band(avs_batch *batch, int32_t idx, double *band, int64_t ldband);
/* All singular values, concatenated in matrix order (n_i values each,
 * descending); runs the reduction first if needed.  For b <= 32 the band ->
 * bidiagonal step runs on the device (one wavefront per matrix, Householder
 * bulge chasing in FP64), and so do the bidiagonals' singular values
 * (bisection on the Golub-Kahan tridiagonal, one lane per value) when
 * 2N - 1 <= 4096; only sigma comes back.  AVSVD_BDSDC=host computes them with
 * LAPACK dbdsdc on the host; AVSVD_STAGE2=host (or dgbbrd, halving) runs both
 * steps on the host.  AVS_ERR_STATE if the device chase leaves its band store
 * (bounds-checked). */
avs_status avs_batch_svdvals(avs_batch *batch, double *sigma);

/* ---- serving: continuous-batching svdvals request loop ----
 * Requests of any n (FP32 or FP64) are accepted as they arrive and admitted
 * to the device first-come first-served while the HBM budget has room for
 * their working set; up to max_inflight band reductions run concurrently, each
 * on the stream of its handle.  A request is zero-padded to its size bucket
 * (multiples of b, eight sizes per octave; the padding adds zero singular
 * values, which are dropped), and runs on a pooled handle of that (dtype,
 * bucket, b): handles, their device buffers and a pinned host staging buffer
 * are created on first use, kept while the budget allows (least recently used
 * ones are released for new shapes) and freed by avs_server_destroy.  One
 * device thread makes every HIP call of the loop; the host_threads workers do
 * the CPU work (band -> bidiagonal, and dbdsdc without device_sigma).  Only
 * the band crosses PCIe back.  Latency = submit -> singular values ready.
 * AVSVD_SERVER_TRACE=1 prints one line per finished request (queueing,
 * set-up, device and host-stage times) on stderr.  No reference counterpart
 * (the reference has no serving path). */
typedef struct avs_server avs_server;
typedef struct avs_server_opts {
    int64_t hbm_budget_bytes; /* device working-set budget for admitted requests */
    int32_t max_inflight;     /* concurrent device reductions (streams) */
    int32_t host_threads;     /* host-stage workers */
    int32_t b_f32, b_f64;     /* tile sizes (0 = the tile-size table below) */
    int32_t update_mode;      /* avs_update_mode for FP32 requests */
    int32_t device_sigma;     /* != 0: the bidiagonal's singular values by bisection on the
                                 device, run by the device thread on max_inflight streams with
                                 buffers allocated at server creation; 0: LAPACK dbdsdc on the
                                 host worker */
    int32_t reserved[6];      /* must be 0 */
} avs_server_opts;
void avs_server_default_opts(avs_server_opts *opts);
avs_status avs_server_create(avs_server **server, const avs_server_opts *opts);
/* Copies A (n x n, column-major, dtype) and queues it; returns its id. */
avs_status avs_server_submit(avs_server *server, int32_t dtype, int64_t n, const void *A, int64_t lda,
                             uint64_t *req_id);
avs_status avs_server_poll(avs_server *server, uint64_t req_id, int32_t *done);
/* Blocks until the request is done; sigma gets n values (descending). */
avs_status avs_server_wait(avs_server *server, uint64_t req_id, double *sigma, double *latency_ms);
avs_status avs_server_stats(avs_server *server, double *p50_ms, double *p99_ms, double *solves_per_s,
                            int64_t *completed);
avs_status avs_server_destroy(avs_server *server);

/* ---- tile-size table (filled by the autotuner) ----
 * Per dtype, entries (n_upto, b): a served request of order n whose tile size
 * is not fixed by avs_server_opts takes the b of the first entry with
 * n <= n_upto, or of the last entry when n exceeds them all; with no entries
 * for its dtype, b = 32 up to n = 2048, else 64.  avs_tune_set replaces the
 * entry with the same n_upto.  The table is process-wide and thread-safe.  No
 * reference counterpart: the reference fixes its tile size by hand before
 * compiling its kernels (KAbasedSVD/src/QRkernelscompile.jl:1-4,
 * tiledalgos.jl:262-267); here every tile size is built ahead of time and the
 * choice is a measured table.  Errors: -1 dtype, -2 n_upto or n < 1, -3 tile
 * size not in {32, 64, 128} (avs_tune_set) or null b (avs_tune_lookup). */
avs_status avs_tune_set(int32_t dtype, int64_t n_upto, int32_t b);
avs_status avs_tune_clear(void);
avs_status avs_tune_lookup(int32_t dtype, int64_t n, int32_t *b);

/* ---- diagnostics ----
 * avs_debug_mfma_layout: runs one v_mfma_f64_16x16x4f64 and one
 * v_mfma_f32_16x16x4f32 in which lane l supplies a_l as its A operand and b_l
 * as its B operand (variant 0: a_l = l+1, b_l = (7l mod 13) - 6; variant 1:
 * a_l = 1, b_l = l; variant 2: a_l = l, b_l = 1) and returns each lane's 4
 * accumulator values (out[lane*4 + m]), so a test can pin the operand and
 * accumulator lane layouts the update kernels rely on.  Variant 3 checks the
 * panel kernels' lane exchanges instead: out64[4l + k] = a[l ^ (4 << k)] and
 * out32[4l + k] = sum of a over the lanes congruent to l modulo 4 << k, with
 * a_l = l.
 * avs_schedule_stats: size of the task DAG and the list-scheduling makespan
 * estimate (microseconds) for an nt x nt tile grid, without a device. */
avs_status avs_debug_mfma_layout(int32_t variant, double *out64, float *out32);
avs_status avs_schedule_stats(int32_t nt, int32_t b, int32_t alg, int32_t kend, int32_t tree, int32_t domain,
                              int32_t workers, int64_t *ntasks, double *est_us);
/* The persistent executor's task list in dispatch order as a flat int32
 * stream, per task: kind, mode, trans, ntiles, vtile, btile, c1tile, cstep,
 * fac, ndep, ndep x (tile, count), nsig, nsig x tile.  alg: 0 Blockbidiag!,
 * 1 tiled_QR!.  *needed gets the stream length; out is filled if cap allows.
 * Lets a host test replay the DAG exactly as the device executes it. */
avs_status avs_schedule_dump(int32_t nt, int32_t b, int32_t alg, int32_t kend, int32_t tree, int32_t domain,
                             int32_t workers, int32_t *out, int64_t cap, int64_t *needed);
/* The stream-ordered executor's phases as a flat int32 stream: nphases, then
 * per phase kind, ncols, nchains and per chain nops followed by nops x (kind,
 * mode, trans, vtile, btile, c1tile, cstep, fac, ncols). */
avs_status avs_phase_dump(int32_t nt, int32_t b, int32_t alg, int32_t kend, int32_t tree, int32_t domain,
                          int32_t *out, int64_t cap, int64_t *needed);

#ifdef __cplusplus
}
#endif
#endif /* AVSVD_H */
