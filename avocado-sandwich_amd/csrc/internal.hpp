// internal.hpp -- helpers shared by the host-side translation units.
#pragma once
#include "avsvd.h"

namespace avs {
// Records a message for avs_last_error() (thread-local) and returns `code`.
avs_status fail(avs_status code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
// host stage with an explicit thread count (1 = sequential sweeps)
avs_status bidiagonalize_band_threads(int64_t n, int32_t bw, const double *band, int64_t ldband, double *d, double *e,
                                      int nthreads);
// stage 2 with LAPACK dgbbrd on one thread
avs_status bidiagonalize_band_lapack(int64_t n, int32_t bw, const double *band, int64_t ldband, double *d, double *e);
// stage 2 for a caller that gives each matrix one host thread (dgbbrd if available)
avs_status bidiagonalize_band_one(int64_t n, int32_t bw, const double *band, int64_t ldband, double *d, double *e,
                                  int nthreads);
avs_status bidiagonalize_band_auto(int64_t n, int32_t bw, const double *band, int64_t ldband, double *d, double *e);
// singular values of the upper bidiagonal (d, e) of order n, the nkeep
// largest, descending: bisection on the device on stream `s` with the
// caller's buffers (dT2: 2n doubles, dSig: n on the device; hT2, hSig the same
// sizes in pinned host memory), or LAPACK dbdsdc on the host for n < 128 or
// with AVSVD_BDSDC=host (bidiag_sigma in api.cpp)
avs_status bidiag_sigma(int64_t n, int64_t nkeep, const double *d, const double *e, double *sigma, double *dT2,
                        double *dSig, double *hT2, double *hSig, void *s);
// the device part of bidiag_sigma without the wait: the nkeep values land in
// hSig once stream `s` reaches this point
avs_status bidiag_sigma_enqueue(int64_t n, int64_t nkeep, const double *d, const double *e, double *dT2,
                                double *dSig, double *hT2, double *hSig, void *s);
// whether bidiag_sigma bisects on the device for order n (n >= 128, unless
// AVSVD_BDSDC=host)
bool bisect_on_device(int64_t n);
// the handle's band, FP64 (b+1) x n upper-band layout, into its pinned host
// buffer (*host_band), enqueued on the handle's stream (no wait)
avs_status band_enqueue(avs_handle *h, const double **host_band);
// the handle's stream (hipStream_t)
void *handle_stream(avs_handle *h);
// the handle's panel streams are ordinary streams of its own, not the
// device's pool of dedicated queues (for the serving loop's handles)
void handle_use_own_panel_stream(avs_handle *h);
// band -> bidiagonal by Householder bulge chasing, sweeps pipelined over
// nthreads host threads (bandbrd.cpp); the result does not depend on nthreads
avs_status bidiagonalize_band_chase(int64_t n, int32_t bw, const double *band, int64_t ldband, double *d, double *e,
                                    int nthreads);
// host threads the library uses (AVSVD_HOST_THREADS, else min(16, cores))
int host_thread_count();
// column tiles one update workgroup takes (its 8 strips of 16 columns)
inline int upd_tiles_for(int b) { return b >= 128 ? 1 : (b == 64 ? 2 : 4); }
} // namespace avs
