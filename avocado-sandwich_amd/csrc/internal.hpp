// internal.hpp -- helpers shared by the host-side translation units.
#pragma once
#include "avsvd.h"

namespace avs {
// Records a message for avs_last_error() (thread-local) and returns `code`.
avs_status fail(avs_status code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
// host stage with an explicit thread count (1 = sequential sweeps)
avs_status bidiagonalize_band_threads(int64_t n, int32_t bw, const double *band, int64_t ldband, double *d, double *e,
                                      int nthreads);
// host threads the library uses (AVSVD_HOST_THREADS, else min(16, cores))
int host_thread_count();
} // namespace avs
