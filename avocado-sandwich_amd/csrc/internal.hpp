// internal.hpp -- helpers shared by the host-side translation units.
#pragma once
#include "avsvd.h"

namespace avs {
// Records a message for avs_last_error() (thread-local) and returns `code`.
avs_status fail(avs_status code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
} // namespace avs
