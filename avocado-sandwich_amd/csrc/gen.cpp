// gen.cpp -- seeded synthetic inputs: Philox4x32-10 (Salmon, Moraes, Dror and
// Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11) with the
// Box-Muller transform.  Element e of a column-major n x n matrix comes from
// counter (e/2 lo, e/2 hi, 0, 0) under key (seed lo, seed hi); one Philox call
// gives two 53-bit uniforms and therefore the normal pair (2p, 2p+1).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "avsvd.h"
#include "internal.hpp"

namespace {

inline void philox10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = (uint32_t)p1;
        c[2] = n2;
        c[3] = (uint32_t)p0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

inline double u53(uint32_t a, uint32_t b)
{
    const uint64_t x = (((uint64_t)a << 32) | b) >> 11;
    return ((double)x + 0.5) * (1.0 / 9007199254740992.0);
}

inline double normal_at(uint64_t seed, uint64_t e)
{
    const uint64_t p = e >> 1;
    uint32_t c[4] = {(uint32_t)p, (uint32_t)(p >> 32), 0u, 0u};
    philox10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const double rad = std::sqrt(-2.0 * std::log(u53(c[0], c[1])));
    const double th = 6.283185307179586476925286766559 * u53(c[2], c[3]);
    return (e & 1) ? rad * std::sin(th) : rad * std::cos(th);
}

} // namespace

extern "C" avs_status avs_gen_gaussian_host(uint64_t seed, int64_t n, double scale, int32_t dtype, void *A,
                                            int64_t lda)
{
    using avs::fail;
    if (n <= 0) return fail(-2, "n must be positive");
    if (dtype != AVS_F64 && dtype != AVS_F32) return fail(-4, "dtype %d not supported", dtype);
    if (!A) return fail(-5, "null output");
    if (lda < n) return fail(-6, "lda < n");
    unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if ((uint64_t)n * n < (1u << 16)) nth = 1;
    auto work = [&](unsigned w) {
        for (int64_t j = w; j < n; j += nth) {
            for (int64_t i = 0; i < n; ++i) {
                const double v = scale * normal_at(seed, (uint64_t)i + (uint64_t)j * n);
                if (dtype == AVS_F64) ((double *)A)[i + j * lda] = v;
                else ((float *)A)[i + j * lda] = (float)v;
            }
        }
    };
    std::vector<std::thread> th;
    for (unsigned w = 1; w < nth; ++w) th.emplace_back(work, w);
    work(0);
    for (auto &t : th) t.join();
    return AVS_OK;
}
