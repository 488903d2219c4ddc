// ooc.cpp -- out-of-core Blockbidiag! and tiled_QR! on a host matrix.
//
// OOC_alg! (KAbasedSVD/src/tiledalgos.jl:232-251): when n > kswitch * b, the
// first nt - kswitch steps run on the host matrix streamed through the GPU,
// the trailing kswitch x kswitch tiles then run in core; otherwise the whole
// matrix runs in core.  The streamed steps follow the reference's
// LargeTiledMatrix algorithm (KAbasedSVD/src/TiledMatrix.jl:121-186,
// tiledalgos.jl:75-225): the flat tree, one unit of tiles at a time against
// a resident pivot unit -- a QR sweep of column k streams the tile rows
// k+1.. (tiles k..nt-1 of each) against tile row k, an LQ sweep of row k
// streams the tile columns k+2.. (tiles k..nt-1 of each) against tile column
// k+1.
//
// On the device a sweep holds the pivot unit and two unit buffers (3 (nt-k)
// tiles).  A unit goes host -> pinned staging (packed tile-major on the CPU)
// -> HBM on an upload stream, through its TSQRT and TSMQR on the compute
// stream, and back on a download stream; the next unit is packed and
// uploaded while the current one computes.  The kernels and the order in
// which every tile receives its updates are those of the in-core flat-tree
// executors, so the result equals the in-core AVS_TREE_FLAT run bit for bit.
#include <hip/hip_runtime.h>

#include <condition_variable>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "avsvd.h"
#include "internal.hpp"
#include "kernels.hpp"
#include "schedule.hpp"
#include "task.hpp"

namespace avs {
namespace {

#define OOCCHK(call)                                                                                                   \
    do {                                                                                                               \
        hipError_t e_ = (call);                                                                                        \
        if (e_ != hipSuccess) return fail(AVS_ERR_HIP, "%s failed: %s (%d)", #call, hipGetErrorString(e_), (int)e_);  \
    } while (0)

// Host threads for packing and unpacking units: run(count, f) calls f(lo, hi)
// on nth disjoint parts of [0, count), part 0 on the calling thread, and
// returns when every part is done.  The threads live for one streamed run.
class Pool {
    std::vector<std::thread> th_;
    std::mutex m_;
    std::condition_variable cv_, done_;
    const std::function<void(int, int)> *job_ = nullptr;
    int count_ = 0, gen_ = 0, pending_ = 0, nth_;
    bool stop_ = false;

    void part(int id) const
    {
        const int lo = (int)((long)count_ * id / nth_), hi = (int)((long)count_ * (id + 1) / nth_);
        if (lo < hi) (*job_)(lo, hi);
    }
    void loop(int id)
    {
        int seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(m_);
                cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
            }
            part(id);
            std::lock_guard<std::mutex> lk(m_);
            if (--pending_ == 0) done_.notify_one();
        }
    }

  public:
    explicit Pool(int n) : nth_(n < 1 ? 1 : n)
    {
        for (int i = 1; i < nth_; ++i) th_.emplace_back([this, i] { loop(i); });
    }
    ~Pool()
    {
        {
            std::lock_guard<std::mutex> lk(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : th_) t.join();
    }
    void run(int count, const std::function<void(int, int)> &f)
    {
        if (nth_ == 1 || count < 2 * nth_) {
            f(0, count);
            return;
        }
        {
            std::lock_guard<std::mutex> lk(m_);
            job_ = &f;
            count_ = count;
            pending_ = nth_ - 1;
            ++gen_;
        }
        cv_.notify_all();
        part(0);
        std::unique_lock<std::mutex> lk(m_);
        done_.wait(lk, [&] { return pending_ == 0; });
    }
};

// device and pinned buffers of one streamed run; released on every exit path
struct OocRes {
    void *dP = nullptr, *dU[2] = {nullptr, nullptr};
    double *dTau = nullptr, *dTjj = nullptr; // [0] GEQRT, [1] TSQRT factors
    void *hP = nullptr, *hU[2] = {nullptr, nullptr};
    hipStream_t cs = nullptr, up = nullptr, down = nullptr;
    hipEvent_t evUp[2] = {}, evComp[2] = {}, evDown[2] = {}, evP = nullptr;
    ~OocRes()
    {
        if (cs) hipStreamSynchronize(cs);
        if (up) hipStreamSynchronize(up);
        if (down) hipStreamSynchronize(down);
        hipFree(dP);
        hipFree(dU[0]);
        hipFree(dU[1]);
        hipFree(dTau);
        hipFree(dTjj);
        hipHostFree(hP);
        hipHostFree(hU[0]);
        hipHostFree(hU[1]);
        for (int s = 0; s < 2; ++s) {
            if (evUp[s]) hipEventDestroy(evUp[s]);
            if (evComp[s]) hipEventDestroy(evComp[s]);
            if (evDown[s]) hipEventDestroy(evDown[s]);
        }
        if (evP) hipEventDestroy(evP);
        if (cs) hipStreamDestroy(cs);
        if (up) hipStreamDestroy(up);
        if (down) hipStreamDestroy(down);
    }
};

template <typename TS, int B> struct Ooc {
    TS *A;
    long lda;
    int nt, split;
    Pool *pool;
    OocRes r;
    static constexpr long TSZ = (long)B * B;

    // unit u of a sweep: tile row u (QR: tiles (u, k..nt-1)) or tile column u
    // (LQ: tiles (k..nt-1, u)), as L = nt - k consecutive b x b column-major
    // tiles
    // (column c of tile t): B contiguous elements on both sides
    void copy_unit(bool lq, int k, int u, TS *tiles, bool to_tiles)
    {
        pool->run((nt - k) * B, [&](int lo, int hi) {
            for (int x = lo; x < hi; ++x) {
                const int t = x / B, c = x % B;
                const long r0 = lq ? (long)(k + t) * B : (long)u * B;
                const long c0 = lq ? (long)u * B : (long)(k + t) * B;
                TS *h = A + r0 + (c0 + c) * lda, *d = tiles + t * TSZ + (long)c * B;
                if (to_tiles) memcpy(d, h, sizeof(TS) * B);
                else memcpy(h, d, sizeof(TS) * B);
            }
        });
    }
    void pack(bool lq, int k, int u, TS *dst) { copy_unit(lq, k, u, dst, true); }
    void unpack(bool lq, int k, int u, TS *src) { copy_unit(lq, k, u, src, false); }

    avs_status setup()
    {
        const size_t ub = sizeof(TS) * TSZ * nt;
        OOCCHK(hipMalloc(&r.dP, ub));
        OOCCHK(hipMalloc(&r.dU[0], ub));
        OOCCHK(hipMalloc(&r.dU[1], ub));
        OOCCHK(hipMalloc(&r.dTau, sizeof(double) * 2 * B));
        OOCCHK(hipMalloc(&r.dTjj, sizeof(double) * 2 * 16 * B));
        OOCCHK(hipHostMalloc(&r.hP, ub, hipHostMallocDefault));
        OOCCHK(hipHostMalloc(&r.hU[0], ub, hipHostMallocDefault));
        OOCCHK(hipHostMalloc(&r.hU[1], ub, hipHostMallocDefault));
        OOCCHK(hipStreamCreateWithFlags(&r.cs, hipStreamNonBlocking));
        OOCCHK(hipStreamCreateWithFlags(&r.up, hipStreamNonBlocking));
        OOCCHK(hipStreamCreateWithFlags(&r.down, hipStreamNonBlocking));
        for (int s = 0; s < 2; ++s) {
            OOCCHK(hipEventCreateWithFlags(&r.evUp[s], hipEventDisableTiming));
            OOCCHK(hipEventCreateWithFlags(&r.evComp[s], hipEventDisableTiming));
            OOCCHK(hipEventCreateWithFlags(&r.evDown[s], hipEventDisableTiming));
        }
        OOCCHK(hipEventCreateWithFlags(&r.evP, hipEventDisableTiming));
        return AVS_OK;
    }

    // GEQRT of the pivot's first tile and UNMQR of the rest of the pivot
    // (GE task and its updates, schedule.cpp flat sweep), or TSQRT of the
    // pivot's first tile against the unit's and TSMQR of both units
    void ops(bool lq, int L, TS *unit)
    {
        TS *P = (TS *)r.dP;
        const int per = upd_tiles_for(B);
        if (!unit) {
            launch_panel<TS, B>(MODE_GE, P, B, P, B, lq, r.dTau, r.dTjj, r.cs);
            if (L > 1)
                launch_update<TS, B>(MODE_GE, true, P, B, lq, r.dTjj, P + TSZ, B, TSZ, P + TSZ, B, TSZ, lq, L - 1, per,
                                     split, r.cs);
        } else {
            launch_panel<TS, B>(MODE_TS, P, B, unit, B, lq, r.dTau + B, r.dTjj + 16 * B, r.cs);
            if (L > 1)
                launch_update<TS, B>(MODE_TS, true, unit, B, lq, r.dTjj + 16 * B, P + TSZ, B, TSZ, unit + TSZ, B, TSZ,
                                     lq, L - 1, per, split, r.cs);
        }
    }

    // one sweep: QR of tile column k (pivot tile row k, units k+1..nt-1) or
    // LQ of tile row k (pivot tile column k+1, units k+2..nt-1)
    avs_status sweep(bool lq, int k)
    {
        const int L = nt - k;
        const int pu = lq ? k + 1 : k;
        if (pu >= nt) return AVS_OK;
        const size_t lb = sizeof(TS) * TSZ * L;
        pack(lq, k, pu, (TS *)r.hP);
        OOCCHK(hipMemcpyAsync(r.dP, r.hP, lb, hipMemcpyHostToDevice, r.up));
        OOCCHK(hipEventRecord(r.evP, r.up));
        OOCCHK(hipStreamWaitEvent(r.cs, r.evP, 0));
        ops(lq, L, nullptr);
        OOCCHK(hipGetLastError());
        const int u0 = pu + 1, nu = nt - u0;
        auto upload = [&](int i) -> avs_status {
            const int s = i & 1;
            pack(lq, k, u0 + i, (TS *)r.hU[s]);
            OOCCHK(hipMemcpyAsync(r.dU[s], r.hU[s], lb, hipMemcpyHostToDevice, r.up));
            OOCCHK(hipEventRecord(r.evUp[s], r.up));
            return AVS_OK;
        };
        avs_status st;
        if (nu > 0 && (st = upload(0))) return st;
        for (int i = 0; i < nu; ++i) {
            const int s = i & 1;
            // compute unit i once it is on the device, then send it back
            OOCCHK(hipStreamWaitEvent(r.cs, r.evUp[s], 0));
            ops(lq, L, (TS *)r.dU[s]);
            OOCCHK(hipGetLastError());
            OOCCHK(hipEventRecord(r.evComp[s], r.cs));
            OOCCHK(hipStreamWaitEvent(r.down, r.evComp[s], 0));
            OOCCHK(hipMemcpyAsync(r.hU[s], r.dU[s], lb, hipMemcpyDeviceToHost, r.down));
            OOCCHK(hipEventRecord(r.evDown[s], r.down));
            // unit i-1 back on the host frees slot s^1 for unit i+1
            if (i > 0) {
                OOCCHK(hipEventSynchronize(r.evDown[s ^ 1]));
                unpack(lq, k, u0 + i - 1, (TS *)r.hU[s ^ 1]);
            }
            if (i + 1 < nu && (st = upload(i + 1))) return st;
        }
        if (nu > 0) {
            const int s = (nu - 1) & 1;
            OOCCHK(hipEventSynchronize(r.evDown[s]));
            unpack(lq, k, u0 + nu - 1, (TS *)r.hU[s]);
        }
        // the pivot comes back after its last update
        OOCCHK(hipEventRecord(r.evP, r.cs));
        OOCCHK(hipStreamWaitEvent(r.down, r.evP, 0));
        OOCCHK(hipMemcpyAsync(r.hP, r.dP, lb, hipMemcpyDeviceToHost, r.down));
        OOCCHK(hipStreamSynchronize(r.down));
        unpack(lq, k, pu, (TS *)r.hP);
        return AVS_OK;
    }

    // the zeros Blockbidiag! leaves in its first ksteps steps (tiledalgos.jl:
    // 46,58): outside the diagonal and super-diagonal tiles, and the strict
    // lower / strict upper part of those, in every tile row or column < ksteps
    void band_rule(int ksteps)
    {
        pool->run(nt * B, [&](int lo, int hi) {
            for (int x = lo; x < hi; ++x) {
                const int j = x / B, c = x % B;
                TS *col = A + (long)x * lda;
                for (int i = 0; i < nt; ++i) {
                    if (i >= ksteps && j >= ksteps) continue;
                    TS *t = col + (long)i * B;
                    if (j == i) {
                        for (int rr = c + 1; rr < B; ++rr) t[rr] = (TS)0;
                    } else if (j == i + 1) {
                        for (int rr = 0; rr < c; ++rr) t[rr] = (TS)0;
                    } else {
                        memset(t, 0, sizeof(TS) * B);
                    }
                }
            }
        });
    }

    avs_status run(int alg, int ksteps)
    {
        avs_status st = setup();
        if (st) return st;
        for (int k = 0; k < ksteps; ++k) {
            if ((st = sweep(false, k))) return st;
            if (alg == ALG_BLOCKBIDIAG && (st = sweep(true, k))) return st;
        }
        if (alg == ALG_BLOCKBIDIAG) band_rule(ksteps);
        return AVS_OK;
    }
};

// the whole matrix, or its trailing m x m block, in core with a handle
avs_status in_core(int alg, int32_t dtype, int64_t m, int32_t b, void *A, int64_t lda, const avs_opts *opts)
{
    avs_handle *h = nullptr;
    avs_status st = avs_create(&h, dtype, m, b, opts);
    if (st) return st;
    st = avs_set_matrix_host(h, A, lda);
    if (!st) st = alg == ALG_BLOCKBIDIAG ? avs_blockbidiag(h, 0) : avs_tiled_qr(h, 0);
    if (!st) st = avs_get_matrix_host(h, A, lda);
    avs_destroy(h);
    return st;
}

avs_status ooc_alg(int alg, int32_t dtype, int64_t n, int32_t b, void *A, int64_t lda, int32_t kswitch,
                   const avs_opts *opts)
{
    if (dtype != AVS_F64 && dtype != AVS_F32) return fail(-1, "dtype %d not supported", dtype);
    if (b != 32 && b != 64 && b != 128) return fail(-3, "Not supported: tile size %d (built: 32, 64, 128)", b);
    if (n <= 0 || n % b != 0)
        return fail(-2, "Not supported: n=%lld is not a positive multiple of b=%d", (long long)n, b);
    if (!A) return fail(-4, "null matrix");
    if (lda < n) return fail(-5, "lda < n");
    if (kswitch < 0) return fail(-6, "kswitch must be >= 0");
    const int nt = (int)(n / b);
    if (n <= (int64_t)kswitch * b) return in_core(alg, dtype, n, b, A, lda, opts);
    const int ksteps = nt - kswitch;
    avs_opts o;
    avs_default_opts(&o);
    if (opts) o = *opts;
    const int split = dtype == AVS_F32 && o.update_mode == AVS_UPD_SPLIT_BF16;
    avs_status st = AVS_ERR_UNSUPPORTED;
    auto go = [&](auto tp, auto bc) {
        typedef std::remove_pointer_t<decltype(tp)> TS;
        constexpr int B = decltype(bc)::value;
        Pool pool(host_thread_count());
        Ooc<TS, B> w{(TS *)A, (long)lda, nt, split, &pool, {}};
        st = w.run(alg, ksteps);
    };
    if (dtype == AVS_F64) {
        if (b == 32) go((double *)0, std::integral_constant<int, 32>());
        else if (b == 64) go((double *)0, std::integral_constant<int, 64>());
        else go((double *)0, std::integral_constant<int, 128>());
    } else {
        if (b == 32) go((float *)0, std::integral_constant<int, 32>());
        else if (b == 64) go((float *)0, std::integral_constant<int, 64>());
        else go((float *)0, std::integral_constant<int, 128>());
    }
    if (st || kswitch == 0) return st;
    const size_t es = dtype == AVS_F64 ? 8 : 4;
    const int64_t off = (int64_t)ksteps * b;
    return in_core(alg, dtype, (int64_t)kswitch * b, b, (char *)A + es * (off + off * lda), lda, opts);
}

} // namespace
} // namespace avs

using namespace avs;

avs_status avs_ooc_blockbidiag(int32_t dtype, int64_t n, int32_t b, void *A, int64_t lda, int32_t kswitch,
                               const avs_opts *opts)
{
    return ooc_alg(ALG_BLOCKBIDIAG, dtype, n, b, A, lda, kswitch, opts);
}

avs_status avs_ooc_tiled_qr(int32_t dtype, int64_t n, int32_t b, void *A, int64_t lda, int32_t kswitch,
                            const avs_opts *opts)
{
    return ooc_alg(ALG_TILED_QR, dtype, n, b, A, lda, kswitch, opts);
}
