// server.cpp -- continuous-batching svdvals request loop.
//
// Requests (any n, FP32 or FP64) are accepted as they arrive and admitted to
// the device first-come first-served.  The device state lives in a pool of
// handles (tile-major working copy, upload staging, factor slots, band
// buffers), one per (dtype, padded order N, tile size b); a handle is created
// when a request of its shape is admitted and none is idle, kept after the
// request and reused by the next request of the same shape, so a steady
// stream of requests allocates nothing.  Every pooled handle counts against
// the HBM budget, idle or busy; when a new shape does not fit, idle handles
// of other shapes are destroyed (least recently used first).  Each of
// `max_inflight` device workers runs one request at a time on its handle's
// streams, so the band reductions of admitted requests overlap on the GPU.
// Only the (b+1) x N band comes back over PCIe; the host stage (bulge chasing
// in FP64, then the bidiagonal's singular values by bisection on the device
// or LAPACK dbdsdc) runs on a pool of host workers, each with its own stream
// and bisection buffers sized once for the largest order the budget admits.
// Latency is arrival -> singular values ready.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "avsvd.h"
#include "internal.hpp"

using avs::fail;

namespace {

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Request {
    uint64_t id = 0;
    int dtype = AVS_F64;
    int64_t n = 0, N = 0;
    int b = 32;
    size_t footprint = 0;
    std::vector<char> A;
    std::vector<double> band, sigma;
    double t_arrive = 0, t_admit = 0, t_band = 0, t_done = 0;
    avs_status status = AVS_OK;
    std::string err;
    bool done = false;
};

// Tile sizes chosen by the autotuner: for each dtype, entries (n_upto, b)
// sorted by n_upto; a request of order n takes the first entry with
// n <= n_upto, or the last entry when n is larger than every n_upto.
struct TuneEntry {
    int64_t n_upto;
    int32_t b;
};
std::mutex g_tune_mu;
std::map<int, std::vector<TuneEntry>> g_tune;

int default_b(int64_t n)
{
    // the host stage cost grows with b; small tiles keep it cheap
    return n <= 2048 ? 32 : 64;
}

int tuned_b(int dtype, int64_t n)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(dtype);
    if (it == g_tune.end() || it->second.empty()) return default_b(n);
    for (const TuneEntry &e : it->second)
        if (n <= e.n_upto) return e.b;
    return it->second.back().b;
}

int pick_b(const avs_server_opts &o, int dtype, int64_t n)
{
    const int want = dtype == AVS_F64 ? o.b_f64 : o.b_f32;
    if (want == 32 || want == 64 || want == 128) return want;
    return tuned_b(dtype, n);
}

size_t footprint_of(int dtype, int64_t N, int b)
{
    const size_t es = dtype == AVS_F64 ? 8 : 4;
    const size_t nt = (size_t)(N / b);
    return 2 * es * (size_t)N * N + 4 * nt * nt * 17 * (size_t)b * 8 + nt * nt * 4 + (size_t)(b + 1) * N * 8 +
           (64u << 20);
}

} // namespace

// one pooled handle
struct Slot {
    avs_handle *h = nullptr;
    int dtype = AVS_F64;
    int64_t N = 0;
    int b = 0;
    size_t bytes = 0;
    bool busy = false;
    uint64_t last_use = 0;
};

// per host worker: stream and bisection buffers (2*nmax and nmax doubles on
// the device and pinned on the host), allocated at server creation
struct HostBufs {
    hipStream_t s = nullptr;
    double *dT2 = nullptr, *dSig = nullptr, *hT2 = nullptr, *hSig = nullptr;
};

struct avs_server {
    avs_server_opts opts{};
    std::vector<Slot> pool;
    uint64_t use_clock = 0;
    int64_t nmax = 0;
    std::vector<HostBufs> hbufs;
    std::mutex mu;
    std::condition_variable cv_dev, cv_host, cv_done;
    std::deque<std::shared_ptr<Request>> pending, host_q;
    std::map<uint64_t, std::shared_ptr<Request>> reqs;
    size_t used = 0;
    bool stop = false;
    uint64_t next_id = 1;
    std::vector<std::thread> dev_workers, host_workers;
    std::vector<double> lat;
    double t_first = -1, t_last = 0;
    int64_t completed = 0;
};

namespace {

void free_host_bufs(avs_server *S)
{
    for (HostBufs &hb : S->hbufs) {
        if (hb.s) hipStreamSynchronize(hb.s);
        hipFree(hb.dT2);
        hipFree(hb.dSig);
        if (hb.hT2) hipHostFree(hb.hT2);
        if (hb.hSig) hipHostFree(hb.hSig);
        if (hb.s) hipStreamDestroy(hb.s);
        hb = HostBufs();
    }
}

// Takes the request at the head of the queue once a pooled handle of its
// shape is idle or the budget has room for a new one (after evicting idle
// handles of other shapes, least recently used first).  Called with S->mu
// held; returns the slot index or -1 (wait).  Handles to destroy are moved to
// `evict` (destroyed by the caller without the lock).
int admit(avs_server *S, const Request &r, std::vector<avs_handle *> &evict)
{
    int idle = -1;
    for (size_t i = 0; i < S->pool.size(); ++i) {
        const Slot &x = S->pool[i];
        if (!x.busy && x.dtype == r.dtype && x.N == r.N && x.b == r.b) {
            idle = (int)i;
            break;
        }
    }
    if (idle >= 0) {
        S->pool[idle].busy = true;
        S->pool[idle].last_use = ++S->use_clock;
        return idle;
    }
    const size_t budget = (size_t)S->opts.hbm_budget_bytes;
    size_t idle_bytes = 0;
    for (const Slot &x : S->pool)
        if (!x.busy) idle_bytes += x.bytes;
    if (S->used - idle_bytes + r.footprint > budget) return -1; // wait for busy handles to finish
    while (S->used + r.footprint > budget) {
        int lru = -1;
        for (size_t i = 0; i < S->pool.size(); ++i)
            if (!S->pool[i].busy && (lru < 0 || S->pool[i].last_use < S->pool[lru].last_use)) lru = (int)i;
        evict.push_back(S->pool[lru].h);
        S->used -= S->pool[lru].bytes;
        S->pool.erase(S->pool.begin() + lru);
    }
    Slot x;
    x.dtype = r.dtype;
    x.N = r.N;
    x.b = r.b;
    x.bytes = r.footprint;
    x.busy = true;
    x.last_use = ++S->use_clock;
    S->used += x.bytes;
    S->pool.push_back(x);
    return (int)S->pool.size() - 1;
}

void device_worker(avs_server *S)
{
    for (;;) {
        std::shared_ptr<Request> r;
        std::vector<avs_handle *> evict;
        avs_handle *h = nullptr;
        size_t key = 0; // slot identity: its position can move when others are evicted
        {
            std::unique_lock<std::mutex> lk(S->mu);
            int slot = -1;
            S->cv_dev.wait(lk, [&] {
                if (S->stop) return true;
                if (S->pending.empty()) return false;
                slot = admit(S, *S->pending.front(), evict);
                return slot >= 0;
            });
            if (S->stop) {
                lk.unlock();
                for (avs_handle *e : evict) avs_destroy(e);
                return;
            }
            r = S->pending.front();
            S->pending.pop_front();
            r->t_admit = now_ms();
            h = S->pool[slot].h;
            key = S->pool[slot].last_use;
        }
        for (avs_handle *e : evict) avs_destroy(e);
        avs_status st = AVS_OK;
        if (!h) {
            avs_opts o;
            avs_default_opts(&o);
            o.update_mode = S->opts.update_mode;
            st = avs_create(&h, r->dtype, r->N, r->b, &o);
        }
        if (!st) st = avs_set_matrix_host_sub(h, r->A.data(), r->n, r->n);
        if (!st) st = avs_blockbidiag(h, 0);
        if (!st) {
            r->band.assign((size_t)(r->b + 1) * r->N, 0.0);
            st = avs_get_band(h, r->band.data(), r->b + 1);
        }
        std::string msg = st ? avs_last_error() : "";
        r->t_band = now_ms();
        std::vector<char>().swap(r->A);
        {
            std::lock_guard<std::mutex> lk(S->mu);
            for (size_t i = 0; i < S->pool.size(); ++i) {
                Slot &x = S->pool[i];
                if (x.busy && x.last_use == key) {
                    if (st) { // a failed handle is not reused
                        if (h) evict.assign(1, h);
                        S->used -= x.bytes;
                        S->pool.erase(S->pool.begin() + i);
                    } else {
                        x.h = h;
                        x.busy = false;
                    }
                    break;
                }
            }
            r->status = st;
            r->err = msg;
            S->host_q.push_back(r);
        }
        if (st)
            for (avs_handle *e : evict) avs_destroy(e);
        S->cv_dev.notify_all();
        S->cv_host.notify_one();
    }
}

void host_worker(avs_server *S, int wid)
{
    const HostBufs &hb = S->hbufs[wid];
    for (;;) {
        std::shared_ptr<Request> r;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_host.wait(lk, [&] { return S->stop || !S->host_q.empty(); });
            if (S->host_q.empty()) return;
            r = S->host_q.front();
            S->host_q.pop_front();
        }
        if (!r->status) {
            const int64_t N = r->N;
            std::vector<double> d(N), e(std::max<int64_t>(N - 1, 1)), sg(N);
            // one host thread per request, four from n = 4096 (the largest
            // requests set the p99; the sweeps pipeline well there)
            avs_status st = avs::bidiagonalize_band_one(N, r->b, r->band.data(), r->b + 1, d.data(), e.data(),
                                                        N >= 4096 ? 4 : 1);
            // the bidiagonal's singular values: bisection on the device with
            // this worker's stream and buffers (nothing allocated here), or
            // LAPACK dbdsdc on this thread
            if (!st) {
                if (S->opts.device_sigma && N <= S->nmax)
                    st = avs::bidiag_sigma(N, N, d.data(), e.data(), sg.data(), hb.dT2, hb.dSig, hb.hT2, hb.hSig,
                                           hb.s);
                else
                    st = avs_bdsdc_values(N, d.data(), e.data(), sg.data());
            }
            if (st) {
                r->status = st;
                r->err = avs_last_error();
            } else {
                r->sigma.assign(sg.begin(), sg.begin() + r->n); // drop the padding's zeros
            }
        }
        std::vector<double>().swap(r->band);
        {
            std::lock_guard<std::mutex> lk(S->mu);
            r->t_done = now_ms();
            r->done = true;
            if (!r->status) {
                S->lat.push_back(r->t_done - r->t_arrive);
                S->t_last = std::max(S->t_last, r->t_done);
                S->completed++;
            }
        }
        S->cv_done.notify_all();
    }
}

} // namespace

extern "C" {

avs_status avs_tune_set(int32_t dtype, int64_t n_upto, int32_t b)
{
    if (dtype != AVS_F32 && dtype != AVS_F64) return fail(-1, "dtype must be AVS_F32 or AVS_F64");
    if (n_upto < 1) return fail(-2, "n_upto must be positive");
    if (b != 32 && b != 64 && b != 128) return fail(-3, "tile size must be 32, 64 or 128");
    std::lock_guard<std::mutex> lk(g_tune_mu);
    std::vector<TuneEntry> &v = g_tune[dtype];
    auto it = std::find_if(v.begin(), v.end(), [&](const TuneEntry &e) { return e.n_upto == n_upto; });
    if (it != v.end()) it->b = b;
    else v.push_back({n_upto, b});
    std::sort(v.begin(), v.end(), [](const TuneEntry &x, const TuneEntry &y) { return x.n_upto < y.n_upto; });
    return AVS_OK;
}

avs_status avs_tune_clear(void)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune.clear();
    return AVS_OK;
}

avs_status avs_tune_lookup(int32_t dtype, int64_t n, int32_t *b)
{
    if (dtype != AVS_F32 && dtype != AVS_F64) return fail(-1, "dtype must be AVS_F32 or AVS_F64");
    if (n < 1) return fail(-2, "n must be positive");
    if (!b) return fail(-3, "null b");
    *b = tuned_b(dtype, n);
    return AVS_OK;
}

void avs_server_default_opts(avs_server_opts *o)
{
    if (!o) return;
    memset(o, 0, sizeof *o);
    o->hbm_budget_bytes = (int64_t)64 << 30;
    o->max_inflight = 4;
    o->host_threads = avs::host_thread_count();
    o->update_mode = AVS_UPD_NATIVE;
}

avs_status avs_server_create(avs_server **sp, const avs_server_opts *opts)
{
    if (!sp) return fail(-1, "null server pointer");
    *sp = nullptr;
    std::unique_ptr<avs_server> S(new avs_server());
    if (opts) S->opts = *opts;
    else avs_server_default_opts(&S->opts);
    if (S->opts.hbm_budget_bytes <= 0) return fail(-2, "hbm_budget_bytes must be positive");
    if (S->opts.max_inflight <= 0) S->opts.max_inflight = 4;
    if (S->opts.host_threads <= 0) S->opts.host_threads = avs::host_thread_count();
    for (int32_t r : S->opts.reserved)
        if (r != 0) return fail(-2, "opts.reserved must be zero");
    // the largest padded order the budget admits (FP32, the smallest
    // footprint per element), which sizes the bisection buffers
    int64_t nmax = 128;
    while (footprint_of(AVS_F32, 2 * nmax, 128) <= (size_t)S->opts.hbm_budget_bytes) nmax *= 2;
    while (footprint_of(AVS_F32, nmax + 128, 128) <= (size_t)S->opts.hbm_budget_bytes) nmax += 128;
    S->nmax = nmax;
    if (S->opts.device_sigma) {
        S->hbufs.resize(S->opts.host_threads);
        for (HostBufs &hb : S->hbufs) {
            hipError_t e = hipStreamCreateWithFlags(&hb.s, hipStreamNonBlocking);
            if (e == hipSuccess) e = hipMalloc(&hb.dT2, sizeof(double) * 2 * (size_t)nmax);
            if (e == hipSuccess) e = hipMalloc(&hb.dSig, sizeof(double) * (size_t)nmax);
            if (e == hipSuccess) e = hipHostMalloc((void **)&hb.hT2, sizeof(double) * 2 * (size_t)nmax, hipHostMallocDefault);
            if (e == hipSuccess) e = hipHostMalloc((void **)&hb.hSig, sizeof(double) * (size_t)nmax, hipHostMallocDefault);
            if (e != hipSuccess) {
                free_host_bufs(S.get());
                return fail(AVS_ERR_HIP, "server bisection buffers: %s", hipGetErrorString(e));
            }
        }
    } else {
        S->hbufs.resize(S->opts.host_threads);
    }
    for (int i = 0; i < S->opts.max_inflight; ++i) S->dev_workers.emplace_back(device_worker, S.get());
    for (int i = 0; i < S->opts.host_threads; ++i) S->host_workers.emplace_back(host_worker, S.get(), i);
    *sp = S.release();
    return AVS_OK;
}

avs_status avs_server_submit(avs_server *S, int32_t dtype, int64_t n, const void *A, int64_t lda, uint64_t *req_id)
{
    if (!S) return fail(-1, "null server");
    if (dtype != AVS_F64 && dtype != AVS_F32) return fail(-2, "dtype %d not supported", dtype);
    if (n <= 0) return fail(-3, "n must be positive");
    if (!A) return fail(-4, "null matrix");
    if (lda < n) return fail(-5, "lda < n");
    if (!req_id) return fail(-6, "null request id");
    auto r = std::make_shared<Request>();
    r->t_arrive = now_ms();
    r->dtype = dtype;
    r->n = n;
    r->b = pick_b(S->opts, dtype, n);
    r->N = (n + r->b - 1) / r->b * r->b;
    r->footprint = footprint_of(dtype, r->N, r->b);
    if (r->footprint > (size_t)S->opts.hbm_budget_bytes)
        return fail(AVS_ERR_BUDGET, "request n=%lld needs %zu bytes, budget %lld", (long long)n, r->footprint,
                    (long long)S->opts.hbm_budget_bytes);
    const size_t es = dtype == AVS_F64 ? 8 : 4;
    r->A.resize(es * (size_t)n * n);
    for (int64_t j = 0; j < n; ++j)
        memcpy(r->A.data() + es * (size_t)j * n, (const char *)A + es * (size_t)j * lda, es * n);
    {
        std::lock_guard<std::mutex> lk(S->mu);
        r->id = S->next_id++;
        if (S->t_first < 0) S->t_first = r->t_arrive;
        S->reqs[r->id] = r;
        S->pending.push_back(r);
    }
    S->cv_dev.notify_all();
    *req_id = r->id;
    return AVS_OK;
}

avs_status avs_server_poll(avs_server *S, uint64_t req_id, int32_t *done)
{
    if (!S) return fail(-1, "null server");
    if (!done) return fail(-3, "null output");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(req_id);
    if (it == S->reqs.end()) return fail(-2, "unknown request %llu", (unsigned long long)req_id);
    *done = it->second->done ? 1 : 0;
    return AVS_OK;
}

avs_status avs_server_wait(avs_server *S, uint64_t req_id, double *sigma, double *latency_ms)
{
    if (!S) return fail(-1, "null server");
    std::shared_ptr<Request> r;
    {
        std::unique_lock<std::mutex> lk(S->mu);
        auto it = S->reqs.find(req_id);
        if (it == S->reqs.end()) return fail(-2, "unknown request %llu", (unsigned long long)req_id);
        r = it->second;
        S->cv_done.wait(lk, [&] { return r->done; });
        S->reqs.erase(it);
    }
    if (r->status) return fail(r->status, "request %llu failed: %s", (unsigned long long)req_id, r->err.c_str());
    if (sigma) memcpy(sigma, r->sigma.data(), sizeof(double) * r->n);
    if (latency_ms) *latency_ms = r->t_done - r->t_arrive;
    return AVS_OK;
}

avs_status avs_server_stats(avs_server *S, double *p50_ms, double *p99_ms, double *solves_per_s, int64_t *completed)
{
    if (!S) return fail(-1, "null server");
    std::lock_guard<std::mutex> lk(S->mu);
    std::vector<double> l = S->lat;
    std::sort(l.begin(), l.end());
    auto pct = [&](double p) {
        if (l.empty()) return 0.0;
        const size_t i = std::min(l.size() - 1, (size_t)(p * (l.size() - 1) + 0.5));
        return l[i];
    };
    if (p50_ms) *p50_ms = pct(0.50);
    if (p99_ms) *p99_ms = pct(0.99);
    if (solves_per_s) *solves_per_s = (S->completed > 0 && S->t_last > S->t_first)
                                          ? 1000.0 * S->completed / (S->t_last - S->t_first) : 0.0;
    if (completed) *completed = S->completed;
    return AVS_OK;
}

avs_status avs_server_destroy(avs_server *S)
{
    if (!S) return AVS_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_dev.notify_all();
    S->cv_host.notify_all();
    for (auto &t : S->dev_workers) t.join();
    for (auto &t : S->host_workers) t.join();
    for (Slot &x : S->pool)
        if (x.h) avs_destroy(x.h);
    free_host_bufs(S);
    delete S;
    return AVS_OK;
}

} // extern "C"
