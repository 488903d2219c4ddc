// server.cpp -- continuous-batching svdvals request loop.
//
// Requests (any n, FP32 or FP64) are accepted as they arrive and admitted to
// the device first-come first-served.  One device thread makes every HIP
// call of the loop: it admits requests, enqueues each one's upload, band
// reduction and band download on the streams of a pooled handle, polls their
// completion events, and enqueues the bisections of finished bidiagonals; up
// to `max_inflight` reductions are in flight at once, so they overlap on the
// GPU.  Host workers do CPU work only: band -> bidiagonal (bulge chasing in
// FP64) and, unless the device bisection is selected, LAPACK dbdsdc.
//
// The device state lives in a pool of handles (tile-major working copy,
// upload staging, factor slots, band buffers), one per (dtype, padded order
// N, tile size b); a handle is created when a request of its shape is
// admitted and none is idle, kept after the request and reused by the next
// request of the same shape, so a steady stream of requests allocates
// nothing.  Every pooled handle counts against the HBM budget, idle or busy;
// when a new shape does not fit, idle handles of other shapes are destroyed
// (least recently used first).  Only the (b+1) x N band comes back over PCIe.
// Latency is arrival -> singular values ready.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
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
    std::vector<double> band, sigma, d, e;
    double t_arrive = 0, t_admit = 0, t_started = 0, t_band = 0, t_done = 0;
    bool new_handle = false;
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

// The order a request is padded to (zero rows and columns, which add exact
// zero singular values that are dropped): 8 sizes per octave, multiples of b,
// so requests of nearby orders share pooled handles.  At most 12.5 % padding.
int64_t bucket_of(int64_t n, int b)
{
    int64_t p = 1;
    while (p * 2 <= n) p *= 2;
    const int64_t step = std::max<int64_t>(b, (p / 8 + b - 1) / b * b);
    return (n + step - 1) / step * step;
}

size_t footprint_of(int dtype, int64_t N, int b)
{
    const size_t es = dtype == AVS_F64 ? 8 : 4;
    const size_t nt = (size_t)(N / b);
    return 2 * es * (size_t)N * N + 4 * nt * nt * 17 * (size_t)b * 8 + nt * nt * 4 + (size_t)(b + 1) * N * 8 +
           (64u << 20);
}

} // namespace

// one pooled handle and its completion event
struct Slot {
    avs_handle *h = nullptr;
    hipEvent_t ev = nullptr;
    void *hpin = nullptr; // pinned host copy of the request's matrix (n x n, ld n), so the upload is a DMA
    int dtype = AVS_F64;
    int64_t N = 0;
    int b = 0;
    size_t bytes = 0;
    bool busy = false;
    uint64_t last_use = 0;
};

// one in-flight bisection: stream, buffers sized for the largest order the
// budget admits, completion event
struct Bisect {
    hipStream_t s = nullptr;
    hipEvent_t ev = nullptr;
    double *dT2 = nullptr, *dSig = nullptr, *hT2 = nullptr, *hSig = nullptr;
    std::shared_ptr<Request> r;
};

struct avs_server {
    avs_server_opts opts{};
    std::vector<Slot> pool;
    uint64_t use_clock = 0;
    int64_t nmax = 0;
    std::mutex mu;
    std::condition_variable cv_dev, cv_host, cv_done;
    std::deque<std::shared_ptr<Request>> pending, host_q, sigma_q;
    std::map<uint64_t, std::shared_ptr<Request>> reqs;
    size_t used = 0;
    bool stop = false;
    uint64_t next_id = 1;
    std::thread dev_thread;
    std::vector<std::thread> host_workers;
    std::vector<double> lat;
    double t_first = -1, t_last = 0;
    int64_t completed = 0;
    std::string dev_err; // a failure of the device thread's own set-up
};

namespace {

// Takes the request at the head of the queue once a pooled handle of its
// shape is idle or the budget has room for a new one (after evicting idle
// handles of other shapes, least recently used first).  Called with S->mu
// held; returns the slot's key (its last_use stamp) or 0 (wait).  Evicted
// slots are moved to `evict`.
uint64_t admit(avs_server *S, const Request &r, std::vector<Slot> &evict)
{
    for (Slot &x : S->pool) {
        if (!x.busy && x.dtype == r.dtype && x.N == r.N && x.b == r.b) {
            x.busy = true;
            x.last_use = ++S->use_clock;
            return x.last_use;
        }
    }
    const size_t budget = (size_t)S->opts.hbm_budget_bytes;
    size_t idle_bytes = 0;
    for (const Slot &x : S->pool)
        if (!x.busy) idle_bytes += x.bytes;
    if (S->used - idle_bytes + r.footprint > budget) return 0; // wait for busy handles to finish
    while (S->used + r.footprint > budget) {
        size_t lru = S->pool.size();
        for (size_t i = 0; i < S->pool.size(); ++i)
            if (!S->pool[i].busy && (lru == S->pool.size() || S->pool[i].last_use < S->pool[lru].last_use)) lru = i;
        evict.push_back(S->pool[lru]);
        S->used -= S->pool[lru].bytes;
        S->pool.erase(S->pool.begin() + (long)lru);
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
    return x.last_use;
}

// The pending request admitted next: the oldest, or with
// AVSVD_SERVER_ORDER=largest the one of largest padded order (ties: the
// oldest).  On 48 mixed requests largest-first raised the saturated rate
// from 14.2-14.7 to 15.3-15.4 solves/s but its median latency from 1.38 to
// 1.82 s (p99 2.58 to 2.70 s; Poisson arrivals unchanged), so the default
// stays first-come first-served.  Called with S->mu held.
size_t next_pending(const avs_server *S)
{
    static const bool largest = [] {
        const char *e = getenv("AVSVD_SERVER_ORDER");
        return e && strcmp(e, "largest") == 0;
    }();
    size_t pick = 0;
    if (largest)
        for (size_t i = 1; i < S->pending.size(); ++i)
            if (S->pending[i]->N > S->pending[pick]->N) pick = i;
    return pick;
}

// memcpy split over up to 8 threads for large copies (the pinned staging copy
// of a large request runs on the device thread)
void copy_parallel(void *dst, const void *src, size_t bytes)
{
    const size_t chunk = (size_t)32 << 20;
    const int nt = (int)std::min<size_t>(8, bytes / chunk);
    if (nt <= 1) {
        memcpy(dst, src, bytes);
        return;
    }
    const size_t per = (bytes / nt + 4095) & ~(size_t)4095;
    std::vector<std::thread> th;
    for (int i = 0; i < nt; ++i) {
        const size_t a = std::min(bytes, per * i), e = std::min(bytes, per * (i + 1));
        if (a >= e) continue;
        try {
            th.emplace_back([=] { memcpy((char *)dst + a, (const char *)src + a, e - a); });
        } catch (...) { // no thread to be had: copy this part here
            memcpy((char *)dst + a, (const char *)src + a, e - a);
        }
    }
    for (std::thread &t : th) t.join();
}

Slot *slot_of(avs_server *S, uint64_t key)
{
    for (Slot &x : S->pool)
        if (x.busy && x.last_use == key) return &x;
    return nullptr;
}

void destroy_slot(Slot &x)
{
    if (x.h) avs_destroy(x.h);
    if (x.ev) hipEventDestroy(x.ev);
    if (x.hpin) hipHostFree(x.hpin);
    x.h = nullptr;
    x.ev = nullptr;
    x.hpin = nullptr;
}

// AVSVD_SERVER_TRACE=1: one line per finished request on stderr
bool trace_on()
{
    static const bool on = [] {
        const char *e = getenv("AVSVD_SERVER_TRACE");
        return e && e[0] == '1';
    }();
    return on;
}

// marks a request done (with its status) and wakes its waiter
void finish(avs_server *S, const std::shared_ptr<Request> &r)
{
    {
        std::lock_guard<std::mutex> lk(S->mu);
        r->t_done = now_ms();
        if (trace_on())
            fprintf(stderr,
                    "avs_server: request %llu n %lld N %lld %s b %d: queued %.1f ms, set-up %.1f ms%s, device %.1f ms, "
                    "host stage %.1f ms, status %d\n",
                    (unsigned long long)r->id, (long long)r->n, (long long)r->N, r->dtype == AVS_F64 ? "f64" : "f32",
                    r->b, r->t_admit - r->t_arrive, r->t_started - r->t_admit, r->new_handle ? " (new handle)" : "",
                    r->t_band - r->t_started, r->t_done - r->t_band, (int)r->status);
        r->done = true;
        if (!r->status) {
            S->lat.push_back(r->t_done - r->t_arrive);
            S->t_last = std::max(S->t_last, r->t_done);
            S->completed++;
        }
    }
    S->cv_done.notify_all();
}

// In flight on the device: the request, its slot key (slots are found by key:
// only this thread changes the pool, but admit's push_back can move it), the
// band's pinned host buffer
struct Running {
    std::shared_ptr<Request> r;
    uint64_t key = 0;
    const double *band = nullptr;
};

// Enqueues request r on slot x: handle (created on first use), upload, band
// reduction, band download, completion event.  Device thread only.
avs_status start(avs_server *S, Slot &x, Running &run)
{
    const Request &r = *run.r;
    avs_status st = AVS_OK;
    run.r->new_handle = !x.h;
    if (!x.h) {
        avs_opts o;
        avs_default_opts(&o);
        o.update_mode = S->opts.update_mode;
        if ((st = avs_create(&x.h, r.dtype, r.N, r.b, &o))) return st;
        avs::handle_use_own_panel_stream(x.h);
        if (hipEventCreateWithFlags(&x.ev, hipEventDisableTiming) != hipSuccess)
            return fail(AVS_ERR_HIP, "server: event creation failed");
        const size_t es = r.dtype == AVS_F64 ? 8 : 4;
        if (hipHostMalloc(&x.hpin, es * (size_t)r.N * r.N, hipHostMallocDefault) != hipSuccess)
            return fail(AVS_ERR_ALLOC, "server: pinned staging of %zu bytes", es * (size_t)r.N * r.N);
    }
    // the slot is idle, so its previous upload from hpin has completed
    copy_parallel(x.hpin, r.A.data(), r.A.size());
    if ((st = avs_set_matrix_host_sub(x.h, x.hpin, r.n, r.n))) return st;
    if ((st = avs_blockbidiag(x.h, 0))) return st;
    if ((st = avs::band_enqueue(x.h, &run.band))) return st;
    if (hipEventRecord(x.ev, (hipStream_t)avs::handle_stream(x.h)) != hipSuccess)
        return fail(AVS_ERR_HIP, "server: event record failed");
    run.r->t_started = now_ms();
    return AVS_OK;
}

void device_thread(avs_server *S)
{
    std::vector<Running> running;
    std::vector<Bisect> bis(S->opts.device_sigma ? std::max(1, (int)S->opts.max_inflight) : 0);
    for (Bisect &q : bis) {
        const size_t nm = (size_t)S->nmax;
        if (hipStreamCreateWithFlags(&q.s, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&q.ev, hipEventDisableTiming) != hipSuccess ||
            hipMalloc(&q.dT2, sizeof(double) * 2 * nm) != hipSuccess ||
            hipMalloc(&q.dSig, sizeof(double) * nm) != hipSuccess ||
            hipHostMalloc((void **)&q.hT2, sizeof(double) * 2 * nm, hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc((void **)&q.hSig, sizeof(double) * nm, hipHostMallocDefault) != hipSuccess) {
            std::lock_guard<std::mutex> lk(S->mu);
            S->dev_err = "server: bisection buffers could not be allocated";
            S->opts.device_sigma = 0; // host workers fall back to dbdsdc
            break;
        }
    }
    for (;;) {
        std::vector<Slot> evict;
        std::vector<Running> starting;
        std::vector<std::shared_ptr<Request>> sig_new;
        bool stop;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            // sleep while there is nothing to admit and nothing in flight
            const bool busy_dev = !running.empty();
            bool busy_bis = false;
            for (const Bisect &q : bis) busy_bis |= (bool)q.r;
            if (!busy_dev && !busy_bis && !S->stop && S->sigma_q.empty()) {
                S->cv_dev.wait(lk, [&] { return S->stop || !S->pending.empty() || !S->sigma_q.empty(); });
            }
            stop = S->stop;
            while (!stop && !S->pending.empty() && (int)(running.size() + starting.size()) < S->opts.max_inflight) {
                const size_t pick = next_pending(S);
                const uint64_t key = admit(S, *S->pending[pick], evict);
                if (!key) break;
                Running run;
                run.r = S->pending[pick];
                run.key = key;
                run.r->t_admit = now_ms();
                S->pending.erase(S->pending.begin() + (long)pick);
                starting.push_back(run);
            }
            for (Bisect &q : bis)
                if (!q.r && !S->sigma_q.empty()) {
                    q.r = S->sigma_q.front();
                    S->sigma_q.pop_front();
                    sig_new.push_back(q.r);
                }
        }
        for (Slot &x : evict) destroy_slot(x);
        if (stop) break;
        for (Running &run : starting) {
            Slot *x;
            {
                std::lock_guard<std::mutex> lk(S->mu);
                x = slot_of(S, run.key);
            }
            avs_status st = start(S, *x, run);
            if (st) {
                run.r->status = st;
                run.r->err = avs_last_error();
                std::vector<Slot> bad;
                {
                    std::lock_guard<std::mutex> lk(S->mu);
                    Slot *y = slot_of(S, run.key);
                    bad.push_back(*y); // a failed handle is not reused
                    S->used -= y->bytes;
                    S->pool.erase(S->pool.begin() + (y - S->pool.data()));
                }
                for (Slot &z : bad) destroy_slot(z);
                std::vector<char>().swap(run.r->A);
                finish(S, run.r);
                continue;
            }
            running.push_back(run);
        }
        // new bisections
        for (Bisect &q : bis) {
            if (!q.r || std::find(sig_new.begin(), sig_new.end(), q.r) == sig_new.end()) continue;
            Request &r = *q.r;
            avs_status st = avs::bidiag_sigma_enqueue(r.N, r.N, r.d.data(), r.e.data(), q.dT2, q.dSig, q.hT2, q.hSig,
                                                      q.s);
            if (!st && hipEventRecord(q.ev, q.s) != hipSuccess) st = fail(AVS_ERR_HIP, "server: event record failed");
            if (st) {
                r.status = st;
                r.err = avs_last_error();
                std::shared_ptr<Request> done = q.r;
                q.r.reset();
                finish(S, done);
            }
        }
        // poll the reductions in flight
        bool progressed = !starting.empty() || !sig_new.empty();
        for (size_t i = 0; i < running.size();) {
            Running &run = running[i];
            Slot *x;
            {
                std::lock_guard<std::mutex> lk(S->mu);
                x = slot_of(S, run.key);
            }
            const hipError_t q = hipEventQuery(x->ev);
            if (q == hipErrorNotReady) {
                ++i;
                continue;
            }
            Request &r = *run.r;
            if (q != hipSuccess) {
                r.status = AVS_ERR_HIP;
                r.err = std::string("server: reduction failed: ") + hipGetErrorString(q);
            } else {
                r.band.assign(run.band, run.band + (size_t)(r.b + 1) * r.N);
            }
            r.t_band = now_ms();
            std::vector<char>().swap(r.A);
            {
                std::lock_guard<std::mutex> lk(S->mu);
                Slot *y = slot_of(S, run.key);
                y->busy = false;
                if (!r.status) S->host_q.push_back(run.r);
            }
            if (r.status) finish(S, run.r);
            else S->cv_host.notify_one();
            running.erase(running.begin() + (long)i);
            progressed = true;
        }
        // poll the bisections in flight
        for (Bisect &q : bis) {
            if (!q.r) continue;
            const hipError_t e = hipEventQuery(q.ev);
            if (e == hipErrorNotReady) continue;
            Request &r = *q.r;
            if (e != hipSuccess) {
                r.status = AVS_ERR_HIP;
                r.err = std::string("server: bisection failed: ") + hipGetErrorString(e);
            } else {
                r.sigma.assign(q.hSig, q.hSig + r.n); // drop the padding's zeros
            }
            std::shared_ptr<Request> done = q.r;
            q.r.reset();
            finish(S, done);
            progressed = true;
        }
        if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(100));
    }
    // stop: wait for what is in flight, then release the pool and buffers
    for (Running &run : running) {
        Slot *x = slot_of(S, run.key);
        if (x && x->ev) hipEventSynchronize(x->ev);
    }
    for (Bisect &q : bis) {
        if (q.s) hipStreamSynchronize(q.s);
        hipFree(q.dT2);
        hipFree(q.dSig);
        if (q.hT2) hipHostFree(q.hT2);
        if (q.hSig) hipHostFree(q.hSig);
        if (q.ev) hipEventDestroy(q.ev);
        if (q.s) hipStreamDestroy(q.s);
    }
    for (Slot &x : S->pool) destroy_slot(x);
    S->pool.clear();
}

void host_worker(avs_server *S)
{
    for (;;) {
        std::shared_ptr<Request> r;
        bool dev_sigma;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_host.wait(lk, [&] { return S->stop || !S->host_q.empty(); });
            if (S->host_q.empty()) return;
            r = S->host_q.front();
            S->host_q.pop_front();
            dev_sigma = S->opts.device_sigma != 0;
        }
        const int64_t N = r->N;
        r->d.assign(N, 0.0);
        r->e.assign(std::max<int64_t>(N - 1, 1), 0.0);
        // one host thread per request, four from n = 4096 and
        // AVSVD_SERVER_BIG_THREADS (default 8) from n = 8192: the largest
        // requests set the p99, and their sweeps pipeline well.  48 mixed
        // requests on a 16-core host: Poisson p99 1.08-1.20 s with 8 threads
        // against 1.60-1.70 s with 4 and 0.96 s with 12, at 13.9-14.7 solves/s
        // saturated against 14.2-15.0 (4) and 13.8 (12); every host thread was
        // slower still (the overlapping large requests oversubscribe the cores).
        static const int big_threads = [] {
            const char *e = getenv("AVSVD_SERVER_BIG_THREADS");
            const int v = e ? atoi(e) : 0;
            return v > 0 ? v : 8;
        }();
        avs_status st = avs::bidiagonalize_band_one(N, r->b, r->band.data(), r->b + 1, r->d.data(), r->e.data(),
                                                    N >= 8192 ? big_threads : (N >= 4096 ? 4 : 1));
        std::vector<double>().swap(r->band);
        if (!st && dev_sigma && N <= S->nmax && avs::bisect_on_device(N)) {
            // the bisection runs on the device thread
            {
                std::lock_guard<std::mutex> lk(S->mu);
                S->sigma_q.push_back(r);
            }
            S->cv_dev.notify_all();
            continue;
        }
        if (!st) {
            std::vector<double> sg(N);
            st = avs_bdsdc_values(N, r->d.data(), r->e.data(), sg.data());
            if (!st) r->sigma.assign(sg.begin(), sg.begin() + r->n); // drop the padding's zeros
        }
        if (st) {
            r->status = st;
            r->err = avs_last_error();
        }
        finish(S, r);
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
    for (int32_t r : S->opts.reserved)
        if (r != 0) return fail(-2, "opts.reserved must be zero");
    if (S->opts.max_inflight <= 0) S->opts.max_inflight = 4;
    if (S->opts.host_threads <= 0) S->opts.host_threads = avs::host_thread_count();
    // the largest padded order the budget admits (FP32, the smallest
    // footprint per element), which sizes the bisection buffers
    int64_t nmax = 128;
    while (footprint_of(AVS_F32, 2 * nmax, 128) <= (size_t)S->opts.hbm_budget_bytes) nmax *= 2;
    while (footprint_of(AVS_F32, nmax + 128, 128) <= (size_t)S->opts.hbm_budget_bytes) nmax += 128;
    S->nmax = nmax;
    S->dev_thread = std::thread(device_thread, S.get());
    for (int i = 0; i < S->opts.host_threads; ++i) S->host_workers.emplace_back(host_worker, S.get());
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
    r->N = bucket_of(n, r->b);
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
    for (auto &t : S->host_workers) t.join();
    S->dev_thread.join();
    delete S;
    return AVS_OK;
}

} // extern "C"
