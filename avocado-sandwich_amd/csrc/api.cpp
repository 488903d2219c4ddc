// api.cpp -- the C ABI of libavsvd.so (declared in include/avsvd.h).
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <atomic>
#include <cmath>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "avsvd.h"
#include "internal.hpp"
#include "kernels.hpp"
#include "schedule.hpp"

using namespace avs;

namespace avs {

thread_local std::string g_err;

avs_status fail(avs_status code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

} // namespace avs

#define HIPCHK(call)                                                                                                   \
    do {                                                                                                               \
        hipError_t e_ = (call);                                                                                        \
        if (e_ != hipSuccess) return fail(AVS_ERR_HIP, "%s failed: %s (%d)", #call, hipGetErrorString(e_), (int)e_); \
    } while (0)

namespace {

bool supported_b(int b) { return b == 32 || b == 64 || b == 128; }
size_t esize(int dtype) { return dtype == AVS_F64 ? 8 : 4; }

template <typename F> avs_status dispatch_tb(int dtype, int b, F &&f)
{
    if (dtype == AVS_F64) {
        if (b == 32) return f((double *)0, std::integral_constant<int, 32>());
        if (b == 64) return f((double *)0, std::integral_constant<int, 64>());
        if (b == 128) return f((double *)0, std::integral_constant<int, 128>());
    } else if (dtype == AVS_F32) {
        if (b == 32) return f((float *)0, std::integral_constant<int, 32>());
        if (b == 64) return f((float *)0, std::integral_constant<int, 64>());
        if (b == 128) return f((float *)0, std::integral_constant<int, 128>());
    }
    return fail(AVS_ERR_UNSUPPORTED, "dtype %d with tile size %d is not built", dtype, b);
}

struct DevProgram {
    Program P;
    Task *dTasks = nullptr;
    int *dDepTile = nullptr, *dDepVal = nullptr, *dSig = nullptr;
    int device = -1;
    ~DevProgram()
    {
        hipFree(dTasks);
        hipFree(dDepTile);
        hipFree(dDepVal);
        hipFree(dSig);
    }
};

// Task DAGs depend only on (device, nt, b, alg, kend, tree, domain, update
// tiles, workers); handles of the same shape share one device copy.
struct ProgKey {
    int device, nt, b, alg, kend, tree, domain, ut, workers;
    bool operator<(const ProgKey &o) const
    {
        return std::tie(device, nt, b, alg, kend, tree, domain, ut, workers) <
               std::tie(o.device, o.nt, o.b, o.alg, o.kend, o.tree, o.domain, o.ut, o.workers);
    }
};
std::mutex g_prog_mu;
std::map<ProgKey, std::shared_ptr<DevProgram>> g_progs;

struct DevPhases {
    PhaseProgram P;
    ChainOp *dOps = nullptr;
    int *dChain = nullptr;
    ~DevPhases()
    {
        hipFree(dOps);
        hipFree(dChain);
    }
};
std::map<ProgKey, std::shared_ptr<DevPhases>> g_phases;

} // namespace

struct avs_handle {
    int dtype = AVS_F64;
    int64_t n = 0;
    int b = 0, nt = 0;
    avs_opts opts{};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    void *dTiles = nullptr;
    void *dStage = nullptr; // column-major staging (n*n elements)
    double *dTau = nullptr, *dTjj = nullptr, *dBand = nullptr;
    double *hBand = nullptr; // pinned host staging for the band (b+1) x n
    // band store ((3b+1) x n), its diagonal and superdiagonal, and the access
    // flag of the chase on the device
    double *dStore = nullptr, *dDE = nullptr;
    int *dChaseErr = nullptr;
    int *dCnt = nullptr, *dTicket = nullptr, *dStatus = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // what the tiles hold: nothing yet, the uploaded matrix, the reduced
    // matrix after band_k steps of Blockbidiag! (band in place, reflectors
    // where the reference writes exact zeros), or R and reflectors after
    // tiled_QR!
    enum { ST_EMPTY, ST_MATRIX, ST_BAND, ST_QR } state = ST_EMPTY;
    int band_k = 0;
    // bisection for the bidiagonal's singular values: device buffers (the
    // Golub-Kahan tridiagonal's squared off-diagonals, 2n; sigma, n) and their
    // pinned host staging, allocated on first use and kept with the handle
    double *dT2 = nullptr, *dSig = nullptr, *hT2 = nullptr, *hSig = nullptr;
    bool timed = false;
    int workers = 0, device = 0;
    long long wallclock_khz = 100000;
    std::shared_ptr<DevProgram> prog;
    std::shared_ptr<DevPhases> phases;
    // per-launch timing of the stream-ordered executor (opts.phase_timing != 0):
    // pev[i] .. pev[i+1] brackets launch i, whose kind is pkind[i]
    std::vector<hipEvent_t> pev;
    std::vector<int> pkind, pmode;
    std::vector<double> pflops;
    int plaunches = 0;
    bool ptimed = false; // the last stream-ordered run recorded per-launch events
    // the stream-ordered run captured once as a HIP graph and replayed, and
    // the phase program it was captured from
    hipGraphExec_t gexec = nullptr;
    std::shared_ptr<DevPhases> gphases;
    // panel stream of the stream-ordered executor and one completion event per
    // phase (the panels of a sweep run ahead of that sweep's updates)
    hipStream_t pstream = nullptr;
    int pslot = -1;  // the pool slot of the last run's panel stream (-1: none, or the handle's own)
    bool panel_pool = true; // panel streams from the device's pool when AVSVD_PANEL_QUEUES > 0
    std::vector<hipEvent_t> done;
    // update stream whose CU mask leaves opts.reserve_cus CUs to the panels,
    // and the events that join it to the handle's stream
    hipStream_t ustream = nullptr;
    hipEvent_t ujoin[2] = {nullptr, nullptr};
};

namespace {

// The panel stream.  By default every handle has an ordinary stream of its
// own.  The HIP runtime maps ordinary streams onto a few shared hardware
// queues (GPU_MAX_HW_QUEUES, 4 by default); once a process holds more streams
// than that, a handle's panel and update streams can land on one queue, where
// every panel waits behind the updates enqueued before it (measured at
// n=16384 FP64 b=128, two handles in one process: the second took 471 ms, the
// one-stream time, instead of 355).  With AVSVD_PANEL_QUEUES=k (1..4) panel
// streams come instead from a per-device pool of k streams created with a
// CU mask that enables every CU, each of which owns a hardware queue.  Each run
// takes an idle pool stream, starting after the one taken last, so up to k runs
// in flight get a queue each (the two handles above: 355 and 356 ms).  With
// every pool stream busy, runs share them in turn; their panels then
// serialise with another run's panels, never with their own updates.  A run's
// first panel waits for every earlier update of its handle, so runs of one
// handle stay ordered whichever streams they take.  The pool's streams live
// until the process ends.
//
// The pool is not the default because dedicated queues slow a process that
// runs many small reductions concurrently, even when they sit idle.  The
// 48-request serving mix ran 9.8–11.0 solves/s with 1–4 pool queues in the
// process (its own handles on ordinary streams or not), against 14.1 without
// any.  Raising GPU_MAX_HW_QUEUES to 8 fixes the two-handle case as well (both
// 357 ms).  The serving loop's handles always take ordinary streams
// (handle_use_own_panel_stream).
// At most 4: a sixth dedicated queue beside the runtime's four made its
// handle slower again (428 ms).
constexpr int kPanelQueuesMax = 4;
int panel_queues()
{
    static const int k = [] {
        const char *e = getenv("AVSVD_PANEL_QUEUES");
        const int v = e ? atoi(e) : 0;
        return v < 0 ? 0 : (v > kPanelQueuesMax ? kPanelQueuesMax : v);
    }();
    return k;
}
struct PanelPool {
    std::vector<hipStream_t> s;
    int next = 0;
};
std::mutex g_ppool_mu;
std::map<int, PanelPool> g_ppool;

// the panel stream of the run about to be enqueued (a handle's own stream,
// pslot < 0, is kept)
avs_status acquire_panel_stream(avs_handle *h)
{
    if (h->pstream && h->pslot < 0) return AVS_OK;
    if (!h->panel_pool || panel_queues() == 0) {
        HIPCHK(hipStreamCreateWithFlags(&h->pstream, hipStreamNonBlocking));
        return AVS_OK;
    }
    std::lock_guard<std::mutex> lk(g_ppool_mu);
    PanelPool &P = g_ppool[h->device];
    const int ns = (int)P.s.size();
    int slot = -1;
    for (int k = 0; k < ns && slot < 0; ++k) {
        const int i = (P.next + k) % ns;
        if (hipStreamQuery(P.s[i]) == hipSuccess) slot = i; // nothing pending
    }
    if (slot < 0 && ns < panel_queues()) {
        int ncu = 0;
        HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, h->device));
        // the size argument counts 32-bit words; ncu words covers every CU
        std::vector<uint32_t> mask(ncu, 0);
        for (int i = 0; i < ncu; ++i) mask[i / 32] |= 1u << (i % 32);
        hipStream_t st = nullptr;
        HIPCHK(hipExtStreamCreateWithCUMask(&st, (uint32_t)ncu, mask.data()));
        P.s.push_back(st);
        slot = ns;
    }
    if (slot < 0) slot = P.next % ns;
    P.next = (slot + 1) % (int)P.s.size();
    h->pstream = P.s[slot];
    h->pslot = slot;
    return AVS_OK;
}

// the handle stops using the pool's streams
void release_panel_stream(avs_handle *h)
{
    if (h->pslot < 0) return;
    h->pslot = -1;
    h->pstream = nullptr;
}

int resolve_tree(const avs_handle *h) { return h->opts.tree; }
// 0 = the automatic per-sweep rule of domain_for: two tiles per domain where
// the panel chain bounds a sweep (the chain is then 2 + log2(m/2) panels long,
// the shortest for m >= 4 up to one), four where its updates do (the update
// flops are the same for every domain size; larger domains put more of them
// in the TSQRT updates, the fastest)
int resolve_domain(const avs_handle *h) { return h->opts.domain > 0 ? h->opts.domain : 0; }

avs_status ensure_program(avs_handle *h, int alg, int kend)
{
    const int tree = resolve_tree(h), dom = resolve_domain(h), ut = upd_tiles_for(h->b);
    const ProgKey key{h->device, h->nt, h->b, alg, kend, tree, dom, ut, h->workers};
    if (h->prog && h->prog->P.alg == alg && h->prog->P.kend == kend && h->prog->P.tree == tree &&
        h->prog->P.domain == dom)
        return AVS_OK;
    std::lock_guard<std::mutex> lk(g_prog_mu);
    auto it = g_progs.find(key);
    if (it != g_progs.end()) {
        h->prog = it->second;
        return AVS_OK;
    }
    auto dp = std::make_shared<DevProgram>();
    dp->device = h->device;
    build_program(dp->P, h->nt, h->b, alg, kend, tree, dom, ut, h->workers);
    Program &P = dp->P;
    std::vector<Task> ordered(P.order.size());
    for (size_t i = 0; i < P.order.size(); ++i) ordered[i] = P.prog[P.order[i]];
    const size_t nd = std::max<size_t>(P.dep_tile.size(), 1), ns = std::max<size_t>(P.sig_tile.size(), 1);
    HIPCHK(hipMalloc(&dp->dTasks, sizeof(Task) * std::max<size_t>(ordered.size(), 1)));
    HIPCHK(hipMalloc(&dp->dDepTile, sizeof(int) * nd));
    HIPCHK(hipMalloc(&dp->dDepVal, sizeof(int) * nd));
    HIPCHK(hipMalloc(&dp->dSig, sizeof(int) * ns));
    if (!ordered.empty())
        HIPCHK(hipMemcpy(dp->dTasks, ordered.data(), sizeof(Task) * ordered.size(), hipMemcpyHostToDevice));
    if (!P.dep_tile.empty()) {
        HIPCHK(hipMemcpy(dp->dDepTile, P.dep_tile.data(), sizeof(int) * P.dep_tile.size(), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dp->dDepVal, P.dep_val.data(), sizeof(int) * P.dep_val.size(), hipMemcpyHostToDevice));
    }
    if (!P.sig_tile.empty())
        HIPCHK(hipMemcpy(dp->dSig, P.sig_tile.data(), sizeof(int) * P.sig_tile.size(), hipMemcpyHostToDevice));
    g_progs[key] = dp;
    h->prog = dp;
    return AVS_OK;
}

avs_status ensure_phases(avs_handle *h, int alg, int kend)
{
    const int tree = resolve_tree(h), dom = resolve_domain(h);
    const ProgKey key{h->device, h->nt, h->b, alg, kend, tree, dom, 0, 0};
    if (h->phases && h->phases->P.alg == alg && h->phases->P.kend == kend && h->phases->P.tree == tree &&
        h->phases->P.domain == dom)
        return AVS_OK;
    std::lock_guard<std::mutex> lk(g_prog_mu);
    auto it = g_phases.find(key);
    if (it != g_phases.end()) {
        h->phases = it->second;
        return AVS_OK;
    }
    auto dp = std::make_shared<DevPhases>();
    build_phase_program(dp->P, h->nt, h->b, alg, kend, tree, dom);
    const auto &ops = dp->P.ops;
    const auto &cs = dp->P.chain_start;
    HIPCHK(hipMalloc(&dp->dOps, sizeof(ChainOp) * std::max<size_t>(ops.size(), 1)));
    HIPCHK(hipMalloc(&dp->dChain, sizeof(int) * std::max<size_t>(cs.size(), 1)));
    if (!ops.empty()) HIPCHK(hipMemcpy(dp->dOps, ops.data(), sizeof(ChainOp) * ops.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dp->dChain, cs.data(), sizeof(int) * cs.size(), hipMemcpyHostToDevice));
    g_phases[key] = dp;
    h->phases = dp;
    return AVS_OK;
}

int split_of(const avs_handle *h) { return h->dtype == AVS_F32 && h->opts.update_mode == AVS_UPD_SPLIT_BF16; }

template <typename TS, int B> avs_status run_host_enqueued(avs_handle *h)
{
    const Program &P = h->prog->P;
    const long tsz = (long)B * B;
    TS *tiles = (TS *)h->dTiles;
    size_t i = 0;
    while (i < P.prog.size()) {
        const Task &t = P.prog[i];
        double *tau = h->dTau + (long)t.fac * B;
        double *tjj = h->dTjj + (long)t.fac * 16 * B;
        if (t.kind == TK_PANEL) {
            launch_panel<TS, B>(t.mode, tiles + t.vtile * tsz, B, tiles + t.btile * tsz, B, t.trans, tau, tjj,
                                h->stream);
            ++i;
            continue;
        }
        // group the consecutive update tasks of one panel op into one launch
        size_t j = i;
        int total = 0;
        while (j < P.prog.size() && P.prog[j].kind == TK_UPDATE && P.prog[j].fac == t.fac) {
            total += P.prog[j].ntiles;
            ++j;
        }
        const long st = (long)t.cstep * tsz;
        launch_update<TS, B>(t.mode, true, tiles + t.vtile * tsz, B, t.trans, tjj, tiles + t.c1tile * tsz, B, st,
                             tiles + t.btile * tsz, B, st, t.trans, total, t.ntiles, split_of(h), h->stream);
        i = j;
    }
    HIPCHK(hipGetLastError());
    return AVS_OK;
}

// Algorithmic flops of an update phase: every op applies B reflectors to
// ncols column tiles; a TS op (full B x B V) costs 4B^3 per tile (W = C1 + V^T C2,
// C2 -= V W), a GE or TT op (triangular V) 2B^3.
double phase_update_flops(const PhaseProgram &P, const Phase &ph)
{
    const double b3 = (double)P.B * P.B * P.B;
    double f = 0.0;
    for (int c = ph.chain0; c < ph.chain0 + ph.nchains; ++c)
        for (int o = P.chain_start[c]; o < P.chain_start[c + 1]; ++o)
            f += (P.ops[o].mode == MODE_TS ? 4.0 : 2.0) * b3 * P.ops[o].ncols;
    return f;
}

// Stream-ordered execution.  Updates run in program order on the handle's
// stream.  Unless opts.single_stream is set, panels run in program order on a
// second stream: within a sweep the panel chain (GEQRT, TSQRT chains, merge
// levels) does not depend on that sweep's updates, so it runs ahead of them;
// each update phase waits for its own panel phase, and the first panel phase
// of a sweep waits for every update before it.  With opts.phase_timing set,
// every launch is bracketed by timing events on its stream.
template <typename TS, int B> avs_status run_streams(avs_handle *h)
{
    const DevPhases &D = *h->phases;
    const int per = upd_tiles_for(B);
    const bool timing = h->opts.phase_timing != 0;
    const bool two = h->opts.single_stream == 0;
    const size_t nph = D.P.phases.size();
    h->plaunches = 0;
    h->ptimed = timing;
    if (timing) {
        while (h->pev.size() < 2 * nph) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            h->pev.push_back(e);
        }
        h->pkind.assign(nph, 0);
        h->pmode.assign(nph, 0);
        h->pflops.assign(nph, 0.0);
    }
    if (two) {
        if (avs_status st = acquire_panel_stream(h)) return st;
        while (h->done.size() < nph + 1) {
            hipEvent_t e;
            HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            h->done.push_back(e);
        }
    }
    hipStream_t us = h->stream, ps = two ? h->pstream : h->stream;
    const int reserve = two ? h->opts.reserve_cus : 0;
    if (reserve > 0) {
        if (!h->ustream) {
            int ncu = 0;
            HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, h->device));
            // at most ncu/8 - ncu/32 CUs (24 of 256): even if every cleared bit
            // fell on one of the 8 XCDs, it keeps a quarter of its CUs
            if (reserve > ncu / 8 - ncu / 32) return fail(AVS_ERR_UNSUPPORTED, "too many reserved CUs");
            // the size argument counts 32-bit words; ncu words (zero past the
            // first ncu bits) is in bounds whichever way it is read
            std::vector<uint32_t> mask(ncu, 0);
            for (int i = 0; i < ncu; ++i) mask[i / 32] |= 1u << (i % 32);
            // mask bit i is a CU of XCD i % 8 (measured on MI355X): the last
            // `reserve` bits take the same number of CUs from every XCD, which
            // matters because workgroups are dealt to the XCDs round-robin
            for (int r = 0; r < reserve; ++r) {
                const int bit = ncu - 1 - r;
                mask[bit / 32] &= ~(1u << (bit % 32));
            }
            HIPCHK(hipExtStreamCreateWithCUMask(&h->ustream, (uint32_t)ncu, mask.data()));
            for (hipEvent_t &e : h->ujoin) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        }
        HIPCHK(hipEventRecord(h->ujoin[0], h->stream));
        HIPCHK(hipStreamWaitEvent(h->ustream, h->ujoin[0], 0));
        us = h->ustream;
    }
    // done[nph]: the latest update (or the matrix upload) the next sweep's panels wait for
    hipEvent_t last_update = two ? h->done[nph] : nullptr;
    if (two) HIPCHK(hipEventRecord(last_update, us));
    // with reserved CUs, the update modes in bit mask opts.masked_modes (bit m:
    // mode m; 0 = all) run on the CU-masked stream, the others on the
    // handle's stream, each after the previous update
    const int rmask = h->opts.masked_modes ? h->opts.masked_modes : 7;
    hipStream_t prev_us = us;
    int last_panel = -1;
    for (size_t i = 0; i < nph; ++i) {
        const Phase &ph = D.P.phases[i];
        hipStream_t ust = (us != h->stream && !((rmask >> ph.mode) & 1)) ? h->stream : us;
        hipStream_t st = ph.kind == TK_PANEL ? ps : ust;
        if (two) {
            if (ph.kind == TK_PANEL && ph.mode == MODE_GE) HIPCHK(hipStreamWaitEvent(ps, last_update, 0));
            if (ph.kind == TK_UPDATE && ust != prev_us) HIPCHK(hipStreamWaitEvent(ust, last_update, 0));
            if (ph.kind == TK_UPDATE && last_panel >= 0) HIPCHK(hipStreamWaitEvent(ust, h->done[last_panel], 0));
        }
        if (ph.kind == TK_UPDATE) prev_us = ust;
        if (timing) {
            h->pkind[i] = ph.kind;
            h->pmode[i] = ph.mode;
            h->pflops[i] = ph.kind == TK_UPDATE ? phase_update_flops(D.P, ph) : 0.0;
            HIPCHK(hipEventRecord(h->pev[2 * i], st));
        }
        if (ph.kind == TK_PANEL) {
            launch_chain_panel<TS, B>(ph.nchains, ph.mode, h->dTiles, h->dTau, h->dTjj, D.dOps, D.dChain, ph.chain0,
                                      st);
            last_panel = (int)i;
        } else {
            const int groups = (ph.ncols + per - 1) / per;
            launch_chain_update<TS, B>(ph.nchains, groups, per, ph.mode, ph.trans, h->dTiles, h->dTjj, D.dOps,
                                       D.dChain, ph.chain0, split_of(h), st);
        }
        if (timing) HIPCHK(hipEventRecord(h->pev[2 * i + 1], st));
        if (two) {
            if (ph.kind == TK_PANEL) HIPCHK(hipEventRecord(h->done[i], ps));
            else HIPCHK(hipEventRecord(last_update, ust));
        }
        ++h->plaunches;
    }
    // the handle's stream ends after the last panel and the last update too
    if (two && last_panel >= 0) HIPCHK(hipStreamWaitEvent(h->stream, h->done[last_panel], 0));
    if (us != h->stream) {
        HIPCHK(hipEventRecord(h->ujoin[1], us));
        HIPCHK(hipStreamWaitEvent(h->stream, h->ujoin[1], 0));
    }
    HIPCHK(hipGetLastError());
    return AVS_OK;
}

// The stream-ordered run replayed from a HIP graph: the first run of a phase
// program is captured from the handle's stream (the panel stream joins the
// capture through the events it waits on, and rejoins before the capture
// ends; the second stream and the events are created before the capture) and
// every run replays it.  The graph has kernel and event nodes only, no
// memset.  Only with opts.graph != 0 and without per-launch timing
// (phase_timing), one stream (single_stream) or the CU-masked update stream
// (reserve_cus): replaying the n=16384, b=128 FP64 graph (3571 kernels) took
// 527-544 ms against 460-464 ms for the eager enqueue on the same MI355X, so
// eager is the default.
template <typename TS, int B> avs_status run_streams_graph(avs_handle *h)
{
    const bool graph = h->opts.phase_timing == 0 && h->opts.single_stream == 0 && h->opts.reserve_cus == 0 &&
                       h->opts.graph != 0;
    if (!graph) return run_streams<TS, B>(h);
    if (h->gphases != h->phases) {
        if (h->gexec) HIPCHK(hipGraphExecDestroy(h->gexec));
        h->gexec = nullptr;
        h->gphases = h->phases;
    }
    if (!h->gexec) {
        const size_t nph = h->phases->P.phases.size();
        // a capture joins the panel stream to the graph: it must be the
        // handle's own (a pooled stream may carry other handles' launches)
        if (h->pslot >= 0) release_panel_stream(h);
        if (!h->pstream) HIPCHK(hipStreamCreateWithFlags(&h->pstream, hipStreamNonBlocking));
        while (h->done.size() < nph + 1) {
            hipEvent_t e;
            HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            h->done.push_back(e);
        }
        HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        const avs_status st = run_streams<TS, B>(h);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(h->stream, &g);
        if (st) {
            if (ec == hipSuccess && g) hipGraphDestroy(g);
            return st;
        }
        HIPCHK(ec);
        const hipError_t ei = hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (ei != hipSuccess) h->gexec = nullptr;
        HIPCHK(ei);
    }
    HIPCHK(hipGraphLaunch(h->gexec, h->stream));
    return AVS_OK;
}

} // namespace

namespace avs {

// Singular values of a bidiagonal by bisection on the device (one lane per
// value): enqueues on stream `sv` the upload of the Golub-Kahan tridiagonal's
// squared off-diagonals (through the caller's pinned hT2, 2n doubles, to dT2),
// the bisection into dSig (n doubles) and the download of the nkeep largest
// into the caller's pinned hSig; valid once the stream reaches this point.
// Allocates nothing.
avs_status bidiag_sigma_enqueue(int64_t n, int64_t nkeep, const double *d, const double *e, double *dT2,
                                double *dSig, double *hT2, double *hSig, void *sv)
{
    hipStream_t s = (hipStream_t)sv;
    if (n <= 0 || nkeep <= 0 || nkeep > n) return fail(-1, "bad sizes n=%lld nkeep=%lld", (long long)n,
                                                       (long long)nkeep);
    // squared off-diagonals of the Golub-Kahan tridiagonal, its Gershgorin
    // bound and LAPACK's pivmin
    double bnd = 0.0, tmx = 0.0, prev = 0.0;
    for (int64_t k = 0; k < 2 * n - 1; ++k) {
        const double t = std::fabs((k & 1) ? e[k >> 1] : d[k >> 1]);
        hT2[k] = t * t;
        bnd = std::max(bnd, prev + t);
        tmx = std::max(tmx, hT2[k]);
        prev = t;
    }
    bnd = std::max(bnd, prev);
    const double hi0 = bnd * (1.0 + 1e-14) + 1e-300;
    const double pivmin = 2.2250738585072014e-308 * std::max(1.0, tmx);
    HIPCHK(hipMemcpyAsync(dT2, hT2, sizeof(double) * (size_t)(2 * n - 1), hipMemcpyHostToDevice, s));
    launch_bisect(dT2, n, nkeep, hi0, pivmin, dSig, s);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(hSig, dSig, sizeof(double) * nkeep, hipMemcpyDeviceToHost, s));
    return AVS_OK;
}

// The same, waited for: sigma gets the nkeep largest, descending.  LAPACK
// dbdsdc on the host instead for n < 128 (the device round trip costs more)
// or with AVSVD_BDSDC=host.
avs_status bidiag_sigma(int64_t n, int64_t nkeep, const double *d, const double *e, double *sigma, double *dT2,
                        double *dSig, double *hT2, double *hSig, void *sv)
{
    if (n <= 0 || nkeep <= 0 || nkeep > n) return fail(-1, "bad sizes n=%lld nkeep=%lld", (long long)n,
                                                       (long long)nkeep);
    if (!bisect_on_device(n)) {
        std::vector<double> sg(n);
        avs_status st = avs_bdsdc_values(n, d, e, sg.data());
        if (st) return st;
        memcpy(sigma, sg.data(), sizeof(double) * nkeep);
        return AVS_OK;
    }
    avs_status st = bidiag_sigma_enqueue(n, nkeep, d, e, dT2, dSig, hT2, hSig, sv);
    if (st) return st;
    HIPCHK(hipStreamSynchronize((hipStream_t)sv));
    memcpy(sigma, hSig, sizeof(double) * nkeep);
    return AVS_OK;
}

bool bisect_on_device(int64_t n)
{
    const char *pol = getenv("AVSVD_BDSDC");
    return n >= 128 && !(pol && strcmp(pol, "host") == 0);
}

// Band -> bidiagonal of one matrix on the device, from a band store (ld 3b+1,
// rows c-2b .. c+b of column c, the band in place and zeros elsewhere, n +
// CHASE_SPARE_COLS columns allocated, err cleared) on stream s: one chase_step launch per time of the pipelined
// schedule (step s of sweep i at time 3i + s; the steps of one time touch
// disjoint parts of the store), then d and e to the host.
avs_status chase_store_on_device(double *dStore, int64_t n, int b, int *dErr, double *dDE, double *d, double *e,
                                 hipStream_t s)
{
    if (b > 128) return fail(AVS_ERR_UNSUPPORTED, "device chase: bandwidth %d > 128", b);
    const long last = n - 1, nsweeps = n - 1;
    if (b >= 2 && nsweeps >= 1) {
        // steps of sweep i: T1 + T2, then one T3 + T4 per r = i+1, i+1+b, ...
        // with r + b < last (Chase::steps of the host routine in closed form)
        std::vector<long> steps(nsweeps);
        for (long i = 0; i < nsweeps; ++i) {
            const long x = last - i - 1 - b;
            steps[i] = 1 + (x > 0 ? (x + b - 1) / b : 0);
        }
        // sweeps ilo .. ihi are active at time t: started (3i <= t) and not
        // finished (t - 3i < steps[i]); finished sweeps form a prefix
        long ilo = 0;
        for (long t = 0; ilo < nsweeps; ++t) {
            const long ihi = std::min(t / 3, nsweeps - 1);
            while (ilo <= ihi && t - 3 * ilo >= steps[ilo]) ++ilo;
            if (ilo <= ihi) launch_chase_step(dStore, n, b, t, ilo, ihi, dErr, s);
        }
    }
    launch_chase_diag(dStore, n, b, dDE, s);
    HIPCHK(hipGetLastError());
    int herr = 0;
    HIPCHK(hipMemcpyAsync(d, dDE, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, s));
    if (n > 1) HIPCHK(hipMemcpyAsync(e, dDE + n, sizeof(double) * (size_t)(n - 1), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&herr, dErr, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (herr) return fail(AVS_ERR_STATE, "device chase: a band store access outside the band");
    return AVS_OK;
}

// Band of the handle's reduced matrix, FP64 (b+1) x n LAPACK upper-band
// layout, into the handle's pinned host buffer (*host_band), enqueued on the
// handle's stream: the buffer is valid once that stream has reached this
// point.  Allocates the band buffers on first use only.
avs_status band_enqueue(avs_handle *h, const double **host_band)
{
    if (h->state == avs_handle::ST_EMPTY) return fail(AVS_ERR_STATE, "no matrix set on the handle");
    const size_t ld = (size_t)h->b + 1;
    if (!h->dBand) HIPCHK(hipMalloc(&h->dBand, sizeof(double) * ld * h->n));
    if (!h->hBand) HIPCHK(hipHostMalloc((void **)&h->hBand, sizeof(double) * ld * h->n, hipHostMallocDefault));
    if (h->dtype == AVS_F64) launch_band<double>(h->dTiles, h->dBand, (long)ld, h->nt, h->b, h->stream);
    else launch_band<float>(h->dTiles, h->dBand, (long)ld, h->nt, h->b, h->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(h->hBand, h->dBand, sizeof(double) * ld * h->n, hipMemcpyDeviceToHost, h->stream));
    *host_band = h->hBand;
    return AVS_OK;
}

void *handle_stream(avs_handle *h) { return h->stream; }
void handle_use_own_panel_stream(avs_handle *h) { h->panel_pool = false; }

} // namespace avs

namespace {

// the band of the handle's reduced matrix into its band store, then the chase
avs_status chase_on_device(avs_handle *h, double *d, double *e)
{
    const int64_t n = h->n;
    const int b = h->b;
    const size_t ld = 3 * (size_t)b + 1, cols = (size_t)n + CHASE_SPARE_COLS;
    if (!h->dStore) HIPCHK(hipMalloc(&h->dStore, sizeof(double) * ld * cols));
    if (!h->dDE) HIPCHK(hipMalloc(&h->dDE, sizeof(double) * 2 * (size_t)n));
    if (!h->dChaseErr) HIPCHK(hipMalloc(&h->dChaseErr, sizeof(int)));
    HIPCHK(hipMemsetAsync(h->dStore, 0, sizeof(double) * ld * cols, h->stream));
    HIPCHK(hipMemsetAsync(h->dChaseErr, 0, sizeof(int), h->stream));
    // band[(b + i - j) + j*ldband] = A(i, j) lands on store element (i, j) at
    // j*ld + 2b + i - j with band = store + b and ldband = ld
    if (h->dtype == AVS_F64) launch_band<double>(h->dTiles, h->dStore + b, (long)ld, h->nt, b, h->stream);
    else launch_band<float>(h->dTiles, h->dStore + b, (long)ld, h->nt, b, h->stream);
    HIPCHK(hipGetLastError());
    return chase_store_on_device(h->dStore, n, b, h->dChaseErr, h->dDE, d, e, h->stream);
}

// Where avs_svdvals runs the band -> bidiagonal stage: AVSVD_STAGE2=device on
// the device; any other value (host, halving, dgbbrd) on the host; unset, on
// the device from n = AVSVD_DEVICE_CHASE_MIN_N up (default 8192 for b > 64,
// 16384 otherwise).  Each time of the device schedule is one launch, about 3n
// of them, of 15 us (b = 32) to 36 us (b = 128) on MI355X, and the host
// chase runs on 16 threads.  After the reduction, device vs host seconds
// (device bisection in both): b = 128: 0.65 vs 0.68 at n = 6144, 0.88 vs 0.97
// at 8192, 1.36 vs 1.80 at 12288, 1.86 vs 2.9 at 16384; b = 64: 0.59 vs 0.43
// at 8192, 0.91 vs 0.90 at 12288, 1.25 vs 1.68 at 16384.
bool svdvals_chase_on_device(int64_t n, int b)
{
    if (b < 2 || b > 128) return false;
    if (const char *p = getenv("AVSVD_STAGE2")) return strcmp(p, "device") == 0;
    int64_t nmin = b > 64 ? 8192 : 16384;
    if (const char *p = getenv("AVSVD_DEVICE_CHASE_MIN_N")) nmin = atoll(p);
    return n >= nmin;
}

template <typename TS, int B> avs_status run_persistent(avs_handle *h)
{
    const Program &P = h->prog->P;
    HIPCHK(hipMemsetAsync(h->dCnt, 0, sizeof(int) * (size_t)h->nt * h->nt, h->stream));
    HIPCHK(hipMemsetAsync(h->dTicket, 0, sizeof(int), h->stream));
    HIPCHK(hipMemsetAsync(h->dStatus, 0, sizeof(int), h->stream));
    // dependency wait bound: opts.wait_ms milliseconds (default 30 s)
    const long long ms = h->opts.wait_ms > 0 ? h->opts.wait_ms : 30000;
    launch_persistent<TS, B>(h->workers, h->dTiles, h->dTau, h->dTjj, h->prog->dTasks, (int)P.order.size(),
                             h->prog->dDepTile, h->prog->dDepVal, h->prog->dSig, h->dCnt, h->dTicket,
                             ms * h->wallclock_khz, h->dStatus, split_of(h), h->stream);
    HIPCHK(hipGetLastError());
    return AVS_OK;
}

avs_status run_alg(avs_handle *h, int alg, int32_t kend)
{
    if (!h) return fail(-1, "null handle");
    if (h->opts.sched == AVS_SCHED_PERSISTENT)
        return fail(AVS_ERR_UNSUPPORTED,
                    "the persistent dependency-counting executor is disabled: it did not complete on the device "
                    "and its cause is not yet identified; use AVS_SCHED_STREAMS or AVS_SCHED_HOST");
    if (h->state == avs_handle::ST_EMPTY) return fail(AVS_ERR_STATE, "no matrix set on the handle");
    if (kend <= 0 || kend > h->nt) kend = h->nt;
    avs_status st = h->opts.sched == AVS_SCHED_STREAMS ? ensure_phases(h, alg, kend) : ensure_program(h, alg, kend);
    if (st) return st;
    // after Blockbidiag! the reference's matrix holds the band with explicit
    // zeros (tiledalgos.jl:46,58); the tiles hold reflectors there, so write
    // the zeros first and run on the same matrix the reference would
    if (h->state == avs_handle::ST_BAND) {
        if (h->dtype == AVS_F64) launch_zero_band<double>(h->dTiles, h->nt, h->b, h->band_k, h->stream);
        else launch_zero_band<float>(h->dTiles, h->nt, h->b, h->band_k, h->stream);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipEventRecord(h->ev0, h->stream));
    st = dispatch_tb(h->dtype, h->b, [&](auto tp, auto bc) -> avs_status {
        typedef std::remove_pointer_t<decltype(tp)> TS;
        constexpr int B = decltype(bc)::value;
        if (h->opts.sched == AVS_SCHED_PERSISTENT) return run_persistent<TS, B>(h);
        if (h->opts.sched == AVS_SCHED_STREAMS) return run_streams_graph<TS, B>(h);
        return run_host_enqueued<TS, B>(h);
    });
    if (st) return st;
    HIPCHK(hipEventRecord(h->ev1, h->stream));
    h->timed = true;
    h->state = alg == ALG_BLOCKBIDIAG ? avs_handle::ST_BAND : avs_handle::ST_QR;
    h->band_k = kend;
    return AVS_OK;
}

avs_status check_status(avs_handle *h)
{
    int s = 0;
    HIPCHK(hipMemcpyAsync(&s, h->dStatus, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (s != 0)
        return fail(AVS_ERR_STATE, "persistent executor: task %d exceeded its dependency wait bound", s - 1);
    return AVS_OK;
}

} // namespace

extern "C" {

const char *avs_last_error(void) { return g_err.c_str(); }

void avs_default_opts(avs_opts *o)
{
    if (!o) return;
    memset(o, 0, sizeof *o);
    o->sched = AVS_SCHED_STREAMS;
    o->tree = AVS_TREE_HIER;
    o->domain = 0;
    o->update_mode = AVS_UPD_NATIVE;
}

avs_status avs_create(avs_handle **hp, int32_t dtype, int64_t n, int32_t b, const avs_opts *opts)
{
    if (!hp) return fail(-1, "null handle pointer");
    *hp = nullptr;
    if (dtype != AVS_F64 && dtype != AVS_F32) return fail(-2, "dtype %d not supported", dtype);
    if (n <= 0 || b <= 0 || n % b != 0) return fail(-3, "Not supported: n=%lld is not a positive multiple of b=%d",
                                                    (long long)n, b);
    if (!supported_b(b)) return fail(-4, "Not supported: tile size %d (built: 32, 64, 128)", b);
    if (opts)
        for (int32_t r : opts->reserved)
            if (r != 0) return fail(-5, "opts.reserved must be zero");
    std::unique_ptr<avs_handle> h(new avs_handle());
    h->dtype = dtype;
    h->n = n;
    h->b = b;
    h->nt = (int)(n / b);
    if (opts) h->opts = *opts;
    else avs_default_opts(&h->opts);
    if (h->opts.stream) h->stream = (hipStream_t)h->opts.stream;
    else {
        HIPCHK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
        h->own_stream = true;
    }
    const size_t nt2 = (size_t)h->nt * h->nt;
    const size_t nslots = 4 * nt2;
    HIPCHK(hipMalloc(&h->dTiles, esize(dtype) * (size_t)n * n));
    HIPCHK(hipMalloc(&h->dTau, sizeof(double) * nslots * b));
    HIPCHK(hipMalloc(&h->dTjj, sizeof(double) * nslots * 16 * b));
    HIPCHK(hipMalloc(&h->dCnt, sizeof(int) * nt2));
    HIPCHK(hipMalloc(&h->dTicket, sizeof(int)));
    HIPCHK(hipMalloc(&h->dStatus, sizeof(int)));
    HIPCHK(hipMemset(h->dStatus, 0, sizeof(int)));
    HIPCHK(hipEventCreate(&h->ev0));
    HIPCHK(hipEventCreate(&h->ev1));
    int dev = 0, ncu = 0;
    HIPCHK(hipGetDevice(&dev));
    h->device = dev;
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    int wclk = 0;
    if (hipDeviceGetAttribute(&wclk, hipDeviceAttributeWallClockRate, dev) == hipSuccess && wclk > 0)
        h->wallclock_khz = wclk;
    int occ = 0;
    avs_status st = dispatch_tb(dtype, b, [&](auto tp, auto bc) -> avs_status {
        typedef std::remove_pointer_t<decltype(tp)> TS;
        constexpr int B = decltype(bc)::value;
        occ = persistent_occupancy<TS, B>();
        return AVS_OK;
    });
    if (st) return st;
    if (occ < 1) return fail(AVS_ERR_HIP, "persistent executor does not fit on a CU");
    const int maxw = ncu * occ;
    h->workers = h->opts.workers > 0 ? std::min(h->opts.workers, maxw) : maxw;
    *hp = h.release();
    return AVS_OK;
}

avs_status avs_destroy(avs_handle *h)
{
    if (!h) return AVS_OK;
    if (h->stream) hipStreamSynchronize(h->stream);
    if (h->gexec) hipGraphExecDestroy(h->gexec);
    h->prog.reset();
    hipFree(h->dTiles);
    hipFree(h->dStage);
    hipFree(h->dTau);
    hipFree(h->dTjj);
    hipFree(h->dBand);
    if (h->hBand) hipHostFree(h->hBand);
    hipFree(h->dStore);
    hipFree(h->dDE);
    hipFree(h->dChaseErr);
    hipFree(h->dCnt);
    hipFree(h->dTicket);
    hipFree(h->dStatus);
    hipFree(h->dT2);
    hipFree(h->dSig);
    if (h->hT2) hipHostFree(h->hT2);
    if (h->hSig) hipHostFree(h->hSig);
    if (h->ev0) hipEventDestroy(h->ev0);
    if (h->ev1) hipEventDestroy(h->ev1);
    for (hipEvent_t e : h->pev) hipEventDestroy(e);
    for (hipEvent_t e : h->done) hipEventDestroy(e);
    // a pooled panel stream goes back to its device's pool (the handle's
    // stream, synchronised above, waited for its last panel)
    if (h->pslot >= 0) release_panel_stream(h);
    else if (h->pstream) hipStreamDestroy(h->pstream);
    if (h->ustream) hipStreamDestroy(h->ustream);
    for (hipEvent_t e : h->ujoin)
        if (e) hipEventDestroy(e);
    if (h->own_stream) hipStreamDestroy(h->stream);
    delete h;
    return AVS_OK;
}

static avs_status ensure_stage(avs_handle *h)
{
    if (!h->dStage) HIPCHK(hipMalloc(&h->dStage, esize(h->dtype) * (size_t)h->n * h->n));
    return AVS_OK;
}

avs_status avs_set_matrix_device(avs_handle *h, const void *dA, int64_t lda)
{
    if (!h) return fail(-1, "null handle");
    if (!dA) return fail(-2, "null matrix");
    if (lda < h->n) return fail(-3, "lda < n");
    if (h->dtype == AVS_F64) launch_pack<double>(dA, lda, h->dTiles, h->nt, h->b, h->stream);
    else launch_pack<float>(dA, lda, h->dTiles, h->nt, h->b, h->stream);
    HIPCHK(hipGetLastError());
    h->state = avs_handle::ST_MATRIX;
    h->band_k = 0;
    return AVS_OK;
}

avs_status avs_set_matrix_host(avs_handle *h, const void *A, int64_t lda)
{
    if (!h) return fail(-1, "null handle");
    if (!A) return fail(-2, "null matrix");
    if (lda < h->n) return fail(-3, "lda < n");
    avs_status st = ensure_stage(h);
    if (st) return st;
    const size_t es = esize(h->dtype);
    HIPCHK(hipMemcpy2DAsync(h->dStage, es * h->n, A, es * lda, es * h->n, h->n, hipMemcpyHostToDevice, h->stream));
    return avs_set_matrix_device(h, h->dStage, h->n);
}

avs_status avs_set_matrix_host_sub(avs_handle *h, const void *A, int64_t m, int64_t lda)
{
    if (!h) return fail(-1, "null handle");
    if (!A) return fail(-2, "null matrix");
    if (m <= 0 || m > h->n) return fail(-3, "m must be in 1..n");
    if (lda < m) return fail(-4, "lda < m");
    avs_status st = ensure_stage(h);
    if (st) return st;
    const size_t es = esize(h->dtype);
    if (m < h->n) HIPCHK(hipMemsetAsync(h->dStage, 0, es * (size_t)h->n * h->n, h->stream));
    HIPCHK(hipMemcpy2DAsync(h->dStage, es * h->n, A, es * lda, es * m, m, hipMemcpyHostToDevice, h->stream));
    return avs_set_matrix_device(h, h->dStage, h->n);
}

avs_status avs_get_matrix_device(avs_handle *h, void *dA, int64_t lda)
{
    if (!h) return fail(-1, "null handle");
    if (!dA) return fail(-2, "null matrix");
    if (lda < h->n) return fail(-3, "lda < n");
    if (h->state == avs_handle::ST_EMPTY) return fail(AVS_ERR_STATE, "no matrix set on the handle");
    const int mode = h->state == avs_handle::ST_BAND ? 1 : 0;
    const int kend = h->state == avs_handle::ST_BAND ? h->band_k : 0;
    if (h->dtype == AVS_F64) launch_unpack<double>(h->dTiles, dA, lda, h->nt, h->b, mode, kend, h->stream);
    else launch_unpack<float>(h->dTiles, dA, lda, h->nt, h->b, mode, kend, h->stream);
    HIPCHK(hipGetLastError());
    return AVS_OK;
}

avs_status avs_get_matrix_host(avs_handle *h, void *A, int64_t lda)
{
    if (!h) return fail(-1, "null handle");
    if (!A) return fail(-2, "null matrix");
    if (lda < h->n) return fail(-3, "lda < n");
    avs_status st = ensure_stage(h);
    if (st) return st;
    st = avs_get_matrix_device(h, h->dStage, h->n);
    if (st) return st;
    const size_t es = esize(h->dtype);
    HIPCHK(hipMemcpy2DAsync(A, es * lda, h->dStage, es * h->n, es * h->n, h->n, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return check_status(h);
}

avs_status avs_blockbidiag(avs_handle *h, int32_t kend) { return run_alg(h, ALG_BLOCKBIDIAG, kend); }
avs_status avs_tiled_qr(avs_handle *h, int32_t kend) { return run_alg(h, ALG_TILED_QR, kend); }

// myormqr! (CAQR/src/qr_kernels.jl:256-265): the reflectors of every panel
// operation of the handle's last tiled_QR!, in program order (Q^T) or reverse
// (Q), applied to the tile rows of C they act on, over every column tile of C.
// The phase program is rebuilt with the run's tree, domain and kend: its
// factor slots are the ones the run filled.
avs_status avs_apply_q(avs_handle *h, void *dC, int64_t ldc, int64_t ncols, int32_t applyt)
{
    if (!h) return fail(-1, "null handle");
    if (!dC) return fail(-2, "null C");
    if (ldc < h->n) return fail(-3, "ldc < n");
    if (ncols < 0 || ncols % h->b != 0) return fail(-4, "ncols must be a non-negative multiple of b");
    if (h->state != avs_handle::ST_QR)
        return fail(AVS_ERR_STATE, "no tiled_QR! factors on the handle (run avs_tiled_qr first)");
    if (ncols == 0) return AVS_OK;
    PhaseProgram P;
    build_phase_program(P, h->nt, h->b, ALG_TILED_QR, h->band_k, resolve_tree(h), resolve_domain(h));
    std::vector<ChainOp> pan;
    for (const Phase &ph : P.phases) {
        if (ph.kind != TK_PANEL) continue;
        for (int c = ph.chain0; c < ph.chain0 + ph.nchains; ++c)
            for (int o = P.chain_start[c]; o < P.chain_start[c + 1]; ++o) pan.push_back(P.ops[o]);
    }
    if (!applyt) std::reverse(pan.begin(), pan.end());
    const int ntc = (int)(ncols / h->b);
    return dispatch_tb(h->dtype, h->b, [&](auto tp, auto bc) -> avs_status {
        typedef std::remove_pointer_t<decltype(tp)> TS;
        constexpr int B = decltype(bc)::value;
        const long tsz = (long)B * B, cstep = (long)B * ldc;
        TS *tiles = (TS *)h->dTiles, *C = (TS *)dC;
        for (const ChainOp &op : pan) {
            // GEQRT reflectors sit in the tile they factor; TSQRT and merge
            // reflectors in the bottom tile.  Tile id = column * nt + row.
            const TS *V = tiles + (long)(op.mode == MODE_GE ? op.vtile : op.btile) * tsz;
            TS *C1 = C + (long)(op.vtile % h->nt) * B;
            TS *C2 = C + (long)(op.btile % h->nt) * B;
            launch_update<TS, B>(op.mode, applyt != 0, V, B, 0, h->dTjj + (long)op.fac * 16 * B, C1, ldc, cstep, C2,
                                 ldc, cstep, 0, ntc, upd_tiles_for(B), split_of(h), h->stream);
        }
        HIPCHK(hipGetLastError());
        return AVS_OK;
    });
}

avs_status avs_synchronize(avs_handle *h)
{
    if (!h) return fail(-1, "null handle");
    HIPCHK(hipStreamSynchronize(h->stream));
    return check_status(h);
}

avs_status avs_device_arch(int32_t device, char *name, int64_t len)
{
    if (!name) return fail(-2, "null name");
    if (len < 2) return fail(-3, "len < 2");
    hipDeviceProp_t p;
    HIPCHK(hipGetDeviceProperties(&p, device));
    int64_t k = 0;
    while (k + 1 < len && p.gcnArchName[k] && p.gcnArchName[k] != ':') {
        name[k] = p.gcnArchName[k];
        ++k;
    }
    name[k] = 0;
    return AVS_OK;
}

avs_status avs_last_time_ms(avs_handle *h, double *ms)
{
    if (!h) return fail(-1, "null handle");
    if (!ms) return fail(-2, "null output");
    if (!h->timed) return fail(AVS_ERR_STATE, "nothing timed yet");
    HIPCHK(hipEventSynchronize(h->ev1));
    float f = 0.f;
    HIPCHK(hipEventElapsedTime(&f, h->ev0, h->ev1));
    *ms = f;
    return check_status(h);
}

avs_status avs_phase_timing(avs_handle *h, double *panel_ms, int64_t *panel_launches, double *update_ms,
                            int64_t *update_launches, double *update_flops)
{
    if (!h) return fail(-1, "null handle");
    if (!h->timed || h->opts.sched != AVS_SCHED_STREAMS || !h->ptimed || h->plaunches == 0)
        return fail(AVS_ERR_STATE, "the last run recorded no per-launch events (opts.phase_timing = 1 or avs_set_phase_timing before it)");
    HIPCHK(hipStreamSynchronize(h->stream));
    double pm = 0, um = 0, uf = 0;
    int64_t pn = 0, un = 0;
    for (int i = 0; i < h->plaunches; ++i) {
        float f = 0.f;
        HIPCHK(hipEventElapsedTime(&f, h->pev[2 * i], h->pev[2 * i + 1]));
        if (h->pkind[i] == TK_PANEL) { pm += f; ++pn; }
        else { um += f; ++un; uf += h->pflops[i]; }
    }
    if (panel_ms) *panel_ms = pm;
    if (panel_launches) *panel_launches = pn;
    if (update_ms) *update_ms = um;
    if (update_launches) *update_launches = un;
    if (update_flops) *update_flops = uf;
    return check_status(h);
}

avs_status avs_set_phase_timing(avs_handle *h, int32_t on)
{
    if (!h) return fail(-1, "null handle");
    h->opts.phase_timing = on != 0;
    return AVS_OK;
}

avs_status avs_phase_trace(avs_handle *h, double *start_ms, double *end_ms, int32_t *kind_mode, int64_t cap,
                           int64_t *count)
{
    if (!h) return fail(-1, "null handle");
    if (!h->timed || h->opts.sched != AVS_SCHED_STREAMS || !h->ptimed || h->plaunches == 0)
        return fail(AVS_ERR_STATE, "the last run recorded no per-launch events (opts.phase_timing = 1 or avs_set_phase_timing before it)");
    if (count) *count = h->plaunches;
    if (!start_ms || !end_ms || !kind_mode) return AVS_OK;
    if (cap < h->plaunches) return fail(-5, "cap < number of launches");
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int i = 0; i < h->plaunches; ++i) {
        float a = 0.f, e = 0.f;
        HIPCHK(hipEventElapsedTime(&a, h->pev[0], h->pev[2 * i]));
        HIPCHK(hipEventElapsedTime(&e, h->pev[0], h->pev[2 * i + 1]));
        start_ms[i] = a;
        end_ms[i] = e;
        kind_mode[i] = h->pkind[i] * 4 + h->pmode[i];
    }
    return check_status(h);
}

avs_status avs_get_band(avs_handle *h, double *band, int64_t ldband)
{
    if (!h) return fail(-1, "null handle");
    if (!band) return fail(-2, "null band");
    if (ldband < h->b + 1) return fail(-3, "ldband < b+1");
    const double *hb = nullptr;
    avs_status st = avs::band_enqueue(h, &hb);
    if (st) return st;
    HIPCHK(hipStreamSynchronize(h->stream));
    const size_t ld = (size_t)h->b + 1;
    for (int64_t j = 0; j < h->n; ++j) memcpy(band + j * ldband, hb + j * ld, sizeof(double) * ld);
    return check_status(h);
}

avs_status avs_bidiagonalize_band_device(int64_t n, int32_t bw, const double *band, int64_t ldband, double *d,
                                         double *e)
{
    if (n <= 0) return fail(-1, "n must be positive");
    if (bw < 1 || bw > 128) return fail(-2, "bandwidth must be in 1 .. 128");
    if (!band) return fail(-3, "null band");
    if (ldband < bw + 1) return fail(-4, "ldband < bw+1");
    if (!d) return fail(-5, "null d");
    if (n > 1 && !e) return fail(-6, "null e");
    const int64_t b = bw, ld = 3 * b + 1;
    // the band store on the host: A(r, c) at c*ld + 2b + r - c for c-b <= r <= c
    std::vector<double> S((size_t)(ld * (n + CHASE_SPARE_COLS)), 0.0);
    for (int64_t c = 0; c < n; ++c)
        for (int64_t r = std::max<int64_t>(0, c - b); r <= c; ++r) S[c * ld + 2 * b + r - c] = band[(b + r - c) + c * ldband];
    hipStream_t s = nullptr;
    double *dS = nullptr, *dDE = nullptr;
    int *dErr = nullptr;
    avs_status st = AVS_OK;
    auto run = [&]() -> avs_status {
        HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        HIPCHK(hipMalloc(&dS, sizeof(double) * S.size()));
        HIPCHK(hipMalloc(&dDE, sizeof(double) * 2 * (size_t)n));
        HIPCHK(hipMalloc(&dErr, sizeof(int)));
        HIPCHK(hipMemcpyAsync(dS, S.data(), sizeof(double) * S.size(), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemsetAsync(dErr, 0, sizeof(int), s));
        return chase_store_on_device(dS, n, bw, dErr, dDE, d, e, s);
    };
    st = run();
    if (s) hipStreamSynchronize(s);
    hipFree(dS);
    hipFree(dDE);
    hipFree(dErr);
    if (s) hipStreamDestroy(s);
    return st;
}

avs_status avs_svdvals(avs_handle *h, double *sigma)
{
    if (!h) return fail(-1, "null handle");
    if (!sigma) return fail(-2, "null sigma");
    if (h->state == avs_handle::ST_EMPTY) return fail(AVS_ERR_STATE, "no matrix set on the handle");
    // tiled_QR! has replaced the matrix by R and its reflectors, whose
    // singular values are not the matrix's: the caller must upload it again
    if (h->state == avs_handle::ST_QR)
        return fail(AVS_ERR_STATE, "the tiles hold tiled_QR!'s R and reflectors, not the matrix: set the matrix again");
    avs_status st;
    // the uploaded matrix, or a partial Blockbidiag! (its band and trailing
    // block are an orthogonal transform of the matrix: same singular values)
    if (h->state != avs_handle::ST_BAND || h->band_k != h->nt) {
        st = avs_blockbidiag(h, 0);
        if (st) return st;
    }
    const int64_t n = h->n;
    const int b = h->b;
    std::vector<double> d(n), e(n > 1 ? n - 1 : 1);
    if (svdvals_chase_on_device(n, b)) {
        st = chase_on_device(h, d.data(), e.data());
    } else {
        std::vector<double> band((size_t)(b + 1) * n);
        st = avs_get_band(h, band.data(), b + 1);
        if (st) return st;
        st = bidiagonalize_band_auto(n, b, band.data(), b + 1, d.data(), e.data());
    }
    if (st) return st;
    if (!h->dT2) {
        HIPCHK(hipMalloc(&h->dT2, sizeof(double) * 2 * (size_t)n));
        HIPCHK(hipMalloc(&h->dSig, sizeof(double) * (size_t)n));
        HIPCHK(hipHostMalloc((void **)&h->hT2, sizeof(double) * 2 * (size_t)n, hipHostMallocDefault));
        HIPCHK(hipHostMalloc((void **)&h->hSig, sizeof(double) * (size_t)n, hipHostMallocDefault));
    }
    return bidiag_sigma(n, n, d.data(), e.data(), sigma, h->dT2, h->dSig, h->hT2, h->hSig, h->stream);
}

// ------------------------------------------------------------ tile operators
avs_status avs_geqrt(int32_t dtype, int32_t b, void *dA, int64_t lda, double *dtau, int32_t trans, void *stream)
{
    if (!dA) return fail(-3, "null tile");
    if (lda < b) return fail(-4, "lda < b");
    if (!dtau) return fail(-5, "null tau");
    return dispatch_tb(dtype, b, [&](auto tp, auto bc) -> avs_status {
        typedef std::remove_pointer_t<decltype(tp)> TS;
        constexpr int B = decltype(bc)::value;
        launch_panel<TS, B>(MODE_GE, dA, lda, dA, lda, trans, dtau, nullptr, (hipStream_t)stream);
        HIPCHK(hipGetLastError());
        return AVS_OK;
    });
}

avs_status avs_tsqrt(int32_t dtype, int32_t b, void *dA1, int64_t lda1, void *dA2, int64_t lda2, double *dtau,
                     int32_t trans, void *stream)
{
    if (!dA1) return fail(-3, "null tile");
    if (lda1 < b) return fail(-4, "lda1 < b");
    if (!dA2) return fail(-5, "null tile");
    if (lda2 < b) return fail(-6, "lda2 < b");
    if (!dtau) return fail(-7, "null tau");
    return dispatch_tb(dtype, b, [&](auto tp, auto bc) -> avs_status {
        typedef std::remove_pointer_t<decltype(tp)> TS;
        constexpr int B = decltype(bc)::value;
        launch_panel<TS, B>(MODE_TS, dA1, lda1, dA2, lda2, trans, dtau, nullptr, (hipStream_t)stream);
        HIPCHK(hipGetLastError());
        return AVS_OK;
    });
}

// T_JJ of the reflectors in dV into dT (16 * b doubles), then the update of
// ncols / b column tiles; both launches on stream s
static avs_status apply_op_ws(int mode, int32_t dtype, int32_t b, const void *dV, int64_t ldv, const double *dtau,
                              void *dC1, int64_t ldc1, void *dC2, int64_t ldc2, int64_t ncols, int32_t applyt,
                              int32_t vtrans, int32_t ctrans, double *dT, hipStream_t s)
{
    const int ntl = (int)(ncols / b);
    return dispatch_tb(dtype, b, [&](auto tp, auto bc) -> avs_status {
        typedef std::remove_pointer_t<decltype(tp)> TS;
        constexpr int B = decltype(bc)::value;
        launch_tjj<TS, B>(mode, dV, ldv, vtrans, dtau, dT, s);
        const long s1 = ctrans ? (long)B : (long)B * ldc1;
        const long s2 = ctrans ? (long)B : (long)B * ldc2;
        launch_update<TS, B>(mode, applyt != 0, dV, ldv, vtrans, dT, dC1, ldc1, s1, dC2, ldc2, s2, ctrans, ntl,
                             upd_tiles_for(B), 0, s);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return fail(AVS_ERR_HIP, "update launch failed: %s", hipGetErrorString(e));
        return AVS_OK;
    });
}

static avs_status apply_op(int mode, int32_t dtype, int32_t b, const void *dV, int64_t ldv, const double *dtau,
                           void *dC1, int64_t ldc1, void *dC2, int64_t ldc2, int64_t ncols, int32_t applyt,
                           int32_t vtrans, int32_t ctrans, void *stream)
{
    const avs_status built = dispatch_tb(dtype, b, [](auto, auto) { return AVS_OK; });
    if (built) return built;
    hipStream_t s = (hipStream_t)stream;
    double *dT = nullptr;
    HIPCHK(hipMalloc(&dT, sizeof(double) * 16 * (size_t)b));
    const avs_status st =
        apply_op_ws(mode, dtype, b, dV, ldv, dtau, dC1, ldc1, dC2, ldc2, ncols, applyt, vtrans, ctrans, dT, s);
    hipStreamSynchronize(s);
    hipFree(dT);
    return st;
}

avs_status avs_unmqr(int32_t dtype, int32_t b, const void *dV, int64_t ldv, const double *dtau, void *dC,
                     int64_t ldc, int64_t ncols, int32_t applyt, int32_t vtrans, int32_t ctrans, void *stream)
{
    if (!dV) return fail(-3, "null V");
    if (ldv < b) return fail(-4, "ldv < b");
    if (!dtau) return fail(-5, "null tau");
    if (!dC) return fail(-6, "null C");
    if (!ctrans && ldc < b) return fail(-7, "ldc < b");
    if (ncols < 0 || ncols % b != 0) return fail(-8, "ncols must be a multiple of b");
    return apply_op(MODE_GE, dtype, b, dV, ldv, dtau, dC, ldc, dC, ldc, ncols, applyt, vtrans, ctrans, stream);
}

avs_status avs_tsmqr(int32_t dtype, int32_t b, void *dC1, int64_t ldc1, void *dC2, int64_t ldc2, const void *dV,
                     int64_t ldv, const double *dtau, int64_t ncols, int32_t applyt, int32_t vtrans, int32_t ctrans,
                     void *stream)
{
    if (!dC1) return fail(-3, "null C1");
    if (!dC2) return fail(-5, "null C2");
    if (!dV) return fail(-7, "null V");
    if (ldv < b) return fail(-8, "ldv < b");
    if (!dtau) return fail(-9, "null tau");
    if (ncols < 0 || ncols % b != 0) return fail(-10, "ncols must be a multiple of b");
    return apply_op(MODE_TS, dtype, b, dV, ldv, dtau, dC1, ldc1, dC2, ldc2, ncols, applyt, vtrans, ctrans, stream);
}

avs_status avs_tsmqr_noalloc(int32_t dtype, int32_t b, void *dC1, int64_t ldc1, void *dC2, int64_t ldc2,
                             const void *dV, int64_t ldv, const double *dtau, int64_t ncols, int32_t applyt,
                             int32_t vtrans, int32_t ctrans, void *stream)
{
    if (b < 1 || 16L * b > NOALLOC_SCRATCH_DOUBLES)
        return fail(-2, "b = %d: the fixed T_JJ workspace holds %d doubles (16 * b for b <= 128)", b,
                    NOALLOC_SCRATCH_DOUBLES);
    if (!dC1) return fail(-3, "null C1");
    if (!dC2) return fail(-5, "null C2");
    if (!dV) return fail(-7, "null V");
    if (ldv < b) return fail(-8, "ldv < b");
    if (!dtau) return fail(-9, "null tau");
    if (ncols < 0 || ncols % b != 0) return fail(-10, "ncols must be a multiple of b");
    hipStream_t s = (hipStream_t)stream;
    int device = -1;
    if (s) HIPCHK(hipStreamGetDevice(s, &device));
    else HIPCHK(hipGetDevice(&device));
    double *ws = nullptr;
    const hipError_t e = noalloc_scratch(device, &ws);
    if (e != hipSuccess)
        return fail(AVS_ERR_HIP, "device %d: address of the T_JJ workspace: %s", device, hipGetErrorString(e));
    return apply_op_ws(MODE_TS, dtype, b, dV, ldv, dtau, dC1, ldc1, dC2, ldc2, ncols, applyt, vtrans, ctrans, ws, s);
}

avs_status avs_debug_mfma_layout(int32_t variant, double *out64, float *out32)
{
    if (variant < 0 || variant > 3) return fail(-1, "variant must be 0, 1, 2 or 3");
    if (!out64 || !out32) return fail(-2, "null output");
    double ha[64], hb[64];
    float fa[64], fb[64];
    for (int i = 0; i < 64; ++i) {
        // lane i supplies ha[i] as its A operand and hb[i] as its B operand
        if (variant == 0) { ha[i] = (double)(i + 1); hb[i] = (double)((i * 7) % 13 - 6); }
        else if (variant == 1) { ha[i] = 1.0; hb[i] = (double)i; }     // sums of B lanes per output
        else { ha[i] = (double)i; hb[i] = 1.0; }                      // sums of A lanes per output
        // variant 3: lane exchanges instead of MFMA (out64[4l+k] = lane l ^ (4<<k) of a,
        // out32[4l+k] = sum of a over lanes congruent to l modulo 4<<k)
        fa[i] = (float)ha[i];
        fb[i] = (float)hb[i];
    }
    double *d = nullptr;
    float *f = nullptr;
    HIPCHK(hipMalloc(&d, sizeof(double) * (128 + 256)));
    HIPCHK(hipMalloc(&f, sizeof(float) * (128 + 256)));
    HIPCHK(hipMemcpy(d, ha, sizeof ha, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d + 64, hb, sizeof hb, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(f, fa, sizeof fa, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(f + 64, fb, sizeof fb, hipMemcpyHostToDevice));
    launch_mfma_probe(d, d + 64, d + 128, f, f + 64, f + 128, variant == 3 ? 1 : 0, 0);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy(out64, d + 128, sizeof(double) * 256, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(out32, f + 128, sizeof(float) * 256, hipMemcpyDeviceToHost));
    hipFree(d);
    hipFree(f);
    return AVS_OK;
}

avs_status avs_schedule_stats(int32_t nt, int32_t b, int32_t alg, int32_t kend, int32_t tree, int32_t domain,
                              int32_t workers, int64_t *ntasks, double *est_us)
{
    if (nt <= 0) return fail(-1, "nt must be positive");
    if (b <= 0) return fail(-2, "b must be positive");
    Program P;
    build_program(P, nt, b, alg, kend <= 0 ? nt : kend, tree, domain, upd_tiles_for(b), workers);
    if (ntasks) *ntasks = (int64_t)P.prog.size();
    if (est_us) *est_us = P.est_makespan_us;
    return AVS_OK;
}

avs_status avs_phase_dump(int32_t nt, int32_t b, int32_t alg, int32_t kend, int32_t tree, int32_t domain,
                          int32_t *out, int64_t cap, int64_t *needed)
{
    if (nt <= 0) return fail(-1, "nt must be positive");
    if (b <= 0) return fail(-2, "b must be positive");
    if (!needed) return fail(-9, "null needed");
    PhaseProgram P;
    build_phase_program(P, nt, b, alg, kend <= 0 ? nt : kend, tree, domain);
    std::vector<int32_t> v;
    v.push_back((int32_t)P.phases.size());
    for (const Phase &ph : P.phases) {
        v.push_back(ph.kind);
        v.push_back(ph.ncols);
        v.push_back(ph.nchains);
        for (int c = ph.chain0; c < ph.chain0 + ph.nchains; ++c) {
            v.push_back(P.chain_start[c + 1] - P.chain_start[c]);
            for (int o = P.chain_start[c]; o < P.chain_start[c + 1]; ++o) {
                const ChainOp &t = P.ops[o];
                const int32_t f[] = {t.kind, t.mode, t.trans, t.vtile, t.btile, t.c1tile, t.cstep, t.fac, t.ncols};
                v.insert(v.end(), f, f + 9);
            }
        }
    }
    *needed = (int64_t)v.size();
    if (out && cap >= (int64_t)v.size()) memcpy(out, v.data(), sizeof(int32_t) * v.size());
    return AVS_OK;
}

avs_status avs_schedule_dump(int32_t nt, int32_t b, int32_t alg, int32_t kend, int32_t tree, int32_t domain,
                             int32_t workers, int32_t *out, int64_t cap, int64_t *needed)
{
    if (nt <= 0) return fail(-1, "nt must be positive");
    if (b <= 0) return fail(-2, "b must be positive");
    if (!needed) return fail(-10, "null needed");
    Program P;
    build_program(P, nt, b, alg, kend <= 0 ? nt : kend, tree, domain, upd_tiles_for(b), workers);
    std::vector<int32_t> v;
    for (int idx : P.order) {
        const Task &t = P.prog[idx];
        const int32_t head[] = {t.kind, t.mode, t.trans, t.ntiles, t.vtile, t.btile, t.c1tile, t.cstep, t.fac, t.ndep};
        v.insert(v.end(), head, head + 10);
        for (int d = 0; d < t.ndep; ++d) {
            v.push_back(P.dep_tile[t.dep0 + d]);
            v.push_back(P.dep_val[t.dep0 + d]);
        }
        v.push_back(t.nsig);
        for (int s2 = 0; s2 < t.nsig; ++s2) v.push_back(P.sig_tile[t.sig0 + s2]);
    }
    *needed = (int64_t)v.size();
    if (out && cap >= (int64_t)v.size()) memcpy(out, v.data(), sizeof(int32_t) * v.size());
    return AVS_OK;
}

} // extern "C"

// ------------------------------------------------------------- batched form
struct avs_batch {
    int dtype = AVS_F64, b = 0, count = 0;
    std::vector<int64_t> n;                                   // true sizes
    std::vector<int> nt;                                      // tiles per side (padded)
    std::vector<long> off, soff, tile0, col0, boff;           // element / tile / column offsets
    long total_tiles = 0, total_cols = 0, arena_elems = 0, band_elems = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    void *dArena = nullptr, *dStage = nullptr;
    double *dBands = nullptr, *dScratch = nullptr;
    long *dOff = nullptr, *dSoff = nullptr, *dTile0 = nullptr, *dCol0 = nullptr, *dBoff = nullptr;
    int *dNts = nullptr, *dOrder = nullptr, *dNext = nullptr;
    int workers = 0, split = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool reduced = false, timed = false;
    double *hBands = nullptr; // pinned host copy of every band (band_elems doubles)
    bool bands_valid = false;
    // device stage 2: band stores ((3b+1) x N per matrix), bidiagonals (2N per
    // matrix), out-of-store flags, pinned host copy of the bidiagonals
    double *dChase = nullptr, *dDE = nullptr, *hDE = nullptr;
    int *dChaseErr = nullptr;
    // device singular values: workgroup table (matrix, first value), true
    // sizes and output offsets per matrix, the values
    int *dWmat = nullptr, *dWfirst = nullptr, nbis = 0;
    long *dNreal = nullptr, *dSigOff = nullptr;
    double *dSigma = nullptr;
};

extern "C" {

avs_status avs_batch_create(avs_batch **bp, int32_t dtype, int32_t count, const int64_t *ns, int32_t b,
                            const avs_opts *opts)
{
    if (!bp) return fail(-1, "null batch pointer");
    *bp = nullptr;
    if (dtype != AVS_F64 && dtype != AVS_F32) return fail(-2, "dtype %d not supported", dtype);
    if (count <= 0) return fail(-3, "count must be positive");
    if (!ns) return fail(-4, "null sizes");
    if (!supported_b(b)) return fail(-5, "Not supported: tile size %d (built: 32, 64, 128)", b);
    std::unique_ptr<avs_batch> B(new avs_batch());
    B->dtype = dtype;
    B->b = b;
    B->count = count;
    B->n.assign(ns, ns + count);
    long toff = 0, soffv = 0, t0 = 0, c0 = 0, bo = 0;
    for (int i = 0; i < count; ++i) {
        if (ns[i] <= 0) return fail(-4, "matrix %d has size %lld", i, (long long)ns[i]);
        const int ntm = (int)((ns[i] + b - 1) / b);
        const long N = (long)ntm * b;
        B->nt.push_back(ntm);
        B->off.push_back(toff);
        B->soff.push_back(soffv);
        B->tile0.push_back(t0);
        B->col0.push_back(c0);
        B->boff.push_back(bo);
        toff += N * N;
        soffv += N * N;
        t0 += (long)ntm * ntm;
        c0 += N;
        bo += (long)(b + 1) * N;
    }
    B->arena_elems = toff;
    B->total_tiles = t0;
    B->total_cols = c0;
    B->band_elems = bo;
    if (opts && opts->stream) B->stream = (hipStream_t)opts->stream;
    else {
        HIPCHK(hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking));
        B->own_stream = true;
    }
    const size_t es = esize(dtype);
    HIPCHK(hipMalloc(&B->dArena, es * (size_t)toff));
    HIPCHK(hipMalloc(&B->dStage, es * (size_t)soffv));
    HIPCHK(hipMemsetAsync(B->dStage, 0, es * (size_t)soffv, B->stream));
    HIPCHK(hipMalloc(&B->dBands, sizeof(double) * (size_t)bo));
    std::vector<int> order(count);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return B->nt[x] > B->nt[y]; });
    auto up = [&](void **d, const void *h, size_t bytes) -> avs_status {
        HIPCHK(hipMalloc(d, bytes));
        HIPCHK(hipMemcpy(*d, h, bytes, hipMemcpyHostToDevice));
        return AVS_OK;
    };
    avs_status st;
    if ((st = up((void **)&B->dOff, B->off.data(), sizeof(long) * count))) return st;
    if ((st = up((void **)&B->dSoff, B->soff.data(), sizeof(long) * count))) return st;
    if ((st = up((void **)&B->dTile0, B->tile0.data(), sizeof(long) * count))) return st;
    if ((st = up((void **)&B->dCol0, B->col0.data(), sizeof(long) * count))) return st;
    if ((st = up((void **)&B->dBoff, B->boff.data(), sizeof(long) * count))) return st;
    if ((st = up((void **)&B->dNts, B->nt.data(), sizeof(int) * count))) return st;
    if ((st = up((void **)&B->dOrder, order.data(), sizeof(int) * count))) return st;
    HIPCHK(hipMalloc(&B->dNext, sizeof(int)));
    int dev = 0, ncu = 0, occ = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    st = dispatch_tb(dtype, b, [&](auto tp, auto bc) -> avs_status {
        typedef std::remove_pointer_t<decltype(tp)> TS;
        constexpr int BB = decltype(bc)::value;
        occ = batch_occupancy<TS, BB>();
        return AVS_OK;
    });
    if (st) return st;
    if (occ < 1) return fail(AVS_ERR_HIP, "batch kernel does not fit on a CU");
    B->workers = std::min(ncu * occ, count);
    B->split = (dtype == AVS_F32 && opts && opts->update_mode == AVS_UPD_SPLIT_BF16) ? 1 : 0;
    if (opts && opts->workers > 0) B->workers = std::min(B->workers, (int)opts->workers);
    HIPCHK(hipMalloc(&B->dScratch, sizeof(double) * 17 * b * (size_t)B->workers));
    HIPCHK(hipEventCreate(&B->ev0));
    HIPCHK(hipEventCreate(&B->ev1));
    *bp = B.release();
    return AVS_OK;
}

avs_status avs_batch_destroy(avs_batch *B)
{
    if (!B) return AVS_OK;
    if (B->stream) hipStreamSynchronize(B->stream);
    void *ptrs[] = {B->dWmat, B->dWfirst, B->dNreal, B->dSigOff, B->dSigma, B->dChase, B->dDE, B->dChaseErr, B->dArena, B->dStage, B->dBands, B->dScratch, B->dOff, B->dSoff, B->dTile0,
                    B->dCol0, B->dBoff, B->dNts, B->dOrder, B->dNext};
    for (void *p : ptrs) hipFree(p);
    if (B->ev0) hipEventDestroy(B->ev0);
    if (B->ev1) hipEventDestroy(B->ev1);
    if (B->hBands) hipHostFree(B->hBands);
    if (B->hDE) hipHostFree(B->hDE);
    if (B->own_stream) hipStreamDestroy(B->stream);
    delete B;
    return AVS_OK;
}

avs_status avs_batch_set_matrix_host(avs_batch *B, int32_t idx, const void *A, int64_t lda)
{
    if (!B) return fail(-1, "null batch");
    if (idx < 0 || idx >= B->count) return fail(-2, "matrix index %d out of range", idx);
    if (!A) return fail(-3, "null matrix");
    const int64_t n = B->n[idx];
    if (lda < n) return fail(-4, "lda < n");
    const size_t es = esize(B->dtype);
    const long N = (long)B->nt[idx] * B->b;
    char *dst = (char *)B->dStage + es * (size_t)B->soff[idx];
    HIPCHK(hipMemcpy2DAsync(dst, es * N, A, es * lda, es * n, n, hipMemcpyHostToDevice, B->stream));
    B->reduced = false;
    B->bands_valid = false;
    return AVS_OK;
}

avs_status avs_batch_blockbidiag(avs_batch *B)
{
    if (!B) return fail(-1, "null batch");
    avs_status st = dispatch_tb(B->dtype, B->b, [&](auto tp, auto bc) -> avs_status {
        typedef std::remove_pointer_t<decltype(tp)> TS;
        constexpr int BB = decltype(bc)::value;
        launch_batch_pack<TS>(B->dStage, B->dSoff, B->dArena, B->dOff, B->dNts, B->dTile0, B->total_tiles, B->count,
                              BB, B->stream);
        HIPCHK(hipMemsetAsync(B->dNext, 0, sizeof(int), B->stream));
        HIPCHK(hipEventRecord(B->ev0, B->stream));
        launch_batch<TS, BB>(B->workers, B->dArena, B->dOff, B->dNts, B->dOrder, B->count, B->dScratch, B->dNext,
                             B->split, B->stream);
        HIPCHK(hipEventRecord(B->ev1, B->stream));
        HIPCHK(hipGetLastError());
        return AVS_OK;
    });
    if (st) return st;
    B->reduced = true;
    B->bands_valid = false;
    B->timed = true;
    return AVS_OK;
}

avs_status avs_batch_last_time_ms(avs_batch *B, double *ms)
{
    if (!B) return fail(-1, "null batch");
    if (!ms) return fail(-2, "null output");
    if (!B->timed) return fail(AVS_ERR_STATE, "nothing timed yet");
    HIPCHK(hipEventSynchronize(B->ev1));
    float f = 0.f;
    HIPCHK(hipEventElapsedTime(&f, B->ev0, B->ev1));
    *ms = f;
    return AVS_OK;
}

static avs_status batch_fetch_bands(avs_batch *B)
{
    avs_status st = dispatch_tb(B->dtype, B->b, [&](auto tp, auto) -> avs_status {
        typedef std::remove_pointer_t<decltype(tp)> TS;
        launch_batch_band<TS>(B->dArena, B->dOff, B->dNts, B->dCol0, B->total_cols, B->dBands, B->dBoff, B->count,
                              B->b, B->stream);
        HIPCHK(hipGetLastError());
        return AVS_OK;
    });
    if (st) return st;
    if (!B->hBands)
        HIPCHK(hipHostMalloc((void **)&B->hBands, sizeof(double) * (size_t)B->band_elems, hipHostMallocDefault));
    HIPCHK(hipMemcpyAsync(B->hBands, B->dBands, sizeof(double) * B->band_elems, hipMemcpyDeviceToHost, B->stream));
    HIPCHK(hipStreamSynchronize(B->stream));
    B->bands_valid = true;
    return AVS_OK;
}

avs_status avs_batch_get_band(avs_batch *B, int32_t idx, double *band, int64_t ldband)
{
    if (!B) return fail(-1, "null batch");
    if (idx < 0 || idx >= B->count) return fail(-2, "matrix index %d out of range", idx);
    if (!band) return fail(-3, "null band");
    if (ldband < B->b + 1) return fail(-4, "ldband < b+1");
    if (!B->reduced) return fail(AVS_ERR_STATE, "batch not reduced yet");
    if (!B->bands_valid) {
        avs_status st = batch_fetch_bands(B);
        if (st) return st;
    }
    const long N = (long)B->nt[idx] * B->b;
    const double *src = B->hBands + B->boff[idx];
    for (long j = 0; j < N; ++j) memcpy(band + j * ldband, src + j * (B->b + 1), sizeof(double) * (B->b + 1));
    return AVS_OK;
}

// Stage 2 on the device: every matrix's band -> bidiagonal in one launch (one
// wavefront per matrix), then only (d, e) comes back (fetch_de), or the
// bidiagonals stay on the device for batch_device_sigma.
static avs_status batch_device_chase(avs_batch *B, bool fetch_de = true)
{
    const size_t cols = (size_t)B->total_cols;
    if (!B->dChase) {
        // 32 spare columns: the device kernel's unrolled loads run up to 31
        // strides past a matrix's last column (their values are not used)
        HIPCHK(hipMalloc(&B->dChase, sizeof(double) * (cols + 32) * (3 * (size_t)B->b + 1)));
        HIPCHK(hipMalloc(&B->dDE, sizeof(double) * 2 * cols));
        HIPCHK(hipMalloc(&B->dChaseErr, sizeof(int) * (size_t)B->count));
        HIPCHK(hipHostMalloc((void **)&B->hDE, sizeof(double) * 2 * cols, hipHostMallocDefault));
    }
    HIPCHK(hipMemsetAsync(B->dChaseErr, 0, sizeof(int) * (size_t)B->count, B->stream));
    avs_status st = dispatch_tb(B->dtype, B->b, [&](auto tp, auto) -> avs_status {
        typedef std::remove_pointer_t<decltype(tp)> TS;
        launch_batch_chase<TS>(B->dArena, B->dOff, B->dNts, B->dOrder, B->dCol0, B->count, B->b, B->dChase, B->dDE,
                               B->dChaseErr, B->stream);
        HIPCHK(hipGetLastError());
        return AVS_OK;
    });
    if (st) return st;
    std::vector<int> errs(B->count);
    if (fetch_de)
        HIPCHK(hipMemcpyAsync(B->hDE, B->dDE, sizeof(double) * 2 * cols, hipMemcpyDeviceToHost, B->stream));
    HIPCHK(hipMemcpyAsync(errs.data(), B->dChaseErr, sizeof(int) * (size_t)B->count, hipMemcpyDeviceToHost,
                          B->stream));
    HIPCHK(hipStreamSynchronize(B->stream));
    for (int m = 0; m < B->count; ++m)
        if (errs[m]) return fail(AVS_ERR_STATE, "device band chase left its band store (matrix %d)", m);
    return AVS_OK;
}

// stage 2 of the batch on the device unless AVSVD_STAGE2 asks for a host
// routine (host, dgbbrd, halving); the device kernel handles b <= 32
static bool batch_stage2_on_device(const avs_batch *B)
{
    const char *pol = getenv("AVSVD_STAGE2");
    return B->b <= 32 && !(pol && *pol);
}

// the bidiagonals' singular values on the device too (bisection) unless
// AVSVD_BDSDC=host asks for LAPACK dbdsdc on the host; the kernel stages
// 2N - 1 <= 4096 values in LDS
static bool batch_sigma_on_device(const avs_batch *B)
{
    const char *pol = getenv("AVSVD_BDSDC");
    if (pol && strcmp(pol, "host") == 0) return false;
    for (int m = 0; m < B->count; ++m)
        if (2L * B->nt[m] * B->b - 1 > 4096) return false;
    return true;
}

// Stage 2 and the singular values on the device; only sigma comes back.
static avs_status batch_device_sigma(avs_batch *B, const std::vector<long> &soff, double *sigma)
{
    avs_status st = batch_device_chase(B, false);
    if (st) return st;
    long total = 0;
    for (int m = 0; m < B->count; ++m) total += B->n[m];
    if (!B->dWmat) {
        // workgroups of the largest matrices first; none for a range that
        // holds only padding values
        std::vector<int> order(B->count), wm, wf;
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return B->nt[x] > B->nt[y]; });
        for (int m : order) {
            const long N = (long)B->nt[m] * B->b;
            for (long j0 = 0; j0 < N; j0 += BATCH_BISECT_THREADS) {
                if (j0 + BATCH_BISECT_THREADS <= N - B->n[m]) continue;
                wm.push_back(m);
                wf.push_back((int)j0);
            }
        }
        std::vector<long> nreal(B->n.begin(), B->n.end());
        B->nbis = (int)wm.size();
        HIPCHK(hipMalloc(&B->dWmat, sizeof(int) * wm.size()));
        HIPCHK(hipMalloc(&B->dWfirst, sizeof(int) * wf.size()));
        HIPCHK(hipMalloc(&B->dNreal, sizeof(long) * B->count));
        HIPCHK(hipMalloc(&B->dSigOff, sizeof(long) * B->count));
        HIPCHK(hipMalloc(&B->dSigma, sizeof(double) * (size_t)std::max(total, 1L)));
        HIPCHK(hipMemcpy(B->dWmat, wm.data(), sizeof(int) * wm.size(), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(B->dWfirst, wf.data(), sizeof(int) * wf.size(), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(B->dNreal, nreal.data(), sizeof(long) * B->count, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(B->dSigOff, soff.data(), sizeof(long) * B->count, hipMemcpyHostToDevice));
    }
    launch_batch_bisect(B->dDE, B->dCol0, B->dNts, B->b, B->dWmat, B->dWfirst, B->nbis, B->dNreal, B->dSigOff,
                        B->dSigma, B->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(sigma, B->dSigma, sizeof(double) * total, hipMemcpyDeviceToHost, B->stream));
    HIPCHK(hipStreamSynchronize(B->stream));
    return AVS_OK;
}

avs_status avs_batch_svdvals(avs_batch *B, double *sigma)
{
    if (!B) return fail(-1, "null batch");
    if (!sigma) return fail(-2, "null sigma");
    avs_status st;
    if (!B->reduced) {
        if ((st = avs_batch_blockbidiag(B))) return st;
    }
    B->bands_valid = false;
    std::vector<long> soff(B->count, 0);
    for (int i = 1; i < B->count; ++i) soff[i] = soff[i - 1] + B->n[i - 1];
    const bool dev2 = batch_stage2_on_device(B);
    if (dev2 && batch_sigma_on_device(B)) return batch_device_sigma(B, soff, sigma);
    if ((st = dev2 ? batch_device_chase(B) : batch_fetch_bands(B))) return st;
    std::vector<int> order(B->count);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return B->nt[x] > B->nt[y]; });
    std::atomic<int> next{0};
    std::atomic<int> first_err{0};
    std::string err_msg;
    std::mutex mu;
    auto work = [&]() {
        std::vector<double> d, e, sg;
        for (;;) {
            const int q = next.fetch_add(1);
            if (q >= B->count) break;
            const int m = order[q];
            const long N = (long)B->nt[m] * B->b;
            d.resize(N);
            e.resize(std::max<long>(N - 1, 1));
            sg.resize(N);
            avs_status s2 = AVS_OK;
            if (dev2) {
                const double *de = B->hDE + 2 * B->col0[m];
                memcpy(d.data(), de, sizeof(double) * N);
                if (N > 1) memcpy(e.data(), de + N, sizeof(double) * (N - 1));
            } else {
                s2 = bidiagonalize_band_one(N, B->b, B->hBands + B->boff[m], B->b + 1, d.data(), e.data(), 1);
            }
            if (!s2) s2 = avs_bdsdc_values(N, d.data(), e.data(), sg.data());
            if (s2) {
                std::lock_guard<std::mutex> lk(mu);
                if (!first_err) {
                    first_err = s2;
                    err_msg = g_err;
                }
                continue;
            }
            // padding adds N - n exact zero singular values: keep the n largest
            memcpy(sigma + soff[m], sg.data(), sizeof(double) * B->n[m]);
        }
    };
    const int nth = std::max(1, std::min(host_thread_count(), B->count));
    std::vector<std::thread> th;
    for (int t = 1; t < nth; ++t) th.emplace_back(work);
    work();
    for (auto &x : th) x.join();
    if (first_err) return fail(first_err, "%s", err_msg.c_str());
    return AVS_OK;
}

} // extern "C"
