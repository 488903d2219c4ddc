// schedule.cpp -- tile-DAG builder and dispatch ordering.
//
// Program order follows the reference's loops (KAbasedSVD/src/tiledalgos.jl:
// 17-34 for Blockbidiag!, :6-15 for tiled_QR!): step k = QR sweep of tile
// column k (GEQRT, UNMQR on row k, then per row: TSQRT + TSMQR), then the LQ
// sweep of tile row k from column k+1 (the same on adjoints, TiledMatrix.jl:
// 98-113).  With the hierarchical tree each sweep is split into domains of
// `domain` panel tiles; every domain runs the reference's flat chain on its
// own head tile, and the domain heads are merged pairwise (binary tree) by
// triangle-on-triangle TSQRT/TSMQR.
//
// Dependencies are expressed as per-tile write counts: a task that writes a
// tile waits for all earlier writers of that tile (program order), a task that
// reads reflectors waits for the write that produced them.  Reads of the
// reflector part of a tile never overlap the part later writers touch (GEQRT
// reflectors sit below the diagonal while later TSQRTs rewrite R above it;
// tree-merge reflectors sit in the upper triangle of a domain head whose lower
// triangle holds that head's GEQRT reflectors), so no write-after-read edges
// are needed.
#include "schedule.hpp"

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <functional>
#include <queue>
#include <string>
#include <utility>

namespace avs {

int domain_for(int m, int b, int domain)
{
    if (domain > 0) return domain;
    // (m_i, d_i) sorted by m_i; parsed once (a function-local static is
    // initialised thread-safely)
    static const std::vector<std::pair<int, int>> rule = [] {
        std::vector<std::pair<int, int>> r;
        if (const char *e = getenv("AVSVD_DOMAIN_RULE")) {
            std::string s(e);
            size_t pos = 0;
            while (pos < s.size()) {
                size_t end = s.find(',', pos);
                if (end == std::string::npos) end = s.size();
                const std::string item = s.substr(pos, end - pos);
                const size_t c = item.find(':');
                if (c != std::string::npos) {
                    const int m0 = atoi(item.substr(0, c).c_str()), d0 = atoi(item.substr(c + 1).c_str());
                    if (m0 >= 0 && d0 >= 1) r.push_back({m0, d0});
                }
                pos = end + 1;
            }
        }
        std::sort(r.begin(), r.end());
        return r;
    }();
    if (rule.empty()) {
        // a long sweep is bound by its updates: larger domains move merge (TT)
        // and GEQRT update work into the TSQRT updates, which run at the
        // highest MFMA rate; short sweeps are bound by their panel chain,
        // shortest with two tiles per domain (n=16384 FP64 b=128: 376 ms with
        // 4 from 10240 rows and 3 from 5120, 379 ms with 4 from 10240, 397 ms
        // with 2 throughout)
        const long rows = (long)m * b;
        return rows >= 10240 ? 4 : (rows >= 5120 ? 3 : 2);
    }
    int d = rule.front().second;
    for (const auto &x : rule)
        if (m >= x.first) d = x.second;
    return d;
}

namespace {

struct Builder {
    Program &P;
    int nt, B, upd_tiles;
    std::vector<std::vector<int>> writers; // per tile: op indices in program order
    std::vector<std::vector<std::pair<int, int>>> deps;
    std::vector<std::vector<int>> sigs;
    std::vector<int> vver; // per op (panels): write count of the reflector tile after it

    Builder(Program &p, int nt_, int B_, int ut) : P(p), nt(nt_), B(B_), upd_tiles(ut), writers((size_t)nt_ * nt_) {}

    int tid(int i, int j) const { return j * nt + i; }

    int add(Task t, const std::vector<int> &wr, int vtile, int vop)
    {
        const int id = (int)P.prog.size();
        std::vector<std::pair<int, int>> d;
        for (int x : wr)
            if (!writers[x].empty()) d.push_back({x, (int)writers[x].size()});
        if (vop >= 0) d.push_back({vtile, vver[vop]});
        for (int x : wr) writers[x].push_back(id);
        P.prog.push_back(t);
        deps.push_back(std::move(d));
        sigs.push_back(wr);
        vver.push_back(t.kind == TK_PANEL ? (int)writers[t.mode == MODE_GE ? t.vtile : t.btile].size() : 0);
        return id;
    }

    int panel(int mode, int trans, int top, int bot, int fac)
    {
        Task t{};
        t.kind = TK_PANEL;
        t.mode = (int8_t)mode;
        t.trans = (int8_t)trans;
        t.vtile = top;
        t.btile = bot;
        t.fac = fac;
        std::vector<int> wr{top};
        if (mode != MODE_GE) wr.push_back(bot);
        P.n_panel++;
        return add(t, wr, 0, -1);
    }

    // reflectors of panel op `pop` (tile vt) applied to `count` column tiles
    // starting at c2first (C2) / c1first (C1), tile step `step`
    void updates(int mode, int trans, int pop, int vt, int c1first, int c2first, int step, int count, int fac)
    {
        for (int c0 = 0; c0 < count; c0 += upd_tiles) {
            const int nti = std::min(upd_tiles, count - c0);
            Task t{};
            t.kind = TK_UPDATE;
            t.mode = (int8_t)mode;
            t.trans = (int8_t)trans;
            t.ntiles = (int8_t)nti;
            t.vtile = vt;
            t.btile = c2first + c0 * step;
            t.c1tile = mode == MODE_GE ? t.btile : c1first + c0 * step;
            t.cstep = step;
            t.fac = fac;
            std::vector<int> wr;
            for (int q = 0; q < nti; ++q) {
                wr.push_back(t.btile + q * step);
                if (mode != MODE_GE) wr.push_back(t.c1tile + q * step);
            }
            P.n_update++;
            add(t, wr, vt, pop);
        }
    }

    // one QR sweep (column k) or LQ sweep (row k)
    void sweep(int k, bool lq, int tree, int domain)
    {
        const int m = lq ? nt - k - 1 : nt - k;
        if (m <= 0) return;
        auto T = [&](int p) { return lq ? tid(k, k + 1 + p) : tid(k + p, k); };
        auto U0 = [&](int p) { return lq ? tid(k + 1, k + 1 + p) : tid(k + p, k + 1); };
        const int step = lq ? 1 : nt;
        const int cnt = nt - k - 1;
        const int nt2 = nt * nt;
        const int famGE = lq ? 2 : 0, famTT = lq ? 3 : 1;
        const int h = tree == 0 ? m : domain_for(m, B, domain);
        std::vector<int> heads;
        for (int d0 = 0; d0 < m; d0 += h) {
            heads.push_back(d0);
            const int d1 = std::min(d0 + h, m);
            int fac = famGE * nt2 + T(d0);
            int op = panel(MODE_GE, lq, T(d0), T(d0), fac);
            if (cnt > 0) updates(MODE_GE, lq, op, T(d0), U0(d0), U0(d0), step, cnt, fac);
            for (int p = d0 + 1; p < d1; ++p) {
                fac = famGE * nt2 + T(p);
                op = panel(MODE_TS, lq, T(d0), T(p), fac);
                if (cnt > 0) updates(MODE_TS, lq, op, T(p), U0(d0), U0(p), step, cnt, fac);
            }
        }
        const int nh = (int)heads.size();
        for (int s = 1; s < nh; s *= 2) {
            for (int a = 0; a + s < nh; a += 2 * s) {
                const int pa = heads[a], pb = heads[a + s];
                const int fac = famTT * nt2 + T(pb);
                const int op = panel(MODE_TT, lq, T(pa), T(pb), fac);
                if (cnt > 0) updates(MODE_TT, lq, op, T(pb), U0(pa), U0(pb), step, cnt, fac);
            }
        }
    }
};

double task_cost_us(const Task &t, int B)
{
    const double b = B / 128.0;
    if (t.kind == TK_PANEL) return 28.0 * b + 2.0;
    double c = 34.0 * b * b * b * t.ntiles;
    if (t.mode == MODE_GE) c *= 0.6;
    return c + 1.0;
}

} // namespace

void build_program(Program &P, int nt, int B, int alg, int kend, int tree, int domain, int upd_tiles, int workers)
{
    P = Program();
    P.nt = nt;
    P.B = B;
    P.alg = alg;
    P.kend = kend;
    P.tree = tree;
    P.domain = domain;
    P.upd_tiles = upd_tiles;
    P.nfactor_slots = 4 * nt * nt;
    Builder bd(P, nt, B, upd_tiles);
    for (int k = 0; k < kend && k < nt; ++k) {
        bd.sweep(k, false, tree, domain);
        if (alg == ALG_TILED_QR) continue;
        if (k == nt - 1) break;
        bd.sweep(k, true, tree, domain);
    }
    const int N = (int)P.prog.size();
    // flatten dependencies/signals in program order
    for (int i = 0; i < N; ++i) {
        Task &t = P.prog[i];
        t.dep0 = (int)P.dep_tile.size();
        t.ndep = (int)bd.deps[i].size();
        for (auto &d : bd.deps[i]) {
            P.dep_tile.push_back(d.first);
            P.dep_val.push_back(d.second);
        }
        t.sig0 = (int)P.sig_tile.size();
        t.nsig = (int)bd.sigs[i].size();
        for (int x : bd.sigs[i]) P.sig_tile.push_back(x);
    }
    // predecessor edges: dep (tile x, count v) -> the v-th writer of x
    std::vector<int> npred(N, 0);
    std::vector<int> soff(N + 1, 0);
    std::vector<std::pair<int, int>> edges;
    edges.reserve(P.dep_tile.size());
    for (int i = 0; i < N; ++i) {
        const Task &t = P.prog[i];
        for (int d = 0; d < t.ndep; ++d) {
            const int x = P.dep_tile[t.dep0 + d], v = P.dep_val[t.dep0 + d];
            const int pred = bd.writers[x][v - 1];
            edges.push_back({pred, i});
        }
    }
    std::sort(edges.begin(), edges.end());
    edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
    std::vector<int> succ(edges.size());
    for (auto &e : edges) soff[e.first + 1]++;
    for (int i = 0; i < N; ++i) soff[i + 1] += soff[i];
    {
        std::vector<int> fill(soff.begin(), soff.end() - 1);
        for (auto &e : edges) {
            succ[fill[e.first]++] = e.second;
            npred[e.second]++;
        }
    }
    // upward rank (critical path to the end), program order is topological
    std::vector<double> cost(N), rank(N);
    for (int i = 0; i < N; ++i) cost[i] = task_cost_us(P.prog[i], B);
    for (int i = N - 1; i >= 0; --i) {
        double r = 0.0;
        for (int s = soff[i]; s < soff[i + 1]; ++s) r = std::max(r, rank[succ[s]]);
        rank[i] = cost[i] + r;
    }
    // list scheduling on `workers` identical workers, highest rank first
    if (workers < 1) workers = 1;
    auto cmp_rank = [&](int a, int b) { return rank[a] < rank[b] || (rank[a] == rank[b] && a > b); };
    std::priority_queue<int, std::vector<int>, decltype(cmp_rank)> ready(cmp_rank);
    typedef std::pair<double, int> Ev;
    std::priority_queue<Ev, std::vector<Ev>, std::greater<Ev>> running;
    for (int i = 0; i < N; ++i)
        if (npred[i] == 0) ready.push(i);
    P.order.clear();
    P.order.reserve(N);
    int free_workers = workers;
    double now = 0.0;
    while ((int)P.order.size() < N) {
        while (free_workers > 0 && !ready.empty()) {
            const int i = ready.top();
            ready.pop();
            P.order.push_back(i);
            running.push({now + cost[i], i});
            --free_workers;
        }
        if (running.empty()) break; // cannot happen for a DAG
        const Ev e = running.top();
        running.pop();
        now = e.first;
        ++free_workers;
        for (int s = soff[e.second]; s < soff[e.second + 1]; ++s)
            if (--npred[succ[s]] == 0) ready.push(succ[s]);
    }
    while (!running.empty()) {
        now = running.top().first;
        running.pop();
    }
    P.est_makespan_us = now;
}

namespace {

struct PhaseBuilder {
    PhaseProgram &P;
    int nt;
    explicit PhaseBuilder(PhaseProgram &p, int nt_) : P(p), nt(nt_) {}
    int tid(int i, int j) const { return j * nt + i; }
    // opens a phase; chains are appended with chain()
    int cur_trans = 0;
    void begin(int kind, int ncols, int mode)
    {
        Phase ph{kind, (int)P.chain_start.size() - 1, 0, ncols, cur_trans, mode};
        P.phases.push_back(ph);
    }
    void chain(const std::vector<ChainOp> &ops)
    {
        for (auto &o : ops) P.ops.push_back(o);
        P.chain_start.push_back((int)P.ops.size());
        P.phases.back().nchains++;
    }
    ChainOp panel(int mode, int trans, int top, int bot, int fac)
    {
        ChainOp o{};
        o.kind = TK_PANEL;
        o.mode = (int8_t)mode;
        o.trans = (int8_t)trans;
        o.vtile = top;
        o.btile = bot;
        o.fac = fac;
        return o;
    }
    ChainOp update(int mode, int trans, int vt, int c1first, int c2first, int step, int cnt, int fac)
    {
        ChainOp o{};
        o.kind = TK_UPDATE;
        o.mode = (int8_t)mode;
        o.trans = (int8_t)trans;
        o.vtile = vt;
        o.btile = c2first;
        o.c1tile = mode == MODE_GE ? c2first : c1first;
        o.cstep = step;
        o.fac = fac;
        o.ncols = cnt;
        return o;
    }
    void sweep(int k, bool lq, int tree, int domain)
    {
        const int m = lq ? nt - k - 1 : nt - k;
        if (m <= 0) return;
        cur_trans = lq ? 1 : 0;
        auto T = [&](int p) { return lq ? tid(k, k + 1 + p) : tid(k + p, k); };
        auto U0 = [&](int p) { return lq ? tid(k + 1, k + 1 + p) : tid(k + p, k + 1); };
        const int step = lq ? 1 : nt;
        const int cnt = nt - k - 1;
        const int nt2 = nt * nt;
        const int famGE = lq ? 2 : 0, famTT = lq ? 3 : 1;
        const int h = tree == 0 ? m : domain_for(m, P.B, domain);
        std::vector<int> heads;
        for (int d0 = 0; d0 < m; d0 += h) heads.push_back(d0);
        // GEQRT of every domain head, then its update
        begin(TK_PANEL, 0, MODE_GE);
        for (int d0 : heads) chain({panel(MODE_GE, lq, T(d0), T(d0), famGE * nt2 + T(d0))});
        if (cnt > 0) {
            begin(TK_UPDATE, cnt, MODE_GE);
            for (int d0 : heads) chain({update(MODE_GE, lq, T(d0), U0(d0), U0(d0), step, cnt, famGE * nt2 + T(d0))});
        }
        // the TSQRT chain of every domain (on its head's R), then its updates
        if (h > 1) {
            begin(TK_PANEL, 0, MODE_TS);
            for (int d0 : heads) {
                const int d1 = std::min(d0 + h, m);
                if (d1 - d0 < 2) continue;
                std::vector<ChainOp> c;
                for (int p = d0 + 1; p < d1; ++p) c.push_back(panel(MODE_TS, lq, T(d0), T(p), famGE * nt2 + T(p)));
                chain(c);
            }
            if (P.phases.back().nchains == 0) P.phases.pop_back();
            else if (cnt > 0) {
                begin(TK_UPDATE, cnt, MODE_TS);
                for (int d0 : heads) {
                    const int d1 = std::min(d0 + h, m);
                    if (d1 - d0 < 2) continue;
                    std::vector<ChainOp> c;
                    for (int p = d0 + 1; p < d1; ++p)
                        c.push_back(update(MODE_TS, lq, T(p), U0(d0), U0(p), step, cnt, famGE * nt2 + T(p)));
                    chain(c);
                }
            }
        }
        // pairwise merges of the domain heads, one level per phase pair
        const int nh = (int)heads.size();
        for (int s2 = 1; s2 < nh; s2 *= 2) {
            std::vector<std::pair<int, int>> pairs;
            for (int a = 0; a + s2 < nh; a += 2 * s2) pairs.push_back({heads[a], heads[a + s2]});
            begin(TK_PANEL, 0, MODE_TT);
            for (auto &pr : pairs) chain({panel(MODE_TT, lq, T(pr.first), T(pr.second), famTT * nt2 + T(pr.second))});
            if (cnt > 0) {
                begin(TK_UPDATE, cnt, MODE_TT);
                for (auto &pr : pairs)
                    chain({update(MODE_TT, lq, T(pr.second), U0(pr.first), U0(pr.second), step, cnt,
                                  famTT * nt2 + T(pr.second))});
            }
        }
    }
};

} // namespace

void build_phase_program(PhaseProgram &P, int nt, int B, int alg, int kend, int tree, int domain)
{
    P = PhaseProgram();
    P.nt = nt;
    P.B = B;
    P.alg = alg;
    P.kend = kend;
    P.tree = tree;
    P.domain = domain;
    P.chain_start.push_back(0);
    PhaseBuilder pb(P, nt);
    for (int k = 0; k < kend && k < nt; ++k) {
        pb.sweep(k, false, tree, domain);
        if (alg == ALG_TILED_QR) continue;
        if (k == nt - 1) break;
        pb.sweep(k, true, tree, domain);
    }
}

} // namespace avs
