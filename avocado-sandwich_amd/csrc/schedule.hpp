// schedule.hpp -- tile-DAG construction for tiled_QR! / Blockbidiag!.
#pragma once
#include <vector>

#include "task.hpp"

namespace avs {

enum { ALG_BLOCKBIDIAG = 0, ALG_TILED_QR = 1 };

struct Program {
    int nt = 0, B = 0, alg = 0, kend = 0, tree = 0, domain = 0, upd_tiles = 1;
    std::vector<Task> prog;  // program order (a valid sequential order)
    std::vector<int> order;  // dispatch order for the persistent executor (indices into prog)
    std::vector<int> dep_tile, dep_val, sig_tile;
    int nfactor_slots = 0;
    double est_makespan_us = 0.0; // list-scheduling estimate
    long n_panel = 0, n_update = 0;
};

// Stream-ordered form of the same reduction: a sequence of phases, each one
// kernel launch whose workgroups run independent chains of operations of a
// single tile mode (a chain runs sequentially inside one workgroup).  Panel
// phases: one workgroup per chain.  Update phases: one workgroup per (chain,
// column group).  Per sweep: the GEQRT of every domain head, the TSQRT chain
// of every domain, then one TT merge level after another; each panel phase
// is followed by the update phase of the same operations.
struct Phase {
    int kind;         // TK_PANEL or TK_UPDATE
    int chain0, nchains;
    int ncols;        // update: column tiles of every op in the phase
    int trans;        // every op of the phase is QR (0) or LQ (1)
    int mode;         // every op of the phase has this MODE_* (one kernel body per launch)
};
struct PhaseProgram {
    int nt = 0, B = 0, alg = 0, kend = 0, tree = 0, domain = 0;
    std::vector<ChainOp> ops;
    std::vector<int> chain_start; // ops of chain c: [chain_start[c], chain_start[c+1])
    std::vector<Phase> phases;
};
void build_phase_program(PhaseProgram &P, int nt, int B, int alg, int kend, int tree, int domain);

// Tiles per domain of the hierarchical tree for a sweep over m panel tiles of
// size b: `domain` when it is positive, otherwise the automatic rule (4 when
// m*b >= 10240, 3 when m*b >= 5120, else 2; or the rule in AVSVD_DOMAIN_RULE = "m0:d0,m1:d1,...":
// d_i for m >= m_i, the largest such m_i applying; read once per process).
int domain_for(int m, int b, int domain);

// Builds the task DAG of Blockbidiag! (tiledalgos.jl:17-34) or tiled_QR!
// (tiledalgos.jl:6-15) over an nt x nt tile grid, for kend steps.
// tree: 0 = flat (reference), 1 = hierarchical with `domain` tiles per domain.
// upd_tiles: column tiles per update task.  workers: executor width used by
// the list-scheduling simulation that fixes the dispatch order.
void build_program(Program &P, int nt, int B, int alg, int kend, int tree, int domain, int upd_tiles, int workers);

} // namespace avs
