// task.hpp -- task descriptor shared by the schedule builder (host) and the
// persistent executor (device).
#pragma once
#include <stdint.h>

namespace avs {

enum { MODE_GE = 0, MODE_TS = 1, MODE_TT = 2 };

// ---------------------------------------------------------------------------
// Task list executed by the persistent kernel.  Tiles are addressed by index
// idx = j*nt + i into a tile-major array (tile (i,j) at idx*B*B, column-major
// inside).  Dependencies are per-tile completion counters: a task waits until
// cnt[tile] >= val for each (tile, val) in its dependency slice, and on
// completion adds 1 to the counter of every tile it wrote.
// ---------------------------------------------------------------------------
enum { TK_PANEL = 0, TK_UPDATE = 1 };

struct Task {
    int8_t kind;    // TK_PANEL / TK_UPDATE
    int8_t mode;    // MODE_GE / MODE_TS / MODE_TT
    int8_t trans;   // LQ twin
    int8_t ntiles;  // update: column tiles covered
    int32_t vtile;  // panel: top tile; update: reflector tile
    int32_t btile;  // panel: bottom tile (TS/TT); update: first C2 tile
    int32_t c1tile; // update: first C1 tile (TS/TT)
    int32_t cstep;  // update: tile-index step between consecutive column tiles
    int32_t fac;    // factor slot (tau + T_JJ)
    int32_t dep0, ndep; // slice of the dependency arrays
    int32_t sig0, nsig; // slice of the signal array (tiles written)
};

// One operation of a chain executed by the stream-ordered executor.  Panel
// ops: vtile = top tile, btile = bottom tile.  Update ops: vtile = reflector
// tile; btile / c1tile = first C2 / C1 tile of the row being updated, cstep =
// tile-index step between consecutive column tiles, ncols = column tiles.
struct ChainOp {
    int8_t kind, mode, trans, pad;
    int32_t vtile, btile, c1tile, cstep, fac, ncols;
};

} // namespace avs
