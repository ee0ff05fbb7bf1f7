// traverse.hpp -- shapes and device-resident control state of the frontier
// traversal kernels (traverse.hip).
#pragma once

#include <cstdint>

#include "internal.hpp"

namespace apspie {
namespace trav {

constexpr int WG = 256;           // threads per expansion workgroup (4 waves of 64)
constexpr int WGC = 1024;         // threads per compaction workgroup (16 waves)
constexpr int VPT = 4;           // vertices per thread in the compaction pass
constexpr int CHUNK = WGC * VPT;  // vertices per compaction workgroup
constexpr int TILE = 2048;        // frontier edges per expansion tile
constexpr int IPT = TILE / WG;    // edges per thread per tile
constexpr uint32_t INF_BITS = 0x7effffeeu;  // bit pattern of 1.70141e38f

enum { MODE_BFS = 0, MODE_SSSP = 1, MODE_STEP = 2 };

// Frontier descriptor written by a compaction, read by the next expansion.
struct Ctl {
    int32_t nF;      // queued frontier vertices (out-degree > 0)
    int32_t E;       // edges of the queued frontier
    int32_t nTiles;  // ceil(E / TILE)
    int32_t pull;    // how the next round expands this frontier: DIR_*
};

enum {
    DIR_PUSH = 0,    // k_expand over the queued frontier
    DIR_PULL = 1,    // pull step over the CSC
    DIR_BINNED = 2,  // SSSP: binned relaxation (k_relax's bin_tiles, k_bin_scan, k_bin_scatter, k_bin_apply)
    DIR_REPICK = 3,  // SSSP near/far: nothing to expand; changed vertices wait at or past the
                     // threshold, which was raised, and the next compaction picks again
};

// Counters every compaction workgroup adds to are spread over SPREAD slots
// (workgroup id mod SPREAD): a single address serialises the atomics of all
// workgroups.  Readers sum the slots.
constexpr int SPREAD = 16;

// A counter alone in a 256-byte slot.  Device-scope atomics that fall in one
// 64-byte line are serialised where memory performs them (about 12 ns each
// on MI355X: 4096 adds to one address take 47 us, to 16 adjacent 8-byte
// slots 28 us, to 16 slots 256 bytes apart 1.7 us), so the slots of a spread
// counter must not share a line.
template <class T>
struct alignas(256) Padded {
    T v;
};

// Totals over one call.
struct Acc {
    unsigned long long edges;             // sum of frontier edge counts
    Padded<unsigned long long> vertices[SPREAD];  // vertices that entered a frontier
    int32_t levels;                       // rounds with a non-empty frontier
    int32_t binned;                       // rounds marked binned (SSSP)
};

constexpr int LOG_ROUNDS = 64;  // rounds whose frontier descriptor is kept for apspie_ctx_rounds

// The frontier a compaction builds is split by chunk group: vertex chunk b
// (CHUNK vertices, one compaction workgroup) belongs to group b % NGROUP, and
// each group fills its own region of the queue and of the tile table.  A
// chunk takes its place in its group with one returning atomic add on the
// group's counter, so no workgroup waits for another and the counters are
// spread over NGROUP addresses (one address would serialise every
// workgroup's atomic).  Entry order inside a group is the order the chunks
// arrived in; every frontier algorithm here is order-independent.
constexpr int NGROUP = SPREAD;

struct Control {
    Ctl ctl[2];                 // by round parity: totals over the groups
    int cur_level;  // device-driven rounds: the round the next compaction runs
    // BFS rounds that queue their own frontier (k_round_bfs): round + 1 of a
    // round whose queue overflowed a group region and is rebuilt by a pass
    // over its bitmap before the next round runs (0: none); rounds exit at
    // once while it is set
    int rebuild;
    Acc acc;  // read with the descriptors once a traversal ends
    // per round parity and group: (entries << 32) | edges, and chunks done
    Padded<unsigned long long> gcnt[2][NGROUP];
    Padded<unsigned int> gdone[2][NGROUP];
    unsigned int alldone[2];    // groups whose chunks are all done
    int32_t gtile[2][NGROUP + 1];  // first tile of each group in the round's tile numbering
    Padded<unsigned int> nnew[2][SPREAD];  // vertices picked by the compaction, by round parity
    // SSSP near/far: a changed vertex enters the frontier only while its
    // distance is below thr (float bits; all ones: no threshold); the others
    // wait, counted (ndefer) with their smallest distance (dmin) by parity
    uint32_t thr;
    unsigned int ndefer[2];
    uint32_t dmin[2];
    // k_round_bfs, by round parity: vertices the round reached (by slot), and
    // whether an append found its group region full
    Padded<unsigned int> rfound[2][SPREAD];
    unsigned int qoverflow[2];
    Ctl log[LOG_ROUNDS];  // the descriptor each compaction wrote, by round
};

struct Bufs {
    int nblk = 0;
    unsigned expand_grid = 0;
    uint8_t *flags = nullptr;
    uint32_t *visited = nullptr;
    uint32_t *prev = nullptr;  // SSSP: distance bits at the start of the round
    // BFS with a CSC: three frontier bitmaps of fbw words each, the frontier
    // of round L in bitmap L mod 3 (round L reads bitmap L - 1, writes L and
    // clears L + 1)
    uint32_t *fbits = nullptr;
    int32_t fbw = 0;
    int32_t *trows = nullptr;   // SSSP pull: CSC row holding the first edge of each tile
    int32_t *queue = nullptr, *qstart = nullptr, *qoff = nullptr, *tiles = nullptr;
    int32_t qcap = 0, tcap = 0;  // entries and tiles per group region
    // BFS with a CSC: the queue of odd rounds (vertex, row start, edge offset,
    // tiles); even rounds use the arrays above
    int32_t *queue1 = nullptr, *qstart1 = nullptr, *qoff1 = nullptr, *tiles1 = nullptr;
    float *qval = nullptr;
    uint64_t *status = nullptr;
    Control *C = nullptr;
};

Bufs get_bufs(apspie_ctx *ctx, int m, int64_t nnz);

// Looks up the fixed scratch array of apspie_sssp_noalloc on `device` (the
// current device), once per device; apspie_ctx_create calls it.
void noalloc_setup(int device);

}  // namespace trav
}  // namespace apspie
