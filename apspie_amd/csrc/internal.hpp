// internal.hpp -- context, error plumbing and scratch pool shared by the
// library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/apspie.h"

namespace apspie {

// Error carried from deep inside an entry point to its C boundary.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &msg) : std::runtime_error(msg), code(c) {}
};

#define APSPIE_HIP(call)                                                                  \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            throw ::apspie::Error(APSPIE_EHIP, std::string(#call) + ": " +                \
                                                   hipGetErrorString(e_) + " (" __FILE__ \
                                                   ":" + std::to_string(__LINE__) + ")"); \
    } while (0)

#define APSPIE_CHECK(cond, code, msg)                         \
    do {                                                      \
        if (!(cond)) throw ::apspie::Error((code), (msg));    \
    } while (0)

constexpr size_t PINNED_BYTES = 16384;  // pinned host staging per context

// Named slots of the grow-only scratch pool.  A slot keeps the largest size
// asked of it; memory is released only when the context is destroyed.
enum Slot : int {
    S_FLAGS = 0,   // uint8 per vertex: "touched this round"
    S_VISITED,     // uint32 bitmap: visited vertices (BFS)
    S_QUEUE,       // int32 frontier vertex ids
    S_QSTART,      // int32 row start of each frontier vertex
    S_QOFF,        // int32 exclusive edge offset of each frontier vertex
    S_QVAL,        // float per frontier vertex (SSSP distance snapshot)
    S_TILES,       // int32 first frontier index of each edge tile
    S_STATUS,      // uint64 look-back status, two parities
    S_CTL,         // small control block
    S_PREV,        // uint32 per vertex: SSSP distance bits at the round start
    S_FBITS,       // uint32 bitmap: the frontier a compaction produced (pull steps)
    S_TROWS,       // int32 per CSC edge tile: row holding the tile's first edge
    S_FVAL,        // uint32 per vertex: SSSP frontier snapshot for pull steps
    S_BINOUT,      // uint2 per frontier edge: binned SSSP candidates, by bin workgroup
    S_BINTBL,      // int32 per (bucket, bin workgroup): candidate counts, then offsets inside the bucket
    S_BINTOT,      // int32 per bucket: binned candidates per bucket
    S_QUEUE1,      // BFS with a CSC: the queue arrays of odd rounds
    S_QSTART1,
    S_QOFF1,
    S_TILES1,
    S_HINT,        // uint8 per vertex: SSSP distance bound codes
    S_BINSORT,     // uint2 per candidate: binned SSSP candidates ordered by bucket
    S_AUX0,
    S_AUX1,
    S_AUX2,
    S_AUX3,
    S_AUX4,
    S_AUX5,
    S_AUX6,
    S_AUX7,
    S_OUTER0,                   // outer-product SpGEMM working arrays: S_OUTER0 + k, k < 40
    S_TC0 = S_OUTER0 + 40,      // triangle counting's partner arrays: S_TC0 + k, k < 12
    S_NSLOTS = S_TC0 + 12
};

}  // namespace apspie

struct apspie_ctx {
    int device = 0;
    int cu_count = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    void *slot_ptr[apspie::S_NSLOTS] = {};
    size_t slot_bytes[apspie::S_NSLOTS] = {};
    void *pinned = nullptr;  // PINNED_BYTES of pinned host staging for counters
    void *hslot_ptr[4] = {};  // growable pinned host staging (hslot)
    void *hslot_dev[4] = {};  // the same memory as kernels address it
    size_t hslot_bytes[4] = {};
    std::mutex mu;

    // Per-kernel-class event timing (off unless enabled).
    bool profiling = false;
    std::vector<hipEvent_t> prof_pool;                 // free events
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> prof_open;  // (class, (start, stop))
    double prof_ms[APSPIE_KCLASS_COUNT] = {};
    int64_t prof_launches[APSPIE_KCLASS_COUNT] = {};

    // Knobs of the traversals (apspie_ctx_set_tuning, apspie_autotune).
    struct {
        // SSSP pull and binned rounds are exact but measured slower than push
        // on RMAT s22 (the push round's distance probes mostly hit L2): off by
        // default
        int bfs_alpha = 14, bfs_beta = 24, sssp_pull_ratio = 0, sssp_bin_ratio = 0;
        int expand_wg_per_cu = 8, round_chunk = 4, batch_chunk = 4, sssp_batch_lanes = 0;
        int sssp_delta = 0;
        int spgemm_long_rows = 1;
        int spgemm_passes = 1;
        int bfs_group_cap = 0;
        int spgemm_esc_min = 1024;
        int sssp_hint = 1;
    } tune;

    // Instantiated hipGraphs of captured round chunks, keyed by everything
    // their kernel arguments depend on; dropped whenever a slot moves.
    std::map<std::string, hipGraphExec_t> graphs;
    void drop_graphs();

    // Device buffer of at least `bytes` bytes for slot `s`.  Growing drops
    // the old contents; callers initialise what they use.
    void *slot(int s, size_t bytes);
    // Pinned host buffer of at least `bytes` bytes for host slot `s` (< 4),
    // so copies from or into it run asynchronously on the stream instead of
    // waiting for the stream; growing synchronises the stream first.
    void *hslot(int s, size_t bytes);
    // Device address of host slot `s` (after hslot(s, ...)): kernels write
    // small results there directly, visible on the host once the stream is
    // synchronised, with no copy engine in between.
    void *hslot_device(int s) const { return hslot_dev[s]; }

    // Bracket one launch of kernel class `cls` (no-ops unless profiling).
    void prof_begin(int cls);
    void prof_end();
    // Fold finished brackets into the totals (stream must be idle).
    void prof_collect();
};

namespace apspie {

// Run `f` and turn exceptions into status codes at the C boundary.
template <class F>
int guarded(apspie_ctx *ctx, F &&f) {
    try {
        if (ctx) APSPIE_HIP(hipSetDevice(ctx->device));
        f();
        return APSPIE_OK;
    } catch (const Error &e) {
        if (ctx) ctx->err = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        if (ctx) ctx->err = "host allocation failed";
        return APSPIE_ENOMEM;
    } catch (const std::exception &e) {
        if (ctx) ctx->err = e.what();
        return APSPIE_EINVAL;
    }
}

void validate_csr(const apspie_csr *g, bool need_val);

namespace trav {
// A batched BFS (64, 128 or 256 lanes) or SSSP (16 or 64 lanes) that more
// sources may join between chunks of captured rounds, for the serving loop's
// continuous batching (batch.hip): begin seeds k sources in lanes 0 .. k-1
// (d_srcs on the device) with the output rows at d_out (K x m depths for BFS;
// SSSP writes its rows at end); step runs one chunk and returns false once
// the union frontier is empty; add seeds more sources in the next lanes,
// only after a step returned true (the rounds go on and take them); end
// writes rows 0 .. lanes-1 and stats.  All throw Error; free releases the
// handle.
struct BatchJoin;
BatchJoin *batch_join_begin(apspie_ctx *ctx, const apspie_csr *g, int kind, const int32_t *d_srcs, int k, int K,
                            void *d_out);
bool batch_join_step(BatchJoin *j);
int batch_join_lanes(const BatchJoin *j);
void batch_join_add(BatchJoin *j, const int32_t *d_srcs, int k);
void batch_join_end(BatchJoin *j, void *d_out, apspie_stats *stats);
void batch_join_free(BatchJoin *j);
}  // namespace trav

inline unsigned ceil_div(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

}  // namespace apspie
