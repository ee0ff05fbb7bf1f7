// traverse.hip -- frontier traversal on gfx950: BFS (Boolean OR-AND SpMSpV)
// and SSSP (min-plus SpMSpV, synchronous Bellman-Ford), on the generic
// compaction / expansion kernels of frontier.hpp.
//
// BFS round:  expand  -- per edge (u -> c): flag[c] = 1 unless c is visited
//             compact -- flagged vertices get depth = level and the visited
//                        bit, and form the next queue; flags are cleared.
// SSSP round: expand  -- per edge: atomicMin(dist[c], d_u + w) when lower
//                        (d_u is the distance snapshot queued with u)
//             compact -- vertices whose distance dropped below the copy taken
//                        at the previous round (dist < prev) form the next
//                        queue with their new distance as snapshot, so the
//                        schedule is the Jacobi one and no per-edge flag is
//                        written.
// The host launches rounds in chunks and reads the frontier size once per
// chunk; kernels of an exhausted traversal exit at once.
//
// Replaces the per-level pipeline of src/spmspvBfs.cuh:174-222 and
// src/mXv.cuh:111-247 / src/spmspvSssp.cuh:92-291 (bitify, Scan with host
// total, streamCompact, IntervalGather/Expand, radix sort, ReduceByKey,
// scatter, updateBfs / addResultSssp).
#include <algorithm>
#include <cstring>

#include "frontier.hpp"
#include "internal.hpp"

namespace apspie {
namespace trav {

// ----------------------------------------------------------------------------
// per-algorithm pieces

struct PickBfs {
    uint8_t *flags;
    uint32_t *visited;
    int32_t *depth;
    int level, seed;
    __device__ bool pick(int v, uint32_t &) const {
        const uint8_t f = flags[v];  // flags are padded to nblk * CHUNK
        if (f) flags[v] = 0;
        return f != 0 || v == seed;
    }
    __device__ void on_round(uint64_t ball, int v0, int lane) const {
        if (lane == 0) {  // this wave alone owns the two bitmap words of v0 .. v0+63
            uint2 *w = (uint2 *)(visited + (v0 >> 5));
            uint2 o = *w;
            o.x |= uint32_t(ball);
            o.y |= uint32_t(ball >> 32);
            *w = o;
        }
    }
    __device__ void on_new(int v) const { depth[v] = level; }
};

struct PickSssp {
    uint32_t *dist, *prev;
    int m, seed;
    __device__ bool pick(int v, uint32_t &payload) const {
        if (v >= m) return false;
        uint32_t d = dist[v];
        const uint32_t p = prev[v];
        if (v == seed) {
            d = 0u;
            dist[v] = 0u;
        }
        payload = d;
        if (d < p) {
            prev[v] = d;
            return true;
        }
        return false;
    }
    __device__ void on_round(uint64_t, int, int) const {}
    __device__ void on_new(int) const {}
};

// frontier of one min-plus step: {u : -1 < x[u] < 1e38} (bitify,
// src/spmspvSssp.cuh:31-36), snapshot x[u]
struct PickStep {
    const float *x;
    int m;
    __device__ bool pick(int v, uint32_t &payload) const {
        if (v >= m) return false;
        const float xv = x[v];
        payload = __float_as_uint(xv);
        return xv > -1.0f && xv < 1e38f;
    }
    __device__ void on_round(uint64_t, int, int) const {}
    __device__ void on_new(int) const {}
};

struct OpBfs {
    static constexpr bool kVert = false, kVal = false;
    using Pre = uint32_t;
    const uint32_t *visited;
    uint8_t *flags;
    __device__ Pre load(int c, int, uint32_t, int) const { return visited[c >> 5]; }
    __device__ void apply(int c, int, uint32_t, int, Pre w) const {
        if (!((w >> (c & 31)) & 1u)) flags[c] = 1;
    }
};

struct OpRelax {
    static constexpr bool kVert = false, kVal = true;
    struct Pre {
        float w;
        uint32_t d;
    };
    const float *val;
    uint32_t *dist;
    __device__ Pre load(int c, int p, uint32_t, int) const { return Pre{val ? val[p] : 1.0f, dist[c]}; }
    __device__ void apply(int c, int, uint32_t du, int, const Pre &pr) const {
        const uint32_t nb = __float_as_uint(__uint_as_float(du) + pr.w);
        if (nb < pr.d) __hip_atomic_fetch_min(&dist[c], nb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
};

__global__ __launch_bounds__(WG) void k_init(int mode, int m, int nblk, int32_t *depth, uint32_t *dist,
                                             uint32_t *prev, uint8_t *flags, uint32_t *visited,
                                             uint64_t *status, Control *C) {
    const int64_t gtid = int64_t(blockIdx.x) * WG + threadIdx.x;
    const int64_t stride = int64_t(gridDim.x) * WG;
    const int64_t mpad = int64_t(nblk) * CHUNK;
    if (mode == MODE_BFS) {
        for (int64_t i = gtid; i < mpad / 16; i += stride) {
            *(uint4 *)(flags + 16 * i) = make_uint4(0, 0, 0, 0);
            if ((i & 1) == 0) visited[i >> 1] = 0u;  // one bitmap word per 32 vertices
        }
        for (int64_t i = gtid; i < m; i += stride) depth[i] = -1;
    } else {
        for (int64_t i = gtid; i < m; i += stride) {
            dist[i] = INF_BITS;
            if (prev) prev[i] = INF_BITS;
        }
    }
    for (int64_t i = gtid; i < 2 * int64_t(nblk); i += stride) status[i] = 0;
    if (gtid == 0) {
        Control z{};
        *C = z;
    }
}

// ----------------------------------------------------------------------------
// host side

Bufs get_bufs(apspie_ctx *ctx, int m, int64_t nnz) {
    Bufs b;
    b.nblk = std::max(1, (int)ceil_div(m, CHUNK));
    const size_t mpad = size_t(b.nblk) * CHUNK;
    b.flags = (uint8_t *)ctx->slot(S_FLAGS, mpad);
    b.visited = (uint32_t *)ctx->slot(S_VISITED, mpad / 8);
    b.prev = (uint32_t *)ctx->slot(S_PREV, (size_t(m) + 1) * 4);
    b.queue = (int32_t *)ctx->slot(S_QUEUE, (size_t(m) + 1) * 4);
    b.qstart = (int32_t *)ctx->slot(S_QSTART, (size_t(m) + 1) * 4);
    b.qoff = (int32_t *)ctx->slot(S_QOFF, (size_t(m) + 1) * 4);
    b.qval = (float *)ctx->slot(S_QVAL, (size_t(m) + 1) * 4);
    b.tiles = (int32_t *)ctx->slot(S_TILES, (size_t(nnz) / TILE + 4) * 4);
    b.status = (uint64_t *)ctx->slot(S_STATUS, size_t(b.nblk) * 2 * 8);
    b.C = (Control *)ctx->slot(S_CTL, sizeof(Control));
    b.expand_grid = std::max(1u, std::min(ceil_div(nnz, TILE) + 1, (unsigned)ctx->cu_count * 8u));
    return b;
}

Queue queue_of(const Bufs &b) { return Queue{b.queue, b.qstart, b.qoff, (uint32_t *)b.qval, b.tiles}; }

unsigned init_grid(const Bufs &b, int m) {
    return std::max(1u, std::min(ceil_div(std::max<int64_t>(m, int64_t(b.nblk) * CHUNK / 16), WG), 4096u));
}

void fill_stats(apspie_ctx *ctx, const Bufs &b, int mode, int m, apspie_stats *stats) {
    Acc acc;
    APSPIE_HIP(hipMemcpyAsync(&acc, &b.C->acc, sizeof(Acc), hipMemcpyDeviceToHost, ctx->stream));
    APSPIE_HIP(hipStreamSynchronize(ctx->stream));
    ctx->prof_collect();
    if (!stats) return;
    float ms = 0.f;
    APSPIE_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    memset(stats, 0, sizeof(*stats));
    stats->edges = (int64_t)acc.edges;
    stats->vertices = (int64_t)acc.vertices;
    stats->iterations = acc.levels;
    stats->kernel_ms = ms;
    if (mode == MODE_BFS) {
        stats->bytes_alg = 8.0 * stats->edges + 20.0 * stats->vertices + 4.0 * m;
    } else if (mode == MODE_SSSP) {
        // improved (vertex, round) pairs: every frontier entry but the source
        stats->improved = stats->vertices > 0 ? stats->vertices - 1 : 0;
        stats->bytes_alg = 12.0 * stats->edges + 4.0 * stats->improved + 20.0 * stats->vertices + 8.0 * m;
    } else {
        stats->iterations = 1;
        stats->bytes_alg = 12.0 * stats->edges + 20.0 * stats->vertices + 8.0 * m;
    }
}

namespace {

template <class Pick>
void launch_compact(apspie_ctx *ctx, const Bufs &b, const apspie_csr *g, const Pick &pk, int level, int always) {
    ctx->prof_begin(APSPIE_KCLASS_COMPACT);
    k_compact<Pick><<<b.nblk, WGC, 0, ctx->stream>>>(pk, level, g->m, b.nblk, always, g->row_ptr, queue_of(b),
                                                      b.status, b.C);
    ctx->prof_end();
}

template <class Op>
void launch_expand(apspie_ctx *ctx, const Bufs &b, const apspie_csr *g, const Op &op, int level) {
    ctx->prof_begin(APSPIE_KCLASS_EXPAND);
    k_expand<Op><<<b.expand_grid, WG, 0, ctx->stream>>>(op, level, g->m, b.C, queue_of(b), g->col_ind);
    ctx->prof_end();
}

template <int MODE>
void run(apspie_ctx *ctx, const apspie_csr *g, int32_t src, int32_t *depth, float *dist_f,
         apspie_stats *stats) {
    const Bufs b = get_bufs(ctx, g->m, g->nnz);
    uint32_t *dist = (uint32_t *)dist_f;
    APSPIE_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    ctx->prof_begin(APSPIE_KCLASS_OTHER);
    k_init<<<init_grid(b, g->m), WG, 0, ctx->stream>>>(MODE, g->m, b.nblk, depth, dist, b.prev, b.flags,
                                                       b.visited, b.status, b.C);
    ctx->prof_end();
    auto compact = [&](int level, int seed) {
        if (MODE == MODE_BFS)
            launch_compact(ctx, b, g, PickBfs{b.flags, b.visited, depth, level, seed}, level, level == 0);
        else
            launch_compact(ctx, b, g, PickSssp{dist, b.prev, g->m, seed}, level, level == 0);
    };
    auto expand = [&](int level) {
        if (MODE == MODE_BFS)
            launch_expand(ctx, b, g, OpBfs{b.visited, b.flags}, level);
        else
            launch_expand(ctx, b, g, OpRelax{g->val, dist}, level);
    };
    compact(0, src);
    int level = 1, chunk = 4;
    volatile int32_t *host_nf = (volatile int32_t *)ctx->pinned;
    for (;;) {
        for (int i = 0; i < chunk; ++i, ++level) {
            expand(level);
            compact(level, -1);
        }
        APSPIE_HIP(hipGetLastError());
        APSPIE_HIP(hipMemcpyAsync((void *)host_nf, &b.C->ctl[(level - 1) & 1].nF, 4, hipMemcpyDeviceToHost,
                                  ctx->stream));
        APSPIE_HIP(hipStreamSynchronize(ctx->stream));
        if (*host_nf == 0) break;
        APSPIE_CHECK(level <= g->m + 2, APSPIE_ENOCONV, "traversal did not converge");
        chunk = std::min(chunk * 2, 32);
    }
    APSPIE_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    fill_stats(ctx, b, MODE, g->m, stats);
}

void validate_traversal(const apspie_csr *g, int32_t src) {
    validate_csr(g, false);
    APSPIE_CHECK(g->m == g->n, APSPIE_EINVAL, "traversal needs a square adjacency matrix");
    APSPIE_CHECK(src >= 0 && src < g->m, APSPIE_EINVAL, "source vertex out of range");
}

}  // namespace
}  // namespace trav
}  // namespace apspie

using namespace apspie;

extern "C" {

int apspie_bfs(apspie_ctx *ctx, const apspie_csr *g, int32_t src, int32_t *d_depth, apspie_stats *stats) {
    if (!ctx) return APSPIE_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    return guarded(ctx, [&] {
        trav::validate_traversal(g, src);
        APSPIE_CHECK(d_depth != nullptr, APSPIE_EINVAL, "null depth");
        trav::run<trav::MODE_BFS>(ctx, g, src, d_depth, nullptr, stats);
    });
}

int apspie_sssp(apspie_ctx *ctx, const apspie_csr *g, int32_t src, float *d_dist, apspie_stats *stats) {
    if (!ctx) return APSPIE_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    return guarded(ctx, [&] {
        trav::validate_traversal(g, src);
        APSPIE_CHECK(d_dist != nullptr, APSPIE_EINVAL, "null dist");
        trav::run<trav::MODE_SSSP>(ctx, g, src, nullptr, d_dist, stats);
    });
}

int apspie_sssp_step(apspie_ctx *ctx, const apspie_csr *g, const float *d_x, float *d_y,
                     apspie_stats *stats) {
    if (!ctx) return APSPIE_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    return guarded(ctx, [&] {
        validate_csr(g, false);
        APSPIE_CHECK(g->m == g->n, APSPIE_EINVAL, "step needs a square adjacency matrix");
        APSPIE_CHECK(d_x && d_y, APSPIE_EINVAL, "null vector");
        using namespace trav;
        const Bufs b = get_bufs(ctx, g->m, g->nnz);
        uint32_t *y = (uint32_t *)d_y;
        APSPIE_HIP(hipEventRecord(ctx->ev0, ctx->stream));
        ctx->prof_begin(APSPIE_KCLASS_OTHER);
        k_init<<<init_grid(b, g->m), WG, 0, ctx->stream>>>(MODE_STEP, g->m, b.nblk, nullptr, y, nullptr,
                                                           b.flags, b.visited, b.status, b.C);
        ctx->prof_end();
        launch_compact(ctx, b, g, PickStep{d_x, g->m}, 0, 1);
        launch_expand(ctx, b, g, OpRelax{g->val, y}, 1);
        APSPIE_HIP(hipGetLastError());
        APSPIE_HIP(hipEventRecord(ctx->ev1, ctx->stream));
        fill_stats(ctx, b, MODE_STEP, g->m, stats);
    });
}

}  // extern "C"
