// batch.hip -- many traversals of one graph at once.
//
// Batched BFS (k <= 256 sources): bit-parallel multi-source BFS.  Per vertex
// W = ceil(k/64) words of visited / next / frontier bits (vertex-major, the
// visited and next words of a vertex side by side in one record); the union
// frontier goes through the generic compaction and expansion, and every edge
// (u -> c) of it serves all sources at once:
//     next[c] |= frontier[u] & ~visited[c]      (64-bit atomicOr per word)
// skipping the atomic when next[c], read with visited[c], already holds the
// bits (most edges of a dense round bring bits another edge brought first).
// The compaction turns next into the new frontier bits, marks them visited and
// writes depth[s * m + c] = level for each new (s, c): per 64-vertex slice of
// a wave, one store instruction per source with a new vertex in the slice,
// so each store covers consecutive vertices of one depth row.
//
// Batched SSSP (k <= 64 sources): synchronous-round Bellman-Ford over the
// union frontier with vertex-major distances dist[v * K + s], so the probes of
// all sources for one edge touch one 256-byte row; the 64 lanes of a wave are
// the 64 sources of an edge.  A vertex enters the next frontier with the mask
// of sources whose distance it improved (atomicOr of a ballot); relaxations
// read live distances, so results equal k separate runs (the fixed point is
// the shortest-path distances) while rounds may be fewer.
//
// Rounds are device-driven (the round number lives in device memory), so a
// chunk of rounds is captured once into a hipGraph and replayed until the
// union frontier is empty.
#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "frontier.hpp"
#include "internal.hpp"

namespace apspie {
namespace trav {
namespace {

// st[v * 2W + w] = visited word w of v, st[v * 2W + W + w] = next word w
template <int W>
struct PickMs {
    uint64_t *st, *fr;
    int32_t *depth;
    const int *lvl;
    int m, k;
    const int32_t *off;  // per source: the level its seed is picked at (a source that joined late)
    __device__ bool pick(int v, uint32_t &) const {
        if (v >= m) return false;
        uint64_t *vis = st + int64_t(v) * 2 * W, *next = vis + W;
        uint64_t n[W];
        bool any = false;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            uint64_t x = next[w];
            if (x) {
                next[w] = 0;
                x &= ~vis[w];
            }
            n[w] = x;
            any |= x != 0;
        }
        if (!any) return false;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            fr[int64_t(v) * W + w] = n[w];
            if (n[w]) vis[w] |= n[w];
        }
        return true;
    }
    __device__ void on_new(int) const {}
    // depths of the new (source, vertex) pairs of the wave's vertices
    // vbase + 64 r + lane, r < rounds (lane r holds round r's ballot of picked
    // vertices): the wave ORs the new-source words of a 64-vertex slice and
    // stores, per source present, the level at the lanes holding it
    __device__ void on_balls(uint64_t ball, int vbase, int lane, int rounds) const {
        const int level = *(volatile const int *)lvl;
        int offw[W];  // lane l: the offset of source w * 64 + l (read by readlane below)
#pragma unroll
        for (int w = 0; w < W; ++w) offw[w] = off[w * 64 + lane];
        for (int r = 0; r < rounds; ++r) {
            const uint64_t picked = __shfl(ball, r, WAVE);
            if (!picked) continue;
            const int v = vbase + r * WAVE + lane;
            const bool mine = (picked >> lane) & 1ull;
#pragma unroll
            for (int w = 0; w < W; ++w) {
                const uint64_t nw = mine ? fr[int64_t(v) * W + w] : 0ull;
                uint64_t any = nw;
#pragma unroll
                for (int d = 1; d < WAVE; d <<= 1) {
                    const uint32_t lo = __shfl_xor(uint32_t(any), d, WAVE);
                    const uint32_t hi = __shfl_xor(uint32_t(any >> 32), d, WAVE);
                    any |= (uint64_t(hi) << 32) | lo;
                }
                while (any) {  // wave-uniform
                    const int b = __ffsll((unsigned long long)any) - 1;
                    const int bu = __builtin_amdgcn_readfirstlane(b);
                    const int src = w * 64 + bu;
                    const int d = level - __builtin_amdgcn_readlane(offw[w], bu);
                    if (((nw >> b) & 1ull) && src < k) depth[int64_t(src) * m + v] = d;
                    any &= any - 1;
                }
            }
        }
    }
};

template <int W>
struct OpMs {
    static constexpr bool kVert = true, kVal = false;
    static constexpr int kBatch = W >= 2 ? 2 : IPT;
    struct Pre {
        uint64_t f[W], v[W];  // frontier words of u; visited | next words of c
    };
    const uint64_t *fr;
    uint64_t *st;
    __device__ Pre load(int c, int, uint32_t, int u) const {
        Pre p;
        const uint64_t *rec = st + int64_t(c) * 2 * W;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            p.f[w] = fr[int64_t(u) * W + w];
            p.v[w] = rec[w] | rec[W + w];
        }
        return p;
    }
    __device__ void apply(int c, int, uint32_t, int, const Pre &p) const {
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const uint64_t x = p.f[w] & ~p.v[w];
            if (x) atomicOr((unsigned long long *)&st[int64_t(c) * 2 * W + W + w], (unsigned long long)x);
        }
    }
};

template <int W>
__global__ __launch_bounds__(256) void k_ms_init(int m, int k, const int32_t *srcs, uint64_t *st, int32_t *depth,
                                                 int nblk, uint64_t *status, Control *C) {
    const int64_t gtid = int64_t(blockIdx.x) * 256 + threadIdx.x, stride = int64_t(gridDim.x) * 256;
    for (int64_t i = gtid; i < int64_t(m) * 2 * W; i += stride) st[i] = 0;
    for (int64_t i = gtid; i < int64_t(m) * k; i += stride) depth[i] = -1;
    for (int64_t i = gtid; i < 2 * int64_t(nblk); i += stride) status[i] = 0;
    control_clear(C, gtid, stride, ~0u);
}

// sources srcs[0, k) in lanes lane0 .. lane0 + k - 1, whose seeds the
// compaction of round `level` picks: their depth rows start at -1 and their
// depths count from that round
template <int W>
__global__ __launch_bounds__(256) void k_ms_seed(int m, int k, int lane0, int level, const int32_t *srcs,
                                                 uint64_t *st, int32_t *depth, int32_t *off, bool rows) {
    const int64_t gtid = int64_t(blockIdx.x) * 256 + threadIdx.x, stride = int64_t(gridDim.x) * 256;
    if (rows)
        for (int64_t i = gtid; i < int64_t(m) * k; i += stride) depth[int64_t(lane0) * m + i] = -1;
    if (gtid < k) {
        const int s = lane0 + int(gtid);
        off[s] = level;
        atomicOr((unsigned long long *)&st[int64_t(srcs[gtid]) * 2 * W + W + s / 64], 1ull << (s % 64));
    }
}

// ---------------------------------------------------------------------------
// batched SSSP

struct PickChanged {
    uint64_t *chg, *fr;
    int m;
    __device__ bool pick(int v, uint32_t &) const {
        if (v >= m) return false;
        const uint64_t x = chg[v];
        if (!x) return false;
        chg[v] = 0;
        fr[v] = x;
        return true;
    }
    __device__ void on_new(int) const {}
};

// Expansion with lanes = (edge, source): a wave step covers E = 64 / K edge
// slots of one round (owners from the shared tile staging) with K lanes per
// edge, one per source, so each edge's relaxations read and update one
// coalesced K-float row; the ballot of improved lanes becomes the edge
// target's changed-source mask.
template <int K>
__global__ __launch_bounds__(WG) void k_expand_ms_sssp(int m, const Control *C, Queue q,
                                                       const int32_t *__restrict__ col_ind,
                                                       const float *__restrict__ val, uint32_t *dist,
                                                       const uint64_t *__restrict__ fr, uint64_t *chg) {
    constexpr int NW = WG / WAVE, SEG = TILE / NW, E = WAVE / K, U = K == 64 ? 16 : 8;
    static_assert(WAVE % K == 0, "K divides the wave");
    __shared__ int32_t s_own[TILE];
    __shared__ int32_t s_base[TILE + 1];
    __shared__ uint32_t s_val[1];
    __shared__ int32_t s_vert[TILE + 1];
    __shared__ int32_t s_wmax[NW];
    const int level = C->cur_level;
    const Round cur = round_of(C, level);
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
    const int sub = lane / K, src = lane % K;
    for (int t = blockIdx.x; t < cur.nTiles; t += gridDim.x) {
        int own[IPT];
        const int elim = tile_owners<false, true>(t, cur, q, s_own, s_base, s_val, s_vert, s_wmax, own);
#pragma unroll 1
        for (int r = 0; r < IPT; ++r) {
            const int s = wid * SEG + r * WAVE + lane;
            int c = -1, u = 0;
            float w = 0.f;
            uint64_t fu = 0;
            if (s < elim) {
                const int p = s_base[own[r]] + s;
                c = col_ind[p];
                w = val ? val[p] : 1.0f;
                u = s_vert[own[r]];
                if ((unsigned)c < (unsigned)m) fu = fr[u];
                else c = -1;
            }
            const uint32_t fu_lo = uint32_t(fu), fu_hi = uint32_t(fu >> 32);
#pragma unroll 1
            for (int j0 = 0; j0 < WAVE; j0 += E * U) {
                int cj[U], uj[U];
                float wj[U];
                bool act[U];
                uint32_t du[U], dc[U];
#pragma unroll
                for (int x = 0; x < U; ++x) {
                    const int j = j0 + x * E + sub;
                    uint32_t lo, hi;
                    if constexpr (K == WAVE) {
                        // one edge per wave step: its lane index is wave-uniform,
                        // so a scalar read of that lane replaces the LDS shuffle
                        cj[x] = __builtin_amdgcn_readlane(c, j);
                        uj[x] = __builtin_amdgcn_readlane(u, j);
                        wj[x] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(w), j));
                        lo = uint32_t(__builtin_amdgcn_readlane(int(fu_lo), j));
                        hi = uint32_t(__builtin_amdgcn_readlane(int(fu_hi), j));
                    } else {
                        cj[x] = __shfl(c, j, WAVE);
                        uj[x] = __shfl(u, j, WAVE);
                        wj[x] = __shfl(w, j, WAVE);
                        lo = __shfl(fu_lo, j, WAVE);
                        hi = __shfl(fu_hi, j, WAVE);
                    }
                    const uint64_t f = (uint64_t(hi) << 32) | lo;
                    act[x] = cj[x] >= 0 && ((f >> src) & 1ull);
                }
#pragma unroll
                for (int x = 0; x < U; ++x)
                    if (act[x]) {
                        du[x] = dist[int64_t(uj[x]) * K + src];
                        dc[x] = dist[int64_t(cj[x]) * K + src];
                    }
#pragma unroll
                for (int x = 0; x < U; ++x) {
                    bool imp = false;
                    if (act[x]) {
                        const uint32_t nb = __float_as_uint(__uint_as_float(du[x]) + wj[x]);
                        if (nb < dc[x]) {
                            __hip_atomic_fetch_min(&dist[int64_t(cj[x]) * K + src], nb, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT);
                            imp = true;
                        }
                    }
                    const uint64_t b = __ballot(imp);
                    if (b && src == 0) {
                        const uint64_t mine = K == 64 ? b : (b >> (sub * K)) & ((1ull << (K % 64)) - 1);
                        if (mine) atomicOr((unsigned long long *)&chg[cj[x]], (unsigned long long)mine);
                    }
                }
            }
        }
        __syncthreads();
    }
}

template <int K>
__global__ __launch_bounds__(256) void k_mss_init(int m, int k, const int32_t *srcs, uint32_t *dist, uint64_t *chg,
                                                  int nblk, uint64_t *status, Control *C) {
    const int64_t gtid = int64_t(blockIdx.x) * 256 + threadIdx.x, stride = int64_t(gridDim.x) * 256;
    for (int64_t i = gtid; i < int64_t(m) * K; i += stride) dist[i] = INF_BITS;
    for (int64_t i = gtid; i < m; i += stride) chg[i] = 0;
    for (int64_t i = gtid; i < 2 * int64_t(nblk); i += stride) status[i] = 0;
    control_clear(C, gtid, stride, ~0u);
}

// sources srcs[0, k) in lanes lane0 .. lane0 + k - 1
template <int K>
__global__ void k_mss_seed(int k, int lane0, const int32_t *srcs, uint32_t *dist, uint64_t *chg) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < k) {
        const int lane = lane0 + s;
        dist[int64_t(srcs[s]) * K + lane] = 0u;
        atomicOr((unsigned long long *)&chg[srcs[s]], 1ull << lane);
    }
}

// vertex-major [m][K] -> source-major [k][m]
template <int K>
__global__ __launch_bounds__(256) void k_mss_transpose(int m, int k, const uint32_t *src, uint32_t *dst) {
    __shared__ uint32_t tile[64][K + 1];
    const int v0 = blockIdx.x * 64;
    for (int i = threadIdx.x; i < 64 * K; i += 256) {
        const int vv = i / K, s = i % K;
        if (v0 + vv < m) tile[vv][s] = src[int64_t(v0 + vv) * K + s];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 64 * k; i += 256) {
        const int s = i / 64, vv = i % 64;
        if (v0 + vv < m) dst[int64_t(s) * m + v0 + vv] = tile[vv][s];
    }
}

// ---------------------------------------------------------------------------
// driver: device-driven rounds captured in chunks into a hipGraph

// The captured chunk of batch_chunk rounds.  The instantiated graph is
// cached in the context under `key`, which names every address and size its
// kernel arguments hold (slot growth drops the cache), so repeated calls on
// one graph skip capture and instantiation.
template <class F>
hipGraphExec_t chunk_exec(apspie_ctx *ctx, const Bufs &b, const std::string &base_key, F &&enqueue_round) {
    const int CH = ctx->tune.batch_chunk;
    // the chunk length and the launch shape are part of what was captured
    const std::string key = base_key + "|ch" + std::to_string(CH) + "|eg" + std::to_string(b.expand_grid);
    auto it = ctx->graphs.find(key);
    if (it != ctx->graphs.end()) return it->second;
    if (ctx->graphs.size() >= 64) ctx->drop_graphs();  // bound the cache
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    APSPIE_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < CH; ++i) enqueue_round();
    APSPIE_HIP(hipStreamEndCapture(ctx->stream, &graph));
    const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    APSPIE_HIP(ie);
    ctx->graphs[key] = exec;
    return exec;
}

// One replay of the chunk; false once the union frontier is empty (the
// frontier written by the last compaction that ran: an exhausted traversal
// stops advancing cur_level, see k_compact).
bool chunk_step(apspie_ctx *ctx, const Bufs &b, hipGraphExec_t exec, int &rounds, int max_rounds) {
    APSPIE_HIP(hipGraphLaunch(exec, ctx->stream));
    rounds += ctx->tune.batch_chunk;
    Control *hc = (Control *)ctx->pinned;
    // the descriptors and the round counter lead the control block
    static_assert(offsetof(Control, cur_level) + sizeof(int) <= 4096, "fits the pinned staging");
    APSPIE_HIP(hipMemcpyAsync(hc, b.C, offsetof(Control, cur_level) + sizeof(int), hipMemcpyDeviceToHost,
                              ctx->stream));
    APSPIE_HIP(hipStreamSynchronize(ctx->stream));
    if (hc->ctl[(hc->cur_level - 1) & 1].nF == 0) return false;
    APSPIE_CHECK(rounds <= max_rounds, APSPIE_ENOCONV, "batched traversal did not converge");
    return true;
}

// Replays the chunk until the union frontier is empty.
template <class F>
int run_rounds(apspie_ctx *ctx, const Bufs &b, const std::string &base_key, int max_rounds, F &&enqueue_round) {
    const hipGraphExec_t exec = chunk_exec(ctx, b, base_key, enqueue_round);
    int rounds = 0;
    while (chunk_step(ctx, b, exec, rounds, max_rounds)) {
    }
    return rounds;
}

std::string graph_key(const char *what, const apspie_csr *g, const Bufs &b, std::initializer_list<const void *> ptrs,
                      int extra) {
    char buf[96];
    std::string k(what);
    snprintf(buf, sizeof(buf), "|%p|%p|%p|%d|%lld|%d|%d|%p|%p", (const void *)g->row_ptr, (const void *)g->col_ind,
             (const void *)g->val, g->m, (long long)g->nnz, b.nblk, extra, (void *)b.C, (void *)b.queue);
    k += buf;
    for (const void *p : ptrs) {
        snprintf(buf, sizeof(buf), "|%p", p);
        k += buf;
    }
    return k;
}

template <int W>
void bfs_batch(apspie_ctx *ctx, const apspie_csr *g, const int32_t *srcs, int k, int32_t *depth,
               apspie_stats *stats) {
    std::unique_ptr<BatchJoin, void (*)(BatchJoin *)> j(
        batch_join_begin(ctx, g, APSPIE_REQ_BFS, srcs, k, 64 * W, depth), batch_join_free);
    while (batch_join_step(j.get())) {
    }
    batch_join_end(j.get(), depth, stats);
}

}  // namespace

// A batch in K lanes (BFS: 64, 128 or 256; SSSP: 16 or 64) that more
// sources may join between chunks of rounds.  SSSP: a joined source starts
// at distance 0 and is picked by the next compaction; the rounds relax live
// distances until nothing changes, so its distances are its shortest paths
// whichever round it joins in.  BFS: a joined source's seed is picked by the
// next compaction, whose round is the source's depth offset, so its depths
// count from there.
struct BatchJoin {
    apspie_ctx *ctx = nullptr;
    const apspie_csr *g = nullptr;
    int kind = 0, K = 0, lanes = 0, rounds = 0, level = 0;
    Bufs b;
    uint32_t *dist = nullptr;
    uint64_t *chg = nullptr, *fr = nullptr, *st = nullptr;
    int32_t *depth = nullptr, *off = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

template <int W>
void join_begin_bfs(BatchJoin &j, const int32_t *srcs, int k) {
    apspie_ctx *ctx = j.ctx;
    const apspie_csr *g = j.g;
    j.b = get_bufs(ctx, g->m, g->nnz);
    const Bufs &b = j.b;
    const size_t words = size_t(g->m) * W;
    j.st = (uint64_t *)ctx->slot(S_AUX0, words * 2 * 8);
    j.fr = (uint64_t *)ctx->slot(S_AUX2, words * 8);
    j.off = (int32_t *)ctx->slot(S_AUX3, 256 * 4);
    APSPIE_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    const unsigned gi = std::max(1u, std::min(ceil_div(int64_t(g->m) * std::max(W, k), 256), 8192u));
    k_ms_init<W><<<gi, 256, 0, ctx->stream>>>(g->m, k, srcs, j.st, j.depth, b.nblk, b.status, b.C);
    k_ms_seed<W><<<1, 256, 0, ctx->stream>>>(g->m, k, 0, 0, srcs, j.st, j.depth, j.off, false);
    const PickMs<W> pk{j.st, j.fr, j.depth, &b.C->cur_level, g->m, 64 * W, j.off};
    const OpMs<W> op{j.fr, j.st};
    // seeding round (round 0) runs eagerly, the rest from the captured graph
    k_compact<PickMs<W>><<<b.nblk, WGC, 0, ctx->stream>>>(pk, -1, g->m, b.nblk, 1, g->row_ptr, queue_of(b), b.C);
    APSPIE_HIP(hipGetLastError());
    const std::string key = graph_key("msbfs", g, b, {j.st, j.fr, j.depth, j.off, b.status, b.tiles}, W);
    j.exec = chunk_exec(ctx, b, key, [&] {
        k_expand<OpMs<W>><<<b.expand_grid, WG, 0, ctx->stream>>>(op, -1, g->m, b.C, queue_of(b), g->col_ind);
        k_compact<PickMs<W>><<<b.nblk, WGC, 0, ctx->stream>>>(pk, -1, g->m, b.nblk, 0, g->row_ptr, queue_of(b),
                                                               b.C);
    });
    j.lanes = k;
}

template <int K>
void join_begin(BatchJoin &j, const int32_t *srcs, int k) {
    apspie_ctx *ctx = j.ctx;
    const apspie_csr *g = j.g;
    j.b = get_bufs(ctx, g->m, g->nnz);
    const Bufs &b = j.b;
    j.dist = (uint32_t *)ctx->slot(S_AUX0, size_t(g->m) * K * 4);
    j.chg = (uint64_t *)ctx->slot(S_AUX1, size_t(g->m) * 8);
    j.fr = (uint64_t *)ctx->slot(S_AUX2, size_t(g->m) * 8);
    APSPIE_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    const unsigned gi = std::max(1u, std::min(ceil_div(int64_t(g->m) * K, 256), 8192u));
    k_mss_init<K><<<gi, 256, 0, ctx->stream>>>(g->m, k, srcs, j.dist, j.chg, b.nblk, b.status, b.C);
    k_mss_seed<K><<<1, 64, 0, ctx->stream>>>(k, 0, srcs, j.dist, j.chg);
    const PickChanged pk{j.chg, j.fr, g->m};
    k_compact<PickChanged><<<b.nblk, WGC, 0, ctx->stream>>>(pk, -1, g->m, b.nblk, 1, g->row_ptr, queue_of(b), b.C);
    APSPIE_HIP(hipGetLastError());
    const std::string key = graph_key("mssssp", g, b, {j.dist, j.chg, j.fr, b.status, b.tiles}, K);
    uint32_t *dist = j.dist;
    uint64_t *chg = j.chg, *fr = j.fr;
    j.exec = chunk_exec(ctx, b, key, [&] {
        k_expand_ms_sssp<K><<<b.expand_grid, WG, 0, ctx->stream>>>(g->m, b.C, queue_of(b), g->col_ind, g->val,
                                                                    dist, fr, chg);
        k_compact<PickChanged><<<b.nblk, WGC, 0, ctx->stream>>>(pk, -1, g->m, b.nblk, 0, g->row_ptr, queue_of(b),
                                                                 b.C);
    });
    j.lanes = k;
}

template <int K>
void join_end(BatchJoin &j, float *out, apspie_stats *stats) {
    apspie_ctx *ctx = j.ctx;
    k_mss_transpose<K><<<ceil_div(j.g->m, 64), 256, 0, ctx->stream>>>(j.g->m, j.lanes, j.dist, (uint32_t *)out);
    APSPIE_HIP(hipGetLastError());
    APSPIE_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    fill_stats(ctx, j.b, MODE_SSSP, j.g->m, stats);
    if (stats) {
        stats->improved = 0;
        stats->bytes_alg = 0.0;
    }
}

}  // namespace

BatchJoin *batch_join_begin(apspie_ctx *ctx, const apspie_csr *g, int kind, const int32_t *d_srcs, int k, int K,
                            void *d_out) {
    if (kind == APSPIE_REQ_BFS)
        APSPIE_CHECK(K == 64 || K == 128 || K == 256, APSPIE_EINVAL, "a joinable BFS batch has 64, 128 or 256 lanes");
    else
        APSPIE_CHECK(K == 16 || K == 64, APSPIE_EINVAL, "a joinable SSSP batch has 16 or 64 lanes");
    APSPIE_CHECK(k >= 1 && k <= K, APSPIE_ELIMIT, "sources beyond the batch's lanes");
    std::unique_ptr<BatchJoin> j(new BatchJoin);
    j->ctx = ctx;
    j->g = g;
    j->kind = kind;
    j->K = K;
    if (kind == APSPIE_REQ_BFS) {
        j->depth = (int32_t *)d_out;
        if (K == 64) join_begin_bfs<1>(*j, d_srcs, k);
        else if (K == 128) join_begin_bfs<2>(*j, d_srcs, k);
        else join_begin_bfs<4>(*j, d_srcs, k);
    } else if (K == 16) {
        join_begin<16>(*j, d_srcs, k);
    } else {
        join_begin<64>(*j, d_srcs, k);
    }
    return j.release();
}

bool batch_join_step(BatchJoin *j) {
    const bool more = chunk_step(j->ctx, j->b, j->exec, j->rounds, j->kind == APSPIE_REQ_BFS ? j->g->m + 2 : j->g->m * 2 + 4);
    j->level = ((const Control *)j->ctx->pinned)->cur_level;  // the round the next compaction runs
    return more;
}

int batch_join_lanes(const BatchJoin *j) { return j->lanes; }

void batch_join_add(BatchJoin *j, const int32_t *d_srcs, int k) {
    APSPIE_CHECK(k >= 1 && j->lanes + k <= j->K, APSPIE_ELIMIT, "sources beyond the batch's lanes");
    hipStream_t st = j->ctx->stream;
    const int m = j->g->m;
    if (j->kind == APSPIE_REQ_BFS) {
        const unsigned gi = std::max(1u, std::min(ceil_div(int64_t(m) * k, 256), 8192u));
        if (j->K == 64) k_ms_seed<1><<<gi, 256, 0, st>>>(m, k, j->lanes, j->level, d_srcs, j->st, j->depth, j->off, true);
        else if (j->K == 128)
            k_ms_seed<2><<<gi, 256, 0, st>>>(m, k, j->lanes, j->level, d_srcs, j->st, j->depth, j->off, true);
        else k_ms_seed<4><<<gi, 256, 0, st>>>(m, k, j->lanes, j->level, d_srcs, j->st, j->depth, j->off, true);
    } else if (j->K == 16) {
        k_mss_seed<16><<<1, 64, 0, st>>>(k, j->lanes, d_srcs, j->dist, j->chg);
    } else {
        k_mss_seed<64><<<1, 64, 0, st>>>(k, j->lanes, d_srcs, j->dist, j->chg);
    }
    APSPIE_HIP(hipGetLastError());
    j->lanes += k;
}

void batch_join_end(BatchJoin *j, void *d_out, apspie_stats *stats) {
    if (j->kind == APSPIE_REQ_BFS) {
        APSPIE_HIP(hipEventRecord(j->ctx->ev1, j->ctx->stream));
        fill_stats(j->ctx, j->b, MODE_BFS, j->g->m, stats);
        if (stats) stats->bytes_alg = 0.0;
    } else if (j->K == 16) {
        join_end<16>(*j, (float *)d_out, stats);
    } else {
        join_end<64>(*j, (float *)d_out, stats);
    }
}

void batch_join_free(BatchJoin *j) { delete j; }

namespace {

template <int K>
void sssp_batch(apspie_ctx *ctx, const apspie_csr *g, const int32_t *srcs, int k, float *out,
                apspie_stats *stats) {
    std::unique_ptr<BatchJoin, void (*)(BatchJoin *)> j(
        batch_join_begin(ctx, g, APSPIE_REQ_SSSP, srcs, k, K, out), batch_join_free);
    while (batch_join_step(j.get())) {
    }
    batch_join_end(j.get(), out, stats);
}

void check_sources(apspie_ctx *ctx, const apspie_csr *g, const int32_t *d_srcs, int k) {
    std::vector<int32_t> h(k);
    APSPIE_HIP(hipMemcpyAsync(h.data(), d_srcs, size_t(k) * 4, hipMemcpyDeviceToHost, ctx->stream));
    APSPIE_HIP(hipStreamSynchronize(ctx->stream));
    for (int s = 0; s < k; ++s) APSPIE_CHECK(h[s] >= 0 && h[s] < g->m, APSPIE_EINVAL, "source out of range");
}

}  // namespace
}  // namespace trav
}  // namespace apspie

using namespace apspie;

extern "C" {

int apspie_bfs_batch(apspie_ctx *ctx, const apspie_csr *g, const int32_t *d_srcs, int32_t k, int32_t *d_depth,
                     apspie_stats *stats) {
    if (!ctx) return APSPIE_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    return guarded(ctx, [&] {
        validate_csr(g, false);
        APSPIE_CHECK(g->m == g->n, APSPIE_EINVAL, "traversal needs a square adjacency matrix");
        APSPIE_CHECK(k >= 1 && k <= 256, APSPIE_ELIMIT, "batched BFS takes 1..256 sources");
        APSPIE_CHECK(d_srcs && d_depth, APSPIE_EINVAL, "null sources or output");
        trav::check_sources(ctx, g, d_srcs, k);
        const int W = (k + 63) / 64;
        if (W == 1) trav::bfs_batch<1>(ctx, g, d_srcs, k, d_depth, stats);
        else if (W == 2) trav::bfs_batch<2>(ctx, g, d_srcs, k, d_depth, stats);
        else trav::bfs_batch<4>(ctx, g, d_srcs, k, d_depth, stats);
    });
}

int apspie_sssp_batch(apspie_ctx *ctx, const apspie_csr *g, const int32_t *d_srcs, int32_t k, float *d_dist,
                      apspie_stats *stats) {
    if (!ctx) return APSPIE_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    return guarded(ctx, [&] {
        validate_csr(g, false);
        APSPIE_CHECK(g->m == g->n, APSPIE_EINVAL, "traversal needs a square adjacency matrix");
        APSPIE_CHECK(k >= 1 && k <= 64, APSPIE_ELIMIT, "batched SSSP takes 1..64 sources");
        APSPIE_CHECK(d_srcs && d_dist, APSPIE_EINVAL, "null sources or output");
        trav::check_sources(ctx, g, d_srcs, k);
        const int lanes = ctx->tune.sssp_batch_lanes;
        if (k <= 16 && lanes != 64) trav::sssp_batch<16>(ctx, g, d_srcs, k, d_dist, stats);
        else trav::sssp_batch<64>(ctx, g, d_srcs, k, d_dist, stats);
    });
}

}  // extern "C"
