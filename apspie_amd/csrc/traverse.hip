// traverse.hip -- frontier traversal on gfx950: BFS (Boolean OR-AND SpMSpV)
// and SSSP (min-plus SpMSpV, synchronous Bellman-Ford).
//
// One round ("level") is two kernels:
//   expand   -- the frontier's edges, cut into equal tiles of TILE edges, are
//               gathered with a load-balanced search (each tile finds its
//               frontier vertices from a tile->vertex table built by the
//               previous compaction); per edge: BFS sets flag[c] if c is not
//               visited, SSSP does atomicMin on the distance bits and sets
//               flag[c] on improvement.  No sort, no dedup pass: the flag
//               byte array is the dedup.
//   compact  -- one pass over the vertex range in id order: flagged vertices
//               become the next frontier (queue, row start, edge offset via a
//               single-pass decoupled look-back scan), BFS writes depth and
//               the visited bitmap, and the tile->vertex table for the next
//               expand is written.  Flags are cleared on the way.
// The host launches rounds in chunks and reads the frontier size once per
// chunk; kernels of an exhausted traversal exit at once.
//
// Replaces the per-level pipeline of src/spmspvBfs.cuh:174-222 and
// src/mXv.cuh:111-247 / src/spmspvSssp.cuh:92-291 (bitify, Scan with host
// total, streamCompact, IntervalGather/Expand, radix sort, ReduceByKey,
// scatter, updateBfs / addResultSssp).
#include <algorithm>
#include <cstring>

#include "device.hpp"
#include "internal.hpp"
#include "traverse.hpp"

namespace apspie {
namespace trav {

using namespace dev;

// ----------------------------------------------------------------------------
// kernels

__global__ __launch_bounds__(WG) void k_init(int mode, int m, int nblk, int32_t *depth,
                                             uint32_t *dist_bits, const float *x, uint8_t *flags,
                                             uint32_t *visited, uint64_t *status, Control *C) {
    const int64_t gtid = int64_t(blockIdx.x) * WG + threadIdx.x;
    const int64_t stride = int64_t(gridDim.x) * WG;
    const int64_t mpad = int64_t(nblk) * CHUNK;
    for (int64_t i = gtid; i < mpad; i += stride) {
        uint8_t f = 0;
        if (i < m) {
            if (mode == MODE_BFS) depth[i] = -1;
            else dist_bits[i] = INF_BITS;
            if (mode == MODE_STEP) {
                const float xv = x[i];
                f = (xv > -1.0f && xv < 1e38f) ? 1 : 0;  // bitify, src/spmspvSssp.cuh:31-36
            }
        }
        flags[i] = f;
        if ((i & 31) == 0 && visited) visited[i >> 5] = 0;
    }
    for (int64_t i = gtid; i < 2 * int64_t(nblk); i += stride) status[i] = 0;
    if (gtid == 0) memset(C, 0, sizeof(Control));
}

__global__ __launch_bounds__(WG) void k_clear_flags(int nblk, uint8_t *flags) {
    const int64_t mpad = int64_t(nblk) * CHUNK;
    for (int64_t i = (int64_t(blockIdx.x) * WG + threadIdx.x) * 8; i < mpad; i += int64_t(gridDim.x) * WG * 8)
        *(uint2 *)(flags + i) = make_uint2(0, 0);
}

template <int MODE>
__global__ __launch_bounds__(WG) void k_compact(int level, int m, int nblk, int first, int seed,
                                                const int32_t *__restrict__ row_ptr, uint8_t *flags,
                                                uint32_t *visited, int32_t *depth, float *vsrc,
                                                int32_t *queue, int32_t *qstart, int32_t *qoff,
                                                float *qval, int32_t *tiles, uint64_t *status,
                                                Control *C) {
    __shared__ int s_bid;
    __shared__ uint64_t s_wsum[WG / WAVE];
    __shared__ uint64_t s_prefix;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
    const int par = level & 1;
    Ctl *next = &C->ctl[par];
    if (!first && C->ctl[par ^ 1].nF == 0) {
        if (blockIdx.x == 0 && tid == 0) *next = Ctl{0, 0, 0, 0};
        return;
    }
    if (tid == 0) s_bid = (int)atomicAdd(&C->blk_counter[par], 1u);
    __syncthreads();
    const int bid = s_bid;
    // state for the launch after this one (its previous user has finished)
    if (tid == 0) status[(par ^ 1) * nblk + bid] = 0;
    if (bid == 0 && tid == 0) C->blk_counter[par ^ 1] = 0;

    const int v0 = bid * CHUNK + tid * VPT;
    const uint2 fw = *(const uint2 *)(flags + v0);
    uint32_t bits = nonzero_bytes8(fw);
    if (seed >= v0 && seed < v0 + VPT) bits |= 1u << (seed - v0);
    if (fw.x | fw.y) *(uint2 *)(flags + v0) = make_uint2(0, 0);

    if (MODE == MODE_BFS) {
        // 4 threads share one 32-bit word of the visited bitmap
        uint32_t w = bits << ((tid & 3) * VPT);
        w |= __shfl_xor(w, 1, WAVE);
        w |= __shfl_xor(w, 2, WAVE);
        if ((tid & 3) == 0 && w) visited[v0 >> 5] |= w;
    }

    int32_t rp[VPT + 1];
    float vv[VPT];
    uint32_t cnt = 0, esum = 0;
    if (bits) {
        if (v0 + VPT <= m) {
            const int4 a = *(const int4 *)(row_ptr + v0);
            const int4 b = *(const int4 *)(row_ptr + v0 + 4);
            rp[0] = a.x; rp[1] = a.y; rp[2] = a.z; rp[3] = a.w;
            rp[4] = b.x; rp[5] = b.y; rp[6] = b.z; rp[7] = b.w;
            rp[8] = row_ptr[v0 + VPT];
        } else {
#pragma unroll
            for (int i = 0; i <= VPT; ++i) rp[i] = (v0 + i <= m) ? row_ptr[v0 + i] : 0;
        }
#pragma unroll
        for (int i = 0; i < VPT; ++i) {
            if (!((bits >> i) & 1u)) continue;
            const int v = v0 + i;
            if (MODE == MODE_BFS) depth[v] = level;
            if (MODE == MODE_SSSP) {
                if (v == seed) {
                    vsrc[v] = 0.0f;
                    vv[i] = 0.0f;
                } else {
                    vv[i] = vsrc[v];
                }
            }
            if (MODE == MODE_STEP) vv[i] = vsrc[v];
            const int d = rp[i + 1] - rp[i];
            if (d > 0) {
                cnt++;
                esum += d;
            }
        }
    }
    const uint32_t nnew = __popc(bits);

    // block-wide exclusive scan of (queued count, edge count)
    const uint64_t mine = (uint64_t(cnt) << 31) | esum;
    const uint64_t incl = wave_incl_scan(mine, lane);
    const uint32_t wnew = wave_sum32(nnew);
    __shared__ uint32_t s_new[WG / WAVE];
    if (lane == WAVE - 1) s_wsum[wid] = incl;
    if (lane == 0) s_new[wid] = wnew;
    __syncthreads();
    uint64_t wpre = 0, total = 0;
    uint32_t bnew = 0;
#pragma unroll
    for (int w = 0; w < WG / WAVE; ++w) {
        if (w < wid) wpre += s_wsum[w];
        total += s_wsum[w];
        bnew += s_new[w];
    }
    if (wid == 0) {
        const uint64_t p = lookback_exclusive(status + par * nblk, bid, total, lane);
        if (lane == 0) s_prefix = p;
    }
    __syncthreads();
    const uint64_t base = s_prefix + wpre + incl - mine;
    int pos = int(base >> 31);
    int off = int(base & 0x7fffffffu);
    if (cnt) {
#pragma unroll
        for (int i = 0; i < VPT; ++i) {
            if (!((bits >> i) & 1u)) continue;
            const int d = rp[i + 1] - rp[i];
            if (d <= 0) continue;
            queue[pos] = v0 + i;
            qstart[pos] = rp[i];
            qoff[pos] = off;
            if (MODE != MODE_BFS) qval[pos] = vv[i];
            // tiles whose first edge falls inside this vertex's edge range
            for (int t = (off + TILE - 1) / TILE; t * TILE < off + d; ++t) tiles[t] = pos;
            pos++;
            off += d;
        }
    }
    if (tid == 0) {
        if (bnew) {
            atomicAdd(&C->acc.vertices, (unsigned long long)bnew);
            atomicMax(&C->acc.levels, level + 1);
        }
        if (bid == nblk - 1) {
            const uint64_t g = s_prefix + total;
            Ctl c;
            c.nF = int(g >> 31);
            c.E = int(g & 0x7fffffffu);
            c.nTiles = (c.E + TILE - 1) / TILE;
            c.pad = 0;
            *next = c;
            atomicAdd(&C->acc.edges, (unsigned long long)c.E);
        }
    }
}

template <int MODE>
__global__ __launch_bounds__(WG) void k_expand(int level, int m, const Control *C,
                                               const int32_t *__restrict__ qstart,
                                               const int32_t *__restrict__ qoff,
                                               const float *__restrict__ qval,
                                               const int32_t *__restrict__ tiles,
                                               const int32_t *__restrict__ col_ind,
                                               const float *__restrict__ val,
                                               const uint32_t *__restrict__ visited,
                                               uint32_t *dist_bits, uint8_t *flags) {
    __shared__ int32_t s_off[TILE + 1];
    __shared__ int32_t s_start[TILE + 1];
    __shared__ float s_val[MODE == MODE_BFS ? 1 : TILE + 1];
    const Ctl cur = C->ctl[(level - 1) & 1];
    const int tid = threadIdx.x;
    for (int t = blockIdx.x; t < cur.nTiles; t += gridDim.x) {
        const int e0 = t * TILE;
        const int f0 = tiles[t];
        const int f1 = (t + 1 < cur.nTiles) ? tiles[t + 1] : cur.nF - 1;
        const int cnt = min(f1 - f0 + 1, TILE + 1);  // every queued vertex has >= 1 edge
        for (int k = tid; k < cnt; k += WG) {
            s_off[k] = qoff[f0 + k] - e0;
            s_start[k] = qstart[f0 + k];
            if (MODE != MODE_BFS) s_val[k] = qval[f0 + k];
        }
        __syncthreads();
        const int elim = min(TILE, cur.E - e0);
        int lo = 0;
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
            const int el = j * WG + tid;
            if (el >= elim) break;
            int hi = cnt - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (s_off[mid] <= el) lo = mid;
                else hi = mid - 1;
            }
            const int p = s_start[lo] + (el - s_off[lo]);
            const int c = col_ind[p];
            if ((unsigned)c >= (unsigned)m) continue;  // malformed column id: ignored
            if (MODE == MODE_BFS) {
                if (!((visited[c >> 5] >> (c & 31)) & 1u)) flags[c] = 1;
            } else {
                const float nd = s_val[lo] + (val ? val[p] : 1.0f);
                const uint32_t nb = __float_as_uint(nd);
                if (nb < dist_bits[c]) {
                    __hip_atomic_fetch_min(&dist_bits[c], nb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    flags[c] = 1;
                }
            }
        }
        __syncthreads();
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

namespace {

template <int MODE>
void launch_round(apspie_ctx *ctx, const Bufs &b, const apspie_csr *g, int level, int32_t *depth,
                  uint32_t *dist_bits, float *vsrc) {
    ctx->prof_begin(APSPIE_KCLASS_EXPAND);
    k_expand<MODE><<<b.expand_grid, WG, 0, ctx->stream>>>(level, g->m, b.C, b.qstart, b.qoff, b.qval, b.tiles,
                                                           g->col_ind, g->val, b.visited, dist_bits,
                                                           b.flags);
    ctx->prof_end();
    ctx->prof_begin(APSPIE_KCLASS_COMPACT);
    k_compact<MODE><<<b.nblk, WG, 0, ctx->stream>>>(level, g->m, b.nblk, 0, -1, g->row_ptr, b.flags,
                                                     b.visited, depth, vsrc, b.queue, b.qstart, b.qoff,
                                                     b.qval, b.tiles, b.status, b.C);
    ctx->prof_end();
}

template <int MODE>
void run(apspie_ctx *ctx, const apspie_csr *g, int32_t src, int32_t *depth, float *dist,
         apspie_stats *stats) {
    const Bufs b = get_bufs(ctx, g->m, g->nnz);
    uint32_t *dist_bits = (uint32_t *)dist;
    const unsigned init_grid = std::max(1u, std::min(ceil_div(int64_t(b.nblk) * CHUNK, WG), 4096u));
    APSPIE_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    ctx->prof_begin(APSPIE_KCLASS_OTHER);
    k_init<<<init_grid, WG, 0, ctx->stream>>>(MODE, g->m, b.nblk, depth, dist_bits, nullptr, b.flags,
                                              MODE == MODE_BFS ? b.visited : nullptr, b.status, b.C);
    ctx->prof_end();
    ctx->prof_begin(APSPIE_KCLASS_COMPACT);
    k_compact<MODE><<<b.nblk, WG, 0, ctx->stream>>>(0, g->m, b.nblk, 1, src, g->row_ptr, b.flags,
                                                     b.visited, depth, dist, b.queue, b.qstart, b.qoff,
                                                     b.qval, b.tiles, b.status, b.C);
    ctx->prof_end();
    int level = 1, chunk = 4;
    volatile int32_t *host_nf = (volatile int32_t *)ctx->pinned;
    for (;;) {
        for (int i = 0; i < chunk; ++i, ++level) launch_round<MODE>(ctx, b, g, level, depth, dist_bits, dist);
        APSPIE_HIP(hipGetLastError());
        APSPIE_HIP(hipMemcpyAsync((void *)host_nf, &b.C->ctl[(level - 1) & 1].nF, 4, hipMemcpyDeviceToHost,
                                  ctx->stream));
        APSPIE_HIP(hipStreamSynchronize(ctx->stream));
        if (*host_nf == 0) break;
        APSPIE_CHECK(level <= g->m + 2, APSPIE_ENOCONV, "traversal did not converge");
        chunk = std::min(chunk * 2, 32);
    }
    APSPIE_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    Acc acc;
    APSPIE_HIP(hipMemcpyAsync(&acc, &b.C->acc, sizeof(Acc), hipMemcpyDeviceToHost, ctx->stream));
    APSPIE_HIP(hipStreamSynchronize(ctx->stream));
    ctx->prof_collect();
    if (stats) {
        float ms = 0.f;
        APSPIE_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        memset(stats, 0, sizeof(*stats));
        stats->edges = (int64_t)acc.edges;
        stats->vertices = (int64_t)acc.vertices;
        stats->iterations = acc.levels;
        stats->kernel_ms = ms;
        if (MODE == MODE_BFS) {
            stats->bytes_alg = 8.0 * stats->edges + 20.0 * stats->vertices + 4.0 * g->m;
        } else {
            // improved (vertex, round) pairs: every frontier entry but the source
            stats->improved = stats->vertices > 0 ? stats->vertices - 1 : 0;
            stats->bytes_alg = 12.0 * stats->edges + 4.0 * stats->improved + 20.0 * stats->vertices +
                               8.0 * g->m;
        }
    }
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
        const unsigned init_grid = std::max(1u, std::min(ceil_div(int64_t(b.nblk) * CHUNK, WG), 4096u));
        APSPIE_HIP(hipEventRecord(ctx->ev0, ctx->stream));
        k_init<<<init_grid, WG, 0, ctx->stream>>>(MODE_STEP, g->m, b.nblk, nullptr, (uint32_t *)d_y, d_x,
                                                  b.flags, nullptr, b.status, b.C);
        k_compact<MODE_STEP><<<b.nblk, WG, 0, ctx->stream>>>(0, g->m, b.nblk, 1, -1, g->row_ptr, b.flags,
                                                              b.visited, nullptr, (float *)d_x, b.queue,
                                                              b.qstart, b.qoff, b.qval, b.tiles, b.status,
                                                              b.C);
        k_expand<MODE_STEP><<<b.expand_grid, WG, 0, ctx->stream>>>(1, g->m, b.C, b.qstart, b.qoff, b.qval, b.tiles,
                                                                    g->col_ind, g->val, nullptr,
                                                                    (uint32_t *)d_y, b.flags);
        k_clear_flags<<<std::min(b.nblk, 1024), WG, 0, ctx->stream>>>(b.nblk, b.flags);
        APSPIE_HIP(hipGetLastError());
        APSPIE_HIP(hipEventRecord(ctx->ev1, ctx->stream));
        Acc acc;
        APSPIE_HIP(hipMemcpyAsync(&acc, &b.C->acc, sizeof(Acc), hipMemcpyDeviceToHost, ctx->stream));
        APSPIE_HIP(hipStreamSynchronize(ctx->stream));
        if (stats) {
            float ms = 0.f;
            APSPIE_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
            memset(stats, 0, sizeof(*stats));
            stats->edges = (int64_t)acc.edges;
            stats->vertices = (int64_t)acc.vertices;
            stats->iterations = 1;
            stats->kernel_ms = ms;
            stats->bytes_alg = 12.0 * stats->edges + 20.0 * stats->vertices + 8.0 * g->m;
        }
    });
}

}  // extern "C"
