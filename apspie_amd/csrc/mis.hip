// mis.hip -- maximal independent set (Luby with fixed priorities) and the
// one-step Boolean SpMSpV, on the generic compaction / expansion kernels.
//
// MIS round (all state in one byte per vertex: undecided / in / out):
//   compact  undecided vertices -> queue U (payload: priority)
//   expand   per edge (v -> c), v in U, c > v (each edge of the symmetric
//            graph once): when c is undecided, the endpoint of smaller key
//            is blocked, key = (priority, -id); one 8-byte load a probe, the
//            key of each undecided vertex (0 once decided)
//   compact  undecided, unblocked vertices join (state = in) -> queue J
//   expand   per edge (v -> c), v in J: undecided c leaves (state = out)
// Every round the undecided vertex of largest key joins, so the loop ends;
// its fixed point is the lexicographically-first MIS in descending key order,
// i.e. the greedy MIS of test/mis/main.cu:26-88 visited by key instead of by
// index (Blelloch, Fineman & Shun, SPAA 2012).  The reference's threshold
// schedule (src/spmspvMis.cuh:232-239, `delta`) only delays vertices and
// reaches the same fixed point, so it is not needed.
//
// Replaces spmspvMis (src/spmspvMis.cuh:150-344: buildVector, Scan,
// IntervalGather/Expand, SortPairs, gather, ReduceByKey(max), scatter,
// updateMis, updateNeighbor) and spmspvCsr (src/spmspvMis.cuh:81-147).
#include <algorithm>
#include <cstring>

#include "frontier.hpp"
#include "internal.hpp"

namespace apspie {
namespace trav {
namespace {

constexpr uint8_t UND = 0, IN = 1, OUT = 2;

struct PickUndecided {
    const uint8_t *state;
    const uint32_t *prio;
    int m;
    __device__ bool pick(int v, uint32_t &payload) const {
        if (v >= m) return false;
        payload = prio[v];
        return state[v] == UND;
    }
    __device__ void on_new(int) const {}
};

// key(v) = priority << 32 | ~v: a larger key wins, ties to the smaller id
__device__ __forceinline__ uint64_t mis_key(uint32_t p, int v) { return (uint64_t(p) << 32) | uint32_t(~uint32_t(v)); }

struct OpBlock {
    static constexpr bool kVert = true, kVal = true;
    using Pre = uint64_t;
    const uint64_t *eff;  // key of each undecided vertex, 0 once decided
    uint8_t *blocked;
    __device__ Pre load(int c, int, uint32_t, int v) const { return c > v ? eff[c] : 0ull; }
    __device__ void apply(int c, int, uint32_t pv, int v, Pre ec) const {
        if (!ec) return;  // c <= v (the other orientation does it), or c decided
        if (ec > mis_key(pv, v)) blocked[v] = 1;
        else blocked[c] = 1;
    }
};

struct PickJoin {
    uint8_t *state, *blocked;
    uint64_t *eff;
    int m;
    __device__ bool pick(int v, uint32_t &) const {
        if (v >= m || state[v] != UND) return false;
        if (blocked[v]) {
            blocked[v] = 0;
            return false;
        }
        state[v] = IN;
        eff[v] = 0;
        return true;
    }
    __device__ void on_new(int) const {}
};

struct OpRemove {
    static constexpr bool kVert = false, kVal = false;
    using Pre = uint8_t;
    uint8_t *state;
    uint64_t *eff;
    __device__ Pre load(int c, int, uint32_t, int) const { return state[c]; }
    __device__ void apply(int c, int, uint32_t, int, Pre s) const {
        if (s == UND) {
            state[c] = OUT;
            eff[c] = 0;
        }
    }
};

struct PickNonzero {
    const int32_t *in;
    int m;
    __device__ bool pick(int v, uint32_t &) const { return v < m && in[v] != 0; }
    __device__ void on_new(int) const {}
};

struct OpSetOne {
    static constexpr bool kVert = false, kVal = false;
    using Pre = int;
    int32_t *out;
    __device__ Pre load(int, int, uint32_t, int) const { return 0; }
    __device__ void apply(int c, int, uint32_t, int, Pre) const { out[c] = 1; }
};

__global__ __launch_bounds__(WG) void k_mis_init(int m, int nblk, uint8_t *state, uint8_t *blocked,
                                                 const uint32_t *prio, uint64_t *eff, int32_t *out,
                                                 uint64_t *status, Control *C) {
    const int64_t gtid = int64_t(blockIdx.x) * WG + threadIdx.x;
    const int64_t stride = int64_t(gridDim.x) * WG;
    const int64_t mpad = int64_t(nblk) * CHUNK;
    for (int64_t i = gtid; i < mpad; i += stride) {
        if (state) state[i] = UND;
        if (blocked) blocked[i] = 0;
        if (out && i < m) out[i] = 0;
        if (eff && i < m) eff[i] = mis_key(prio[i], int(i));
    }
    for (int64_t i = gtid; i < 2 * int64_t(nblk); i += stride) status[i] = 0;
    control_clear(C, gtid, stride, ~0u);
}

__global__ __launch_bounds__(WG) void k_mis_finish(int m, const uint8_t *state, int32_t *out) {
    for (int64_t i = int64_t(blockIdx.x) * WG + threadIdx.x; i < m; i += int64_t(gridDim.x) * WG)
        out[i] = state[i] == IN ? 1 : 0;
}

template <class Pick>
void compact(apspie_ctx *ctx, const Bufs &b, const apspie_csr *g, const Pick &pk, int level) {
    ctx->prof_begin(APSPIE_KCLASS_COMPACT);
    k_compact<Pick><<<b.nblk, WGC, 0, ctx->stream>>>(pk, level, g->m, b.nblk, 1, g->row_ptr, queue_of(b),
                                                      b.C);
    ctx->prof_end();
}

template <class Op>
void expand(apspie_ctx *ctx, const Bufs &b, const apspie_csr *g, const Op &op, int level) {
    ctx->prof_begin(APSPIE_KCLASS_EXPAND);
    k_expand<Op><<<b.expand_grid, WG, 0, ctx->stream>>>(op, level, g->m, b.C, queue_of(b), g->col_ind);
    ctx->prof_end();
}

}  // namespace
}  // namespace trav
}  // namespace apspie

using namespace apspie;
using namespace apspie::trav;

extern "C" {

int apspie_mis(apspie_ctx *ctx, const apspie_csr *g, const uint32_t *d_prio, int32_t *d_mis,
               apspie_stats *stats) {
    if (!ctx) return APSPIE_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    return guarded(ctx, [&] {
        validate_csr(g, false);
        APSPIE_CHECK(g->m == g->n, APSPIE_EINVAL, "MIS needs a square (symmetric) adjacency matrix");
        APSPIE_CHECK(d_prio && d_mis, APSPIE_EINVAL, "null priority or output");
        const Bufs b = get_bufs(ctx, g->m, g->nnz);
        uint8_t *state = (uint8_t *)ctx->slot(S_AUX0, size_t(b.nblk) * CHUNK);
        uint64_t *eff = (uint64_t *)ctx->slot(S_AUX1, size_t(std::max(g->m, 1)) * 8);
        uint8_t *blocked = b.flags;
        APSPIE_HIP(hipEventRecord(ctx->ev0, ctx->stream));
        ctx->prof_begin(APSPIE_KCLASS_OTHER);
        k_mis_init<<<init_grid(b, g->m), WG, 0, ctx->stream>>>(g->m, b.nblk, state, blocked, d_prio, eff, nullptr,
                                                               b.status, b.C);
        ctx->prof_end();
        static_assert(sizeof(Control::nnew[0]) <= 4096, "the joined counts fit the pinned staging");
        volatile Padded<uint32_t> *host = (volatile Padded<uint32_t> *)ctx->pinned;
        // rounds between host checks: 6 (RMAT s16-s21 converge in 5-7), then
        // 2, 4, 8, 16 -- a round after convergence is four launches of
        // nothing, so the chunk that passes convergence is kept short
        int level = 0, rounds = 0, chunk = 6, next_chunk = 2;
        for (;;) {
            for (int i = 0; i < chunk; ++i, ++rounds) {
                compact(ctx, b, g, PickUndecided{state, d_prio, g->m}, level++);
                expand(ctx, b, g, OpBlock{eff, blocked}, level);
                compact(ctx, b, g, PickJoin{state, blocked, eff, g->m}, level++);
                expand(ctx, b, g, OpRemove{state, eff}, level);
            }
            APSPIE_HIP(hipGetLastError());
            // vertices that joined in the last round; none => nothing undecided
            APSPIE_HIP(hipMemcpyAsync((void *)host, &b.C->nnew[(level - 1) & 1][0], sizeof(b.C->nnew[0]),
                                      hipMemcpyDeviceToHost, ctx->stream));
            APSPIE_HIP(hipStreamSynchronize(ctx->stream));
            uint32_t joined = 0;
            for (int i = 0; i < SPREAD; ++i) joined += host[i].v;
            if (joined == 0) break;
            APSPIE_CHECK(rounds <= g->m + 1, APSPIE_ENOCONV, "MIS did not converge");
            chunk = next_chunk;
            next_chunk = std::min(next_chunk * 2, 16);
        }
        k_mis_finish<<<init_grid(b, g->m), WG, 0, ctx->stream>>>(g->m, state, d_mis);
        APSPIE_HIP(hipGetLastError());
        APSPIE_HIP(hipEventRecord(ctx->ev1, ctx->stream));
        fill_stats(ctx, b, MODE_BFS, g->m, stats);
        if (stats) {
            // two compactions per round; the last one with picks ends the count
            stats->iterations = (stats->iterations + 1) / 2;
            stats->bytes_alg = 0.0;
        }
    });
}

int apspie_spmspv_or(apspie_ctx *ctx, const apspie_csr *g, const int32_t *d_in, int32_t *d_out) {
    if (!ctx) return APSPIE_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    return guarded(ctx, [&] {
        validate_csr(g, false);
        APSPIE_CHECK(g->m == g->n, APSPIE_EINVAL, "SpMSpV here needs a square matrix");
        APSPIE_CHECK(d_in && d_out, APSPIE_EINVAL, "null vector");
        const Bufs b = get_bufs(ctx, g->m, g->nnz);
        APSPIE_HIP(hipEventRecord(ctx->ev0, ctx->stream));
        k_mis_init<<<init_grid(b, g->m), WG, 0, ctx->stream>>>(g->m, b.nblk, nullptr, nullptr, nullptr, nullptr,
                                                               d_out, b.status, b.C);
        compact(ctx, b, g, PickNonzero{d_in, g->m}, 0);
        expand(ctx, b, g, OpSetOne{d_out}, 1);
        APSPIE_HIP(hipGetLastError());
        APSPIE_HIP(hipEventRecord(ctx->ev1, ctx->stream));
        APSPIE_HIP(hipStreamSynchronize(ctx->stream));
        ctx->prof_collect();
    });
}

}  // extern "C"
