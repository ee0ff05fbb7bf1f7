// frontier.hpp -- the two generic gfx950 kernels every frontier algorithm is
// built from (BFS, SSSP, MIS, Boolean SpMSpV):
//
//   k_compact<Pick>  one coalesced pass over the vertex range in id order.
//                    Pick decides which vertices enter the new frontier (and
//                    may update per-vertex state on the way).  Entries with
//                    out-degree > 0 are queued with (vertex, row start, edge
//                    offset, 32-bit payload) through a single-pass decoupled
//                    look-back scan, and the tile -> first-entry table for the
//                    expansion is written.
//   k_expand<Op>     the queued entries' edges, cut into equal tiles of TILE
//                    edges; each tile finds its entries from the table, an
//                    inclusive max-scan over entry start marks gives every edge
//                    slot its owner, and Op runs on every edge in two phases
//                    (load, then apply) so a thread issues all of its loads
//                    before consuming any.
//
// Control state (frontier sizes, dynamic workgroup counters) lives in device
// memory; a compaction at round L writes ctl[L & 1], the expansion of round
// L + 1 reads it.  Compactions that follow an empty frontier exit at once
// (unless told to always run), so the host can launch rounds speculatively.
#pragma once

#include <type_traits>

#include "device.hpp"
#include "traverse.hpp"

namespace apspie {
namespace trav {

using namespace dev;

struct Queue {
    int32_t *vert;    // vertex id
    int32_t *start;   // row_ptr[v]
    int32_t *off;     // exclusive offset of the entry's first edge
    uint32_t *val;    // payload (distance bits, priority, ...)
    int32_t *tiles;   // first entry of each edge tile
};

// host helpers shared by the frontier algorithms (traverse.hip)
Queue queue_of(const Bufs &b);
unsigned init_grid(const Bufs &b, int m);
void fill_stats(apspie_ctx *ctx, const Bufs &b, int mode, int m, apspie_stats *stats);

// Compaction.  Wave w of workgroup b owns vertices [b*CHUNK + w*SEGV, +SEGV)
// and walks them in ROUNDS rounds of 64 consecutive vertices; in-wave ranks
// come from ballots (mbcnt) and wave scans, the workgroup offset from the
// look-back.
//
// Pick interface, either one phase
//   bool pick(int v, uint32_t &payload)       v may be >= m (padding lanes)
// or two (all ROUNDS loads are issued before any is used)
//   typedef Pre;  Pre load(int v);  bool pick(int v, const Pre &, uint32_t &payload)
// plus
//   void on_new(int v)                        for every picked vertex
// and optionally
//   void on_balls(uint64_t ball, int vbase, int lane, int rounds)
//        lane r < rounds holds the ballot of picked lanes of vertices
//        vbase + 64r .. +63 (visited / frontier bitmaps, one word pair each)
//   int direction(const Ctl &c, unsigned long long edges_so_far, int prev_pull)
//                                             1: run the next round as a pull
//                                             step (written into Ctl::pull)
//   uint32_t defer_key(int v, const Pre &)    a changed vertex left out of the
//                                             frontier to wait (its distance
//                                             bits; all ones: not waiting)
//   uint32_t raised(uint32_t dmin)            the threshold that admits the
//                                             smallest waiting distance
// A negative level_arg means "device-driven": the round number is read from
// C->cur_level and the compaction advances it when it completes, so a
// captured sequence of launches can be replayed unchanged (hipGraph).
template <class P, class = void>
struct pick_phases {
    struct Pre {};
    static __device__ __forceinline__ Pre load(const P &, int) { return Pre{}; }
    static __device__ __forceinline__ bool pick(const P &p, int v, const Pre &, uint32_t &payload) {
        return p.pick(v, payload);
    }
};
template <class P>
struct pick_phases<P, std::void_t<typename P::Pre>> {
    using Pre = typename P::Pre;
    static __device__ __forceinline__ Pre load(const P &p, int v) { return p.load(v); }
    static __device__ __forceinline__ bool pick(const P &p, int v, const Pre &x, uint32_t &payload) {
        return p.pick(v, x, payload);
    }
};
template <class P, class = void>
struct has_on_balls : std::false_type {};
template <class P>
struct has_on_balls<P, std::void_t<decltype(&P::on_balls)>> : std::true_type {};
template <class P, class = void>
struct has_direction : std::false_type {};
template <class P>
struct has_direction<P, std::void_t<decltype(&P::direction)>> : std::true_type {};
template <class P, class = void>
struct has_defer : std::false_type {};
template <class P>
struct has_defer<P, std::void_t<decltype(&P::defer_key)>> : std::true_type {};

template <class Pick>
__global__ __launch_bounds__(WGC) void k_compact(Pick pk, int level_arg, int m, int nblk, int always,
                                                 const int32_t *__restrict__ row_ptr, Queue q,
                                                 uint64_t *status, Control *C) {
    constexpr int NW = WGC / WAVE, SEGV = CHUNK / NW, ROUNDS = SEGV / WAVE;
    static_assert(ROUNDS <= 32, "round bits must fit a 32-bit mask");
    __shared__ int s_bid;
    __shared__ uint64_t s_wsum[NW];
    __shared__ uint32_t s_new[NW];
    __shared__ uint32_t s_def[NW], s_dmin[NW];
    __shared__ uint64_t s_prefix;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
    // every block reads the round before it takes a dynamic id, so the last
    // id's block may advance cur_level once it has finished
    // (an atomic load keeps level_arg in a register: a select between the
    // two addresses of a volatile load would put it in scratch memory)
    const int level = level_arg >= 0 ? level_arg
                                     : __hip_atomic_load(&C->cur_level, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int par = level & 1;
    Ctl *next = &C->ctl[par];
    if (!always && C->ctl[par ^ 1].nF == 0 && C->ctl[par ^ 1].pull != DIR_REPICK) {
        if (blockIdx.x == 0 && tid == 0) *next = Ctl{0, 0, 0, 0};
        return;
    }
    if (tid == 0) s_bid = (int)atomicAdd(&C->blk_counter[par], 1u);
    __syncthreads();
    const int bid = s_bid;
    // state for the launch after this one (its previous user has finished)
    if (tid == 0) status[(par ^ 1) * nblk + bid] = 0;
    if (bid == 0 && tid < SPREAD) {
        if (tid == 0) {
            C->blk_counter[par ^ 1] = 0;
            C->ndefer[par ^ 1] = 0;
            C->dmin[par ^ 1] = ~0u;
        }
        C->nnew[par ^ 1][tid] = 0;
    }

    const int vbase = bid * CHUNK + wid * SEGV;
    using PP = pick_phases<Pick>;
    typename PP::Pre pre[ROUNDS];
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) pre[r] = PP::load(pk, vbase + r * WAVE + lane);
    uint32_t newbits = 0;
    uint64_t myball = 0;
    uint32_t vv[ROUNDS];
    uint32_t ndef = 0, dmin = ~0u;
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
        const int v = vbase + r * WAVE + lane;
        uint32_t payload = 0;
        if constexpr (has_defer<Pick>::value) {
            const uint32_t key = pk.defer_key(v, pre[r]);
            ndef += key != ~0u;
            dmin = min(dmin, key);
        }
        const bool isnew = PP::pick(pk, v, pre[r], payload);
        const uint64_t ball = __ballot(isnew);
        if (lane == r) myball = ball;
        vv[r] = payload;
        if (isnew) {
            newbits |= 1u << r;
            pk.on_new(v);
        }
    }
    if constexpr (has_on_balls<Pick>::value) pk.on_balls(myball, vbase, lane, ROUNDS);
    // row extents of the picked vertices (picked vertices are < m, so
    // row_ptr[v + 1] exists); every load is issued before any is used
    int32_t rs[ROUNDS], dg[ROUNDS];  // dg holds the row end until it becomes the degree
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
        rs[r] = 0;
        dg[r] = 0;
        if ((newbits >> r) & 1u) {
            const int v = vbase + r * WAVE + lane;
            rs[r] = row_ptr[v];
            dg[r] = row_ptr[v + 1];
        }
    }
    uint32_t qb = 0, lane_cnt = 0, lane_es = 0;
    const uint32_t lane_new = __popc(newbits);
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
        const int d = dg[r] - rs[r];
        dg[r] = d;
        if (d > 0) {
            qb |= 1u << r;
            lane_cnt++;
            lane_es += d;
        }
    }
    const uint32_t wcnt = wave_total32(lane_cnt), wes = wave_total32(lane_es), wnew = wave_total32(lane_new);
    if constexpr (has_defer<Pick>::value) {
        ndef = wave_total32(ndef);
        dmin = wave_min_u32(dmin);
    }
    if (lane == 0) {
        s_wsum[wid] = (uint64_t(wcnt) << 31) | wes;
        s_new[wid] = wnew;
        s_def[wid] = ndef;
        s_dmin[wid] = dmin;
    }
    __syncthreads();
    uint64_t wpre = 0, total = 0;
    uint32_t bnew = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        if (w < wid) wpre += s_wsum[w];
        total += s_wsum[w];
        bnew += s_new[w];
    }
    if (wid == 0) {
        if constexpr (has_defer<Pick>::value) {
            // waiting vertices are counted before this workgroup publishes its
            // prefix, so the last workgroup sees every count
            if (lane == 0) {
                uint32_t bdef = 0, bmin = ~0u;
#pragma unroll
                for (int w = 0; w < NW; ++w) {
                    bdef += s_def[w];
                    bmin = min(bmin, s_dmin[w]);
                }
                if (bdef) {
                    atomicAdd(&C->ndefer[par], bdef);
                    atomicMin(&C->dmin[par], bmin);
                }
            }
            __threadfence();
        }
        const uint64_t p = lookback_exclusive(status + par * nblk, bid, total, lane);
        if (lane == 0) s_prefix = p;
    }
    __syncthreads();
    const uint64_t base = s_prefix + wpre;
    int pos = int(base >> 31);
    int off = int(base & 0x7fffffffu);
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
        const bool qd = (qb >> r) & 1u;
        const uint64_t qball = __ballot(qd);
        if (!qball) continue;
        const int rank =
            __builtin_amdgcn_mbcnt_hi(uint32_t(qball >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(qball), 0));
        const uint32_t d = qd ? uint32_t(dg[r]) : 0u;
        const uint32_t dincl = wave_incl_sum_dpp(d);
        if (qd) {
            const int p = pos + rank;
            const int o = off + int(dincl - d);
            q.vert[p] = vbase + r * WAVE + lane;
            q.start[p] = rs[r];
            q.off[p] = o;
            q.val[p] = vv[r];
            // tiles whose first edge falls inside this entry's edge range
            for (int t = (o + TILE - 1) / TILE; t * TILE < o + int(d); ++t) q.tiles[t] = p;
        }
        pos += __popcll(qball);
        off += wave_last(int(dincl));
    }
    if (tid == 0) {
        if (bnew) {
            atomicAdd(&C->nnew[par][bid % SPREAD], bnew);
            atomicAdd(&C->acc.vertices[bid % SPREAD], (unsigned long long)bnew);
            // levels only grows: most workgroups see it already set
            if (*(volatile int *)&C->acc.levels < level + 1) atomicMax(&C->acc.levels, level + 1);
        }
        if (bid == nblk - 1) {
            const uint64_t g = s_prefix + total;
            Ctl c;
            c.nF = int(g >> 31);
            c.E = int(g & 0x7fffffffu);
            c.nTiles = (c.E + TILE - 1) / TILE;
            const unsigned long long before = atomicAdd(&C->acc.edges, (unsigned long long)c.E);
            c.pull = 0;
            if constexpr (has_direction<Pick>::value) c.pull = pk.direction(c, before + c.E, C->ctl[par ^ 1].pull);
            if constexpr (has_defer<Pick>::value) {
                __threadfence();
                // nothing left below the threshold: raise it past the smallest
                // waiting distance and pick again next round
                if (c.nF == 0 && __hip_atomic_load(&C->ndefer[par], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                    C->thr = pk.raised(__hip_atomic_load(&C->dmin[par], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
                    c.pull = DIR_REPICK;
                }
            }
            *next = c;
            if (level < LOG_ROUNDS) C->log[level] = c;
            if (level_arg < 0) C->cur_level = level + 1;
        }
    }
}

// Round 0 of a single-source traversal: what k_compact<Pick> computes when
// only the seed is picked (the pick and its side effects, its bitmap words,
// the queue entry, the tile table, the counters and the round descriptor
// with its direction), without a pass over the vertex range.  State that
// pass rewrites for every other vertex (frontier bitmap words, pull
// snapshots) is cleared by the traversal's initialisation, and the control
// block starts zeroed.  One workgroup.
template <class Pick>
__global__ __launch_bounds__(WG) void k_seed(Pick pk, int seed, const int32_t *__restrict__ row_ptr, Queue q,
                                             Control *C) {
    const int start = row_ptr[seed], deg = row_ptr[seed + 1] - start;
    const int ntiles = (deg + TILE - 1) / TILE;
    for (int t = threadIdx.x; t < ntiles; t += WG) q.tiles[t] = 0;
    if (threadIdx.x != 0) return;
    using PP = pick_phases<Pick>;
    uint32_t payload = 0;
    const typename PP::Pre pre = PP::load(pk, seed);
    PP::pick(pk, seed, pre, payload);  // the seed is always picked
    pk.on_new(seed);
    if constexpr (has_on_balls<Pick>::value) pk.on_balls(1ull << (seed & (WAVE - 1)), seed & ~(WAVE - 1), 0, 1);
    if (deg > 0) {
        q.vert[0] = seed;
        q.start[0] = start;
        q.off[0] = 0;
        q.val[0] = payload;
    }
    C->nnew[0][0] = 1;
    C->acc.vertices[0] = 1;
    C->acc.levels = 1;
    C->acc.edges = (unsigned long long)deg;
    Ctl c;
    c.nF = deg > 0 ? 1 : 0;
    c.E = deg;
    c.nTiles = ntiles;
    c.pull = 0;
    if constexpr (has_direction<Pick>::value) c.pull = pk.direction(c, (unsigned long long)deg, C->ctl[1].pull);
    C->ctl[0] = c;
    C->log[0] = c;
}

// Inclusive max-scan of x over the TILE slots of a workgroup in slot order
// (slot = wid*SEG + r*64 + lane); s_wmax holds one int per wave.  Contains a
// workgroup barrier.
__device__ __forceinline__ void slot_max_scan(int (&x)[IPT], int32_t *s_wmax) {
    constexpr int NW = WG / WAVE;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
    int mx = 0;
#pragma unroll
    for (int r = 0; r < IPT; ++r) mx = max(mx, x[r]);
    mx = wave_last(wave_incl_max_nonneg(mx));
    if (lane == 0) s_wmax[wid] = mx;
    __syncthreads();
    int carry = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w)
        if (w < wid) carry = max(carry, s_wmax[w]);
#pragma unroll
    for (int r = 0; r < IPT; ++r) {
        const int xo = wave_incl_max_nonneg(max(x[r], carry));
        x[r] = xo;
        carry = wave_last(xo);
    }
}

// Expansion over the entries queued by the previous compaction.
//
// Op interface:
//   typedef Pre;                                   per-edge loaded state
//   Pre load(int c, int p, uint32_t payload, int v)   issue loads for edge
//                                                  (v -> c) at col_ind[p]
//   void apply(int c, int p, uint32_t payload, int v, const Pre &)
//   static constexpr bool kVert                    needs the entry's vertex id
//   static constexpr bool kVal                     needs the entry's payload
// Columns outside [0, m) are skipped.
// Stages tile t of the queued entries' edges into LDS (entry start marks,
// col_ind base, optional payload and vertex id) and gives this thread the
// owner entry of its IPT edge slots (slot = wid*SEG + r*64 + lane) through an
// inclusive max-scan.  Returns the number of valid slots in the tile.
template <bool KVAL, bool KVERT>
__device__ __forceinline__ int tile_owners(int t, const Ctl &cur, const Queue &q, int32_t *s_own, int32_t *s_base,
                                           uint32_t *s_val, int32_t *s_vert, int32_t *s_wmax, int (&own)[IPT]) {
    constexpr int NW = WG / WAVE, SEG = TILE / NW;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
    const int e0 = t * TILE;
    const int f0 = q.tiles[t];
    const int f1 = (t + 1 < cur.nTiles) ? q.tiles[t + 1] : cur.nF - 1;
    const int cnt = min(f1 - f0 + 1, TILE + 1);  // every queued entry has >= 1 edge
    for (int i = tid; i < TILE; i += WG) s_own[i] = 0;
    __syncthreads();
    for (int k = tid; k < cnt; k += WG) {
        const int o = q.off[f0 + k] - e0;
        s_base[k] = q.start[f0 + k] - o;
        if (KVAL) s_val[k] = q.val[f0 + k];
        if (KVERT) s_vert[k] = q.vert[f0 + k];
        if (o > 0 && o < TILE) s_own[o] = k;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < IPT; ++r) own[r] = s_own[wid * SEG + r * WAVE + lane];
    slot_max_scan(own, s_wmax);
    return min(TILE, cur.E - e0);
}

// Slots whose loads an op issues before applying any (Op::kBatch when the
// op declares it, else all IPT); ops with a large Pre lower it to bound
// register use.
template <class T, class = void>
struct op_batch {
    static constexpr int value = IPT;
};
template <class T>
struct op_batch<T, std::void_t<decltype(T::kBatch)>> {
    static constexpr int value = T::kBatch;
};

// The tiles t = bid, bid + nblk, ... of the frontier `cur` (block-uniform
// arguments; every thread of the workgroup calls it).
template <class Op>
__device__ __forceinline__ void expand_tiles(const Op &op, int m, const Ctl &cur, const Queue &q,
                                             const int32_t *__restrict__ col_ind, int bid, int nblk) {
    constexpr int NW = WG / WAVE, SEG = TILE / NW;
    __shared__ int32_t s_own[TILE];
    __shared__ int32_t s_base[TILE + 1];  // start - (off - e0): col_ind index of slot 0
    __shared__ uint32_t s_val[Op::kVal ? TILE + 1 : 1];
    __shared__ int32_t s_vert[Op::kVert ? TILE + 1 : 1];
    __shared__ int32_t s_wmax[NW];
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
    for (int t = bid; t < cur.nTiles; t += nblk) {
        int own[IPT];
        const int elim = tile_owners<Op::kVal, Op::kVert>(t, cur, q, s_own, s_base, s_val, s_vert, s_wmax, own);
        int pos[IPT], col[IPT];
#pragma unroll
        for (int r = 0; r < IPT; ++r) {
            const int s = wid * SEG + r * WAVE + lane;
            pos[r] = s < elim ? s_base[own[r]] + s : -1;
        }
#pragma unroll
        for (int r = 0; r < IPT; ++r) col[r] = pos[r] >= 0 ? col_ind[pos[r]] : -1;
        constexpr int B = op_batch<Op>::value;
        static_assert(IPT % B == 0, "batch divides IPT");
#pragma unroll
        for (int r0 = 0; r0 < IPT; r0 += B) {
            typename Op::Pre pre[B];
#pragma unroll
            for (int r = 0; r < B; ++r)
                if ((unsigned)col[r0 + r] < (unsigned)m)
                    pre[r] = op.load(col[r0 + r], pos[r0 + r], Op::kVal ? s_val[own[r0 + r]] : 0u,
                                     Op::kVert ? s_vert[own[r0 + r]] : 0);
#pragma unroll
            for (int r = 0; r < B; ++r)
                if ((unsigned)col[r0 + r] < (unsigned)m)
                    op.apply(col[r0 + r], pos[r0 + r], Op::kVal ? s_val[own[r0 + r]] : 0u,
                             Op::kVert ? s_vert[own[r0 + r]] : 0, pre[r]);
        }
        __syncthreads();
    }
}

template <class Op>
__global__ __launch_bounds__(WG) void k_expand(Op op, int level_arg, int m, const Control *C, Queue q,
                                               const int32_t *__restrict__ col_ind) {
    const int level = level_arg >= 0 ? level_arg : C->cur_level;  // round of the next compaction
    const Ctl cur = C->ctl[(level - 1) & 1];
    if (cur.pull) return;  // this round runs as a pull step
    expand_tiles(op, m, cur, q, col_ind, blockIdx.x, gridDim.x);
}

}  // namespace trav
}  // namespace apspie
