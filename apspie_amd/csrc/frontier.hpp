// frontier.hpp -- the two generic gfx950 kernels every frontier algorithm is
// built from (BFS, SSSP, MIS, Boolean SpMSpV):
//
//   k_compact<Pick>  one coalesced pass over the vertex range, a chunk of
//                    CHUNK vertices per workgroup.  Pick decides which
//                    vertices enter the new frontier (and may update
//                    per-vertex state on the way).  Entries with out-degree
//                    > 0 are queued with (vertex, row start, edge offset,
//                    32-bit payload) in their chunk group's region (one
//                    returning atomic add per chunk on the group's counter
//                    gives its place; nothing waits on another workgroup),
//                    and the tile -> first-entry table for the expansion is
//                    written.
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

#include <cstddef>
#include <type_traits>

#include "device.hpp"
#include "traverse.hpp"

namespace apspie {
namespace trav {

using namespace dev;

// Queue regions: group g's entries sit at [g * qcap, g * qcap + entries of
// g), its tile table at [g * tcap, ...); offsets and tile numbers inside a
// group are the group's own.
struct Queue {
    int32_t *vert;    // vertex id
    int32_t *start;   // row_ptr[v]
    int32_t *off;     // exclusive offset of the entry's first edge within its group
    uint32_t *val;    // payload (distance bits, priority, ...)
    int32_t *tiles;   // first entry of each edge tile of each group
    int32_t qcap, tcap;
};

// The frontier a round expands: the totals and, per group, its first tile in
// the round's tile numbering and its (entries << 32) | edges.
struct Round : Ctl {
    const int32_t *gtile;
    const Padded<unsigned long long> *gcnt;
};
__device__ __forceinline__ Round round_of(const Control *C, int level) {
    const int pp = (level - 1) & 1;
    Round r;
    static_cast<Ctl &>(r) = C->ctl[pp];
    r.gtile = C->gtile[pp];
    r.gcnt = C->gcnt[pp];
    return r;
}

// Clears the control block with every thread of an initialisation launch
// (thread gtid of stride, over 16-byte words), then the near/far threshold is
// thr0 (all ones: none) and the smallest waiting distances all ones.
__device__ __forceinline__ void control_clear(Control *C, int64_t gtid, int64_t stride, uint32_t thr0) {
    static_assert(sizeof(Control) % 16 == 0, "the control block is a whole number of 16-byte words");
    constexpr int64_t n = int64_t(sizeof(Control) / 16);
    constexpr int64_t ithr = int64_t(offsetof(Control, thr) / 4), idmin = int64_t(offsetof(Control, dmin) / 4);
    uint4 *w = (uint4 *)C;
    for (int64_t i = gtid; i < n; i += stride) {
        uint32_t x[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t k = 4 * i + j;
            x[j] = k == ithr ? thr0 : (k == idmin || k == idmin + 1) ? ~0u : 0u;
        }
        w[i] = make_uint4(x[0], x[1], x[2], x[3]);
    }
}

// host helpers shared by the frontier algorithms (traverse.hip)
Queue queue_of(const Bufs &b);
unsigned init_grid(const Bufs &b, int m);
// (host_acc: the totals already copied to the host; else they are copied here)
void fill_stats(apspie_ctx *ctx, const Bufs &b, int mode, int m, apspie_stats *stats, const Acc *host_acc = nullptr);

// Compaction.  Wave w of workgroup b owns vertices [b*CHUNK + w*SEGV, +SEGV)
// and walks them in ROUNDS rounds of 64 consecutive vertices; in-wave ranks
// come from ballots (mbcnt) and wave scans, the workgroup's place in its
// chunk group from one returning atomic add of its (entries, edges).
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
// The workgroup that finishes last (counted per group, then over the groups)
// writes the round's totals, tile numbering, direction and log entry.
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
// An op with weight(p) has its edge weight loaded by the expansion beside the
// column and stored into Pre::w after load(), which then fetches only the
// target's data: the weight load no longer waits for the column.
template <class P, class = void>
struct has_weight : std::false_type {};
template <class P>
struct has_weight<P, std::void_t<decltype(&P::weight)>> : std::true_type {};

template <class P, class = void>
struct has_mid : std::false_type {};
template <class P>
struct has_mid<P, std::void_t<decltype(&P::mid)>> : std::true_type {};
template <class P, class = void>
struct has_defer : std::false_type {};
template <class P>
struct has_defer<P, std::void_t<decltype(&P::defer_key)>> : std::true_type {};

template <class Pick>
__global__ __launch_bounds__(WGC) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_compact(Pick pk, int level_arg, int m, int nblk, int always,
                                                 const int32_t *__restrict__ row_ptr, Queue q, Control *C) {
    constexpr int NW = WGC / WAVE, SEGV = CHUNK / NW, ROUNDS = SEGV / WAVE;
    static_assert(ROUNDS <= 32, "round bits must fit a 32-bit mask");
    __shared__ uint64_t s_wsum[NW];
    __shared__ uint32_t s_new[NW];
    __shared__ uint32_t s_def[NW], s_dmin[NW];
    __shared__ uint64_t s_base;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
    // every workgroup reads the round before its completion is counted, so
    // the one that completes last may advance cur_level
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
    const int bid = blockIdx.x, grp = bid % NGROUP;
    // counters of the round after this one (their last readers, the
    // expansion of this round among them, have finished)
    if (bid == 0 && tid < NGROUP) {
        C->gcnt[par ^ 1][tid].v = 0;
        C->gdone[par ^ 1][tid].v = 0;
        C->nnew[par ^ 1][tid].v = 0;
        C->rfound[par ^ 1][tid].v = 0;
        if (tid == 0) {
            C->qoverflow[par ^ 1] = 0;
            C->alldone[par ^ 1] = 0;
            C->ndefer[par ^ 1] = 0;
            C->dmin[par ^ 1] = ~0u;
        }
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
    // degrees of the picked vertices (picked vertices are < m, so
    // row_ptr[v + 1] exists); every load is issued before any is used.  The
    // row extents are not kept past the workgroup's reduction (they are read
    // again, from cache, when the entries are written): fewer live registers
    // let two 1024-thread workgroups share a compute unit.
    uint32_t qb = 0, lane_cnt = 0, lane_es = 0;
    const uint32_t lane_new = __popc(newbits);
    {
        int32_t rs[ROUNDS], re[ROUNDS];
#pragma unroll
        for (int r = 0; r < ROUNDS; ++r) {
            rs[r] = 0;
            re[r] = 0;
            if ((newbits >> r) & 1u) {
                const int v = vbase + r * WAVE + lane;
                rs[r] = row_ptr[v];
                re[r] = row_ptr[v + 1];
            }
        }
#pragma unroll
        for (int r = 0; r < ROUNDS; ++r) {
            const int d = re[r] - rs[r];
            if (d > 0) {
                qb |= 1u << r;
                lane_cnt++;
                lane_es += d;
            }
        }
    }
    const uint32_t wcnt = wave_total32(lane_cnt), wes = wave_total32(lane_es), wnew = wave_total32(lane_new);
    if constexpr (has_defer<Pick>::value) {
        ndef = wave_total32(ndef);
        dmin = wave_min_u32(dmin);
    }
    if (lane == 0) {
        s_wsum[wid] = (uint64_t(wcnt) << 32) | wes;
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
    // every count of this chunk is a returning atomic whose value is awaited
    // before the chunk counts itself done, so the last workgroup sees them all
    uint32_t dep = 0;
    if (tid == 0) {
        if constexpr (has_defer<Pick>::value) {
            uint32_t bdef = 0, bmin = ~0u;
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                bdef += s_def[w];
                bmin = min(bmin, s_dmin[w]);
            }
            if (bdef) {
                dep ^= atomicAdd(&C->ndefer[par], bdef);
                dep ^= atomicMin(&C->dmin[par], bmin);
            }
        }
        if (bnew) {
            dep ^= atomicAdd(&C->nnew[par][bid % SPREAD].v, bnew);
            dep ^= uint32_t(atomicAdd(&C->acc.vertices[bid % SPREAD].v, (unsigned long long)bnew));
        }
        // this chunk's place in its group: the entries and edges before it
        // (a chunk that queues nothing needs none)
        s_base = total ? atomicAdd(&C->gcnt[par][grp].v, (unsigned long long)total) : 0ull;
    }
    __syncthreads();
    // count the chunk done (the return is awaited only after this wave's
    // writes, which hide the round trip)
    unsigned done = 0;
    if (tid == 0) {
        after_values(dep);
        done = atomicAdd(&C->gdone[par][grp].v, 1u) + 1u;
    }
    const uint64_t base = s_base + wpre;
    int pos = int(base >> 32);
    int off = int(base & 0xffffffffu);
    const size_t qo = size_t(grp) * q.qcap;
    int32_t *const tl = q.tiles + size_t(grp) * q.tcap;
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
        const bool qd = (qb >> r) & 1u;
        const uint64_t qball = __ballot(qd);
        if (!qball) continue;
        const int v = vbase + r * WAVE + lane;
        const int r0 = qd ? row_ptr[v] : 0;
        const int d = qd ? row_ptr[v + 1] - r0 : 0;
        const int rank =
            __builtin_amdgcn_mbcnt_hi(uint32_t(qball >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(qball), 0));
        const uint32_t dincl = wave_incl_sum_dpp(uint32_t(d));
        if (qd) {
            const int p = pos + rank;
            const int o = off + int(dincl) - d;
            q.vert[qo + p] = v;
            q.start[qo + p] = r0;
            q.off[qo + p] = o;
            q.val[qo + p] = vv[r];
            // tiles whose first edge falls inside this entry's edge range
            for (int t = (o + TILE - 1) / TILE; t * TILE < o + d; ++t) tl[t] = p;
        }
        pos += __popcll(qball);
        off += wave_last(int(dincl));
    }
    if (tid != 0 || done != unsigned((nblk - 1 - grp) / NGROUP + 1)) return;
    // the last chunk of its group; the last such group finishes the round
    if (atomicAdd(&C->alldone[par], 1u) + 1u != unsigned(min(nblk, NGROUP))) return;
    // every count this round made, loaded together (atomic loads: from where
    // the atomics were performed)
    unsigned long long gc[NGROUP];
    unsigned nn[SPREAD];
#pragma unroll
    for (int g = 0; g < NGROUP; ++g) gc[g] = __hip_atomic_load(&C->gcnt[par][g].v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
    for (int g = 0; g < SPREAD; ++g) nn[g] = __hip_atomic_load(&C->nnew[par][g].v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // the running edge total has one writer per round: the round's last chunk
    const unsigned long long before = __hip_atomic_load(&C->acc.edges, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int prev_pull = C->ctl[par ^ 1].pull;
    unsigned nd = 0;
    uint32_t dm = ~0u;
    if constexpr (has_defer<Pick>::value) {
        nd = __hip_atomic_load(&C->ndefer[par], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        dm = __hip_atomic_load(&C->dmin[par], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    Ctl c{0, 0, 0, 0};
    int tile0 = 0;
    unsigned picked = 0;  // a round that picked anything counts as a level
#pragma unroll
    for (int g = 0; g < SPREAD; ++g) picked += nn[g];
#pragma unroll
    for (int g = 0; g < NGROUP; ++g) {
        const int eg = int(gc[g] & 0xffffffffu);
        C->gtile[par][g] = tile0;
        tile0 += (eg + TILE - 1) / TILE;
        c.nF += int(gc[g] >> 32);
        c.E += eg;
    }
    C->gtile[par][NGROUP] = tile0;
    c.nTiles = tile0;
    __hip_atomic_store(&C->acc.edges, before + (unsigned long long)c.E, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (picked && C->acc.levels < level + 1) C->acc.levels = level + 1;
    if constexpr (has_direction<Pick>::value) {
        c.pull = pk.direction(c, before + c.E, prev_pull);
        if (c.pull == DIR_BINNED) C->acc.binned += 1;  // the round's one writer
    }
    if constexpr (has_defer<Pick>::value) {
        // nothing left below the threshold: raise it past the smallest
        // waiting distance and pick again next round
        if (c.nF == 0 && nd) {
            C->thr = pk.raised(dm);
            c.pull = DIR_REPICK;
        }
    }
    *next = c;
    if (level < LOG_ROUNDS) C->log[level] = c;
    if (level_arg < 0) C->cur_level = level + 1;
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
    // the entry sits in group 0; every later group starts past its tiles
    if (threadIdx.x <= NGROUP) C->gtile[0][threadIdx.x] = threadIdx.x == 0 ? 0 : ntiles;
    if (threadIdx.x != 0) return;
    C->gcnt[0][0].v = deg > 0 ? (1ull << 32) | (unsigned long long)deg : 0ull;
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
    C->nnew[0][0].v = 1;
    C->acc.vertices[0].v = 1;
    C->acc.levels = 1;
    C->acc.binned = 0;
    C->acc.edges = (unsigned long long)deg;
    Ctl c;
    c.nF = deg > 0 ? 1 : 0;
    c.E = deg;
    c.nTiles = ntiles;
    c.pull = 0;
    if constexpr (has_direction<Pick>::value) c.pull = pk.direction(c, (unsigned long long)deg, C->ctl[1].pull);
    // the source's round runs push: the host then launches none of the
    // binned round's kernels for it
    if (c.pull == DIR_BINNED) c.pull = DIR_PUSH;
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
//   optionally
//   void mid(int c, int p, uint32_t payload, int v, Pre &)
//                                                  loads that depend on the
//                                                  first ones, issued for all
//                                                  slots before any apply
//   static constexpr bool kVert                    needs the entry's vertex id
//   static constexpr bool kVal                     needs the entry's payload
// Columns outside [0, m) are skipped.
// Stages tile t (in the round's numbering over the groups) of the queued
// entries' edges into LDS (entry start marks,
// col_ind base, optional payload and vertex id) and gives this thread the
// owner entry of its IPT edge slots (slot = wid*SEG + r*64 + lane) through an
// inclusive max-scan.  Returns the number of valid slots in the tile.
template <bool KVAL, bool KVERT>
__device__ __forceinline__ int tile_owners(int t, const Round &cur, const Queue &q, int32_t *s_own, int32_t *s_base,
                                           uint32_t *s_val, int32_t *s_vert, int32_t *s_wmax, int (&own)[IPT]) {
    constexpr int NW = WG / WAVE, SEG = TILE / NW;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
    // the group holding tile t: the last whose first tile is <= t (empty
    // groups share their first tile with the next one)
    int g = 0;
#pragma unroll
    for (int k = 1; k < NGROUP; ++k) g += cur.gtile[k] <= t ? 1 : 0;
    const int lt = t - cur.gtile[g], ntg = cur.gtile[g + 1] - cur.gtile[g];
    const unsigned long long gc = cur.gcnt[g].v;
    const int nfg = int(gc >> 32), eg = int(gc & 0xffffffffu);
    const size_t qo = size_t(g) * q.qcap;
    const int32_t *tl = q.tiles + size_t(g) * q.tcap;
    const int e0 = lt * TILE;
    const int f0 = tl[lt];
    const int f1 = (lt + 1 < ntg) ? tl[lt + 1] : nfg - 1;
    const int cnt = min(f1 - f0 + 1, TILE + 1);  // every queued entry has >= 1 edge
    for (int i = tid; i < TILE; i += WG) s_own[i] = 0;
    __syncthreads();
    for (int k = tid; k < cnt; k += WG) {
        const size_t e = qo + f0 + k;
        const int o = q.off[e] - e0;
        s_base[k] = q.start[e] - o;
        if (KVAL) s_val[k] = q.val[e];
        if (KVERT) s_vert[k] = q.vert[e];
        if (o > 0 && o < TILE) s_own[o] = k;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < IPT; ++r) own[r] = s_own[wid * SEG + r * WAVE + lane];
    slot_max_scan(own, s_wmax);
    return min(TILE, eg - e0);
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
// (the tile staging in LDS given by the caller: s_own[TILE], s_base[TILE +
// 1], s_val[TILE + 1] when Op::kVal, s_vert[TILE + 1] when Op::kVert,
// s_wmax[WG / WAVE])
template <class Op>
__device__ __forceinline__ void expand_tiles_in(const Op &op, int m, const Round &cur, const Queue &q,
                                                const int32_t *__restrict__ col_ind, int bid, int nblk,
                                                int32_t *s_own, int32_t *s_base, uint32_t *s_val, int32_t *s_vert,
                                                int32_t *s_wmax) {
    constexpr int NW = WG / WAVE, SEG = TILE / NW;
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
        float wt[IPT];  // an op's edge weight, loaded beside the column (it needs only the position)
#pragma unroll
        for (int r = 0; r < IPT; ++r) {
            col[r] = pos[r] >= 0 ? col_ind[pos[r]] : -1;
            if constexpr (has_weight<Op>::value) wt[r] = pos[r] >= 0 ? op.weight(pos[r]) : 0.0f;
        }
        constexpr int B = op_batch<Op>::value;
        static_assert(IPT % B == 0, "batch divides IPT");
#pragma unroll
        for (int r0 = 0; r0 < IPT; r0 += B) {
            typename Op::Pre pre[B];
#pragma unroll
            for (int r = 0; r < B; ++r)
                if ((unsigned)col[r0 + r] < (unsigned)m) {
                    pre[r] = op.load(col[r0 + r], pos[r0 + r], Op::kVal ? s_val[own[r0 + r]] : 0u,
                                     Op::kVert ? s_vert[own[r0 + r]] : 0);
                    if constexpr (has_weight<Op>::value) pre[r].w = wt[r0 + r];
                }
            if constexpr (has_mid<Op>::value) {
#pragma unroll
                for (int r = 0; r < B; ++r)
                    if ((unsigned)col[r0 + r] < (unsigned)m)
                        op.mid(col[r0 + r], pos[r0 + r], Op::kVal ? s_val[own[r0 + r]] : 0u,
                               Op::kVert ? s_vert[own[r0 + r]] : 0, pre[r]);
            }
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
__device__ __forceinline__ void expand_tiles(const Op &op, int m, const Round &cur, const Queue &q,
                                             const int32_t *__restrict__ col_ind, int bid, int nblk) {
    constexpr int NW = WG / WAVE;
    __shared__ int32_t s_own[TILE];
    __shared__ int32_t s_base[TILE + 1];  // start - (off - e0): col_ind index of slot 0
    __shared__ uint32_t s_val[Op::kVal ? TILE + 1 : 1];
    __shared__ int32_t s_vert[Op::kVert ? TILE + 1 : 1];
    __shared__ int32_t s_wmax[NW];
    expand_tiles_in(op, m, cur, q, col_ind, bid, nblk, s_own, s_base, s_val, s_vert, s_wmax);
}

template <class Op>
__global__ __launch_bounds__(WG) void k_expand(Op op, int level_arg, int m, const Control *C, Queue q,
                                               const int32_t *__restrict__ col_ind) {
    const int level = level_arg >= 0 ? level_arg : C->cur_level;  // round of the next compaction
    const Round cur = round_of(C, level);
    if (cur.pull) return;  // this round runs as a pull step
    expand_tiles(op, m, cur, q, col_ind, blockIdx.x, gridDim.x);
}

}  // namespace trav
}  // namespace apspie
