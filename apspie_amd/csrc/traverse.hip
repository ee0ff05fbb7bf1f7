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

#include "binned.hpp"
#include "frontier.hpp"
#include "internal.hpp"

namespace apspie {
namespace trav {

// ----------------------------------------------------------------------------
// per-algorithm pieces

// Direction rule (Beamer, Asanovic, Patterson, "Direction-Optimizing
// Breadth-First Search", SC 2012): go bottom-up once the frontier's edges
// exceed the unexplored edges / alpha, stay while the frontier holds at least
// m / beta vertices.
struct PullRuleBfs {
    uint32_t *fbits;  // null: no CSC, push only
    int64_t nnz;
    int m, alpha, beta;
    __device__ int decide(const Ctl &c, unsigned long long seen, int prev) const {
        if (!fbits || alpha <= 0 || c.nF == 0) return 0;
        if (!prev) return int64_t(c.E) * alpha > nnz - int64_t(seen) ? 1 : 0;
        return int64_t(c.nF) * beta >= m ? 1 : 0;
    }
};

struct PickBfs {
    uint8_t *flags;
    uint32_t *visited;
    int32_t *depth;
    int level, seed;
    PullRuleBfs rule;
    using Pre = uint8_t;
    // flags are padded to nblk * CHUNK; null when the rounds are k_round_bfs
    // ones (only the seed kernel runs this pick then)
    __device__ Pre load(int v) const { return flags ? flags[v] : Pre(0); }
    __device__ bool pick(int v, Pre f, uint32_t &) const {
        if (f) flags[v] = 0;
        return f != 0 || v == seed;
    }
    __device__ void on_new(int v) const { depth[v] = level; }
    // this wave alone owns the bitmap words of its vertices: lane r ORs round
    // r's ballot into the visited words and writes it as the frontier words
    __device__ void on_balls(uint64_t ball, int vbase, int lane, int rounds) const {
        if (lane >= rounds) return;
        const int wi = (vbase >> 5) + 2 * lane;
        if (ball) {
            uint2 *w = (uint2 *)(visited + wi);
            uint2 o = *w;
            o.x |= uint32_t(ball);
            o.y |= uint32_t(ball >> 32);
            *w = o;
        }
        if (rule.fbits) *(uint2 *)(rule.fbits + wi) = make_uint2(uint32_t(ball), uint32_t(ball >> 32));
    }
    __device__ int direction(const Ctl &c, unsigned long long seen, int prev) const {
        return rule.decide(c, seen, prev);
    }
};

constexpr uint32_t NOT_IN_FRONTIER = 0xffffffffu;

struct PickSssp {
    uint32_t *dist, *prev;
    uint8_t *hint;  // distance bound codes, rewritten when a distance dropped (null: none)
    int m, seed;
    // null: no CSC, push only.  Else fval[v] = the distance snapshot of v if
    // v entered the frontier, NOT_IN_FRONTIER otherwise (what a pull step reads)
    uint32_t *fval;
    int64_t nnz;
    int ratio;      // pull when the frontier's edges * ratio exceed nnz
    int bin_ratio;  // binned relaxation when the frontier's edges * bin_ratio exceed nnz (0: off)
    struct Pre {
        uint32_t d, p;
    };
    __device__ Pre load(int v) const { return v < m ? Pre{dist[v], prev[v]} : Pre{0u, 0u}; }
    __device__ bool pick(int v, const Pre &x, uint32_t &payload) const {
        if (v >= m) return false;
        uint32_t d = x.d;
        if (v == seed) {
            d = 0u;
            dist[v] = 0u;
        }
        payload = d;
        const bool isnew = d < x.p;
        if (isnew) prev[v] = d;
        if (hint && isnew) hint[v] = uint8_t(hint_code(d));
        if (fval) fval[v] = isnew ? d : NOT_IN_FRONTIER;
        return isnew;
    }
    __device__ void on_new(int) const {}
    __device__ int direction(const Ctl &c, unsigned long long, int) const {
        if (bin_ratio > 0 && int64_t(c.E) * bin_ratio > nnz) return DIR_BINNED;
        if (fval && ratio > 0 && int64_t(c.E) * ratio > nnz) return DIR_PULL;
        return DIR_PUSH;
    }
};

// Near/far SSSP: a changed vertex whose distance is at or past *thr waits
// (keeps its old snapshot, so it stays changed) until the frontier below the
// threshold runs dry and the threshold rises by delta.  (A type of its own:
// the plain compaction keeps fewer registers without the waiting counts.)
struct PickSsspNF : PickSssp {
    const uint32_t *thr;
    float delta;
    __device__ uint32_t threshold() const { return __hip_atomic_load(thr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
    __device__ bool pick(int v, const Pre &x, uint32_t &payload) const {
        if (v >= m) return false;
        uint32_t d = x.d;
        if (v == seed) {
            d = 0u;
            dist[v] = 0u;
        }
        payload = d;
        const bool changed = d < x.p, isnew = changed && d < threshold();
        if (isnew) prev[v] = d;
        if (hint && changed) hint[v] = uint8_t(hint_code(d));  // a waiting vertex's bound falls too
        if (fval) fval[v] = isnew ? d : NOT_IN_FRONTIER;
        return isnew;
    }
    __device__ uint32_t defer_key(int v, const Pre &x) const {
        if (v >= m) return ~0u;
        const uint32_t d = v == seed ? 0u : x.d;
        return d < x.p && d >= threshold() ? d : ~0u;
    }
    // dmin + delta, but at least the next float above dmin: where delta is at
    // most half an ulp of dmin the sum rounds back to dmin, and a threshold
    // that does not rise would admit nothing
    __device__ uint32_t raised(uint32_t dmin) const {
        return max(__float_as_uint(__uint_as_float(dmin) + delta), dmin + 1u);
    }
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
    __device__ float weight(int p) const { return val ? val[p] : 1.0f; }
    __device__ Pre load(int c, int, uint32_t, int) const { return Pre{0.0f, dist[c]}; }  // w: weight()
    __device__ void apply(int c, int, uint32_t du, int, const Pre &pr) const {
        const uint32_t nb = __float_as_uint(__uint_as_float(du) + pr.w);
        if (nb < pr.d) __hip_atomic_fetch_min(&dist[c], nb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
};

// OpRelax behind the distance bound codes: the loads fetch the weight and the
// target's code (the weight beside the column), the middle phase probes the distance of the targets whose
// bound the candidate is below (the others get 0, which no candidate is
// below), and the apply issues the atomic as OpRelax does.
struct OpRelaxHint {
    static constexpr bool kVert = false, kVal = true;
    struct Pre {
        float w;
        uint32_t h, d;
    };
    const float *val;
    uint32_t *dist;
    const uint8_t *hint;
    __device__ float weight(int p) const { return val ? val[p] : 1.0f; }
    __device__ Pre load(int c, int, uint32_t, int) const { return Pre{0.0f, hint[c], 0u}; }  // w: weight()
    __device__ void mid(int c, int, uint32_t du, int, Pre &pr) const {
        const uint32_t nb = __float_as_uint(__uint_as_float(du) + pr.w);
        pr.d = nb < hint_bound(pr.h) ? dist[c] : 0u;
    }
    __device__ void apply(int c, int, uint32_t du, int, const Pre &pr) const {
        const uint32_t nb = __float_as_uint(__uint_as_float(du) + pr.w);
        if (nb < pr.d) __hip_atomic_fetch_min(&dist[c], nb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
};

// The relaxation of an SSSP round with binned rounds enabled, in one launch
// whichever way the round runs: the push tiles with Op, or the binning
// expansion of a binned round (bin_tiles), both on one LDS staging (a push
// expansion and a binning one as separate kernels cost a launch that exits
// at once in every round).  Pull rounds exit.
template <class Op, bool HINT>
__global__ __launch_bounds__(WG) void k_relax(Op op, int level_arg, int m, int nb, const Control *C, Queue q,
                                              const int32_t *__restrict__ col_ind, const float *__restrict__ val,
                                              const uint8_t *__restrict__ hint, uint2 *__restrict__ out,
                                              int32_t *__restrict__ tbl, int32_t *__restrict__ wcount) {
    constexpr int NW = WG / WAVE;
    static_assert(Op::kVal && !Op::kVert, "the staging holds payloads, not vertex ids");
    __shared__ int32_t s_own[TILE];
    __shared__ int32_t s_base[TILE + 1];
    __shared__ uint32_t s_val[TILE + 1];
    __shared__ int32_t s_vert[1];
    __shared__ int32_t s_wmax[NW];
    __shared__ uint32_t s_hist[MAX_BUCKETS];
    __shared__ uint32_t s_wcnt[NW];
    const int level = level_arg >= 0 ? level_arg : C->cur_level;
    const Round cur = round_of(C, level);
    if (cur.pull == DIR_PUSH)
        expand_tiles_in(op, m, cur, q, col_ind, blockIdx.x, gridDim.x, s_own, s_base, s_val, s_vert, s_wmax);
    else if (cur.pull == DIR_BINNED)
        bin_tiles<HINT>(cur, m, nb, q, col_ind, val, hint, out, tbl, wcount, s_own, s_base, s_val, s_vert, s_wmax,
                        s_hist, s_wcnt);
}

// ----------------------------------------------------------------------------
// pull steps (rounds whose compaction set Ctl::pull), over the CSC: cp[m+1]
// column pointers, ri[nnz] source rows, cv[nnz] values

// BFS with a CSC (rounds may run bottom-up).  One launch per round
// (k_round_bfs) and no pass over the vertex range: a vertex the round reaches
// gets its depth and visited bit there, enters the round's frontier bitmap
// (the input of a bottom-up successor) and is appended to the round's queue
// (the input of a top-down successor); the workgroup that finishes last
// writes the round's descriptor.  Queues alternate between two sets of
// arrays by round parity, since a round reads its predecessor's queue while
// it writes its own.
struct BfsRound {
    uint32_t *visited;
    int32_t *depth;
    uint32_t *fb;  // three frontier bitmaps of fbw words, round L's in bitmap L mod 3
    int fbw;
    const int32_t *row_ptr;
    PullRuleBfs rule;
    int cap;  // entries a round may append to one group region (<= the region's size)
};

// Lanes of a wave with equal adjacent keys form runs.  Returns the OR of x
// over this lane's run up to this lane; excl gets the OR over the run's lanes
// before this one, tail the run's last lane.  Every lane of the wave calls it.
__device__ __forceinline__ uint32_t run_or(int key, uint32_t x, int lane, uint32_t &excl, int &tail) {
    const int kprev = __shfl_up(key, 1);
    const uint64_t heads = __ballot(lane == 0 || kprev != key);
    const int h = 63 - __clzll((long long)(heads & (~0ull >> (63 - lane))));  // this run's first lane
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        const uint32_t y = __shfl_up(x, d);
        if (lane - d >= h) x |= y;
    }
    const uint32_t p = __shfl_up(x, 1);
    excl = lane - 1 >= h ? p : 0u;
    const uint64_t tails = (heads >> 1) | (1ull << 63);
    tail = __ffsll((long long)(tails & (~0ull << lane))) - 1;
    return x;
}

// Appends the vertices this thread holds (get(j), j < K, -1: none) with
// out-degree > 0 to group grp's region of queue qn: one returning atomic add
// of the workgroup's (entries, edges) on the group's counter, entries in
// thread order.  A workgroup whose entries do not fit the region's first cap
// entries writes none and raises the overflow flag (its counts still go to
// the counter: the totals stay exact; the flag's atomic is folded into dep).
// s_scan holds NW + 1 words; contains barriers.
template <int K, class Get>
__device__ __forceinline__ void append_entries(const Get &get, const int32_t *__restrict__ row_ptr, const Queue &qn,
                                               int cap, Padded<unsigned long long> *gcnt, int grp,
                                               unsigned *overflow, uint64_t *s_scan, uint32_t &dep) {
    constexpr int NW = WG / WAVE;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
    uint32_t cnt = 0, es = 0;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const int v = get(j);
        if (v >= 0) {
            const int d = row_ptr[v + 1] - row_ptr[v];
            cnt += d > 0;
            es += d;
        }
    }
    const uint64_t x = (uint64_t(cnt) << 32) | es;
    const uint64_t incl = wave_incl_scan(x, lane);
    if (lane == WAVE - 1) s_scan[wid] = incl;
    __syncthreads();
    uint64_t wpre = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        if (w < wid) wpre += s_scan[w];
        tot += s_scan[w];
    }
    __syncthreads();
    if (tid == 0) s_scan[NW] = tot ? atomicAdd(&gcnt[grp].v, tot) : 0ull;
    __syncthreads();
    const uint64_t gb = s_scan[NW];
    __syncthreads();  // s_scan is free again for the caller
    if (!tot) return;
    if ((gb >> 32) + (tot >> 32) > uint64_t(cap)) {
        if (tid == 0) dep ^= atomicOr(overflow, 1u);
        return;
    }
    const uint64_t base = gb + wpre + incl - x;
    int pos = int(base >> 32), off = int(base & 0xffffffffu);
    const size_t qo = size_t(grp) * qn.qcap;
    int32_t *const tl = qn.tiles + size_t(grp) * qn.tcap;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const int v = get(j);
        if (v < 0) continue;
        const int r0 = row_ptr[v], d = row_ptr[v + 1] - r0;
        if (d <= 0) continue;
        qn.vert[qo + pos] = v;
        qn.start[qo + pos] = r0;
        qn.off[qo + pos] = off;
        for (int t = (off + TILE - 1) / TILE; t * TILE < off + d; ++t) tl[t] = pos;
        ++pos;
        off += d;
    }
}

constexpr int PULL_CHUNK = 1024;  // vertices per bottom-up chunk (32 bitmap words, 4 per thread)

// Bottom-up: chunks of PULL_CHUNK consecutive vertices, one workgroup each at
// a time.  Every unvisited vertex scans its in-edges for a parent in the
// previous frontier's bitmap and stops at the first one (on RMAT graphs
// nearly all find one at their first in-edge: the hubs, lowest ids, come
// first); each thread walks its vertices on its own and marks the ones it
// reaches in an LDS bitmap of the chunk, which the workgroup then writes
// whole into the visited and frontier words of the chunk and appends.
//
// Top-down: the queued frontier's edges in TILE-edge tiles (tile_owners).
// An unvisited target is claimed with an atomic OR on its visited word; the
// lanes of a wave whose targets share a word (a source's sorted columns put
// them side by side) OR their bits together first and the run's last lane
// issues one atomic for all, since atomics on one line are serialised.  A
// lane owns its vertex when the bit was clear before the run's atomic and no
// earlier lane of the run holds the same vertex; owned vertices are listed in
// LDS and appended per tile.
//
// (eight waves per SIMD: the bottom-up scan is bound by the latency of its
// dependent loads)
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_round_bfs(
    BfsRound rb, int level, int m, int expand_blocks, Control *C, Queue qc, Queue qn,
    const int32_t *__restrict__ col_ind, const int32_t *__restrict__ cp, const int32_t *__restrict__ ri) {
    constexpr int NW = WG / WAVE, SEG = TILE / NW;
    static_assert(PULL_CHUNK == 4 * WG && PULL_CHUNK / 32 <= TILE, "chunk shape");
    __shared__ int32_t s_own[TILE];  // tile owners; then this tile's claims, or the chunk bitmap
    __shared__ int32_t s_base[TILE + 1];
    __shared__ int32_t s_wmax[NW];
    __shared__ uint64_t s_scan[NW + 1];
    __shared__ uint32_t s_found[NW];
    __shared__ int s_ncl;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
    if (C->rebuild) return;  // an earlier round's queue is being rebuilt: this launch is void
    const int par = level & 1;
    const Round cur = round_of(C, level);
    if (cur.nF == 0) {  // nothing left: the descriptor of this round is empty too
        if (blockIdx.x == 0 && tid == 0) C->ctl[par] = Ctl{0, 0, 0, 0};
        return;
    }
    const uint32_t *fcur = rb.fb + size_t((level + 2) % 3) * rb.fbw;
    uint32_t *fnext = rb.fb + size_t(level % 3) * rb.fbw;
    // the bitmap of the round after this one (last read by the previous round)
    uint32_t *fclear = rb.fb + size_t((level + 1) % 3) * rb.fbw;
    for (int i = blockIdx.x * WG + tid; i < rb.fbw; i += gridDim.x * WG) fclear[i] = 0u;
    const bool pull = cur.pull == DIR_PULL;
    const int nchunks = (m + PULL_CHUNK - 1) / PULL_CHUNK;
    const int parts = pull ? min(int(gridDim.x), nchunks) : min(expand_blocks, cur.nTiles);
    const int bid = blockIdx.x, grp = bid % NGROUP;
    if (bid >= parts) return;
    Padded<unsigned long long> *gcnt = C->gcnt[par];
    uint32_t found = 0;
    uint32_t dep = 0;  // returning atomics of thread 0 awaited before the completion count
    if (pull) {
        uint32_t *s_bits = (uint32_t *)s_own;
        for (int ch = bid; ch < nchunks; ch += gridDim.x) {
            const int base = ch * PULL_CHUNK, nv = min(PULL_CHUNK, m - base), nw = (nv + 31) / 32;
            for (int i = tid; i < PULL_CHUNK / 32; i += WG) s_bits[i] = 0u;  // every word append_entries reads
            __syncthreads();
            for (int i = tid; i < nv; i += WG) {
                const int v = base + i;
                if ((rb.visited[v >> 5] >> (v & 31)) & 1u) continue;
                const int a = cp[v], b = cp[v + 1];
                for (int e = a; e < b; ++e) {
                    const int u = ri[e];
                    if ((unsigned)u < (unsigned)m && ((fcur[u >> 5] >> (u & 31)) & 1u)) {
                        rb.depth[v] = level;
                        found++;
                        atomicOr(&s_bits[i >> 5], 1u << (i & 31));
                        break;
                    }
                }
            }
            __syncthreads();
            for (int i = tid; i < nw; i += WG) {  // base is a multiple of 32
                const uint32_t bits = s_bits[i];
                const int wi = (base >> 5) + i;
                if (bits) rb.visited[wi] |= bits;
                fnext[wi] = bits;
            }
            // this thread's four vertices of the chunk, 4 * tid .. 4 * tid + 3
            const uint32_t nib = (uint32_t(s_bits[tid >> 3]) >> (4 * (tid & 7))) & 0xfu;
            append_entries<4>([&](int j) { return (nib >> j) & 1u ? base + 4 * tid + j : -1; }, rb.row_ptr, qn,
                              rb.cap, gcnt, grp, &C->qoverflow[par], s_scan, dep);
            __syncthreads();  // the next chunk clears s_bits
        }
    } else {
        int32_t *s_list = s_own;
        if (tid == 0) s_ncl = 0;
        for (int t = bid; t < cur.nTiles; t += expand_blocks) {
            int own[IPT];
            const int elim = tile_owners<false, false>(t, cur, qc, s_own, s_base, nullptr, nullptr, s_wmax, own);
            int col[IPT];
#pragma unroll
            for (int r = 0; r < IPT; ++r) {
                const int s = wid * SEG + r * WAVE + lane;
                col[r] = s < elim ? col_ind[s_base[own[r]] + s] : -1;
                if ((unsigned)col[r] >= (unsigned)m) col[r] = -1;
            }
            __syncthreads();  // s_own becomes the claim list
            // claims, in batches of CB slots: every run's atomic of a batch is
            // issued before any result is used
            constexpr int CB = 2;
#pragma unroll
            for (int r0 = 0; r0 < IPT; r0 += CB) {
                uint32_t w[CB], old[CB], mine[CB];
                int tail[CB];
#pragma unroll
                for (int j = 0; j < CB; ++j) w[j] = col[r0 + j] >= 0 ? rb.visited[col[r0 + j] >> 5] : ~0u;
#pragma unroll
                for (int j = 0; j < CB; ++j) {
                    const int c = col[r0 + j];
                    const uint32_t bit = c >= 0 && !((w[j] >> (c & 31)) & 1u) ? 1u << (c & 31) : 0u;
                    uint32_t excl;
                    const uint32_t incl = run_or(c >= 0 ? c >> 5 : -1, bit, lane, excl, tail[j]);
                    mine[j] = bit & ~excl;  // no earlier lane of the run holds this vertex
                    old[j] = ~0u;
                    if (tail[j] == lane && incl) old[j] = atomicOr(&rb.visited[c >> 5], incl);
                    if (tail[j] == lane) w[j] = incl;  // the run's bits, kept by its last lane
                }
#pragma unroll
                for (int j = 0; j < CB; ++j) {
                    const uint32_t before = __shfl(old[j], tail[j]);
                    const int c = col[r0 + j];
                    // the run's new bits enter the frontier bitmap in one atomic
                    if (tail[j] == lane && (w[j] & ~before)) atomicOr(&fnext[c >> 5], w[j] & ~before);
                    if (mine[j] & ~before) {
                        rb.depth[c] = level;
                        found++;
                        s_list[atomicAdd(&s_ncl, 1)] = c;  // at most TILE claims per tile
                    }
                }
            }
            __syncthreads();
            const int ncl = s_ncl;
            append_entries<IPT>(
                [&](int j) {
                    const int i = tid * IPT + j;
                    return i < ncl ? s_list[i] : -1;
                },
                rb.row_ptr, qn, rb.cap, gcnt, grp, &C->qoverflow[par], s_scan, dep);
            __syncthreads();  // every thread is done with s_list and s_ncl
            if (tid == 0) s_ncl = 0;  // (the next tile's first barrier orders this)
        }
    }
    // reached vertices of this workgroup, then its completion
    found = wave_total32(found);
    if (lane == 0) s_found[wid] = found;
    __syncthreads();
    if (tid != 0) return;
    uint32_t f = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) f += s_found[w];
    if (f) dep ^= atomicAdd(&C->rfound[par][bid % SPREAD].v, f);
    // (the group counters' returns were awaited inside append_entries)
    after_values(dep);
    const unsigned done = atomicAdd(&C->gdone[par][grp].v, 1u) + 1u;
    if (done != unsigned((parts - 1 - grp) / NGROUP + 1)) return;
    if (atomicAdd(&C->alldone[par], 1u) + 1u != unsigned(min(parts, NGROUP))) return;
    // the last workgroup: the round's descriptor, from atomic loads
    unsigned long long gc[NGROUP];
#pragma unroll
    for (int g = 0; g < NGROUP; ++g) gc[g] = __hip_atomic_load(&gcnt[g].v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned nfound = 0;
#pragma unroll
    for (int g = 0; g < SPREAD; ++g)
        nfound += __hip_atomic_load(&C->rfound[par][g].v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned over = __hip_atomic_load(&C->qoverflow[par], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long before = C->acc.edges;
    Ctl c{0, 0, 0, 0};
    int gt[NGROUP + 1];
    int tile0 = 0;
#pragma unroll
    for (int g = 0; g < NGROUP; ++g) {
        const int eg = int(gc[g] & 0xffffffffu);
        gt[g] = tile0;
        tile0 += (eg + TILE - 1) / TILE;
        c.nF += int(gc[g] >> 32);
        c.E += eg;
    }
    gt[NGROUP] = tile0;
    c.nTiles = tile0;
    c.pull = rb.rule.decide(c, before + (unsigned long long)c.E, cur.pull);
    if (over && c.nF > 0 && c.pull != DIR_PULL) {
        // the next round expands a queue this round could not hold: a pass
        // over this round's bitmap (k_compact<PickBfsBits>) queues it and
        // writes the descriptor; its counters start from zero
        for (int g = 0; g < NGROUP; ++g) {
            gcnt[g].v = 0;
            C->gdone[par][g].v = 0;
            C->nnew[par][g].v = 0;
            C->rfound[par][g].v = 0;
        }
        C->alldone[par] = 0;
        C->qoverflow[par] = 0;
        C->rebuild = level + 1;
        return;
    }
    for (int g = 0; g <= NGROUP; ++g) C->gtile[par][g] = gt[g];
    C->ctl[par] = c;
    if (level < LOG_ROUNDS) C->log[level] = c;
    C->acc.edges = before + (unsigned long long)c.E;
    if (nfound) {
        C->acc.vertices[0].v += nfound;
        if (C->acc.levels < level + 1) C->acc.levels = level + 1;
    }
    // counters of the next round (their last users, the previous round and
    // this round's readers of its queue, have finished)
    for (int g = 0; g < NGROUP; ++g) {
        C->gcnt[par ^ 1][g].v = 0;
        C->gdone[par ^ 1][g].v = 0;
        C->rfound[par ^ 1][g].v = 0;
    }
    C->alldone[par ^ 1] = 0;
    C->qoverflow[par ^ 1] = 0;
}

// Rebuilds the queue of a round whose appends overflowed: the vertices of the
// round's frontier bitmap (depth, visited and bitmap are already written) go
// through the generic compaction, which writes the descriptor.
struct PickBfsBits {
    const uint32_t *bits;  // the round's frontier bitmap
    PullRuleBfs rule;
    using Pre = uint32_t;
    __device__ Pre load(int v) const { return bits[v >> 5]; }  // the bitmap covers nblk * CHUNK vertices
    __device__ bool pick(int v, Pre w, uint32_t &) const { return (w >> (v & 31)) & 1u; }
    __device__ void on_new(int) const {}
    __device__ int direction(const Ctl &c, unsigned long long seen, int prev) const {
        return rule.decide(c, seen, prev);
    }
};

// Row holding the first edge of every TILE-edge tile of the CSC.
__global__ __launch_bounds__(WG) void k_tile_rows(int m, const int32_t *__restrict__ cp, int32_t *trows) {
    for (int v = blockIdx.x * WG + threadIdx.x; v < m; v += gridDim.x * WG) {
        const int a = cp[v], b = cp[v + 1];
        for (int t = (a + TILE - 1) / TILE; t * TILE < b; ++t) trows[t] = v;
    }
}

// SSSP pull: X[v] = min(X[v], min over in-edges (u, v) with u in the frontier
// of Q[u] + w), Q[u] = u's distance when the compaction put it in the
// frontier (NOT_IN_FRONTIER for the rest), so an in-edge costs one random
// 4-byte read.  (X[v] is compared live: a vertex waiting past the near/far
// threshold already holds a distance below its snapshot.)  Equal TILE-edge tiles of the CSC; row starts are marked in
// LDS and a max-scan gives every slot the start slot of its row; a segmented
// wave min per 64 slots, then one LDS atomicMin per segment, then one write
// per row: a plain store for rows inside the tile, an atomicMin for the (at
// most two) rows it shares with other tiles.  Only the improved vertices are
// written, so X below the snapshot marks the next frontier exactly as a push
// round leaves it.
__global__ __launch_bounds__(WG) void k_pull_relax(int level_arg, int m, int nnz, const Control *C,
                                                   const int32_t *__restrict__ cp, const int32_t *__restrict__ ri,
                                                   const float *__restrict__ cv, const int32_t *__restrict__ trows,
                                                   const uint32_t *__restrict__ Q, uint32_t *X) {
    constexpr int NW = WG / WAVE, SEG = TILE / NW;
    constexpr uint32_t NONE = NOT_IN_FRONTIER;
    __shared__ int32_t s_rid[TILE];   // row (relative to the tile's first row) starting at a slot, 0 if none
    __shared__ uint32_t s_best[TILE];  // min candidate of the row starting at a slot
    __shared__ int32_t s_wmax[NW];
    const int level = level_arg >= 0 ? level_arg : C->cur_level;
    if (C->ctl[(level - 1) & 1].pull != DIR_PULL) return;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
    const int ntiles = (nnz + TILE - 1) / TILE;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int e0 = t * TILE;
        const int elim = min(TILE, nnz - e0);
        const int r0 = trows[t];
        const int r1 = t + 1 < ntiles ? trows[t + 1] : m - 1;
        for (int i = tid; i < TILE; i += WG) {
            s_rid[i] = 0;
            s_best[i] = NONE;
        }
        __syncthreads();
        // an empty row shares its start with the next row: the larger index wins
        for (int i = 1 + tid; i <= r1 - r0; i += WG) {
            const int o = cp[r0 + i] - e0;
            if (o > 0 && o < elim) atomicMax(&s_rid[o], i);
        }
        __syncthreads();
        int st[IPT], u[IPT];
        uint32_t cand[IPT];
#pragma unroll
        for (int r = 0; r < IPT; ++r) {
            const int s = wid * SEG + r * WAVE + lane;
            st[r] = (s == 0 || s_rid[s] > 0) ? s : 0;
            u[r] = s < elim ? ri[e0 + s] : -1;
        }
        slot_max_scan(st, s_wmax);
#pragma unroll
        for (int r = 0; r < IPT; ++r) cand[r] = (unsigned)u[r] < (unsigned)m ? Q[u[r]] : NONE;
#pragma unroll
        for (int r = 0; r < IPT; ++r)
            if (cand[r] != NONE) {
                const int s = wid * SEG + r * WAVE + lane;
                const float w = cv ? cv[e0 + s] : 1.0f;
                cand[r] = __float_as_uint(__uint_as_float(cand[r]) + w);
            }
#pragma unroll
        for (int r = 0; r < IPT; ++r) {
            uint32_t x = cand[r];
            const int sr = st[r];
#pragma unroll
            for (int d = 1; d < WAVE; d <<= 1) {
                const uint32_t y = __shfl_up(x, d, WAVE);
                const int ys = __shfl_up(sr, d, WAVE);
                if (lane >= d && ys == sr) x = min(x, y);
            }
            const int nxt = __shfl_down(sr, 1, WAVE);
            if ((lane == WAVE - 1 || nxt != sr) && x != NONE) atomicMin(&s_best[sr], x);
        }
        __syncthreads();
        for (int s = tid; s < elim; s += WG) {
            if (s != 0 && s_rid[s] == 0) continue;
            const uint32_t b = s_best[s];
            if (b == NONE) continue;
            const int v = r0 + s_rid[s];
            if (b < X[v]) {  // a row inside the tile has no other writer
                if (cp[v] >= e0 && cp[v + 1] <= e0 + elim)
                    X[v] = b;
                else
                    __hip_atomic_fetch_min(&X[v], b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        __syncthreads();
    }
}

// (flags: BFS without a CSC; fbits: the three frontier bitmaps of BFS with
// one; fval: SSSP pull snapshots)
__global__ __launch_bounds__(WG) void k_init(int mode, int m, int nblk, int32_t *depth, uint32_t *dist,
                                             uint32_t *prev, uint8_t *flags, uint32_t *visited,
                                             uint64_t *status, Control *C, uint32_t thr0,
                                             uint32_t *fbits, uint32_t *fval, uint8_t *hint) {
    const int64_t gtid = int64_t(blockIdx.x) * WG + threadIdx.x;
    const int64_t stride = int64_t(gridDim.x) * WG;
    const int64_t mpad = int64_t(nblk) * CHUNK;
    if (mode == MODE_BFS) {
        for (int64_t i = gtid; i < mpad / 16; i += stride) {
            if (flags) *(uint4 *)(flags + 16 * i) = make_uint4(0, 0, 0, 0);
            if ((i & 1) == 0) {  // one bitmap word per 32 vertices
                const int64_t w = i >> 1, words = mpad / 32;
                visited[w] = 0u;
                if (fbits) fbits[w] = fbits[words + w] = fbits[2 * words + w] = 0u;
            }
        }
        for (int64_t i = gtid; i < m; i += stride) depth[i] = -1;
    } else {
        // 16-byte stores where the arrays allow (the caller's distances may
        // be only 4-byte aligned; prev and fval are pool slots)
        const int64_t m4 = m / 4;
        const uint4 inf4 = make_uint4(INF_BITS, INF_BITS, INF_BITS, INF_BITS);
        const bool dvec = (reinterpret_cast<uintptr_t>(dist) & 15) == 0;
        for (int64_t i = gtid; i < m4; i += stride) {
            if (dvec) ((uint4 *)dist)[i] = inf4;
            if (prev) ((uint4 *)prev)[i] = inf4;
            if (fval) ((uint4 *)fval)[i] = make_uint4(NOT_IN_FRONTIER, NOT_IN_FRONTIER, NOT_IN_FRONTIER, NOT_IN_FRONTIER);
        }
        for (int64_t i = (dvec ? 4 * m4 : 0) + gtid; i < m; i += stride) dist[i] = INF_BITS;
        for (int64_t i = 4 * m4 + gtid; i < m; i += stride) {
            if (prev) prev[i] = INF_BITS;
            if (fval) fval[i] = NOT_IN_FRONTIER;
        }
        // no bound yet (the slot holds nblk * CHUNK bytes, 16-byte aligned)
        if (hint)
            for (int64_t i = gtid; i < mpad / 16; i += stride) *(uint4 *)(hint + 16 * i) = make_uint4(~0u, ~0u, ~0u, ~0u);
    }
    for (int64_t i = gtid; i < 2 * int64_t(nblk); i += stride) status[i] = 0;
    control_clear(C, gtid, stride, thr0);
}

// ----------------------------------------------------------------------------
// host side

// Entries per group region of the queue: a group's chunks hold at most
// CHUNK picked vertices each.  Tiles per group region: a group holds at most
// every edge.
int32_t group_entries(int nblk) { return int32_t(ceil_div(nblk, NGROUP)) * CHUNK; }
int32_t group_tiles(int64_t nnz) { return int32_t(nnz / TILE + 2); }

Bufs get_bufs(apspie_ctx *ctx, int m, int64_t nnz) {
    Bufs b;
    b.nblk = std::max(1, (int)ceil_div(m, CHUNK));
    const size_t mpad = size_t(b.nblk) * CHUNK;
    b.qcap = group_entries(b.nblk);
    b.tcap = group_tiles(nnz);
    const size_t qn = size_t(NGROUP) * size_t(b.qcap);
    b.flags = (uint8_t *)ctx->slot(S_FLAGS, mpad);
    b.visited = (uint32_t *)ctx->slot(S_VISITED, mpad / 8);
    b.prev = (uint32_t *)ctx->slot(S_PREV, (size_t(m) + 1) * 4);
    b.queue = (int32_t *)ctx->slot(S_QUEUE, qn * 4);
    b.qstart = (int32_t *)ctx->slot(S_QSTART, qn * 4);
    b.qoff = (int32_t *)ctx->slot(S_QOFF, qn * 4);
    b.qval = (float *)ctx->slot(S_QVAL, qn * 4);
    b.tiles = (int32_t *)ctx->slot(S_TILES, size_t(NGROUP) * size_t(b.tcap) * 4);
    b.status = (uint64_t *)ctx->slot(S_STATUS, size_t(b.nblk) * 2 * 8);
    b.C = (Control *)ctx->slot(S_CTL, sizeof(Control));
    b.expand_grid =
        std::max(1u, std::min(ceil_div(nnz, TILE) + 1, unsigned(ctx->cu_count * ctx->tune.expand_wg_per_cu)));
    return b;
}

Queue queue_of(const Bufs &b) { return Queue{b.queue, b.qstart, b.qoff, (uint32_t *)b.qval, b.tiles, b.qcap, b.tcap}; }

// ----------------------------------------------------------------------------
// fixed scratch of the allocation-free SSSP (apspie_sssp_noalloc): one array
// per device, loaded with the library's code object, never allocated or freed
// by a call.  Host code reaches it only through the address
// hipGetSymbolAddress gives for each device (the array's name is a host-side
// stand-in, not a device address).

constexpr size_t NOALLOC_BYTES = size_t(48) << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_noalloc_mu[MAX_DEVICES];     // serialises the calls that share a device's array
void *g_noalloc_ptr[MAX_DEVICES] = {};    // the array's address on each device
std::mutex g_noalloc_setup_mu;

// Looks the array up for the current device (once per device).  Called with
// `device` current, by apspie_ctx_create.
void noalloc_setup(int device) {
    APSPIE_CHECK(device >= 0 && device < MAX_DEVICES, APSPIE_ELIMIT, "device id beyond the fixed scratch table");
    std::lock_guard<std::mutex> lk(g_noalloc_setup_mu);
    if (g_noalloc_ptr[device]) return;
    void *ptr = nullptr;
    APSPIE_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    APSPIE_CHECK(ptr != nullptr, APSPIE_EHIP, "no device address for the fixed scratch");
    g_noalloc_ptr[device] = ptr;
}

// Bytes a single-source SSSP over (m, nnz) carves from one block: distance
// snapshots, the four queue arrays (a region per chunk group), the tile
// tables, the look-back status and the control block, each 256-byte aligned.
size_t sssp_carve_bytes(int m, int64_t nnz) {
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    const int nblk = std::max(1, (int)ceil_div(m, CHUNK));
    const size_t qn = size_t(NGROUP) * size_t(group_entries(nblk));
    return al((size_t(m) + 1) * 4) + 4 * al(qn * 4) + al(size_t(NGROUP) * size_t(group_tiles(nnz)) * 4) +
           al(size_t(nblk) * 2 * 8) + al(sizeof(Control));
}

// The SSSP buffers carved from `base` (sssp_carve_bytes(m, nnz) bytes); the
// BFS-only flags and bitmaps stay null.
Bufs sssp_carve(apspie_ctx *ctx, unsigned char *base, int m, int64_t nnz) {
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    size_t at = 0;
    auto take = [&](size_t bytes) {
        void *p = base + at;
        at += al(bytes);
        return p;
    };
    Bufs b;
    b.nblk = std::max(1, (int)ceil_div(m, CHUNK));
    b.qcap = group_entries(b.nblk);
    b.tcap = group_tiles(nnz);
    const size_t qn = size_t(NGROUP) * size_t(b.qcap);
    b.prev = (uint32_t *)take((size_t(m) + 1) * 4);
    b.queue = (int32_t *)take(qn * 4);
    b.qstart = (int32_t *)take(qn * 4);
    b.qoff = (int32_t *)take(qn * 4);
    b.qval = (float *)take(qn * 4);
    b.tiles = (int32_t *)take(size_t(NGROUP) * size_t(b.tcap) * 4);
    b.status = (uint64_t *)take(size_t(b.nblk) * 2 * 8);
    b.C = (Control *)take(sizeof(Control));
    b.expand_grid =
        std::max(1u, std::min(ceil_div(nnz, TILE) + 1, unsigned(ctx->cu_count * ctx->tune.expand_wg_per_cu)));
    return b;
}

unsigned init_grid(const Bufs &b, int m) {
    return std::max(1u, std::min(ceil_div(std::max<int64_t>(m, int64_t(b.nblk) * CHUNK / 16), WG), 4096u));
}

void fill_stats(apspie_ctx *ctx, const Bufs &b, int mode, int m, apspie_stats *stats, const Acc *host_acc) {
    Acc acc;
    if (host_acc) {
        acc = *host_acc;
    } else {
        APSPIE_HIP(hipMemcpyAsync(&acc, &b.C->acc, sizeof(Acc), hipMemcpyDeviceToHost, ctx->stream));
        APSPIE_HIP(hipStreamSynchronize(ctx->stream));
    }
    ctx->prof_collect();
    if (!stats) return;
    float ms = 0.f;
    APSPIE_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    memset(stats, 0, sizeof(*stats));
    stats->edges = (int64_t)acc.edges;
    for (int i = 0; i < SPREAD; ++i) stats->vertices += (int64_t)acc.vertices[i].v;
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
                                                      b.C);
    ctx->prof_end();
}

template <class Op>
void launch_expand(apspie_ctx *ctx, const Bufs &b, const apspie_csr *g, const Op &op, int level) {
    ctx->prof_begin(APSPIE_KCLASS_EXPAND);
    k_expand<Op><<<b.expand_grid, WG, 0, ctx->stream>>>(op, level, g->m, b.C, queue_of(b), g->col_ind);
    ctx->prof_end();
}

// fixed: buffers carved from the fixed scratch (apspie_sssp_noalloc); the
// traversal then takes no slot of the pool, so it runs push rounds only
// (pull and binned rounds need buffers of their own; the results are the same)
template <int MODE>
void run(apspie_ctx *ctx, const apspie_csr *g, int32_t src, int32_t *depth, float *dist_f,
         apspie_stats *stats, const Bufs *fixed = nullptr) {
    Bufs b = fixed ? *fixed : get_bufs(ctx, g->m, g->nnz);
    const bool pull = !fixed && g->csc_col_ptr != nullptr && g->nnz > 0 &&
                      (MODE == MODE_BFS ? ctx->tune.bfs_alpha > 0 : ctx->tune.sssp_pull_ratio > 0);
    uint32_t *fval = nullptr;
    if (pull) {  // growing a slot leaves the others in place
        if (MODE == MODE_BFS) {
            b.fbw = int32_t(size_t(b.nblk) * CHUNK / 32);
            b.fbits = (uint32_t *)ctx->slot(S_FBITS, 3 * size_t(b.fbw) * 4);
            b.flags = nullptr;  // the rounds keep depth, visited and the frontier themselves
            const size_t qn = size_t(NGROUP) * size_t(b.qcap);
            b.queue1 = (int32_t *)ctx->slot(S_QUEUE1, qn * 4);
            b.qstart1 = (int32_t *)ctx->slot(S_QSTART1, qn * 4);
            b.qoff1 = (int32_t *)ctx->slot(S_QOFF1, qn * 4);
            b.tiles1 = (int32_t *)ctx->slot(S_TILES1, size_t(NGROUP) * size_t(b.tcap) * 4);
        } else {
            fval = (uint32_t *)ctx->slot(S_FVAL, size_t(g->m) * 4);
            b.trows = (int32_t *)ctx->slot(S_TROWS, (size_t(g->nnz) / TILE + 2) * 4);
        }
    }
    // binned relaxation of dense SSSP rounds: the bin workgroups' candidate
    // regions (a round has at most max_tiles + NGROUP tiles, a workgroup
    // ceil(tiles / G) of them), the bucket-ordered candidates (at most one per
    // frontier edge), the per-(bucket, workgroup) table with the workgroups'
    // counts after it, and the bucket totals
    const int nb = (int)ceil_div(g->m, BUCKET);
    const bool binned =
        !fixed && MODE == MODE_SSSP && ctx->tune.sssp_bin_ratio > 0 && g->nnz > 0 && nb <= MAX_BUCKETS;
    uint2 *bin_out = nullptr, *bin_sorted = nullptr;
    int32_t *bin_tbl = nullptr, *bin_wcnt = nullptr, *bin_tot = nullptr;
    const int max_tiles = (int)ceil_div(g->nnz, TILE);
    const int bin_grid = int(b.expand_grid);  // k_relax's grid: the push expansion's
    if (binned) {
        bin_out = (uint2 *)ctx->slot(S_BINOUT, size_t(max_tiles + NGROUP + bin_grid) * TILE * sizeof(uint2));
        bin_sorted = (uint2 *)ctx->slot(S_BINSORT, size_t(g->nnz) * sizeof(uint2));
        bin_tbl = (int32_t *)ctx->slot(S_BINTBL, (size_t(nb) + 1) * bin_grid * sizeof(int32_t));
        bin_wcnt = bin_tbl + size_t(nb) * bin_grid;
        bin_tot = (int32_t *)ctx->slot(S_BINTOT, MAX_BUCKETS * sizeof(int32_t));
    }
    uint32_t *dist = (uint32_t *)dist_f;
    const int nnz32 = int(g->nnz);
    const unsigned pull_grid = std::max(1u, std::min(ceil_div(g->nnz, TILE), (unsigned)ctx->cu_count * 8u));
    const unsigned vert_grid = std::max(1u, std::min(ceil_div(g->m, WG), (unsigned)ctx->cu_count * 16u));
    const unsigned apply_grid = unsigned(nb + g->nnz / PART_ENTRIES + 1);
    // binned rounds may still come: the rounds launched now carry the binned
    // kernels and their compactions may choose binned rounds.  Off once a push
    // round follows the binned ones (frontiers rise, then fall): the later
    // rounds then launch without them, and their compactions cannot choose them.
    bool bin_on = binned;
    int relax_rounds = -1;  // rounds launched with bin_on, once it turns off
    APSPIE_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    ctx->prof_begin(APSPIE_KCLASS_OTHER);
    // near/far threshold: the first delta of distance (none for BFS or delta 0)
    const float delta = MODE == MODE_SSSP ? float(ctx->tune.sssp_delta) : 0.0f;
    uint32_t thr0 = ~0u;
    if (delta > 0.0f) memcpy(&thr0, &delta, 4);
    // distance bound codes (push rounds of a traversal with its own buffers)
    uint8_t *hint = nullptr;
    if (MODE == MODE_SSSP && !fixed && ctx->tune.sssp_hint)
        hint = (uint8_t *)ctx->slot(S_HINT, size_t(b.nblk) * CHUNK);
    k_init<<<init_grid(b, g->m), WG, 0, ctx->stream>>>(MODE, g->m, b.nblk, depth, dist, b.prev, b.flags,
                                                       b.visited, b.status, b.C, thr0,
                                                       MODE == MODE_BFS ? b.fbits : nullptr, fval, hint);
    if (pull && MODE == MODE_SSSP)
        k_tile_rows<<<vert_grid, WG, 0, ctx->stream>>>(g->m, g->csc_col_ptr, b.trows);
    ctx->prof_end();
    auto pick_for = [&](int level, int seed) {
        if constexpr (MODE == MODE_BFS)
            return PickBfs{b.flags, b.visited, depth, level, seed,
                           PullRuleBfs{b.fbits, g->nnz, g->m, ctx->tune.bfs_alpha, ctx->tune.bfs_beta}};
        else
            return PickSssp{dist, b.prev, hint, g->m, seed, fval, g->nnz, ctx->tune.sssp_pull_ratio,
                            bin_on ? ctx->tune.sssp_bin_ratio : 0};
    };
    auto near_far = [&](const PickSssp &pk) {
        PickSsspNF nf;
        static_cast<PickSssp &>(nf) = pk;
        nf.thr = &b.C->thr;
        nf.delta = delta;
        return nf;
    };
    auto compact = [&](int level) {
        if constexpr (MODE == MODE_SSSP)
            if (delta > 0.0f) return launch_compact(ctx, b, g, near_far(pick_for(level, -1)), level, 0);
        launch_compact(ctx, b, g, pick_for(level, -1), level, 0);
    };
    // BFS with a CSC: one k_round_bfs launch per round, queues by round parity
    auto queue_par = [&](int p) {
        Queue q = queue_of(b);
        if (p) {
            q.vert = b.queue1;
            q.start = b.qstart1;
            q.off = b.qoff1;
            q.tiles = b.tiles1;
        }
        return q;
    };
    const PullRuleBfs bfs_rule{b.fbits, g->nnz, g->m, ctx->tune.bfs_alpha, ctx->tune.bfs_beta};
    const int bfs_cap = ctx->tune.bfs_group_cap > 0 ? std::min(ctx->tune.bfs_group_cap, b.qcap) : b.qcap;
    auto bfs_round = [&](int level) {
        ctx->prof_begin(APSPIE_KCLASS_ROUND);
        const BfsRound rb{b.visited, depth, b.fbits, b.fbw, g->row_ptr, bfs_rule, bfs_cap};
        k_round_bfs<<<std::max(b.expand_grid, vert_grid), WG, 0, ctx->stream>>>(
            rb, level, g->m, int(b.expand_grid), b.C, queue_par((level - 1) & 1), queue_par(level & 1), g->col_ind,
            g->csc_col_ptr, g->csc_row_ind);
        ctx->prof_end();
    };
    // a round whose queue overflowed its group regions: the pass over its bitmap
    auto bfs_rebuild = [&](int level) {
        ctx->prof_begin(APSPIE_KCLASS_COMPACT);
        k_compact<PickBfsBits><<<b.nblk, WGC, 0, ctx->stream>>>(
            PickBfsBits{b.fbits + size_t(level % 3) * b.fbw, bfs_rule}, level, g->m, b.nblk, 0, g->row_ptr,
            queue_par(level & 1), b.C);
        ctx->prof_end();
        APSPIE_HIP(hipMemsetAsync(&b.C->rebuild, 0, sizeof(int), ctx->stream));
    };
    auto expand = [&](int level) {
        if (bin_on) {
            // one relaxation launch a round (push or binning), then the
            // binned round's scan, scatter and apply (which exit otherwise)
            ctx->prof_begin(APSPIE_KCLASS_BIN);
            if (hint)
                k_relax<OpRelaxHint, true><<<b.expand_grid, WG, 0, ctx->stream>>>(
                    OpRelaxHint{g->val, dist, hint}, level, g->m, nb, b.C, queue_of(b), g->col_ind, g->val, hint,
                    bin_out, bin_tbl, bin_wcnt);
            else
                k_relax<OpRelax, false><<<b.expand_grid, WG, 0, ctx->stream>>>(
                    OpRelax{g->val, dist}, level, g->m, nb, b.C, queue_of(b), g->col_ind, g->val, nullptr, bin_out,
                    bin_tbl, bin_wcnt);
            ctx->prof_end();
            if (level > 1) {  // level 1 expands the source's round, which is never binned
                ctx->prof_begin(APSPIE_KCLASS_APPLY);
                k_bin_scan<<<nb, 256, 0, ctx->stream>>>(level, nb, bin_grid, b.C, bin_tbl, bin_tot);
                k_bin_scatter<<<bin_grid, WG, 0, ctx->stream>>>(level, nb, b.C, bin_out, bin_tbl, bin_tot, bin_wcnt,
                                                                bin_sorted);
                k_bin_apply<<<apply_grid, WGA, 0, ctx->stream>>>(level, g->m, nb, b.C, bin_sorted, bin_tot, dist);
                ctx->prof_end();
            }
        } else if (MODE == MODE_BFS) {
            launch_expand(ctx, b, g, OpBfs{b.visited, b.flags}, level);
        } else if (hint) {
            launch_expand(ctx, b, g, OpRelaxHint{g->val, dist, hint}, level);
        } else {
            launch_expand(ctx, b, g, OpRelax{g->val, dist}, level);
        }
        if (!pull) return;
        ctx->prof_begin(APSPIE_KCLASS_PULL);
        k_pull_relax<<<pull_grid, WG, 0, ctx->stream>>>(level, g->m, nnz32, b.C, g->csc_col_ptr, g->csc_row_ind,
                                                        g->csc_val, b.trows, fval, dist);
        ctx->prof_end();
    };
    // round 0: the source alone, without a pass over the vertices
    {
        const auto pk0 = pick_for(0, src);
        ctx->prof_begin(APSPIE_KCLASS_COMPACT);
        bool seeded = false;
        if constexpr (MODE == MODE_SSSP)
            if (delta > 0.0f) {
                k_seed<PickSsspNF><<<1, WG, 0, ctx->stream>>>(near_far(pk0), src, g->row_ptr, queue_of(b), b.C);
                seeded = true;
            }
        if (!seeded)
            k_seed<std::decay_t<decltype(pk0)>><<<1, WG, 0, ctx->stream>>>(pk0, src, g->row_ptr, queue_of(b), b.C);
        ctx->prof_end();
    }
    // chunks of 4, 2, 2, 4, 8, ... rounds between frontier checks: RMAT
    // traversals end within 6-12 rounds, and a round launched past the end
    // costs its kernels' launch
    int level = 1, chunk = ctx->tune.round_chunk, checks = 0;
    // the descriptors, the round counter, the rebuild mark and the totals lead
    // the control block: one copy after each chunk of rounds reads them all
    // (the totals of the last chunk are the call's, with no further copy)
    constexpr size_t LEAD = offsetof(Control, acc) + sizeof(Acc);
    static_assert(LEAD <= PINNED_BYTES, "fits the pinned staging");
    volatile Control *host_c = (volatile Control *)ctx->pinned;
    const bool bfs_rounds = MODE == MODE_BFS && pull;
    for (;;) {
        for (int i = 0; i < chunk; ++i, ++level) {
            if (bfs_rounds) {
                bfs_round(level);
                continue;
            }
            expand(level);
            compact(level);
        }
        APSPIE_HIP(hipGetLastError());
        APSPIE_HIP(hipEventRecord(ctx->ev1, ctx->stream));
        APSPIE_HIP(hipMemcpyAsync((void *)host_c, b.C, LEAD, hipMemcpyDeviceToHost, ctx->stream));
        APSPIE_HIP(hipStreamSynchronize(ctx->stream));
        if (bfs_rounds && host_c->rebuild) {  // the later launches of the chunk were void
            level = host_c->rebuild - 1;
            bfs_rebuild(level);
            ++level;
            continue;
        }
        const Ctl last{host_c->ctl[(level - 1) & 1].nF, host_c->ctl[(level - 1) & 1].E,
                       host_c->ctl[(level - 1) & 1].nTiles, host_c->ctl[(level - 1) & 1].pull};
        // done: an empty frontier with nothing waiting past a threshold
        if (last.nF == 0 && last.pull != DIR_REPICK) break;
        // every round settles a vertex or, repicking, admits a waiting one
        APSPIE_CHECK(level <= 2 * int64_t(g->m) + 4, APSPIE_ENOCONV, "traversal did not converge");
        if (bin_on && host_c->acc.binned > 0 && last.pull != DIR_BINNED) {
            bin_on = false;  // past the dense rounds: a fresh chunk of plain rounds
            relax_rounds = level - 1;
            chunk = std::max(4, ctx->tune.round_chunk);
            continue;
        }
        ++checks;
        chunk = checks <= 2 ? 2 : std::min(chunk * 2, 32);
    }
    Acc acc;
    memcpy(&acc, (const void *)&host_c->acc, sizeof(Acc));
    fill_stats(ctx, b, MODE, g->m, stats, &acc);
    if (stats && binned) stats->relax_rounds = relax_rounds >= 0 ? relax_rounds : level - 1;
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
        APSPIE_CHECK(!g->csc_col_ptr || (g->csc_val != nullptr) == (g->val != nullptr), APSPIE_EINVAL,
                     "the CSC must carry weights exactly when the CSR does");
        trav::run<trav::MODE_SSSP>(ctx, g, src, nullptr, d_dist, stats);
    });
}

int apspie_sssp_noalloc(apspie_ctx *ctx, const apspie_csr *g, int32_t src, float *d_dist, apspie_stats *stats) {
    if (!ctx) return APSPIE_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    return guarded(ctx, [&] {
        using namespace trav;
        validate_traversal(g, src);
        APSPIE_CHECK(d_dist != nullptr, APSPIE_EINVAL, "null dist");
        const size_t need = sssp_carve_bytes(g->m, g->nnz);
        APSPIE_CHECK(need <= NOALLOC_BYTES, APSPIE_ELIMIT,
                     "apspie_sssp_noalloc: the graph needs " + std::to_string(need) + " bytes of scratch, the fixed "
                     "scratch holds " + std::to_string(NOALLOC_BYTES) + " (about 20 B per vertex + 32 B per 1024 "
                     "edges); use apspie_sssp");
        APSPIE_CHECK(ctx->device < MAX_DEVICES && g_noalloc_ptr[ctx->device] != nullptr, APSPIE_EHIP,
                     "fixed scratch not set up on this device");
        std::lock_guard<std::mutex> dev(g_noalloc_mu[ctx->device]);
        const Bufs b = sssp_carve(ctx, (unsigned char *)g_noalloc_ptr[ctx->device], g->m, g->nnz);
        // per-kernel timing creates events on demand: off for this call
        const bool prof = ctx->profiling;
        ctx->profiling = false;
        try {
            run<MODE_SSSP>(ctx, g, src, nullptr, d_dist, stats, &b);
        } catch (...) {
            ctx->profiling = prof;
            throw;
        }
        ctx->profiling = prof;
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
                                                           b.flags, b.visited, b.status, b.C, ~0u,
                                                           nullptr, nullptr, nullptr);
        ctx->prof_end();
        launch_compact(ctx, b, g, PickStep{d_x, g->m}, 0, 1);
        launch_expand(ctx, b, g, OpRelax{g->val, y}, 1);
        APSPIE_HIP(hipGetLastError());
        APSPIE_HIP(hipEventRecord(ctx->ev1, ctx->stream));
        fill_stats(ctx, b, MODE_STEP, g->m, stats);
    });
}

int apspie_ctx_rounds(apspie_ctx *ctx, apspie_round *out, int max, int *n) {
    if (!ctx || !n || max < 0 || (max > 0 && !out)) return APSPIE_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    *n = 0;
    return guarded(ctx, [&] {
        using namespace trav;
        if (!ctx->slot_ptr[S_CTL]) return;
        Control c;
        APSPIE_HIP(hipMemcpyAsync(&c, ctx->slot_ptr[S_CTL], sizeof(Control), hipMemcpyDeviceToHost, ctx->stream));
        APSPIE_HIP(hipStreamSynchronize(ctx->stream));
        const int k = std::min({max, LOG_ROUNDS, std::max(c.acc.levels, 0)});
        for (int i = 0; i < k; ++i) out[i] = apspie_round{c.log[i].nF, c.log[i].pull, c.log[i].E};
        *n = k;
    });
}

}  // extern "C"
