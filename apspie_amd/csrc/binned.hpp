// binned.hpp -- binned relaxation of dense SSSP rounds (propagation
// blocking: Beamer, Asanovic, Patterson, "Reducing PageRank Communication via
// Propagation Blocking", IPDPS 2017), which replaces the random 4-byte
// distance probes and atomics a push round makes per edge with streams.
//
//   bin_tiles      (in k_relax, traverse.hip, the round's one relaxation
//                  launch) G workgroups over the round's tiles (t = w, w + G, ...).
//                  Each edge's candidate d_u + w is kept when it is below its
//                  target's bound code (all are kept without codes); the kept
//                  (target, candidate) pairs are appended, wave-compacted, to
//                  the workgroup's own region of the candidate buffer (no
//                  global atomics), and counted per bucket (target >>
//                  BUCKET_SHIFT) in LDS; the counts go to a bucket-major
//                  table row.
//   k_bin_scan     one workgroup per bucket: exclusive scan of the bucket's
//                  counts over the G workgroups, and the bucket's total.
//   k_bin_scatter  the G regions again, in chunks sorted by bucket in LDS,
//                  written to their places in the bucket-ordered buffer (one
//                  contiguous run per bucket a chunk).
//   k_bin_apply    one workgroup per (bucket, part): the bucket's slice of the
//                  distance array (BUCKET vertices, 64 KB) sits in LDS, the
//                  part's candidates are folded in with LDS atomicMin, and
//                  the improved distances are written back (atomicMin when
//                  other parts share the bucket).  Heavy buckets (the RMAT
//                  hubs) are split into parts of about PART_ENTRIES.
//
// The round computes exactly what the push round does (every candidate of the
// frontier snapshot, min-reduced into the distances; a candidate at or past
// its target's bound could not lower it), so the schedule, the counters and
// the result are the same.
#pragma once

#include "frontier.hpp"
#include "hint.hpp"

namespace apspie {
namespace trav {

constexpr int BUCKET_SHIFT = 14;
constexpr int BUCKET = 1 << BUCKET_SHIFT;  // vertices per bucket (64 KB of distances)
constexpr int MAX_BUCKETS = 512;           // binned rounds need m <= MAX_BUCKETS * BUCKET
constexpr int WGA = 1024;                  // threads per apply workgroup
constexpr int PART_ENTRIES = 1 << 18;      // candidates per apply workgroup (heavy buckets split)
constexpr int SCH = 4096;                  // entries per scatter chunk (16 per thread)

// Entries of each bin workgroup's region for a round of ntiles tiles over G
// workgroups: a workgroup expands at most ceil(ntiles / G) tiles.  Region w
// starts at w * bin_cap.
__device__ __forceinline__ int64_t bin_cap(int ntiles, int G) { return int64_t((ntiles + G - 1) / G) * TILE; }

// Exclusive scan of x[0..n) (n <= MAX_BUCKETS) by one full wave into dst;
// returns the total.
__device__ __forceinline__ uint32_t bucket_scan(const int32_t *x, int n, uint32_t *dst, int lane) {
    uint32_t carry = 0;
    for (int b0 = 0; b0 < n; b0 += WAVE) {
        const int b = b0 + lane;
        const uint32_t h = b < n ? uint32_t(x[b]) : 0u;
        const uint32_t incl = wave_incl_sum_dpp(h);
        if (b < n) dst[b] = carry + incl - h;
        carry += uint32_t(wave_last(int(incl)));
    }
    return carry;
}

// The binning expansion of a binned round by workgroup w = blockIdx.x of G =
// gridDim.x.  Candidate regions: G * bin_cap(nTiles, G) uint2 (target,
// candidate bits); table: nb * G int32 per-bucket counts (bucket-major);
// wcount: G int32.  LDS: the tile staging (s_own[TILE], s_base[TILE + 1],
// s_val[TILE + 1], s_vert[1], s_wmax[WG / WAVE]), s_hist[MAX_BUCKETS] and
// s_wcnt[WG / WAVE].
template <bool HINT>
__device__ __forceinline__ void bin_tiles(const Round &cur, int m, int nb, const Queue &q,
                                          const int32_t *__restrict__ col_ind, const float *__restrict__ val,
                                          const uint8_t *__restrict__ hint, uint2 *__restrict__ out,
                                          int32_t *__restrict__ tbl, int32_t *__restrict__ wcount, int32_t *s_own,
                                          int32_t *s_base, uint32_t *s_val, int32_t *s_vert, int32_t *s_wmax,
                                          uint32_t *s_hist, uint32_t *s_wcnt) {
    constexpr int NW = WG / WAVE, SEG = TILE / NW;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
    const int G = gridDim.x, w = blockIdx.x;
    uint2 *reg = out + size_t(w) * size_t(bin_cap(cur.nTiles, G));
    for (int b = tid; b < nb; b += WG) s_hist[b] = 0;
    __syncthreads();
    int cursor = 0;  // the same in every thread
    for (int t = w; t < cur.nTiles; t += G) {
        int own[IPT];
        const int elim = tile_owners<true, false>(t, cur, q, s_own, s_base, s_val, s_vert, s_wmax, own);
        int pos[IPT], col[IPT];
#pragma unroll
        for (int r = 0; r < IPT; ++r) {
            const int s = wid * SEG + r * WAVE + lane;
            pos[r] = s < elim ? s_base[own[r]] + s : -1;
        }
        float wt[IPT];  // loaded beside the columns: it needs only the edge's position
#pragma unroll
        for (int r = 0; r < IPT; ++r) {
            col[r] = pos[r] >= 0 ? col_ind[pos[r]] : -1;
            wt[r] = (val && pos[r] >= 0) ? val[pos[r]] : 1.0f;
        }
        uint32_t hc[IPT];
#pragma unroll
        for (int r = 0; r < IPT; ++r) hc[r] = (HINT && (unsigned)col[r] < (unsigned)m) ? uint32_t(hint[col[r]]) : 255u;
        uint32_t cand[IPT];
        uint32_t keep = 0;
#pragma unroll
        for (int r = 0; r < IPT; ++r) {
            cand[r] = __float_as_uint(__uint_as_float(s_val[own[r]]) + wt[r]);
            if ((unsigned)col[r] < (unsigned)m && cand[r] < hint_bound(hc[r])) keep |= 1u << r;
        }
        const uint32_t wn = wave_total32(uint32_t(__popc(keep)));
        if (lane == 0) s_wcnt[wid] = wn;
        __syncthreads();
        int wbase = 0, tn = 0;
#pragma unroll
        for (int x = 0; x < NW; ++x) {
            if (x < wid) wbase += int(s_wcnt[x]);
            tn += int(s_wcnt[x]);
        }
        int at = cursor + wbase;
#pragma unroll
        for (int r = 0; r < IPT; ++r) {
            const bool k = (keep >> r) & 1u;
            const uint64_t ball = __ballot(k);
            if (k) {
                const int rank =
                    __builtin_amdgcn_mbcnt_hi(uint32_t(ball >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(ball), 0));
                reg[at + rank] = make_uint2(uint32_t(col[r]), cand[r]);
                atomicAdd(&s_hist[col[r] >> BUCKET_SHIFT], 1u);
            }
            at += __popcll(ball);
        }
        cursor += tn;
        __syncthreads();  // s_wcnt and the tile staging are rewritten next tile
    }
    for (int b = tid; b < nb; b += WG) tbl[size_t(b) * G + w] = int32_t(s_hist[b]);
    if (tid == 0) wcount[w] = cursor;
}

// Bucket b's counts over the G workgroups become their exclusive offsets
// inside the bucket; tot[b] gets the bucket's total.  Grid: nb workgroups.
__global__ __launch_bounds__(256) void k_bin_scan(int level_arg, int nb, int G, const Control *C,
                                                  int32_t *__restrict__ tbl, int32_t *__restrict__ tot) {
    __shared__ uint32_t s_w[4];
    const int level = level_arg >= 0 ? level_arg : C->cur_level;
    if (C->ctl[(level - 1) & 1].pull != DIR_BINNED) return;
    const int b = blockIdx.x;
    if (b >= nb) return;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
    int32_t *row = tbl + size_t(b) * G;
    uint32_t carry = 0;
    for (int i0 = 0; i0 < G; i0 += 256) {
        const int i = i0 + tid;
        const uint32_t x = i < G ? uint32_t(row[i]) : 0u;
        const uint32_t incl = wave_incl_sum_dpp(x);
        if (lane == WAVE - 1) s_w[wid] = incl;
        __syncthreads();
        uint32_t wpre = 0, all = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (k < wid) wpre += s_w[k];
            all += s_w[k];
        }
        if (i < G) row[i] = int32_t(carry + wpre + incl - x);
        carry += all;
        __syncthreads();
    }
    if (tid == 0) tot[b] = int32_t(carry);
}

// Region w's candidates, SCH at a time, sorted by bucket in LDS and written
// to bucket b's run at (first entry of b) + (w's offset inside b) + (what w
// already wrote to b).  Grid: the G workgroups of the round's k_relax.
__global__ __launch_bounds__(WG) void k_bin_scatter(int level_arg, int nb, const Control *C,
                                                    const uint2 *__restrict__ out, const int32_t *__restrict__ tbl,
                                                    const int32_t *__restrict__ tot,
                                                    const int32_t *__restrict__ wcount, uint2 *__restrict__ sorted) {
    constexpr int K = SCH / WG;
    __shared__ uint32_t s_cur[MAX_BUCKETS];  // where this workgroup's next entry of each bucket goes
    __shared__ uint32_t s_cnt[MAX_BUCKETS];  // this chunk's entries per bucket
    __shared__ uint32_t s_off[MAX_BUCKETS];  // their first place in the chunk's sorted order
    __shared__ uint2 s_sorted[SCH];
    const int level = level_arg >= 0 ? level_arg : C->cur_level;
    const Ctl cur = C->ctl[(level - 1) & 1];
    if (cur.pull != DIR_BINNED) return;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
    const int G = gridDim.x, w = blockIdx.x;
    if (wid == 0) bucket_scan(tot, nb, s_cur, lane);
    for (int b = tid; b < nb; b += WG) s_cnt[b] = 0;
    __syncthreads();
    for (int b = tid; b < nb; b += WG) s_cur[b] += uint32_t(tbl[size_t(b) * G + w]);
    const int n = wcount[w];
    const uint2 *reg = out + size_t(w) * size_t(bin_cap(cur.nTiles, G));
    for (int c0 = 0; c0 < n; c0 += SCH) {
        const int cn = min(SCH, n - c0);
        uint2 e[K];
        uint32_t rk[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int i = k * WG + tid;
            if (i < cn) {
                e[k] = reg[c0 + i];
                rk[k] = atomicAdd(&s_cnt[e[k].x >> BUCKET_SHIFT], 1u);
            }
        }
        __syncthreads();
        if (wid == 0) bucket_scan((const int32_t *)s_cnt, nb, s_off, lane);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int i = k * WG + tid;
            if (i < cn) s_sorted[s_off[e[k].x >> BUCKET_SHIFT] + rk[k]] = e[k];
        }
        __syncthreads();
        for (int i = tid; i < cn; i += WG) {
            const uint2 x = s_sorted[i];
            const int b = int(x.x >> BUCKET_SHIFT);
            sorted[s_cur[b] + (uint32_t(i) - s_off[b])] = x;
        }
        __syncthreads();
        for (int b = tid; b < nb; b += WG) {
            s_cur[b] += s_cnt[b];
            s_cnt[b] = 0;
        }
        __syncthreads();
    }
}

// Workgroup (bucket b, part p of S_b): S_b = ceil(tot[b] / PART_ENTRIES),
// parts split the bucket's run evenly.  Launch at least nb + nnz /
// PART_ENTRIES + 1 workgroups; the surplus exits.
__global__ __launch_bounds__(WGA) void k_bin_apply(int level_arg, int m, int nb, const Control *C,
                                                   const uint2 *__restrict__ sorted, const int32_t *__restrict__ tot,
                                                   uint32_t *dist) {
    __shared__ uint32_t s_d[BUCKET];
    __shared__ int32_t s_parts[MAX_BUCKETS];
    __shared__ uint32_t s_first[MAX_BUCKETS + 1];  // first workgroup of each bucket
    __shared__ uint32_t s_base[MAX_BUCKETS + 1];   // first entry of each bucket
    const int level = level_arg >= 0 ? level_arg : C->cur_level;
    if (C->ctl[(level - 1) & 1].pull != DIR_BINNED) return;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
    for (int b = tid; b < nb; b += WGA) s_parts[b] = (tot[b] + PART_ENTRIES - 1) / PART_ENTRIES;
    __syncthreads();
    if (wid == 0) s_first[nb] = bucket_scan(s_parts, nb, s_first, lane);
    if (wid == 1) s_base[nb] = bucket_scan(tot, nb, s_base, lane);
    __syncthreads();
    const int item = blockIdx.x;
    if (item >= int(s_first[nb])) return;  // uniform
    int lo = 0, hi = nb;                   // last bucket whose first workgroup <= item
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (int(s_first[mid]) <= item) lo = mid;
        else hi = mid;
    }
    const int b = lo;
    const int nparts = s_parts[b];
    const int part = item - int(s_first[b]);
    const int64_t tb = tot[b];
    const int64_t e0 = s_base[b] + tb * part / nparts, e1 = s_base[b] + tb * (part + 1) / nparts;
    const int v0 = b << BUCKET_SHIFT;
    const int vn = min(BUCKET, m - v0);
    for (int i = tid; i < vn; i += WGA) s_d[i] = dist[v0 + i];
    __syncthreads();
    // eight loads in flight per thread before their LDS updates
    constexpr int U = 8;
    for (int64_t j0 = e0; j0 < e1; j0 += int64_t(U) * WGA) {
        uint2 e[U];
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const int64_t j = j0 + k * WGA + tid;
            e[k] = j < e1 ? sorted[j] : make_uint2(uint32_t(v0), ~0u);  // ~0u lowers nothing
        }
#pragma unroll
        for (int k = 0; k < U; ++k) {
            uint32_t *slot = &s_d[e[k].x - uint32_t(v0)];
            if (e[k].y < *slot) atomicMin(slot, e[k].y);
        }
    }
    __syncthreads();
    for (int i = tid; i < vn; i += WGA) {
        const uint32_t nv = s_d[i];
        if (nv < dist[v0 + i]) {
            if (nparts == 1)
                dist[v0 + i] = nv;
            else
                __hip_atomic_fetch_min(&dist[v0 + i], nv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

}  // namespace trav
}  // namespace apspie
