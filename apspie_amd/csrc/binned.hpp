// binned.hpp -- binned relaxation of dense SSSP rounds (propagation blocking:
// Beamer, Asanovic, Patterson, "Reducing PageRank Communication via
// Propagation Blocking", IPDPS 2017), the gfx950 replacement for the random
// 4-byte distance probes and atomics a push round makes per edge.
//
//   k_bin_relax    a persistent grid of G workgroups; workgroup w owns the
//                  frontier tiles [w T / G, (w + 1) T / G) and the same stretch
//                  of the entry buffer.  Pass 1 counts its candidates per
//                  bucket (target >> BUCKET_SHIFT) in LDS; pass 2 writes every
//                  candidate (target, d_u + w) into its bucket's run inside the
//                  stretch.  One row of G x (nb + 1) run offsets per
//                  workgroup; per-bucket totals for the apply's split.  All
//                  global traffic is streams (the column ids are read twice).
//   k_apply_relax  one workgroup per (bucket, part): the bucket's slice of the
//                  distance array (BUCKET vertices, 64 KB) sits in LDS, the
//                  bucket's runs of the part's bin workgroups are folded in
//                  with LDS atomicMin, and the improved distances are written
//                  back (atomicMin when other parts share the bucket).  Heavy
//                  buckets (the RMAT hubs) are split into parts of about
//                  PART_ENTRIES candidates.
//
// The round computes exactly what the push round does (every candidate of the
// frontier snapshot, min-reduced into the live distances), so the schedule,
// the counters and the result are the same.
#pragma once

#include "frontier.hpp"

namespace apspie {
namespace trav {

constexpr int BUCKET_SHIFT = 14;
constexpr int BUCKET = 1 << BUCKET_SHIFT;  // vertices per bucket (64 KB of distances)
constexpr int MAX_BUCKETS = 512;           // binned rounds need m <= MAX_BUCKETS * BUCKET
constexpr int WGA = 1024;                  // threads per apply workgroup
constexpr int PART_ENTRIES = 1 << 18;      // candidates per apply workgroup (heavy buckets split)

// Frontier tiles of bin workgroup w among G: [tile_lo(w), tile_lo(w + 1)).
__device__ __forceinline__ int bin_tile_lo(int w, int G, int ntiles) { return int(int64_t(w) * ntiles / G); }

// Entry buffer: >= ntiles * TILE uint2 (target, candidate bits); table: G *
// (nb + 1) int32 run offsets; totals: 2 parities * MAX_BUCKETS.
__global__ __launch_bounds__(WG) void k_bin_relax(int level_arg, int m, int nb, const Control *C, Queue q,
                                                  const int32_t *__restrict__ col_ind,
                                                  const float *__restrict__ val, uint2 *__restrict__ out,
                                                  int32_t *__restrict__ tbl, uint32_t *totals) {
    constexpr int NW = WG / WAVE, SEG = TILE / NW;
    __shared__ int32_t s_own[TILE];
    __shared__ int32_t s_base[TILE + 1];
    __shared__ uint32_t s_val[TILE + 1];
    __shared__ int32_t s_vert[1];
    __shared__ int32_t s_wmax[NW];
    __shared__ uint32_t s_hist[MAX_BUCKETS];  // counts, then fill cursors
    __shared__ uint32_t s_off[MAX_BUCKETS + 1];
    __shared__ uint32_t s_tcnt[MAX_BUCKETS];   // pass 2: this tile's candidates per bucket
    __shared__ uint32_t s_toff[MAX_BUCKETS];   // pass 2: their offset in the tile's sorted order
    __shared__ uint32_t s_gpos[MAX_BUCKETS];   // pass 2: their position in the entry buffer
    __shared__ uint2 s_sorted[TILE];           // pass 2: the tile's candidates sorted by bucket
    __shared__ uint32_t s_tn;
    const int level = level_arg >= 0 ? level_arg : C->cur_level;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
    // every round clears the other parity's totals for the next round (their
    // last reader, an earlier apply, has finished); the first two rounds'
    // totals start cleared by the traversal's initialisation
    if (blockIdx.x == 0)
        for (int b = tid; b < MAX_BUCKETS; b += WG) totals[((level & 1) ^ 1) * MAX_BUCKETS + b] = 0;
    const Round cur = round_of(C, level);
    if (cur.pull != DIR_BINNED) return;
    const int G = gridDim.x, w = blockIdx.x;
    const int t0 = bin_tile_lo(w, G, cur.nTiles), t1 = bin_tile_lo(w + 1, G, cur.nTiles);
    for (int b = tid; b < nb; b += WG) s_hist[b] = 0;
    // pass 1: count
    for (int t = t0; t < t1; ++t) {
        int own[IPT];
        const int elim = tile_owners<false, false>(t, cur, q, s_own, s_base, s_val, s_vert, s_wmax, own);
        int col[IPT];
#pragma unroll
        for (int r = 0; r < IPT; ++r) {
            const int s = wid * SEG + r * WAVE + lane;
            col[r] = s < elim ? col_ind[s_base[own[r]] + s] : -1;
        }
#pragma unroll
        for (int r = 0; r < IPT; ++r)
            if ((unsigned)col[r] < (unsigned)m) atomicAdd(&s_hist[col[r] >> BUCKET_SHIFT], 1u);
        __syncthreads();  // the next tile_owners rewrites the staging
    }
    __syncthreads();
    if (wid == 0) {
        uint32_t carry = 0;
        for (int b0 = 0; b0 < nb; b0 += WAVE) {
            const int b = b0 + lane;
            const uint32_t h = b < nb ? s_hist[b] : 0u;
            const uint32_t incl = wave_incl_scan32(h, lane);
            if (b < nb) s_off[b] = carry + incl - h;
            carry += __shfl(incl, WAVE - 1, WAVE);
        }
        if (lane == 0) s_off[nb] = carry;
    }
    __syncthreads();
    int32_t *row = tbl + size_t(w) * (nb + 1);
    uint32_t *tot = totals + (level & 1) * MAX_BUCKETS;
    for (int b = tid; b <= nb; b += WG) row[b] = int32_t(s_off[b]);
    for (int b = tid; b < nb; b += WG) {
        if (s_hist[b]) atomicAdd(&tot[b], s_hist[b]);
        s_hist[b] = 0;
    }
    __syncthreads();
    // pass 2: sort each tile's candidates by bucket in LDS, then write them
    // out in that order, so a wave's stores land as runs of its buckets
    uint2 *dst = out + size_t(t0) * TILE;
    for (int b = tid; b < nb; b += WG) s_tcnt[b] = 0;
    for (int t = t0; t < t1; ++t) {
        int own[IPT];
        const int elim = tile_owners<true, false>(t, cur, q, s_own, s_base, s_val, s_vert, s_wmax, own);
        int pos[IPT], col[IPT];
#pragma unroll
        for (int r = 0; r < IPT; ++r) {
            const int s = wid * SEG + r * WAVE + lane;
            pos[r] = s < elim ? s_base[own[r]] + s : -1;
        }
#pragma unroll
        for (int r = 0; r < IPT; ++r) col[r] = pos[r] >= 0 ? col_ind[pos[r]] : -1;
        float wt[IPT];
#pragma unroll
        for (int r = 0; r < IPT; ++r) wt[r] = (val && (unsigned)col[r] < (unsigned)m) ? val[pos[r]] : 1.0f;
        uint32_t cand[IPT], rank[IPT];
#pragma unroll
        for (int r = 0; r < IPT; ++r)
            if ((unsigned)col[r] < (unsigned)m) {
                cand[r] = __float_as_uint(__uint_as_float(s_val[own[r]]) + wt[r]);
                rank[r] = atomicAdd(&s_tcnt[col[r] >> BUCKET_SHIFT], 1u);
            }
        __syncthreads();
        if (wid == 0) {  // tile offsets and reserved buffer positions per bucket
            uint32_t carry = 0;
            for (int b0 = 0; b0 < nb; b0 += WAVE) {
                const int b = b0 + lane;
                const uint32_t h = b < nb ? s_tcnt[b] : 0u;
                const uint32_t incl = wave_incl_scan32(h, lane);
                if (b < nb) {
                    s_toff[b] = carry + incl - h;
                    s_gpos[b] = s_off[b] + s_hist[b];
                    s_hist[b] += h;
                    s_tcnt[b] = 0;  // ready for the next tile
                }
                carry += __shfl(incl, WAVE - 1, WAVE);
            }
            if (lane == 0) s_tn = carry;
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < IPT; ++r)
            if ((unsigned)col[r] < (unsigned)m)
                s_sorted[s_toff[col[r] >> BUCKET_SHIFT] + rank[r]] = make_uint2(uint32_t(col[r]), cand[r]);
        __syncthreads();
        const int tn = int(s_tn);
        for (int i = tid; i < tn; i += WG) {
            const uint2 e = s_sorted[i];
            const int b = int(e.x >> BUCKET_SHIFT);
            dst[s_gpos[b] + (uint32_t(i) - s_toff[b])] = e;
        }
        __syncthreads();
    }
}

// Workgroup (bucket b, part p of S_b): S_b = ceil(total_b / PART_ENTRIES),
// parts split the G bin workgroups' runs evenly.  Launch at least
// nb + nnz / PART_ENTRIES + 1 workgroups; the surplus exits.
__global__ __launch_bounds__(WGA) void k_apply_relax(int level_arg, int m, int nb, int G, const Control *C,
                                                     const uint2 *__restrict__ out,
                                                     const int32_t *__restrict__ tbl,
                                                     const uint32_t *__restrict__ totals, uint32_t *dist) {
    constexpr int NWA = WGA / WAVE;
    __shared__ uint32_t s_d[BUCKET];
    __shared__ int32_t s_first[MAX_BUCKETS + 1];  // first workgroup of each bucket
    __shared__ int32_t s_pre[WGA + 1];            // exclusive prefix of the part's run lengths
    __shared__ int64_t s_at[WGA];                 // the runs' first entries
    __shared__ int32_t s_wsum[NWA];
    const int level = level_arg >= 0 ? level_arg : C->cur_level;
    const Ctl cur = C->ctl[(level - 1) & 1];
    if (cur.pull != DIR_BINNED) return;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
    const uint32_t *tot = totals + (level & 1) * MAX_BUCKETS;
    if (wid == 0) {
        int carry = 0;
        for (int b0 = 0; b0 < nb; b0 += WAVE) {
            const int b = b0 + lane;
            const uint32_t t = b < nb ? tot[b] : 0u;
            const uint32_t parts = (t + PART_ENTRIES - 1) / PART_ENTRIES;
            const uint32_t incl = wave_incl_scan32(parts, lane);
            if (b < nb) s_first[b] = carry + int(incl - parts);
            carry += int(__shfl(incl, WAVE - 1, WAVE));
        }
        if (lane == 0) s_first[nb] = carry;
    }
    __syncthreads();
    const int item = blockIdx.x;
    if (item >= s_first[nb]) return;  // uniform
    int lo = 0, hi = nb;              // last bucket whose first workgroup <= item
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (s_first[mid] <= item) lo = mid;
        else hi = mid;
    }
    const int b = lo;
    const int nparts = s_first[b + 1] - s_first[b];
    const int part = item - s_first[b];
    const int w0 = int(int64_t(part) * G / nparts), w1 = int(int64_t(part + 1) * G / nparts);
    const int v0 = b << BUCKET_SHIFT;
    const int vn = min(BUCKET, m - v0);
    for (int i = tid; i < vn; i += WGA) s_d[i] = dist[v0 + i];
    __syncthreads();
    // thread i takes the run of bin workgroup w0 + i (G <= WGA); the
    // workgroup walks the concatenation of the runs, a candidate finding its
    // run by binary search over the runs' prefix
    const int nr = w1 - w0;
    int len = 0;
    int64_t at = 0;
    if (tid < nr) {
        const int32_t *row = tbl + size_t(w0 + tid) * (nb + 1);
        const int r0 = row[b];
        len = row[b + 1] - r0;
        at = int64_t(bin_tile_lo(w0 + tid, G, cur.nTiles)) * TILE + r0;
    }
    const uint32_t incl = wave_incl_scan32(uint32_t(len), lane);
    if (lane == WAVE - 1) s_wsum[wid] = int(incl);
    __syncthreads();
    int wpre = 0, total = 0;
#pragma unroll
    for (int x = 0; x < NWA; ++x) {
        if (x < wid) wpre += s_wsum[x];
        total += s_wsum[x];
    }
    s_pre[tid] = wpre + int(incl) - len;
    s_at[tid] = at;
    __syncthreads();
    for (int j = tid; j < total; j += WGA) {
        int a = 0, z = nr - 1;  // last run whose prefix <= j
        while (a < z) {
            const int mid = (a + z + 1) >> 1;
            if (s_pre[mid] <= j) a = mid;
            else z = mid - 1;
        }
        const uint2 e = out[s_at[a] + (j - s_pre[a])];
        uint32_t *slot = &s_d[e.x - uint32_t(v0)];
        if (e.y < *slot) atomicMin(slot, e.y);
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
