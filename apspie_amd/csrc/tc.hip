// tc.hip -- triangle counting as masked SpGEMM, count = sum over (i,k) in L of
// |L(i,:) & L(k,:)| with L the strict lower triangle (row i holds sorted
// j < i), exact int64 result.
//
//   1. ew[e] = nnz(L(k,:)) for every edge e = (i,k) of L (the wedges it
//      opens), and gpre = its int64 exclusive scan (single-pass look-back);
//      row i owns wedges [gpre[rp[i]], gpre[rp[i+1]]).
//   2. Work items (i, c) cover wedges [c*CW, (c+1)*CW) of row i and are binned
//      by nnz(L(i,:)): <= 256 (one wave per item, 512-slot LDS set), <= 4096
//      (256 threads, 8192-slot LDS set), larger (binary search of the sorted
//      row in global memory).
//   3. Per item: the mask row L(i,:) goes into an LDS hash set behind a bit
//      filter (most probes are of absent columns and stop there); the item's
//      first edge comes from a binary search of gpre; its wedges (k, j) are
//      enumerated in blocks of entries with a block scan of nnz(L(k,:)) and a
//      binary search for each wedge's k; each j found in the set closes a
//      triangle j < k < i, counted once.  Counts reduce per wave into one
//      int64.
//
// Replaces the TC pipeline of test/tc/main.cu:148-261: cuSPARSE csrgemm of
// L*L^T materialised in full (src/tc.cuh:62-165), then ewiseMultTc's masked
// merge (src/tc.cuh:168-199) and a float mgpu::Reduce (inexact above 2^24).
#include <algorithm>
#include <cstring>

#include "device.hpp"
#include "internal.hpp"

namespace apspie {
namespace tcx {

using namespace dev;

constexpr int CW = 16384;  // wedges per work item
constexpr int NB = 3;
constexpr int BIN_DEG[NB - 1] = {256, 4096};

__global__ __launch_bounds__(256) void k_setup(int64_t nnz, const int32_t *__restrict__ rp,
                                               const int32_t *__restrict__ col, int32_t *ew, int nscan,
                                               uint64_t *status, int32_t *counts, unsigned long long *acc,
                                               int32_t *bad) {
    const int64_t stride = int64_t(gridDim.x) * 256;
    for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < nnz; e += stride) {
        const int k = col[e];
        ew[e] = rp[k + 1] - rp[k];
    }
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < nscan + 2; i += stride) status[i] = 0;
    if (blockIdx.x == 0 && threadIdx.x < 16) counts[threadIdx.x] = 0;
    if (blockIdx.x == 0 && threadIdx.x < 4) acc[threadIdx.x] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) *bad = 0;  // the slot keeps earlier calls' bytes
}

// exclusive scan of n int32 into int64 out[0..n] (out[n] = total), one pass
__global__ __launch_bounds__(256) void k_scan64(const int32_t *in, int64_t *out, int64_t n, uint64_t *status,
                                                unsigned int *counter) {
    constexpr int ITEMS = 16, TILE_N = 256 * ITEMS;
    __shared__ int s_bid;
    __shared__ uint64_t s_w[4];
    __shared__ uint64_t s_pre;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    if (tid == 0) s_bid = (int)atomicAdd(counter, 1u);
    __syncthreads();
    const int bid = s_bid;
    const int64_t b0 = int64_t(bid) * TILE_N + int64_t(tid) * ITEMS;
    uint32_t v[ITEMS];
    uint64_t sum = 0;
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        v[k] = b0 + k < n ? uint32_t(in[b0 + k]) : 0u;
        sum += v[k];
    }
    const uint64_t incl = wave_incl_scan(sum, lane);
    if (lane == 63) s_w[wid] = incl;
    __syncthreads();
    uint64_t wpre = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        if (w < wid) wpre += s_w[w];
        tot += s_w[w];
    }
    if (wid == 0) {
        const uint64_t p = lookback_exclusive(status, bid, tot, lane);
        if (lane == 0) s_pre = p;
    }
    __syncthreads();
    uint64_t run = s_pre + wpre + incl - sum;
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        if (b0 + k < n) out[b0 + k] = int64_t(run);
        run += v[k];
        if (b0 + k == n - 1) out[n] = int64_t(run);
    }
    if (n == 0 && bid == 0 && tid == 0) out[0] = 0;
}

__device__ __forceinline__ int bin_of(int deg) {
    int b = NB - 1;
#pragma unroll
    for (int k = NB - 2; k >= 0; --k)
        if (deg <= BIN_DEG[k]) b = k;
    return b;
}

// pass 0: count items per bin; pass 1: write them
__global__ __launch_bounds__(256) void k_items(int pass, int m, const int32_t *__restrict__ rp,
                                               const int64_t *__restrict__ gpre, int32_t *counts, int2 *lists,
                                               const int64_t *bin_base, int32_t *bad) {
    // whole waves walk the rows together and claim each bin's slots with one
    // atomic per wave (a counter every row hits serialises otherwise)
    const int lane = threadIdx.x & (WAVE - 1);
    const int m_up = (m + WAVE - 1) / WAVE * WAVE;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < m_up; i += gridDim.x * 256) {
        int n = 0, b = -1;
        if (i < m) {
            const int64_t wi = gpre[rp[i + 1]] - gpre[rp[i]];
            if (wi >= (int64_t(1) << 31)) {  // per-row wedge offsets are int32
                *bad = 1;
            } else if (wi > 0) {
                n = int((wi + CW - 1) / CW);
                b = bin_of(rp[i + 1] - rp[i]);
            }
        }
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            const int mine = b == k ? n : 0;
            const uint32_t incl = wave_incl_scan32(uint32_t(mine), lane);
            const uint32_t wtot = __shfl(incl, WAVE - 1, WAVE);
            if (!wtot) continue;
            int base = 0;
            if (lane == WAVE - 1) base = atomicAdd(&counts[(pass ? NB : 0) + k], int(wtot));
            base = __shfl(base, WAVE - 1, WAVE) + int(incl) - mine;
            if (pass && mine)
                for (int c = 0; c < n; ++c) lists[bin_base[k] + base + c] = make_int2(i, c);
        }
    }
}

__device__ __forceinline__ uint32_t hslot(int j, uint32_t mask) { return (uint32_t(j) * 2654435761u) & mask; }

// Bit filter in front of the LDS hash set: a row whose set has sz slots sets
// one of 32 sz bits per entry (index from the high bits of another
// multiplicative hash), so at most 1 in 64 probes of an absent column passes
// it; most wedges do not close a triangle and end at one LDS read.
__device__ __forceinline__ uint32_t fslot(int j, int shift) { return (uint32_t(j) * 0x85ebca6bu) >> shift; }

template <int T, int S, bool LDSMASK>
__global__ __launch_bounds__(T) void k_count(const int2 *__restrict__ items, int nitems, const int32_t *__restrict__ rp,
                                             const int32_t *__restrict__ col, const int64_t *__restrict__ gpre,
                                             unsigned long long *total) {
    __shared__ int32_t s_set[LDSMASK ? S : 1];
    __shared__ uint32_t s_filt[LDSMASK ? S : 1];
    __shared__ int32_t s_bs[T];
    __shared__ int32_t s_pre[T];
    __shared__ int32_t s_tmp[T / WAVE + 1];
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
    unsigned long long cnt = 0;
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const int i = items[it].x, c = items[it].y;
        const int ib = rp[i], ie = rp[i + 1], deg = ie - ib;
        const int64_t g0 = gpre[ib];
        const int wi = int(gpre[ie] - g0);
        const int lo = c * CW, hi = min(lo + CW, wi);
        uint32_t mask = 0;
        int fshift = 0;  // filter of 32 sz bits: index = top log2(32 sz) bits of fslot's product
        if (LDSMASK) {
            int sz = 64;
            while (sz < 2 * deg) sz <<= 1;
            mask = uint32_t(sz - 1);
            fshift = __clz(sz) - 4;  // 32 - log2(32 sz)
            for (int s = tid; s < sz; s += T) s_set[s] = -1;
            for (int s = tid; s < sz; s += T) s_filt[s] = 0u;
            __syncthreads();
            for (int p = ib + tid; p < ie; p += T) {
                const int j = col[p];
                uint32_t s = hslot(j, mask);
                while (atomicCAS(&s_set[s], -1, j) != -1) s = (s + 1) & mask;
                const uint32_t f = fslot(j, fshift);
                atomicOr(&s_filt[f >> 5], 1u << (f & 31));
            }
            __syncthreads();
        }
        // first edge whose wedge range reaches lo: last e with gpre[e] - g0 <= lo
        int a = ib, b = ie - 1;
        while (a < b) {
            const int mid = (a + b + 1) >> 1;
            if (int(gpre[mid] - g0) <= lo) a = mid;
            else b = mid - 1;
        }
        int base = int(gpre[a] - g0);  // wedges before the current block of entries
        for (int eb = a; eb < ie && base < hi; eb += T) {
            const int e = eb + tid;
            int L = 0, bs = 0;
            if (e < ie) {
                const int k = col[e];
                bs = rp[k];
                L = rp[k + 1] - bs;
            }
            const int incl = (int)wave_incl_scan32((uint32_t)L, lane);
            if (lane == WAVE - 1) s_tmp[wid] = incl;
            __syncthreads();
            int pre = 0, tot = 0;
#pragma unroll
            for (int q = 0; q < T / WAVE; ++q) {
                if (q < wid) pre += s_tmp[q];
                tot += s_tmp[q];
            }
            s_bs[tid] = bs;
            s_pre[tid] = base + pre + incl - L;
            __syncthreads();
            const int ne = min(T, ie - eb);
            const int t0 = max(lo, base), t1 = min(hi, base + tot);
            // U wedges per thread in flight: their entries come from one
            // binary search and a forward walk (a thread's wedges only move
            // right), their column loads are issued together
            constexpr int U = 8;
            int x = 0;
            for (int tb = t0 + tid; tb < t1; tb += U * T) {
                int j[U];
#pragma unroll
                for (int q = 0; q < U; ++q) {
                    const int t = tb + q * T;
                    j[q] = -1;
                    if (t < t1) {
                        if (q == 0) {
                            int y = ne - 1;  // last entry of the block whose prefix <= t
                            while (x < y) {
                                const int mid = (x + y + 1) >> 1;
                                if (s_pre[mid] <= t) x = mid;
                                else y = mid - 1;
                            }
                        } else {
                            while (x + 1 < ne && s_pre[x + 1] <= t) ++x;
                        }
                        j[q] = col[s_bs[x] + (t - s_pre[x])];
                    }
                }
#pragma unroll
                for (int q = 0; q < U; ++q) {
                    if (j[q] < 0) continue;
                    bool hit = false;
                    if (LDSMASK) {
                        const uint32_t f = fslot(j[q], fshift);
                        if (!((s_filt[f >> 5] >> (f & 31)) & 1u)) continue;  // absent
                        uint32_t s = hslot(j[q], mask);
                        for (;;) {
                            const int kk = s_set[s];
                            if (kk == j[q]) {
                                hit = true;
                                break;
                            }
                            if (kk == -1) break;
                            s = (s + 1) & mask;
                        }
                    } else {
                        int u = ib, v = ie - 1;
                        while (u <= v) {
                            const int mid = (u + v) >> 1;
                            const int val = col[mid];
                            if (val == j[q]) {
                                hit = true;
                                break;
                            }
                            if (val < j[q]) u = mid + 1;
                            else v = mid - 1;
                        }
                    }
                    cnt += hit ? 1u : 0u;
                }
            }
            base += tot;
            __syncthreads();
        }
    }
    cnt = wave_sum(cnt);
    if (lane == 0 && cnt) atomicAdd(total, cnt);
}

}  // namespace tcx
}  // namespace apspie

using namespace apspie;

extern "C" int apspie_tc(apspie_ctx *ctx, const apspie_csr *l, int64_t *count, apspie_stats *stats) {
    if (!ctx || !count) return APSPIE_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    return guarded(ctx, [&] {
        using namespace tcx;
        validate_csr(l, false);
        APSPIE_CHECK(l->m == l->n, APSPIE_EINVAL, "L must be square");
        const int m = l->m;
        const int64_t nnz = l->nnz;
        APSPIE_HIP(hipEventRecord(ctx->ev0, ctx->stream));
        int32_t *ew = (int32_t *)ctx->slot(S_AUX0, size_t(nnz + 1) * 4);
        int64_t *gpre = (int64_t *)ctx->slot(S_AUX1, size_t(nnz + 1) * 8);
        int32_t *counts = (int32_t *)ctx->slot(S_AUX3, 512);
        unsigned long long *acc = (unsigned long long *)(counts + 16);  // [0] triangles
        int64_t *bin_base = (int64_t *)(counts + 32);
        int32_t *bad = counts + 48;
        const int nscan = std::max(1, (int)ceil_div(nnz, 256 * 16));
        uint64_t *status = (uint64_t *)ctx->slot(S_STATUS, size_t(nscan + 2) * 8);
        const unsigned g = std::max(1u, std::min(ceil_div(std::max<int64_t>(nnz, m) + 2, 256), 8192u));
        k_setup<<<g, 256, 0, ctx->stream>>>(nnz, l->row_ptr, l->col_ind, ew, nscan, status, counts, acc, bad);
        k_scan64<<<nscan, 256, 0, ctx->stream>>>(ew, gpre, nnz, status, (unsigned *)(status + nscan));
        const unsigned gm = std::max(1u, std::min(ceil_div(std::max(m, 1), 256), 4096u));
        k_items<<<gm, 256, 0, ctx->stream>>>(0, m, l->row_ptr, gpre, counts, nullptr, nullptr, bad);
        APSPIE_HIP(hipGetLastError());
        int hc[NB + 1];
        int64_t wedges = 0;
        APSPIE_HIP(hipMemcpyAsync(hc, counts, sizeof(int) * NB, hipMemcpyDeviceToHost, ctx->stream));
        APSPIE_HIP(hipMemcpyAsync(&hc[NB], bad, 4, hipMemcpyDeviceToHost, ctx->stream));
        APSPIE_HIP(hipMemcpyAsync(&wedges, gpre + nnz, 8, hipMemcpyDeviceToHost, ctx->stream));
        APSPIE_HIP(hipStreamSynchronize(ctx->stream));
        APSPIE_CHECK(hc[NB] == 0, APSPIE_ELIMIT, "a row opens 2^31 or more wedges");
        int64_t hb[NB], tot = 0;
        for (int b = 0; b < NB; ++b) {
            hb[b] = tot;
            tot += hc[b];
        }
        int2 *lists = (int2 *)ctx->slot(S_AUX2, size_t(std::max<int64_t>(tot, 1)) * sizeof(int2));
        APSPIE_HIP(hipMemcpyAsync(bin_base, hb, sizeof(hb), hipMemcpyHostToDevice, ctx->stream));
        k_items<<<gm, 256, 0, ctx->stream>>>(1, m, l->row_ptr, gpre, counts, lists, bin_base, bad);
        const unsigned cu = (unsigned)ctx->cu_count;
        if (hc[0])
            k_count<64, 512, true><<<std::min((unsigned)hc[0], cu * 16), 64, 0, ctx->stream>>>(
                lists + hb[0], hc[0], l->row_ptr, l->col_ind, gpre, acc);
        if (hc[1])
            k_count<256, 8192, true><<<std::min((unsigned)hc[1], cu * 4), 256, 0, ctx->stream>>>(
                lists + hb[1], hc[1], l->row_ptr, l->col_ind, gpre, acc);
        if (hc[2])
            k_count<256, 1, false><<<std::min((unsigned)hc[2], cu * 8), 256, 0, ctx->stream>>>(
                lists + hb[2], hc[2], l->row_ptr, l->col_ind, gpre, acc);
        APSPIE_HIP(hipGetLastError());
        APSPIE_HIP(hipEventRecord(ctx->ev1, ctx->stream));
        unsigned long long tri = 0;
        APSPIE_HIP(hipMemcpyAsync(&tri, acc, 8, hipMemcpyDeviceToHost, ctx->stream));
        APSPIE_HIP(hipStreamSynchronize(ctx->stream));
        *count = (int64_t)tri;
        if (stats) {
            float ms = 0.f;
            APSPIE_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
            memset(stats, 0, sizeof(*stats));
            stats->edges = wedges;
            stats->vertices = nnz;
            stats->iterations = 1;
            stats->kernel_ms = ms;
            stats->bytes_alg = 4.0 * (double(m) + 1) + 12.0 * double(nnz) + 4.0 * double(wedges) + 8.0;
        }
    });
}
