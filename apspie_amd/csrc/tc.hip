// tc.hip -- triangle counting as masked SpGEMM, count = sum over (i,k) in L of
// |L(i,:) & L(k,:)| with L the strict lower triangle (row i holds sorted
// j < i), exact int64 result.
//
//   0. Partners: each edge (i,k) of L is one intersection, done by the row of
//      the two whose L row is longer (i when nnz(L(k,:)) <= nnz(L(i,:)), else
//      k): that row's L row is the set, the other's is scanned, so the work
//      is sum over edges of min(nnz(L(i,:)), nnz(L(k,:))) probes instead of
//      sum of nnz(L(k,:)) (2.27 G instead of 8.07 G on RMAT s20).  Row x's
//      partners (P, a CSR) are its L entries with a row no longer than its
//      own and the rows i > x holding x in L (U = the transpose of L) with a
//      shorter one.
//   1. ew[e] = nnz(L(y,:)) for every partner entry e = (x,y) of P (the wedges
//      it opens), and gpre = its int64 exclusive scan (single-pass look-back);
//      row x owns wedges [gpre[prp[x]], gpre[prp[x+1]]).
//   2. Work items (i, c) cover wedges [c*CW, (c+1)*CW) of row i (CW by bin:
//      an item rebuilds its row's set, so bins of larger sets take larger
//      items) and are binned by nnz(L(i,:)): <= 256 (one wave per item,
//      512-slot LDS set), <= 1024 / 2048 / 4096 (256 threads, 2048 / 4096 /
//      8192 slots; the smaller the set, the more workgroups share a CU, and
//      the faster they probe), larger (binary search of the sorted row in
//      global memory: a 16384-slot set leaves one workgroup a CU and probed
//      slower).
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

#include "convert.hpp"
#include "device.hpp"
#include "internal.hpp"

namespace apspie {
namespace tcx {

using namespace dev;

// wedges per work item, by bin
__host__ __device__ constexpr int cw_of(int bin) { return bin == 3 ? 65536 : bin == 1 || bin == 2 ? 32768 : 16384; }
constexpr int NB = 5;
constexpr int BIN_DEG[NB - 1] = {256, 1024, 2048, 4096};

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

// Partners, edge-parallel.  Row ids of L's entries (rowL) and of U's
// (rowU, the column of L each U entry came from), and deg = nnz(L(x,:)), one
// wave per row.
__global__ __launch_bounds__(256) void k_rowids(int m, const int32_t *__restrict__ rp, const int32_t *__restrict__ ucp,
                                                int32_t *deg, int32_t *rowL, int32_t *rowU) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int64_t nw = int64_t(gridDim.x) * (256 / WAVE);
    for (int64_t x = int64_t(blockIdx.x) * (256 / WAVE) + threadIdx.x / WAVE; x < m; x += nw) {
        if (lane == 0) deg[x] = rp[x + 1] - rp[x];
        for (int e = rp[x] + lane; e < rp[x + 1]; e += WAVE) rowL[e] = int32_t(x);
        for (int e = ucp[x] + lane; e < ucp[x + 1]; e += WAVE) rowU[e] = int32_t(x);
    }
}

// which side keeps each edge: L entry (x, k) stays with x when deg[k] <=
// deg[x] (fa); U entry (x, i) goes to x when deg[i] < deg[x] (fb).  Every
// edge of L is kept by exactly one side.
__global__ __launch_bounds__(256) void k_keep(int64_t nnz, const int32_t *__restrict__ col,
                                              const int32_t *__restrict__ rowL, const int32_t *__restrict__ uri,
                                              const int32_t *__restrict__ rowU, const int32_t *__restrict__ deg,
                                              int32_t *fa, int32_t *fb) {
    for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < nnz; e += int64_t(gridDim.x) * 256) {
        fa[e] = deg[col[e]] <= deg[rowL[e]] ? 1 : 0;
        fb[e] = deg[uri[e]] < deg[rowU[e]] ? 1 : 0;
    }
}

// prp[x] = the kept L entries of rows < x plus the kept U entries of rows
// < x (sa, sb: exclusive scans of fa, fb); also sums the plain form's
// wedges, sum over x of nnz(L(x,:)) * nnz(U(x,:)) (edge (i,x) of L opens
// nnz(L(x,:)) of them), into *wedges (one atomic per workgroup).
__global__ __launch_bounds__(256) void k_prp(int m, const int32_t *__restrict__ rp, const int32_t *__restrict__ ucp,
                                             const int32_t *__restrict__ sa, const int32_t *__restrict__ sb,
                                             int32_t *prp, unsigned long long *wedges) {
    __shared__ unsigned long long s_w[256 / WAVE];
    const int lane = threadIdx.x & (WAVE - 1);
    unsigned long long w = 0;
    for (int x = blockIdx.x * 256 + threadIdx.x; x <= m; x += gridDim.x * 256) {
        prp[x] = sa[rp[x]] + sb[ucp[x]];
        if (x < m) w += (unsigned long long)(rp[x + 1] - rp[x]) * (unsigned long long)(ucp[x + 1] - ucp[x]);
    }
    w = wave_sum(w);
    if (lane == 0) s_w[threadIdx.x / WAVE] = w;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
#pragma unroll
        for (int q = 0; q < 256 / WAVE; ++q) t += s_w[q];
        if (t) atomicAdd(wedges, t);
    }
}

// P(x,:): x's kept L entries in order, then its kept U entries
__global__ __launch_bounds__(256) void k_pfill(int64_t nnz, const int32_t *__restrict__ rp,
                                               const int32_t *__restrict__ col, const int32_t *__restrict__ rowL,
                                               const int32_t *__restrict__ ucp, const int32_t *__restrict__ uri,
                                               const int32_t *__restrict__ rowU, const int32_t *__restrict__ fa,
                                               const int32_t *__restrict__ fb, const int32_t *__restrict__ sa,
                                               const int32_t *__restrict__ sb, const int32_t *__restrict__ prp,
                                               int32_t *pcol) {
    for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < nnz; e += int64_t(gridDim.x) * 256) {
        if (fa[e]) {
            const int x = rowL[e];
            pcol[prp[x] + sa[e] - sa[rp[x]]] = col[e];
        }
        if (fb[e]) {
            const int x = rowU[e];
            const int na = sa[rp[x + 1]] - sa[rp[x]];
            pcol[prp[x] + na + sb[e] - sb[ucp[x]]] = uri[e];
        }
    }
}

// pass 0: count items per bin; pass 1: write them.  Row i's wedges are its
// partner entries' [prp[i], prp[i + 1]); its bin is its mask size nnz(L(i,:)).
__global__ __launch_bounds__(256) void k_items(int pass, int m, const int32_t *__restrict__ prp,
                                               const int32_t *__restrict__ rp, const int64_t *__restrict__ gpre,
                                               int32_t *counts, int2 *lists, const int64_t *bin_base,
                                               int32_t *bad) {
    // whole waves walk the rows together and claim each bin's slots with one
    // atomic per wave (a counter every row hits serialises otherwise)
    const int lane = threadIdx.x & (WAVE - 1);
    const int m_up = (m + WAVE - 1) / WAVE * WAVE;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < m_up; i += gridDim.x * 256) {
        int n = 0, b = -1;
        if (i < m) {
            const int64_t wi = gpre[prp[i + 1]] - gpre[prp[i]];
            if (wi >= (int64_t(1) << 31)) {  // per-row wedge offsets are int32
                *bad = 1;
            } else if (wi > 0) {
                b = bin_of(rp[i + 1] - rp[i]);
                n = int((wi + cw_of(b) - 1) / cw_of(b));
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
// one of FB sz bits per entry (index from the high bits of another
// multiplicative hash; FB = 32, or 16 for the largest sets), so at most 1 in
// 2 FB probes of an absent column passes it; most wedges do not close a
// triangle and end at one LDS read.
__device__ __forceinline__ uint32_t fslot(int j, int shift) { return (uint32_t(j) * 0x85ebca6bu) >> shift; }

// Item (i, c): the wedges [c * CW, (c + 1) * CW) of row i's partners (prp,
// pcol), each partner k's L row scanned against the set of L(i,:) (rp, col).
template <int T, int S, bool LDSMASK, int CW>
__global__ __launch_bounds__(T) void k_count(const int2 *__restrict__ items, int nitems, const int32_t *__restrict__ rp,
                                             const int32_t *__restrict__ col, const int32_t *__restrict__ prp,
                                             const int32_t *__restrict__ pcol, const int64_t *__restrict__ gpre,
                                             unsigned long long *total) {
    __shared__ int32_t s_set[LDSMASK ? S : 1];
    // the filter: FB bits per set slot (16 for the sets of 4096 slots and
    // more, whose LDS footprint sets how many workgroups share a CU)
    constexpr int FB = S >= 4096 ? 16 : 32;
    __shared__ uint32_t s_filt[LDSMASK ? S * FB / 32 : 1];
    __shared__ int32_t s_bs[T];
    __shared__ int32_t s_pre[T];
    __shared__ int32_t s_tmp[T / WAVE + 1];
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
    unsigned long long cnt = 0;
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const int i = items[it].x, c = items[it].y;
        const int mb = rp[i], me = rp[i + 1], deg = me - mb;  // the mask row L(i,:)
        const int ib = prp[i], ie = prp[i + 1];               // the partners
        const int64_t g0 = gpre[ib];
        const int wi = int(gpre[ie] - g0);
        const int lo = c * CW, hi = min(lo + CW, wi);
        uint32_t mask = 0;
        int fshift = 0;  // filter of FB sz bits: index = top log2(FB sz) bits of fslot's product
        if (LDSMASK) {
            int sz = 64;
            while (sz < 2 * deg) sz <<= 1;
            mask = uint32_t(sz - 1);
            fshift = __clz(sz) - (FB == 32 ? 4 : 3);  // 32 - log2(FB sz)
            for (int s = tid; s < sz; s += T) s_set[s] = -1;
            for (int s = tid; s < sz * FB / 32; s += T) s_filt[s] = 0u;
            __syncthreads();
            for (int p = mb + tid; p < me; p += T) {
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
                const int k = pcol[e];
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
                        int u = mb, v = me - 1;
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
        // 0. partners: U = the transpose of L, then P, edge-parallel (every
        // array in slots of its own: the transpose uses S_AUX0..5)
        auto S = [&](int k, int64_t n4) { return (int32_t *)ctx->slot(S_TC0 + k, size_t(std::max<int64_t>(n4, 1)) * 4); };
        int32_t *ucp = S(0, int64_t(m) + 1), *uri = S(1, nnz), *deg = S(2, m), *rowL = S(3, nnz), *rowU = S(4, nnz),
                *fa = S(5, nnz), *fb = S(6, nnz), *sa = S(7, nnz + 1), *sb = S(8, nnz + 1), *prp = S(9, int64_t(m) + 1),
                *pcol = S(10, nnz);
        cvt::csr2csc_into(ctx, l, ucp, uri, nullptr);
        // (the scratch slots the transpose grew are taken only now)
        int32_t *ew = (int32_t *)ctx->slot(S_AUX0, size_t(nnz + 1) * 4);
        int64_t *gpre = (int64_t *)ctx->slot(S_AUX1, size_t(nnz + 1) * 8);
        int32_t *counts = (int32_t *)ctx->slot(S_AUX3, 512);
        unsigned long long *acc = (unsigned long long *)(counts + 16);  // [0] triangles
        int64_t *bin_base = (int64_t *)(counts + 32);
        int32_t *bad = counts + 48;
        // the plain form's wedge count (stats), summed by k_prp
        unsigned long long *wcnt = (unsigned long long *)(counts + 64);
        APSPIE_HIP(hipMemsetAsync(wcnt, 0, 8, ctx->stream));
        const unsigned gr = std::max(1u, std::min(ceil_div(std::max(m, 1), 4), 16384u));
        k_rowids<<<gr, 256, 0, ctx->stream>>>(m, l->row_ptr, ucp, deg, rowL, rowU);
        const unsigned ge = cvt::grid_for(ctx, nnz, 256);
        k_keep<<<ge, 256, 0, ctx->stream>>>(nnz, l->col_ind, rowL, uri, rowU, deg, fa, fb);
        cvt::scan32(ctx, fa, sa, nnz);
        cvt::scan32(ctx, fb, sb, nnz);
        k_prp<<<std::max(1u, std::min(ceil_div(m + 1, 256), 1024u)), 256, 0, ctx->stream>>>(m, l->row_ptr, ucp, sa, sb,
                                                                                            prp, wcnt);
        k_pfill<<<ge, 256, 0, ctx->stream>>>(nnz, l->row_ptr, l->col_ind, rowL, ucp, uri, rowU, fa, fb, sa, sb, prp,
                                             pcol);
        // 1. the wedges each partner entry opens (its set probes), their
        // prefix (its look-back status words come from the slot the scans
        // above used, cleared here)
        const int nscan = std::max(1, (int)ceil_div(nnz, 256 * 16));
        uint64_t *status = (uint64_t *)ctx->slot(S_STATUS, size_t(nscan + 2) * 8);
        const unsigned g = std::max(1u, std::min(ceil_div(std::max<int64_t>(nnz, m) + 2, 256), 8192u));
        k_setup<<<g, 256, 0, ctx->stream>>>(nnz, l->row_ptr, pcol, ew, nscan, status, counts, acc, bad);
        k_scan64<<<nscan, 256, 0, ctx->stream>>>(ew, gpre, nnz, status, (unsigned *)(status + nscan));
        const unsigned gm = std::max(1u, std::min(ceil_div(std::max(m, 1), 256), 4096u));
        k_items<<<gm, 256, 0, ctx->stream>>>(0, m, prp, l->row_ptr, gpre, counts, nullptr, nullptr, bad);
        APSPIE_HIP(hipGetLastError());
        int hc[NB + 1];
        int64_t probes = 0;
        unsigned long long wedges = 0;
        APSPIE_HIP(hipMemcpyAsync(hc, counts, sizeof(int) * NB, hipMemcpyDeviceToHost, ctx->stream));
        APSPIE_HIP(hipMemcpyAsync(&hc[NB], bad, 4, hipMemcpyDeviceToHost, ctx->stream));
        APSPIE_HIP(hipMemcpyAsync(&probes, gpre + nnz, 8, hipMemcpyDeviceToHost, ctx->stream));
        APSPIE_HIP(hipMemcpyAsync(&wedges, wcnt, 8, hipMemcpyDeviceToHost, ctx->stream));
        APSPIE_HIP(hipStreamSynchronize(ctx->stream));
        APSPIE_CHECK(hc[NB] == 0, APSPIE_ELIMIT, "a row opens 2^31 or more wedges");
        int64_t hb[NB], tot = 0;
        for (int b = 0; b < NB; ++b) {
            hb[b] = tot;
            tot += hc[b];
        }
        int2 *lists = (int2 *)ctx->slot(S_AUX2, size_t(std::max<int64_t>(tot, 1)) * sizeof(int2));
        APSPIE_HIP(hipMemcpyAsync(bin_base, hb, sizeof(hb), hipMemcpyHostToDevice, ctx->stream));
        k_items<<<gm, 256, 0, ctx->stream>>>(1, m, prp, l->row_ptr, gpre, counts, lists, bin_base, bad);
        const unsigned cu = (unsigned)ctx->cu_count;
        if (hc[0])
            k_count<64, 512, true, cw_of(0)><<<std::min((unsigned)hc[0], cu * 16), 64, 0, ctx->stream>>>(
                lists + hb[0], hc[0], l->row_ptr, l->col_ind, prp, pcol, gpre, acc);
        if (hc[1])
            k_count<256, 2048, true, cw_of(1)><<<std::min((unsigned)hc[1], cu * 8), 256, 0, ctx->stream>>>(
                lists + hb[1], hc[1], l->row_ptr, l->col_ind, prp, pcol, gpre, acc);
        if (hc[2])
            k_count<256, 4096, true, cw_of(2)><<<std::min((unsigned)hc[2], cu * 8), 256, 0, ctx->stream>>>(
                lists + hb[2], hc[2], l->row_ptr, l->col_ind, prp, pcol, gpre, acc);
        if (hc[3])
            k_count<256, 8192, true, cw_of(3)><<<std::min((unsigned)hc[3], cu * 4), 256, 0, ctx->stream>>>(
                lists + hb[3], hc[3], l->row_ptr, l->col_ind, prp, pcol, gpre, acc);
        if (hc[4])
            k_count<256, 1, false, cw_of(4)><<<std::min((unsigned)hc[4], cu * 8), 256, 0, ctx->stream>>>(
                lists + hb[4], hc[4], l->row_ptr, l->col_ind, prp, pcol, gpre, acc);
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
            stats->edges = int64_t(wedges);
            stats->vertices = nnz;
            stats->improved = probes;
            stats->iterations = 1;
            stats->kernel_ms = ms;
            stats->bytes_alg = 4.0 * (double(m) + 1) + 12.0 * double(nnz) + 4.0 * double(wedges) + 8.0;
        }
    });
}
