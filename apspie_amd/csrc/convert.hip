// convert.hip -- sparse format conversions on the device.
//
//   apspie_coo_to_csr   row offsets of row-sorted COO row indices
//                       (cusparseXcoo2csr in coo2csr, src/util.cuh:456-471)
//   apspie_csr2csc      the transpose (CSC) of a CSR, rows ascending inside
//                       each column, values moved with them
//                       (cusparseScsr2csc in csr2csc, src/util.cuh:473-511)
//
// csr2csc is a stable least-significant-digit radix sort of the entries by
// column, 8 bits per pass: entries enter in CSR order (rows ascending), so a
// stable sort by column leaves each column's rows ascending.  A pass is a
// per-tile digit histogram, one exclusive scan of the histograms (digit-major,
// so every tile's range for a digit follows the earlier tiles'), and a
// scatter in which each wave ranks its lanes among equal digits with eight
// ballots and the waves of a tile add up their counts in LDS, round by round
// in item order, so equal digits keep their order.  The sorted columns give
// the column offsets the same way coo2csr gives row offsets; row indices and
// values are gathered through the original entry positions.
#include <algorithm>
#include <cstring>
#include <vector>

#include "convert.hpp"
#include "device.hpp"
#include "internal.hpp"

namespace apspie {
namespace cvt {

using namespace dev;

constexpr int T = 256, ITEMS = 16, TILE_N = T * ITEMS, RADIX = 256;

// off[r] = first position e with key[e] >= r, r in [0, m]; keys must be
// non-decreasing and in [0, m) (else *bad = 1: unsorted 1, out of range 2)
__global__ __launch_bounds__(256) void k_compress(const int32_t *__restrict__ key, int64_t n, int m, int32_t *off,
                                                  int32_t *bad) {
    for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e <= n; e += int64_t(gridDim.x) * 256) {
        int prev = e > 0 ? key[e - 1] : -1;
        int cur = e < n ? key[e] : m;
        if (e < n && (cur < 0 || cur >= m)) {
            *bad = 2;
            continue;
        }
        if (e > 0 && (prev < 0 || prev >= m)) continue;  // reported by its own thread
        if (cur < prev) {
            *bad = 1;
            continue;
        }
        for (int r = prev + 1; r <= cur; ++r) off[r] = int32_t(e);
    }
}

// exclusive scan of n int32 into out[0..n] (out[n] = total), one pass; O is
// int32 (totals below 2^31) or int64 (each thread then sums 16 consecutive
// items in 64 bits)
template <class O>
__global__ __launch_bounds__(256) void k_scan(const int32_t *in, O *out, int64_t n, uint64_t *status,
                                              unsigned int *counter) {
    __shared__ int s_bid;
    __shared__ uint64_t s_w[4];
    __shared__ uint64_t s_pre;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    if (tid == 0) s_bid = (int)atomicAdd(counter, 1u);
    __syncthreads();
    const int bid = s_bid;
    if constexpr (sizeof(O) == 4) {
        // 32-bit totals: wave w takes the tile's items [w * 1024, (w + 1) *
        // 1024) as 16 rows of 64 (coalesced loads and stores), each row
        // scanned on the VALU (DPP) with the rows before it carried in
        const int64_t wb = int64_t(bid) * TILE_N + int64_t(wid) * (TILE_N / 4);
        uint32_t v[ITEMS], ex[ITEMS];
        uint32_t carry = 0;
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) {
            const int64_t e = wb + k * WAVE + lane;
            v[k] = e < n ? uint32_t(in[e]) : 0u;
            const uint32_t incl = wave_incl_sum_dpp(v[k]);
            ex[k] = carry + incl - v[k];
            carry += uint32_t(wave_last(int(incl)));
        }
        if (lane == 0) s_w[wid] = carry;
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
        const uint32_t base = uint32_t(s_pre + wpre);
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) {
            const int64_t e = wb + k * WAVE + lane;
            if (e < n) out[e] = O(base + ex[k]);
            if (e == n - 1) out[n] = O(base + ex[k] + v[k]);
        }
        if (n == 0 && bid == 0 && tid == 0) out[0] = 0;
        return;
    }
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
        if (b0 + k < n) out[b0 + k] = O(run);
        run += v[k];
        if (b0 + k == n - 1) out[n] = O(run);
    }
    if (n == 0 && bid == 0 && tid == 0) out[0] = 0;
}

__global__ __launch_bounds__(256) void k_zero64(uint64_t *p, int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) p[i] = 0;
}

// key[e] = col[e] (the first sort pass takes idx[e] = e itself)
__global__ __launch_bounds__(256) void k_seed(const int32_t *__restrict__ col, int64_t n, int32_t *key) {
    for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < n; e += int64_t(gridDim.x) * 256) key[e] = col[e];
}

// hist[d * ntiles + tile] = items of the tile with digit d.  Each wave counts
// its quarter of the tile into its own LDS row (a quarter of the atomics
// per counter), full tiles read 16-byte vectors (the key arrays are
// scratch-pool slots, so tiles of 4096 keys start 16-byte aligned).
__global__ __launch_bounds__(T) void k_hist(const int32_t *__restrict__ key, int64_t n, int shift, int ntiles,
                                            int32_t *hist) {
    constexpr int NW = T / WAVE;
    __shared__ int32_t s_h[NW][RADIX];
    const int tid = threadIdx.x, wid = tid / WAVE;
    auto count = [&](int k) { atomicAdd(&s_h[wid][(uint32_t(k) >> shift) & (RADIX - 1)], 1); };
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
#pragma unroll
        for (int w = 0; w < NW; ++w) s_h[w][tid] = 0;
        __syncthreads();
        const int64_t t0 = int64_t(tile) * TILE_N;
        if (t0 + TILE_N <= n) {
            const int4 *kv = reinterpret_cast<const int4 *>(key + t0);
#pragma unroll
            for (int j = 0; j < ITEMS / 4; ++j) {
                const int4 q = kv[j * T + tid];
                count(q.x);
                count(q.y);
                count(q.z);
                count(q.w);
            }
        } else {
            for (int j = 0; j < ITEMS; ++j) {
                const int64_t t = t0 + j * T + tid;
                if (t < n) count(key[t]);
            }
        }
        __syncthreads();
        int c = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) c += s_h[w][tid];
        hist[int64_t(tid) * ntiles + tile] = c;
        __syncthreads();
    }
}

// stable scatter of one pass: off[d * ntiles + tile] = where the tile's
// first item of digit d goes.  Wave w of a workgroup owns the tile's items
// [w * 1024, (w + 1) * 1024) and keeps them in registers: it counts their
// digits into its own LDS row, one thread per digit turns the rows into each
// wave's tile-local starting position per digit (the digits before d, then
// digit d in the waves before w), and the wave walks its items in order,
// ranking lanes of equal digit with eight ballots and advancing its row (the
// LDS operations of one wave execute in order, so no barrier is needed
// between rounds).  The items land in LDS in the tile's digit order; then
// consecutive threads write consecutive items, so each digit's run of the
// tile goes out as one contiguous store instead of one line per lane.
// Equal digits keep their input order.
__global__ __launch_bounds__(T) void k_scatter(const int32_t *__restrict__ key, const int32_t *__restrict__ idx,
                                               bool ident, int64_t n, int shift, int ntiles,
                                               const int32_t *__restrict__ off, int32_t *key_out, int32_t *idx_out) {
    constexpr int NW = T / WAVE, PER = TILE_N / NW, R = PER / WAVE;
    static_assert(RADIX == T, "one thread per digit");
    __shared__ int32_t s_cnt[NW][RADIX];
    __shared__ int32_t s_goff[RADIX];  // global position of the tile's digit-d run, less its tile-local start
    __shared__ int32_t s_wsum[NW];
    __shared__ int32_t s_k[TILE_N], s_x[TILE_N];
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
    const uint64_t lt = (1ull << lane) - 1;
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
#pragma unroll
        for (int w = 0; w < NW; ++w) s_cnt[w][tid] = 0;
        __syncthreads();
        const int64_t t0 = int64_t(tile) * TILE_N;
        const int items = int(min<int64_t>(TILE_N, n - t0));
        const int64_t wb = t0 + int64_t(wid) * PER;
        int k[R], x[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int64_t t = wb + r * WAVE + lane;
            k[r] = t < n ? key[t] : -1;
            x[r] = t < n ? (ident ? int32_t(t) : idx[t]) : 0;  // ident: the items' own positions
        }
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (k[r] >= 0) atomicAdd(&s_cnt[wid][(uint32_t(k[r]) >> shift) & (RADIX - 1)], 1);
        __syncthreads();
        {  // thread d: digit d's tile-local start, then each wave's start
            int tot = 0;
#pragma unroll
            for (int w = 0; w < NW; ++w) tot += s_cnt[w][tid];
            const int incl = int(wave_incl_scan32(uint32_t(tot), lane));
            if (lane == WAVE - 1) s_wsum[wid] = incl;
            __syncthreads();
            int run = incl - tot;
#pragma unroll
            for (int w = 0; w < NW; ++w) run += w < wid ? s_wsum[w] : 0;
            s_goff[tid] = off[int64_t(tid) * ntiles + tile] - run;
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                const int c = s_cnt[w][tid];
                s_cnt[w][tid] = run;
                run += c;
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const bool valid = k[r] >= 0;
            const int d = valid ? (uint32_t(k[r]) >> shift) & (RADIX - 1) : 0;
            uint64_t peers = __ballot(valid);
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                const uint64_t ones = __ballot((d >> b) & 1);
                peers &= ((d >> b) & 1) ? ones : ~ones;
            }
            const int rank = __popcll(peers & lt);
            const int pos = valid ? s_cnt[wid][d] + rank : 0;
            if (valid && rank == 0) s_cnt[wid][d] = pos + __popcll(peers);
            if (valid) {
                s_k[pos] = k[r];
                s_x[pos] = x[r];
            }
        }
        __syncthreads();
        for (int q = tid; q < items; q += T) {
            const int kk = s_k[q];
            const int g = s_goff[(uint32_t(kk) >> shift) & (RADIX - 1)] + q;
            key_out[g] = kk;
            idx_out[g] = s_x[q];
        }
        __syncthreads();  // the next tile clears the rows and refills the stage
    }
}

// Segmented passes (seg_radix_kv): tiles of at most TILE_N items that never
// cross a segment, tile j given as (t0, t1, hbase, hstride): its count of
// digit d sits at hist[hbase + d * hstride].  hist[...] = items of the tile
// with digit d, counted per wave as in k_hist.
__global__ __launch_bounds__(T) void k_seg_hist(const int32_t *__restrict__ key, const int4 *__restrict__ tiles,
                                                int ntiles, int shift, int32_t *hist) {
    constexpr int NW = T / WAVE;
    __shared__ int32_t s_h[NW][RADIX];
    const int tid = threadIdx.x, wid = tid / WAVE;
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int4 tl = tiles[tile];
#pragma unroll
        for (int w = 0; w < NW; ++w) s_h[w][tid] = 0;
        __syncthreads();
        for (int t = tl.x + tid; t < tl.y; t += T) atomicAdd(&s_h[wid][(uint32_t(key[t]) >> shift) & (RADIX - 1)], 1);
        __syncthreads();
        int c = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) c += s_h[w][tid];
        hist[tl.z + tid * tl.w] = c;
        __syncthreads();
    }
}

// The stable scatter of k_scatter over the segmented tiles, moving an FP64
// value with each key (staged in LDS beside it) instead of an index.
__global__ __launch_bounds__(T) void k_seg_scatter(const int32_t *__restrict__ key, const double *__restrict__ val,
                                                   const int4 *__restrict__ tiles, int ntiles, int shift,
                                                   const int32_t *__restrict__ off, int32_t *key_out,
                                                   double *val_out) {
    constexpr int NW = T / WAVE, PER = TILE_N / NW, R = PER / WAVE;
    static_assert(RADIX == T, "one thread per digit");
    __shared__ int32_t s_cnt[NW][RADIX];
    __shared__ int32_t s_goff[RADIX];  // global position of the tile's digit-d run, less its tile-local start
    __shared__ int32_t s_wsum[NW];
    __shared__ int32_t s_k[TILE_N];
    __shared__ double s_v[TILE_N];
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
    const uint64_t lt = (1ull << lane) - 1;
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int4 tl = tiles[tile];
#pragma unroll
        for (int w = 0; w < NW; ++w) s_cnt[w][tid] = 0;
        __syncthreads();
        const int items = tl.y - tl.x;
        const int wb = tl.x + wid * PER;
        int k[R];
        double v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int t = wb + r * WAVE + lane;
            k[r] = t < tl.y ? key[t] : -1;
            v[r] = t < tl.y ? val[t] : 0.0;
        }
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (k[r] >= 0) atomicAdd(&s_cnt[wid][(uint32_t(k[r]) >> shift) & (RADIX - 1)], 1);
        __syncthreads();
        {  // thread d: digit d's tile-local start, then each wave's start
            int tot = 0;
#pragma unroll
            for (int w = 0; w < NW; ++w) tot += s_cnt[w][tid];
            const int incl = int(wave_incl_scan32(uint32_t(tot), lane));
            if (lane == WAVE - 1) s_wsum[wid] = incl;
            __syncthreads();
            int run = incl - tot;
#pragma unroll
            for (int w = 0; w < NW; ++w) run += w < wid ? s_wsum[w] : 0;
            s_goff[tid] = off[tl.z + tid * tl.w] - run;
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                const int c = s_cnt[w][tid];
                s_cnt[w][tid] = run;
                run += c;
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const bool valid = k[r] >= 0;
            const int d = valid ? (uint32_t(k[r]) >> shift) & (RADIX - 1) : 0;
            uint64_t peers = __ballot(valid);
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                const uint64_t ones = __ballot((d >> b) & 1);
                peers &= ((d >> b) & 1) ? ones : ~ones;
            }
            const int rank = __popcll(peers & lt);
            const int pos = valid ? s_cnt[wid][d] + rank : 0;
            if (valid && rank == 0) s_cnt[wid][d] = pos + __popcll(peers);
            if (valid) {
                s_k[pos] = k[r];
                s_v[pos] = v[r];
            }
        }
        __syncthreads();
        for (int q = tid; q < items; q += T) {
            const int kk = s_k[q];
            const int g = s_goff[(uint32_t(kk) >> shift) & (RADIX - 1)] + q;
            key_out[g] = kk;
            val_out[g] = s_v[q];
        }
        __syncthreads();  // the next tile clears the rows and refills the stage
    }
}

// rowid[e] = the row holding entry e (one wave per row)
__global__ __launch_bounds__(256) void k_row_ids(int m, const int32_t *__restrict__ rp, int32_t *rowid) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int64_t nw = int64_t(gridDim.x) * (256 / WAVE);
    for (int64_t r = int64_t(blockIdx.x) * (256 / WAVE) + threadIdx.x / WAVE; r < m; r += nw)
        for (int e = rp[r] + lane; e < rp[r + 1]; e += WAVE) rowid[e] = int32_t(r);
}

__global__ __launch_bounds__(256) void k_gather(const int32_t *__restrict__ idx, int64_t n,
                                                const int32_t *__restrict__ rowid, const float *__restrict__ val,
                                                int32_t *row_ind, float *val_out) {
    for (int64_t p = int64_t(blockIdx.x) * 256 + threadIdx.x; p < n; p += int64_t(gridDim.x) * 256) {
        const int32_t e = idx[p];
        row_ind[p] = rowid[e];
        if (val_out) val_out[p] = val[e];
    }
}

unsigned grid_for(apspie_ctx *ctx, int64_t n, unsigned per) {
    return std::max(1u, std::min(ceil_div(std::max<int64_t>(n, 1), per), unsigned(ctx->cu_count) * 32));
}

void compress(apspie_ctx *ctx, const int32_t *key, int64_t n, int m, int32_t *off) {
    int32_t *bad = (int32_t *)ctx->slot(S_AUX7, 256);
    APSPIE_HIP(hipMemsetAsync(bad, 0, 4, ctx->stream));
    k_compress<<<grid_for(ctx, n + 1, 256), 256, 0, ctx->stream>>>(key, n, m, off, bad);
    APSPIE_HIP(hipGetLastError());
    int32_t h = 0;
    APSPIE_HIP(hipMemcpyAsync(&h, bad, 4, hipMemcpyDeviceToHost, ctx->stream));
    APSPIE_HIP(hipStreamSynchronize(ctx->stream));
    APSPIE_CHECK(h != 1, APSPIE_EINVAL, "row indices are not sorted");
    APSPIE_CHECK(h != 2, APSPIE_EINVAL, "row index out of range");
}

template <class O>
void scan_to(apspie_ctx *ctx, const int32_t *in, O *out, int64_t n) {
    const int nscan = std::max(1, (int)ceil_div(n, TILE_N));
    uint64_t *status = (uint64_t *)ctx->slot(S_STATUS, size_t(nscan + 2) * 8);
    k_zero64<<<grid_for(ctx, nscan + 2, 256), 256, 0, ctx->stream>>>(status, nscan + 2);
    k_scan<O><<<nscan, 256, 0, ctx->stream>>>(in, out, n, status, (unsigned *)(status + nscan));
    APSPIE_HIP(hipGetLastError());
}

void scan32(apspie_ctx *ctx, const int32_t *in, int32_t *out, int64_t n) { scan_to(ctx, in, out, n); }
void scan_wide(apspie_ctx *ctx, const int32_t *in, int64_t *out, int64_t n) { scan_to(ctx, in, out, n); }

int radix_pairs(apspie_ctx *ctx, int32_t *key[2], int32_t *idx[2], int64_t n, int bits, int start, bool ident) {
    if (n == 0) return start;
    if (ident && bits == 0) bits = 1;  // one pass writes the identity the caller reads
    const int ntiles = (int)ceil_div(n, TILE_N);
    const int64_t nh = int64_t(RADIX) * ntiles;
    int32_t *hist = (int32_t *)ctx->slot(S_AUX4, size_t(nh) * 4);
    int32_t *offs = (int32_t *)ctx->slot(S_AUX5, size_t(nh + 1) * 4);
    const unsigned gt = std::max(1u, std::min((unsigned)ntiles, unsigned(ctx->cu_count) * 8));
    int cur = start;
    for (int shift = 0; shift < bits; shift += 8, cur ^= 1) {
        k_hist<<<gt, T, 0, ctx->stream>>>(key[cur], n, shift, ntiles, hist);
        scan32(ctx, hist, offs, nh);
        k_scatter<<<gt, T, 0, ctx->stream>>>(key[cur], idx[cur], ident && shift == 0, n, shift, ntiles, offs,
                                             key[cur ^ 1], idx[cur ^ 1]);
    }
    APSPIE_HIP(hipGetLastError());
    return cur;
}

int seg_radix_kv(apspie_ctx *ctx, int32_t *key[2], double *val[2], int64_t n, const std::vector<int32_t> &seg,
                 int bits, std::vector<int32_t> &tiles, int first_shift) {
    int cur = 0;
    if (n == 0 || bits == 0) return cur;
    // the tile table: tiles of a segment are consecutive, histogram slots of
    // a segment are digit-major over its tiles
    tiles.clear();
    const int nseg = int(seg.size()) - 1;
    int64_t nh = 0;
    for (int r = 0; r < nseg; ++r) {
        const int64_t a = seg[r], b = seg[size_t(r) + 1];
        const int nt = int(ceil_div(b - a, TILE_N));
        for (int j = 0; j < nt; ++j) {
            tiles.push_back(int32_t(a + int64_t(j) * TILE_N));
            tiles.push_back(int32_t(std::min<int64_t>(b, a + int64_t(j + 1) * TILE_N)));
            tiles.push_back(int32_t(nh + j));
            tiles.push_back(nt);
        }
        nh += int64_t(nt) * RADIX;
    }
    const int ntiles = int(tiles.size() / 4);
    APSPIE_CHECK(nh < (int64_t(1) << 31), APSPIE_ELIMIT, "segmented sort histogram too large");
    int4 *dt = (int4 *)ctx->slot(S_AUX6, tiles.size() * 4);
    APSPIE_HIP(hipMemcpyAsync(dt, tiles.data(), tiles.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    int32_t *hist = (int32_t *)ctx->slot(S_AUX4, size_t(nh) * 4);
    int32_t *offs = (int32_t *)ctx->slot(S_AUX5, size_t(nh + 1) * 4);
    const unsigned gt = std::max(1u, std::min((unsigned)ntiles, unsigned(ctx->cu_count) * 8));
    for (int shift = std::max(0, first_shift); shift < bits; shift += 8, cur ^= 1) {
        k_seg_hist<<<gt, T, 0, ctx->stream>>>(key[cur], dt, ntiles, shift, hist);
        scan32(ctx, hist, offs, nh);
        k_seg_scatter<<<gt, T, 0, ctx->stream>>>(key[cur], val[cur], dt, ntiles, shift, offs, key[cur ^ 1],
                                                 val[cur ^ 1]);
    }
    APSPIE_HIP(hipGetLastError());
    return cur;
}

void csr2csc_into(apspie_ctx *ctx, const apspie_csr *a, int32_t *d_col_ptr, int32_t *d_row_ind, float *d_val) {
    const int64_t n = a->nnz;
    if (n == 0) {
        APSPIE_HIP(hipMemsetAsync(d_col_ptr, 0, (size_t(a->n) + 1) * 4, ctx->stream));
        APSPIE_HIP(hipStreamSynchronize(ctx->stream));
        return;
    }
    int32_t *key[2] = {(int32_t *)ctx->slot(S_AUX0, size_t(n) * 4), (int32_t *)ctx->slot(S_AUX1, size_t(n) * 4)};
    int32_t *idx[2] = {(int32_t *)ctx->slot(S_AUX2, size_t(n) * 4), (int32_t *)ctx->slot(S_AUX3, size_t(n) * 4)};
    k_seed<<<grid_for(ctx, n, 256), 256, 0, ctx->stream>>>(a->col_ind, n, key[0]);
    int bits = 1;
    while (bits < 31 && (int64_t(1) << bits) < int64_t(a->n)) ++bits;
    const int cur = radix_pairs(ctx, key, idx, n, bits, 0, true);
    compress(ctx, key[cur], n, a->n, d_col_ptr);
    int32_t *rowid = key[cur ^ 1];  // free after the last pass
    k_row_ids<<<grid_for(ctx, int64_t(a->m) * WAVE, 256), 256, 0, ctx->stream>>>(a->m, a->row_ptr, rowid);
    k_gather<<<grid_for(ctx, n, 256), 256, 0, ctx->stream>>>(idx[cur], n, rowid, a->val, d_row_ind, d_val);
    APSPIE_HIP(hipGetLastError());
    APSPIE_HIP(hipStreamSynchronize(ctx->stream));
}

}  // namespace cvt
}  // namespace apspie

using namespace apspie;

extern "C" {

int apspie_coo_to_csr(apspie_ctx *ctx, const int32_t *d_rows, int64_t nnz, int32_t m, int32_t *d_row_ptr) {
    if (!ctx) return APSPIE_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    return guarded(ctx, [&] {
        APSPIE_CHECK(m >= 0 && nnz >= 0 && nnz < (int64_t(1) << 31), APSPIE_EINVAL, "bad COO shape");
        APSPIE_CHECK(d_row_ptr && (nnz == 0 || d_rows), APSPIE_EINVAL, "null COO arrays");
        cvt::compress(ctx, d_rows, nnz, m, d_row_ptr);
    });
}

int apspie_csr2csc(apspie_ctx *ctx, const apspie_csr *a, int32_t *d_col_ptr, int32_t *d_row_ind, float *d_val) {
    if (!ctx) return APSPIE_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    return guarded(ctx, [&] {
        using namespace cvt;
        validate_csr(a, false);
        APSPIE_CHECK(d_col_ptr && (a->nnz == 0 || d_row_ind), APSPIE_EINVAL, "null CSC arrays");
        APSPIE_CHECK(!d_val || a->val || a->nnz == 0, APSPIE_EINVAL, "values asked of a matrix without values");
        csr2csc_into(ctx, a, d_col_ptr, d_row_ind, d_val);
    });
}

int apspie_csr_attach_csc_device(apspie_ctx *ctx, apspie_csr_owned *c) {
    if (!ctx || !c) return APSPIE_EINVAL;
    int32_t *cp = nullptr, *ri = nullptr;
    float *cv = nullptr;
    const int rc = guarded(ctx, [&] {
        APSPIE_CHECK(c->csc_col_ptr == nullptr, APSPIE_EINVAL, "the matrix already has a CSC");
        APSPIE_HIP(hipMalloc(&cp, (size_t(c->n) + 1) * 4));
        APSPIE_HIP(hipMalloc(&ri, size_t(std::max<int64_t>(c->nnz, 1)) * 4));
        if (c->val) APSPIE_HIP(hipMalloc(&cv, size_t(std::max<int64_t>(c->nnz, 1)) * 4));
    });
    if (rc == APSPIE_OK) {
        const apspie_csr v = apspie_csr_view(c);
        const int r2 = apspie_csr2csc(ctx, &v, cp, ri, cv);
        if (r2 == APSPIE_OK) {
            c->csc_col_ptr = cp;
            c->csc_row_ind = ri;
            c->csc_val = cv;
            return APSPIE_OK;
        }
        (void)hipFree(cp);
        (void)hipFree(ri);
        if (cv) (void)hipFree(cv);
        return r2;
    }
    if (cp) (void)hipFree(cp);
    if (ri) (void)hipFree(ri);
    if (cv) (void)hipFree(cv);
    return rc;
}

}  // extern "C"
