// spgemm.hip -- C = A * B on gfx950: row-wise Gustavson with row bins:
// register rows, LDS hash tables, LDS sorts up to 1024 products
// (spgemm_esc_min; the sorts reach 2048), expand-sort-compress beyond.
//
// Pipeline (one context stream):
//   1. ub[i]   = sum over k in A(i,:) of nnz(B(k,:))  (products of row i),
//                a wave per 64 rows summing its entries' B row lengths.
//   2. bins    rows by ub: <= 16 (four rows a wave) and <= 64 (a wave per
//              row), both in registers; <= 128, <= 256 (LDS hash tables of
//              512 and 1024 slots); <= 512, <= 1024, <= 2048 (the row's
//              products sorted in LDS, a workgroup a row); > 2048 (the long rows,
//              joined by the bins above spgemm_esc_min, 1024 by default).
//   3. LDS rows in one pass: the products (column, a * b) with FP64 sums
//      (products of FP32 inputs are exact in FP64; a sum is rounded to FP32
//      once), the sorted row written to a staging area at the prefix of the
//      rows' ub, with its nnz.  A sorted row sums each run of equal columns
//      in k order, as Gustavson's row accumulator does.  Long rows: a count
//      of each row's products by the top 8 bits of their column, then the
//      products written once, each straight into its row's group of equal
//      top bits in k order; each window of a row's groups is sorted by column
//      in LDS and each run of equal columns is one entry (FP64 sum in k
//      order); nnz = the runs.
//   4. row_ptr(C) = exclusive scan (decoupled look-back).
//   5. the staged rows are copied into C; the long rows' runs are copied
//      into C.
// Tuning keeps the older forms selectable (and they take over when int32
// offsets would overflow): a counting pass before the product pass for the
// LDS rows (spgemm_passes = 2), global hash tables of 4 * pow2ceil(ub) slots
// for the long rows (spgemm_long_rows = 0).
// The hash table uses an order-preserving home slot, home(c) = c * H / n, with
// linear probing and no wrap-around (2H slots for at most H / 2 keys): a key's
// slot lies in the cluster (run of occupied slots) holding its home, so all
// keys of an earlier cluster are smaller and the row comes out sorted after
// an insertion sort inside each cluster.  Correct for any column
// distribution; fast when columns spread over [0, n).
//
// Replaces spgemm (src/spgemm.cuh:131-342, 2D-partitioned DCSC outer product
// with per-partition global hash tables, src/spgemmKernel.cuh:164-738) and the
// cuSPARSE csrgemm oracle path (src/spgemm.cuh:344-445).  Structural zeros
// (products that cancel) are kept, as the reference's hash does
// (src/spgemmKernel.cuh:194-198).
#include <algorithm>
#include <climits>
#include <cstring>
#include <utility>
#include <vector>

#include "convert.hpp"
#include "esc.hpp"
#include "device.hpp"
#include "internal.hpp"

namespace apspie {
namespace sg {

using namespace dev;

struct CsrD {
    const int32_t *rp, *col;
    const float *val;
    int m, n;
};

// A table of S slots has H = S / 2 home slots and takes rows of at most S / 4
// products, so keys fill at most half of the home range (linear probing with
// order-preserving homes clusters badly when that range is nearly full).
constexpr int NBINS = 8;
constexpr int BIN_UB[NBINS - 1] = {16, 64, 128, 256, 512, 1024, 2048};
constexpr int LONG = NBINS - 1;  // the bin of rows beyond the LDS tables

// Product bounds ub[i] = sum over A(i,:) of nnz(B(k,:)) (0 for a column
// outside B).  A wave takes 64 consecutive rows: their entries are
// contiguous, so its lanes stride over them, find each entry's row by a
// binary search over the rows' pointers in LDS (an entry's row only grows)
// and add the entry's B row length into the row's 64-bit LDS sum; four
// entries a lane have their loads in flight together.  The waves
// share no LDS, so they synchronise by wavefront-scope fences only.  *bad =
// 1 for a row of 2^31 - 1 or more products.
__global__ __launch_bounds__(256) void k_row_ub(int m, const int32_t *__restrict__ rp, const int32_t *__restrict__ col,
                                                const int32_t *__restrict__ brp, int bm, int32_t *ub, int32_t *bad) {
    constexpr int NW = 256 / WAVE;
    __shared__ int32_t s_rp[NW][WAVE];
    __shared__ unsigned long long s_acc[NW][WAVE];
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    int32_t *srp = s_rp[wid];
    unsigned long long *acc = s_acc[wid];
    for (int64_t r0 = (int64_t(blockIdx.x) * NW + wid) * WAVE; r0 < m; r0 += int64_t(gridDim.x) * NW * WAVE) {
        const int nr = int(std::min<int64_t>(WAVE, m - r0));
        const int e0 = rp[r0], e1 = rp[r0 + nr];
        wave_lds_sync();  // the previous rows' reads are done
        srp[lane] = rp[r0 + std::min(lane, nr)];
        acc[lane] = 0;
        wave_lds_sync();
        constexpr int U = 16;  // entries per lane in flight
        int row = 0;          // the last row x with srp[x] <= e
        for (int eb = e0 + lane; eb < e1; eb += U * WAVE) {
            int k[U], b0[U], b1[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int e = eb + u * WAVE;
                k[u] = e < e1 ? col[e] : -1;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                b0[u] = b1[u] = 0;
                if ((unsigned)k[u] < (unsigned)bm) {
                    b0[u] = brp[k[u]];
                    b1[u] = brp[k[u] + 1];
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int e = eb + u * WAVE;
                if (e >= e1 || b1[u] == b0[u]) continue;
                int top = nr - 1;
                while (row < top) {
                    const int mid = (row + top + 1) >> 1;
                    if (srp[mid] <= e) row = mid;
                    else top = mid - 1;
                }
                atomicAdd(&acc[row], (unsigned long long)(b1[u] - b0[u]));
            }
        }
        wave_lds_sync();
        if (lane < nr) {
            const unsigned long long u = acc[lane];
            if (u >= (1ull << 31) - 1) *bad = 1;
            ub[r0 + lane] = int32_t(std::min<unsigned long long>(u, (1ull << 31) - 2));
        }
    }
}

// dst[0, n) = src[0, n): small results written straight into pinned host
// memory (its device address), read by the host after a stream
// synchronisation.
__global__ __launch_bounds__(256) void k_to_host(const int32_t *__restrict__ src, int n, int32_t *dst) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) dst[i] = src[i];
}

inline void to_host(apspie_ctx *ctx, const int32_t *src, int n, int hs, size_t byte_off) {
    int32_t *dst = (int32_t *)((char *)ctx->hslot_device(hs) + byte_off);
    k_to_host<<<std::max(1u, std::min(ceil_div(std::max(n, 1), 256), 1024u)), 256, 0, ctx->stream>>>(src, n, dst);
}

__global__ __launch_bounds__(256) void k_zero_i32(int32_t *p, int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) p[i] = 0;
}

// rows -> bin lists; rows without products get nnz 0 directly.  Each
// workgroup takes a contiguous chunk of rows: it counts its rows per bin in
// LDS, claims each bin's range with one global atomic, then walks the chunk
// again placing rows through LDS cursors (counters every wave hit serialise
// otherwise).
__device__ __forceinline__ int bin_of_ub(int u) {
    if (u == 0) return -1;
    int b = NBINS - 1;
#pragma unroll
    for (int k = NBINS - 2; k >= 0; --k)
        if (u <= BIN_UB[k]) b = k;
    return b;
}

__global__ __launch_bounds__(256) void k_bin(int m, const int32_t *ub, int32_t *nnzc, int32_t *lists,
                                             int32_t *counts, unsigned long long *products) {
    __shared__ int s_cnt[NBINS], s_cur[NBINS];
    __shared__ unsigned long long s_prod;
    const int lane = threadIdx.x & (WAVE - 1);
    const int chunk = (m + gridDim.x - 1) / gridDim.x;
    const int r0 = min(m, int(blockIdx.x) * chunk), r1 = min(m, r0 + chunk);
    if (threadIdx.x < NBINS) s_cnt[threadIdx.x] = 0;
    if (threadIdx.x == 0) s_prod = 0;
    __syncthreads();
    unsigned long long mine = 0;
    for (int i = r0 + threadIdx.x; i < r1; i += 256) {
        const int u = ub[i];
        mine += (unsigned long long)u;
        const int b = bin_of_ub(u);
        if (b < 0) nnzc[i] = 0;
#pragma unroll
        for (int k = 0; k < NBINS; ++k) {
            const uint64_t ball = __ballot(b == k);
            if (ball && lane == __ffsll((unsigned long long)ball) - 1) atomicAdd(&s_cnt[k], __popcll(ball));
        }
    }
    mine = wave_sum(mine);
    if (lane == 0 && mine) atomicAdd(&s_prod, mine);
    __syncthreads();
    if (threadIdx.x < NBINS) {
        const int n = s_cnt[threadIdx.x];
        s_cur[threadIdx.x] = n ? atomicAdd(&counts[threadIdx.x], n) : 0;
    }
    if (threadIdx.x == 0 && s_prod) atomicAdd(products, s_prod);
    __syncthreads();
    for (int i = r0 + threadIdx.x; i < r1; i += 256) {
        const int b = bin_of_ub(ub[i]);
#pragma unroll
        for (int k = 0; k < NBINS; ++k) {
            const uint64_t ball = __ballot(b == k);
            if (!ball) continue;
            const int lead = __ffsll((unsigned long long)ball) - 1;
            int base = 0;
            if (lane == lead) base = atomicAdd(&s_cur[k], __popcll(ball));
            base = __shfl(base, lead, WAVE);
            if (b == k) {
                const int rank = __builtin_amdgcn_mbcnt_hi(uint32_t(ball >> 32),
                                                           __builtin_amdgcn_mbcnt_lo(uint32_t(ball), 0));
                lists[int64_t(k) * m + base + rank] = i;
            }
        }
    }
}

// out[r] = x[idx[r]]
__global__ __launch_bounds__(256) void k_gather(const int32_t *__restrict__ idx, int n, const int32_t *__restrict__ x,
                                                int32_t *out) {
    for (int r = blockIdx.x * 256 + threadIdx.x; r < n; r += gridDim.x * 256) out[r] = x[idx[r]];
}

// exclusive scan of n non-negative int32 into out[0..n], out[n] = total
// (single pass, 64-bit sums); the exact total also goes to *total64, so the
// host can refuse a total that out[] cannot hold
__global__ __launch_bounds__(256) void k_scan(const int32_t *in, int32_t *out, int n, uint64_t *status,
                                              unsigned int *counter, unsigned long long *total64) {
    constexpr int ITEMS = 16, TILE_N = 256 * ITEMS;
    __shared__ int s_bid;
    __shared__ uint64_t s_w[4];
    __shared__ uint64_t s_pre;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    if (tid == 0) s_bid = (int)atomicAdd(counter, 1u);
    __syncthreads();
    const int bid = s_bid;
    const int64_t b0 = int64_t(bid) * TILE_N + int64_t(tid) * ITEMS;
    int32_t v[ITEMS];
    uint64_t sum = 0;
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        v[k] = b0 + k < n ? in[b0 + k] : 0;
        sum += uint64_t(uint32_t(v[k]));
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
        if (b0 + k < n) out[b0 + k] = int32_t(run);
        run += uint64_t(uint32_t(v[k]));
        if (b0 + k == n - 1) {
            out[n] = int32_t(run);
            *total64 = run;
        }
    }
    if (n == 0 && bid == 0 && tid == 0) {
        out[0] = 0;
        *total64 = 0;
    }
}

template <int T>
__device__ __forceinline__ int block_excl_scan(int x, int *s_tmp, int &total) {
    constexpr int NW = T / WAVE;
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    const int incl = (int)wave_incl_scan32((uint32_t)x, lane);
    if (lane == WAVE - 1) s_tmp[wid] = incl;
    __syncthreads();
    int pre = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        if (w < wid) pre += s_tmp[w];
        tot += s_tmp[w];
    }
    __syncthreads();
    total = tot;
    return pre + incl - x;
}

// The lanes of the wave holding the same BITS-bit digit d as this lane (and
// valid), from a ballot per bit.
template <int BITS>
__device__ __forceinline__ uint64_t digit_peers(bool valid, int d) {
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < BITS; ++b) {
        const uint64_t ones = __ballot((d >> b) & 1);
        peers &= ((d >> b) & 1) ? ones : ~ones;
    }
    return peers;
}

template <bool GLOBAL>
__device__ __forceinline__ int ld_key(const int32_t *key, int s) {
    if (GLOBAL) return __hip_atomic_load(key + s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return key[s];
}

template <bool GLOBAL>
__device__ __forceinline__ double ld_val(const double *val, int s) {
    if (GLOBAL) return __hip_atomic_load(val + s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return val[s];
}

// insertion sort of the keys (and their values) in [a, b) of an LDS table
__device__ __forceinline__ void sort_run(int32_t *key, double *val, int a, int b) {
    for (int x = a + 1; x < b; ++x) {
        const int c = key[x];
        const double v = val[x];
        int w = x;
        while (w > a && key[w - 1] > c) {
            key[w] = key[w - 1];
            val[w] = val[w - 1];
            --w;
        }
        key[w] = c;
        val[w] = v;
    }
}

// One row per workgroup iteration.  T threads, S table slots (LDS) or the
// row's own slice of a global table.
template <int T, int S, bool NUMERIC, bool GLOBAL>
__global__ __launch_bounds__(T) void k_rows(const int32_t *__restrict__ rows, int nrows, CsrD A, CsrD B,
                                            int32_t *nnzc, const int32_t *__restrict__ crp, int32_t *ccol,
                                            float *cval, int32_t *gkeys, double *gvals,
                                            const int64_t *__restrict__ goff, const int32_t *__restrict__ gslots) {
    __shared__ int32_t s_key[GLOBAL ? 1 : S];
    __shared__ double s_val[(NUMERIC && !GLOBAL) ? S : 1];
    __shared__ int32_t s_bs[T];
    __shared__ int32_t s_pre[T];
    __shared__ float s_a[T];
    __shared__ int32_t s_tmp[T / WAVE + 1];
    __shared__ int32_t s_cnt;
    constexpr int TS = (GLOBAL && NUMERIC) ? 4 * T : 1;  // global tables stream through LDS in tiles
    __shared__ int32_t s_tk[TS];
    __shared__ double s_tv[TS];
    // A global table outlives its kernel: the counting pass already
    // accumulates the values, and the numeric pass only extracts the row.
    constexpr bool VALS = NUMERIC || GLOBAL, INSERT = !(NUMERIC && GLOBAL);
    const int tid = threadIdx.x;
    for (int ri = blockIdx.x; ri < nrows; ri += gridDim.x) {
        const int i = rows[ri];
        int32_t *key = s_key;
        double *val = s_val;
        int slots = S;
        if (GLOBAL) {
            key = gkeys + goff[ri];
            val = gvals + goff[ri];
            slots = gslots[ri];
        }
        const int H = slots / 2;
        if (INSERT) {
            for (int s = tid; s < slots; s += T) {
                if (GLOBAL) {
                    __hip_atomic_store(key + s, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (VALS) __hip_atomic_store(val + s, 0.0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                } else {
                    key[s] = -1;
                    if (VALS) val[s] = 0.0;
                }
            }
            if (tid == 0) s_cnt = 0;
            __syncthreads();
            const int ab = A.rp[i], ae = A.rp[i + 1];
            for (int jb = ab; jb < ae; jb += T) {
                const int j = jb + tid;
                int L = 0, bs = 0;
                float a = 0.f;
                if (j < ae) {
                    const int k = A.col[j];
                    if ((unsigned)k < (unsigned)B.m) {
                        a = A.val ? A.val[j] : 1.0f;
                        bs = B.rp[k];
                        L = B.rp[k + 1] - bs;
                    }
                }
                int tot;
                const int pre = block_excl_scan<T>(L, s_tmp, tot);
                s_bs[tid] = bs;
                s_pre[tid] = pre;
                s_a[tid] = a;
                const int ne = min(T, ae - jb);
                __syncthreads();
                // U products per thread in flight: their entries come from one
                // binary search and a forward walk (a thread's products only move
                // right), their loads are issued before any insertion
                constexpr int U = 4;
                int lo = 0;
                for (int tb = tid; tb < tot; tb += U * T) {
                    int cq[U];
                    double pq[U];
    #pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const int t = tb + u * T;
                        cq[u] = -1;
                        pq[u] = 0.0;
                        if (t < tot) {
                            if (u == 0) {
                                int hi = ne - 1;  // last entry whose prefix <= t
                                while (lo < hi) {
                                    const int mid = (lo + hi + 1) >> 1;
                                    if (s_pre[mid] <= t) lo = mid;
                                    else hi = mid - 1;
                                }
                            } else {
                                while (lo + 1 < ne && s_pre[lo + 1] <= t) ++lo;
                            }
                            const int q = s_bs[lo] + (t - s_pre[lo]);
                            cq[u] = B.col[q];
                            if (VALS) pq[u] = double(s_a[lo]) * double(B.val ? B.val[q] : 1.0f);
                        }
                    }
    #pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const int c = cq[u];
                        if ((unsigned)c >= (unsigned)B.n) continue;
                        int s = int((uint64_t(uint32_t(c)) * uint64_t(H)) / uint64_t(uint32_t(B.n)));
                        for (;;) {
                            const int old = atomicCAS(&key[s], -1, c);
                            if (old == -1 || old == c) {
                                if (VALS) atomicAdd(&val[s], pq[u]);
                                if (!NUMERIC && old == -1) atomicAdd(&s_cnt, 1);
                                break;
                            }
                            ++s;  // < slots: at most H / 2 keys, homes < H
                        }
                    }
                }
                __syncthreads();
            }
        }  // INSERT
        if (!NUMERIC) {
            if (tid == 0) nnzc[i] = s_cnt;
            __syncthreads();
            continue;
        }
        if (GLOBAL) {
            // a global table streams through LDS in tiles of 4 T slots
            // (coalesced loads); a thread owns 4 consecutive slots of a tile
            // and emits the clusters that start in them, finishing a cluster
            // that runs past the tile from global memory
            int run = crp[i];
            for (int t0 = 0; t0 < slots; t0 += TS) {
                for (int s = tid; s < TS; s += T) {
                    const int q = t0 + s;
                    s_tk[s] = q < slots ? ld_key<GLOBAL>(key, q) : -1;
                    s_tv[s] = q < slots ? ld_val<GLOBAL>(val, q) : 0.0;
                }
                const int left = (t0 > 0 && tid == 0) ? ld_key<GLOBAL>(key, t0 - 1) : -1;
                __syncthreads();
                auto key_at = [&](int q) { return q - t0 < TS ? s_tk[q - t0] : ld_key<GLOBAL>(key, q); };
                auto starts = [&](int s) {  // slot t0 + s holds the first key of a cluster
                    return t0 + s < slots && s_tk[s] != -1 && (s > 0 ? s_tk[s - 1] : left) == -1;
                };
                int cnt = 0;
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    if (starts(tid * 4 + r))
                        for (int q = t0 + tid * 4 + r; q < slots && key_at(q) != -1; ++q) ++cnt;
                int tot;
                int out = run + block_excl_scan<T>(cnt, s_tmp, tot);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int s0 = tid * 4 + r;
                    if (!starts(s0)) continue;
                    int s1 = s0;
                    while (s1 < TS && s_tk[s1] != -1) ++s1;
                    if (s1 < TS || t0 + TS == slots) {
                        // the cluster ends inside the tile: sort it in LDS (its
                        // slots are this thread's alone), then write it out
                        sort_run(s_tk, s_tv, s0, s1);
                        for (int a = s0; a < s1; ++a, ++out) {
                            ccol[out] = s_tk[a];
                            cval[out] = float(s_tv[a]);
                        }
                        continue;
                    }
                    const int o0 = out;
                    for (int q = t0 + s0; q < slots; ++q) {
                        const int c = key_at(q);
                        if (c == -1) break;
                        // insertion into the sorted run [o0, out)
                        const float v = float(q - t0 < TS ? s_tv[q - t0] : ld_val<GLOBAL>(val, q));
                        int w = out;
                        while (w > o0 && ccol[w - 1] > c) {
                            ccol[w] = ccol[w - 1];
                            cval[w] = cval[w - 1];
                            --w;
                        }
                        ccol[w] = c;
                        cval[w] = v;
                        ++out;
                    }
                }
                run += tot;
                __syncthreads();
            }
            continue;
        }
        // extraction: each thread owns a segment of slots and emits the
        // clusters that start in it (a cluster running in from the left
        // belongs to the previous segment)
        const int seg = slots / T;
        const int sb = tid * seg, se = sb + seg;
        int p = sb;
        if (p > 0 && ld_key<GLOBAL>(key, p - 1) != -1)
            while (p < slots && ld_key<GLOBAL>(key, p) != -1) ++p;
        int cnt = 0;
        for (int q = p; q < se;) {
            if (ld_key<GLOBAL>(key, q) == -1) {
                ++q;
                continue;
            }
            while (q < slots && ld_key<GLOBAL>(key, q) != -1) {
                ++cnt;
                ++q;
            }
        }
        int tot;
        int out = crp[i] + block_excl_scan<T>(cnt, s_tmp, tot);
        if (tid == 0) nnzc[i] = tot;  // (a one-pass product learns nnz here)
        for (int q = p; q < se;) {
            if (key[q] == -1) {
                ++q;
                continue;
            }
            // the cluster [q0, q) is this thread's alone: sort it in LDS,
            // then write it out
            const int q0 = q;
            while (q < slots && key[q] != -1) ++q;
            sort_run(key, val, q0, q);
            for (int a = q0; a < q; ++a, ++out) {
                ccol[out] = key[a];
                cval[out] = float(val[a]);
            }
        }
        __syncthreads();
    }
}

// Rows of at most 16 products, four to a wave: a 16-lane group per row, in
// registers, no LDS.  Lane e takes A entry e of each chunk of 16, the group
// scans the B row lengths, and product p of the row lands on lane p (the
// row's products, ub <= 16).  A sorting network on (column, p) puts equal
// columns side by side in increasing k; each run's first lane sums the run
// in FP64 in that order and the run starts are written at their rank: at
// crp[i] (NUMERIC) after nnzc[i] (always).  Every exchange inside a 16-lane
// row is a DPP move (row shifts, row broadcasts, quad permutes, mirrors),
// which never leaves the row, so the four groups stay apart.
template <int CTRL>
__device__ __forceinline__ int row_dpp(int x) {
    return __builtin_amdgcn_update_dpp(0, x, CTRL, 0xf, 0xf, true);
}

constexpr int DPP_XOR1 = 0xb1;         // quad_perm [1,0,3,2]
constexpr int DPP_XOR2 = 0x4e;         // quad_perm [2,3,0,1]
constexpr int DPP_XOR3 = 0x1b;         // quad_perm [3,2,1,0]
constexpr int DPP_HALF_MIRROR = 0x141;  // lane i of each 8 <- lane 7 - i
constexpr int DPP_MIRROR = 0x140;       // lane i of the row <- lane 15 - i
constexpr int DPP_SHL0 = 0x100;         // + n: lane i <- lane i + n (0 past the row)
constexpr int DPP_SHR0 = 0x110;         // + n: lane i <- lane i - n (0 before the row)
constexpr int DPP_BCAST0 = 0x150;       // + n: every lane of the row <- lane n

__device__ __forceinline__ int row16_incl_sum(int x) {
    x += row_dpp<DPP_SHR0 + 1>(x);
    x += row_dpp<DPP_SHR0 + 2>(x);
    x += row_dpp<DPP_SHR0 + 4>(x);
    x += row_dpp<DPP_SHR0 + 8>(x);
    return x;
}

// The entry of the chunk that holds product t: the number of entries e < ne
// whose inclusive end incl[e] <= t (the ends do not decrease, so these are a
// prefix and the next entry, of length > 0, holds t).  ne is wave-uniform.
template <int... E>
__device__ __forceinline__ int row16_owner(int incl, int t, int ne, std::integer_sequence<int, E...>) {
    int own = 0;
    ((own += (E < ne && row_dpp<DPP_BCAST0 + E>(incl) <= t) ? 1 : 0), ...);
    return own;
}

// One compare-exchange of the network: the lane takes its partner's key
// when it keeps the smaller key (lower) and the partner's is smaller, or it
// keeps the larger (upper) and the partner's is larger.  Keys are distinct.
template <class Partner>
__device__ __forceinline__ void row16_cx(int &col, int &prod, bool lower, Partner partner) {
    const int oc = partner(col), op = partner(prod);
    const bool smaller = oc < col || (oc == col && op < prod);
    if (lower == smaller) {
        col = oc;
        prod = op;
    }
}

// Sum of the run that starts on this lane, over the next lanes with the same
// column, left to right; mp bounds every group's products (wave-uniform).
// The moves run with every lane active, the sums under the lane's own flag.
template <int... D>
__device__ __forceinline__ double row16_run_sum(int col, double v, int lane, int mp, std::integer_sequence<int, D...>) {
    double acc = v;
    bool run = true;
    auto step = [&](auto dc) {
        constexpr int d = decltype(dc)::value;
        if (d < mp) {
            const int nc = row_dpp<DPP_SHL0 + d>(col);
            const int lo = row_dpp<DPP_SHL0 + d>(__double2loint(v)), hi = row_dpp<DPP_SHL0 + d>(__double2hiint(v));
            run = run && lane + d < 16 && nc == col;
            if (run) acc += __hiloint2double(hi, lo);
        }
    };
    (step(std::integral_constant<int, D + 1>{}), ...);
    return acc;
}

template <bool NUMERIC>
__global__ __launch_bounds__(256) void k_rows_tiny(const int32_t *__restrict__ rows, int nrows, CsrD A, CsrD B,
                                                   int32_t *nnzc, const int32_t *__restrict__ crp, int32_t *ccol,
                                                   float *cval) {
    const int lane = threadIdx.x & 15, grp = (threadIdx.x & (WAVE - 1)) >> 4;
    const int wid = (blockIdx.x * 256 + threadIdx.x) / WAVE, nw = gridDim.x * (256 / WAVE);
    // the loop bound is wave-uniform (four rows a step): every DPP move and
    // shuffle below runs with all 64 lanes active
    for (int rb = wid * 4; rb < nrows; rb += nw * 4) {
        const int r = rb + grp;
        const bool live = r < nrows;
        const int i = live ? rows[r] : 0;
        const int ab = live ? A.rp[i] : 0, ae = live ? A.rp[i + 1] : 0;
        int col = INT_MAX, prod = lane;  // this lane's product: column, index in the row
        double val = 0.0;
        int pbase = 0;
        for (int cb = ab;; cb += 16) {
            const uint64_t have = __ballot(cb + lane < ae);
            if (!have) break;
            // entries in the wave's fullest group (each group's are a prefix)
            const uint32_t g16 = uint32_t(have | (have >> 16) | (have >> 32) | (have >> 48)) & 0xffffu;
            const int ne = 32 - __clz(g16);
            const int j = cb + lane;
            int L = 0, bs = 0;
            float a = 0.f;
            if (j < ae) {
                const int k = A.col[j];
                if ((unsigned)k < (unsigned)B.m) {
                    a = A.val ? A.val[j] : 1.0f;
                    bs = B.rp[k];
                    L = B.rp[k + 1] - bs;
                }
            }
            const int incl = row16_incl_sum(L), excl = incl - L;
            const int tot = row_dpp<DPP_BCAST0 + 15>(incl);
            const int t = lane - pbase;  // products pbase .. pbase + tot - 1 sit on those lanes
            const int own = min(row16_owner(incl, t, ne, std::make_integer_sequence<int, 16>{}), 15);
            const int obs = __shfl(bs, own, 16), oex = __shfl(excl, own, 16), ol = __shfl(L, own, 16);
            const float oa = __shfl(a, own, 16);
            if (t >= 0 && t < tot && t - oex >= 0 && t - oex < ol) {
                const int q = obs + (t - oex);
                const int cq = B.col[q];
                if ((unsigned)cq < (unsigned)B.n) {
                    col = cq;
                    val = double(oa) * double(B.val ? B.val[q] : 1.0f);
                }
            }
            pbase += tot;
        }
        // the most products of any group of the wave
        const int mp = max(max(__builtin_amdgcn_readlane(pbase, 0), __builtin_amdgcn_readlane(pbase, 16)),
                           max(__builtin_amdgcn_readlane(pbase, 32), __builtin_amdgcn_readlane(pbase, 48)));
        // sorting network on (column, product): each merge of two sorted
        // blocks of k/2 compares mirrored lanes, then halves; the stages
        // for blocks of k are needed only when some group has more than
        // k/2 products (the rest of the lanes hold (INT_MAX, lane), in order)
        auto x1 = [](int x) { return row_dpp<DPP_XOR1>(x); };
        auto x2 = [](int x) { return row_dpp<DPP_XOR2>(x); };
        auto x4 = [lane](int x) {
            const int up = row_dpp<DPP_SHR0 + 4>(x), down = row_dpp<DPP_SHL0 + 4>(x);
            return (lane & 4) ? up : down;
        };
        if (mp > 1) row16_cx(col, prod, (lane & 1) == 0, x1);
        if (mp > 2) {
            row16_cx(col, prod, (lane & 2) == 0, [](int x) { return row_dpp<DPP_XOR3>(x); });
            row16_cx(col, prod, (lane & 1) == 0, x1);
        }
        if (mp > 4) {
            row16_cx(col, prod, (lane & 4) == 0, [](int x) { return row_dpp<DPP_HALF_MIRROR>(x); });
            row16_cx(col, prod, (lane & 2) == 0, x2);
            row16_cx(col, prod, (lane & 1) == 0, x1);
        }
        if (mp > 8) {
            row16_cx(col, prod, (lane & 8) == 0, [](int x) { return row_dpp<DPP_MIRROR>(x); });
            row16_cx(col, prod, (lane & 4) == 0, x4);
            row16_cx(col, prod, (lane & 2) == 0, x2);
            row16_cx(col, prod, (lane & 1) == 0, x1);
        }
        val = __shfl(val, prod, 16);  // the products stayed on their lanes
        const int prev = row_dpp<DPP_SHR0 + 1>(col);
        const bool start = col != INT_MAX && (lane == 0 || prev != col);
        const double acc = row16_run_sum(col, val, lane, mp, std::make_integer_sequence<int, 15>{});
        const uint64_t ball = __ballot(start);
        const uint32_t mine = uint32_t(ball >> (grp * 16)) & 0xffffu;
        const int rank = __popc(mine & ((1u << lane) - 1u));
        if (live && lane == 0) nnzc[i] = __popc(mine);
        if (NUMERIC && live && start && rank < pbase) {  // runs <= products <= ub
            const int o = crp[i] + rank;
            ccol[o] = col;
            cval[o] = float(acc);
        }
    }
}

// Rows of 17 .. 64 products, one to a wave, in registers: the same steps
// as k_rows_tiny over 64 lanes.  The owner search reads the scan through
// readlane (wave-uniform), the sorting network's steps inside a 16-lane row
// are DPP moves and its three steps across rows (mirrors over 32 and 64
// lanes, xor 16) are lane permutes; the run sums stop once no run goes on.
template <bool NUMERIC>
__global__ __launch_bounds__(256) void k_rows_wave(const int32_t *__restrict__ rows, int nrows, CsrD A, CsrD B,
                                                   int32_t *nnzc, const int32_t *__restrict__ crp, int32_t *ccol,
                                                   float *cval) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int wid = (blockIdx.x * 256 + threadIdx.x) / WAVE, nw = gridDim.x * (256 / WAVE);
    for (int r = wid; r < nrows; r += nw) {  // wave-uniform
        const int i = rows[r];
        const int ab = A.rp[i], ae = A.rp[i + 1];
        int col = INT_MAX, prod = lane;  // this lane's product: column, index in the row
        double val = 0.0;
        int pbase = 0;
        for (int cb = ab;; cb += WAVE) {
            const uint64_t have = __ballot(cb + lane < ae);
            if (!have) break;
            const int ne = __popcll(have);
            const int j = cb + lane;
            int L = 0, bs = 0;
            float a = 0.f;
            if (j < ae) {
                const int k = A.col[j];
                if ((unsigned)k < (unsigned)B.m) {
                    a = A.val ? A.val[j] : 1.0f;
                    bs = B.rp[k];
                    L = B.rp[k + 1] - bs;
                }
            }
            const int incl = int(wave_incl_sum_dpp(uint32_t(L))), excl = incl - L;
            const int tot = wave_last(incl);
            const int t = lane - pbase;  // products pbase .. pbase + tot - 1 sit on those lanes
            int own = 0;
            for (int e = 0; e < ne; ++e) own += __builtin_amdgcn_readlane(incl, e) <= t ? 1 : 0;
            own = min(own, WAVE - 1);
            const int obs = __shfl(bs, own), oex = __shfl(excl, own), ol = __shfl(L, own);
            const float oa = __shfl(a, own);
            if (t >= 0 && t < tot && t - oex >= 0 && t - oex < ol) {
                const int q = obs + (t - oex);
                const int cq = B.col[q];
                if ((unsigned)cq < (unsigned)B.n) {
                    col = cq;
                    val = double(oa) * double(B.val ? B.val[q] : 1.0f);
                }
            }
            pbase += tot;
        }
        const int mp = pbase;  // the row's products (wave-uniform)
        auto x1 = [](int x) { return row_dpp<DPP_XOR1>(x); };
        auto x2 = [](int x) { return row_dpp<DPP_XOR2>(x); };
        auto x4 = [lane](int x) {
            const int up = row_dpp<DPP_SHR0 + 4>(x), down = row_dpp<DPP_SHL0 + 4>(x);
            return (lane & 4) ? up : down;
        };
        auto x8 = [lane](int x) {
            const int up = row_dpp<DPP_SHR0 + 8>(x), down = row_dpp<DPP_SHL0 + 8>(x);
            return (lane & 8) ? up : down;
        };
        auto perm = [lane](int mask) { return [lane, mask](int x) { return __shfl(x, lane ^ mask); }; };
        if (mp > 1) row16_cx(col, prod, (lane & 1) == 0, x1);
        if (mp > 2) {
            row16_cx(col, prod, (lane & 2) == 0, [](int x) { return row_dpp<DPP_XOR3>(x); });
            row16_cx(col, prod, (lane & 1) == 0, x1);
        }
        if (mp > 4) {
            row16_cx(col, prod, (lane & 4) == 0, [](int x) { return row_dpp<DPP_HALF_MIRROR>(x); });
            row16_cx(col, prod, (lane & 2) == 0, x2);
            row16_cx(col, prod, (lane & 1) == 0, x1);
        }
        if (mp > 8) {
            row16_cx(col, prod, (lane & 8) == 0, [](int x) { return row_dpp<DPP_MIRROR>(x); });
            row16_cx(col, prod, (lane & 4) == 0, x4);
            row16_cx(col, prod, (lane & 2) == 0, x2);
            row16_cx(col, prod, (lane & 1) == 0, x1);
        }
        if (mp > 16) {
            row16_cx(col, prod, (lane & 16) == 0, perm(31));
            row16_cx(col, prod, (lane & 8) == 0, x8);
            row16_cx(col, prod, (lane & 4) == 0, x4);
            row16_cx(col, prod, (lane & 2) == 0, x2);
            row16_cx(col, prod, (lane & 1) == 0, x1);
        }
        if (mp > 32) {
            row16_cx(col, prod, (lane & 32) == 0, perm(63));
            row16_cx(col, prod, (lane & 16) == 0, perm(16));
            row16_cx(col, prod, (lane & 8) == 0, x8);
            row16_cx(col, prod, (lane & 4) == 0, x4);
            row16_cx(col, prod, (lane & 2) == 0, x2);
            row16_cx(col, prod, (lane & 1) == 0, x1);
        }
        val = __shfl(val, prod);  // the products stayed on their lanes
        const int prev = __shfl_up(col, 1);
        const bool start = col != INT_MAX && (lane == 0 || prev != col);
        double acc = val;
        bool run = true;
        for (int d = 1; d < mp; ++d) {  // wave-uniform: every lane shuffles
            const int nc = __shfl_down(col, d);
            const double nv = __shfl_down(val, d);
            run = run && lane + d < WAVE && nc == col;
            if (run) acc += nv;
            if (!__ballot(run)) break;
        }
        const uint64_t ball = __ballot(start);
        const int rank = __builtin_amdgcn_mbcnt_hi(uint32_t(ball >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(ball), 0));
        if (lane == 0) nnzc[i] = __popcll(ball);
        if (NUMERIC && start && rank < mp) {  // runs <= products <= ub
            const int o = crp[i] + rank;
            ccol[o] = col;
            cval[o] = float(acc);
        }
    }
}

// The rows with the most products, by expand-sort-compress instead of
// global hash tables (whose inserts are device-scope atomics, a few G/s):
// row r of the list writes its products at poff[r] in the order of A's
// entries, then of B's row: key = column (B.n for a column outside B,
// dropped later as the hash tables skip it), value a * b in FP64.  A stable
// sort by column inside each row's range (cvt::seg_radix_kv, the values
// moving with the keys) leaves each output entry's products in increasing k,
// so the FP64 run sums are the tables' and Gustavson's.  (The full radix
// sort's path; the windowed path expands through k_long_dscatter.)
template <int T>
__global__ __launch_bounds__(T) void k_long_expand(const int32_t *__restrict__ rows, int nrows, CsrD A, CsrD B,
                                                   const int32_t *__restrict__ poff, int32_t *key, double *pval) {
    __shared__ int32_t s_bs[T];
    __shared__ int32_t s_pre[T];
    __shared__ float s_a[T];
    __shared__ int32_t s_tmp[T / WAVE + 1];
    const int tid = threadIdx.x;
    for (int r = blockIdx.x; r < nrows; r += gridDim.x) {
        const int i = rows[r];
        int base = poff[r];
        const int ab = A.rp[i], ae = A.rp[i + 1];
        for (int jb = ab; jb < ae; jb += T) {
            const int j = jb + tid;
            int L = 0, bs = 0;
            float a = 0.f;
            if (j < ae) {
                const int k = A.col[j];
                if ((unsigned)k < (unsigned)B.m) {
                    a = A.val ? A.val[j] : 1.0f;
                    bs = B.rp[k];
                    L = B.rp[k + 1] - bs;
                }
            }
            int tot;
            const int pre = block_excl_scan<T>(L, s_tmp, tot);
            s_bs[tid] = bs;
            s_pre[tid] = pre;
            s_a[tid] = a;
            const int ne = min(T, ae - jb);
            __syncthreads();
            int lo = 0;
            for (int t = tid; t < tot; t += T) {
                int hi = ne - 1;  // last entry whose prefix <= t (t only grows)
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (s_pre[mid] <= t) lo = mid;
                    else hi = mid - 1;
                }
                const int q = s_bs[lo] + (t - s_pre[lo]);
                const int c = B.col[q];
                const int o = base + t;
                key[o] = (unsigned)c < (unsigned)B.n ? c : B.n;
                pval[o] = double(s_a[lo]) * double(B.val ? B.val[q] : 1.0f);
            }
            base += tot;
            __syncthreads();
        }
    }
}

// order[0, ng) = the long rows by decreasing bit length of their product
// count (rows of one length in any order), one workgroup; also clears the
// three row counters of the windowed path.
__global__ __launch_bounds__(1024) void k_long_order(int ng, const int32_t *__restrict__ gub, int32_t *order,
                                                     unsigned *next) {
    __shared__ int32_t s_cnt[33];
    const int tid = threadIdx.x;
    if (tid < 33) s_cnt[tid] = 0;
    if (tid < 3) next[tid] = 0;
    __syncthreads();
    for (int r = tid; r < ng; r += 1024) atomicAdd(&s_cnt[32 - __clz(gub[r])], 1);
    __syncthreads();
    if (tid == 0) {
        int run = 0;
        for (int b = 32; b >= 0; --b) {
            const int c = s_cnt[b];
            s_cnt[b] = run;
            run += c;
        }
    }
    __syncthreads();
    for (int r = tid; r < ng; r += 1024) order[atomicAdd(&s_cnt[32 - __clz(gub[r])], 1)] = r;
}

// A chunk of up to T * EPT entries of A(i,:) from jb (thread t takes the EPT
// consecutive entries from jb + t * EPT, their loads issued together): each
// entry's B row start, its products' prefix within the chunk and its A
// value into LDS; returns the chunk's products.  Contains barriers; the
// caller synchronises before reading the arrays.
template <int T, int EPT>
__device__ __forceinline__ int chunk_entries(const CsrD &A, const CsrD &B, int jb, int ae, int32_t *s_bs,
                                             int32_t *s_pre, float *s_a, int32_t *s_tmp) {
    const int tid = threadIdx.x;
    int k[EPT], bs[EPT], L[EPT];
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const int j = jb + tid * EPT + e;
        k[e] = j < ae ? A.col[j] : -1;
    }
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        bs[e] = 0;
        L[e] = 0;
        if ((unsigned)k[e] < (unsigned)B.m) {
            bs[e] = B.rp[k[e]];
            L[e] = B.rp[k[e] + 1] - bs[e];
        }
    }
    int sum = 0;
#pragma unroll
    for (int e = 0; e < EPT; ++e) sum += L[e];
    int tot;
    int pre = block_excl_scan<T>(sum, s_tmp, tot);
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const int j = jb + tid * EPT + e;
        s_bs[tid * EPT + e] = bs[e];
        s_pre[tid * EPT + e] = pre;
        if (s_a) s_a[tid * EPT + e] = (j < ae && A.val) ? A.val[j] : 1.0f;
        pre += L[e];
    }
    return tot;
}

// The long rows' expansion fused with the first radix pass on the top bits
// of the column (digit = key >> shift, below 256 since key <= ncol <
// 2^(shift + 8)): k_long_dhist counts each row's digits from B's columns
// alone, k_long_dscatter expands the row again in k order and writes each
// product (key, FP64 value) straight to its place among the row's digit
// groups, stable.  A workgroup a row, rows taken from a counter in `order`
// (largest first); both kernels make the same products in the same order.
template <int T, int U, int EPT>
__global__ __launch_bounds__(T) void k_long_dhist(int nrows, const int32_t *__restrict__ order,
                                                  const int32_t *__restrict__ rows, CsrD A, CsrD B, int shift,
                                                  int32_t *hist, unsigned *next) {
    constexpr int NW = T / WAVE;
    static_assert(T == 256, "one thread per digit");
    __shared__ int32_t s_h[NW][256];
    __shared__ int32_t s_bs[T * EPT], s_pre[T * EPT];
    __shared__ int32_t s_tmp[NW + 1];
    __shared__ int32_t s_row;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
    const int ncol = B.n;
    for (;;) {
        if (tid == 0) s_row = int(atomicAdd(next, 1u));
#pragma unroll
        for (int w = 0; w < NW; ++w) s_h[w][tid] = 0;
        __syncthreads();
        const int o = s_row;
        __syncthreads();  // s_row is read before it is written again
        if (o >= nrows) break;
        const int r = order[o];
        const int i = rows[r];
        const int ab = A.rp[i], ae = A.rp[i + 1];
        for (int jb = ab; jb < ae; jb += T * EPT) {
            const int tot = chunk_entries<T, EPT>(A, B, jb, ae, s_bs, s_pre, nullptr, s_tmp);
            const int ne = min(T * EPT, ae - jb);
            __syncthreads();
            int lo = 0;
            for (int tb = 0; tb < tot; tb += U * T) {  // uniform over the workgroup
                int cq[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int t = tb + u * T + tid;
                    cq[u] = -1;
                    if (t < tot) {
                        int hi = ne - 1;  // last entry whose prefix <= t (t only grows)
                        while (lo < hi) {
                            const int mid = (lo + hi + 1) >> 1;
                            if (s_pre[mid] <= t) lo = mid;
                            else hi = mid - 1;
                        }
                        cq[u] = B.col[s_bs[lo] + (t - s_pre[lo])];
                    }
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {  // one add per digit a wave holds
                    const bool valid = tb + u * T + tid < tot;
                    const int d = valid ? ((unsigned)cq[u] < (unsigned)ncol ? cq[u] : ncol) >> shift : 0;
                    const uint64_t peers = digit_peers<8>(valid, d);
                    if (valid && __popcll(peers & ((1ull << lane) - 1)) == 0)
                        atomicAdd(&s_h[wid][d], __popcll(peers));
                }
            }
            __syncthreads();  // the next chunk's scan rewrites s_bs / s_pre
        }
        int c = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) c += s_h[w][tid];
        hist[int64_t(r) * 256 + tid] = c;
        __syncthreads();  // the next row clears the counts
    }
}

template <int T, int U, int EPT>
__global__ __launch_bounds__(T) void k_long_dscatter(int nrows, const int32_t *__restrict__ order,
                                                     const int32_t *__restrict__ rows, CsrD A, CsrD B,
                                                     const int32_t *__restrict__ poff, int shift,
                                                     const int32_t *__restrict__ hist, int32_t *key_out,
                                                     double *val_out, unsigned *next) {
    constexpr int NW = T / WAVE, C = T * U;  // C products a step; wave w holds [w * 64U, (w + 1) * 64U)
    static_assert(T == 256, "one thread per digit");
    __shared__ int32_t s_cnt[NW][256];
    __shared__ int32_t s_cur[256];   // the digit's next place in the row
    __shared__ int32_t s_goff[256];  // the digit's place less its start in the step's staging
    __shared__ int32_t s_k[C];
    __shared__ double s_v[C];
    __shared__ int32_t s_bs[T * EPT], s_pre[T * EPT];
    __shared__ float s_a[T * EPT];
    __shared__ int32_t s_tmp[NW + 1];
    __shared__ int32_t s_row;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
    const uint64_t lt = (1ull << lane) - 1;
    const int ncol = B.n;
    for (;;) {
        if (tid == 0) s_row = int(atomicAdd(next, 1u));
        __syncthreads();
        const int o = s_row;
        __syncthreads();  // s_row is read before it is written again
        if (o >= nrows) break;
        const int r = order[o];
        const int i = rows[r];
        {
            const int h = hist[int64_t(r) * 256 + tid];
            int tot;
            const int ex = block_excl_scan<T>(h, s_tmp, tot);
            s_cur[tid] = poff[r] + ex;
        }
        const int ab = A.rp[i], ae = A.rp[i + 1];
        for (int jb = ab; jb < ae; jb += T * EPT) {
            const int tot = chunk_entries<T, EPT>(A, B, jb, ae, s_bs, s_pre, s_a, s_tmp);
            const int ne = min(T * EPT, ae - jb);
            int lo = 0;
            for (int tb = 0; tb < tot; tb += C) {  // uniform over the workgroup
#pragma unroll
                for (int w = 0; w < NW; ++w) s_cnt[w][tid] = 0;
                __syncthreads();  // (also publishes s_bs / s_pre / s_a)
                int k[U];
                double v[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int t = tb + wid * (WAVE * U) + u * WAVE + lane;
                    k[u] = -1;
                    v[u] = 0.0;
                    if (t < tot) {
                        int hi = ne - 1;  // last entry whose prefix <= t (t only grows)
                        while (lo < hi) {
                            const int mid = (lo + hi + 1) >> 1;
                            if (s_pre[mid] <= t) lo = mid;
                            else hi = mid - 1;
                        }
                        const int q = s_bs[lo] + (t - s_pre[lo]);
                        k[u] = B.col[q];
                        v[u] = double(s_a[lo]) * double(B.val ? B.val[q] : 1.0f);
                    }
                }
                uint64_t pe[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {  // one add per digit a wave holds
                    const bool valid = tb + wid * (WAVE * U) + u * WAVE + lane < tot;
                    if (valid) k[u] = (unsigned)k[u] < (unsigned)ncol ? k[u] : ncol;
                    pe[u] = digit_peers<8>(valid, valid ? k[u] >> shift : 0);
                    if (valid && __popcll(pe[u] & lt) == 0) atomicAdd(&s_cnt[wid][k[u] >> shift], __popcll(pe[u]));
                }
                __syncthreads();
                {  // thread d: the digit's start in the staging, then each wave's
                    int dt = 0;
#pragma unroll
                    for (int w = 0; w < NW; ++w) dt += s_cnt[w][tid];
                    int t2;
                    int run = block_excl_scan<T>(dt, s_tmp, t2);
                    s_goff[tid] = s_cur[tid] - run;
                    s_cur[tid] += dt;
#pragma unroll
                    for (int w = 0; w < NW; ++w) {
                        const int c = s_cnt[w][tid];
                        s_cnt[w][tid] = run;
                        run += c;
                    }
                }
                __syncthreads();
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const bool valid = k[u] >= 0;
                    const int d = valid ? k[u] >> shift : 0;
                    const uint64_t peers = pe[u];
                    const int rank = __popcll(peers & lt);
                    const int pos = valid ? s_cnt[wid][d] + rank : 0;
                    if (valid && rank == 0) s_cnt[wid][d] = pos + __popcll(peers);
                    if (valid) {
                        s_k[pos] = k[u];
                        s_v[pos] = v[u];
                    }
                }
                __syncthreads();
                const int cnt = min(C, tot - tb);
                for (int q = tid; q < cnt; q += T) {
                    const int kk = s_k[q];
                    const int g = s_goff[kk >> shift] + q;
                    key_out[g] = kk;
                    val_out[g] = s_v[q];
                }
                __syncthreads();  // the next step clears the counts and refills the staging
            }
            __syncthreads();  // the next chunk's scan rewrites s_bs / s_pre / s_a
        }
    }
}

// One stable LSD radix pass in LDS over items [0, cnt) of a window: digit =
// (key >> sh) & 127, from (ik, iv) to (ok, ov).  Wave w takes items [w *
// PER, (w + 1) * PER) in rounds of 64; per-wave digit counts (one LDS add
// per digit a round holds), a digit-major scan over (digit, wave), then
// ranks inside a round from the same ballots over the digit's 7 bits.  Contains barriers; every thread of the workgroup calls it.
template <int T, int W>
__device__ __forceinline__ void lds_radix_pass(const int32_t *ik, const double *iv, int32_t *ok, double *ov, int cnt,
                                               int sh, int32_t (*s_cnt)[128], int32_t *s_tot) {
    constexpr int NW = T / WAVE, PER = W / NW, RR = PER / WAVE;
    static_assert(T >= 128 && PER % WAVE == 0, "pass shape");
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
    const uint64_t lt = (1ull << lane) - 1;
    for (int x = tid; x < NW * 128; x += T) s_cnt[x / 128][x % 128] = 0;
    __syncthreads();
    int k[RR];
    double v[RR];
#pragma unroll
    for (int r = 0; r < RR; ++r) {
        const int t = wid * PER + r * WAVE + lane;
        k[r] = t < cnt ? ik[t] : -1;
        v[r] = t < cnt ? iv[t] : 0.0;
    }
    uint64_t pe[RR];
#pragma unroll
    for (int r = 0; r < RR; ++r) {  // one add per digit a wave holds
        const bool valid = k[r] >= 0;
        pe[r] = digit_peers<7>(valid, valid ? (k[r] >> sh) & 127 : 0);
        if (valid && __popcll(pe[r] & lt) == 0) atomicAdd(&s_cnt[wid][(k[r] >> sh) & 127], __popcll(pe[r]));
    }
    __syncthreads();
    // thread d < 128: digit d's start (a DPP scan over the two waves of
    // digits), then each wave's start inside it
    int tot = 0;
    if (tid < 128) {
#pragma unroll
        for (int w = 0; w < NW; ++w) tot += s_cnt[w][tid];
    }
    const uint32_t incl = wave_incl_sum_dpp(uint32_t(tot));
    if (lane == WAVE - 1 && wid < 2) s_tot[wid] = int(incl);
    __syncthreads();
    if (tid < 128) {
        int run = int(incl) - tot + (wid == 1 ? s_tot[0] : 0);
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            const int c = s_cnt[w][tid];
            s_cnt[w][tid] = run;
            run += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RR; ++r) {
        const bool valid = k[r] >= 0;
        const int d = valid ? (k[r] >> sh) & 127 : 0;
        const uint64_t peers = pe[r];
        const int rank = __popcll(peers & lt);
        const int pos = valid ? s_cnt[wid][d] + rank : 0;
        if (valid && rank == 0) s_cnt[wid][d] = pos + __popcll(peers);
        if (valid) {
            ok[pos] = k[r];
            ov[pos] = v[r];
        }
    }
    __syncthreads();
}

// After one segmented radix pass on the top 8 bits of the column key
// (digit = key >> shift), a long row's products form groups of equal digit
// in increasing digit order.  A workgroup at a time per row (rows taken from
// a counter, largest first) sorts the row stably by the whole key, in place,
// one window of up to W products at a time: a window ends at a group
// boundary unless the row ends there, so the windows' keys are ordered among
// themselves and sorting each window alone sorts the row.  A window is
// sorted in LDS by LSD radix passes of 7 bits; its runs of equal columns
// (never split between windows: a window ends where the digit changes) are
// summed in FP64 in k order, and each run's column and sum are written to
// the row's staging at poff[r] + the runs of its earlier windows, in order
// (columns outside B, key ncol, are dropped); nnzc[rows[r]] gets the row's
// run count.  A group larger than a window (columns clustered in one digit
// range) raises *ovf and leaves the rest of its row undone: the caller then
// sorts the long rows by the full radix passes instead.
template <int T, int W>
__global__ __launch_bounds__(T) void k_long_lsort(int nrows, const int32_t *__restrict__ order,
                                                  const int32_t *__restrict__ rows, const int32_t *__restrict__ poff,
                                                  int shift, int bits, int ncol, const int32_t *__restrict__ key,
                                                  const double *__restrict__ val, int32_t *scol, float *sval,
                                                  int32_t *nnzc, int32_t *ovf, unsigned *next) {
    constexpr int NW = T / WAVE, R = W / T;
    __shared__ int32_t s_k[2][W];
    __shared__ double s_v[2][W];
    __shared__ int32_t s_cnt[NW][128];
    __shared__ int32_t s_tot[128];
    __shared__ int32_t s_cut, s_row;
    __shared__ int32_t s_wh[NW];
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
    for (;;) {
        if (tid == 0) s_row = int(atomicAdd(next, 1u));
        __syncthreads();
        const int o = s_row;
        __syncthreads();  // s_row is read before it is written again
        if (o >= nrows) break;
        const int r = order[o];
        const int rb = poff[r], re = poff[r + 1];
        int runs = 0;  // the row's runs so far (the same in every thread)
        bool done = true;
        for (int w0 = rb; w0 < re;) {
            const int n = min(W, re - w0);
            {  // every load issued before the stores
                int lk[R];
                double lv[R];
#pragma unroll
                for (int q = 0; q < R; ++q) {
                    const int i = q * T + tid;
                    if (i < n) {
                        lk[q] = key[w0 + i];
                        lv[q] = val[w0 + i];
                    }
                }
#pragma unroll
                for (int q = 0; q < R; ++q) {
                    const int i = q * T + tid;
                    if (i < n) {
                        s_k[0][i] = lk[q];
                        s_v[0][i] = lv[q];
                    }
                }
            }
            if (tid == 0) s_cut = w0 + n < re ? 0 : n;  // the row goes on: the last group's start
            __syncthreads();
            if (w0 + n < re) {
                const int dl = s_k[0][n - 1] >> shift;
                for (int i = tid; i < n; i += T)
                    if (i > 0 && (s_k[0][i] >> shift) == dl && (s_k[0][i - 1] >> shift) != dl) s_cut = i;
            }
            __syncthreads();
            const int cut = s_cut;
            if (cut == 0) {  // uniform: one group fills the window and goes on
                if (tid == 0) atomicOr(ovf, 1);
                done = false;
                break;
            }
            int cur = 0;
            for (int sh = 0; sh < bits; sh += 7, cur ^= 1)
                lds_radix_pass<T, W>(s_k[cur], s_v[cur], s_k[cur ^ 1], s_v[cur ^ 1], cut, sh, s_cnt, s_tot);
            // runs: thread t takes the sorted items [t * R, t * R + R)
            const int32_t *sk = s_k[cur];
            const double *sv = s_v[cur];
            uint32_t heads = 0;
            int nh = 0;
#pragma unroll
            for (int q = 0; q < R; ++q) {
                const int i = tid * R + q;
                if (i < cut && sk[i] < ncol && (i == 0 || sk[i] != sk[i - 1])) {
                    heads |= 1u << q;
                    ++nh;
                }
            }
            const uint32_t incl = wave_incl_sum_dpp(uint32_t(nh));
            if (lane == WAVE - 1) s_wh[wid] = int(incl);
            __syncthreads();
            int pos = int(incl) - nh, tot = 0;
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                if (w < wid) pos += s_wh[w];
                tot += s_wh[w];
            }
            const int64_t ob = int64_t(rb) + runs;
#pragma unroll
            for (int q = 0; q < R; ++q) {
                if (!((heads >> q) & 1u)) continue;
                const int i = tid * R + q;
                const int kk = sk[i];
                double acc = sv[i];
                for (int j = i + 1; j < cut && sk[j] == kk; ++j) acc += sv[j];  // in k order
                scol[ob + pos] = kk;
                sval[ob + pos] = float(acc);
                ++pos;
            }
            runs += tot;
            __syncthreads();  // the next window refills the staging
            w0 += cut;
        }
        if (done && tid == 0) nnzc[rows[r]] = runs;
    }
}

// Rows of at most W products by sorting instead of hashing: one workgroup a
// row (taken from a counter) expands the row's products into LDS in k order
// (column key, FP64 product; a column outside B gets the key ncol), sorts
// them stably by column in LDS (7-bit LSD passes), and sums each run of equal
// columns in FP64 in k order: nnzc[i] = the runs; NUMERIC also writes them
// at crp[i] in column order.  The sums are Gustavson's, so the rows come out
// identical to the oracle's.
template <int T, int W, bool NUMERIC>
__global__ __launch_bounds__(T) void k_rows_sort(const int32_t *__restrict__ rows, int nrows, CsrD A, CsrD B,
                                                 int bits, int32_t *nnzc, const int32_t *__restrict__ crp,
                                                 int32_t *ccol, float *cval, unsigned *next) {
    constexpr int NW = T / WAVE, R = W / T;
    __shared__ int32_t s_k[2][W];
    __shared__ double s_v[2][W];
    __shared__ int32_t s_cnt[NW][128];
    __shared__ int32_t s_tot[128];
    __shared__ int32_t s_bs[T], s_pre[T];
    __shared__ float s_a[T];
    __shared__ int32_t s_tmp[NW + 1], s_wh[NW];
    __shared__ int32_t s_row;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
    const int ncol = B.n;
    for (;;) {
        if (tid == 0) s_row = int(atomicAdd(next, 1u));
        __syncthreads();
        const int ri = s_row;
        __syncthreads();  // s_row is read before it is written again
        if (ri >= nrows) break;
        const int i = rows[ri];
        // 1. the products in k order
        int n = 0;
        const int ab = A.rp[i], ae = A.rp[i + 1];
        for (int jb = ab; jb < ae; jb += T) {
            const int j = jb + tid;
            int L = 0, bs = 0;
            float a = 0.f;
            if (j < ae) {
                const int k = A.col[j];
                if ((unsigned)k < (unsigned)B.m) {
                    a = A.val ? A.val[j] : 1.0f;
                    bs = B.rp[k];
                    L = B.rp[k + 1] - bs;
                }
            }
            int tot;
            const int pre = block_excl_scan<T>(L, s_tmp, tot);
            s_bs[tid] = bs;
            s_pre[tid] = pre;
            s_a[tid] = a;
            const int ne = min(T, ae - jb);
            __syncthreads();
            constexpr int U = 4;  // products per thread in flight
            int lo = 0;
            for (int tb = tid; tb < tot; tb += U * T) {
                int cq[U];
                double pq[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int t = tb + u * T;
                    cq[u] = -1;
                    if (t < tot) {
                        if (u == 0) {
                            int hi = ne - 1;  // last entry whose prefix <= t
                            while (lo < hi) {
                                const int mid = (lo + hi + 1) >> 1;
                                if (s_pre[mid] <= t) lo = mid;
                                else hi = mid - 1;
                            }
                        } else {
                            while (lo + 1 < ne && s_pre[lo + 1] <= t) ++lo;
                        }
                        const int q = s_bs[lo] + (t - s_pre[lo]);
                        cq[u] = B.col[q];
                        pq[u] = double(s_a[lo]) * double(B.val ? B.val[q] : 1.0f);
                    }
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int t = tb + u * T;
                    if (t < tot) {
                        s_k[0][n + t] = (unsigned)cq[u] < (unsigned)ncol ? cq[u] : ncol;
                        s_v[0][n + t] = pq[u];
                    }
                }
            }
            n += tot;
            __syncthreads();  // the next chunk's scan rewrites s_bs / s_pre
        }
        // 2. the sort (n <= W)
        int cur = 0;
        for (int sh = 0; sh < bits; sh += 7, cur ^= 1)
            lds_radix_pass<T, W>(s_k[cur], s_v[cur], s_k[cur ^ 1], s_v[cur ^ 1], n, sh, s_cnt, s_tot);
        // 3. the runs: thread t takes the sorted items [t * R, t * R + R)
        const int32_t *sk = s_k[cur];
        const double *sv = s_v[cur];
        uint32_t heads = 0;
        int nh = 0;
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int x = tid * R + q;
            if (x < n && sk[x] < ncol && (x == 0 || sk[x] != sk[x - 1])) {
                heads |= 1u << q;
                ++nh;
            }
        }
        const uint32_t incl = wave_incl_sum_dpp(uint32_t(nh));
        if (lane == WAVE - 1) s_wh[wid] = int(incl);
        __syncthreads();
        int pos = int(incl) - nh, tot = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            if (w < wid) pos += s_wh[w];
            tot += s_wh[w];
        }
        if (tid == 0) nnzc[i] = tot;
        if (NUMERIC) {
            const int o = crp[i];
#pragma unroll
            for (int q = 0; q < R; ++q) {
                if (!((heads >> q) & 1u)) continue;
                const int x = tid * R + q;
                const int kk = sk[x];
                double acc = sv[x];
                for (int y = x + 1; y < n && sk[y] == kk; ++y) acc += sv[y];  // in k order
                ccol[o + pos] = kk;
                cval[o + pos] = float(acc);
                ++pos;
            }
        }
        __syncthreads();  // the next row refills the staging
    }
}

// C(i,:) of the long rows from their staging (k_long_lsort): row r's runs
// at poff[r] go to crp[rows[r]], one workgroup a row.
__global__ __launch_bounds__(256) void k_long_copy(int nrows, const int32_t *__restrict__ rows,
                                                   const int32_t *__restrict__ poff, const int32_t *__restrict__ crp,
                                                   const int32_t *__restrict__ scol, const float *__restrict__ sval,
                                                   int32_t *ccol, float *cval) {
    constexpr int U = 4;  // loads in flight per thread
    for (int r = blockIdx.x; r < nrows; r += gridDim.x) {
        const int i = rows[r];
        const int c0 = crp[i], nz = crp[i + 1] - c0, s0 = poff[r];
        for (int xb = threadIdx.x; xb < nz; xb += U * 256) {
            int32_t c[U];
            float v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int x = xb + u * 256;
                if (x < nz) {
                    c[u] = scol[s0 + x];
                    v[u] = sval[s0 + x];
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int x = xb + u * 256;
                if (x < nz) {
                    ccol[c0 + x] = c[u];
                    cval[c0 + x] = v[u];
                }
            }
        }
    }
}

// flag[s] = 1 where a run of equal columns starts inside a long row, after
// the sort: s is its row's first product or its column differs from s - 1's.
// Products outside B's columns (key ncol, sorted last in their row) start
// none.  Row starts are marked by k_long_rowstart afterwards.
__global__ __launch_bounds__(256) void k_long_flags(int64_t n, const int32_t *__restrict__ key, int ncol,
                                                    int32_t *flag) {
    for (int64_t s = int64_t(blockIdx.x) * 256 + threadIdx.x; s < n; s += int64_t(gridDim.x) * 256) {
        const int k = key[s];
        flag[s] = k < ncol && (s == 0 || k != key[s - 1]) ? 1 : 0;
    }
}
__global__ __launch_bounds__(256) void k_long_rowstart(int nrows, const int32_t *__restrict__ poff,
                                                       const int32_t *__restrict__ key, int ncol, int32_t *flag) {
    for (int r = blockIdx.x * 256 + threadIdx.x; r < nrows; r += gridDim.x * 256) {
        const int s = poff[r];
        if (s < poff[r + 1]) flag[s] = key[s] < ncol ? 1 : 0;
    }
}

// nnz(C(i,:)) of the long rows: the runs inside their product ranges
__global__ __launch_bounds__(256) void k_long_nnz(int nrows, const int32_t *__restrict__ rows,
                                                  const int32_t *__restrict__ poff,
                                                  const int32_t *__restrict__ runpos, int32_t *nnzc) {
    for (int r = blockIdx.x * 256 + threadIdx.x; r < nrows; r += gridDim.x * 256)
        nnzc[rows[r]] = runpos[poff[r + 1]] - runpos[poff[r]];
}

// one output entry per run of a long row: its products (consecutive after
// the sort, in increasing k) summed in FP64 in that order, written at the
// row's offset in C.  The row of a run start s: the row holding product
// (s >> 11) << 11 (brow), then forward (every row holds over 256 products,
// so at most eight steps; two with the default threshold of 1024).
__global__ __launch_bounds__(256) void k_long_runs(int64_t n, int nrows, const int32_t *__restrict__ rows,
                                                   const int32_t *__restrict__ poff,
                                                   const int32_t *__restrict__ brow,
                                                   const int32_t *__restrict__ flag,
                                                   const int32_t *__restrict__ runpos,
                                                   const int32_t *__restrict__ key, const double *__restrict__ val,
                                                   const int32_t *__restrict__ crp, int32_t *ccol, float *cval) {
    for (int64_t s = int64_t(blockIdx.x) * 256 + threadIdx.x; s < n; s += int64_t(gridDim.x) * 256) {
        if (!flag[s]) continue;
        int r = brow[s >> 11];
        while (r + 1 < nrows && poff[r + 1] <= s) ++r;
        const int k = key[s];
        const int64_t end = poff[r + 1];
        double acc = val[s];
        for (int64_t x = s + 1; x < end && key[x] == k; ++x) acc += val[x];
        const int j = crp[rows[r]] + runpos[s] - runpos[poff[r]];
        ccol[j] = k;
        cval[j] = float(acc);
    }
}

// ubs[i] = the products of row i when the LDS tables take it (ub <= lub),
// else 0: the rows' room in the staging area of the one-pass product
__global__ __launch_bounds__(256) void k_stage_ub(int m, const int32_t *__restrict__ ub, int lub, int32_t *ubs) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < m; i += gridDim.x * 256) ubs[i] = ub[i] <= lub ? ub[i] : 0;
}

// C(i,:) of the rows of the LDS and register bins from the staging area.
// A workgroup takes 256 consecutive rows: their outputs [crp[r0], crp[r1])
// are contiguous, so its threads stride over them (coalesced writes) and
// find each output's row by a binary search over the block's row pointers
// in LDS; the outputs of long rows (written by their own path) are skipped.
__global__ __launch_bounds__(256) void k_unstage(int m, const int32_t *__restrict__ ub, int lub,
                                                 const int32_t *__restrict__ ubpre, const int32_t *__restrict__ crp,
                                                 const int32_t *__restrict__ scol, const float *__restrict__ sval,
                                                 int32_t *ccol, float *cval) {
    __shared__ int32_t s_rp[257], s_src[256];
    for (int r0 = blockIdx.x * 256; r0 < m; r0 += gridDim.x * 256) {
        const int nr = min(256, m - r0);
        __syncthreads();  // the previous block's searches are done
        for (int x = threadIdx.x; x <= nr; x += 256) s_rp[x] = crp[r0 + x];
        if (threadIdx.x < nr) {
            const int i = r0 + threadIdx.x;
            s_src[threadIdx.x] = ub[i] > lub ? -1 : ubpre[i];
        }
        __syncthreads();
        const int hi = s_rp[nr];
        int row = 0;  // the last row x with s_rp[x] <= o (o only grows)
        for (int o = s_rp[0] + threadIdx.x; o < hi;) {
            int top = nr - 1;
            while (row < top) {
                const int mid = (row + top + 1) >> 1;
                if (s_rp[mid] <= o) row = mid;
                else top = mid - 1;
            }
            const int src = s_src[row];
            if (src < 0) {  // a long row: past its outputs, same stride class
                o += (s_rp[row + 1] - o + 255) / 256 * 256;
                continue;
            }
            const int q = src + (o - s_rp[row]);
            ccol[o] = scol[q];
            cval[o] = sval[q];
            o += 256;
        }
    }
}

// Device arrays of the long rows' expand-sort-compress, kept from the
// symbolic to the numeric phase.
struct LongRows {
    int nrows = 0;
    int64_t n = 0;  // products
    const int32_t *rows = nullptr;
    int32_t *poff = nullptr, *brow = nullptr, *key = nullptr, *flag = nullptr, *runpos = nullptr;
    double *val = nullptr;
    // the windowed sort's staging (k_long_lsort): each row's runs at poff[r]
    bool fused = false;
    int32_t *scol = nullptr;
    float *sval = nullptr;
    // host sources of the uploaded tables, kept until the stream syncs
    std::vector<int32_t> hpo, hbrow, htiles;
};

// Symbolic phase of the long rows: products, the sort, runs, nnz(C(i,:)).
// gub[r] = the products of row r of the list (their sum < 2^31).
// Workgroups of `kernel` resident on the device at once (its register and
// LDS use, from the runtime's occupancy query), capped by `want`: the grid of
// a kernel whose workgroups take rows from a counter.
template <class K>
unsigned resident_grid(const apspie_ctx *ctx, K kernel, int threads, int64_t want) {
    int per = 0;
    APSPIE_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, 0));
    return unsigned(std::max<int64_t>(1, std::min<int64_t>(want, int64_t(ctx->cu_count) * std::max(per, 1))));
}

LongRows long_symbolic(apspie_ctx *ctx, const int32_t *rows, const std::vector<int32_t> &gub, const int32_t *dgub,
                       const CsrD &A, const CsrD &B, int32_t *nnzc, int32_t *ovf) {
    LongRows L;
    const int ng = int(gub.size());
    int64_t n = 0;
    for (int r = 0; r < ng; ++r) n += gub[r];
    L.nrows = ng;
    L.n = n;
    L.rows = rows;
    auto S = [&](int k, int64_t bytes) { return ctx->slot(S_OUTER0 + k, size_t(std::max<int64_t>(bytes, 16))); };
    L.poff = (int32_t *)S(0, (int64_t(ng) + 1) * 4);
    int32_t *key[2] = {(int32_t *)S(2, n * 4), (int32_t *)S(3, n * 4)};
    double *val[2] = {(double *)S(4, n * 8), (double *)S(5, n * 8)};
    // ovf (device): the expansion fused with one radix pass on the top 8
    // bits, then each row's digit groups sorted in LDS (k_long_lsort; a group
    // too large for it raises *ovf); null: the expansion, then the full
    // radix sort
    const int bits = esc::bits_for(int64_t(B.n) + 1);
    if (ovf != nullptr && bits > 8) {
        // the rows' product offsets and their order by decreasing product
        // count (a row is one workgroup's work) made on the device, so
        // nothing waits for a copy from the host
        cvt::scan32(ctx, dgub, L.poff, ng);
        int32_t *dord = (int32_t *)S(7, int64_t(ng) * 4 + 256);
        unsigned *next = (unsigned *)(dord + ((ng + 63) / 64) * 64);  // three row counters
        k_long_order<<<1, 1024, 0, ctx->stream>>>(ng, dgub, dord, next);
        int32_t *hist = (int32_t *)S(6, int64_t(ng) * 256 * 4);
        auto kh = k_long_dhist<256, 8, 4>;
        auto ks = k_long_dscatter<256, 4, 4>;
        kh<<<resident_grid(ctx, kh, 256, ng), 256, 0, ctx->stream>>>(ng, dord, rows, A, B, bits - 8, hist, next + 1);
        ks<<<resident_grid(ctx, ks, 256, ng), 256, 0, ctx->stream>>>(ng, dord, rows, A, B, L.poff, bits - 8, hist,
                                                                      key[1], val[1], next + 2);
        // the staging: the buffers the expansion left free
        L.fused = true;
        L.scol = key[0];
        L.sval = (float *)val[0];
        k_long_lsort<256, 2048><<<std::max(1u, std::min((unsigned)ng, unsigned(ctx->cu_count) * 3)), 256, 0,
                                 ctx->stream>>>(ng, dord, rows, L.poff, bits - 8, bits, B.n, key[1], val[1], L.scol,
                                                L.sval, nnzc, ovf, next);
        APSPIE_HIP(hipGetLastError());
        return L;
    }
    // the row offsets and brow[b] = the row holding product b * 2048 from the
    // host (pinned staging)
    std::vector<int32_t> &po = L.hpo;
    po.assign(size_t(ng) + 1, 0);
    for (int r = 0; r < ng; ++r) po[r + 1] = po[r] + gub[r];
    std::vector<int32_t> &brow = L.hbrow;
    brow.resize(size_t(n >> 11) + 1);
    for (size_t b = 0, r = 0; b < brow.size(); ++b) {
        while (r + 1 < size_t(ng) && po[r + 1] <= (int64_t(b) << 11)) ++r;
        brow[b] = int32_t(r);
    }
    L.brow = (int32_t *)S(1, int64_t(brow.size()) * 4);
    int32_t *hs = (int32_t *)ctx->hslot(2, (size_t(ng) + 1 + brow.size()) * 4);
    memcpy(hs, po.data(), (size_t(ng) + 1) * 4);
    memcpy(hs + ng + 1, brow.data(), brow.size() * 4);
    APSPIE_HIP(hipMemcpyAsync(L.poff, hs, (size_t(ng) + 1) * 4, hipMemcpyHostToDevice, ctx->stream));
    APSPIE_HIP(hipMemcpyAsync(L.brow, hs + ng + 1, brow.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    k_long_expand<1024><<<std::max(1u, std::min((unsigned)ng, unsigned(ctx->cu_count) * 2)), 1024, 0,
                          ctx->stream>>>(rows, ng, A, B, L.poff, key[0], val[0]);
    APSPIE_HIP(hipGetLastError());
    const int c = cvt::seg_radix_kv(ctx, key, val, n, po, bits, L.htiles, 0);
    L.key = key[c];
    L.val = val[c];
    L.flag = key[c ^ 1];  // free after the last pass
    L.runpos = (int32_t *)S(6, (n + 1) * 4);
    const unsigned g = cvt::grid_for(ctx, n, 256);
    k_long_flags<<<g, 256, 0, ctx->stream>>>(n, L.key, B.n, L.flag);
    const unsigned gr = std::max(1u, std::min(ceil_div(ng, 256), 1024u));
    k_long_rowstart<<<gr, 256, 0, ctx->stream>>>(ng, L.poff, L.key, B.n, L.flag);
    cvt::scan32(ctx, L.flag, L.runpos, n);
    k_long_nnz<<<gr, 256, 0, ctx->stream>>>(ng, rows, L.poff, L.runpos, nnzc);
    APSPIE_HIP(hipGetLastError());
    return L;
}

void long_numeric(apspie_ctx *ctx, const LongRows &L, const int32_t *crp, int32_t *ccol, float *cval) {
    if (L.fused) {
        k_long_copy<<<std::max(1u, std::min((unsigned)L.nrows, unsigned(ctx->cu_count) * 8)), 256, 0, ctx->stream>>>(
            L.nrows, L.rows, L.poff, crp, L.scol, L.sval, ccol, cval);
        APSPIE_HIP(hipGetLastError());
        return;
    }
    k_long_runs<<<cvt::grid_for(ctx, L.n, 256), 256, 0, ctx->stream>>>(L.n, L.nrows, L.rows, L.poff, L.brow,
                                                                       L.flag, L.runpos, L.key, L.val, crp, ccol,
                                                                       cval);
    APSPIE_HIP(hipGetLastError());
}

struct Plan {
    int counts[NBINS] = {};
    int32_t *lists = nullptr;
    std::vector<int32_t> grows;
    int64_t gslots_total = 0;
};

template <bool NUMERIC>
void launch_bins(apspie_ctx *ctx, const Plan &P, const CsrD &A, const CsrD &B, int32_t *nnzc, const int32_t *crp,
                 int32_t *ccol, float *cval, int32_t *gkeys, double *gvals, const int64_t *goff,
                 const int32_t *gslots) {
    const int m = A.m;
    auto L = [&](int k) { return P.lists + k * int64_t(m); };
    // the register and hash-table bins stride over their rows; workgroups a
    // CU as measured on the 1M power-law A * A (the hash bins are slower
    // with all of their resident workgroups: 14 of 16 and 10 of 12)
    auto grid = [&](int64_t n, unsigned per_cu) {
        return unsigned(std::max<int64_t>(1, std::min<int64_t>(n, int64_t(ctx->cu_count) * per_cu)));
    };
    if (P.counts[0])
        k_rows_tiny<NUMERIC><<<grid(ceil_div(P.counts[0], 16), 16), 256, 0, ctx->stream>>>(L(0), P.counts[0], A, B,
                                                                                          nnzc, crp, ccol, cval);
    if (P.counts[1])
        k_rows_wave<NUMERIC><<<grid(ceil_div(P.counts[1], 256 / WAVE), 16), 256, 0, ctx->stream>>>(
            L(1), P.counts[1], A, B, nnzc, crp, ccol, cval);
    if (P.counts[2])
        k_rows<128, 512, NUMERIC, false><<<grid(P.counts[2], 14), 128, 0, ctx->stream>>>(
            L(2), P.counts[2], A, B, nnzc, crp, ccol, cval, nullptr, nullptr, nullptr, nullptr);
    if (P.counts[3])
        k_rows<128, 1024, NUMERIC, false><<<grid(P.counts[3], 10), 128, 0, ctx->stream>>>(
            L(3), P.counts[3], A, B, nnzc, crp, ccol, cval, nullptr, nullptr, nullptr, nullptr);
    // rows of 257 to 2048 products: sorted in LDS, a workgroup a row taken
    // from a counter (each bin its own)
    unsigned *next = nullptr;
    if (P.counts[4] || P.counts[5] || P.counts[6]) {
        next = (unsigned *)ctx->slot(S_OUTER0 + 9, 64);
        k_zero_i32<<<1, 64, 0, ctx->stream>>>((int32_t *)next, 16);
    }
    const int bits = esc::bits_for(int64_t(B.n) + 1);
    if (P.counts[4]) {
        auto k = k_rows_sort<128, 512, NUMERIC>;
        k<<<resident_grid(ctx, k, 128, P.counts[4]), 128, 0, ctx->stream>>>(L(4), P.counts[4], A, B, bits, nnzc, crp,
                                                                            ccol, cval, next + 4);
    }
    if (P.counts[5]) {
        auto k = k_rows_sort<256, 1024, NUMERIC>;
        k<<<resident_grid(ctx, k, 256, P.counts[5]), 256, 0, ctx->stream>>>(L(5), P.counts[5], A, B, bits, nnzc, crp,
                                                                            ccol, cval, next + 5);
    }
    if (P.counts[6]) {
        auto k = k_rows_sort<256, 2048, NUMERIC>;
        k<<<resident_grid(ctx, k, 256, P.counts[6]), 256, 0, ctx->stream>>>(L(6), P.counts[6], A, B, bits, nnzc, crp,
                                                                            ccol, cval, next + 6);
    }
    if (P.counts[LONG])
        k_rows<1024, 1, NUMERIC, true><<<grid(P.counts[LONG], 2), 1024, 0, ctx->stream>>>(
            P.lists + LONG * int64_t(m), P.counts[LONG], A, B, nnzc, crp, ccol, cval, gkeys, gvals, goff, gslots);
    APSPIE_HIP(hipGetLastError());
}

// row_ptr[0..n] = exclusive scan of n non-negative counts; returns the total.
// A total at or past the CSR limit (2^31 - 4096 entries, what every entry
// point accepts) runs on_fail and throws APSPIE_ELIMIT: row_ptr then holds
// wrapped offsets, and nothing may be allocated or written by them.
// Synchronises the stream.
template <class F>
int64_t rowptr_scan(apspie_ctx *ctx, const int32_t *counts, int n, int32_t *row_ptr, F &&on_fail) {
    const int nb = std::max(1, (int)ceil_div(n, 256 * 16));
    uint64_t *st = (uint64_t *)ctx->slot(S_STATUS, size_t(nb) * 8 + 64);
    unsigned *ctr = (unsigned *)(st + nb);
    unsigned long long *tot = (unsigned long long *)(st + nb + 1);
    k_zero_i32<<<std::max(1u, std::min(ceil_div(nb * 2 + 4, 256), 1024u)), 256, 0, ctx->stream>>>(
        (int32_t *)st, int64_t(nb) * 2 + 4);
    k_scan<<<nb, 256, 0, ctx->stream>>>(counts, row_ptr, n, st, ctr, tot);
    APSPIE_HIP(hipGetLastError());
    unsigned long long *htot = (unsigned long long *)((char *)ctx->hslot(0, 512) + 256);
    to_host(ctx, (const int32_t *)tot, 2, 0, 256);
    APSPIE_HIP(hipStreamSynchronize(ctx->stream));
    const unsigned long long total = *htot;
    if (total >= (1ull << 31) - 4096) {
        on_fail();
        throw Error(APSPIE_ELIMIT, "the result would hold " + std::to_string(total) +
                                       " entries; a CSR holds fewer than 2^31 - 4096");
    }
    return int64_t(total);
}

}  // namespace sg
}  // namespace apspie

using namespace apspie;

extern "C" int apspie_rowptr_from_counts(apspie_ctx *ctx, const int32_t *d_counts, int32_t n, int32_t *d_row_ptr,
                                         int64_t *total) {
    if (!ctx) return APSPIE_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    return guarded(ctx, [&] {
        APSPIE_CHECK(n >= 0 && d_row_ptr != nullptr && (n == 0 || d_counts != nullptr), APSPIE_EINVAL,
                     "bad arguments");
        const int64_t t = sg::rowptr_scan(ctx, d_counts, n, d_row_ptr, [] {});
        if (total) *total = t;
    });
}

extern "C" int apspie_spgemm(apspie_ctx *ctx, const apspie_csr *a, const apspie_csr *b, apspie_csr_owned *c,
                             apspie_stats *stats) {
    if (!ctx || !c) return APSPIE_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    memset(c, 0, sizeof(*c));
    return guarded(ctx, [&] {
        using namespace sg;
        validate_csr(a, false);
        validate_csr(b, false);
        APSPIE_CHECK(a->n == b->m, APSPIE_EINVAL, "inner dimensions differ");
        const int m = a->m;
        const CsrD A{a->row_ptr, a->col_ind, a->val, a->m, a->n};
        const CsrD B{b->row_ptr, b->col_ind, b->val, b->m, b->n};
        APSPIE_HIP(hipEventRecord(ctx->ev0, ctx->stream));
        // 1. product bounds (the products of each row of C)
        int32_t *ub = (int32_t *)ctx->slot(S_AUX0, (size_t(m) + 1) * 4);
        int32_t *nnzc = (int32_t *)ctx->slot(S_AUX1, (size_t(m) + 1) * 4);
        int32_t *lists = (int32_t *)ctx->slot(S_AUX2, size_t(NBINS) * (size_t(m) + 1) * 4);
        int32_t *counts = (int32_t *)ctx->slot(S_AUX3, 256);
        unsigned long long *products = (unsigned long long *)(counts + 16);
        const unsigned g256 = std::max(1u, std::min(ceil_div(std::max(m, 1), 256), 4096u));
        int32_t *bad = counts + 10;  // counts[0, NBINS), bad, products: 128 bytes
        k_zero_i32<<<1, 256, 0, ctx->stream>>>(counts, 64);
        k_row_ub<<<std::max(1u, std::min(ceil_div(std::max(m, 1), 256), (unsigned)ctx->cu_count * 16)), 256, 0,
                   ctx->stream>>>(m, a->row_ptr, a->col_ind, b->row_ptr, b->m, ub, bad);
        // 2. bins
        k_bin<<<std::max(1u, std::min(ceil_div(std::max(m, 1), 4096), (unsigned)ctx->cu_count * 2)), 256, 0,
                ctx->stream>>>(m, ub, nnzc, lists, counts, products);
        APSPIE_HIP(hipGetLastError());
        // the counts, the bad flag and the product total in one copy into
        // pinned memory; C's row pointers allocated while the kernels run
        int32_t *hc = (int32_t *)ctx->hslot(0, 512);
        to_host(ctx, counts, 32, 0, 0);
        struct FreeOnThrow {  // C's arrays go if anything below throws
            apspie_ctx *ctx;
            apspie_csr_owned *c;
            bool keep = false;
            ~FreeOnThrow() {
                if (!keep) apspie_csr_free(ctx, c);
            }
        } out_guard{ctx, c};
        APSPIE_HIP(hipMalloc(&c->row_ptr, (size_t(m) + 1) * 4));
        c->m = m;
        c->n = b->n;
        APSPIE_HIP(hipStreamSynchronize(ctx->stream));
        Plan P;
        P.lists = lists;
        int hcounts[NBINS];
        memcpy(hcounts, hc, sizeof(hcounts));
        const unsigned long long prod_total = *(const unsigned long long *)(hc + 16);
        APSPIE_CHECK(hc[10] == 0, APSPIE_ELIMIT, "a row of C has 2^31 or more products");
        for (int k = 0; k < NBINS; ++k) P.counts[k] = hcounts[k];
        // rows beyond lub products leave the LDS tables for the long rows'
        // path: the bins above spgemm_esc_min join the long list (each bin's
        // list region holds m + 1 rows, so the long region has room)
        int lub = BIN_UB[NBINS - 2];
        if (ctx->tune.spgemm_long_rows == 1)
            for (int k = LONG - 1; k > 0 && BIN_UB[k - 1] >= ctx->tune.spgemm_esc_min; --k) {
                lub = BIN_UB[k - 1];
                if (!P.counts[k]) continue;
                APSPIE_HIP(hipMemcpyAsync(lists + LONG * int64_t(m) + P.counts[LONG], lists + k * int64_t(m),
                                          size_t(P.counts[k]) * 4, hipMemcpyDeviceToDevice, ctx->stream));
                P.counts[LONG] += P.counts[k];
                P.counts[k] = 0;
            }
        // global tables for the largest rows: 2 * pow2ceil(ub) slots each
        int64_t *goff = nullptr;
        int32_t *gslots = nullptr, *gkeys = nullptr;
        double *gvals = nullptr;
        std::vector<int32_t> gub;
        bool esc_rows = false;
        int64_t gtot = 0;  // products of the long rows
        if (P.counts[LONG]) {
            const int ng = P.counts[LONG];
            // the product bounds of these rows only (gathered into the slot
            // that later holds their table sizes)
            gslots = (int32_t *)ctx->slot(S_AUX5, size_t(ng) * 4);
            k_gather<<<std::max(1u, std::min(ceil_div(ng, 256), 1024u)), 256, 0, ctx->stream>>>(
                lists + LONG * int64_t(m), ng, ub, gslots);
            int32_t *hg = (int32_t *)ctx->hslot(1, size_t(ng) * 4);
            to_host(ctx, gslots, ng, 1, 0);
            APSPIE_HIP(hipStreamSynchronize(ctx->stream));
            gub.assign(hg, hg + ng);
            for (int r = 0; r < ng; ++r) gtot += gub[r];
            // expand-sort-compress while the products index with int32
            esc_rows = ctx->tune.spgemm_long_rows == 1 && gtot < (int64_t(1) << 31) - 1;
        }
        if (P.counts[LONG] && !esc_rows) {
            const int ng = P.counts[LONG];
            std::vector<int64_t> off(ng);
            std::vector<int32_t> sl(ng);
            int64_t tot = 0;
            for (int r = 0; r < ng; ++r) {
                int64_t h = 1;
                while (h < gub[r]) h <<= 1;
                APSPIE_CHECK(4 * h < (int64_t(1) << 31), APSPIE_ELIMIT, "row with too many products");
                sl[r] = int32_t(4 * h);  // at most a quarter full, as the LDS tables
                off[r] = tot;
                tot += 4 * h;
            }
            goff = (int64_t *)ctx->slot(S_AUX4, size_t(ng) * 8);
            gkeys = (int32_t *)ctx->slot(S_AUX6, size_t(tot) * 4);
            gvals = (double *)ctx->slot(S_AUX7, size_t(tot) * 8);
            char *hs = (char *)ctx->hslot(3, size_t(ng) * 12);
            memcpy(hs, off.data(), size_t(ng) * 8);
            memcpy(hs + size_t(ng) * 8, sl.data(), size_t(ng) * 4);
            APSPIE_HIP(hipMemcpyAsync(goff, hs, size_t(ng) * 8, hipMemcpyHostToDevice, ctx->stream));
            APSPIE_HIP(hipMemcpyAsync(gslots, hs + size_t(ng) * 8, size_t(ng) * 4, hipMemcpyHostToDevice,
                                      ctx->stream));
        }
        // 3. the rows of the LDS tables in one pass while their products fit
        // int32 offsets: each sorted row goes to a staging area at the prefix
        // of the rows' products, with its nnz (else a counting pass now and
        // the product pass after the row pointers); the long rows by
        // expand-sort-compress when their products fit int32 indices, else in
        // global hash tables
        Plan P03 = P, P4 = P;
        P03.counts[LONG] = 0;
        for (int k = 0; k < LONG; ++k) P4.counts[k] = 0;
        const int64_t staged = int64_t(prod_total) - gtot;
        const bool one_pass = ctx->tune.spgemm_passes == 1 && staged < (int64_t(1) << 31) - 1;
        int32_t *ubpre = nullptr, *scol = nullptr;
        float *sval = nullptr;
        if (one_pass) {
            ubpre = (int32_t *)ctx->slot(S_OUTER0 + 12, (size_t(m) + 1) * 4);
            int32_t *ubs = (int32_t *)ctx->slot(S_OUTER0 + 13, size_t(std::max(m, 1)) * 4);
            scol = (int32_t *)ctx->slot(S_OUTER0 + 10, size_t(std::max<int64_t>(staged, 1)) * 4);
            sval = (float *)ctx->slot(S_OUTER0 + 11, size_t(std::max<int64_t>(staged, 1)) * 4);
            k_stage_ub<<<g256, 256, 0, ctx->stream>>>(m, ub, lub, ubs);
            cvt::scan32(ctx, ubs, ubpre, m);
            launch_bins<true>(ctx, P03, A, B, nnzc, ubpre, scol, sval, nullptr, nullptr, nullptr, nullptr);
        } else {
            launch_bins<false>(ctx, P03, A, B, nnzc, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
        }
        LongRows lr;
        int32_t *ovf = nullptr;
        int32_t *hovf = (int32_t *)((char *)ctx->hslot(0, 512) + 264);
        *hovf = 0;
        if (esc_rows) {
            ovf = (int32_t *)ctx->slot(S_OUTER0 + 8, 256);
            k_zero_i32<<<1, 256, 0, ctx->stream>>>(ovf, 1);
            lr = long_symbolic(ctx, lists + LONG * int64_t(m), gub, gslots, A, B, nnzc, ovf);
            to_host(ctx, ovf, 1, 0, 264);
        } else if (P.counts[LONG]) {
            launch_bins<false>(ctx, P4, A, B, nnzc, nullptr, nullptr, nullptr, gkeys, gvals, goff, gslots);
        }
        // 4. row pointers (the scan's synchronisation also brings the long
        // rows' overflow flag; on overflow they are sorted again by the full
        // radix passes and the scan repeated)
        int32_t total = int32_t(rowptr_scan(ctx, nnzc, m, c->row_ptr, [&] { apspie_csr_free(ctx, c); }));
        if (*hovf) {
            lr = long_symbolic(ctx, lists + LONG * int64_t(m), gub, gslots, A, B, nnzc, nullptr);
            total = int32_t(rowptr_scan(ctx, nnzc, m, c->row_ptr, [&] { apspie_csr_free(ctx, c); }));
        }
        c->nnz = total;
        if (hipMalloc(&c->col_ind, size_t(std::max(total, 1)) * 4) != hipSuccess ||
            hipMalloc(&c->val, size_t(std::max(total, 1)) * 4) != hipSuccess) {
            (void)hipGetLastError();
            throw Error(APSPIE_ENOMEM, "output allocation failed");
        }
        // 5. numeric
        if (one_pass)
            k_unstage<<<std::max(1u, std::min(ceil_div(std::max(m, 1), 256), (unsigned)ctx->cu_count * 8)), 256, 0,
                        ctx->stream>>>(m, ub, lub, ubpre, c->row_ptr, scol, sval, c->col_ind, c->val);
        else
            launch_bins<true>(ctx, P03, A, B, nnzc, c->row_ptr, c->col_ind, c->val, nullptr, nullptr, nullptr,
                              nullptr);
        if (esc_rows)
            long_numeric(ctx, lr, c->row_ptr, c->col_ind, c->val);
        else if (P.counts[LONG])
            launch_bins<true>(ctx, P4, A, B, nnzc, c->row_ptr, c->col_ind, c->val, gkeys, gvals, goff, gslots);
        APSPIE_HIP(hipGetLastError());
        APSPIE_HIP(hipEventRecord(ctx->ev1, ctx->stream));
        APSPIE_HIP(hipStreamSynchronize(ctx->stream));
        if (stats) {
            float ms = 0.f;
            APSPIE_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
            memset(stats, 0, sizeof(*stats));
            stats->edges = (int64_t)prod_total;  // products
            stats->vertices = total;             // nnz(C)
            stats->iterations = 1;
            stats->kernel_ms = ms;
            stats->bytes_alg = 8.0 * (double(m) + 1) + 16.0 * double(a->nnz) + 8.0 * double(prod_total) +
                               8.0 * double(total);
        }
        out_guard.keep = true;
    });
}
