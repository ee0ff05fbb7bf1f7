// semiring.hip -- one sparse-vector x sparse-matrix product over a chosen
// semiring (the reference's general mXv, src/mXv.cuh:110-389):
//
//   y[c] = (+)_{u present in x, (u, c) in A} x[u] (*) A(u, c)
//
//   APSPIE_SR_PLUS_TIMES  (+, *), identity 0; sums accumulate in FP64 (the
//                         products of FP32 inputs are exact there) and round
//                         to FP32 once
//   APSPIE_SR_MIN_PLUS    (min, +), identity 1.70141e38 (the reference's
//                         sentinel); present: -1 < x < 1e38 (bitify,
//                         src/spmspvSssp.cuh:31-36), as apspie_sssp_step
//   APSPIE_SR_MAX_TIMES   (max, *), identity -inf
//   APSPIE_SR_OR_AND      (or, and) on 0 / 1, identity 0
//
// The dense form takes x[m] (present: differs from the identity, MIN_PLUS by
// spmspvSssp's bitify rule; or, asked for, x > 0.5, the bitify of mXv and
// bfs, src/mXv.cuh:50-55) and writes y[n] with the identity where nothing
// reaches.
// The sparse form takes (index, value) pairs -- every listed index is present
// -- and returns the reached columns in ascending order with their values,
// as the reference's sparse mXv returns its reduced keys.
//
// Both run the traversal kernels' frontier compaction and edge-balanced
// expansion over the present entries; min / max combine through atomic
// minima / maxima on an order-preserving integer image of the floats, sums
// through FP64 atomic adds.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "frontier.hpp"
#include "internal.hpp"

namespace apspie {
namespace srg {

using namespace dev;
using trav::Bufs;
using trav::Control;
using trav::WG;
using trav::WGC;

// float -> uint32 whose unsigned order is the floats' order (no NaN)
__device__ __forceinline__ uint32_t ord(float f) {
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float unord(uint32_t o) {
    return __uint_as_float((o & 0x80000000u) ? (o & 0x7fffffffu) : ~o);
}

__device__ __forceinline__ float identity(int sr) {
    switch (sr) {
    case APSPIE_SR_MIN_PLUS: return APSPIE_SSSP_INF;
    case APSPIE_SR_MAX_TIMES: return -INFINITY;
    default: return 0.0f;
    }
}

struct PickPresent {
    const float *x;
    const uint8_t *listed;  // sparse form: the listed indices (x holds their values)
    int m, sr, presence;
    __device__ bool pick(int v, uint32_t &payload) const {
        if (v >= m) return false;
        const float xv = x[v];
        payload = __float_as_uint(xv);
        if (listed) return listed[v] != 0;
        if (presence == APSPIE_MXV_ABOVE_HALF) return xv > 0.5f;
        switch (sr) {
        case APSPIE_SR_MIN_PLUS: return xv > -1.0f && xv < 1e38f;
        case APSPIE_SR_MAX_TIMES: return xv != -INFINITY;
        default: return xv != 0.0f;
        }
    }
    __device__ void on_new(int) const {}
};

struct OpSemiring {
    static constexpr bool kVert = false, kVal = true;
    using Pre = float;  // the edge's value
    const float *val;
    int sr;
    double *ysum;     // PLUS_TIMES
    uint32_t *yord;   // MIN_PLUS / MAX_TIMES (order image), OR_AND (0 / 1)
    uint8_t *reached;  // sparse form: columns any present entry reaches
    __device__ Pre load(int, int p, uint32_t, int) const { return val ? val[p] : 1.0f; }
    __device__ void apply(int c, int, uint32_t xb, int, Pre w) const {
        const float x = __uint_as_float(xb);
        if (reached) reached[c] = 1;
        switch (sr) {
        case APSPIE_SR_PLUS_TIMES: atomicAdd(&ysum[c], double(x) * double(w)); break;
        case APSPIE_SR_MIN_PLUS: {
            const uint32_t o = ord(x + w);
            if (o < yord[c]) atomicMin(&yord[c], o);
            break;
        }
        case APSPIE_SR_MAX_TIMES: {
            const uint32_t o = ord(x * w);
            if (o > yord[c]) atomicMax(&yord[c], o);
            break;
        }
        default:
            if (x != 0.0f && w != 0.0f) yord[c] = 1u;
        }
    }
};

__global__ __launch_bounds__(256) void k_init(int sr, int n, double *ysum, uint32_t *yord, uint8_t *reached,
                                              int nblk, uint64_t *status, Control *C) {
    const int64_t gtid = int64_t(blockIdx.x) * 256 + threadIdx.x, stride = int64_t(gridDim.x) * 256;
    const uint32_t o0 = sr == APSPIE_SR_OR_AND ? 0u : ord(identity(sr));
    for (int64_t i = gtid; i < n; i += stride) {
        if (sr == APSPIE_SR_PLUS_TIMES) ysum[i] = 0.0;
        else yord[i] = o0;
        if (reached) reached[i] = 0;
    }
    for (int64_t i = gtid; i < 2 * int64_t(nblk); i += stride) status[i] = 0;
    control_clear(C, gtid, stride, ~0u);
}

__global__ __launch_bounds__(256) void k_finish(int sr, int n, const double *ysum, const uint32_t *yord, float *y) {
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
        if (sr == APSPIE_SR_PLUS_TIMES) y[i] = float(ysum[i]);
        else if (sr == APSPIE_SR_OR_AND) y[i] = yord[i] ? 1.0f : 0.0f;
        else y[i] = unord(yord[i]);
    }
}

// sparse input: x dense = identity, listed = 0, then the pairs scattered
__global__ __launch_bounds__(256) void k_fill_x(int sr, int m, float *x, uint8_t *listed) {
    const float id = identity(sr);
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < m; i += int64_t(gridDim.x) * 256) {
        x[i] = id;
        listed[i] = 0;
    }
}

__global__ __launch_bounds__(256) void k_scatter_x(const int32_t *__restrict__ ind, const float *__restrict__ val,
                                                   int nx, int m, float *x, uint8_t *listed, int32_t *bad) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < nx; i += gridDim.x * 256) {
        const int u = ind[i];
        if (u < 0 || u >= m) {
            *bad = 1;
            continue;
        }
        x[u] = val[i];
        listed[u] = 1;
    }
}

// reached columns in ascending order with their values: per 4096-column
// block a local scan, the block's offset by decoupled look-back
__global__ __launch_bounds__(256) void k_collect(int n, const uint8_t *__restrict__ reached,
                                                 const float *__restrict__ y, int32_t *yind, float *yval,
                                                 int32_t *ny, uint64_t *status, unsigned int *counter) {
    constexpr int ITEMS = 16, TILE_N = 256 * ITEMS;
    __shared__ int s_bid;
    __shared__ uint64_t s_w[4];
    __shared__ uint64_t s_pre;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    if (tid == 0) s_bid = (int)atomicAdd(counter, 1u);
    __syncthreads();
    const int bid = s_bid;
    const int64_t b0 = int64_t(bid) * TILE_N + int64_t(tid) * ITEMS;
    uint32_t f[ITEMS];
    uint64_t sum = 0;
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        f[k] = b0 + k < n ? uint32_t(reached[b0 + k] != 0) : 0u;
        sum += f[k];
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
        if (f[k]) {
            yind[run] = int32_t(b0 + k);
            yval[run] = y[b0 + k];
        }
        run += f[k];
    }
    if (bid == (int)gridDim.x - 1 && tid == 255) *ny = int32_t(run);
}

void check_sr(int sr) {
    APSPIE_CHECK(sr >= APSPIE_SR_PLUS_TIMES && sr <= APSPIE_SR_OR_AND, APSPIE_EINVAL, "unknown semiring");
}

// y = A^T x over the semiring; x dense in HBM (listed: the sparse form's
// index marks, else presence by value); fills `reached` when asked
void mxv(apspie_ctx *ctx, const apspie_csr *a, int sr, int presence, const float *x, const uint8_t *listed,
         float *y, uint8_t *reached, apspie_stats *stats) {
    const Bufs b = trav::get_bufs(ctx, a->m, a->nnz);
    double *ysum = (double *)ctx->slot(S_AUX4, size_t(std::max(a->n, 1)) * 8);
    uint32_t *yord = (uint32_t *)ysum;  // one of the two is used
    APSPIE_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    const unsigned gi = std::max(1u, std::min(ceil_div(std::max(std::max(a->n, a->m), 2 * b.nblk), 256), 4096u));
    k_init<<<gi, 256, 0, ctx->stream>>>(sr, a->n, ysum, yord, reached, b.nblk, b.status, b.C);
    trav::k_compact<PickPresent><<<b.nblk, WGC, 0, ctx->stream>>>(PickPresent{x, listed, a->m, sr, presence}, 0, a->m,
                                                                    b.nblk, 1, a->row_ptr, trav::queue_of(b),
                                                                    b.C);
    trav::k_expand<OpSemiring><<<b.expand_grid, WG, 0, ctx->stream>>>(OpSemiring{a->val, sr, ysum, yord, reached},
                                                                      1, a->n, b.C, trav::queue_of(b), a->col_ind);
    k_finish<<<gi, 256, 0, ctx->stream>>>(sr, a->n, ysum, yord, y);
    APSPIE_HIP(hipGetLastError());
    APSPIE_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    trav::fill_stats(ctx, b, trav::MODE_STEP, a->m, stats);
}

}  // namespace srg
}  // namespace apspie

using namespace apspie;

extern "C" {

int apspie_mxv(apspie_ctx *ctx, const apspie_csr *a, int semiring, int presence, const float *d_x, float *d_y,
               apspie_stats *stats) {
    if (!ctx) return APSPIE_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    return guarded(ctx, [&] {
        validate_csr(a, false);
        srg::check_sr(semiring);
        APSPIE_CHECK(d_x && d_y, APSPIE_EINVAL, "null vectors");
        APSPIE_CHECK(presence == APSPIE_MXV_NOT_IDENTITY || presence == APSPIE_MXV_ABOVE_HALF, APSPIE_EINVAL,
                     "unknown presence rule");
        srg::mxv(ctx, a, semiring, presence, d_x, nullptr, d_y, nullptr, stats);
    });
}

int apspie_mxv_sparse(apspie_ctx *ctx, const apspie_csr *a, int semiring, const int32_t *d_xind,
                      const float *d_xval, int32_t nx, int32_t *d_yind, float *d_yval, int32_t *ny,
                      apspie_stats *stats) {
    if (!ctx || !ny) return APSPIE_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    return guarded(ctx, [&] {
        using namespace srg;
        validate_csr(a, false);
        check_sr(semiring);
        APSPIE_CHECK(nx >= 0 && (nx == 0 || (d_xind && d_xval)) && d_yind && d_yval, APSPIE_EINVAL,
                     "null vectors");
        const int m = a->m, n = a->n;
        float *x = (float *)ctx->slot(S_AUX0, size_t(std::max(m, 1)) * 4);
        uint8_t *listed = (uint8_t *)ctx->slot(S_AUX1, size_t(std::max(m, 1)));
        float *y = (float *)ctx->slot(S_AUX2, size_t(std::max(n, 1)) * 4);
        uint8_t *reached = (uint8_t *)ctx->slot(S_AUX3, size_t(std::max(n, 1)) + 64);
        int32_t *bad = (int32_t *)ctx->slot(S_AUX5, 256);
        const unsigned gm = std::max(1u, std::min(ceil_div(std::max(m, 1), 256), 4096u));
        APSPIE_HIP(hipMemsetAsync(bad, 0, 8, ctx->stream));
        k_fill_x<<<gm, 256, 0, ctx->stream>>>(semiring, m, x, listed);
        if (nx)
            k_scatter_x<<<std::max(1u, std::min(ceil_div(nx, 256), 4096u)), 256, 0, ctx->stream>>>(
                d_xind, d_xval, nx, m, x, listed, bad);
        APSPIE_HIP(hipGetLastError());
        int32_t hbad = 0;
        APSPIE_HIP(hipMemcpyAsync(&hbad, bad, 4, hipMemcpyDeviceToHost, ctx->stream));
        APSPIE_HIP(hipStreamSynchronize(ctx->stream));
        APSPIE_CHECK(hbad == 0, APSPIE_EINVAL, "vector index out of range");
        mxv(ctx, a, semiring, APSPIE_MXV_NOT_IDENTITY, x, listed, y, reached, stats);
        const int nc = std::max(1, (int)ceil_div(n, 256 * 16));
        uint64_t *status = (uint64_t *)ctx->slot(S_AUX6, size_t(nc + 2) * 8);
        int32_t *dny = bad + 1;
        APSPIE_HIP(hipMemsetAsync(status, 0, size_t(nc + 2) * 8, ctx->stream));
        APSPIE_HIP(hipMemsetAsync(dny, 0, 4, ctx->stream));
        k_collect<<<nc, 256, 0, ctx->stream>>>(n, reached, y, d_yind, d_yval, dny, status,
                                               (unsigned *)(status + nc));
        APSPIE_HIP(hipGetLastError());
        APSPIE_HIP(hipMemcpyAsync(ny, dny, 4, hipMemcpyDeviceToHost, ctx->stream));
        APSPIE_HIP(hipStreamSynchronize(ctx->stream));
    });
}

}  // extern "C"
