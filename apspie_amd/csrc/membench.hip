// membench.hip -- STREAM-like device copy used to report the measured HBM
// bandwidth beside the nominal peak, and as the known-size kernel that
// calibrates the FETCH_SIZE / WRITE_SIZE counters.
#include <algorithm>

#include "internal.hpp"

namespace apspie {
namespace {

__global__ __launch_bounds__(256) void k_copy(const uint4 *__restrict__ a, uint4 *__restrict__ b, size_t n) {
    for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += size_t(gridDim.x) * 256) b[i] = a[i];
}

}  // namespace
}  // namespace apspie

using namespace apspie;

extern "C" int apspie_hbm_copy_bandwidth(apspie_ctx *ctx, size_t bytes, int iters, double *gbps) {
    if (!ctx || !gbps || iters < 1 || bytes < 16) return APSPIE_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    return guarded(ctx, [&] {
        const size_t n = bytes / 16;
        uint4 *a = nullptr, *b = nullptr;
        if (hipMalloc(&a, n * 16) != hipSuccess || hipMalloc(&b, n * 16) != hipSuccess) {
            (void)hipGetLastError();
            if (a) (void)hipFree(a);
            throw Error(APSPIE_ENOMEM, "copy buffers");
        }
        double best = 0.0;
        try {
            APSPIE_HIP(hipMemsetAsync(a, 0x5a, n * 16, ctx->stream));
            const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, size_t(ctx->cu_count) * 16);
            for (int it = 0; it < iters; ++it) {
                APSPIE_HIP(hipEventRecord(ctx->ev0, ctx->stream));
                k_copy<<<grid, 256, 0, ctx->stream>>>(a, b, n);
                APSPIE_HIP(hipEventRecord(ctx->ev1, ctx->stream));
                APSPIE_HIP(hipEventSynchronize(ctx->ev1));
                float ms = 0.f;
                APSPIE_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
                best = std::max(best, 2.0 * double(n * 16) / (double(ms) * 1e-3) / 1e9);
            }
        } catch (...) {
            (void)hipFree(a);
            (void)hipFree(b);
            throw;
        }
        APSPIE_HIP(hipFree(a));
        APSPIE_HIP(hipFree(b));
        *gbps = best;
    });
}
