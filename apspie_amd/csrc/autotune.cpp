// autotune.cpp -- the project's autotuner: picks the traversal knobs of
// apspie_tuning (direction thresholds, expansion launch shape, round chunks,
// batched lane width) for one graph by timing candidates on it.
//
// Coordinate descent: the knobs are visited in a fixed order, each candidate
// value is timed (median wall time of three calls after one warm-up) with the
// other knobs at their current best, and the best value is kept before the
// next knob.  Knobs that several kinds share are scored on the sum of their
// times.  Results never depend on the knobs, only speed does.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <functional>
#include <vector>

#include "internal.hpp"

using namespace apspie;

namespace {

double wall_ms(apspie_ctx *ctx, const std::function<int()> &call) {
    const auto t0 = std::chrono::steady_clock::now();
    const int rc = call();
    const auto t1 = std::chrono::steady_clock::now();
    if (rc != APSPIE_OK) throw Error(rc, "autotune: " + ctx->err);
    return std::chrono::duration<double, std::milli>(t1 - t0).count();
}

double median3(apspie_ctx *ctx, const std::function<int()> &call) {
    wall_ms(ctx, call);  // warm-up: buffers, cached graphs
    double t[3];
    for (double &x : t) x = wall_ms(ctx, call);
    std::sort(t, t + 3);
    return t[1];
}

}  // namespace

extern "C" int apspie_autotune(apspie_ctx *ctx, const apspie_csr *g, int32_t src, int32_t k, uint32_t kinds,
                               apspie_tuning *out) {
    if (!ctx || !g || !out) return APSPIE_EINVAL;
    void *d_out = nullptr, *d_batch = nullptr, *d_srcs = nullptr;
    apspie_tuning best{}, orig{};
    apspie_ctx_get_tuning(ctx, &orig);
    best = orig;
    const int rc = guarded(ctx, [&] {
        validate_csr(g, false);
        APSPIE_CHECK(g->m == g->n && g->m > 0, APSPIE_EINVAL, "autotune needs a square, non-empty graph");
        APSPIE_CHECK(src >= 0 && src < g->m, APSPIE_EINVAL, "source vertex out of range");
        const bool batch = kinds & (APSPIE_TUNE_BFS_BATCH | APSPIE_TUNE_SSSP_BATCH);
        const int kb = std::min(k, (kinds & APSPIE_TUNE_BFS_BATCH) ? 256 : 64);
        APSPIE_CHECK(!batch || (k >= 1 && k <= 256 && (!(kinds & APSPIE_TUNE_SSSP_BATCH) || k <= 64)),
                     APSPIE_ELIMIT, "batched tuning takes 1..256 sources (1..64 with SSSP)");
        APSPIE_CHECK(!(kinds & (APSPIE_TUNE_SSSP | APSPIE_TUNE_SSSP_BATCH)) || g->val, APSPIE_EINVAL,
                     "SSSP tuning needs weights");
        auto alloc = [&](void **p, size_t bytes) {
            if (hipMalloc(p, std::max<size_t>(bytes, 4)) != hipSuccess) {
                (void)hipGetLastError();
                throw Error(APSPIE_ENOMEM, "autotune buffer allocation failed");
            }
        };
        alloc(&d_out, size_t(g->m) * 4);
        if (batch) {
            // k sources with out-edges, spread over the id range by a fixed
            // multiplicative stride
            std::vector<int32_t> rp(size_t(g->m) + 1), srcs;
            APSPIE_HIP(hipMemcpyAsync(rp.data(), g->row_ptr, rp.size() * 4, hipMemcpyDeviceToHost, ctx->stream));
            APSPIE_HIP(hipStreamSynchronize(ctx->stream));
            const uint64_t m = uint64_t(g->m);
            for (uint64_t i = 0; i < 4 * m && int(srcs.size()) < kb; ++i) {
                const int32_t v = int32_t((i * 2654435761ull + uint64_t(src)) % m);
                if (rp[v + 1] > rp[v]) srcs.push_back(v);
            }
            while (int(srcs.size()) < kb) srcs.push_back(src);
            alloc(&d_srcs, size_t(kb) * 4);
            alloc(&d_batch, size_t(kb) * g->m * 4);
            APSPIE_HIP(hipMemcpyAsync(d_srcs, srcs.data(), size_t(kb) * 4, hipMemcpyHostToDevice, ctx->stream));
            APSPIE_HIP(hipStreamSynchronize(ctx->stream));
        }
        const bool csc = g->csc_col_ptr != nullptr;
        auto run_kinds = [&](uint32_t which) {
            double t = 0;
            if (which & kinds & APSPIE_TUNE_BFS)
                t += median3(ctx, [&] { return apspie_bfs(ctx, g, src, (int32_t *)d_out, nullptr); });
            if (which & kinds & APSPIE_TUNE_SSSP)
                t += median3(ctx, [&] { return apspie_sssp(ctx, g, src, (float *)d_out, nullptr); });
            if (which & kinds & APSPIE_TUNE_BFS_BATCH)
                t += median3(ctx, [&] {
                    return apspie_bfs_batch(ctx, g, (const int32_t *)d_srcs, kb, (int32_t *)d_batch, nullptr);
                });
            if (which & kinds & APSPIE_TUNE_SSSP_BATCH)
                t += median3(ctx, [&] {
                    return apspie_sssp_batch(ctx, g, (const int32_t *)d_srcs, std::min(kb, 64), (float *)d_batch,
                                             nullptr);
                });
            return t;
        };
        // one knob: try each value, keep the fastest
        auto tune = [&](int32_t apspie_tuning::*knob, std::vector<int32_t> values, uint32_t which) {
            if (!(which & kinds)) return;
            double best_t = 1e300;
            int32_t best_v = best.*knob;
            for (int32_t v : values) {
                apspie_tuning t = best;
                t.*knob = v;
                if (apspie_ctx_set_tuning(ctx, &t) != APSPIE_OK) continue;
                const double ms = run_kinds(which);
                if (ms < best_t) {
                    best_t = ms;
                    best_v = v;
                }
            }
            best.*knob = best_v;
            APSPIE_CHECK(apspie_ctx_set_tuning(ctx, &best) == APSPIE_OK, APSPIE_EINVAL, "autotune: bad tuning");
        };
        const uint32_t plain = APSPIE_TUNE_BFS | APSPIE_TUNE_SSSP;
        const uint32_t batched = APSPIE_TUNE_BFS_BATCH | APSPIE_TUNE_SSSP_BATCH;
        if (csc) {
            tune(&apspie_tuning::bfs_alpha, {0, 6, 14, 30}, APSPIE_TUNE_BFS);
            if (best.bfs_alpha > 0) tune(&apspie_tuning::bfs_beta, {12, 24, 48}, APSPIE_TUNE_BFS);
            if (g->csc_val) tune(&apspie_tuning::sssp_pull_ratio, {0, 2}, APSPIE_TUNE_SSSP);
        }
        if (g->m <= (1 << 23)) tune(&apspie_tuning::sssp_bin_ratio, {0, 2, 4}, APSPIE_TUNE_SSSP);
        tune(&apspie_tuning::sssp_delta, {0, 16, 32, 64, 128}, APSPIE_TUNE_SSSP);
        tune(&apspie_tuning::sssp_hint, {0, 1}, APSPIE_TUNE_SSSP);
        tune(&apspie_tuning::expand_wg_per_cu, {4, 8, 16, 32}, plain | batched);
        tune(&apspie_tuning::round_chunk, {3, 4, 5, 6, 8, 12}, plain);
        tune(&apspie_tuning::batch_chunk, {2, 4, 8}, batched);
        if (kb <= 16) tune(&apspie_tuning::sssp_batch_lanes, {16, 64}, APSPIE_TUNE_SSSP_BATCH);
        *out = best;
    });
    if (d_out) (void)hipFree(d_out);
    if (d_batch) (void)hipFree(d_batch);
    if (d_srcs) (void)hipFree(d_srcs);
    if (rc != APSPIE_OK) {  // leave the context as it was
        const std::string err = ctx->err;
        apspie_ctx_set_tuning(ctx, &orig);
        ctx->err = err;
    }
    return rc;
}
