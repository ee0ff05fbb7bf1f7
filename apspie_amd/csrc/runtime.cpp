// runtime.cpp -- context lifetime, scratch pool and memory helpers.
#include <algorithm>
#include <cstring>
#include <new>

#include "internal.hpp"
#include "traverse.hpp"

using namespace apspie;

void apspie_ctx::drop_graphs() {
    for (auto &kv : graphs) (void)hipGraphExecDestroy(kv.second);
    graphs.clear();
}

void *apspie_ctx::slot(int s, size_t bytes) {
    if (bytes == 0) bytes = 256;
    if (slot_bytes[s] >= bytes) return slot_ptr[s];
    APSPIE_HIP(hipStreamSynchronize(stream));
    drop_graphs();  // captured kernels may hold the old address
    if (slot_ptr[s]) {
        APSPIE_HIP(hipFree(slot_ptr[s]));
        slot_ptr[s] = nullptr;
        slot_bytes[s] = 0;
    }
    // round up so a growing graph does not reallocate every call
    size_t want = (bytes + (1u << 20) - 1) & ~size_t((1u << 20) - 1);
    void *p = nullptr;
    if (hipMalloc(&p, want) != hipSuccess) {
        (void)hipGetLastError();
        throw Error(APSPIE_ENOMEM, "scratch allocation of " + std::to_string(want) + " bytes failed");
    }
    slot_ptr[s] = p;
    slot_bytes[s] = want;
    return p;
}

void *apspie_ctx::hslot(int s, size_t bytes) {
    if (hslot_bytes[s] >= bytes && hslot_ptr[s]) return hslot_ptr[s];
    APSPIE_HIP(hipStreamSynchronize(stream));  // copies in flight may use the old buffer
    if (hslot_ptr[s]) {
        APSPIE_HIP(hipHostFree(hslot_ptr[s]));
        hslot_ptr[s] = hslot_dev[s] = nullptr;
        hslot_bytes[s] = 0;
    }
    const size_t want = (std::max<size_t>(bytes, 1) + (1u << 16) - 1) & ~size_t((1u << 16) - 1);
    void *p = nullptr;
    if (hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        throw Error(APSPIE_ENOMEM, "pinned host allocation of " + std::to_string(want) + " bytes failed");
    }
    void *d = nullptr;
    if (hipHostGetDevicePointer(&d, p, 0) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipHostFree(p);
        throw Error(APSPIE_EHIP, "no device address for pinned host memory");
    }
    hslot_ptr[s] = p;
    hslot_dev[s] = d;
    hslot_bytes[s] = want;
    return p;
}

void apspie_ctx::prof_begin(int cls) {
    if (!profiling) return;
    auto take = [&] {
        hipEvent_t e;
        if (!prof_pool.empty()) {
            e = prof_pool.back();
            prof_pool.pop_back();
        } else {
            APSPIE_HIP(hipEventCreate(&e));
        }
        return e;
    };
    hipEvent_t a = take(), b = take();
    APSPIE_HIP(hipEventRecord(a, stream));
    prof_open.push_back({cls, {a, b}});
}

void apspie_ctx::prof_end() {
    if (!profiling || prof_open.empty()) return;
    APSPIE_HIP(hipEventRecord(prof_open.back().second.second, stream));
}

void apspie_ctx::prof_collect() {
    for (auto &o : prof_open) {
        float ms = 0.f;
        APSPIE_HIP(hipEventSynchronize(o.second.second));
        APSPIE_HIP(hipEventElapsedTime(&ms, o.second.first, o.second.second));
        prof_ms[o.first] += ms;
        prof_launches[o.first] += 1;
        prof_pool.push_back(o.second.first);
        prof_pool.push_back(o.second.second);
    }
    prof_open.clear();
}

void apspie::validate_csr(const apspie_csr *g, bool need_val) {
    APSPIE_CHECK(g != nullptr, APSPIE_EINVAL, "null csr");
    APSPIE_CHECK(g->m >= 0 && g->n >= 0 && g->nnz >= 0, APSPIE_EINVAL, "negative csr size");
    APSPIE_CHECK(g->nnz < (int64_t(1) << 31) - 4096, APSPIE_ELIMIT, "nnz must stay below 2^31");
    APSPIE_CHECK(g->m == 0 || g->row_ptr != nullptr, APSPIE_EINVAL, "null row_ptr");
    APSPIE_CHECK(g->nnz == 0 || g->col_ind != nullptr, APSPIE_EINVAL, "null col_ind");
    APSPIE_CHECK((reinterpret_cast<uintptr_t>(g->row_ptr) & 15) == 0, APSPIE_EINVAL,
                 "row_ptr must be 16-byte aligned");
    APSPIE_CHECK(!need_val || g->nnz == 0 || g->val != nullptr, APSPIE_EINVAL, "weights required");
    if (g->csc_col_ptr) {
        APSPIE_CHECK(g->nnz == 0 || g->csc_row_ind != nullptr, APSPIE_EINVAL, "null csc_row_ind");
        APSPIE_CHECK((reinterpret_cast<uintptr_t>(g->csc_col_ptr) & 15) == 0, APSPIE_EINVAL,
                     "csc_col_ptr must be 16-byte aligned");
    }
}

extern "C" {

int apspie_ctx_create(int device, apspie_ctx **out) {
    if (!out) return APSPIE_EINVAL;
    *out = nullptr;
    apspie_ctx *ctx = new (std::nothrow) apspie_ctx();
    if (!ctx) return APSPIE_ENOMEM;
    ctx->device = device;
    int rc = guarded(ctx, [&] {
        int n = 0;
        APSPIE_HIP(hipGetDeviceCount(&n));
        APSPIE_CHECK(device >= 0 && device < n, APSPIE_EINVAL, "no such device");
        APSPIE_HIP(hipSetDevice(device));
        APSPIE_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        APSPIE_HIP(hipEventCreate(&ctx->ev0));
        APSPIE_HIP(hipEventCreate(&ctx->ev1));
        APSPIE_HIP(hipDeviceGetAttribute(&ctx->cu_count, hipDeviceAttributeMultiprocessorCount, device));
        APSPIE_HIP(hipHostMalloc(&ctx->pinned, PINNED_BYTES, hipHostMallocDefault));
        trav::noalloc_setup(device);
    });
    if (rc != APSPIE_OK) {
        fprintf(stderr, "apspie_ctx_create: %s\n", ctx->err.c_str());
        apspie_ctx_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return APSPIE_OK;
}

int apspie_ctx_destroy(apspie_ctx *ctx) {
    if (!ctx) return APSPIE_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    ctx->drop_graphs();
    for (int s = 0; s < S_NSLOTS; ++s)
        if (ctx->slot_ptr[s]) (void)hipFree(ctx->slot_ptr[s]);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    for (auto p : ctx->hslot_ptr)
        if (p) (void)hipHostFree(p);
    for (auto &o : ctx->prof_open) {
        (void)hipEventDestroy(o.second.first);
        (void)hipEventDestroy(o.second.second);
    }
    for (auto e : ctx->prof_pool) (void)hipEventDestroy(e);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return APSPIE_OK;
}

const char *apspie_last_error(const apspie_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

void *apspie_ctx_stream(apspie_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int apspie_ctx_synchronize(apspie_ctx *ctx) {
    return guarded(ctx, [&] { APSPIE_HIP(hipStreamSynchronize(ctx->stream)); });
}

int apspie_ctx_cu_count(const apspie_ctx *ctx) { return ctx ? ctx->cu_count : 0; }

int apspie_ctx_set_tuning(apspie_ctx *ctx, const apspie_tuning *t) {
    if (!ctx || !t) return APSPIE_EINVAL;
    if (t->expand_wg_per_cu < 1 || t->expand_wg_per_cu > 32 || t->round_chunk < 1 || t->round_chunk > 32 ||
        t->batch_chunk < 1 || t->batch_chunk > 16 ||
        (t->sssp_batch_lanes != 0 && t->sssp_batch_lanes != 16 && t->sssp_batch_lanes != 64) ||
        t->sssp_delta < 0 || t->sssp_delta > (1 << 24) || (t->spgemm_long_rows != 0 && t->spgemm_long_rows != 1) ||
        (t->spgemm_passes != 1 && t->spgemm_passes != 2) || t->bfs_group_cap < 0 || t->bfs_group_cap > (1 << 30) ||
        (t->spgemm_esc_min != 256 && t->spgemm_esc_min != 512 && t->spgemm_esc_min != 1024 &&
         t->spgemm_esc_min != 2048) ||
        (t->sssp_hint != 0 && t->sssp_hint != 1)) {
        ctx->err = "tuning knob out of range";
        return APSPIE_EINVAL;
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->tune.bfs_alpha = t->bfs_alpha;
    ctx->tune.bfs_beta = t->bfs_beta;
    ctx->tune.sssp_pull_ratio = t->sssp_pull_ratio;
    ctx->tune.sssp_bin_ratio = t->sssp_bin_ratio;
    ctx->tune.expand_wg_per_cu = t->expand_wg_per_cu;
    ctx->tune.round_chunk = t->round_chunk;
    ctx->tune.batch_chunk = t->batch_chunk;
    ctx->tune.sssp_batch_lanes = t->sssp_batch_lanes;
    ctx->tune.sssp_delta = t->sssp_delta;
    ctx->tune.spgemm_long_rows = t->spgemm_long_rows;
    ctx->tune.spgemm_passes = t->spgemm_passes;
    ctx->tune.bfs_group_cap = t->bfs_group_cap;
    ctx->tune.spgemm_esc_min = t->spgemm_esc_min;
    ctx->tune.sssp_hint = t->sssp_hint;
    return APSPIE_OK;
}

int apspie_ctx_get_tuning(const apspie_ctx *ctx, apspie_tuning *t) {
    if (!ctx || !t) return APSPIE_EINVAL;
    memset(t, 0, sizeof(*t));
    t->bfs_alpha = ctx->tune.bfs_alpha;
    t->bfs_beta = ctx->tune.bfs_beta;
    t->sssp_pull_ratio = ctx->tune.sssp_pull_ratio;
    t->sssp_bin_ratio = ctx->tune.sssp_bin_ratio;
    t->expand_wg_per_cu = ctx->tune.expand_wg_per_cu;
    t->round_chunk = ctx->tune.round_chunk;
    t->batch_chunk = ctx->tune.batch_chunk;
    t->sssp_batch_lanes = ctx->tune.sssp_batch_lanes;
    t->sssp_delta = ctx->tune.sssp_delta;
    t->spgemm_long_rows = ctx->tune.spgemm_long_rows;
    t->spgemm_passes = ctx->tune.spgemm_passes;
    t->bfs_group_cap = ctx->tune.bfs_group_cap;
    t->spgemm_esc_min = ctx->tune.spgemm_esc_min;
    t->sssp_hint = ctx->tune.sssp_hint;
    return APSPIE_OK;
}

int apspie_ctx_set_profiling(apspie_ctx *ctx, int enable) {
    if (!ctx) return APSPIE_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->profiling = enable != 0;
    return APSPIE_OK;
}

int apspie_ctx_profile_read(apspie_ctx *ctx, int kclass, double *ms, int64_t *launches) {
    if (!ctx || kclass < 0 || kclass >= APSPIE_KCLASS_COUNT) return APSPIE_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    return guarded(ctx, [&] {
        ctx->prof_collect();
        if (ms) *ms = ctx->prof_ms[kclass];
        if (launches) *launches = ctx->prof_launches[kclass];
    });
}

int apspie_ctx_profile_reset(apspie_ctx *ctx) {
    if (!ctx) return APSPIE_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->mu);
    return guarded(ctx, [&] {
        ctx->prof_collect();
        for (int i = 0; i < APSPIE_KCLASS_COUNT; ++i) {
            ctx->prof_ms[i] = 0;
            ctx->prof_launches[i] = 0;
        }
    });
}

int apspie_malloc(apspie_ctx *ctx, void **ptr, size_t bytes) {
    if (!ctx || !ptr) return APSPIE_EINVAL;
    return guarded(ctx, [&] {
        *ptr = nullptr;
        if (hipMalloc(ptr, bytes ? bytes : 1) != hipSuccess) {
            (void)hipGetLastError();
            throw Error(APSPIE_ENOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
        }
    });
}

int apspie_free(apspie_ctx *ctx, void *ptr) {
    if (!ctx) return APSPIE_EINVAL;
    return guarded(ctx, [&] {
        if (ptr) {
            APSPIE_HIP(hipStreamSynchronize(ctx->stream));
            APSPIE_HIP(hipFree(ptr));
        }
    });
}

int apspie_memcpy_h2d(apspie_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return APSPIE_EINVAL;
    return guarded(ctx, [&] {
        if (!bytes) return;
        APSPIE_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        APSPIE_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int apspie_memcpy_d2h(apspie_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return APSPIE_EINVAL;
    return guarded(ctx, [&] {
        if (!bytes) return;
        APSPIE_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        APSPIE_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int apspie_csr_upload(apspie_ctx *ctx, const apspie_host_csr *h, apspie_csr_owned *out) {
    if (!ctx || !h || !out) return APSPIE_EINVAL;
    memset(out, 0, sizeof(*out));
    return guarded(ctx, [&] {
        APSPIE_CHECK(h->m >= 0 && h->nnz >= 0, APSPIE_EINVAL, "negative size");
        APSPIE_CHECK(h->nnz < (int64_t(1) << 31) - 4096, APSPIE_ELIMIT, "nnz must stay below 2^31");
        out->m = h->m;
        out->n = h->n;
        out->nnz = h->nnz;
        auto alloc = [&](void **p, size_t b) {
            if (hipMalloc(p, b ? b : 4) != hipSuccess) {
                (void)hipGetLastError();
                throw Error(APSPIE_ENOMEM, "csr upload allocation failed");
            }
        };
        try {
            alloc((void **)&out->row_ptr, (size_t(h->m) + 1) * 4);
            alloc((void **)&out->col_ind, size_t(h->nnz) * 4);
            if (h->val) alloc((void **)&out->val, size_t(h->nnz) * 4);
            APSPIE_HIP(hipMemcpyAsync(out->row_ptr, h->row_ptr, (size_t(h->m) + 1) * 4,
                                      hipMemcpyHostToDevice, ctx->stream));
            if (h->nnz)
                APSPIE_HIP(hipMemcpyAsync(out->col_ind, h->col_ind, size_t(h->nnz) * 4,
                                          hipMemcpyHostToDevice, ctx->stream));
            if (h->val && h->nnz)
                APSPIE_HIP(hipMemcpyAsync(out->val, h->val, size_t(h->nnz) * 4, hipMemcpyHostToDevice,
                                          ctx->stream));
            APSPIE_HIP(hipStreamSynchronize(ctx->stream));
        } catch (...) {
            apspie_csr_free(ctx, out);
            throw;
        }
    });
}

int apspie_csr_free(apspie_ctx *ctx, apspie_csr_owned *c) {
    if (!ctx || !c) return APSPIE_EINVAL;
    return guarded(ctx, [&] {
        APSPIE_HIP(hipStreamSynchronize(ctx->stream));
        // CSC arrays that alias the CSR ones (symmetric matrix) are freed once
        if (c->csc_col_ptr && c->csc_col_ptr != c->row_ptr) APSPIE_HIP(hipFree(c->csc_col_ptr));
        if (c->csc_row_ind && c->csc_row_ind != c->col_ind) APSPIE_HIP(hipFree(c->csc_row_ind));
        if (c->csc_val && c->csc_val != c->val) APSPIE_HIP(hipFree(c->csc_val));
        if (c->row_ptr) APSPIE_HIP(hipFree(c->row_ptr));
        if (c->col_ind) APSPIE_HIP(hipFree(c->col_ind));
        if (c->val) APSPIE_HIP(hipFree(c->val));
        memset(c, 0, sizeof(*c));
    });
}

apspie_csr apspie_csr_view(const apspie_csr_owned *c) {
    apspie_csr v{};
    if (!c) return v;
    v.m = c->m;
    v.n = c->n;
    v.nnz = c->nnz;
    v.row_ptr = c->row_ptr;
    v.col_ind = c->col_ind;
    v.val = c->val;
    v.csc_col_ptr = c->csc_col_ptr;
    v.csc_row_ind = c->csc_row_ind;
    v.csc_val = c->csc_val;
    return v;
}

int apspie_csr_attach_csc(apspie_ctx *ctx, apspie_csr_owned *c, const apspie_host_csr *csc) {
    if (!ctx || !c) return APSPIE_EINVAL;
    return guarded(ctx, [&] {
        APSPIE_CHECK(c->csc_col_ptr == nullptr, APSPIE_EINVAL, "the matrix already has a CSC");
        if (!csc) {
            APSPIE_CHECK(c->m == c->n, APSPIE_EINVAL, "only a square matrix can be its own transpose");
            c->csc_col_ptr = c->row_ptr;
            c->csc_row_ind = c->col_ind;
            c->csc_val = c->val;
            return;
        }
        APSPIE_CHECK(csc->m == c->n && csc->n == c->m && csc->nnz == c->nnz, APSPIE_EINVAL,
                     "the CSC must be the transpose of the matrix (n x m, same nnz)");
        APSPIE_CHECK((csc->val != nullptr) == (c->val != nullptr), APSPIE_EINVAL,
                     "the CSC must carry values exactly when the CSR does");
        apspie_csr_owned t{};
        const int rc = apspie_csr_upload(ctx, csc, &t);
        if (rc != APSPIE_OK) throw Error(rc, ctx->err);
        c->csc_col_ptr = t.row_ptr;
        c->csc_row_ind = t.col_ind;
        c->csc_val = t.val;
    });
}

}  // extern "C"
