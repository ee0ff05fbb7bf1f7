// serve.cpp -- serving loop over a device-resident graph pool.
//
// Requests (BFS or SSSP from one source on one pooled graph) enter a FIFO.
// A batch of at most plain_bfs_upto / plain_sssp_upto requests runs as plain
// single-source launches instead (one source: the batched kernels' lanes
// would idle).
// `lanes` worker threads each own a context (HIP stream + scratch) and loop:
// take the oldest pending request, gather every other pending request for the
// same (graph, kind) up to the batch limit while the device-memory budget
// admits them, run one batched traversal (apspie_bfs_batch /
// apspie_sssp_batch), copy each requested row back and complete the tickets.
// Requests that arrive while batches run are picked up by the next lane that
// frees, and batches of different lanes overlap on the device (one stream
// each), so the device is kept busy without waiting for a full batch.
//
// A source without out-edges reaches only itself: its row (depth 0 / distance
// 0 at the source, -1 / APSPIE_SSSP_INF elsewhere) is known without a
// traversal.  Such requests take no lane of a batch and no device memory; the
// worker that meets them in the queue answers them on the host.  A host byte
// per vertex (the bit length of its out-degree, 0 without out-edges) decides
// this at submission and orders the SSSP batches (below).
//
// Latency = completion time - arrival time on one steady clock.  A request
// whose footprint exceeds the whole budget is refused at submission.
#include <algorithm>
#include <cstdlib>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <vector>

#include "internal.hpp"

using namespace apspie;

namespace {

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Request {
    uint64_t id;
    int kind, graph;
    int32_t src;
    void *out;
    double arrive, done_at = 0;
    size_t footprint;
    bool done = false, trivial = false, picked = false;
    int status = 0;
    uint8_t dclass = 0;  // bit length of the source's out-degree
};

constexpr size_t TRIVIAL_PER_PASS = 4096;  // host-answered requests per worker pass

}  // namespace

struct apspie_server {
    apspie_server_opts opts{};
    std::vector<apspie_csr> graphs;
    std::vector<std::vector<uint8_t>> dclass;  // per graph and vertex: bit length of the out-degree (0: none)
    int device = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> pending;
    std::map<uint64_t, std::unique_ptr<Request>> all;
    uint64_t next_id = 1;
    size_t used = 0;
    uint64_t released = 0;  // bumped whenever a batch returns its memory to the budget
    bool stop = false;
    int inflight = 0;  // worker passes that took requests and have not completed them all
    std::vector<apspie_ctx *> ctxs;  // one per lane, created with the server
    std::vector<void *> d_srcs;      // per lane: the batch's source ids (256 int32)
    std::vector<std::thread> workers;
    std::vector<double> lat_ms;
    double first_arrive = 0, last_done = 0;
    int64_t submitted = 0, completed = 0, rejected = 0, batches = 0, batched_reqs = 0, trivial = 0, joined = 0;
    std::string err;

    size_t footprint(int kind, int g) const {
        const size_t m = size_t(graphs[g].m);
        // output row + share of the batch state (vertex-major masks or
        // distances) + the host-visible copy staging
        return kind == 0 ? m * 5 + 4096 : m * 13 + 4096;
    }
    void worker(int lane);
    void answer_trivial(std::vector<Request *> &rs);
    int run_joined(apspie_ctx *ctx, int kind, int gi, int K, std::vector<Request *> &batch, int32_t *d_src,
                   void *d_out);
};

// A batch with continuous batching: after each chunk of rounds, while the
// union frontier is not empty, requests of the same kind and graph that wait
// in the queue (oldest first, as the budget admits them) take the batch's
// free lanes; their sources start at the next round (SSSP distances are the
// same whatever round a source starts in; BFS depths count from the round
// that picks the source), and their rows come out with the batch's.  The
// lane's context is held for the whole batch (new knobs apply from its next
// batch).
int apspie_server::run_joined(apspie_ctx *ctx, int kind, int gi, int K, std::vector<Request *> &batch,
                              int32_t *d_src, void *d_out) {
    std::lock_guard<std::mutex> clk(ctx->mu);
    trav::BatchJoin *j = nullptr;
    std::vector<Request *> more;
    std::vector<int32_t> add;
    const int rc = guarded(ctx, [&] {
        j = trav::batch_join_begin(ctx, &graphs[size_t(gi)], kind, d_src, int(batch.size()), K, d_out);
        const int limit = std::min(K, kind == APSPIE_REQ_BFS ? opts.max_bfs_batch : opts.max_sssp_batch);
        while (trav::batch_join_step(j)) {
            const int lanes = trav::batch_join_lanes(j);
            if (lanes >= limit) continue;
            more.clear();
            {
                std::lock_guard<std::mutex> lk(mu);
                for (Request *r : pending) {
                    if ((int)more.size() == limit - lanes) break;
                    if (r->trivial || r->kind != kind || r->graph != gi) continue;
                    if (used + r->footprint > opts.budget_bytes) break;
                    used += r->footprint;
                    r->picked = true;
                    more.push_back(r);
                }
                if (!more.empty()) {
                    pending.erase(std::remove_if(pending.begin(), pending.end(),
                                                 [](const Request *r) { return r->picked; }),
                                  pending.end());
                    joined += int64_t(more.size());
                }
            }
            if (more.empty()) continue;
            batch.insert(batch.end(), more.begin(), more.end());
            add.resize(more.size());
            for (size_t i = 0; i < more.size(); ++i) add[i] = more[i]->src;
            APSPIE_HIP(hipMemcpyAsync(d_src + lanes, add.data(), add.size() * 4, hipMemcpyHostToDevice, ctx->stream));
            trav::batch_join_add(j, d_src + lanes, int(more.size()));
        }
        apspie_stats st;
        trav::batch_join_end(j, d_out, &st);
        APSPIE_HIP(hipStreamSynchronize(ctx->stream));
    });
    if (j) trav::batch_join_free(j);
    return rc;
}

void apspie_server::answer_trivial(std::vector<Request *> &rs) {
    for (Request *r : rs) {
        if (!r->out) continue;
        const size_t m = size_t(graphs[r->graph].m);
        if (r->kind == APSPIE_REQ_BFS) {
            int32_t *o = (int32_t *)r->out;
            std::fill(o, o + m, -1);
            o[r->src] = 0;
        } else {
            float *o = (float *)r->out;
            std::fill(o, o + m, APSPIE_SSSP_INF);
            o[r->src] = 0.0f;
        }
    }
    const double t = now_s();
    {
        std::lock_guard<std::mutex> lk(mu);
        for (Request *r : rs) {
            r->done = true;
            r->done_at = t;
            r->status = APSPIE_OK;
            completed++;
            trivial++;
            lat_ms.push_back((t - r->arrive) * 1e3);
            last_done = std::max(last_done, t);
        }
    }
    cv_done.notify_all();
    rs.clear();
}

void apspie_server::worker(int lane) {
    apspie_ctx *ctx = ctxs[size_t(lane)];
    void *d_src = d_srcs[size_t(lane)], *d_out = nullptr;
    size_t out_cap = 0;
    std::vector<Request *> batch, easy, cand;
    std::vector<int32_t> srcs;
    for (;;) {
        batch.clear();
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] { return stop || !pending.empty(); });
            if (stop && pending.empty()) break;
            // in arrival order: take host-answered requests of any graph, and
            // let the oldest traversal request decide (graph, kind) and gather
            // more of that kind while budget and batch limit allow.  When more
            // SSSP requests wait than a batch takes, the oldest one is joined
            // by those whose sources have the closest out-degree class (ties
            // in arrival order): sources alike in degree reach the graph's
            // hubs in the same rounds, so the batch's union frontier repeats
            // less (measured on RMAT s22: 7 % less device time per source).
            Request *head = nullptr;
            int limit = 0;
            size_t cap = 0;
            cand.clear();
            for (auto it = pending.begin(); it != pending.end();) {
                Request *r = *it;
                if (r->trivial) {
                    if (easy.size() < TRIVIAL_PER_PASS) {
                        easy.push_back(r);
                        it = pending.erase(it);
                        continue;
                    }
                } else {
                    if (!head) {
                        head = r;
                        limit = r->kind == APSPIE_REQ_BFS ? opts.max_bfs_batch : opts.max_sssp_batch;
                        cap = size_t(limit) * (r->kind == APSPIE_REQ_SSSP ? 8 : 1);
                    }
                    if (cand.size() < cap && r->kind == head->kind && r->graph == head->graph) cand.push_back(r);
                }
                if (head && cand.size() == cap && easy.size() == TRIVIAL_PER_PASS) break;
                ++it;
            }
            if (head && head->kind == APSPIE_REQ_SSSP && (int)cand.size() > limit) {
                const int hc = head->dclass;
                std::stable_sort(cand.begin() + 1, cand.end(), [hc](const Request *x, const Request *y) {
                    return std::abs(int(x->dclass) - hc) < std::abs(int(y->dclass) - hc);
                });
            }
            for (Request *r : cand) {  // cand[0] is the oldest
                if ((int)batch.size() == limit || used + r->footprint > opts.budget_bytes) break;
                used += r->footprint;
                r->picked = true;
                batch.push_back(r);
            }
            if (!batch.empty())
                pending.erase(std::remove_if(pending.begin(), pending.end(), [](const Request *r) { return r->picked; }),
                              pending.end());
            if (batch.empty() && easy.empty()) {
                // budget full: wait until some lane returns memory (or none
                // holds any, when the oldest request fits by itself)
                const uint64_t seen = released;
                cv_work.wait(lk, [&] { return stop || released != seen || used == 0; });
                continue;
            }
            // counted until every request taken here is complete, so drain
            // cannot return while their rows are still being written
            inflight++;
        }
        if (!easy.empty()) answer_trivial(easy);
        if (batch.empty()) {
            {
                std::lock_guard<std::mutex> lk(mu);
                inflight--;
            }
            cv_done.notify_all();
            cv_work.notify_all();
            continue;
        }
        const int kind = batch[0]->kind, gi = batch[0]->graph;
        const apspie_csr &g = graphs[gi];
        const int k = (int)batch.size();
        srcs.resize(k);
        for (int i = 0; i < k; ++i) srcs[i] = batch[i]->src;
        int rc = apspie_memcpy_h2d(ctx, d_src, srcs.data(), size_t(k) * 4);
        // a batch keeps rows for every lane that sources may join: SSSP 16
        // or 64 lanes (as apspie_sssp_batch picks), BFS 64, 128 or 256
        const int plain_upto = kind == APSPIE_REQ_BFS ? opts.plain_bfs_upto : opts.plain_sssp_upto;
        const bool joinable = k > plain_upto;
        int K = k;
        if (joinable && kind == APSPIE_REQ_SSSP) K = k <= 16 && ctx->tune.sssp_batch_lanes != 64 ? 16 : 64;
        if (joinable && kind == APSPIE_REQ_BFS) K = k <= 64 ? 64 : k <= 128 ? 128 : 256;
        const size_t need = size_t(K) * size_t(g.m) * 4;
        if (rc == APSPIE_OK && need > out_cap) {
            if (d_out) apspie_free(ctx, d_out);
            d_out = nullptr;
            out_cap = 0;
            rc = apspie_malloc(ctx, &d_out, need);
            if (rc == APSPIE_OK) out_cap = need;
        }
        // a batch of very few sources runs faster as plain launches
        if (rc == APSPIE_OK && k <= plain_upto) {
            for (int i = 0; i < k && rc == APSPIE_OK; ++i) {
                void *row = (char *)d_out + size_t(i) * g.m * 4;
                rc = kind == APSPIE_REQ_BFS ? apspie_bfs(ctx, &g, srcs[i], (int32_t *)row, nullptr)
                                            : apspie_sssp(ctx, &g, srcs[i], (float *)row, nullptr);
            }
        } else if (rc == APSPIE_OK) {
            rc = run_joined(ctx, kind, gi, K, batch, (int32_t *)d_src, d_out);
        }
        for (size_t i = 0; i < batch.size() && rc == APSPIE_OK; ++i)  // joined requests too
            if (batch[i]->out)
                rc = apspie_memcpy_d2h(ctx, batch[i]->out, (const char *)d_out + size_t(i) * g.m * 4,
                                       size_t(g.m) * 4);
        const double t = now_s();
        {
            std::lock_guard<std::mutex> lk(mu);
            for (Request *r : batch) {
                r->done = true;
                r->done_at = t;
                r->status = rc;
                used -= r->footprint;
                if (rc == APSPIE_OK) {
                    completed++;
                    lat_ms.push_back((t - r->arrive) * 1e3);
                    last_done = std::max(last_done, t);
                }
            }
            if (rc != APSPIE_OK) err = apspie_last_error(ctx);
            batches++;
            batched_reqs += int64_t(batch.size());
            released++;
            inflight--;
        }
        cv_done.notify_all();
        cv_work.notify_all();
    }
    if (d_out) apspie_free(ctx, d_out);
}

extern "C" {

double apspie_now(void) { return now_s(); }

int apspie_server_create(int device, const apspie_csr *graphs, int ngraphs, const apspie_server_opts *opts,
                         apspie_server **out) {
    if (!out || !graphs || ngraphs < 1) return APSPIE_EINVAL;
    *out = nullptr;
    auto *s = new (std::nothrow) apspie_server();
    if (!s) return APSPIE_ENOMEM;
    s->device = device;
    apspie_server_opts o{};
    o.lanes = 4;
    o.max_bfs_batch = 256;
    o.max_sssp_batch = 64;
    o.plain_bfs_upto = 3;
    o.plain_sssp_upto = 1;
    o.budget_bytes = size_t(32) << 30;
    if (opts) {
        if (opts->lanes > 0) o.lanes = opts->lanes;
        if (opts->max_bfs_batch > 0) o.max_bfs_batch = std::min(opts->max_bfs_batch, 256);
        if (opts->max_sssp_batch > 0) o.max_sssp_batch = std::min(opts->max_sssp_batch, 64);
        if (opts->plain_bfs_upto != 0) o.plain_bfs_upto = opts->plain_bfs_upto;
        if (opts->plain_sssp_upto != 0) o.plain_sssp_upto = opts->plain_sssp_upto;
        if (opts->budget_bytes > 0) o.budget_bytes = opts->budget_bytes;
    }
    s->opts = o;
    for (int i = 0; i < ngraphs; ++i) {
        if (graphs[i].m != graphs[i].n || graphs[i].m <= 0) {
            delete s;
            return APSPIE_EINVAL;
        }
        s->graphs.push_back(graphs[i]);
    }
    // which vertices have out-edges, from each graph's row offsets
    if (hipSetDevice(device) != hipSuccess) {
        delete s;
        return APSPIE_EHIP;
    }
    std::vector<int32_t> rp;
    for (const apspie_csr &g : s->graphs) {
        rp.resize(size_t(g.m) + 1);
        if (hipMemcpy(rp.data(), g.row_ptr, rp.size() * 4, hipMemcpyDeviceToHost) != hipSuccess) {
            delete s;
            return APSPIE_EHIP;
        }
        std::vector<uint8_t> dc(size_t(g.m), 0);
        for (int v = 0; v < g.m; ++v) {
            const uint32_t d = uint32_t(rp[v + 1] - rp[v]);
            dc[size_t(v)] = uint8_t(d ? 32 - __builtin_clz(d) : 0);
        }
        s->dclass.push_back(std::move(dc));
    }
    // every lane's context and source buffer exist before any request is
    // taken, so a server that starts can serve
    for (int l = 0; l < o.lanes; ++l) {
        apspie_ctx *c = nullptr;
        void *d = nullptr;
        int rc = apspie_ctx_create(device, &c);
        if (rc == APSPIE_OK) {
            rc = apspie_malloc(c, &d, 256 * 4);
            if (rc != APSPIE_OK) apspie_ctx_destroy(c);
        }
        if (rc != APSPIE_OK) {
            apspie_server_destroy(s);
            return rc;
        }
        s->ctxs.push_back(c);
        s->d_srcs.push_back(d);
    }
    for (int l = 0; l < o.lanes; ++l) s->workers.emplace_back([s, l] { s->worker(l); });
    *out = s;
    return APSPIE_OK;
}

int apspie_server_submit(apspie_server *s, int kind, int graph, int32_t src, void *host_out, double arrival,
                         uint64_t *ticket) {
    if (!s || !ticket || (kind != 0 && kind != 1) || graph < 0 || graph >= (int)s->graphs.size()) return APSPIE_EINVAL;
    if (src < 0 || src >= s->graphs[graph].m) return APSPIE_EINVAL;
    auto r = std::make_unique<Request>();
    r->kind = kind;
    r->graph = graph;
    r->src = src;
    r->out = host_out;
    r->arrive = arrival >= 0 ? arrival : now_s();
    r->dclass = s->dclass[graph][size_t(src)];
    r->trivial = r->dclass == 0;  // no out-edges: answered on the host
    r->footprint = r->trivial ? 0 : s->footprint(kind, graph);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (r->footprint > s->opts.budget_bytes) {
            s->rejected++;
            return APSPIE_EBUSY;
        }
        r->id = s->next_id++;
        *ticket = r->id;
        if (s->submitted == 0 || r->arrive < s->first_arrive) s->first_arrive = r->arrive;
        s->submitted++;
        s->pending.push_back(r.get());
        s->all[r->id] = std::move(r);
    }
    s->cv_work.notify_one();
    return APSPIE_OK;
}

int apspie_server_wait(apspie_server *s, uint64_t ticket, double *latency_ms) {
    if (!s) return APSPIE_EINVAL;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return APSPIE_EINVAL;
    Request *r = it->second.get();
    s->cv_done.wait(lk, [&] { return r->done; });
    if (latency_ms) *latency_ms = (r->done_at - r->arrive) * 1e3;
    const int rc = r->status;
    s->all.erase(it);
    return rc;
}

int apspie_server_drain(apspie_server *s) {
    if (!s) return APSPIE_EINVAL;
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->pending.empty() && s->inflight == 0; });
    for (auto it = s->all.begin(); it != s->all.end();)
        it = it->second->done ? s->all.erase(it) : std::next(it);
    return APSPIE_OK;
}

int apspie_server_stats(apspie_server *s, apspie_server_report *st) {
    if (!s || !st) return APSPIE_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    memset(st, 0, sizeof(*st));
    st->submitted = s->submitted;
    st->completed = s->completed;
    st->rejected = s->rejected;
    st->batches = s->batches;
    st->trivial = s->trivial;
    st->joined = s->joined;
    st->mean_batch = s->batches ? double(s->batched_reqs) / double(s->batches) : 0.0;
    if (!s->lat_ms.empty()) {
        std::vector<double> v = s->lat_ms;
        std::sort(v.begin(), v.end());
        auto q = [&](double p) { return v[std::min(v.size() - 1, size_t(p * double(v.size() - 1) + 0.5))]; };
        st->p50_ms = q(0.50);
        st->p99_ms = q(0.99);
        st->max_ms = v.back();
        double sum = 0;
        for (double x : v) sum += x;
        st->mean_ms = sum / double(v.size());
        const double span = s->last_done - s->first_arrive;
        st->qps = span > 0 ? double(s->completed) / span : 0.0;
    }
    return APSPIE_OK;
}

int apspie_server_reset_stats(apspie_server *s) {
    if (!s) return APSPIE_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    s->lat_ms.clear();
    s->submitted = s->completed = s->rejected = s->batches = s->batched_reqs = s->trivial = s->joined = 0;
    s->first_arrive = s->last_done = 0;
    return APSPIE_OK;
}

int apspie_server_set_tuning(apspie_server *s, const apspie_tuning *t) {
    if (!s || !t) return APSPIE_EINVAL;
    for (apspie_ctx *c : s->ctxs) {  // each waits for its lane's current call
        const int rc = apspie_ctx_set_tuning(c, t);
        if (rc != APSPIE_OK) {
            std::lock_guard<std::mutex> lk(s->mu);
            s->err = apspie_last_error(c);
            return rc;
        }
    }
    return APSPIE_OK;
}

const char *apspie_server_last_error(apspie_server *s) { return s ? s->err.c_str() : "null server"; }

int apspie_server_destroy(apspie_server *s) {
    if (!s) return APSPIE_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->workers) t.join();
    for (size_t l = 0; l < s->ctxs.size(); ++l) {
        apspie_free(s->ctxs[l], s->d_srcs[l]);
        apspie_ctx_destroy(s->ctxs[l]);
    }
    delete s;
    return APSPIE_OK;
}

}  // extern "C"
