"""GPU tests of the serving loop: every served result equals a solo plain
launch of the same request (alone and in a mixed stream), admission by
memory footprint refuses what can never fit and queues what must wait, and
the report counts what it served."""
import numpy as np
import pytest

import apspie_amd as A
from apspie_amd import serving

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pool(ctx):
    hosts = serving.build_pool([10, 11, 12, 13])
    return hosts, [ctx.upload(h) for h in hosts]


def solo(ctx, graphs, r):
    g = graphs[r.graph]
    return ctx.bfs(g, r.src)[0] if r.kind == A.REQ_BFS else ctx.sssp(g, r.src)[0]


@pytest.mark.parametrize("plain", [0, -1, 1000])
def test_served_equals_solo_in_a_mix(ctx, pool, plain):
    """plain: batches up to that size as plain launches (0 default, -1
    never, 1000 always)."""
    hosts, graphs = pool
    reqs = serving.make_stream([h.m for h in hosts], n=600, rate=50000.0, seed=7)
    outs = {i: np.empty(hosts[r.graph].m, np.int32 if r.kind == A.REQ_BFS else np.float32)
            for i, r in enumerate(reqs)}
    with A.Server(graphs, lanes=3, plain_bfs_upto=plain, plain_sssp_upto=plain) as srv:
        rep = serving.run_served(srv, reqs, open_loop=True, outs=outs)
    assert rep["completed"] == len(reqs) and rep["rejected"] == 0
    assert rep["batches"] < len(reqs)  # requests were batched together
    assert 0 < rep["trivial"] < len(reqs)  # sources without out-edges answered on the host
    assert rep["p50_ms"] <= rep["p99_ms"]
    for i, r in enumerate(reqs):
        assert np.array_equal(outs[i], solo(ctx, graphs, r)), i


def test_alone_equals_mix(ctx, pool):
    hosts, graphs = pool
    r = serving.Req(A.REQ_SSSP, 3, 5, 0.0)
    alone = np.empty(hosts[3].m, np.float32)
    with A.Server(graphs, lanes=1) as srv:
        srv.wait(srv.submit(r.kind, r.graph, r.src, alone))
    assert np.array_equal(alone, solo(ctx, graphs, r))


def test_sources_without_out_edges(ctx, pool):
    hosts, graphs = pool
    h = hosts[2]
    dead = np.nonzero(h.degrees() == 0)[0][:3]
    live = np.nonzero(h.degrees() > 0)[0][:3]
    assert len(dead) == 3 and len(live) == 3
    with A.Server(graphs, lanes=2) as srv:
        outs, ts = [], []
        for kind in (A.REQ_BFS, A.REQ_SSSP):
            for v in list(dead) + list(live):
                o = np.empty(h.m, np.int32 if kind == A.REQ_BFS else np.float32)
                outs.append((kind, int(v), o))
                ts.append(srv.submit(kind, 2, int(v), o))
        for t in ts:
            srv.wait(t)
        assert srv.report()["trivial"] == 6
    for kind, v, o in outs:
        assert np.array_equal(o, solo(ctx, graphs, serving.Req(kind, 2, v, 0.0))), (kind, v)


def test_admission_by_footprint(ctx, pool):
    hosts, graphs = pool
    m3 = hosts[3].m
    # room for about two SSSP requests on graph 3 at a time
    with A.Server(graphs, lanes=4, budget_bytes=2 * (13 * m3 + 4096) + 100) as srv:
        outs = [np.empty(m3, np.float32) for _ in range(12)]
        ts = [srv.submit(A.REQ_SSSP, 3, 7 * i, outs[i]) for i in range(12)]
        for t in ts:
            srv.wait(t)
        rep = srv.report()
        assert rep["completed"] == 12 and rep["batches"] >= 6
        for i in range(12):
            assert np.array_equal(outs[i], ctx.sssp(graphs[3], 7 * i)[0])
    with A.Server(graphs, lanes=1, budget_bytes=1000) as srv:
        with pytest.raises(A.ApspieError) as e:
            srv.submit(A.REQ_BFS, 0, 0)
        assert e.value.code == -7
        assert srv.report()["rejected"] == 1


def test_lane_tuning_keeps_results(ctx, pool):
    """apspie_server_set_tuning: the lanes take any valid knob setting with
    results unchanged, refuse an out-of-range one, and the Python side
    insists on every knob."""
    hosts, graphs = pool
    reqs = serving.make_stream([h.m for h in hosts], n=200, rate=50000.0, seed=11)
    outs = {i: np.empty(hosts[r.graph].m, np.int32 if r.kind == A.REQ_BFS else np.float32)
            for i, r in enumerate(reqs)}
    knobs = ctx.tuning()
    knobs.update(sssp_bin_ratio=2, sssp_hint=0, expand_wg_per_cu=3, round_chunk=1, batch_chunk=2, bfs_alpha=1)
    with A.Server(graphs, lanes=2, plain_bfs_upto=-1, plain_sssp_upto=2) as srv:
        srv.set_tuning(knobs)
        rep = serving.run_served(srv, reqs, open_loop=False, outs=outs)
        bad = dict(knobs, round_chunk=0)
        with pytest.raises(A.ApspieError):
            srv.set_tuning(bad)
        with pytest.raises(KeyError):
            srv.set_tuning({"round_chunk": 4})
    assert rep["completed"] == len(reqs)
    for i, r in enumerate(reqs):
        assert np.array_equal(outs[i], solo(ctx, graphs, r)), i


def oracle_sssp(h, src):
    import oracle as O
    d = O.sssp_dijkstra(O.Csr(h.m, h.n, h.row_ptr, h.col_ind, h.val), src)
    d[d == -1.0] = np.float32(1.70141e38)
    return d


@pytest.mark.parametrize("lanes_knob", [0, 64])
def test_sssp_requests_join_running_batches(ctx, lanes_knob):
    """Continuous batching: with one lane and every SSSP batch run batched
    (no plain launches), SSSP requests submitted while a batch of their graph
    runs take its free lanes at the next chunk of rounds (16-lane batches by
    default, 64 with sssp_batch_lanes = 64).  Some requests join, and every
    row equals the oracle's Dijkstra distances, whatever round it joined in."""
    import time
    h = A.rmat(15, 16, 9115, undirected=False, weight_seed=9116)
    g = ctx.upload(h)
    live = np.nonzero(h.degrees() > 0)[0]
    srcs = [int(x) for x in np.random.RandomState(3).choice(live, 40, replace=False)]
    outs = [np.empty(h.m, np.float32) for _ in srcs]
    with A.Server([g], lanes=1, plain_sssp_upto=-1) as srv:
        t = ctx.tuning()
        t.update(batch_chunk=1, sssp_batch_lanes=lanes_knob)
        srv.set_tuning(t)
        srv.reset_stats()
        tickets = []
        for s, o in zip(srcs, outs):
            tickets.append(srv.submit(A.REQ_SSSP, 0, s, o))
            time.sleep(0.0003)
        for tk in tickets:
            srv.wait(tk)
        rep = srv.report()
    assert rep["completed"] == len(srcs) and rep["joined"] > 0, rep
    for s, o in zip(srcs, outs):
        assert np.array_equal(o, oracle_sssp(h, s)), s


def test_bfs_requests_join_running_batches(ctx):
    """Continuous batching for BFS: requests submitted while a BFS batch of
    their graph runs take its free lanes at the next chunk of rounds (every
    batch batched, one lane); each joined source's depths count from the round
    that picks it, so every row equals the oracle's BFS."""
    import time
    import oracle as O
    h = A.rmat(15, 16, 9215, undirected=False)
    g = ctx.upload(h)
    live = np.nonzero(h.degrees() > 0)[0]
    srcs = [int(x) for x in np.random.RandomState(4).choice(live, 60, replace=False)]
    outs = [np.empty(h.m, np.int32) for _ in srcs]
    with A.Server([g], lanes=1, plain_bfs_upto=-1) as srv:
        t = ctx.tuning()
        t.update(batch_chunk=1)
        srv.set_tuning(t)
        srv.reset_stats()
        tickets = []
        for s, o in zip(srcs, outs):
            tickets.append(srv.submit(A.REQ_BFS, 0, s, o))
            time.sleep(0.0002)
        for tk in tickets:
            srv.wait(tk)
        rep = srv.report()
    assert rep["completed"] == len(srcs) and rep["joined"] > 0, rep
    oc = O.Csr(h.m, h.n, h.row_ptr, h.col_ind, None)
    for s, o in zip(srcs, outs):
        assert np.array_equal(o, O.bfs(oc, s)[0]), s
