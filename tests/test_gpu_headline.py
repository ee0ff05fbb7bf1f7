"""GPU parity at the headline sizes, against the CPU oracle.

- SSSP on the directed RMAT scale-22 graph (edge factor 16, weights 1..64)
  from its highest-degree vertex: distances bit-exact against the heap
  Dijkstra of src/testSssp.cpp:33-115, and the call's counters against the
  synchronous Bellman-Ford schedule the GPU follows.
- BFS on the undirected RMAT scale-22 graph from vertex 0, push only and
  direction-optimised over the CSC: depths and counters against the FIFO BFS
  of src/testBfs.cpp:6-72.
- Luby MIS on the undirected RMAT scale-21 graph with MT19937 priorities
  against the priority-order greedy MIS (test/mis/main.cu:26-88 in priority
  order), plus the XOR check of src/util.cuh:277.
- The served stream over the 16-graph pool of scales 16-22: the first
  requests of the Poisson stream the benchmark replays, every served row
  checked against the oracle (not against plain launches).

The graphs are the ones the benchmark times (same generators and seeds).
"""
import concurrent.futures as cf

import numpy as np
import pytest

import apspie_amd as A
import oracle as O
from apspie_amd import serving

pytestmark = pytest.mark.gpu

INF = np.float32(1.70141e38)
SSSP_SEED, SSSP_WSEED = 2022, 2022 + 1_000_000
BFS_SEED = 1022
MIS_SEED = 4021


def ocsr(h):
    return O.Csr(h.m, h.n, h.row_ptr, h.col_ind, h.val)


def oracle_dist(o, src):
    d = O.sssp_dijkstra(o, src)
    d[d == -1.0] = INF
    return d


@pytest.fixture(scope="module")
def sssp_s22():
    h = A.rmat(22, 16, SSSP_SEED, undirected=False, weight_seed=SSSP_WSEED)
    o = ocsr(h)
    src = h.argmax_degree()
    assert src == O.argmax_degree(o)
    print(f"sssp s22: m={h.m} nnz={h.nnz} src={src}", flush=True)
    return h, o, src, oracle_dist(o, src), O.sssp_bf(o, src)


@pytest.mark.parametrize("mode", ["push", "pull", "near_far"])
def test_sssp_rmat_s22_directed(ctx, sssp_s22, mode):
    """Headline workload: bit-exact distances, and (iterations, relaxed
    edges, frontier entries, improvements) equal to the oracle's synchronous
    schedule in every mode (pull rounds and near/far only move time)."""
    h, o, src, want, bf = sssp_s22
    g = ctx.upload(h, csc=A.transpose(h) if mode == "pull" else None)
    saved = ctx.tuning()
    try:
        if mode == "pull":
            ctx.set_tuning(sssp_pull_ratio=2)
        if mode == "near_far":
            ctx.set_tuning(sssp_delta=32)
        dist, st = ctx.sssp(g, src)
    finally:
        ctx.set_tuning(**saved)
    assert np.array_equal(dist, want)
    _, it, er, sf, ri = bf
    if mode != "near_far":  # near/far rounds follow another schedule
        assert (st.iterations, st.edges, st.vertices, st.improved) == (it, er, sf, ri)
    assert st.bytes_alg > 0


@pytest.fixture(scope="module")
def bfs_s22():
    h = A.rmat(22, 16, BFS_SEED, undirected=True)
    print(f"bfs s22: m={h.m} nnz={h.nnz}", flush=True)
    return h, O.bfs(ocsr(h), 0)


@pytest.mark.parametrize("csc", [None, "symmetric"])
def test_bfs_rmat_s22_undirected(ctx, bfs_s22, csc):
    h, (want, sd, e_exp, vr) = bfs_s22
    g = ctx.upload(h, csc=csc)
    depth, st = ctx.bfs(g, 0)
    assert np.array_equal(depth, want)
    assert (st.edges, st.vertices, st.iterations) == (e_exp, vr, sd)
    assert st.bytes_alg == 8 * e_exp + 20 * vr + 4 * h.m
    if csc:  # the dense levels ran bottom-up
        assert any(p for _, _, p in ctx.rounds())


def test_bfs_rmat_s22_on_the_sssp_graph(ctx, sssp_s22):
    """BFS over the directed s22 graph (the survey's second BFS figure)."""
    h, o, src, _, _ = sssp_s22
    g = ctx.upload(h)
    depth, st = ctx.bfs(g, src)
    want, sd, e_exp, vr = O.bfs(o, src)
    assert np.array_equal(depth, want)
    assert (st.edges, st.vertices, st.iterations) == (e_exp, vr, sd)


def test_mis_rmat_s21_mt19937(ctx):
    h = A.rmat(21, 16, MIS_SEED, undirected=True)
    prio = serving.mt19937_stream(MIS_SEED, h.m).astype(np.uint32)
    assert np.array_equal(prio, O.mt_fill(MIS_SEED, h.m))
    g = ctx.upload(h)
    got, st = ctx.mis(g, prio)
    want = O.mis_greedy(ocsr(h), prio)
    assert np.array_equal(got, want)
    has = ctx.spmspv_or(g, got)
    assert np.all(got + has == 1)  # src/util.cuh:277: in the set XOR a neighbour in it
    assert st.iterations >= 1


@pytest.fixture(scope="module")
def served_pool(ctx):
    hosts = serving.build_pool()
    print("served pool: " + " ".join(f"s{s}:{h.nnz}" for s, h in zip(serving.POOL_SCALES, hosts)), flush=True)
    return hosts, [ctx.upload(h) for h in hosts]


@pytest.mark.parametrize("open_loop", [False, True])
def test_served_stream_on_the_pool_against_the_oracle(served_pool, open_loop):
    """The first 400 requests of the benchmark's stream (seed 5555, 50/50
    BFS/SSSP, uniform graph and source over the s16-s22 pool), all released at
    once (batches form) or at 20,000 req/s (requests join while others run);
    every row equals the oracle's answer for that request."""
    hosts, graphs = served_pool
    reqs = serving.make_stream([h.m for h in hosts], n=400, rate=20000.0)
    outs = {i: np.empty(hosts[r.graph].m, np.int32 if r.kind == A.REQ_BFS else np.float32)
            for i, r in enumerate(reqs)}
    with A.Server(graphs, lanes=4) as srv:
        rep = serving.run_served(srv, reqs, open_loop=open_loop, outs=outs)
    assert rep["completed"] == len(reqs) and rep["rejected"] == 0
    assert 0 < rep["trivial"] < len(reqs)
    if not open_loop:
        assert rep["mean_batch"] > 1.5  # released together, they were batched
    ocs = [ocsr(h) for h in hosts]

    def check(i):
        r = reqs[i]
        if r.kind == A.REQ_BFS:
            return np.array_equal(outs[i], O.bfs(ocs[r.graph], r.src)[0])
        return np.array_equal(outs[i], oracle_dist(ocs[r.graph], r.src))

    with cf.ThreadPoolExecutor(8) as ex:  # the oracle releases the GIL in C
        ok = list(ex.map(check, range(len(reqs))))
    bad = [i for i, v in enumerate(ok) if not v]
    assert not bad, [(i, reqs[i]) for i in bad[:5]]
