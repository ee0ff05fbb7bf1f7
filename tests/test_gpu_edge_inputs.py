"""GPU behaviour of every entry point on degenerate inputs: a graph with no
edges, a single vertex, a single edge, and a self-contained pair; results
against the oracle (or the obvious closed form), no hangs, no faults."""
import numpy as np
import pytest

import apspie_amd as A
import oracle as O

pytestmark = pytest.mark.gpu
INF = np.float32(1.70141e38)


def empty(m):
    return A.HostCsr(m, m, np.zeros(m + 1, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32))


def one_edge(m, a, b, w=3.0):
    rp = np.zeros(m + 1, np.int32)
    rp[a + 1:] = 1
    return A.HostCsr(m, m, rp, np.array([b], np.int32), np.array([w], np.float32))


@pytest.mark.parametrize("m", [1, 7])
def test_traversals_without_edges(ctx, m):
    g = ctx.upload(empty(m))
    for src in (0, m - 1):
        d, _ = ctx.bfs(g, src)
        want = np.full(m, -1, np.int32)
        want[src] = 0
        assert np.array_equal(d, want)
        dist, _ = ctx.sssp(g, src)
        wd = np.full(m, INF, np.float32)
        wd[src] = 0
        assert np.array_equal(dist, wd)
    srcs = np.arange(m, dtype=np.int32)
    bd, _ = ctx.bfs_batch(g, srcs)
    sd, _ = ctx.sssp_batch(g, srcs)
    assert np.array_equal(bd, np.where(np.eye(m, dtype=bool), 0, -1).astype(np.int32))
    assert np.array_equal(sd, np.where(np.eye(m, dtype=bool), 0, INF).astype(np.float32))


def test_mis_tc_spmspv_without_edges(ctx):
    m = 9
    g = ctx.upload(empty(m))
    mis, _ = ctx.mis(g, np.arange(m, dtype=np.uint32))
    assert np.array_equal(mis, np.ones(m, np.int32))  # every vertex is independent
    t, st = ctx.tc(g)
    assert t == 0 and st.edges == 0
    y = ctx.spmspv_or(g, np.ones(m, np.int32))
    assert not y.any()


def test_spgemm_with_empty_operand(ctx):
    m = 6
    a = ctx.upload(one_edge(m, 1, 2))
    z = ctx.upload(empty(m))
    for x, y in ((a, z), (z, a), (z, z)):
        c, st = ctx.spgemm(x, y)
        assert c.nnz == 0 and np.array_equal(c.row_ptr, np.zeros(m + 1, np.int32))


def test_single_edge(ctx):
    m = 5
    h = one_edge(m, 3, 1, 7.0)
    g = ctx.upload(h)
    d, _ = ctx.bfs(g, 3)
    assert list(d) == [-1, 1, -1, 0, -1]
    dist, _ = ctx.sssp(g, 3)
    assert dist[1] == 7.0 and dist[3] == 0 and (dist[[0, 2, 4]] == INF).all()
    o = O.Csr(h.m, h.n, h.row_ptr, h.col_ind, h.val)
    sd, _ = ctx.sssp_batch(g, np.array([3, 1, 3], np.int32))
    for row, s in zip(sd, (3, 1, 3)):
        want = O.sssp_dijkstra(o, s)
        want[want == -1.0] = INF
        assert np.array_equal(row, want)


def test_tc_after_scratch_reuse(ctx):
    """TC's row-limit flag lives in a scratch slot that other entry points fill
    (the device CSR->CSC keeps row indices there): a count that follows such a
    call must still see the flag cleared and return the oracle's count."""
    h = A.rmat(10, 8, seed=5, undirected=True)
    ctx.csr2csc(ctx.upload(h), values=True)  # leaves nonzero row indices in the shared slot
    lo = A.strict_lower(h)
    t, _ = ctx.tc(ctx.upload(lo))
    assert t == O.tc(O.Csr(lo.m, lo.n, lo.row_ptr, lo.col_ind, None))[0]
    t0, _ = ctx.tc(ctx.upload(empty(9)))
    assert t0 == 0
