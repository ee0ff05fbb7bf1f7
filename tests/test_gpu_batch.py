"""GPU parity of the batched path: every row of apspie_bfs_batch /
apspie_sssp_batch must equal the oracle's single-source result for that
source (bit-exact depths and distances), including duplicate sources,
partial 64-bit words and the RMAT scale-20 batch configuration."""
import numpy as np
import pytest

import apspie_amd as A
import oracle as O

pytestmark = pytest.mark.gpu
INF = np.float32(1.70141e38)


def ocsr(h):
    return O.Csr(h.m, h.n, h.row_ptr, h.col_ind, h.val)


def sources(h, k, seed):
    """k sources drawn with mt19937 from vertices with degree > 0."""
    cand = np.nonzero(h.degrees() > 0)[0]
    draws = O.mt_fill(seed, k)
    return cand[(draws.astype(np.uint64) * np.uint64(len(cand))) >> np.uint64(32)].astype(np.int32)


@pytest.mark.parametrize("k", [1, 63, 64, 65, 130, 256])
def test_bfs_batch_rows_match_single(ctx, k):
    h = A.rmat(12, 16, 4012, undirected=True)
    g = ctx.upload(h)
    srcs = sources(h, k, 77 + k)
    srcs[-1] = srcs[0]  # a duplicate source
    depth, st = ctx.bfs_batch(g, srcs)
    o = ocsr(h)
    for s, v in enumerate(srcs):
        assert np.array_equal(depth[s], O.bfs(o, int(v))[0]), (k, s)


def test_bfs_batch_directed(ctx):
    h = A.rmat(11, 8, 4111, undirected=False)
    g = ctx.upload(h)
    srcs = np.arange(0, 2048, 16, dtype=np.int32)
    depth, _ = ctx.bfs_batch(g, srcs)
    o = ocsr(h)
    for s, v in enumerate(srcs):
        assert np.array_equal(depth[s], O.bfs(o, int(v))[0])


@pytest.mark.parametrize("k", [1, 17, 64])
def test_sssp_batch_rows_match_single(ctx, k):
    h = A.rmat(12, 16, 4212, undirected=False, weight_seed=5212)
    g = ctx.upload(h)
    srcs = sources(h, k, 99 + k)
    srcs[-1] = srcs[0]
    dist, _ = ctx.sssp_batch(g, srcs)
    o = ocsr(h)
    for s, v in enumerate(srcs):
        want = O.sssp_dijkstra(o, int(v))
        want[want == -1.0] = INF
        assert np.array_equal(dist[s], want), (k, s)


def path_graph(m):
    """Undirected path 0 - 1 - ... - (m - 1)."""
    a = np.arange(m - 1, dtype=np.int64)
    u = np.concatenate([a, a + 1])
    v = np.concatenate([a + 1, a])
    order = np.lexsort((v, u))
    u, v = u[order], v[order]
    rp = np.zeros(m + 1, np.int32)
    np.add.at(rp, u + 1, 1)
    rp = np.cumsum(rp).astype(np.int32)
    return A.HostCsr(m, m, rp, v.astype(np.int32), np.ones(len(v), np.float32))


@pytest.mark.parametrize("k", [3, 70, 200])
def test_bfs_batch_on_a_long_path(ctx, k):
    """A traversal of hundreds of rounds (many captured chunks, the last one
    partly idle): a 700-vertex path from its ends and spread vertices,
    depths up to 699 (past what one byte holds)."""
    h = path_graph(700)
    g = ctx.upload(h)
    srcs = np.linspace(0, 699, k).astype(np.int32)
    srcs[0], srcs[-1] = 0, 699
    depth, st = ctx.bfs_batch(g, srcs)
    o = ocsr(h)
    for s, v in enumerate(srcs):
        assert np.array_equal(depth[s], O.bfs(o, int(v))[0]), (k, s)
    assert depth.max() == 699


def test_batch_limits(ctx):
    h = A.rmat(8, 4, 1, undirected=True)
    g = ctx.upload(h)
    with pytest.raises(A.ApspieError):
        ctx.bfs_batch(g, np.zeros(257, np.int32))
    with pytest.raises(A.ApspieError):
        ctx.sssp_batch(g, np.zeros(65, np.int32))
    with pytest.raises(A.ApspieError):
        ctx.bfs_batch(g, np.array([0, h.m], np.int32))


def test_batch_config_s20(ctx):
    """256-source BFS and 64-source SSSP on RMAT scale 20 (ef 16, weights
    1..64), sources from vertices with degree > 0; 8 rows of each checked in
    full against the oracle."""
    hb = A.rmat(20, 16, 4020, undirected=True)
    gb = ctx.upload(hb)
    sb = sources(hb, 256, 4020)
    depth, _ = ctx.bfs_batch(gb, sb)
    o = ocsr(hb)
    for s in range(0, 256, 37):
        assert np.array_equal(depth[s], O.bfs(o, int(sb[s]))[0])
    del gb
    hs = A.rmat(20, 16, 4120, undirected=False, weight_seed=4120 + 1_000_000)
    gs = ctx.upload(hs)
    ss = sources(hs, 64, 4120)
    dist, _ = ctx.sssp_batch(gs, ss)
    o = ocsr(hs)
    for s in range(0, 64, 9):
        want = O.sssp_dijkstra(o, int(ss[s]))
        want[want == -1.0] = INF
        assert np.array_equal(dist[s], want)
