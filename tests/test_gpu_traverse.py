"""GPU parity of the plain traversal path (apspie_bfs, apspie_sssp,
apspie_sssp_step) against the CPU oracle: bit-exact depths and distances,
and the per-call counters against the oracle's schedule."""
import os

import numpy as np
import pytest

import apspie_amd as A
import oracle as O
from conftest import MTX

pytestmark = pytest.mark.gpu

NAMES = ["test_cc", "test_bc", "test_pr", "test_mesh", "test_mis", "test_spgemm", "chesapeake"]
INF = np.float32(1.70141e38)


def ocsr(h):
    return O.Csr(h.m, h.n, h.row_ptr, h.col_ind, h.val)


def dist_from_oracle(d):
    d = d.copy()
    d[d == -1.0] = INF
    return d


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("und", [False, True])
def test_bfs_small_files(ctx, name, und):
    h = A.load_mtx(os.path.join(MTX, name + ".mtx"), undirected=und)
    g = ctx.upload(h)
    for src in range(h.m):
        depth, st = ctx.bfs(g, src)
        want, sd, e_exp, vr = O.bfs(ocsr(h), src)
        assert np.array_equal(depth, want), (name, und, src)
        assert st.edges == e_exp and st.vertices == vr and st.iterations == sd


@pytest.mark.parametrize("scale,und,seed", [(10, True, 1010), (12, False, 2012), (14, True, 1014)])
def test_bfs_rmat(ctx, scale, und, seed):
    h = A.rmat(scale, 16, seed, undirected=und)
    g = ctx.upload(h)
    rng = np.random.RandomState(seed)
    srcs = [0, h.argmax_degree(), h.m - 1] + list(rng.randint(0, h.m, 3))
    for src in srcs:
        depth, st = ctx.bfs(g, int(src))
        want, sd, e_exp, vr = O.bfs(ocsr(h), int(src))
        assert np.array_equal(depth, want), (scale, src)
        assert (st.edges, st.vertices, st.iterations) == (e_exp, vr, sd)
        assert st.bytes_alg == 8 * e_exp + 20 * vr + 4 * h.m


def test_bfs_repeat_is_deterministic(ctx):
    h = A.rmat(13, 16, 77, undirected=True)
    g = ctx.upload(h)
    first, _ = ctx.bfs(g, 0)
    for _ in range(3):
        again, _ = ctx.bfs(g, 0)
        assert np.array_equal(first, again)


def test_bfs_edge_cases(ctx):
    # isolated source, zero-degree vertices, a single vertex, a long path
    rp = np.array([0, 0, 1, 2, 2], dtype=np.int32)
    ci = np.array([2, 3], dtype=np.int32)
    h = A.HostCsr(4, 4, rp, ci)
    g = ctx.upload(h)
    for src in range(4):
        d, _ = ctx.bfs(g, src)
        assert np.array_equal(d, O.bfs(ocsr(h), src)[0])
    one = ctx.upload(A.HostCsr(1, 1, np.zeros(2, np.int32), np.zeros(0, np.int32)))
    d, st = ctx.bfs(one, 0)
    assert list(d) == [0] and st.iterations == 1
    n = 5000  # path longer than any launch chunk
    path = A.HostCsr(n, n, np.concatenate([np.arange(n, dtype=np.int32), [n - 1]]).astype(np.int32),
                     np.arange(1, n, dtype=np.int32))
    gp = ctx.upload(path)
    d, st = ctx.bfs(gp, 0)
    assert np.array_equal(d, np.arange(n, dtype=np.int32)) and st.iterations == n
    with pytest.raises(A.ApspieError):
        ctx.bfs(gp, n)
    with pytest.raises(A.ApspieError):
        ctx.bfs(gp, -1)


@pytest.mark.parametrize("m", [3, 64, 65, 100, 16384, 16385])
def test_last_vertex_is_expanded(ctx, m):
    """0 -> m-1 -> 1 -> ... : the last vertex is the only bridge (its row end
    comes from row_ptr[m]); checks BFS, SSSP and their counters."""
    src = np.array([0, m - 1] + list(range(1, m - 2)), dtype=np.int64)
    dst = np.array([m - 1, 1] + list(range(2, m - 1)), dtype=np.int64)
    order = np.argsort(src, kind="stable")
    rp = np.zeros(m + 1, np.int32)
    np.add.at(rp, src + 1, 1)
    rp = np.cumsum(rp).astype(np.int32)
    h = A.HostCsr(m, m, rp, dst[order].astype(np.int32), np.full(len(src), 2.0, np.float32))
    g = ctx.upload(h)
    depth, st = ctx.bfs(g, 0)
    want, sd, e_exp, vr = O.bfs(ocsr(h), 0)
    assert np.array_equal(depth, want) and st.edges == e_exp and st.vertices == vr
    assert depth[1] == 2
    dist, _ = ctx.sssp(g, 0)
    assert np.array_equal(dist, dist_from_oracle(O.sssp_dijkstra(ocsr(h), 0)))


@pytest.mark.parametrize("scale,seed", [(8, 2008), (12, 2012), (14, 2014)])
def test_sssp_rmat(ctx, scale, seed):
    h = A.rmat(scale, 16, seed, undirected=False, weight_seed=seed + 1000)
    g = ctx.upload(h)
    o = ocsr(h)
    for src in [h.argmax_degree(), 0, h.m // 3]:
        dist, st = ctx.sssp(g, src)
        want = O.sssp_dijkstra(o, src)
        assert np.array_equal(dist, dist_from_oracle(want)), (scale, src)
        bf, it, er, sf, ri = O.sssp_bf(o, src)
        assert (st.iterations, st.edges, st.vertices, st.improved) == (it, er, sf, ri)


def test_sssp_small_files_unit_weights(ctx):
    for name in NAMES:
        h = A.load_mtx(os.path.join(MTX, name + ".mtx"), undirected=True)
        h.val = np.ones(h.nnz, np.float32)
        g = ctx.upload(h)
        for src in range(min(h.m, 8)):
            dist, _ = ctx.sssp(g, src)
            b = O.bfs(ocsr(h), src)[0].astype(np.float32)
            b[b < 0] = INF
            assert np.array_equal(dist, b), (name, src)


def test_sssp_unweighted_graph_is_bfs(ctx):
    h = A.rmat(12, 16, 5, undirected=True)
    g = ctx.upload(h)
    dist, _ = ctx.sssp(g, 0)
    b = O.bfs(ocsr(h), 0)[0].astype(np.float32)
    b[b < 0] = INF
    assert np.array_equal(dist, b)


@pytest.mark.parametrize("frac", [0.0, 0.05, 1.0])
def test_sssp_step(ctx, frac):
    h = A.rmat(11, 16, 99, undirected=False, weight_seed=98)
    g = ctx.upload(h)
    rng = np.random.RandomState(int(frac * 100))
    x = np.where(rng.rand(h.m) < frac, rng.randint(0, 1000, h.m), INF).astype(np.float32)
    x[:5] = -1.0
    y, st = ctx.sssp_step(g, x)
    assert np.array_equal(y, O.sssp_step(ocsr(h), x))
    # the step leaves the flag scratch clean for the next traversal
    d, _ = ctx.bfs(g, 0)
    assert np.array_equal(d, O.bfs(ocsr(h), 0)[0])
