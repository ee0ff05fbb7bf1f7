"""The tunable knobs and the autotuner (apspie_ctx_set_tuning,
apspie_autotune): every setting leaves results bit-exact against the oracle,
out-of-range knobs are refused, and the autotuner returns a setting the
context then uses."""
import numpy as np
import pytest

import apspie_amd as A
import oracle as O

pytestmark = pytest.mark.gpu

INF = np.float32(1.70141e38)


def ocsr(h):
    return O.Csr(h.m, h.n, h.row_ptr, h.col_ind, h.val)


@pytest.fixture
def tuned(ctx):
    saved = ctx.tuning()
    yield ctx
    ctx.set_tuning(**saved)


@pytest.fixture(scope="module")
def graph():
    h = A.rmat(13, 16, 913, undirected=False, weight_seed=914)
    return h, A.transpose(h)


def sources(h, k, seed):
    cand = np.nonzero(h.degrees() > 0)[0]
    return np.random.RandomState(seed).choice(cand, k, replace=False).astype(np.int32)


@pytest.mark.parametrize("knobs", [
    dict(expand_wg_per_cu=1), dict(expand_wg_per_cu=32), dict(round_chunk=1), dict(round_chunk=32),
    dict(batch_chunk=1), dict(batch_chunk=16), dict(sssp_batch_lanes=64), dict(sssp_batch_lanes=16),
    dict(bfs_alpha=1, bfs_beta=1000), dict(sssp_pull_ratio=2, sssp_bin_ratio=0), dict(sssp_bin_ratio=2),
    dict(sssp_delta=1), dict(sssp_delta=40), dict(sssp_delta=40, sssp_pull_ratio=2),
    dict(sssp_delta=40, sssp_bin_ratio=2), dict(sssp_hint=0), dict(sssp_hint=1, sssp_delta=40),
])
def test_every_knob_setting_is_exact(tuned, graph, knobs):
    ctx = tuned
    ctx.set_tuning(**knobs)
    h, t = graph
    g = ctx.upload(h, csc=t)
    o = ocsr(h)
    src = h.argmax_degree()
    depth, _ = ctx.bfs(g, src)
    assert np.array_equal(depth, O.bfs(o, src)[0])
    dist, _ = ctx.sssp(g, src)
    want = O.sssp_dijkstra(o, src)
    want[want == -1.0] = INF
    assert np.array_equal(dist, want)
    srcs = sources(h, 8, 3)
    dd, _ = ctx.sssp_batch(g, srcs)
    bd, _ = ctx.bfs_batch(g, srcs)
    for i, s in enumerate(srcs):
        w = O.sssp_dijkstra(o, int(s))
        w[w == -1.0] = INF
        assert np.array_equal(dd[i], w)
        assert np.array_equal(bd[i], O.bfs(o, int(s))[0])


def test_out_of_range_knobs_are_refused(tuned):
    ctx = tuned
    before = ctx.tuning()
    for bad in (dict(expand_wg_per_cu=0), dict(round_chunk=33), dict(batch_chunk=0), dict(sssp_batch_lanes=32),
                dict(sssp_delta=-1), dict(sssp_delta=(1 << 24) + 1), dict(spgemm_long_rows=2),
                dict(spgemm_long_rows=-1), dict(spgemm_passes=0), dict(spgemm_passes=3), dict(bfs_group_cap=-1),
                dict(bfs_group_cap=(1 << 30) + 1), dict(spgemm_esc_min=300), dict(spgemm_esc_min=4096),
                dict(sssp_hint=-1), dict(sssp_hint=2)):
        with pytest.raises(A.ApspieError):
            ctx.set_tuning(**bad)
        assert ctx.tuning() == before
    with pytest.raises(KeyError):
        ctx.set_tuning(no_such_knob=1)


def test_autotune_returns_the_setting_in_use_and_stays_exact(tuned, graph):
    ctx = tuned
    h, t = graph
    g = ctx.upload(h, csc=t)
    src = h.argmax_degree()
    kinds = A.TUNE_BFS | A.TUNE_SSSP | A.TUNE_BFS_BATCH | A.TUNE_SSSP_BATCH
    best = ctx.autotune(g, src, kinds, k=16)
    assert best == ctx.tuning()
    assert 1 <= best["expand_wg_per_cu"] <= 32 and best["sssp_batch_lanes"] in (0, 16, 64)
    o = ocsr(h)
    depth, _ = ctx.bfs(g, src)
    assert np.array_equal(depth, O.bfs(o, src)[0])
    dist, _ = ctx.sssp(g, src)
    want = O.sssp_dijkstra(o, src)
    want[want == -1.0] = INF
    assert np.array_equal(dist, want)


def test_autotune_errors_leave_the_context_unchanged(tuned):
    ctx = tuned
    before = ctx.tuning()
    h = A.rmat(10, 8, 77)  # unweighted: SSSP tuning is refused
    g = ctx.upload(h)
    with pytest.raises(A.ApspieError):
        ctx.autotune(g, 0, A.TUNE_SSSP)
    with pytest.raises(A.ApspieError):
        ctx.autotune(g, h.m, A.TUNE_BFS)
    assert ctx.tuning() == before
