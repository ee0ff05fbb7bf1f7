"""GPU parity of the SSSP distance bound codes (tuning knob sssp_hint): the
push rounds skip a target's distance probe when the edge's candidate is not
below the byte-coded bound of the target's distance.  Distances stay
bit-exact against the oracle, and the schedule (rounds, frontier sizes,
relaxed edges) is the one without the codes, for weights whose distances
fall in every code range: 0, the exact integers up to 127, the stepped range
up to 504 and beyond it (no bound), fractional and sub-unit values."""
import numpy as np
import pytest

import apspie_amd as A
import oracle as O

pytestmark = pytest.mark.gpu

INF = np.float32(1.70141e38)


def oracle_dist(h, src):
    d = O.sssp_dijkstra(O.Csr(h.m, h.n, h.row_ptr, h.col_ind, h.val), src)
    d[d == -1.0] = INF
    return d


@pytest.fixture
def tuned(ctx):
    saved = ctx.tuning()
    yield ctx
    ctx.set_tuning(**saved)


def weighted(kind):
    h = A.rmat(13, 16, 7113, undirected=False, weight_seed=8113)
    rs = np.random.RandomState(11)
    if kind == "zero":
        h.val = np.where(rs.rand(h.nnz) < 0.3, 0.0, h.val).astype(np.float32)
    elif kind == "fractional":
        h.val = (h.val + 0.25).astype(np.float32)
    elif kind == "stepped":  # distances mostly between 128 and 504, not on the code grid
        h.val = (h.val * 3.0 + 0.5).astype(np.float32)
    elif kind == "wide":  # distances across all three ranges and past 504
        h.val = (h.val * 9.0).astype(np.float32)
    elif kind == "subunit":  # every distance below 1: one code, no filtering
        h.val = (h.val * 1e-3).astype(np.float32)
    elif kind == "large":
        h.val = (h.val * 1e6).astype(np.float32)
    return h


@pytest.mark.parametrize("kind", ["integer", "zero", "fractional", "stepped", "wide", "subunit", "large"])
def test_hint_exact_and_same_schedule(tuned, kind):
    h = weighted(kind)
    g = tuned.upload(h)
    for src in (h.argmax_degree(), 17, 4000):
        want = oracle_dist(h, src)
        got = {}
        for hint in (0, 1):
            tuned.set_tuning(sssp_hint=hint)
            dist, st = tuned.sssp(g, src)
            assert np.array_equal(dist, want), (kind, src, hint)
            got[hint] = (st.edges, st.vertices, st.iterations, tuned.rounds())
        assert got[0] == got[1], (kind, src)


@pytest.mark.parametrize("mode", [dict(sssp_delta=16), dict(sssp_delta=3), dict(sssp_pull_ratio=2),
                                  dict(sssp_bin_ratio=2), dict(sssp_delta=24, sssp_pull_ratio=2)])
def test_hint_with_other_schedules(tuned, mode):
    h = weighted("stepped")
    g = tuned.upload(h, csc=A.transpose(h))
    src = h.argmax_degree()
    tuned.set_tuning(sssp_hint=1, **mode)
    dist, _ = tuned.sssp(g, src)
    assert np.array_equal(dist, oracle_dist(h, src)), mode


def test_hint_reused_buffers_across_graphs(tuned):
    """The codes of one traversal never leak into the next: a second graph
    with smaller distances, then the first graph again, on one context."""
    tuned.set_tuning(sssp_hint=1)
    a = weighted("wide")
    b = A.rmat(12, 8, 7212, undirected=False, weight_seed=8212)
    ga, gb = tuned.upload(a), tuned.upload(b)
    for h, g in ((a, ga), (b, gb), (a, ga)):
        src = h.argmax_degree()
        dist, _ = tuned.sssp(g, src)
        assert np.array_equal(dist, oracle_dist(h, src))


def fan_into_one_bucket(n_mid=300_000, per=2, bucket=20):
    """Vertex 0 reaches n_mid middle vertices; each middle vertex has `per`
    edges into one 16K-vertex bucket, so the second round sends
    n_mid * per candidates (more than one apply part) to that bucket."""
    lo = bucket << 14
    m = lo + (1 << 14)
    rs = np.random.RandomState(21)
    deg = np.zeros(m, np.int64)
    deg[0] = n_mid
    deg[1:n_mid + 1] = per
    rp = np.concatenate([[0], np.cumsum(deg)]).astype(np.int32)
    col = np.empty(rp[-1], np.int32)
    col[:n_mid] = np.arange(1, n_mid + 1)
    tgt = lo + rs.randint(0, 1 << 14, size=(n_mid, per))
    tgt.sort(axis=1)
    col[n_mid:] = tgt.ravel()
    val = rs.randint(1, 65, size=rp[-1]).astype(np.float32)
    return A.HostCsr(m, m, rp, col, val)


@pytest.mark.parametrize("hint", [0, 1])
def test_binned_heavy_bucket_split_into_parts(tuned, hint):
    h = fan_into_one_bucket()
    g = tuned.upload(h)
    want = oracle_dist(h, 0)
    tuned.set_tuning(sssp_hint=hint, sssp_bin_ratio=2)
    dist, st = tuned.sssp(g, 0)
    assert np.array_equal(dist, want)
    assert tuned.rounds()[1] == (300_000, 600_000, 2)  # the fan-in round ran binned
    tuned.set_tuning(sssp_bin_ratio=0)
    dist0, st0 = tuned.sssp(g, 0)
    assert np.array_equal(dist0, want) and (st0.edges, st0.vertices) == (st.edges, st.vertices)


@pytest.mark.parametrize("hint", [0, 1])
@pytest.mark.parametrize("ratio", [2, 4])
def test_binned_rounds_rmat_s18(tuned, hint, ratio):
    """Binned dense rounds on RMAT s18 (a few buckets, hub buckets split),
    against the oracle and against push rounds' schedule."""
    h = A.rmat(18, 16, 7318, undirected=False, weight_seed=8318)
    g = tuned.upload(h)
    src = h.argmax_degree()
    want = oracle_dist(h, src)
    tuned.set_tuning(sssp_hint=hint, sssp_bin_ratio=ratio)
    dist, st = tuned.sssp(g, src)
    rounds = tuned.rounds()
    assert np.array_equal(dist, want)
    assert any(p == 2 for _, _, p in rounds)
    tuned.set_tuning(sssp_bin_ratio=0)
    dist0, st0 = tuned.sssp(g, src)
    assert [r[:2] for r in tuned.rounds()] == [r[:2] for r in rounds]
    assert (st0.edges, st0.vertices, st0.iterations) == (st.edges, st.vertices, st.iterations)


@pytest.mark.parametrize("kind", ["integer", "stepped", "wide", "fractional", "zero"])
@pytest.mark.parametrize("k", [16, 64])
def test_batched_rows_exact_for_weight_kinds(tuned, kind, k):
    """Batched SSSP (16- and 64-lane forms) on the same weight kinds: every
    row equals the oracle's single-source distances."""
    h = weighted(kind)
    g = tuned.upload(h)
    cand = np.nonzero(h.degrees() > 0)[0]
    srcs = np.random.RandomState(k).choice(cand, k, replace=False).astype(np.int32)
    want = np.stack([oracle_dist(h, int(s)) for s in srcs])
    dist, _ = tuned.sssp_batch(g, srcs)
    assert np.array_equal(dist, want), (kind, k)


@pytest.mark.parametrize("chunk", [1, 2, 3, 5, 12])
def test_binned_launches_stop_after_the_dense_rounds(tuned, chunk):
    """With binned rounds on, the binned kernels are launched until a push
    round follows binned ones at a chunk boundary; the later rounds run the
    plain push expansion.  Results and the schedule are the same for every
    chunk length, and relax_rounds covers every binned round."""
    h = A.rmat(18, 16, 7418, undirected=False, weight_seed=8418)
    g = tuned.upload(h)
    src = h.argmax_degree()
    want = oracle_dist(h, src)
    tuned.set_tuning(sssp_bin_ratio=0)
    _, st0 = tuned.sssp(g, src)
    r0 = tuned.rounds()
    assert st0.relax_rounds == 0
    tuned.set_tuning(sssp_bin_ratio=3, round_chunk=chunk)
    dist, st = tuned.sssp(g, src)
    rounds = tuned.rounds()
    assert np.array_equal(dist, want)
    assert [r[:2] for r in rounds] == [r[:2] for r in r0]
    binned = [i for i, (_, _, p) in enumerate(rounds) if p == 2]
    assert binned and st.relax_rounds >= binned[-1] + 1
    assert st.relax_rounds <= len(rounds) + max(chunk, 4)
