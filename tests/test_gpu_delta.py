"""GPU parity of the near/far SSSP schedule (tuning knob sssp_delta): a
changed vertex waits while its distance is at or past a threshold, and the
threshold rises by delta once the frontier below it runs dry.  Distances stay
bit-exact against the oracle for every delta, with zero, fractional and large
weights, and with the pull and binned rounds."""
import numpy as np
import pytest

import apspie_amd as A
import oracle as O

pytestmark = pytest.mark.gpu

INF = np.float32(1.70141e38)
REPICK = 3  # round whose frontier was empty while vertices waited


def oracle_dist(h, src):
    d = O.sssp_dijkstra(O.Csr(h.m, h.n, h.row_ptr, h.col_ind, h.val), src)
    d[d == -1.0] = INF
    return d


@pytest.fixture
def tuned(ctx):
    saved = ctx.tuning()
    yield ctx
    ctx.set_tuning(**saved)


@pytest.mark.parametrize("delta", [1, 7, 32, 64, 1000])
def test_delta_rmat_exact(tuned, delta):
    tuned.set_tuning(sssp_delta=delta)
    h = A.rmat(13, 16, 4713, undirected=False, weight_seed=5713)
    g = tuned.upload(h)
    for src in (h.argmax_degree(), 5, 1234):
        dist, _ = tuned.sssp(g, src)
        assert np.array_equal(dist, oracle_dist(h, src)), (delta, src)


def test_delta_repicks_and_counts(tuned):
    """A small delta raises the threshold many times; the round log shows
    the repick rounds, and the distances are still exact."""
    tuned.set_tuning(sssp_delta=4)
    h = A.rmat(12, 8, 4812, undirected=False, weight_seed=5812)
    g = tuned.upload(h)
    src = h.argmax_degree()
    dist, st = tuned.sssp(g, src)
    assert np.array_equal(dist, oracle_dist(h, src))
    assert any(p == REPICK for _, _, p in tuned.rounds())


@pytest.mark.parametrize("kind", ["zero", "fractional", "large"])
def test_delta_weight_kinds(tuned, kind):
    tuned.set_tuning(sssp_delta=16)
    h = A.rmat(12, 8, 4912, undirected=False, weight_seed=5912)
    rs = np.random.RandomState(4)
    if kind == "zero":
        h.val = np.where(rs.rand(h.nnz) < 0.3, 0.0, h.val).astype(np.float32)
    elif kind == "fractional":
        h.val = (h.val + 0.25).astype(np.float32)
    else:
        h.val = (h.val * 1000.0).astype(np.float32)
    g = tuned.upload(h)
    for src in (h.argmax_degree(), 17):
        dist, _ = tuned.sssp(g, src)
        assert np.array_equal(dist, oracle_dist(h, src)), (kind, src)


@pytest.mark.parametrize("mode", [dict(sssp_pull_ratio=2), dict(sssp_bin_ratio=2), dict(sssp_pull_ratio=1 << 20)])
def test_delta_with_pull_and_binned_rounds(tuned, mode):
    tuned.set_tuning(sssp_delta=24, **mode)
    h = A.rmat(13, 16, 5013, undirected=False, weight_seed=6013)
    g = tuned.upload(h, csc=A.transpose(h))
    src = h.argmax_degree()
    dist, _ = tuned.sssp(g, src)
    assert np.array_equal(dist, oracle_dist(h, src))


def test_delta_unreachable_and_isolated_sources(tuned):
    """A source without out-edges and one whose reach is small end cleanly."""
    tuned.set_tuning(sssp_delta=8)
    h = A.rmat(11, 4, 5111, undirected=False, weight_seed=6111)
    g = tuned.upload(h)
    deg = h.degrees()
    for src in (int(np.nonzero(deg == 0)[0][0]), int(np.nonzero(deg == 1)[0][0])):
        dist, _ = tuned.sssp(g, src)
        assert np.array_equal(dist, oracle_dist(h, src)), src


def test_delta_below_half_an_ulp_of_the_distances(tuned):
    """Distances past 2^28 with delta 16: there dmin + delta rounds back to
    dmin, so the threshold must still rise to the next float (else no round
    admits anything and the traversal stalls)."""
    tuned.set_tuning(sssp_delta=16)
    n = 40  # a path 0 -> 1 -> ... with weights 2^27: dist[k] = k * 2^27
    path = A.HostCsr(n, n, np.concatenate([np.arange(n, dtype=np.int32), [n - 1]]).astype(np.int32),
                     np.arange(1, n, dtype=np.int32), np.full(n - 1, 2.0 ** 27, np.float32))
    dist, _ = tuned.sssp(tuned.upload(path), 0)
    assert np.array_equal(dist, (np.arange(n) * 2.0 ** 27).astype(np.float32))
    # RMAT with weights (1..64) * 2^22: rounded sums, the same least fixed point
    h = A.rmat(10, 8, 4410, undirected=False, weight_seed=5410)
    h.val = (h.val * np.float32(2.0 ** 22)).astype(np.float32)
    g = tuned.upload(h)
    for src in (h.argmax_degree(), 3):
        d, _ = tuned.sssp(g, src)
        want = oracle_dist(h, src)
        assert want[want < INF].max() > 2.0 ** 28
        assert np.array_equal(d, want), src
