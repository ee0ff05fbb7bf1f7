"""GPU parity of the pull (direction-optimised) rounds of apspie_bfs and
apspie_sssp, taken when the graph carries its CSC: depths and distances
bit-exact against the oracle, and the per-call counters equal to the push
schedule's (a pull round computes the same Jacobi step)."""
import os

import numpy as np
import pytest

import apspie_amd as A
import oracle as O
from conftest import MTX

pytestmark = pytest.mark.gpu

INF = np.float32(1.70141e38)
NAMES = ["test_cc", "test_bc", "test_pr", "test_mesh", "test_mis", "test_spgemm", "chesapeake"]
ALWAYS = dict(bfs_alpha=1 << 20, bfs_beta=1 << 20, sssp_pull_ratio=1 << 20, sssp_bin_ratio=0)  # pull every round
BINNED = dict(sssp_pull_ratio=0, sssp_bin_ratio=1 << 20)  # binned relaxation every SSSP round
PUSH = dict(bfs_alpha=0, sssp_pull_ratio=0, sssp_bin_ratio=0)


def ocsr(h):
    return O.Csr(h.m, h.n, h.row_ptr, h.col_ind, h.val)


def oracle_dist(h, src):
    d = O.sssp_dijkstra(ocsr(h), src)
    d[d == -1.0] = INF
    return d


@pytest.fixture
def tuned(ctx):
    """Restores the context's direction thresholds after the test."""
    saved = ctx.tuning()
    yield ctx
    ctx.set_tuning(**saved)


def with_csc(ctx, h, und):
    # a weighted undirected RMAT has w(u, v) != w(v, u): only its structure is symmetric
    return ctx.upload(h, csc="symmetric" if und and h.val is None else A.transpose(h))


@pytest.mark.parametrize("cap", [1, 16, 700])
@pytest.mark.parametrize("scale,und,seed", [(12, True, 1212), (14, False, 2214)])
def test_bfs_rounds_whose_queue_overflows_are_rebuilt(tuned, cap, scale, und, seed):
    """BFS with a CSC queues each round's frontier inside the round; with a
    group region capped at `cap` entries the larger frontiers overflow, and a
    top-down successor's queue is rebuilt by a pass over the round's bitmap
    (k_compact launches).  alpha = 1 keeps most rounds top-down.  Depths and
    counters equal the oracle's either way."""
    ctx = tuned
    ctx.set_tuning(bfs_alpha=1, bfs_beta=1 << 20, bfs_group_cap=cap)
    h = A.rmat(scale, 16, seed, undirected=und)
    g = with_csc(ctx, h, und)
    rebuilt = 0
    for src in [0, h.argmax_degree(), h.m // 3]:
        ctx.set_profiling(True)
        ctx.profile_reset()
        depth, st = ctx.bfs(g, int(src))
        rebuilt += ctx.profile()["compact"][1] - 1  # less the seed kernel's launch
        ctx.set_profiling(False)
        want, sd, e_exp, vr = O.bfs(ocsr(h), int(src))
        assert np.array_equal(depth, want), (scale, und, src, cap)
        assert (st.edges, st.vertices, st.iterations) == (e_exp, vr, sd)
        # the round log is the oracle's level by level
        rounds = ctx.rounds()
        assert len(rounds) == sd
    if cap <= 16:
        assert rebuilt > 0  # the capped regions overflowed


@pytest.mark.parametrize("mode", ["default", "always"])
@pytest.mark.parametrize("scale,und,seed", [(10, True, 1010), (12, False, 2012), (14, True, 1014),
                                            (15, False, 2015)])
def test_bfs_pull_rmat(tuned, mode, scale, und, seed):
    ctx = tuned
    if mode == "always":
        ctx.set_tuning(**ALWAYS)
    h = A.rmat(scale, 16, seed, undirected=und)
    g = with_csc(ctx, h, und)
    rng = np.random.RandomState(seed)
    for src in [0, h.argmax_degree(), h.m - 1] + list(rng.randint(0, h.m, 2)):
        depth, st = ctx.bfs(g, int(src))
        want, sd, e_exp, vr = O.bfs(ocsr(h), int(src))
        assert np.array_equal(depth, want), (scale, und, src, mode)
        assert (st.edges, st.vertices, st.iterations) == (e_exp, vr, sd)


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("und", [False, True])
def test_bfs_pull_small_files(tuned, name, und):
    ctx = tuned
    ctx.set_tuning(**ALWAYS)
    h = A.load_mtx(os.path.join(MTX, name + ".mtx"), undirected=und)
    g = with_csc(ctx, h, und)
    for src in range(h.m):
        depth, _ = ctx.bfs(g, src)
        assert np.array_equal(depth, O.bfs(ocsr(h), src)[0]), (name, und, src)


MODES = {"default": {}, "always": ALWAYS, "binned": BINNED, "push": PUSH}


@pytest.mark.parametrize("mode", ["default", "always", "binned", "push"])
@pytest.mark.parametrize("scale,und,seed", [(8, False, 2008), (12, False, 2012), (13, True, 3013),
                                            (15, False, 2015), (17, False, 2017)])
def test_sssp_pull_rmat(tuned, mode, scale, und, seed):
    ctx = tuned
    ctx.set_tuning(**MODES[mode])
    h = A.rmat(scale, 16, seed, undirected=und, weight_seed=seed + 1000)
    g = with_csc(ctx, h, und)
    o = ocsr(h)
    for src in [h.argmax_degree(), 0, h.m // 3]:
        dist, st = ctx.sssp(g, src)
        assert np.array_equal(dist, oracle_dist(h, src)), (scale, src, mode)
        _, it, er, sf, ri = O.sssp_bf(o, src)
        assert (st.iterations, st.edges, st.vertices, st.improved) == (it, er, sf, ri)


@pytest.mark.parametrize("mode", ["always", "binned"])
def test_sssp_pull_unweighted_is_bfs(tuned, mode):
    ctx = tuned
    ctx.set_tuning(**MODES[mode])
    h = A.rmat(12, 16, 5, undirected=True)
    g = with_csc(ctx, h, True)
    dist, _ = ctx.sssp(g, 0)
    b = O.bfs(ocsr(h), 0)[0].astype(np.float32)
    b[b < 0] = INF
    assert np.array_equal(dist, b)


def test_binned_rounds_without_csc_and_round_log(tuned):
    """Binned relaxation needs no CSC; the round log records its direction
    (every round but the source's, which always runs push)."""
    ctx = tuned
    ctx.set_tuning(**BINNED)
    h = A.rmat(14, 16, 714, undirected=False, weight_seed=715)
    g = ctx.upload(h)
    src = h.argmax_degree()
    dist, st = ctx.sssp(g, src)
    assert np.array_equal(dist, oracle_dist(h, src))
    rounds = ctx.rounds()
    assert len(rounds) == st.iterations and rounds[0][2] == 0
    assert all(p == 2 for f, e, p in rounds[1:] if e > 0)
    _, it, er, sf, ri = O.sssp_bf(ocsr(h), src)
    assert (st.iterations, st.edges, st.vertices, st.improved) == (it, er, sf, ri)


def _star_with_gaps(n_leaves, n_isolated):
    """Leaves 1..L point at hub 0 (in-degree L, spanning many CSC tiles) and at
    their successor; a run of isolated vertices follows, then a chain."""
    L, Z = n_leaves, n_isolated
    m = 1 + L + Z + 3
    rows = [[] for _ in range(m)]
    rows[0] = [1]
    for i in range(1, L + 1):
        rows[i] = [0] + ([i + 1] if i < L else [1 + L + Z])
    c0 = 1 + L + Z
    rows[c0] = [c0 + 1]
    rows[c0 + 1] = [c0 + 2]
    rp = np.zeros(m + 1, np.int32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    ci = np.array([c for r in rows for c in sorted(r)], np.int32)
    w = (np.arange(len(ci)) % 7 + 1).astype(np.float32)
    return A.HostCsr(m, m, rp, ci, w)


@pytest.mark.parametrize("leaves,isolated", [(5000, 0), (5000, 20000), (3, 9000), (2047, 1), (2048, 4096)])
@pytest.mark.parametrize("mode", ["always", "binned"])
def test_pull_tiles_with_hubs_and_empty_rows(tuned, leaves, isolated, mode):
    ctx = tuned
    ctx.set_tuning(**MODES[mode])
    h = _star_with_gaps(leaves, isolated)
    g = with_csc(ctx, h, False)
    for src in [1, 0, leaves, h.m - 3]:
        d, _ = ctx.bfs(g, src)
        assert np.array_equal(d, O.bfs(ocsr(h), src)[0]), (leaves, isolated, src)
        dist, st = ctx.sssp(g, src)
        assert np.array_equal(dist, oracle_dist(h, src)), (leaves, isolated, src)
        _, it, er, sf, ri = O.sssp_bf(ocsr(h), src)
        assert (st.iterations, st.edges, st.vertices, st.improved) == (it, er, sf, ri)


def test_binned_bucket_split_across_workgroups(tuned):
    """600,000 candidates land in the first 16,384-vertex bucket in round 2,
    so several apply workgroups share it and merge with atomicMin."""
    ctx = tuned
    ctx.set_tuning(**BINNED)
    n = 600_000
    m = n + 1  # vertex 0 has n out-edges, every other vertex one
    rp = np.zeros(m + 1, np.int64)
    rp[1] = n
    rp[2:] = n + np.arange(1, m)
    ci = np.concatenate([np.arange(1, m), np.arange(1, m) % 97 + 3]).astype(np.int32)
    w = np.concatenate([np.arange(n) % 64 + 1, np.arange(n) % 5 + 1]).astype(np.float32)
    h = A.HostCsr(m, m, rp.astype(np.int32), ci, w)
    g = ctx.upload(h)
    dist, st = ctx.sssp(g, 0)
    assert np.array_equal(dist, oracle_dist(h, 0))
    _, it, er, sf, ri = O.sssp_bf(ocsr(h), 0)
    assert (st.iterations, st.edges, st.vertices, st.improved) == (it, er, sf, ri)


def test_push_and_pull_agree_and_thresholds_round_trip(tuned):
    ctx = tuned
    h = A.rmat(14, 16, 4242, undirected=False, weight_seed=4243)
    g = with_csc(ctx, h, False)
    src = h.argmax_degree()
    ctx.set_tuning(bfs_alpha=0, sssp_pull_ratio=0)  # push only
    assert ctx.tuning()["bfs_alpha"] == 0 and ctx.tuning()["sssp_pull_ratio"] == 0
    d_push, s_push = ctx.sssp(g, src)
    b_push, t_push = ctx.bfs(g, src)
    ctx.set_tuning(**ALWAYS)
    d_pull, s_pull = ctx.sssp(g, src)
    b_pull, t_pull = ctx.bfs(g, src)
    assert np.array_equal(d_push, d_pull) and np.array_equal(b_push, b_pull)
    assert s_push.as_dict() | {"kernel_ms": 0} == s_pull.as_dict() | {"kernel_ms": 0}
    assert (t_push.edges, t_push.vertices, t_push.iterations) == (t_pull.edges, t_pull.vertices, t_pull.iterations)


def test_pull_rounds_are_timed_separately(tuned):
    """BFS with a CSC runs one launch per round (class "round": the push tiles
    or the bottom-up step, which also queue the next frontier) and no
    compaction after the seed's; SSSP with pull rounds launches the expansion
    and the pull step side by side, each timed in its own class."""
    ctx = tuned
    ctx.set_tuning(**ALWAYS)
    h = A.rmat(14, 16, 77, undirected=True, weight_seed=5)
    g = with_csc(ctx, h, True)
    ctx.set_profiling(True)
    ctx.profile_reset()
    try:
        _, st = ctx.bfs(g, 0)
        prof_bfs = ctx.profile()
        ctx.profile_reset()
        ctx.sssp(g, 0)
        prof_sssp = ctx.profile()
    finally:
        ctx.set_profiling(False)
    assert prof_bfs["round"][1] > 0 and prof_bfs["expand"][1] == 0 and prof_bfs["pull"][1] == 0
    assert prof_bfs["compact"][1] == 1  # the seed kernel
    assert prof_bfs["round"][1] >= st.iterations  # every level, and the launches past the end
    assert prof_sssp["pull"][1] > 0 and prof_sssp["expand"][1] == prof_sssp["pull"][1]
    assert prof_sssp["round"][1] == 0
