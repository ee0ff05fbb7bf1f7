"""Pins the CPU oracle before it is trusted as the checker.

Checks: the reference's own MT19937 (oracle/_ref, compiled from
src/mt19937ar.cpp) and numpy's; the golden answers for the reference's
dataset/small inputs; closed forms (K_n triangles, trace(A^3)/6, unit-weight
SSSP == BFS); brute force written independently in Python on tiny inputs;
invariants (BFS parent layers, SSSP edge tightness, MIS independence and
maximality).
"""
import itertools
import json
import os

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN, MTX

EXP = json.load(open(os.path.join(GOLDEN, "expected.json")))
SMALL = ["test_cc", "test_bc", "test_pr", "test_mesh", "test_mis", "test_spgemm"]


def mtx(name):
    return os.path.join(MTX, name + ".mtx")


def text_edges(name):
    """Independent parse of a .mtx file: (m, [(row, col)]) with the reference's
    convention (line 'a b' -> row b-1, column a-1)."""
    lines = [l.split() for l in open(mtx(name)) if not l.startswith("%") and l.strip()]
    m, n, nnz = map(int, lines[0])
    return m, [(int(b) - 1, int(a) - 1) for a, b, *_ in lines[1 : 1 + nnz]]


def py_bfs(m, adj, src):
    d = [-1] * m
    d[src] = 0
    q = [src]
    for u in q:
        for v in adj[u]:
            if d[v] < 0:
                d[v] = d[u] + 1
                q.append(v)
    return d


def adj_of(a):
    return [list(a.col_ind[a.row_ptr[i] : a.row_ptr[i + 1]]) for i in range(a.m)]


# --- MT19937 ---------------------------------------------------------------

def test_mt19937_known_answer():
    assert int(O.mt_fill(5489, 10000)[-1]) == EXP["mt19937_seed5489_10000th"]


@pytest.mark.parametrize("seed", [0, 1, 5489, 2022, 0xDEADBEEF])
def test_mt19937_matches_numpy_and_reference(seed):
    ours = O.mt_fill(seed, 20000)
    npy = np.random.RandomState(seed).randint(0, 2**32, size=20000, dtype=np.uint64).astype(np.uint32)
    assert np.array_equal(ours, npy)
    ref = O.ref_mt_fill(seed, 20000)
    if ref is not None:
        assert np.array_equal(ours, ref)


# --- loader + BFS golden -------------------------------------------------------

@pytest.mark.parametrize("name", SMALL)
@pytest.mark.parametrize("und", [False, True])
def test_bfs_golden(name, und):
    a = O.read_mtx(mtx(name), und)
    depth, sd, e_exp, vr = O.bfs(a, 0)
    want = EXP["bfs_from_0"][name]["und" if und else "dir"]
    assert list(depth) == want
    assert sd == max(want) + 1
    reached = depth >= 0
    assert vr == int(reached.sum())
    assert e_exp == int(a.degrees()[reached].sum())


def test_loader_matches_text_parse():
    for name in SMALL + ["chesapeake"]:
        m, edges = text_edges(name)
        a = O.read_mtx(mtx(name), False)
        got = sorted((r, int(c)) for r in range(a.m) for c in a.col_ind[a.row_ptr[r] : a.row_ptr[r + 1]])
        assert got == sorted(edges)
        # undirected: both directions, repeats dropped, self-loops kept once
        u = O.read_mtx(mtx(name), True)
        want = set(edges) | {(c, r) for r, c in edges}
        got = [(r, int(c)) for r in range(u.m) for c in u.col_ind[u.row_ptr[r] : u.row_ptr[r + 1]]]
        assert len(got) == len(set(got)) and set(got) == want


def test_chesapeake():
    m, edges = text_edges("chesapeake")
    adj = [[] for _ in range(m)]
    for r, c in edges:
        adj[r].append(c)
    py = py_bfs(m, adj, 0)
    a = O.read_mtx(mtx("chesapeake"), False)
    d = O.bfs(a, 0)[0]
    assert list(d) == py
    assert int((d == -1).sum()) == EXP["chesapeake"]["dir_unreached"]
    u = O.read_mtx(mtx("chesapeake"), True)
    du = O.bfs(u, 0)[0]
    assert int(du.max()) == EXP["chesapeake"]["und_max_depth"]
    assert int((du == -1).sum()) == EXP["chesapeake"]["und_unreached"]


# --- triangles ---------------------------------------------------------------

def dense(a, drop_loops=True):
    A = np.zeros((a.m, a.n), dtype=np.int64)
    for r in range(a.m):
        for c in a.col_ind[a.row_ptr[r] : a.row_ptr[r + 1]]:
            if not (drop_loops and r == c):
                A[r, c] = 1
    return A


@pytest.mark.parametrize("name", SMALL + ["chesapeake"])
def test_tc_golden_and_trace(name):
    u = O.read_mtx(mtx(name), True)
    t, w = O.tc(O.strict_lower(u))
    assert t == EXP["triangles_und"][name]
    A = dense(u)
    assert t == int(np.trace(A @ A @ A)) // 6


@pytest.mark.parametrize("n", [3, 4, 7, 12])
def test_tc_complete_graph(n):
    rp = np.arange(0, n * (n - 1) + 1, n - 1, dtype=np.int32)
    ci = np.array([j for i in range(n) for j in range(n) if j != i], dtype=np.int32)
    k = O.Csr(n, n, rp, ci)
    t, w = O.tc(O.strict_lower(k))
    assert t == n * (n - 1) * (n - 2) // 6
    # wedges: sum over (i,k) in L of |L(k,:)| = sum_i sum_{k<i} k
    assert w == sum(k for i in range(n) for k in range(i))


def test_tc_rmat_vs_trace():
    a = O.rmat(9, 8, 77, True)
    A = dense(a)
    assert O.tc(O.strict_lower(a))[0] == int(np.trace(A @ A @ A)) // 6


def test_tc_rows_split_adds_up():
    """The row-range count (the CPU baseline's bounded sample) over a split
    of the rows adds up to the whole count and wedges; a range of one row
    counts the triangles whose largest vertex is that row (pairs of its
    smaller neighbours that are adjacent), by brute force."""
    a = O.rmat(10, 8, 78, True)
    lo = O.strict_lower(a)
    t, w = O.tc(lo)
    cuts = [0, 100, 101, 517, lo.m]
    parts = [O.tc_rows(lo, x, y) for x, y in zip(cuts, cuts[1:])]
    assert sum(p[0] for p in parts) == t and sum(p[1] for p in parts) == w
    A = dense(a)
    i = int(np.argmax(np.diff(lo.row_ptr)))
    nb = lo.col_ind[lo.row_ptr[i]:lo.row_ptr[i + 1]]
    brute = sum(int(A[k, j]) for x, k in enumerate(nb) for j in nb[:x])
    assert O.tc_rows(lo, i, i + 1)[0] == brute


# --- MIS ---------------------------------------------------------------------

@pytest.mark.parametrize("name", SMALL)
def test_mis_index_golden(name):
    u = O.read_mtx(mtx(name), True)
    got = O.mis_index(u, 0)
    assert list(got) == EXP["mis_index_order_src0_und"][name]
    # priority-greedy with the source highest and index order after it
    prio = np.full(u.m, 7, dtype=np.uint32)
    prio[0] = 8
    assert list(O.mis_greedy(u, prio)) == list(got)


def test_mis_chesapeake_size():
    u = O.read_mtx(mtx("chesapeake"), True)
    assert int(O.mis_index(u, 0).sum()) == EXP["chesapeake_mis_size"]


def luby_fixed_point(adj, prio):
    """Independent restatement: repeat 'join iff key beats every undecided
    neighbour' with (prio, -v) keys until nothing is undecided."""
    m = len(adj)
    state = [-1] * m
    key = [(int(prio[v]), -v) for v in range(m)]
    while -1 in state:
        join = [v for v in range(m) if state[v] == -1 and all(
            state[u] != -1 or key[u] < key[v] for u in adj[v] if u != v)]
        assert join
        for v in join:
            state[v] = 1
        for v in join:
            for u in adj[v]:
                if state[u] == -1:
                    state[u] = 0
    return state


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_mis_greedy_equals_luby_and_is_maximal(seed):
    a = O.rmat(8, 8, 100 + seed, True)
    prio = O.mt_fill(seed, a.m)
    prio[: a.m // 4] = prio[0]  # force ties: the -v tie-break must decide
    got = O.mis_greedy(a, prio)
    adj = adj_of(a)
    assert list(got) == luby_fixed_point(adj, prio)
    for v in range(a.m):
        nb = [u for u in adj[v] if u != v]
        if got[v] == 1:
            assert all(got[u] == 0 for u in nb)
        else:
            assert any(got[u] == 1 for u in nb)
    # the reference's XOR check (src/util.cuh:277): inMIS + hasMISNeighbour == 1
    has = O.spmspv_or(a, got)
    assert np.all(got + has == 1)


@pytest.mark.parametrize("und", [False, True])
def test_spmspv_or_brute_force(und):
    a = O.rmat(8, 8, 31, und)
    rng = np.random.RandomState(5)
    for frac in (0.0, 0.02, 0.5, 1.0):
        x = (rng.rand(a.m) < frac).astype(np.int32) * rng.randint(-3, 4, a.m)
        want = np.zeros(a.m, np.int32)
        for u in range(a.m):
            if x[u] != 0:
                want[a.col_ind[a.row_ptr[u]:a.row_ptr[u + 1]]] = 1
        D = dense(a, drop_loops=False)  # y = A^T x over (OR, AND)
        assert np.array_equal(((D.T @ (x != 0).astype(np.int64)) > 0).astype(np.int32), want)
        assert np.array_equal(O.spmspv_or(a, x), want)


# --- SSSP --------------------------------------------------------------------

def py_bellman_ford(a, src):
    d = [None] * a.m
    d[src] = 0.0
    for _ in range(a.m):
        changed = False
        for u in range(a.m):
            if d[u] is None:
                continue
            for p in range(a.row_ptr[u], a.row_ptr[u + 1]):
                v, w = int(a.col_ind[p]), float(a.val[p])
                if d[v] is None or d[u] + w < d[v]:
                    d[v] = d[u] + w
                    changed = True
        if not changed:
            break
    return [-1.0 if x is None else x for x in d]


@pytest.mark.parametrize("seed", [11, 12])
def test_sssp_small_brute_force(seed):
    a = O.rmat(7, 4, seed, False, weight_seed=seed + 1)
    src = O.argmax_degree(a)
    want = py_bellman_ford(a, src)
    assert list(O.sssp_dijkstra(a, src)) == want
    d, it, er, sf, ri = O.sssp_bf(a, src)
    assert list(d) == want


def test_sssp_unit_weights_equal_bfs():
    a = O.rmat(11, 16, 5, False)
    a.val = np.ones(a.nnz, dtype=np.float32)
    src = O.argmax_degree(a)
    d = O.sssp_dijkstra(a, src)
    b = O.bfs(a, src)[0]
    assert np.array_equal(d, b.astype(np.float32))


def test_sssp_invariants_rmat():
    a = O.rmat(12, 16, 2012, False, weight_seed=3012)
    src = O.argmax_degree(a)
    d = O.sssp_dijkstra(a, src)
    bf, it, er, sf, ri = O.sssp_bf(a, src)
    assert np.array_equal(d, bf)
    assert set(np.unique(a.val)) <= set(np.arange(1, 65, dtype=np.float32))
    rows = np.repeat(np.arange(a.m), a.degrees())
    du, dv = d[rows], d[a.col_ind]
    reached = du >= 0
    assert np.all(dv[reached] >= 0)
    assert np.all(dv[reached] <= du[reached] + a.val[reached])
    # every reached vertex but the source has a tight in-edge
    tight = np.zeros(a.m, dtype=bool)
    tight[a.col_ind[reached & (dv == du + a.val)]] = True
    assert np.all(tight[(d >= 0) & (np.arange(a.m) != src)])
    assert np.all(d[d >= 0] < 2**24) and np.all(d == np.round(d))
    # schedule counts: sum_f counts the source once plus every improvement
    assert sf == ri + 1 and er >= a.degrees()[d >= 0].sum()


@pytest.mark.parametrize("seed,scale,ef", [(21, 6, 4), (22, 7, 8), (23, 7, 3)])
def test_sssp_bf_counters_by_hop_limited_paths(seed, scale, ef):
    """The schedule counters of orc_sssp_bf, from the characterisation of its
    rounds rather than its loop: after round t every distance is the lightest
    walk of at most t edges (D_t, by dense min-plus products), the frontier
    F_t is {v : D_t[v] < D_t-1[v]} (F_0 = {source}), e_relax sums the
    out-degrees of every F_t, sum_f their sizes, r_imp the improvements."""
    a = O.rmat(scale, ef, seed, False, weight_seed=seed + 7)
    src = O.argmax_degree(a)
    m = a.m
    W = np.full((m, m), np.inf)
    for u in range(m):
        for p in range(a.row_ptr[u], a.row_ptr[u + 1]):
            W[u, a.col_ind[p]] = min(W[u, a.col_ind[p]], a.val[p])
    deg = a.degrees()
    D = np.full(m, np.inf)
    D[src] = 0.0
    F = np.zeros(m, bool)
    F[src] = True
    iters = er = sf = ri = 0
    while F.any():
        iters += 1
        sf += int(F.sum())
        er += int(deg[F].sum())
        nxt = np.minimum(D, (D[:, None] + W).min(axis=0))  # walks of one more edge
        F = nxt < D
        ri += int(F.sum())
        D = nxt
    d, it, e_relax, sum_f, r_imp = O.sssp_bf(a, src)
    assert (it, e_relax, sum_f, r_imp) == (iters, er, sf, ri)
    assert np.array_equal(d, np.where(np.isinf(D), -1.0, D).astype(np.float32))


def test_sssp_step_brute_force():
    a = O.rmat(8, 8, 9, False, weight_seed=10)
    rng = np.random.RandomState(0)
    x = np.where(rng.rand(a.m) < 0.2, rng.randint(0, 50, a.m), 1.70141e38).astype(np.float32)
    x[3] = -1.0  # not in the frontier (bitify: -1 < x < 1e38)
    y = O.sssp_step(a, x)
    want = np.full(a.m, np.float32(1.70141e38))
    for u in range(a.m):
        if -1 < x[u] < 1e38:
            for p in range(a.row_ptr[u], a.row_ptr[u + 1]):
                v = a.col_ind[p]
                want[v] = min(want[v], x[u] + a.val[p])
    assert np.array_equal(y, want)


# --- SpGEMM ------------------------------------------------------------------

def test_spgemm_golden():
    a = O.read_mtx(mtx("test_spgemm"), False)
    c, p = O.spgemm(a, a)
    g = EXP["spgemm_test_spgemm_ones"]
    assert list(c.row_ptr) == g["row_ptr"] and list(c.col_ind) == g["col_ind"]
    assert list(c.val) == g["val"]


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_spgemm_vs_dense(seed):
    rng = np.random.RandomState(seed)
    m, k, n = 40, 30, 50
    A = (rng.rand(m, k) < 0.15) * rng.uniform(-1, 1, (m, k))
    B = (rng.rand(k, n) < 0.15) * rng.uniform(-1, 1, (k, n))
    A[0, :] = 0  # empty row
    def csr(D):
        rp = np.concatenate([[0], np.cumsum((D != 0).sum(1))]).astype(np.int32)
        r, c = np.nonzero(D)
        return O.Csr(D.shape[0], D.shape[1], rp, c.astype(np.int32), D[r, c].astype(np.float32))
    a, b = csr(A.astype(np.float32)), csr(B.astype(np.float32))
    c, p = O.spgemm(a, b)
    Af = np.zeros_like(A); Af[A != 0] = a.val
    Bf = np.zeros_like(B); Bf[B != 0] = b.val
    P = (A != 0).astype(int) @ (B != 0).astype(int)
    assert p == int(P.sum())
    Cd = Af.astype(np.float64) @ Bf.astype(np.float64)
    for i in range(m):
        cols = c.col_ind[c.row_ptr[i] : c.row_ptr[i + 1]]
        assert list(cols) == list(np.nonzero(P[i])[0])  # structural zeros kept
        assert np.allclose(c.val[c.row_ptr[i] : c.row_ptr[i + 1]], Cd[i, cols], rtol=1e-6, atol=1e-7)


# --- generator -----------------------------------------------------------------

def test_rmat_shape_and_symmetry():
    d = O.rmat(10, 16, 1010, False)
    u = O.rmat(10, 16, 1010, True)
    for a in (d, u):
        assert a.m == 1024 and a.row_ptr[-1] == a.nnz
        for r in range(a.m):
            row = a.col_ind[a.row_ptr[r] : a.row_ptr[r + 1]]
            assert np.all(np.diff(row) > 0) and r not in row
    S = dense(u)
    assert np.array_equal(S, S.T)
    D = dense(d)
    assert np.array_equal(S, ((D + D.T) > 0).astype(np.int64))
    # vertex 0 is the hub of an unpermuted RMAT
    assert O.argmax_degree(u) == 0


def test_rmat_quadrant_rule_brute_force():
    """Re-derive a tiny RMAT edge list directly from the MT stream."""
    scale, ef, seed = 4, 2, 321
    draws = O.mt_fill(seed, scale * ef * 16)
    edges = set()
    for e in range(ef * 16):
        s = d = 0
        for l in range(scale):
            u = draws[e * scale + l] / 2**32
            bit = 1 << (scale - 1 - l)
            if u < 0.57:
                pass
            elif u < 0.76:
                d |= bit
            elif u < 0.95:
                s |= bit
            else:
                s |= bit
                d |= bit
        if s != d:
            edges.add((s, d))
    a = O.rmat(scale, ef, seed, False)
    got = {(r, int(c)) for r in range(a.m) for c in a.col_ind[a.row_ptr[r] : a.row_ptr[r + 1]]}
    assert got == edges


# --- general-semiring SpMSpV ---------------------------------------------------

def test_mxv_semirings_pinned():
    """orc_mxv against independent computations: plus-times = dense A^T x in
    double; min-plus = the min-plus step oracle; max-times and or-and by
    brute force over the edge list; listed (sparse) presence by reach."""
    a = O.rmat(9, 8, 79, False, 80)
    rs = np.random.RandomState(6)
    x = np.where(rs.rand(a.m) < 0.3, rs.randint(1, 9, a.m), 0).astype(np.float32)
    D = np.zeros((a.m, a.n))
    rows = np.repeat(np.arange(a.m), np.diff(a.row_ptr))
    np.add.at(D, (rows, a.col_ind), a.val.astype(np.float64))
    y, r = O.mxv(a, 0, x)
    assert np.array_equal(y, (D.T @ x.astype(np.float64)).astype(np.float32))
    xs = np.where(x > 0, x, np.float32(1.70141e38)).astype(np.float32)
    assert np.array_equal(O.mxv(a, 1, xs)[0], O.sssp_step(a, xs))
    want_max = np.full(a.n, -np.inf, np.float32)
    want_or = np.zeros(a.n, np.float32)
    for u, c, w in zip(rows, a.col_ind, a.val):
        if x[u] != 0:
            want_max[c] = max(want_max[c], x[u] * w)
            want_or[c] = 1.0
    xm = np.where(x != 0, x, -np.inf).astype(np.float32)
    assert np.array_equal(O.mxv(a, 2, xm)[0], want_max)
    assert np.array_equal(O.mxv(a, 3, x)[0], want_or)
    listed = (rs.rand(a.m) < 0.05).astype(np.uint8)
    _, reach = O.mxv(a, 0, np.zeros(a.m, np.float32), listed)
    assert np.array_equal(reach.astype(bool), (D[listed.astype(bool)] != 0).any(axis=0))
