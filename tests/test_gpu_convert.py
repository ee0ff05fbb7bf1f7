"""GPU parity of the device format conversions: apspie_csr2csc equals the
host transpose (rows ascending in each column, values moved with them) on
square, rectangular, duplicate-holding, empty and multi-pass (n > 2^16)
inputs; a CSC built on the device drives the pull rounds exactly; and
apspie_coo_to_csr equals the row offsets of sorted COO rows, refusing unsorted
or out-of-range rows."""
import numpy as np
import pytest

import apspie_amd as A
import oracle as O

pytestmark = pytest.mark.gpu
INF = np.float32(1.70141e38)


def vals(h):
    return np.zeros(0, np.float32) if h.val is None and h.nnz == 0 else h.val


def same(a, b):
    va, vb = vals(a), vals(b)
    return (a.m, a.n) == (b.m, b.n) and np.array_equal(a.row_ptr, b.row_ptr) and \
        np.array_equal(a.col_ind, b.col_ind) and \
        ((va is None and vb is None) or (va is not None and vb is not None and np.array_equal(va, vb)))


@pytest.mark.parametrize("case", ["rmat12w", "rmat18", "rect", "dups", "empty_rows", "no_edges"])
def test_csr2csc_equals_host_transpose(ctx, case):
    if case == "rmat12w":
        h = A.rmat(12, 16, 6012, weight_seed=7012)
    elif case == "rmat18":
        h = A.rmat(18, 8, 6018, weight_seed=7018)  # 18-bit columns: three radix passes
    elif case == "rect":
        h = A.powerlaw(3000, 70000, 6.0, 300, 6013)
    elif case == "dups":
        rp = np.array([0, 3, 3, 6], np.int32)
        h = A.HostCsr(3, 4, rp, np.array([1, 1, 3, 0, 3, 3], np.int32), np.arange(6, dtype=np.float32))
    elif case == "empty_rows":
        rp = np.array([0, 0, 2, 2, 2, 4, 4], np.int32)
        h = A.HostCsr(6, 6, rp, np.array([5, 0, 0, 5], np.int32), np.array([1, 2, 3, 4], np.float32))
    else:
        h = A.HostCsr(4, 7, np.zeros(5, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32))
    g = ctx.upload(h)
    got = ctx.csr2csc(g)
    assert same(got, A.transpose(h)), case
    if h.nnz:
        s = ctx.csr2csc(g, values=False)
        assert s.val is None and np.array_equal(s.col_ind, got.col_ind)


def test_device_csc_drives_pull_rounds(ctx):
    saved = ctx.tuning()
    try:
        ctx.set_tuning(bfs_alpha=1 << 20, bfs_beta=1 << 20, sssp_pull_ratio=1 << 20, sssp_bin_ratio=0)
        h = A.rmat(13, 16, 6113, weight_seed=7113)
        g = ctx.upload(h, csc="device")
        assert g.has_csc
        o = O.Csr(h.m, h.n, h.row_ptr, h.col_ind, h.val)
        src = h.argmax_degree()
        d, _ = ctx.bfs(g, src)
        assert np.array_equal(d, O.bfs(o, src)[0])
        dist, _ = ctx.sssp(g, src)
        want = O.sssp_dijkstra(o, src)
        want[want == -1.0] = INF
        assert np.array_equal(dist, want)
        assert any(p for _, _, p in ctx.rounds())
    finally:
        ctx.set_tuning(**saved)


def test_coo_to_csr(ctx):
    rows = np.array([1, 1, 1, 4, 4, 6, 9, 9], np.int32)
    m = 11
    want = np.searchsorted(rows, np.arange(m + 1), side="left").astype(np.int32)
    assert np.array_equal(ctx.coo_to_csr(rows, m), want)
    assert np.array_equal(ctx.coo_to_csr(np.zeros(0, np.int32), 5), np.zeros(6, np.int32))
    big = np.sort(np.random.RandomState(5).randint(0, 50000, 300000)).astype(np.int32)
    assert np.array_equal(ctx.coo_to_csr(big, 50000),
                          np.searchsorted(big, np.arange(50001), side="left").astype(np.int32))
    with pytest.raises(A.ApspieError):
        ctx.coo_to_csr(np.array([0, 2, 1], np.int32), 3)
    with pytest.raises(A.ApspieError):
        ctx.coo_to_csr(np.array([0, 3], np.int32), 3)
