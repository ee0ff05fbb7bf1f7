"""Host-side product code against the oracle (no GPU compute): the
MatrixMarket loader, the RMAT generator, and the C ABI's exported symbols."""
import ctypes
import os

import numpy as np
import pytest

import apspie_amd as A
import oracle as O
from conftest import MTX, ROOT

NAMES = ["test_cc", "test_bc", "test_pr", "test_mesh", "test_mis", "test_spgemm", "chesapeake"]


def test_library_exports_every_declared_symbol():
    lib = A.lib()
    names = A.exported_symbols(os.path.join(ROOT, "include", "apspie.h"))
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("und", [False, True])
def test_loader_matches_oracle(name, und):
    p = os.path.join(MTX, name + ".mtx")
    got = A.load_mtx(p, undirected=und)
    want = O.read_mtx(p, und)
    assert (got.m, got.n) == (want.m, want.n)
    assert np.array_equal(got.row_ptr, want.row_ptr)
    assert np.array_equal(got.col_ind, want.col_ind)
    assert got.val is None or np.array_equal(got.val, want.val)


def test_loader_weighted_and_sorted(tmp_path):
    f = tmp_path / "w.mtx"
    f.write_text("%%MatrixMarket matrix coordinate real general\n% c\n4 4 5\n2 1 3.5\n1 2 2\n4 1 7\n3 1 1\n1 1 9\n")
    g = A.load_mtx(str(f))
    o = O.read_mtx(str(f), False)
    assert np.array_equal(g.row_ptr, o.row_ptr) and np.array_equal(g.col_ind, o.col_ind)
    assert np.array_equal(g.val, o.val)
    s = A.load_mtx(str(f), sort_columns=True)
    for r in range(s.m):
        row = s.col_ind[s.row_ptr[r]:s.row_ptr[r + 1]]
        assert np.all(np.diff(row) >= 0)
        # values travel with their columns
        for c, v in zip(row, s.val[s.row_ptr[r]:s.row_ptr[r + 1]]):
            k = list(g.col_ind[g.row_ptr[r]:g.row_ptr[r + 1]]).index(c)
            assert g.val[g.row_ptr[r] + k] == v


def test_loader_errors(tmp_path):
    with pytest.raises(A.ApspieError):
        A.load_mtx(str(tmp_path / "missing.mtx"))
    f = tmp_path / "short.mtx"
    f.write_text("3 3 4\n1 2\n2 3\n")
    with pytest.raises(A.ApspieError):
        A.load_mtx(str(f))
    f2 = tmp_path / "range.mtx"
    f2.write_text("3 3 1\n5 1\n")
    with pytest.raises(A.ApspieError):
        A.load_mtx(str(f2))


@pytest.mark.parametrize("scale,und", [(8, False), (8, True), (11, False), (12, True)])
def test_rmat_matches_oracle(scale, und):
    seed = 1000 + scale
    got = A.rmat(scale, 16, seed, undirected=und, weight_seed=seed + 7)
    want = O.rmat(scale, 16, seed, und, weight_seed=seed + 7)
    assert np.array_equal(got.row_ptr, want.row_ptr)
    assert np.array_equal(got.col_ind, want.col_ind)
    assert np.array_equal(got.val, want.val)


def test_rmat_limits():
    with pytest.raises(A.ApspieError):
        A.rmat(0, 16, 1)
    with pytest.raises(A.ApspieError):
        A.rmat(28, 16, 1)  # 2^32 edges: beyond int32 offsets


def _dense_transpose_check(h, t):
    """t must be h^T by rows with each row's entries in increasing source order."""
    assert (t.m, t.n, t.nnz) == (h.n, h.m, h.nnz)
    rows = np.repeat(np.arange(h.m, dtype=np.int64), np.diff(h.row_ptr))
    order = np.lexsort((np.arange(h.nnz), rows, h.col_ind))  # by column, then source row, then position
    assert np.array_equal(t.row_ptr, np.concatenate([[0], np.cumsum(np.bincount(h.col_ind, minlength=h.n))]))
    assert np.array_equal(t.col_ind, rows[order])
    if h.val is not None:
        assert np.array_equal(t.val, h.val[order])


@pytest.mark.parametrize("name", NAMES)
def test_transpose_small_files(name):
    h = A.load_mtx(os.path.join(MTX, name + ".mtx"))
    t = A.transpose(h)
    _dense_transpose_check(h, t)
    back = A.transpose(t)
    assert np.array_equal(back.row_ptr, h.row_ptr)
    # a row of h keeps its file order; the double transpose sorts it
    for i in range(h.m):
        a, b = h.row_ptr[i], h.row_ptr[i + 1]
        assert np.array_equal(back.col_ind[a:b], np.sort(h.col_ind[a:b], kind="stable"))


def test_transpose_rmat_weighted_and_rectangular():
    h = A.rmat(13, 16, 31, weight_seed=32)
    _dense_transpose_check(h, A.transpose(h))
    p = A.powerlaw(3000, 500, 6.0, 200, 9)
    _dense_transpose_check(p, A.transpose(p))
    u = A.rmat(11, 8, 33, undirected=True)
    t = A.transpose(u)  # symmetric: the transpose is the matrix itself
    assert np.array_equal(t.row_ptr, u.row_ptr) and np.array_equal(t.col_ind, u.col_ind)


def test_transpose_duplicates_and_bad_columns():
    h = A.HostCsr(3, 4, np.array([0, 3, 3, 5], np.int32), np.array([2, 2, 0, 3, 2], np.int32),
                  np.array([1, 2, 3, 4, 5], np.float32))
    t = A.transpose(h)
    _dense_transpose_check(h, t)
    assert list(t.col_ind[t.row_ptr[2]:t.row_ptr[3]]) == [0, 0, 2]
    assert list(t.val[t.row_ptr[2]:t.row_ptr[3]]) == [1, 2, 5]
    bad = A.HostCsr(2, 2, np.array([0, 1, 2], np.int32), np.array([0, 2], np.int32))
    with pytest.raises(A.ApspieError):
        A.transpose(bad)


BANNER = os.path.join(ROOT, "tests", "golden", "mtx_banner")


def dense_of(h):
    d = np.zeros((h.m, h.n), np.float32)
    for r in range(h.m):
        for p in range(h.row_ptr[r], h.row_ptr[r + 1]):
            d[r, h.col_ind[p]] += 1.0 if h.val is None else h.val[p]
    return d


def test_banner_symmetric_and_skew():
    """Banner symmetry mirrors the stored triangle (skew: negated); entries
    keep readMtx's reading (line "a b" is column a-1 of row b-1)."""
    h = A.load_mtx(os.path.join(BANNER, "sym_real.mtx"), banner=True)
    want = np.array([[2.5, 1.5, -2, 0], [1.5, 0, 0, 0.25], [-2, 0, 0, 3], [0, 0.25, 3, 0]], np.float32)
    assert np.array_equal(dense_of(h), want) and h.nnz == 9
    k = A.load_mtx(os.path.join(BANNER, "skew_int.mtx"), banner=True)
    assert np.array_equal(dense_of(k), np.array([[0, 4, 0], [-4, 0, -7], [0, 7, 0]], np.float32))
    # without the flag: the reference's reading, no mirroring
    plain = A.load_mtx(os.path.join(BANNER, "sym_real.mtx"))
    assert plain.nnz == 5 and plain.val is not None


def test_banner_field():
    p = A.load_mtx(os.path.join(BANNER, "pattern_general.mtx"), banner=True)
    assert p.val is None and np.array_equal(dense_of(p), np.array([[0, 0, 1], [1, 0, 0], [0, 1, 0]], np.float32))
    # a tab after the first pair: the banner reads the values; the reference's
    # rule (values iff a space follows) reads none and then trips on them
    t = A.load_mtx(os.path.join(BANNER, "real_general_tab.mtx"), banner=True)
    assert np.array_equal(dense_of(t), np.array([[0, 0, -4], [0.5, 0, 0], [0, 1.25, 0]], np.float32))
    with pytest.raises(A.ApspieError):
        A.load_mtx(os.path.join(BANNER, "real_general_tab.mtx"))
    with pytest.raises(A.ApspieError):
        A.load_mtx(os.path.join(BANNER, "array.mtx"), banner=True)
    with pytest.raises(A.ApspieError):  # no banner line
        A.load_mtx(os.path.join(MTX, "test_cc.mtx"), banner=True)


@pytest.mark.parametrize("weighted", [False, True])
def test_csr_cache_round_trip(tmp_path, weighted):
    h = A.rmat(10, 8, 91, undirected=False, weight_seed=92 if weighted else None)
    f = str(tmp_path / "g.csr")
    A.save_csr(f, h)
    g = A.load_csr(f)
    assert (g.m, g.n) == (h.m, h.n)
    assert np.array_equal(g.row_ptr, h.row_ptr) and np.array_equal(g.col_ind, h.col_ind)
    assert (g.val is None) == (not weighted)
    if weighted:
        assert np.array_equal(g.val, h.val)
    e = A.HostCsr(5, 3, np.zeros(6, np.int32), np.zeros(0, np.int32))
    A.save_csr(f, e)
    z = A.load_csr(f)
    assert z.m == 5 and z.n == 3 and z.nnz == 0


def test_csr_cache_rejects_damage(tmp_path):
    h = A.rmat(9, 8, 93, weight_seed=94)
    f = str(tmp_path / "g.csr")
    A.save_csr(f, h)
    raw = bytearray(open(f, "rb").read())
    for pos in (3, 40, len(raw) // 2, len(raw) - 3):  # magic, header, payload, check word
        bad = bytearray(raw)
        bad[pos] ^= 0x10
        open(f, "wb").write(bytes(bad))
        with pytest.raises(A.ApspieError):
            A.load_csr(f)
    open(f, "wb").write(bytes(raw[:-20]))  # truncated
    with pytest.raises(A.ApspieError):
        A.load_csr(f)
    with pytest.raises(A.ApspieError):
        A.load_csr(str(tmp_path / "missing.csr"))


def _read_coo(path):
    """The .mtx entries as (col, row, val) the way readMtx stores them (a - 1
    into the column array, b - 1 into the row array), parsed independently."""
    lines = [ln for ln in open(path).read().splitlines() if ln.strip() and not ln.startswith("%")]
    nnz = int(lines[0].split()[2])
    ent = [ln.split() for ln in lines[1:1 + nnz]]
    col = np.array([int(e[0]) - 1 for e in ent], np.int32)
    row = np.array([int(e[1]) - 1 for e in ent], np.int32)
    val = np.array([float(e[2]) if len(e) > 2 else 1.0 for e in ent], np.float32)
    return col, row, val


@pytest.mark.parametrize("name", NAMES)
def test_make_symmetric_and_build_match_the_loader(name):
    """makeSymmetric + buildMatrix (the reference's host calls, apspie_host_coo_*)
    give the CSR the one-call loader and the oracle give."""
    p = os.path.join(MTX, name + ".mtx")
    col, row, val = _read_coo(p)
    m = int([ln for ln in open(p) if not ln.startswith("%")][0].split()[0])
    c2, r2, v2 = A.coo_make_symmetric(col, row, val)
    # symmetric, no repeats, self-loops kept, sorted by (row, col)
    pairs = list(zip(r2.tolist(), c2.tolist()))
    assert pairs == sorted(set(pairs))
    assert set(pairs) == {(int(r), int(c)) for r, c in zip(row, col)} | {(int(c), int(r)) for r, c in zip(row, col)}
    h = A.coo_to_csr_host(m, r2, c2, v2)
    want = O.read_mtx(p, True)
    assert np.array_equal(h.row_ptr, want.row_ptr) and np.array_equal(h.col_ind, want.col_ind)
    # directed: a stable counting sort of the file order by row
    hd = A.coo_to_csr_host(m, row, col, val)
    wd = O.read_mtx(p, False)
    assert np.array_equal(hd.row_ptr, wd.row_ptr) and np.array_equal(hd.col_ind, wd.col_ind)
    order = np.argsort(row, kind="stable")
    assert np.array_equal(hd.col_ind, col[order]) and np.array_equal(hd.val, val[order])


def test_make_symmetric_keeps_values_with_entries():
    col = np.array([1, 0, 2, 2], np.int32)
    row = np.array([0, 1, 2, 0], np.int32)
    val = np.array([5.0, 5.0, 7.0, 3.0], np.float32)
    c, r, v = A.coo_make_symmetric(col, row, val)
    assert list(zip(r.tolist(), c.tolist(), v.tolist())) == [(0, 1, 5.0), (0, 2, 3.0), (1, 0, 5.0), (2, 0, 3.0),
                                                            (2, 2, 7.0)]
    with pytest.raises(A.ApspieError):  # a row outside [0, m)
        A.coo_to_csr_host(2, np.array([0, 2], np.int32), np.array([1, 1], np.int32))


def test_server_submit_checks_the_output_array():
    """The row is written from the server's threads after submit returns, so
    a short, mistyped or non-contiguous array is refused before it is queued."""
    srv = A.Server.__new__(A.Server)
    srv._graphs = [type("G", (), {"m": 10})()]
    srv._held = {}
    srv.h = None
    with pytest.raises(TypeError):
        srv.submit(A.REQ_BFS, 0, 0, np.empty(10, np.float32))
    with pytest.raises(TypeError):
        srv.submit(A.REQ_SSSP, 0, 0, np.empty(10, np.int32))
    with pytest.raises(ValueError):
        srv.submit(A.REQ_BFS, 0, 0, np.empty(9, np.int32))
    with pytest.raises(ValueError):
        srv.submit(A.REQ_BFS, 0, 0, np.empty(20, np.int32)[::2])
    with pytest.raises(ValueError):
        srv.submit(A.REQ_BFS, 1, 0, np.empty(10, np.int32))
    assert srv._held == {}
