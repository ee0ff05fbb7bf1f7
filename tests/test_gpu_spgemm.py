"""GPU parity of apspie_spgemm against the oracle's Gustavson product: the
structure (row_ptr, sorted col_ind, structural zeros kept) must be identical
and values within 1e-5 relative.  Covers every row bin (LDS tables and the
global-memory fallback), skewed columns (RMAT), rectangular shapes, empty
rows, and the full 1M x 1M power-law configuration."""
import json
import os

import numpy as np
import pytest

import apspie_amd as A
import oracle as O
from conftest import GOLDEN, MTX

pytestmark = pytest.mark.gpu

EXP = json.load(open(os.path.join(GOLDEN, "expected.json")))
RTOL = 1e-5


def ocsr(h):
    return O.Csr(h.m, h.n, h.row_ptr, h.col_ind, h.val)


def check(got, want):
    assert (got.m, got.n) == (want.m, want.n)
    assert np.array_equal(got.row_ptr, want.row_ptr)
    assert np.array_equal(got.col_ind, want.col_ind)
    err = np.abs(got.val.astype(np.float64) - want.val.astype(np.float64))
    assert np.all(err <= RTOL * np.abs(want.val.astype(np.float64))), float(err.max())


def run(ctx, ha, hb):
    c, st = ctx.spgemm(ctx.upload(ha), ctx.upload(hb))
    want, prods = O.spgemm(ocsr(ha), ocsr(hb))
    check(c, want)
    assert st.edges == prods and st.vertices == want.nnz
    return c, st


def test_spgemm_golden(ctx):
    h = A.load_mtx(os.path.join(MTX, "test_spgemm.mtx"))
    c, _ = ctx.spgemm(ctx.upload(h), ctx.upload(h))
    g = EXP["spgemm_test_spgemm_ones"]
    assert list(c.row_ptr) == g["row_ptr"] and list(c.col_ind) == g["col_ind"]
    assert list(c.val) == g["val"]


@pytest.mark.parametrize("m,maxlen,seed", [(2000, 64, 1), (30000, 4096, 2), (60000, 2048, 3)])
def test_spgemm_powerlaw_square(ctx, m, maxlen, seed):
    h = A.powerlaw(m, m, 8.0, maxlen, seed)
    run(ctx, h, h)


def test_spgemm_rectangular_and_empty_rows(ctx):
    a = A.powerlaw(3000, 2000, 6.0, 500, 11)
    b = A.powerlaw(2000, 5000, 5.0, 1000, 12)
    # empty some rows of A and of B
    for h, rows in ((a, [0, 5, 2999]), (b, [1, 7, 1999])):
        keep = np.ones(h.nnz, bool)
        for r in rows:
            keep[h.row_ptr[r]:h.row_ptr[r + 1]] = False
        cnt = np.diff(h.row_ptr).copy()
        cnt[rows] = 0
        h.col_ind, h.val = h.col_ind[keep], h.val[keep]
        h.row_ptr = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int32)
    run(ctx, a, b)


def test_spgemm_rmat_skewed_columns(ctx):
    h = A.rmat(12, 16, 1212, undirected=True)
    h.val = np.random.RandomState(0).uniform(-1, 1, h.nnz).astype(np.float32)
    run(ctx, h, h)


def test_spgemm_cancellation_keeps_structure(ctx):
    # row 0 of C = 1*B(0,:) + (-1)*B(1,:): values cancel to 0 but stay stored
    a = A.HostCsr(1, 2, np.array([0, 2], np.int32), np.array([0, 1], np.int32), np.array([1, -1], np.float32))
    b = A.HostCsr(2, 3, np.array([0, 2, 4], np.int32), np.array([0, 2, 0, 2], np.int32),
                  np.array([0.5, 2, 0.5, 2], np.float32))
    c, _ = ctx.spgemm(ctx.upload(a), ctx.upload(b))
    assert list(c.row_ptr) == [0, 2] and list(c.col_ind) == [0, 2] and list(c.val) == [0.0, 0.0]


def test_spgemm_full_config(ctx):
    """1M x 1M, power-law rows (mean 8, max 4096), uniform columns, C = A * A."""
    h = A.powerlaw(1 << 20, 1 << 20, 8.0, 4096, 3003)
    run(ctx, h, h)


def long_row_products(ha, hb):
    """Products of each row of A * B (the bound the row bins use)."""
    deg = np.diff(hb.row_ptr).astype(np.int64)
    per = deg[ha.col_ind]
    return np.add.reduceat(per, ha.row_ptr[:-1]) * (np.diff(ha.row_ptr) > 0)


@pytest.mark.parametrize("m,maxlen,seed", [(30000, 4096, 2), (200000, 4096, 7)])
def test_spgemm_long_rows_both_ways(ctx, m, maxlen, seed):
    """Rows beyond the LDS tables (over 2048 products) by expand-sort-compress
    (default) and by global hash tables (spgemm_long_rows=0), the other rows in
    one pass through a staging area (default) or in a counting and a product
    pass (spgemm_passes=2): every combination matches the oracle."""
    h = A.powerlaw(m, m, 8.0, maxlen, seed)
    assert (long_row_products(h, h) > 2048).sum() > 10  # the long-row bin is exercised
    saved = ctx.tuning()
    try:
        got = {}
        for long_rows in (0, 1):
            for passes in (1, 2):  # the LDS-table rows in one pass (staging) or two
                ctx.set_tuning(spgemm_long_rows=long_rows, spgemm_passes=passes)
                got[long_rows, passes], _ = run(ctx, h, h)
    finally:
        ctx.set_tuning(**saved)
    ref = got[1, 1]
    for c in got.values():
        assert np.array_equal(c.row_ptr, ref.row_ptr) and np.array_equal(c.col_ind, ref.col_ind)


@pytest.mark.parametrize("esc_min", [256, 512, 2048])
def test_spgemm_esc_threshold_moves_rows_off_the_lds_tables(ctx, esc_min):
    """spgemm_esc_min moves the LDS-table bins above it to the
    expand-sort-compress rows (and out of the staging area): the product
    matches the oracle at every threshold, and equals the default split
    (1024) bit for bit."""
    h = A.powerlaw(30000, 30000, 8.0, 4096, 5)
    lp = long_row_products(h, h)
    lo, hi = min(esc_min, 1024), max(esc_min, 1024)
    assert ((lp > lo) & (lp <= hi)).sum() > 10  # rows that change side
    assert ctx.tuning()["spgemm_esc_min"] == 1024
    saved = ctx.tuning()
    try:
        base, _ = run(ctx, h, h)
        ctx.set_tuning(spgemm_esc_min=esc_min)
        got, _ = run(ctx, h, h)
    finally:
        ctx.set_tuning(**saved)
    assert np.array_equal(got.row_ptr, base.row_ptr) and np.array_equal(got.col_ind, base.col_ind)
    assert np.array_equal(got.val, base.val)


def test_spgemm_long_rows_with_columns_outside_b(ctx):
    """A long row whose B rows hold columns outside B's range: those products
    are dropped by both paths (validate_csr does not check column ranges)."""
    h = A.powerlaw(3000, 3000, 40.0, 2000, 9)
    assert (long_row_products(h, h) > 2048).any()
    bad = h.col_ind.copy()
    bad[::97] = 3000 + (np.arange(bad[::97].size) % 5)  # columns n .. n+4
    hb = A.HostCsr(h.m, h.n, h.row_ptr, bad, h.val)
    saved = ctx.tuning()
    try:
        out = {}
        for mode in (0, 1):
            ctx.set_tuning(spgemm_long_rows=mode)
            out[mode], _ = ctx.spgemm(ctx.upload(h), ctx.upload(hb))
    finally:
        ctx.set_tuning(**saved)
    assert np.array_equal(out[0].row_ptr, out[1].row_ptr) and np.array_equal(out[0].col_ind, out[1].col_ind)
    assert np.allclose(out[0].val, out[1].val, rtol=1e-5, atol=0)
    assert out[1].col_ind.size == 0 or out[1].col_ind.max() < 3000


def _rows_of(rng, lens, draw):
    """CSR arrays of rows with len[r] distinct columns each, drawn by draw(r, l)."""
    rp = np.zeros(len(lens) + 1, np.int32)
    rp[1:] = np.cumsum(lens)
    col = np.concatenate([np.sort(draw(r, l)) for r, l in enumerate(lens)]).astype(np.int32)
    val = rng.uniform(-2.0, 2.0, col.size).astype(np.float32)
    return rp, col, val


def test_spgemm_register_bins_exact(ctx):
    """Rows of at most 64 products (the register bins: four rows a wave up to
    16 products, a wave per row up to 64) at every sorting-stage boundary
    (1, 2, 15, 16, 17, 31, 32, 33, 63, 64 products), with duplicate columns
    (runs of several terms), columns outside B (dropped), and A rows of over
    64 entries most of whose B rows are empty (several chunks per row).
    These rows sum each entry in FP64 in k order, as the oracle does, so C is
    identical to the oracle's, values included."""
    rng = np.random.default_rng(23)
    n, mb = 40, 3000
    blen = rng.choice([0, 0, 1, 2, 3, 4], mb)
    brp = np.zeros(mb + 1, np.int32)
    brp[1:] = np.cumsum(blen)
    bcol = np.concatenate([np.sort(rng.choice(n + 3, l, replace=False)) for l in blen]).astype(np.int32)
    bval = rng.uniform(-2, 2, bcol.size).astype(np.float32)
    hb = A.HostCsr(mb, n, brp, bcol, bval)  # columns n .. n+2 lie outside B
    full = np.nonzero(blen > 0)[0]
    empty = np.nonzero(blen == 0)[0]
    rows = []
    for target in [1, 2, 15, 16, 17, 31, 32, 33, 63, 64] * 6:
        ks, tot = [], 0
        for k in rng.permutation(full):
            if tot + blen[k] <= target:
                ks.append(k)
                tot += blen[k]
            if tot == target:
                break
        if rng.random() < 0.3:  # pad past 64 entries with empty B rows
            ks += list(rng.choice(empty, 70, replace=False))
        rows.append(np.sort(np.array(ks, np.int32)))
    arp = np.zeros(len(rows) + 1, np.int32)
    arp[1:] = np.cumsum([r.size for r in rows])
    acol = np.concatenate(rows).astype(np.int32)
    ha = A.HostCsr(len(rows), mb, arp, acol, rng.uniform(-2, 2, acol.size).astype(np.float32))
    ub = long_row_products(ha, hb)
    assert ub.max() <= 64 and {16, 17, 64} <= set(ub.tolist())
    # the oracle takes columns inside B only: its B drops the others, as
    # the product does
    keep = bcol < n
    krp = np.zeros(mb + 1, np.int32)
    krp[1:] = np.cumsum(np.bincount(np.repeat(np.arange(mb), blen)[keep], minlength=mb))
    want, _ = O.spgemm(ocsr(ha), O.Csr(mb, n, krp, bcol[keep], bval[keep]))
    c, st = ctx.spgemm(ctx.upload(ha), ctx.upload(hb))
    assert st.edges == ub.sum()  # products counted before the columns outside B drop
    assert np.array_equal(c.row_ptr, want.row_ptr) and np.array_equal(c.col_ind, want.col_ind)
    assert np.array_equal(c.val, want.val)


@pytest.mark.parametrize("esc_min,passes", [(1024, 1), (2048, 1), (2048, 2)])
def test_spgemm_sorted_bins_exact(ctx, esc_min, passes):
    """Rows of 128 to 2048 products (LDS hash tables up to 256, then sorted
    in LDS, a workgroup a row) at every bin boundary (128/129, 256/257,
    512/513, 1024/1025, 2048), with
    runs of equal columns (a narrow column range), columns outside B
    (dropped), and A rows of several hundred entries most of whose B rows are
    empty (several expansion chunks).  With spgemm_esc_min 2048 the rows of
    1025 to 2048 products are sorted too; two passes count before they write.
    Every run is summed in FP64 in k order, as the oracle does, so C is
    identical to the oracle's, values included."""
    rng = np.random.default_rng(29)
    n, mb = 700, 6000
    blen = rng.choice([0, 0, 0, 1, 3, 6, 9, 12], mb)
    brp = np.zeros(mb + 1, np.int32)
    brp[1:] = np.cumsum(blen)
    bcol = np.concatenate([np.sort(rng.choice(n + 3, l, replace=False)) for l in blen]).astype(np.int32)
    bval = rng.uniform(-2, 2, bcol.size).astype(np.float32)
    hb = A.HostCsr(mb, n, brp, bcol, bval)  # columns n .. n+2 lie outside B
    full = np.nonzero(blen > 0)[0]
    empty = np.nonzero(blen == 0)[0]
    rows = []
    for target in [128, 129, 200, 256, 257, 400, 512, 513, 777, 1024, 1025, 1500, 2047, 2048] * 3:
        ks, tot = [], 0
        for k in rng.permutation(full):
            if tot + blen[k] <= target:
                ks.append(k)
                tot += blen[k]
            if tot == target:
                break
        assert tot == target
        if rng.random() < 0.5:  # empty B rows: more A entries than a chunk
            ks += list(rng.choice(empty, 300, replace=False))
        rows.append(np.sort(np.array(ks, np.int32)))
    arp = np.zeros(len(rows) + 1, np.int32)
    arp[1:] = np.cumsum([r.size for r in rows])
    acol = np.concatenate(rows).astype(np.int32)
    ha = A.HostCsr(len(rows), mb, arp, acol, rng.uniform(-2, 2, acol.size).astype(np.float32))
    ub = long_row_products(ha, hb)
    assert ub.min() == 128 and ub.max() == 2048
    keep = bcol < n
    krp = np.zeros(mb + 1, np.int32)
    krp[1:] = np.cumsum(np.bincount(np.repeat(np.arange(mb), blen)[keep], minlength=mb))
    want, _ = O.spgemm(ocsr(ha), O.Csr(mb, n, krp, bcol[keep], bval[keep]))
    assert (np.diff(want.row_ptr) < ub).all()  # every row has runs
    saved = ctx.tuning()
    try:
        ctx.set_tuning(spgemm_esc_min=esc_min, spgemm_passes=passes)
        c, st = ctx.spgemm(ctx.upload(ha), ctx.upload(hb))
    finally:
        ctx.set_tuning(**saved)
    assert st.edges == ub.sum()
    assert np.array_equal(c.row_ptr, want.row_ptr) and np.array_equal(c.col_ind, want.col_ind)
    assert np.array_equal(c.val, want.val)


def skewed_long_rows():
    """A * B whose rows are all over 2048 products: rows over uniform B rows
    (up to 21000 products), rows whose B rows crowd a band of 300 low columns,
    and two rows with 11500 and 12000 products on column 7 alone."""
    rng = np.random.default_rng(17)
    n, mb = 60000, 20000
    uni = lambda k: rng.choice(np.arange(300, n), k, replace=False)

    def b_row(r, l):
        if r < 4000:  # uniform
            return rng.choice(n, l, replace=False)
        if r < 8000:  # 20 in the band [0, 300), 20 beyond it
            return np.concatenate([rng.choice(300, 20, replace=False), uni(20)])
        return np.concatenate([[7], uni(4)])  # column 7 and 4 beyond the band
    blens = np.where(np.arange(mb) < 4000, 30, np.where(np.arange(mb) < 8000, 40, 5))
    brp, bcol, bval = _rows_of(rng, blens, b_row)
    hb = A.HostCsr(mb, n, brp, bcol, bval)
    alens = np.concatenate([rng.integers(80, 700, 20), rng.integers(200, 600, 10), [11500, 12000]])

    def a_row(r, l):
        if r < 20:
            return rng.choice(4000, l, replace=False)
        if r < 30:
            return 4000 + rng.choice(4000, l, replace=False)
        return 8000 + rng.choice(12000, l, replace=False)
    arp, acol, aval = _rows_of(rng, alens, a_row)
    ha = A.HostCsr(len(alens), mb, arp, acol, aval)
    assert long_row_products(ha, hb).min() > 2048
    return ha, hb


def test_spgemm_long_rows_skewed_exact(ctx):
    """Every row of C from the long-row paths, over skewed columns (a crowded
    band, one column with 12000 products in a row).  Expand-sort-compress
    (default) sums each entry in FP64 in k order, as the oracle does, so its
    C is identical, values included; the global hash tables match within
    1e-5."""
    ha, hb = skewed_long_rows()
    want, prods = O.spgemm(ocsr(ha), ocsr(hb))
    saved = ctx.tuning()
    try:
        got = {}
        for mode in (1, 0):
            ctx.set_tuning(spgemm_long_rows=mode)
            got[mode], st = ctx.spgemm(ctx.upload(ha), ctx.upload(hb))
            assert st.edges == prods and st.vertices == want.nnz
    finally:
        ctx.set_tuning(**saved)
    c = got[1]
    assert np.array_equal(c.row_ptr, want.row_ptr) and np.array_equal(c.col_ind, want.col_ind)
    assert np.array_equal(c.val, want.val)
    check(got[0], want)


def test_spgemm_long_rows_windowed_runs_exact(ctx):
    """Long rows that the windowed LDS sort takes whole (every group of one
    top-digit range stays under a window): rows of about 18000 products over
    several windows, and runs of about 12 equal columns from a set of 512
    spread over the whole column range, summed in FP64 in k order.  C is the
    oracle's, values included."""
    rng = np.random.default_rng(23)
    n, mb = 1 << 20, 50000
    common = np.sort(rng.choice(n, 512, replace=False))

    def b_row(r, l):
        cols = set(rng.choice(common, 4, replace=False).tolist())
        while len(cols) < l:
            cols.add(int(rng.integers(n)))
        return np.array(sorted(cols))
    brp, bcol, bval = _rows_of(rng, np.full(mb, 12), b_row)
    hb = A.HostCsr(mb, n, brp, bcol, bval)
    alens = np.concatenate([np.full(24, 1500), rng.integers(200, 400, 8)])
    arp, acol, aval = _rows_of(rng, alens, lambda r, l: rng.choice(mb, l, replace=False))
    ha = A.HostCsr(len(alens), mb, arp, acol, aval)
    prods = long_row_products(ha, hb)
    assert prods.max() > 3 * 2048 and prods.min() > 1024
    c, _ = run(ctx, ha, hb)
    want, _ = O.spgemm(ocsr(ha), ocsr(hb))
    assert np.array_equal(c.val, want.val)


@pytest.mark.skip(reason="the LDS-sorted long-row path (a workgroup per long row, column parts radix-sorted "
                  "in LDS, spgemm_long_rows=2) faulted on its first GPU run; neither reading nor a CPU "
                  "emulation of it showed a cause, so the path is not in the library")
def test_spgemm_long_rows_sorted_in_lds_exact(ctx):
    """The LDS-sorted long rows: C identical to the oracle's, values included,
    on the skewed rows (column parts, a band halved until it fits LDS, and a
    column with more products than LDS holds)."""
    ha, hb = skewed_long_rows()
    want, _ = O.spgemm(ocsr(ha), ocsr(hb))
    saved = ctx.tuning()
    try:
        ctx.set_tuning(spgemm_long_rows=2)
        c, _ = ctx.spgemm(ctx.upload(ha), ctx.upload(hb))
    finally:
        ctx.set_tuning(**saved)
    assert np.array_equal(c.row_ptr, want.row_ptr) and np.array_equal(c.col_ind, want.col_ind)
    assert np.array_equal(c.val, want.val)

def test_rowptr_scan_refuses_a_total_a_csr_cannot_hold(ctx):
    """C's row pointers come from one 64-bit-exact scan of the row counts; a
    total at or past 2^31 - 4096 is refused (apspie_spgemm checks it before
    allocating C) instead of wrapping."""
    rng = np.random.RandomState(11)
    counts = rng.randint(0, 1000, 100_000).astype(np.int32)
    rp, tot = ctx.rowptr_from_counts(counts)
    assert tot == int(counts.sum())
    assert np.array_equal(rp, np.concatenate([[0], np.cumsum(counts)]).astype(np.int32))
    rp0, tot0 = ctx.rowptr_from_counts(np.zeros(0, np.int32))
    assert tot0 == 0 and list(rp0) == [0]
    big = np.array([2 ** 30, 2 ** 30 - 4096, 0, 5], np.int32)  # sums to 2^31 - 4091
    with pytest.raises(A.ApspieError) as e:
        ctx.rowptr_from_counts(big)
    assert e.value.code == -5
    edge = np.array([2 ** 30, 2 ** 30 - 4097], np.int32)  # one below the limit: accepted
    rp2, tot2 = ctx.rowptr_from_counts(edge)
    assert tot2 == 2 ** 31 - 4097 and rp2[-1] == 2 ** 31 - 4097
