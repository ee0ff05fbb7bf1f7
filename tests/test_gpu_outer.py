"""GPU parity of apspie_spgemm_outer, the reference's strip-pair outer-product
SpGEMM (C = A * B^T), against the oracle's Gustavson product of A and B^T.
Both sum every output entry in FP64 over k ascending, so the comparison is
exact: identical structure (sorted columns, structural zeros kept) and
identical FP32 values.  Covers part sizes from 1 row to one strip, rectangular
and unweighted operands, empty and disjoint operands, the reference's own
test_spgemm.mtx, the statistics, the argument checks, and agreement with the
row-wise apspie_spgemm on a 1M x 1M power-law product."""
import json
import os

import numpy as np
import pytest

import apspie_amd as A
import oracle as O
from conftest import GOLDEN, MTX

pytestmark = pytest.mark.gpu

EXP = json.load(open(os.path.join(GOLDEN, "expected.json")))


def ocsr(h):
    return O.Csr(h.m, h.n, h.row_ptr, h.col_ind, h.val)


def transposed(h):
    """h^T by rows (numpy, stable in the source row), as a HostCsr."""
    rows = np.repeat(np.arange(h.m, dtype=np.int32), np.diff(h.row_ptr))
    order = np.argsort(h.col_ind, kind="stable")
    rp = np.zeros(h.n + 1, np.int64)
    np.add.at(rp, h.col_ind.astype(np.int64) + 1, 1)
    rp = np.cumsum(rp).astype(np.int32)
    val = None if h.val is None else np.ascontiguousarray(h.val[order])
    return A.HostCsr(h.n, h.m, rp, np.ascontiguousarray(rows[order]), val)


def oracle_abt(ha, hb):
    return O.spgemm(ocsr(ha), ocsr(transposed(hb)))


def matches(ha, hb, part):
    """Strip-pair column matches: sum over k of (A strips holding column k) x
    (B strips holding column k)."""
    def strips(h):
        rows = np.repeat(np.arange(h.m), np.diff(h.row_ptr))
        pairs = np.unique(np.stack([h.col_ind.astype(np.int64), rows // part]), axis=1)
        return np.bincount(pairs[0], minlength=h.n)
    return int((strips(ha) * strips(hb)).sum())


def check_exact(got, want):
    assert (got.m, got.n) == (want.m, want.n)
    assert np.array_equal(got.row_ptr, want.row_ptr)
    assert np.array_equal(got.col_ind, want.col_ind)
    wv = want.val if want.val is not None else np.ones(want.nnz, np.float32)
    assert np.array_equal(got.val, wv.astype(np.float32))


@pytest.mark.parametrize("part", [1, 3, 64, 1000, 65536])
def test_outer_bit_exact_against_oracle(ctx, part):
    ha = A.powerlaw(600, 500, 5.0, 200, 21)
    hb = A.powerlaw(700, 500, 4.0, 300, 22)
    c, st = ctx.spgemm_outer(ctx.upload(ha), ctx.upload(hb), part)
    want, prods = oracle_abt(ha, hb)
    check_exact(c, want)
    assert st.edges == prods and st.vertices == want.nnz
    assert st.improved == matches(ha, hb, part)


def test_outer_golden(ctx):
    # the reference's test_spgemm.mtx, all ones: C = A * A = outer(A, A^T)
    h = A.load_mtx(os.path.join(MTX, "test_spgemm.mtx"))
    c, _ = ctx.spgemm_outer(ctx.upload(h), ctx.upload(transposed(h)), 2)
    g = EXP["spgemm_test_spgemm_ones"]
    assert list(c.row_ptr) == g["row_ptr"] and list(c.col_ind) == g["col_ind"]
    assert list(c.val) == g["val"]


def test_outer_unweighted_rectangular(ctx):
    ha = A.powerlaw(900, 300, 6.0, 250, 31)
    hb = A.powerlaw(200, 300, 7.0, 300, 32)
    ha = A.HostCsr(ha.m, ha.n, ha.row_ptr, ha.col_ind, None)
    hb = A.HostCsr(hb.m, hb.n, hb.row_ptr, hb.col_ind, None)
    c, _ = ctx.spgemm_outer(ctx.upload(ha), ctx.upload(hb), 128)
    want, _ = oracle_abt(ha, hb)
    check_exact(c, want)
    # unweighted: every value counts the common columns of the two rows
    dense = np.zeros((ha.m, hb.m))
    ra = np.repeat(np.arange(ha.m), np.diff(ha.row_ptr))
    rb = np.repeat(np.arange(hb.m), np.diff(hb.row_ptr))
    ma = np.zeros((ha.m, ha.n))
    mb = np.zeros((hb.m, hb.n))
    ma[ra, ha.col_ind] = 1
    mb[rb, hb.col_ind] = 1
    dense = ma @ mb.T
    got = np.zeros_like(dense)
    rc = np.repeat(np.arange(c.m), np.diff(c.row_ptr))
    got[rc, c.col_ind] = c.val
    assert np.array_equal(got, dense)


def test_outer_empty_and_disjoint(ctx):
    m, n = 50, 40
    empty = A.HostCsr(m, n, np.zeros(m + 1, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32))
    full = A.powerlaw(m, n, 3.0, 10, 41)
    for ha, hb in ((empty, full), (full, empty)):
        c, st = ctx.spgemm_outer(ctx.upload(ha), ctx.upload(hb), 16)
        assert (c.m, c.n, c.nnz) == (ha.m, hb.m, 0) and not c.row_ptr.any()
        assert st.edges == 0
    # A uses columns 0..19 only, B columns 20..39 only: no column matches
    rp = np.arange(0, 2 * m + 1, 2, dtype=np.int32)
    ca = np.tile(np.array([3, 7], np.int32), m)
    cb = np.tile(np.array([23, 31], np.int32), m)
    ones = np.ones(2 * m, np.float32)
    c, st = ctx.spgemm_outer(ctx.upload(A.HostCsr(m, n, rp, ca, ones)), ctx.upload(A.HostCsr(m, n, rp, cb, ones)), 8)
    assert c.nnz == 0 and st.improved == 0


def test_outer_argument_checks(ctx):
    ha = A.powerlaw(100, 50, 3.0, 10, 51)
    hb = A.powerlaw(100, 60, 3.0, 10, 52)
    with pytest.raises(A.ApspieError) as e:
        ctx.spgemm_outer(ctx.upload(ha), ctx.upload(hb), 16)  # column counts differ
    assert e.value.code == -1  # APSPIE_EINVAL
    with pytest.raises(A.ApspieError) as e:
        ctx.spgemm_outer(ctx.upload(ha), ctx.upload(ha), 0)
    assert e.value.code == -1  # APSPIE_EINVAL
    big = A.powerlaw(70000, 10, 1.0, 2, 53)
    with pytest.raises(A.ApspieError) as e:  # 70000^2 strip pairs
        ctx.spgemm_outer(ctx.upload(big), ctx.upload(big), 1)
    assert e.value.code == -1  # APSPIE_EINVAL


def test_outer_repeated_calls_and_rowwise_agreement(ctx):
    """The 1M x 1M power-law A * A (mean row 8, rows up to 4096) through the
    outer products of A and A^T, against the row-wise apspie_spgemm: the same
    structure and values (both FP64 over k ascending); a second, smaller call
    on the same context reuses the scratch slots."""
    h = A.powerlaw(1 << 20, 1 << 20, 8.0, 4096, 3)
    g, gt = ctx.upload(h), ctx.upload(A.transpose(h))
    want, _ = ctx.spgemm(g, g)
    got, st = ctx.spgemm_outer(g, gt, 65536)
    assert np.array_equal(got.row_ptr, want.row_ptr) and np.array_equal(got.col_ind, want.col_ind)
    err = np.abs(got.val.astype(np.float64) - want.val.astype(np.float64))
    assert np.all(err <= 1e-5 * np.abs(want.val.astype(np.float64))), float(err.max())
    small_a = A.powerlaw(300, 200, 4.0, 50, 61)
    small_b = A.powerlaw(250, 200, 4.0, 50, 62)
    c, _ = ctx.spgemm_outer(ctx.upload(small_a), ctx.upload(small_b), 100)
    check_exact(c, oracle_abt(small_a, small_b)[0])
