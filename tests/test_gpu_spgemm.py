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
