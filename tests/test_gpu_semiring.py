"""GPU parity of the general-semiring SpMSpV (apspie_mxv, apspie_mxv_sparse)
against the oracle's restatement of the reference's mXv: every semiring,
dense and sparse inputs, weighted and unweighted, rectangular matrices and
empty inputs.  Min, max and or are exact; plus-times sums in FP64 and rounds
once, so it is exact on integer data and within 1e-6 relative otherwise."""
import numpy as np
import pytest

import apspie_amd as A
import oracle as O

pytestmark = pytest.mark.gpu
SR = [A.SR_PLUS_TIMES, A.SR_MIN_PLUS, A.SR_MAX_TIMES, A.SR_OR_AND]


def ocsr(h):
    return O.Csr(h.m, h.n, h.row_ptr, h.col_ind, h.val)


def dense_x(sr, m, seed, frac=0.2, integer=True):
    rs = np.random.RandomState(seed)
    v = rs.randint(1, 9, m).astype(np.float32) if integer else rs.uniform(-1, 1, m).astype(np.float32)
    on = rs.rand(m) < frac
    absent = {A.SR_PLUS_TIMES: 0.0, A.SR_MIN_PLUS: 1.70141e38, A.SR_MAX_TIMES: -np.inf, A.SR_OR_AND: 0.0}[sr]
    if sr == A.SR_OR_AND:
        v = np.ones(m, np.float32)
    return np.where(on, v, absent).astype(np.float32)


@pytest.mark.parametrize("sr", SR)
@pytest.mark.parametrize("weighted", [True, False])
def test_mxv_dense(ctx, sr, weighted):
    h = A.rmat(12, 16, 6212, weight_seed=7212 if weighted else None)
    g = ctx.upload(h)
    x = dense_x(sr, h.m, 3 + sr)
    y, st = ctx.mxv(g, x, sr)
    want, _ = O.mxv(ocsr(h), sr, x)
    assert np.array_equal(y, want), sr


def test_mxv_plus_times_fractional_and_rectangular(ctx):
    h = A.powerlaw(5000, 3000, 6.0, 200, 6214)
    g = ctx.upload(h)
    x = dense_x(A.SR_PLUS_TIMES, h.m, 9, frac=0.5, integer=False)
    y, _ = ctx.mxv(g, x)
    want, _ = O.mxv(ocsr(h), A.SR_PLUS_TIMES, x)
    assert np.allclose(y, want, rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("sr", SR)
def test_mxv_sparse(ctx, sr):
    h = A.rmat(11, 16, 6311, weight_seed=7311)
    g = ctx.upload(h)
    rs = np.random.RandomState(sr)
    ind = np.sort(rs.choice(h.m, 150, replace=False)).astype(np.int32)
    val = rs.randint(0, 5, 150).astype(np.float32)  # zeros too: listed entries are present anyway
    yi, yv = ctx.mxv_sparse(g, ind, val, sr)
    x = np.zeros(h.m, np.float32)
    x[ind] = val
    listed = np.zeros(h.m, np.uint8)
    listed[ind] = 1
    y, reached = O.mxv(ocsr(h), sr, x, listed)
    want_i = np.nonzero(reached)[0].astype(np.int32)
    assert np.array_equal(yi, want_i) and np.array_equal(yv, y[want_i])


def test_mxv_empty_inputs(ctx):
    h = A.rmat(10, 8, 6410, weight_seed=7410)
    g = ctx.upload(h)
    yi, yv = ctx.mxv_sparse(g, np.zeros(0, np.int32), np.zeros(0, np.float32))
    assert len(yi) == 0 and len(yv) == 0
    y, _ = ctx.mxv(g, np.zeros(h.m, np.float32))
    assert not y.any()
    with pytest.raises(A.ApspieError):
        ctx.mxv_sparse(g, np.array([h.m], np.int32), np.array([1.0], np.float32))
    with pytest.raises(A.ApspieError):
        ctx.mxv(g, np.zeros(h.m, np.float32), 7)


def test_mxv_reference_frontier_rule(ctx):
    """The reference mXv's frontier is x > 0.5 (bitify, src/mXv.cuh:50-55):
    entries of 0.25 are left out, as if they were 0."""
    h = A.rmat(11, 8, 6511, weight_seed=7511)
    g = ctx.upload(h)
    rs = np.random.RandomState(8)
    x = rs.choice(np.array([0.0, 0.25, 0.5, 1.0, 3.0], np.float32), h.m)
    y, st = ctx.mxv(g, x, A.SR_PLUS_TIMES, above_half=True)
    want, _ = O.mxv(ocsr(h), A.SR_PLUS_TIMES, np.where(x > 0.5, x, 0).astype(np.float32))
    assert np.array_equal(y, want)
    assert st.edges == int(np.diff(h.row_ptr)[x > 0.5].sum())
