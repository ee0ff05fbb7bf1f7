"""GPU parity of apspie_mis and apspie_spmspv_or against the oracle:
bit-exact MIS membership under mt19937 priorities (greedy by (priority, -id)),
the reference's index-order MIS on its own inputs, and the reference's XOR
validity check (src/util.cuh:277) computed with the GPU SpMSpV."""
import json
import os

import numpy as np
import pytest

import apspie_amd as A
import oracle as O
from conftest import GOLDEN, MTX

pytestmark = pytest.mark.gpu

SMALL = ["test_cc", "test_bc", "test_pr", "test_mesh", "test_mis", "test_spgemm"]
EXP = json.load(open(os.path.join(GOLDEN, "expected.json")))


def ocsr(h):
    return O.Csr(h.m, h.n, h.row_ptr, h.col_ind, h.val)


@pytest.mark.parametrize("name", SMALL)
def test_mis_index_order_golden(ctx, name):
    h = A.load_mtx(os.path.join(MTX, name + ".mtx"), undirected=True)
    g = ctx.upload(h)
    prio = np.full(h.m, 7, dtype=np.uint32)
    prio[0] = 8  # source first, then index order (SimpleReferenceMis)
    got, _ = ctx.mis(g, prio)
    assert list(got) == EXP["mis_index_order_src0_und"][name]


def test_mis_chesapeake(ctx):
    h = A.load_mtx(os.path.join(MTX, "chesapeake.mtx"), undirected=True)
    g = ctx.upload(h)
    prio = np.full(h.m, 1, dtype=np.uint32)
    prio[0] = 2
    got, _ = ctx.mis(g, prio)
    assert int(got.sum()) == EXP["chesapeake_mis_size"]
    assert np.array_equal(got, O.mis_index(ocsr(h), 0))


@pytest.mark.parametrize("scale,seed,ties", [(10, 4010, False), (14, 4014, False), (14, 4114, True), (18, 4018, False)])
def test_mis_rmat_mt19937(ctx, scale, seed, ties):
    h = A.rmat(scale, 16, seed, undirected=True)
    g = ctx.upload(h)
    prio = O.mt_fill(seed, h.m)
    if ties:
        prio = prio >> 28  # 16 distinct values: the -id tie-break decides
    got, st = ctx.mis(g, prio)
    want = O.mis_greedy(ocsr(h), prio)
    assert np.array_equal(got, want)
    # invariant: in the set XOR has a neighbour in the set
    has = ctx.spmspv_or(g, got)
    assert np.all(got + has == 1)
    assert st.iterations >= 1


def test_mis_isolated_and_self_loops(ctx):
    # 0-1 edge, 2 isolated, 3 with a self-loop, 4-5 edge
    rp = np.array([0, 1, 2, 2, 3, 4, 5], dtype=np.int32)
    ci = np.array([1, 0, 3, 5, 4], dtype=np.int32)
    h = A.HostCsr(6, 6, rp, ci)
    g = ctx.upload(h)
    prio = np.array([5, 9, 1, 3, 2, 2], dtype=np.uint32)
    got, _ = ctx.mis(g, prio)
    assert np.array_equal(got, O.mis_greedy(ocsr(h), prio))
    assert list(got) == [0, 1, 1, 1, 1, 0]


@pytest.mark.parametrize("und", [False, True])
def test_spmspv_or(ctx, und):
    h = A.rmat(13, 16, 77, undirected=und)
    g = ctx.upload(h)
    rng = np.random.RandomState(3)
    for frac in (0.0, 0.001, 0.3, 1.0):
        x = (rng.rand(h.m) < frac).astype(np.int32) * rng.randint(1, 5, h.m)
        got = ctx.spmspv_or(g, x)
        assert np.array_equal(got, O.spmspv_or(ocsr(h), x))
