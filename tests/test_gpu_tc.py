"""GPU parity of apspie_tc: exact int64 triangle counts against the oracle's
masked intersection count, closed forms (K_n -> C(n,3), triangle-free mesh),
the golden counts of the reference's own inputs, and the RMAT scale-20 count
recorded in the golden fixture."""
import json
import os

import numpy as np
import pytest

import apspie_amd as A
import oracle as O
from conftest import GOLDEN, MTX

pytestmark = pytest.mark.gpu

EXP = json.load(open(os.path.join(GOLDEN, "expected.json")))
NAMES = ["test_cc", "test_bc", "test_pr", "test_mesh", "test_mis", "test_spgemm", "chesapeake"]


def count(ctx, h):
    lo = A.strict_lower(h)
    return ctx.tc(ctx.upload(lo))


@pytest.mark.parametrize("name", NAMES)
def test_tc_golden(ctx, name):
    h = A.load_mtx(os.path.join(MTX, name + ".mtx"), undirected=True)
    t, _ = count(ctx, h)
    assert t == EXP["triangles_und"][name]


@pytest.mark.parametrize("n", [3, 10, 200, 1500])
def test_tc_complete_graph(ctx, n):
    rp = np.arange(0, n * (n - 1) + 1, n - 1, dtype=np.int32)
    ci = np.array([j for i in range(n) for j in range(n) if j != i], dtype=np.int32)
    t, st = count(ctx, A.HostCsr(n, n, rp, ci))
    assert t == n * (n - 1) * (n - 2) // 6
    assert st.edges == sum(k for i in range(n) for k in range(i))  # wedges


@pytest.mark.parametrize("scale,seed", [(10, 3010), (14, 3014), (16, 3016)])
def test_tc_rmat_vs_oracle(ctx, scale, seed):
    h = A.rmat(scale, 16, seed, undirected=True)
    lo = A.strict_lower(h)
    t, st = ctx.tc(ctx.upload(lo))
    want, wedges = O.tc(O.Csr(lo.m, lo.n, lo.row_ptr, lo.col_ind, None))
    assert t == want and st.edges == wedges


def test_tc_rmat_s20_golden(ctx):
    """RMAT scale 20, edge factor 16, seed 3020, symmetrized: the count the
    oracle computed (tests/golden/make_golden.py)."""
    h = A.rmat(20, 16, 3020, undirected=True)
    t, st = count(ctx, h)
    g = EXP["tc_rmat_s20_seed3020"]
    assert t == g["triangles"] and st.edges == g["wedges"]
