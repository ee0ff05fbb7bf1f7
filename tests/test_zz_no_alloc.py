"""apspie_sssp_noalloc (scratch in the library's fixed device array, no
allocation in the call) against apspie_sssp: bit-exact distances and equal
counters on two calls in a row, the second on a smaller graph (the array keeps
the first call's contents), and a clear refusal past the array's size."""
import numpy as np
import pytest

import apspie_amd as A

pytestmark = pytest.mark.gpu


def test_noalloc_equals_sssp_twice_then_refuses_too_large(ctx):
    big = A.rmat(18, 16, 7018, undirected=False, weight_seed=8018)
    small = A.rmat(12, 8, 7012, undirected=False, weight_seed=8012)
    for h in (big, small):
        g = ctx.upload(h)
        src = h.argmax_degree()
        print(f"noalloc m={h.m} nnz={h.nnz} src={src}", flush=True)
        want, st = ctx.sssp(g, src)
        d = ctx.alloc(4 * h.m)
        d.upload(np.full(h.m, 7.0, np.float32))  # stale output must be overwritten
        st2 = ctx.sssp_noalloc_into(g, src, d)
        got = d.download(np.float32, h.m)
        assert np.array_equal(got, want)
        assert (st2.edges, st2.vertices, st2.improved, st2.iterations) == (st.edges, st.vertices, st.improved,
                                                                           st.iterations)
        d.free()
    # 48 MiB of fixed scratch hold about 2.4 M vertices: 3 M is refused
    m = 3_000_000
    huge = ctx.upload(A.HostCsr(m, m, np.zeros(m + 1, np.int32), np.zeros(0, np.int32),
                                np.zeros(0, np.float32)))
    d = ctx.alloc(4 * m)
    with pytest.raises(A.ApspieError) as e:
        ctx.sssp_noalloc_into(huge, 0, d)
    assert e.value.code == -5 and "scratch" in str(e.value)
    d.free()
