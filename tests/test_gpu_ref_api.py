"""Runs the C++ driver of the reference-named host API
(tests/cpp/test_ref_api.cpp, built by __graft_entry__.build()): BFS, SSSP,
MIS (+ the XOR validity check), SpGEMM (A * D^T as the reference's driver
calls it) and TC through bfs / sssp / spmspvMis / spmspvCsr / spgemm / tc,
each checked against the oracle inside the driver."""
import os
import subprocess

import pytest

from conftest import MTX, ROOT

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "tests", "cpp", "_build", "test_ref_api")


@pytest.mark.parametrize("name", ["chesapeake", "test_mesh", "test_cc"])
def test_reference_named_api(name):
    assert os.path.exists(BIN), "build the driver with __graft_entry__.build()"
    r = subprocess.run([BIN, os.path.join(MTX, name + ".mtx")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "REF_API OK" in r.stdout
