"""Regenerates the oracle-computed entries of tests/golden/expected.json.

The hand brute-force answers for the reference's dataset/small inputs are
kept as they are; this script adds values the CPU oracle computes on seeded
synthetic inputs that are too slow to recompute in every test run.
Run from the repository root: python tests/golden/make_golden.py
"""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import apspie_amd as A  # noqa: E402  (host-side generator only)
import oracle as O  # noqa: E402

PATH = os.path.join(ROOT, "tests", "golden", "expected.json")


def main():
    exp = json.load(open(PATH))
    h = A.rmat(20, 16, 3020, undirected=True)
    lo = O.strict_lower(O.Csr(h.m, h.n, h.row_ptr, h.col_ind, None))
    t, w = O.tc(lo)
    exp["tc_rmat_s20_seed3020"] = {"triangles": t, "wedges": w, "nnz_lower": lo.nnz}
    with open(PATH, "w") as f:
        json.dump(exp, f, indent=2)
        f.write("\n")
    print(exp["tc_rmat_s20_seed3020"])


if __name__ == "__main__":
    main()
