"""Pins the oracle's expression analysis and reshape to golden vectors that
the reference's own headers produced (tests/golden/gen_contract_info.py runs
src/Basics/ContractInfo.h and Utils.h compiled from the reference tree)."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as orc

GOLD = os.path.join(os.path.dirname(__file__), "golden")
CASES = json.load(open(os.path.join(GOLD, "contract_info.json")))["cases"]
RESHAPES = json.load(open(os.path.join(GOLD, "reshape.json")))["cases"]

# The reference checks the output index count only (ContractInfo.h:257) and so
# accepts an output index that neither input has; the build rejects it.
DIVERGENT = {"ij,jk->il"}


@pytest.mark.parametrize("case", CASES, ids=[c["expr"] or "<empty>" for c in CASES])
def test_contract_info_matches_reference(case):
    expr = case["expr"]
    if expr in DIVERGENT:
        assert case["ok"]  # reference accepted it
        with pytest.raises(orc.OracleError, match="Dimension mismatch"):
            orc.contract_info(expr, case["sa"], case["sb"], case["sc"])
        return
    if not case["ok"]:
        with pytest.raises(orc.OracleError) as ei:
            orc.contract_info(expr, case["sa"], case["sb"], case["sc"])
        assert str(ei.value) == case["err"]
        return
    m = orc.contract_info(expr, case["sa"], case["sb"], case["sc"])
    for k in ("a_row", "a_col", "b_row", "b_col", "c_row", "c_col"):
        assert m[k] == case[k], k
    for t, sizes in (("a", case["sa"]), ("b", case["sb"]), ("c", case["sc"])):
        assert orc.scatter(m[t + "_row"], sizes).tolist() == case[t + "_srow"]
        assert orc.scatter(m[t + "_col"], sizes).tolist() == case[t + "_scol"]


@pytest.mark.parametrize("case", RESHAPES, ids=[str(c["order"]) for c in RESHAPES])
def test_reshape_matches_reference(case):
    n = int(np.prod(case["sizes"]))
    out = orc.reshape(np.arange(n, dtype=np.float64), case["sizes"], case["order"])
    assert out.astype(np.int64).tolist() == case["out"]


def test_reshape_is_transpose():
    # closed form: reshape with order p of a column-major array is numpy's
    # transpose(p) of the Fortran view.
    rng = np.random.default_rng(1)
    sizes, order = [3, 4, 5, 2], [2, 0, 3, 1]
    x = rng.standard_normal(int(np.prod(sizes)))
    want = np.transpose(x.reshape(sizes, order="F"), order).ravel(order="F")
    assert np.array_equal(orc.reshape(x, sizes, order), want)
