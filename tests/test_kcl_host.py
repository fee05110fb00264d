"""KCL constraint sets larger than the small Schur solve's 32 rows (and
junctioned structures with more than 64 ports): the oracle against the
reference's solve (tests/golden/chain*.npz), and the host port plan."""

import pytest

from conftest import golden, rel_err
from momwire_amd import workloads as W
from momwire_amd.bspline import BSplineSolver
from oracle.bspline_oracle import OracleSolver

CASES = {"chain40": (40, None), "chain70_feeds": (70, 1)}


@pytest.mark.parametrize("case", sorted(CASES))
def test_chain_oracle_matches_reference(case):
    d = golden(case)
    y, X = OracleSolver(**W.wire_chain(CASES[case][0], 3, CASES[case][1])).compute_port_solution()
    assert rel_err(y, d["y"]) < 1e-10
    assert rel_err(X, d["coeffs"]) < 1e-10


@pytest.mark.parametrize("case", sorted(CASES))
def test_chain_port_plan(case):
    n, every = CASES[case]
    cols, con = BSplineSolver(**W.wire_chain(n, 3, every))._port_plan()
    assert len(con) == n - 1
    assert len(cols) == (n if every else 1)
