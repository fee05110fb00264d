"""GPU parity of the spread gap feeds (segment gap, cos^2-smoothed source):
Y and coefficients against the reference's (tests/golden/feed_*.npz),
FP64 within 1e-10, mixed within 1e-8."""

import pytest

from conftest import golden, rel_err
from momwire_amd.bspline import BSplineSolver
from test_feeds_host import CASES, feed_case

pytestmark = pytest.mark.gpu
pytest.importorskip("torch")


@pytest.mark.parametrize("case", CASES)
def test_spread_feeds_match_reference(case):
    d = golden(case)
    for precision in ("fp64", "mixed"):
        print("[gpu-test] spread feed", case, precision, flush=True)
        sol = BSplineSolver(**feed_case(case), precision=precision).compute_port_solution()
        assert rel_err(sol.y, d["y"]) < (1e-10 if precision == "fp64" else 1e-8)
        if precision == "fp64":
            assert rel_err(sol.coeffs, d["coeffs"]) < 1e-10
