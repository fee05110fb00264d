"""GPU parity of the distributed wire loading: the merged Z'_w S_w entries
added into Z on the device (mw_zadd_sparse) after the fill; Y and
coefficients against the reference's (tests/golden/loading_*.npz), FP64
within 1e-10 and mixed within 1e-8; a graph replay equals the eager solve."""

import pytest

from conftest import golden, rel_err
from momwire_amd.bspline import BSplineSolver
from test_loading_host import CASES, loading_case

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.mark.parametrize("case", CASES)
def test_loading_matches_reference(case):
    d = golden(case)
    for precision in ("fp64", "mixed"):
        print("[gpu-test] wire loading", case, precision, flush=True)
        s = BSplineSolver(**loading_case(case), precision=precision)
        sol = s.compute_port_solution()
        assert rel_err(sol.y, d["y"]) < (1e-10 if precision == "fp64" else 1e-8)
        if precision == "fp64":
            assert rel_err(sol.coeffs, d["coeffs"]) < 1e-10
            Y, _W = s.solve_graph()
            torch.cuda.synchronize()
            assert rel_err(Y.cpu().numpy(), sol.y) < 1e-13
