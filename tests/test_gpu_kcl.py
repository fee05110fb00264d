"""GPU parity of the KCL Schur step past the small solve: 39 and 69
constraint rows (and 70 ports), solved by the blocked LU on [S | T], against
the reference's Y and coefficients (tests/golden/chain*.npz); FP64 within
1e-10, mixed within 1e-8; graph replays equal the eager solve."""

import pytest

from conftest import golden, rel_err
from momwire_amd import workloads as W
from momwire_amd.bspline import BSplineSolver

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

CASES = {"chain40": (40, None), "chain70_feeds": (70, 1)}


def _say(*a):
    print("[gpu-test]", *a, flush=True)


@pytest.mark.parametrize("case", sorted(CASES))
def test_large_kcl_matches_reference(case):
    n, every = CASES[case]
    d = golden(case)
    for precision in ("fp64", "mixed"):
        _say("large KCL", case, precision)
        s = BSplineSolver(**W.wire_chain(n, 3, every), precision=precision)
        sol = s.compute_port_solution()
        _say("  Y err", rel_err(sol.y, d["y"]))
        assert rel_err(sol.y, d["y"]) < (1e-10 if precision == "fp64" else 1e-8)
        if precision == "fp64":
            assert rel_err(sol.coeffs, d["coeffs"]) < 1e-10
            assert not s._dev.kcl_small


def test_large_kcl_graph_replay():
    _say("large KCL graph replay")
    s = BSplineSolver(**W.wire_chain(40, 3, None))
    y = s.compute_port_solution().y
    for _ in range(3):
        Y, _X = s.solve_graph()
        torch.cuda.synchronize()
        assert rel_err(Y.cpu().numpy(), y) < 1e-13
        s.compute_port_solution()
