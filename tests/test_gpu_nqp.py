"""GPU parity at other Gauss orders per segment pair (n_qp_pair 2, 3, 7 on
the junction tee; 5 over PEC ground): Z and Y against the reference's
(tests/golden/nqp*.npz) within 1e-10."""

import pytest

from conftest import golden, rel_err
from momwire_amd import _lib
from momwire_amd.bspline import BSplineSolver
from test_nqp_host import CASES, nqp_case

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.mark.parametrize("case", CASES)
def test_nqp_matches_reference(case):
    print("[gpu-test] n_qp_pair", case, flush=True)
    d = golden(case)
    s = BSplineSolver(**nqp_case(case))
    lib = _lib.load()
    D = s._device()
    st = torch.cuda.current_stream().cuda_stream
    _lib.check(lib.mw_same_edge_table(D.mw, s.k, st), "same-edge")
    s._fill(D, s.k, s.omega, D.Z.data_ptr(), D.ld, st)
    torch.cuda.synchronize()
    assert rel_err(D.Z[:, :D.nb].cpu().numpy(), d["Z"]) < 1e-10
    sol = s.compute_port_solution()
    assert rel_err(sol.y, d["y"]) < 1e-10 and rel_err(sol.coeffs, d["coeffs"]) < 1e-10


@pytest.mark.parametrize("nqs", [2, 5])
def test_nqs_matches_reference(nqs):
    from momwire_amd import _ground, workloads as W
    print("[gpu-test] n_qp_sommerfeld", nqs, flush=True)
    d = golden(f"nqs{nqs}_ground_tee")
    kw = W.grounded_tee()
    kw["n_qp_sommerfeld"] = nqs
    _ground.clear_cache()
    sol = BSplineSolver(**kw).compute_port_solution()
    assert rel_err(sol.y, d["y"]) < 1e-10 and rel_err(sol.coeffs, d["coeffs"]) < 1e-10
