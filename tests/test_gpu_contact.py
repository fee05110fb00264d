"""GPU parity of ground contact: quarter-wave verticals standing on PEC and
on the Sommerfeld ground (base-fed and segment-fed), and a vertical with
three radials meeting at a grounded junction; Z (PEC), Y and coefficients
against the reference's (tests/golden/contact_*.npz): Z and Y within 1e-10
in FP64, Y within 1e-8 in mixed precision."""

import pytest

from conftest import golden, rel_err
from momwire_amd import _ground, _lib, workloads as W
from momwire_amd.bspline import BSplineSolver
from test_contact_host import CASES

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _say(*a):
    print("[gpu-test]", *a, flush=True)


def test_contact_Z_matches_reference():
    _say("ground contact Z (PEC)")
    d = golden("contact_pec_base")
    s = BSplineSolver(**W.grounded_monopole("pec", True))
    lib = _lib.load()
    D = s._device()
    st = torch.cuda.current_stream().cuda_stream
    _lib.check(lib.mw_same_edge_table(D.mw, s.k, st), "same-edge")
    s._fill(D, s.k, s.omega, D.Z.data_ptr(), D.ld, st)
    torch.cuda.synchronize()
    Z = D.Z[:, :D.nb].cpu().numpy()
    _say("  Z err", rel_err(Z, d["Z"]))
    assert rel_err(Z, d["Z"]) < 1e-10


@pytest.mark.parametrize("case", sorted(CASES))
def test_contact_port_solution_matches_reference(case):
    d = golden(case)
    for precision in ("fp64", "mixed"):
        _say("ground contact", case, precision)
        _ground.clear_cache()
        sol = BSplineSolver(**W.grounded_monopole(*CASES[case]),
                            precision=precision).compute_port_solution()
        _say("  Y err", rel_err(sol.y, d["y"]))
        assert rel_err(sol.y, d["y"]) < (1e-10 if precision == "fp64" else 1e-8)
        if precision == "fp64":
            assert rel_err(sol.coeffs, d["coeffs"]) < 1e-10
