"""GPU parity of the extended thin-wire kernel (``extended_kernel=True``):
the static closed forms, the remainder and the off-edge and image factor on
coaxial equal-radius pairs, against the reference's Z, Y and coefficients
(tests/golden/ek_*.npz) for a straight dipole, collinear wires across a
junction, a bent tee, a vertical wire over PEC ground (image coaxial with its
wire), a dipole over the reflection-coefficient ground and a tee over the
Sommerfeld ground.  Bars: Z, Y and coefficients within 1e-10 in FP64, Y
within 1e-8 in mixed precision (as for every other structure).
"""

import numpy as np
import pytest

from conftest import golden, rel_err
from momwire_amd import _ground, _lib, workloads as W
from momwire_amd.bspline import BSplineSolver

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

CASES = ("dipole", "collinear", "tee", "pec_vertical", "refl_dipole", "somm_tee")


def _say(*a):
    print("[gpu-test]", *a, flush=True)


def _device_Z(s):
    lib = _lib.load()
    D = s._device()
    st = torch.cuda.current_stream().cuda_stream
    _lib.check(lib.mw_same_edge_table(D.mw, s.k, st), "same-edge")
    s._fill(D, s.k, s.omega, D.Z.data_ptr(), D.ld, st)
    torch.cuda.synchronize()
    return D.Z[:, :D.nb].cpu().numpy()


@pytest.mark.parametrize("case", CASES)
def test_ek_Z_matches_reference(case):
    _say("extended-kernel Z", case)
    d = golden(f"ek_{case}")
    _ground.clear_cache()
    Z = _device_Z(BSplineSolver(**W.ek_structure(case)))
    _say("  Z err", rel_err(Z, d["Z"]))
    assert rel_err(Z, d["Z"]) < 1e-10


@pytest.mark.parametrize("case", CASES)
def test_ek_port_solution_matches_reference(case):
    _say("extended-kernel port solution", case)
    d = golden(f"ek_{case}")
    for precision in ("fp64", "mixed"):
        sol = BSplineSolver(**W.ek_structure(case), precision=precision).compute_port_solution()
        _say(f"  {precision}: Y err", rel_err(sol.y, d["y"]))
        assert rel_err(sol.y, d["y"]) < (1e-10 if precision == "fp64" else 1e-8)
        if precision == "fp64":
            assert rel_err(sol.coeffs, d["coeffs"]) < 1e-10


def test_ek_reduces_to_reduced_kernel_for_thin_wires():
    """With a radius 1e-7 of the segment length the extended kernel's
    corrections (O(a / h) on same-edge moments, O((a / R)^2) elsewhere) fall
    below 1e-6 of Z: Z equals the reduced kernel's to that bar."""
    _say("extended kernel, thin-wire limit")
    kw = W.ek_structure("collinear")
    kw["wire_radius"] = 1e-9
    Ze = _device_Z(BSplineSolver(**kw))
    kw["extended_kernel"] = False
    Zr = _device_Z(BSplineSolver(**kw))
    assert rel_err(Ze, Zr) < 1e-6
