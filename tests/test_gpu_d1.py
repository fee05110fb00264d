"""GPU parity of degree-1 (tent) bases: Z, Y and coefficients against the
reference's (tests/golden/d1_*.npz; Z and Y within 1e-10, mixed Y within
1e-8) and the reference's pinned tent impedances (relative 1e-6, the bar of
its own tests/test_tent_parity.py) through compute_impedance."""

import numpy as np
import pytest

from conftest import golden, rel_err
from momwire_amd import _lib
from momwire_amd.bspline import BSplineSolver
from test_d1_host import CASES, PINS, d1_case

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _say(*a):
    print("[gpu-test]", *a, flush=True)


def test_d1_Z_matches_reference():
    _say("degree-1 Z")
    d = golden("d1_dipole")
    s = BSplineSolver(**d1_case("d1_dipole"))
    lib = _lib.load()
    D = s._device()
    st = torch.cuda.current_stream().cuda_stream
    _lib.check(lib.mw_same_edge_table(D.mw, s.k, st), "same-edge")
    s._fill(D, s.k, s.omega, D.Z.data_ptr(), D.ld, st)
    torch.cuda.synchronize()
    Z = D.Z[:, :D.nb].cpu().numpy()
    _say("  Z err", rel_err(Z, d["Z"]))
    assert rel_err(Z, d["Z"]) < 1e-10


@pytest.mark.parametrize("case", CASES)
def test_d1_port_solution_matches_reference(case):
    d = golden(case)
    for precision in ("fp64", "mixed"):
        _say("degree-1", case, precision)
        sol = BSplineSolver(**d1_case(case), precision=precision).compute_port_solution()
        assert rel_err(sol.y, d["y"]) < (1e-10 if precision == "fp64" else 1e-8)
        if precision == "fp64":
            assert rel_err(sol.coeffs, d["coeffs"]) < 1e-10


def test_d1_reference_pins():
    for wires, nsegs, z_pin in PINS:
        z, coeffs = BSplineSolver(wires=[np.asarray(w, float) for w in wires], nsegs=nsegs,
                                  degree=1).compute_impedance()
        _say(f"  pin {z_pin:.6f}: {z:.6f}")
        assert abs(z - z_pin) <= 1e-6 * abs(z_pin)
        assert coeffs.ndim == 1
