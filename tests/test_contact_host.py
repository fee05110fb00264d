"""Ground contact on the CPU (reference bspline.py:683-698, 1108-1169,
1296-1316, _ground_spec.py:127-200): wire ends in the plane keep their
value-1 basis, grounded junctions lose their KCL row.  The oracle against
the reference's Z, Y and coefficients (tests/golden/contact_*.npz); the
host mesh's kept bases and KCL plan; the refusals."""

import numpy as np
import pytest

from conftest import golden, rel_err
from momwire_amd import workloads as W
from momwire_amd.bspline import BSplineSolver
from oracle.bspline_oracle import OracleSolver

CASES = {"contact_pec_base": ("pec", True), "contact_pec_seg": ("pec", False),
         "contact_somm_base": ("sommerfeld", True), "contact_radials": ("radials", False)}


@pytest.mark.parametrize("case", sorted(CASES))
def test_contact_oracle_matches_reference(case):
    d = golden(case)
    o = OracleSolver(**W.grounded_monopole(*CASES[case]))
    if "Z" in d.files:
        assert rel_err(o.Z(), d["Z"]) < 1e-10
    y, X = o.compute_port_solution()
    assert rel_err(y, d["y"]) < 1e-10
    assert rel_err(X, d["coeffs"]) < 1e-10


def test_contact_basis_and_kcl_plan():
    s = BSplineSolver(**W.grounded_monopole("pec", True))
    kept = s._build().per_wire[0]["kept"]
    assert kept[0][1] == "gnd" and kept[-1][1] == "int"
    # base feed drives the grounded basis alone (value 1 at s = 0)
    cols, con = s._port_plan()
    assert cols[0] == ([0], [1.0]) and con == []
    r = BSplineSolver(**W.grounded_monopole("radials", False))
    assert r._grounded_junctions() == frozenset({0})
    assert r._port_plan()[1] == []


def test_contact_refusals():
    kw = W.grounded_monopole("pec", True)
    with pytest.raises(NotImplementedError, match="refl-coef"):
        BSplineSolver(**dict(kw, ground_eps=(13.0, 0.005)))
    bad = dict(kw, wires=[[(0.0, 0.0, -0.5), (0.0, 0.0, 5.0)]])
    with pytest.raises(ValueError, match="dips below"):
        BSplineSolver(**bad)
    flat = dict(kw, wires=[[(0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (1.0, 0.0, 3.0)]],
                n_per_edge_per_wire=[[5, 5]])
    with pytest.raises(ValueError, match="edge lying"):
        BSplineSolver(**flat)
    r = W.grounded_monopole("radials", False)
    r["junction_ports"] = [0]
    with pytest.raises(ValueError, match="grounded and a junction port"):
        BSplineSolver(**r)._port_plan()
