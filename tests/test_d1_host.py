"""Degree-1 (tent) bases on the CPU: the oracle against the reference's Z,
Y and coefficients (tests/golden/d1_*.npz) and against the reference's own
pinned impedances (tests/test_tent_parity.py:43-80 of the reference: the
retired triangular solver's values, relative 1e-6); the host tables' padding
of a two-wing basis into the three-wing layout."""

import numpy as np
import pytest

from conftest import golden, rel_err
from momwire_amd import workloads as W
from momwire_amd.bspline import BSplineSolver
from oracle.bspline_oracle import OracleSolver

HD = 0.962 * 22 / 4
C30, S30 = np.cos(np.pi / 6), np.sin(np.pi / 6)
PINS = [  # (wires, nsegs, z) -- reference tests/test_tent_parity.py:43-80
    ([[[0, 0, -5.29], [0, 0, 5.29]]], 20, 69.069365 - 20.348461j),
    ([[[0, 0, -5.29], [0, 0, 5.29]]], 40, 69.403082 - 19.237949j),
    ([[[0, 0, -2.65], [0, 0, 2.65]]], 20, 11.984894 - 942.538771j),
    ([[[0, 0, -2.65], [0, 0, 2.65]]], 40, 11.965807 - 941.022355j),
    ([[[0, 0, -HD], [0, 0, HD]], [[-2.2, 0, -1.05 * HD], [-2.2, 0, 1.05 * HD]]], 20,
     27.793081 + 2.327531j),
    ([[[-HD * C30, 0, HD * S30], [0, 0, 0], [HD * C30, 0, HD * S30]]], 10,
     53.636373 - 32.936517j),
]


def d1_case(name):
    if name == "d1_dipole":
        kw = dict(wires=[[(0.0, 0.0, -5.29), (0.0, 0.0, 5.29)]], nsegs=20)
    elif name == "d1_tee":
        kw = W.junction_tee("junction_port")
    else:
        kw = W.grounded_monopole("pec", True)
    kw["degree"] = 1
    return kw


CASES = ("d1_dipole", "d1_tee", "d1_contact")


@pytest.mark.parametrize("case", CASES)
def test_d1_oracle_matches_reference(case):
    d = golden(case)
    o = OracleSolver(**d1_case(case))
    if "Z" in d.files:
        assert rel_err(o.Z(), d["Z"]) < 1e-10
    y, X = o.compute_port_solution()
    assert rel_err(y, d["y"]) < 1e-10
    assert rel_err(X, d["coeffs"]) < 1e-10


@pytest.mark.parametrize("i", range(len(PINS)))
def test_d1_oracle_meets_reference_pins(i):
    wires, nsegs, z_pin = PINS[i]
    y, _X = OracleSolver(wires=[np.asarray(w, float) for w in wires], nsegs=nsegs,
                         degree=1).compute_port_solution()
    z = 1.0 / y[0, 0]
    assert abs(z - z_pin) <= 1e-6 * abs(z_pin)


def test_d1_tables_pad_to_three_wings():
    m = BSplineSolver(**d1_case("d1_dipole"))._build()
    assert m.supp.shape == (m.n_basis, 3) and m.polys.shape == (m.n_basis, 3, 3)
    assert np.all(m.supp[:, 2] == -1) and np.all(m.polys[:, :, 2] == 0.0)
    assert np.all(m.band_basis[:, 2] == -1)
