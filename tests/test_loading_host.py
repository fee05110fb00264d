"""Distributed wire loading on the CPU (reference _wire_loading.py,
bspline.py:2667-2800): the oracle against the reference's Y, coefficients
and loss power (tests/golden/loading_*.npz); the host Gram and merged
entries against the oracle's; the series impedance against its closed-form
limits; the constructor's checks."""

import numpy as np
import pytest

from conftest import golden, rel_err
from momwire_amd import _loading, workloads as W
from momwire_amd.bspline import BSplineSolver
from oracle.bspline_oracle import OracleSolver

NAN = float("nan")


def loading_case(name):
    if name == "loading_dipole":
        kw = W.straight_dipole()
        kw.update(wire_conductivity=5.8e7)
    elif name == "loading_tee":
        kw = W.junction_tee("junction_port")
        kw.update(insulation_radius=2.5e-3, insulation_eps_r=3.2, wire_conductivity=3.5e7)
    else:
        kw = W.dipole_array(2, 2, 11, seed=5)
        kw.update(wire_conductivity=[5.8e7, NAN, 1e6, 3e7],
                  insulation_radius=[NAN, 2e-3, 3e-3, NAN],
                  insulation_eps_r=[NAN, 2.3, 4.0, NAN])
    return kw


CASES = ("loading_dipole", "loading_tee", "loading_array2x2")


@pytest.mark.parametrize("case", CASES)
def test_loading_oracle_matches_reference(case):
    d = golden(case)
    kw = loading_case(case)
    y, X = OracleSolver(**kw).compute_port_solution()
    assert rel_err(y, d["y"]) < 1e-10
    assert rel_err(X, d["coeffs"]) < 1e-10
    # the loading is visible in Y
    kw0 = {k: v for k, v in kw.items() if k not in ("wire_conductivity", "insulation_radius",
                                                     "insulation_eps_r")}
    y0, _X0 = OracleSolver(**kw0).compute_port_solution()
    assert rel_err(y0, d["y"]) > 1e-6


@pytest.mark.parametrize("case", CASES)
def test_host_loading_term_matches_oracle(case):
    d = golden(case)
    kw = loading_case(case)
    s = BSplineSolver(**kw)
    term = _loading.LoadingTerm(s._build(), s._loading)
    L = np.zeros((s._build().n_basis,) * 2, dtype=complex)
    L[term.rows, term.cols] = term.entries(s.omega)
    o = OracleSolver(**kw)
    geom = o.geometry()
    supp, polys = o.basis(geom)[:2]
    Lo = o.loading_term(geom, supp, polys)
    assert rel_err(L, Lo) < 1e-13
    assert np.abs(L - L.T).max() < 1e-13 * np.abs(L).max()
    p_tot, p_wire = s.wire_loss_power(d["coeffs"][:, 0])
    assert abs(p_tot - float(d["loss"])) < 1e-12 * abs(float(d["loss"]))
    assert np.abs(p_wire - d["loss_per_wire"]).max() < 1e-12 * abs(float(d["loss"]))


def test_series_impedance_limits():
    """DC: Z_int -> 1/(sigma pi a^2) + j omega mu0/(8 pi); strong skin effect: (1 + j)/(2 pi a
    sigma delta), delta = sqrt(2/(omega mu0 sigma)); insulation: j omega
    mu0/(2 pi) (1 - 1/eps_r) ln(b/a)."""
    a, sig = 1e-3, 5.8e7
    z_dc = _loading.wire_internal_impedance(1e-3, a, sig)
    r_dc = 1 / (sig * np.pi * a * a)
    assert abs(z_dc.real - r_dc) < 1e-12 * r_dc
    # the low-frequency internal inductance mu0 / (8 pi)
    assert abs(z_dc.imag - 1e-3 * _loading.MU0 / (8 * np.pi)) < 1e-6 * z_dc.imag
    w = 2 * np.pi * 3e9
    delta = np.sqrt(2 / (w * _loading.MU0 * sig))
    z_hf = _loading.wire_internal_impedance(w, a, sig)
    exact = (1 + 1j) / (2 * np.pi * a * sig * delta)
    assert abs(z_hf - exact) < 2e-3 * abs(exact)  # next term is O(delta / a)
    L = _loading.insulation_inductance(a, 3e-3, 2.5)
    assert abs(L - _loading.MU0 / (2 * np.pi) * 0.6 * np.log(3.0)) < 1e-22


def test_loading_validation():
    kw = W.straight_dipole()
    with pytest.raises(ValueError, match="given together"):
        BSplineSolver(**kw, insulation_radius=2e-3)
    with pytest.raises(ValueError, match="must exceed"):
        BSplineSolver(**kw, insulation_radius=1e-4, insulation_eps_r=2.0)
    with pytest.raises(ValueError, match="must be > 0"):
        BSplineSolver(**kw, wire_conductivity=-1.0)
    with pytest.raises(ValueError, match="length-1"):
        BSplineSolver(**kw, wire_conductivity=[1.0, 2.0])
