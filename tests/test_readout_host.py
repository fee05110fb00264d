"""The readout oracle on the CPU: against the reference's own element
walk, far-field moments and near fields (tests/golden/readout_*.npz, written
by the reference's functions), and against closed forms; plus the host
tables the GPU kernels read."""

import numpy as np
import pytest

from conftest import golden, rel_err
from momwire_amd import workloads as W
from momwire_amd.bspline import BSplineSolver
from oracle import readout_oracle as ro
from oracle.bspline_oracle import OracleSolver

CASES = {"readout_tee": lambda: W.junction_tee("junction_port"),
         "readout_array2x2": lambda: W.dipole_array(2, 2, 11, seed=5)}


@pytest.mark.parametrize("case", sorted(CASES))
def test_readout_oracle_matches_reference(case):
    d = golden(case)
    o = OracleSolver(**CASES[case]())
    X = d["coeffs"]
    p = X.shape[1]
    k, gz, f = float(d["k"]), float(d["ground_z"]), float(d["freq_hz"])
    for sub in (1, 8):
        for j in range(p):
            mid, mom, nodes, dlt = ro.element_currents(o, X[:, j], sub)
            assert np.array_equal(mid, d[f"mid{sub}"]) and np.array_equal(nodes, d[f"nodes{sub}"])
            assert rel_err(mom, d[f"moment{sub}"][..., j]) < 1e-13
            assert rel_err(dlt, d[f"delta{sub}"][..., j]) < 1e-12
    th, ph = d["theta"], d["phi"]
    for g in ("free", "pec", "finite"):
        for j in range(p):
            mt, mp = ro.far_moments(d["mid1"], d["moment1"][..., j], k, th, ph, g, gz, 13.0,
                                    0.005, f)
            ref_t, ref_p = d[f"far_{g}_theta"][..., j], d[f"far_{g}_phi"][..., j]
            scale = max(np.abs(ref_t).max(), np.abs(ref_p).max())
            assert np.abs(mt - ref_t).max() < 1e-12 * scale, g
            assert np.abs(mp - ref_p).max() < 1e-12 * scale, g
    pts, a = d["points"], float(d["radius"])
    img_mid = d["mid8"].copy()
    img_mid[:, 2] = 2 * gz - img_mid[:, 2]
    img_nodes = d["nodes8"].copy()
    img_nodes[:, 2] = 2 * gz - img_nodes[:, 2]
    for magnetic, tag in ((False, "E"), (True, "H")):
        for j in range(p):
            els = (d["mid8"], d["moment8"][..., j], d["nodes8"], d["delta8"][..., j])
            f0 = ro.element_fields(pts, els, k, a, magnetic)
            assert rel_err(f0, d[f"near_{tag}"][..., j]) < 1e-12
            img = ro.element_fields(pts, (img_mid, els[1] * [-1.0, -1.0, 1.0], img_nodes,
                                          -els[3]), k, a, magnetic)
            assert rel_err(f0 + img, d[f"near_{tag}_pec"][..., j]) < 1e-12


def test_far_moments_closed_forms():
    """One z-directed element I dl at height h: M_theta = -sin(t) I dl
    exp(jk h cos t), M_phi = 0; over PEC the image doubles it to
    -2 sin(t) I dl cos(k h cos t) (ground at z = 0); at normal incidence the
    Fresnel multipliers are rho_v = -rho_h = (sqrt(eps) - 1)/(sqrt(eps) + 1)."""
    k, h, idl = 2 * np.pi, 0.3, 0.7 - 0.2j
    th = np.linspace(0.05, np.pi - 0.05, 11)
    ph = np.linspace(0.0, 2 * np.pi, 7)
    mid = np.array([[0.0, 0.0, h]])
    mom = np.array([[0.0, 0.0, idl]])
    mt, mp = ro.far_moments(mid, mom, k, th, ph)
    exact = (-np.sin(th) * idl * np.exp(1j * k * h * np.cos(th)))[:, None]
    assert np.abs(mt - exact).max() < 1e-15 and np.abs(mp).max() < 1e-15
    mt, mp = ro.far_moments(mid, mom, k, th, ph, "pec", 0.0)
    exact = (-2 * np.sin(th) * idl * np.cos(k * h * np.cos(th)))[:, None]
    assert np.abs(mt - exact).max() < 1e-14 and np.abs(mp).max() < 1e-15
    eps_r, sigma, f = 13.0, 0.005, 3e8
    rho_h, rho_v = ro.image_coeffs(eps_r, sigma, f, 0.0, 0.0, 1.0)
    e = eps_r - 1j * sigma / (2 * np.pi * f * ro.EPS0)
    n = np.sqrt(e)
    assert abs(rho_v - (n - 1) / (n + 1)) < 1e-15 and abs(rho_h + rho_v) < 1e-15


def test_near_field_radiation_zone_limit():
    """Far from the structure the mixed-potential near field tends to
    -j eta k exp(-jkr) / (4 pi r) M_perp: the two readouts are one physics."""
    d = golden("readout_array2x2")
    k = float(d["k"])
    els = (d["mid8"], d["moment8"][..., 0], d["nodes8"], d["delta8"][..., 0])
    th, ph = np.array([0.7, 1.3]), np.array([0.4, 2.9])
    mt, mp = ro.far_moments(d["mid8"], els[1], k, th, ph)
    r = 2e5
    for i, t in enumerate(th):
        for j, q in enumerate(ph):
            u = np.array([np.sin(t) * np.cos(q), np.sin(t) * np.sin(q), np.cos(t)])
            E = ro.element_fields(r * u[None], els, k, float(d["radius"]), False)[0]
            pref = -1j * ro.ETA0 * k * np.exp(-1j * k * r) / (4 * np.pi * r)
            th_hat = np.array([np.cos(t) * np.cos(q), np.cos(t) * np.sin(q), -np.sin(t)])
            ph_hat = np.array([-np.sin(q), np.cos(q), 0.0])
            scale = abs(pref) * max(abs(mt[i, j]), abs(mp[i, j]))
            assert abs(E @ th_hat - pref * mt[i, j]) < 1e-4 * scale
            assert abs(E @ ph_hat - pref * mp[i, j]) < 1e-4 * scale


@pytest.mark.parametrize("case", sorted(CASES))
def test_readout_tables_reproduce_reference_walk(case):
    """The host tables of the GPU kernels, applied on the CPU: taps x
    coefficients at the samples, elements and knot steps, against the
    reference's walk for every column at subdiv 1 and 8."""
    from momwire_amd.readout import ReadoutMesh
    d = golden(case)
    s = BSplineSolver(**CASES[case]())
    X = d["coeffs"]
    for sub in (1, 8):
        rm = ReadoutMesh(s, sub)
        assert np.array_equal(rm.mid, d[f"mid{sub}"]) and np.array_equal(rm.nodes, d[f"nodes{sub}"])
        Xp = np.vstack([X, np.zeros((1, X.shape[1]))])  # index -1 -> zero row
        cur = np.einsum("st,stc->sc", rm.tap_v, Xp[rm.tap_i])
        el = 0.5 * (cur[rm.el_tab[:, 1]] + cur[rm.el_tab[:, 0]])
        mom = el[:, None, :] * rm.dl[:, :, None]
        assert rel_err(mom, d[f"moment{sub}"]) < 1e-13
        elp = np.vstack([el, np.zeros((1, X.shape[1]))])
        delta = elp[rm.node_tab[:, 0]] - elp[rm.node_tab[:, 1]]
        assert rel_err(delta, d[f"delta{sub}"]) < 1e-12
