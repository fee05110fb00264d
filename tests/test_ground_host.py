"""Host-side ground logic (no GPU): constructor validation, the grid layout and
bucketing against the reference's grids, and the oracle against the
reference's compiled engine."""

import numpy as np
import pytest

from conftest import golden
from momwire_amd import _ground, workloads as W
from momwire_amd.bspline import BSplineSolver
from oracle import sommerfeld_oracle as so


def _kw(**over):
    kw = W.grounded_array(1, 1, 21, params=(0.3, 12.0, 0.01, 10e6))
    kw.update(over)
    return kw


def test_ground_options_validation():
    with pytest.raises(ValueError):
        BSplineSolver(**_kw(ground_z=None))  # ground_eps without ground_z
    with pytest.raises(ValueError):
        BSplineSolver(**_kw(ground_model="bogus"))
    with pytest.raises(ValueError):
        BSplineSolver(**_kw(ground_eps=None))  # sommerfeld needs eps
    with pytest.raises(ValueError):
        BSplineSolver(**_kw(ground_model="refl-coef", ground_phi_mode="bogus"))
    s = BSplineSolver(**_kw(ground_model="refl-coef"))
    assert s.ground_phi_mode == "normal"
    with pytest.raises(NotImplementedError):
        BSplineSolver(**_kw(n_qp_sommerfeld=9))
    with pytest.raises(ValueError):
        BSplineSolver(**_kw(n_qp_sommerfeld=0))
    lam = _kw()["wavelength"]
    with pytest.raises(ValueError):
        BSplineSolver(**_kw(ground_z=0.5 * lam))  # wire below the plane
    with pytest.raises(ValueError, match="edge lying"):
        BSplineSolver(**_kw(ground_z=0.3 * lam))  # horizontal wire lying in the plane
    s = BSplineSolver(**_kw())
    assert s.ground_model == "sommerfeld" and s.ground_z == 0.0


@pytest.mark.parametrize("case", ["ground_dipole", "ground_tee", "c5_ground4x4"])
def test_grid_layout_matches_reference(case):
    d = golden(case)
    kw = {"ground_dipole": _kw(), "ground_tee": W.grounded_tee(),
          "c5_ground4x4": W.grounded_array(4, 4, 21, seed=4)}[case]
    s = BSplineSolver(**kw)
    mesh = s._build()
    r1 = _ground.max_image_distance(mesh.seg[:, 0:3], mesh.seg[:, 3:6], 0.0, rows=7)
    assert abs(r1 - float(d["r1_max"])) < 1e-12 * r1
    eps_t = _ground.eps_tilde(s.ground_eps, s.omega, s.eps)
    assert abs(eps_t - complex(d["eps_t"])) < 1e-14 * abs(eps_t)
    lam = 2 * np.pi / s.k
    r1b = _ground.r1_bucket_wl(min(r1 / lam, _ground.R1_CAP_WL))
    eb = _ground.eps_bucket(eps_t)
    assert eb == complex(d["grid_eps"])
    regions, scal = _ground.grid_layout(eb, _ground.K2_REF, r1b)
    sc = lam
    np.testing.assert_allclose(np.array(scal) * sc, d["grid_scalars"][:4], rtol=1e-13,
                               atol=1e-300)
    shapes, R, T = _ground.grid_nodes(regions, scal[3])
    lay = d["grid_layout"]
    for i, ((r0, _r1, dr, th0, _t1, dth), (n_r, n_th)) in enumerate(zip(regions, shapes)):
        assert (n_r, n_th) == (int(lay[i, 2]), int(lay[i, 5]))
        np.testing.assert_allclose([r0 * sc, dr * sc, np.radians(th0), np.radians(dth)],
                                   lay[i, [0, 1, 3, 4]], rtol=1e-13, atol=1e-300)
    assert R.size == sum(a * b for a, b in shapes) == T.size


def test_far_grid_layout_matches_reference():
    d = golden("somm_grid_far")
    regions, scal = _ground.grid_layout(complex(15, -20), _ground.K2_REF,
                                        _ground.r1_bucket_wl(6.0))
    shapes, _R, _T = _ground.grid_nodes(regions, scal[3])
    lay = d["grid_layout"]
    assert len(regions) == 6
    for i, (n_r, n_th) in enumerate(shapes):
        assert (n_r, n_th) == (int(lay[i, 2]), int(lay[i, 5]))


def test_bucketing_matches_oracle():
    rng = np.random.default_rng(0)
    for _ in range(200):
        e = complex(rng.uniform(1, 40), -np.exp(rng.uniform(-3, 6)))
        assert _ground.eps_bucket(e) == so.eps_bucket(e)
        x = float(np.exp(rng.uniform(-3, 3)))
        assert _ground.r1_bucket_wl(x) == so.r1_bucket_wl(x)
    assert _ground.eps_bucket(complex(5, 0)) == complex(5, 0)


def test_oracle_engine_matches_reference_compiled_engine(ref_accel):
    """The reference's C++ engine (its own Bessel series, built from its
    sources into oracle/_ref) against the oracle's scipy-based restatement."""
    g = golden("somm_engine")
    for eps in g["grounds"]:
        ref = ref_accel.somm_six_integrals_batch(complex(eps), 2 * np.pi, g["rho"], g["h"],
                                                 1e-9, 0, 0)
        mine = np.array([so.six_integrals(eps, 2 * np.pi, r, h, 1e-9)
                         for r, h in zip(g["rho"], g["h"])])
        err = np.abs(mine - ref).max(1) / np.abs(ref).max(1)
        assert err.max() < 1e-8


# -- reflection-coefficient ground (reference _ground_refl.py) --------------
REFL = {
    "refl_dipole": lambda: W.grounded_array(1, 1, 21, params=(0.3, 12.0, 0.01, 10e6),
                                            ground_model="refl-coef"),
    "refl_tee": lambda: W.grounded_tee(ground_model="refl-coef", phi_mode="rho_v"),
    "refl_tee_blend": lambda: W.grounded_tee(eps_r=25.0, sigma=2e-2, freq_hz=7e6,
                                             ground_model="refl-coef", phi_mode="blend"),
}


@pytest.mark.parametrize("case", sorted(REFL))
def test_refl_oracle_matches_reference(case):
    """The oracle's Fresnel weight tables, Z and Y against the reference's own
    (tests/golden/refl_*.npz, written by the reference's solver)."""
    from oracle import refl_oracle as ro
    from oracle.bspline_oracle import EPS0, OracleSolver
    d = golden(case)
    kw = REFL[case]()
    o = OracleSolver(**kw)
    geom = o.geometry()
    eps_t = ro.eps_tilde(kw["ground_eps"], o.omega, EPS0)
    assert eps_t == complex(d["eps_t"])
    cen = 0.5 * (geom["seg_l"] + geom["seg_r"])
    w_A, w_Phi = ro.weights(cen, geom["tangents"], kw["ground_z"], eps_t, o.ground_phi_mode)
    assert np.abs(w_A - d["w_A"]).max() < 1e-15
    assert np.abs(np.broadcast_to(w_Phi, w_A.shape) - d["w_Phi"]).max() < 1e-15
    Z = o.Z(geom)
    assert np.abs(Z - d["Z"]).max() < 1e-10 * np.abs(d["Z"]).max()
    y, coeffs = o.compute_port_solution()
    assert np.abs(y - d["y"]).max() < 1e-10 * np.abs(d["y"]).max()
    assert np.abs(coeffs - d["coeffs"]).max() < 1e-10 * np.abs(d["coeffs"]).max()


def test_refl_weights_closed_forms():
    """What the mathematics fixes: at normal incidence rho_v = -rho_h =
    (sqrt(eps) - 1)/(sqrt(eps) + 1); at grazing incidence rho_v = rho_h = -1;
    as eps grows every weight tends to the PEC image's (w_A = t . M t,
    w_Phi = 1 for every phi mode); each phi mode's (c0, c1) against its
    definition; and the host's coefficients against the oracle's."""
    from oracle import refl_oracle as ro
    eps = 13.0 - 7.5j
    rv, rh = ro.fresnel_rho(eps, np.array([1.0, 0.0]))
    r = np.sqrt(eps)
    assert abs(rv[0] - (r - 1) / (r + 1)) < 1e-15 and abs(rh[0] + rv[0]) < 1e-15
    assert abs(rv[1] + 1) < 1e-15 and abs(rh[1] + 1) < 1e-15
    rng = np.random.default_rng(3)
    cen = rng.uniform(0.2, 2.0, (12, 3))
    tan = rng.standard_normal((12, 3))
    tan /= np.linalg.norm(tan, axis=1, keepdims=True)
    td = tan @ (tan * [1.0, 1.0, -1.0]).T
    for mode in ro.PHI_MODES:
        w_A, w_Phi = ro.weights(cen, tan, 0.0, 1e14 - 1e13j, mode)
        assert np.abs(w_A - td).max() < 1e-6
        assert np.abs(w_Phi - 1).max() < 1e-6
        c0, c1 = _ground.phi_coefficients(mode, eps)
        assert (c0, c1) == ro.phi_coeffs(mode, eps)
    rv, _rh = ro.fresnel_rho(eps, 0.4)
    c0, c1 = ro.phi_coeffs("blend", eps)
    assert abs(c0 + c1 * rv - 0.5 * (rv + (eps - 1) / (eps + 1))) < 1e-15
    # a pair and its transpose see the same weights (the symmetric fill)
    w_A, w_Phi = ro.weights(cen, tan, 0.0, eps, "rho_v")
    assert np.abs(w_A - w_A.T).max() < 1e-15 and np.abs(w_Phi - w_Phi.T).max() < 1e-15
