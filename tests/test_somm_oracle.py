"""The Sommerfeld-ground oracle against the reference's own values and against
what the mathematics fixes (CPU only)."""

import numpy as np
import pytest

from conftest import golden, rel_err
from momwire_amd import workloads
from oracle import sommerfeld_oracle as so
from oracle.bspline_oracle import EPS0, OracleSolver

K2 = 2 * np.pi
OMEGA = K2 * so.C_LIGHT


def _scale_err(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


@pytest.fixture(scope="module")
def engine():
    return golden("somm_engine")


def test_six_integrals_match_reference(engine):
    """Production tolerance (rtol 1e-6) and tight tolerance (1e-9) sets."""
    rho, h = engine["rho"], engine["h"]
    sel = engine["six9_idx"]
    for g, eps in enumerate(engine["grounds"]):
        six = np.array([so.six_integrals(eps, K2, r, hh, 1e-6) for r, hh in zip(rho, h)])
        ref = engine[f"six_{g}"]
        # per node, relative to that node's largest integral
        err = np.abs(six - ref).max(1) / np.abs(ref).max(1)
        assert err.max() < 1e-12, (eps, err.max())
        six9 = np.array([so.six_integrals(eps, K2, rho[i], h[i], 1e-9) for i in sel])
        assert _scale_err(six9, engine[f"six9_{g}"]) < 1e-12


def test_bessel_and_hankel_contours_agree(engine):
    """Both contours converge in the overlap band; each matches the reference's."""
    for g, eps in enumerate(engine["grounds"]):
        for i, (r, hh) in enumerate(zip(engine["form_rho"], engine["form_h"])):
            vj = so.six_integrals(eps, K2, r, hh, 1e-9, form="J")
            vh = so.six_integrals(eps, K2, r, hh, 1e-9, form="H")
            assert np.abs(vj - engine[f"formJ_{g}"][i]).max() < 1e-12 * np.abs(vj).max()
            assert np.abs(vh - engine[f"formH_{g}"][i]).max() < 1e-12 * np.abs(vh).max()
            assert np.abs(vj - vh).max() < 1e-6 * np.abs(vj).max()


@pytest.mark.parametrize("bessel,rho,h", [(True, 0.1, 0.3), (True, 0.0, 0.5),
                                          (False, 0.8, 0.1), (False, 2.0, 0.4)])
def test_sommerfeld_identity(bessel, rho, h):
    num, exact = so.identity_check(K2, rho, h, bessel)
    assert abs(num - exact) < 1e-7 * abs(exact)


def test_laplacian_identity_pointwise():
    """J0'' + J0'/x + J0 = 0 makes integrand 0 + 3 + lam^2 * 4 vanish."""
    lam = np.array([0.3 + 0.2j, 2.0 + 0.5j, 7.0 - 0.3j, 15.0 + 0.0j])
    k1 = K2 * np.sqrt(10 - 2j)
    for bessel in (True, False):
        f = so.integrands(lam, 0.4, 0.2, np.conj(k1) if k1.imag > 0 else k1, K2, bessel)
        resid = f[0] + f[3] + lam ** 2 * f[4]
        assert np.abs(resid).max() < 1e-12 * np.abs(f[0]).max()


def test_surfaces_and_limits_match_reference(engine):
    r1 = np.hypot(engine["rho"], engine["h"])
    th = np.arctan2(engine["h"], engine["rho"])
    for g, eps in enumerate(engine["grounds"]):
        s = so.surfaces(eps, K2, r1, th, 1e-6, OMEGA, so.MU0_EXACT)
        assert _scale_err(s, engine[f"surf_{g}"]) < 1e-12
        lim = so.limits_r1_zero(eps, K2, engine["lim_theta"], OMEGA, so.MU0_EXACT)
        assert _scale_err(lim, engine[f"lim_{g}"]) < 1e-14


def test_limits_are_the_small_r1_values():
    """Eqs 169-172 are the R1 -> 0 limits of the direct evaluation (steep band,
    where the interface layer is resolved at R1 ~ 1e-4 lambda)."""
    eps = complex(15, -20)
    th = np.radians([50.0, 70.0, 89.0])
    lim = so.limits_r1_zero(eps, K2, th, OMEGA, so.MU0_EXACT)
    near = so.surfaces(eps, K2, np.full(3, 2e-5), th, 1e-9, OMEGA, so.MU0_EXACT)
    assert _scale_err(near, lim) < 2e-2


def test_image_coefficient_limits():
    assert abs(so.image_coefficient(1.0 + 0j)) == 0.0
    assert abs(so.image_coefficient(1e12 - 1e12j) - 1.0) < 1e-11


@pytest.mark.parametrize("case", ["ground_dipole", "ground_tee"])
def test_grid_tables_and_eval_match_reference(case):
    d = golden(case)
    kw = workloads.grounded_tee() if case == "ground_tee" else \
        workloads.grounded_array(1, 1, 21, params=(0.3, 12.0, 0.01, 10e6))
    o = OracleSolver(**kw)
    geom = o.geometry()
    eps_t = so.eps_tilde(o.ground_eps, o.omega, EPS0)
    assert abs(eps_t - complex(d["eps_t"])) < 1e-14 * abs(eps_t)
    grid = o.ground_grid(geom, eps_t)
    np.testing.assert_allclose([grid.r1_max, grid.r_break, grid.r_near, grid.r0_fill],
                               d["grid_scalars"][:4], rtol=1e-14)
    for i, r in enumerate(grid.regions):
        ref = d[f"grid_vals{i}"]
        assert r["vals"].shape == ref.shape
        assert _scale_err(r["vals"], ref) < 1e-12
    ev = grid.eval(d["grid_eval_r1"], d["grid_eval_th"])
    assert _scale_err(ev, d["grid_eval"]) < 1e-12


def test_far_zone_grid_matches_reference():
    d = golden("somm_grid_far")
    g = so.get_grid(complex(15, -20), K2, 6.0, OMEGA, so.MU0_EXACT)
    assert len(g.regions) == 6
    lay = d["grid_layout"]
    for i, r in enumerate(g.regions):
        assert (r["n_r"], r["n_th"]) == (int(lay[i, 2]), int(lay[i, 5]))
        np.testing.assert_allclose([r["r0"], r["dr"], r["th0"], r["dth"]], lay[i, [0, 1, 3, 4]],
                                   rtol=1e-14, atol=1e-300)
    ev = g.eval(d["grid_eval_r1"], d["grid_eval_th"])
    assert _scale_err(ev, d["grid_eval"]) < 1e-12


@pytest.mark.parametrize("case", ["ground_dipole", "ground_tee"])
def test_grounded_port_solution_matches_reference(case):
    d = golden(case)
    kw = workloads.grounded_tee() if case == "ground_tee" else \
        workloads.grounded_array(1, 1, 21, params=(0.3, 12.0, 0.01, 10e6))
    o = OracleSolver(**kw)
    geom = o.geometry()
    basis = o.basis(geom)
    grid = o.ground_grid(geom, so.eps_tilde(o.ground_eps, o.omega, EPS0))
    Q = so.remainder_block(geom, basis[0], basis[1], o.ground_z, o.k, grid)
    assert _scale_err(Q, d["Q"]) < 1e-12
    Z = o.Z(geom, basis)
    assert _scale_err(Z, d["Z"]) < 1e-10
    y, X = o.compute_port_solution()
    assert rel_err(y, d["y"]) < 1e-10
    assert rel_err(X, d["coeffs"]) < 1e-10


def test_remainder_is_small_but_not_negligible():
    """Q changes Y at a level well above the parity tolerance, so the parity
    tests above do exercise it."""
    d = golden("ground_dipole")
    kw = workloads.grounded_array(1, 1, 21, params=(0.3, 12.0, 0.01, 10e6))
    o = OracleSolver(**kw)
    geom = o.geometry()
    basis = o.basis(geom)
    Z = o.Z(geom, basis, remainder=False)
    B, _ = o.port_columns(geom, basis)
    y0 = B.T @ np.linalg.solve(Z, B)
    assert rel_err(y0, d["y"]) > 1e-6


def test_c5_array_matches_reference():
    d = golden("c5_ground4x4")
    o = OracleSolver(**workloads.grounded_array(4, 4, 21, seed=4))
    y, X = o.compute_port_solution()
    assert rel_err(y, d["y"]) < 1e-10
    assert rel_err(X, d["coeffs"]) < 1e-10
