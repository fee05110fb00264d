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
    with pytest.raises(NotImplementedError):
        BSplineSolver(**_kw(ground_model="refl-coef"))
    with pytest.raises(NotImplementedError):
        BSplineSolver(**_kw(n_qp_sommerfeld=5))
    lam = _kw()["wavelength"]
    with pytest.raises(ValueError):
        BSplineSolver(**_kw(ground_z=0.5 * lam))  # wire below the plane
    with pytest.raises(NotImplementedError):
        BSplineSolver(**_kw(ground_z=0.3 * lam))  # wire on the plane
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
