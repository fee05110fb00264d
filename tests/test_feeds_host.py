"""Spread gap feeds on the CPU: the segment gap and the cos^2-smoothed
source (reference bspline.py:2806-2942) -- the oracle against the
reference's Y and coefficients (tests/golden/feed_*.npz), the host port
columns against the oracle's, and what the definitions fix: each column sums
to 1 over a wire's bases (partition of unity: the spread drive has unit
integral), the smoothing tends to the delta gap as alpha -> 0, and the two
spread forms refuse to combine."""

import numpy as np
import pytest

from conftest import golden, rel_err
from momwire_amd import _mesh, workloads as W
from momwire_amd.bspline import BSplineSolver
from oracle.bspline_oracle import OracleSolver


def feed_case(name):
    if name == "feed_segment":
        kw = W.junction_tee("junction_port")
        kw["feeds"] = [(0, None, 1.0), (2, 0.1, 1.0), (1, 0.13, 1.0)]
        kw["feed_model"] = "segment"
    elif name == "feed_smooth_tee":
        kw = W.junction_tee("feed")
        kw["feeds"] = [(0, None, 1.0), (1, 0.21, 1.0)]
        kw["feed_smoothing_factor"] = 2.5
    else:
        kw = W.straight_dipole()
        kw["feed_smoothing_factor"] = 3.0
    return kw


CASES = ("feed_segment", "feed_smooth_tee", "feed_smooth_dipole")


@pytest.mark.parametrize("case", CASES)
def test_feed_oracle_matches_reference(case):
    d = golden(case)
    y, X = OracleSolver(**feed_case(case)).compute_port_solution()
    assert rel_err(y, d["y"]) < 1e-10
    assert rel_err(X, d["coeffs"]) < 1e-10


@pytest.mark.parametrize("case", CASES)
def test_host_feed_columns_match_oracle(case):
    kw = feed_case(case)
    s = BSplineSolver(**kw)
    cols, _con = s._port_plan()
    o = OracleSolver(**kw)
    geom = o.geometry()
    B, _k = o.port_columns(geom, o.basis(geom))
    for f in range(len(kw["feeds"])):
        dense = np.zeros(B.shape[0])
        idx, val = cols[f]
        dense[idx] = val
        assert np.abs(dense - B[:, f].real).max() < 1e-14
        # unit drive: a feed away from free ends sums to 1 over the bases
        assert abs(dense.sum() - 1.0) < 1e-12


def test_smoothing_tends_to_delta_gap():
    mesh = BSplineSolver(**W.straight_dipole())._build()
    i0, v0 = _mesh.feed_column(mesh, 0, None)
    i1, v1 = _mesh.feed_column(mesh, 0, None, smoothing=1e-6)
    d0, d1 = np.zeros(mesh.n_basis), np.zeros(mesh.n_basis)
    d0[i0], d1[i1] = v0, v1
    assert np.abs(d0 - d1).max() < 1e-9


def test_spread_feed_options_validate():
    kw = W.straight_dipole()
    with pytest.raises(ValueError):
        BSplineSolver(**kw, feed_model="segment", feed_smoothing_factor=2.0)
    with pytest.raises(ValueError):
        BSplineSolver(**kw, feed_model="cell")
