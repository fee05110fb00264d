"""The Sommerfeld ground remainder at points in space on the CPU: the
oracle's vector form against the reference's own reflected_field_at /
reflected_h_field_at outputs (tests/golden/fieldpt_dipole.npz, written by the
reference), against the fill's projected remainder (which the reference's
Galerkin goldens pin), the central-difference curl against the free-space
dipole's exact H, and the host API's refusals."""

import ctypes
import math
import os

import numpy as np
import pytest

from conftest import golden, rel_err
from oracle import readout_oracle as ro
from oracle import sommerfeld_oracle as so


def _golden_grid(d):
    pts, mid = d["points"], d["src_mid"]
    ex = np.concatenate([pts, mid])
    gz = float(d["ground_z"])
    r1 = so.max_image_distance(ex, ex, gz)
    return so.get_grid(complex(d["eps_t"]), float(d["k"]), r1, float(d["omega"]), so.MU0_EXACT)


def _scale_err(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max()


def test_oracle_matches_reference_field_point():
    d = golden("fieldpt_dipole")
    grid = _golden_grid(d)
    gz, k, om = float(d["ground_z"]), float(d["k"]), float(d["omega"])
    e = so.reflected_field_at(d["points"], d["src_mid"], d["src_moment"], gz, k, grid)
    assert _scale_err(e, d["e"]) < 1e-13
    step = 1e-3 * 2.0 * math.pi / k
    h = so.curl_fd(lambda q: so.reflected_field_at(q, d["src_mid"], d["src_moment"], gz, k, grid),
                   d["points"], step) / (-1j * om * so.MU0_EXACT)
    assert _scale_err(h, d["h"]) < 1e-11


def _synthetic_grid(rng):
    """A grid with random surface values: the projection identity below is
    algebra, true for any values away from rho = 0."""
    regions, scalars = so.grid_layout(complex(12.0, -3.0), 2.0 * math.pi, 1.1)
    vals = []
    for i, reg in enumerate(regions):
        rr, _tt = so.Grid.region_nodes(reg, i, scalars[3])
        vals.append(rng.standard_normal((4,) + rr.shape) + 1j * rng.standard_normal((4,) + rr.shape))
    return so.Grid(regions, scalars, vals, 2.0 * math.pi, 2.0 * math.pi * so.C_LIGHT,
                   so.MU0_EXACT, complex(12.0, -3.0))


def test_vector_form_projects_to_the_fill_remainder():
    """t_i . E(p = t_j) equals remainder_proj(t_obs = t_i, t_src = t_j) for
    Cartesian unit t_i, t_j: the nine projections fix the dyad, and the vector
    form's linearity in the complex moment does the rest (checked on a
    complex moment as the sum of its three components)."""
    rng = np.random.default_rng(5)
    grid = _synthetic_grid(rng)
    k, gz = 2.0 * math.pi, 0.0
    pts = np.column_stack([rng.uniform(-0.4, 0.4, 9), rng.uniform(-0.4, 0.4, 9),
                           rng.uniform(0.05, 0.4, 9)])
    src = np.column_stack([rng.uniform(-0.4, 0.4, 7), rng.uniform(-0.4, 0.4, 7),
                           rng.uniform(0.05, 0.3, 7)])
    eye = np.eye(3)
    for j in range(3):
        mom = np.zeros((7, 3), dtype=complex)
        mom[:, j] = 1.0
        e = so.reflected_field_at(pts, src, mom, gz, k, grid)
        for i in range(3):
            ref = so.remainder_proj(pts, np.tile(eye[i], (9, 1)), src, np.tile(eye[j], (7, 1)), gz,
                                    k, grid).sum(axis=1)
            assert np.abs(e[:, i] - ref).max() <= 1e-13 * np.abs(ref).max(), (i, j)
    p = rng.standard_normal((7, 3)) + 1j * rng.standard_normal((7, 3))
    parts = sum(so.reflected_field_at(pts, src, p * eye[j], gz, k, grid) for j in range(3))
    assert rel_err(so.reflected_field_at(pts, src, p, gz, k, grid), parts) < 1e-14


def test_curl_fd_on_the_free_dipole():
    """curl E / (-j omega mu) of a few free-space elements equals their
    exact H (radius 0), within the central difference's (k step)^2 / 6."""
    rng = np.random.default_rng(8)
    k = 2.0 * math.pi
    mid = rng.uniform(-0.3, 0.3, (5, 3))
    mom = rng.standard_normal((5, 3)) + 1j * rng.standard_normal((5, 3))
    nodes = rng.uniform(-0.3, 0.3, (4, 3))
    delta = rng.standard_normal(4) + 1j * rng.standard_normal(4)
    pts = rng.uniform(0.8, 1.5, (6, 3))
    els = (mid, mom, nodes, delta)
    step = 1e-3 * 2.0 * math.pi / k
    curl = so.curl_fd(lambda q: ro.element_fields(q, els, k, 0.0, False), pts, step)
    h = ro.element_fields(pts, els, k, 0.0, True)
    # omega mu = eta k
    assert rel_err(curl / (-1j * ro.ETA0 * k), h) < 1e-4


@pytest.mark.parametrize("key", [("below", "below"), ("below", "above"), ("above", "below"),
                                 ("air", "above")])
def test_regimes_other_than_above_above_refuse(key):
    from momwire_amd import readout
    with pytest.raises(ValueError, match="regime"):
        readout.reflected_field_at(np.zeros((1, 3)), np.zeros((1, 3)), np.ones((1, 3)),
                                   complex(10, -1), 0.0, 1.0, 1.0, source_medium=key[0],
                                   observer_medium=key[1])
    with pytest.raises(ValueError, match="regime"):
        readout.reflected_h_field_at(np.zeros((1, 3)), np.zeros((1, 3)), np.ones((1, 3)),
                                     complex(10, -1), 0.0, 1.0, 1.0, source_medium=key[0],
                                     observer_medium=key[1])


def test_c_abi_ground_fields_rejects_bad_arguments():
    """Checked before anything reaches the device (no GPU needed)."""
    from momwire_amd import _ground, _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libmomwire_hip.so not built")
    lib = _lib.load()
    rc = lib.mw_ground_fields(None, 0.0, 1.0, 1, 64, 1, 64, 64, 3, 1, 0, 64, 3, None)
    assert rc == _lib.MW_EINVAL
    g = _ground.MwSommGrid()
    g.n_reg = 4
    g.r1_max, g.r_break, g.r_near = 1.0, 0.2, 1.0
    g.th_split = math.radians(30.0)
    g.value_scale = 1.0
    for i in range(4):
        g.dr[i], g.dth[i], g.n_r[i], g.n_th[i] = 0.01, 0.1, 8, 8
        g.off[i] = 64 * i
    g.vals = 64
    rc = lib.mw_ground_fields(ctypes.byref(g), 0.0, 1.0, 1, 64, 1, 64, 64, 2, 1, 0, 64, 3, None)
    assert rc == _lib.MW_EINVAL and b"3 p" in lib.mw_last_error()
    rc = lib.mw_ground_fields(ctypes.byref(g), 0.0, 0.0, 1, 64, 1, 64, 64, 3, 1, 0, 64, 3, None)
    assert rc == _lib.MW_EINVAL and b"positive" in lib.mw_last_error()
