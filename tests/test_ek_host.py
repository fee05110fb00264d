"""The extended thin-wire kernel on the CPU: the oracle against the
reference's own labels, static moments, Z and Y (tests/golden/ek_*.npz,
written by the reference's solver), what the mathematics fixes (the static
closed forms against mpmath, the static + remainder split of the full
kernel, the a -> 0 limit), and the host's coaxial labels against the
reference's."""

import numpy as np
import pytest

from conftest import golden, rel_err
from momwire_amd import _mesh, workloads as W
from momwire_amd.bspline import BSplineSolver
from oracle import ek_oracle as eo
from oracle.bspline_oracle import OracleSolver
from oracle.static_moments import static_moment, static_moments_uniform

CASES = ("dipole", "collinear", "tee", "pec_vertical", "refl_dipole", "somm_tee")
GROUNDED = ("pec_vertical", "refl_dipole", "somm_tee")


@pytest.mark.parametrize("case", CASES)
def test_ek_oracle_matches_reference(case):
    d = golden(f"ek_{case}")
    o = OracleSolver(**W.ek_structure(case))
    geom = o.geometry()
    gi, gj = o.ek_labels(geom)
    assert np.array_equal(gi, d["labels"]) and np.array_equal(gj, d["labels"])
    if case in GROUNDED:
        gi, gj = o.ek_labels(geom, mirror=True)
        assert np.array_equal(gi, d["labels_img_obs"])
        assert np.array_equal(gj, d["labels_img_src"])
    # same-edge static moments of the first edge, extended and reduced
    pw = geom["per_wire"][0]
    N = d["static_ek"].shape[-1]
    h, a = float(pw["h"][0]), float(o.radius[0])
    assert np.allclose(pw["h"][:N], h, rtol=1e-14, atol=0)
    # Near pairs (|i - j| <= 2, the closed forms) agree to rounding; far
    # pairs carry the reference's closed-form cancellation, which the
    # Gauss-Legendre rule here avoids (pinned against mpmath below).
    near = np.abs(np.arange(N)[None, :] - np.arange(N)[:, None]) <= 2
    for key, ek in (("static_ek", True), ("static", False)):
        S = static_moments_uniform(h, a, N, 2, ek=ek)
        r = np.abs(S - d[key]) / np.abs(S)
        assert r[:, :, near].max() < 1e-12, key
        assert rel_err(S, d[key]) < 1e-9, key
    Z = o.Z(geom)
    assert rel_err(Z, d["Z"]) < 1e-10
    y, coeffs = o.compute_port_solution()
    assert rel_err(y, d["y"]) < 1e-10
    assert rel_err(coeffs, d["coeffs"]) < 1e-10


def test_ek_changes_the_answer():
    """The fat-wire structures are chosen so the extended factor shows: the
    reference's extended Y differs from its reduced Y well above the parity
    tolerance (else the tests above would pin nothing)."""
    d = golden("ek_dipole")
    kw = W.ek_structure("dipole")
    kw["extended_kernel"] = False
    y0, _c = OracleSolver(**kw).compute_port_solution()
    assert rel_err(y0, d["y"]) > 1e-4


def test_ek_static_moments_vs_mpmath():
    """The extended static kernel 1/R + a^2 (a^2 - 2 x^2) / (4 R^5) against
    mpmath quadrature, near (closed forms) and far (Gauss-Legendre) pairs."""
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 20
    h, a = 0.02, 4e-3
    for p, q in ((0, 0), (1, 2), (2, 1), (2, 2)):
        for d in (0, 1, -1, 3):
            H, A, c = mp.mpf(h), mp.mpf(a), d * mp.mpf(h)

            def inner(u, p=p, q=q, c=c, H=H, A=A):
                def f(v):
                    x = u - v - c
                    R = mp.sqrt(x * x + A * A)
                    return u**p * v**q * (1 / R + A * A * (A * A - 2 * x * x) / (4 * R**5))
                pts = [0, u - c, H] if 0 < u - c < H else [0, H]
                return mp.quad(f, pts)

            exact = float(mp.quad(inner, [0, H]))
            got = static_moment(p, q, d, h, a, ek=True)
            assert abs(got - exact) <= 1e-12 * abs(exact), (p, q, d)


def test_ek_kernel_split_and_limits():
    """exp(-jkR) fac / (4 pi R) = remainder + fac_static / (4 pi R) with
    fac_static = 1 + 3 T1 - T2 (the static kernel the closed forms integrate);
    the remainder is finite as R -> a; every factor tends to 1 as a -> 0."""
    k = 2 * np.pi
    a = 4e-3
    R = np.sqrt(np.linspace(0.0, 0.3, 50) ** 2 + a * a)
    t1, t2 = a**4 / (4 * R**4), a * a / (2 * R * R)
    full = np.exp(-1j * k * R) * eo.factor(R, a, k) / (4 * np.pi * R)
    static = (1 + 3 * t1 - t2) / (4 * np.pi * R)
    assert np.abs(eo.reg_kernel(R, a, k) + static - full).max() < 1e-12 * np.abs(full).max()
    assert np.isfinite(eo.reg_kernel(np.array([a]), a, k)).all()
    assert np.abs(eo.factor(R, 1e-9, k) - 1).max() < 1e-12


def test_axis_group_rules():
    """Collinear equal-radius segments share a label (either orientation);
    a bend, an offset beyond tol * a, or another radius starts a new one;
    the first representative wins."""
    z = np.array([0.0, 0.0, 1.0])
    x = np.array([1.0, 0.0, 0.0])
    L = np.array([[0, 0, 0], [0, 0, 1], [0, 0, 2], [0, 0, 3], [1e-3, 0, 4], [0, 0, 5]], float)
    T = np.array([z, z, -z, x, z, z])
    R = L + T
    a = np.array([1e-3, 1e-3, 1e-3, 1e-3, 1e-3, 1.1e-3])
    lab = eo.axis_groups(L, R, T, a)
    assert list(lab) == [0, 0, 0, 1, 2, 3]
    assert np.array_equal(_mesh.coaxial_labels(L, R, T, a), lab)


@pytest.mark.parametrize("case", CASES)
def test_host_labels_match_reference(case):
    """The host tables the GPU kernels read: labels of the real segments and
    of the image block, bit for bit the reference's."""
    d = golden(f"ek_{case}")
    s = BSplineSolver(**W.ek_structure(case))
    lab, lo, ls = _mesh.ek_tables(s._build(), s.ground_z)
    assert np.array_equal(lab, d["labels"])
    if case in GROUNDED:
        assert np.array_equal(lo, d["labels_img_obs"])
        assert np.array_equal(ls, d["labels_img_src"])
    else:
        assert lo is None and ls is None


def test_c_abi_rejects_ek_without_labels():
    """The C-ABI refuses an extended-kernel mesh whose label tables are
    missing, before anything reaches the device (so no GPU is needed)."""
    import ctypes
    import os

    from momwire_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libmomwire_hip.so not built")
    lib = _lib.load()
    m = _lib.MwMesh()
    m.n_seg, m.n_edge, m.n_basis, m.n_tile, m.n_se, m.degree, m.n_qp = 4, 1, 5, 1, 1, 2, 4
    for name in ("seg", "seg_edge", "edge", "edge_i", "se_edge", "se_tab", "band_poly",
                 "bw_loc", "bw_slot", "tile_b0", "tile_seg", "tile_nseg", "gl"):
        setattr(m, name, 64)  # never dereferenced: the checks run first
    m.ek = 1
    rc = lib.mw_same_edge_table(ctypes.byref(m), 1.0, None)
    assert rc == _lib.MW_EINVAL and b"coaxial labels" in lib.mw_last_error()
    m.ek_label = 64
    rc = lib.mw_fill_ground(ctypes.byref(m), 1.0, 1.0, 1.0, 1.0, 0.0, 1.0, 0.0, None, 3, 64, 5,
                            None, 0, 1, None)
    assert rc == _lib.MW_EINVAL and b"image block" in lib.mw_last_error()
    rc = lib.mw_fill_ground_refl(ctypes.byref(m), 1.0, 1.0, 1.0, 1.0, 0.0, 4.0, -1.0, 0.3, 0.0,
                                 0.0, 0.0, 64, 5, None, 0, 1, None)
    assert rc == _lib.MW_EINVAL and b"image block" in lib.mw_last_error()
