"""CPU checks of the host-side tables the HIP kernels read (no GPU needed).

The fill kernel never sees supp/polys directly: it reads per-segment band
polynomials, per-basis (tile-local segment, band slot) pairs and tile segment
lists.  Here the kernel's data flow is replayed in numpy on the oracle's
moments -- band blocks B_ij, then the per-entry gather -- and the resulting Z
must equal the oracle's Galerkin assembly.  The basis polynomials of the host
set-up (Cox-de Boor on polynomials) are checked against the oracle's (the
reference's design-matrix + Vandermonde route).
"""

import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT, rel_err
from momwire_amd import _lib, workloads as W
from momwire_amd._mesh import build_mesh, bspline_values, feed_column
from momwire_amd.bspline import BSplineSolver
from oracle.bspline_oracle import OracleSolver, EPS0, MU0

CASES = [
    ("c1", W.straight_dipole()),
    ("array2x2", W.dipole_array(2, 2, 11, seed=5)),
    ("tee", W.junction_tee("feed")),
    ("tee_jport", W.junction_tee("junction_port")),
]


def _host(kw):
    s = BSplineSolver(**kw)
    return s, s._build()


@pytest.mark.parametrize("name,kw", CASES, ids=[c[0] for c in CASES])
def test_basis_polys_match_oracle(name, kw):
    s, mesh = _host(kw)
    o = OracleSolver(**kw)
    supp, polys = o.basis(o.geometry())[:2]
    assert mesh.n_basis == supp.shape[0]
    for m in range(mesh.n_basis):
        for a in range(3):
            if mesh.supp[m, a] >= 0:
                assert mesh.supp[m, a] == supp[m, a]
                assert np.allclose(mesh.polys[m, a], polys[m, a], rtol=1e-12, atol=1e-12 * np.abs(polys[m]).max())
            else:
                assert not np.any(polys[m, a])


@pytest.mark.parametrize("name,kw", CASES, ids=[c[0] for c in CASES])
def test_band_gather_replay_equals_oracle_Z(name, kw):
    s, mesh = _host(kw)
    o = OracleSolver(**kw)
    geom = o.geometry()
    J = o.J_blocks(geom, o.k)
    Zref = o.Z(geom)
    wmu, inv_weps = o.omega * MU0, 1.0 / (o.omega * EPS0)
    N = mesh.n_seg
    t = mesh.seg[:, 6:9]
    td = t @ t.T
    C = mesh.band_poly                      # [seg][slot][p]
    Cd = C[:, :, 1:] * np.array([1.0, 2.0])  # derivative coefficients
    A = np.einsum("iap,pqij,jbq->ijab", C, J, C)
    Phi = np.einsum("iap,pqij,jbq->ijab", Cd, J[:2, :2], Cd)
    B = 1j * wmu * td[:, :, None, None] * A + Phi / (1j * o.omega * EPS0) * (o.omega * EPS0) * inv_weps
    Z = np.zeros((mesh.n_basis, mesh.n_basis), dtype=complex)
    for ti in range(len(mesh.tile_nseg)):
        rs = mesh.tile_seg[ti, :mesh.tile_nseg[ti]]
        for tj in range(len(mesh.tile_nseg)):
            cs = mesh.tile_seg[tj, :mesh.tile_nseg[tj]]
            for m in range(mesh.tile_b0[ti], mesh.tile_b0[ti + 1]):
                for n in range(mesh.tile_b0[tj], mesh.tile_b0[tj + 1]):
                    z = 0
                    for a in range(3):
                        if mesh.bw_loc[m, a] < 0:
                            continue
                        for b in range(3):
                            if mesh.bw_loc[n, b] < 0:
                                continue
                            i, j = rs[mesh.bw_loc[m, a]], cs[mesh.bw_loc[n, b]]
                            z += B[i, j, mesh.bw_slot[m, a], mesh.bw_slot[n, b]]
                    Z[m, n] = z
    assert rel_err(Z, Zref) < 1e-13
    assert mesh.tile_nseg.max() <= _lib.TSMAX
    assert np.all(np.diff(mesh.tile_b0) <= _lib.TB)


def test_feed_column_is_basis_value():
    kw = W.straight_dipole()
    s, mesh = _host(kw)
    o = OracleSolver(**kw)
    geom = o.geometry()
    B, _ = o.port_columns(geom, o.basis(geom))
    idx, val = feed_column(mesh, 0, None)
    dense = np.zeros(mesh.n_basis)
    dense[idx] = val
    assert np.allclose(dense, B[:, 0].real, atol=1e-15)


def test_bspline_values_partition_of_unity():
    knots = np.concatenate([[0, 0], np.linspace(0, 2.0, 8), [2.0, 2.0]])
    for x in np.linspace(0, 2.0, 37):
        v = bspline_values(knots, 2, len(knots) - 3, x)
        assert abs(v.sum() - 1.0) < 1e-14
        assert np.all(v >= -1e-15)


def test_c_abi_exports_every_declared_symbol():
    """The built library exports every function include/momwire_hip.h declares."""
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libmomwire_hip.so not built")
    hdr = open(os.path.join(ROOT, "include", "momwire_hip.h")).read()
    names = set(re.findall(r"^\s*(?:int|int64_t|double|const char\*)\s+(mw_\w+)\s*\(", hdr, re.M))
    assert names, "no declarations parsed"
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for nm in names:
        assert hasattr(lib, nm), nm
    assert names <= set(_lib.SIGNATURES)
    lib2 = _lib.load()
    assert lib2.mw_abi_version() == 1
    assert lib2.mw_getrf_workspace_bytes(2624) > 0


@pytest.mark.parametrize("degree", [1, 2])
def test_current_slopes_match_spline_derivative(degree):
    """current_slopes against scipy's exact B-spline derivative (the
    reference's route, bspline.py:4523-4584) at the knots, and against a
    central difference of currents_at_knots at element centres."""
    from scipy.interpolate import BSpline

    from momwire_amd.bspline import BSplineSolver
    kw = W.junction_tee("junction_port")
    kw["degree"] = degree
    s = BSplineSolver(**kw)
    m = s._build()
    c = np.random.default_rng(degree).standard_normal(m.n_basis) * (1 + 0.5j)
    slopes = s.current_slopes(c)
    for w, pw in enumerate(m.per_wire):
        cb = np.zeros(pw["nb_w"], dtype=complex)
        for kidx, (j, *_r) in enumerate(pw["kept"]):
            cb[j] = c[pw["l2g"][kidx]]
        ref = BSpline(pw["knots"], cb, degree, extrapolate=False).derivative(1)(
            np.clip(pw["arc"], pw["knots"][0], pw["knots"][-1]))
        assert np.abs(slopes[w] - ref).max() <= 1e-13 * np.abs(ref).max()
        mids = 0.5 * (pw["arc"][1:] + pw["arc"][:-1])
        eps = 1e-6 * pw["arc"][-1]
        hi = s.currents_at_knots(c, [mids + eps if i == w else [] for i in range(len(m.per_wire))])
        lo = s.currents_at_knots(c, [mids - eps if i == w else [] for i in range(len(m.per_wire))])
        fd = (hi[w] - lo[w]) / (2 * eps)
        mid_sl = s.current_slopes(c, [mids if i == w else [] for i in range(len(m.per_wire))])[w]
        assert np.abs(fd - mid_sl).max() <= 1e-6 * np.abs(mid_sl).max()


def test_sysv_workspace_holds_the_block_inverses():
    """mw_sysv_workspace_bytes covers the LU workspace plus, for every size
    the symmetric path factors (past 4,096 unknowns, 512-column outer
    blocks), two 512 x 512 complex slots per outer block and the 512 x 1024
    staging block; a workspace sized for the LU alone would let the
    block-inverse kernels write past its end."""
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libmomwire_hip.so not built")
    lib = _lib.load()
    for n in (100, 1024, 1025, 2000, 2624, 4096, 4097, 15360, 41984):
        getrf = lib.mw_getrf_workspace_bytes(n)
        bz = 512 if n > 4096 else 0
        need = getrf + ((2 * -(-n // bz) * 512 * 512 * 16 + 512 * 1024 * 16) if bz else 0)
        assert lib.mw_sysv_workspace_bytes(n) >= need, n


def test_sweep_plan_rejects_nan_pivot_bound():
    """The sweep plan refuses a NaN pivot threshold before touching anything
    (so no GPU is needed)."""
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libmomwire_hip.so not built")
    lib = _lib.load()
    h = ctypes.c_void_p()
    rc = lib.mw_sweep_plan_create(1, None, None, 1, None, 0, float("nan"), ctypes.byref(h))
    assert rc == _lib.MW_EINVAL and b"NaN" in lib.mw_last_error() and not h.value


def test_noalloc_gesv_refuses_oversize_on_the_host():
    """mw_zgesv_noalloc's size limit is checked before any HIP call: a system
    past its static workspace (n > 16384) and bad sizes are refused on a
    machine with no GPU at all."""
    from momwire_amd import _lib

    lib = _lib.load()
    assert lib.mw_zgesv_noalloc(16385, 1, None, 16386, None, None, None, None) == _lib.MW_ERANGE
    assert b"n <= 16384" in lib.mw_last_error()
    assert lib.mw_zgesv_noalloc(0, 1, None, 1, None, None, None, None) == _lib.MW_EINVAL


def test_batched_sweep_refuses_ground():
    """The batched fill is free space only: a grounded structure is refused
    before anything touches the GPU, never solved without its ground."""
    from momwire_amd.sweep import BatchedSweep

    with pytest.raises(NotImplementedError, match="ground"):
        BatchedSweep([W.straight_dipole(), W.grounded_monopole("pec")], [1.0])
