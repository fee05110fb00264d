"""GPU parity of the Sommerfeld ground path: the HIP engine, the device grid,
the interpolation and the grounded fill, against the reference's values
(tests/golden) and the oracle.

Bars: the engine's six integrals within 1e-8 of the reference's at the same
quadrature tolerance (the adaptive and tail stop tests take the same
decisions unless a panel sits within ~1e-12 of a threshold; typically the
agreement is ~1e-12); grid tables and interpolated values within 1e-8;
grounded Z within 1e-10 and grounded Y and coeffs within 1e-10 of the
reference's solver.
"""

import ctypes

import numpy as np
import pytest

from conftest import golden, rel_err
from momwire_amd import _ground, _lib, workloads as W
from momwire_amd.bspline import BSplineSolver
from oracle import sommerfeld_oracle as so
from oracle.bspline_oracle import OracleSolver

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
K2 = 2 * np.pi


def _say(*a):
    print("[gpu-test]", *a, flush=True)


def _st():
    return torch.cuda.current_stream().cuda_stream


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _six(eps, rho, h, rtol, form=0):
    lib = _lib.load()
    out = torch.empty((len(rho), 6), dtype=torch.complex128, device="cuda")
    dr, dh = _dev(np.asarray(rho, float)), _dev(np.asarray(h, float))
    _lib.check(lib.mw_somm_six_integrals(eps.real, eps.imag, K2, dr.data_ptr(), dh.data_ptr(),
                                         len(rho), rtol, form, out.data_ptr(), None, _st()), "six")
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_engine_six_integrals_match_reference():
    g = golden("somm_engine")
    rho, h, sel = g["rho"], g["h"], g["six9_idx"]
    for i, eps in enumerate(g["grounds"]):
        _say("six integrals, ground", eps)
        six = _six(eps, rho, h, 1e-6)
        ref = g[f"six_{i}"]
        err = np.abs(six - ref).max(1) / np.abs(ref).max(1)
        _say("  max node error", err.max())
        assert err.max() < 1e-8
        six9 = _six(eps, rho[sel], h[sel], 1e-9)
        assert rel_err(six9, g[f"six9_{i}"]) < 1e-9


def test_engine_forced_forms_match_reference():
    g = golden("somm_engine")
    for i, eps in enumerate(g["grounds"]):
        _say("forced forms, ground", eps)
        vj = _six(eps, g["form_rho"], g["form_h"], 1e-9, form=1)
        vh = _six(eps, g["form_rho"], g["form_h"], 1e-9, form=2)
        assert rel_err(vj, g[f"formJ_{i}"]) < 1e-9
        assert rel_err(vh, g[f"formH_{i}"]) < 1e-9


def test_engine_surfaces_and_limits():
    g = golden("somm_engine")
    lib = _lib.load()
    r1 = np.concatenate([np.hypot(g["rho"], g["h"]), np.zeros(len(g["lim_theta"]))])
    th = np.concatenate([np.arctan2(g["h"], g["rho"]), g["lim_theta"]])
    dR, dT = _dev(r1), _dev(th)
    for i, eps in enumerate(g["grounds"]):
        _say("surfaces, ground", eps)
        out = torch.empty((r1.size, 4), dtype=torch.complex128, device="cuda")
        _lib.check(lib.mw_somm_surfaces(eps.real, eps.imag, K2, K2 * so.C_LIGHT, so.MU0_EXACT,
                                        dR.data_ptr(), dT.data_ptr(), r1.size, 1e-6,
                                        out.data_ptr(), None, _st()), "surfaces")
        torch.cuda.synchronize()
        s = out.cpu().numpy().T
        n = g["rho"].size
        assert rel_err(s[:, :n], g[f"surf_{i}"]) < 1e-8
        assert rel_err(s[:, n:], g[f"lim_{i}"]) < 1e-13


def _grid_case(case):
    if case == "ground_tee":
        return W.grounded_tee()
    return W.grounded_array(1, 1, 21, params=(0.3, 12.0, 0.01, 10e6))


@pytest.mark.parametrize("case", ["ground_dipole", "ground_tee"])
def test_device_grid_and_eval_match_reference(case):
    _say("device grid", case)
    d = golden(case)
    s = BSplineSolver(**_grid_case(case))
    s._build()
    _ground.clear_cache()
    gs = s._ground_grid(s.k, s.omega)
    m = s._grid_master
    np.testing.assert_allclose([gs.r1_max, gs.r_break, gs.r_near], d["grid_scalars"][:3],
                               rtol=1e-13)
    vals = m.vals.cpu().numpy() * gs.value_scale
    for i, (n_r, n_th) in enumerate(m.shapes):
        ref = d[f"grid_vals{i}"]  # (4, n_r, n_th)
        got = vals[m.off[i]:m.off[i] + n_r * n_th].reshape(n_r, n_th, 4).transpose(2, 0, 1)
        assert got.shape == ref.shape
        assert rel_err(got, ref) < 1e-8, i
    lib = _lib.load()
    r1, th = d["grid_eval_r1"], d["grid_eval_th"]
    dR, dT = _dev(r1), _dev(th)
    out = torch.empty((r1.size, 4), dtype=torch.complex128, device="cuda")
    _lib.check(lib.mw_somm_eval(ctypes.byref(gs), dR.data_ptr(), dT.data_ptr(), r1.size,
                                out.data_ptr(), _st()), "eval")
    torch.cuda.synchronize()
    assert rel_err(out.cpu().numpy().T, d["grid_eval"]) < 1e-8


def test_interpolation_matches_oracle_on_reference_tables():
    """mw_somm_eval on the reference's own tables: isolates the stencil,
    the region select and the clamps from the engine."""
    _say("eval on reference tables")
    d = golden("ground_dipole")
    lay = d["grid_layout"]
    tabs = [d[f"grid_vals{i}"] for i in range(lay.shape[0])]
    flat = np.concatenate([t.transpose(1, 2, 0).reshape(-1, 4) for t in tabs])
    dv = _dev(flat)
    g = _ground.MwSommGrid()
    g.n_reg = lay.shape[0]
    g.r1_max, g.r_break, g.r_near = d["grid_scalars"][:3]
    g.th_split = np.radians(30.0)
    g.value_scale = 1.0
    off = 0
    for i in range(g.n_reg):
        g.r0[i], g.dr[i], g.th0[i], g.dth[i] = lay[i, 0], lay[i, 1], lay[i, 3], lay[i, 4]
        g.n_r[i], g.n_th[i] = int(lay[i, 2]), int(lay[i, 5])
        g.off[i] = off
        off += g.n_r[i] * g.n_th[i]
    g.vals = dv.data_ptr()
    rng = np.random.default_rng(3)
    r1 = np.concatenate([d["grid_eval_r1"], rng.uniform(0, 1.3 * g.r1_max, 50)])  # clamp too
    th = np.concatenate([d["grid_eval_th"], rng.uniform(-0.1, 1.7, 50)])
    dR, dT = _dev(r1), _dev(th)
    out = torch.empty((r1.size, 4), dtype=torch.complex128, device="cuda")
    lib = _lib.load()
    _lib.check(lib.mw_somm_eval(ctypes.byref(g), dR.data_ptr(), dT.data_ptr(), r1.size,
                                out.data_ptr(), _st()), "eval")
    torch.cuda.synchronize()
    got = out.cpu().numpy().T
    assert rel_err(got[:, :300], d["grid_eval"]) < 1e-13
    grid = so.Grid([(lay[i, 0], 0, lay[i, 1], np.degrees(lay[i, 3]), 0, np.degrees(lay[i, 4]))
                    for i in range(g.n_reg)], tuple(d["grid_scalars"][:4]), tabs, 0, 0, 0, 0)
    assert rel_err(got, grid.eval(r1, th)) < 1e-13


def _device_Z(s):
    """Run only the fill (free space + ground) and return Z."""
    lib = _lib.load()
    D = s._device()
    st = _st()
    _lib.check(lib.mw_same_edge_table(D.mw, s.k, st), "same-edge")
    s._fill(D, s.k, s.omega, D.Z.data_ptr(), D.ld, st)
    torch.cuda.synchronize()
    return D.Z[:, :D.nb].cpu().numpy()


@pytest.mark.parametrize("case", ["ground_dipole", "ground_tee"])
def test_grounded_Z_matches_reference(case):
    _say("grounded Z", case)
    d = golden(case)
    s = BSplineSolver(**_grid_case(case))
    Z = _device_Z(s)
    assert rel_err(Z, d["Z"]) < 1e-10


def test_image_only_and_pec_match_oracle():
    """grid = NULL: Z_free - C2 * Z_image (C2 = 1 is the PEC ground)."""
    _say("PEC image")
    kw = _grid_case("ground_tee")
    kw.pop("ground_eps")
    kw.pop("ground_model")
    s = BSplineSolver(**kw)
    Z = _device_Z(s)
    o = OracleSolver(**kw)
    geom = o.geometry()
    supp, polys = o.basis(geom)[:2]
    Zf = o.assemble(o.J_blocks(geom, o.k), supp, polys, geom)
    mir = np.array([1.0, 1.0, -1.0])
    src = (geom["seg_l"] * mir, geom["seg_r"] * mir)
    Ji = o.full_moments(geom, o.seg_radius(geom), o.k, src=src)
    Zi = o.assemble(Ji, supp, polys, geom, td=geom["tangents"] @ (geom["tangents"] * mir).T)
    assert rel_err(Z, Zf - Zi) < 1e-11


@pytest.mark.parametrize("case", ["ground_dipole", "ground_tee", "c5_ground4x4"])
def test_grounded_port_solution_matches_reference(case):
    _say("grounded port solution", case)
    d = golden(case)
    kw = W.grounded_array(4, 4, 21, seed=4) if case == "c5_ground4x4" else _grid_case(case)
    sol = BSplineSolver(**kw).compute_port_solution()
    _say("  Y err", rel_err(sol.y, d["y"]))
    assert rel_err(sol.y, d["y"]) < 1e-10
    assert rel_err(sol.coeffs, d["coeffs"]) < 1e-10


REFL = {
    "refl_dipole": lambda: W.grounded_array(1, 1, 21, params=(0.3, 12.0, 0.01, 10e6),
                                            ground_model="refl-coef"),
    "refl_tee": lambda: W.grounded_tee(ground_model="refl-coef", phi_mode="rho_v"),
    "refl_tee_blend": lambda: W.grounded_tee(eps_r=25.0, sigma=2e-2, freq_hz=7e6,
                                             ground_model="refl-coef", phi_mode="blend"),
    "refl_array4x4": lambda: W.grounded_array(4, 4, 21, seed=4, ground_model="refl-coef"),
}


@pytest.mark.parametrize("case", ["refl_dipole", "refl_tee", "refl_tee_blend"])
def test_refl_Z_matches_reference(case):
    """Reflection-coefficient ground: the Fresnel weights formed in the fill
    kernel, the whole grounded Z against the reference's."""
    _say("refl-coef Z", case)
    d = golden(case)
    Z = _device_Z(BSplineSolver(**REFL[case]()))
    assert rel_err(Z, d["Z"]) < 1e-10


@pytest.mark.parametrize("case", sorted(REFL))
def test_refl_port_solution_matches_reference(case):
    _say("refl-coef port solution", case)
    d = golden(case)
    for precision in ("fp64", "mixed"):
        sol = BSplineSolver(**REFL[case](), precision=precision).compute_port_solution()
        _say(f"  {precision}: Y err", rel_err(sol.y, d["y"]))
        assert rel_err(sol.y, d["y"]) < (1e-10 if precision == "fp64" else 1e-8)
        if precision == "fp64":
            assert rel_err(sol.coeffs, d["coeffs"]) < 1e-10


def test_refl_pec_limit_matches_pec_image():
    """As eps grows the Fresnel-weighted image tends to the PEC image (the
    same geometry with no ground_eps)."""
    _say("refl-coef PEC limit")
    kw = REFL["refl_tee"]()
    kw["ground_eps"] = complex(1e16, -1e15)
    Zr = _device_Z(BSplineSolver(**kw))
    kw.pop("ground_eps")
    kw.pop("ground_model")
    kw.pop("ground_phi_mode")
    Zp = _device_Z(BSplineSolver(**kw))
    assert rel_err(Zr, Zp) < 1e-7


def test_grounded_graph_replay_matches_eager():
    _say("grounded graph replay")
    s = BSplineSolver(**W.grounded_array(2, 2, 21, seed=6))
    y_eager = s.compute_port_solution().y
    for _ in range(3):
        Y, _X = s.solve_graph()
        torch.cuda.synchronize()
        assert rel_err(Y.cpu().numpy(), y_eager) < 1e-13
        s.compute_port_solution()


def test_engine_refuses_pageable_cancel_flag():
    """The engine's kernels poll the cancel flag on the device, so an
    ordinary host address is refused on the host before any launch (the
    call asks for zero nodes, so nothing would be launched either way)."""
    _say("engine cancel: pageable host flag refused")
    lib = _lib.load()
    flag = np.zeros(1, dtype=np.int32)
    out = torch.empty((1, 4), dtype=torch.complex128, device="cuda")
    r = _dev(np.ones(1))
    for call in (
        lambda: lib.mw_somm_surfaces(12.0, -1.0, K2, K2 * 299792458.0, 4e-7 * np.pi,
                                     r.data_ptr(), r.data_ptr(), 0, 1e-6, out.data_ptr(),
                                     flag.ctypes.data, _st()),
        lambda: lib.mw_somm_six_integrals(12.0, -1.0, K2, r.data_ptr(), r.data_ptr(), 0, 1e-6, 0,
                                          out.data_ptr(), flag.ctypes.data, _st()),
    ):
        rc = call()
        assert rc == _lib.MW_EINVAL, rc
        assert b"device-accessible" in lib.mw_last_error()
    torch.cuda.synchronize()


def test_grounded_solve_with_cancel_token():
    """A cold grounded solve carrying a (pinned, device-readable) token:
    untripped it matches the reference's Y; tripped before the solve the
    engine stops and the solve raises SolveAborted, caching no grid."""
    from momwire_amd.bspline import CancelToken, SolveAborted
    _say("engine cancel: grounded dipole with a token")
    d = golden("ground_dipole")
    tok = CancelToken()
    _ground.clear_cache()
    s = BSplineSolver(**_grid_case("ground_dipole"), cancel=tok)
    sol = s.compute_port_solution()
    assert rel_err(sol.y, d["y"]) < 1e-10
    assert rel_err(sol.coeffs, d["coeffs"]) < 1e-10
    _say("engine cancel: pre-tripped token")
    _ground.clear_cache()
    tok.cancel()
    s2 = BSplineSolver(**_grid_case("ground_dipole"), cancel=tok)
    with pytest.raises(SolveAborted):
        s2.compute_port_solution()
    assert not _ground._MASTERS
    tok.reset()
    assert rel_err(s2.compute_port_solution().y, d["y"]) < 1e-10
