"""GPU parity of the HIP path against the oracle and the reference's goldens.

All calls go through the C ABI (momwire_amd._lib).  Bars: Y and coeffs within
1e-10 relative of the reference's own solver (tests/golden); Z within 1e-11
relative of the oracle; linear-algebra kernels within a few ulps times the
condition number of numpy on the same inputs.
"""

import numpy as np
import pytest

from conftest import golden, rel_err
from momwire_amd import _lib, workloads as W
from momwire_amd.bspline import BSplineSolver

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _say(*a):
    print("[gpu-test]", *a, flush=True)


def _stream():
    return torch.cuda.current_stream().cuda_stream


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("dtype", ["z", "c"])
@pytest.mark.parametrize("m,n,k", [(64, 32, 8), (100, 70, 33), (257, 129, 64), (31, 5, 3)])
def test_gemm_sub(dtype, m, n, k):
    _say("gemm", dtype, m, n, k)
    rng = np.random.default_rng(m * 7 + n)
    ct = np.complex128 if dtype == "z" else np.complex64
    A = (rng.standard_normal((m, k)) + 1j * rng.standard_normal((m, k))).astype(ct)
    B = (rng.standard_normal((k, n)) + 1j * rng.standard_normal((k, n))).astype(ct)
    C = (rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))).astype(ct)
    dA, dB, dC = _dev(A), _dev(B), _dev(C)
    lib = _lib.load()
    fn = lib.mw_zgemm_sub if dtype == "z" else lib.mw_cgemm_sub
    _lib.check(fn(m, n, k, dA.data_ptr(), k, dB.data_ptr(), n, dC.data_ptr(), n, _stream()), "gemm")
    torch.cuda.synchronize()
    ref = C.astype(np.complex128) - A.astype(np.complex128) @ B.astype(np.complex128)
    tol = 1e-13 if dtype == "z" else 2e-5
    assert rel_err(dC.cpu().numpy(), ref) < tol


def _lu_solve(M, RHS, dtype="z"):
    n, p = RHS.shape
    lib = _lib.load()
    ct = torch.complex128 if dtype == "z" else torch.complex64
    dM = torch.from_numpy(M).to(ct).cuda()
    dX = torch.from_numpy(RHS).to(ct).cuda()
    ipiv = torch.empty(n, dtype=torch.int32, device="cuda")
    work = torch.empty(int(lib.mw_getrf_workspace_bytes(n)), dtype=torch.uint8, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    f, s = (lib.mw_zgetrf, lib.mw_zgetrs) if dtype == "z" else (lib.mw_cgetrf, lib.mw_cgetrs)
    _lib.check(f(n, dM.data_ptr(), n, ipiv.data_ptr(), work.data_ptr(), info.data_ptr(), None, _stream()), "getrf")
    _lib.check(s(n, p, dM.data_ptr(), n, ipiv.data_ptr(), work.data_ptr(), dX.data_ptr(), p, _stream()), "getrs")
    torch.cuda.synchronize()
    return dX.cpu().numpy().astype(np.complex128), int(info.cpu()[0])


@pytest.mark.parametrize("n,p", [(1, 1), (5, 2), (33, 3), (64, 64), (100, 7), (257, 40), (1000, 17), (2113, 64)])
def test_lu_solve_random(n, p):
    _say("lu random", n, p)
    rng = np.random.default_rng(n)
    M = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    RHS = rng.standard_normal((n, p)) + 1j * rng.standard_normal((n, p))
    X, info = _lu_solve(M, RHS)
    assert info == 0
    ref = np.linalg.solve(M, RHS)
    cond = np.linalg.cond(M)
    assert rel_err(X, ref) < 1e-15 * cond * 50
    # backward error no worse than LAPACK's on the same system
    res = np.abs(M @ X - RHS).max()
    res_ref = np.abs(M @ ref - RHS).max()
    assert res <= 20 * res_ref + 1e-13


@pytest.mark.parametrize("dtype", ["z", "c"])
@pytest.mark.parametrize("n,p", [(1100, 300), (1024, 257), (33, 260), (4500, 300), (4160, 260)])
def test_lu_solve_many_rhs(dtype, n, p):
    """More than 256 right-hand sides: the back substitution forms each panel
    pair's solution once and updates the rows above with a plain GEMM (odd and
    even panel counts, a ragged last panel); past n = 4096 both triangular
    solves run two levels (rows inside a 512-column block per panel, the rest
    once per block)."""
    _say("lu many rhs", dtype, n, p)
    rng = np.random.default_rng(n + p)
    M = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    RHS = rng.standard_normal((n, p)) + 1j * rng.standard_normal((n, p))
    X, info = _lu_solve(M, RHS, dtype)
    assert info == 0
    ref = np.linalg.solve(M, RHS)
    eps = 1.1e-16 if dtype == "z" else 6e-8
    assert rel_err(X, ref) < eps * np.linalg.cond(M) * 200


@pytest.mark.parametrize("dtype", ["z", "c"])
@pytest.mark.parametrize("n,p", [(1, 1), (40, 3), (300, 64), (1111, 9), (4500, 70), (5000, 300)])
def test_gesv_augmented(dtype, n, p):
    """mw_?gesv: the right-hand sides ride through the factorisation as extra
    columns of [A | B]; the answer must match LAPACK's."""
    _say("gesv", dtype, n, p)
    rng = np.random.default_rng(n + p)
    M = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + (3 * np.eye(n) if dtype == "c" else 0)
    RHS = rng.standard_normal((n, p)) + 1j * rng.standard_normal((n, p))
    lib = _lib.load()
    ct = torch.complex128 if dtype == "z" else torch.complex64
    aug = torch.from_numpy(np.hstack([M, RHS])).to(ct).cuda()
    ipiv = torch.empty(n, dtype=torch.int32, device="cuda")
    work = torch.empty(int(lib.mw_getrf_workspace_bytes(n)), dtype=torch.uint8, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    fn = lib.mw_zgesv if dtype == "z" else lib.mw_cgesv
    _lib.check(fn(n, p, aug.data_ptr(), n + p, ipiv.data_ptr(), work.data_ptr(), info.data_ptr(), None, _stream()), "gesv")
    torch.cuda.synchronize()
    X = aug[:, n:].cpu().numpy().astype(np.complex128)
    ref = np.linalg.solve(M, RHS)
    eps = 1.1e-16 if dtype == "z" else 6e-8
    tol = eps * np.linalg.cond(M) * 200
    assert int(info.cpu()[0]) == 0
    assert rel_err(X, ref) < tol


def _gesv_dev(fn, n, p, aug, ipiv, work, info, *extra):
    _lib.check(fn(n, p, aug.data_ptr(), n + p, ipiv.data_ptr(), work.data_ptr(), info.data_ptr(),
                  None, *extra, _stream()), "gesv")


def test_gesv_two_level_graph_replay():
    """Past n = 4096 (two-level blocking: 512-column outer blocks), the solve
    captured in a graph must replay to the eager answer bit for bit, replay
    after replay (the matrix re-uploaded between replays, synchronisations in
    between); the random matrix takes the tournament fallback on many
    panels."""
    _say("gesv two-level graph replay")
    n, p = 4200, 40
    rng = np.random.default_rng(77)
    M = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    RHS = rng.standard_normal((n, p)) + 1j * rng.standard_normal((n, p))
    lib = _lib.load()
    host = torch.from_numpy(np.hstack([M, RHS])).cuda()
    aug = host.clone()
    ipiv = torch.empty(n, dtype=torch.int32, device="cuda")
    work = torch.empty(int(lib.mw_getrf_workspace_bytes(n)), dtype=torch.uint8, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    _gesv_dev(lib.mw_zgesv, n, p, aug, ipiv, work, info)
    torch.cuda.synchronize()
    X_eager = aug[:, n:].cpu().numpy()
    assert int(info.cpu()[0]) == 0
    ref = np.linalg.solve(M, RHS)
    assert rel_err(X_eager, ref) < 1.1e-16 * np.linalg.cond(M) * 200
    aug.copy_(host)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        _gesv_dev(lib.mw_zgesv, n, p, aug, ipiv, work, info)
    for _ in range(3):
        aug.copy_(host)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(aug[:, n:].cpu().numpy(), X_eager)
        assert int(info.cpu()[0]) == 0


def test_lu_needs_pivoting():
    _say("lu pivoting")
    n = 300
    rng = np.random.default_rng(3)
    M = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    M[np.arange(n), np.arange(n)] = 1e-14  # tiny diagonal: unpivoted LU would fail
    RHS = rng.standard_normal((n, 3)) + 0j
    X, info = _lu_solve(M, RHS)
    assert info == 0
    assert rel_err(M @ X, RHS) < 1e-11


def test_lu_singular_reports_info():
    _say("lu singular")
    n = 70
    M = np.zeros((n, n), dtype=complex)
    M[:, :40] = np.random.default_rng(1).standard_normal((n, 40))
    _X, info = _lu_solve(M, np.ones((n, 1), dtype=complex))
    assert info > 0


def test_lu_nan_input_completes():
    """Regression: NaN magnitudes in a panel must still yield an in-range pivot
    row (a NaN-only column once left the pivot search without a candidate)."""
    _say("lu nan input")
    n = 150
    rng = np.random.default_rng(5)
    M = rng.standard_normal((n, n)) + 0j
    M[40:, 33:] = np.nan
    X, _info = _lu_solve(M, np.ones((n, 2), dtype=complex))
    assert X.shape == (n, 2)
    # the library is still healthy afterwards
    M2 = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    R2 = rng.standard_normal((n, 1)) + 0j
    X2, info2 = _lu_solve(M2, R2)
    assert info2 == 0 and rel_err(M2 @ X2, R2) < 1e-10


@pytest.mark.parametrize("name,kw", [("array2x2", W.dipole_array(2, 2, 11, seed=5)),
                                     ("tee_jport", W.junction_tee("junction_port"))])
def test_threshold_and_full_pivoting_agree(name, kw):
    """Threshold pivoting (diagonal kept while multipliers stay <= 8) and
    tournament partial pivoting on every panel give the same port solution."""
    _say("pivot modes", name)
    lib = _lib.load()
    g = golden(name)
    default = lib.mw_get_pivot_threshold()
    try:
        ys = []
        for bound in (default, 0.0):
            _lib.check(lib.mw_set_pivot_threshold(bound), "threshold")
            sol = BSplineSolver(**kw).compute_port_solution()
            assert rel_err(sol.y, g["y"]) < 1e-10
            assert rel_err(sol.coeffs, g["coeffs"]) < 1e-10
            ys.append(sol.y)
        assert rel_err(ys[0], ys[1]) < 1e-12
    finally:
        lib.mw_set_pivot_threshold(default)


def test_lu_solve_fp32():
    _say("lu fp32")
    n = 700
    rng = np.random.default_rng(9)
    M = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + 30 * np.eye(n)
    RHS = rng.standard_normal((n, 5)) + 0j
    X, info = _lu_solve(M, RHS, "c")
    assert info == 0
    assert rel_err(X, np.linalg.solve(M, RHS)) < 1e-4


CASES = [
    ("c1_dipole", W.straight_dipole()),
    ("array2x2", W.dipole_array(2, 2, 11, seed=5)),
    ("tee_feed", W.junction_tee("feed")),
    ("tee_jport", W.junction_tee("junction_port")),
    ("tee_nodegap", W.junction_tee("node_gap")),
]


@pytest.mark.parametrize("name,kw", CASES, ids=[c[0] for c in CASES])
def test_port_solution_matches_reference_golden(name, kw):
    _say("port solution", name)
    g = golden(name)
    s = BSplineSolver(**kw)
    sol = s.compute_port_solution()
    assert sol.port_currents is sol.y
    assert rel_err(sol.y, g["y"]) < 1e-10
    assert rel_err(sol.coeffs, g["coeffs"]) < 1e-10
    if "knot_currents" in g:
        kc = np.concatenate(s.currents_at_knots(sol.coeffs[:, 0]))
        assert rel_err(kc, g["knot_currents"]) < 1e-10
    if "sweep_y" in g:
        ys = s.compute_y_matrix_swept(g["sweep_k"])
        assert rel_err(ys, g["sweep_y"]) < 1e-10


def test_fill_matches_oracle_Z():
    from oracle.bspline_oracle import OracleSolver
    for name, kw in CASES[:3]:
        _say("fill Z", name)
        s = BSplineSolver(**kw)
        s._solve_device()  # leaves the factorised Z; refill below
        D = s._dev
        lib = _lib.load()
        _lib.check(lib.mw_same_edge_table(D.mw, s.k, _stream()), "se")
        for sym in (1, 0):
            D.Z.zero_()
            pairs = D.keep["pairs"]
            _lib.check(lib.mw_fill_z(D.mw, s.k, s.omega, s.eps, s.mu, D.Z.data_ptr(), D.ld,
                                     pairs.data_ptr(), int(pairs.shape[0]), sym, _stream()), "fill")
            torch.cuda.synchronize()
            Z = D.Z[:, :D.nb].cpu().numpy()
            o = OracleSolver(**kw)
            assert rel_err(Z, o.Z()) < 1e-11, (name, sym)


def test_c2_array_8x8_matches_reference_golden():
    _say("port solution c2 8x8")
    g = golden("c2_array8x8")
    s = BSplineSolver(**W.dipole_array(8, 8, 41, seed=1))
    sol = s.compute_port_solution()
    assert rel_err(sol.y, g["y"]) < 1e-10
    assert rel_err(sol.coeffs[:, g["coeffs_cols"]], g["coeffs"]) < 1e-10
    assert rel_err(sol.coeffs.sum(axis=1), g["coeffs_colsum"]) < 1e-10
    assert rel_err(sol.y, sol.y.T) < 1e-12  # reciprocity


def test_c2_every_coefficient_column():
    """The golden holds 4 of C2's 64 coefficient columns and the column sums;
    every column is pinned here by the system itself: Z (pinned to the
    oracle at 1e-11 by the fill tests) times the GPU's coefficients gives
    each port's delta-gap column back to a backward error of a few ulps
    times n, and Y = B^T X reads the same coefficients out."""
    _say("c2 every coefficient column")
    from momwire_amd import _lib

    s = BSplineSolver(**W.dipole_array(8, 8, 41, seed=1))
    sol = s.compute_port_solution()
    D = s._device()
    lib = _lib.load()
    sym = int(s._mesh.symmetric)
    pairs = D.keep["pairs"]
    _lib.check(lib.mw_same_edge_table(D.mw, s.k, _stream()), "same-edge table")
    _lib.check(lib.mw_fill_z(D.mw, s.k, s.omega, s.eps, s.mu, D.Z.data_ptr(), D.ld,
                             pairs.data_ptr(), int(pairs.shape[0]), sym, _stream()), "fill")
    torch.cuda.synchronize()
    Z = D.Z[:, :D.nb].cpu().numpy()
    X = sol.coeffs
    assert X.shape == (2624, 64)
    B = np.zeros_like(X)
    cols = D.col_idx.cpu().numpy().reshape(D.ncols, -1)
    vals = D.col_val.cpu().numpy().reshape(D.ncols, -1)
    for j in range(64):
        for r, v in zip(cols[j], vals[j]):
            if r >= 0:
                B[r, j] += v
    # backward error per column, no worse than LAPACK's on the same system
    scale = np.abs(Z).max() * np.abs(X).max(axis=0)
    res = np.abs(Z @ X - B).max(axis=0) / scale
    X_ref = np.linalg.solve(Z, B)
    res_ref = np.abs(Z @ X_ref - B).max(axis=0) / scale
    assert (res <= 20 * res_ref + 1e-15).all(), (res.max(), res_ref.max())
    assert rel_err(X, X_ref) < 1e-11
    # Y is B^T X for every column pair
    assert rel_err(sol.y, B.T @ X) < 1e-12


def test_graph_replay_matches_eager():
    """The captured solve replays to the eager answer, replay after replay,
    with eager work and host synchronisations in between."""
    _say("graph replay")
    s = BSplineSolver(**W.dipole_array(2, 3, 17, seed=7))
    Y0, W0 = s._solve_device()
    y0, w0 = Y0.cpu().numpy().copy(), W0.cpu().numpy().copy()
    for rep in range(4):
        Yg, Wg = s.solve_graph()
        torch.cuda.synchronize()
        assert rel_err(Yg.cpu().numpy(), y0) < 1e-14, rep
        assert rel_err(Wg.cpu().numpy(), w0) < 1e-14, rep
        s._solve_device()  # eager work between replays
        torch.cuda.synchronize()


def test_cancel_token_stops_factorisation():
    from momwire_amd.bspline import CancelToken, SolveAborted
    _say("cancel")
    tok = CancelToken()
    s = BSplineSolver(**W.dipole_array(2, 2, 11, seed=5), cancel=tok)
    tok.cancel()
    with pytest.raises(SolveAborted):
        s.compute_port_solution()
    tok.reset()
    g = golden("array2x2")
    assert rel_err(s.compute_port_solution().y, g["y"]) < 1e-10


@pytest.mark.parametrize("dtype", ["z", "c"])
def test_two_level_lu_large(dtype):
    """n > 4096 takes the two-level blocking (512-column blocks, K = 512
    trailing updates); random entries exercise the tournament fallback and a
    dominant diagonal the threshold path."""
    for dom in (0.0, 1.0):
        n, p = 4500, 24
        _say("two-level lu", dtype, n, "diag-dominant" if dom else "random")
        rng = np.random.default_rng(11 + int(dom))
        M = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        M += dom * 8 * np.sqrt(n) * np.eye(n)
        RHS = rng.standard_normal((n, p)) + 1j * rng.standard_normal((n, p))
        X, info = _lu_solve(M, RHS, dtype)
        assert info == 0
        ref = np.linalg.solve(M, RHS)
        eps = 1.1e-16 if dtype == "z" else 6e-8
        assert rel_err(X, ref) < eps * np.linalg.cond(M) * 200


@pytest.mark.parametrize("name,kw", [("array2x2", W.dipole_array(2, 2, 11, seed=5)),
                                     ("tee_jport", W.junction_tee("junction_port")),
                                     ("c2_array8x8", W.dipole_array(8, 8, 41, seed=1))])
def test_mixed_precision_refines_to_fp64(name, kw):
    """FP32 factorisation + FP64-residual refinement reaches the FP64 answer:
    Y within 1e-8 of the reference golden after the default two corrections,
    and the error shrinks with every correction."""
    _say("mixed", name)
    g = golden(name)
    s = BSplineSolver(**kw, precision="mixed", refine_steps=3)
    hist = []
    s._solve_device_mixed(history=hist)
    plain = [h for h in hist if not isinstance(h, tuple)]
    adj = [h[1] for h in hist if isinstance(h, tuple)]
    errs = [rel_err(h.cpu().numpy(), g["y"]) for h in plain]
    aerrs = [rel_err(h.cpu().numpy(), g["y"]) for h in adj]
    _say("mixed Y errors per step", name, ["%.2e" % e for e in errs],
         "adjoint-corrected", ["%.2e" % e for e in aerrs])
    assert errs[0] < 1e-2
    assert errs[-1] < 1e-8
    assert errs[2] <= errs[1] <= errs[0] or errs[2] < 1e-12
    # default: one correction plus the adjoint-corrected readout (three
    # corrections when KCL rows are present)
    sol = BSplineSolver(**kw, precision="mixed").compute_port_solution()
    assert rel_err(sol.y, g["y"]) < 1e-8
    assert rel_err(sol.y, sol.y.T) < 1e-9


def test_readout_check_kernel():
    """mw_zreadout_check against numpy: Y + T, delta = max_j |C_j| / |X0_j|,
    max|T|, max|Y + T|, the estimate delta max|T| and the gate on both sides
    of the tolerance; a NaN in T sets the gate."""
    _say("readout check kernel")
    lib = _lib.load()
    rng = np.random.default_rng(11)
    for n, p in [(700, 5), (3000, 70), (257, 300)]:
        X0 = rng.standard_normal((n, p)) + 1j * rng.standard_normal((n, p))
        C = 1e-3 * (rng.standard_normal((n, p)) + 1j * rng.standard_normal((n, p)))
        C[:, p // 2] *= 7.0
        T = 1e-6 * (rng.standard_normal((p, p)) + 1j * rng.standard_normal((p, p)))
        Y = rng.standard_normal((p, p)) + 1j * rng.standard_normal((p, p))
        delta = np.max(np.linalg.norm(C, axis=0) / np.linalg.norm(X0, axis=0))
        est = delta * np.abs(T).max()
        ymax = np.abs(Y + T).max()
        for tol, want in ((0.5 * est / ymax, 1), (2.0 * est / ymax, 0)):
            dX, dC, dT, dY = _dev(X0), _dev(C), _dev(T), _dev(Y)
            scratch = torch.empty(16 * (p + 1), dtype=torch.uint8, device="cuda")
            stats = torch.zeros(4, dtype=torch.float64, device="cuda")
            gate = torch.full((1,), 7, dtype=torch.int32, device="cuda")
            _lib.check(lib.mw_zreadout_check(n, p, dX.data_ptr(), p, dC.data_ptr(), p,
                                             dT.data_ptr(), p, dY.data_ptr(), p, tol,
                                             scratch.data_ptr(), stats.data_ptr(),
                                             gate.data_ptr(), _stream()), "check")
            torch.cuda.synchronize()
            assert rel_err(dY.cpu().numpy(), Y + T) < 1e-15
            st = stats.cpu().numpy()
            assert abs(st[0] - delta) < 1e-12 * delta
            assert abs(st[1] - np.abs(T).max()) < 1e-14 * st[1]
            assert abs(st[2] - ymax) < 1e-14 * ymax
            assert abs(st[3] - est) < 1e-12 * est
            assert int(gate.cpu()[0]) == want, (n, p, tol)
    T[1, 2] = np.nan
    dX, dC, dT, dY = _dev(X0), _dev(C), _dev(T), _dev(Y)
    _lib.check(lib.mw_zreadout_check(n, p, dX.data_ptr(), p, dC.data_ptr(), p, dT.data_ptr(), p,
                                     dY.data_ptr(), p, 1.0, scratch.data_ptr(), stats.data_ptr(),
                                     gate.data_ptr(), _stream()), "check NaN")
    torch.cuda.synchronize()
    assert int(gate.cpu()[0]) == 1


@pytest.mark.parametrize("name,kw", [("array2x2", W.dipole_array(2, 2, 11, seed=5)),
                                     ("c2_array8x8", W.dipole_array(8, 8, 41, seed=1))])
def test_mixed_one_residual_readout(name, kw, monkeypatch):
    """The default mixed form on a symmetric Z: one FP64 residual, one FP32
    correction and the readout B^T X0 + X0^T R + C^T R, whose error the device
    estimates as delta max|C^T R|.  The estimate must not fall short of the
    readout's actual error (max-entry, against the reference's golden Y) by
    more than 2x; past READOUT_TOL a fresh residual reads out again, and the
    default Y is within 1e-8 of the golden either way.  Forced past the check
    (tolerance 0), graph replays (compute_port_solution from its second
    call) equal the eager solve."""
    from momwire_amd import bspline as B
    _say("mixed one-residual readout", name)
    g = golden(name)
    gmax = np.abs(g["y"]).max()
    s = BSplineSolver(**kw, precision="mixed")
    y1 = s._solve_device_mixed()[0].cpu().numpy()
    info = s.mixed_readout
    _say("  check", info, "Y error", rel_err(y1, g["y"]))
    assert info["second_residual"] == (info["estimate"] > B.READOUT_TOL * info["max_y"])
    assert rel_err(y1, g["y"]) < 1e-8
    monkeypatch.setattr(B, "READOUT_TOL", 1.0)
    s1 = BSplineSolver(**kw, precision="mixed")
    y_one = s1._solve_device_mixed()[0].cpu().numpy()
    est = s1.mixed_readout["estimate"] / s1.mixed_readout["max_y"]
    err = np.abs(y_one - g["y"]).max() / gmax
    _say("  one-residual error %.3e, estimate %.3e" % (err, est))
    assert not s1.mixed_readout["second_residual"]
    assert err <= 2.0 * est + 1e-11
    monkeypatch.setattr(B, "READOUT_TOL", 0.0)
    s2 = BSplineSolver(**kw, precision="mixed")
    y2 = s2._solve_device_mixed()[0].cpu().numpy()
    assert s2.mixed_readout["second_residual"]
    assert rel_err(y2, g["y"]) < 1e-8
    sols = [s2.compute_port_solution() for _ in range(3)]
    assert s2._graph_entry is not None and s2._graph_entry[4]
    for sol in sols:
        assert rel_err(sol.y, y2) < 1e-13
        assert rel_err(sol.coeffs, sols[0].coeffs) < 1e-13
    assert s2.mixed_readout["second_residual"]


def test_gemm_tn_add():
    _say("gemm tn add")
    rng = np.random.default_rng(4)
    for m, n, k in [(64, 64, 2624), (200, 130, 517), (1024, 1024, 700)]:
        A = rng.standard_normal((k, m)) + 1j * rng.standard_normal((k, m))
        B = rng.standard_normal((k, n)) + 1j * rng.standard_normal((k, n))
        C = rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))
        dA, dB, dC = _dev(A), _dev(B), _dev(C)
        _lib.check(_lib.load().mw_zgemm_tn_add(m, n, k, dA.data_ptr(), m, dB.data_ptr(), n,
                                               dC.data_ptr(), n, _stream()), "tn")
        torch.cuda.synchronize()
        assert rel_err(dC.cpu().numpy(), C + A.T @ B) < 1e-13


def test_cgemm_tn_add():
    """The FP32 form (the mixed readout's C^T R): against the product of the
    same complex64 inputs in FP64.  FP32 sums over k = 2,624 terms leave
    ~2.4e-6 of the largest entry (measured); a transposed operand or a wrong
    index would leave O(1), so the bar is 1e-5."""
    _say("cgemm tn add")
    rng = np.random.default_rng(6)
    for m, n, k in [(64, 64, 2624), (200, 130, 517), (1024, 1024, 700)]:
        A = (rng.standard_normal((k, m)) + 1j * rng.standard_normal((k, m))).astype(np.complex64)
        B = (rng.standard_normal((k, n)) + 1j * rng.standard_normal((k, n))).astype(np.complex64)
        C = (rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))).astype(np.complex64)
        dA, dB, dC = (torch.from_numpy(x).to("cuda") for x in (A, B, C))
        _lib.check(_lib.load().mw_cgemm_tn_add(m, n, k, dA.data_ptr(), m, dB.data_ptr(), n,
                                               dC.data_ptr(), n, _stream()), "tn32")
        torch.cuda.synchronize()
        ref = C.astype(np.complex128) + A.astype(np.complex128).T @ B.astype(np.complex128)
        assert rel_err(dC.cpu().numpy().astype(np.complex128), ref) < 1e-5


def test_compute_port_solution_replays_graph_across_k():
    """compute_port_solution: eager on the first solve at a k, the captured
    graph from the second on.  Moving to another k and back replays the
    graph captured before, whose per-frequency loading values must still be
    alive; every call matches an eager solve at the same k."""
    _say("compute_port_solution graph across k")
    kw = dict(W.dipole_array(2, 2, 13, seed=3), wire_conductivity=5.8e7)
    s = BSplineSolver(**kw)
    ref = BSplineSolver(**kw)
    k1 = s.k
    k2 = 1.05 * k1
    for kk in (k1, k1, k1, k2, k2, k1, k1, k2):
        s._set_k(kk)
        ref._set_k(kk)
        sol = s.compute_port_solution()
        Yr, Wr = ref._solve_device()
        torch.cuda.synchronize()
        assert rel_err(sol.y, Yr.cpu().numpy()) < 1e-13, kk
        assert rel_err(sol.coeffs, Wr.cpu().numpy()) < 1e-13, kk
    assert s._graph_entry is not None


@pytest.mark.parametrize("dtype", ["z", "c"])
@pytest.mark.parametrize("n,p", [(4500, 70), (5000, 1100)])
def test_sysv_symmetric_two_level(dtype, n, p):
    """mw_?sysv on a complex symmetric system past n = 4096: no row exchanges,
    each block's U rows from its L columns, the trailing update on the blocks
    on or below the block diagonal, the right-hand sides by block products
    with the 512 x 512 diagonal blocks' inverses (a partial last block; more
    than 1,024 right-hand sides in two column chunks).  A well-scaled
    symmetric system matches LAPACK, with and without the first-row hint of
    staircase right-hand sides, and mw_?sytrs with the same factors
    solves a second set of right-hand sides to the same accuracy; a symmetric
    system whose unpivoted multipliers exceed the threshold bound reports
    info = -1 (result not valid)."""
    _say("sysv", dtype, n, p)
    rng = np.random.default_rng(11)
    R = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    lib = _lib.load()
    ct = torch.complex128 if dtype == "z" else torch.complex64
    fn, trs = (lib.mw_zsysv, lib.mw_zsytrs) if dtype == "z" else (lib.mw_csysv, lib.mw_csytrs)
    RHS = rng.standard_normal((n, p)) + 1j * rng.standard_normal((n, p))
    eps = 1.1e-16 if dtype == "z" else 6e-8
    # a staircase of first rows (the first three columns dense), as a port
    # solve's drive columns: zero above each column's first row
    first = np.sort(rng.integers(0, n, p)).astype(np.int32)
    first[:3] = 0
    RHS_st = np.where(np.arange(n)[:, None] >= first[None, :], RHS, 0)
    for dominant in (True, False):
        M = R + R.T + (4.0 * np.sqrt(n) * np.eye(n) if dominant else 0.0)
        ipiv = torch.empty(n, dtype=torch.int32, device="cuda")
        work = torch.empty(int(lib.mw_sysv_workspace_bytes(n)), dtype=torch.uint8,
                           device="cuda")
        info = torch.zeros(1, dtype=torch.int32, device="cuda")
        if dominant:
            aug = torch.from_numpy(np.hstack([M, RHS_st])).to(ct).cuda()
            _gesv_dev(fn, n, p, aug, ipiv, work, info, first.ctypes.data)
            torch.cuda.synchronize()
            assert int(info.cpu()[0]) == 0
            X = aug[:, n:].cpu().numpy().astype(np.complex128)
            assert rel_err(X, np.linalg.solve(M, RHS_st)) < eps * 2e4
        aug = torch.from_numpy(np.hstack([M, RHS])).to(ct).cuda()
        _gesv_dev(fn, n, p, aug, ipiv, work, info, None)
        torch.cuda.synchronize()
        if dominant:
            assert int(info.cpu()[0]) == 0
            X = aug[:, n:].cpu().numpy().astype(np.complex128)
            ref = np.linalg.solve(M, RHS)
            # the dominant system is well conditioned: |M| |M^-1| is O(10)
            assert rel_err(X, ref) < eps * 2e4, rel_err(X, ref)
            assert np.array_equal(ipiv.cpu().numpy(), np.arange(n))
            RHS2 = rng.standard_normal((n, 37)) + 1j * rng.standard_normal((n, 37))
            dX2 = torch.from_numpy(RHS2).to(ct).cuda()
            _lib.check(trs(n, 37, aug.data_ptr(), n + p, ipiv.data_ptr(), work.data_ptr(),
                           dX2.data_ptr(), 37, _stream()), "sytrs")
            torch.cuda.synchronize()
            X2 = dX2.cpu().numpy().astype(np.complex128)
            assert rel_err(X2, np.linalg.solve(M, RHS2)) < eps * 2e4
        else:
            assert int(info.cpu()[0]) == -1


def test_symmetric_solver_path_and_fallback():
    """A 15 x 20 array (4,500 unknowns, one radius): the solver takes the
    symmetric factorisation and gives the pivoting path's Y; with a pivot
    bound of 0 (every panel fails the threshold test) compute_port_solution
    falls back to the pivoting path and still returns the right Y."""
    _say("symmetric solver path")
    kw = W.dipole_array(15, 20, 15, seed=3)
    s_sym = BSplineSolver(**kw)
    y_sym = s_sym.compute_port_solution().y
    assert s_sym._sym_used and not s_sym._sym_failed
    s_piv = BSplineSolver(**kw)
    s_piv._sym_failed = True  # the pivoting path only
    y_piv = s_piv.compute_port_solution().y
    assert not s_piv._sym_used
    assert rel_err(y_sym, y_piv) < 1e-12
    lib = _lib.load()
    default = lib.mw_get_pivot_threshold()
    try:
        _lib.check(lib.mw_set_pivot_threshold(0.0), "threshold")
        s_fb = BSplineSolver(**kw)
        sol = s_fb.compute_port_solution()
        assert s_fb._sym_failed and not s_fb._sym_used
        assert rel_err(sol.y, y_piv) < 1e-12
    finally:
        lib.mw_set_pivot_threshold(default)
