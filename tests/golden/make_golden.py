"""Write the end-to-end golden vectors for the port-solution path.

Runs ONLY in the development container, where the reference checkout is
mounted read-only at /root/reference.  It imports the reference's pure-Python
solver (its numpy fallback: no compiled accelerator is loaded) and records,
for the structures in ``momwire_amd.workloads``, what
``BSplineSolver.compute_port_solution()`` returns (reference
src/momwire/bspline.py:4053-4179): the port admittance matrix ``y`` and the
coefficient columns ``coeffs``; for the smallest case also the dense Z.

Nothing else pins Y and coeffs end to end at 1e-10, so these vectors are what
the oracle and the GPU path are both held to.  Outputs are small ``.npz``
files next to this script.

    python tests/golden/make_golden.py [name ...]
"""

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, "/root/reference/src")

import momwire  # noqa: E402
from momwire import BSplineSolver  # noqa: E402

from momwire_amd import workloads  # noqa: E402


def _solve(kw):
    s = BSplineSolver(**kw)
    t0 = time.time()
    sol = s.compute_port_solution()
    return s, sol, time.time() - t0


def case_c1():
    kw = workloads.straight_dipole()
    s, sol, dt = _solve(kw)
    geom = s._build_geometry()
    supp, polys, kcl, _wk, _wb = s._build_basis_polynomials(geom)
    Z = s._compute_Z_operator(geom, supp, polys)
    return dict(y=sol.y, coeffs=sol.coeffs, Z=np.asarray(Z), k=s.k,
                omega=s.omega), dt


def case_c2():
    kw = workloads.dipole_array(8, 8, 41, seed=1)
    s, sol, dt = _solve(kw)
    cols = np.array([0, 9, 27, 63])
    return dict(y=sol.y, coeffs_cols=cols, coeffs=sol.coeffs[:, cols],
                coeffs_colsum=sol.coeffs.sum(axis=1), k=s.k), dt


def case_small_array():
    kw = workloads.dipole_array(2, 2, 11, seed=5)
    s, sol, dt = _solve(kw)
    return dict(y=sol.y, coeffs=sol.coeffs), dt


def _tee(ports):
    kw = workloads.junction_tee(ports)
    s, sol, dt = _solve(kw)
    out = dict(y=sol.y, coeffs=sol.coeffs)
    if ports == "feed":
        ks = np.array([2 * np.pi * 0.9, 2 * np.pi * 1.0, 2 * np.pi * 1.15])
        out["sweep_k"] = ks
        out["sweep_y"] = s.compute_y_matrix_swept(ks)
        geom = s._build_geometry()
        supp, polys, kcl, _wk, _wb = s._build_basis_polynomials(geom)
        out["Z"] = np.asarray(s._compute_Z_operator(geom, supp, polys))
        out["knot_currents"] = np.concatenate(s.currents_at_knots(sol.coeffs[:, 0]))
    return out, dt


def case_sweep_struct(index):
    """One random structure of the batched sweep, swept over 3 k."""
    kw = workloads.random_structure(3 * 100003 + index)
    s = BSplineSolver(**kw)
    ks = workloads.sweep_wavenumbers()[[0, 7, 15]]
    t0 = time.time()
    ys = s.compute_y_matrix_swept(ks)
    sol = s.compute_port_solution()
    return dict(sweep_k=ks, sweep_y=ys, y=sol.y, coeffs=sol.coeffs), time.time() - t0


K2N = 2 * np.pi           # normalised wavenumber (lambda = 1)
OMEGA_N = K2N * 299792458.0
MU0 = 4e-7 * np.pi
GROUNDS = [complex(5.0, -0.6), complex(15.0, -20.0), complex(30.0, -300.0)]
RHO = np.array([0.0, 0.004, 0.03, 0.1, 0.3, 0.7, 1.5, 3.0])
HH = np.array([0.002, 0.03, 0.12, 0.45, 1.3, 2.0])


def case_somm_engine():
    """Six lambda-integrals, the four surfaces and the R1 -> 0 limits at a
    lattice of (rho, h) nodes for three grounds (reference
    src/momwire/_sommerfeld.py:342-584)."""
    from momwire import _sommerfeld as S
    t0 = time.time()
    rr, hh = np.meshgrid(RHO, HH, indexing="ij")
    rho, h = rr.ravel(), hh.ravel()
    out = dict(grounds=np.array(GROUNDS), rho=rho, h=h)
    for g, eps in enumerate(GROUNDS):
        out[f"six_{g}"] = S._six_integrals_batch(eps, K2N, rho, h, rtol=1e-6)
        sel = np.arange(0, rho.size, 3)
        out[f"six9_idx"] = sel
        out[f"six9_{g}"] = S._six_integrals_batch(eps, K2N, rho[sel], h[sel], rtol=1e-9)
        # forced forms in the overlap band rho ~ h .. 2h
        ro = np.array([0.05, 0.2, 0.6, 1.0])
        ho = np.array([0.04, 0.15, 0.4, 0.6])
        out["form_rho"], out["form_h"] = ro, ho
        out[f"formJ_{g}"] = S._six_integrals_batch(eps, K2N, ro, ho, rtol=1e-9, form="J")
        out[f"formH_{g}"] = S._six_integrals_batch(eps, K2N, ro, ho, rtol=1e-9, form="H")
        r1 = np.hypot(rho, h)
        th = np.arctan2(h, rho)
        surf = S.iv_surfaces_direct(eps, K2N, r1, th, rtol=1e-6, omega=OMEGA_N, mu=MU0)
        out[f"surf_{g}"] = np.stack([surf[k] for k in S._SURF_KEYS])
        th0 = np.radians([0.0, 15.0, 45.0, 89.9, 90.0])
        lim = S.iv_surfaces_direct(eps, K2N, np.zeros_like(th0), th0, omega=OMEGA_N, mu=MU0)
        out["lim_theta"] = th0
        out[f"lim_{g}"] = np.stack([lim[k] for k in S._SURF_KEYS])
    return out, time.time() - t0


def _grid_dump(grid, prefix, rng, n_eval=300):
    from momwire import _sommerfeld as S
    out = {}
    regs = grid._regions
    out[prefix + "scalars"] = np.array([grid.r1_max, grid.r_break, grid.r_near, grid._r0_fill,
                                        grid.k2, grid.omega, grid.mu])
    out[prefix + "eps"] = np.array(grid.eps_t)
    out[prefix + "layout"] = np.array([[r["r0"], r["dr"], r["n_r"], r["th0"], r["dth"], r["n_th"]]
                                       for r in regs])
    for i, r in enumerate(regs):
        out[prefix + f"vals{i}"] = r["vals"]
    R1 = rng.uniform(0.0, grid.r1_max, n_eval)
    R1[:20] = rng.uniform(0.0, 0.05 * grid.r1_max, 20)
    th = rng.uniform(0.0, 0.5 * np.pi, n_eval)
    ev = grid.eval(R1, th)
    out[prefix + "eval_r1"] = R1
    out[prefix + "eval_th"] = th
    out[prefix + "eval"] = np.stack([ev[k] for k in S._SURF_KEYS])
    return out


def _ground_solve(kw, extra=True):
    from momwire import _sommerfeld as S
    s, sol, dt = _solve(kw)
    out = dict(y=sol.y, coeffs=sol.coeffs, k=s.k, omega=s.omega)
    geom = s._build_geometry()
    supp, polys, kcl, _wk, _wb = s._build_basis_polynomials(geom)
    from momwire import _ground_refl
    eps_t = _ground_refl.eps_tilde(s.ground_eps, s.omega, s.eps)
    out["eps_t"] = np.array(eps_t)
    r1_max = S.max_image_distance(geom["seg_l"], geom["seg_r"], s.ground_z)
    out["r1_max"] = np.array(r1_max)
    grid = s._somm_grid(eps_t, r1_max)
    out.update(_grid_dump(grid, "grid_", np.random.default_rng(11)))
    if extra:
        out["Z"] = np.asarray(s._compute_Z_operator(geom, supp, polys))
        out["Q"] = np.asarray(s._Z_sommerfeld_remainder(geom, supp, polys, eps_t))
    return out, dt


def case_somm_grid_far():
    """A master grid whose extent reaches the far zone (six regions)."""
    from momwire import _sommerfeld as S
    t0 = time.time()
    grid = S.get_grid(GROUNDS[1], K2N, 6.0, omega=OMEGA_N, mu=MU0)
    out = _grid_dump(grid, "grid_", np.random.default_rng(12), n_eval=400)
    for i in range(len(grid._regions)):
        del out[f"grid_vals{i}"]
    return out, time.time() - t0


def case_ground_dipole():
    kw = workloads.grounded_array(1, 1, 21, params=(0.3, 12.0, 0.01, 10e6))
    return _ground_solve(kw)


def case_ground_tee():
    return _ground_solve(workloads.grounded_tee())


def case_c5():
    out, dt = _ground_solve(workloads.grounded_array(4, 4, 21, seed=4), extra=False)
    for k in [k for k in out if k.startswith("grid_vals")]:
        del out[k]
    return out, dt


def _refl_solve(kw, extra=True):
    """Reflection-coefficient ground (reference _ground_refl.py): Y, coeffs,
    and for small cases Z and the per-segment-pair weight tables."""
    from momwire import _ground_refl
    s, sol, dt = _solve(kw)
    out = dict(y=sol.y, coeffs=sol.coeffs, k=s.k, omega=s.omega)
    eps_t = _ground_refl.eps_tilde(s.ground_eps, s.omega, s.eps)
    out["eps_t"] = np.array(eps_t)
    if extra:
        geom = s._build_geometry()
        supp, polys, _kcl, _wk, _wb = s._build_basis_polynomials(geom)
        out["Z"] = np.asarray(s._compute_Z_operator(geom, supp, polys))
        seg_c = 0.5 * (geom["seg_l"] + geom["seg_r"])
        cos_th, td_img, P = _ground_refl.specular_pair_tables(seg_c, geom["tangents"], s.ground_z)
        rho_v, rho_h = _ground_refl.fresnel_rho(eps_t, cos_th)
        out["w_A"] = _ground_refl.a_term_weights(rho_v, rho_h, td_img, P)
        out["w_Phi"] = np.broadcast_to(
            _ground_refl.phi_term_weights(s.ground_phi_mode, eps_t, rho_v), cos_th.shape).copy()
        out["cos_th"] = cos_th
    return out, dt


def case_refl_dipole():
    kw = workloads.grounded_array(1, 1, 21, params=(0.3, 12.0, 0.01, 10e6),
                                  ground_model="refl-coef")
    return _refl_solve(kw)


def case_refl_tee():
    return _refl_solve(workloads.grounded_tee(ground_model="refl-coef", phi_mode="rho_v"))


def case_refl_tee_blend():
    return _refl_solve(workloads.grounded_tee(eps_r=25.0, sigma=2e-2, freq_hz=7e6,
                                              ground_model="refl-coef", phi_mode="blend"))


def case_refl_array4x4():
    return _refl_solve(workloads.grounded_array(4, 4, 21, seed=4, ground_model="refl-coef"),
                       extra=False)


def case_ek(name):
    """Extended thin-wire kernel (reference bspline.py:393-450,
    _bspline_kernels.py:28-80, 198-322): Y, coeffs, Z and the coaxial labels."""
    from momwire import _bspline_kernels as K
    kw = workloads.ek_structure(name)
    s, sol, dt = _solve(kw)
    geom = s._build_geometry()
    supp, polys, _kcl, _wk, _wb = s._build_basis_polynomials(geom)
    out = dict(y=sol.y, coeffs=sol.coeffs, k=s.k, omega=s.omega,
               Z=np.asarray(s._compute_Z_operator(geom, supp, polys)))
    gi, _gj = s._ek_axis_labels(geom, False)
    out["labels"] = np.asarray(gi)
    if s.ground_z is not None:
        gi, gj = s._ek_axis_labels(geom, True)
        out["labels_img_obs"] = np.asarray(gi)
        out["labels_img_src"] = np.asarray(gj)
    # EK same-edge static moments of the first edge (h, a) for the closed forms
    ed_arc = geom["per_wire"][0]["edge_arc_edges"][0]
    a0 = float(s._radius_per_wire[0])
    out["static_ek"] = np.asarray(K._seg_seg_static_moments(ed_arc, a0, max_d=2,
                                                            ek=K._EK(a=None, group_i=None,
                                                                     group_j=None)))
    out["static"] = np.asarray(K._seg_seg_static_moments(ed_arc, a0, max_d=2))
    return out, dt


def case_readout(kind):
    """Readout of every port column: the reference's element_currents at
    subdiv 1 and 8 (_element_currents.py:33-120), _far_moments over free
    space, PEC and a finite ground (portal/_portal.py:1384-1423, 1520-1598)
    and _element_fields E and H, alone and with the PEC image added the way
    the portal adds it (portal/_portal.py:1601-1651, 3160-3204)."""
    from types import SimpleNamespace

    from momwire.portal import _portal as P
    if kind == "tee":
        kw = workloads.junction_tee("junction_port")
        gz = 0.0
    else:
        kw = workloads.dipole_array(2, 2, 11, seed=5)
        gz = -0.3
    s, sol, dt = _solve(kw)
    X = np.asarray(sol.coeffs)
    p = X.shape[1]
    out = dict(coeffs=X, k=s.k, freq_hz=s.freq, ground_z=gz)
    theta = np.radians(np.arange(0.0, 181.0, 15.0))
    phi = np.radians(np.arange(0.0, 360.0, 40.0))
    out.update(theta=theta, phi=phi)
    for sub in (1, 8):
        els = [s.element_currents(X[:, j], subdiv=sub) for j in range(p)]
        out[f"mid{sub}"] = els[0][0]
        out[f"nodes{sub}"] = els[0][2]
        out[f"moment{sub}"] = np.stack([e[1] for e in els], axis=-1)
        out[f"delta{sub}"] = np.stack([e[3] for e in els], axis=-1)
    grounds = dict(free=SimpleNamespace(kind="free"), pec=SimpleNamespace(kind="pec"),
                   finite=SimpleNamespace(kind="refl", eps_r=13.0, sigma=0.005))
    for name, g in grounds.items():
        mt, mp = zip(*[P._far_moments(out["mid1"], out["moment1"][..., j], s.k, theta, phi, g, gz,
                                      s.freq) for j in range(p)])
        out[f"far_{name}_theta"] = np.stack(mt, axis=-1)
        out[f"far_{name}_phi"] = np.stack(mp, axis=-1)
    rng = np.random.default_rng(17)
    lo = np.min(np.vstack(kw["wires"] if kind == "tee" else
                          [np.asarray(w) for w in kw["wires"]]), axis=0)
    hi = np.max(np.vstack([np.asarray(w) for w in kw["wires"]]), axis=0)
    pts = rng.uniform(lo - 0.3 * s.wavelength, hi + 0.3 * s.wavelength, size=(24, 3))
    pts[:, 2] = np.maximum(pts[:, 2], gz + 0.02 * s.wavelength)
    # one point 0.01 wavelength off the first wire's midpoint
    w0 = np.asarray(kw["wires"][0], dtype=float)
    pts[0] = 0.5 * (w0[0] + w0[1]) + np.array([0.01, 0.0, 0.0]) * s.wavelength
    radius = float(np.min(np.atleast_1d(kw["wire_radius"])))
    out.update(points=pts, radius=radius)
    mid, nodes = out["mid8"], out["nodes8"]
    img_mid = mid.copy()
    img_mid[:, 2] = 2.0 * gz - mid[:, 2]
    img_nodes = nodes.copy()
    img_nodes[:, 2] = 2.0 * gz - nodes[:, 2]
    for magnetic, tag in ((False, "E"), (True, "H")):
        free, pec = [], []
        for j in range(p):
            mom, dlt = out["moment8"][..., j], out["delta8"][..., j]
            f = P._element_fields(pts, (mid, mom, nodes, dlt), s.k, radius, magnetic)
            img = P._element_fields(pts, (img_mid, mom * np.array([-1.0, -1.0, 1.0]), img_nodes,
                                          -dlt), s.k, radius, magnetic)
            free.append(f)
            pec.append(f + img)
        out[f"near_{tag}"] = np.stack(free, axis=-1)
        out[f"near_{tag}_pec"] = np.stack(pec, axis=-1)
    return out, dt


def case_chain(n_wires, feed_every):
    """More KCL constraint rows than the small Schur solve takes (and, with a
    feed on every wire, more ports than 64): reference bspline.py:2971-2989."""
    s, sol, dt = _solve(workloads.wire_chain(n_wires, 3, feed_every))
    return dict(y=sol.y, coeffs=sol.coeffs), dt


def case_feeds(kind):
    """Spread gap feeds (reference bspline.py:2806-2942): the segment gap and
    the cos^2-smoothed source, on the junction tee (three and two feeds,
    one on the bent arm) and on a dipole."""
    if kind == "segment":
        kw = workloads.junction_tee("junction_port")
        kw["feeds"] = [(0, None, 1.0), (2, 0.1, 1.0), (1, 0.13, 1.0)]
        kw["feed_model"] = "segment"
    elif kind == "smooth_tee":
        kw = workloads.junction_tee("feed")
        kw["feeds"] = [(0, None, 1.0), (1, 0.21, 1.0)]
        kw["feed_smoothing_factor"] = 2.5
    else:
        kw = workloads.straight_dipole()
        kw["feed_smoothing_factor"] = 3.0
    s, sol, dt = _solve(kw)
    return dict(y=sol.y, coeffs=sol.coeffs), dt


def case_loading(kind):
    """Distributed wire loading (reference _wire_loading.py,
    bspline.py:2667-2800): conductor loss on a 14 MHz dipole, insulation on
    the tee, a per-wire mix (NaN switches a wire off) on a 2x2 array; Y,
    coeffs and the wire loss power of the first column."""
    if kind == "dipole":
        kw = workloads.straight_dipole()
        kw.update(wire_conductivity=5.8e7)
    elif kind == "tee":
        kw = workloads.junction_tee("junction_port")
        kw.update(insulation_radius=2.5e-3, insulation_eps_r=3.2, wire_conductivity=3.5e7)
    else:
        kw = workloads.dipole_array(2, 2, 11, seed=5)
        kw.update(wire_conductivity=[5.8e7, float("nan"), 1e6, 3e7],
                  insulation_radius=[float("nan"), 2e-3, 3e-3, float("nan")],
                  insulation_eps_r=[float("nan"), 2.3, 4.0, float("nan")])
    s, sol, dt = _solve(kw)
    p_tot, p_wire = s.wire_loss_power(sol.coeffs[:, 0])
    return dict(y=sol.y, coeffs=sol.coeffs, loss=p_tot, loss_per_wire=p_wire), dt


def case_contact(kind, base_feed):
    """Ground contact (reference bspline.py:1108-1169, 1296-1316): a base-
    or segment-fed quarter-wave vertical over PEC and over the Sommerfeld
    ground, and a vertical with three radials meeting it at a grounded
    junction; Y, coeffs and (PEC) Z."""
    kw = workloads.grounded_monopole(kind, base_feed)
    s, sol, dt = _solve(kw)
    out = dict(y=sol.y, coeffs=sol.coeffs)
    if kind == "pec":
        geom = s._build_geometry()
        supp, polys, _kcl, _wk, _wb = s._build_basis_polynomials(geom)
        out["Z"] = np.asarray(s._compute_Z_operator(geom, supp, polys))
    return out, dt


def case_d1(kind):
    """Degree-1 (tent) bases (reference bspline.py:1216-1414 at d = 1): a
    dipole at the tent-parity pin (half 5.29 m, 20 segments), the junction
    tee with a junction port, and a grounded vertical over PEC."""
    if kind == "dipole":
        kw = dict(wires=[[(0.0, 0.0, -5.29), (0.0, 0.0, 5.29)]], nsegs=20)
    elif kind == "tee":
        kw = workloads.junction_tee("junction_port")
    else:
        kw = workloads.grounded_monopole("pec", True)
    kw["degree"] = 1
    s, sol, dt = _solve(kw)
    out = dict(y=sol.y, coeffs=sol.coeffs)
    if kind == "dipole":
        geom = s._build_geometry()
        supp, polys, _kcl, _wk, _wb = s._build_basis_polynomials(geom)
        out["Z"] = np.asarray(s._compute_Z_operator(geom, supp, polys))
        out["z_in"] = s.compute_impedance()[0]
    return out, dt


def case_nqp(kind, nqp):
    """Other Gauss orders per segment pair (reference bspline.py:390-394,
    781): the off-edge, same-edge remainder and image moments at n_qp_pair
    points -- the junction tee in free space, the grounded tee over PEC."""
    kw = workloads.junction_tee("junction_port")
    if kind == "pec":
        kw = workloads.grounded_tee()
        kw.pop("ground_eps", None)
        kw.pop("ground_model", None)
    kw["n_qp_pair"] = nqp
    s, sol, dt = _solve(kw)
    geom = s._build_geometry()
    supp, polys, _kcl, _wk, _wb = s._build_basis_polynomials(geom)
    return dict(y=sol.y, coeffs=sol.coeffs,
                Z=np.asarray(s._compute_Z_operator(geom, supp, polys))), dt


def case_nqs(nqs):
    """The Sommerfeld remainder at another Gauss order per segment
    (n_qp_sommerfeld, reference bspline.py:600-616, 1791-1874): the grounded
    tee's Y and coeffs."""
    kw = workloads.grounded_tee()
    kw["n_qp_sommerfeld"] = nqs
    s, sol, dt = _solve(kw)
    return dict(y=sol.y, coeffs=sol.coeffs), dt


def case_fieldpt():
    """The Sommerfeld ground remainder at points in space (reference
    src/momwire/_field_point.py:292-432): E by reflected_field_at and H by
    reflected_h_field_at (central-difference curl) for seeded complex
    moments on the grounded dipole's element midpoints, at points around it
    (one directly above a source: the rho -> 0 branch)."""
    from momwire import _field_point as F
    from momwire import _ground_refl
    t0 = time.time()
    kw = workloads.grounded_array(1, 1, 21, params=(0.3, 12.0, 0.01, 10e6))
    s = BSplineSolver(**kw)
    eps_t = complex(_ground_refl.eps_tilde(s.ground_eps, s.omega, s.eps))
    (p0, p1), = kw["wires"]
    t = (np.arange(20) + 0.5) / 20
    mid = np.asarray(p0)[None, :] + t[:, None] * (np.asarray(p1) - np.asarray(p0))[None, :]
    rng = np.random.default_rng(21)
    mom = rng.standard_normal((20, 3)) + 1j * rng.standard_normal((20, 3))
    mom[:, 2] *= 0.5
    pts = np.column_stack([rng.uniform(-15.0, 15.0, 24), rng.uniform(-10.0, 10.0, 24),
                           rng.uniform(0.5, 15.0, 24)])
    pts[0] = [mid[7, 0], mid[7, 1], 4.0]
    e = F.reflected_field_at(pts, mid, mom, eps_t, s.ground_z, s.k, s.omega)
    h = F.reflected_h_field_at(pts, mid, mom, eps_t, s.ground_z, s.k, s.omega)
    return dict(points=pts, src_mid=mid, src_moment=mom, eps_t=np.array(eps_t),
                ground_z=np.array(s.ground_z), k=np.array(s.k), omega=np.array(s.omega),
                e=e, h=h), time.time() - t0


CASES = {
    "fieldpt_dipole": case_fieldpt,
    "nqs2_ground_tee": lambda: case_nqs(2),
    "nqs5_ground_tee": lambda: case_nqs(5),
    "nqp2_tee": lambda: case_nqp("free", 2),
    "nqp3_tee": lambda: case_nqp("free", 3),
    "nqp7_tee": lambda: case_nqp("free", 7),
    "nqp5_pec_tee": lambda: case_nqp("pec", 5),
    "d1_dipole": lambda: case_d1("dipole"),
    "d1_tee": lambda: case_d1("tee"),
    "d1_contact": lambda: case_d1("contact"),
    "contact_pec_base": lambda: case_contact("pec", True),
    "contact_pec_seg": lambda: case_contact("pec", False),
    "contact_somm_base": lambda: case_contact("sommerfeld", True),
    "contact_radials": lambda: case_contact("radials", False),
    "loading_dipole": lambda: case_loading("dipole"),
    "loading_tee": lambda: case_loading("tee"),
    "loading_array2x2": lambda: case_loading("array"),
    "feed_segment": lambda: case_feeds("segment"),
    "feed_smooth_tee": lambda: case_feeds("smooth_tee"),
    "feed_smooth_dipole": lambda: case_feeds("smooth_dipole"),
    "chain40": lambda: case_chain(40, None),
    "chain70_feeds": lambda: case_chain(70, 1),
    "readout_tee": lambda: case_readout("tee"),
    "readout_array2x2": lambda: case_readout("array"),
    "ek_dipole": lambda: case_ek("dipole"),
    "ek_collinear": lambda: case_ek("collinear"),
    "ek_tee": lambda: case_ek("tee"),
    "ek_pec_vertical": lambda: case_ek("pec_vertical"),
    "ek_refl_dipole": lambda: case_ek("refl_dipole"),
    "ek_somm_tee": lambda: case_ek("somm_tee"),
    "refl_dipole": case_refl_dipole,
    "refl_tee": case_refl_tee,
    "refl_tee_blend": case_refl_tee_blend,
    "refl_array4x4": case_refl_array4x4,
    "somm_engine": case_somm_engine,
    "somm_grid_far": case_somm_grid_far,
    "ground_dipole": case_ground_dipole,
    "ground_tee": case_ground_tee,
    "c5_ground4x4": case_c5,
    "sweep_struct4": lambda: case_sweep_struct(4),
    "sweep_struct8": lambda: case_sweep_struct(8),
    "c1_dipole": case_c1,
    "c2_array8x8": case_c2,
    "array2x2": case_small_array,
    "tee_feed": lambda: _tee("feed"),
    "tee_jport": lambda: _tee("junction_port"),
    "tee_nodegap": lambda: _tee("node_gap"),
}


def main(names):
    assert not momwire.accelerated, "expected the reference's pure-Python path"
    for name in names or CASES:
        out, dt = CASES[name]()
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **out)
        print(f"{name}: {dt:.1f} s -> {os.path.relpath(path)}", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
