"""CPU oracle for the readout of solved port columns -- TEST INFRASTRUCTURE
ONLY.

A numpy restatement, one solution column at a time, of
* the element walk (reference _element_currents.py:33-120): knots by
  ``np.linspace`` along each polyline edge, currents at the knots (the
  solver's own knot arcs at subdiv 1, chord arcs otherwise), element current
  = mean of its end currents, moment = current x vector length, the step
  across each knot = arriving minus leaving current;
* the far-field moments (portal/_portal.py:1384-1423, 1520-1598):
  M = sum_e m_e exp(+jk r^ . r_e), the PEC image with horizontal components
  negated, and over a finite ground the Fresnel image multipliers
  (rho_h, rho_v) = ((cos t - q)/(cos t + q), (eps cos t - q)/(eps cos t + q)),
  q = sqrt(eps - sin^2 t), applied as rho_v (M_img . v^) v^ - rho_h
  (M_img . h^) h^; projected on NEC's theta^ and phi^;
* the near field in mixed-potential form (portal/_portal.py:1601-1651).

Pinned by the reference's own outputs for two structures
(tests/golden/readout_*.npz, written by the reference's functions) and by
closed forms: a single z-directed element (M_theta = -sin(theta) I dl,
M_phi = 0), the PEC image of a vertical element (factor 2 cos(k h cos t)),
normal incidence (rho_v = -rho_h = (sqrt(eps) - 1)/(sqrt(eps) + 1)), and
the radiation-zone limit of the near field, E -> -j eta k e^{-jkr}/(4 pi r)
M_perp.
"""

import math

import numpy as np
from scipy.interpolate import BSpline

ETA0 = 376.730_313_668      # portal/_portal.py:347
EPS0 = 8.854_187_817e-12    # portal/_portal.py:346


def currents_at(osolver, coeffs, s_arrays=None):
    """B-spline current on each wire at its knot arcs or at ``s_arrays``
    (bspline.py:4603-4685 without enrichment)."""
    geom = osolver.geometry()
    _s, _p, _k, knots, wbg = osolver.basis(geom)
    d = osolver.degree
    out = []
    for w, pw in enumerate(geom["per_wire"]):
        kv = knots[w]
        s = pw["arc"] if s_arrays is None else np.asarray(s_arrays[w], dtype=float)
        s = np.clip(s, kv[0], kv[-1])
        c = np.zeros(len(kv) - d - 1, dtype=complex)
        kept, l2g = wbg[w]
        for kidx, (j, *_r) in enumerate(kept):
            c[j] = coeffs[l2g[kidx]]
        out.append(BSpline.design_matrix(s, kv, d).toarray() @ c)
    return out


def element_currents(osolver, coeffs, subdiv=1):
    """(mid, moment, nodes, delta) of one column (_element_currents.py:86-120)."""
    fine, arcs = [], []
    for pl, npe in zip(osolver.wires, osolver.npe):
        parts = []
        for i, n_e in enumerate(npe):
            seg = np.linspace(pl[i], pl[i + 1], n_e * subdiv + 1)
            parts.append(seg if i == 0 else seg[1:])
        kn = np.vstack(parts)
        fine.append(kn)
        step = np.linalg.norm(kn[1:] - kn[:-1], axis=1)
        arcs.append(np.concatenate([[0.0], np.cumsum(step)]))
    cur = currents_at(osolver, coeffs, None if subdiv == 1 else arcs)
    mids, moms, nodes, deltas = [], [], [], []
    for kn, c in zip(fine, cur):
        el = 0.5 * (c[1:] + c[:-1])
        mids.append(0.5 * (kn[1:] + kn[:-1]))
        moms.append(el[:, None] * (kn[1:] - kn[:-1]))
        nodes.append(kn)
        z = np.zeros(1, dtype=complex)
        deltas.append(np.concatenate([z, el]) - np.concatenate([el, z]))
    return (np.concatenate(mids), np.concatenate(moms), np.concatenate(nodes),
            np.concatenate(deltas))


def image_coeffs(eps_r, sigma, freq_hz, rx, ry, rz):
    """(rho_h, rho_v) image-current multipliers (portal/_portal.py:1397-1423)."""
    eps_c = eps_r - 1j * sigma / (2.0 * math.pi * freq_hz * EPS0)
    q = np.sqrt(eps_c - rx * rx - ry * ry)
    return (rz - q) / (rz + q), (eps_c * rz - q) / (eps_c * rz + q)


def far_moments(mid, moment, k, theta, phi, ground="free", ground_z=0.0, eps_r=None,
                sigma=None, freq_hz=None):
    """(M_theta, M_phi) of one column on the theta x phi grid; ground "free",
    "pec" or "finite" (portal/_portal.py:1520-1598)."""
    st, ct = np.sin(theta)[:, None], np.cos(theta)[:, None]
    sp, cp = np.sin(phi)[None, :], np.cos(phi)[None, :]
    r = np.stack(np.broadcast_arrays(st * cp, st * sp, ct), axis=-1)
    th_hat = np.stack(np.broadcast_arrays(ct * cp, ct * sp, -st), axis=-1)
    ph_hat = np.stack(np.broadcast_arrays(-sp, cp, 0.0 * st), axis=-1)

    def moments(c, m):
        return np.exp(1j * k * (r @ c.T)) @ m

    total = moments(mid, moment)
    if ground != "free":
        mid_i = mid.copy()
        mid_i[:, 2] = 2.0 * ground_z - mid[:, 2]
        m_img = moments(mid_i, moment * np.array([-1.0, -1.0, 1.0]))
        if ground == "pec":
            total = total + m_img
        else:
            v_hat = -th_hat
            rho_h, rho_v = image_coeffs(eps_r, sigma, freq_hz, r[..., 0], r[..., 1], r[..., 2])
            mv = np.sum(m_img * v_hat, axis=-1)
            mh = np.sum(m_img * ph_hat, axis=-1)
            total = total + (rho_v * mv)[..., None] * v_hat - (rho_h * mh)[..., None] * ph_hat
    return np.sum(total * th_hat, axis=-1), np.sum(total * ph_hat, axis=-1)


def element_fields(points, elements, k, radius, magnetic):
    """E or H at points from one column's elements (portal/_portal.py:1601-1651)."""
    mid, moment, nodes, delta = elements

    def geometry(src):
        rv = points[:, None, :] - src[None, :, :]
        rr = np.sqrt(np.sum(rv * rv, axis=-1) + radius * radius)
        return rv / rr[..., None], rr, np.exp(-1j * k * rr) / (4.0 * math.pi * rr)

    u, rr, g = geometry(mid)
    if magnetic:
        cr = np.cross(np.broadcast_to(moment, (len(points),) + moment.shape), u)
        return np.sum(((1j * k + 1.0 / rr) * g)[..., None] * cr, axis=1)
    e = -1j * ETA0 * k * np.sum(g[..., None] * moment[None], axis=1)
    qu, qr, qg = geometry(nodes)
    sc = -1j * ETA0 / k * (delta[None, :] * (1.0 + 1j * k * qr) / qr * qg)
    return e + np.sum(sc[..., None] * qu, axis=1)
