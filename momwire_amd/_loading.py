"""Distributed series wire impedance: conductor loss and insulation loading
(reference _wire_loading.py:1-318, BSplineSolver._loading_gram /
_apply_loading at bspline.py:2667-2753).

Each wire w carries a series impedance per unit length

    Z'_w(omega) = Z'_int + j omega L'_ins,
    Z'_int = k_c / (2 pi a sigma) I0(k_c a) / I1(k_c a),  k_c = sqrt(j omega mu0 sigma),
    L'_ins = mu0 / (2 pi) (1 - 1/eps_r) ln(b / a),

and the Galerkin term is Z += sum_w Z'_w S_w with the k-independent Gram
S[m, n] = int Phi_m Phi_n dl over the segments two bases share on wire w
(closed-form polynomial moments h^(p+q+1)/(p+q+1)).  The Gram is built once
per structure; per frequency the host forms the merged complex entries
Z'_w S_w (a few per unknown) and the device adds them into Z after the
fill (mw_zadd_sparse).  The Bessel ratio uses scipy's exponentially scaled
``ive``, whose common factor cancels.
"""

from __future__ import annotations

import numpy as np
from scipy.special import ive

MU0 = 1.25663706127e-6   # _wire_loading.py:85


def wire_internal_impedance(omega, radius, sigma):
    """Internal impedance of a round solid conductor [Ohm/m] (_wire_loading.py:88-104)."""
    omega = np.asarray(omega, dtype=float)
    if radius <= 0.0:
        raise ValueError(f"wire radius must be > 0, got {radius}")
    if sigma <= 0.0:
        raise ValueError(f"conductivity must be > 0 S/m, got {sigma}")
    kc = np.sqrt(1j * omega * MU0 * sigma)
    z = kc * radius
    return kc / (2.0 * np.pi * radius * sigma) * (ive(0, z) / ive(1, z))


def insulation_inductance(radius, ins_radius, eps_r):
    """Series inductance of a dielectric jacket [H/m] (_wire_loading.py:107-117)."""
    if ins_radius <= radius:
        raise ValueError(f"insulation_radius ({ins_radius}) must exceed the conductor "
                         f"radius ({radius})")
    if eps_r < 1.0:
        raise ValueError(f"insulation_eps_r must be >= 1, got {eps_r}")
    return MU0 / (2.0 * np.pi) * (1.0 - 1.0 / eps_r) * np.log(ins_radius / radius)


def normalize_per_wire(value, n_wires, name):
    """None | scalar | length-n_wires sequence -> None | (n_wires,) floats
    (NaN switches a wire off) (_wire_loading.py:120-137)."""
    if value is None:
        return None
    arr = np.asarray(value, dtype=float)
    if arr.ndim == 0:
        arr = np.full(n_wires, float(arr))
    if arr.shape != (n_wires,):
        raise ValueError(f"{name}: expected a scalar or a length-{n_wires} sequence "
                         f"(one entry per wire), got shape {arr.shape}")
    return arr


class WireLoading:
    """Validated loading options of one solver (the checks and their order
    of _wire_loading.configure_loading, _wire_loading.py:186-240)."""

    def __init__(self, radius_per_wire, conductivity, insulation_radius, insulation_eps_r):
        n_w = len(radius_per_wire)
        self.radius = np.asarray(radius_per_wire, dtype=float)
        self.conductivity = normalize_per_wire(conductivity, n_w, "wire_conductivity")
        self.insulation_radius = normalize_per_wire(insulation_radius, n_w, "insulation_radius")
        self.insulation_eps_r = normalize_per_wire(insulation_eps_r, n_w, "insulation_eps_r")
        if (self.insulation_radius is None) != (self.insulation_eps_r is None):
            raise ValueError("insulation_radius and insulation_eps_r must be given together")
        if self.insulation_radius is not None:
            fin = np.isfinite(self.insulation_radius)
            if not np.array_equal(fin, np.isfinite(self.insulation_eps_r)):
                raise ValueError("insulation_radius and insulation_eps_r must be finite on the "
                                 "same wires (NaN switches a wire off in both)")
            for w in np.nonzero(fin)[0]:
                insulation_inductance(self.radius[w], self.insulation_radius[w],
                                      self.insulation_eps_r[w])
        if self.conductivity is not None:
            for w in np.nonzero(np.isfinite(self.conductivity))[0]:
                if self.conductivity[w] <= 0.0:
                    raise ValueError(f"wire_conductivity[{w}] must be > 0 S/m, "
                                     f"got {self.conductivity[w]}")
        self.active = self.conductivity is not None or self.insulation_radius is not None

    def z_wire(self, omega):
        """Z'_w(omega) per wire [Ohm/m] (_wire_loading.py:140-178)."""
        n_w = len(self.radius)
        out = np.zeros(n_w, dtype=np.complex128)
        for w in range(n_w):
            if self.conductivity is not None and np.isfinite(self.conductivity[w]):
                out[w] += wire_internal_impedance(omega, self.radius[w], self.conductivity[w])
            if self.insulation_radius is not None and np.isfinite(self.insulation_radius[w]):
                L = insulation_inductance(self.radius[w], self.insulation_radius[w],
                                          self.insulation_eps_r[w])
                out[w] += 1j * omega * L
        return out


def loading_gram(mesh):
    """(rows, cols, vals, wire) of S: per segment, the moments of every
    pair of bases living on it, C_m H C_n^T with H[p, q] = h^(p+q+1)/(p+q+1)
    (bspline.py:2667-2730); duplicate (row, col) entries are summed later."""
    seg_len = mesh.seg[:, 9]
    wire_of_seg = np.empty(mesh.n_seg, dtype=np.int64)
    for w, pw in enumerate(mesh.per_wire):
        wire_of_seg[pw["seg_off"]:pw["seg_off"] + pw["n"]] = w
    on_seg: dict[int, list[tuple[int, int]]] = {}
    nonzero = np.any(mesh.polys != 0.0, axis=2)
    for m in range(mesh.n_basis):
        for a in range(mesh.supp.shape[1]):
            if mesh.supp[m, a] >= 0 and nonzero[m, a]:
                on_seg.setdefault(int(mesh.supp[m, a]), []).append((m, a))
    pq = np.arange(mesh.polys.shape[2])
    pq_sum = pq[:, None] + pq[None, :] + 1
    rows, cols, vals, wires = [], [], [], []
    for s, ent in on_seg.items():
        H = seg_len[s] ** pq_sum / pq_sum
        C = mesh.polys[[m for m, _a in ent], [a for _m, a in ent], :]
        M = C @ H @ C.T
        for i, (mi, _a) in enumerate(ent):
            for j, (mj, _b) in enumerate(ent):
                rows.append(mi)
                cols.append(mj)
                vals.append(M[i, j])
                wires.append(wire_of_seg[s])
    return (np.asarray(rows, dtype=np.int64), np.asarray(cols, dtype=np.int64),
            np.asarray(vals, dtype=np.float64), np.asarray(wires, dtype=np.int64))


class LoadingTerm:
    """The Gram of one mesh with its unique (row, col) pattern, and the
    merged complex entries sum_w Z'_w S_w at a frequency."""

    def __init__(self, mesh, loading: WireLoading):
        self.loading = loading
        self.n_basis = mesh.n_basis
        self.r_full, self.c_full, self.vals, self.wires = loading_gram(mesh)
        key = self.r_full * mesh.n_basis + self.c_full
        pairs, self.inverse = np.unique(key, return_inverse=True)
        self.rows = (pairs // mesh.n_basis).astype(np.int32)
        self.cols = (pairs % mesh.n_basis).astype(np.int32)

    def entries(self, omega):
        """complex (nnz,) values aligned with (rows, cols)."""
        zw = self.loading.z_wire(omega)
        out = np.zeros(len(self.rows), dtype=np.complex128)
        np.add.at(out, self.inverse, zw[self.wires] * self.vals)
        return out

    def loss_power(self, coeffs, omega):
        """(P_wire, per wire): 1/2 sum_w Re Z'_w c^H S_w c, trailing entries
        past the unknowns ignored (bspline.py:2775-2800)."""
        r_w = np.real(self.loading.z_wire(omega))
        c = np.asarray(coeffs)[:self.n_basis]
        per = np.zeros(len(r_w))
        contrib = 0.5 * r_w[self.wires] * np.real(np.conj(c[self.r_full]) * c[self.c_full]) * self.vals
        np.add.at(per, self.wires, contrib)
        return float(per.sum()), per
