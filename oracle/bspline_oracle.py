"""numpy restatement of the reference's B-spline port-solution path (oracle).

Test infrastructure only (see oracle/__init__.py).  Each function names the
reference lines it follows; paths are relative to the reference checkout.

Scope: free space and the PEC, reflection-coefficient and Sommerfeld grounds,
the reduced or the extended thin-wire kernel, point delta-gap feeds,
junctions with KCL rows, junction ports and node gaps, degree 1 or 2 -- the
configuration space of the port-solution path.
"""

from math import comb  # noqa: F401  (kept for readers cross-checking weights)

import numpy as np
import scipy.linalg
from scipy.interpolate import BSpline

from .static_moments import static_moments_uniform

EPS0 = 8.8541878188e-12  # bspline.py:480
MU0 = 1.25663706127e-6   # bspline.py:481

_V_UNIT_INV = {  # bspline.py:112-115
    1: np.array([[1.0, 0.0], [-1.0, 1.0]]),
    2: np.array([[1.0, 0.0, 0.0], [-3.0, 4.0, -1.0], [2.0, -4.0, 2.0]]),
}


def gl01(n):
    """Gauss-Legendre nodes/weights mapped to [0, 1] (_bspline_kernels.py:634-636)."""
    xi, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (xi + 1.0), 0.5 * w


class OracleSolver:
    """Mirror of BSplineSolver's constructor surface for the path (bspline.py:511-982)."""

    def __init__(self, *, wires, n_per_edge_per_wire=None, degree=2, feeds=None,
                 feed_wire_index=0, feed_arclength=None, junctions=None,
                 junction_ports=None, node_gaps=None, n_qp_pair=4, wavelength=22,
                 wire_radius=0.0005, nsegs=101, ground_z=None, ground_eps=None,
                 ground_model="refl-coef", ground_phi_mode="normal", n_qp_sommerfeld=3,
                 extended_kernel=False, feed_model="point", feed_smoothing_factor=None,
                 n_qp_source=16, wire_conductivity=None, insulation_radius=None,
                 insulation_eps_r=None):
        self.degree = int(degree)
        self.extended_kernel = bool(extended_kernel)
        self.feed_model, self.feed_smoothing_factor = feed_model, feed_smoothing_factor
        self.n_qp_source = int(n_qp_source)
        self.loading_args = (wire_conductivity, insulation_radius, insulation_eps_r)
        self.ground_z, self.ground_eps = ground_z, ground_eps
        self.ground_model, self.n_qp_sommerfeld = ground_model, int(n_qp_sommerfeld)
        self.ground_phi_mode = ground_phi_mode
        self.wires = [np.asarray(w, dtype=float) for w in wires]
        n_w = len(self.wires)
        r = np.asarray(wire_radius, dtype=float)
        self.radius = np.full(n_w, float(r)) if r.ndim == 0 else r.copy()
        if n_per_edge_per_wire is None:
            n_per_edge_per_wire = [None] * n_w
        self.npe = []
        for pl, npe in zip(self.wires, n_per_edge_per_wire):
            ne = pl.shape[0] - 1
            if npe is None:
                npe = nsegs
            if np.isscalar(npe):
                npe = [int(npe)] * ne
            self.npe.append([int(x) for x in npe])
        if feeds is None:
            feeds = [(feed_wire_index, feed_arclength, 1.0)]
        self.feeds = [(int(w), None if s is None else float(s), complex(v)) for w, s, v in feeds]
        self.junctions = [list(j) for j in (junctions or [])]
        self.junction_ports = [(int(p), 0j) if np.isscalar(p) else (int(p[0]), complex(p[1]))
                               for p in (junction_ports or [])]
        self.node_gaps = [(int(w), e, complex(v)) for w, e, v in (node_gaps or [])]
        self.n_qp_pair = int(n_qp_pair)
        # bspline.py:672-676
        self.c = 1.0 / np.sqrt(EPS0 * MU0)
        self.set_wavelength(wavelength)

    def set_wavelength(self, wavelength):
        self.wavelength = wavelength
        self.freq = self.c / wavelength
        self.omega = 2 * np.pi * self.freq
        self.k = self.omega / self.c

    def set_k(self, k):  # _port_solution.py:101-105
        self.k = float(k)
        self.omega = self.k * self.c
        self.wavelength = self.c / (self.omega / (2 * np.pi))

    # -- geometry: bspline.py:988-1093 ------------------------------------
    def geometry(self):
        per_wire, seg_off = [], [0]
        seg_l, seg_r, tan, hs, edge_id, edge_idx = [], [], [], [], [], []
        eid = 0
        for pl, npe in zip(self.wires, self.npe):
            sl_w, sr_w, h_w, ed_off, ed_arc = [], [], [], [0], []
            for e in range(pl.shape[0] - 1):
                p0, p1 = pl[e], pl[e + 1]
                L = float(np.linalg.norm(p1 - p0))
                n_e = npe[e]
                t = np.linspace(0.0, 1.0, n_e + 1)
                pts = (1 - t[:, None]) * p0[None, :] + t[:, None] * p1[None, :]
                sl_w.append(pts[:-1])
                sr_w.append(pts[1:])
                tan.append(np.tile((p1 - p0) / L, (n_e, 1)))
                h_w.append(np.full(n_e, L / n_e))
                ed_arc.append(np.linspace(0.0, L, n_e + 1))
                ed_off.append(ed_off[-1] + n_e)
                edge_id.append(np.full(n_e, eid))
                edge_idx.append(np.arange(n_e))
                eid += 1
            h_w = np.concatenate(h_w)
            per_wire.append(dict(h=h_w, arc=np.concatenate([[0.0], np.cumsum(h_w)]),
                                 ed_off=ed_off, ed_arc=ed_arc, n=len(h_w)))
            seg_l += sl_w
            seg_r += sr_w
            hs.append(h_w)
            seg_off.append(seg_off[-1] + len(h_w))
        return dict(per_wire=per_wire, seg_off=seg_off, seg_l=np.vstack(seg_l),
                    seg_r=np.vstack(seg_r), tangents=np.vstack(tan),
                    h=np.concatenate(hs), n_segs=seg_off[-1])

    def seg_radius(self, geom):  # bspline.py:2032-2038
        return np.repeat(self.radius, np.diff(geom["seg_off"]))

    # -- basis: bspline.py:1216-1414 ----------------------------------------
    @staticmethod
    def touch_tol(pl):  # _ground_spec.py:127-153
        return 1e-6 * max(float(np.sum(np.linalg.norm(np.diff(pl, axis=0), axis=1))), 1e-30)

    def endpoint_status(self):  # bspline.py:1108-1151
        """'free', a junction index, or 'ground' (an unjunctioned end in
        the plane)."""
        n_w = len(self.wires)
        st, en = ["free"] * n_w, ["free"] * n_w
        for j, jw in enumerate(self.junctions):
            for w, end in jw:
                if end == "start":
                    st[w] = j
                else:
                    en[w] = j
        if self.ground_z is not None:
            for w, pl in enumerate(self.wires):
                at = np.abs(pl[:, 2] - self.ground_z) <= self.touch_tol(pl)
                if st[w] == "free" and at[0]:
                    st[w] = "ground"
                if en[w] == "free" and at[-1]:
                    en[w] = "ground"
        return st, en

    def grounded_junctions(self):  # bspline.py:1152-1169
        if self.ground_z is None:
            return set()
        out = set()
        for j, jw in enumerate(self.junctions):
            w, end = jw[0]
            pl = self.wires[w]
            pt = pl[0] if end == "start" else pl[-1]
            if abs(pt[2] - self.ground_z) <= self.touch_tol(pl):
                out.add(j)
        return out

    def basis(self, geom):
        d = self.degree
        st, en = self.endpoint_status()
        supp_all, polys_all, knots_all, wbg = [], [], [], []
        jdirs = {j: [] for j in range(len(self.junctions))}
        m = 0
        for w, pw in enumerate(geom["per_wire"]):
            arc = pw["arc"]
            knots = np.concatenate([np.zeros(d), arc, np.full(d, arc[-1])])
            knots_all.append(knots)
            nb_w = len(knots) - d - 1
            kept = []
            if st[w] == "ground":
                kept.append((0, "gnd", None, "start"))
            elif st[w] != "free":
                kept.append((0, "dir", st[w], "start"))
            for j in range(1, nb_w - 1):
                kept.append((j, "int", None, None))
            if en[w] == "ground":
                kept.append((nb_w - 1, "gnd", None, "end"))
            elif en[w] != "free":
                kept.append((nb_w - 1, "dir", en[w], "end"))
            n_w = pw["n"]
            unit = np.linspace(0.0, 1.0, d + 1)
            u_glob = (arc[:-1, None] + pw["h"][:, None] * unit[None, :]).reshape(-1)
            DM = BSpline.design_matrix(u_glob, knots, d).toarray().reshape(n_w, d + 1, nb_w)
            l2g = {}
            for kidx, (j, kind, jidx, endp) in enumerate(kept):
                lo, hi = max(0, j - d), min(n_w, j + 1)
                s_m = np.zeros(d + 1, dtype=np.int64)
                p_m = np.zeros((d + 1, d + 1))
                for a, s in enumerate(range(lo, hi)):
                    s_m[a] = geom["seg_off"][w] + s
                    coef = _V_UNIT_INV[d] @ DM[s, :, j]
                    p_m[a] = coef * pw["h"][s] ** (-np.arange(d + 1))
                supp_all.append(s_m)
                polys_all.append(p_m)
                l2g[kidx] = m
                if kind == "dir":
                    jdirs[jidx].append((m, 1.0 if endp == "start" else -1.0))
                m += 1
            wbg.append((kept, l2g))
        supp = np.array(supp_all, dtype=np.int64)
        polys = np.array(polys_all)
        kcl = np.zeros((len(self.junctions), m))
        for j in range(len(self.junctions)):
            for mg, sgn in jdirs[j]:
                kcl[j, mg] = sgn
        return supp, polys, kcl, knots_all, wbg

    # -- moments ------------------------------------------------------------
    def full_moments(self, geom, a_row, k, il=None, jl=None, src=None, ek=None):
        """_bspline_kernels.py:716-747 (numpy body of the off-edge moments);
        ``src`` = (seg_l, seg_r) of other source segments (the mirrored ones);
        ``ek`` = (labels_i, labels_j): pairs with equal labels take the
        extended kernel's factor (_bspline_kernels.py:599-747)."""
        d = self.degree
        t01, w01 = gl01(self.n_qp_pair)
        sl, sr = geom["seg_l"], geom["seg_r"]
        il = slice(None) if il is None else il
        jl = slice(None) if jl is None else jl
        sli, sri = sl[il], sr[il]
        slj, srj = (sl[jl], sr[jl]) if src is None else (src[0][jl], src[1][jl])
        a = np.asarray(a_row)[il]
        Li = np.linalg.norm(sri - sli, axis=1)
        Lj = np.linalg.norm(srj - slj, axis=1)
        pi = (1 - t01[None, :, None]) * sli[:, None, :] + t01[None, :, None] * sri[:, None, :]
        pj = (1 - t01[None, :, None]) * slj[:, None, :] + t01[None, :, None] * srj[:, None, :]
        diff = pi[:, :, None, None, :] - pj[None, None, :, :, :]
        R = np.sqrt((diff * diff).sum(-1) + (a * a)[:, None, None, None])
        G = np.exp(-1j * k * R) / (4 * np.pi * R)
        if ek is not None:
            from . import ek_oracle
            gi, gj = np.asarray(ek[0])[il], np.asarray(ek[1])[jl]
            mask = (gi[:, None] == gj[None, :])[:, None, :, None]
            G = np.where(mask, G * ek_oracle.factor(R, a[:, None, None, None], k), G)
        ui, uj = t01[None, :] * Li[:, None], t01[None, :] * Lj[:, None]
        wi, wj = w01[None, :] * Li[:, None], w01[None, :] * Lj[:, None]
        wui = np.stack([wi * ui ** p for p in range(d + 1)])
        wuj = np.stack([wj * uj ** p for p in range(d + 1)])
        return np.einsum("piq,iqjr,Pjr->pPij", wui, G, wuj)

    def reg_moments(self, ed_arc, a, k):
        """_bspline_kernels.py:425-517 (smooth remainder on one edge; the
        extended kernel's remainder under extended_kernel)."""
        d = self.degree
        t01, w01 = gl01(self.n_qp_pair)
        sl = ed_arc[:-1]
        h = ed_arc[1:] - sl
        N = len(sl)
        s = (sl[:, None] + t01[None, :] * h[:, None]).ravel()
        u = t01[None, :] * h[:, None]
        w = w01[None, :] * h[:, None]
        R = np.sqrt((s[:, None] - s[None, :]) ** 2 + a * a)
        if self.extended_kernel:
            from . import ek_oracle
            G = ek_oracle.reg_kernel(R, a, k).reshape(N, -1, N, len(t01))
        else:
            G = ((np.exp(-1j * k * R) - 1.0) / (4 * np.pi * R)).reshape(N, -1, N, len(t01))
        wu = np.stack([w * u ** p for p in range(d + 1)])
        return np.einsum("piq,iqjr,Pjr->pPij", wu, G, wu)

    def J_blocks(self, geom, k):
        """bspline.py:2222-2283: all pairs, then same-edge overwrite."""
        d = self.degree
        a_row = self.seg_radius(geom)
        J = self.full_moments(geom, a_row, k, ek=self.ek_labels(geom))
        for w, pw in enumerate(geom["per_wire"]):
            base = geom["seg_off"][w]
            for e in range(len(pw["ed_off"]) - 1):
                sl = slice(base + pw["ed_off"][e], base + pw["ed_off"][e + 1])
                arc = pw["ed_arc"][e]
                n_e = len(arc) - 1
                h = float(arc[1] - arc[0])
                A_st = static_moments_uniform(h, float(self.radius[w]), n_e, d,
                                              ek=self.extended_kernel)
                A_reg = self.reg_moments(arc, float(self.radius[w]), k)
                J[:, :, sl, sl] = A_st + A_reg
        return J

    def assemble(self, J, supp, polys, geom, td=None, wphi=None):
        """bspline.py:2322-2348 (numpy einsum assembly); with per-segment-pair
        weights td on the A term and wphi on the charge term, the weighted
        image assembly of bspline.py:1700-1740."""
        d = self.degree
        if td is None:
            td = geom["tangents"] @ geom["tangents"].T
        nb = supp.shape[0]
        ZA = np.zeros((nb, nb), dtype=complex)
        ZP = np.zeros((nb, nb), dtype=complex)
        pv = np.arange(1, d + 1, dtype=float)
        for a in range(d + 1):
            sm = supp[:, a]
            for b in range(d + 1):
                sn = supp[:, b]
                Jb = J[:, :, sm[:, None], sn[None, :]]
                ZA += td[sm[:, None], sn[None, :]] * np.einsum(
                    "mp,pPmn,nP->mn", polys[:, a, :], Jb, polys[:, b, :])
                dm = polys[:, a, 1:] * pv[None, :]
                dn = polys[:, b, 1:] * pv[None, :]
                zp = np.einsum("mp,pPmn,nP->mn", dm, Jb[:d, :d], dn)
                ZP += zp if wphi is None else wphi[sm[:, None], sn[None, :]] * zp
        return 1j * self.omega * MU0 * ZA + ZP / (1j * self.omega * EPS0)

    def Z(self, geom=None, basis=None, remainder=True):
        geom = geom or self.geometry()
        supp, polys = (basis or self.basis(geom))[:2]
        Z = self.assemble(self.J_blocks(geom, self.k), supp, polys, geom)
        if self.ground_z is not None:
            Z = Z - self.ground_terms(geom, supp, polys, remainder)
        if any(a is not None for a in self.loading_args):
            Z = Z + self.loading_term(geom, supp, polys)
        return Z

    # -- wire loading: bspline.py:2667-2753, _wire_loading.py:88-178 ---------
    def z_wire(self, omega):
        """Series impedance per unit length of each wire (NaN: off)."""
        from scipy.special import ive
        mu0 = 1.25663706127e-6
        n_w = len(self.wires)
        per = lambda v: None if v is None else (np.full(n_w, float(v)) if np.ndim(v) == 0  # noqa: E731
                                                else np.asarray(v, dtype=float))
        sig, b, er = (per(v) for v in self.loading_args)
        out = np.zeros(n_w, dtype=complex)
        for w in range(n_w):
            a = self.radius[w]
            if sig is not None and np.isfinite(sig[w]):
                kc = np.sqrt(1j * omega * mu0 * sig[w])
                out[w] += kc / (2 * np.pi * a * sig[w]) * ive(0, kc * a) / ive(1, kc * a)
            if b is not None and np.isfinite(b[w]):
                out[w] += 1j * omega * mu0 / (2 * np.pi) * (1 - 1 / er[w]) * np.log(b[w] / a)
        return out

    def loading_gram(self, geom, supp, polys):
        """(rows, cols, vals, wire) of int Phi_m Phi_n over shared segments."""
        on_seg = {}
        nz = np.any(polys != 0.0, axis=2)
        for m in range(supp.shape[0]):
            for a in range(supp.shape[1]):
                if nz[m, a]:
                    on_seg.setdefault(int(supp[m, a]), []).append((m, a))
        pq = np.arange(polys.shape[2])
        ps = pq[:, None] + pq[None, :] + 1
        out = []
        for sg, ent in on_seg.items():
            M = polys[[m for m, _ in ent], [a for _, a in ent]] @ (geom["h"][sg] ** ps / ps) \
                @ polys[[m for m, _ in ent], [a for _, a in ent]].T
            w = int(np.searchsorted(geom["seg_off"], sg, side="right") - 1)
            out += [(mi, mj, M[i, j], w) for i, (mi, _a) in enumerate(ent)
                    for j, (mj, _b) in enumerate(ent)]
        r, c, v, w = (np.array(x) for x in zip(*out))
        return r.astype(int), c.astype(int), v.astype(float), w.astype(int)

    def loading_term(self, geom, supp, polys):
        r, c, v, w = self.loading_gram(geom, supp, polys)
        L = np.zeros((supp.shape[0], supp.shape[0]), dtype=complex)
        np.add.at(L, (r, c), self.z_wire(self.omega)[w] * v)
        return L

    # -- ground: bspline.py:3730-3770, 2552-2661, 1654-1740, 1791-1874 -------
    def ground_terms(self, geom, supp, polys, remainder=True):
        """The block the grounded fill subtracts: the PEC image (no
        ground_eps), C2 * image + Q (Sommerfeld), or the Fresnel-weighted
        image (reflection coefficient, _ground_refl.py)."""
        from . import refl_oracle as ro
        from . import sommerfeld_oracle as so
        gz = self.ground_z
        mir = np.array([1.0, 1.0, -1.0])
        img_l, img_r = geom["seg_l"] * mir, geom["seg_r"] * mir
        img_l[:, 2] += 2 * gz
        img_r[:, 2] += 2 * gz
        J_img = self.full_moments(geom, self.seg_radius(geom), self.k, src=(img_l, img_r),
                                  ek=self.ek_labels(geom, mirror=True))
        td_img = geom["tangents"] @ (geom["tangents"] * mir).T
        if self.ground_eps is None:
            return self.assemble(J_img, supp, polys, geom, td=td_img)
        if self.ground_model == "refl-coef":
            eps_t = ro.eps_tilde(self.ground_eps, self.omega, EPS0)
            centers = 0.5 * (geom["seg_l"] + geom["seg_r"])
            w_A, w_Phi = ro.weights(centers, geom["tangents"], gz, eps_t, self.ground_phi_mode)
            return self.assemble(J_img, supp, polys, geom, td=w_A, wphi=w_Phi)
        eps_t = so.eps_tilde(self.ground_eps, self.omega, EPS0)
        c2 = so.image_coefficient(eps_t)
        out = c2 * self.assemble(J_img, supp, polys, geom, td=td_img)
        if remainder:
            grid = self.ground_grid(geom, eps_t)
            out = out + so.remainder_block(geom, supp, polys, gz, self.k, grid,
                                           self.n_qp_sommerfeld)
        return out

    def ek_labels(self, geom, mirror=False):
        """bspline.py:2040-2104: coaxial labels (observer, source), or None
        for the reduced kernel.  For the image block one scan over the real
        and the mirrored segments keeps both sides' labels comparable."""
        if not self.extended_kernel:
            return None
        from . import ek_oracle
        a = self.seg_radius(geom)
        if not mirror:
            g = ek_oracle.axis_groups(geom["seg_l"], geom["seg_r"], geom["tangents"], a)
            return g, g
        mir = np.array([1.0, 1.0, -1.0])
        img_l, img_r = geom["seg_l"] * mir, geom["seg_r"] * mir
        img_l[:, 2] += 2 * self.ground_z
        img_r[:, 2] += 2 * self.ground_z
        n = len(a)
        g = ek_oracle.axis_groups(np.vstack([geom["seg_l"], img_l]),
                                  np.vstack([geom["seg_r"], img_r]),
                                  np.vstack([geom["tangents"], geom["tangents"] * mir]),
                                  np.concatenate([a, a]))
        return g[:n], g[n:]

    def ground_grid(self, geom, eps_t):
        from . import sommerfeld_oracle as so
        r1 = so.max_image_distance(geom["seg_l"], geom["seg_r"], self.ground_z)
        return so.get_grid(eps_t, self.k, r1, self.omega, MU0)

    # -- ports: bspline.py:2806-2870, 3784-3867, 1170-1210 ------------------
    def source_weights(self, arc, knots, s_f):
        """bspline.py:2806-2942: the feed's weights on every basis of its wire
        -- point delta gap, segment gap (mean over the cell holding s_f) or
        the cos^2 bump of width alpha h, by n_qp_source-point Gauss-Legendre."""
        d = self.degree
        if self.feed_smoothing_factor is None and self.feed_model == "point":
            return BSpline.design_matrix(np.array([s_f]), knots, d).toarray()[0]
        cell = int(np.searchsorted(arc, s_f, side="right")) - 1
        xi, wq = np.polynomial.legendre.leggauss(self.n_qp_source)
        if self.feed_model == "segment":
            cell = max(0, min(cell, len(arc) - 2))
            lo, hi = float(arc[cell]), float(arc[cell + 1])
            t = 0.5 * (hi + lo) + 0.5 * (hi - lo) * xi
            return BSpline.design_matrix(t, knots, d).toarray().T @ ((0.5 * (hi - lo) * wq) / (hi - lo))
        cell = max(0, min(cell, len(arc) - 2))
        w = float(self.feed_smoothing_factor) * float(arc[cell + 1] - arc[cell])
        lo, hi = max(0.0, s_f - w / 2), min(float(arc[-1]), s_f + w / 2)
        if lo >= hi:
            raise ValueError("feed_smoothing_factor too large for wire")
        t = 0.5 * (hi + lo) + 0.5 * (hi - lo) * xi
        g = np.where(np.abs(t - s_f) < w / 2, (2.0 / w) * np.cos(np.pi * (t - s_f) / w) ** 2, 0.0)
        return np.einsum("qj,q,q->j", BSpline.design_matrix(t, knots, d).toarray(), g,
                         0.5 * (hi - lo) * wq)

    def port_columns(self, geom, basis):
        supp, polys, kcl, knots, wbg = basis
        d = self.degree
        nb = supp.shape[0]
        B = np.zeros((nb, len(self.feeds)), dtype=complex)
        for f, (w, s_f, _v) in enumerate(self.feeds):
            arc = geom["per_wire"][w]["arc"]
            s_f = arc[-1] / 2.0 if s_f is None else s_f
            DM = self.source_weights(arc, knots[w], s_f)
            kept, l2g = wbg[w]
            for kidx, (j, *_r) in enumerate(kept):
                B[l2g[kidx], f] = DM[j]
        port_rows = [j for j, _v in self.junction_ports]
        grounded = self.grounded_junctions()
        keep = [r for r in range(kcl.shape[0]) if r not in set(port_rows) and r not in grounded]
        kcl_con, port_A = kcl[keep], kcl[port_rows]
        gap = np.zeros((nb, len(self.node_gaps)))
        end_to_j = {tuple(mm): j for j, jw in enumerate(self.junctions) for mm in jw}
        for p, (w, end, _v) in enumerate(self.node_gaps):
            jidx = end_to_j[(w, end)]
            kept, l2g = wbg[w]
            for kidx, (_j, kind, ji, endp) in enumerate(kept):
                if kind == "dir" and ji == jidx and endp == end:
                    gap[l2g[kidx], p] = 1.0 if end == "start" else -1.0
        return np.hstack([B, port_A.T.astype(complex), gap.astype(complex)]), kcl_con

    @staticmethod
    def solve_kcl(Z, V, A):  # bspline.py:2971-2989
        if A.shape[0] == 0:
            return scipy.linalg.solve(Z, V)
        sol = scipy.linalg.solve(Z, np.hstack([V, A.T.astype(complex)]))
        W, X = sol[:, :V.shape[1]], sol[:, V.shape[1]:]
        return W - X @ scipy.linalg.solve(A @ X, A @ W)

    def compute_port_solution(self):
        """bspline.py:4053-4179 -> (y, coeffs)."""
        geom = self.geometry()
        basis = self.basis(geom)
        Z = self.Z(geom, basis)
        B, kcl_con = self.port_columns(geom, basis)
        X = self.solve_kcl(Z, B, kcl_con)
        return B.T @ X, X

    def currents_at_knots(self, coeffs):  # bspline.py:4603-4685 (no enrichment)
        geom = self.geometry()
        _s, _p, _k, knots, wbg = self.basis(geom)
        d = self.degree
        out = []
        for w, pw in enumerate(geom["per_wire"]):
            kv = knots[w]
            s_eval = np.clip(pw["arc"], kv[0], kv[-1])
            c = np.zeros(len(kv) - d - 1, dtype=complex)
            kept, l2g = wbg[w]
            for kidx, (j, *_r) in enumerate(kept):
                c[j] = coeffs[l2g[kidx]]
            out.append(BSpline.design_matrix(s_eval, kv, d).toarray() @ c)
        return out
