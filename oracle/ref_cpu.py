"""The reference's CPU port solve, driven through its compiled accelerator.

Test / baseline infrastructure only.  Loads the reference's own C++ module
built from its sources by oracle/Makefile (oracle/_ref/_accelerators*.so) and
runs the reference's tensor route for compute_port_solution (bspline.py:
2222-2348, 2971-2989): all-pairs moments seg_seg_full_moments_bspline, the
same-edge overwrite with seg_seg_static_moments_bspline_uniform +
seg_seg_reg_moments_bspline_swept, assemble_Z_bspline, then LAPACK
(scipy.linalg.solve) over all port columns.  Geometry, basis and port columns
come from the oracle's restatement of the reference's Python set-up.

Used as bench.py's timed CPU baseline (kind "reference") and as a checker.
"""

import glob
import importlib.util
import os

import numpy as np
import scipy.linalg

from .bspline_oracle import EPS0, MU0, OracleSolver, gl01

HERE = os.path.dirname(os.path.abspath(__file__))


def load_ref_accel():
    paths = glob.glob(os.path.join(HERE, "_ref", "_accelerators*.so"))
    if not paths:
        return None
    spec = importlib.util.spec_from_file_location("_accelerators", paths[0])
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


class RefCpuSolver(OracleSolver):
    """OracleSolver whose fill and assembly run in the reference's C++."""

    def __init__(self, acc, **kw):
        super().__init__(**kw)
        self.acc = acc

    def J_blocks(self, geom, k):
        d = self.degree
        t01, w01 = gl01(self.n_qp_pair)
        a = float(self.radius[0])
        if not np.all(self.radius == a):
            raise NotImplementedError("uniform radius only")
        J = self.acc.seg_seg_full_moments_bspline(geom["seg_l"], geom["seg_r"], geom["seg_l"],
                                                  geom["seg_r"], a * a, float(k), d, t01, w01)
        for w, pw in enumerate(geom["per_wire"]):
            base = geom["seg_off"][w]
            for e in range(len(pw["ed_off"]) - 1):
                sl = slice(base + pw["ed_off"][e], base + pw["ed_off"][e + 1])
                arc = pw["ed_arc"][e]
                n_e = len(arc) - 1
                h = float(arc[1] - arc[0])
                A_st = self.acc.seg_seg_static_moments_bspline_uniform(h, a, n_e, d)
                sq = (arc[:-1, None] + t01[None, :] * (arc[1:] - arc[:-1])[:, None]).ravel()
                R = np.sqrt((sq[:, None] - sq[None, :]) ** 2 + a * a)
                hh = (arc[1:] - arc[:-1])[:, None]
                u, ww = t01[None, :] * hh, w01[None, :] * hh
                wu = np.stack([ww * u ** p for p in range(d + 1)])
                A_reg = self.acc.seg_seg_reg_moments_bspline_swept(R, wu, np.array([float(k)]))[0]
                J[:, :, sl, sl] = A_st + A_reg
        return J

    def assemble(self, J, supp, polys, geom, td=None):
        if td is None:
            td = geom["tangents"] @ geom["tangents"].T
        return self.acc.assemble_Z_bspline(J, supp.astype(np.int64), polys, td, self.omega,
                                           EPS0, MU0, self.degree, 0)

    # -- Sommerfeld ground through the reference's C++ (bspline.py:3730-3770) --
    def ground_terms(self, geom, supp, polys, remainder=True):
        """C2 * image + Q: mirrored-source moments and assembly in the
        reference's kernels, the grid from its six-integral engine
        (somm_six_integrals_batch), Q from sommerfeld_remainder_bspline_Q."""
        from . import sommerfeld_oracle as so
        d = self.degree
        t01, w01 = gl01(self.n_qp_pair)
        a = float(self.radius[0])
        eps_t = so.eps_tilde(self.ground_eps, self.omega, EPS0)
        c2 = so.image_coefficient(eps_t)
        gz = self.ground_z
        mir = np.array([1.0, 1.0, -1.0])
        img_l, img_r = geom["seg_l"] * mir, geom["seg_r"] * mir
        img_l[:, 2] += 2 * gz
        img_r[:, 2] += 2 * gz
        J = self.acc.seg_seg_full_moments_bspline(geom["seg_l"], geom["seg_r"], img_l, img_r,
                                                  a * a, float(self.k), d, t01, w01)
        td = geom["tangents"] @ (geom["tangents"] * mir).T
        out = c2 * self.assemble(J, supp, polys, geom, td=td)
        if not remainder:
            return out
        grid = self.ground_grid(geom, eps_t)
        q = self.n_qp_sommerfeld
        xg, wg = np.polynomial.legendre.leggauss(q)
        tq = 0.5 * (xg + 1.0)
        sl, sr, h = geom["seg_l"], geom["seg_r"], geom["h"]
        nodes = np.ascontiguousarray(sl[:, None, :] + tq[None, :, None] * (sr - sl)[:, None, :])
        W = np.ascontiguousarray((0.5 * h[:, None] * wg[None, :])[None]
                                 * (h[:, None] * tq[None, :])[None] ** np.arange(d + 1)[:, None, None])
        tang = np.ascontiguousarray(geom["tangents"])
        regs = grid.regions
        args = (float(grid.r1_max), float(grid.r_break), float(np.radians(30.0)), float(grid.r_near),
                np.array([r["r0"] for r in regs]), np.array([r["dr"] for r in regs]),
                np.array([r["th0"] for r in regs]), np.array([r["dth"] for r in regs]),
                [np.ascontiguousarray(r["vals"]) for r in regs])
        s64 = supp.astype(np.int64)
        Q = self.acc.sommerfeld_remainder_bspline_Q(nodes, tang, W, nodes, tang, W, s64, polys, s64,
                                                    polys, float(gz), float(self.k), *args, 0)
        return out + Q

    _GRIDS = {}

    def ground_grid(self, geom, eps_t):
        """The master grid filled by the reference's engine, cached like the
        reference's module cache (_sommerfeld.py:1081-1116)."""
        from . import sommerfeld_oracle as so
        lam = 2.0 * np.pi / self.k
        r1 = so.max_image_distance(geom["seg_l"], geom["seg_r"], self.ground_z)
        r1b = so.r1_bucket_wl(min(r1 / lam, so.R1_CAP_WL))
        eb = so.eps_bucket(complex(eps_t))
        key = (eb, r1b)
        if key not in self._GRIDS:
            self._GRIDS[key] = self.build_master(eb, r1b)
        return self._GRIDS[key].scaled(self.k, self.omega, MU0)

    def build_master(self, eps_b, r1b):
        from . import sommerfeld_oracle as so
        k2, om = so.K2_REF, so.K2_REF * so.C_LIGHT
        regions, scal = so.grid_layout(eps_b, k2, r1b)
        vals = []
        for i, reg in enumerate(regions):
            rr, tt = so.Grid.region_nodes(reg, i, scal[3])
            R, T = rr.ravel(), tt.ravel()
            S = np.zeros((4, R.size), dtype=complex)
            z = R == 0.0
            if z.any():
                S[:, z] = so.limits_r1_zero(eps_b, k2, T[z], om, so.MU0_EXACT)
            nz = ~z
            rho = np.maximum(R[nz] * np.cos(T[nz]), 0.0)
            hh = np.maximum(R[nz] * np.sin(T[nz]), 0.0)
            six = self.acc.somm_six_integrals_batch(complex(eps_b), k2, rho, hh, 1e-6, 0, 0)
            v_rr, v_zz, v_rz, v_r1, v, u = six.T
            k1s, k2s = k2 * k2 * eps_b, k2 * k2
            ph = so.c1_norm(k2, om, so.MU0_EXACT) * R[nz] * np.exp(1j * k2 * R[nz])
            S[:, nz] = [ph * k1s * v_rz, ph * k1s * (v_zz + k2s * v), ph * k2s * (v_rr + u),
                        -ph * k2s * (v_r1 + u)]
            vals.append(S.reshape((4,) + rr.shape))
        return so.Grid(regions, scal, vals, k2, om, so.MU0_EXACT, eps_b)

    def compute_port_solution(self):
        geom = self.geometry()
        basis = self.basis(geom)
        Z = self.Z(geom, basis)
        B, kcl = self.port_columns(geom, basis)
        if kcl.shape[0] == 0:
            X = scipy.linalg.solve(Z, B, overwrite_a=True)
        else:
            X = self.solve_kcl(Z, B, kcl)
        return B.T @ X, X
