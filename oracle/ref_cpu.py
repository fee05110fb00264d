"""The reference's CPU port solve, driven through its compiled accelerator.

Test / baseline infrastructure only.  Loads the reference's own C++ module
built from its sources by oracle/Makefile (oracle/_ref/_accelerators*.so) and
runs the reference's tensor route for compute_port_solution (bspline.py:
2222-2348, 2971-2989): all-pairs moments seg_seg_full_moments_bspline, the
same-edge overwrite with seg_seg_static_moments_bspline_uniform +
seg_seg_reg_moments_bspline_swept, assemble_Z_bspline, then LAPACK
(scipy.linalg.solve) over all port columns.  Geometry, basis and port columns
come from the oracle's restatement of the reference's Python set-up.

Where the reference itself would not hold the whole moment tensor -- its
(d+1)^2 N^2 complex entries over the swept_mem_mb budget, 256 MiB by default
(bspline.py:3695-3700) -- it takes the chunked route instead
(bspline.py:2386-2550): observer-row chunks of the moments accumulated into a
Fortran-order Z by assemble_Z_bspline_windowed, then per-edge corrections
(static + smooth remainder minus the full-kernel block), and LAPACK factors Z
in place.  ``Z`` follows the same dispatch.

Used as bench.py's timed CPU baseline (kind "reference") and as a checker.
"""

import glob
import importlib.util
import os
import time

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

    def __init__(self, acc, swept_mem_mb=256, **kw):
        super().__init__(**kw)
        self.acc = acc
        self.swept_mem_mb = int(swept_mem_mb)  # the reference's default (bspline.py:511-982)

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

    def tensor_fits_budget(self, n_segs):
        """The reference's route choice (bspline.py:3695-3700)."""
        return (self.degree + 1) ** 2 * int(n_segs) ** 2 * 16 <= (self.swept_mem_mb << 20)

    def Z(self, geom=None, basis=None, remainder=True):
        geom = geom or self.geometry()
        basis = basis or self.basis(geom)
        if self.tensor_fits_budget(len(geom["seg_l"])):
            return super().Z(geom, basis, remainder)
        supp, polys = basis[:2]
        Z = self.Z_chunked(geom, supp, polys)
        if self.ground_z is not None:
            Z -= self.ground_terms(geom, supp, polys, remainder)
        if any(a is not None for a in self.loading_args):
            Z += self.loading_term(geom, supp, polys)
        return Z

    def Z_chunked(self, geom, supp, polys):
        """Free-space Z on the reference's chunked route (bspline.py:2386-2550
        -> assemble_Z_bspline_windowed, _accelerators.cpp:1325-1443)."""
        d = self.degree
        t01, w01 = gl01(self.n_qp_pair)
        a = float(self.radius[0])
        if not np.all(self.radius == a):
            raise NotImplementedError("uniform radius only")
        seg_l, seg_r = geom["seg_l"], geom["seg_r"]
        n_segs = len(seg_l)
        n_b = supp.shape[0]
        Z = np.zeros((n_b, n_b), dtype=np.complex128, order="F")
        supp_c = np.ascontiguousarray(supp, dtype=np.int64)
        polys_c = np.ascontiguousarray(polys, dtype=np.float64)
        tan_c = np.ascontiguousarray(geom["tangents"], dtype=np.float64)
        all_n = np.arange(n_b, dtype=np.int64)
        om, acc = float(self.omega), self.acc

        def touching(lo, hi):
            return np.nonzero(((supp_c >= lo) & (supp_c < hi)).any(axis=1))[0].astype(np.int64)

        def accumulate(Jw, i0, i1, j0, j1, m_idx, n_idx):
            acc.assemble_Z_bspline_windowed(Jw, supp_c, polys_c, tan_c, m_idx, n_idx, int(i0),
                                            int(i1), int(j0), int(j1), om, EPS0, MU0, Z, 0)

        row_bytes = (d + 1) ** 2 * n_segs * 16
        chunk = max(1, int(self.swept_mem_mb * 1024 * 1024 // row_bytes))
        for i0 in range(0, n_segs, chunk):
            i1 = min(i0 + chunk, n_segs)
            Jc = acc.seg_seg_full_moments_bspline(seg_l[i0:i1], seg_r[i0:i1], seg_l, seg_r, a * a,
                                                  float(self.k), d, t01, w01)
            accumulate(Jc, i0, i1, 0, n_segs, touching(i0, i1), all_n)
            del Jc
        for w, pw in enumerate(geom["per_wire"]):
            base = geom["seg_off"][w]
            for e in range(len(pw["ed_off"]) - 1):
                lo, hi = base + pw["ed_off"][e], base + pw["ed_off"][e + 1]
                arc = pw["ed_arc"][e]
                n_e = len(arc) - 1
                h = float(arc[1] - arc[0])
                A_st = acc.seg_seg_static_moments_bspline_uniform(h, a, n_e, d)
                sq = (arc[:-1, None] + t01[None, :] * (arc[1:] - arc[:-1])[:, None]).ravel()
                R = np.sqrt((sq[:, None] - sq[None, :]) ** 2 + a * a)
                hh = (arc[1:] - arc[:-1])[:, None]
                u, ww = t01[None, :] * hh, w01[None, :] * hh
                wu = np.stack([ww * u ** p for p in range(d + 1)])
                A_reg = acc.seg_seg_reg_moments_bspline_swept(R, wu, np.array([float(self.k)]))[0]
                J_edge = acc.seg_seg_full_moments_bspline(seg_l[lo:hi], seg_r[lo:hi], seg_l[lo:hi],
                                                          seg_r[lo:hi], a * a, float(self.k), d,
                                                          t01, w01)
                corr = np.ascontiguousarray((A_st + A_reg) - J_edge)
                idx = touching(lo, hi)
                accumulate(corr, lo, hi, lo, hi, idx, idx)
        return Z

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

    def prepared(self):
        """Geometry, basis and port columns: k-independent, built once and
        reused across solves as the reference caches them (bspline.py:1004-1014)."""
        if getattr(self, "_prep", None) is None:
            geom = self.geometry()
            basis = self.basis(geom)
            self._prep = (geom, basis, self.port_columns(geom, basis))
        return self._prep

    def compute_port_solution(self, timings=None):
        geom, basis, (B, kcl) = self.prepared()
        t0 = time.perf_counter()
        Z = self.Z(geom, basis)
        if timings is not None:
            timings["fill_s"] = time.perf_counter() - t0
        if kcl.shape[0] == 0:
            # the reference solves in place on its Fortran-order Z (bspline.py:2984)
            X = scipy.linalg.solve(Z, B, overwrite_a=True)
        else:
            X = self.solve_kcl(Z, B, kcl)
        Y = B.T @ X
        if timings is not None:
            timings["solve_s"] = time.perf_counter() - t0 - timings["fill_s"]
        return Y, X
