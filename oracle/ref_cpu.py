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

    def assemble(self, J, supp, polys, geom):
        td = geom["tangents"] @ geom["tangents"].T
        return self.acc.assemble_Z_bspline(J, supp.astype(np.int64), polys, td, self.omega,
                                           EPS0, MU0, self.degree, 0)

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
