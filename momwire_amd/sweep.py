"""Batched port solves over many structures and frequencies in one graph.

``BatchedSweep(structures, k_array)`` builds one problem per (structure,
wavenumber) -- uneven sizes allowed -- and solves them all with the batched
kernels of the C ABI (``mw_sweep_plan_create`` / ``mw_sweep_run``): every
fill, factorisation and port solve of the sweep is one fixed sequence of
launches, captured once as a HIP graph and replayed.  Each problem computes
exactly what ``BSplineSolver.compute_port_solution()`` computes at that
wavenumber (reference src/momwire/bspline.py:4053-4179); the reference has
no cross-structure batching (it sweeps one structure at a time,
bspline.py:4181-4439).

Two forms of the factorisations exist, and a small autotuner picks the form
and its tile shape: every candidate is timed on the sweep itself and the
fastest is kept (per process, keyed by the sweep's size profile).
  0, 1  per-step launches over every active problem (about 4 per 32-column
        panel of the largest problem), with 64 x 32 or 128 x 64 update tiles;
  2..5  after the fill, ONE persistent launch in which each workgroup takes
        one problem at a time (largest first) through its whole solve:
        one-level or two-level (128-column outer blocks) blocking (bit 0),
        two or one workgroups per compute unit (bit 1).  Its cancel token is
        read by the kernel between problems.

Threshold pivoting keeps a panel's diagonal pivots while every multiplier
stays within ``pivot_threshold`` (|re| + |im|); 32 by default here, against
8 for one structure at a time.  In a batch, one problem whose panel falls
back to the tournament holds up the next step of all of them, and the
unpivoted multipliers of Galerkin wire matrices sit in the tens (on the
sweep's random structures, 50 % of panels exceed 8 and 7 % exceed 32, with
the same solution error).
"""

import ctypes
import time

import numpy as np

from . import _lib, _mesh
from ._lib import MwPorts, MwSweepProb
from .bspline import BSplineSolver, _Device, _torch

VARIANTS = (0, 1, 2, 3, 4, 5)

_TUNED = {}


class BatchedSweep:
    """All port solutions of ``structures`` x ``k_array`` on the GPU."""

    def __init__(self, structures, k_array, gemm_variant=None, cancel=None, pivot_threshold=32.0):
        self.solvers = [BSplineSolver(**kw) for kw in structures]
        for i, s in enumerate(self.solvers):
            if s.n_qp_pair != 4:
                raise NotImplementedError(f"structure {i}: n_qp_pair != 4 in the batched sweep")
            if s._loading_active:
                raise NotImplementedError(f"structure {i}: wire loading in the batched sweep "
                                          "(solve it with BSplineSolver one frequency at a time)")
            if s.ground_z is not None:
                raise NotImplementedError(f"structure {i}: a ground plane in the batched sweep "
                                          "(its fill is free space only; solve it with "
                                          "BSplineSolver one frequency at a time)")
        torch = _torch()
        lib = _lib.load()
        self.lib = lib
        self.torch = torch
        self.k_array = np.asarray(k_array, dtype=float)
        self.cancel = cancel
        self.pivot_threshold = float(pivot_threshold)
        dev = torch.device("cuda")
        n_mesh = len(self.solvers)
        meshes = (_lib.MwMesh * n_mesh)()
        ports = (MwPorts * n_mesh)()
        self._keep = []
        self.devs = []
        for i, s in enumerate(self.solvers):
            s._build()
            cols, con = s._port_plan()
            ek = _mesh.ek_tables(s._mesh) if s.extended_kernel else None
            D = _Device(s._mesh, cols + con, con, cols, ek=ek)
            if D.n_ports > 32 or D.n_con > 32 or D.n_ports + D.n_con > 64:
                raise NotImplementedError(
                    f"structure {i}: the batched sweep serves up to 32 ports and 32 KCL rows "
                    f"(64 together); got {D.n_ports} ports, {D.n_con} rows -- solve it with "
                    "BSplineSolver.port_solutions_swept one frequency at a time")
            self.devs.append(D)
            meshes[i] = D.mw
            ports[i] = MwPorts(
                n_ports=D.n_ports, n_con=D.n_con, col_max=D.col_max,
                con_max=D.con_max if D.n_con else 0, port_max=D.port_max, pad=0,
                col_idx=D.col_idx.data_ptr(), col_val=D.col_val.data_ptr(),
                con_idx=D.con_idx.data_ptr() if D.n_con else None,
                con_val=D.con_val.data_ptr() if D.n_con else None,
                port_idx=D.port_idx.data_ptr(), port_val=D.port_val.data_ptr())
        self.problems = []
        probs = (MwSweepProb * (n_mesh * len(self.k_array)))()
        idx = 0
        for i, (s, D) in enumerate(zip(self.solvers, self.devs)):
            n, ncols = D.nb, D.ncols
            npt = max(D.n_ports, 1)
            for kk in self.k_array:
                buf = dict(
                    A=torch.empty((n, n + ncols), dtype=torch.complex128, device=dev),
                    ipiv=torch.empty(n, dtype=torch.int32, device=dev),
                    work=torch.empty(int(lib.mw_getrf_workspace_bytes(n)), dtype=torch.uint8,
                                     device=dev),
                    info=torch.zeros(2, dtype=torch.int32, device=dev),
                    Y=torch.empty((npt, npt), dtype=torch.complex128, device=dev),
                    S=torch.empty((max(D.n_con, 1), max(D.n_con, 1)), dtype=torch.complex128,
                                  device=dev),
                    T=torch.empty((max(D.n_con, 1), npt), dtype=torch.complex128, device=dev),
                    se=torch.empty((max(len(s._mesh.se_edge), 1), 9), dtype=torch.complex128,
                                   device=dev),
                )
                pairs = D.keep["pairs"]
                if not s._mesh.symmetric:
                    # mixed radii: Z is not symmetric, every tile pair is filled
                    if "pairs_full" not in D.keep:
                        nt = len(s._mesh.tile_nseg)
                        full = np.stack(np.meshgrid(np.arange(nt), np.arange(nt), indexing="ij"),
                                        axis=-1).reshape(-1, 2).astype(np.int32)
                        D.keep["pairs_full"] = torch.from_numpy(full).to(dev)
                    pairs = D.keep["pairs_full"]
                probs[idx] = MwSweepProb(
                    mesh=i, symmetric=int(s._mesh.symmetric), k=float(kk), omega=float(kk * s.c),
                    A=buf["A"].data_ptr(), lda=n + ncols, n=n, ncols=ncols,
                    se_tab=buf["se"].data_ptr(), tile_pairs=pairs.data_ptr(),
                    n_pairs=int(pairs.shape[0]), n_se=len(s._mesh.se_edge),
                    ipiv=buf["ipiv"].data_ptr(), work=buf["work"].data_ptr(),
                    info=buf["info"].data_ptr(), Y=buf["Y"].data_ptr(), S=buf["S"].data_ptr(),
                    T=buf["T"].data_ptr())
                self.problems.append((i, float(kk), buf))
                idx += 1
        self._meshes, self._ports, self._probs = meshes, ports, probs
        self.eps, self.mu = self.solvers[0].eps, self.solvers[0].mu
        self._plan = {}
        self._graph = None
        sizes = tuple(sorted(D.nb for D in self.devs))
        key = (sizes, len(self.k_array))
        if gemm_variant is None:
            gemm_variant = _TUNED.get(key)
        if gemm_variant is None:
            gemm_variant = self.autotune()
            _TUNED[key] = gemm_variant
        self.gemm_variant = gemm_variant

    # ------------------------------------------------------------------
    def _plan_for(self, variant):
        if variant not in self._plan:
            h = ctypes.c_void_p()
            _lib.check(self.lib.mw_sweep_plan_create(len(self.solvers), ctypes.addressof(self._meshes),
                                                     ctypes.addressof(self._ports), len(self.problems),
                                                     ctypes.addressof(self._probs), int(variant),
                                                     self.pivot_threshold, ctypes.byref(h)),
                        "sweep plan")
            self._plan[variant] = h
        return self._plan[variant]

    def _launch(self, variant):
        st = self.torch.cuda.current_stream().cuda_stream
        cancel = self.cancel.address if self.cancel is not None else None
        _lib.check(self.lib.mw_sweep_run(self._plan_for(variant), self.eps, self.mu, cancel, st),
                   "sweep")

    def autotune(self, reps=2):
        """Time each form and tile variant on this sweep; keep the fastest."""
        torch = self.torch
        best, self.tuning = None, {}
        for variant in VARIANTS:
            self._launch(variant)  # warm
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                self._launch(variant)
            torch.cuda.synchronize()
            dt = (time.perf_counter() - t0) / reps
            self.tuning[variant] = dt
            if best is None or dt < self.tuning[best]:
                best = variant
        return best

    def run(self, graph=True):
        """Solve every problem; returns nothing (results stay on the device)."""
        if not graph:
            self._launch(self.gemm_variant)
            return
        torch = self.torch
        if self._graph is None:
            side = torch.cuda.Stream()
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):
                self._launch(self.gemm_variant)
            torch.cuda.current_stream().wait_stream(side)
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._launch(self.gemm_variant)
            self._graph = g
        self._graph.replay()

    def fill(self):
        """Only the fill: every problem's Z into its buffer (what the
        single-structure fill writes at that wavenumber); see z_matrix."""
        st = self.torch.cuda.current_stream().cuda_stream
        _lib.check(self.lib.mw_sweep_fill(self._plan_for(self.gemm_variant), self.eps, self.mu,
                                          st), "sweep fill")

    def z_matrix(self, index):
        """Z of problem ``index`` (after fill(); run() factors it in place)."""
        i, kk, buf = self.problems[index]
        n = self.devs[i].nb
        return buf["A"][:, :n].cpu().numpy()

    def y_matrices(self):
        """[(structure index, k, Y as numpy)] in problem order; raises on a
        singular system."""
        out = []
        for i, kk, buf in self.problems:
            info = buf["info"].cpu().numpy()
            if info[0] < 0:
                raise _lib.SolveAborted(f"structure {i} at k={kk}: sweep cancelled before it")
            if info[0] or info[1]:
                raise np.linalg.LinAlgError(f"structure {i} at k={kk}: singular system")
            npt = self.devs[i].n_ports
            out.append((i, kk, buf["Y"][:npt, :npt].cpu().numpy()))
        return out

    def close(self):
        for h in self._plan.values():
            self.lib.mw_sweep_plan_destroy(h)
        self._plan = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
