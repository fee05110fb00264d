"""BSplineSolver: the reference's degree-2 B-spline Galerkin port solve, on MI355X.

Drop-in for ``momwire.BSplineSolver(...).compute_port_solution()`` (reference
src/momwire/bspline.py:4053-4179): the same constructor keywords for the
supported configuration space, the same port order ([feeds..., junction
ports..., node gaps...]), the same ``PortSolution`` result.  The Z fill, the
LU factorisation, the all-port solve, the KCL Schur step and the Y readout run
as HIP kernels on the GPU (momwire_amd/csrc); Python only prepares the
k-independent tables and moves results.  There is no CPU fallback.

Supported: free space, PEC ground, the reflection-coefficient finite ground
(``ground_model="refl-coef"``, the default with ``ground_eps``: the image
weighted per segment pair by Fresnel coefficients, any ``ground_phi_mode``)
and the Sommerfeld ground (``ground_model="sommerfeld"``: C2-scaled exact
image plus the interpolated remainder); ground contact (wire ends and
junctions in the plane, PEC or Sommerfeld; refused under the reflection-
coefficient ground as the reference refuses it); degree 2 and 1 (tents);
n_qp_pair 1 .. 8;
point delta-gap, segment-gap and cos^2-smoothed feeds; K-wire junctions (KCL
rows, any number), junction ports and node gaps; per-wire radii;
distributed wire loading (conductor loss, insulation).
The extended thin-wire kernel (``extended_kernel=True``) on coaxial
equal-radius pairs, in free space and over every ground (the Sommerfeld
remainder stays reduced, as in the reference).
Refused with NotImplementedError: singular enrichment, degree > 2,
n_qp_pair or n_qp_sommerfeld > 8.
"""

import ctypes

import numpy as np

from . import _ground, _lib
from . import _loading, _mesh
from ._mesh import (EPS0, MU0, build_mesh, bspline_slopes, bspline_values, feed_column,
                    sparse_pack)
from ._port_solution import PortSolution, _SweptPortSolutions
from ._lib import SolveAborted  # noqa: F401  (re-exported, reference _cancel.py)


# far_fill="fp32": pairs within this many segment lengths stay FP64 (the
# near-neighbour integrals of the mixed form)
FAR_NEAR_RATIO = 3.0
# The mixed form's one-residual readout is kept when its error estimate
# delta max|C^T R| stays within this fraction of max|Y| (half the mixed
# form's 1e-8 bar; the estimate has tracked the readout's actual max-entry
# error within 15 % on dipole arrays from 8x8 to 32x32); otherwise a fresh
# FP64 residual at the corrected iterate reads out again.
READOUT_TOL = 5e-9


class CancelToken:
    """Cooperative cancel flag (reference src/momwire/_cancel.py:20-79).

    The host polls it between LU panels, sweep steps and refinement steps;
    the Sommerfeld engine's kernels poll it while they run.  So the flag
    lives in mapped pinned host memory (``mw_cancel_create``), which the host
    and the device both read.  It is allocated when a solve first asks for
    its address; until then a plain host int32 holds the state."""

    def __init__(self):
        self._flag = np.zeros(1, dtype=np.int32)
        self._pinned = None

    def _ensure_pinned(self):
        if self._pinned is None:
            import weakref

            lib = _lib.load()
            p = ctypes.POINTER(ctypes.c_int32)()
            _lib.check(lib.mw_cancel_create(ctypes.byref(p)), "cancel flag")
            arr = np.ctypeslib.as_array(p, shape=(1,))
            arr[0] = self._flag[0]
            self._flag, self._pinned = arr, p
            weakref.finalize(self, lib.mw_cancel_destroy, p)

    def cancel(self):
        self._flag[0] = 1

    def reset(self):
        self._flag[0] = 0

    @property
    def cancelled(self):
        return bool(self._flag[0])

    def raise_if_cancelled(self):
        """Raise SolveAborted once cancel() has been called."""
        if self._flag[0]:
            raise _lib.SolveAborted()

    @property
    def address(self):
        """Address of the flag's int32, readable by the host and the device."""
        self._ensure_pinned()
        return self._flag.ctypes.data

    ptr = address


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise _lib.AcceleratorUnavailable("no GPU visible to torch (HIP path needs an MI355X)")
    return torch


def _to_host(*tensors):
    """Device tensors to numpy through page-locked host buffers: one DMA
    each at full PCIe rate, then one wait on the stream.  Each array is a
    view of its own page-locked buffer (kept alive by the array); torch's
    host allocator reuses the buffer once the array is released."""
    torch = _torch()
    out = []
    for t in tensors:
        h = torch.empty(tuple(t.shape), dtype=t.dtype, pin_memory=True)
        h.copy_(t, non_blocking=True)
        out.append(h)
    torch.cuda.current_stream().synchronize()
    return [h.numpy() for h in out]


class _Device:
    """Device copies of the mesh tables plus reusable work buffers."""

    def __init__(self, mesh, cols, rows_con, rows_port, ek=None):
        torch = _torch()
        self.torch = torch
        dev = torch.device("cuda")
        self.dev = dev

        def up(a, dt=None):
            t = torch.from_numpy(np.ascontiguousarray(a if dt is None else a.astype(dt)))
            return t.to(dev)

        self.keep = dict(
            seg=up(mesh.seg, np.float64), seg_edge=up(mesh.seg_edge, np.int32),
            edge=up(mesh.edge, np.float64), edge_i=up(mesh.edge_i, np.int32),
            se_edge=up(mesh.se_edge, np.int32),
            se_tab=torch.empty((max(len(mesh.se_edge), 1), 9), dtype=torch.complex128, device=dev),
            band_poly=up(mesh.band_poly, np.float64), bw_loc=up(mesh.bw_loc, np.int32),
            bw_slot=up(mesh.bw_slot, np.int32), tile_b0=up(mesh.tile_b0, np.int32),
            tile_seg=up(mesh.tile_seg, np.int32), tile_nseg=up(mesh.tile_nseg, np.int32),
            gl=up(mesh.gl, np.float64), pairs=up(mesh.tile_pairs, np.int32),
        )
        # extended kernel: coaxial labels (free space; image observer, source)
        ek_ptr = [None, None, None]
        if ek is not None:
            for i, lab in enumerate(ek):
                if lab is not None:
                    self.keep[f"ek{i}"] = up(lab, np.int32)
                    ek_ptr[i] = self.keep[f"ek{i}"].data_ptr()
        k = self.keep
        self.mw = _lib.MwMesh(
            n_seg=mesh.n_seg, n_edge=len(mesh.edge), n_basis=mesh.n_basis,
            n_tile=len(mesh.tile_nseg), n_se=len(mesh.se_edge), degree=mesh.degree, n_qp=mesh.n_qp,
            ek=int(ek is not None),
            seg=k["seg"].data_ptr(), seg_edge=k["seg_edge"].data_ptr(), edge=k["edge"].data_ptr(),
            edge_i=k["edge_i"].data_ptr(), se_edge=k["se_edge"].data_ptr(),
            se_tab=k["se_tab"].data_ptr(), band_poly=k["band_poly"].data_ptr(),
            bw_loc=k["bw_loc"].data_ptr(), bw_slot=k["bw_slot"].data_ptr(),
            tile_b0=k["tile_b0"].data_ptr(), tile_seg=k["tile_seg"].data_ptr(),
            tile_nseg=k["tile_nseg"].data_ptr(), gl=k["gl"].data_ptr(),
            ek_label=ek_ptr[0], ek_img_obs=ek_ptr[1], ek_img_src=ek_ptr[2])
        # the k-independent static half of the same-edge table, once per mesh
        n_se = len(mesh.se_edge)
        if n_se:
            self.keep["se_static"] = torch.empty((n_se, 9), dtype=torch.float64, device=dev)
            ptr = self.keep["se_static"].data_ptr()
            _lib.check(_lib.load().mw_same_edge_static(
                self.mw, ptr, torch.cuda.current_stream().cuda_stream), "same-edge static table")
            self.mw.se_static = ptr
        nb = mesh.n_basis
        self.nb = nb
        self.n_ports = len(rows_port)
        self.n_con = len(rows_con)
        self.ncols = len(cols)
        ci, cv, cmax = sparse_pack(cols)
        self.col_idx, self.col_val, self.col_max = up(ci), up(cv), cmax
        # first row of each right-hand side (host): the symmetric solve's
        # forward pass skips the columns that have not started yet
        self.first_row = np.ascontiguousarray(
            np.where(ci >= 0, ci, nb).min(axis=1).astype(np.int32) if self.ncols
            else np.zeros(0, dtype=np.int32))
        pi, pv, pmax = sparse_pack(rows_port)
        self.port_idx, self.port_val, self.port_max = up(pi), up(pv), pmax
        if self.n_con:
            ai, av, amax = sparse_pack(rows_con)
            self.con_idx, self.con_val, self.con_max = up(ai), up(av), amax
        self.Z = None
        self.X = None

    def buffers(self):
        torch = self.torch
        if self.Z is None:
            nb = self.nb
            # [Z | B]: the right-hand sides ride through the factorisation as
            # extra columns (mw_zgesv), so X is a column block of the same buffer
            self.ld = nb + self.ncols
            self.Z = torch.empty((nb, self.ld), dtype=torch.complex128, device=self.dev)
            self.X = self.Z[:, nb:]
            self.ipiv = torch.empty(nb, dtype=torch.int32, device=self.dev)
            lib = _lib.load()
            # the symmetric factorisation's workspace (its block inverses) past
            # n = 4096; the LU's below
            self.work = torch.empty(int(lib.mw_sysv_workspace_bytes(nb)), dtype=torch.uint8,
                                    device=self.dev)
            self.post_buffers()
        return self

    def post_buffers(self):
        """What the KCL Schur step and the readout need (no Z)."""
        torch = self.torch
        if getattr(self, "Y", None) is None:
            lib = _lib.load()
            self.info = torch.zeros(2, dtype=torch.int32, device=self.dev)
            self.Y = torch.empty((max(self.n_ports, 1), max(self.n_ports, 1)),
                                 dtype=torch.complex128, device=self.dev)
            if self.n_con:
                # [S | T] = [A X_con | A W]: the Schur system and its right-hand
                # sides side by side, so a large one rides through mw_zgesv
                npt = max(self.n_ports, 1)
                self.ld_st = self.n_con + npt
                self.ST = torch.empty((self.n_con, self.ld_st), dtype=torch.complex128,
                                      device=self.dev)
                self.S = self.ST[:, :self.n_con]
                self.T = self.ST[:, self.n_con:]
                self.kcl_small = self.n_con <= 32 and npt <= 64
                if not self.kcl_small:
                    self.ipiv_con = torch.empty(self.n_con, dtype=torch.int32, device=self.dev)
                    self.work_con = torch.empty(int(lib.mw_getrf_workspace_bytes(self.n_con)),
                                                dtype=torch.uint8, device=self.dev)
        return self

    def buffers_mixed(self):
        """Extra buffers of the mixed form: [Z32 | B32] factor block, the FP64
        iterate and residual, and the FP32 correction."""
        torch = self.torch
        self.buffers()
        if getattr(self, "Z32", None) is None:
            nb, nc = self.nb, max(self.ncols, 1)
            self.ld32 = nb + self.ncols
            self.Z32 = torch.empty((nb, self.ld32), dtype=torch.complex64, device=self.dev)
            self.X32 = self.Z32[:, nb:]
            self.X64 = torch.empty((nb, nc), dtype=torch.complex128, device=self.dev)
            self.R64 = torch.empty((nb, nc), dtype=torch.complex128, device=self.dev)
            self.D32 = torch.empty((nb, nc), dtype=torch.complex64, device=self.dev)
            # the one-residual readout: C in FP64, T = C^T R, its check
            npt = max(self.n_ports, 1)
            self.C64 = torch.empty((nb, nc), dtype=torch.complex128, device=self.dev)
            self.T64 = torch.empty((npt, npt), dtype=torch.complex128, device=self.dev)
            self.T32 = torch.empty((npt, npt), dtype=torch.complex64, device=self.dev)
            self.gscratch = torch.empty(16 * (npt + 1), dtype=torch.uint8, device=self.dev)
            self.gstats = torch.zeros(4, dtype=torch.float64, device=self.dev)
            self.gate = torch.zeros(1, dtype=torch.int32, device=self.dev)
            self.ipiv32 = torch.empty(nb, dtype=torch.int32, device=self.dev)
            lib = _lib.load()
            self.work32 = torch.empty(int(lib.mw_sysv_workspace_bytes(nb)), dtype=torch.uint8,
                                      device=self.dev)
        return self


class BSplineSolver(_SweptPortSolutions):
    """Degree-2 B-spline Galerkin MoM port solve on the GPU (see module doc)."""

    eps = EPS0
    mu = MU0

    def __init__(self, *, wires, n_per_edge_per_wire=None, degree=2, feed_wire_index=0,
                 feed_arclength=None, feeds=None, feed_model="point",
                 feed_smoothing_factor=None, junctions=None, junction_ports=None,
                 node_gaps=None, n_qp_pair=4, n_qp_source=16, extended_kernel=False,
                 wavelength=22, halfdriver_factor=0.962, wire_radius=0.0005,
                 wire_conductivity=None, insulation_radius=None, insulation_eps_r=None,
                 nsegs=101, ground_z=None, ground_eps=None, ground_phi_mode="normal",
                 ground_model="refl-coef", n_qp_sommerfeld=3, use_singular_enrichment=False,
                 swept_mem_mb=256, cancel=None, precision="fp64", refine_steps=None,
                 somm_surrogate=None, far_fill="fp64", **unsupported):
        if unsupported:
            raise NotImplementedError(f"options not served on the GPU path: {sorted(unsupported)}")
        if degree < 1:
            raise ValueError(f"degree must be >= 1, got {degree}")
        if degree > 2:
            raise NotImplementedError("the GPU path is built for degrees 1 and 2")
        if not wires:
            raise ValueError("wires must be non-empty")
        if int(n_qp_pair) < 1:
            raise ValueError(f"n_qp_pair must be >= 1, got {n_qp_pair}")
        if int(n_qp_pair) > 8:
            raise NotImplementedError("the GPU fill is built for n_qp_pair 1 .. 8")
        if feed_model not in ("point", "segment"):
            raise ValueError(f"feed_model must be 'point' or 'segment', got {feed_model!r}")
        if feed_model == "segment" and feed_smoothing_factor is not None:
            raise ValueError("feed_model='segment' and feed_smoothing_factor are mutually "
                             "exclusive -- both replace the point drive with a spread one")
        if use_singular_enrichment:
            raise NotImplementedError("singular enrichment is not served")
        if ground_eps is not None and ground_z is None:
            raise ValueError("ground_eps requires ground_z to be set")
        if ground_model not in ("refl-coef", "sommerfeld"):
            raise ValueError(f"ground_model must be 'refl-coef' or 'sommerfeld', got {ground_model!r}")
        if ground_model == "sommerfeld" and ground_eps is None:
            raise ValueError("ground_model='sommerfeld' requires ground_eps")
        if ground_phi_mode not in _ground.PHI_MODES:
            raise ValueError(f"ground_phi_mode must be one of {_ground.PHI_MODES}, "
                             f"got {ground_phi_mode!r}")
        if int(n_qp_sommerfeld) < 1:
            raise ValueError(f"n_qp_sommerfeld must be >= 1, got {n_qp_sommerfeld}")
        if int(n_qp_sommerfeld) > 8:
            raise NotImplementedError("the remainder block is built for n_qp_sommerfeld 1 .. 8")
        if int(swept_mem_mb) < 1:
            raise ValueError(f"swept_mem_mb must be >= 1, got {swept_mem_mb}")

        if precision not in ("fp64", "mixed"):
            raise ValueError(f"precision must be 'fp64' or 'mixed', got {precision!r}")
        if refine_steps is not None and int(refine_steps) < 0:
            raise ValueError("refine_steps must be >= 0")
        if far_fill not in ("fp64", "fp32"):
            raise ValueError(f"far_fill must be 'fp64' or 'fp32', got {far_fill!r}")
        if far_fill == "fp32" and precision != "mixed":
            raise ValueError("far_fill='fp32' is an option of precision='mixed'")
        # far_fill="fp32": segment pairs more than 3 segment lengths apart take
        # the kernel's phase in FP32 (the rest of the fill stays FP64)
        self.far_fill = far_fill
        # "mixed": FP64 fill, FP32 LU and solves on MFMA, iterative refinement
        # with the residual B - Z X in FP64 (refine_steps corrections; default
        # 1 plus the adjoint-corrected readout when Z is symmetric, else 3).
        self.precision = precision
        self.refine_steps = None if refine_steps is None else int(refine_steps)
        self.degree = int(degree)
        self.n_qp_pair = int(n_qp_pair)
        self.feed_model = feed_model
        self.feed_smoothing_factor = feed_smoothing_factor
        self.n_qp_source = int(n_qp_source)
        self.wavelength = wavelength
        self.halfdriver_factor = halfdriver_factor
        self.wire_radius = wire_radius
        self.nsegs = nsegs
        self.ground_z = None if ground_z is None else float(ground_z)
        self.ground_eps = ground_eps
        self.ground_model = ground_model
        self.extended_kernel = bool(extended_kernel)
        self.ground_phi_mode = ground_phi_mode
        self.n_qp_sommerfeld = int(n_qp_sommerfeld)
        # a trained momwire_amd.surrogate.SommerfeldSurrogate fills the grid
        # nodes of its domain in place of the exact engine
        self.somm_surrogate = somm_surrogate
        self.swept_mem_mb = int(swept_mem_mb)
        self._cancel = cancel
        # bspline.py:672-676
        self.c = 1 / np.sqrt(self.eps * self.mu)
        self.freq = self.c / self.wavelength
        self.omega = 2 * np.pi * self.freq
        self.k = self.omega / self.c

        self.wires_polylines = [np.asarray(w, dtype=float) for w in wires]
        n_w = len(self.wires_polylines)
        for i, pl in enumerate(self.wires_polylines):
            if pl.ndim != 2 or pl.shape[0] < 2 or pl.shape[1] != 3:
                raise ValueError(f"wire {i}: polyline must be (M, 3) with M >= 2")
        r = np.asarray(wire_radius, dtype=float)
        if r.ndim == 0:
            r = np.full(n_w, float(r))
        elif r.shape != (n_w,):
            raise ValueError(f"wire_radius: expected a scalar or a length-{n_w} sequence")
        if not np.all(np.isfinite(r)) or np.any(r <= 0):
            raise ValueError(f"wire_radius entries must be positive and finite, got {r}")
        self._radius_per_wire = r
        # distributed series loading (conductor loss, insulation): validated here
        self._loading = _loading.WireLoading(r, wire_conductivity, insulation_radius,
                                             insulation_eps_r)
        self.wire_conductivity = self._loading.conductivity
        self.insulation_radius = self._loading.insulation_radius
        self.insulation_eps_r = self._loading.insulation_eps_r
        self._loading_active = self._loading.active
        # ground contact (reference bspline.py:683-698, 1108-1169): a wire END
        # within 1e-6 of its own length of the plane touches it
        self._ground_ends = ()
        if self.ground_z is not None:
            gz = self.ground_z
            ends = []
            for w, pl in enumerate(self.wires_polylines):
                tol = 1e-6 * max(float(np.sum(np.linalg.norm(np.diff(pl, axis=0), axis=1))), 1e-30)
                if float(pl[:, 2].min()) < gz - tol:
                    raise ValueError(f"wire {w} dips below the ground plane (min z = "
                                     f"{pl[:, 2].min():.6g} < ground_z = {gz:g})")
                at = np.abs(pl[:, 2] - gz) <= tol
                if np.any(at[:-1] & at[1:]):
                    raise ValueError(f"wire {w} has an edge lying in the ground plane (both "
                                     "endpoints at ground_z) -- degenerate over a conducting "
                                     "ground")
                if at[0]:
                    ends.append((w, "start"))
                if at[-1]:
                    ends.append((w, "end"))
            if ends and ground_eps is not None and ground_model == "refl-coef":
                where = ", ".join(f"wire {w} {e}" for w, e in ends)
                raise NotImplementedError(
                    f"{where} lies in the ground plane: ground contact under "
                    "ground_model='refl-coef' is refused (the reflection-coefficient "
                    "charge weight has no validity at zero clearance); use "
                    "ground_model='sommerfeld' or a PEC ground")
            self._ground_ends = tuple(ends)

        if n_per_edge_per_wire is None:
            n_per_edge_per_wire = [None] * n_w
        if len(n_per_edge_per_wire) != n_w:
            raise ValueError(f"n_per_edge_per_wire length {len(n_per_edge_per_wire)} != n_wires {n_w}")
        self.n_per_edge_per_wire = []
        for i, (pl, npe) in enumerate(zip(self.wires_polylines, n_per_edge_per_wire)):
            ne = pl.shape[0] - 1
            if npe is None:
                npe = self.nsegs
            if np.isscalar(npe):
                npe = [int(npe)] * ne
            npe = list(npe)
            if len(npe) != ne:
                raise ValueError(f"wire {i}: n_per_edge length {len(npe)} != n_edges {ne}")
            self.n_per_edge_per_wire.append([int(x) for x in npe])

        if feeds is None:
            if not (0 <= feed_wire_index < n_w):
                raise ValueError(f"feed_wire_index {feed_wire_index} out of range")
            self.feeds = [(int(feed_wire_index), feed_arclength, 1.0 + 0.0j)]
        else:
            if len(feeds) == 0 and not junction_ports and not node_gaps:
                raise ValueError("feeds must contain at least one entry")
            norm = []
            for i, f in enumerate(feeds):
                if len(f) != 3:
                    raise ValueError(f"feeds[{i}]: expected (wire_index, arclength, voltage), got {f!r}")
                w_i, arc_i, v_i = f
                if not (0 <= w_i < n_w):
                    raise ValueError(f"feeds[{i}]: wire_index {w_i} out of range [0, {n_w})")
                norm.append((int(w_i), None if arc_i is None else float(arc_i), complex(v_i)))
            self.feeds = norm

        self.junctions = []
        for j, jw in enumerate(junctions or []):
            if len(jw) < 1:
                raise ValueError(f"junction {j}: need >= 1 wire-end")
            grp = []
            for w, end in jw:
                if not (0 <= w < n_w):
                    raise ValueError(f"junction {j}: wire_idx {w} out of range [0, {n_w})")
                if end not in ("start", "end"):
                    raise ValueError(f"junction {j}: end must be 'start' or 'end', got {end!r}")
                grp.append((int(w), end))
            self.junctions.append(grp)
        self._check_junction_coincidence()

        self.junction_ports = []
        seen = set()
        for p in junction_ports or []:
            j_idx, volt = (int(p), 0j) if isinstance(p, (int, np.integer)) else (int(p[0]), complex(p[1]))
            if not (0 <= j_idx < len(self.junctions)):
                raise ValueError(f"junction_ports: junction index {j_idx} out of range")
            if j_idx in seen:
                raise ValueError(f"junction_ports: junction {j_idx} listed twice")
            seen.add(j_idx)
            self.junction_ports.append((j_idx, volt))

        self.node_gaps = []
        end_to_j = {m: j for j, jw in enumerate(self.junctions) for m in jw}
        used_j, used_m = set(), set()
        ported = {j for j, _v in self.junction_ports}
        for i, g in enumerate(node_gaps or []):
            if len(g) != 3:
                raise ValueError(f"node_gaps[{i}]: expected (wire_index, 'start'|'end', voltage)")
            w_i, end_i, v_i = g
            if not (0 <= w_i < n_w):
                raise ValueError(f"node_gaps[{i}]: wire_index {w_i} out of range")
            if end_i not in ("start", "end"):
                raise ValueError(f"node_gaps[{i}]: end must be 'start' or 'end'")
            member = (int(w_i), end_i)
            j_idx = end_to_j.get(member)
            if j_idx is None:
                raise ValueError(f"node_gaps[{i}]: wire {w_i} {end_i!r} is not a member of any junction")
            if len(self.junctions[j_idx]) < 2:
                raise ValueError(f"node_gaps[{i}]: junction {j_idx} has a single member")
            if member in used_m or j_idx in used_j:
                raise ValueError(f"node_gaps[{i}]: wire-end or junction already carries a node gap")
            if j_idx in ported:
                raise ValueError(f"node_gaps[{i}]: junction {j_idx} is also a junction port")
            used_m.add(member)
            used_j.add(j_idx)
            self.node_gaps.append((int(w_i), end_i, complex(v_i)))

        self._mesh = None
        self._ports = None
        self._dev = None
        # the symmetric factorisation (mw_?sysv) used by the last solve, and
        # whether one has needed row exchanges (then the pivoting path only)
        self._sym_used = False
        self._sym_failed = False

    # ------------------------------------------------------------------
    def _check_junction_coincidence(self):
        """Refuse junction groups whose ends do not meet (reference _wire_spec.py:62-104)."""
        for j, grp in enumerate(self.junctions):
            pts, lens = [], []
            for w, end in grp:
                pl = self.wires_polylines[w]
                npe = self.n_per_edge_per_wire[w]
                if end == "start":
                    pts.append(pl[0])
                    lens.append(np.linalg.norm(pl[1] - pl[0]) / npe[0])
                else:
                    pts.append(pl[-1])
                    lens.append(np.linalg.norm(pl[-1] - pl[-2]) / npe[-1])
            tol = max(1e-3 * min(lens), 1e-5)
            for p in pts[1:]:
                dist = float(np.linalg.norm(p - pts[0]))
                if dist > tol:
                    raise ValueError(f"junction {j}: member ends are {dist:.3g} apart (tolerance {tol:.3g})")

    def _build(self):
        if self._mesh is None:
            jends = {(w, e) for jw in self.junctions for w, e in jw}
            self._mesh = build_mesh(self.wires_polylines, self.n_per_edge_per_wire,
                                    self._radius_per_wire, self.degree, self.junctions,
                                    n_qp=self.n_qp_pair,
                                    ground_ends=[g for g in self._ground_ends if g not in jends])
        return self._mesh

    def _grounded_junctions(self):
        """Junctions whose shared point lies in the ground plane: their KCL
        row is dropped (current may leave through the image; reference
        bspline.py:1152-1169, 1399-1406)."""
        if self.ground_z is None:
            return frozenset()
        out = set()
        for j, jw in enumerate(self.junctions):
            w, end = jw[0]
            pl = self.wires_polylines[w]
            pt = pl[0] if end == "start" else pl[-1]
            tol = 1e-6 * max(float(np.sum(np.linalg.norm(np.diff(pl, axis=0), axis=1))), 1e-30)
            if abs(pt[2] - self.ground_z) <= tol:
                out.add(j)
        return frozenset(out)

    def _port_count(self):
        return len(self.feeds) + len(self.junction_ports) + len(self.node_gaps)

    def _port_plan(self):
        """Sparse port columns (drive == readout) and remaining KCL rows
        (bspline.py:3784-3867, 1170-1210)."""
        if self._ports is None:
            mesh = self._build()
            cols = [feed_column(mesh, w, s, self.feed_model, self.feed_smoothing_factor,
                                self.n_qp_source) for w, s, _v in self.feeds]
            rows = {j: ([m for m, _s in dirs], [s for _m, s in dirs]) for j, dirs in mesh.kcl}
            port_js = [j for j, _v in self.junction_ports]
            grounded = self._grounded_junctions()
            for j in port_js:
                if j in grounded:
                    raise ValueError(f"junction {j} is both grounded and a junction port -- a "
                                     "grounded node's voltage is pinned by the ground image, "
                                     "so it cannot also be a driven port")
            cols += [rows[j] for j in port_js]
            for w, end, _v in self.node_gaps:
                j = next(jj for jj, jw in enumerate(self.junctions) if (w, end) in jw)
                pw = mesh.per_wire[w]
                m = None
                for kidx, (_jl, kind, ji, endp) in enumerate(pw["kept"]):
                    if kind == "dir" and ji == j and endp == end:
                        m = pw["l2g"][kidx]
                cols.append(([m], [1.0 if end == "start" else -1.0]))
            con = [rows[j] for j in range(len(self.junctions))
                   if j not in set(port_js) and j not in grounded]
            self._ports = (cols, con)
        return self._ports

    def _device_base(self):
        """The device mesh tables and port columns (no solve buffers)."""
        if self._dev is None:
            cols, con = self._port_plan()
            mesh = self._build()
            ek = _mesh.ek_tables(mesh, self.ground_z) if self.extended_kernel else None
            self._dev = _Device(mesh, cols + con, con, cols, ek=ek)
        return self._dev

    def _device(self):
        return self._device_base().buffers()

    # ------------------------------------------------------------------
    def _solve_device(self, k=None, timers=None):
        """One fill + one factorisation + all-port solve on the GPU.

        Returns device tensors (Y, W): Y (n_ports x n_ports) and W, whose first
        n_ports columns are the port coefficient columns.  Nothing is copied
        to the host and nothing is synchronised.  ``timers``: optional dict
        that collects (start, end) HIP events per stage on the launch stream."""
        lib = _lib.load()
        D = self._device()
        torch = D.torch
        st = torch.cuda.current_stream().cuda_stream

        def mark(stage, first):
            if timers is not None:
                ev = torch.cuda.Event(enable_timing=True)
                ev.record()
                if first:
                    timers.setdefault(stage, []).append([ev, None])
                else:
                    timers[stage][-1][1] = ev

        kk = self.k if k is None else float(k)
        omega = kk * self.c
        mesh = self._mesh
        nb, ncols = D.nb, D.ncols
        cancel = self._cancel.address if self._cancel is not None else None
        mark("fill", True)
        _lib.check(lib.mw_same_edge_table(D.mw, kk, st), "same-edge table")
        ld = D.ld
        self._fill(D, kk, omega, D.Z.data_ptr(), ld, st)
        mark("fill", False)
        _lib.check(lib.mw_zscatter_cols(nb, ncols, D.col_max, D.col_idx.data_ptr(),
                                        D.col_val.data_ptr(), D.X.data_ptr(), ld, st),
                   "right-hand sides")
        mark("solve", True)
        self._sym_used = self._sym_ok(D)
        if self._sym_used:
            _lib.check(lib.mw_zsysv(nb, ncols, D.Z.data_ptr(), ld, D.ipiv.data_ptr(),
                                    D.work.data_ptr(), D.info.data_ptr(), cancel,
                                    D.first_row.ctypes.data, st), "factorisation and solve")
        else:
            _lib.check(lib.mw_zgesv(nb, ncols, D.Z.data_ptr(), ld, D.ipiv.data_ptr(),
                                    D.work.data_ptr(), D.info.data_ptr(), cancel, st),
                       "factorisation and solve")
        mark("solve", False)
        self._post(D, D.X.data_ptr(), ld, D.Y.data_ptr(), st)
        return D.Y[:D.n_ports, :D.n_ports], D.X[:, :D.n_ports]

    def _sym_ok(self, D):
        """Whether the solve takes the symmetric factorisation (mw_?sysv, no
        row exchanges, half the trailing update): Z = Z^T exactly (one radius
        for every wire, free space, no wire loading: the fill computes the
        upper tiles and mirrors them), unless a solve of this solver has
        already needed row exchanges."""
        return (bool(self._mesh.symmetric) and self.ground_z is None
                and not self._loading_active and not self.extended_kernel
                and not self._sym_failed)

    def _solve_valid(self, graph=False):
        """The device solve (a graph replay when ``graph``), repeated with the
        pivoting factorisation when the symmetric one met a panel that needs
        row exchanges (info = -1): that result is not valid."""
        Y, W = self.solve_graph(warm=False) if graph else self._solve()
        if self._sym_used:
            info, = _to_host(self._dev.info)
            if info[0] == -1:
                self._sym_failed = True
                self._graph_entry = None
                Y, W = self._solve()
        return Y, W

    def _fill(self, D, k, omega, zptr, ld, st):
        """Free-space fill, then the ground terms: Z -= C2 * image + Q
        (PEC: C2 = 1, no Q), or Z -= the Fresnel-weighted image for the
        reflection-coefficient ground."""
        self._fill_launcher(D, k, omega, zptr, ld, st)()

    def _fill_launcher(self, D, k, omega, zptr, ld, st):
        """Resolve everything the fill needs on the host (loading values,
        ground grid, weights) and return a function that only launches its
        kernels."""
        lib = _lib.load()
        mesh = self._mesh
        pairs = D.keep["pairs"]
        sym = int(mesh.symmetric)
        if self.far_fill == "fp32":
            calls = [lambda: _lib.check(lib.mw_fill_z_far32(
                D.mw, k, omega, self.eps, self.mu, zptr, ld, pairs.data_ptr(),
                int(pairs.shape[0]), sym, FAR_NEAR_RATIO, st), "fill (FP32 far pairs)")]
        else:
            calls = [lambda: _lib.check(lib.mw_fill_z(D.mw, k, omega, self.eps, self.mu, zptr, ld,
                                                      pairs.data_ptr(), int(pairs.shape[0]), sym,
                                                      st), "fill")]
        if self._loading_active:
            rows, cols, vals = self._loading_device(D, omega)
            calls.append(lambda: _lib.check(lib.mw_zadd_sparse(
                D.nb, int(rows.shape[0]), rows.data_ptr(), cols.data_ptr(), vals.data_ptr(), zptr,
                ld, st), "wire loading"))
        if self.ground_z is not None and self.ground_eps is not None \
                and self.ground_model == "refl-coef":
            eps_t = _ground.eps_tilde(self.ground_eps, omega, self.eps)
            c0, c1 = _ground.phi_coefficients(self.ground_phi_mode, eps_t)
            calls.append(lambda: _lib.check(lib.mw_fill_ground_refl(
                D.mw, k, omega, self.eps, self.mu, self.ground_z, eps_t.real, eps_t.imag, c0.real,
                c0.imag, c1.real, c1.imag, zptr, ld, pairs.data_ptr(), int(pairs.shape[0]), sym,
                st), "reflection-coefficient ground fill"))
        elif self.ground_z is not None:
            grid = self._ground_grid(k, omega)
            eps_t = None if self.ground_eps is None else _ground.eps_tilde(self.ground_eps, omega,
                                                                          self.eps)
            c2 = _ground.image_coefficient(eps_t)
            gref = ctypes.byref(grid) if grid is not None else None
            calls.append(lambda: _lib.check(lib.mw_fill_ground(
                D.mw, k, omega, self.eps, self.mu, self.ground_z, c2.real, c2.imag, gref,
                self.n_qp_sommerfeld, zptr, ld, pairs.data_ptr(), int(pairs.shape[0]), sym, st),
                "ground fill"))

        def launch():
            for c in calls:
                c()
        launch.grid = grid if self.ground_z is not None and self.ground_model == "sommerfeld" \
            and self.ground_eps is not None else None  # kept alive with the launcher
        return launch

    def _loading_device(self, D, omega):
        """Device (rows, cols, merged Z'_w S_w values) of the loading term at
        omega, uploaded once per frequency (a graph replay reuses them)."""
        cache = D.__dict__.setdefault("loading", {})
        if "term" not in cache:
            term = _loading.LoadingTerm(self._build(), self._loading)
            torch = D.torch
            cache["term"] = term
            cache["rows"] = torch.from_numpy(term.rows).to(D.dev)
            cache["cols"] = torch.from_numpy(term.cols).to(D.dev)
        key = float(omega)
        if cache.get("omega") != key:
            cache["vals"] = D.torch.from_numpy(cache["term"].entries(key)).to(D.dev)
            cache["omega"] = key
        return cache["rows"], cache["cols"], cache["vals"]

    def wire_loss_power(self, coeffs, omega=None):
        """Ohmic power in the wire metal, 1/2 sum_w Re Z'_w c^H S_w c, with
        per-wire parts (reference bspline.py:2775-2800)."""
        n_w = len(self.wires_polylines)
        if not self._loading_active:
            return 0.0, np.zeros(n_w)
        term = _loading.LoadingTerm(self._build(), self._loading)
        return term.loss_power(coeffs, self.omega if omega is None else omega)

    def _ground_grid(self, k, omega):
        """The Sommerfeld grid view at this frequency (None for PEC)."""
        if self.ground_model != "sommerfeld" or self.ground_eps is None:
            return None
        mesh = self._build()
        seg = mesh.seg
        eps_t = _ground.eps_tilde(self.ground_eps, omega, self.eps)
        r1 = _ground.max_image_distance(seg[:, 0:3], seg[:, 3:6], self.ground_z)
        filler = None if self.somm_surrogate is None else self.somm_surrogate.grid_filler()
        key = None if self.somm_surrogate is None else id(self.somm_surrogate)
        master = _ground.get_master(eps_t, k, r1, cancel=self._cancel, filler=filler, key_extra=key)
        self._grid_master = master
        return master.struct(k, omega, self.mu)

    def _post(self, D, xptr, ldx, yptr, st):
        """KCL Schur step (bspline.py:2981-2989) and the readout Y = B^T X
        (bspline.py:4178-4179), both on the device, in place on X."""
        lib = _lib.load()
        nb = D.nb
        npt = D.n_ports
        if D.n_con:
            Xc = xptr + 16 * npt
            _lib.check(lib.mw_zsparse_rows(D.n_con, D.con_max, D.con_idx.data_ptr(),
                                           D.con_val.data_ptr(), Xc, ldx, D.n_con,
                                           D.S.data_ptr(), D.ld_st, st), "KCL A.X")
            _lib.check(lib.mw_zsparse_rows(D.n_con, D.con_max, D.con_idx.data_ptr(),
                                           D.con_val.data_ptr(), xptr, ldx, npt,
                                           D.T.data_ptr(), D.ld_st, st), "KCL A.W")
            if D.kcl_small:
                _lib.check(lib.mw_zsmall_solve(D.n_con, npt, D.S.data_ptr(), D.ld_st,
                                               D.T.data_ptr(), D.ld_st,
                                               D.info.data_ptr() + 4, st), "KCL Schur")
            else:
                # more than 32 constraint rows or 64 ports: the blocked LU on [S | T]
                _lib.check(lib.mw_zgesv(D.n_con, npt, D.ST.data_ptr(), D.ld_st,
                                        D.ipiv_con.data_ptr(), D.work_con.data_ptr(),
                                        D.info.data_ptr() + 4, None, st), "KCL Schur")
            _lib.check(lib.mw_zgemm_sub(nb, npt, D.n_con, Xc, ldx, D.T.data_ptr(),
                                        D.ld_st, xptr, ldx, st), "KCL update")
        _lib.check(lib.mw_zsparse_rows(npt, D.port_max, D.port_idx.data_ptr(),
                                       D.port_val.data_ptr(), xptr, ldx, npt,
                                       yptr, max(npt, 1), st), "port readout")

    def _solve_device_mixed(self, k=None, timers=None, history=None):
        """Mixed form on the GPU: FP64 fill; [Z | B] factored and solved in
        FP32 on MFMA; then refine_steps corrections X += Z32^{-1} (B - Z X)
        with the residual in FP64.  ``history``: optional list that receives
        a device Y after the FP32 solve and after each correction."""
        lib = _lib.load()
        D = self._device().buffers_mixed()
        torch = D.torch
        st = torch.cuda.current_stream().cuda_stream

        def mark(stage, first):
            if timers is not None:
                ev = torch.cuda.Event(enable_timing=True)
                ev.record()
                if first:
                    timers.setdefault(stage, []).append([ev, None])
                else:
                    timers[stage][-1][1] = ev

        kk = self.k if k is None else float(k)
        omega = kk * self.c
        mesh = self._mesh
        nb, ncols = D.nb, D.ncols
        nc = max(ncols, 1)
        cancel = self._cancel.address if self._cancel is not None else None
        pairs = D.keep["pairs"]
        mark("fill", True)
        _lib.check(lib.mw_same_edge_table(D.mw, kk, st), "same-edge table")
        self._fill(D, kk, omega, D.Z.data_ptr(), nb, st)
        mark("fill", False)
        mark("solve32", True)
        _lib.check(lib.mw_zscatter_cols(nb, ncols, D.col_max, D.col_idx.data_ptr(),
                                        D.col_val.data_ptr(), D.R64.data_ptr(), nc, st), "rhs")
        _lib.check(lib.mw_z2c(nb, nb, D.Z.data_ptr(), nb, D.Z32.data_ptr(), D.ld32, st), "Z to FP32")
        _lib.check(lib.mw_z2c(nb, ncols, D.R64.data_ptr(), nc, D.X32.data_ptr(), D.ld32, st),
                   "B to FP32")
        self._sym_used = self._sym_ok(D)
        if self._sym_used:
            _lib.check(lib.mw_csysv(nb, ncols, D.Z32.data_ptr(), D.ld32, D.ipiv32.data_ptr(),
                                    D.work32.data_ptr(), D.info.data_ptr(), cancel,
                                    D.first_row.ctypes.data, st), "FP32 factorisation and solve")
        else:
            _lib.check(lib.mw_cgesv(nb, ncols, D.Z32.data_ptr(), D.ld32, D.ipiv32.data_ptr(),
                                    D.work32.data_ptr(), D.info.data_ptr(), cancel, st),
                       "FP32 factorisation and solve")
        _lib.check(lib.mw_c2z(nb, ncols, D.X32.data_ptr(), D.ld32, D.X64.data_ptr(), nc, 0, st),
                   "X to FP64")
        mark("solve32", False)

        def record():
            if history is None:
                return
            Xt = D.X64.clone()
            Yt = torch.empty((max(D.n_ports, 1), max(D.n_ports, 1)), dtype=torch.complex128,
                             device=D.dev)
            self._post(D, Xt.data_ptr(), nc, Yt.data_ptr(), st)
            history.append(Yt[:D.n_ports, :D.n_ports])

        npt = D.n_ports
        # The adjoint-corrected readout Y = B^T X + X^T (B - Z X) has an error
        # quadratic in the iterate's only when Z = Z^T: every wire shares one
        # radius (the fill uses the observer row's radius, so per-wire radii
        # make Z unsymmetric) and no KCL rows are eliminated afterwards.
        # Otherwise: the plain readout after three corrections.
        adjoint = D.n_con == 0 and bool(mesh.symmetric)
        one_residual = adjoint and self.refine_steps is None
        steps = self.refine_steps if self.refine_steps is not None else (1 if adjoint else 3)
        trs32 = lib.mw_csytrs if self._sym_used else lib.mw_cgetrs

        def record_adjoint():
            if history is None or not adjoint:
                return
            Yt = torch.empty((max(npt, 1), max(npt, 1)), dtype=torch.complex128, device=D.dev)
            self._adjoint_readout(D, Yt.data_ptr(), st)
            history.append(("adjoint", Yt[:npt, :npt]))

        def correction(accumulate_into):
            # D32 <- Z32^{-1} (B - Z X) with the residual in FP64 (left in R64)
            self._fp64_residual(D, st)
            record_adjoint()
            _lib.check(lib.mw_z2c(nb, ncols, D.R64.data_ptr(), nc, D.D32.data_ptr(), nc, st),
                       "residual to FP32")
            _lib.check(trs32(nb, ncols, D.Z32.data_ptr(), D.ld32, D.ipiv32.data_ptr(),
                             D.work32.data_ptr(), D.D32.data_ptr(), nc, st), "FP32 correction")
            if accumulate_into is not None:
                _lib.check(lib.mw_c2z(nb, ncols, D.D32.data_ptr(), nc, accumulate_into, nc, 1, st),
                           "X += correction")

        record()
        if one_residual:
            # One FP64 residual R = B - Z X0.  C = Z32^{-1} R corrects the
            # coefficients (X1 = X0 + C), and Y = B^T X0 + X0^T R + C^T R reads
            # out B^T Z^{-1} B = B^T X0 + X0^T R + R^T Z^{-1} R (Z = Z^T) with
            # the error (Z^{-1} R - C)^T R, estimated on the device as
            # delta max|C^T R| (delta: the FP32 solve's relative error
            # |C| / |X0|).  Past READOUT_TOL max|Y|, a fresh FP64 residual at
            # X1 reads out again (_mixed_check).
            if self._cancel is not None and self._cancel.cancelled:
                raise _lib.SolveAborted("refinement cancelled")
            mark("refine", True)
            correction(None)
            _lib.check(lib.mw_c2z(nb, ncols, D.D32.data_ptr(), nc, D.C64.data_ptr(), nc, 0, st),
                       "C to FP64")
            mark("refine", False)
            mark("readout", True)
            self._adjoint_readout(D, D.Y.data_ptr(), st)
            # T = C^T R in FP32 (C from the FP32 solve, R rounded into the
            # factor block's right-hand-side columns, free after the solve):
            # T is ~1e-6 of Y, so single precision leaves ~1e-11 of Y
            _lib.check(lib.mw_z2c(nb, npt, D.R64.data_ptr(), nc, D.X32.data_ptr(), D.ld32, st),
                       "R to FP32")
            D.T32.zero_()
            _lib.check(lib.mw_cgemm_tn_add(npt, npt, nb, D.D32.data_ptr(), nc, D.X32.data_ptr(),
                                           D.ld32, D.T32.data_ptr(), max(npt, 1), st), "C^T R")
            _lib.check(lib.mw_c2z(npt, npt, D.T32.data_ptr(), max(npt, 1), D.T64.data_ptr(),
                                  max(npt, 1), 0, st), "T to FP64")
            _lib.check(lib.mw_zreadout_check(nb, npt, D.X64.data_ptr(), nc, D.C64.data_ptr(), nc,
                                             D.T64.data_ptr(), max(npt, 1), D.Y.data_ptr(),
                                             max(npt, 1), READOUT_TOL, D.gscratch.data_ptr(),
                                             D.gstats.data_ptr(), D.gate.data_ptr(), st),
                       "readout check")
            _lib.check(lib.mw_c2z(nb, ncols, D.D32.data_ptr(), nc, D.X64.data_ptr(), nc, 1, st),
                       "X += correction")
            mark("readout", False)
            record()
            if torch.cuda.is_current_stream_capturing():
                self._capture_needs_check = True   # solve_graph checks after each replay
            else:
                self._mixed_check(D, st, mark)
            if history is not None:
                history.append(("adjoint", D.Y[:npt, :npt].clone()))
            return D.Y[:npt, :npt], D.X64[:, :npt]
        for _step in range(steps):
            if self._cancel is not None and self._cancel.cancelled:
                raise _lib.SolveAborted("refinement cancelled")
            mark("refine", True)
            correction(D.X64.data_ptr())
            mark("refine", False)
            record()
        if adjoint:
            mark("readout", True)
            self._fp64_residual(D, st)
            self._adjoint_readout(D, D.Y.data_ptr(), st)
            mark("readout", False)
            if history is not None:
                history.append(("adjoint", D.Y[:npt, :npt].clone()))
        else:
            self._post(D, D.X64.data_ptr(), nc, D.Y.data_ptr(), st)
        return D.Y[:npt, :npt], D.X64[:, :npt]

    def _fp64_residual(self, D, st):
        """R64 = B - Z X64 in FP64 (the port and constraint columns together)."""
        lib = _lib.load()
        nb, ncols, nc = D.nb, D.ncols, max(D.ncols, 1)
        _lib.check(lib.mw_zscatter_cols(nb, ncols, D.col_max, D.col_idx.data_ptr(),
                                        D.col_val.data_ptr(), D.R64.data_ptr(), nc, st), "rhs")
        _lib.check(lib.mw_zgemm_sub(nb, ncols, nb, D.Z.data_ptr(), nb, D.X64.data_ptr(), nc,
                                    D.R64.data_ptr(), nc, st), "FP64 residual")

    def _adjoint_readout(self, D, yptr, st):
        """Y = B^T X + X^T R with X = X64, R = R64 = B - Z X: the
        adjoint-corrected readout, whose error is quadratic in X's when
        Z = Z^T."""
        lib = _lib.load()
        npt, nc = D.n_ports, max(D.ncols, 1)
        self._post(D, D.X64.data_ptr(), nc, yptr, st)
        _lib.check(lib.mw_zgemm_tn_add(npt, npt, D.nb, D.X64.data_ptr(), nc, D.R64.data_ptr(),
                                       nc, yptr, max(npt, 1), st), "adjoint readout")

    def _mixed_check(self, D, st, mark=None):
        """Read the one-residual readout's check (a host synchronisation);
        when it failed, read out again from a fresh FP64 residual at the
        corrected coefficients (an error quadratic in theirs; timed as the
        stage "readout2" when ``mark`` is the solve's stage marker)."""
        (gate,), stats = _to_host(D.gate, D.gstats)
        self.mixed_readout = {"delta": float(stats[0]), "max_ctr": float(stats[1]),
                              "max_y": float(stats[2]), "estimate": float(stats[3]),
                              "tol": READOUT_TOL, "second_residual": bool(gate)}
        if gate:
            if mark is not None:
                mark("readout2", True)
            self._fp64_residual(D, st)
            self._adjoint_readout(D, D.Y.data_ptr(), st)
            if mark is not None:
                mark("readout2", False)

    def _solve(self, k=None, timers=None):
        if self.precision == "mixed":
            return self._solve_device_mixed(k, timers=timers)
        return self._solve_device(k, timers=timers)

    def _capture(self, warm=True):
        """Capture the whole device solve at the current k as a HIP graph
        (``warm``: run it once eagerly first, unless a solve at this k has
        already filled the grid and the loading values).

        Everything the captured kernels read stays referenced by the entry:
        the Sommerfeld master (the grid cache may evict it) and the loading
        values of this frequency (replaced when the frequency changes)."""
        torch = _torch()
        self._device()
        if warm:
            side = torch.cuda.Stream()
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):
                self._solve()
            torch.cuda.current_stream().wait_stream(side)
        g = torch.cuda.CUDAGraph()
        self._capture_needs_check = False
        with torch.cuda.graph(g):
            out = self._solve()
        keep = (getattr(self, "_grid_master", None),
                self._dev.__dict__.get("loading", {}).get("vals"))
        self._graph_entry = (self.k, g, out, keep, self._capture_needs_check)

    def solve_graph(self, warm=True):
        """The whole device solve captured once as a HIP graph for the current
        k and replayed on each call (launch-bound panels cost one graph launch
        instead of ~10 kernel launches per 32 columns).  Returns the device
        (Y, W) views the replay writes; cancellation is not polled inside a
        replay.  Re-captured when k changes."""
        entry = getattr(self, "_graph_entry", None)
        if entry is None or entry[0] != self.k:
            self._capture(warm)
            entry = self._graph_entry
        entry[1].replay()
        if entry[4]:
            self._mixed_check(self._dev, _torch().cuda.current_stream().cuda_stream)
        return entry[2]

    def _check_info(self):
        info, = _to_host(self._dev.info)
        if info[0] == -1:
            raise np.linalg.LinAlgError("the symmetric factorisation met a panel that needs row "
                                        "exchanges; its result is not valid (compute_port_"
                                        "solution repeats such a solve with pivoting)")
        if info[0] != 0:
            raise np.linalg.LinAlgError(f"Z is singular: zero pivot at column {info[0] - 1}")
        if info[1] != 0:
            raise np.linalg.LinAlgError("KCL Schur complement is singular")

    def compute_port_solution(self, same_edge_prep=None) -> PortSolution:
        """Solve every port from ONE fill and ONE factorisation (bspline.py:4053-4179).

        The first solve at a wavenumber launches its kernels one by one; from
        the second solve at the same k on, the captured graph of the whole
        solve is replayed (one launch instead of thousands).  A solver with a
        cancel token always launches eagerly, so that the flag is polled
        between panels and refinement steps."""
        graph = self._cancel is None and getattr(self, "_last_k", None) == self.k
        Y, W = self._solve_valid(graph)
        self._last_k = self.k
        y, coeffs = _to_host(Y, W)
        self._check_info()
        return PortSolution(y=y, coeffs=coeffs, port_currents=y,
                            basis=("bspline-gpu", self._mesh, id(self)))

    def compute_y_matrix(self):
        return self.compute_port_solution().y

    def _port_voltages(self):
        """Configured port voltages in port order: [gap feeds..., junction
        ports..., node gaps...] (reference bspline.py:3881-3904)."""
        return np.array([v for _w, _s, v in self.feeds] + [v for _j, v in self.junction_ports]
                        + [v for _w, _e, v in self.node_gaps], dtype=np.complex128)

    def compute_impedance(self, same_edge_prep=None):
        """Driving-point impedance under the configured voltages, and the
        coefficient vector of that excitation: (V_i / I_i per port -- a
        scalar for one port --, coeffs), I = Y V and coeffs = W V from the one
        all-port solve (reference bspline.py:3919-4017)."""
        sol = self.compute_port_solution()
        V = self._port_voltages()
        z = V / (sol.y @ V)
        return (z[0] if len(V) == 1 else z), sol.coeffs @ V

    def compute_impedance_swept(self, k_array):
        """Driving-point impedance at each wavenumber (reference
        bspline.py:4479-4520): an array over k (of per-port arrays when there
        are several ports)."""
        with self._k_restored():
            out = []
            for kk in np.asarray(k_array, dtype=float):
                self._set_k(kk)
                out.append(self.compute_impedance()[0])
        return np.array(out)

    def _port_solutions_swept(self, k_array):
        with self._k_restored():
            for kk in np.asarray(k_array, dtype=float):
                self._set_k(kk)
                yield self.compute_port_solution()

    # ------------------------------------------------------------------
    # Readout (bspline.py:4603-4685; _element_currents.py:33-120)
    # ------------------------------------------------------------------
    def currents_at_knots(self, coeffs, s_array=None):
        coeffs = np.asarray(coeffs)
        mesh = self._build()
        d = self.degree
        out = []
        for w, pw in enumerate(mesh.per_wire):
            kv = pw["knots"]
            s_eval = pw["arc"] if s_array is None else np.asarray(s_array[w], dtype=float)
            s_eval = np.clip(s_eval, kv[0], kv[-1])
            c = np.zeros(pw["nb_w"], dtype=np.complex128)
            for kidx, (j, *_r) in enumerate(pw["kept"]):
                c[j] = coeffs[pw["l2g"][kidx]]
            vals = np.array([bspline_values(kv, d, pw["nb_w"], s) for s in s_eval])
            out.append(vals @ c if len(s_eval) else np.zeros(0, dtype=np.complex128))
        return out

    # -- readout of every port column on the GPU ------------------------
    def readout_mesh(self, subdiv: int = 1):
        """The readout tables at ``subdiv`` elements per segment (cached)."""
        from . import readout
        cache = self.__dict__.setdefault("_readout_meshes", {})
        if subdiv not in cache:
            cache[subdiv] = readout.ReadoutMesh(self, subdiv)
        return cache[subdiv]

    def _device_coeffs(self, coeffs):
        """Solution columns on the device: the port columns of a fresh solve
        (None) or the given (n_b, p) array."""
        import torch
        if coeffs is None:
            _Y, W = self._solve_valid()
            self._check_info()
            return W
        if isinstance(coeffs, torch.Tensor):
            return coeffs.to(device="cuda", dtype=torch.complex128)
        c = np.asarray(coeffs, dtype=np.complex128)
        if c.ndim == 1:
            c = c[:, None]
        return torch.from_numpy(np.ascontiguousarray(c)).cuda()

    def port_element_moments(self, coeffs=None, subdiv: int = 1):
        """``(mid, moment (n_el, 3, p), nodes, delta (n_node, p))`` for every
        column at once on the GPU (``element_currents`` per column,
        reference _element_currents.py:33-120); moment and delta stay on the
        device."""
        from . import readout
        rm = self.readout_mesh(subdiv)
        mom, delta = readout.element_moments(rm, self._device_coeffs(coeffs))
        return rm.mid, mom, rm.nodes, delta

    def port_far_moments(self, theta, phi, coeffs=None, as_numpy=True):
        """Far-field moments ``(M_theta, M_phi)``, each (n_theta, n_phi, p),
        of every port column (or of ``coeffs``) over this solver's ground:
        free space, the PEC image, or the Fresnel image of a finite ground
        (reference portal/_portal.py:1520-1598)."""
        from . import readout
        rm = self.readout_mesh(1)
        mom, _delta = readout.element_moments(rm, self._device_coeffs(coeffs))
        gz = 0.0 if self.ground_z is None else self.ground_z
        mt, mp = readout.far_moments(rm.device()["mid"], mom, self.k, theta, phi,
                                     readout.solver_ground(self), gz, self.freq)
        return (mt.cpu().numpy(), mp.cpu().numpy()) if as_numpy else (mt, mp)

    def port_near_fields(self, points, magnetic=False, coeffs=None,
                         subdiv: int = 8, as_numpy=True):
        """E (or H) at ``points`` (n, 3) for every port column, (n, 3, p),
        from the mesh resampled at ``subdiv`` elements per segment; over PEC
        the image is added (reference portal/_portal.py:3160-3204, which
        refuses a finite ground: so does this)."""
        from . import readout
        if self.ground_z is not None and self.ground_eps is not None:
            raise NotImplementedError("near fields over a finite ground (the reference's "
                                      "readout refuses it too)")
        rm = self.readout_mesh(subdiv)
        mom, delta = readout.element_moments(rm, self._device_coeffs(coeffs))
        dev = rm.device()
        out = readout.element_fields(points, (dev["mid"], mom, dev["nodes"], delta), self.k,
                                     float(np.min(self._radius_per_wire)), magnetic,
                                     image_z=self.ground_z)
        return out.cpu().numpy() if as_numpy else out

    def port_ground_remainder(self, points, magnetic=False, coeffs=None, subdiv: int = 8,
                              as_numpy=True):
        """The Sommerfeld ground's REMAINDER of E (or H, by the reference's
        central-difference curl) at ``points`` for every port column, (n, 3,
        p): ``reflected_field_at`` / ``reflected_h_field_at`` (reference
        _field_point.py:292-432) fed with this solve's element moments at
        ``subdiv`` elements per segment.  The full field over the ground is
        E_direct + C2 E_image + this remainder, C2 = (eps~ - 1)/(eps~ + 1);
        composing it is the caller's, as in the reference."""
        from . import _ground, readout
        if self.ground_model != "sommerfeld" or self.ground_eps is None:
            raise ValueError("the ground remainder needs a Sommerfeld ground")
        rm = self.readout_mesh(subdiv)
        mom, _delta = readout.element_moments(rm, self._device_coeffs(coeffs))
        eps_t = _ground.eps_tilde(self.ground_eps, self.omega, self.eps)
        fn = readout.reflected_h_field_at if magnetic else readout.reflected_field_at
        out = fn(points, rm.mid, mom, eps_t, self.ground_z, self.k, self.omega, mu=self.mu)
        if as_numpy:
            return out if isinstance(out, np.ndarray) else out.cpu().numpy()
        return out

    def current_slopes(self, coeffs, s_array=None):
        """Per-wire dI/ds at the mesh knots or at s_array per wire: the
        B-spline sum's exact arc-length derivative (reference
        bspline.py:4523-4584)."""
        coeffs = np.asarray(coeffs)
        mesh = self._build()
        d = self.degree
        out = []
        for w, pw in enumerate(mesh.per_wire):
            kv = pw["knots"]
            s_eval = pw["arc"] if s_array is None else np.asarray(s_array[w], dtype=float)
            s_eval = np.clip(s_eval, kv[0], kv[-1])
            c = np.zeros(pw["nb_w"], dtype=np.complex128)
            for kidx, (j, *_r) in enumerate(pw["kept"]):
                c[j] = coeffs[pw["l2g"][kidx]]
            vals = np.array([bspline_slopes(kv, d, pw["nb_w"], x) for x in s_eval])
            out.append(vals @ c if len(s_eval) else np.zeros(0, dtype=np.complex128))
        return out

    def element_currents(self, coeffs, subdiv: int = 1):
        subdiv = int(subdiv)
        if subdiv < 1:
            raise ValueError(f"subdiv must be >= 1, got {subdiv}")
        coeffs = np.asarray(coeffs)
        if coeffs.ndim != 1:
            raise ValueError("element_currents takes ONE solution vector")
        fine, arcs = [], []
        for pl, npe in zip(self.wires_polylines, self.n_per_edge_per_wire):
            parts = []
            for i, n_e in enumerate(npe):
                s = np.linspace(pl[i], pl[i + 1], n_e * subdiv + 1)
                parts.append(s if i == 0 else s[1:])
            kn = np.vstack(parts)
            fine.append(kn)
            step = np.linalg.norm(kn[1:] - kn[:-1], axis=1)
            arcs.append(np.concatenate([[0.0], np.cumsum(step)]))
        cur = self.currents_at_knots(coeffs, None if subdiv == 1 else arcs)
        mids, moms, nodes, deltas = [], [], [], []
        for kn, c in zip(fine, cur):
            el = 0.5 * (c[1:] + c[:-1])
            mids.append(0.5 * (kn[1:] + kn[:-1]))
            moms.append(el[:, None] * (kn[1:] - kn[:-1]))
            nodes.append(kn)
            z = np.zeros(1, dtype=np.complex128)
            deltas.append(np.concatenate([z, el]) - np.concatenate([el, z]))
        return (np.concatenate(mids), np.concatenate(moms), np.concatenate(nodes),
                np.concatenate(deltas))
