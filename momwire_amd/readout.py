"""Readout of solved port columns on the GPU, every column at once.

The reference reads one solution column at a time on the host:
``element_currents`` walks the mesh knot to knot (_element_currents.py:33-120),
``_far_moments`` sums the element moments into a far-field pattern
(portal/_portal.py:1520-1598, image terms 1384-1423) and ``_element_fields``
evaluates the near field in mixed-potential form (portal/_portal.py:1601-1651).
Here the same three steps run as HIP kernels over all port columns of a
solve, with the solution left where the solver put it:

* ``ReadoutMesh`` holds the k-independent tables of a structure's readout
  mesh: per knot (or sub-knot) sample the B-spline taps onto the solver's
  unknowns, per element its end samples, midpoint and vector length, per knot
  the elements arriving and leaving.
* ``element_moments`` -> ``mw_element_moments``;
  ``far_moments`` -> ``mw_far_moments`` (FP64 MFMA, phase tiles generated in
  LDS); ``element_fields`` -> ``mw_near_fields``;
  ``reflected_field_at`` / ``reflected_h_field_at`` (the Sommerfeld ground
  remainder at points, _field_point.py) -> ``mw_ground_fields``.

``far_moments`` and ``element_fields`` take the reference's arguments (one
column: ``moment`` (n, 3), ``delta`` (n,)) or a trailing column axis
(``moment`` (n, 3, p), ``delta`` (n, p)); numpy in gives numpy out, device
tensors in give device tensors out.  Nothing here falls back to the host: a
missing library or GPU raises.
"""

from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib
from ._mesh import bspline_values

ETA0 = 376.730_313_668      # portal/_portal.py:347
EPS0 = 8.854_187_817e-12    # portal/_portal.py:346 (the Fresnel step's eps0)
NEAR_FIELD_SUBDIV = 8       # portal/_portal.py:639

_GROUND_CODES = {"free": 0, "pec": 1}


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("the readout kernels need a GPU (torch.cuda is not available)")
    return torch


def _stream(torch):
    return torch.cuda.current_stream().cuda_stream


class ReadoutMesh:
    """Host and device tables of one solver's readout mesh at ``subdiv``
    elements per mesh segment (reference _element_currents.py:86-120: knots
    from ``np.linspace`` along each polyline edge, currents at the solver's
    own knot arcs for ``subdiv = 1`` and at chord arcs otherwise)."""

    def __init__(self, solver, subdiv: int = 1):
        subdiv = int(subdiv)
        if subdiv < 1:
            raise ValueError(f"subdiv must be >= 1, got {subdiv}")
        self.subdiv = subdiv
        mesh = solver._build()
        d = solver.degree
        tap_i, tap_v, el_tab, mids, dls, nodes, node_tab = [], [], [], [], [], [], []
        n_samp = n_el = 0
        for w, (pl, npe) in enumerate(zip(solver.wires_polylines, solver.n_per_edge_per_wire)):
            parts = []
            for i, n_e in enumerate(npe):
                seg = np.linspace(pl[i], pl[i + 1], n_e * subdiv + 1)
                parts.append(seg if i == 0 else seg[1:])
            kn = np.vstack(parts)
            pw = mesh.per_wire[w]
            if subdiv == 1:
                s_eval = np.asarray(pw["arc"], dtype=float)
            else:
                step = np.linalg.norm(kn[1:] - kn[:-1], axis=1)
                s_eval = np.concatenate([[0.0], np.cumsum(step)])
            if len(s_eval) != len(kn):
                raise ValueError(f"wire {w}: {len(s_eval)} sample arcs for {len(kn)} knots")
            kv = pw["knots"]
            s_eval = np.clip(s_eval, kv[0], kv[-1])
            glob = np.full(pw["nb_w"], -1, dtype=np.int64)
            for kidx, (j, *_r) in enumerate(pw["kept"]):
                glob[j] = pw["l2g"][kidx]
            for s in s_eval:
                vals = bspline_values(kv, d, pw["nb_w"], s)
                nz = [j for j in np.nonzero(vals)[0] if glob[j] >= 0]
                if len(nz) > 3:
                    raise ValueError("more than three B-spline taps at a sample (degree > 2)")
                ti = [int(glob[j]) for j in nz] + [-1] * (3 - len(nz))
                tv = [float(vals[j]) for j in nz] + [0.0] * (3 - len(nz))
                tap_i.append(ti)
                tap_v.append(tv)
            n_k = len(kn)
            for i in range(n_k - 1):
                el_tab.append((n_samp + i, n_samp + i + 1))
            mids.append(0.5 * (kn[1:] + kn[:-1]))
            dls.append(kn[1:] - kn[:-1])
            nodes.append(kn)
            for i in range(n_k):
                node_tab.append((n_el + i - 1 if i >= 1 else -1, n_el + i if i < n_k - 1 else -1))
            n_samp += n_k
            n_el += n_k - 1
        self.n_el, self.n_node, self.n_samp = n_el, len(node_tab), n_samp
        self.n_basis = mesh.n_basis
        self.mid = np.ascontiguousarray(np.concatenate(mids, axis=0))
        self.dl = np.ascontiguousarray(np.concatenate(dls, axis=0))
        self.nodes = np.ascontiguousarray(np.concatenate(nodes, axis=0))
        self.tap_i = np.asarray(tap_i, dtype=np.int32)
        self.tap_v = np.asarray(tap_v, dtype=np.float64)
        self.el_tab = np.asarray(el_tab, dtype=np.int32).reshape(-1, 2)
        self.node_tab = np.asarray(node_tab, dtype=np.int32).reshape(-1, 2)
        if self.tap_i.size and self.tap_i.max() >= self.n_basis:
            raise ValueError("tap index past the solver's unknowns")
        self._dev = None

    def device(self):
        """The tables as device tensors (uploaded once)."""
        if self._dev is None:
            torch = _torch()
            up = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
            self._dev = dict(tap_i=up(self.tap_i), tap_v=up(self.tap_v), el_tab=up(self.el_tab),
                             dl=up(self.dl), node_tab=up(self.node_tab), mid=up(self.mid),
                             nodes=up(self.nodes))
        return self._dev


def element_moments(rmesh: ReadoutMesh, X):
    """Element moments (n_el, 3, p) and continuity steps (n_node, p) of the
    solution columns ``X`` (a complex128 device tensor (n_b, p), e.g. the
    solver's device solution, read in place when its columns are unit-stride)
    -- the batched form of
    ``element_currents`` (reference _element_currents.py:33-120)."""
    torch = _torch()
    lib = _lib.load()
    if X.dim() != 2 or X.shape[0] != rmesh.n_basis:
        raise ValueError(f"X must be (n_b={rmesh.n_basis}, p), got {tuple(X.shape)}")
    if X.dtype != torch.complex128 or not X.is_cuda:
        raise ValueError("X must be a complex128 device tensor")
    if X.stride(1) != 1:
        X = X.contiguous()
    p = int(X.shape[1])
    D = rmesh.device()
    mom = torch.empty((rmesh.n_el, 3, p), dtype=torch.complex128, device=X.device)
    delta = torch.empty((rmesh.n_node, p), dtype=torch.complex128, device=X.device)
    _lib.check(lib.mw_element_moments(rmesh.n_el, rmesh.n_node, p, D["tap_i"].data_ptr(),
                                      D["tap_v"].data_ptr(), D["el_tab"].data_ptr(),
                                      D["dl"].data_ptr(), D["node_tab"].data_ptr(), X.data_ptr(),
                                      X.stride(0), mom.data_ptr(), 3 * p, delta.data_ptr(), p,
                                      _stream(torch)), "element moments")
    return mom, delta


def _ground_args(ground, freq_hz):
    """(code, eps_re, eps_im): 0 free, 1 PEC, 2 finite (eps_c = eps_r -
    j sigma / (omega eps0), reference portal/_portal.py:1397-1423)."""
    if ground is None:
        return 0, 0.0, 0.0
    kind = getattr(ground, "kind", ground)
    if kind in _GROUND_CODES:
        return _GROUND_CODES[kind], 0.0, 0.0
    if freq_hz is None:
        raise ValueError("a finite ground needs freq_hz")
    omega = 2.0 * math.pi * freq_hz
    eps_c = ground.eps_r - 1j * ground.sigma / (omega * EPS0)
    return 2, float(eps_c.real), float(eps_c.imag)


def _as_device(a, torch, dtype):
    """(device tensor, was it one already)."""
    if isinstance(a, torch.Tensor):
        return a.to(device="cuda", dtype=dtype).contiguous(), True
    np_dtype = np.complex128 if dtype == torch.complex128 else np.float64
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np_dtype))).cuda(), False


def far_moments(mid, moment, k, theta, phi, ground=None, ground_z=0.0, freq_hz=None,
                cliff=None):
    """``(M_theta, M_phi)`` on the theta x phi grid (radians) -- the
    reference's ``_far_moments`` (portal/_portal.py:1520-1598) for one column
    (``moment`` (n, 3) -> (n_theta, n_phi)) or many (``moment`` (n, 3, p) ->
    (n_theta, n_phi, p)).  ``ground``: None or an object with ``kind``
    ("free", "pec", or a finite kind with ``eps_r`` and ``sigma``)."""
    if cliff is not None:
        raise NotImplementedError("cliff far-field modes (RP 2/3 with a second medium)")
    torch = _torch()
    lib = _lib.load()
    theta = np.atleast_1d(np.asarray(theta, dtype=float))
    phi = np.atleast_1d(np.asarray(phi, dtype=float))
    mid_d, _ = _as_device(mid, torch, torch.float64)
    mom_d, on_dev = _as_device(moment, torch, torch.complex128)
    single = mom_d.dim() == 2
    if single:
        mom_d = mom_d.unsqueeze(-1)
    n_el, p = int(mom_d.shape[0]), int(mom_d.shape[2])
    if mom_d.shape[1] != 3 or tuple(mid_d.shape) != (n_el, 3):
        raise ValueError("mid must be (n, 3) and moment (n, 3) or (n, 3, p)")
    trig_t = torch.from_numpy(np.stack([np.sin(theta), np.cos(theta)], axis=1)).cuda()
    trig_p = torch.from_numpy(np.stack([np.sin(phi), np.cos(phi)], axis=1)).cuda()
    code, er, ei = _ground_args(ground, freq_hz)
    mt = torch.empty((len(theta), len(phi), p), dtype=torch.complex128, device="cuda")
    mp = torch.empty_like(mt)
    _lib.check(lib.mw_far_moments(n_el, mid_d.data_ptr(), mom_d.data_ptr(), 3 * p, p, len(theta),
                                  len(phi), trig_t.data_ptr(), trig_p.data_ptr(), float(k), code,
                                  float(ground_z), er, ei, mt.data_ptr(), mp.data_ptr(), p,
                                  _stream(torch)), "far-field moments")
    if single:
        mt, mp = mt[..., 0], mp[..., 0]
    if on_dev:
        return mt, mp
    return mt.cpu().numpy(), mp.cpu().numpy()


def element_fields(points, elements, k, radius, magnetic, image_z=None):
    """E (or H, ``magnetic``) at ``points`` from ``elements = (mid, moment,
    nodes, delta)`` -- the reference's ``_element_fields``
    (portal/_portal.py:1601-1651) for one column (``moment`` (n, 3),
    ``delta`` (n_nodes,) -> (n_points, 3)) or many (``moment`` (n, 3, p),
    ``delta`` (n_nodes, p) -> (n_points, 3, p)).  ``image_z``: add the PEC
    image about z = image_z (moments mirrored with horizontal components
    negated, charges negated; portal/_portal.py:3187-3204)."""
    torch = _torch()
    lib = _lib.load()
    mid, moment, nodes, delta = elements
    pts_d, _ = _as_device(np.asarray(points, dtype=float).reshape(-1, 3)
                          if not isinstance(points, torch.Tensor) else points, torch, torch.float64)
    mid_d, _ = _as_device(mid, torch, torch.float64)
    nod_d, _ = _as_device(nodes, torch, torch.float64)
    mom_d, on_dev = _as_device(moment, torch, torch.complex128)
    del_d, _ = _as_device(delta, torch, torch.complex128)
    single = mom_d.dim() == 2
    if single:
        mom_d = mom_d.unsqueeze(-1)
        del_d = del_d.reshape(-1, 1)
    n_el, p = int(mom_d.shape[0]), int(mom_d.shape[2])
    n_node, n_pts = int(nod_d.shape[0]), int(pts_d.shape[0])
    if tuple(del_d.shape) != (n_node, p) or tuple(mid_d.shape) != (n_el, 3):
        raise ValueError("elements must be mid (n, 3), moment (n, 3[, p]), nodes (m, 3), "
                         "delta (m[, p])")
    out = torch.empty((n_pts, 3, p), dtype=torch.complex128, device="cuda")
    st = _stream(torch)

    def call(m_mid, m_mom, m_nodes, sign, acc):
        _lib.check(lib.mw_near_fields(n_pts, pts_d.data_ptr(), n_el, m_mid.data_ptr(),
                                      m_mom.data_ptr(), 3 * p, n_node, m_nodes.data_ptr(),
                                      del_d.data_ptr(), p, sign, p, float(k), float(radius), ETA0,
                                      int(bool(magnetic)), acc, out.data_ptr(), 3 * p, st),
                   "near fields")

    call(mid_d, mom_d, nod_d, 1.0, 0)
    if image_z is not None:
        flip = torch.tensor([-1.0, -1.0, 1.0], dtype=torch.float64, device="cuda")
        mid_i = mid_d.clone()
        mid_i[:, 2] = 2.0 * image_z - mid_d[:, 2]
        nod_i = nod_d.clone()
        nod_i[:, 2] = 2.0 * image_z - nod_d[:, 2]
        mom_i = (mom_d * flip.view(1, 3, 1)).contiguous()
        call(mid_i, mom_i, nod_i, -1.0, 1)
    if single:
        out = out[..., 0]
    if on_dev:
        return out
    return out.cpu().numpy()


def solver_ground(solver):
    """The far-field ground of a solver: None (free space), "pec", or a
    finite ground with (eps_r, sigma) (a complex ground_eps is folded back to
    sigma at the solver's frequency with the same eps0)."""
    if solver.ground_z is None:
        return None
    if solver.ground_eps is None:
        return "pec"
    ge = solver.ground_eps
    if isinstance(ge, (tuple, list)):
        eps_r, sigma = float(ge[0]), float(ge[1])
    else:
        ge = complex(ge)
        eps_r, sigma = ge.real, -ge.imag * solver.omega * EPS0

    class _Finite:
        kind = "finite"

    g = _Finite()
    g.eps_r, g.sigma = eps_r, sigma
    return g


# ---------------------------------------------------------------------------
# Ground remainder at points in space (reference _field_point.py)
# ---------------------------------------------------------------------------

CURL_STEP_WAVELENGTHS = 1e-3   # _field_point.py:94
_REGIMES = {
    ("above", "above"): None,
    ("below", "below"): "a source below the interface read at an observer below it",
    ("below", "above"): "a source below the interface read at an observer above it",
    ("above", "below"): "a source above the interface read at an observer below it",
}


def _check_regime(source_medium, observer_medium):
    """The ("above", "above") reflected wave is the one regime with an
    evaluator; the others raise ValueError naming themselves
    (_field_point.py:266-289)."""
    key = (source_medium, observer_medium)
    if key not in _REGIMES:
        raise ValueError(f"unknown ground regime {key!r}: source_medium and observer_medium "
                         "must each be 'above' or 'below'")
    if _REGIMES[key] is not None:
        raise ValueError(f"the {key!r} ground regime -- {_REGIMES[key]} -- has no field "
                         "evaluator here: its field is carried by the transmitted-field "
                         "Sommerfeld surfaces, a different integral family. Only the "
                         "('above', 'above') reflected-wave regime is implemented.")


def ground_grid_for(points, src_mid, eps_t, ground_z, k, omega, mu=None, cancel=None,
                    filler=None):
    """(device master, physical view) of the Sommerfeld grid for these
    observer/source pairs, sized as the reference sizes it: the largest
    observer-to-image distance over points and sources together
    (_field_point.py:122-142, _sommerfeld.py:1081-1116)."""
    from . import _ground

    mu = _ground.MU0_EXACT if mu is None else float(mu)
    ex = np.concatenate([np.asarray(points, dtype=float).reshape(-1, 3),
                         np.asarray(src_mid, dtype=float).reshape(-1, 3)])
    r1 = _ground.max_image_distance(ex, ex, float(ground_z))
    master = _ground.get_master(complex(eps_t), float(k), r1, cancel=cancel, filler=filler)
    return master, master.struct(float(k), float(omega), mu)


def reflected_field_at(points, src_mid, src_moment, eps_t, ground_z, k, omega, *,
                       source_medium="above", observer_medium="above", mu=None, grid=None):
    """The ground REMAINDER of E at ``points``: (n_points, 3) complex for
    ``src_moment`` (n_src, 3), or (n_points, 3, p) for (n_src, 3, p) -- the
    reference's ``reflected_field_at`` (_field_point.py:292-331) batched over
    port columns, on the GPU (``mw_ground_fields``).  ``eps_t`` is the complex
    relative permittivity (e^{+jwt}, Im <= 0), ``omega`` the radian frequency
    the surfaces are normalised at, ``grid`` an optional (master, view) pair
    from ``ground_grid_for`` so that related evaluations share one fill.
    E(point) = E_direct + C2 E_image + this; composing is the caller's."""
    _check_regime(source_medium, observer_medium)
    torch = _torch()
    lib = _lib.load()
    pts_d, _ = _as_device(np.asarray(points, dtype=float).reshape(-1, 3)
                          if not isinstance(points, torch.Tensor) else points, torch, torch.float64)
    mid_d, _ = _as_device(src_mid, torch, torch.float64)
    mom_d, on_dev = _as_device(src_moment, torch, torch.complex128)
    single = mom_d.dim() == 2
    if single:
        mom_d = mom_d.unsqueeze(-1)
    n_src, p, n_pts = int(mom_d.shape[0]), int(mom_d.shape[2]), int(pts_d.shape[0])
    if tuple(mid_d.shape) != (n_src, 3) or mom_d.shape[1] != 3 or pts_d.shape[-1] != 3:
        raise ValueError("expected points (n, 3), src_mid (s, 3), src_moment (s, 3[, p])")
    if grid is None:
        grid = ground_grid_for(pts_d.cpu().numpy(), mid_d.cpu().numpy(), eps_t, ground_z, k,
                               omega, mu)
    view = grid[1]
    out = torch.empty((n_pts, 3, p), dtype=torch.complex128, device="cuda")
    _lib.check(lib.mw_ground_fields(ctypes.byref(view), float(ground_z), float(k), n_pts,
                                    pts_d.data_ptr(), n_src, mid_d.data_ptr(), mom_d.data_ptr(),
                                    3 * p, p, 0, out.data_ptr(), 3 * p, _stream(torch)),
               "ground remainder field")
    if single:
        out = out[..., 0]
    return out if on_dev else out.cpu().numpy()


def reflected_h_field_at(points, src_mid, src_moment, eps_t, ground_z, k, omega, *,
                         source_medium="above", observer_medium="above", mu=None, step=None):
    """The ground REMAINDER of H: curl E_rem / (-j omega mu) by central
    differences at ``step`` (default 1e-3 wavelength), six evaluations of
    ``reflected_field_at`` on ONE grid sized over all offset points, as the
    reference does (_field_point.py:334-432)."""
    from . import _ground

    _check_regime(source_medium, observer_medium)
    mu_v = _ground.MU0_EXACT if mu is None else float(mu)
    pts = np.asarray(points, dtype=float).reshape(-1, 3)
    if step is None:
        step = CURL_STEP_WAVELENGTHS * (2.0 * math.pi / k)
    offsets = []
    for axis in range(3):
        for sign in (1.0, -1.0):
            shift = np.zeros(3)
            shift[axis] = sign * step
            offsets.append(pts + shift)
    grid = ground_grid_for(np.concatenate(offsets), np.asarray(src_mid, dtype=float), eps_t,
                           ground_z, k, omega, mu_v)
    torch = _torch()
    on_dev = isinstance(src_moment, torch.Tensor)
    f = [reflected_field_at(o, src_mid, src_moment, eps_t, ground_z, k, omega, mu=mu_v, grid=grid)
         for o in offsets]
    if on_dev:
        f = [x.cpu().numpy() for x in f]
    d = [(f[2 * a] - f[2 * a + 1]) / (2.0 * step) for a in range(3)]
    curl = np.stack([d[1][:, 2] - d[2][:, 1], d[2][:, 0] - d[0][:, 2], d[0][:, 1] - d[1][:, 0]],
                    axis=1)
    h = curl / (-1j * omega * mu_v)
    return torch.from_numpy(h).cuda() if on_dev else h
