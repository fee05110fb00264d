"""Ground planes for the GPU port solve: the ground decision and the
Sommerfeld interpolation grid, built on the device by the HIP engine.

PEC ground (``ground_z`` without ``ground_eps``) is the exact image with
coefficient 1; the Sommerfeld ground (``ground_model="sommerfeld"``) is the
exact image scaled by C2 = (eps - 1)/(eps + 1) plus the smooth remainder Q
interpolated from four tabulated surfaces (reference
src/momwire/_ground_spec.py:90-125, _potential_ground.py:779-883,
bspline.py:3730-3770).  The grounded fill (``mw_fill_ground``) subtracts both
from the free-space Z.

The grid follows the reference's SommerfeldGrid (_sommerfeld.py:642-914): a
normalised master at lambda = 1 per (eps bucket, extent bucket), tabulated on
4 or 6 (R1, theta) regions, rescaled to each frequency by a length scale and
one factor omega*mu.  Here every node of a master is one wavefront of the HIP
engine (``mw_somm_surfaces``), so a cold grid costs milliseconds of GPU time;
masters stay resident in HBM in a small FIFO cache.
"""

import ctypes
import cmath
import math

import numpy as np

from . import _lib

C_LIGHT = 299792458.0
MU0_EXACT = 4e-7 * np.pi  # the normalised masters' mu (reference _sommerfeld.py:133)
K2_REF = 2.0 * np.pi      # lambda = 1
R1_CAP_WL = 15.0          # extent cap, wavelengths (_sommerfeld.py:67)
R1_NEAR_WL = 4.0          # near/far split (_sommerfeld.py:99)
DR_FAR_WL, DTH_FAR_DEG = 0.2, 2.5
N0_CAP = 513              # steep-band node cap (_sommerfeld.py:110)
TH_SPLIT_DEG = 30.0
EPS_IM_STEP = 0.01        # Im(eps) ladder (_sommerfeld.py:133)
GRID_RTOL = 1e-6          # the grid fill's quadrature tolerance (_sommerfeld.py:688)
_MASTER_CACHE_MAX = 32


class MwSommGrid(ctypes.Structure):
    _fields_ = [
        ("n_reg", ctypes.c_int32), ("pad", ctypes.c_int32),
        ("r1_max", ctypes.c_double), ("r_break", ctypes.c_double),
        ("th_split", ctypes.c_double), ("r_near", ctypes.c_double),
        ("value_scale", ctypes.c_double),
        ("r0", ctypes.c_double * 6), ("dr", ctypes.c_double * 6),
        ("th0", ctypes.c_double * 6), ("dth", ctypes.c_double * 6),
        ("n_r", ctypes.c_int32 * 6), ("n_th", ctypes.c_int32 * 6),
        ("off", ctypes.c_int64 * 6),
        ("vals", ctypes.c_void_p),
    ]


def eps_tilde(ground_eps, omega, eps0):
    """Complex relative permittivity: a complex value as given, or
    (eps_r, sigma) -> eps_r - j sigma/(omega eps0) (reference _ground_refl.py:46-60)."""
    if isinstance(ground_eps, (tuple, list)):
        if len(ground_eps) != 2:
            raise ValueError(f"ground_eps tuple must be (eps_r, sigma), got {ground_eps!r}")
        return complex(float(ground_eps[0]), -float(ground_eps[1]) / (omega * eps0))
    return complex(ground_eps)


def image_coefficient(eps_t):
    """C2 = (eps - 1)/(eps + 1); 1 for PEC (eps_t None)."""
    return 1.0 + 0j if eps_t is None else (eps_t - 1.0) / (eps_t + 1.0)


PHI_MODES = ("rho_v", "image", "normal", "blend")


def phi_coefficients(mode, eps_t):
    """The reflection-coefficient ground's charge weight as w_Phi = c0 +
    c1 rho_v (reference _ground_refl.py:276-289): "rho_v" the TM coefficient
    at the pair's specular angle, "image" the quasi-static (eps-1)/(eps+1),
    "normal" rho_v at normal incidence (the default), "blend" the mean of
    "rho_v" and "image"."""
    if mode == "rho_v":
        return 0j, 1 + 0j
    if mode == "image":
        return (eps_t - 1.0) / (eps_t + 1.0), 0j
    if mode == "normal":
        r = cmath.sqrt(eps_t)
        return (r - 1.0) / (r + 1.0), 0j
    if mode == "blend":
        return 0.5 * (eps_t - 1.0) / (eps_t + 1.0), 0.5 + 0j
    raise ValueError(f"ground_phi_mode must be one of {PHI_MODES}, got {mode!r}")


def eps_bucket(eps_t):
    """Im(eps) rounded onto the 1 % geometric ladder; Re kept exactly
    (reference _sommerfeld.py:1064-1078)."""
    step = 1.0 + EPS_IM_STEP
    re, im = eps_t.real, eps_t.imag
    if not (re > 0.0) or im >= 0.0:
        return eps_t
    return complex(re, -(step ** round(math.log(-im, step))))


def r1_bucket_wl(r1_wl):
    """Extent rounded up to the next 1.25^n wavelengths, floor 0.1
    (reference _sommerfeld.py:1048-1055)."""
    x = max(float(r1_wl), 0.1)
    b = 1.25 ** math.ceil(math.log(x, 1.25) - 1e-12)
    return float(b * 1.25 if b < r1_wl else b)


def max_image_distance(seg_l, seg_r, gz, rows=None):
    """Largest distance from a segment endpoint to an image endpoint, x 1.001
    (reference _sommerfeld.py:154-190); row-chunked."""
    ex = np.concatenate([seg_l, seg_r])
    n = ex.shape[0]
    rows = rows or max(1, 4_000_000 // (16 * n))
    x, y, z = ex[:, 0], ex[:, 1], ex[:, 2] - gz
    best = 0.0
    for i0 in range(0, n, rows):
        i1 = min(i0 + rows, n)
        d2 = ((x[i0:i1, None] - x[None]) ** 2 + (y[i0:i1, None] - y[None]) ** 2
              + (z[i0:i1, None] + z[None]) ** 2)
        best = max(best, float(np.sqrt(d2.max())))
    return best * 1.001


def grid_layout(eps_t, k2, r1_max):
    """Regions (r0, r1, dr, th0_deg, th1_deg, dth_deg) and (r1_max, r_break,
    r_near, r0_fill) of a grid (reference _sommerfeld.py:688-812): fine
    near-zone lattices keyed to the lateral-wave beat and the interface
    boundary layer, coarse far zones past 4 wavelengths, every zone split at
    30 degrees."""
    lam = 2.0 * np.pi / k2
    r1_max = min(max(float(r1_max), 0.35 * lam), R1_CAP_WL * lam)
    k1 = k2 * np.sqrt(complex(eps_t))
    if k1.imag > 0:
        k1 = np.conj(k1)
    beat = 2.0 * np.pi / max(abs(k1 - k2), 1e-30) if abs(k1) <= 12.0 * k2 else np.inf
    r_near = min(r1_max, R1_NEAR_WL * lam)
    dth2_t = min(5.0, np.degrees(0.07 * lam / r_near))
    n_th2 = int(np.ceil(TH_SPLIT_DEG / dth2_t)) + 1
    dth2 = TH_SPLIT_DEG / (n_th2 - 1)
    r_break = 0.2 * lam
    layer = 1.0 / max(abs(k1), 1e-300)
    dr0 = min(0.01 * lam, max(layer / 4.0, r_break / (N0_CAP - 1)))
    r0_fill = 0.5 * dr0 if layer < 0.25 * dr0 else 0.0
    sp = TH_SPLIT_DEG
    regions = [(0.0, r_break, 0.01 * lam, 0.0, sp, 10.0),
               (0.0, r_break, dr0, sp, 90.0, 10.0),
               (r_break, r_near, min(0.05 * lam, beat / 6.0), 0.0, sp, dth2),
               (r_break, r_near, min(0.1 * lam, beat / 6.0), sp, 90.0, 10.0)]
    if r1_max > r_near * (1.0 + 1e-9):
        regions += [(r_near, r1_max, DR_FAR_WL * lam, 0.0, sp, DTH_FAR_DEG),
                    (r_near, r1_max, DR_FAR_WL * lam, sp, 90.0, 10.0)]
    return regions, (r1_max, r_break, r_near, r0_fill)


def grid_nodes(regions, r0_fill):
    """Per region (n_r, n_th) and the flat node list (R1, theta), region
    after region, R1-major; region 1's first row sits at r0_fill when the
    boundary layer is unresolvable (reference _sommerfeld.py:786-796)."""
    shapes, R, T = [], [], []
    for idx, (r0, r1, dr, th0, th1, dth) in enumerate(regions):
        n_r = max(int(np.ceil((r1 - r0) / dr)) + 1, 4)
        n_th = int(round((th1 - th0) / dth)) + 1
        rr, tt = np.meshgrid(r0 + dr * np.arange(n_r), np.radians(th0 + dth * np.arange(n_th)),
                             indexing="ij")
        if idx == 1 and r0_fill > 0.0:
            rr = rr.copy()
            rr[0, :] = r0_fill
        shapes.append((n_r, n_th))
        R.append(rr.ravel())
        T.append(tt.ravel())
    return shapes, np.concatenate(R), np.concatenate(T)


class GridMaster:
    """A normalised (lambda = 1) grid resident on the device.

    ``fill`` is the node filler: ``"engine"`` runs the exact HIP engine on
    every node; any callable ``fill(R1, theta, eps_t) -> (n, 4) complex
    device tensor or None`` may replace it (the learned surrogate), with the
    engine still filling the nodes it leaves to it (see ``learned``).
    """

    def __init__(self, eps_t, r1_wl, cancel=None, filler=None):
        import torch

        lib = _lib.load()
        self.eps_t = complex(eps_t)
        regions, scalars = grid_layout(self.eps_t, K2_REF, r1_wl)
        self.regions = regions
        self.r1_max, self.r_break, self.r_near, self.r0_fill = scalars
        self.shapes, R, T = grid_nodes(regions, self.r0_fill)
        self.n_nodes = R.size
        dev = torch.device("cuda")
        self.R1 = torch.from_numpy(R).to(dev)
        self.theta = torch.from_numpy(T).to(dev)
        self.vals = torch.empty((self.n_nodes, 4), dtype=torch.complex128, device=dev)
        self.engine_nodes = self.n_nodes
        st = torch.cuda.current_stream().cuda_stream
        addr = cancel.address if cancel is not None else None
        if filler is None:
            _lib.check(lib.mw_somm_surfaces(self.eps_t.real, self.eps_t.imag, K2_REF,
                                            K2_REF * C_LIGHT, MU0_EXACT, self.R1.data_ptr(),
                                            self.theta.data_ptr(), self.n_nodes, GRID_RTOL,
                                            self.vals.data_ptr(), addr, st), "Sommerfeld grid")
        else:
            self.engine_nodes = filler(self, lib, st, addr)
        if cancel is not None:
            # the engine's nodes stop early once the flag is seen, leaving
            # their values undefined: decide after the fill has finished
            torch.cuda.current_stream().synchronize()
            if cancel.cancelled:
                raise _lib.SolveAborted("Sommerfeld grid fill cancelled")
        off = np.concatenate([[0], np.cumsum([a * b for a, b in self.shapes])])[:-1]
        self.off = off

    def struct(self, k, omega, mu):
        """The physical-units view at wavenumber k (reference
        _sommerfeld.py:878-914): lengths x lambda, values x omega mu / (2 pi c mu0)."""
        sc = K2_REF / k
        g = MwSommGrid()
        g.n_reg = len(self.regions)
        g.r1_max, g.r_break, g.r_near = self.r1_max * sc, self.r_break * sc, self.r_near * sc
        g.th_split = math.radians(TH_SPLIT_DEG)
        g.value_scale = (omega * mu) / (K2_REF * C_LIGHT * MU0_EXACT)
        for i, ((r0, _r1, dr, th0, _th1, dth), (n_r, n_th)) in enumerate(zip(self.regions,
                                                                            self.shapes)):
            g.r0[i], g.dr[i] = r0 * sc, dr * sc
            g.th0[i], g.dth[i] = math.radians(th0), math.radians(dth)
            g.n_r[i], g.n_th[i] = n_r, n_th
            g.off[i] = int(self.off[i])
        g.vals = self.vals.data_ptr()
        return g


_MASTERS = {}


def get_master(eps_t, k, r1_max, cancel=None, filler=None, key_extra=None):
    """The cached normalised master for (eps bucket, extent bucket)
    (reference _sommerfeld.py:1081-1116)."""
    lam = 2.0 * np.pi / k
    r1b = r1_bucket_wl(min(float(r1_max) / lam, R1_CAP_WL))
    eb = eps_bucket(complex(eps_t))
    key = (eb, r1b, key_extra)
    m = _MASTERS.get(key)
    if m is None:
        while len(_MASTERS) >= _MASTER_CACHE_MAX:
            _MASTERS.pop(next(iter(_MASTERS)))
        m = GridMaster(eb, r1b, cancel=cancel, filler=filler)
        _MASTERS[key] = m
    return m


def clear_cache():
    _MASTERS.clear()
