"""CPU oracle for the reflection-coefficient finite ground -- TEST
INFRASTRUCTURE ONLY.

A numpy restatement of the reference's per-segment-pair Fresnel weights
(src/momwire/_ground_refl.py): the image block of a grounded Z is the PEC
image (mirrored sources, tangent dot t_m . M t_n) with the A term weighted by

    w_A = rho_v (t_m . M t_n - P) - rho_h P,   P = (t_m . p)(t_n . p)

and the charge term by w_Phi (per phi mode), where rho_v, rho_h are the
Fresnel coefficients at the specular angle of the ray from the source's image
midpoint to the observer's midpoint and p is the horizontal unit vector
normal to the plane of incidence.  Pinned by the reference's own weight
tables and Y (tests/golden/refl_*.npz) and by closed forms: at normal
incidence rho_v = -rho_h = (sqrt(eps) - 1)/(sqrt(eps) + 1); as eps -> inf
every weight goes to the PEC image's.
"""

import numpy as np

TINY = 1e-30  # _ground_refl.py:33
PHI_MODES = ("rho_v", "image", "normal", "blend")  # _ground_refl.py:242


def eps_tilde(ground_eps, omega, eps0):
    """_ground_refl.py:47-62: complex eps, or (eps_r, sigma S/m) folded as
    eps_r - j sigma / (omega eps0)."""
    if isinstance(ground_eps, (tuple, list)):
        eps_r, sigma = ground_eps
        return complex(float(eps_r), -float(sigma) / (omega * eps0))
    return complex(ground_eps)


def fresnel_rho(eps_t, cos_th):
    """_ground_refl.py:65-79: (rho_v, rho_h), principal-branch sqrt."""
    cos_th = np.asarray(cos_th, dtype=float)
    root = np.sqrt(eps_t - (1.0 - cos_th * cos_th) + 0j)
    rho_v = (eps_t * cos_th - root) / (eps_t * cos_th + root)
    rho_h = (cos_th - root) / (cos_th + root)
    return rho_v, rho_h


def specular_tables(centers, tangents, ground_z):
    """_ground_refl.py:82-224: per (observer m, source n) pair, cos of the
    incidence angle of the ray from the image midpoint of n to the midpoint
    of m, the PEC tangent dot t_m . M t_n, and P = (t_m . p)(t_n . p)."""
    c = np.asarray(centers, dtype=float)
    t = np.asarray(tangents, dtype=float)
    dx = c[:, 0][:, None] - c[:, 0][None, :]
    dy = c[:, 1][:, None] - c[:, 1][None, :]
    dz = c[:, 2][:, None] + c[:, 2][None, :] - 2.0 * ground_z
    hyp = np.hypot(dx, dy)
    rmag = np.maximum(np.sqrt(dx * dx + dy * dy + dz * dz), TINY)
    cos_th = dz / rmag
    safe = hyp > TINY
    inv = 1.0 / np.where(safe, hyp, 1.0)
    px = np.where(safe, -dy * inv, 1.0)   # p = z x d_horizontal; x for a vertical ray
    py = np.where(safe, dx * inv, 0.0)
    tm_p = t[:, 0][:, None] * px + t[:, 1][:, None] * py
    tn_p = t[:, 0][None, :] * px + t[:, 1][None, :] * py
    td_img = t @ (t * np.array([1.0, 1.0, -1.0])).T
    return cos_th, td_img, tm_p * tn_p


def phi_coeffs(mode, eps_t):
    """_ground_refl.py:276-289: w_Phi = c0 + c1 rho_v."""
    if mode == "rho_v":
        return 0j, 1 + 0j
    if mode == "image":
        return (eps_t - 1.0) / (eps_t + 1.0), 0j
    if mode == "normal":
        r = np.sqrt(complex(eps_t))
        return (r - 1.0) / (r + 1.0), 0j
    if mode == "blend":
        return 0.5 * (eps_t - 1.0) / (eps_t + 1.0), 0.5 + 0j
    raise ValueError(f"unknown ground_phi_mode {mode!r}; expected one of {PHI_MODES}")


def weights(centers, tangents, ground_z, eps_t, phi_mode):
    """(w_A, w_Phi) per segment pair (_ground_refl.py:227-273)."""
    cos_th, td_img, P = specular_tables(centers, tangents, ground_z)
    rho_v, rho_h = fresnel_rho(eps_t, cos_th)
    w_A = rho_v * (td_img - P) - rho_h * P
    c0, c1 = phi_coeffs(phi_mode, eps_t)
    return w_A, c0 + c1 * rho_v
