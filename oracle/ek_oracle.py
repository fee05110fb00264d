"""CPU oracle for the extended thin-wire kernel -- TEST INFRASTRUCTURE ONLY.

A numpy restatement of the reference's extended kernel for the B-spline
solver (src/momwire/_bspline_kernels.py:28-80, 198-322, 464-477): on segment
pairs that are coaxial and of equal radius the Green's function is

    G_ek = exp(-jkR) / (4 pi R) * fac(R),
    fac  = 1 + T1 C2 - T2 C1,  T1 = a^4/(4 R^4),  T2 = a^2/(2 R^2),
    C1   = 1 + jkR,  C2 = 3 C1 - (kR)^2,

with R = sqrt(zeta^2 + a^2) the reduced kernel's own distance.  Same-edge
blocks split it into the static part (static_moments.py, ``ek=True``) and the
smooth remainder [(exp(-jkR) - 1) fac + extra] / (4 pi R), extra = fac -
fac_static.  Pinned by the reference's labels, static moments, Z and Y
(tests/golden/ek_*.npz) and by the a -> 0 limit, where every factor is 1.
"""

import numpy as np


def factor(R, a, k):
    """_bspline_kernels.py:198-222."""
    r2 = R * R
    kr = k * R
    c1 = 1.0 + 1j * kr
    c2 = 3.0 * c1 - kr * kr
    t1 = 0.25 * a ** 4 / (r2 * r2)
    t2 = 0.5 * a * a / r2
    return t1 * c2 - t2 * c1 + 1.0


def reg_extra(R, a, k):
    """fac - fac_static, term by term (_bspline_kernels.py:225-246)."""
    r2 = R * R
    kr = k * R
    t1 = 0.25 * a ** 4 / (r2 * r2)
    t2 = 0.5 * a * a / r2
    return t1 * (3j * kr - kr * kr) - t2 * (1j * kr)


def reg_kernel(R, a, k):
    """[(exp(-jkR) - 1) fac + extra] / (4 pi R) (_bspline_kernels.py:464-477)."""
    return ((np.exp(-1j * k * R) - 1.0) * factor(R, a, k) + reg_extra(R, a, k)) / (4 * np.pi * R)


def axis_groups(seg_l, seg_r, tangents, seg_a, tol=1e-6):
    """Coaxial-and-equal-radius labels (_bspline_kernels.py:249-322): a
    segment joins the first earlier representative with |t.t_rep| >= 1 - tol,
    |a/a_rep - 1| <= tol and centre offset perpendicular to its own axis
    <= tol * a_rep; otherwise it starts a new label."""
    centers = 0.5 * (np.asarray(seg_l, float) + np.asarray(seg_r, float))
    t = np.asarray(tangents, float)
    a = np.asarray(seg_a, float)
    labels = np.full(len(a), -1, dtype=np.int64)
    reps = []
    for i in range(len(a)):
        for g, r in enumerate(reps):
            if abs(t[r] @ t[i]) < 1.0 - tol or abs(a[i] / a[r] - 1.0) > tol:
                continue
            d = centers[i] - centers[r]
            perp = d - (d @ t[i]) * t[i]
            if np.linalg.norm(perp) <= tol * a[r]:
                labels[i] = g
                break
        else:
            labels[i] = len(reps)
            reps.append(i)
    return labels
