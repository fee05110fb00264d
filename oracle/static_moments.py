"""Same-edge static moments J_pq^static on a uniform straight edge (oracle).

Restates what reference src/momwire/_bspline_kernels.py:333-422 computes
(``_seg_seg_static_moments``, uniform-h Toeplitz path, the 1/(4 pi) prefactor
folded in; C++ twin _accelerators.cpp:2822-2951):

    J_pq[i, j] = 1/(4 pi) * int_0^h du int_0^h dv  u^p v^q / sqrt((u - v - c)^2 + a^2),
    c = (j - i) h,

which depends on (i, j) only through j - i (Toeplitz, 2N-1 values per moment).

The reference evaluates sympy-expanded closed forms.  This restatement uses its
own reduction: with y = u - v the double integral becomes a single integral
of a piecewise polynomial weight w(y) (degree p+q+1 on [-h, 0] and on [0, h])
against K(y - c) = 1/sqrt((y - c)^2 + a^2).

* |j - i| <= 1: exact antiderivatives F_n(x) = int x^n K(x) dx via
  F_0 = asinh(x/a), F_1 = sqrt(x^2+a^2), n F_n = x^{n-1} F_1 - (n-1) a^2 F_{n-2},
  which carries the log singularity of the self and touching terms exactly;
* |j - i| >= 2: the kernel is analytic over each piece (its poles sit at
  least one segment away), so a 16-point Gauss-Legendre rule per piece is
  exact to rounding and avoids the cancellation the expanded closed forms
  suffer on far pairs.

Pinned by tests against the reference's own C++ closed forms (oracle/_ref) and
against mpmath quadrature.

Extended thin-wire kernel (``ek=True``; reference _bspline_kernels.py:28-80,
333-422, its closed forms ``D_ek_moment``): the static kernel gains
-a^2/(2 R^3) + 3 a^4/(4 R^5) = a^2 (a^2 - 2 x^2) / (4 R^5), whose x^n
moments have the antiderivative (a^2 / 4) (x^{n+1} / R^3 - n G_n) with
G_n = int x^n / R^3 dx (G_1 = -1/R, G_0 = x/(a^2 R),
G_n = F_{n-2} - a^2 G_{n-2}): one integrand, so the two O(1) halves of the
correction, which cancel to O((a/h)^2), never meet as separate terms.
"""

from math import comb

import numpy as np
from numpy.polynomial import polynomial as P

_GL_FAR = 16


def weight_polys(p, q, h):
    """Ascending-power coefficients of w(y) on [-h, 0] and on [0, h].

    w(y) = int int delta(y - (u - v)) u^p v^q du dv over [0, h]^2.
    """
    plus = np.zeros(1)
    for k in range(p + 1):
        # C(p,k) y^{p-k} (h - y)^{k+q+1} / (k+q+1)
        term = P.polymul(P.polypow([0.0, 1.0], p - k), P.polypow([h, -1.0], k + q + 1))
        plus = P.polyadd(plus, comb(p, k) * term / (k + q + 1))
    minus = np.zeros(1)
    for k in range(q + 1):
        # C(q,k) (-y)^{q-k} (h + y)^{p+k+1} / (p+k+1)
        term = P.polymul(P.polypow([0.0, -1.0], q - k), P.polypow([h, 1.0], p + k + 1))
        minus = P.polyadd(minus, comb(q, k) * term / (p + k + 1))
    return minus, plus


def _taylor_shift(coef, c):
    """Coefficients of P(x + c) in powers of x."""
    n = len(coef)
    out = np.zeros(n)
    for m in range(n):
        out[m] = sum(coef[k] * comb(k, m) * c ** (k - m) for k in range(m, n))
    return out


def _antiderivs(x, a, nmax):
    F = [np.arcsinh(x / a), np.sqrt(x * x + a * a)]
    for n in range(2, nmax + 1):
        F.append((x ** (n - 1) * F[1] - (n - 1) * a * a * F[n - 2]) / n)
    return F


def _ek_antiderivs(x, a, nmax):
    """E_n(x) = int x^n a^2 (a^2 - 2 x^2) / (4 R^5) dx, n = 0..nmax."""
    F = _antiderivs(x, a, nmax)
    R = np.sqrt(x * x + a * a)
    G = [x / (a * a * R), -1.0 / R]
    for n in range(2, nmax + 1):
        G.append(F[n - 2] - a * a * G[n - 2])
    return [0.25 * a * a * (x ** (n + 1) / R ** 3 - n * G[n]) for n in range(nmax + 1)]


def _piece_closed(coef, c, y0, y1, a, ek=False):
    b = _taylor_shift(coef, c)
    x0, x1 = y0 - c, y1 - c
    F0 = _antiderivs(np.float64(x0), a, len(b) - 1)
    F1 = _antiderivs(np.float64(x1), a, len(b) - 1)
    out = float(sum(b[n] * (F1[n] - F0[n]) for n in range(len(b))))
    if ek:
        E0 = _ek_antiderivs(np.float64(x0), a, len(b) - 1)
        E1 = _ek_antiderivs(np.float64(x1), a, len(b) - 1)
        out += float(sum(b[n] * (E1[n] - E0[n]) for n in range(len(b))))
    return out


def _piece_gl(coef, c, y0, y1, a, ek=False):
    xi, wi = np.polynomial.legendre.leggauss(_GL_FAR)
    y = 0.5 * (y0 + y1) + 0.5 * (y1 - y0) * xi
    w = 0.5 * (y1 - y0) * wi
    x = y - c
    R = np.sqrt(x * x + a * a)
    K = 1.0 / R
    if ek:
        K = K + a * a * (a * a - 2.0 * x * x) / (4.0 * R ** 5)
    return float(np.sum(w * P.polyval(y, coef) * K))


def static_moment(p, q, delta, h, a, ek=False):
    """J_pq^static for source offset c = delta*h (no 1/(4 pi) factor)."""
    minus, plus = weight_polys(p, q, h)
    c = delta * h
    piece = _piece_closed if abs(delta) <= 1 else _piece_gl
    return piece(minus, c, -h, 0.0, a, ek) + piece(plus, c, 0.0, h, a, ek)


def static_toeplitz_values(h, a, N, max_d, ek=False):
    """(max_d+1, max_d+1, 2N-1) values; index delta + N - 1, delta = j - i."""
    n_d = max_d + 1
    out = np.empty((n_d, n_d, 2 * N - 1))
    for p in range(n_d):
        for q in range(n_d):
            for t, delta in enumerate(range(-(N - 1), N)):
                out[p, q, t] = static_moment(p, q, delta, h, a, ek)
    return out / (4.0 * np.pi)


def static_moments_uniform(h, a, N, max_d, ek=False):
    """(max_d+1, max_d+1, N, N) static moments of one uniform edge."""
    vals = static_toeplitz_values(h, a, N, max_d, ek)
    gather = (np.arange(N)[None, :] - np.arange(N)[:, None]) + (N - 1)
    return vals[:, :, gather]
