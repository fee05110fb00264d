// Shared device helpers for the MI355X (gfx950) B-spline port-solution path.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/momwire_hip.h"

#define MW_DEV __device__ __forceinline__

// Interleaved complex scalars: (re, im) pairs, 16 B (double) / 8 B (float).
template <typename T> struct cplx_t;
template <> struct cplx_t<double> { typedef double2 type; };
template <> struct cplx_t<float> { typedef float2 type; };

template <typename C> MW_DEV C cmk(double re, double im) { C z; z.x = re; z.y = im; return z; }
template <typename C> MW_DEV C cadd(C a, C b) { C z; z.x = a.x + b.x; z.y = a.y + b.y; return z; }
template <typename C> MW_DEV C csub(C a, C b) { C z; z.x = a.x - b.x; z.y = a.y - b.y; return z; }
template <typename C> MW_DEV C cmul(C a, C b) {
  C z; z.x = a.x * b.x - a.y * b.y; z.y = a.x * b.y + a.y * b.x; return z;
}
// a - b*c, as two fused multiply-adds per component
template <typename C> MW_DEV C cfms(C a, C b, C c) {
  C z;
  z.x = fma(-b.x, c.x, fma(b.y, c.y, a.x));
  z.y = fma(-b.x, c.y, fma(-b.y, c.x, a.y));
  return z;
}
// a + b*c, as two fused multiply-adds per component
template <typename C> MW_DEV C cfma(C a, C b, C c) {
  C z;
  z.x = fma(b.x, c.x, fma(-b.y, c.y, a.x));
  z.y = fma(b.x, c.y, fma(b.y, c.x, a.y));
  return z;
}
template <typename C> MW_DEV C cdiv(C a, C b) {
  // Smith's algorithm: no overflow for well-scaled inputs.
  typedef decltype(a.x) T;
  C z;
  if (fabs(b.x) >= fabs(b.y)) {
    T r = b.y / b.x, den = b.x + b.y * r;
    z.x = (a.x + a.y * r) / den; z.y = (a.y - a.x * r) / den;
  } else {
    T r = b.x / b.y, den = b.y + b.x * r;
    z.x = (a.x * r + a.y) / den; z.y = (a.y * r - a.x) / den;
  }
  return z;
}
template <typename C> MW_DEV C crecip(C b) { C one; one.x = 1; one.y = 0; return cdiv(one, b); }
// |re| + |im|: the pivot magnitude LAPACK's izamax uses.
template <typename C> MW_DEV double cabs1(C a) { return fabs((double)a.x) + fabs((double)a.y); }

// The value the neighbouring lane (lane ^ 1) holds: two DPP quad-permute
// moves (quad_perm [1,0,3,2]) on the halves, plain VALU work where a
// generic shuffle goes through the LDS crossbar.
MW_DEV double lane_swap1(double x) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)b, 0xB1, 0xF, 0xF, false);
  const unsigned hi =
      (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), 0xB1, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// A frequency group of a batched sweep: problems group_prob[first ..
// first + nk) share one mesh, tile list and symmetry, and their wavenumbers
// are k0 + f dk in that order (to a few ulp).
struct mw_kgroup {
  int32_t mesh, nk, first, pad;
  double k0, dk;
};

// Extended thin-wire kernel factor of a coaxial equal-radius pair (NEC's
// Eq 89 with the observer on its own wire's surface: rho = b = a):
//   1 + T1 C2 - T2 C1,  T1 = a^4/(4R^4),  T2 = a^2/(2R^2),  C1 = 1 + jkR,
//   C2 = 3 C1 - (kR)^2,
// multiplying exp(-jkR)/(4 pi R) at the reduced kernel's own R
// (reference _bspline_kernels.py:198-222).  Exactly 1 at a = 0.
MW_DEV double2 ek_factor(double R, double a, double k) {
  const double r2 = R * R, kr = k * R, a2 = a * a;
  const double t1 = 0.25 * a2 * a2 / (r2 * r2), t2 = 0.5 * a2 / r2;
  return make_double2(1.0 + t1 * (3.0 - kr * kr) - t2, kr * (3.0 * t1 - t2));
}
// ek_factor - 1, without forming the 1
MW_DEV double2 ek_factor_minus1(double R, double a, double k) {
  const double r2 = R * R, kr = k * R, a2 = a * a;
  const double t1 = 0.25 * a2 * a2 / (r2 * r2), t2 = 0.5 * a2 / r2;
  return make_double2(t1 * (3.0 - kr * kr) - t2, kr * (3.0 * t1 - t2));
}

// Status codes (MW_OK ...) come from the public header.

// R = sqrt(s) and 1 / R together: the hardware reciprocal square root and two
// Newton steps y <- y (3/2 - (s/2) y^2) (s > 0: a^2 is part of s).
MW_DEV void r_and_rinv(double s, double& R, double& rinv) {
  const double h = 0.5 * s;
  double y = __builtin_amdgcn_rsq(s);
  y = y * fma(-h * y, y, 1.5);
  y = y * fma(-h * y, y, 1.5);
  R = s * y;
  rinv = y;
}

// cos x and sin x for |x| <= 0.6 from their Taylor series through x^14 and
// x^15 (the first omitted terms are below 1.3e-17 there).
MW_DEV void cos_sin_small(double x, double& c, double& sn) {
  const double x2 = x * x;
  double pc = -1.0 / 87178291200.0;  // -1/14!
  pc = fma(pc, x2, 1.0 / 479001600.0);
  pc = fma(pc, x2, -1.0 / 3628800.0);
  pc = fma(pc, x2, 1.0 / 40320.0);
  pc = fma(pc, x2, -1.0 / 720.0);
  pc = fma(pc, x2, 1.0 / 24.0);
  pc = fma(pc, x2, -0.5);
  c = fma(pc, x2, 1.0);
  double ps = -1.0 / 1307674368000.0;  // -1/15!
  ps = fma(ps, x2, 1.0 / 6227020800.0);
  ps = fma(ps, x2, -1.0 / 39916800.0);
  ps = fma(ps, x2, 1.0 / 362880.0);
  ps = fma(ps, x2, -1.0 / 5040.0);
  ps = fma(ps, x2, 1.0 / 120.0);
  ps = fma(ps, x2, -1.0 / 6.0);
  sn = x * fma(ps, x2, 1.0);
}

// Largest |k (R - Rc)| that cos_sin_small takes, Rc the distance between the
// segments' midpoints: |R - Rc| <= (L_i + L_j) / 2 for points on the segments.
constexpr double KDR_MAX = 0.6;
