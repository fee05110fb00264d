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

// Status codes (MW_OK ...) come from the public header.
