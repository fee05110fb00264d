// Sommerfeld ground on gfx950: the exact lambda-integral engine, the
// interpolated remainder field and the grounded Z fill.
//
// Engine.  For one (rho, h) the six NEC lambda-integrals (theory manual eqs
// 148-153; reference src/momwire/_sommerfeld.py:328-465,
// _accelerators.cpp:5725-5769) are contour integrals evaluated by adaptive
// 24-point Gauss-Legendre bisection plus panel tails.  One WAVEFRONT owns one
// node: its 64 lanes evaluate the integrand at the 24 Gauss points of two
// segments at once (lanes 0-23 and 32-55), the two 32-lane halves reduce with
// xor shuffles, and every lane then holds both segment sums, so the control
// flow of the adaptive recursion is wave-uniform.  The recursion runs on an
// explicit stack distributed over the lanes (entry e lives in lane e's
// registers; push = one lane writes, pop = a broadcast from that lane), so it
// needs neither LDS nor scratch.  Tails evaluate two consecutive panels per
// step and apply the stop test to them in order, which keeps the sequential
// stopping rule exactly.
//
// Bessel/Hankel functions of complex argument (orders 0 and 1) use the
// ascending series (Abramowitz & Stegun 9.1.10-9.1.13) for |x| <= 13 and the
// Hankel asymptotic expansions (A&S 9.2.5-9.2.8) with optimal truncation
// beyond; both errors are ~1e-12 at the switch radius.
//
// Ground fill.  Z -= C2 * Z_image + Q over basis tiles (reference
// bspline.py:3730-3770): per segment pair, the image moments (4x4 Gauss on
// the mirrored source segment, exp(-jkR)/(4 pi R)) and the remainder moments
// (3x3 Gauss of t_m . F . t_n, F interpolated from the tabulated surfaces by
// the 4x4 Lagrange stencil and combined with the eqs 143-147 azimuth
// factors, reference _accelerators.cpp:5906-5963) are contracted into band
// blocks in LDS and gathered into Z exactly as the free-space fill does.
#include <math.h>

#include "mw_common.h"

namespace {

typedef double2 cd;

MW_DEV cd cz(double r, double i) { return make_double2(r, i); }
MW_DEV cd zadd(cd a, cd b) { return make_double2(a.x + b.x, a.y + b.y); }
MW_DEV cd zsub(cd a, cd b) { return make_double2(a.x - b.x, a.y - b.y); }
MW_DEV cd zmul(cd a, cd b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
MW_DEV cd zsc(cd a, double s) { return make_double2(a.x * s, a.y * s); }
MW_DEV cd zdiv(cd a, cd b) { return cdiv(a, b); }
MW_DEV cd zi(cd a) { return make_double2(-a.y, a.x); }     // j a
MW_DEV cd zmi(cd a) { return make_double2(a.y, -a.x); }    // -j a
MW_DEV double znorm(cd a) { return a.x * a.x + a.y * a.y; }
MW_DEV double zabs(cd a) { return hypot(a.x, a.y); }
MW_DEV cd zexp(cd a) {
  double s, c;
  sincos(a.y, &s, &c);
  const double e = exp(a.x);
  return make_double2(e * c, e * s);
}
MW_DEV cd zlog(cd a) { return make_double2(log(hypot(a.x, a.y)), atan2(a.y, a.x)); }
// principal square root (cut on the negative real axis, sign of Im kept)
MW_DEV cd zsqrt(cd a) {
  if (a.x == 0.0 && a.y == 0.0) return make_double2(0.0, a.y);
  const double r = hypot(a.x, a.y);
  if (a.x >= 0.0) {
    const double t = sqrt(0.5 * (r + a.x));
    return make_double2(t, a.y / (2.0 * t));
  }
  const double t = sqrt(0.5 * (r - a.x));
  return make_double2(fabs(a.y) / (2.0 * t), copysign(t, a.y));
}

constexpr double PI = 3.141592653589793;
constexpr double EULER = 0.5772156649015329;
constexpr int ADAPT_DEPTH = 14;
constexpr int MAX_PANELS = 800;
constexpr double SER_TOL2 = 1e-34;     // (1e-17)^2: series/asymptotic stop test
constexpr double SWITCH2 = 169.0;      // series for |x| <= 13

// 24-point Gauss-Legendre on [-1, 1]: the positive half (nodes are symmetric).
__constant__ double GLX[12] = {
    0.06405689286260563, 0.1911188674736163, 0.3150426796961634, 0.4337935076260451,
    0.5454214713888396,  0.6480936519369755, 0.7401241915785544, 0.820001985973903,
    0.886415527004401,   0.9382745520027328, 0.9747285559713095, 0.9951872199970213};
__constant__ double GLW[12] = {
    0.12793819534675221, 0.1258374563468283,  0.12167047292780342, 0.11550566805372561,
    0.1074442701159656,  0.09761865210411406, 0.08619016153195329, 0.07334648141108041,
    0.05929858491543674, 0.04427743881741955, 0.028531388628933743, 0.012341229799987091};

MW_DEV void gl24(int q, double& x, double& w) {
  const int h = q < 12 ? 11 - q : q - 12;
  x = q < 12 ? -GLX[h] : GLX[h];
  w = GLW[h];
}

// ---------------------------------------------------------------------------
// Bessel / Hankel functions of orders 0 and 1
// ---------------------------------------------------------------------------

// Ascending series.  With q = -(x/2)^2 the ladders a_k = q^k/(k!)^2 and
// b_k = q^k/(2 k! (k+1)!) give J0 = sum a_k and J1/x = sum b_k; when
// want_y, the harmonic-number sums give Y0 and Y1 (A&S 9.1.11, 9.1.13):
//   Y0 = (2/pi) [(ln(x/2) + gamma) J0 - sum_{k>=1} H_k a_k]
//   Y1 = (2/pi) (ln(x/2) + gamma) J1 - 2/(pi x) - (x/pi) sum_k (H_k + H_{k+1}) b_k
MW_DEV void bessel_series(cd x, bool want_y, cd& j0, cd& j1x, cd& y0, cd& y1) {
  const cd q = zsc(zmul(x, x), -0.25);
  cd a = cz(1.0, 0.0), b = cz(0.5, 0.0);
  cd sa = a, sb = b;
  cd ha = cz(0.0, 0.0), hb = b;  // sum H_k a_k (k>=1), sum (H_k + H_{k+1}) b_k
  double hk = 0.0;
  for (int k = 1; k <= 60; ++k) {
    a = zsc(zmul(a, q), 1.0 / double(k * k));
    b = zsc(zmul(b, q), 1.0 / double(k * (k + 1)));
    sa = zadd(sa, a);
    sb = zadd(sb, b);
    bool done = znorm(a) <= SER_TOL2 * znorm(sa) && znorm(b) <= SER_TOL2 * znorm(sb);
    if (want_y) {
      hk += 1.0 / double(k);
      const double hk1 = hk + 1.0 / double(k + 1);
      const cd ta = zsc(a, hk), tb = zsc(b, hk + hk1);
      ha = zadd(ha, ta);
      hb = zadd(hb, tb);
      done = done && znorm(ta) <= SER_TOL2 * (znorm(ha) + 1.0) &&
             znorm(tb) <= SER_TOL2 * (znorm(hb) + 1.0);
    }
    if (done) break;
  }
  j0 = sa;
  j1x = sb;
  if (want_y) {
    const cd lg = zadd(zlog(zsc(x, 0.5)), cz(EULER, 0.0));
    y0 = zsc(zsub(zmul(lg, sa), ha), 2.0 / PI);
    const cd j1 = zmul(sb, x);
    y1 = zsub(zsub(zsc(zmul(lg, j1), 2.0 / PI), zsc(zdiv(cz(1.0, 0.0), x), 2.0 / PI)),
              zsc(zmul(x, hb), 1.0 / PI));
  }
}

// Asymptotic sums S_nu(-) = sum_k c_k(nu) (-j/x)^k and S_nu(+) = sum_k c_k(nu) (j/x)^k,
// c_0 = 1, c_{k+1} = c_k (4 nu^2 - (2k+1)^2) / (8 (k+1)), each order truncated
// where its terms stop decreasing.
MW_DEV void hankel_sums(cd x, bool plus, cd& s0m, cd& s1m, cd& s0p, cd& s1p) {
  const cd r = zmi(zdiv(cz(1.0, 0.0), x));  // -j / x
  cd t0 = cz(1.0, 0.0), t1 = cz(1.0, 0.0);
  s0m = s1m = s0p = s1p = cz(1.0, 0.0);
  double last0 = 1e300, last1 = 1e300;
  bool on0 = true, on1 = true;
  for (int k = 0; k < 40 && (on0 || on1); ++k) {
    const double o2 = double(2 * k + 1) * double(2 * k + 1);
    const double den = 1.0 / (8.0 * double(k + 1));
    const double sg = (k & 1) ? 1.0 : -1.0;  // (-1)^(k+1)
    if (on0) {
      t0 = zmul(t0, zsc(r, -o2 * den));
      const double m = znorm(t0);
      if (m >= last0) {
        on0 = false;
      } else {
        s0m = zadd(s0m, t0);
        if (plus) s0p = zadd(s0p, zsc(t0, sg));
        last0 = m;
        if (m <= SER_TOL2 * znorm(s0m)) on0 = false;
      }
    }
    if (on1) {
      t1 = zmul(t1, zsc(r, (4.0 - o2) * den));
      const double m = znorm(t1);
      if (m >= last1) {
        on1 = false;
      } else {
        s1m = zadd(s1m, t1);
        if (plus) s1p = zadd(s1p, zsc(t1, sg));
        last1 = m;
        if (m <= SER_TOL2 * znorm(s1m)) on1 = false;
      }
    }
  }
}

// Bessel form: (J0(x), J1(x)/x).
MW_DEV void bessel_pair(cd x, cd& b0, cd& b1x) {
  if (znorm(x) <= SWITCH2) {
    cd y0, y1;
    bessel_series(x, false, b0, b1x, y0, y1);
    return;
  }
  // J = (H1 + H2)/2 with H2_0 = W E S0-, H2_1 = j W E S1-, H1_0 = (W/E) S0+,
  // H1_1 = -j (W/E) S1+, W = sqrt(2/(pi x)), E = exp(-j (x - pi/4)).
  cd s0m, s1m, s0p, s1p;
  hankel_sums(x, true, s0m, s1m, s0p, s1p);
  const cd w = zsqrt(zdiv(cz(2.0 / PI, 0.0), x));
  const cd e = zexp(zmi(zsub(x, cz(0.25 * PI, 0.0))));
  const cd we = zmul(w, e), wie = zdiv(w, e);
  b0 = zsc(zadd(zmul(wie, s0p), zmul(we, s0m)), 0.5);
  b1x = zdiv(zsc(zi(zsub(zmul(we, s1m), zmul(wie, s1p))), 0.5), x);
}

// Hankel form: (H2_0(x)/2, H2_1(x)/(2x)).
MW_DEV void hankel_pair(cd x, cd& b0, cd& b1x) {
  if (znorm(x) <= SWITCH2) {
    cd j0, j1x, y0, y1;
    bessel_series(x, true, j0, j1x, y0, y1);
    b0 = zsc(zsub(j0, zi(y0)), 0.5);                          // (J0 - j Y0)/2
    b1x = zdiv(zsc(zsub(zmul(j1x, x), zi(y1)), 0.5), x);       // (J1 - j Y1)/(2x)
    return;
  }
  cd s0m, s1m, s0p, s1p;
  hankel_sums(x, false, s0m, s1m, s0p, s1p);
  const cd p = zmul(zsqrt(zdiv(cz(2.0 / PI, 0.0), x)), zexp(zmi(zsub(x, cz(0.25 * PI, 0.0)))));
  b0 = zsc(zmul(p, s0m), 0.5);
  b1x = zdiv(zsc(zi(zmul(p, s1m)), 0.5), x);
}

// ---------------------------------------------------------------------------
// the six integrands (eqs 148-153 with the D1/D2 kernels of eqs 154-155)
// ---------------------------------------------------------------------------
struct Node {
  double rho, h;
  cd k1, k1s, k2s;
  double k2;
  bool bessel;
};

// (lam^2 - k^2)^(1/2) with vertical branch cuts: sqrt(-j(lam-k)) sqrt(j(lam+k))
MW_DEV cd gam(cd lam, cd k) { return zmul(zsqrt(zmi(zsub(lam, k))), zsqrt(zi(zadd(lam, k)))); }

MW_DEV void integrand6(const Node& c, cd lam, cd f[6]) {
  const cd g1 = gam(lam, c.k1), g2 = gam(lam, cz(c.k2, 0.0));
  const cd ks = zadd(c.k1s, c.k2s);
  const cd g2ks = zmul(g2, ks);
  const cd d1 = zsub(zdiv(cz(2.0, 0.0), zadd(g1, g2)), zdiv(zsc(c.k2s, 2.0), g2ks));
  const cd d2 = zsub(zdiv(cz(2.0, 0.0), zadd(zmul(c.k1s, g2), zmul(c.k2s, g1))),
                     zdiv(cz(2.0, 0.0), g2ks));
  const cd e = zexp(zsc(g2, -c.h));
  const cd x = zsc(lam, c.rho);
  cd b0, b1x;
  if (c.bessel)
    bessel_pair(x, b0, b1x);
  else
    hankel_pair(x, b0, b1x);
  const cd l2 = zmul(lam, lam), l3 = zmul(l2, lam);
  const cd cm = zmul(d2, e);
  f[0] = zmul(zmul(cm, zsub(b1x, b0)), l3);
  f[1] = zmul(zmul(zmul(zmul(cm, g2), g2), b0), lam);
  f[2] = zmul(zmul(zmul(cm, g2), zmul(b1x, x)), l2);
  f[3] = zsc(zmul(zmul(cm, b1x), l3), -1.0);
  f[4] = zmul(zmul(cm, b0), lam);
  f[5] = zmul(zmul(zmul(d1, e), b0), lam);
}

// ---------------------------------------------------------------------------
// wave-level quadrature
// ---------------------------------------------------------------------------
struct Six { cd v[6]; };

MW_DEV Six six_zero() {
  Six s;
#pragma unroll
  for (int i = 0; i < 6; ++i) s.v[i] = cz(0.0, 0.0);
  return s;
}
MW_DEV void six_add(Six& a, const Six& b) {
#pragma unroll
  for (int i = 0; i < 6; ++i) a.v[i] = zadd(a.v[i], b.v[i]);
}
MW_DEV double six_absmax(const Six& a) {
  double m = 0.0;
#pragma unroll
  for (int i = 0; i < 6; ++i) m = fmax(m, zabs(a.v[i]));
  return m;
}
MW_DEV double six_absmax_diff(const Six& a, const Six& b) {
  double m = 0.0;
#pragma unroll
  for (int i = 0; i < 6; ++i) m = fmax(m, zabs(zsub(a.v[i], b.v[i])));
  return m;
}

// 24-point Gauss sums over segments [a0, a1] (lanes 0-23) and [b0, b1]
// (lanes 32-55); every lane returns both.
MW_DEV void gauss2(const Node& c, cd a0, cd a1, cd b0, cd b1, Six& A, Six& B) {
  const int lane = threadIdx.x & 63;
  const int q = lane & 31;
  const bool second = lane >= 32;
  const cd z0 = second ? b0 : a0, z1 = second ? b1 : a1;
  double acc[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) acc[i] = 0.0;
  if (q < 24) {
    double gx, gw;
    gl24(q, gx, gw);
    const cd mid = zsc(zadd(z0, z1), 0.5), half = zsc(zsub(z1, z0), 0.5);
    const cd lam = zadd(mid, zsc(half, gx));
    cd f[6];
    integrand6(c, lam, f);
    const cd w = zsc(half, gw);
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      const cd t = zmul(f[i], w);
      acc[2 * i] = t.x;
      acc[2 * i + 1] = t.y;
    }
  }
#pragma unroll
  for (int off = 16; off >= 1; off >>= 1)
#pragma unroll
    for (int i = 0; i < 12; ++i) acc[i] += __shfl_xor(acc[i], off, 64);
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    A.v[i] = cz(__shfl(acc[2 * i], 0, 64), __shfl(acc[2 * i + 1], 0, 64));
    B.v[i] = cz(__shfl(acc[2 * i], 32, 64), __shfl(acc[2 * i + 1], 32, 64));
  }
}

// Adaptive bisection on [z0, z1] (reference _sommerfeld.py:268-288): a
// segment is accepted when max|left + right - whole| <= rtol * max|left + right|
// or its depth is exhausted; otherwise its halves are pushed, left on top.
MW_DEV void adaptive(const Node& c, cd z0, cd z1, double rtol, Six& total,
                     const volatile int32_t* cancel) {
  const int lane = threadIdx.x & 63;
  // stack entry held by lane e: (a, b, depth, whole)
  cd sa = cz(0.0, 0.0), sb = cz(0.0, 0.0);
  int sd = 0;
  Six sw = six_zero();
  Six whole, unused;
  gauss2(c, z0, z1, z0, z1, whole, unused);
  int sp = 0;
  auto push = [&](cd a, cd b, int d, const Six& w) {
    if (lane == sp) { sa = a; sb = b; sd = d; sw = w; }
    ++sp;
  };
  push(z0, z1, ADAPT_DEPTH, whole);
  int guard = 0;
  while (sp > 0) {
    --sp;
    const cd a = cz(__shfl(sa.x, sp, 64), __shfl(sa.y, sp, 64));
    const cd b = cz(__shfl(sb.x, sp, 64), __shfl(sb.y, sp, 64));
    const int d = __shfl(sd, sp, 64);
    Six w;
#pragma unroll
    for (int i = 0; i < 6; ++i) w.v[i] = cz(__shfl(sw.v[i].x, sp, 64), __shfl(sw.v[i].y, sp, 64));
    const cd mid = zsc(zadd(a, b), 0.5);
    Six L, R;
    gauss2(c, a, mid, mid, b, L, R);
    Six both = L;
    six_add(both, R);
    const double err = six_absmax_diff(both, w);
    const double sc = six_absmax(both);
    if (d <= 0 || err <= rtol * fmax(sc, 1e-300)) {
      six_add(total, both);
    } else {
      push(mid, b, d - 1, R);
      push(a, mid, d - 1, L);
    }
    // a full binary tree of depth 14 has 32767 segments; the guard can only
    // trip on a cancelled solve
    if (((++guard) & 255) == 0 && cancel && *cancel) break;
  }
}

// Panel tail from z0 toward dir * infinity (reference _sommerfeld.py:291-325).
MW_DEV void tail(const Node& c, cd z0, cd dir, double panel, double rtol, double ref,
                 double panel0, Six& tot) {
  tot = six_zero();
  int quiet = 0;
  cd z = z0;
  double step = panel0 > 0.0 ? fmin(panel0, panel) : panel;
  for (int i = 0; i < MAX_PANELS; i += 2) {
    const double s1 = step, s2 = fmin(2.0 * s1, panel);
    const cd za = zadd(z, zsc(dir, s1));
    const cd zb = zadd(za, zsc(dir, s2));
    Six A, B;
    gauss2(c, z, za, za, zb, A, B);
    six_add(tot, A);
    z = za;
    step = s2;
    double cm = six_absmax(A);
    if (cm == 0.0 || cm < rtol * fmax(six_absmax(tot), ref)) {
      if (++quiet >= 2) break;
    } else {
      quiet = 0;
    }
    if (i + 1 >= MAX_PANELS) break;
    six_add(tot, B);
    z = zb;
    step = fmin(2.0 * s2, panel);
    cm = six_absmax(B);
    if (cm == 0.0 || cm < rtol * fmax(six_absmax(tot), ref)) {
      if (++quiet >= 2) break;
    } else {
      quiet = 0;
    }
  }
}

// The six integrals at one (rho, h) (reference _sommerfeld.py:328-465).
// form: 0 automatic (Bessel for rho < 2h), 1 Bessel, 2 Hankel.
MW_DEV Six six_integrals(cd eps, double k2, double rho, double h, double rtol, int form,
                         const volatile int32_t* cancel) {
  Six total = six_zero();
  if (eps.x == 1.0 && eps.y == 0.0) return total;
  Node c;
  c.rho = rho;
  c.h = h;
  c.k2 = k2;
  cd k1 = zsc(zsqrt(eps), k2);
  if (k1.y > 0.0) k1.y = -k1.y;
  c.k1 = k1;
  c.k1s = zmul(k1, k1);
  c.k2s = cz(k2 * k2, 0.0);
  const double scale = fmax(rho, h);
  const double panel = 0.2 * PI / scale;
  const double kabs = zabs(k1);
  const double kmax = fmax(kabs, k2);
  const double kcap = 1.2 * k2 + 50.0 / scale;
  const double qtol = fmin(rtol, 1e-11);
  c.bessel = form == 0 ? (rho < 2.0 * h) : (form == 1);
  if (c.bessel) {
    const double p = fmin(rho > 0.0 ? 1.0 / rho : INFINITY, 1.0 / h);
    const cd brk = cz(p, p);
    const cd end = cz(1.3 * fmin(kmax, kcap) + 3.0 * p, p);
    adaptive(c, cz(0.0, 0.0), brk, qtol, total, cancel);
    cd start = brk;
    if (end.x > brk.x) {
      adaptive(c, brk, end, qtol, total, cancel);
      start = end;
    }
    Six t;
    tail(c, start, cz(1.0, 0.0), panel, rtol, six_absmax(total), -1.0, t);
    six_add(total, t);
    return total;
  }
  const double r1 = hypot(rho, h);
  const cd right = cz(h / r1, -rho / r1), left = cz(-h / r1, -rho / r1);
  const cd a = cz(0.0, -0.4 * k2), b = cz(0.6 * k2, 0.2 * k2), cc = cz(1.02 * k2, 0.2 * k2);
  // waypoint d clears the k1 branch point unless its neighbourhood is dead
  const bool dead = k1.x * h - k1.y * rho >= 50.0;
  const double cap_d = (dead || kabs > 200.0 * k2) ? kcap : fmax(kcap, 1.01 * k1.x);
  cd d = (1.01 * k1.x <= cap_d) ? cz(1.01 * k1.x, 0.99 * fmax(k1.y, -cap_d)) : cz(cap_d, 0.0);
  if (d.x < 1.1 * k2) d.x = 1.1 * k2;
  adaptive(c, a, b, qtol, total, cancel);
  adaptive(c, b, cc, qtol, total, cancel);
  adaptive(c, cc, d, qtol, total, cancel);
  const double ref = six_absmax(total);
  Six t;
  tail(c, d, right, panel, rtol, ref, 0.5 * kmax, t);
  six_add(total, t);
  tail(c, a, left, panel, rtol, ref, 0.5 * kmax, t);
#pragma unroll
  for (int i = 0; i < 6; ++i) total.v[i] = zsub(total.v[i], t.v[i]);
  return total;
}

// one wavefront per node
__global__ __launch_bounds__(256) void k_somm_six(double er, double ei, double k2,
                                                  const double* __restrict__ rho,
                                                  const double* __restrict__ h, int64_t n,
                                                  double rtol, int form, double2* __restrict__ out,
                                                  const volatile int32_t* cancel) {
  const int64_t node = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (node >= n) return;
  if (cancel && *cancel) return;
  const Six s = six_integrals(cz(er, ei), k2, rho[node], h[node], rtol, form, cancel);
  const int lane = threadIdx.x & 63;
  if (lane < 6) {
    cd v = s.v[0];
#pragma unroll
    for (int i = 1; i < 6; ++i)
      if (lane == i) v = s.v[i];
    out[node * 6 + lane] = v;
  }
}

// The four surfaces I_rho^V, I_z^V, I_rho^H, I_phi^H at (R1, theta)
// (reference _sommerfeld.py:535-584), R1 = 0 by the limits of eqs 169-172.
__global__ __launch_bounds__(256) void k_somm_surfaces(double er, double ei,
                                                       const double* __restrict__ eps_node,
                                                       double k2, double omega, double mu,
                                                       const double* __restrict__ R1,
                                                       const double* __restrict__ theta,
                                                       int64_t n, double rtol,
                                                       double2* __restrict__ out,
                                                       const volatile int32_t* cancel) {
  const int64_t node = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (node >= n) return;
  if (cancel && *cancel) return;
  const int lane = threadIdx.x & 63;
  const cd eps = eps_node ? cz(eps_node[2 * node], eps_node[2 * node + 1]) : cz(er, ei);
  const double k2s = k2 * k2;
  const cd k1s = zsc(eps, k2s);
  const cd c1 = cz(0.0, -omega * mu / (4.0 * PI * k2s));
  const double r1 = R1[node], th = theta[node];
  cd surf[4];
  if (r1 == 0.0) {
    const cd ks = zadd(k1s, cz(k2s, 0.0));
    const cd c2 = zdiv(zsub(k1s, cz(k2s, 0.0)), ks);
    const cd c3 = zdiv(zsc(zsub(k1s, cz(k2s, 0.0)), k2s), zmul(ks, ks));
    double s, co;
    sincos(th, &s, &co);
    const bool near = fabs(co) < 1e-8;
    const double cs = near ? 1.0 : co;
    const double q1 = near ? 0.0 : (1.0 - s) / cs;
    const double q2 = near ? 0.5 : (1.0 - s) / (cs * cs);
    const cd c13 = zmul(zmul(c1, c3), k1s);
    surf[0] = zsc(c13, q1);
    surf[1] = c13;
    surf[2] = zsc(zmul(c1, zadd(zsub(c2, c3), zsc(c3, q2))), k2s);
    surf[3] = zsc(zmul(c1, zsub(c2, zsc(c3, q2))), -k2s);
  } else {
    double sn, cs;
    sincos(th, &sn, &cs);
    const double rho = fmax(r1 * cs, 0.0), h = fmax(r1 * sn, 0.0);
    const Six v = six_integrals(eps, k2, rho, h, rtol, 0, cancel);
    // v = (V_rr, V_zz, V_rz, V_r1, V, U)
    double ps, pc;
    sincos(k2 * r1, &ps, &pc);
    const cd ph = zmul(c1, cz(r1 * pc, r1 * ps));
    surf[0] = zmul(zmul(ph, k1s), v.v[2]);
    surf[1] = zmul(zmul(ph, k1s), zadd(v.v[1], zsc(v.v[4], k2s)));
    surf[2] = zmul(zsc(ph, k2s), zadd(v.v[0], v.v[5]));
    surf[3] = zsc(zmul(zsc(ph, k2s), zadd(v.v[3], v.v[5])), -1.0);
  }
  if (lane < 4) {
    cd o = surf[0];
#pragma unroll
    for (int i = 1; i < 4; ++i)
      if (lane == i) o = surf[i];
    out[node * 4 + lane] = o;
  }
}

// ---------------------------------------------------------------------------
// interpolation and the projected remainder field
// ---------------------------------------------------------------------------
MW_DEV void lagrange4(double u, double* w) {
  const double u1 = u - 1.0, u2 = u - 2.0, u3 = u - 3.0;
  w[0] = -u1 * u2 * u3 / 6.0;
  w[1] = u * u2 * u3 / 2.0;
  w[2] = -u * u1 * u3 / 2.0;
  w[3] = u * u1 * u2 / 6.0;
}

// Interpolated surfaces (scaled to physical units) at (r1, theta); tables are
// [n_r][n_th][4 surfaces] per region.
MW_DEV void grid_eval(const mw_somm_grid& G, double r1, double theta, cd s[4]) {
  theta = fmin(fmax(theta, 0.0), 0.5 * PI);
  const double rc = fmin(r1, G.r1_max);
  const bool graze = theta <= G.th_split;
  const int reg = rc <= G.r_break ? (graze ? 0 : 1)
                                   : (rc <= G.r_near || G.n_reg == 4 ? (graze ? 2 : 3)
                                                                     : (graze ? 4 : 5));
  const double fr = (rc - G.r0[reg]) / G.dr[reg];
  const double ft = (theta - G.th0[reg]) / G.dth[reg];
  const int nr = G.n_r[reg], nt = G.n_th[reg];
  int i0 = (int)floor(fr) - 1, j0 = (int)floor(ft) - 1;
  i0 = min(max(i0, 0), nr - 4);
  j0 = min(max(j0, 0), nt - 4);
  double wr[4], wt[4];
  lagrange4(fr - i0, wr);
  lagrange4(ft - j0, wt);
  const double2* V = reinterpret_cast<const double2*>(G.vals) + G.off[reg] * 4;
  double ar[4] = {0, 0, 0, 0}, ai[4] = {0, 0, 0, 0};
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const double2* row = V + ((int64_t)(i0 + i) * nt + j0) * 4;
    double rr[4] = {0, 0, 0, 0}, ri[4] = {0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const double2 v = row[j * 4 + t];
        rr[t] += v.x * wt[j];
        ri[t] += v.y * wt[j];
      }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      ar[t] += rr[t] * wr[i];
      ai[t] += ri[t] * wr[i];
    }
  }
#pragma unroll
  for (int t = 0; t < 4; ++t) s[t] = cz(ar[t] * G.value_scale, ai[t] * G.value_scale);
}

struct SrcDir { double ux, uy, th, tz; };

MW_DEV SrcDir src_dir(double tx, double ty, double tz) {
  SrcDir d;
  d.th = hypot(tx, ty);
  const bool ok = d.th > 1e-12;
  d.ux = ok ? tx / d.th : 1.0;
  d.uy = ok ? ty / d.th : 0.0;
  d.tz = tz;
  return d;
}

// t_obs . F(obs, src) . t_src for a unit current moment (eqs 143-147).
MW_DEV cd proj_one(const mw_somm_grid& G, double gz, double k, const double* o,
                   const double* to, const double* s, const SrcDir& sd) {
  const double dx = o[0] - s[0], dy = o[1] - s[1];
  const double rho = hypot(dx, dy);
  const double hh = (o[2] - gz) + (s[2] - gz);
  const double r1 = sqrt(rho * rho + hh * hh);
  cd f[4];
  grid_eval(G, r1, atan2(hh, rho), f);
  double gs, gc;
  sincos(k * r1, &gs, &gc);
  const cd g = cz(gc / r1, -gs / r1);
  const bool safe = rho > 1e-12 * G.r1_max;
  const double inv = safe ? 1.0 / rho : 0.0;
  const double dhx = safe ? dx * inv : sd.ux, dhy = safe ? dy * inv : sd.uy;
  const double cphi = sd.ux * dhx + sd.uy * dhy;
  const double sphi = sd.ux * dhy - sd.uy * dhx;
  const cd e_rho = zmul(g, zadd(zsc(f[0], sd.tz), zsc(f[2], sd.th * cphi)));
  const cd e_phi = zmul(g, zsc(f[3], sd.th * sphi));
  const cd e_z = zmul(g, zsub(zsc(f[1], sd.tz), zsc(f[0], sd.th * cphi)));
  return zadd(zadd(zsc(zsub(zsc(e_rho, dhx), zsc(e_phi, dhy)), to[0]),
                   zsc(zadd(zsc(e_rho, dhy), zsc(e_phi, dhx)), to[1])),
              zsc(e_z, to[2]));
}

__global__ __launch_bounds__(256) void k_somm_eval(mw_somm_grid G, const double* __restrict__ R1,
                                                   const double* __restrict__ th, int64_t n,
                                                   double2* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  cd s[4];
  grid_eval(G, R1[i], th[i], s);
#pragma unroll
  for (int t = 0; t < 4; ++t) out[i * 4 + t] = s[t];
}

// ---------------------------------------------------------------------------
// Ground remainder of E at points in space (reference _field_point.py:145-233,
// 292-331): the eqs 143-147 combination kept as a vector and summed over the
// sources.  A complex moment p enters through cph = p_x dx + p_y dy and
// sph = p_x dy - p_y dx (d = horizontal unit vector from the source's ground
// projection to the point's; (1, 0) when they coincide, where
// I_rho^H = -I_phi^H makes any direction serve).  A workgroup takes GF_PTS
// points and GF_NCB port columns; its threads split each LDS stage of sources
// GF_SPLIT ways (thread: point tid % GF_PTS, sources tid / GF_PTS + GF_SPLIT
// i) and the partial sums are added in a fixed order at the end, so a launch
// has GF_SPLIT times the threads of one thread per point and the result does
// not depend on scheduling.  Layouts as in the
// near-field kernel: mom [n_src][ldm] with column c of component d at
// d * p + c; out [n_pts][ldo] likewise.
// ---------------------------------------------------------------------------
constexpr int GF_NCB = 4, GF_NSRC = 64, GF_PTS = 64, GF_SPLIT = 4;
static_assert(GF_PTS * GF_SPLIT == 256, "one 256-thread workgroup");

struct GfLds {
  double pos[GF_NSRC][3];
  double2 val[GF_NSRC][3 * GF_NCB];
  double2 red[GF_SPLIT][GF_PTS][3];
};

__global__ __launch_bounds__(256) void k_ground_fields(
    mw_somm_grid G, double gz, double k, int n_pts, const double* __restrict__ pts, int n_src,
    const double* __restrict__ mid, const double2* __restrict__ mom, int64_t ldm, int p,
    int accumulate, double2* __restrict__ out, int64_t ldo) {
  __shared__ GfLds L;
  const int tid = threadIdx.x, lane = tid % GF_PTS, w = tid / GF_PTS;
  const int ip = blockIdx.x * GF_PTS + lane;
  const int col0 = blockIdx.y * GF_NCB;
  const bool live = ip < n_pts;
  double px = 0.0, py = 0.0, pz = 0.0;
  if (live) { px = pts[3 * ip]; py = pts[3 * ip + 1]; pz = pts[3 * ip + 2]; }
  cd acc[GF_NCB][3];
#pragma unroll
  for (int c = 0; c < GF_NCB; ++c)
#pragma unroll
    for (int d = 0; d < 3; ++d) acc[c][d] = cz(0.0, 0.0);
  const double tiny = 1e-12 * G.r1_max;
  for (int s0 = 0; s0 < n_src; s0 += GF_NSRC) {
    __syncthreads();
    for (int i = tid; i < GF_NSRC * 3 * GF_NCB; i += 256) {
      const int s = i / (3 * GF_NCB), j = i % (3 * GF_NCB), d = j / GF_NCB, c = col0 + j % GF_NCB;
      double2 v = make_double2(0.0, 0.0);
      if (s0 + s < n_src && c < p) v = mom[(int64_t)(s0 + s) * ldm + (int64_t)d * p + c];
      L.val[s][j] = v;
    }
    for (int i = tid; i < GF_NSRC * 3; i += 256)
      L.pos[i / 3][i % 3] = (s0 + i / 3 < n_src) ? mid[3 * s0 + i] : 0.0;
    __syncthreads();
    if (!live) continue;
    const int ns = min(GF_NSRC, n_src - s0);
    for (int s = w; s < ns; s += GF_SPLIT) {
      const double dx = px - L.pos[s][0], dy = py - L.pos[s][1];
      const double rho = hypot(dx, dy);
      const double hh = (pz - gz) + (L.pos[s][2] - gz);
      const double r1 = sqrt(rho * rho + hh * hh);
      cd f[4];
      grid_eval(G, r1, atan2(hh, rho), f);
      double gs, gc;
      sincos(k * r1, &gs, &gc);
      const cd g = cz(gc / r1, -gs / r1);
      const bool safe = rho > tiny;
      const double inv = safe ? 1.0 / rho : 0.0;
      const double dhx = safe ? dx * inv : 1.0, dhy = safe ? dy * inv : 0.0;
      const cd irv = zmul(g, f[0]), izv = zmul(g, f[1]), irh = zmul(g, f[2]), iph = zmul(g, f[3]);
#pragma unroll
      for (int c = 0; c < GF_NCB; ++c) {
        const double2 mx = L.val[s][c], my = L.val[s][GF_NCB + c], mz = L.val[s][2 * GF_NCB + c];
        const cd cph = cz(mx.x * dhx + my.x * dhy, mx.y * dhx + my.y * dhy);
        const cd sph = cz(mx.x * dhy - my.x * dhx, mx.y * dhy - my.y * dhx);
        const cd pzc = cz(mz.x, mz.y);
        const cd e_rho = zadd(zmul(pzc, irv), zmul(cph, irh));
        const cd e_phi = zmul(sph, iph);
        const cd e_z = zsub(zmul(pzc, izv), zmul(cph, irv));
        acc[c][0] = zadd(acc[c][0], zsub(zsc(e_rho, dhx), zsc(e_phi, dhy)));
        acc[c][1] = zadd(acc[c][1], zadd(zsc(e_rho, dhy), zsc(e_phi, dhx)));
        acc[c][2] = zadd(acc[c][2], e_z);
      }
    }
  }
  // the waves' partial sums, column by column, added in wave order
#pragma unroll
  for (int c = 0; c < GF_NCB; ++c) {
    __syncthreads();
#pragma unroll
    for (int d = 0; d < 3; ++d) L.red[w][lane][d] = acc[c][d];
    __syncthreads();
    if (w == 0 && live && col0 + c < p) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        double2 v = L.red[0][lane][d];
#pragma unroll
        for (int q = 1; q < GF_SPLIT; ++q) {
          v.x += L.red[q][lane][d].x;
          v.y += L.red[q][lane][d].y;
        }
        double2* o = out + (int64_t)ip * ldo + (int64_t)d * p + col0 + c;
        if (accumulate) { v.x += o->x; v.y += o->y; }
        *o = v;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// grounded fill: Z -= C2 * Z_image + Q
// ---------------------------------------------------------------------------
// NQ (image moments): the mesh's n_qp, a template parameter; the remainder
// rule's order (n_qp_sommerfeld, default 3) is a runtime count <= NQS_MAX.
constexpr int NM = 3, NQS_MAX = 8;
constexpr double INV4PI = 0.079577471545947667884441881686257181017229822870228;

template <int NQ>
struct GStage {
  double cen[3];         // segment midpoint (never mirrored)
  double pos[NQ][3];     // NQ-point nodes (image moments; mirrored for sources)
  double wu[NM][NQ];
  double sp[NQS_MAX][3];  // remainder-rule nodes
  double sw[NM][NQS_MAX]; // w_q u_q^p of the remainder rule
  double tan[3];         // (mirrored for sources)
  double a;
  double poly[NM][NM];
  SrcDir dir;            // horizontal/vertical split of the (unmirrored) tangent
  int ekl;               // coaxial label in the image block (extended kernel), -1 off
};

struct GroundArgs {
  double k, wmu, inv_weps, gz;
  double c2r, c2i;
  double t3[NQS_MAX], w3[NQS_MAX];  // remainder rule on [0, 1]
  int nqs;                          // its order
  int remainder;
  // reflection-coefficient ground: eps~ and the charge weight c0 + c1 rho_v
  int refl;
  double er, ei, p0r, p0i, p1r, p1i;
};

template <int NQ>
MW_DEV void stage_ground(const mw_mesh& M, int s, bool source, const GroundArgs& A,
                         GStage<NQ>& S) {
  const double* g = M.seg + (size_t)s * 12;
  const double L = g[9];
  for (int c = 0; c < 3; ++c) S.cen[c] = 0.5 * (g[c] + g[3 + c]);
  for (int q = 0; q < NQ; ++q) {
    const double t = M.gl[q];
    for (int c = 0; c < 3; ++c) S.pos[q][c] = (1.0 - t) * g[c] + t * g[3 + c];
    if (source) S.pos[q][2] = 2.0 * A.gz - S.pos[q][2];
    const double w = M.gl[NQ + q] * L, u = t * L;
    S.wu[0][q] = w; S.wu[1][q] = w * u; S.wu[2][q] = w * u * u;
  }
  for (int q = 0; q < A.nqs; ++q) {
    const double t = A.t3[q];
    for (int c = 0; c < 3; ++c) S.sp[q][c] = g[c] + t * (g[3 + c] - g[c]);
    const double w = A.w3[q] * L, u = t * L;
    S.sw[0][q] = w; S.sw[1][q] = w * u; S.sw[2][q] = w * u * u;
  }
  S.tan[0] = g[6]; S.tan[1] = g[7]; S.tan[2] = source ? -g[8] : g[8];
  S.a = g[10];
  const double* bp = M.band_poly + (size_t)s * 9;
  for (int i = 0; i < 9; ++i) S.poly[i / 3][i % 3] = bp[i];
  S.dir = src_dir(g[6], g[7], g[8]);
  S.ekl = M.ek ? (source ? M.ek_img_src[s] : M.ek_img_obs[s]) : -1;
}

// Fresnel weights of one segment pair for the reflection-coefficient ground
// (reference _ground_refl.py:65-273): the specular ray runs from the
// source's image midpoint to the observer's midpoint; rho_v, rho_h at its
// incidence angle (principal-branch sqrt); p the horizontal unit vector
// normal to the plane of incidence (x for a vertical ray), P = (t_i.p)(t_j.p);
// w_A = rho_v (td - P) - rho_h P, w_Phi = c0 + c1 rho_v.  Symmetric in the
// pair, so the symmetric tile fill stays valid.
template <int NQ>
MW_DEV void refl_weights(const GStage<NQ>& I, const GStage<NQ>& J, const GroundArgs& A, double td,
                         cd& wa, cd& wp) {
  constexpr double TINY = 1e-30;
  const double dx = I.cen[0] - J.cen[0], dy = I.cen[1] - J.cen[1];
  const double dz = I.cen[2] + J.cen[2] - 2.0 * A.gz;
  const double hyp = hypot(dx, dy);
  const double rmag = fmax(sqrt(dx * dx + dy * dy + dz * dz), TINY);
  const double ct = dz / rmag;
  const bool safe = hyp > TINY;
  const double ih = 1.0 / (safe ? hyp : 1.0);
  const double px = safe ? -dy * ih : 1.0, py = safe ? dx * ih : 0.0;
  // x and y of the (mirrored) source tangent are the unmirrored ones
  const double P = (I.tan[0] * px + I.tan[1] * py) * (J.tan[0] * px + J.tan[1] * py);
  const cd e = cz(A.er, A.ei);
  const cd root = zsqrt(cz(A.er - (1.0 - ct * ct), A.ei));
  const cd ec = zsc(e, ct);
  const cd rv = zdiv(zsub(ec, root), zadd(ec, root));
  const cd rh = zdiv(zsub(cz(ct, 0.0), root), zadd(cz(ct, 0.0), root));
  wa = zsub(zsc(rv, td - P), zsc(rh, P));
  wp = zadd(cz(A.p0r, A.p0i), zmul(cz(A.p1r, A.p1i), rv));
}

template <bool SYM, int NQ>
MW_DEV void ground_body(int ti, int tj, const mw_mesh& M, const GroundArgs& A,
                        const mw_somm_grid& G, double2* __restrict__ Z, int64_t ldz) {
  __shared__ GStage<NQ> rs[MW_TSMAX], cs[MW_TSMAX];
  __shared__ double2 Bt[MW_TSMAX * MW_TSMAX * 9];

  const int tid = threadIdx.x;
  const int nr = M.tile_nseg[ti], nc = M.tile_nseg[tj];
  if (tid < nr) stage_ground<NQ>(M, M.tile_seg[ti * MW_TSMAX + tid], false, A, rs[tid]);
  if (tid >= 64 && tid - 64 < nc)
    stage_ground<NQ>(M, M.tile_seg[tj * MW_TSMAX + tid - 64], true, A, cs[tid - 64]);
  constexpr int QH = (NQ + 1) / 2;  // image-moment observer points per half
  __syncthreads();

  const int npair = nr * nc;
  for (int item = tid; item < 2 * ((npair + 127) / 128) * 128; item += 256) {
    const int pair = item >> 1, half = item & 1;
    const bool live = pair < npair;
    const int li = live ? pair / nc : 0, lj = live ? pair % nc : 0;
    const GStage<NQ>& I = rs[li];
    const GStage<NQ>& J = cs[lj];
    const int be0 = half == 0 ? 0 : 2, be1 = half == 0 ? 2 : 3;
    double* bt = reinterpret_cast<double*>(Bt + (li * nc + lj) * 9);
    {
      // image moments: observer points QH*half .. against the mirrored source
      double jr[NM][NM], jm[NM][NM];
#pragma unroll
      for (int p = 0; p < NM; ++p)
#pragma unroll
        for (int P = 0; P < NM; ++P) jr[p][P] = jm[p][P] = 0.0;
      if (live) {
        const double a2 = I.a * I.a;
        // extended kernel on image pairs coaxial with the observer (a vertical
        // wire and its own image); the remainder below stays reduced
        const bool ext = I.ekl >= 0 && I.ekl == J.ekl;
#pragma unroll
        for (int qq = 0; qq < QH; ++qq) {
          const int q = QH * half + qq;
          if (q >= NQ) break;
          const double x = I.pos[q][0], y = I.pos[q][1], z = I.pos[q][2];
#pragma unroll
          for (int r = 0; r < NQ; ++r) {
            const double dx = x - J.pos[r][0], dy = y - J.pos[r][1], dz = z - J.pos[r][2];
            const double R = sqrt(dx * dx + dy * dy + dz * dz + a2);
            double sn, cn;
            sincos(A.k * R, &sn, &cn);
            const double inv = INV4PI / R;
            double gr = cn * inv, gi = -sn * inv;
            if (ext) {
              const double2 f = ek_factor(R, I.a, A.k);
              const double g0 = gr;
              gr = g0 * f.x - gi * f.y;
              gi = g0 * f.y + gi * f.x;
            }
#pragma unroll
            for (int p = 0; p < NM; ++p)
#pragma unroll
              for (int P = 0; P < NM; ++P) {
                const double w = I.wu[p][q] * J.wu[P][r];
                jr[p][P] += w * gr;
                jm[p][P] += w * gi;
              }
          }
        }
      }
#pragma unroll
      for (int p = 0; p < NM; ++p)
#pragma unroll
        for (int P = 0; P < NM; ++P) {
          jr[p][P] += __shfl_xor(jr[p][P], 1);
          jm[p][P] += __shfl_xor(jm[p][P], 1);
        }
      if (live) {
        // t_i . M t_j: the source stage holds the mirrored tangent
        const double td = I.tan[0] * J.tan[0] + I.tan[1] * J.tan[1] + I.tan[2] * J.tan[2];
        cd wa = zsc(cz(A.c2r, A.c2i), td), wp = cz(A.c2r, A.c2i);
        if (A.refl) refl_weights(I, J, A, td, wa, wp);
        for (int be = be0; be < be1; ++be) {
          const double cj0 = J.poly[be][0], cj1 = J.poly[be][1], cj2 = J.poly[be][2];
          double tr[NM], tm[NM];
#pragma unroll
          for (int p = 0; p < NM; ++p) {
            tr[p] = jr[p][0] * cj0 + jr[p][1] * cj1 + jr[p][2] * cj2;
            tm[p] = jm[p][0] * cj0 + jm[p][1] * cj1 + jm[p][2] * cj2;
          }
          const double dj0 = cj1, dj1 = 2.0 * cj2;
          double pr[2], pm[2];
#pragma unroll
          for (int p = 0; p < 2; ++p) {
            pr[p] = jr[p][0] * dj0 + jr[p][1] * dj1;
            pm[p] = jm[p][0] * dj0 + jm[p][1] * dj1;
          }
#pragma unroll
          for (int al = 0; al < NM; ++al) {
            const double ci0 = I.poly[al][0], ci1 = I.poly[al][1], ci2 = I.poly[al][2];
            const double Ar = ci0 * tr[0] + ci1 * tr[1] + ci2 * tr[2];
            const double Am = ci0 * tm[0] + ci1 * tm[1] + ci2 * tm[2];
            const double Pr = ci1 * pr[0] + 2.0 * ci2 * pr[1];
            const double Pm = ci1 * pm[0] + 2.0 * ci2 * pm[1];
            // image entry -(j w mu w_A A + w_Phi Phi / (j w eps))
            const double Wr = wa.x * Ar - wa.y * Am, Wi = wa.x * Am + wa.y * Ar;
            const double Vr = wp.x * Pr - wp.y * Pm, Vi = wp.x * Pm + wp.y * Pr;
            bt[2 * (al * 3 + be)] = A.wmu * Wi - Vi * A.inv_weps;
            bt[2 * (al * 3 + be) + 1] = -A.wmu * Wr + Vr * A.inv_weps;
          }
        }
      }
    }
    if (A.remainder) {
      // remainder moments: node pairs t = qi nqs + rj, the first half (rounded
      // up) on lane half 0, the rest on half 1
      double fr[NM][NM], fm[NM][NM];
#pragma unroll
      for (int p = 0; p < NM; ++p)
#pragma unroll
        for (int P = 0; P < NM; ++P) fr[p][P] = fm[p][P] = 0.0;
      if (live) {
        const int nt = A.nqs * A.nqs, mid = (nt + 1) / 2;
        const int t0 = half ? mid : 0, t1 = half ? nt : mid;
        for (int t = t0; t < t1; ++t) {
          const int qi = t / A.nqs, rj = t % A.nqs;
          const cd v = proj_one(G, A.gz, A.k, I.sp[qi], I.tan, J.sp[rj], J.dir);
#pragma unroll
          for (int p = 0; p < NM; ++p)
#pragma unroll
            for (int P = 0; P < NM; ++P) {
              const double w = I.sw[p][qi] * J.sw[P][rj];
              fr[p][P] += w * v.x;
              fm[p][P] += w * v.y;
            }
        }
      }
#pragma unroll
      for (int p = 0; p < NM; ++p)
#pragma unroll
        for (int P = 0; P < NM; ++P) {
          fr[p][P] += __shfl_xor(fr[p][P], 1);
          fm[p][P] += __shfl_xor(fm[p][P], 1);
        }
      if (live) {
        for (int be = be0; be < be1; ++be) {
          const double cj0 = J.poly[be][0], cj1 = J.poly[be][1], cj2 = J.poly[be][2];
          double qr[NM], qm[NM];
#pragma unroll
          for (int p = 0; p < NM; ++p) {
            qr[p] = fr[p][0] * cj0 + fr[p][1] * cj1 + fr[p][2] * cj2;
            qm[p] = fm[p][0] * cj0 + fm[p][1] * cj1 + fm[p][2] * cj2;
          }
#pragma unroll
          for (int al = 0; al < NM; ++al) {
            const double ci0 = I.poly[al][0], ci1 = I.poly[al][1], ci2 = I.poly[al][2];
            bt[2 * (al * 3 + be)] -= ci0 * qr[0] + ci1 * qr[1] + ci2 * qr[2];
            bt[2 * (al * 3 + be) + 1] -= ci0 * qm[0] + ci1 * qm[1] + ci2 * qm[2];
          }
        }
      }
    }
  }
  __syncthreads();

  const int m0 = M.tile_b0[ti], m1 = M.tile_b0[ti + 1];
  const int n0 = M.tile_b0[tj], n1 = M.tile_b0[tj + 1];
  const int ml = tid / MW_TB, nl = tid % MW_TB;
  const int m = m0 + ml, n = n0 + nl;
  if (m < m1 && n < n1) {
    double zr = 0.0, zm = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int la = M.bw_loc[3 * m + a], sa = M.bw_slot[3 * m + a];
      if (la < 0) continue;
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        const int lb = M.bw_loc[3 * n + b], sb = M.bw_slot[3 * n + b];
        if (lb < 0) continue;
        const double2 v = Bt[(la * nc + lb) * 9 + sa * 3 + sb];
        zr += v.x;
        zm += v.y;
      }
    }
    double2* zp = Z + (int64_t)m * ldz + n;
    *zp = make_double2(zp->x + zr, zp->y + zm);
    if (SYM && ti != tj) {
      double2* zq = Z + (int64_t)n * ldz + m;
      *zq = make_double2(zq->x + zr, zq->y + zm);
    }
  }
}

template <bool SYM, int NQ>
__global__ __launch_bounds__(256) void k_fill_ground(mw_mesh M, GroundArgs A, mw_somm_grid G,
                                                     double2* __restrict__ Z, int64_t ldz,
                                                     const int32_t* __restrict__ pairs) {
  if (SYM)
    ground_body<true, NQ>(pairs[2 * blockIdx.x], pairs[2 * blockIdx.x + 1], M, A, G, Z, ldz);
  else
    ground_body<false, NQ>(blockIdx.y, blockIdx.x, M, A, G, Z, ldz);
}

template <int NQ>
hipError_t launch_ground_nq(const mw_mesh& M, const GroundArgs& A, const mw_somm_grid& g,
                            double2* Z, int64_t ldz, const int32_t* pairs, int32_t n_pairs,
                            bool symmetric, hipStream_t st) {
  if (symmetric) {
    if (n_pairs == 0) return hipSuccess;
    hipLaunchKernelGGL((k_fill_ground<true, NQ>), dim3(n_pairs), dim3(256), 0, st, M, A, g, Z,
                       ldz, pairs);
  } else {
    hipLaunchKernelGGL((k_fill_ground<false, NQ>), dim3(M.n_tile, M.n_tile), dim3(256), 0, st, M,
                       A, g, Z, ldz, pairs);
  }
  return hipGetLastError();
}

}  // namespace

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
hipError_t mw_launch_somm_six(double er, double ei, double k2, const double* rho, const double* h,
                              int64_t n, double rtol, int form, double2* out,
                              const volatile int32_t* cancel, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_somm_six, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, st, er, ei, k2, rho,
                     h, n, rtol, form, out, cancel);
  return hipGetLastError();
}

hipError_t mw_launch_somm_surfaces(double er, double ei, double k2, double omega, double mu,
                                   const double* R1, const double* th, int64_t n, double rtol,
                                   double2* out, const volatile int32_t* cancel, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_somm_surfaces, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, st, er, ei,
                     (const double*)nullptr, k2, omega, mu, R1, th, n, rtol, out, cancel);
  return hipGetLastError();
}

hipError_t mw_launch_somm_surfaces_varied(const double* eps, double k2, double omega, double mu,
                                          const double* R1, const double* th, int64_t n, double rtol,
                                          double2* out, const volatile int32_t* cancel,
                                          hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_somm_surfaces, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, st, 0.0, 0.0, eps,
                     k2, omega, mu, R1, th, n, rtol, out, cancel);
  return hipGetLastError();
}

hipError_t mw_launch_somm_eval(const mw_somm_grid& G, const double* R1, const double* th, int64_t n,
                               double2* out, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_somm_eval, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, G, R1, th,
                     n, out);
  return hipGetLastError();
}

hipError_t mw_launch_ground_fields(const mw_somm_grid& G, double gz, double k, int n_pts,
                                   const double* pts, int n_src, const double* mid,
                                   const double2* mom, int64_t ldm, int p, int accumulate,
                                   double2* out, int64_t ldo, hipStream_t st) {
  if (n_pts <= 0 || p <= 0) return hipSuccess;
  dim3 grid((unsigned)((n_pts + GF_PTS - 1) / GF_PTS), (unsigned)((p + GF_NCB - 1) / GF_NCB));
  hipLaunchKernelGGL(k_ground_fields, grid, dim3(GF_PTS * GF_SPLIT), 0, st, G, gz, k, n_pts, pts,
                     n_src, mid,
                     mom, ldm, p, accumulate, out, ldo);
  return hipGetLastError();
}

// n-point Gauss-Legendre on [0, 1], n <= NQS_MAX: Newton on P_n from the
// Tricomi starting values, P_n and P_n' by the three-term recurrence.
static void gauss_legendre01(int n, double* t, double* w) {
  for (int i = 0; i < n; ++i) {
    double x = cos(3.14159265358979323846 * (i + 0.75) / (n + 0.5));
    double dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = x;
      for (int j = 2; j <= n; ++j) {
        const double p2 = ((2 * j - 1) * x * p1 - (j - 1) * p0) / j;
        p0 = p1;
        p1 = p2;
      }
      if (n == 1) { p0 = 1.0; p1 = x; }
      dp = n * (x * p1 - p0) / (x * x - 1.0);
      const double dx = p1 / dp;
      x -= dx;
      if (fabs(dx) < 1e-16) break;
    }
    // x descends with i; t ascends
    t[i] = 0.5 * (1.0 - x);
    w[i] = 1.0 / ((1.0 - x * x) * dp * dp);
  }
}

hipError_t mw_launch_fill_ground(const mw_mesh& M, double k, double omega, double eps, double mu,
                                 double gz, double c2r, double c2i, const mw_somm_grid* G,
                                 const double* refl, int nqs, double2* Z, int64_t ldz,
                                 const int32_t* pairs, int32_t n_pairs, bool symmetric,
                                 hipStream_t st) {
  if (nqs < 1 || nqs > NQS_MAX) return hipErrorInvalidValue;
  GroundArgs A;
  A.refl = refl != nullptr;
  A.er = A.ei = A.p0r = A.p0i = A.p1r = A.p1i = 0.0;
  if (refl) {
    A.er = refl[0]; A.ei = refl[1];
    A.p0r = refl[2]; A.p0i = refl[3];
    A.p1r = refl[4]; A.p1i = refl[5];
  }
  A.k = k;
  A.wmu = omega * mu;
  A.inv_weps = 1.0 / (omega * eps);
  A.gz = gz;
  A.c2r = c2r;
  A.c2i = c2i;
  A.nqs = nqs;
  for (int q = nqs; q < NQS_MAX; ++q) A.t3[q] = A.w3[q] = 0.0;
  gauss_legendre01(nqs, A.t3, A.w3);
  A.remainder = G != nullptr;
  mw_somm_grid g{};
  if (G) g = *G;
  switch (M.n_qp) {  // the mesh check holds n_qp to 1 .. 8
    case 1: return launch_ground_nq<1>(M, A, g, Z, ldz, pairs, n_pairs, symmetric, st);
    case 2: return launch_ground_nq<2>(M, A, g, Z, ldz, pairs, n_pairs, symmetric, st);
    case 3: return launch_ground_nq<3>(M, A, g, Z, ldz, pairs, n_pairs, symmetric, st);
    case 4: return launch_ground_nq<4>(M, A, g, Z, ldz, pairs, n_pairs, symmetric, st);
    case 5: return launch_ground_nq<5>(M, A, g, Z, ldz, pairs, n_pairs, symmetric, st);
    case 6: return launch_ground_nq<6>(M, A, g, Z, ldz, pairs, n_pairs, symmetric, st);
    case 7: return launch_ground_nq<7>(M, A, g, Z, ldz, pairs, n_pairs, symmetric, st);
    case 8: return launch_ground_nq<8>(M, A, g, Z, ldz, pairs, n_pairs, symmetric, st);
    default: return hipErrorInvalidValue;
  }
}
