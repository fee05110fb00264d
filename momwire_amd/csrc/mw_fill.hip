// Galerkin Z fill for the degree-2 B-spline thin-wire EFIE, free space, on
// gfx950.
//
//   Z[m][n] = j w mu  sum_{a,b} (t_i . t_j) sum_{p,q} C[m,a,p] C[n,b,q] J_pq(i,j)
//           + 1/(j w eps) sum_{a,b} sum_{p,q>=1} p q C[m,a,p] C[n,b,q] J_{p-1,q-1}(i,j)
//   (reference _accelerators.cpp:1196-1309 / bspline.py:2289-2348), with
//   i = supp(m,a), j = supp(n,b) and the segment-pair moments
//   J_pq(i,j) = int int u^p u'^q exp(-jkR)/(4 pi R),  R^2 = |r_i(u)-r_j(u')|^2 + a_i^2
//   by 4x4 Gauss-Legendre (_accelerators.cpp:349-542) on every pair except
//   pairs on one straight edge, which take static closed form + smooth
//   remainder instead (bspline.py:2254-2281).
//
// Layout of the work: the basis index range is cut into tiles of <= 16 bases;
// a workgroup owns one 16x16 tile of Z.  It stages the <= 18 segments each
// tile side touches, evaluates every segment pair once (two lanes per pair,
// eight quadrature points each, one shuffle to combine), contracts the nine
// moments straight into the 3x3 "band block" B_ij -- the contribution of the
// pair to the (<= 3) bases living on segment i times those on segment j --
// and keeps B in LDS.  Each thread then sums the nine band blocks its Z entry
// needs.  The (d+1)^2 N^2 moment tensor of the reference never exists.
#include "mw_common.h"

namespace {

constexpr int NM = 3;       // moments per axis (degree 2)
constexpr int NQ_MAX = 8;   // Gauss-Legendre points per segment per axis: 1 .. NQ_MAX
static_assert(NQ_MAX == 8, "MW_NQ_SWITCH below instantiates NQ = 1 .. 8");
constexpr double INV4PI = 0.079577471545947667884441881686257181017229822870228;

// ---------------------------------------------------------------------------
// Same-edge table
// ---------------------------------------------------------------------------

// Coefficients (ascending powers of eta) of the unit-length weight
//   w(eta) = int int delta(eta - (U - V)) U^p V^q dU dV over [0,1]^2
// on the piece eta in [0,1] (plus) or [-1,0] (minus); degree <= p+q+1 <= 5.
MW_DEV void weight_poly(int p, int q, bool plus, double* c /*[6]*/) {
  for (int n = 0; n < 6; ++n) c[n] = 0.0;
  const int top = plus ? p : q;
  for (int k = 0; k <= top; ++k) {
    // plus : C(p,k) eta^{p-k} (1-eta)^{k+q+1} / (k+q+1)
    // minus: C(q,k) (-eta)^{q-k} (1+eta)^{p+k+1} / (p+k+1)
    double binom = 1.0;
    for (int t = 0; t < k; ++t) binom = binom * (top - t) / (t + 1);
    const int e1 = top - k;                        // power of (+-eta)
    const int e2 = plus ? (k + q + 1) : (p + k + 1);
    const double s1 = (plus || (e1 % 2 == 0)) ? 1.0 : -1.0;
    const double sg = plus ? -1.0 : 1.0;           // (1 + sg*eta)^{e2}
    double b2 = 1.0;                               // C(e2, t)
    for (int t = 0; t <= e2; ++t) {
      const double term = binom * s1 * b2 * ((t % 2 == 1) ? sg : 1.0) / e2;
      c[e1 + t] += term;
      b2 = b2 * (e2 - t) / (t + 1);
    }
  }
}

// int_{x0}^{x1} P(x) K(x) dx for P of degree <= 5, exactly, with
// K = 1 / R, R = sqrt(x^2 + ah^2), or under the extended kernel (ek)
// K = 1/R - ah^2/(2R^3) + 3ah^4/(4R^5) = 1/R + ah^2 (ah^2 - 2x^2)/(4R^5).
// The correction's x^n moments have the antiderivative
// (ah^2/4) (x^{n+1}/R^3 - n G_n), G_n = int x^n/R^3 (G_0 = x/(ah^2 R),
// G_1 = -1/R, G_n = F_{n-2} - ah^2 G_{n-2}): one integrand, so its two
// O(1) halves, which cancel to O(ah^2), never meet as separate terms.
MW_DEV double closed_piece(const double* b, double x0, double x1, double ah, bool ek) {
  const double a2 = ah * ah;
  double Fa[6], Fb[6];
  Fa[0] = asinh(x0 / ah); Fb[0] = asinh(x1 / ah);
  const double Ra = sqrt(x0 * x0 + a2), Rb = sqrt(x1 * x1 + a2);
  Fa[1] = Ra; Fb[1] = Rb;
  double xa = 1.0, xb = 1.0;  // x^{n-1}
  for (int n = 2; n < 6; ++n) {
    xa *= x0; xb *= x1;
    Fa[n] = (xa * Ra - (n - 1) * a2 * Fa[n - 2]) / n;
    Fb[n] = (xb * Rb - (n - 1) * a2 * Fb[n - 2]) / n;
  }
  double s = 0.0;
  for (int n = 0; n < 6; ++n) s += b[n] * (Fb[n] - Fa[n]);
  if (ek) {
    double Ga[6], Gb[6];
    Ga[0] = x0 / (a2 * Ra); Gb[0] = x1 / (a2 * Rb);
    Ga[1] = -1.0 / Ra; Gb[1] = -1.0 / Rb;
    for (int n = 2; n < 6; ++n) {
      Ga[n] = Fa[n - 2] - a2 * Ga[n - 2];
      Gb[n] = Fb[n - 2] - a2 * Gb[n - 2];
    }
    const double ia3 = 1.0 / (Ra * Ra * Ra), ib3 = 1.0 / (Rb * Rb * Rb);
    double pa = x0, pb = x1;  // x^{n+1}
    double e = 0.0;
    for (int n = 0; n < 6; ++n) {
      e += b[n] * ((pb * ib3 - n * Gb[n]) - (pa * ia3 - n * Ga[n]));
      pa *= x0; pb *= x1;
    }
    s += 0.25 * a2 * e;
  }
  return s;
}

// Unit-length static moment I(p,q,delta; ah): source offset delta segments
// (ek: with the extended kernel's static correction).
MW_DEV double static_unit(int p, int q, int delta, double ah, const double* glx,
                          const double* glw, bool ek) {
  double total = 0.0;
  for (int piece = 0; piece < 2; ++piece) {
    const bool plus = piece == 1;
    double c[6];
    weight_poly(p, q, plus, c);
    const double e0 = plus ? 0.0 : -1.0, e1 = plus ? 1.0 : 0.0;
    if (delta >= -1 && delta <= 1) {
      // shift eta = x + delta: b_m = sum_{n>=m} c_n C(n,m) delta^{n-m}
      double b[6];
      for (int m = 0; m < 6; ++m) {
        double s = 0.0, bin = 1.0;  // C(n,m) for n = m..5
        double dp = 1.0;            // delta^{n-m}
        for (int n = m; n < 6; ++n) {
          s += c[n] * bin * dp;
          bin = bin * (n + 1) / (n + 1 - m);
          dp *= (double)delta;
        }
        b[m] = s;
      }
      total += closed_piece(b, e0 - delta, e1 - delta, ah, ek);
    } else {
      // The kernel's poles are >= one segment away: 16-point Gauss-Legendre
      // is exact to rounding and avoids the far-pair cancellation of the
      // expanded closed forms.
      const double mid = 0.5 * (e0 + e1), half = 0.5 * (e1 - e0);
      double s = 0.0;
      for (int g = 0; g < 16; ++g) {
        const double eta = mid + half * glx[g];
        double w = c[5];
        for (int n = 4; n >= 0; --n) w = w * eta + c[n];
        const double x = eta - delta;
        const double r2 = x * x + ah * ah, ir = 1.0 / sqrt(r2);
        double kern = ir;
        if (ek) kern += ah * ah * (ah * ah - 2.0 * x * x) * 0.25 * ir * ir * ir * ir * ir;
        s += glw[g] * w * kern;
      }
      total += half * s;
    }
  }
  return total;
}

// The k-independent part of a same-edge entry: static moment (p, q) of entry e,
// h^{p+q+1} I_unit(p, q, delta; a/h) / (4 pi).
MW_DEV double same_edge_static(int e, int p, int q, const mw_mesh& M, int nq) {
  const int ed = M.se_edge[e];
  const int n_e = M.edge_i[2 * ed], off = M.edge_i[2 * ed + 1];
  const int delta = e - off - (n_e - 1);
  const double h = M.edge[2 * ed], a = M.edge[2 * ed + 1];
  const double* glx = M.gl + 2 * nq;
  const double* glw = glx + 16;
  double hp = h;
  for (int t = 0; t < p + q; ++t) hp *= h;
  return hp * static_unit(p, q, delta, a / h, glx, glw, M.ek != 0) * INV4PI;
}

// One thread per (entry, p, q): the static table, once per mesh.
__global__ __launch_bounds__(256) void k_same_edge_static(mw_mesh M, int nq, double* out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)M.n_se * 9) return;
  const int e = (int)(t / 9), pq = (int)(t % 9);
  out[t] = same_edge_static(e, pq / 3, pq % 3, M, nq);
}

// Same-edge entry e at wavenumber k: the static part (from M.se_static when
// the mesh carries it, else computed here) plus the smooth remainder
// (exp(-jkR) - 1)/(4 pi R) on the arc distance.  A group of LG lanes owns an
// entry: lane j takes Gauss point pair j (q = j / NQ, r = j % NQ), the
// group's nine complex partial moments are summed by shuffles, and lane
// j < 9 writes moment (j / 3, j % 3).
template <int NQ>
struct SeLanes {
  static constexpr int LG = NQ * NQ <= 16 ? 16 : 64;
};

template <int NQ>
MW_DEV void same_edge_body(int64_t t, const mw_mesh& M, double k) {
  constexpr int LG = SeLanes<NQ>::LG;
  const int e = (int)(t / LG), j = (int)(t % LG);
  const bool live = e < M.n_se;
  const int ec = live ? e : 0;
  const int ed = M.se_edge[ec];
  const int n_e = M.edge_i[2 * ed], off = M.edge_i[2 * ed + 1];
  const int delta = ec - off - (n_e - 1);
  const double h = M.edge[2 * ed], a = M.edge[2 * ed + 1];
  const double* t01 = M.gl;
  const double* w01 = M.gl + NQ;
  double jr[NM * NM], ji[NM * NM];
#pragma unroll
  for (int i = 0; i < NM * NM; ++i) jr[i] = ji[i] = 0.0;
  if (live && j < NQ * NQ) {
    const int q = j / NQ, r = j % NQ;
    double wq[NM], wr[NM];
    const double uq = t01[q] * h, ur = t01[r] * h;
    wq[0] = w01[q] * h; wq[1] = wq[0] * uq; wq[2] = wq[1] * uq;
    wr[0] = w01[r] * h; wr[1] = wr[0] * ur; wr[2] = wr[1] * ur;
    const double ds = (t01[q] - t01[r] - (double)delta) * h;
    const double R = sqrt(ds * ds + a * a);
    double sn, cs, sh, chh;
    sincos(k * R, &sn, &cs);
    sincos(0.5 * k * R, &sh, &chh);
    const double inv = INV4PI / R;
    double gr = -2.0 * sh * sh * inv;  // cos(kR) - 1
    double gi = -sn * inv;
    if (M.ek) {
      // extended kernel: [(exp(-jkR) - 1) fac + extra] / (4 pi R),
      // extra = fac - fac_static = T1 (3jkR - (kR)^2) - T2 jkR, term by term
      // (reference _bspline_kernels.py:225-246, 464-477)
      const double2 f = ek_factor(R, a, k);
      const double r2 = R * R, kr = k * R;
      const double t1 = 0.25 * a * a * a * a / (r2 * r2), t2 = 0.5 * a * a / r2;
      const double er = -t1 * kr * kr, ei = kr * (3.0 * t1 - t2);
      const double g0 = gr, g1 = gi;
      gr = g0 * f.x - g1 * f.y + er * inv;
      gi = g0 * f.y + g1 * f.x + ei * inv;
    }
#pragma unroll
    for (int p = 0; p < NM; ++p)
#pragma unroll
      for (int P = 0; P < NM; ++P) {
        const double w = wq[p] * wr[P];
        jr[p * NM + P] = w * gr;
        ji[p * NM + P] = w * gi;
      }
  }
  // sum over the group's lanes (LG is a power of two; groups are aligned)
#pragma unroll
  for (int o = LG / 2; o > 0; o >>= 1)
#pragma unroll
    for (int i = 0; i < NM * NM; ++i) {
      jr[i] += __shfl_xor(jr[i], o, 64);
      ji[i] += __shfl_xor(ji[i], o, 64);
    }
  if (!live || j >= NM * NM) return;
  double rr = 0.0, ri = 0.0;
#pragma unroll
  for (int i = 0; i < NM * NM; ++i)
    if (i == j) { rr = jr[i]; ri = ji[i]; }
  const double st = M.se_static ? M.se_static[(size_t)e * 9 + j]
                                : same_edge_static(e, j / NM, j % NM, M, NQ);
  reinterpret_cast<double2*>(M.se_tab)[(size_t)e * 9 + j] = make_double2(st + rr, ri);
}

template <int NQ>
__global__ __launch_bounds__(256) void k_same_edge(mw_mesh M, double k) {
  same_edge_body<NQ>((int64_t)blockIdx.x * blockDim.x + threadIdx.x, M, k);
}

// ---------------------------------------------------------------------------
// Fused tile fill
// ---------------------------------------------------------------------------

template <int NQ>
struct SegStage {
  double pos[NQ][3];   // quadrature points
  double wu[NM][NQ];   // w_q L (t_q L)^p
  double tan[3];
  double a;            // observer-row radius
  double poly[NM][NM]; // band polynomials [slot][power]
  double mid[3];       // segment midpoint
  double len;          // segment length
  int edge, eidx;
  int ekl;             // coaxial label (extended kernel), -1 when off
};

template <int NQ>
MW_DEV void stage_segment(const mw_mesh& M, int s, SegStage<NQ>& S) {
  const double* g = M.seg + (size_t)s * 12;
  const double L = g[9];
  for (int q = 0; q < NQ; ++q) {
    const double t = M.gl[q];
    for (int c = 0; c < 3; ++c) S.pos[q][c] = (1.0 - t) * g[c] + t * g[3 + c];
    const double w = M.gl[NQ + q] * L, u = t * L;
    S.wu[0][q] = w; S.wu[1][q] = w * u; S.wu[2][q] = w * u * u;
  }
  S.tan[0] = g[6]; S.tan[1] = g[7]; S.tan[2] = g[8];
  S.a = g[10];
  for (int c = 0; c < 3; ++c) S.mid[c] = 0.5 * (g[c] + g[3 + c]);
  S.len = L;
  const double* bp = M.band_poly + (size_t)s * 9;
  for (int i = 0; i < 9; ++i) S.poly[i / 3][i % 3] = bp[i];
  S.edge = M.seg_edge[2 * s];
  S.eidx = M.seg_edge[2 * s + 1];
  S.ekl = M.ek ? M.ek_label[s] : -1;
}

// Contract a segment pair's moments J[p][P] (both halves' sums) to its band
// block B[al][be] = j w mu td (C_i J C_j^T)[al][be] + (C_i' J C_j'^T)[al][be]
// / (j w eps), the derivative polynomials being C'[p'][s] = (p'+1) C[s][p'+1].
// The pair's two lanes share the 9 entries evenly: column be = half (3 rows
// each), then rows 2 half .. 2 half + 1 of column 2.
template <int NQ>
MW_DEV void contract_band(const double (&jr)[NM][NM], const double (&jim)[NM][NM],
                          const SegStage<NQ>& I, const SegStage<NQ>& J, double td, double wmu,
                          double inv_weps, int half, double2* __restrict__ out) {
#pragma unroll
  for (int step = 0; step < 2; ++step) {
    const int be = step == 0 ? half : 2;
    const double cj0 = J.poly[be][0], cj1 = J.poly[be][1], cj2 = J.poly[be][2];
    double tr[NM], tm[NM];
#pragma unroll
    for (int a = 0; a < NM; ++a) {
      tr[a] = jr[a][0] * cj0 + jr[a][1] * cj1 + jr[a][2] * cj2;
      tm[a] = jim[a][0] * cj0 + jim[a][1] * cj1 + jim[a][2] * cj2;
    }
    const double dj0 = cj1, dj1 = 2.0 * cj2;  // derivative coefficients
    double fr[2], fi[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      fr[a] = jr[a][0] * dj0 + jr[a][1] * dj1;
      fi[a] = jim[a][0] * dj0 + jim[a][1] * dj1;
    }
#pragma unroll
    for (int t = 0; t < (step == 0 ? NM : 2); ++t) {
      const int al = step == 0 ? t : 2 * half + t;
      if (al >= NM) break;
      const double ci0 = I.poly[al][0], ci1 = I.poly[al][1], ci2 = I.poly[al][2];
      const double Ar = ci0 * tr[0] + ci1 * tr[1] + ci2 * tr[2];
      const double Ai = ci0 * tm[0] + ci1 * tm[1] + ci2 * tm[2];
      const double di0 = ci1, di1 = 2.0 * ci2;
      const double Pr = di0 * fr[0] + di1 * fr[1];
      const double Pi = di0 * fi[0] + di1 * fi[1];
      // j w mu td A + Phi / (j w eps)
      const double Br = -wmu * td * Ai + Pi * inv_weps;
      const double Bi = wmu * td * Ar - Pr * inv_weps;
      out[al * 3 + be] = make_double2(Br, Bi);
    }
  }
}

// `sym` is a run-time flag (a uniform branch at the store) rather than a
// template parameter: the static LDS arrays below exist once per
// instantiation, so a kernel holding both forms would carry twice the LDS
// and lose half its occupancy.
// The phase of a far pair in FP32: kR (up to ~10^2 rad) is reduced to
// [-pi, pi] in FP64 (Cody-Waite with 2 pi split in two), then sin and cos are
// taken in single precision and 1/(4 pi R) with an FP32 reciprocal.
MW_DEV void far_phase32(double k, double R, double& gr, double& gi) {
  constexpr double INV_2PI = 0.15915494309189533577;
  constexpr double TWO_PI_HI = 6.28318530717958623200;  // 2 pi rounded to double
  constexpr double TWO_PI_LO = 2.44929359829470635445e-16;
  const double kr = k * R;
  const double nn = rint(kr * INV_2PI);
  const double t = fma(-nn, TWO_PI_LO, fma(-nn, TWO_PI_HI, kr));
  float s32, c32;
  sincosf((float)t, &s32, &c32);
  const float inv = (float)INV4PI / (float)R;
  gr = (double)(c32 * inv);
  gi = -(double)(s32 * inv);
}

// FAR32: segment pairs whose midpoints lie more than sqrt(near2) x the longer
// segment apart take the kernel's phase in FP32 (far_phase32); the others,
// and every same-edge, extended-kernel or image term, stay FP64.
template <int NQ, bool FAR32 = false>
MW_DEV void fill_body(bool sym, int ti, int tj, const mw_mesh& M, double k, double wmu,
                      double inv_weps, double2* __restrict__ Z, int64_t ldz, double near2 = 0.0) {
  __shared__ SegStage<NQ> rs[MW_TSMAX], cs[MW_TSMAX];
  __shared__ double2 Bt[MW_TSMAX * MW_TSMAX * 9];

  const int tid = threadIdx.x;
  const int nr = M.tile_nseg[ti], nc = M.tile_nseg[tj];
  if (tid < nr) stage_segment<NQ>(M, M.tile_seg[ti * MW_TSMAX + tid], rs[tid]);
  if (tid >= 64 && tid - 64 < nc)
    stage_segment<NQ>(M, M.tile_seg[tj * MW_TSMAX + tid - 64], cs[tid - 64]);
  __syncthreads();

  const double2* se = reinterpret_cast<const double2*>(M.se_tab);
  const int npair = nr * nc;
  // Two lanes per pair: half 0 takes the first QH observer points, half 1
  // the rest (q = 0,1 and 2,3 at the default four).
  constexpr int QH = (NQ + 1) / 2;
  for (int item = tid; item < 2 * ((npair + 127) / 128) * 128; item += 256) {
    const int pair = item >> 1, half = item & 1;
    const bool live = pair < npair;
    const int li = live ? pair / nc : 0, lj = live ? pair % nc : 0;
    const SegStage<NQ>& I = rs[li];
    const SegStage<NQ>& J = cs[lj];
    double jr[NM][NM], jim[NM][NM];
#pragma unroll
    for (int p = 0; p < NM; ++p)
#pragma unroll
      for (int P = 0; P < NM; ++P) { jr[p][P] = 0.0; jim[p][P] = 0.0; }
    const bool same_edge = I.edge == J.edge;
    if (live && same_edge) {
      if (half == 0) {
        const int ed = I.edge;
        const int n_e = M.edge_i[2 * ed], off = M.edge_i[2 * ed + 1];
        const double2* tab = se + (size_t)(off + (J.eidx - I.eidx) + n_e - 1) * 9;
#pragma unroll
        for (int p = 0; p < NM; ++p)
#pragma unroll
          for (int P = 0; P < NM; ++P) {
            const double2 v = tab[p * NM + P];
            jr[p][P] = v.x; jim[p][P] = v.y;
          }
      }
    } else if (live) {
      const double a2 = I.a * I.a;
      const bool ext = I.ekl >= 0 && I.ekl == J.ekl;  // extended kernel on this pair
      const double mx = I.mid[0] - J.mid[0], my = I.mid[1] - J.mid[1], mz = I.mid[2] - J.mid[2];
      const double md2 = mx * mx + my * my + mz * mz;
      bool far = false;
      if (FAR32 && !ext) {
        const double lm = fmax(I.len, J.len);
        far = md2 > near2 * lm * lm;
      }
      // exp(-jkR) = exp(-jk Rc) exp(-jk (R - Rc)): one full sincos per pair
      // at the midpoints' distance Rc, a short series per point pair
      const bool series = !(FAR32 && far) && k * 0.5 * (I.len + J.len) <= KDR_MAX;
      double Rc = 0.0, sc = 0.0, cc = 1.0;
      if (series) {
        double rci;
        r_and_rinv(md2 + a2, Rc, rci);
        sincos(k * Rc, &sc, &cc);
      }
#pragma unroll
      for (int qq = 0; qq < QH; ++qq) {
        const int q = QH * half + qq;
        if (q >= NQ) break;
        const double x = I.pos[q][0], y = I.pos[q][1], z = I.pos[q][2];
        const double wq0 = I.wu[0][q], wq1 = I.wu[1][q], wq2 = I.wu[2][q];
        double tr[NM] = {0.0, 0.0, 0.0}, ti[NM] = {0.0, 0.0, 0.0};  // sum_r w_P(r) g(q, r)
#pragma unroll
        for (int r = 0; r < NQ; ++r) {
          const double dx = x - J.pos[r][0], dy = y - J.pos[r][1], dz = z - J.pos[r][2];
          double R, rinv;
          r_and_rinv(dx * dx + dy * dy + dz * dz + a2, R, rinv);
          double gr, gi;  // exp(-jkR)/(4 pi R)
          if (FAR32 && far) {
            far_phase32(k, R, gr, gi);
          } else if (series) {
            double cd, sd;
            cos_sin_small(k * (R - Rc), cd, sd);
            const double inv = INV4PI * rinv;
            gr = fma(cc, cd, -sc * sd) * inv;
            gi = -fma(sc, cd, cc * sd) * inv;
          } else {
            double sn, cn;
            sincos(k * R, &sn, &cn);
            const double inv = INV4PI * rinv;
            gr = cn * inv;
            gi = -sn * inv;
          }
          if (ext) {
            const double2 f = ek_factor(R, I.a, k);
            const double g0 = gr;
            gr = g0 * f.x - gi * f.y;
            gi = g0 * f.y + gi * f.x;
          }
#pragma unroll
          for (int P = 0; P < NM; ++P) {
            tr[P] = fma(J.wu[P][r], gr, tr[P]);
            ti[P] = fma(J.wu[P][r], gi, ti[P]);
          }
        }
        // the observer point's weights times its source sums
#pragma unroll
        for (int P = 0; P < NM; ++P) {
          jr[0][P] = fma(wq0, tr[P], jr[0][P]); jim[0][P] = fma(wq0, ti[P], jim[0][P]);
          jr[1][P] = fma(wq1, tr[P], jr[1][P]); jim[1][P] = fma(wq1, ti[P], jim[1][P]);
          jr[2][P] = fma(wq2, tr[P], jr[2][P]); jim[2][P] = fma(wq2, ti[P], jim[2][P]);
        }
      }
    }
    // combine the two halves (partner lane = item ^ 1 = lane ^ 1)
#pragma unroll
    for (int p = 0; p < NM; ++p)
#pragma unroll
      for (int P = 0; P < NM; ++P) {
        jr[p][P] += lane_swap1(jr[p][P]);
        jim[p][P] += lane_swap1(jim[p][P]);
      }
    if (!live) continue;
    const double td = I.tan[0] * J.tan[0] + I.tan[1] * J.tan[1] + I.tan[2] * J.tan[2];
    contract_band<NQ>(jr, jim, I, J, td, wmu, inv_weps, half, Bt + (li * nc + lj) * 9);
  }
  __syncthreads();

  // Gather: Z[m][n] = sum over wing pairs of the band-block entries.
  const int m0 = M.tile_b0[ti], m1 = M.tile_b0[ti + 1];
  const int n0 = M.tile_b0[tj], n1 = M.tile_b0[tj + 1];
  const int ml = tid / MW_TB, nl = tid % MW_TB;
  const int m = m0 + ml, n = n0 + nl;
  if (m < m1 && n < n1) {
    double zr = 0.0, zi = 0.0;
    int la[3], sa[3], lb[3], sb[3];
#pragma unroll
    for (int w = 0; w < 3; ++w) {
      la[w] = M.bw_loc[3 * m + w]; sa[w] = M.bw_slot[3 * m + w];
      lb[w] = M.bw_loc[3 * n + w]; sb[w] = M.bw_slot[3 * n + w];
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (la[a] < 0) continue;
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        if (lb[b] < 0) continue;
        const double2 v = Bt[(la[a] * nc + lb[b]) * 9 + sa[a] * 3 + sb[b]];
        zr += v.x; zi += v.y;
      }
    }
    Z[(int64_t)m * ldz + n] = make_double2(zr, zi);
    if (sym && ti != tj) Z[(int64_t)n * ldz + m] = make_double2(zr, zi);
  }
}

template <bool SYM, int NQ, bool FAR32 = false>
__global__ __launch_bounds__(256) void k_fill(mw_mesh M, double k, double wmu,
                                              double inv_weps, double2* __restrict__ Z,
                                              int64_t ldz, const int32_t* __restrict__ pairs,
                                              double near2 = 0.0) {
  if (SYM)
    fill_body<NQ, FAR32>(true, pairs[2 * blockIdx.x], pairs[2 * blockIdx.x + 1], M, k, wmu,
                         inv_weps, Z, ldz, near2);
  else
    fill_body<NQ, FAR32>(false, blockIdx.y, blockIdx.x, M, k, wmu, inv_weps, Z, ldz, near2);
}

// ---------------------------------------------------------------------------
// Batched fill over many (structure, wavenumber) problems
// ---------------------------------------------------------------------------
// one workgroup per (problem, block of entries) work item
__global__ __launch_bounds__(256) void kb_same_edge(const mw_sweep_prob* P, const mw_mesh* meshes,
                                                    const int32_t* __restrict__ items) {
  const mw_sweep_prob& p = P[items[2 * blockIdx.x]];
  mw_mesh M = meshes[p.mesh];
  M.se_tab = p.se_tab;
  same_edge_body<4>((int64_t)items[2 * blockIdx.x + 1] * blockDim.x + threadIdx.x, M, p.k);
}

// Batched fill swept over wavenumber.  A frequency group is a run of
// problems on one mesh whose wavenumbers step evenly, k_f = k0 + f dk.  One
// workgroup per (group, tile pair) stages the two tiles' segments and each
// segment pair's point-to-point distances once, then walks the group's
// wavenumbers: exp(-j k_f R)/(4 pi R) advances by one complex product with
// exp(-j dk R) per step, where the single-problem fill takes a sincos, a
// square root and a division per point pair and wavenumber.  Pairs under the
// extended kernel multiply in its factor, a polynomial in k R; same-edge
// pairs read each problem's own table.  Two lanes per segment pair (half 0: observer
// points 0-1, half 1: points 2-3), up to MW_TSMAX^2 pairs in one pass; the
// band blocks of consecutive wavenumbers alternate between two LDS buffers,
// so one barrier per wavenumber separates the gather of one from the
// contraction of the next.
constexpr int KF_THREADS = 2 * ((MW_TSMAX * MW_TSMAX + 127) / 128) * 128;
static_assert(KF_THREADS == 768, "two lanes per segment pair of an 18 x 18 tile pair");

// EXT: the groups' meshes carry the extended kernel (the plan launches the
// two forms over separate item lists).
template <bool EXT>
__global__ __launch_bounds__(KF_THREADS) void kb_fill_k(const mw_sweep_prob* __restrict__ P,
                                                        const mw_mesh* __restrict__ meshes,
                                                        const mw_kgroup* __restrict__ groups,
                                                        const int32_t* __restrict__ group_prob,
                                                        const int32_t* __restrict__ item_group,
                                                        const int32_t* __restrict__ item_pair,
                                                        double eps, double mu) {
  constexpr int NQ = 4, QH = 2;
  __shared__ SegStage<NQ> rs[MW_TSMAX], cs[MW_TSMAX];
  __shared__ double2 Bt[2][MW_TSMAX * MW_TSMAX * 9];

  const mw_kgroup G = groups[item_group[blockIdx.x]];
  const mw_mesh& M = meshes[G.mesh];
  const mw_sweep_prob& p0 = P[group_prob[G.first]];
  const bool sym = p0.symmetric != 0;
  const int qp = item_pair[blockIdx.x];
  const int ti = p0.tile_pairs[2 * qp], tj = p0.tile_pairs[2 * qp + 1];
  const int tid = threadIdx.x;
  const int nr = M.tile_nseg[ti], nc = M.tile_nseg[tj];
  if (tid < nr) stage_segment<NQ>(M, M.tile_seg[ti * MW_TSMAX + tid], rs[tid]);
  if (tid >= 64 && tid - 64 < nc)
    stage_segment<NQ>(M, M.tile_seg[tj * MW_TSMAX + tid - 64], cs[tid - 64]);
  __syncthreads();

  const int npair = nr * nc;
  const int pair = tid >> 1, half = tid & 1;
  const bool live = pair < npair;
  const int li = live ? pair / nc : 0, lj = live ? pair % nc : 0;
  const SegStage<NQ>& I = rs[li];
  const SegStage<NQ>& J = cs[lj];
  const bool same_edge = live && I.edge == J.edge;
  const bool ext = EXT && I.ekl >= 0 && I.ekl == J.ekl;
  const bool smooth = live && !same_edge;

  // the lane's QH x NQ point pairs: phasor exp(-j k R)/(4 pi R) at the
  // current wavenumber and its step exp(-j dk R)
  double er[QH][NQ], ei[QH][NQ], sr[QH][NQ], si[QH][NQ];
#pragma unroll
  for (int qq = 0; qq < QH; ++qq)
#pragma unroll
    for (int r = 0; r < NQ; ++r) { er[qq][r] = ei[qq][r] = 0.0; sr[qq][r] = 1.0; si[qq][r] = 0.0; }
  if (smooth) {
    const double a2 = I.a * I.a;
#pragma unroll
    for (int qq = 0; qq < QH; ++qq) {
      const int q = QH * half + qq;
#pragma unroll
      for (int r = 0; r < NQ; ++r) {
        const double dx = I.pos[q][0] - J.pos[r][0], dy = I.pos[q][1] - J.pos[r][1];
        const double dz = I.pos[q][2] - J.pos[r][2];
        const double R = sqrt(dx * dx + dy * dy + dz * dz + a2);
        double sn, cn;
        sincos(G.k0 * R, &sn, &cn);
        const double inv = INV4PI / R;
        er[qq][r] = cn * inv;
        ei[qq][r] = -sn * inv;
        if (G.nk > 1) {
          sincos(G.dk * R, &sn, &cn);
          sr[qq][r] = cn;
          si[qq][r] = -sn;
        }
      }
    }
  }
  int se_off = 0;
  if (same_edge) {
    const int ed = I.edge;
    const int n_e = M.edge_i[2 * ed], off = M.edge_i[2 * ed + 1];
    se_off = (off + (J.eidx - I.eidx) + n_e - 1) * 9;
  }
  const double td = I.tan[0] * J.tan[0] + I.tan[1] * J.tan[1] + I.tan[2] * J.tan[2];

  // gather: output (m, n) of the tile pair per thread of the first MW_TB^2;
  // the band-block offsets of its up to 9 wing pairs, 16 bits each
  // (0xffff: absent), packed once for every wavenumber
  const int m0 = M.tile_b0[ti], n0 = M.tile_b0[tj];
  const int gm = m0 + tid / MW_TB, gn = n0 + tid % MW_TB;
  const bool gather = tid < MW_TB * MW_TB && gm < M.tile_b0[ti + 1] && gn < M.tile_b0[tj + 1];
  uint32_t goff[5] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu};
  if (gather) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int la = M.bw_loc[3 * gm + a], sa = M.bw_slot[3 * gm + a];
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        const int lb = M.bw_loc[3 * gn + b], sb = M.bw_slot[3 * gn + b];
        const uint32_t o = (la < 0 || lb < 0) ? 0xffffu
                                               : (uint32_t)((la * nc + lb) * 9 + sa * 3 + sb);
        const int i = 3 * a + b, sh = 16 * (i & 1);
        goff[i >> 1] = (goff[i >> 1] & ~(0xffffu << sh)) | (o << sh);
      }
    }
  }

#pragma unroll 2
  for (int f = 0; f < G.nk; ++f) {
    const mw_sweep_prob& p = P[group_prob[G.first + f]];
    const double wmu = p.omega * mu, inv_weps = 1.0 / (p.omega * eps);
    double2* __restrict__ bt = Bt[f & 1];
    double jr[NM][NM], jim[NM][NM];
    if (smooth) {
      // T[b] = sum_r w_b(r) g(q, r) per observer point, then J += w_a(q) T[b];
      // each phasor is used at this wavenumber, then rotated to the next
#pragma unroll
      for (int qq = 0; qq < QH; ++qq) {
        const int q = QH * half + qq;
        double tr[NM], tm[NM];
#pragma unroll
        for (int r = 0; r < NQ; ++r) {
          const double gr = er[qq][r], gi = ei[qq][r];
#pragma unroll
          for (int b = 0; b < NM; ++b) {
            const double w = J.wu[b][r];
            tr[b] = r == 0 ? w * gr : fma(w, gr, tr[b]);
            tm[b] = r == 0 ? w * gi : fma(w, gi, tm[b]);
          }
          if (ext) {
            // the extended kernel's factor, a polynomial in k R: its
            // departure from 1 is added on
            const double dx = I.pos[q][0] - J.pos[r][0], dy = I.pos[q][1] - J.pos[r][1];
            const double dz = I.pos[q][2] - J.pos[r][2];
            const double R = sqrt(dx * dx + dy * dy + dz * dz + I.a * I.a);
            const double2 d = ek_factor_minus1(R, I.a, p.k);
            const double hr = gr * d.x - gi * d.y, hi = gr * d.y + gi * d.x;
#pragma unroll
            for (int b = 0; b < NM; ++b) {
              tr[b] = fma(J.wu[b][r], hr, tr[b]);
              tm[b] = fma(J.wu[b][r], hi, tm[b]);
            }
          }
          er[qq][r] = gr * sr[qq][r] - gi * si[qq][r];
          ei[qq][r] = gr * si[qq][r] + gi * sr[qq][r];
        }
#pragma unroll
        for (int a = 0; a < NM; ++a) {
          const double w = I.wu[a][q];
#pragma unroll
          for (int b = 0; b < NM; ++b) {
            jr[a][b] = qq == 0 ? w * tr[b] : fma(w, tr[b], jr[a][b]);
            jim[a][b] = qq == 0 ? w * tm[b] : fma(w, tm[b], jim[a][b]);
          }
        }
      }
    } else {
#pragma unroll
      for (int a = 0; a < NM; ++a)
#pragma unroll
        for (int b = 0; b < NM; ++b) { jr[a][b] = 0.0; jim[a][b] = 0.0; }
      if (same_edge && half == 0) {
        const double2* tab = reinterpret_cast<const double2*>(p.se_tab) + se_off;
#pragma unroll
        for (int a = 0; a < NM; ++a)
#pragma unroll
          for (int b = 0; b < NM; ++b) {
            const double2 v = tab[a * NM + b];
            jr[a][b] = v.x; jim[a][b] = v.y;
          }
      }
    }
#pragma unroll
    for (int a = 0; a < NM; ++a)
#pragma unroll
      for (int b = 0; b < NM; ++b) {
        jr[a][b] += lane_swap1(jr[a][b]);
        jim[a][b] += lane_swap1(jim[a][b]);
      }
    if (live) contract_band<NQ>(jr, jim, I, J, td, wmu, inv_weps, half, bt + (li * nc + lj) * 9);
    __syncthreads();
    if (gather) {
      double zr = 0.0, zi = 0.0;
#pragma unroll
      for (int i = 0; i < 9; ++i) {
        const uint32_t o = (goff[i >> 1] >> (16 * (i & 1))) & 0xffffu;
        if (o != 0xffffu) {
          const double2 v = bt[o];
          zr += v.x; zi += v.y;
        }
      }
      double2* Z = reinterpret_cast<double2*>(p.A);
      Z[(int64_t)gm * p.lda + gn] = make_double2(zr, zi);
      if (sym && ti != tj) Z[(int64_t)gn * p.lda + gm] = make_double2(zr, zi);
    }
  }
}

template <int NQ, bool FAR32 = false>
hipError_t launch_fill_nq(const mw_mesh& M, double k, double wmu, double inv_weps, double2* Z,
                          int64_t ldz, const int32_t* pairs, int32_t n_pairs, bool symmetric,
                          hipStream_t st, double near2 = 0.0) {
  if (symmetric) {
    if (n_pairs == 0) return hipSuccess;
    hipLaunchKernelGGL((k_fill<true, NQ, FAR32>), dim3(n_pairs), dim3(256), 0, st, M, k, wmu,
                       inv_weps, Z, ldz, pairs, near2);
  } else {
    hipLaunchKernelGGL((k_fill<false, NQ, FAR32>), dim3(M.n_tile, M.n_tile), dim3(256), 0, st, M,
                       k, wmu, inv_weps, Z, ldz, pairs, near2);
  }
  return hipGetLastError();
}

}  // namespace

// Dispatch on the mesh's point count (the mesh check holds it to 1 .. NQ_MAX).
#define MW_NQ_SWITCH(nq, CALL) \
  switch (nq) {                \
    case 1: CALL(1);           \
    case 2: CALL(2);           \
    case 3: CALL(3);           \
    case 4: CALL(4);           \
    case 5: CALL(5);           \
    case 6: CALL(6);           \
    case 7: CALL(7);           \
    case 8: CALL(8);           \
    default: return hipErrorInvalidValue; \
  }

// ---------------------------------------------------------------------------
// Host launchers (called from mw_capi.cpp)
// ---------------------------------------------------------------------------
hipError_t mw_launch_same_edge(const mw_mesh& M, double k, hipStream_t st) {
  if (M.n_se == 0) return hipSuccess;
#define MW_SE(NQ)                                                                          \
  {                                                                                        \
    const int64_t thr = (int64_t)M.n_se * SeLanes<NQ>::LG;                                 \
    hipLaunchKernelGGL(k_same_edge<NQ>, dim3((unsigned)((thr + 255) / 256)), dim3(256), 0, \
                       st, M, k);                                                          \
    return hipGetLastError();                                                              \
  }
  MW_NQ_SWITCH(M.n_qp, MW_SE)
#undef MW_SE
}

hipError_t mw_launch_same_edge_static(const mw_mesh& M, double* out, hipStream_t st) {
  if (M.n_se == 0) return hipSuccess;
  const int64_t thr = (int64_t)M.n_se * 9;
  hipLaunchKernelGGL(k_same_edge_static, dim3((unsigned)((thr + 255) / 256)), dim3(256), 0, st, M,
                     (int)M.n_qp, out);
  return hipGetLastError();
}

hipError_t mw_launch_fill(const mw_mesh& M, double k, double omega, double eps, double mu,
                          double2* Z, int64_t ldz, const int32_t* pairs, int32_t n_pairs,
                          bool symmetric, hipStream_t st) {
  const double wmu = omega * mu, inv_weps = 1.0 / (omega * eps);
#define MW_FILL(NQ) return launch_fill_nq<NQ>(M, k, wmu, inv_weps, Z, ldz, pairs, n_pairs, symmetric, st);
  MW_NQ_SWITCH(M.n_qp, MW_FILL)
#undef MW_FILL
}

// 256-thread blocks of a batched same-edge table (n_qp_pair = 4)
int64_t mw_same_edge_blocks(int32_t n_se) {
  return ((int64_t)n_se * SeLanes<4>::LG + 255) / 256;
}

// The mixed form's fill: far pairs with an FP32 phase (near_ratio: the
// distance, in segment lengths, inside which pairs stay FP64).
hipError_t mw_launch_fill_far32(const mw_mesh& M, double k, double omega, double eps, double mu,
                                double2* Z, int64_t ldz, const int32_t* pairs, int32_t n_pairs,
                                bool symmetric, double near_ratio, hipStream_t st) {
  const double wmu = omega * mu, inv_weps = 1.0 / (omega * eps);
  const double near2 = near_ratio * near_ratio;
#define MW_FILL32(NQ) return launch_fill_nq<NQ, true>(M, k, wmu, inv_weps, Z, ldz, pairs, n_pairs, symmetric, st, near2);
  MW_NQ_SWITCH(M.n_qp, MW_FILL32)
#undef MW_FILL32
}

// Items [0, n_plain) are tile pairs of meshes without the extended kernel,
// [n_plain, n_items) of meshes with it.
hipError_t mw_launch_batch_fill(const mw_sweep_prob* P, const mw_mesh* meshes, int nprob,
                                const int32_t* se_items, int64_t n_se_items,
                                const mw_kgroup* groups, const int32_t* group_prob,
                                const int32_t* item_group, const int32_t* item_pair,
                                int64_t n_items, int64_t n_plain, double eps, double mu,
                                hipStream_t st) {
  if (nprob <= 0) return hipSuccess;
  if (n_se_items > 0)
    hipLaunchKernelGGL(kb_same_edge, dim3((unsigned)n_se_items), dim3(256), 0, st, P, meshes,
                       se_items);
  if (n_plain > 0)
    hipLaunchKernelGGL(kb_fill_k<false>, dim3((unsigned)n_plain), dim3(KF_THREADS), 0, st, P,
                       meshes, groups, group_prob, item_group, item_pair, eps, mu);
  if (n_items > n_plain)
    hipLaunchKernelGGL(kb_fill_k<true>, dim3((unsigned)(n_items - n_plain)), dim3(KF_THREADS), 0,
                       st, P, meshes, groups, group_prob, item_group + n_plain,
                       item_pair + n_plain, eps, mu);
  return hipGetLastError();
}
