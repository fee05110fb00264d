// Readout of solved port columns on gfx950: element current moments and
// continuity steps, far-field current moments on a (theta, phi) grid, and
// near fields at points, every port column at once.
//
// Far field (reference portal/_portal.py:1520-1598): per direction r^ and
// port column c,
//   P_d(r^, c) = sum_e exp(+jk r^ . r_e) m_e(c)              (3 components)
//   P_i(r^, c) = sum_e exp(+jk r^ . r_e') m_e(c)             (r_e' mirrored)
//   M_theta = theta^ . P_d + rho_v  theta^ . S P_i
//   M_phi   = phi^   . P_d - rho_h  phi^   . S P_i,   S = diag(-1, -1, 1),
// with (rho_h, rho_v) = (-1, 1) over PEC and the Fresnel image multipliers of
// portal/_portal.py:1397-1423 over a finite ground.  The sums are complex
// GEMMs (directions x elements) x (elements x 3 columns) on FP64 MFMA whose
// left factor is generated in LDS, a tile at a time, and never stored.
#include "mw_common.h"
#include "mw_mfma.h"

namespace {

constexpr double PI_D = 3.14159265358979323846;

// ---------------------------------------------------------------------------
// Element moments m_e = 0.5 (I_a + I_b) dl_e and continuity steps
// delta_n = el_in - el_out (reference _element_currents.py:86-120) from the
// solution columns X (n_b x p, row-major, ldx).  The current at a sample is
// the B-spline sum over at most three taps (basis index, value).
// ---------------------------------------------------------------------------
MW_DEV double2 sample_current(int s, int c, const int32_t* tap_i, const double* tap_v,
                              const double2* X, int64_t ldx) {
  double2 acc = make_double2(0.0, 0.0);
#pragma unroll
  for (int t = 0; t < 3; ++t) {
    const int j = tap_i[3 * s + t];
    if (j >= 0) {
      const double v = tap_v[3 * s + t];
      const double2 x = X[(int64_t)j * ldx + c];
      acc.x += v * x.x;
      acc.y += v * x.y;
    }
  }
  return acc;
}

MW_DEV double2 element_current(int e, int c, const int32_t* el_tab, const int32_t* tap_i,
                               const double* tap_v, const double2* X, int64_t ldx) {
  const double2 ia = sample_current(el_tab[2 * e], c, tap_i, tap_v, X, ldx);
  const double2 ib = sample_current(el_tab[2 * e + 1], c, tap_i, tap_v, X, ldx);
  return make_double2(0.5 * (ib.x + ia.x), 0.5 * (ib.y + ia.y));
}

__global__ __launch_bounds__(256) void k_element_moments(
    int n_el, int n_node, int p, const int32_t* __restrict__ tap_i,
    const double* __restrict__ tap_v, const int32_t* __restrict__ el_tab,
    const double* __restrict__ dl, const int32_t* __restrict__ node_tab,
    const double2* __restrict__ X, int64_t ldx, double2* __restrict__ mom, int64_t ldm,
    double2* __restrict__ delta, int64_t ldd) {
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t row = idx / p;
  const int c = (int)(idx - row * p);
  if (row < n_el) {
    const int e = (int)row;
    const double2 el = element_current(e, c, el_tab, tap_i, tap_v, X, ldx);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double w = dl[3 * e + d];
      mom[e * ldm + (int64_t)d * p + c] = make_double2(el.x * w, el.y * w);
    }
  } else if (row < (int64_t)n_el + n_node) {
    const int n = (int)(row - n_el);
    const int ein = node_tab[2 * n], eout = node_tab[2 * n + 1];
    double2 a = make_double2(0.0, 0.0), b = make_double2(0.0, 0.0);
    if (ein >= 0) a = element_current(ein, c, el_tab, tap_i, tap_v, X, ldx);
    if (eout >= 0) b = element_current(eout, c, el_tab, tap_i, tap_v, X, ldx);
    delta[n * ldd + c] = make_double2(a.x - b.x, a.y - b.y);
  }
}

// ---------------------------------------------------------------------------
// Far-field moments.  Workgroup: FD = 64 directions x FC = 16 port columns
// (48 complex product columns: 3 components x 16), four waves, wave w the
// 16 directions 16w..16w+15 across all six 16-wide real column tiles, so
// each lane holds all three components of its (direction, column) and the
// projections need no exchange beyond the (re, im) lane pair.
// ---------------------------------------------------------------------------
constexpr int FD = 64;   // directions per workgroup
constexpr int FE = 16;   // elements per k chunk (32 real k)
constexpr int FC = 32;   // port columns per workgroup
constexpr int FB = 3 * FC;
constexpr int HT = FC / 8;  // 16-wide real column tiles per component

struct FarLds {
  double A[FD][2 * FE + 1];   // phase tile, (re, im) pairs along the row
  double2 B[FE][FB + 1];      // moments: [element][component * FC + column]
  double r[FD][3];            // r^ of the tile's directions
  double m[FE][3];            // the chunk's element midpoints (z mirrored for the image)
};

MW_DEV double2 csqrt_principal(double x, double y) {
  const double m = hypot(x, y);
  if (m == 0.0) return make_double2(0.0, y);
  if (x >= 0.0) {
    const double t = sqrt(0.5 * (m + x));
    return make_double2(t, y / (2.0 * t));
  }
  const double t = sqrt(0.5 * (m - x));
  return make_double2(fabs(y) / (2.0 * t), copysign(t, y));
}

// Stage the phase tile exp(+jk r^ . r_e) for elements e0 .. e0 + FE - 1 (image:
// z mirrored about ground_z) and the matching moment rows.
MW_DEV void far_stage(FarLds& L, int e0, bool image, int n_el, int ndir_t, const double* mid,
                      double k, double gz2, const double2* mom, int64_t ldm, int p, int col0,
                      bool stage_b) {
  const int tid = threadIdx.x;
  if (tid < FE * 3) {
    const int e = tid / 3, c = tid % 3, eg = e0 + e;
    double v = 0.0;
    if (eg < n_el) v = (image && c == 2) ? gz2 - mid[3 * eg + 2] : mid[3 * eg + c];
    L.m[e][c] = v;
  }
  __syncthreads();
#pragma unroll 1
  for (int i = tid; i < FD * FE; i += 256) {
    const int r = i / FE, e = i % FE, eg = e0 + e;
    double cs = 0.0, sn = 0.0;
    if (r < ndir_t && eg < n_el) {
      const double ph = k * (L.r[r][0] * L.m[e][0] + L.r[r][1] * L.m[e][1] + L.r[r][2] * L.m[e][2]);
      sincos(ph, &sn, &cs);
    }
    L.A[r][2 * e] = cs;
    L.A[r][2 * e + 1] = sn;
  }
  if (stage_b)
    for (int i = tid; i < FE * FB; i += 256) {
      const int e = i / FB, j = i % FB, comp = j / FC, c = col0 + j % FC, eg = e0 + e;
      double2 v = make_double2(0.0, 0.0);
      if (eg < n_el && c < p) v = mom[eg * ldm + (int64_t)comp * p + c];
      L.B[e][j] = v;
    }
}

// acc += A B_exp over the staged chunk.  Wave w owns the 16 directions
// 16w..16w+15 and the six real column tiles.  B_exp expands each complex b
// to [[br, bi], [-bi, br]]: at real row parity kpar and real column parity
// odd it is br (kpar == odd), bi (kpar = 0, odd = 1) or -bi (kpar = 1,
// odd = 0); both parities are fixed per lane, so each lane reads one real.
MW_DEV void far_mfma(const FarLds& L, Mfma<double>::acc_t (&acc)[TileGrid<FD, FB>::TPW]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int arow = 16 * wave + (lane & 15);
  const int kq = lane >> 4, odd = lane & 1, kpar = kq & 1;
  const int sel = kpar ^ odd;                     // 1: take b.y
  const double sgn = (kpar && !odd) ? -1.0 : 1.0;  // -bi on (odd k, real column)
  const double* Bd = reinterpret_cast<const double*>(&L.B[0][0]);
#pragma unroll 2
  for (int ks = 0; ks < 2 * FE; ks += 4) {
    const int kr = ks + kq;
    const double a = L.A[arow][kr];
    const double* brow = Bd + 2 * ((kr >> 1) * (FB + 1)) + sel;
#pragma unroll
    for (int j = 0; j < TileGrid<FD, FB>::TPW; ++j) {
      const int cc = (16 * j + (lane & 15)) >> 1;
      acc[j] = Mfma<double>::mma(a, sgn * brow[2 * cc], acc[j]);
    }
  }
}

// ground: 0 free space, 1 PEC, 2 finite (Fresnel multipliers from eps_c).
__global__ __launch_bounds__(256, 2) void k_far_moments(
    int n_el, const double* __restrict__ mid, const double2* __restrict__ mom, int64_t ldm,
    int p, int n_theta, int n_phi, const double* __restrict__ trig_t,
    const double* __restrict__ trig_p, double k, int ground, double ground_z, double eps_re,
    double eps_im, double2* __restrict__ m_theta, double2* __restrict__ m_phi, int64_t ldo) {
  __shared__ FarLds L;
  typedef Mfma<double>::acc_t acc_t;
  constexpr int TPW = TileGrid<FD, FB>::TPW;  // 6: one row tile, all column tiles
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t n_dir = (int64_t)n_theta * n_phi;
  const int64_t d0 = (int64_t)blockIdx.x * FD;
  const int col0 = blockIdx.y * FC;
  const int ndir_t = (int)min<int64_t>(FD, n_dir - d0);
  if (tid < FD) {
    double rx = 0.0, ry = 0.0, rz = 0.0;
    if (tid < ndir_t) {
      const int64_t dg = d0 + tid;
      const int it = (int)(dg / n_phi), ip = (int)(dg % n_phi);
      const double st = trig_t[2 * it], ct = trig_t[2 * it + 1];
      const double sp = trig_p[2 * ip], cp = trig_p[2 * ip + 1];
      rx = st * cp; ry = st * sp; rz = ct;
    }
    L.r[tid][0] = rx; L.r[tid][1] = ry; L.r[tid][2] = rz;
  }
  const double gz2 = 2.0 * ground_z;
  acc_t acc[TPW];
  // One product pass over every element chunk: acc = sum_e A_e B_e, A the
  // direct (image = false) or mirrored (image = true) phase tile.
  auto pass = [&](bool image) {
#pragma unroll
    for (int j = 0; j < TPW; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[j][q] = 0.0;
    for (int e0 = 0; e0 < n_el; e0 += FE) {
      __syncthreads();
      far_stage(L, e0, image, n_el, ndir_t, mid, k, gz2, mom, ldm, p, col0, true);
      __syncthreads();
      far_mfma(L, acc);
    }
  };
  // Lane l holds real column 16 tc + l % 16 of tile tc = HT comp + h: complex
  // column 8 h + (l % 16) / 2 of component comp, the real part on even lanes
  // and the imaginary part on odd ones; rows Mfma::row(l, rr) of the wave's
  // 16 directions.
  const int part = lane & 1;
  auto frame = [&](int rr, double (&th)[3], double (&ph)[3], double& st, double& ct, double& sp,
                   double& cp) {
    const int rloc = 16 * wave + Mfma<double>::row(lane, rr);
    st = ct = sp = cp = 0.0;
    if (rloc < ndir_t) {
      const int64_t dg = d0 + rloc;
      const int it = (int)(dg / n_phi), ip = (int)(dg % n_phi);
      st = trig_t[2 * it]; ct = trig_t[2 * it + 1];
      sp = trig_p[2 * ip]; cp = trig_p[2 * ip + 1];
    }
    th[0] = ct * cp; th[1] = ct * sp; th[2] = -st;
    ph[0] = -sp; ph[1] = cp; ph[2] = 0.0;
  };
  // Image pass first; only this lane's part of the reflected contribution
  //   rho_v theta^ . S P_i,  -rho_h phi^ . S P_i
  // is kept across the direct pass.
  double img_t[4][HT], img_p[4][HT];
#pragma unroll
  for (int rr = 0; rr < 4; ++rr)
#pragma unroll
    for (int h = 0; h < HT; ++h) { img_t[rr][h] = 0.0; img_p[rr][h] = 0.0; }
  if (ground) {
    pass(true);
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      double th[3], ph[3], st, ct, sp, cp;
      frame(rr, th, ph, st, ct, sp, cp);
      double2 rho_v = make_double2(1.0, 0.0), mrho_h = make_double2(1.0, 0.0);  // PEC
      if (ground == 2) {
        const double rx = st * cp, ry = st * sp, rz = ct;
        const double2 q = csqrt_principal(eps_re - rx * rx - ry * ry, eps_im);
        const double2 rh = cdiv(make_double2(rz - q.x, -q.y), make_double2(rz + q.x, q.y));
        const double2 ez = make_double2(eps_re * rz, eps_im * rz);
        rho_v = cdiv(make_double2(ez.x - q.x, ez.y - q.y), make_double2(ez.x + q.x, ez.y + q.y));
        mrho_h = make_double2(-rh.x, -rh.y);
      }
#pragma unroll
      for (int h = 0; h < HT; ++h) {
        double ti = 0.0, pi = 0.0;
#pragma unroll
        for (int comp = 0; comp < 3; ++comp) {
          const double vi = (comp == 2 ? 1.0 : -1.0) * acc[HT * comp + h][rr];
          ti += th[comp] * vi;
          pi += ph[comp] * vi;
        }
        // complete the sums to complex values through the (re, im) lane pair
        const double ti_o = __shfl_xor(ti, 1), pi_o = __shfl_xor(pi, 1);
        const double2 tiz = part ? make_double2(ti_o, ti) : make_double2(ti, ti_o);
        const double2 piz = part ? make_double2(pi_o, pi) : make_double2(pi, pi_o);
        const double2 a = cmul(rho_v, tiz), b = cmul(mrho_h, piz);
        img_t[rr][h] = part ? a.y : a.x;
        img_p[rr][h] = part ? b.y : b.x;
      }
    }
  }
  pass(false);
  double* out_t = reinterpret_cast<double*>(m_theta);
  double* out_p = reinterpret_cast<double*>(m_phi);
#pragma unroll
  for (int rr = 0; rr < 4; ++rr) {
    double th[3], ph[3], st, ct, sp, cp;
    frame(rr, th, ph, st, ct, sp, cp);
    const int rloc = 16 * wave + Mfma<double>::row(lane, rr);
    const int64_t dg = d0 + rloc;
#pragma unroll
    for (int h = 0; h < HT; ++h) {
      const int cc = col0 + 8 * h + ((lane & 15) >> 1);
      double mt = 0.0, mp = 0.0;
#pragma unroll
      for (int comp = 0; comp < 3; ++comp) {
        mt += th[comp] * acc[HT * comp + h][rr];
        mp += ph[comp] * acc[HT * comp + h][rr];
      }
      mt += img_t[rr][h];
      mp += img_p[rr][h];
      if (rloc < ndir_t && cc < p) {
        out_t[2 * (dg * ldo + cc) + part] = mt;
        out_p[2 * (dg * ldo + cc) + part] = mp;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Near fields in mixed-potential form (reference portal/_portal.py:1601-1651):
//   E = -j eta k sum_e G_e m_e - j eta / k sum_n delta_n (1 + jkR_n)/R_n G_n R^_n
//   H = sum_e (jk + 1/R_e) G_e (m_e x R^_e),   G = exp(-jkR) / (4 pi R),
// R = sqrt(|r - r_s|^2 + a^2).  One thread per point and NCB port columns;
// sources are staged through LDS in chunks of NSRC.
// ---------------------------------------------------------------------------
constexpr int NCB = 4;     // port columns per thread
constexpr int NSRC = 64;   // sources per LDS chunk

struct NearLds {
  double pos[NSRC][3];
  double2 val[NSRC][3 * NCB];  // element moments (3 comps x NCB) or node steps (NCB)
};

__global__ __launch_bounds__(256) void k_near_fields(
    int n_pts, const double* __restrict__ pts, int n_el, const double* __restrict__ mid,
    const double2* __restrict__ mom, int64_t ldm, int n_node, const double* __restrict__ nodes,
    const double2* __restrict__ delta, int64_t ldd, double node_sign, int p, double k,
    double radius, double eta, int magnetic, int accumulate, double2* __restrict__ out,
    int64_t ldo) {
  __shared__ NearLds L;
  const int tid = threadIdx.x;
  const int ip = blockIdx.x * 256 + tid;
  const int col0 = blockIdx.y * NCB;
  const bool live = ip < n_pts;
  double px = 0.0, py = 0.0, pz = 0.0;
  if (live) { px = pts[3 * ip]; py = pts[3 * ip + 1]; pz = pts[3 * ip + 2]; }
  const double a2 = radius * radius;
  // sum_e G m (E) or sum_e w G (m x R^) (H), per column and component
  double2 acc[NCB][3];
#pragma unroll
  for (int c = 0; c < NCB; ++c)
#pragma unroll
    for (int d = 0; d < 3; ++d) acc[c][d] = make_double2(0.0, 0.0);
  for (int s0 = 0; s0 < n_el; s0 += NSRC) {
    __syncthreads();
    for (int i = tid; i < NSRC * 3 * NCB; i += 256) {
      const int s = i / (3 * NCB), j = i % (3 * NCB), d = j / NCB, c = col0 + j % NCB;
      double2 v = make_double2(0.0, 0.0);
      if (s0 + s < n_el && c < p) v = mom[(int64_t)(s0 + s) * ldm + (int64_t)d * p + c];
      L.val[s][j] = v;
    }
    for (int i = tid; i < NSRC * 3; i += 256)
      L.pos[i / 3][i % 3] = (s0 + i / 3 < n_el) ? mid[3 * s0 + i] : 0.0;
    __syncthreads();
    const int ns = min(NSRC, n_el - s0);
    for (int s = 0; s < ns; ++s) {
      const double rx = px - L.pos[s][0], ry = py - L.pos[s][1], rz = pz - L.pos[s][2];
      const double r = sqrt(rx * rx + ry * ry + rz * rz + a2);
      double sn, cs;
      sincos(k * r, &sn, &cs);
      const double den = (4.0 * PI_D) * r;
      const double2 g = make_double2(cs / den, -sn / den);
      if (!magnetic) {
#pragma unroll
        for (int c = 0; c < NCB; ++c)
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            const double2 m = L.val[s][d * NCB + c];
            acc[c][d].x += g.x * m.x - g.y * m.y;
            acc[c][d].y += g.x * m.y + g.y * m.x;
          }
      } else {
        const double ux = rx / r, uy = ry / r, uz = rz / r;
        // w G = (jk + 1/R) G
        const double ir = 1.0 / r;
        const double2 wg = make_double2(ir * g.x - k * g.y, ir * g.y + k * g.x);
#pragma unroll
        for (int c = 0; c < NCB; ++c) {
          const double2 mx = L.val[s][c], my = L.val[s][NCB + c], mz = L.val[s][2 * NCB + c];
          const double2 cr[3] = {
              make_double2(my.x * uz - mz.x * uy, my.y * uz - mz.y * uy),
              make_double2(mz.x * ux - mx.x * uz, mz.y * ux - mx.y * uz),
              make_double2(mx.x * uy - my.x * ux, mx.y * uy - my.y * ux)};
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            acc[c][d].x += wg.x * cr[d].x - wg.y * cr[d].y;
            acc[c][d].y += wg.x * cr[d].y + wg.y * cr[d].x;
          }
        }
      }
    }
  }
  double2 res[NCB][3];
  if (!magnetic) {
    // -j eta k (sum G m)
#pragma unroll
    for (int c = 0; c < NCB; ++c)
#pragma unroll
      for (int d = 0; d < 3; ++d)
        res[c][d] = make_double2(eta * k * acc[c][d].y, -eta * k * acc[c][d].x);
    double2 q[NCB][3];
#pragma unroll
    for (int c = 0; c < NCB; ++c)
#pragma unroll
      for (int d = 0; d < 3; ++d) q[c][d] = make_double2(0.0, 0.0);
    for (int s0 = 0; s0 < n_node; s0 += NSRC) {
      __syncthreads();
      for (int i = tid; i < NSRC * NCB; i += 256) {
        const int s = i / NCB, c = col0 + i % NCB;
        double2 v = make_double2(0.0, 0.0);
        if (s0 + s < n_node && c < p) {
          v = delta[(int64_t)(s0 + s) * ldd + c];
          v.x *= node_sign;
          v.y *= node_sign;
        }
        L.val[s][i % NCB] = v;
      }
      for (int i = tid; i < NSRC * 3; i += 256)
        L.pos[i / 3][i % 3] = (s0 + i / 3 < n_node) ? nodes[3 * s0 + i] : 0.0;
      __syncthreads();
      const int ns = min(NSRC, n_node - s0);
      for (int s = 0; s < ns; ++s) {
        const double rx = px - L.pos[s][0], ry = py - L.pos[s][1], rz = pz - L.pos[s][2];
        const double r = sqrt(rx * rx + ry * ry + rz * rz + a2);
        double sn, cs;
        sincos(k * r, &sn, &cs);
        const double den = (4.0 * PI_D) * r;
        const double2 g = make_double2(cs / den, -sn / den);
        // (1 + jkR)/R G, then times -j eta / k
        const double2 f = make_double2((g.x - k * r * g.y) / r, (g.y + k * r * g.x) / r);
        const double2 sf = make_double2(eta / k * f.y, -eta / k * f.x);
        const double ux = rx / r, uy = ry / r, uz = rz / r;
#pragma unroll
        for (int c = 0; c < NCB; ++c) {
          const double2 dv = L.val[s][c];
          const double2 sc = make_double2(sf.x * dv.x - sf.y * dv.y, sf.x * dv.y + sf.y * dv.x);
          q[c][0].x += sc.x * ux; q[c][0].y += sc.y * ux;
          q[c][1].x += sc.x * uy; q[c][1].y += sc.y * uy;
          q[c][2].x += sc.x * uz; q[c][2].y += sc.y * uz;
        }
      }
    }
#pragma unroll
    for (int c = 0; c < NCB; ++c)
#pragma unroll
      for (int d = 0; d < 3; ++d) { res[c][d].x += q[c][d].x; res[c][d].y += q[c][d].y; }
  } else {
#pragma unroll
    for (int c = 0; c < NCB; ++c)
#pragma unroll
      for (int d = 0; d < 3; ++d) res[c][d] = acc[c][d];
  }
  if (!live) return;
#pragma unroll
  for (int c = 0; c < NCB; ++c) {
    if (col0 + c >= p) break;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double2* o = out + (int64_t)ip * ldo + (int64_t)d * p + col0 + c;
      double2 v = res[c][d];
      if (accumulate) { v.x += o->x; v.y += o->y; }
      *o = v;
    }
  }
}

}  // namespace

hipError_t mw_launch_element_moments(int n_el, int n_node, int p, const int32_t* tap_i,
                                     const double* tap_v, const int32_t* el_tab, const double* dl,
                                     const int32_t* node_tab, const double2* X, int64_t ldx,
                                     double2* mom, int64_t ldm, double2* delta, int64_t ldd,
                                     hipStream_t st) {
  const int64_t work = ((int64_t)n_el + n_node) * p;
  if (work == 0) return hipSuccess;
  hipLaunchKernelGGL(k_element_moments, dim3((unsigned)((work + 255) / 256)), dim3(256), 0, st,
                     n_el, n_node, p, tap_i, tap_v, el_tab, dl, node_tab, X, ldx, mom, ldm, delta,
                     ldd);
  return hipGetLastError();
}

hipError_t mw_launch_far_moments(int n_el, const double* mid, const double2* mom, int64_t ldm,
                                 int p, int n_theta, int n_phi, const double* trig_t,
                                 const double* trig_p, double k, int ground, double ground_z,
                                 double eps_re, double eps_im, double2* m_theta, double2* m_phi,
                                 int64_t ldo, hipStream_t st) {
  const int64_t n_dir = (int64_t)n_theta * n_phi;
  if (n_dir == 0 || p == 0) return hipSuccess;
  dim3 grid((unsigned)((n_dir + FD - 1) / FD), (unsigned)((p + FC - 1) / FC));
  hipLaunchKernelGGL(k_far_moments, grid, dim3(256), 0, st, n_el, mid, mom, ldm, p, n_theta,
                     n_phi, trig_t, trig_p, k, ground, ground_z, eps_re, eps_im, m_theta, m_phi,
                     ldo);
  return hipGetLastError();
}

hipError_t mw_launch_near_fields(int n_pts, const double* pts, int n_el, const double* mid,
                                 const double2* mom, int64_t ldm, int n_node, const double* nodes,
                                 const double2* delta, int64_t ldd, double node_sign, int p,
                                 double k, double radius, double eta, int magnetic,
                                 int accumulate, double2* out, int64_t ldo, hipStream_t st) {
  if (n_pts == 0 || p == 0) return hipSuccess;
  dim3 grid((unsigned)((n_pts + 255) / 256), (unsigned)((p + NCB - 1) / NCB));
  hipLaunchKernelGGL(k_near_fields, grid, dim3(256), 0, st, n_pts, pts, n_el, mid, mom, ldm,
                     n_node, nodes, delta, ldd, node_sign, p, k, radius, eta, magnetic,
                     accumulate, out, ldo);
  return hipGetLastError();
}
