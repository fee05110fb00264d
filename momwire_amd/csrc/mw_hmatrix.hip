// Block-low-rank (H-matrix) operator of the B-spline impedance matrix on
// MI355X (gfx950): Z entries at arbitrary basis pairs, the dense leaf blocks,
// partial-pivoted ACA of every admissible block at once (one workgroup per
// block), the operator's product with a block of vectors, and the batched
// inverses of the leaf diagonal blocks that precondition the Krylov solve.
//
// The entries are the dense fill's (mw_fill.hip): per basis pair, the sum
// over its wing segment pairs of the band entry
//   j w mu (t_i . t_j) A + Phi / (j w eps),
//   A   = sum_qr w_q L_i P_i(u_q) w_r L_j P_j(u_r) g(R_qr),
//   Phi = the same with the basis derivatives P_i', P_j',
// g = exp(-jkR)/(4 pi R) at R^2 = |x_q - y_r|^2 + a_i^2 (times the extended
// kernel's factor on coaxial pairs); same-edge pairs of dense blocks take the
// analytic same-edge table instead (reference hmatrix.py:404-596).
#include "mw_common.h"

namespace {

constexpr int NM = 3;
constexpr double INV4PI = 0.079577471545947667884441881686257181017229822870228;

struct HSeg {
  double l[3], r[3], t[3];
  double L, a;
  int edge, eidx, ekl;
};

MW_DEV void hseg(const mw_mesh& M, int s, HSeg& S) {
  const double* g = M.seg + (size_t)s * 12;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    S.l[c] = g[c];
    S.r[c] = g[3 + c];
    S.t[c] = g[6 + c];
  }
  S.L = g[9];
  S.a = g[10];
  S.edge = M.seg_edge[2 * s];
  S.eidx = M.seg_edge[2 * s + 1];
  S.ekl = M.ek ? M.ek_label[s] : -1;
}

// Band entry B_ij[sa][sb] of segment pair (i, j); `table`: same-edge pairs
// from the same-edge table (dense blocks), else every pair by quadrature
// (the admissible blocks' off-edge kernel, reference hmatrix.py:1097-1155).
MW_DEV double2 band_entry(const mw_mesh& M, int si, int sa, int sj, int sb, double k, double wmu,
                          double inv_weps, bool table) {
  HSeg I, J;
  hseg(M, si, I);
  hseg(M, sj, J);
  const double* ci = M.band_poly + (size_t)si * 9 + sa * 3;
  const double* cj = M.band_poly + (size_t)sj * 9 + sb * 3;
  double Ar = 0.0, Ai = 0.0, Pr = 0.0, Pi = 0.0;
  if (table && I.edge == J.edge) {
    const int ed = I.edge;
    const int n_e = M.edge_i[2 * ed], off = M.edge_i[2 * ed + 1];
    const double2* tab =
        reinterpret_cast<const double2*>(M.se_tab) + (size_t)(off + (J.eidx - I.eidx) + n_e - 1) * 9;
    const double di[2] = {ci[1], 2.0 * ci[2]}, dj[2] = {cj[1], 2.0 * cj[2]};
#pragma unroll
    for (int p = 0; p < NM; ++p)
#pragma unroll
      for (int P = 0; P < NM; ++P) {
        const double2 v = tab[p * NM + P];
        const double w = ci[p] * cj[P];
        Ar += w * v.x;
        Ai += w * v.y;
        if (p < 2 && P < 2) {
          const double wd = di[p] * dj[P];
          Pr += wd * v.x;
          Pi += wd * v.y;
        }
      }
  } else {
    const int nq = M.n_qp;
    const double* t01 = M.gl;
    const double* w01 = M.gl + nq;
    const bool ext = I.ekl >= 0 && I.ekl == J.ekl;
    const double a2 = I.a * I.a;
    // exp(-jkR) = exp(-jk Rc) exp(-jk (R - Rc)), Rc between the midpoints: one
    // full sincos per pair, a short series per point pair (as the dense fill)
    const bool series = k * 0.5 * (I.L + J.L) <= KDR_MAX;
    double Rc = 0.0, sc = 0.0, cc = 1.0;
    if (series) {
      double m2 = a2, rci;
      for (int c = 0; c < 3; ++c) {
        const double d = 0.5 * (I.l[c] + I.r[c] - J.l[c] - J.r[c]);
        m2 += d * d;
      }
      r_and_rinv(m2, Rc, rci);
      sincos(k * Rc, &sc, &cc);
    }
    for (int q = 0; q < nq; ++q) {
      const double tq = t01[q], uq = tq * I.L, wq = w01[q] * I.L;
      const double Pq = wq * (ci[0] + uq * (ci[1] + uq * ci[2]));
      const double Dq = wq * (ci[1] + 2.0 * ci[2] * uq);
      const double x = I.l[0] + tq * (I.r[0] - I.l[0]);
      const double y = I.l[1] + tq * (I.r[1] - I.l[1]);
      const double z = I.l[2] + tq * (I.r[2] - I.l[2]);
      double sAr = 0.0, sAi = 0.0, sPr = 0.0, sPi = 0.0;
      for (int r = 0; r < nq; ++r) {
        const double tr = t01[r], ur = tr * J.L, wr = w01[r] * J.L;
        const double Qr = wr * (cj[0] + ur * (cj[1] + ur * cj[2]));
        const double Er = wr * (cj[1] + 2.0 * cj[2] * ur);
        const double dx = x - (J.l[0] + tr * (J.r[0] - J.l[0]));
        const double dy = y - (J.l[1] + tr * (J.r[1] - J.l[1]));
        const double dz = z - (J.l[2] + tr * (J.r[2] - J.l[2]));
        double R, rinv;
        r_and_rinv(dx * dx + dy * dy + dz * dz + a2, R, rinv);
        const double inv = INV4PI * rinv;
        double gr, gi;
        if (series) {
          double cd, sd;
          cos_sin_small(k * (R - Rc), cd, sd);
          gr = fma(cc, cd, -sc * sd) * inv;
          gi = -fma(sc, cd, cc * sd) * inv;
        } else {
          double sn, cn;
          sincos(k * R, &sn, &cn);
          gr = cn * inv;
          gi = -sn * inv;
        }
        if (ext) {
          const double2 f = ek_factor(R, I.a, k);
          const double g0 = gr;
          gr = g0 * f.x - gi * f.y;
          gi = g0 * f.y + gi * f.x;
        }
        sAr += Qr * gr;
        sAi += Qr * gi;
        sPr += Er * gr;
        sPi += Er * gi;
      }
      Ar += Pq * sAr;
      Ai += Pq * sAi;
      Pr += Dq * sPr;
      Pi += Dq * sPi;
    }
  }
  const double td = I.t[0] * J.t[0] + I.t[1] * J.t[1] + I.t[2] * J.t[2];
  return make_double2(-wmu * td * Ai + Pi * inv_weps, wmu * td * Ar - Pr * inv_weps);
}

struct EntryCtx {
  mw_mesh M;
  const int32_t* wing_seg;   // [n_basis][3], -1: absent wing
  const int32_t* wing_slot;  // [n_basis][3]
  double k, wmu, inv_weps;
};

MW_DEV double2 z_entry(const EntryCtx& C, int m, int n, bool table) {
  double zr = 0.0, zi = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int si = C.wing_seg[3 * m + a];
    if (si < 0) continue;
    const int sa = C.wing_slot[3 * m + a];
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      const int sj = C.wing_seg[3 * n + b];
      if (sj < 0) continue;
      const double2 v = band_entry(C.M, si, sa, sj, C.wing_slot[3 * n + b], C.k, C.wmu,
                                   C.inv_weps, table);
      zr += v.x;
      zi += v.y;
    }
  }
  return make_double2(zr, zi);
}

__global__ __launch_bounds__(256) void k_zentries(EntryCtx C, const int32_t* __restrict__ rows,
                                                  const int32_t* __restrict__ cols, int64_t n,
                                                  int table, double2* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[t] = z_entry(C, rows[t], cols[t], table != 0);
}

// Dense blocks: work item (block, first entry) of 256 entries each.
__global__ __launch_bounds__(256) void k_hm_dense(EntryCtx C, const int32_t* __restrict__ rows_pool,
                                                  const int32_t* __restrict__ cols_pool,
                                                  const mw_hblock* __restrict__ blk,
                                                  const int64_t* __restrict__ items, int table,
                                                  double2* __restrict__ store) {
  const int64_t it = items[blockIdx.x];
  const mw_hblock B = blk[it >> 32];
  const int64_t e = (it & 0xffffffffLL) + threadIdx.x;
  if (e >= (int64_t)B.m * B.n) return;
  const int p = (int)(e / B.n), q = (int)(e % B.n);
  store[B.data_off + e] =
      z_entry(C, rows_pool[B.row_off + p], cols_pool[B.col_off + q], table != 0);
}

// ---------------------------------------------------------------------------
// Batched ACA
// ---------------------------------------------------------------------------
// argmax of key over the workgroup of AT threads (ties: smallest index);
// key < 0 never wins
template <int AT>
MW_DEV void wg_argmax(double key, int idx, double* sv, int* si, double& best, int& bi) {
  const int tid = threadIdx.x;
  sv[tid] = key;
  si[tid] = idx;
  __syncthreads();
  for (int s = AT / 2; s > 0; s >>= 1) {
    if (tid < s) {
      const double o = sv[tid + s];
      const int oi = si[tid + s];
      if (o > sv[tid] || (o == sv[tid] && oi < si[tid])) {
        sv[tid] = o;
        si[tid] = oi;
      }
    }
    __syncthreads();
  }
  best = sv[0];
  bi = si[0];
  __syncthreads();
}

template <int AT>
MW_DEV double wg_sum(double v, double* sv) {
  const int tid = threadIdx.x;
  sv[tid] = v;
  __syncthreads();
  for (int s = AT / 2; s > 0; s >>= 1) {
    if (tid < s) sv[tid] += sv[tid + s];
    __syncthreads();
  }
  const double r = sv[0];
  __syncthreads();
  return r;
}

template <int AT>
MW_DEV double2 wg_csum(double2 v, double* sv, double* sw) {
  const int tid = threadIdx.x;
  sv[tid] = v.x;
  sw[tid] = v.y;
  __syncthreads();
  for (int s = AT / 2; s > 0; s >>= 1) {
    if (tid < s) {
      sv[tid] += sv[tid + s];
      sw[tid] += sw[tid + s];
    }
    __syncthreads();
  }
  const double2 r = make_double2(sv[0], sw[0]);
  __syncthreads();
  return r;
}

MW_DEV double cabs(double2 z) { return hypot(z.x, z.y); }

// One admissible block per workgroup: A ~ U V from ~r sampled rows and
// columns (reference _aca.py:137-211, the same pivoting, hop and stopping
// rules): rows and columns in the block's member order, the first row first,
// the column pivot the largest residual entry of the row among unused
// columns (ties: the first), the next row the largest residual entry of the
// column among unused rows; stop when |u| |v| <= tol |U V|_F.  The factors
// land in the block's store as U^T (cap x m) and V (cap x n).  rank on
// return: the factors' rank, or -1 when the block is better stored dense
// (rank (m + n) >= m n) or outgrew its capacity.  AT threads per workgroup:
// 512 for the large blocks, whose sequential rank steps set the launch's
// length, 256 for the rest.
template <int AT>
__global__ __launch_bounds__(AT) void k_hm_aca(EntryCtx C, const int32_t* __restrict__ rows_pool,
                                               const int32_t* __restrict__ cols_pool,
                                               mw_hblock* __restrict__ blk,
                                               const int32_t* __restrict__ list, double tol,
                                               double2* __restrict__ store,
                                               uint8_t* __restrict__ flags) {
  __shared__ double sv[AT], sw[AT];
  __shared__ int si[AT];
  mw_hblock& B = blk[list[blockIdx.x]];
  const int m = B.m, n = B.n, cap = B.cap;
  const int tid = threadIdx.x;
  const int32_t* I = rows_pool + B.row_off;
  const int32_t* Jc = cols_pool + B.col_off;
  double2* Ut = store + B.data_off;       // [cap][m]
  double2* V = Ut + (int64_t)cap * m;     // [cap][n]
  uint8_t* used_r = flags + B.aux_off;
  uint8_t* used_c = used_r + m;
  for (int i = tid; i < m; i += AT) used_r[i] = 0;
  for (int j = tid; j < n; j += AT) used_c[j] = 0;
  __syncthreads();

  const int maxr = m < n ? m : n;
  int rank = 0, istar = 0;
  double approx = 0.0;
  bool overflow = false;
  for (int it = 0; it < maxr; ++it) {
    if (rank == cap) {
      overflow = true;
      break;
    }
    // residual of row istar
    double2* vr = V + (int64_t)rank * n;
    const int gi = I[istar];
    double key = -1.0;
    int kidx = n;
    for (int j = tid; j < n; j += AT) {
      double2 v = z_entry(C, gi, Jc[j], false);
      for (int kk = 0; kk < rank; ++kk) v = cfms(v, Ut[(int64_t)kk * m + istar], V[(int64_t)kk * n + j]);
      vr[j] = v;
      const double a = used_c[j] ? -1.0 : cabs(v);
      if (a > key) {
        key = a;
        kidx = j;
      }
    }
    __syncthreads();
    if (tid == 0) used_r[istar] = 1;
    double best;
    int jb;
    wg_argmax<AT>(key, kidx, sv, si, best, jb);
    const double2 delta = jb < n ? vr[jb] : make_double2(0.0, 0.0);
    if (jb >= n || cabs(delta) < 1e-300 || best <= 0.0) {
      // the residual row vanished: the first unused row, else done
      int first = m;
      for (int i = tid; i < m; i += AT)
        if (!used_r[i] && i < first) first = i;
      double dummy;
      int fi;
      wg_argmax<AT>(first < m ? (double)(m - first) : -1.0, first, sv, si, dummy, fi);
      if (dummy <= 0.0) break;
      istar = fi;
      continue;
    }
    for (int j = tid; j < n; j += AT) vr[j] = cdiv(vr[j], delta);
    // residual of column jb
    double2* ur = Ut + (int64_t)rank * m;
    const int gj = Jc[jb];
    double un2 = 0.0;
    for (int i = tid; i < m; i += AT) {
      double2 u = z_entry(C, I[i], gj, false);
      for (int kk = 0; kk < rank; ++kk) u = cfms(u, V[(int64_t)kk * n + jb], Ut[(int64_t)kk * m + i]);
      ur[i] = u;
      un2 += u.x * u.x + u.y * u.y;
    }
    double vn2 = 0.0;
    for (int j = tid; j < n; j += AT) vn2 += vr[j].x * vr[j].x + vr[j].y * vr[j].y;
    __syncthreads();
    if (tid == 0) used_c[jb] = 1;
    un2 = wg_sum<AT>(un2, sv);
    vn2 = wg_sum<AT>(vn2, sv);
    double cross = 0.0;
    for (int kk = 0; kk < rank; ++kk) {
      double2 a = make_double2(0.0, 0.0), b = make_double2(0.0, 0.0);
      for (int i = tid; i < m; i += AT) {
        const double2 x = Ut[(int64_t)kk * m + i], y = ur[i];
        a.x += x.x * y.x + x.y * y.y;  // conj(x) y
        a.y += x.x * y.y - x.y * y.x;
      }
      for (int j = tid; j < n; j += AT) {
        const double2 x = V[(int64_t)kk * n + j], y = vr[j];
        b.x += x.x * y.x + x.y * y.y;
        b.y += x.x * y.y - x.y * y.x;
      }
      a = wg_csum<AT>(a, sv, sw);
      b = wg_csum<AT>(b, sv, sw);
      cross += a.x * b.x - a.y * b.y;
    }
    const double unvn = sqrt(un2) * sqrt(vn2);
    approx += 2.0 * cross + unvn * unvn;
    ++rank;
    if (approx <= 0.0 || unvn <= tol * sqrt(approx)) break;
    // next row: the largest residual entry of the column among unused rows
    key = -1.0;
    kidx = m;
    for (int i = tid; i < m; i += AT) {
      const double a = used_r[i] ? -1.0 : cabs(ur[i]);
      if (a > key) {
        key = a;
        kidx = i;
      }
    }
    __syncthreads();
    double bc;
    int ib;
    wg_argmax<AT>(key, kidx, sv, si, bc, ib);
    if (ib >= m) break;  // every row used
    istar = ib;
  }
  if (tid == 0) {
    const bool dense = overflow || (int64_t)rank * (m + n) >= (int64_t)m * n;
    B.rank = dense ? -1 : rank;
  }
}

// ---------------------------------------------------------------------------
// Operator product Y = H X in leaf order (X, Y: n x s, row strides ldx/ldy)
// ---------------------------------------------------------------------------
// Leaf order lists the cluster tree's leaves one after another, so every
// cluster is a contiguous range of rows.  Dense blocks are stored in leaf
// order; compressed blocks keep the ACA's member order (ascending basis id),
// reached through two maps: sigma (a row's position among the block's rows
// in that order) and colpos (a column member's leaf-order position).
//
// Phase 1, T = V X[cols]: the V rows of every compressed block that shares
// a column cluster form one product; a workgroup computes 64 of those rows
// x 64 columns, 4 x 4 complex per thread, K in chunks of 16 through LDS.
constexpr int TR = 64, TC = 64, TKC = 16;

__global__ __launch_bounds__(256) void k_hm_t(const mw_hblock* __restrict__ blk,
                                              const mw_hm_titem* __restrict__ items,
                                              const int32_t* __restrict__ t_rows,
                                              const int32_t* __restrict__ colpos,
                                              const double2* __restrict__ store,
                                              const double2* __restrict__ X, int64_t ldx, int s,
                                              double2* __restrict__ T) {
  __shared__ double2 As[TKC][TR];
  __shared__ double2 Bs[TKC][TC];
  __shared__ int64_t vrow[TR];  // complex offset of the V row, -1: padding
  __shared__ int64_t trow[TR];  // row of T
  const mw_hm_titem it = items[blockIdx.x];
  const int c0 = blockIdx.y * TC;
  const int tid = threadIdx.x, rg = tid >> 4, cg = tid & 15;
  if (tid < TR) {
    vrow[tid] = -1;
    if (tid < it.nrows) {
      const int b = t_rows[2 * (it.row0 + tid)], k = t_rows[2 * (it.row0 + tid) + 1];
      const mw_hblock B = blk[b];
      vrow[tid] = B.data_off + (int64_t)B.cap * B.m + (int64_t)k * B.n;
      trow[tid] = B.t_off + k;
    }
  }
  __syncthreads();
  double2 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = make_double2(0.0, 0.0);
  const int K = it.k;
  for (int k0 = 0; k0 < K; k0 += TKC) {
    for (int idx = tid; idx < TR * TKC; idx += 256) {
      const int r = idx / TKC, kk = idx % TKC;
      const int64_t v = vrow[r];
      As[kk][r] = (v >= 0 && k0 + kk < K) ? store[v + k0 + kk] : make_double2(0.0, 0.0);
    }
    for (int idx = tid; idx < TKC * TC; idx += 256) {
      const int kk = idx / TC, c = idx % TC;
      Bs[kk][c] = (k0 + kk < K && c0 + c < s)
                      ? X[(int64_t)colpos[it.col_off + k0 + kk] * ldx + c0 + c]
                      : make_double2(0.0, 0.0);
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < TKC; ++kk) {
      double2 a[4], bv[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[kk][rg + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; ++j) bv[j] = Bs[kk][cg + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          acc[i][j].x = fma(a[i].x, bv[j].x, fma(-a[i].y, bv[j].y, acc[i][j].x));
          acc[i][j].y = fma(a[i].x, bv[j].y, fma(a[i].y, bv[j].x, acc[i][j].y));
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int r = rg + 16 * i;
    if (r >= it.nrows) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int c = c0 + cg + 16 * j;
      if (c < s) T[trow[r] * (int64_t)s + c] = acc[i][j];
    }
  }
}

// Phase 2: 32 leaf-order rows x 64 columns per workgroup, summed over every
// block whose rows hold them (the item's entries: block, offset of the
// item's first row among the block's leaf-order rows), in block order:
// dense D[rows][:] X[cols], compressed U[rows][:] T_b.  4 x 4 complex per
// thread.
constexpr int RR = 32, RC = 64, RKC = 16;

__global__ __launch_bounds__(128) void k_hm_r(const mw_hblock* __restrict__ blk,
                                              const mw_hm_ritem* __restrict__ items,
                                              const int32_t* __restrict__ ent,
                                              const int32_t* __restrict__ sigma,
                                              const double2* __restrict__ store,
                                              const double2* __restrict__ X, int64_t ldx,
                                              const double2* __restrict__ T, int s,
                                              double2* __restrict__ Y, int64_t ldy) {
  __shared__ double2 As[RKC][RR];
  __shared__ double2 Bs[RKC][RC];
  const mw_hm_ritem it = items[blockIdx.x];
  const int c0 = blockIdx.y * RC;
  const int tid = threadIdx.x, rg = tid >> 4, cg = tid & 15;
  double2 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = make_double2(0.0, 0.0);
  for (int e = it.ent0; e < it.ent0 + it.nent; ++e) {
    const mw_hblock B = blk[ent[2 * e]];
    const int ro = ent[2 * e + 1];
    const bool dense = B.rank < 0;
    const int K = dense ? B.n : B.rank;
    const double2* D = store + B.data_off;
    for (int k0 = 0; k0 < K; k0 += RKC) {
      for (int idx = tid; idx < RR * RKC; idx += 128) {
        double2 v = make_double2(0.0, 0.0);
        if (dense) {
          const int t = idx / RKC, kk = idx % RKC;
          if (t < it.nrows && k0 + kk < K) v = D[(int64_t)(ro + t) * B.n + k0 + kk];
          As[kk][t] = v;
        } else {
          const int kk = idx / RR, t = idx % RR;
          if (t < it.nrows && k0 + kk < K)
            v = D[(int64_t)(k0 + kk) * B.m + sigma[B.row_off + ro + t]];
          As[kk][t] = v;
        }
      }
      for (int idx = tid; idx < RKC * RC; idx += 128) {
        const int kk = idx / RC, c = idx % RC;
        double2 v = make_double2(0.0, 0.0);
        if (k0 + kk < K && c0 + c < s)
          v = dense ? X[(B.col_start + k0 + kk) * ldx + c0 + c]
                    : T[(B.t_off + k0 + kk) * (int64_t)s + c0 + c];
        Bs[kk][c] = v;
      }
      __syncthreads();
#pragma unroll 4
      for (int kk = 0; kk < RKC; ++kk) {
        double2 a[4], bv[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = As[kk][rg + 8 * i];
#pragma unroll
        for (int j = 0; j < 4; ++j) bv[j] = Bs[kk][cg + 16 * j];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            acc[i][j].x = fma(a[i].x, bv[j].x, fma(-a[i].y, bv[j].y, acc[i][j].x));
            acc[i][j].y = fma(a[i].x, bv[j].y, fma(a[i].y, bv[j].x, acc[i][j].y));
          }
      }
      __syncthreads();
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int t = rg + 8 * i;
    if (t >= it.nrows) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int c = c0 + cg + 16 * j;
      if (c < s) Y[(it.row0 + t) * ldy + c] = acc[i][j];
    }
  }
}

// ---------------------------------------------------------------------------
// Batched inverses of square dense blocks (n <= 64): Gauss-Jordan with
// partial pivoting on [A | I] in the workgroup's scratch (2 n^2 complex,
// work + blk.aux_off complex), the inverse written over the block's store.
// info[b] = 1 when a pivot is exactly zero.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_hm_inv(const mw_hblock* __restrict__ blk,
                                                const int32_t* __restrict__ list,
                                                double2* __restrict__ store,
                                                double2* __restrict__ work,
                                                int32_t* __restrict__ info) {
  __shared__ double sv[256];
  __shared__ int si[256];
  __shared__ double2 piv;
  const int b = list[blockIdx.x];
  const mw_hblock B = blk[b];
  const int n = B.m, w = 2 * n, tid = threadIdx.x;
  double2* A = work + B.aux_off;  // [n][2n]
  double2* D = store + B.data_off;
  for (int e = tid; e < n * w; e += 256) {
    const int r = e / w, c = e % w;
    A[e] = c < n ? D[(int64_t)r * n + c] : make_double2(c - n == r ? 1.0 : 0.0, 0.0);
  }
  __syncthreads();
  int bad = 0;
  for (int col = 0; col < n; ++col) {
    double key = -1.0;
    int kr = n;
    for (int r = col + tid; r < n; r += 256) {
      const double a = cabs1(A[(int64_t)r * w + col]);
      if (a > key) {
        key = a;
        kr = r;
      }
    }
    sv[tid] = key;
    si[tid] = kr;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
      if (tid < st) {
        if (sv[tid + st] > sv[tid] || (sv[tid + st] == sv[tid] && si[tid + st] < si[tid])) {
          sv[tid] = sv[tid + st];
          si[tid] = si[tid + st];
        }
      }
      __syncthreads();
    }
    const int pr = si[0];
    if (sv[0] <= 0.0) {
      bad = 1;
      break;
    }
    __syncthreads();
    if (pr != col)
      for (int c = tid; c < w; c += 256) {
        const double2 t = A[(int64_t)col * w + c];
        A[(int64_t)col * w + c] = A[(int64_t)pr * w + c];
        A[(int64_t)pr * w + c] = t;
      }
    __syncthreads();
    if (tid == 0) piv = crecip(A[(int64_t)col * w + col]);
    __syncthreads();
    for (int c = tid; c < w; c += 256) A[(int64_t)col * w + c] = cmul(A[(int64_t)col * w + c], piv);
    __syncthreads();
    for (int e = tid; e < n * w; e += 256) {
      const int r = e / w, c = e % w;
      if (r == col || c == col) continue;
      A[e] = cfms(A[e], A[(int64_t)r * w + col], A[(int64_t)col * w + c]);
    }
    __syncthreads();
    for (int r = tid; r < n; r += 256)
      if (r != col) A[(int64_t)r * w + col] = make_double2(0.0, 0.0);
    __syncthreads();
  }
  if (tid == 0) info[blockIdx.x] = bad;
  if (bad) return;
  for (int e = tid; e < n * n; e += 256) {
    const int r = e / n, c = e % n;
    D[e] = A[(int64_t)r * w + n + c];
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Host launchers (called from mw_capi.hip)
// ---------------------------------------------------------------------------
static EntryCtx entry_ctx(const mw_mesh& M, const int32_t* wing_seg, const int32_t* wing_slot,
                          double k, double omega, double eps, double mu) {
  EntryCtx C;
  C.M = M;
  C.wing_seg = wing_seg;
  C.wing_slot = wing_slot;
  C.k = k;
  C.wmu = omega * mu;
  C.inv_weps = 1.0 / (omega * eps);
  return C;
}

hipError_t mw_launch_zentries(const mw_mesh& M, const int32_t* wing_seg, const int32_t* wing_slot,
                              double k, double omega, double eps, double mu, const int32_t* rows,
                              const int32_t* cols, int64_t n, int table, void* out,
                              hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_zentries, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                     entry_ctx(M, wing_seg, wing_slot, k, omega, eps, mu), rows, cols, n, table,
                     reinterpret_cast<double2*>(out));
  return hipGetLastError();
}

hipError_t mw_launch_hm_dense(const mw_mesh& M, const int32_t* wing_seg, const int32_t* wing_slot,
                              double k, double omega, double eps, double mu,
                              const int32_t* rows_pool, const int32_t* cols_pool,
                              const mw_hblock* blk, const int64_t* items, int64_t n_items,
                              int table, void* store, hipStream_t st) {
  if (n_items <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_hm_dense, dim3((unsigned)n_items), dim3(256), 0, st,
                     entry_ctx(M, wing_seg, wing_slot, k, omega, eps, mu), rows_pool, cols_pool,
                     blk, items, table, reinterpret_cast<double2*>(store));
  return hipGetLastError();
}

// list[0 .. n_large) go to 512-thread workgroups, the rest to 256-thread ones.
hipError_t mw_launch_hm_aca(const mw_mesh& M, const int32_t* wing_seg, const int32_t* wing_slot,
                            double k, double omega, double eps, double mu,
                            const int32_t* rows_pool, const int32_t* cols_pool, mw_hblock* blk,
                            const int32_t* list, int32_t n_list, int32_t n_large, double tol,
                            void* store, void* flags, hipStream_t st) {
  if (n_list <= 0) return hipSuccess;
  const EntryCtx C = entry_ctx(M, wing_seg, wing_slot, k, omega, eps, mu);
  if (n_large > 0)
    hipLaunchKernelGGL(k_hm_aca<512>, dim3((unsigned)n_large), dim3(512), 0, st, C, rows_pool,
                       cols_pool, blk, list, tol, reinterpret_cast<double2*>(store),
                       reinterpret_cast<uint8_t*>(flags));
  if (n_list > n_large)
    hipLaunchKernelGGL(k_hm_aca<256>, dim3((unsigned)(n_list - n_large)), dim3(256), 0, st, C,
                       rows_pool, cols_pool, blk, list + n_large, tol,
                       reinterpret_cast<double2*>(store), reinterpret_cast<uint8_t*>(flags));
  return hipGetLastError();
}

hipError_t mw_launch_hm_apply(const mw_hblock* blk, const void* store, const mw_hm_titem* t_items,
                              int32_t n_t, const int32_t* t_rows, const int32_t* colpos,
                              const mw_hm_ritem* r_items, int32_t n_r, const int32_t* r_ent,
                              const int32_t* sigma, const void* X, int64_t ldx, int32_t s,
                              void* T, void* Y, int64_t ldy, hipStream_t st) {
  if (s <= 0) return hipSuccess;
  if (n_t > 0)
    hipLaunchKernelGGL(k_hm_t, dim3((unsigned)n_t, (unsigned)((s + TC - 1) / TC)), dim3(256), 0,
                       st, blk, t_items, t_rows, colpos, reinterpret_cast<const double2*>(store),
                       reinterpret_cast<const double2*>(X), ldx, (int)s,
                       reinterpret_cast<double2*>(T));
  if (n_r > 0)
    hipLaunchKernelGGL(k_hm_r, dim3((unsigned)n_r, (unsigned)((s + RC - 1) / RC)), dim3(128), 0,
                       st, blk, r_items, r_ent, sigma, reinterpret_cast<const double2*>(store),
                       reinterpret_cast<const double2*>(X), ldx,
                       reinterpret_cast<const double2*>(T), (int)s, reinterpret_cast<double2*>(Y),
                       ldy);
  return hipGetLastError();
}

hipError_t mw_launch_hm_invert(const mw_hblock* blk, const int32_t* list, int32_t n_list,
                               void* store, void* work, int32_t* info, hipStream_t st) {
  if (n_list <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_hm_inv, dim3((unsigned)n_list), dim3(256), 0, st, blk, list,
                     reinterpret_cast<double2*>(store), reinterpret_cast<double2*>(work), info);
  return hipGetLastError();
}
