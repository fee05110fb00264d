// Dense complex linear algebra for the port solve on gfx950: blocked LU with
// tournament row pivoting, multi-RHS triangular solves, complex GEMM on MFMA,
// and the small kernels of the port algebra.  Templated on the real type
// (double: FP64 path, float: the FP32 factor of the mixed path).
//
// Storage: row-major, leading dimension in complex elements.
#include "mw_common.h"

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int NB = 32;   // LU panel width
constexpr int CH = 64;   // rows per tournament leaf (= 2 NB: a merge node's size)

// ---------------------------------------------------------------------------
// Complex GEMM  C <- C - A B  on MFMA.
//
// A complex product is a real product of twice the size:
//   C_r (m x 2n) -= A_r (m x 2k) . B_r (2k x 2n),
// where A_r / C_r are A / C's memory read as reals (re, im interleaved along
// the row) and B_r expands each complex b into [[br, bi], [-bi, br]].  The
// expansion happens when B fragments are formed from LDS, so B is staged
// once as complex.  One workgroup computes a 64 x 64 block of C_r (64 rows x
// 32 complex columns) with four waves in a 2 x 2 arrangement, each wave a
// 2 x 2 grid of 16 x 16 MFMA tiles.
// ---------------------------------------------------------------------------
template <typename T> struct Mfma;
template <> struct Mfma<double> {
  typedef v4d acc_t;
  static MW_DEV acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // output element r of lane l sits at row (l/16) + 4 r, column l % 16
  static MW_DEV int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <> struct Mfma<float> {
  typedef v4f acc_t;
  static MW_DEV acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // output element r of lane l sits at row 4 (l/16) + r, column l % 16
  static MW_DEV int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

constexpr int GBM = 64;      // C rows per workgroup
constexpr int GBN = 32;      // C complex columns per workgroup (64 real)
constexpr int GKC = 8;       // complex k per LDS stage (16 real)

template <typename T>
__global__ __launch_bounds__(256) void k_gemm_sub(int64_t m, int64_t n, int64_t kdim,
                                                  const typename cplx_t<T>::type* __restrict__ A,
                                                  int64_t lda,
                                                  const typename cplx_t<T>::type* __restrict__ B,
                                                  int64_t ldb, typename cplx_t<T>::type* C,
                                                  int64_t ldc) {
  typedef typename cplx_t<T>::type CT;
  typedef typename Mfma<T>::acc_t acc_t;
  __shared__ T As[GBM][2 * GKC + 1];
  __shared__ CT Bs[GKC][GBN + 1];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;           // wave's 32x32 (real) quadrant
  const int64_t row0 = (int64_t)blockIdx.y * GBM;
  const int64_t col0 = (int64_t)blockIdx.x * GBN;    // complex columns

  acc_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  for (int64_t k0 = 0; k0 < kdim; k0 += GKC) {
    // stage A: 64 rows x 8 complex
    for (int e = tid; e < GBM * GKC; e += 256) {
      const int r = e / GKC, c = e % GKC;
      CT v;
      v.x = 0; v.y = 0;
      if (row0 + r < m && k0 + c < kdim) v = A[(row0 + r) * lda + k0 + c];
      As[r][2 * c] = v.x;
      As[r][2 * c + 1] = v.y;
    }
    // stage B: 8 complex rows x 32 complex columns
    {
      const int r = tid / GBN, c = tid % GBN;
      CT v;
      v.x = 0; v.y = 0;
      if (k0 + r < kdim && col0 + c < n) v = B[(k0 + r) * ldb + col0 + c];
      Bs[r][c] = v;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < 2 * GKC; ks += 4) {
      const int kr = ks + (lane >> 4);       // real k of this lane
      const int kc = kr >> 1, kpar = kr & 1; // complex row of B, re/im half
      T a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) a[i] = As[wr * 32 + i * 16 + (lane & 15)][kr];
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int cr = wc * 32 + j * 16 + (lane & 15);  // real column
        const CT bv = Bs[kc][cr >> 1];
        // B_r[2l+s][2j+t] = s==0 ? (t==0 ? br : bi) : (t==0 ? -bi : br)
        b[j] = (kpar == 0) ? ((cr & 1) ? bv.y : bv.x) : ((cr & 1) ? bv.x : -bv.y);
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = Mfma<T>::mma(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
  // C_r -= acc
  T* Cr = reinterpret_cast<T*>(C);
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gr = row0 + wr * 32 + i * 16 + Mfma<T>::row(lane, r);
        const int64_t gcr = 2 * col0 + wc * 32 + j * 16 + (lane & 15);  // real column
        if (gr < m && gcr < 2 * n) Cr[gr * 2 * ldc + gcr] -= acc[i][j][r];
      }
}

template <typename T>
hipError_t launch_gemm_sub(int64_t m, int64_t n, int64_t k, const void* A, int64_t lda,
                           const void* B, int64_t ldb, void* C, int64_t ldc, hipStream_t st) {
  typedef typename cplx_t<T>::type CT;
  if (m <= 0 || n <= 0 || k <= 0) return hipSuccess;
  dim3 grid((unsigned)((n + GBN - 1) / GBN), (unsigned)((m + GBM - 1) / GBM));
  hipLaunchKernelGGL(k_gemm_sub<T>, grid, dim3(256), 0, st, m, n, k, (const CT*)A, lda,
                     (const CT*)B, ldb, (CT*)C, ldc);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Tournament pivot selection for one 32-column panel (columns k0 .. k0+bw).
//
// Leaves: each workgroup takes 128 rows of the panel, factors them with
// partial pivoting and nominates its bw pivot rows.  Merges: a workgroup
// reloads the ORIGINAL rows nominated by up to four siblings (<= 128 rows),
// factors them the same way and nominates again; one kernel launch per tree
// level, so kernel boundaries carry the nominees between levels.  The last
// level's nominees are the panel's pivot rows, and its factorisation of them
// is L11\U11.  The matrix is read-only here.
//
// The 128 x 32 block lives in registers: wave w holds columns 8w..8w+7, lane
// l holds rows l and l+64.  Rows are never moved: a pivoted row is masked out
// of later searches and updates, so one barrier per column suffices (the
// owner wave publishes the pivot and the multipliers through LDS).
// ---------------------------------------------------------------------------
constexpr int RPL = 2;              // rows per lane
constexpr int TROWS = 64 * RPL;     // rows per tournament block
constexpr int FANIN = TROWS / NB;   // nominee sets merged per block

MW_DEV double bcast_d(double x, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(x), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(x), lane);
  return __hiloint2double(hi, lo);
}
MW_DEV float bcast_f(float x, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), lane));
}
MW_DEV double2 bcast(double2 v, int lane) { return make_double2(bcast_d(v.x, lane), bcast_d(v.y, lane)); }
MW_DEV float2 bcast(float2 v, int lane) { return make_float2(bcast_f(v.x, lane), bcast_f(v.y, lane)); }

template <typename CT>
struct TourLds {
  CT mult[2][TROWS];     // multipliers of the current column (double-buffered)
  int p[2];              // pivot row of the current column
  int piv[NB];           // pivot row (block-local) of each column
  int rank[TROWS];       // column at which a block-local row was pivoted
  int32_t rid[TROWS];    // global row id of each block-local row
  int rows;
};

// Factor the register block; on return L.piv[0..steps) holds the pivot rows.
template <typename CT>
MW_DEV void reg_gepp(CT (&S)[RPL][8], int rows, int cols, TourLds<CT>& L, int32_t* info,
                     int64_t k0) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int steps = rows < cols ? rows : cols;
  unsigned done = 0;  // bit t: row lane + 64 t already pivoted
  for (int c = 0; c < steps; ++c) {
    const int buf = c & 1;
    if (w == (c >> 3)) {
      const int jj = c & 7;
      // Magnitudes rank NaN as +inf and rows without a candidate as -1 with
      // index -1, so the reduction always returns an unpivoted row of this
      // block (steps <= rows guarantees one exists) whatever the data.
      double best = -1.0;
      int bidx = -1;
      // column c of my rows (register select: no dynamic indexing)
      // (masked sums, not selects: the compiler must not turn a uniform
      // register choice into an indexed access, which would move S to scratch)
      CT colv[RPL];
#pragma unroll
      for (int t = 0; t < RPL; ++t) {
        colv[t].x = 0; colv[t].y = 0;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          colv[t].x += (j == jj) ? S[t][j].x : 0;
          colv[t].y += (j == jj) ? S[t][j].y : 0;
        }
      }
#pragma unroll
      for (int t = 0; t < RPL; ++t) {
        const int r = lane + 64 * t;
        if (r < rows && !((done >> t) & 1)) {
          double v = cabs1(colv[t]);
          if (v != v) v = HUGE_VAL;
          if (bidx < 0 || v > best) { best = v; bidx = r; }
        }
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(best, off);
        const int oi = __shfl_xor(bidx, off);
        if (oi >= 0 && (bidx < 0 || ov > best || (ov == best && oi < bidx))) {
          best = ov;
          bidx = oi;
        }
      }
      int p = __builtin_amdgcn_readfirstlane(bidx);
      p = p < 0 ? 0 : (p >= rows ? rows - 1 : p);
      const int pl = p & 63, pt = p >> 6;
      CT src;
      src.x = 0; src.y = 0;
#pragma unroll
      for (int t = 0; t < RPL; ++t) {
        src.x += (t == pt) ? colv[t].x : 0;
        src.y += (t == pt) ? colv[t].y : 0;
      }
      const CT apc = bcast(src, pl);
      const bool zero = (apc.x == 0 && apc.y == 0);
      if (zero && info && lane == 0 && *info == 0) *info = (int32_t)(k0 + c + 1);
      CT inv;
      inv.x = 0; inv.y = 0;
      if (!zero) inv = crecip(apc);
#pragma unroll
      for (int t = 0; t < RPL; ++t) {
        const int r = lane + 64 * t;
        CT mlt;
        mlt.x = 0; mlt.y = 0;
        if (r < rows && !((done >> t) & 1) && r != p) {
          mlt = cmul(colv[t], inv);
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            S[t][j].x = (j == jj) ? mlt.x : S[t][j].x;
            S[t][j].y = (j == jj) ? mlt.y : S[t][j].y;
          }
        }
        L.mult[buf][r] = mlt;
      }
      if (lane == 0) {
        L.p[buf] = p;
        L.piv[c] = p;
        L.rank[p] = c;
      }
    }
    __syncthreads();
    const int p = __builtin_amdgcn_readfirstlane(L.p[buf]);
    const int pl = p & 63, pt = p >> 6;
    if (lane == pl) done |= 1u << pt;
    CT u[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      CT src;
      src.x = 0; src.y = 0;
#pragma unroll
      for (int t = 0; t < RPL; ++t) {
        src.x += (t == pt) ? S[t][j].x : 0;
        src.y += (t == pt) ? S[t][j].y : 0;
      }
      u[j] = bcast(src, pl);
    }
#pragma unroll
    for (int t = 0; t < RPL; ++t) {
      const CT mlt = L.mult[buf][lane + 64 * t];
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (8 * w + j > c) S[t][j] = cfms(S[t][j], mlt, u[j]);
    }
  }
}

// Load rows L.rid[0..rows) (global ids) of panel columns k0..k0+bw.
template <typename CT>
MW_DEV void load_block(CT (&S)[RPL][8], const CT* A, int64_t lda, int64_t k0, int bw,
                       const TourLds<CT>& L, int rows) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
#pragma unroll
  for (int t = 0; t < RPL; ++t) {
    const int r = lane + 64 * t;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int c = 8 * w + j;
      CT v;
      v.x = 0; v.y = 0;
      if (r < rows && c < bw) v = A[(int64_t)L.rid[r] * lda + k0 + c];
      S[t][j] = v;
    }
  }
}

// Root: write L11\U11 of the pivot rows in pivot order and the LAPACK-style
// interchange sequence.
template <typename CT>
MW_DEV void finish_root(const CT (&S)[RPL][8], int rows, int bw, TourLds<CT>& L, CT* lu11,
                        int32_t* ipiv, int64_t k0, int32_t* info) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int steps = rows < bw ? rows : bw;
#pragma unroll
  for (int t = 0; t < RPL; ++t) {
    const int r = lane + 64 * t;
    if (r < rows) {
      bool pivoted = false;
      for (int i = 0; i < steps; ++i)
        if (L.piv[i] == r) pivoted = true;
      if (pivoted) {
        const int i = L.rank[r];
#pragma unroll
        for (int j = 0; j < 8; ++j)
          if (8 * w + j < bw) lu11[i * NB + 8 * w + j] = S[t][j];
      }
    }
  }
  if (tid == 0) {
    __shared__ int32_t pos_s[2 * NB], val_s[2 * NB];
    int nt = 0;
    for (int i = 0; i < steps; ++i) {
      const int32_t g = L.rid[L.piv[i]];
      int32_t pg = g;
      for (int t = 0; t < nt; ++t)
        if (val_s[t] == g) { pg = pos_s[t]; break; }
      const int32_t p1 = (int32_t)(k0 + i);
      int32_t id1 = p1;
      int t1 = -1, t2 = -1;
      for (int t = 0; t < nt; ++t) {
        if (pos_s[t] == p1) { id1 = val_s[t]; t1 = t; }
        if (pos_s[t] == pg) t2 = t;
      }
      ipiv[p1] = pg;
      if (pg != p1) {
        if (t1 < 0) { t1 = nt; pos_s[nt] = p1; ++nt; }
        val_s[t1] = g;
        if (t2 < 0) { t2 = nt; pos_s[nt] = pg; ++nt; }
        val_s[t2] = id1;
      }
    }
    if (steps < bw && info && *info == 0) *info = (int32_t)(k0 + steps + 1);
  }
}

// level < 0: leaf over the panel rows; otherwise merge nominee sets of the
// previous level.  is_root: this launch has one block and finishes the panel.
template <typename T>
__global__ __launch_bounds__(256) void k_tournament(const typename cplx_t<T>::type* A, int64_t lda,
                                                    int64_t n, int64_t k0, int bw, int leaf,
                                                    const int32_t* in_cand, const int32_t* in_n,
                                                    int n_in, int32_t* out_cand, int32_t* out_n,
                                                    int is_root, void* lu11, int32_t* ipiv,
                                                    int32_t* info) {
  typedef typename cplx_t<T>::type CT;
  __shared__ TourLds<CT> L;
  const int tid = threadIdx.x;
  const int b = blockIdx.x;
  if (tid == 0) {
    int rows = 0;
    if (leaf) {
      const int64_t r0 = k0 + (int64_t)b * TROWS;
      const int64_t left = n - r0;
      rows = (int)(left < TROWS ? left : TROWS);
      for (int r = 0; r < rows; ++r) L.rid[r] = (int32_t)(r0 + r);
    } else {
      for (int s = 0; s < FANIN; ++s) {
        const int src = b * FANIN + s;
        if (src >= n_in) break;
        const int cnt = in_n[src];
        for (int t = 0; t < cnt; ++t) L.rid[rows++] = in_cand[src * NB + t];
      }
    }
    L.rows = rows;
  }
  __syncthreads();
  const int rows = L.rows;
  CT S[RPL][8];
  load_block(S, A, lda, k0, bw, L, rows);
  reg_gepp(S, rows, bw, L, is_root ? info : (int32_t*)nullptr, k0);
  __syncthreads();
  const int steps = rows < bw ? rows : bw;
  if (is_root) {
    finish_root(S, rows, bw, L, reinterpret_cast<CT*>(lu11), ipiv, k0, info);
  } else {
    if (tid < steps) out_cand[b * NB + tid] = L.rid[L.piv[tid]];
    if (tid == 0) out_n[b] = steps;
  }
}

// Row interchanges ipiv[k0 .. k0+bw) on columns [c0, c1).
template <typename T>
__global__ void k_laswp(typename cplx_t<T>::type* A, int64_t lda, int64_t k0, int bw,
                        const int32_t* ipiv, int64_t c0, int64_t c1) {
  typedef typename cplx_t<T>::type CT;
  const int64_t c = c0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= c1) return;
  for (int i = 0; i < bw; ++i) {
    const int64_t p1 = k0 + i, p2 = ipiv[k0 + i];
    if (p2 != p1) {
      const CT t = A[p1 * lda + c];
      A[p1 * lda + c] = A[p2 * lda + c];
      A[p2 * lda + c] = t;
    }
  }
}

// Diagonal blocks are inverted once per panel (L11^{-1}, U11^{-1}, padded to
// NB with the identity) and applied as small LDS-tiled products: the usual
// GPU substitute for a triangular solve with many right-hand sides.
template <typename T>
__global__ __launch_bounds__(64) void k_tri_inv(const typename cplx_t<T>::type* lu, int64_t ldl,
                                                int bw, typename cplx_t<T>::type* out) {
  typedef typename cplx_t<T>::type CT;
  __shared__ CT M[NB][NB + 1], Li[NB][NB + 1], Ui[NB][NB + 1];
  const int tid = threadIdx.x;
  for (int e = tid; e < NB * NB; e += blockDim.x) {
    const int i = e / NB, j = e % NB;
    CT v;
    v.x = (i == j) ? 1 : 0; v.y = 0;
    if (i < bw && j < bw) v = lu[(int64_t)i * ldl + j];
    M[i][j] = v;
  }
  __syncthreads();
  if (tid < NB) {
    // column j of L^{-1}: L x = e_j, x_i = -sum_{l=j}^{i-1} L[i][l] x_l
    const int j = tid;
    for (int i = 0; i < NB; ++i) {
      CT v;
      v.x = (i == j) ? 1 : 0; v.y = 0;
      if (i > j) {
        for (int l = j; l < i; ++l) v = cfms(v, M[i][l], Li[l][j]);
      }
      Li[i][j] = v;
    }
  } else if (tid < 2 * NB) {
    // column j of U^{-1}: U x = e_j, back substitution
    const int j = tid - NB;
    for (int i = NB - 1; i >= 0; --i) {
      CT v;
      v.x = 0; v.y = 0;
      if (i <= j) {
        if (i == j) v.x = 1;
        for (int l = i + 1; l <= j; ++l) v = cfms(v, M[i][l], Ui[l][j]);
        // an exactly zero pivot (already reported through info) maps to a
        // zero inverse entry rather than inf/NaN
        if (M[i][i].x != 0 || M[i][i].y != 0) v = cdiv(v, M[i][i]);
        else { v.x = 0; v.y = 0; }
      }
      Ui[i][j] = v;
    }
  }
  __syncthreads();
  for (int e = tid; e < NB * NB; e += blockDim.x) {
    out[e] = Li[e / NB][e % NB];
    out[NB * NB + e] = Ui[e / NB][e % NB];
  }
}

constexpr int AW = 64;  // rows (or columns) per workgroup in the block products

template <typename CT>
struct ApplyLds {
  CT Tm[NB][NB + 1];
  CT B[AW][NB + 1];     // rows mode: AW rows x NB; cols mode: used as [NB][AW] below
};

// out[i][c] = sum_l Tm[i][l] X[k0+l][c] for the NB x AW block at columns c0.
template <typename CT>
MW_DEV void apply_left(ApplyLds<CT>& L, CT* X, int64_t ldx, int64_t k0, int bw, int64_t c0,
                       int64_t ncols) {
  const int tid = threadIdx.x;
  CT (*Bc)[AW + 1] = reinterpret_cast<CT (*)[AW + 1]>(&L.B[0][0]);
  for (int e = tid; e < NB * AW; e += blockDim.x) {
    const int i = e / AW, cc = e % AW;
    CT v;
    v.x = 0; v.y = 0;
    if (i < bw && c0 + cc < ncols) v = X[(k0 + i) * ldx + c0 + cc];
    Bc[i][cc] = v;
  }
  __syncthreads();
  const int cc = tid % AW, i0 = (tid / AW) * 8;
  CT acc[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) { acc[q].x = 0; acc[q].y = 0; }
#pragma unroll 4
  for (int l = 0; l < NB; ++l) {
    const CT b = Bc[l][cc];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const CT t = L.Tm[i0 + q][l];
      acc[q].x += t.x * b.x - t.y * b.y;
      acc[q].y += t.x * b.y + t.y * b.x;
    }
  }
  if (c0 + cc < ncols) {
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (i0 + q < bw) X[(k0 + i0 + q) * ldx + c0 + cc] = acc[q];
  }
}

// After the interchanges: A11 <- L11\U11, L21 <- A21 U11^{-1}, U12 <- L11^{-1} A12.
// Blocks [0, nrb): 64-row blocks of L21; [nrb, nrb+ncb): 64-column blocks of
// U12; last: A11.
template <typename T>
__global__ __launch_bounds__(256) void k_panel_apply(int64_t n, typename cplx_t<T>::type* A,
                                                     int64_t lda, int64_t k0, int bw,
                                                     const typename cplx_t<T>::type* inv,
                                                     const typename cplx_t<T>::type* lu11,
                                                     int nrb, int ncb) {
  typedef typename cplx_t<T>::type CT;
  __shared__ ApplyLds<CT> L;
  const int tid = threadIdx.x;
  const int b = blockIdx.x;
  if (b < nrb) {
    const CT* Ui = inv + NB * NB;
    for (int e = tid; e < NB * NB; e += blockDim.x) L.Tm[e / NB][e % NB] = Ui[e];
    const int64_t r0 = k0 + bw + (int64_t)b * AW;
    for (int e = tid; e < AW * NB; e += blockDim.x) {
      const int rr = e / NB, l = e % NB;
      CT v;
      v.x = 0; v.y = 0;
      if (r0 + rr < n && l < bw) v = A[(r0 + rr) * lda + k0 + l];
      L.B[rr][l] = v;
    }
    __syncthreads();
    const int rr = tid / 4, j0 = (tid % 4) * 8;
    CT acc[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) { acc[q].x = 0; acc[q].y = 0; }
#pragma unroll 4
    for (int l = 0; l < NB; ++l) {
      const CT a = L.B[rr][l];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const CT t = L.Tm[l][j0 + q];
        acc[q].x += a.x * t.x - a.y * t.y;
        acc[q].y += a.x * t.y + a.y * t.x;
      }
    }
    if (r0 + rr < n) {
#pragma unroll
      for (int q = 0; q < 8; ++q)
        if (j0 + q < bw) A[(r0 + rr) * lda + k0 + j0 + q] = acc[q];
    }
  } else if (b < nrb + ncb) {
    for (int e = tid; e < NB * NB; e += blockDim.x) L.Tm[e / NB][e % NB] = inv[e];
    __syncthreads();
    apply_left(L, A, lda, k0, bw, k0 + bw + (int64_t)(b - nrb) * AW, n);
  } else {
    for (int e = tid; e < bw * bw; e += blockDim.x)
      A[(k0 + e / bw) * lda + k0 + e % bw] = lu11[(e / bw) * NB + e % bw];
  }
}

// Solve blocks on AW right-hand-side columns per workgroup.
// fwd: interchanges of the panel, then X_p <- L11^{-1} X_p; bwd: X_p <- U11^{-1} X_p.
template <typename T>
__global__ __launch_bounds__(256) void k_solve_apply(int64_t nrhs, const typename cplx_t<T>::type* inv,
                                                     int64_t k0, int bw, const int32_t* ipiv,
                                                     typename cplx_t<T>::type* X, int64_t ldx,
                                                     int fwd) {
  typedef typename cplx_t<T>::type CT;
  __shared__ ApplyLds<CT> L;
  const int tid = threadIdx.x;
  const int64_t c0 = (int64_t)blockIdx.x * AW;
  const CT* Tsrc = fwd ? inv : inv + NB * NB;
  for (int e = tid; e < NB * NB; e += blockDim.x) L.Tm[e / NB][e % NB] = Tsrc[e];
  if (fwd && tid < AW && c0 + tid < nrhs) {
    const int64_t c = c0 + tid;
    for (int i = 0; i < bw; ++i) {
      const int64_t p1 = k0 + i, p2 = ipiv[k0 + i];
      if (p2 != p1) {
        const CT t = X[p1 * ldx + c]; X[p1 * ldx + c] = X[p2 * ldx + c]; X[p2 * ldx + c] = t;
      }
    }
  }
  __syncthreads();
  apply_left(L, X, ldx, k0, bw, c0, nrhs);
}

// Zero a small int32 buffer (used instead of hipMemsetAsync so the whole
// factorisation can be captured in a graph).
__global__ void k_zero_i32(int32_t* p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------------------
// Port algebra helpers (FP64)
// ---------------------------------------------------------------------------
__global__ void k_sparse_rows(int64_t nrows, int maxnz, const int32_t* idx, const double* val,
                              const double2* X, int64_t ldx, int64_t ncols, double2* out,
                              int64_t ldo) {
  const int64_t i = blockIdx.y;
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrows || j >= ncols) return;
  double sr = 0.0, si = 0.0;
  for (int t = 0; t < maxnz; ++t) {
    const int32_t r = idx[i * maxnz + t];
    if (r < 0) continue;
    const double v = val[i * maxnz + t];
    const double2 x = X[(int64_t)r * ldx + j];
    sr += v * x.x;
    si += v * x.y;
  }
  out[i * ldo + j] = make_double2(sr, si);
}

__global__ void k_zero_rows(int64_t n, int64_t ncols, double2* X, int64_t ldx) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * ncols) return;
  X[(e / ncols) * ldx + e % ncols] = make_double2(0.0, 0.0);
}

__global__ void k_scatter_cols(int64_t ncols, int maxnz, const int32_t* idx, const double* val,
                               double2* X, int64_t ldx) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ncols * maxnz) return;
  const int64_t c = e / maxnz;
  const int32_t r = idx[e];
  if (r < 0) return;
  X[(int64_t)r * ldx + c] = make_double2(val[e], 0.0);
}

// T <- S^{-1} T, one workgroup, Gauss-Jordan with partial pivoting in LDS
// (ns <= 32, nt <= 64).
__global__ __launch_bounds__(256) void k_small_solve(int ns, int nt, double2* S, int64_t lds,
                                                     double2* Tm, int64_t ldt, int32_t* info) {
  __shared__ double2 M[32][97];
  __shared__ int piv_s;
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = ns + nt;
  for (int e = tid; e < ns * w; e += blockDim.x) {
    const int r = e / w, c = e % w;
    M[r][c] = c < ns ? S[r * lds + c] : Tm[r * ldt + (c - ns)];
  }
  __syncthreads();
  for (int c = 0; c < ns; ++c) {
    if (tid < 64) {
      const bool cand = lane >= c && lane < ns;
      double v = cand ? cabs1(M[lane][c]) : -1.0;
      if (v != v) v = HUGE_VAL;
      int idx = cand ? lane : -1;
      for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(v, off);
        const int oi = __shfl_xor(idx, off);
        if (oi >= 0 && (idx < 0 || ov > v || (ov == v && oi < idx))) { v = ov; idx = oi; }
      }
      if (lane == 0) piv_s = idx < c ? c : (idx >= ns ? ns - 1 : idx);
    }
    __syncthreads();
    const int p = piv_s;
    if (p != c)
      for (int j = tid; j < w; j += blockDim.x) { const double2 t = M[c][j]; M[c][j] = M[p][j]; M[p][j] = t; }
    __syncthreads();
    const double2 d = M[c][c];
    if (d.x == 0 && d.y == 0) {
      if (tid == 0 && info) *info = c + 1;
      return;
    }
    const double2 inv = crecip(d);
    for (int j = c + tid; j < w; j += blockDim.x) M[c][j] = cmul(M[c][j], inv);
    __syncthreads();
    for (int e = tid; e < ns * w; e += blockDim.x) {
      const int r = e / w, j = e % w;
      if (r != c && j > c) M[r][j] = cfms(M[r][j], M[r][c], M[c][j]);
    }
    __syncthreads();
    for (int r = tid; r < ns; r += blockDim.x)
      if (r != c) M[r][c] = make_double2(0.0, 0.0);
    __syncthreads();
  }
  for (int e = tid; e < ns * nt; e += blockDim.x) Tm[(e / nt) * ldt + e % nt] = M[e / nt][ns + e % nt];
}

}  // namespace

// ---------------------------------------------------------------------------
// Host drivers (called from mw_capi.hip)
// ---------------------------------------------------------------------------
// workspace: lu11 [NB*NB complex double] | diagonal-block inverses
// [n_panels][2][NB*NB complex double] | cand ping/pong [2][nleaf*NB] | counts [2][nleaf]
int64_t mw_getrf_ws_bytes(int64_t n) {
  const int64_t nleaf = (n + TROWS - 1) / TROWS;
  const int64_t npan = (n + NB - 1) / NB;
  const int64_t bytes = (int64_t)NB * NB * 16 * (1 + 2 * npan) + 2 * nleaf * NB * 4 + 2 * nleaf * 4;
  return (bytes + 255) / 256 * 256;
}

static void* panel_inv(void* work, int64_t k0) {
  return reinterpret_cast<char*>(work) + (int64_t)NB * NB * 16 * (1 + 2 * (k0 / NB));
}

template <typename T>
int mw_getrf_t(int64_t n, void* Av, int64_t lda, int32_t* ipiv, void* work, int32_t* info,
               const volatile int32_t* cancel, hipStream_t st, hipError_t* herr) {
  typedef typename cplx_t<T>::type CT;
  CT* A = reinterpret_cast<CT*>(Av);
  const int64_t nleaf_max = (n + TROWS - 1) / TROWS;
  char* wp = reinterpret_cast<char*>(work);
  void* lu11 = wp;
  const int64_t npan = (n + NB - 1) / NB;
  char* cbase = wp + (int64_t)NB * NB * 16 * (1 + 2 * npan);
  int32_t* cand[2] = {(int32_t*)cbase, (int32_t*)cbase + nleaf_max * NB};
  int32_t* cnt[2] = {cand[1] + nleaf_max * NB, cand[1] + nleaf_max * NB + nleaf_max};
  hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, st, info, (int64_t)1);
  for (int64_t k0 = 0; k0 < n; k0 += NB) {
    if (cancel && *cancel) return MW_EABORT;
    const int bw = (int)((n - k0) < NB ? (n - k0) : NB);
    int nodes = (int)((n - k0 + TROWS - 1) / TROWS);
    int cur = 0;
    hipLaunchKernelGGL(k_tournament<T>, dim3(nodes), dim3(256), 0, st, A, lda, n, k0, bw, 1,
                       (const int32_t*)nullptr, (const int32_t*)nullptr, 0, cand[0], cnt[0],
                       nodes == 1 ? 1 : 0, lu11, ipiv, info);
    while (nodes > 1) {
      const int out = (nodes + FANIN - 1) / FANIN;
      hipLaunchKernelGGL(k_tournament<T>, dim3(out), dim3(256), 0, st, A, lda, n, k0, bw, 0,
                         cand[cur], cnt[cur], nodes, cand[cur ^ 1], cnt[cur ^ 1],
                         out == 1 ? 1 : 0, lu11, ipiv, info);
      cur ^= 1;
      nodes = out;
    }
    const int64_t ncol = n - k0;
    hipLaunchKernelGGL(k_laswp<T>, dim3((unsigned)((ncol + 255) / 256)), dim3(256), 0, st, A, lda,
                       k0, bw, ipiv, k0, n);
    const int64_t rest = n - k0 - bw;
    const int nrb = (int)((rest + AW - 1) / AW);
    CT* inv = reinterpret_cast<CT*>(panel_inv(work, k0));
    hipLaunchKernelGGL(k_tri_inv<T>, dim3(1), dim3(64), 0, st, (const CT*)lu11, (int64_t)NB, bw,
                       inv);
    hipLaunchKernelGGL(k_panel_apply<T>, dim3(2 * nrb + 1), dim3(256), 0, st, n, A, lda, k0, bw,
                       (const CT*)inv, (const CT*)lu11, nrb, nrb);
    if (rest > 0) {
      *herr = launch_gemm_sub<T>(rest, rest, bw, A + (k0 + bw) * lda + k0, lda,
                                 A + k0 * lda + k0 + bw, lda, A + (k0 + bw) * lda + k0 + bw, lda,
                                 st);
      if (*herr != hipSuccess) return MW_EHIP;
    }
    *herr = hipGetLastError();
    if (*herr != hipSuccess) return MW_EHIP;
  }
  return MW_OK;
}

template <typename T>
int mw_getrs_t(int64_t n, int64_t nrhs, const void* Av, int64_t lda, const int32_t* ipiv,
               const void* work, void* Xv, int64_t ldx, hipStream_t st, hipError_t* herr) {
  typedef typename cplx_t<T>::type CT;
  const CT* A = reinterpret_cast<const CT*>(Av);
  CT* X = reinterpret_cast<CT*>(Xv);
  const unsigned gx = (unsigned)((nrhs + AW - 1) / AW);
  for (int64_t k0 = 0; k0 < n; k0 += NB) {
    const int bw = (int)((n - k0) < NB ? (n - k0) : NB);
    const CT* inv = reinterpret_cast<const CT*>(panel_inv(const_cast<void*>(work), k0));
    hipLaunchKernelGGL(k_solve_apply<T>, dim3(gx), dim3(256), 0, st, nrhs, inv, k0, bw, ipiv, X,
                       ldx, 1);
    const int64_t rest = n - k0 - bw;
    if (rest > 0) {
      *herr = launch_gemm_sub<T>(rest, nrhs, bw, A + (k0 + bw) * lda + k0, lda, X + k0 * ldx,
                                 ldx, X + (k0 + bw) * ldx, ldx, st);
      if (*herr != hipSuccess) return MW_EHIP;
    }
  }
  const int64_t nlast = ((n + NB - 1) / NB - 1) * NB;
  for (int64_t k0 = nlast; k0 >= 0; k0 -= NB) {
    const int bw = (int)((n - k0) < NB ? (n - k0) : NB);
    const CT* inv = reinterpret_cast<const CT*>(panel_inv(const_cast<void*>(work), k0));
    hipLaunchKernelGGL(k_solve_apply<T>, dim3(gx), dim3(256), 0, st, nrhs, inv, k0, bw,
                       (const int32_t*)nullptr, X, ldx, 0);
    if (k0 > 0) {
      *herr = launch_gemm_sub<T>(k0, nrhs, bw, A + k0, lda, X + k0 * ldx, ldx, X, ldx, st);
      if (*herr != hipSuccess) return MW_EHIP;
    }
  }
  *herr = hipGetLastError();
  return *herr == hipSuccess ? MW_OK : MW_EHIP;
}

template int mw_getrf_t<double>(int64_t, void*, int64_t, int32_t*, void*, int32_t*,
                                const volatile int32_t*, hipStream_t, hipError_t*);
template int mw_getrf_t<float>(int64_t, void*, int64_t, int32_t*, void*, int32_t*,
                               const volatile int32_t*, hipStream_t, hipError_t*);
template int mw_getrs_t<double>(int64_t, int64_t, const void*, int64_t, const int32_t*,
                                const void*, void*, int64_t, hipStream_t, hipError_t*);
template int mw_getrs_t<float>(int64_t, int64_t, const void*, int64_t, const int32_t*,
                               const void*, void*, int64_t, hipStream_t, hipError_t*);

hipError_t mw_gemm_sub_z(int64_t m, int64_t n, int64_t k, const void* A, int64_t lda,
                         const void* B, int64_t ldb, void* C, int64_t ldc, hipStream_t st) {
  return launch_gemm_sub<double>(m, n, k, A, lda, B, ldb, C, ldc, st);
}
hipError_t mw_gemm_sub_c(int64_t m, int64_t n, int64_t k, const void* A, int64_t lda,
                         const void* B, int64_t ldb, void* C, int64_t ldc, hipStream_t st) {
  return launch_gemm_sub<float>(m, n, k, A, lda, B, ldb, C, ldc, st);
}

hipError_t mw_launch_sparse_rows(int64_t nrows, int maxnz, const int32_t* idx, const double* val,
                                 const void* X, int64_t ldx, int64_t ncols, void* out,
                                 int64_t ldo, hipStream_t st) {
  if (nrows <= 0 || ncols <= 0) return hipSuccess;
  dim3 grid((unsigned)((ncols + 255) / 256), (unsigned)nrows);
  hipLaunchKernelGGL(k_sparse_rows, grid, dim3(256), 0, st, nrows, maxnz, idx, val,
                     (const double2*)X, ldx, ncols, (double2*)out, ldo);
  return hipGetLastError();
}

hipError_t mw_launch_scatter_cols(int64_t n, int64_t ncols, int maxnz, const int32_t* idx,
                                  const double* val, void* X, int64_t ldx, hipStream_t st) {
  if (n <= 0 || ncols <= 0) return hipSuccess;
  const int64_t tot = n * ncols;
  hipLaunchKernelGGL(k_zero_rows, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, n, ncols,
                     (double2*)X, ldx);
  const int64_t ne = ncols * maxnz;
  hipLaunchKernelGGL(k_scatter_cols, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, st, ncols,
                     maxnz, idx, val, (double2*)X, ldx);
  return hipGetLastError();
}

hipError_t mw_launch_small_solve(int ns, int nt, void* S, int64_t lds, void* T, int64_t ldt,
                                 int32_t* info, hipStream_t st) {
  hipLaunchKernelGGL(k_small_solve, dim3(1), dim3(256), 0, st, ns, nt, (double2*)S, lds,
                     (double2*)T, ldt, info);
  return hipGetLastError();
}
