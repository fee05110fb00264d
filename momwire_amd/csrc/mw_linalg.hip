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
// Tournament pivot selection for one panel (columns k0 .. k0+bw).
//
// Leaves: each workgroup factors a 64-row chunk of the panel in LDS with
// partial pivoting and nominates its bw pivot rows.  Merges: the last of two
// sibling workgroups to finish (counted with a device-scope atomic) reloads
// the two nominee sets' ORIGINAL rows (<= 64 rows), factors them the same
// way and nominates again; the root's nominees are the panel's pivot rows,
// and its in-LDS factorisation of them is L11\U11.  The matrix itself is
// read-only here.
// ---------------------------------------------------------------------------
struct PivWs {
  int32_t* cand;     // [levels][nodes][NB] nominated row ids
  int32_t* ncand;    // [levels][nodes]
  int32_t* counter;  // [levels][nodes] arrival counts (root resets them)
  void* lu11;        // [NB][NB] complex
};

template <typename CT>
MW_DEV void lds_gepp(CT (*S)[NB + 1], int32_t* rid, int rows, int cols, int* piv_s,
                     int32_t* info, int k0) {
  const int tid = threadIdx.x, lane = tid & 63;
  const int steps = rows < cols ? rows : cols;
  for (int c = 0; c < steps; ++c) {
    if (tid < 64) {
      double v = (lane >= c && lane < rows) ? cabs1(S[lane][c]) : -1.0;
      int idx = lane;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(v, off);
        const int oi = __shfl_xor(idx, off);
        if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
      }
      if (lane == 0) *piv_s = idx;
    }
    __syncthreads();
    const int p = *piv_s;
    if (p != c) {
      for (int j = tid; j < cols; j += blockDim.x) {
        const CT t = S[c][j]; S[c][j] = S[p][j]; S[p][j] = t;
      }
      if (tid == 0) { const int32_t t = rid[c]; rid[c] = rid[p]; rid[p] = t; }
    }
    __syncthreads();
    const CT d = S[c][c];
    const bool zero = (d.x == 0 && d.y == 0);
    if (zero && info && tid == 0 && *info == 0) *info = k0 + c + 1;
    if (!zero) {
      const CT inv = crecip(d);
      for (int r = c + 1 + tid; r < rows; r += blockDim.x) S[r][c] = cmul(S[r][c], inv);
    }
    __syncthreads();
    if (!zero) {
      const int nr = rows - c - 1, ncl = cols - c - 1;
      for (int e = tid; e < nr * ncl; e += blockDim.x) {
        const int r = c + 1 + e / ncl, j = c + 1 + e % ncl;
        S[r][j] = cfms(S[r][j], S[r][c], S[c][j]);
      }
    }
    __syncthreads();
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_tournament(int64_t n, typename cplx_t<T>::type* A,
                                                    int64_t lda, int64_t k0, int bw, int nleaf,
                                                    PivWs ws, int32_t* ipiv, int32_t* info) {
  typedef typename cplx_t<T>::type CT;
  __shared__ CT S[CH][NB + 1];
  __shared__ int32_t rid[CH];
  __shared__ int piv_s, last_s, rows_s;
  __shared__ int32_t pos_s[2 * NB], val_s[2 * NB];

  const int tid = threadIdx.x;
  const int64_t m = n - k0;
  int level = 0, node = blockIdx.x, nodes = nleaf;
  int lvl_off = 0;  // node offset of this level in the flattened arrays

  // leaf: rows k0 + node*CH ...
  int rows = (int)((m - (int64_t)node * CH) < CH ? (m - (int64_t)node * CH) : CH);
  for (int e = tid; e < rows * bw; e += blockDim.x) {
    const int r = e / bw, c = e % bw;
    S[r][c] = A[(k0 + (int64_t)node * CH + r) * lda + k0 + c];
  }
  if (tid < rows) rid[tid] = (int32_t)(k0 + (int64_t)node * CH + tid);
  __syncthreads();

  while (true) {
    lds_gepp(S, rid, rows, bw, &piv_s, nullptr, k0);
    const int nom = rows < bw ? rows : bw;
    if (nodes == 1) break;  // root
    // publish nominees
    if (tid < nom) ws.cand[(size_t)(lvl_off + node) * NB + tid] = rid[tid];
    if (tid == 0) ws.ncand[lvl_off + node] = nom;
    __threadfence();
    __syncthreads();
    const int parent = node >> 1;
    const int pnodes = (nodes + 1) >> 1;
    const int nchild = (2 * parent + 1 < nodes) ? 2 : 1;
    const int plvl_off = lvl_off + nodes;
    if (tid == 0) {
      const int old = __hip_atomic_fetch_add(&ws.counter[plvl_off + parent], 1, __ATOMIC_ACQ_REL,
                                             __HIP_MEMORY_SCOPE_AGENT);
      last_s = (old == nchild - 1);
    }
    __syncthreads();
    if (!last_s) return;
    __threadfence();
    // gather both children's nominees (original rows)
    if (tid == 0) {
      int cnt = 0;
      for (int ch = 0; ch < nchild; ++ch) {
        const int cn = 2 * parent + ch;
        const int nc = __hip_atomic_load(&ws.ncand[lvl_off + cn], __ATOMIC_ACQUIRE,
                                         __HIP_MEMORY_SCOPE_AGENT);
        for (int t = 0; t < nc; ++t)
          rid[cnt++] = __hip_atomic_load(&ws.cand[(size_t)(lvl_off + cn) * NB + t],
                                         __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
      }
      rows_s = cnt;
    }
    __syncthreads();
    rows = rows_s;
    for (int e = tid; e < rows * bw; e += blockDim.x) {
      const int r = e / bw, c = e % bw;
      S[r][c] = A[(int64_t)rid[r] * lda + k0 + c];
    }
    __syncthreads();
    level += 1;
    node = parent;
    lvl_off = plvl_off;
    nodes = pnodes;
  }

  // Root: rid[0..bw) are the pivot rows in order; S holds L11\U11.
  CT* lu = reinterpret_cast<CT*>(ws.lu11);
  for (int e = tid; e < bw * bw; e += blockDim.x) lu[e] = S[e / bw][e % bw];
  for (int c = 0; c < bw; ++c) {
    const CT d = S[c][c];
    if (d.x == 0 && d.y == 0 && tid == 0 && *info == 0) *info = (int32_t)(k0 + c + 1);
  }
  if (tid == 0) {
    // Sequential interchanges that bring rid[i] to position k0+i.
    int nt = 0;  // tracked positions whose content changed
    for (int i = 0; i < bw; ++i) {
      const int32_t g = rid[i];
      int32_t pg = g;  // current position of row id g
      for (int t = 0; t < nt; ++t)
        if (val_s[t] == g) { pg = pos_s[t]; break; }
      const int32_t p1 = (int32_t)(k0 + i);
      int32_t id1 = p1;  // current content of position p1
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
  }
  // reset the arrival counters for the next panel
  int total = 0;
  for (int nd = nleaf; nd > 1; nd = (nd + 1) >> 1) total += nd;
  total += 1;
  for (int e = tid; e < total; e += blockDim.x) ws.counter[e] = 0;
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

// After the interchanges: A11 <- L11\U11, L21 <- A21 U11^{-1}, U12 <- L11^{-1} A12.
// Blocks [0, nrb): rows of L21; [nrb, nrb+ncb): columns of U12; last: A11.
template <typename T>
__global__ __launch_bounds__(256) void k_panel_trsm(int64_t n, typename cplx_t<T>::type* A,
                                                    int64_t lda, int64_t k0, int bw,
                                                    const void* lu11, int nrb, int ncb) {
  typedef typename cplx_t<T>::type CT;
  __shared__ CT L[NB][NB + 1];
  __shared__ CT dinv[NB];
  const int tid = threadIdx.x;
  const CT* lu = reinterpret_cast<const CT*>(lu11);
  for (int e = tid; e < bw * bw; e += blockDim.x) L[e / bw][e % bw] = lu[e];
  __syncthreads();
  if (tid < bw) dinv[tid] = crecip(L[tid][tid]);
  __syncthreads();
  const int64_t rest = n - k0 - bw;
  const int b = blockIdx.x;
  if (b < nrb) {
    // one row of L21 per thread: x U11 = a
    const int64_t r = k0 + bw + (int64_t)b * 256 + tid;
    if (r >= n) return;
    CT* row = A + r * lda + k0;
    CT x[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j)
      if (j < bw) x[j] = row[j];
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      if (j >= bw) break;
      CT s = x[j];
#pragma unroll
      for (int l = 0; l < NB; ++l) {
        if (l >= j) break;
        s = cfms(s, x[l], L[l][j]);
      }
      x[j] = cmul(s, dinv[j]);
    }
#pragma unroll
    for (int j = 0; j < NB; ++j)
      if (j < bw) row[j] = x[j];
  } else if (b < nrb + ncb) {
    // one column of U12 per thread: L11 x = a (unit lower)
    const int64_t c = k0 + bw + (int64_t)(b - nrb) * 256 + tid;
    if (c >= n) return;
    CT x[NB];
#pragma unroll
    for (int i = 0; i < NB; ++i)
      if (i < bw) x[i] = A[(k0 + i) * lda + c];
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      if (i >= bw) break;
      CT s = x[i];
#pragma unroll
      for (int l = 0; l < NB; ++l) {
        if (l >= i) break;
        s = cfms(s, L[i][l], x[l]);
      }
      x[i] = s;
    }
#pragma unroll
    for (int i = 0; i < NB; ++i)
      if (i < bw) A[(k0 + i) * lda + c] = x[i];
  } else {
    for (int e = tid; e < bw * bw; e += blockDim.x)
      A[(k0 + e / bw) * lda + k0 + e % bw] = L[e / bw][e % bw];
  }
  (void)rest;
}

// Forward block of the solve: per RHS column, apply the panel's interchanges
// then X_p <- L11^{-1} X_p.
template <typename T>
__global__ __launch_bounds__(256) void k_solve_fwd_block(int64_t nrhs, const typename cplx_t<T>::type* A,
                                                         int64_t lda, int64_t k0, int bw,
                                                         const int32_t* ipiv,
                                                         typename cplx_t<T>::type* X, int64_t ldx) {
  typedef typename cplx_t<T>::type CT;
  __shared__ CT L[NB][NB + 1];
  const int tid = threadIdx.x;
  for (int e = tid; e < bw * bw; e += blockDim.x) L[e / bw][e % bw] = A[(k0 + e / bw) * lda + k0 + e % bw];
  __syncthreads();
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + tid;
  if (c >= nrhs) return;
  for (int i = 0; i < bw; ++i) {
    const int64_t p1 = k0 + i, p2 = ipiv[k0 + i];
    if (p2 != p1) {
      const CT t = X[p1 * ldx + c]; X[p1 * ldx + c] = X[p2 * ldx + c]; X[p2 * ldx + c] = t;
    }
  }
  CT x[NB];
#pragma unroll
  for (int i = 0; i < NB; ++i)
    if (i < bw) x[i] = X[(k0 + i) * ldx + c];
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    if (i >= bw) break;
    CT s = x[i];
#pragma unroll
    for (int l = 0; l < NB; ++l) {
      if (l >= i) break;
      s = cfms(s, L[i][l], x[l]);
    }
    x[i] = s;
  }
#pragma unroll
  for (int i = 0; i < NB; ++i)
    if (i < bw) X[(k0 + i) * ldx + c] = x[i];
}

// Backward block: X_p <- U11^{-1} X_p.
template <typename T>
__global__ __launch_bounds__(256) void k_solve_bwd_block(int64_t nrhs, const typename cplx_t<T>::type* A,
                                                         int64_t lda, int64_t k0, int bw,
                                                         typename cplx_t<T>::type* X, int64_t ldx) {
  typedef typename cplx_t<T>::type CT;
  __shared__ CT U[NB][NB + 1];
  __shared__ CT dinv[NB];
  const int tid = threadIdx.x;
  for (int e = tid; e < bw * bw; e += blockDim.x) U[e / bw][e % bw] = A[(k0 + e / bw) * lda + k0 + e % bw];
  __syncthreads();
  if (tid < bw) dinv[tid] = crecip(U[tid][tid]);
  __syncthreads();
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + tid;
  if (c >= nrhs) return;
  CT x[NB];
#pragma unroll
  for (int i = 0; i < NB; ++i)
    if (i < bw) x[i] = X[(k0 + i) * ldx + c];
#pragma unroll
  for (int ii = NB - 1; ii >= 0; --ii) {
    if (ii >= bw) continue;
    CT s = x[ii];
#pragma unroll
    for (int l = NB - 1; l >= 0; --l) {
      if (l <= ii) break;
      if (l < bw) s = cfms(s, U[ii][l], x[l]);
    }
    x[ii] = cmul(s, dinv[ii]);
  }
#pragma unroll
  for (int i = 0; i < NB; ++i)
    if (i < bw) X[(k0 + i) * ldx + c] = x[i];
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
      double v = (lane >= c && lane < ns) ? cabs1(M[lane][c]) : -1.0;
      int idx = lane;
      for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(v, off);
        const int oi = __shfl_xor(idx, off);
        if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
      }
      if (lane == 0) piv_s = idx;
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
// Host drivers (called from mw_capi.cpp)
// ---------------------------------------------------------------------------
int64_t mw_getrf_ws_bytes(int64_t n) {
  int64_t nleaf = (n + CH - 1) / CH;
  int64_t nodes = 0;
  for (int64_t nd = nleaf; nd > 1; nd = (nd + 1) / 2) nodes += nd;
  nodes += 1;
  int64_t bytes = nodes * NB * 4 + nodes * 4 + nodes * 4 + (int64_t)NB * NB * 16;
  return (bytes + 255) / 256 * 256 + 1024;
}

static PivWs carve_ws(void* work, int64_t n) {
  int64_t nleaf = (n + CH - 1) / CH;
  int64_t nodes = 0;
  for (int64_t nd = nleaf; nd > 1; nd = (nd + 1) / 2) nodes += nd;
  nodes += 1;
  char* p = reinterpret_cast<char*>(work);
  PivWs ws;
  ws.lu11 = p;                 p += (int64_t)NB * NB * 16;
  ws.counter = (int32_t*)p;    p += nodes * 4;
  ws.ncand = (int32_t*)p;      p += nodes * 4;
  ws.cand = (int32_t*)p;
  return ws;
}

template <typename T>
int mw_getrf_t(int64_t n, void* Av, int64_t lda, int32_t* ipiv, void* work, int32_t* info,
               const volatile int32_t* cancel, hipStream_t st, hipError_t* herr) {
  typedef typename cplx_t<T>::type CT;
  CT* A = reinterpret_cast<CT*>(Av);
  PivWs ws = carve_ws(work, n);
  *herr = hipMemsetAsync(work, 0, mw_getrf_ws_bytes(n), st);
  if (*herr != hipSuccess) return MW_EHIP;
  *herr = hipMemsetAsync(info, 0, sizeof(int32_t), st);
  if (*herr != hipSuccess) return MW_EHIP;
  for (int64_t k0 = 0; k0 < n; k0 += NB) {
    if (cancel && *cancel) return MW_EABORT;
    const int bw = (int)((n - k0) < NB ? (n - k0) : NB);
    const int nleaf = (int)((n - k0 + CH - 1) / CH);
    hipLaunchKernelGGL(k_tournament<T>, dim3(nleaf), dim3(256), 0, st, n, A, lda, k0, bw, nleaf,
                       ws, ipiv, info);
    const int64_t ncol = n - k0;
    hipLaunchKernelGGL(k_laswp<T>, dim3((unsigned)((ncol + 255) / 256)), dim3(256), 0, st, A, lda,
                       k0, bw, ipiv, k0, n);
    const int64_t rest = n - k0 - bw;
    const int nrb = (int)((rest + 255) / 256);
    hipLaunchKernelGGL(k_panel_trsm<T>, dim3(2 * nrb + 1), dim3(256), 0, st, n, A, lda, k0, bw,
                       ws.lu11, nrb, nrb);
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
               void* Xv, int64_t ldx, hipStream_t st, hipError_t* herr) {
  typedef typename cplx_t<T>::type CT;
  const CT* A = reinterpret_cast<const CT*>(Av);
  CT* X = reinterpret_cast<CT*>(Xv);
  const unsigned gx = (unsigned)((nrhs + 255) / 256);
  for (int64_t k0 = 0; k0 < n; k0 += NB) {
    const int bw = (int)((n - k0) < NB ? (n - k0) : NB);
    hipLaunchKernelGGL(k_solve_fwd_block<T>, dim3(gx), dim3(256), 0, st, nrhs, A, lda, k0, bw,
                       ipiv, X, ldx);
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
    hipLaunchKernelGGL(k_solve_bwd_block<T>, dim3(gx), dim3(256), 0, st, nrhs, A, lda, k0, bw, X,
                       ldx);
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
template int mw_getrs_t<double>(int64_t, int64_t, const void*, int64_t, const int32_t*, void*,
                                int64_t, hipStream_t, hipError_t*);
template int mw_getrs_t<float>(int64_t, int64_t, const void*, int64_t, const int32_t*, void*,
                               int64_t, hipStream_t, hipError_t*);

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
