// Dense complex linear algebra for the port solve on gfx950: blocked LU with
// tournament row pivoting, multi-RHS triangular solves, complex GEMM on MFMA,
// and the small kernels of the port algebra.  Templated on the real type
// (double: FP64 path, float: the FP32 factor of the mixed path).
//
// Storage: row-major, leading dimension in complex elements.
#include <cfloat>
#include <algorithm>
#include <vector>

#include "mw_common.h"
#include "mw_mfma.h"

namespace {

constexpr int NB = 32;   // LU panel width

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
// C tile per workgroup: 128 rows x 64 complex columns (128 real columns);
// four waves in 2 x 2, each a 4 x 4 grid of 16 x 16 MFMA tiles (64 x 64 real).
// The accumulators start from C itself and MFMA accumulates (-A) B, so C is
// read once and written once.  K advances 8 complex (16 real) per LDS stage,
// double-buffered: the next stage's global loads are issued before the MFMAs
// of the current one.
constexpr int GBM = 128;     // C rows per workgroup
constexpr int GBN = 64;      // C complex columns per workgroup (128 real)
constexpr int GKC = 8;       // complex k per LDS stage (16 real)

template <typename T>
struct GemmLds {
  T As[2][GBM][2 * GKC + 1];                     // A as reals, padded rows
  typename cplx_t<T>::type Bs[2][GKC][GBN + 1];  // B as complex
};

// TA = false: C <- C - A B.  TA = true: C <- C + A^T B with A stored k x m.
template <typename T, bool TA>
MW_DEV void gemm_big_tile(GemmLds<T>& L, int bx, int by, int64_t m, int64_t n, int64_t kdim,
                          const typename cplx_t<T>::type* __restrict__ A, int64_t lda,
                          const typename cplx_t<T>::type* __restrict__ B, int64_t ldb,
                          typename cplx_t<T>::type* C, int64_t ldc) {
  typedef typename cplx_t<T>::type CT;
  typedef typename Mfma<T>::acc_t acc_t;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;  // wave's 64x64 (real) quadrant
  const int64_t row0 = (int64_t)by * GBM;
  const int64_t col0 = (int64_t)bx * GBN;  // complex columns
  T* Cr = reinterpret_cast<T*>(C);

  // accumulators <- C
  acc_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gr = row0 + wr * 64 + i * 16 + Mfma<T>::row(lane, r);
        const int64_t gcr = 2 * col0 + wc * 64 + j * 16 + (lane & 15);
        acc[i][j][r] = (gr < m && gcr < 2 * n) ? Cr[gr * 2 * ldc + gcr] : (T)0;
      }

  // staging: A 128 x 8 complex (4 per thread), B 8 x 64 complex (2 per thread)
  CT ra[4], rb[2];
  auto gload = [&](int64_t k0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = tid + 256 * q;
      const int r = TA ? (e & 127) : (e >> 3), c = TA ? (e >> 7) : (e & 7);
      CT v;
      v.x = 0; v.y = 0;
      if (row0 + r < m && k0 + c < kdim)
        v = TA ? A[(k0 + c) * lda + row0 + r] : A[(row0 + r) * lda + k0 + c];
      ra[q] = v;
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int e = tid + 256 * q;
      const int r = e >> 6, c = e & 63;
      CT v;
      v.x = 0; v.y = 0;
      if (k0 + r < kdim && col0 + c < n) v = B[(k0 + r) * ldb + col0 + c];
      rb[q] = v;
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = tid + 256 * q;
      const int r = TA ? (e & 127) : (e >> 3), c = TA ? (e >> 7) : (e & 7);
      L.As[buf][r][2 * c] = TA ? ra[q].x : -ra[q].x;  // C - A B  =  C + (-A) B
      L.As[buf][r][2 * c + 1] = TA ? ra[q].y : -ra[q].y;
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int e = tid + 256 * q;
      L.Bs[buf][e >> 6][e & 63] = rb[q];
    }
  };

  gload(0);
  sstore(0);
  __syncthreads();
  int buf = 0;
  for (int64_t k0 = 0; k0 < kdim; k0 += GKC) {
    const bool more = k0 + GKC < kdim;
    if (more) gload(k0 + GKC);
#pragma unroll
    for (int ks = 0; ks < 2 * GKC; ks += 4) {
      const int kr = ks + (lane >> 4);        // real k of this lane
      const int kc = kr >> 1, kpar = kr & 1;  // complex row of B, re/im half
      T a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = L.As[buf][wr * 64 + i * 16 + (lane & 15)][kr];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int cr = wc * 64 + j * 16 + (lane & 15);  // real column
        const CT bv = L.Bs[buf][kc][cr >> 1];
        // B_r[2l+s][2j+t] = s==0 ? (t==0 ? br : bi) : (t==0 ? -bi : br)
        b[j] = (kpar == 0) ? ((cr & 1) ? bv.y : bv.x) : ((cr & 1) ? bv.x : -bv.y);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = Mfma<T>::mma(a[i], b[j], acc[i][j]);
    }
    if (more) {
      sstore(buf ^ 1);
      __syncthreads();
      buf ^= 1;
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gr = row0 + wr * 64 + i * 16 + Mfma<T>::row(lane, r);
        const int64_t gcr = 2 * col0 + wc * 64 + j * 16 + (lane & 15);
        if (gr < m && gcr < 2 * n) Cr[gr * 2 * ldc + gcr] = acc[i][j][r];
      }
}

template <typename T, bool TA>
MW_DEV void gemm_big_body(int bx, int by, int64_t m, int64_t n, int64_t kdim,
                          const typename cplx_t<T>::type* __restrict__ A, int64_t lda,
                          const typename cplx_t<T>::type* __restrict__ B, int64_t ldb,
                          typename cplx_t<T>::type* C, int64_t ldc) {
  __shared__ GemmLds<T> L;
  gemm_big_tile<T, TA>(L, bx, by, m, n, kdim, A, lda, B, ldb, C, ldc);
}

template <typename T, bool TA>
__global__ __launch_bounds__(256) void k_gemm_sub(int64_t m, int64_t n, int64_t kdim,
                                                  const typename cplx_t<T>::type* __restrict__ A,
                                                  int64_t lda,
                                                  const typename cplx_t<T>::type* __restrict__ B,
                                                  int64_t ldb, typename cplx_t<T>::type* C,
                                                  int64_t ldc) {
  gemm_big_body<T, TA>(blockIdx.x, blockIdx.y, m, n, kdim, A, lda, B, ldb, C, ldc);
}

// Small-tile variant (64 rows x 32 complex columns, each wave 2 x 2 MFMA
// tiles) for products too small to occupy the chip with 128 x 64 tiles.
// SKC complex k per stage; the next stage's operands load into registers
// while a stage's MFMAs run (8 complex k per stage with no prefetch left the
// long-K readout products waiting on their loads: 34 TF/s at 1024 x 1024 x
// 15360).
constexpr int SBM = 64, SBN = 32, SKC = 32;

template <typename T>
struct GemmSmallLds {
  T As[SBM][2 * SKC + 1];
  typename cplx_t<T>::type Bs[SKC][SBN + 1];
};

template <typename T, bool TA>
MW_DEV void gemm_small_tile(GemmSmallLds<T>& L, int bx, int by, int64_t m, int64_t n,
                            int64_t kdim, const typename cplx_t<T>::type* __restrict__ A,
                            int64_t lda, const typename cplx_t<T>::type* __restrict__ B,
                            int64_t ldb, typename cplx_t<T>::type* C, int64_t ldc) {
  typedef typename cplx_t<T>::type CT;
  typedef typename Mfma<T>::acc_t acc_t;
  T (*As)[2 * SKC + 1] = L.As;
  CT (*Bs)[SBN + 1] = L.Bs;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int64_t row0 = (int64_t)by * SBM;
  const int64_t col0 = (int64_t)bx * SBN;
  T* Cr = reinterpret_cast<T*>(C);
  acc_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gr = row0 + wr * 32 + i * 16 + Mfma<T>::row(lane, r);
        const int64_t gcr = 2 * col0 + wc * 32 + j * 16 + (lane & 15);
        acc[i][j][r] = (gr < m && gcr < 2 * n) ? Cr[gr * 2 * ldc + gcr] : (T)0;
      }
  constexpr int AU = SBM * SKC / 256, BU = SKC * SBN / 256;
  CT ra[AU], rb[BU];
  auto gload = [&](int64_t k0) {
#pragma unroll
    for (int u = 0; u < AU; ++u) {
      const int e = tid + 256 * u;
      const int r = TA ? (e % SBM) : (e / SKC), c = TA ? (e / SBM) : (e % SKC);
      CT v;
      v.x = 0; v.y = 0;
      if (row0 + r < m && k0 + c < kdim)
        v = TA ? A[(k0 + c) * lda + row0 + r] : A[(row0 + r) * lda + k0 + c];
      ra[u] = v;
    }
#pragma unroll
    for (int u = 0; u < BU; ++u) {
      const int e = tid + 256 * u, r = e / SBN, c = e % SBN;
      CT v;
      v.x = 0; v.y = 0;
      if (k0 + r < kdim && col0 + c < n) v = B[(k0 + r) * ldb + col0 + c];
      rb[u] = v;
    }
  };
  gload(0);
  for (int64_t k0 = 0; k0 < kdim; k0 += SKC) {
#pragma unroll
    for (int u = 0; u < AU; ++u) {
      const int e = tid + 256 * u;
      const int r = TA ? (e % SBM) : (e / SKC), c = TA ? (e / SBM) : (e % SKC);
      As[r][2 * c] = TA ? ra[u].x : -ra[u].x;
      As[r][2 * c + 1] = TA ? ra[u].y : -ra[u].y;
    }
#pragma unroll
    for (int u = 0; u < BU; ++u) {
      const int e = tid + 256 * u;
      Bs[e / SBN][e % SBN] = rb[u];
    }
    __syncthreads();
    if (k0 + SKC < kdim) gload(k0 + SKC);
#pragma unroll
    for (int ks = 0; ks < 2 * SKC; ks += 4) {
      const int kr = ks + (lane >> 4);
      const int kc = kr >> 1, kpar = kr & 1;
      T a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) a[i] = As[wr * 32 + i * 16 + (lane & 15)][kr];
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int cr = wc * 32 + j * 16 + (lane & 15);
        const CT bv = Bs[kc][cr >> 1];
        b[j] = (kpar == 0) ? ((cr & 1) ? bv.y : bv.x) : ((cr & 1) ? bv.x : -bv.y);
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = Mfma<T>::mma(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gr = row0 + wr * 32 + i * 16 + Mfma<T>::row(lane, r);
        const int64_t gcr = 2 * col0 + wc * 32 + j * 16 + (lane & 15);
        if (gr < m && gcr < 2 * n) Cr[gr * 2 * ldc + gcr] = acc[i][j][r];
      }
}

template <typename T, bool TA>
MW_DEV void gemm_small_body(int bx, int by, int64_t m, int64_t n, int64_t kdim,
                            const typename cplx_t<T>::type* __restrict__ A, int64_t lda,
                            const typename cplx_t<T>::type* __restrict__ B, int64_t ldb,
                            typename cplx_t<T>::type* C, int64_t ldc) {
  __shared__ GemmSmallLds<T> L;
  gemm_small_tile<T, TA>(L, bx, by, m, n, kdim, A, lda, B, ldb, C, ldc);
}

template <typename T, bool TA>
__global__ __launch_bounds__(256) void k_gemm_sub_s(int64_t m, int64_t n, int64_t kdim,
                                                    const typename cplx_t<T>::type* __restrict__ A,
                                                    int64_t lda,
                                                    const typename cplx_t<T>::type* __restrict__ B,
                                                    int64_t ldb, typename cplx_t<T>::type* C,
                                                    int64_t ldc) {
  gemm_small_body<T, TA>(blockIdx.x, blockIdx.y, m, n, kdim, A, lda, B, ldb, C, ldc);
}

// ---------------------------------------------------------------------------
// Large products C <- C - A B (the LU's K = 512 trailing updates, the
// residuals of the mixed form): the same 128 x 64-complex C tile per
// workgroup, with the operands laid out in LDS for 16-byte reads.
//
//  - B is expanded to its real form [[br, bi], [-bi, br]] when it is staged,
//    and stored transposed (real column major), so a B fragment is a plain
//    LDS load with no per-step selects.
//  - Within a stage the real k index is permuted: position (kr % G) S + kr / G,
//    G the k groups of the MFMA's lanes, S = KR / G.  A lane then finds the k
//    values of its successive MFMA steps side by side and reads them 16 bytes
//    at a time (A and B use the same permutation, so the sum is unchanged).
//  - FP64 (k_gemm_sub_pd): 16x16x4 MFMA, each wave a 4 x 4 grid of 16 x 16
//    tiles, KR = 16 real k per stage.  FP32 (k_gemm_sub_pf): 32x32x2 MFMA
//    (half the instructions and LDS reads of 16x16x4 for the same work),
//    each wave a 2 x 2 grid of 32 x 32 tiles, KR = 32; its 32 x 32 output
//    element r of lane l sits at row 8 (r / 4) + 4 (l / 32) + r % 4, column
//    l % 32.
//  - Tiles are visited in an XCD-aware grouped order (pgemm_tile): workgroup
//    p runs on XCD p % 8, each XCD takes a contiguous range of the order, and
//    the order walks 8 row tiles down a column of tiles before moving right,
//    so the A rows and B columns a CU pair works on are shared in its XCD's L2.
// ---------------------------------------------------------------------------
typedef float v16f __attribute__((ext_vector_type(16)));

// tile (bx, by) of workgroup p over an nbx x nby tile grid: a bijection
MW_DEV void pgemm_tile(int p, int nbx, int nby, int& bx, int& by) {
  const int T = nbx * nby;
  const int x = p & 7, local = p >> 3;
  const int t = x * (T >> 3) + min(x, T & 7) + local;  // XCD x's range starts here
  constexpr int GM = 8;
  const int gsz = GM * nbx;
  const int gid = t / gsz, first = gid * GM;
  const int gm = min(nby - first, GM);
  const int r = t - gid * gsz;
  by = first + r % gm;
  bx = r / gm;
}

// Tile (bx, by) of workgroup p when only the blocks on or below the block
// diagonal are wanted (the symmetric factorisation's trailing update): the
// matrix in 512 x 512 blocks, block row g holding blocks h = 0 .. g, each
// block 4 x 8 tiles of 128 x 64.  The blocks are numbered row by row and
// their tiles consecutively, and the XCD ranges split that numbering as in
// pgemm_tile.  nbr: block rows.
MW_DEV void pgemm_stair_tile(int p, int nbr, int& bx, int& by) {
  const int T = 16 * nbr * (nbr + 1);  // 32 tiles per block, nbr (nbr + 1) / 2 blocks
  const int x = p & 7, local = p >> 3;
  const int t = x * (T >> 3) + min(x, T & 7) + local;
  const int b = t >> 5, r = t & 31;
  int g = (int)((sqrt(8.0 * b + 1.0) - 1.0) * 0.5);
  while ((g + 1) * (g + 2) / 2 <= b) ++g;
  while (g * (g + 1) / 2 > b) --g;
  const int h = b - g * (g + 1) / 2;
  by = 4 * g + (r & 3);
  bx = 8 * h + (r >> 2);
}

// FP64: 16x16x4 MFMA, KR = 16 real k per stage (8 complex), LDS rows of 18
// doubles (144 bytes: 16 consecutive rows hit distinct banks with 16-byte
// accesses).  Permutation: position (kr % 4) 4 + kr / 4.  A is staged by
// row (thread: row tid / 2, complex k 4 (tid & 1) + q) and written as
// 16-byte pairs; B by REAL column (thread: real column tid & 127, complex
// rows par + 2 q, par = tid >> 7), so consecutive lanes write consecutive
// LDS rows.  Output element r of lane l: row l / 16 + 4 r, column l % 16.
__global__ __launch_bounds__(256, 2) void k_gemm_sub_pd(int64_t m, int64_t n, int64_t kdim,
                                                       const double2* __restrict__ A, int64_t lda,
                                                       const double2* __restrict__ B, int64_t ldb,
                                                       double2* C, int64_t ldc, int nbx, int nby,
                                                       int stair) {
  constexpr int LD = 18, S = 4, KC = 8;
  __shared__ double As[2][128][LD];
  __shared__ double Bt[2][128][LD];
  int bx, by;
  if (stair)
    pgemm_stair_tile(blockIdx.x, stair, bx, by);
  else
    pgemm_tile(blockIdx.x, nbx, nby, bx, by);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;  // the wave's 64 x 64 real quadrant
  const int64_t row0 = (int64_t)by * 128, col0 = (int64_t)bx * 64;
  double* Cr = reinterpret_cast<double*>(C);
  v4d acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gr = row0 + wr * 64 + i * 16 + (lane >> 4) + 4 * r;
        const int64_t gcr = 2 * col0 + wc * 64 + j * 16 + (lane & 15);
        acc[i][j][r] = (gr < m && gcr < 2 * n) ? Cr[gr * 2 * ldc + gcr] : 0.0;
      }
  double2 ra[4], rb[4];
  const int a_row = tid >> 1, a_h = tid & 1;
  const int cr = tid & 127, b_par = tid >> 7;
  auto gload = [&](int64_t k0) {
    const int64_t gr = row0 + a_row;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int64_t kc = k0 + a_h * 4 + q;
      double2 v = make_double2(0.0, 0.0);
      if (gr < m && kc < kdim) v = A[gr * lda + kc];
      ra[q] = v;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int64_t kc = k0 + b_par + 2 * q;
      double2 v = make_double2(0.0, 0.0);
      if (kc < kdim && col0 + (cr >> 1) < n) v = B[kc * ldb + col0 + (cr >> 1)];
      rb[q] = v;
    }
  };
  auto sstore = [&](int buf) {
    // A's reals kr = 8 a_h + 2 q + part sit at ((2q + part) % 4) 4 + 2 a_h +
    // (2q + part) / 4: complex q and q + 2 of one part are neighbours.
    // C - A B = C + (-A) B.
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      *reinterpret_cast<double2*>(&As[buf][a_row][(2 * q) * S + 2 * a_h]) =
          make_double2(-ra[q].x, -ra[q + 2].x);
      *reinterpret_cast<double2*>(&As[buf][a_row][(2 * q + 1) * S + 2 * a_h]) =
          make_double2(-ra[q].y, -ra[q + 2].y);
    }
    // B_r[2kc][cr] = cr even ? br : bi and B_r[2kc + 1][cr] = cr even ? -bi : br,
    // at k groups 2 par and 2 par + 1, step q
    const bool im = cr & 1;
    double v0[4], v1[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      v0[q] = im ? rb[q].y : rb[q].x;
      v1[q] = im ? rb[q].x : -rb[q].y;
    }
    *reinterpret_cast<double2*>(&Bt[buf][cr][(2 * b_par) * S]) = make_double2(v0[0], v0[1]);
    *reinterpret_cast<double2*>(&Bt[buf][cr][(2 * b_par) * S + 2]) = make_double2(v0[2], v0[3]);
    *reinterpret_cast<double2*>(&Bt[buf][cr][(2 * b_par + 1) * S]) = make_double2(v1[0], v1[1]);
    *reinterpret_cast<double2*>(&Bt[buf][cr][(2 * b_par + 1) * S + 2]) = make_double2(v1[2], v1[3]);
  };
  gload(0);
  sstore(0);
  __syncthreads();
  int buf = 0;
  const int g = lane >> 4;
  for (int64_t k0 = 0; k0 < kdim; k0 += KC) {
    const bool more = k0 + KC < kdim;
    if (more) gload(k0 + KC);
#pragma unroll
    for (int h = 0; h < S; h += 2) {
      double a[4][2], b[4][2];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const double2 t =
            *reinterpret_cast<const double2*>(&As[buf][wr * 64 + i * 16 + (lane & 15)][g * S + h]);
        a[i][0] = t.x;
        a[i][1] = t.y;
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const double2 t =
            *reinterpret_cast<const double2*>(&Bt[buf][wc * 64 + j * 16 + (lane & 15)][g * S + h]);
        b[j][0] = t.x;
        b[j][1] = t.y;
      }
#pragma unroll
      for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i][s], b[j][s], acc[i][j], 0, 0, 0);
      if (h == 0 && more) sstore(buf ^ 1);  // between this stage's MFMAs
    }
    if (more) {
      __syncthreads();
      buf ^= 1;
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gr = row0 + wr * 64 + i * 16 + (lane >> 4) + 4 * r;
        const int64_t gcr = 2 * col0 + wc * 64 + j * 16 + (lane & 15);
        if (gr < m && gcr < 2 * n) Cr[gr * 2 * ldc + gcr] = acc[i][j][r];
      }
}

// The FP32 large product written out for the 32x32x2 MFMA (the same
// scheme as k_gemm_sub_pd).  Within a stage, a wave reads the LDS fragments
// of k step h + 1 before it issues the MFMAs of step h, and writes the next
// stage's operands to the other buffer after the second of its four steps;
// tiles wholly inside the matrix load their operands and C with no bounds
// checks (one branch per stage instead of one per element).
typedef float v4f __attribute__((ext_vector_type(4)));

template <int KR, int MINB>
__global__ __launch_bounds__(256, MINB) void k_gemm_sub_pf(int64_t m, int64_t n, int64_t kdim,
                                                          const float2* __restrict__ A, int64_t lda,
                                                          const float2* __restrict__ B, int64_t ldb,
                                                          float2* C, int64_t ldc, int nbx, int nby,
                                                          int stair) {
  constexpr int LD = KR + 4, HALF = KR / 2, KC = KR / 2, BQ = KC / 4;
  __shared__ __attribute__((aligned(16))) float As[2][128][LD];
  __shared__ __attribute__((aligned(16))) float Bt[2][128][LD];
  int bx, by;
  if (stair)
    pgemm_stair_tile(blockIdx.x, stair, bx, by);
  else
    pgemm_tile(blockIdx.x, nbx, nby, bx, by);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int64_t row0 = (int64_t)by * 128, col0 = (int64_t)bx * 64;
  const bool inside = row0 + 128 <= m && col0 + 64 <= n;
  float* Cr = reinterpret_cast<float*>(C);
  v16f acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int64_t gr = row0 + wr * 64 + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const int64_t gcr = 2 * col0 + wc * 64 + j * 32 + (lane & 31);
        acc[i][j][r] = (inside || (gr < m && gcr < 2 * n)) ? Cr[gr * 2 * ldc + gcr] : 0.f;
      }
  constexpr int AQ = KC / 2;  // 8 complex per thread: row tid / 2, half tid & 1
  float2 ra[AQ];
  float2 rb[BQ];              // BQ consecutive complex rows of one column
  const int a_row = tid >> 1, a_h = tid & 1;
  const int b_col = tid & 63, b_k0 = (tid >> 6) * BQ;
  auto gload = [&](int64_t k0) {
    const int64_t gr = row0 + a_row;
    if (inside && k0 + KC <= kdim) {
      const float2* pa = A + gr * lda + k0 + a_h * AQ;
      const float2* pb = B + (k0 + b_k0) * ldb + col0 + b_col;
#pragma unroll
      for (int q = 0; q < AQ; ++q) ra[q] = pa[q];
#pragma unroll
      for (int q = 0; q < BQ; ++q) rb[q] = pb[q * ldb];
      return;
    }
#pragma unroll
    for (int q = 0; q < AQ; ++q) {
      const int64_t kc = k0 + a_h * AQ + q;
      float2 v = make_float2(0.f, 0.f);
      if (gr < m && kc < kdim) v = A[gr * lda + kc];
      ra[q] = v;
    }
#pragma unroll
    for (int q = 0; q < BQ; ++q) {
      const int64_t kc = k0 + b_k0 + q;
      float2 v = make_float2(0.f, 0.f);
      if (kc < kdim && col0 + b_col < n) v = B[kc * ldb + col0 + b_col];
      rb[q] = v;
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int q = 0; q < AQ; ++q) {
      As[buf][a_row][a_h * AQ + q] = -ra[q].x;         // real k 2 kc: position kc
      As[buf][a_row][HALF + a_h * AQ + q] = -ra[q].y;  // real k 2 kc + 1: HALF + kc
    }
    // B's BQ consecutive k of a column, 16 bytes per store (scalar stores two
    // rows apart fell on a quarter of the banks)
#pragma unroll
    for (int q0 = 0; q0 < BQ; q0 += 4) {
      v4f re, im, nim;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        re[q] = rb[q0 + q].x;
        im[q] = rb[q0 + q].y;
        nim[q] = -rb[q0 + q].y;
      }
      const int kc = b_k0 + q0;
      *reinterpret_cast<v4f*>(&Bt[buf][2 * b_col][kc]) = re;
      *reinterpret_cast<v4f*>(&Bt[buf][2 * b_col + 1][kc]) = im;
      *reinterpret_cast<v4f*>(&Bt[buf][2 * b_col][HALF + kc]) = nim;
      *reinterpret_cast<v4f*>(&Bt[buf][2 * b_col + 1][HALF + kc]) = re;
    }
  };
  const int g = lane >> 5;
  // fragments of four consecutive (permuted) real k of one k step
  auto lread = [&](int buf, int h, v4f (&fa)[2], v4f (&fb)[2]) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
      fa[i] = *reinterpret_cast<const v4f*>(&As[buf][wr * 64 + i * 32 + (lane & 31)][g * HALF + h]);
#pragma unroll
    for (int j = 0; j < 2; ++j)
      fb[j] = *reinterpret_cast<const v4f*>(&Bt[buf][wc * 64 + j * 32 + (lane & 31)][g * HALF + h]);
  };
  gload(0);
  sstore(0);
  __syncthreads();
  int buf = 0;
  v4f fa[2][2], fb[2][2];  // [set][tile]: two k steps in flight
  lread(0, 0, fa[0], fb[0]);
  for (int64_t k0 = 0; k0 < kdim; k0 += KC) {
    const bool more = k0 + KC < kdim;
    if (more) gload(k0 + KC);
#pragma unroll
    for (int hh = 0; hh < HALF / 4; ++hh) {
      const int cs = hh & 1;
      if (hh + 1 < HALF / 4) lread(buf, 4 * (hh + 1), fa[cs ^ 1], fb[cs ^ 1]);
#pragma unroll
      for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j)
            acc[i][j] =
                __builtin_amdgcn_mfma_f32_32x32x2f32(fa[cs][i][s], fb[cs][j][s], acc[i][j], 0, 0, 0);
      if (hh == HALF / 8 - 1 && more) sstore(buf ^ 1);
    }
    if (more) {
      __syncthreads();
      buf ^= 1;
      lread(buf, 0, fa[0], fb[0]);
    }
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int64_t gr = row0 + wr * 64 + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const int64_t gcr = 2 * col0 + wc * 64 + j * 32 + (lane & 31);
        if (inside || (gr < m && gcr < 2 * n)) Cr[gr * 2 * ldc + gcr] = acc[i][j][r];
      }
}

// FP32 large products: 16 complex k per stage, two workgroups per CU (8
// complex per stage measured 2-4 % slower, with or without a third workgroup
// per CU).
static void launch_pf(dim3 grid, hipStream_t st, int64_t m, int64_t n, int64_t k, const float2* A,
                      int64_t lda, const float2* B, int64_t ldb, float2* C, int64_t ldc, int nbx,
                      int nby, int stair) {
  hipLaunchKernelGGL((k_gemm_sub_pf<32, 2>), grid, dim3(256), 0, st, m, n, k, A, lda, B, ldb, C,
                     ldc, nbx, nby, stair);
}

template <typename T>
hipError_t launch_gemm_p(int64_t m, int64_t n, int64_t k, const void* A, int64_t lda,
                         const void* B, int64_t ldb, void* C, int64_t ldc, hipStream_t st) {
  const int nbx = (int)((n + GBN - 1) / GBN), nby = (int)((m + GBM - 1) / GBM);
  if constexpr (sizeof(T) == 4)
    launch_pf(dim3((unsigned)(nbx * nby)), st, m, n, k, (const float2*)A, lda, (const float2*)B,
              ldb, (float2*)C, ldc, nbx, nby, 0);
  else
    hipLaunchKernelGGL(k_gemm_sub_pd, dim3((unsigned)(nbx * nby)), dim3(256), 0, st, m, n, k,
                       (const double2*)A, lda, (const double2*)B, ldb, (double2*)C, ldc, nbx, nby,
                       0);
  return hipGetLastError();
}

// C -= A B over the blocks of an m x m product (512 x 512 blocks) on or below
// the block diagonal.
template <typename T>
hipError_t launch_gemm_stair(int64_t m, int64_t k, const void* A, int64_t lda, const void* B,
                             int64_t ldb, void* C, int64_t ldc, hipStream_t st) {
  if (m <= 0 || k <= 0) return hipSuccess;
  const int nbr = (int)((m + 511) / 512);
  const unsigned grid = (unsigned)(16 * nbr * (nbr + 1));
  if constexpr (sizeof(T) == 4)
    launch_pf(dim3(grid), st, m, m, k, (const float2*)A, lda, (const float2*)B, ldb, (float2*)C,
              ldc, 0, 0, nbr);
  else
    hipLaunchKernelGGL(k_gemm_sub_pd, dim3(grid), dim3(256), 0, st, m, m, k, (const double2*)A,
                       lda, (const double2*)B, ldb, (double2*)C, ldc, 0, 0, nbr);
  return hipGetLastError();
}

template <typename T, bool TA = false>
hipError_t launch_gemm_sub(int64_t m, int64_t n, int64_t k, const void* A, int64_t lda,
                           const void* B, int64_t ldb, void* C, int64_t ldc, hipStream_t st) {
  typedef typename cplx_t<T>::type CT;
  if (m <= 0 || n <= 0 || k <= 0) return hipSuccess;
  const int64_t big_tiles = ((n + GBN - 1) / GBN) * ((m + GBM - 1) / GBM);
  // the 128 x 64 tiles when there are enough of them: 1024 at any K; at
  // K >= 256, 256 (one per CU), since each tile's K loop is then long
  // enough to hide its own loads
  const bool big = big_tiles >= 1024 || (k >= 256 && big_tiles >= 256);
  if (big && !TA) {
    return launch_gemm_p<T>(m, n, k, A, lda, B, ldb, C, ldc, st);
  } else if (big) {
    dim3 grid((unsigned)((n + GBN - 1) / GBN), (unsigned)((m + GBM - 1) / GBM));
    hipLaunchKernelGGL((k_gemm_sub<T, TA>), grid, dim3(256), 0, st, m, n, k, (const CT*)A, lda,
                       (const CT*)B, ldb, (CT*)C, ldc);
  } else {
    dim3 grid((unsigned)((n + SBN - 1) / SBN), (unsigned)((m + SBM - 1) / SBM));
    hipLaunchKernelGGL((k_gemm_sub_s<T, TA>), grid, dim3(256), 0, st, m, n, k, (const CT*)A, lda,
                       (const CT*)B, ldb, (CT*)C, ldc);
  }
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

// Broadcast of lane k's value (k wave-uniform): v_readlane into scalar
// registers, so every lane gets it without an LDS round trip.
MW_DEV double lane_bcast(double x, int k) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(x), k);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(x), k);
  return __hiloint2double(hi, lo);
}
MW_DEV float lane_bcast(float x, int k) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), k));
}
template <typename CT> MW_DEV CT lane_bcast(CT v, int k) {
  CT o;
  o.x = lane_bcast(v.x, k);
  o.y = lane_bcast(v.y, k);
  return o;
}

template <typename CT>
struct PanelLds {
  CT M[NB][NB + 1];   // L11\U11 (checking workgroups reuse M and Li for 64 rows of A21)
  CT Li[NB][NB + 1];
  CT Ui[NB][NB + 1];
  CT prow[2][NB];     // the elimination's pivot rows (double-buffered)
  CT T[2][16 * NB];   // products of the inverses' doubling steps (L, U)
  int bad;
  int piv_s;
};

// 1/d for a complex d != 0: conj(d) / |d|^2 with the real reciprocal from
// v_rcp and two Newton steps (a short dependency chain; the elimination step
// waits on it).
MW_DEV double rcp_nr(double a) {
  double r = __builtin_amdgcn_rcp(a);
  double e = fma(-a, r, 1.0);
  r = fma(r, e, r);
  e = fma(-a, r, 1.0);
  return fma(r, e, r);
}
MW_DEV float rcp_nr(float a) { return 1.0f / a; }
template <typename CT> MW_DEV CT crecip_fast(CT d) {
  const auto m = rcp_nr(d.x * d.x + d.y * d.y);
  CT o;
  o.x = d.x * m;
  o.y = -d.y * m;
  return o;
}

// Lane (lane & ~7) | K of a group of 8 lanes (K constant), for every lane:
// ds_swizzle in bit mode (and 0x18, or K) within each half-wave.
template <int K> MW_DEV int swz8(int v) {
  return __builtin_amdgcn_ds_swizzle(v, 0x18 | (K << 5));
}
template <int K> MW_DEV double swz8(double v) {
  return __hiloint2double(swz8<K>(__double2hiint(v)), swz8<K>(__double2loint(v)));
}
template <int K> MW_DEV float swz8(float v) { return __int_as_float(swz8<K>(__float_as_int(v))); }
template <int K, typename CT> MW_DEV CT swz8c(CT v) {
  CT o;
  o.x = swz8<K>(v.x);
  o.y = swz8<K>(v.y);
  return o;
}

// One elimination step k of lu32_eliminate (K = k & 7 and J = k >> 3 as
// template constants, so the slot and the swizzle are static).
template <int k, typename CT>
MW_DEV void lu32_step(PanelLds<CT>& P, CT (&a)[4], int r, int g, double lbound, bool& bad) {
  constexpr int K = k & 7, J = k >> 3;
  const CT* pr = P.prow[k & 1];
  const CT d = pr[k];
  const bool zero = d.x == 0 && d.y == 0;
  CT inv = crecip_fast(d);
  inv.x = zero ? 0 : inv.x;
  inv.y = zero ? 0 : inv.y;
  bad |= zero;
  // A[r][k] lives in lane (r, K) of this row's group of 8, slot J
  const CT ark = swz8c<K>(a[J]);
  if (r > k) {
    const CT l = cmul(ark, inv);
    if (!(cabs1(l) <= lbound)) bad = true;
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (g + 8 * j > k) a[j] = cfms(a[j], l, pr[g + 8 * j]);
    if (g == K) a[J] = l;
  }
  if (r == k + 1) {
#pragma unroll
    for (int j = 0; j < 4; ++j) P.prow[(k + 1) & 1][g + 8 * j] = a[j];
  }
  __syncthreads();
}

template <int k, typename CT>
MW_DEV void lu32_steps(PanelLds<CT>& P, CT (&a)[4], int r, int g, double lbound, bool& bad) {
  if constexpr (k < NB - 1) {
    lu32_step<k>(P, a, r, g, lbound, bad);
    lu32_steps<k + 1>(P, a, r, g, lbound, bad);
  }
}

// LU without row exchanges of the NB x NB block in P.M (identity-padded),
// all 256 threads: thread t holds row t / 8, columns t % 8 + 8 j (j < 4, a
// cyclic split so the shrinking trailing block stays spread over every
// thread).  Step k: the pivot row comes from LDS (written by its 8 threads
// at the end of step k - 1, double-buffered, so one barrier per step), each
// row's multiplier from the group's lane holding column k (ds_swizzle, no
// LDS round trip).  The entries and their update order are the textbook
// right-looking elimination's.  A zero pivot zeroes the step's multipliers;
// it, or a multiplier beyond lbound or not finite, sets P.bad.  Ends with a
// barrier; P.bad must be initialised before the call.
template <typename CT>
MW_DEV void lu32_eliminate(PanelLds<CT>& P, double lbound) {
  const int tid = threadIdx.x;
  const int r = tid >> 3, g = tid & 7;
  CT a[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) a[j] = P.M[r][g + 8 * j];
  if (r == 0) {
#pragma unroll
    for (int j = 0; j < 4; ++j) P.prow[0][g + 8 * j] = a[j];
  }
  __syncthreads();
  bool bad = false;
  lu32_steps<0>(P, a, r, g, lbound, bad);
  const CT last = P.prow[(NB - 1) & 1][NB - 1];
  bad |= (last.x == 0 && last.y == 0);
#pragma unroll
  for (int j = 0; j < 4; ++j) P.M[r][g + 8 * j] = a[j];
  if (__any(bad) && (tid & 63) == 0) P.bad = 1;
  __syncthreads();
}

// P.Li = L^{-1} and P.Ui = U^{-1} of the factors in P.M (L\U, padded to NB
// with the identity), by recursive doubling over diagonal blocks of size s =
// 1, 2, 4, 8, 16: for each pair of s-blocks,
//   [[A, 0], [C, B]]^{-1} = [[A^{-1}, 0], [-B^{-1} (C A^{-1}), B^{-1}]]   (L)
//   [[A, C], [0, B]]^{-1} = [[A^{-1}, -A^{-1} (C B^{-1})], [0, B^{-1}]]   (U)
// two dependent products per level (the bracket into P.T, then the block),
// so five short levels replace the 32-step substitutions.  A zero pivot
// leaves a zero diagonal inverse.  Called by every thread; ends with a
// barrier.
template <typename CT>
MW_DEV void lu32_inverses(PanelLds<CT>& P) {
  const int tid = threadIdx.x;
  for (int e = tid; e < NB * NB; e += blockDim.x) {
    const int i = e >> 5, j = e & 31;
    CT li, ui;
    li.x = (i == j) ? 1 : 0;
    li.y = 0;
    ui.x = 0;
    ui.y = 0;
    if (i == j) {
      const CT d = P.M[i][i];
      if (d.x != 0 || d.y != 0) ui = crecip(d);
    }
    P.Li[i][j] = li;
    P.Ui[i][j] = ui;
  }
  __syncthreads();
#pragma unroll
  for (int s = 1; s < NB; s *= 2) {
    const int cnt = (NB / 2) * s;  // entries of one bracket matrix, all pairs
    for (int e = tid; e < 2 * cnt; e += blockDim.x) {
      const int upper = e >= cnt, q = e - upper * cnt;
      const int p = q / (s * s), i = (q / s) % s, j = q % s;
      const int c0 = 2 * s * p, r0 = c0 + s;
      CT acc;
      acc.x = 0;
      acc.y = 0;
      if (!upper) {
        for (int m = j; m < s; ++m) acc = cfma(acc, P.M[r0 + i][c0 + m], P.Li[c0 + m][c0 + j]);
      } else {
        for (int m = 0; m <= j; ++m) acc = cfma(acc, P.M[c0 + i][r0 + m], P.Ui[r0 + m][r0 + j]);
      }
      P.T[upper][q] = acc;
    }
    __syncthreads();
    for (int e = tid; e < 2 * cnt; e += blockDim.x) {
      const int upper = e >= cnt, q = e - upper * cnt;
      const int p = q / (s * s), i = (q / s) % s, j = q % s;
      const int c0 = 2 * s * p, r0 = c0 + s;
      const CT* Tp = &P.T[upper][p * s * s];
      CT acc;
      acc.x = 0;
      acc.y = 0;
      if (!upper) {
        for (int m = 0; m <= i; ++m) acc = cfma(acc, P.Li[r0 + i][r0 + m], Tp[m * s + j]);
        P.Li[r0 + i][c0 + j].x = -acc.x;
        P.Li[r0 + i][c0 + j].y = -acc.y;
      } else {
        for (int m = i; m < s; ++m) acc = cfma(acc, P.Ui[c0 + i][c0 + m], Tp[m * s + j]);
        P.Ui[c0 + i][r0 + j].x = -acc.x;
        P.Ui[c0 + i][r0 + j].y = -acc.y;
      }
    }
    __syncthreads();
  }
}

constexpr int AW = 64;  // rows (or columns) per workgroup in the block products

// Block products with a 32 x 32 complex factor T on MFMA, one workgroup:
//   ROWS (RA = 64, CB = 32): Out (64 x 32) = In (64 x 32) T
//   COLS (RA = 32, CB = 64): Out (32 x 64) = T In (32 x 64)
// as the real product A_r (RA x 64) B_exp (64 x 2 CB), A_r the left factor's
// (re, im) pairs and B_exp the right factor with each b expanded to
// [[br, bi], [-bi, br]] when fragments are formed.  The 16 output tiles of
// 16 x 16 are four per wave.  In is read whole into LDS before anything is
// written, so Out may alias In.  Rows/columns of In beyond (nrow, ncol) read
// as zero; Out is written only inside them (or not at all when check: then
// the result is returned as "some |entry| > lbound").
template <typename T> struct BlkLds {
  T As[64][65];                       // left factor as reals
  typename cplx_t<T>::type Bs[32][65];  // right factor as complex
};

template <typename T, int RA, int CB, bool ROWS = (RA == 64)>
MW_DEV bool block_product(T (*As)[65], typename cplx_t<T>::type* Bs, int ldbs,
                          const typename cplx_t<T>::type* Tg, int ldt,
                          typename cplx_t<T>::type* X, int64_t ldx, int nrow, int ncol,
                          bool check, double lbound) {
  typedef typename cplx_t<T>::type CT;
  typedef typename Mfma<T>::acc_t acc_t;
  constexpr int TPW = TileGrid<RA, CB>::TPW;
  static_assert(ROWS ? CB == 32 : RA == 32, "T is 32 x 32");
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // stage: left factor (reals) and right factor (complex)
  if (ROWS) {
    for (int e = tid; e < RA * 32; e += 256) {
      const int r = e >> 5, c = e & 31;
      CT v;
      v.x = 0; v.y = 0;
      if (r < nrow && c < ncol) v = X[(int64_t)r * ldx + c];
      As[r][2 * c] = v.x;
      As[r][2 * c + 1] = v.y;
    }
    if (Tg)  // else T is already in Bs
      for (int e = tid; e < 32 * 32; e += 256) Bs[(e >> 5) * ldbs + (e & 31)] = Tg[(e >> 5) * ldt + (e & 31)];
  } else {
    for (int e = tid; e < 32 * 32; e += 256) {
      const CT v = Tg[(e >> 5) * ldt + (e & 31)];
      As[e >> 5][2 * (e & 31)] = v.x;
      As[e >> 5][2 * (e & 31) + 1] = v.y;
    }
    for (int e = tid; e < 32 * CB; e += 256) {
      const int r = e / CB, c = e % CB;
      CT v;
      v.x = 0; v.y = 0;
      if (r < nrow && c < ncol) v = X[(int64_t)r * ldx + c];
      Bs[r * ldbs + c] = v;
    }
  }
  __syncthreads();
  constexpr int TC = TileGrid<RA, CB>::TC;
  acc_t acc[TPW];
  lds_mfma<T, RA, CB>(&As[0][0], Bs, ldbs, acc);
  bool over = false;
  T* Xr = reinterpret_cast<T*>(X);
#pragma unroll
  for (int j = 0; j < TPW; ++j) {
    const int t = TPW * wave + j;
    const int tr = t / TC, tc = t % TC;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = 16 * tr + Mfma<T>::row(lane, r);
      const int cr = 16 * tc + (lane & 15);
      const bool in = row < nrow && (cr >> 1) < ncol;
      if (check) {
        // |re| + |im| of the complex entry: its two parts sit in lanes l, l ^ 1
        const double mag = fabs((double)acc[j][r]) + fabs((double)__shfl_xor(acc[j][r], 1));
        if (in && !(mag <= lbound)) over = true;
      } else if (in) {
        Xr[(int64_t)row * 2 * ldx + cr] = acc[j][r];
      }
    }
  }
  return over;
}

// Write L11\U11 (from M) and the two inverses for panel use and for the solve.
template <typename CT>
MW_DEV void store_factors(const PanelLds<CT>& P, int bw, CT* lu11, CT* inv) {
  const int tid = threadIdx.x;
  for (int e = tid; e < NB * NB; e += blockDim.x) {
    const int i = e / NB, j = e % NB;
    lu11[e] = P.M[i][j];
    inv[e] = P.Li[i][j];
    inv[NB * NB + e] = P.Ui[i][j];
  }
  (void)bw;
}

// Fast path of a panel: factor A11 without row exchanges (every workgroup
// redundantly, in LDS), invert its factors, and check that every multiplier
// of the panel -- those of A11 and L21 = A21 U11^{-1} -- stays within lbound
// (threshold pivoting with tau = 1/lbound).  Nothing in A is written; a
// violated bound (or a zero pivot) sets flag for the tournament fallback.
// Block 0 stores L11\U11, the inverses and identity interchanges; block b > 0
// checks 64 rows of L21, formed on MFMA from U11^{-1}.
template <typename T>
MW_DEV void panel_nopiv_body(PanelLds<typename cplx_t<T>::type>& P, int b, int64_t n,
                             const typename cplx_t<T>::type* A, int64_t lda, int64_t k0, int bw,
                             typename cplx_t<T>::type* lu11, typename cplx_t<T>::type* inv,
                             int32_t* ipiv, int32_t* flag, double lbound) {
  typedef typename cplx_t<T>::type CT;
  typedef typename Mfma<T>::acc_t acc_t;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // the factor is a latency chain: let it issue ahead of update waves
  // sharing its SIMD (when it overlaps a trailing update)
  __builtin_amdgcn_s_setprio(3);
  for (int e = tid; e < NB * NB; e += blockDim.x) {
    const int i = e / NB, j = e % NB;
    CT v;
    v.x = (i == j) ? 1 : 0; v.y = 0;
    if (i < bw && j < bw) v = A[(k0 + i) * lda + k0 + j];
    P.M[i][j] = v;
  }
  if (tid == 0) P.bad = (lbound <= 0.0) ? 1 : 0;
  // a checking block's 64 rows of A21, loaded now so that they arrive while
  // the factor runs
  const int64_t r0 = k0 + bw + (int64_t)(b - 1) * 64;
  CT a21[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int e = tid + 256 * q, r = e >> 5, c = e & 31;
    a21[q].x = 0;
    a21[q].y = 0;
    if (b > 0 && r0 + r < n && c < bw) a21[q] = A[(r0 + r) * lda + k0 + c];
  }
  __syncthreads();
  lu32_eliminate(P, lbound);
  lu32_inverses(P);
  if (b == 0) {
    store_factors(P, bw, lu11, inv);
    if (tid < bw) ipiv[k0 + tid] = (int32_t)(k0 + tid);
    if (tid == 0 && P.bad) atomicOr(flag, 1);
    return;
  }
  // L21 = A21 U11^{-1} for the 64 rows; A21 as reals over M and Li (64 rows
  // of 65), U11^{-1} read from Ui
  static_assert(offsetof(PanelLds<CT>, Li) == sizeof(P.M) &&
                    sizeof(P.M) + sizeof(P.Li) >= sizeof(T) * 64 * 65,
                "64 rows of 65 reals fit over M and Li");
  T* As = reinterpret_cast<T*>(&P.M[0][0]);
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int e = tid + 256 * q, r = e >> 5, c = e & 31;
    As[r * 65 + 2 * c] = a21[q].x;
    As[r * 65 + 2 * c + 1] = a21[q].y;
  }
  __syncthreads();
  constexpr int TC = TileGrid<64, 32>::TC, TPW = TileGrid<64, 32>::TPW;
  acc_t acc[TPW];
  lds_mfma<T, 64, 32>(As, &P.Ui[0][0], NB + 1, acc);
  bool over = false;
#pragma unroll
  for (int j = 0; j < TPW; ++j) {
    const int t = TPW * wave + j, tr = t / TC, tc = t % TC;
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      const int row = 16 * tr + Mfma<T>::row(lane, rr);
      const int cr = 16 * tc + (lane & 15);
      // |re| + |im| of the complex entry: its two parts sit in lanes l, l ^ 1
      const double mag = fabs((double)acc[j][rr]) + fabs((double)__shfl_xor(acc[j][rr], 1));
      if (r0 + row < n && (cr >> 1) < bw && !(mag <= lbound)) over = true;
    }
  }
  if (__any(over) && lane == 0) atomicOr(flag, 1);
}

template <typename T>
__global__ __launch_bounds__(256, 2) void k_panel_nopiv(int64_t n, const typename cplx_t<T>::type* A,
                                                     int64_t lda, int64_t k0, int bw,
                                                     typename cplx_t<T>::type* lu11,
                                                     typename cplx_t<T>::type* inv,
                                                     int32_t* ipiv, int32_t* flag, double lbound) {
  __shared__ PanelLds<typename cplx_t<T>::type> P;
  panel_nopiv_body<T>(P, blockIdx.x, n, A, lda, k0, bw, lu11, inv, ipiv, flag, lbound);
}

// Tournament fallback for a panel whose diagonal failed the threshold test:
// one launch.  Leaves factor 128-row chunks with partial pivoting in
// registers; the last of up to four siblings to finish (device-scope arrival
// counter) merges their nominees' original rows, and so on up to the root,
// which writes L11\U11, the inverses and the interchanges, and applies the
// interchanges to columns [k0, n).  Exits at once when the flag is clear.
struct TourWs {
  int32_t* cand;     // [nodes][NB]
  int32_t* ncand;    // [nodes]
  int32_t* counter;  // [nodes], zero on entry, reset by the root
};

// LDS of the tournament fallback (from the caller, so that a kernel running
// several phases can share one buffer between them).
template <typename CT>
struct PivotLds {
  TourLds<CT> tour;
  PanelLds<CT> panel;
  int last;
  int32_t pos[2 * NB], val[2 * NB];
  int nt;
};

template <typename T>
MW_DEV void panel_pivot_body(PivotLds<typename cplx_t<T>::type>& SL, int node0,
                             typename cplx_t<T>::type* A, int64_t lda, int64_t n,
                             int64_t c_lo, int64_t nc, int64_t k0, int bw, int nleaf, TourWs ws,
                             const int32_t* flag, typename cplx_t<T>::type* lu11,
                             typename cplx_t<T>::type* inv, int32_t* ipiv, int32_t* info) {
  typedef typename cplx_t<T>::type CT;
  if (*flag == 0) return;
  TourLds<CT>& L = SL.tour;
  PanelLds<CT>& P = SL.panel;
  int& last_s = SL.last;
  const int tid = threadIdx.x;
  int node = node0, nodes = nleaf, lvl_off = 0;
  if (tid == 0) {
    const int64_t r0 = k0 + (int64_t)node * TROWS;
    const int64_t left = n - r0;
    const int rows = (int)(left < TROWS ? left : TROWS);
    for (int r = 0; r < rows; ++r) L.rid[r] = (int32_t)(r0 + r);
    L.rows = rows;
  }
  __syncthreads();
  CT S[RPL][8];
  while (true) {
    const int rows = L.rows;
    load_block(S, A, lda, k0, bw, L, rows);
    reg_gepp(S, rows, bw, L, nodes == 1 ? info : (int32_t*)nullptr, k0);
    __syncthreads();
    const int nom = rows < bw ? rows : bw;
    if (nodes == 1) break;
    if (tid < nom) ws.cand[(size_t)(lvl_off + node) * NB + tid] = L.rid[L.piv[tid]];
    if (tid == 0) ws.ncand[lvl_off + node] = nom;
    __threadfence();
    __syncthreads();
    const int parent = node / FANIN;
    const int pnodes = (nodes + FANIN - 1) / FANIN;
    const int nchild = (parent * FANIN + FANIN <= nodes) ? FANIN : nodes - parent * FANIN;
    const int plvl_off = lvl_off + nodes;
    if (tid == 0) {
      const int old = __hip_atomic_fetch_add(&ws.counter[plvl_off + parent], 1, __ATOMIC_ACQ_REL,
                                             __HIP_MEMORY_SCOPE_AGENT);
      last_s = (old == nchild - 1);
    }
    __syncthreads();
    if (!last_s) return;
    __threadfence();
    if (tid == 0) {
      int cnt = 0;
      for (int ch = 0; ch < nchild; ++ch) {
        const int cn = parent * FANIN + ch;
        const int nc = __hip_atomic_load(&ws.ncand[lvl_off + cn], __ATOMIC_ACQUIRE,
                                         __HIP_MEMORY_SCOPE_AGENT);
        for (int t = 0; t < nc; ++t)
          L.rid[cnt++] = __hip_atomic_load(&ws.cand[(size_t)(lvl_off + cn) * NB + t],
                                           __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
      }
      L.rows = cnt;
    }
    __syncthreads();
    node = parent;
    lvl_off = plvl_off;
    nodes = pnodes;
  }
  // Root.  P.M <- L11\U11 of the pivot rows in pivot order (identity-padded).
  const int rows = L.rows;
  const int steps = rows < bw ? rows : bw;
  const int lane = tid & 63, w = tid >> 6;
  for (int e = tid; e < NB * NB; e += blockDim.x) {
    CT v;
    v.x = (e / NB == e % NB) ? 1 : 0; v.y = 0;
    P.M[e / NB][e % NB] = v;
  }
  __syncthreads();
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
          if (8 * w + j < bw) P.M[i][8 * w + j] = S[t][j];
      }
    }
  }
  // final placement of the rows this panel moves: position pos_s[t] receives
  // the row that stood at val_s[t] (a permutation of at most 2 bw rows)
  int32_t* pos_s = SL.pos;
  int32_t* val_s = SL.val;
  int& nt_s = SL.nt;
  if (tid == 0) {
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
    if (steps < bw && *info == 0) *info = (int32_t)(k0 + steps + 1);
    nt_s = nt;
  }
  __syncthreads();
  lu32_inverses(P);
  store_factors(P, bw, lu11, inv);
  // reset the arrival counters for the next fallback
  int total = 0;
  for (int nd = nleaf; nd > 1; nd = (nd + FANIN - 1) / FANIN) total += nd;
  total += 1;
  for (int e = tid; e < total; e += blockDim.x) ws.counter[e] = 0;
  // interchanges on columns [c_lo, nc): from the start of the current outer
  // block, so the block's multiplier columns end up in the block's final row
  // order; nobody else reads A any more.  The table applies them as one
  // gather and one scatter per 32-column strip (a thread holds 8 of the
  // strip's <= 64 x 32 entries between the two), not as bw dependent swaps.
  __syncthreads();
  const int nt = nt_s;
  if (nt == 0) return;
  for (int64_t c0 = c_lo; c0 < nc; c0 += 32) {
    CT v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int e = tid + 256 * j, t = e >> 5;
      const int64_t c = c0 + (e & 31);
      if (t < nt && c < nc) v[j] = A[(int64_t)val_s[t] * lda + c];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int e = tid + 256 * j, t = e >> 5;
      const int64_t c = c0 + (e & 31);
      if (t < nt && c < nc) A[(int64_t)pos_s[t] * lda + c] = v[j];
    }
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_panel_pivot(typename cplx_t<T>::type* A, int64_t lda,
                                                     int64_t n, int64_t c_lo, int64_t nc,
                                                     int64_t k0, int bw, int nleaf,
                                                     TourWs ws, const int32_t* flag,
                                                     typename cplx_t<T>::type* lu11,
                                                     typename cplx_t<T>::type* inv,
                                                     int32_t* ipiv, int32_t* info) {
  __shared__ PivotLds<typename cplx_t<T>::type> S;
  panel_pivot_body<T>(S, blockIdx.x, A, lda, n, c_lo, nc, k0, bw, nleaf, ws, flag, lu11, inv,
                      ipiv, info);
}

// After the interchanges: A11 <- L11\U11, L21 <- A21 U11^{-1}, U12 <- L11^{-1} A12.
// Blocks [0, nrb): 64-row blocks of L21; [nrb, nrb+ncb): 64-column blocks of
// U12; last: A11.
template <typename T>
MW_DEV void panel_apply_body(BlkLds<T>& L, int b, int64_t n, int64_t nc,
                             typename cplx_t<T>::type* A, int64_t lda, int64_t k0, int bw,
                             const typename cplx_t<T>::type* inv,
                             const typename cplx_t<T>::type* lu11, int nrb, int ncb) {
  const int tid = threadIdx.x;
  if (b < nrb) {
    const int64_t r0 = k0 + bw + (int64_t)b * AW;
    const int rows = (int)((n - r0) < AW ? (n - r0) : AW);
    block_product<T, 64, 32>(L.As, &L.Bs[0][0], 65, inv + NB * NB, NB, A + r0 * lda + k0, lda, rows,
                             bw, false, 0.0);
  } else if (b < nrb + ncb) {
    const int64_t c0 = k0 + bw + (int64_t)(b - nrb) * AW;
    const int cols = (int)((nc - c0) < AW ? (nc - c0) : AW);
    block_product<T, 32, 64>(L.As, &L.Bs[0][0], 65, inv, NB, A + k0 * lda + c0, lda, bw, cols, false,
                             0.0);
  } else {
    for (int e = tid; e < bw * bw; e += blockDim.x)
      A[(k0 + e / bw) * lda + k0 + e % bw] = lu11[(e / bw) * NB + e % bw];
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_panel_apply(int64_t n, int64_t nc, typename cplx_t<T>::type* A,
                                                     int64_t lda, int64_t k0, int bw,
                                                     const typename cplx_t<T>::type* inv,
                                                     const typename cplx_t<T>::type* lu11,
                                                     int nrb, int ncb) {
  __shared__ BlkLds<T> L;
  panel_apply_body<T>(L, blockIdx.x, n, nc, A, lda, k0, bw, inv, lu11, nrb, ncb);
}

// ---------------------------------------------------------------------------
// Look-ahead for the one-level factorisation.  The step after panel k
// (columns k0 .. kn = k0 + bw) splits its trailing update in two:
//  - k_panel_apply_la: besides L21, U12 and A11, each 64-row L21 block
//    updates its own rows of the next panel's columns [kn, kn + bwn):
//    A -= L21 U12n, with U12n = L11^{-1} A12n formed in the block from the
//    untouched A12n.  The U12 blocks start right of the next panel.
//  - k_lookahead: the next panel's factor and threshold check (which read
//    only the next panel's columns, final by now) run in the same launch as
//    the rest of the trailing update (columns right of the next panel), with
//    one more block that stores U12n in A.
// The next panel's 32 x 32 elimination, a latency chain of one wave, thus
// overlaps the bulk of the update instead of following it.
// ---------------------------------------------------------------------------
constexpr int LAW = 32;  // rows of an L21 block, columns of a U12 block

template <typename T>
struct LaLds {
  T As[LAW][65];                               // A21, then L21, as reals
  typename cplx_t<T>::type Bs[2][NB][NB + 1];  // U11^{-1} then U12n | A12n
  T Ls[NB][65];                                // L11^{-1} (symmetric: A21 top) as reals
  typename cplx_t<T>::type D[NB];              // symmetric: U11's diagonal
};

//
// SYM (Z = Z^T, no row exchanges: U = D L^T): the trailing update keeps only
// the lower part, the panel-diagonal blocks and the block right of each (the
// next panel's A12) up to date; so every block but the first, whose rows are
// the next panel's, also stores the U12 columns of its own rows, D L21^T, in
// place of the stale upper part (the first block's columns are the next
// panel's U12, which k_lookahead forms from A12 as without symmetry).
template <typename T, bool SYM = false>
MW_DEV void l21_lookahead_body(LaLds<T>& L, int b, int64_t n, typename cplx_t<T>::type* A,
                               int64_t lda, int64_t k0, int bw, int bwn,
                               const typename cplx_t<T>::type* inv,
                               const typename cplx_t<T>::type* lu11 = nullptr) {
  typedef typename cplx_t<T>::type CT;
  typedef typename Mfma<T>::acc_t acc_t;
  // 32 x 32 products: tiles t = 2 wave + j of a 2 x 4 grid
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t r0 = k0 + bw + (int64_t)b * LAW;
  const int rows = (int)((n - r0) < LAW ? (n - r0) : LAW);
  const int64_t kn = k0 + bw;
  T* Ar = reinterpret_cast<T*>(A);
  // this block's rows of the next panel, loaded with the operands (no load
  // waits on a store later on; every global store comes at the end)
  acc_t nxt[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int t = 2 * wave + j, tr = t >> 2, tc = t & 3;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = 16 * tr + Mfma<T>::row(lane, r);
      const int cr = 16 * tc + (lane & 15);
      nxt[j][r] = (row < rows && (cr >> 1) < bwn) ? Ar[(r0 + row) * 2 * lda + 2 * kn + cr] : (T)0;
    }
  }
  for (int e = tid; e < NB * NB; e += 256) {
    const int r = e >> 5, c = e & 31;
    CT v;
    v.x = 0; v.y = 0;
    if (r < rows && c < bw) v = A[(r0 + r) * lda + k0 + c];
    L.As[r][2 * c] = v.x;
    L.As[r][2 * c + 1] = v.y;
    L.Bs[0][r][c] = inv[NB * NB + e];
    const CT li = inv[e];
    L.Ls[r][2 * c] = li.x;
    L.Ls[r][2 * c + 1] = li.y;
    CT w;
    w.x = 0; w.y = 0;
    if (r < bw && c < bwn) w = A[(k0 + r) * lda + kn + c];
    L.Bs[1][r][c] = w;
  }
  if (SYM && tid < NB) {
    CT d;
    d.x = 0; d.y = 0;
    if (tid < bw) d = lu11[tid * NB + tid];
    L.D[tid] = d;
  }
  __syncthreads();
  acc_t l21[2], u12[2];
  lds_mfma<T, LAW, 32>(&L.As[0][0], &L.Bs[0][0][0], NB + 1, l21);
  lds_mfma<T, 32, 32>(&L.Ls[0][0], &L.Bs[1][0][0], NB + 1, u12);
  __syncthreads();
  // L21 as reals into As (its real columns are As's layout), U12n into Bs[0]
  T* Bt = reinterpret_cast<T*>(&L.Bs[0][0][0]);
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int t = 2 * wave + j, tr = t >> 2, tc = t & 3;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = 16 * tr + Mfma<T>::row(lane, r);
      const int cr = 16 * tc + (lane & 15);
      L.As[row][cr] = l21[j][r];
      Bt[row * 2 * (NB + 1) + cr] = u12[j][r];
    }
  }
  __syncthreads();
  if (SYM && b > 0) {
    // this block's U12 columns: U[k0 + i][r0 + row] = D_i L21[row][i]
    for (int e = tid; e < NB * LAW; e += 256) {
      const int i = e / LAW, row = e % LAW;
      if (i < bw && row < rows) {
        CT l;
        l.x = L.As[row][2 * i];
        l.y = L.As[row][2 * i + 1];
        A[(k0 + i) * lda + r0 + row] = cmul(L.D[i], l);
      }
    }
  }
  acc_t upd[2];
  lds_mfma<T, LAW, 32>(&L.As[0][0], &L.Bs[0][0][0], NB + 1, upd);
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int t = 2 * wave + j, tr = t >> 2, tc = t & 3;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = 16 * tr + Mfma<T>::row(lane, r);
      const int cr = 16 * tc + (lane & 15);
      if (row < rows && (cr >> 1) < bw) Ar[(r0 + row) * 2 * lda + 2 * k0 + cr] = l21[j][r];
      if (row < rows && (cr >> 1) < bwn)
        Ar[(r0 + row) * 2 * lda + 2 * kn + cr] = nxt[j][r] - upd[j][r];
    }
  }
}

// Blocks [0, nrb): L21 with the look-ahead update, LAW rows each;
// [nrb, nrb + ncb): U12 on LAW-column blocks from kn + bwn; last: A11.  Small
// blocks: each one's MFMA work is on the panel chain's critical path.
// SYM: the L21 blocks also store U12 (D L21^T) for the columns of their rows,
// so the U12 blocks cover only the right-hand-side columns [n, nc).
template <typename T, bool SYM = false>
__global__ __launch_bounds__(256) void k_panel_apply_la(int64_t n, int64_t nc,
                                                        typename cplx_t<T>::type* A, int64_t lda,
                                                        int64_t k0, int bw, int bwn,
                                                        const typename cplx_t<T>::type* inv,
                                                        const typename cplx_t<T>::type* lu11,
                                                        int nrb, int ncb) {
  __shared__ LaLds<T> L;
  const int b = blockIdx.x;
  if (b < nrb) {
    l21_lookahead_body<T, SYM>(L, b, n, A, lda, k0, bw, bwn, inv, lu11);
  } else if (b < nrb + ncb) {
    const int64_t cfirst = SYM ? ((k0 + bw + bwn) > n ? (k0 + bw + bwn) : n) : k0 + bw + bwn;
    const int64_t c0 = cfirst + (int64_t)(b - nrb) * LAW;
    const int cols = (int)((nc - c0) < LAW ? (nc - c0) : LAW);
    block_product<T, 32, LAW, false>(L.As, &L.Bs[0][0][0], NB + 1, inv, NB, A + k0 * lda + c0, lda,
                                     bw, cols, false, 0.0);
  } else {
    for (int e = threadIdx.x; e < bw * bw; e += blockDim.x)
      A[(k0 + e / bw) * lda + k0 + e % bw] = lu11[(e / bw) * NB + e % bw];
  }
}

// C -= A B on 64 x 32 (complex) tiles with kdim <= 32: the trailing update
// inside the look-ahead launch, whose occupancy the panel factor's registers
// hold to two workgroups per CU.  So each update workgroup walks tiles
// b, b + G, b + 2G, ... and loads the next tile's C, A and B into registers
// while the MFMAs of the current one run: a tile's global latency hides
// behind the previous tile's work instead of behind other workgroups.
template <typename T>
struct K32Lds {
  T As[64][65];                          // -A as reals
  typename cplx_t<T>::type Bs[NB][NB + 1];
};

template <typename T>
MW_DEV void gemm_k32_tiles(K32Lds<T>& L, int b, int G, int ntiles, int gxn, int64_t m, int64_t n,
                           int kdim, const typename cplx_t<T>::type* __restrict__ A, int64_t lda,
                           const typename cplx_t<T>::type* __restrict__ B, int64_t ldb,
                           typename cplx_t<T>::type* C, int64_t ldc) {
  typedef typename cplx_t<T>::type CT;
  typedef typename Mfma<T>::acc_t acc_t;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  T* Cr = reinterpret_cast<T*>(C);
  acc_t cn[4];
  CT ra[8], rb[4];
  // tile t: rows 64 (t / gxn) .., complex columns 32 (t % gxn) ..; acc[j][r]
  // at row 16 wave + Mfma row(lane, r), real column 16 j + lane % 16
  auto load = [&](int t) {
    const int64_t row0 = (int64_t)(t / gxn) * 64, col0 = (int64_t)(t % gxn) * NB;
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gr = row0 + 16 * wave + Mfma<T>::row(lane, r);
        const int cr = 16 * j + (lane & 15);
        cn[j][r] = (gr < m && col0 + (cr >> 1) < n) ? Cr[gr * 2 * ldc + 2 * col0 + cr] : (T)0;
      }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int e = tid + 256 * q, r = e >> 5, c = e & 31;
      CT v;
      v.x = 0; v.y = 0;
      if (row0 + r < m && c < kdim) v = A[(row0 + r) * lda + c];
      ra[q] = v;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = tid + 256 * q, r = e >> 5, c = e & 31;
      CT v;
      v.x = 0; v.y = 0;
      if (r < kdim && col0 + c < n) v = B[r * ldb + col0 + c];
      rb[q] = v;
    }
  };
  int t = b;
  if (t >= ntiles) return;
  load(t);
  while (true) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int e = tid + 256 * q, r = e >> 5, c = e & 31;
      L.As[r][2 * c] = -ra[q].x;
      L.As[r][2 * c + 1] = -ra[q].y;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = tid + 256 * q;
      L.Bs[e >> 5][e & 31] = rb[q];
    }
    acc_t acc[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[j] = cn[j];
    const int cur = t;
    t += G;
    __syncthreads();
    if (t < ntiles) load(t);
    lds_mfma<T, 64, 32, false>(&L.As[0][0], &L.Bs[0][0], NB + 1, acc);
    const int64_t row0 = (int64_t)(cur / gxn) * 64, col0 = (int64_t)(cur % gxn) * NB;
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gr = row0 + 16 * wave + Mfma<T>::row(lane, r);
        const int cr = 16 * j + (lane & 15);
        if (gr < m && col0 + (cr >> 1) < n) Cr[gr * 2 * ldc + 2 * col0 + cr] = acc[j][r];
      }
    if (t >= ntiles) break;
    __syncthreads();  // LDS free for the next tile's operands
  }
}

// C -= A B on 64 x 32 (complex) tiles with any kdim, by chunks of 32: one
// tile's C stays in registers while its K chunks pass through LDS, and the
// next chunk's operands (and the next tile's C) load while the current
// chunk's MFMAs run.  One workgroup walks every tile.
template <typename T>
MW_DEV void gemm_kc_tiles(K32Lds<T>& L, int ntiles, int gxn, int64_t m, int64_t n, int kdim,
                          const typename cplx_t<T>::type* __restrict__ A, int64_t lda,
                          const typename cplx_t<T>::type* __restrict__ B, int64_t ldb,
                          typename cplx_t<T>::type* C, int64_t ldc) {
  typedef typename cplx_t<T>::type CT;
  typedef typename Mfma<T>::acc_t acc_t;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nch = (kdim + NB - 1) / NB;
  const int total = ntiles * nch;
  if (total <= 0) return;
  T* Cr = reinterpret_cast<T*>(C);
  acc_t cn[4], acc[4];
  CT ra[8], rb[4];
  // item it: tile it / nch, K chunk it % nch; a tile's first chunk also
  // brings its C
  auto load = [&](int it) {
    const int t = it / nch, q = it % nch;
    const int64_t row0 = (int64_t)(t / gxn) * 64, col0 = (int64_t)(t % gxn) * NB;
    const int64_t k0 = (int64_t)q * NB;
    if (q == 0) {
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t gr = row0 + 16 * wave + Mfma<T>::row(lane, r);
          const int cr = 16 * j + (lane & 15);
          cn[j][r] = (gr < m && col0 + (cr >> 1) < n) ? Cr[gr * 2 * ldc + 2 * col0 + cr] : (T)0;
        }
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = tid + 256 * u, r = e >> 5, c = e & 31;
      CT v;
      v.x = 0; v.y = 0;
      if (row0 + r < m && k0 + c < kdim) v = A[(row0 + r) * lda + k0 + c];
      ra[u] = v;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = tid + 256 * u, r = e >> 5, c = e & 31;
      CT v;
      v.x = 0; v.y = 0;
      if (k0 + r < kdim && col0 + c < n) v = B[(k0 + r) * ldb + col0 + c];
      rb[u] = v;
    }
  };
  load(0);
  for (int it = 0;;) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = tid + 256 * u, r = e >> 5, c = e & 31;
      L.As[r][2 * c] = -ra[u].x;
      L.As[r][2 * c + 1] = -ra[u].y;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = tid + 256 * u;
      L.Bs[e >> 5][e & 31] = rb[u];
    }
    if (it % nch == 0) {
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[j] = cn[j];
    }
    const int cur = it++;
    __syncthreads();
    if (it < total) load(it);
    lds_mfma<T, 64, 32, false>(&L.As[0][0], &L.Bs[0][0], NB + 1, acc);
    if (cur % nch == nch - 1) {
      const int t = cur / nch;
      const int64_t row0 = (int64_t)(t / gxn) * 64, col0 = (int64_t)(t % gxn) * NB;
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t gr = row0 + 16 * wave + Mfma<T>::row(lane, r);
          const int cr = 16 * j + (lane & 15);
          if (gr < m && col0 + (cr >> 1) < n) Cr[gr * 2 * ldc + 2 * col0 + cr] = acc[j][r];
        }
    }
    if (it >= total) break;
    __syncthreads();  // LDS free for the next chunk's operands
  }
}

// The symmetric look-ahead's trailing update: rows [kn, n) against the
// columns [kn + bwn, n) of the square part, only the 64 x 32 tiles whose first
// column is at most 32 past their last row (so the lower part, the
// panel-diagonal blocks and the block right of each stay up to date; the
// rest lies further above and is replaced by U rows later), then every tile
// of the right-hand-side columns [n, nc).  Square row tr holds
// min(gxs, 2 tr + q + 1) tiles (q = floor((95 - bwn) / 32): 1 for a full
// panel, 2 for a narrower one), so the first tstar rows hold tr^2 + q tr
// tiles before row tr and the rest gxs each.
struct SymTiles {
  int64_t kn, c0s, n, nc;
  int64_t gxs, gxr, nrt, q, tstar, s_star, s_sq;
  __host__ __device__ SymTiles(int64_t n_, int64_t nc_, int64_t kn_, int bwn) {
    n = n_;
    nc = nc_;
    kn = kn_;
    c0s = kn_ + bwn;
    const int64_t ns = n_ - c0s;
    gxs = ns > 0 ? (ns + NB - 1) / NB : 0;
    gxr = nc_ > n_ ? (nc_ - n_ + NB - 1) / NB : 0;
    nrt = (n_ - kn_ + 63) / 64;
    q = (95 - bwn) / 32;
    tstar = gxs > q ? (gxs - q) / 2 : 0;
    if (tstar > nrt) tstar = nrt;
    s_star = tstar * tstar + q * tstar;
    s_sq = gxs > 0 ? s_star + (nrt - tstar) * gxs : 0;
  }
  __host__ __device__ int64_t total() const { return s_sq + nrt * gxr; }
  // tile t: rows row0 .. row0 + 64, columns col0 .. col0 + 32, below colend
  __device__ void tile(int64_t t, int64_t& row0, int64_t& col0, int64_t& colend) const {
    int64_t tr, tc;
    if (t < s_star) {
      tr = (int64_t)((sqrt((double)(q * q + 4 * t)) - (double)q) * 0.5);
      while ((tr + 1) * (tr + 1) + q * (tr + 1) <= t) ++tr;
      while (tr > 0 && tr * tr + q * tr > t) --tr;
      tc = t - (tr * tr + q * tr);
      col0 = c0s + NB * tc;
      colend = n;
    } else if (t < s_sq) {
      const int64_t u = t - s_star;
      tr = tstar + u / gxs;
      tc = u % gxs;
      col0 = c0s + NB * tc;
      colend = n;
    } else {
      const int64_t u = t - s_sq;
      tr = u / gxr;
      tc = u % gxr;
      col0 = n + NB * tc;
      colend = nc;
    }
    row0 = kn + 64 * tr;
  }
};

// gemm_k32_tiles over SymTiles: C = A[row0.., col0..] -= A[row0.., k0..k0+bw)
// A[k0.., col0..) for each tile, workgroup b walking tiles b, b + G, ...
template <typename T>
MW_DEV void gemm_k32_sym(K32Lds<T>& L, int b, int G, const SymTiles& S, int kdim,
                         typename cplx_t<T>::type* A, int64_t lda, int64_t k0) {
  typedef typename cplx_t<T>::type CT;
  typedef typename Mfma<T>::acc_t acc_t;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  T* Ar = reinterpret_cast<T*>(A);
  const int64_t ntiles = S.total();
  acc_t cn[4];
  CT ra[8], rb[4];
  int64_t crow = 0, ccol = 0, cend = 0;  // the tile whose operands are loaded
  auto load = [&](int64_t t) {
    S.tile(t, crow, ccol, cend);
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gr = crow + 16 * wave + Mfma<T>::row(lane, r);
        const int cr = 16 * j + (lane & 15);
        cn[j][r] = (gr < S.n && ccol + (cr >> 1) < cend) ? Ar[gr * 2 * lda + 2 * ccol + cr] : (T)0;
      }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int e = tid + 256 * q, r = e >> 5, c = e & 31;
      CT v;
      v.x = 0; v.y = 0;
      if (crow + r < S.n && c < kdim) v = A[(crow + r) * lda + k0 + c];
      ra[q] = v;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = tid + 256 * q, r = e >> 5, c = e & 31;
      CT v;
      v.x = 0; v.y = 0;
      if (r < kdim && ccol + c < cend) v = A[(k0 + r) * lda + ccol + c];
      rb[q] = v;
    }
  };
  int64_t t = b;
  if (t >= ntiles) return;
  load(t);
  while (true) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int e = tid + 256 * q, r = e >> 5, c = e & 31;
      L.As[r][2 * c] = -ra[q].x;
      L.As[r][2 * c + 1] = -ra[q].y;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = tid + 256 * q;
      L.Bs[e >> 5][e & 31] = rb[q];
    }
    acc_t acc[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[j] = cn[j];
    const int64_t row0 = crow, col0 = ccol, colend = cend;
    t += G;
    __syncthreads();
    if (t < ntiles) load(t);
    lds_mfma<T, 64, 32, false>(&L.As[0][0], &L.Bs[0][0], NB + 1, acc);
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gr = row0 + 16 * wave + Mfma<T>::row(lane, r);
        const int cr = 16 * j + (lane & 15);
        if (gr < S.n && col0 + (cr >> 1) < colend) Ar[gr * 2 * lda + 2 * col0 + cr] = acc[j][r];
      }
    if (t >= ntiles) break;
    __syncthreads();  // LDS free for the next tile's operands
  }
}

// C -= A B on 64 x 64 (complex) tiles, FP64, any kdim by chunks of 32, for
// one workgroup walking every tile (the persistent sweep's updates).  Waves
// 2 x 2: wave (wr, wc) holds rows 32 wr .. + 32 and real columns 64 wc .. +
// 64, so a k step reads 2 A and 4 B fragments for 8 MFMAs (the 64 x 32 tile
// of gemm_kc_tiles reads 1 and 4 for 4).  Operands of the next chunk load
// into registers while a chunk's MFMAs run; a tile's C loads into the
// accumulators as soon as the previous tile is stored.
struct K64Lds {
  double As[64][65];     // -A as reals: row r, real k
  double2 Bs[NB][65];    // k row, complex column
};

MW_DEV void gemm_kc64_tiles(K64Lds& L, int ntiles, int gxn, int64_t m, int64_t n, int kdim,
                            const double2* __restrict__ A, int64_t lda,
                            const double2* __restrict__ B, int64_t ldb, double2* C,
                            int64_t ldc) {
  typedef Mfma<double>::acc_t acc_t;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int nch = (kdim + NB - 1) / NB;
  const int total = ntiles * nch;
  if (total <= 0) return;
  double* Cr = reinterpret_cast<double*>(C);
  acc_t acc[2][4];
  double2 ra[8], rb[8];
  auto load_c = [&](int t) {
    const int64_t row0 = (int64_t)(t / gxn) * 64, col0 = (int64_t)(t % gxn) * 64;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t gr = row0 + 32 * wr + 16 * i + Mfma<double>::row(lane, r);
          const int cr = 64 * wc + 16 * j + (lane & 15);
          acc[i][j][r] =
              (gr < m && col0 + (cr >> 1) < n) ? Cr[gr * 2 * ldc + 2 * col0 + cr] : 0.0;
        }
  };
  auto load = [&](int it) {
    const int t = it / nch, q = it % nch;
    const int64_t row0 = (int64_t)(t / gxn) * 64, col0 = (int64_t)(t % gxn) * 64;
    const int64_t k0 = (int64_t)q * NB;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = tid + 256 * u, r = e >> 5, c = e & 31;
      double2 v = make_double2(0.0, 0.0);
      if (row0 + r < m && k0 + c < kdim) v = A[(row0 + r) * lda + k0 + c];
      ra[u] = v;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = tid + 256 * u, r = e >> 6, c = e & 63;
      double2 v = make_double2(0.0, 0.0);
      if (k0 + r < kdim && col0 + c < n) v = B[(k0 + r) * ldb + col0 + c];
      rb[u] = v;
    }
  };
  load_c(0);
  load(0);
  for (int it = 0;;) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = tid + 256 * u, r = e >> 5, c = e & 31;
      L.As[r][2 * c] = -ra[u].x;
      L.As[r][2 * c + 1] = -ra[u].y;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = tid + 256 * u;
      L.Bs[e >> 6][e & 63] = rb[u];
    }
    const int cur = it++;
    __syncthreads();
    if (it < total) load(it);
#pragma unroll
    for (int ks = 0; ks < 2 * NB; ks += 4) {
      const int kr = ks + (lane >> 4), kpar = kr & 1;
      double a[2], b[4];
#pragma unroll
      for (int i = 0; i < 2; ++i) a[i] = L.As[32 * wr + 16 * i + (lane & 15)][kr];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        // real row kr, real column cr of [[br, bi], [-bi, br]]: one double
        const int cr = 64 * wc + 16 * j + (lane & 15), odd = cr & 1;
        const double v = reinterpret_cast<const double*>(&L.Bs[kr >> 1][cr >> 1])[odd ^ kpar];
        b[j] = (kpar && !odd) ? -v : v;
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = Mfma<double>::mma(a[i], b[j], acc[i][j]);
    }
    if (cur % nch == nch - 1) {
      const int t = cur / nch;
      const int64_t row0 = (int64_t)(t / gxn) * 64, col0 = (int64_t)(t % gxn) * 64;
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int64_t gr = row0 + 32 * wr + 16 * i + Mfma<double>::row(lane, r);
            const int cr = 64 * wc + 16 * j + (lane & 15);
            if (gr < m && col0 + (cr >> 1) < n) Cr[gr * 2 * ldc + 2 * col0 + cr] = acc[i][j][r];
          }
      if (it < total) load_c(t + 1);
    }
    if (it >= total) break;
    __syncthreads();  // LDS free for the next chunk's operands
  }
}

template <typename T>
struct NextLds {
  typedef typename cplx_t<T>::type CT;
  static constexpr size_t mx(size_t a, size_t b) { return a > b ? a : b; }
  static constexpr size_t bytes =
      mx(mx(sizeof(PanelLds<CT>), sizeof(K32Lds<T>)), sizeof(T) * NB * 65 + sizeof(CT) * NB * 65);
};

// Blocks [0, nnp): the next panel's k_panel_nopiv; nnp: U12n into A; then G
// blocks for the trailing update of rows [kn, n) x columns [kn + bwn, nc)
// (ntiles 64 x 32 tiles, gxn per row of tiles).  One LDS buffer serves
// whichever part a block runs.  The factor's blocks come first so that they
// are dispatched first.
// SYM: the update over SymTiles.
template <typename T, bool SYM = false>
__global__ __launch_bounds__(256, 2) void k_lookahead(int64_t n, int64_t nc,
                                                   typename cplx_t<T>::type* A, int64_t lda,
                                                   int64_t k0, int bw, int bwn,
                                                   const typename cplx_t<T>::type* inv,
                                                   typename cplx_t<T>::type* lu11,
                                                   typename cplx_t<T>::type* inv_next,
                                                   int32_t* ipiv, int32_t* flag_next,
                                                   double lbound, int nnp, int gxn,
                                                   int ntiles, int G) {
  typedef typename cplx_t<T>::type CT;
  __shared__ __align__(16) unsigned char smem[NextLds<T>::bytes];
  const int64_t kn = k0 + bw;
  int b = blockIdx.x;
  if (b < nnp) {
    panel_nopiv_body<T>(*reinterpret_cast<PanelLds<CT>*>(smem), b, n, A, lda, kn, bwn, lu11,
                        inv_next, ipiv, flag_next, lbound);
    return;
  }
  b -= nnp;
  if (b == 0) {
    block_product<T, 32, 64>(reinterpret_cast<T (*)[65]>(smem),
                             reinterpret_cast<CT*>(smem + sizeof(T) * NB * 65), 65, inv, NB,
                             A + k0 * lda + kn, lda, bw, bwn, false, 0.0);
    return;
  }
  b -= 1;
  if (SYM) {
    gemm_k32_sym<T>(*reinterpret_cast<K32Lds<T>*>(smem), b, G, SymTiles(n, nc, kn, bwn), bw, A,
                    lda, k0);
    return;
  }
  const int64_t c0 = kn + bwn;
  gemm_k32_tiles<T>(*reinterpret_cast<K32Lds<T>*>(smem), b, G, ntiles, gxn, n - kn, nc - c0, bw,
                    A + kn * lda + k0, lda, A + k0 * lda + c0, lda, A + kn * lda + c0, lda);
}

// Solve blocks on AW right-hand-side columns per workgroup:
// lower != 0: X_p <- L11^{-1} X_p; lower == 0: X_p <- U11^{-1} X_p.
template <typename T>
MW_DEV void solve_apply_body(BlkLds<T>& L, int bx, int64_t nrhs,
                             const typename cplx_t<T>::type* inv, int64_t k0, int bw,
                             typename cplx_t<T>::type* X, int64_t ldx, int lower) {
  const int64_t c0 = (int64_t)bx * AW;
  const int cols = (int)((nrhs - c0) < AW ? (nrhs - c0) : AW);
  block_product<T, 32, 64>(L.As, &L.Bs[0][0], 65, lower ? inv : inv + NB * NB, NB,
                           X + k0 * ldx + c0, ldx, bw, cols, false, 0.0);
}

template <typename T>
__global__ __launch_bounds__(256) void k_solve_apply(int64_t nrhs, const typename cplx_t<T>::type* inv,
                                                     int64_t k0, int bw,
                                                     typename cplx_t<T>::type* X, int64_t ldx,
                                                     int lower) {
  __shared__ BlkLds<T> L;
  solve_apply_body<T>(L, blockIdx.x, nrhs, inv, k0, bw, X, ldx, lower);
}

// One step of the back substitution for a single panel p (an odd last panel
// of the two-panel scheme below): every workgroup forms X_p' = U_pp^{-1} X_p
// for its 32 right-hand-side columns in LDS (MFMA, from the panel's stored
// inverse) and then updates its 64-row tile of the rows above,
// X[rows][cols] -= U[rows][p] X_p'.  X_p itself is left as it is: nothing
// later changes it, and a launch at the end replaces it by U_pp^{-1} X_p.
template <typename T>
__global__ __launch_bounds__(256) void k_bwd_fused(int64_t k0, int bw, int64_t nrhs,
                                                   const typename cplx_t<T>::type* __restrict__ A,
                                                   int64_t lda,
                                                   const typename cplx_t<T>::type* __restrict__ uinv,
                                                   typename cplx_t<T>::type* X, int64_t ldx) {
  typedef typename cplx_t<T>::type CT;
  typedef typename Mfma<T>::acc_t acc_t;
  __shared__ T As[64][65];     // U_pp^{-1} (phase 1), then the A tile (phase 2), as reals
  __shared__ CT Bs[32][33];    // X_p (phase 1), then X_p' (phase 2)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t c0 = (int64_t)blockIdx.x * 32;
  const int64_t r0 = (int64_t)blockIdx.y * 64;
  const int cols = (int)((nrhs - c0) < 32 ? (nrhs - c0) : 32);
  const int rows = (int)((k0 - r0) < 64 ? (k0 - r0) : 64);
  // the phase-2 tiles of U and X, loaded with the phase-1 operands
  T* Xr = reinterpret_cast<T*>(X);
  acc_t cx[4];
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = 16 * wave + Mfma<T>::row(lane, r);
      const int cr = 16 * j + (lane & 15);
      cx[j][r] = (row < rows && (cr >> 1) < cols) ? Xr[(r0 + row) * 2 * ldx + 2 * c0 + cr] : (T)0;
    }
  CT ra[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int e = tid + 256 * q, r = e >> 5, c = e & 31;
    CT v;
    v.x = 0; v.y = 0;
    if (r < rows && c < bw) v = A[(r0 + r) * lda + k0 + c];
    ra[q] = v;
  }
  for (int e = tid; e < 32 * 32; e += 256) {
    const int r = e >> 5, c = e & 31;
    const CT u = uinv[e];
    As[r][2 * c] = u.x;
    As[r][2 * c + 1] = u.y;
    CT v;
    v.x = 0; v.y = 0;
    if (r < bw && c < cols) v = X[(k0 + r) * ldx + c0 + c];
    Bs[r][c] = v;
  }
  __syncthreads();
  // phase 1: X_p' (32 x 64 reals, 8 tiles: two per wave)
  acc_t acc[4];
#pragma unroll
  for (int j = 0; j < 2; ++j)
#pragma unroll
    for (int r = 0; r < 4; ++r) acc[j][r] = 0;
#pragma unroll
  for (int ks = 0; ks < 64; ks += 4) {
    const int kr = ks + (lane >> 4), kc = kr >> 1, kpar = kr & 1;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int t = 2 * wave + j, tr = t >> 2, tc = t & 3;
      const T a = As[16 * tr + (lane & 15)][kr];
      const int cr = 16 * tc + (lane & 15);
      const CT bv = Bs[kc][cr >> 1];
      const T bb = (kpar == 0) ? ((cr & 1) ? bv.y : bv.x) : ((cr & 1) ? bv.x : -bv.y);
      acc[j] = Mfma<T>::mma(a, bb, acc[j]);
    }
  }
  __syncthreads();  // every wave is done reading U^{-1} and X_p
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int t = 2 * wave + j, tr = t >> 2, tc = t & 3;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = 16 * tr + Mfma<T>::row(lane, r);
      const int cr = 16 * tc + (lane & 15);
      T* dst = (cr & 1) ? &Bs[row][cr >> 1].y : &Bs[row][cr >> 1].x;
      *dst = (row < bw) ? acc[j][r] : (T)0;
    }
  }
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int e = tid + 256 * q, r = e >> 5, c = e & 31;
    As[r][2 * c] = ra[q].x;
    As[r][2 * c + 1] = ra[q].y;
  }
  __syncthreads();
  // phase 2: tile (64 x 64 reals, 16 tiles: four per wave) -= A_tile X_p'
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int r = 0; r < 4; ++r) acc[j][r] = 0;
#pragma unroll
  for (int ks = 0; ks < 64; ks += 4) {
    const int kr = ks + (lane >> 4), kc = kr >> 1, kpar = kr & 1;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const T a = As[16 * wave + (lane & 15)][kr];
      const int cr = 16 * j + (lane & 15);
      const CT bv = Bs[kc][cr >> 1];
      const T bb = (kpar == 0) ? ((cr & 1) ? bv.y : bv.x) : ((cr & 1) ? bv.x : -bv.y);
      acc[j] = Mfma<T>::mma(a, bb, acc[j]);
    }
  }
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = 16 * wave + Mfma<T>::row(lane, r);
      const int cr = 16 * j + (lane & 15);
      if (row < rows && (cr >> 1) < cols)
        Xr[(r0 + row) * 2 * ldx + 2 * c0 + cr] = cx[j][r] - acc[j][r];
    }
}

// The panel inverses in the workspace (the host's panel_inv): L^{-1}, then
// U^{-1}, NB x NB each, 16-byte slots.
template <typename T>
MW_DEV const typename cplx_t<T>::type* panel_uinv(const void* work, int64_t k0) {
  return reinterpret_cast<const typename cplx_t<T>::type*>(
             reinterpret_cast<const char*>(work) + (int64_t)NB * NB * 16 * (1 + 2 * (k0 / NB))) +
         NB * NB;
}

// ---------------------------------------------------------------------------
// Back substitution, two panels per step.  Panels pair up from the top,
// (0, 1), (2, 3), ...; an odd last panel steps alone first (k_bwd_fused).
// For a pair (q, p = q + 1) at rows [kq, kq + NB + bwp):
//   X_p' = U_pp^{-1} X_p,   X_q' = U_qq^{-1} (X_q - U_qp X_p'),
// formed in LDS by every workgroup of the step for its 32 columns; each then
// updates a 64-row tile of the rows above with K = 64:
//   X[rows] -= [U_rq U_rp] [X_q'; X_p'].
// The pair's own rows are left as they are (nothing later changes them);
// k_bwd2_final forms every pair's X' at the end.  Half the launches of a
// one-panel step, each a chain of global loads, three 32 x 32 products and
// the update.
// ---------------------------------------------------------------------------
template <typename T>
struct PairLds {
  T U0[NB][65];                                // U_pp^{-1}, then U_qq^{-1}, as reals
  T U1[NB][65];                                // U_qp as reals
  typename cplx_t<T>::type Xs[2 * NB][NB + 1];  // [X_q; X_p] -> [X_q'; X_p']
};

template <typename T>
MW_DEV void pair_stage_x(PairLds<T>& L, int64_t kq, int bwp, const typename cplx_t<T>::type* X,
                         int64_t ldx, int64_t c0, int cols) {
  typedef typename cplx_t<T>::type CT;
  for (int e = threadIdx.x; e < 2 * NB * NB; e += 256) {
    const int r = e >> 5, c = e & 31;
    CT v;
    v.x = 0; v.y = 0;
    if (c < cols && (r < NB || r - NB < bwp)) v = X[(kq + r) * ldx + c0 + c];
    L.Xs[r][c] = v;
  }
}

// [X_q; X_p] in L.Xs -> [X_q'; X_p'].  Starts with a barrier-closed staging
// round, so the caller's own staging needs no barrier of its own.
template <typename T>
MW_DEV void pair_solve(PairLds<T>& L, const void* work, const typename cplx_t<T>::type* A,
                       int64_t lda, int64_t kq, int bwp) {
  typedef typename cplx_t<T>::type CT;
  typedef typename Mfma<T>::acc_t acc_t;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const CT* uq = panel_uinv<T>(work, kq);
  const CT* up = panel_uinv<T>(work, kq + NB);
  // every global operand is loaded here, so no later step waits on memory
  CT rq[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int e = tid + 256 * q, r = e >> 5, c = e & 31;
    const CT u = up[e];
    L.U0[r][2 * c] = u.x;
    L.U0[r][2 * c + 1] = u.y;
    CT w;
    w.x = 0; w.y = 0;
    if (c < bwp) w = A[(kq + r) * lda + kq + NB + c];
    L.U1[r][2 * c] = w.x;
    L.U1[r][2 * c + 1] = w.y;
    rq[q] = uq[e];
  }
  __syncthreads();
  T* Xt = reinterpret_cast<T*>(&L.Xs[0][0]);
  constexpr int LX = 2 * (NB + 1);  // row stride of Xs in reals
  acc_t acc[2];
  lds_mfma<T, 32, 32>(&L.U0[0][0], &L.Xs[NB][0], NB + 1, acc);  // X_p'
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int t = 2 * wave + j, tr = t >> 2, tc = t & 3;
#pragma unroll
    for (int r = 0; r < 4; ++r)
      Xt[(NB + 16 * tr + Mfma<T>::row(lane, r)) * LX + 16 * tc + (lane & 15)] = acc[j][r];
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int e = tid + 256 * q;
    L.U0[e >> 5][2 * (e & 31)] = rq[q].x;
    L.U0[e >> 5][2 * (e & 31) + 1] = rq[q].y;
  }
  __syncthreads();
  lds_mfma<T, 32, 32>(&L.U1[0][0], &L.Xs[NB][0], NB + 1, acc);  // U_qp X_p'
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int t = 2 * wave + j, tr = t >> 2, tc = t & 3;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      T& x = Xt[(16 * tr + Mfma<T>::row(lane, r)) * LX + 16 * tc + (lane & 15)];
      x = x - acc[j][r];
    }
  }
  __syncthreads();
  lds_mfma<T, 32, 32>(&L.U0[0][0], &L.Xs[0][0], NB + 1, acc);  // X_q'
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int t = 2 * wave + j, tr = t >> 2, tc = t & 3;
#pragma unroll
    for (int r = 0; r < 4; ++r)
      Xt[(16 * tr + Mfma<T>::row(lane, r)) * LX + 16 * tc + (lane & 15)] = acc[j][r];
  }
  __syncthreads();
}

template <typename T>
struct Bwd2Lds {
  PairLds<T> P;
  T As[64][4 * NB + 1];  // -[U_rq U_rp] of the tile's rows, as reals
};

// One step: pair (q, q + 1) at kq; blockIdx.x: 32 columns, blockIdx.y: a
// 64-row tile of the rows above kq.
template <typename T>
__global__ __launch_bounds__(256) void k_bwd2(int64_t kq, int bwp, int64_t nrhs,
                                              const typename cplx_t<T>::type* __restrict__ A,
                                              int64_t lda, const void* work,
                                              typename cplx_t<T>::type* X, int64_t ldx) {
  typedef typename cplx_t<T>::type CT;
  typedef typename Mfma<T>::acc_t acc_t;
  __shared__ Bwd2Lds<T> L;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t c0 = (int64_t)blockIdx.x * NB;
  const int64_t r0 = (int64_t)blockIdx.y * 64;
  const int cols = (int)((nrhs - c0) < NB ? (nrhs - c0) : NB);
  const int rows = (int)((kq - r0) < 64 ? (kq - r0) : 64);
  const int kw = NB + bwp;  // complex k of the update
  // the tile to update first, in accumulator layout (tile t = wave, j)
  T* Xr = reinterpret_cast<T*>(X);
  acc_t c[4];
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = 16 * wave + Mfma<T>::row(lane, r);
      const int cr = 16 * j + (lane & 15);
      c[j][r] = (row < rows && (cr >> 1) < cols) ? Xr[(r0 + row) * 2 * ldx + 2 * c0 + cr] : (T)0;
    }
  for (int e = tid; e < 64 * 2 * NB; e += 256) {
    const int r = e >> 6, k = e & 63;
    CT v;
    v.x = 0; v.y = 0;
    if (r < rows && k < kw) v = A[(r0 + r) * lda + kq + k];
    L.As[r][2 * k] = -v.x;
    L.As[r][2 * k + 1] = -v.y;
  }
  pair_stage_x(L.P, kq, bwp, X, ldx, c0, cols);
  pair_solve(L.P, work, A, lda, kq, bwp);
  lds_mfma<T, 64, 32, false, 4 * NB, 4 * NB + 1>(&L.As[0][0], &L.P.Xs[0][0], NB + 1, c);
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = 16 * wave + Mfma<T>::row(lane, r);
      const int cr = 16 * j + (lane & 15);
      if (row < rows && (cr >> 1) < cols) Xr[(r0 + row) * 2 * ldx + 2 * c0 + cr] = c[j][r];
    }
}

// One pair's X' (blockIdx.x: 32 columns), written over the pair's rows.
template <typename T>
__global__ __launch_bounds__(256) void k_bwd2_pair(int64_t kq, int bwp, int64_t nrhs,
                                                   const typename cplx_t<T>::type* __restrict__ A,
                                                   int64_t lda, const void* work,
                                                   typename cplx_t<T>::type* X, int64_t ldx) {
  __shared__ PairLds<T> L;
  const int64_t c0 = (int64_t)blockIdx.x * NB;
  const int cols = (int)((nrhs - c0) < NB ? (nrhs - c0) : NB);
  pair_stage_x(L, kq, bwp, X, ldx, c0, cols);
  pair_solve(L, work, A, lda, kq, bwp);
  for (int e = threadIdx.x; e < 2 * NB * NB; e += 256) {
    const int r = e >> 5, c = e & 31;
    if (c < cols && (r < NB || r - NB < bwp)) X[(kq + r) * ldx + c0 + c] = L.Xs[r][c];
  }
}

// Every pair's X' (blockIdx.y: pair, blockIdx.x: 32 columns).
template <typename T>
__global__ __launch_bounds__(256) void k_bwd2_final(int64_t n, int64_t nrhs,
                                                    const typename cplx_t<T>::type* __restrict__ A,
                                                    int64_t lda, const void* work,
                                                    typename cplx_t<T>::type* X, int64_t ldx) {
  __shared__ PairLds<T> L;
  const int64_t kq = (int64_t)blockIdx.y * 2 * NB;
  const int bwp = (int)((n - kq - NB) < NB ? (n - kq - NB) : NB);
  const int64_t c0 = (int64_t)blockIdx.x * NB;
  const int cols = (int)((nrhs - c0) < NB ? (nrhs - c0) : NB);
  pair_stage_x(L, kq, bwp, X, ldx, c0, cols);
  pair_solve(L, work, A, lda, kq, bwp);
  for (int e = threadIdx.x; e < 2 * NB * NB; e += 256) {
    const int r = e >> 5, c = e & 31;
    if (c < cols && (r < NB || r - NB < bwp)) X[(kq + r) * ldx + c0 + c] = L.Xs[r][c];
  }
}

// Row interchanges ipiv[r0 .. r1) applied in order to columns [0, ncols) of X.
template <typename T>
__global__ void k_laswp_rows(typename cplx_t<T>::type* X, int64_t ldx, int64_t ncols,
                             const int32_t* ipiv, int64_t r0, int64_t r1) {
  typedef typename cplx_t<T>::type CT;
  // the interchanges staged in LDS 1024 at a time (one read each, not one
  // per thread and row)
  __shared__ int32_t ps[1024];
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int64_t i0 = r0; i0 < r1; i0 += 1024) {
    const int cnt = (int)((r1 - i0) < 1024 ? (r1 - i0) : 1024);
    __syncthreads();
    for (int e = threadIdx.x; e < cnt; e += blockDim.x) ps[e] = ipiv[i0 + e];
    __syncthreads();
    if (c >= ncols) continue;
    for (int e = 0; e < cnt; ++e) {
      const int64_t i = i0 + e, p = ps[e];
      if (p != i) {
        const CT t = X[i * ldx + c]; X[i * ldx + c] = X[p * ldx + c]; X[p * ldx + c] = t;
      }
    }
  }
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
MW_DEV void sparse_rows_body(int64_t i, int64_t j, int64_t nrows, int maxnz, const int32_t* idx,
                             const double* val, const double2* X, int64_t ldx, int64_t ncols,
                             double2* out, int64_t ldo) {
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

__global__ void k_sparse_rows(int64_t nrows, int maxnz, const int32_t* idx, const double* val,
                              const double2* X, int64_t ldx, int64_t ncols, double2* out,
                              int64_t ldo) {
  sparse_rows_body(blockIdx.y, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, nrows, maxnz, idx,
                   val, X, ldx, ncols, out, ldo);
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

// Z[rows[e], cols[e]] += vals[e] for distinct (row, col) pairs: the wire
// loading term, merged per frequency on the host.
__global__ void k_add_sparse(int64_t nnz, const int32_t* rows, const int32_t* cols,
                             const double2* vals, double2* Z, int64_t ldz) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nnz) return;
  double2* z = Z + (int64_t)rows[e] * ldz + cols[e];
  const double2 v = vals[e];
  *z = make_double2(z->x + v.x, z->y + v.y);
}

// Precision changes between the FP64 system and the FP32 factorisation.
__global__ void k_z2c(int64_t rows, int64_t cols, const double2* src, int64_t lds, float2* dst,
                      int64_t ldd) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= rows * cols) return;
  const int64_t r = e / cols, c = e % cols;
  const double2 v = src[r * lds + c];
  dst[r * ldd + c] = make_float2((float)v.x, (float)v.y);
}

__global__ void k_c2z(int64_t rows, int64_t cols, const float2* src, int64_t lds, double2* dst,
                      int64_t ldd, int accumulate) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= rows * cols) return;
  const int64_t r = e / cols, c = e % cols;
  const float2 v = src[r * lds + c];
  double2 o = make_double2((double)v.x, (double)v.y);
  if (accumulate) {
    const double2 d = dst[r * ldd + c];
    o.x += d.x;
    o.y += d.y;
  }
  dst[r * ldd + c] = o;
}

// The mixed form's one-residual readout, Y = B^T X0 + X0^T R + C^T R with
// R = B - Z X0 (FP64) and C = Z32^{-1} R (the FP32 factors): its error is
// (Z^{-1} R - C)^T R, about delta |C^T R| with delta = max_j |C_j| / |X0_j|
// (the FP32 solve's relative error, the same for C as for X0).  Three
// kernels: per-column sums of squares, Y += T with max |T| and max |Y|, and
// the verdict.  Maxima of non-negative doubles are kept as their bit
// patterns (ordered as the values), so they can be combined by atomicMax.
__global__ __launch_bounds__(256) void k_colsq2(int64_t n, int64_t p, const double2* X,
                                               int64_t ldx, const double2* C, int64_t ldc,
                                               int64_t rows_per_blk, double* acc) {
  __shared__ double red[2][4][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t j = (int64_t)blockIdx.x * 64 + lane;
  const int64_t r0 = (int64_t)blockIdx.y * rows_per_blk;
  const int64_t r1 = (r0 + rows_per_blk) < n ? r0 + rows_per_blk : n;
  double sx = 0.0, sc = 0.0;
  if (j < p) {
    for (int64_t r = r0 + w; r < r1; r += 4) {
      const double2 x = X[r * ldx + j], c = C[r * ldc + j];
      sx += x.x * x.x + x.y * x.y;
      sc += c.x * c.x + c.y * c.y;
    }
  }
  red[0][w][lane] = sx;
  red[1][w][lane] = sc;
  __syncthreads();
  if (w == 0 && j < p) {
    sx = (red[0][0][lane] + red[0][1][lane]) + (red[0][2][lane] + red[0][3][lane]);
    sc = (red[1][0][lane] + red[1][1][lane]) + (red[1][2][lane] + red[1][3][lane]);
    atomicAdd(acc + 2 * j, sx);
    atomicAdd(acc + 2 * j + 1, sc);
  }
}

__global__ __launch_bounds__(256) void k_add_maxabs(int64_t p, const double2* T, int64_t ldt,
                                                   double2* Y, int64_t ldy,
                                                   unsigned long long* mx) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double t2 = 0.0, y2 = 0.0;
  if (e < p * p) {
    const int64_t r = e / p, c = e % p;
    const double2 t = T[r * ldt + c];
    double2 y = Y[r * ldy + c];
    y.x += t.x;
    y.y += t.y;
    Y[r * ldy + c] = y;
    t2 = t.x * t.x + t.y * t.y;
    y2 = y.x * y.x + y.y * y.y;
  }
  // the quiet-NaN pattern orders above every finite value and infinity, so a
  // NaN survives the maxima and fails the verdict's comparison
  const unsigned long long qnan = 0x7FF8000000000000ULL;
  unsigned long long tb = (t2 == t2) ? (unsigned long long)__double_as_longlong(t2) : qnan;
  unsigned long long yb = (y2 == y2) ? (unsigned long long)__double_as_longlong(y2) : qnan;
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long to = __shfl_xor(tb, off), yo = __shfl_xor(yb, off);
    tb = to > tb ? to : tb;
    yb = yo > yb ? yo : yb;
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMax(mx, tb);
    atomicMax(mx + 1, yb);
  }
}

__global__ __launch_bounds__(256) void k_readout_verdict(int64_t p, const double* acc,
                                                        const unsigned long long* mx, double tol,
                                                        double* stats, int32_t* gate) {
  __shared__ double red[256];
  double q = 0.0;
  for (int64_t j = threadIdx.x; j < p; j += 256) {
    const double x = acc[2 * j], c = acc[2 * j + 1];
    double r = x > 0.0 ? c / x : (c == 0.0 ? 0.0 : INFINITY);
    if (!(r <= DBL_MAX)) r = INFINITY;  // NaN or overflow: no estimate
    q = fmax(q, r);
  }
  red[threadIdx.x] = q;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double delta = sqrt(red[0]);
    const double tmax = sqrt(__longlong_as_double((long long)mx[0]));
    const double ymax = sqrt(__longlong_as_double((long long)mx[1]));
    const double est = delta * tmax;
    stats[0] = delta;
    stats[1] = tmax;
    stats[2] = ymax;
    stats[3] = est;
    // NaN anywhere fails the comparison: the caller then takes a fresh residual
    gate[0] = (est <= tol * ymax) ? 0 : 1;
  }
}

// T <- S^{-1} T, one workgroup, Gauss-Jordan with partial pivoting in LDS
// (ns <= 32, nt <= 64).
struct SmallLds {
  double2 M[32][97];
  int piv;
};

MW_DEV void small_solve_body(SmallLds& SL, int ns, int nt, double2* S, int64_t lds, double2* Tm,
                             int64_t ldt, int32_t* info) {
  double2 (*M)[97] = SL.M;
  int& piv_s = SL.piv;
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

__global__ __launch_bounds__(256) void k_small_solve(int ns, int nt, double2* S, int64_t lds,
                                                     double2* Tm, int64_t ldt, int32_t* info) {
  __shared__ SmallLds SL;
  small_solve_body(SL, ns, nt, S, lds, Tm, ldt, info);
}

// ---------------------------------------------------------------------------
// Batched sweep: one launch per step serves every (structure, wavenumber)
// problem; grid z (or y) indexes the problem, problems are sorted by size so
// a step launches only the prefix still active.  Single-level blocking
// (problems stay below 4096 unknowns), FP64.
// ---------------------------------------------------------------------------
MW_DEV int64_t d_tour_nodes(int64_t n) {
  int64_t nodes = 0;
  for (int64_t nd = (n + TROWS - 1) / TROWS; nd > 1; nd = (nd + FANIN - 1) / FANIN) nodes += nd;
  return nodes + 1;
}
MW_DEV double2* ws_inv(void* work, int64_t k0) {
  return reinterpret_cast<double2*>(reinterpret_cast<char*>(work) +
                                    (int64_t)NB * NB * 16 * (1 + 2 * (k0 / NB)));
}
MW_DEV int32_t* ws_flags(void* work, int64_t n) {
  const int64_t npan = (n + NB - 1) / NB;
  return reinterpret_cast<int32_t*>(reinterpret_cast<char*>(work) +
                                    (int64_t)NB * NB * 16 * (1 + 2 * npan));
}
MW_DEV TourWs ws_tour(void* work, int64_t n) {
  TourWs tw;
  tw.counter = ws_flags(work, n) + (n + NB - 1) / NB;
  const int64_t nodes = d_tour_nodes(n);
  tw.ncand = tw.counter + nodes;
  tw.cand = tw.ncand + nodes;
  return tw;
}

__global__ void kb_zero_ws(const mw_sweep_prob* P) {
  const mw_sweep_prob& p = P[blockIdx.y];
  const int64_t cnt = (p.n + NB - 1) / NB + d_tour_nodes(p.n);
  int32_t* f = ws_flags(p.work, p.n);
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < cnt;
       e += (int64_t)gridDim.x * blockDim.x)
    f[e] = 0;
  if (blockIdx.x == 0 && threadIdx.x < 2) p.info[threadIdx.x] = 0;
}

__global__ void kb_rhs(const mw_sweep_prob* P, const mw_ports* ports, int scatter) {
  const mw_sweep_prob& p = P[blockIdx.y];
  const mw_ports& q = ports[p.mesh];
  double2* X = reinterpret_cast<double2*>(p.A) + p.n;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (!scatter) {
    if (e < p.n * p.ncols) X[(e / p.ncols) * p.lda + e % p.ncols] = make_double2(0.0, 0.0);
  } else if (e < p.ncols * q.col_max) {
    const int32_t r = q.col_idx[e];
    if (r >= 0) X[(int64_t)r * p.lda + e / q.col_max] = make_double2(q.col_val[e], 0.0);
  }
}

// Batched threshold check, in two launches.  The single-problem kernel lets
// every workgroup refactor A11 so that no workgroup waits for another; over
// thousands of problems that repeated 32-step chain is the whole cost, so
// here one workgroup per problem factors A11 (storing L11\U11, the inverses
// and identity interchanges, A11's own multipliers checked on the way) ...
__global__ __launch_bounds__(256, 2) void kb_panel_a11(const mw_sweep_prob* P, int64_t k0,
                                                    double lbound) {
  const mw_sweep_prob& p = P[blockIdx.x];
  if (k0 >= p.n) return;
  const int bw = (int)((p.n - k0) < NB ? (p.n - k0) : NB);
  __shared__ PanelLds<double2> L;
  panel_nopiv_body<double>(L, 0, p.n, (const double2*)p.A, p.lda, k0, bw, (double2*)p.work,
                           ws_inv(p.work, k0), p.ipiv, ws_flags(p.work, p.n) + k0 / NB, lbound);
}

// ... and then each workgroup forms 64 rows of L21 = A21 U11^{-1} from the
// stored inverse (a product, no chain) and flags the panel if a multiplier
// exceeds lbound or is not finite.  Nothing in A is written.
__global__ __launch_bounds__(256) void kb_panel_check(const mw_sweep_prob* P, int64_t k0,
                                                      double lbound, const uint32_t* items) {
  const uint32_t it = items[blockIdx.x];
  const mw_sweep_prob& p = P[it >> 16];
  if (k0 >= p.n) return;
  const int bw = (int)((p.n - k0) < NB ? (p.n - k0) : NB);
  const int64_t r0 = k0 + bw + (int64_t)(it & 0xffff) * 64;
  if (r0 >= p.n) return;
  int32_t* flag = ws_flags(p.work, p.n) + k0 / NB;
  if (*flag) return;  // A11 already failed: the tournament runs anyway
  __shared__ double2 a[64][NB + 1];
  __shared__ double2 u[NB][NB + 1];
  const double2* A = (const double2*)p.A;
  const double2* ui = ws_inv(p.work, k0) + NB * NB;
  const int tid = threadIdx.x;
  for (int e = tid; e < 64 * NB; e += blockDim.x) {
    const int r = e >> 5, c = e & 31;
    double2 v = make_double2(0.0, 0.0);
    if (r0 + r < p.n && c < bw) v = A[(r0 + r) * p.lda + k0 + c];
    a[r][c] = v;
  }
  for (int e = tid; e < NB * NB; e += blockDim.x) u[e >> 5][e & 31] = ui[e];
  __syncthreads();
  // thread: row tid / 4, columns tid % 4 + 4 j (interleaved, so the four
  // threads of a row share the triangle's work evenly); U11^{-1} is upper
  // triangular, so column c sums k = 0 .. c
  const int r = tid >> 2;
  bool over = false;
  if (r0 + r < p.n) {
    for (int j = 0; j < 8; ++j) {
      const int c = (tid & 3) + 4 * j;
      if (c >= bw) break;
      double sr = 0.0, si = 0.0;
      for (int k = 0; k <= c; ++k) {
        const double2 x = a[r][k], y = u[k][c];
        sr += x.x * y.x - x.y * y.y;
        si += x.x * y.y + x.y * y.x;
      }
      if (!(fabs(sr) + fabs(si) <= lbound)) over = true;
    }
  }
  if (__any(over) && (tid & 63) == 0) atomicOr(flag, 1);
}

// One workgroup per (tournament leaf, problem): problems whose panel passed
// the threshold test (nearly all) exit at once; the leaves of a flagged one
// run side by side, and the arrival counters make the last leaf of each
// group the merger, as in the single-problem factorisation.
__global__ __launch_bounds__(256) void kb_panel_pivot(const mw_sweep_prob* P, int64_t k0,
                                                      const uint32_t* items) {
  const uint32_t it = items[blockIdx.x];
  const mw_sweep_prob& p = P[it >> 16];
  if (k0 >= p.n) return;
  const int32_t* flag = ws_flags(p.work, p.n) + k0 / NB;
  if (*flag == 0) return;
  const int bw = (int)((p.n - k0) < NB ? (p.n - k0) : NB);
  const int nleaf = (int)((p.n - k0 + TROWS - 1) / TROWS);
  const int leaf = (int)(it & 0xffff);
  if (leaf >= nleaf) return;
  __shared__ PivotLds<double2> S;
  panel_pivot_body<double>(S, leaf, (double2*)p.A, p.lda, p.n, k0, p.n + p.ncols, k0, bw, nleaf,
                           ws_tour(p.work, p.n), flag, (double2*)p.work, ws_inv(p.work, k0),
                           p.ipiv, p.info);
}

__global__ __launch_bounds__(256) void kb_panel_apply(const mw_sweep_prob* P, int64_t k0,
                                                      const uint32_t* items) {
  const uint32_t it = items[blockIdx.x];
  const mw_sweep_prob& p = P[it >> 16];
  if (k0 >= p.n) return;
  const int bw = (int)((p.n - k0) < NB ? (p.n - k0) : NB);
  const int64_t nc = p.n + p.ncols;
  const int nrb = (int)((p.n - k0 - bw + AW - 1) / AW);
  const int ncb = (int)((nc - k0 - bw + AW - 1) / AW);
  const int b = (int)(it & 0xffff);
  if (b > nrb + ncb) return;
  __shared__ BlkLds<double> L;
  panel_apply_body<double>(L, b, p.n, nc, (double2*)p.A, p.lda, k0, bw, ws_inv(p.work, k0),
                           (const double2*)p.work, nrb, ncb);
}

// mode 0: trailing update after panel k0; mode 1: back-substitution update above panel k0
template <bool BIG>
// items (mode 0): one entry per workgroup, problem << 16 | tile row << 8 |
// tile column; NULL (mode 1): grid (tile column, tile row, problem).
__global__ __launch_bounds__(256) void kb_gemm(const mw_sweep_prob* P, int64_t k0, int mode,
                                               const uint32_t* items) {
  uint32_t pi = blockIdx.z, bx = blockIdx.x, by = blockIdx.y;
  if (items) {
    const uint32_t it = items[blockIdx.x];
    pi = it >> 16;
    by = (it >> 8) & 0xff;
    bx = it & 0xff;
  }
  const mw_sweep_prob& p = P[pi];
  if (k0 >= p.n) return;
  const int bw = (int)((p.n - k0) < NB ? (p.n - k0) : NB);
  double2* A = reinterpret_cast<double2*>(p.A);
  const int64_t lda = p.lda;
  int64_t m, ncol;
  const double2 *a, *b;
  double2* c;
  if (mode == 0) {
    m = p.n - k0 - bw;
    ncol = p.n + p.ncols - k0 - bw;
    a = A + (k0 + bw) * lda + k0;
    b = A + k0 * lda + k0 + bw;
    c = A + (k0 + bw) * lda + k0 + bw;
  } else {
    m = k0;
    ncol = p.ncols;
    a = A + k0;
    b = A + k0 * lda + p.n;
    c = A + p.n;
  }
  const int tm = BIG ? GBM : SBM, tn = BIG ? GBN : SBN;
  if (m <= 0 || ncol <= 0 || (int64_t)by * tm >= m || (int64_t)bx * tn >= ncol) return;
  if (BIG)
    gemm_big_body<double, false>(bx, by, m, ncol, bw, a, lda, b, lda, c, lda);
  else
    gemm_small_body<double, false>(bx, by, m, ncol, bw, a, lda, b, lda, c, lda);
}

__global__ __launch_bounds__(256) void kb_bwd_apply(const mw_sweep_prob* P, int64_t k0) {
  const mw_sweep_prob& p = P[blockIdx.y];
  if (k0 >= p.n || (int64_t)blockIdx.x * AW >= p.ncols) return;
  const int bw = (int)((p.n - k0) < NB ? (p.n - k0) : NB);
  __shared__ BlkLds<double> L;
  solve_apply_body<double>(L, blockIdx.x, p.ncols, ws_inv(p.work, k0), k0, bw,
                           reinterpret_cast<double2*>(p.A) + p.n, p.lda, 0);
}

// KCL Schur products: S = A Xc (ncon x ncon), T = A W (ncon x n_ports).
__global__ void kb_schur_ax(const mw_sweep_prob* P, const mw_ports* ports) {
  const mw_sweep_prob& p = P[blockIdx.y];
  const mw_ports& q = ports[p.mesh];
  const int i = blockIdx.x;
  const int j = threadIdx.x;
  if (i >= q.n_con || j >= q.n_con + q.n_ports) return;
  const double2* X = reinterpret_cast<const double2*>(p.A) + p.n;
  const int col = j < q.n_con ? q.n_ports + j : j - q.n_con;
  double sr = 0.0, si = 0.0;
  for (int t = 0; t < q.con_max; ++t) {
    const int32_t r = q.con_idx[i * q.con_max + t];
    if (r < 0) continue;
    const double v = q.con_val[i * q.con_max + t];
    const double2 x = X[(int64_t)r * p.lda + col];
    sr += v * x.x;
    si += v * x.y;
  }
  double2* out = j < q.n_con ? reinterpret_cast<double2*>(p.S) + i * q.n_con + j
                             : reinterpret_cast<double2*>(p.T) + i * q.n_ports + (j - q.n_con);
  *out = make_double2(sr, si);
}

__global__ __launch_bounds__(256) void kb_small_solve(const mw_sweep_prob* P, const mw_ports* ports) {
  const mw_sweep_prob& p = P[blockIdx.x];
  const mw_ports& q = ports[p.mesh];
  if (q.n_con == 0) return;
  __shared__ SmallLds SL;
  small_solve_body(SL, q.n_con, q.n_ports, (double2*)p.S, q.n_con, (double2*)p.T, q.n_ports,
                   p.info + 1);
}

// W -= Xc Lambda, one row per thread.
__global__ void kb_schur_update(const mw_sweep_prob* P, const mw_ports* ports) {
  const mw_sweep_prob& p = P[blockIdx.y];
  const mw_ports& q = ports[p.mesh];
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q.n_con == 0 || r >= p.n) return;
  double2* X = reinterpret_cast<double2*>(p.A) + p.n + r * p.lda;
  const double2* Lm = reinterpret_cast<const double2*>(p.T);
  for (int j = 0; j < q.n_ports; ++j) {
    double2 acc = X[j];
    for (int c = 0; c < q.n_con; ++c) acc = cfms(acc, X[q.n_ports + c], Lm[c * q.n_ports + j]);
    X[j] = acc;
  }
}

__global__ void kb_readout(const mw_sweep_prob* P, const mw_ports* ports) {
  const mw_sweep_prob& p = P[blockIdx.y];
  const mw_ports& q = ports[p.mesh];
  const int i = blockIdx.x, j = threadIdx.x;
  if (i >= q.n_ports || j >= q.n_ports) return;
  sparse_rows_body(i, j, q.n_ports, q.port_max, q.port_idx, q.port_val,
                   reinterpret_cast<const double2*>(p.A) + p.n, p.lda, q.n_ports,
                   reinterpret_cast<double2*>(p.Y), q.n_ports);
}

// ---------------------------------------------------------------------------
// Persistent batched solve: one workgroup takes one problem at a time and
// runs its whole port solve -- right-hand sides, the threshold-pivoted
// factorisation with its tournament fallback, the back substitution, the KCL
// Schur step and the Y readout.  Problems are sorted largest first and
// claimed through one counter, so the largest start first and the small ones
// fill in behind them.  No workgroup ever waits for another: the launch needs
// no co-residency and no cross-workgroup synchronisation.  Within a
// workgroup the phases are ordered by barriers (one phase's global writes are
// read by other threads of the same workgroup in the next).
//
// Each phase is a separate (not inlined) function, so each gets the whole
// register budget: inlined into one kernel body, the phases' uniform values
// overflowed the scalar registers into vector lanes and spilled.  They share
// one LDS buffer declared at namespace scope, which every phase addresses
// directly (LDS instructions, not generic ones).
// ---------------------------------------------------------------------------
constexpr size_t kWgLds =
    sizeof(BlkLds<double>) > sizeof(PivotLds<double2>) ? sizeof(BlkLds<double>)
                                                       : sizeof(PivotLds<double2>);
static_assert(kWgLds >= sizeof(PanelLds<double2>) && kWgLds >= sizeof(SmallLds) &&
                  kWgLds >= sizeof(K32Lds<double>) && kWgLds >= sizeof(K64Lds),
              "the shared buffer holds every phase's LDS");
__shared__ __align__(16) unsigned char g_wg_lds[kWgLds];
__shared__ int g_wg_flag;

#define MW_PHASE __device__ __attribute__((noinline))

// Workspace flags, tournament counters and info; the right-hand sides.
MW_PHASE void wg_setup(const mw_sweep_prob* pp, const mw_ports* qq) {
  const mw_sweep_prob& p = *pp;
  const mw_ports& q = *qq;
  const int tid = threadIdx.x;
  const int64_t n = p.n, ncols = p.ncols, lda = p.lda;
  double2* X = reinterpret_cast<double2*>(p.A) + n;
  int32_t* flags = ws_flags(p.work, n);
  const int64_t cnt = (n + NB - 1) / NB + d_tour_nodes(n);
  for (int64_t e = tid; e < cnt; e += blockDim.x) flags[e] = 0;
  if (tid < 2) p.info[tid] = 0;
  for (int64_t e = tid; e < n * ncols; e += blockDim.x)
    X[(e / ncols) * lda + e % ncols] = make_double2(0.0, 0.0);
  __syncthreads();
  for (int64_t e = tid; e < ncols * q.col_max; e += blockDim.x) {
    const int32_t r = q.col_idx[e];
    if (r >= 0) X[(int64_t)r * lda + e / q.col_max] = make_double2(q.col_val[e], 0.0);
  }
  __syncthreads();
}

// A11 of the panel at k0 without row exchanges, both inverses, A11's own
// multipliers checked, then L21 = A21 U11^{-1} checked against the bound
// (nothing written into A).  Returns whether the panel needs the fallback.
MW_PHASE int wg_factor_check(const mw_sweep_prob* pp, int64_t k0, double lbound) {
  const mw_sweep_prob& p = *pp;
  const int64_t n = p.n, lda = p.lda;
  const int bw = (int)((n - k0) < NB ? (n - k0) : NB);
  double2* A = reinterpret_cast<double2*>(p.A);
  double2* inv = ws_inv(p.work, k0);
  int32_t* flag = ws_flags(p.work, n) + k0 / NB;
  PanelLds<double2>& P = *reinterpret_cast<PanelLds<double2>*>(g_wg_lds);
  panel_nopiv_body<double>(P, 0, n, A, lda, k0, bw, reinterpret_cast<double2*>(p.work), inv,
                           p.ipiv, flag, lbound);
  __syncthreads();
  const int bad = P.bad;
  __syncthreads();
  if (bad) return 1;
  BlkLds<double>& B = *reinterpret_cast<BlkLds<double>*>(g_wg_lds);
  const int nrb = (int)((n - k0 - bw + AW - 1) / AW);
  bool over = false;
  for (int b = 0; b < nrb; ++b) {
    const int64_t r0 = k0 + bw + (int64_t)b * AW;
    const int rows = (int)((n - r0) < AW ? (n - r0) : AW);
    over |= block_product<double, 64, 32>(B.As, &B.Bs[0][0], 65, inv + NB * NB, NB,
                                          A + r0 * lda + k0, lda, rows, bw, true, lbound);
    __syncthreads();
  }
  if (threadIdx.x == 0) g_wg_flag = 0;
  __syncthreads();
  if (over) g_wg_flag = 1;
  __syncthreads();
  return g_wg_flag;
}

// Tournament fallback of the panel at k0: the leaves one after the other
// (the last child of each merge group, as the arrival counters count, merges);
// its interchanges apply to columns [c_lo, n + ncols).
MW_PHASE void wg_pivot(const mw_sweep_prob* pp, int64_t k0, int64_t c_lo) {
  const mw_sweep_prob& p = *pp;
  const int64_t n = p.n, nc = n + p.ncols;
  const int bw = (int)((n - k0) < NB ? (n - k0) : NB);
  int32_t* flag = ws_flags(p.work, n) + k0 / NB;
  if (threadIdx.x == 0) *flag = 1;
  __syncthreads();
  const int nleaf = (int)((n - k0 + TROWS - 1) / TROWS);
  PivotLds<double2>& S = *reinterpret_cast<PivotLds<double2>*>(g_wg_lds);
  for (int leaf = 0; leaf < nleaf; ++leaf) {
    panel_pivot_body<double>(S, leaf, reinterpret_cast<double2*>(p.A), p.lda, n, c_lo, nc, k0, bw,
                             nleaf, ws_tour(p.work, n), flag, reinterpret_cast<double2*>(p.work),
                             ws_inv(p.work, k0), p.ipiv, p.info);
    __syncthreads();
  }
}

// L21, U12 for columns [k0 + bw, nc) and A11 of the panel at k0 into A.
MW_PHASE void wg_apply(const mw_sweep_prob* pp, int64_t k0, int64_t nc) {
  const mw_sweep_prob& p = *pp;
  const int64_t n = p.n;
  const int bw = (int)((n - k0) < NB ? (n - k0) : NB);
  const int nrb = (int)((n - k0 - bw + AW - 1) / AW);
  const int ncb = (int)((nc - k0 - bw + AW - 1) / AW);
  BlkLds<double>& B = *reinterpret_cast<BlkLds<double>*>(g_wg_lds);
  for (int b = 0; b <= nrb + ncb; ++b) {
    panel_apply_body<double>(B, b, n, nc, reinterpret_cast<double2*>(p.A), p.lda, k0, bw,
                             ws_inv(p.work, k0), reinterpret_cast<const double2*>(p.work), nrb,
                             ncb);
    __syncthreads();
  }
}

// C -= A B over an m x ncol block (any K), 64 x 32 tiles one after the other.
MW_PHASE void wg_gemm(int64_t m, int64_t ncol, int kdim, const double2* a, const double2* b,
                      double2* c, int64_t ld) {
  const int gxn = (int)((ncol + 63) / 64);
  const int ntiles = gxn * (int)((m + 63) / 64);
  gemm_kc64_tiles(*reinterpret_cast<K64Lds*>(g_wg_lds), ntiles, gxn, m, ncol, kdim, a, ld, b, ld,
                  c, ld);
  __syncthreads();
}

// X[k0 .. k0 + bw) <- L11^{-1} X[k0 .. k0 + bw) over ncols columns (the
// panel at k0; X rows of ldx).
MW_PHASE void wg_lower_rows(const mw_sweep_prob* pp, int64_t k0, double2* X, int64_t ldx,
                            int64_t ncols) {
  const mw_sweep_prob& p = *pp;
  const int bw = (int)((p.n - k0) < NB ? (p.n - k0) : NB);
  BlkLds<double>& B = *reinterpret_cast<BlkLds<double>*>(g_wg_lds);
  for (int bx = 0; bx < (int)((ncols + AW - 1) / AW); ++bx) {
    solve_apply_body<double>(B, bx, ncols, ws_inv(p.work, k0), k0, bw, X, ldx, 1);
    __syncthreads();
  }
}

// X_k <- U_kk^{-1} X_k for the panel at k0 (all right-hand-side columns).
MW_PHASE void wg_back_step(const mw_sweep_prob* pp, int64_t k0) {
  const mw_sweep_prob& p = *pp;
  const int64_t n = p.n, ncols = p.ncols;
  const int bw = (int)((n - k0) < NB ? (n - k0) : NB);
  BlkLds<double>& B = *reinterpret_cast<BlkLds<double>*>(g_wg_lds);
  for (int bx = 0; bx < (int)((ncols + AW - 1) / AW); ++bx) {
    solve_apply_body<double>(B, bx, ncols, ws_inv(p.work, k0), k0, bw,
                             reinterpret_cast<double2*>(p.A) + n, p.lda, 0);
    __syncthreads();
  }
}

// KCL Schur step (S = A Xc, T = A W; Lambda = S^{-1} T; W -= Xc Lambda) and
// the Y readout.
MW_PHASE void wg_finish(const mw_sweep_prob* pp, const mw_ports* qq) {
  const mw_sweep_prob& p = *pp;
  const mw_ports& q = *qq;
  const int tid = threadIdx.x;
  const int64_t n = p.n, lda = p.lda;
  double2* X = reinterpret_cast<double2*>(p.A) + n;
  if (q.n_con > 0) {
    const int nsc = q.n_con + q.n_ports;
    for (int e = tid; e < q.n_con * nsc; e += blockDim.x) {
      const int i = e / nsc, j = e % nsc;
      const int col = j < q.n_con ? q.n_ports + j : j - q.n_con;
      double sr = 0.0, si = 0.0;
      for (int t = 0; t < q.con_max; ++t) {
        const int32_t r = q.con_idx[i * q.con_max + t];
        if (r < 0) continue;
        const double v = q.con_val[i * q.con_max + t];
        const double2 x = X[(int64_t)r * lda + col];
        sr += v * x.x;
        si += v * x.y;
      }
      double2* out = j < q.n_con ? reinterpret_cast<double2*>(p.S) + i * q.n_con + j
                                 : reinterpret_cast<double2*>(p.T) + i * q.n_ports + (j - q.n_con);
      *out = make_double2(sr, si);
    }
    __syncthreads();
    small_solve_body(*reinterpret_cast<SmallLds*>(g_wg_lds), q.n_con, q.n_ports,
                     reinterpret_cast<double2*>(p.S), q.n_con, reinterpret_cast<double2*>(p.T),
                     q.n_ports, p.info + 1);
    __syncthreads();
    const double2* Lm = reinterpret_cast<const double2*>(p.T);
    for (int64_t r = tid; r < n; r += blockDim.x) {
      double2* Xr = X + r * lda;
      for (int j = 0; j < q.n_ports; ++j) {
        double2 acc = Xr[j];
        for (int c = 0; c < q.n_con; ++c) acc = cfms(acc, Xr[q.n_ports + c], Lm[c * q.n_ports + j]);
        Xr[j] = acc;
      }
    }
    __syncthreads();
  }
  for (int e = tid; e < q.n_ports * q.n_ports; e += blockDim.x)
    sparse_rows_body(e / q.n_ports, e % q.n_ports, q.n_ports, q.port_max, q.port_idx, q.port_val,
                     X, lda, q.n_ports, reinterpret_cast<double2*>(p.Y), q.n_ports);
  __syncthreads();
}

// One level (TWO = false): per 32-column panel the check (tournament if it
// fails), the apply and the K = 32 update of everything right of it.  Two
// levels: a 128-column outer block is factored panel by panel with the
// updates kept inside the block, then the block's U rows for the columns
// right of it, then one K = 128 update of the rest -- the trailing matrix is
// read and written once per outer block instead of once per panel.
constexpr int kWgOuter = 128;

template <bool TWO>
MW_DEV void wg_port_solve(const mw_sweep_prob* pp, const mw_ports* qq, double lbound) {
  wg_setup(pp, qq);
  const int64_t n = pp->n, nc = n + pp->ncols, lda = pp->lda;
  double2* A = reinterpret_cast<double2*>(pp->A);
  const int64_t nbo = TWO ? kWgOuter : NB;
  for (int64_t K0 = 0; K0 < n; K0 += nbo) {
    const int64_t kend = (n - K0) < nbo ? n : K0 + nbo;
    // one level: the apply and the update reach every column at once
    const int64_t cblk = TWO ? kend : nc;
    for (int64_t k0 = K0; k0 < kend; k0 += NB) {
      const int bw = (int)((n - k0) < NB ? (n - k0) : NB);
      if (wg_factor_check(pp, k0, lbound)) wg_pivot(pp, k0, K0);
      wg_apply(pp, k0, cblk);
      const int64_t rest = n - k0 - bw, restc = cblk - k0 - bw;
      if (rest > 0 && restc > 0)
        wg_gemm(rest, restc, bw, A + (k0 + bw) * lda + k0, A + k0 * lda + k0 + bw,
                A + (k0 + bw) * lda + k0 + bw, lda);
    }
    const int64_t ncr = nc - kend;
    if (TWO && ncr > 0) {
      double2* X = A + kend;  // the block's U rows, columns [kend, nc)
      for (int64_t k0 = K0; k0 < kend; k0 += NB) {
        const int bw = (int)((n - k0) < NB ? (n - k0) : NB);
        wg_lower_rows(pp, k0, X, lda, ncr);
        const int64_t below = kend - k0 - bw;
        if (below > 0)
          wg_gemm(below, ncr, bw, A + (k0 + bw) * lda + k0, X + k0 * lda, X + (k0 + bw) * lda,
                  lda);
      }
      if (n > kend)
        wg_gemm(n - kend, ncr, (int)(kend - K0), A + kend * lda + K0, A + K0 * lda + kend,
                A + kend * lda + kend, lda);
    }
  }
  double2* X = A + n;
  for (int64_t k0 = ((n + NB - 1) / NB - 1) * NB; k0 >= 0; k0 -= NB) {
    wg_back_step(pp, k0);
    const int bw = (int)((n - k0) < NB ? (n - k0) : NB);
    if (k0 > 0) wg_gemm(k0, pp->ncols, bw, A + k0, X + k0 * lda, X, lda);
  }
  wg_finish(pp, qq);
}

// next[0]: the claim counter (zero on entry).  A set cancel flag (device-
// readable) stops the solves: every problem claimed after it is seen gets
// info[0] = -1 instead.
template <bool TWO>
__global__ __launch_bounds__(256, 2) void kb_lu_wg(const mw_sweep_prob* P, const mw_ports* ports,
                                                   int nprob, int32_t* next, double lbound,
                                                   const volatile int32_t* cancel) {
  __shared__ int claim_s, stop_s;
  // Every branch after the claim is on wave-uniform (scalar) values and no
  // path back to the barrier depends on the thread index: with a divergent
  // "stop" branch and a thread-0-only write before the back edge, the
  // compiler once sent lanes 1-63 of wave 0 straight back to the barrier
  // while lane 0, the only one that claims, waited behind them.
  while (true) {
    if (threadIdx.x == 0) {
      claim_s = atomicAdd(next, 1);
      stop_s = (cancel != nullptr && *cancel != 0) ? 1 : 0;
    }
    __syncthreads();
    const int pi = __builtin_amdgcn_readfirstlane(claim_s);
    const int stop = __builtin_amdgcn_readfirstlane(stop_s);
    __syncthreads();
    if (pi >= nprob) return;
    if (stop)
      P[pi].info[0] = -1;  // every thread stores the same value
    else
      wg_port_solve<TWO>(P + pi, ports + P[pi].mesh, lbound);
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Symmetric factorisation helpers.  For Z = Z^T factored without row
// exchanges, U = D L^T (D: U's diagonal), so an outer block's U rows right of
// it follow from its L columns below it:
//   A[K0 + i][c] = A[K0 + i][K0 + i] A[c][K0 + i],  c in [kend, n).
// 64 x 64 tiles transposed through LDS (coalesced reads along i, writes
// along c).  blockIdx.x: 64 columns i of the block, blockIdx.y: 64 rows c.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_sym_u_rows(int64_t K0, int bs, int64_t kend, int64_t n,
                                                   typename cplx_t<T>::type* A, int64_t lda) {
  typedef typename cplx_t<T>::type CT;
  __shared__ CT tile[64][65];
  __shared__ CT dg[64];
  const int tid = threadIdx.x;
  const int i0 = blockIdx.x * 64;
  const int64_t c0 = kend + (int64_t)blockIdx.y * 64;
  const int ni = (bs - i0) < 64 ? (bs - i0) : 64;
  const int nc = (int)((n - c0) < 64 ? (n - c0) : 64);
  if (tid < ni) dg[tid] = A[(K0 + i0 + tid) * lda + K0 + i0 + tid];
  for (int e = tid; e < 64 * 64; e += 256) {
    const int r = e >> 6, j = e & 63;  // row c0 + r, column K0 + i0 + j
    if (r < nc && j < ni) tile[r][j] = A[(c0 + r) * lda + K0 + i0 + j];
  }
  __syncthreads();
  for (int e = tid; e < 64 * 64; e += 256) {
    const int j = e >> 6, r = e & 63;  // row K0 + i0 + j, column c0 + r
    if (r < nc && j < ni) A[(K0 + i0 + j) * lda + c0 + r] = cmul(dg[j], tile[r][j]);
  }
}

// ---------------------------------------------------------------------------
// Inverses of the symmetric factorisation's 512 x 512 diagonal blocks, for
// triangular solves by block products.  Per outer block b (rows K0 = 512 b,
// bs = min(512, n - K0)): N_b = -L_bb^{-1} (L_bb the unit lower diagonal
// block of the factors) and V_b = -U_bb^{-1} = N_b^T D_b^{-1} (U = D L^T),
// both stored 512 x 512 (zero beyond bs).  N_b is built by recursive
// doubling from the panels' 32 x 32 inverses: for a 2s x 2s diagonal block
// [[P, 0], [C, Q]], the lower-left block of its inverse is -Q^{-1} C P^{-1};
// with the negated blocks N_P = -P^{-1} and N_Q = -Q^{-1} it is
// N_Q C N_P, formed as T = 0 - C N_P, then 0 - N_Q T.  T lives in V_b (the
// lower-left quadrants of different levels do not overlap) until V_b is
// written at the end.
// ---------------------------------------------------------------------------
constexpr int SYB = 512;  // largest outer block of the symmetric factorisation (storage slot)

// Outer block of the symmetric factorisation: 512 past 4096 unknowns (the
// LU's two-level block); none below, where the one-level look-ahead LU is
// faster (2,624 unknowns: 5.4 ms per port solve against 6.7 ms for the
// symmetric path with 256-column blocks).
static int64_t sym_block(int64_t n) { return n > 4096 ? 512 : 0; }

// N_b <- the negated panel inverses on its 32 x 32 diagonal blocks, zero
// elsewhere; V_b <- 0.  Grid: (512 rows / 1, nblk); 256 threads, 2 columns each.
template <typename T>
__global__ __launch_bounds__(256) void k_blkinv_init(int64_t n, int bz, const void* work,
                                                    typename cplx_t<T>::type* N,
                                                    typename cplx_t<T>::type* V) {
  typedef typename cplx_t<T>::type CT;
  const int r = blockIdx.x, b = blockIdx.y;
  const int64_t K0 = (int64_t)b * bz;
  const int bs = (int)((n - K0) < bz ? (n - K0) : bz);
  CT* Nr = N + ((int64_t)b * SYB + r) * SYB;
  CT* Vr = V + ((int64_t)b * SYB + r) * SYB;
  const CT* li = reinterpret_cast<const CT*>(reinterpret_cast<const char*>(work) +
                                             (int64_t)NB * NB * 16 * (1 + 2 * ((K0 + r) / NB)));
  for (int c = threadIdx.x; c < SYB; c += 256) {
    CT v;
    v.x = 0;
    v.y = 0;
    if (r < bs && c < bs && (c / NB) == (r / NB) && c <= r) {
      const CT l = li[(r % NB) * NB + (c % NB)];
      v.x = -l.x;
      v.y = -l.y;
    }
    Nr[c] = v;
    CT z;
    z.x = 0;
    z.y = 0;
    Vr[c] = z;
  }
}

// One doubling level (s = 32, 64, 128, 256) over every block and pair:
// PHASE 1: T = 0 - C N_P into V_b; PHASE 2: N_b's lower-left block = 0 - N_Q T.
// Grid: nblk * (256 / s) pairs * tiles of the s x s output (64 x 32 tiles).
template <typename T, int PHASE>
__global__ __launch_bounds__(256) void k_blkinv_level(int s, int64_t n, int bz,
                                                     const typename cplx_t<T>::type* A,
                                                     int64_t lda, typename cplx_t<T>::type* N,
                                                     typename cplx_t<T>::type* V) {
  __shared__ GemmSmallLds<T> L;
  const int tx = (s + SBN - 1) / SBN, ty = (s + SBM - 1) / SBM;
  const int pairs = bz / (2 * s);
  int id = blockIdx.x;
  const int tile = id % (tx * ty);
  id /= tx * ty;
  const int p = id % pairs, b = id / pairs;
  const int64_t K0 = (int64_t)b * bz;
  const int bs = (int)((n - K0) < bz ? (n - K0) : bz);
  const int a0 = 2 * p * s, d0 = a0 + s;
  if (d0 >= bs) return;
  const int mD = (bs - d0) < s ? (bs - d0) : s;
  const int bx = tile % tx, by = tile / tx;
  auto* Nb = N + (int64_t)b * SYB * SYB;
  auto* Vb = V + (int64_t)b * SYB * SYB;
  if (PHASE == 1)
    gemm_small_tile<T, false>(L, bx, by, mD, s, s, A + (K0 + d0) * lda + K0 + a0, lda,
                              Nb + (int64_t)a0 * SYB + a0, SYB, Vb + (int64_t)d0 * SYB + a0, SYB);
  else
    gemm_small_tile<T, false>(L, bx, by, mD, s, mD, Nb + (int64_t)d0 * SYB + d0, SYB,
                              Vb + (int64_t)d0 * SYB + a0, SYB, Nb + (int64_t)d0 * SYB + a0, SYB);
}

// V_b[i][j] = N_b[j][i] / D_j (D_j = U's diagonal, A[K0 + j][K0 + j]): 64 x 64
// tiles through LDS.  Grid: (8, 8, nblk).
template <typename T>
__global__ __launch_bounds__(256) void k_blkinv_vt(int64_t n, int bz,
                                                  const typename cplx_t<T>::type* A, int64_t lda,
                                                  const typename cplx_t<T>::type* N,
                                                  typename cplx_t<T>::type* V) {
  typedef typename cplx_t<T>::type CT;
  __shared__ CT tile[64][65];
  __shared__ CT dr[64];
  const int tid = threadIdx.x;
  const int j0 = blockIdx.x * 64, i0 = blockIdx.y * 64, b = blockIdx.z;
  const int64_t K0 = (int64_t)b * bz;
  const int bs = (int)((n - K0) < bz ? (n - K0) : bz);
  const CT* Nb = N + (int64_t)b * SYB * SYB;
  CT* Vb = V + (int64_t)b * SYB * SYB;
  if (tid < 64) {
    CT d;
    d.x = 0;
    d.y = 0;
    if (j0 + tid < bs) d = crecip(A[(K0 + j0 + tid) * lda + K0 + j0 + tid]);
    dr[tid] = d;
  }
  for (int e = tid; e < 64 * 64; e += 256) {
    const int jj = e >> 6, ii = e & 63;  // N_b[j0 + jj][i0 + ii]
    tile[jj][ii] = Nb[(int64_t)(j0 + jj) * SYB + i0 + ii];
  }
  __syncthreads();
  for (int e = tid; e < 64 * 64; e += 256) {
    const int ii = e >> 6, jj = e & 63;  // V_b[i0 + ii][j0 + jj]
    Vb[(int64_t)(i0 + ii) * SYB + j0 + jj] = cmul(tile[jj][ii], dr[jj]);
  }
}

// tmp[r][c] <- X[r][c], X[r][c] <- 0 over rows x cols (the right-hand sides
// of one block, moved aside for an out-of-place block product).
template <typename T>
__global__ void k_move_zero(int64_t rows, int64_t cols, typename cplx_t<T>::type* X, int64_t ldx,
                            typename cplx_t<T>::type* tmp, int64_t ldt) {
  typedef typename cplx_t<T>::type CT;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= rows * cols) return;
  const int64_t r = e / cols, c = e % cols;
  CT* x = X + r * ldx + c;
  tmp[r * ldt + c] = *x;
  CT z;
  z.x = 0;
  z.y = 0;
  *x = z;
}

// X_b <- X_b - M_b B for one block's rows (bs <= 512) and cw <= SYW columns,
// M_b a 512 x 512 block inverse: lower triangular (N_b, LOWER) or upper
// (V_b), so a 64-row tile's K range stops at its last row or starts at its
// first.  One 64 x 32 tile per workgroup, K in 32-wide chunks with the next
// chunk's operands prefetched into registers (gemm_kc_tiles).
template <typename T, bool LOWER>
__global__ __launch_bounds__(256) void k_blk_trmm(int bs, int cw,
                                                 const typename cplx_t<T>::type* __restrict__ M,
                                                 const typename cplx_t<T>::type* __restrict__ B,
                                                 typename cplx_t<T>::type* X, int64_t ldx) {
  __shared__ K32Lds<T> L;
  const int gxn = (cw + NB - 1) / NB;
  const int r0 = (int)(blockIdx.x / gxn) * 64, c0 = (int)(blockIdx.x % gxn) * NB;
  if (r0 >= bs) return;
  const int kb = LOWER ? 0 : r0;
  const int ke = LOWER ? ((r0 + 64) < bs ? r0 + 64 : bs) : bs;
  const int m = (bs - r0) < 64 ? bs - r0 : 64;
  const int n = (cw - c0) < NB ? cw - c0 : NB;
  gemm_kc_tiles<T>(L, 1, 1, m, n, ke - kb, M + (int64_t)r0 * SYB + kb, SYB,
                   B + (int64_t)kb * cw + c0, cw, X + (int64_t)r0 * ldx + c0, ldx);
}

// info[0] <- -1 (unless already set) when any panel's threshold check failed:
// the symmetric factorisation does not pivot, so its result is not valid.
__global__ void k_sym_info(const int32_t* flags, int64_t npan, int32_t* info) {
  __shared__ int any;
  if (threadIdx.x == 0) any = 0;
  __syncthreads();
  int mine = 0;
  for (int64_t e = threadIdx.x; e < npan; e += blockDim.x) mine |= flags[e];
  if (mine) any = 1;
  __syncthreads();
  if (threadIdx.x == 0 && any && info[0] == 0) info[0] = -1;
}

// ---------------------------------------------------------------------------
// Host drivers (called from mw_capi.hip)
// ---------------------------------------------------------------------------
// workspace: lu11 [NB*NB complex double] | diagonal-block inverses
// [n_panels][2][NB*NB complex double] | flags [n_panels] | counters [nodes] |
// ncand [nodes] | cand [nodes][NB]
static int64_t tour_nodes(int64_t n) {
  int64_t nodes = 0;
  for (int64_t nd = (n + TROWS - 1) / TROWS; nd > 1; nd = (nd + FANIN - 1) / FANIN) nodes += nd;
  return nodes + 1;
}

int64_t mw_getrf_ws_bytes(int64_t n) {
  const int64_t npan = (n + NB - 1) / NB;
  const int64_t nodes = tour_nodes(n);
  const int64_t bytes = (int64_t)NB * NB * 16 * (1 + 2 * npan) + 4 * npan + 4 * nodes * (2 + NB);
  return (bytes + 255) / 256 * 256;
}

static void* panel_inv(void* work, int64_t k0) {
  return reinterpret_cast<char*>(work) + (int64_t)NB * NB * 16 * (1 + 2 * (k0 / NB));
}

// Threshold-pivoting multiplier bound (1/tau; <= 0: always pivot), one per
// calling thread: solvers driven from different threads never see each
// other's setting, and each factorisation reads it once when it starts.
static thread_local double g_lbound = 8.0;

// Outer block width: trailing updates use K = nbo.  Small systems are bound
// by the panel chain rather than by the updates, so they keep one level.
// 512 measured best at 15,360 unknowns (FP64 solve 389 ms at 128, 353 at 256,
// 338 at 512, 340 at 768, 344 at 1024).
static int64_t outer_block(int64_t n) { return n > 4096 ? 512 : NB; }

// Workgroups of k_lookahead<T> resident on the whole device at once (per
// device, queried on first use).
template <typename T>
static hipError_t lookahead_resident(int* out) {
  constexpr int kMaxDev = 64;
  static int cached[kMaxDev] = {0};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDev) return hipErrorInvalidDevice;
  if (cached[dev] == 0) {
    int cus = 0, per_cu = 0;
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess)
      return e;
    if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_lookahead<T>, 256, 0)) !=
        hipSuccess)
      return e;
    cached[dev] = std::max(1, cus * per_cu);
  }
  *out = cached[dev];
  return hipSuccess;
}

// One-level factorisation with look-ahead (see k_lookahead): per panel the
// tournament fallback, the apply with the next panel's update, then the next
// panel's factor together with the rest of the trailing update, in one
// launch.  (The same overlap from two streams joined by events ran slower:
// the update's workgroups, dispatched first, held the CUs and the factor's
// blocks waited behind them.)
// SYM: Z = Z^T without row exchanges (no tournament fallback: a failed
// threshold check leaves its flag for k_sym_info), the update on the lower
// part and the panel-diagonal blocks only (k_lookahead<T, true>).
template <typename T, bool SYM = false>
static int getrf_lookahead(int64_t n, int64_t nc, typename cplx_t<T>::type* A, int64_t lda,
                           int32_t* ipiv, void* work, int32_t* flags, const TourWs& tw,
                           int32_t* info, const volatile int32_t* cancel, hipStream_t st,
                           hipError_t* herr) {
  typedef typename cplx_t<T>::type CT;
  const double lbound = g_lbound;
  CT* lu11 = reinterpret_cast<CT*>(work);
  int resident = 0;
  if ((*herr = lookahead_resident<T>(&resident)) != hipSuccess) return MW_EHIP;
  {
    const int bw = (int)(n < NB ? n : NB);
    const int nrb = (int)((n - bw + AW - 1) / AW);
    hipLaunchKernelGGL(k_panel_nopiv<T>, dim3(nrb + 1), dim3(256), 0, st, n, (const CT*)A, lda,
                       (int64_t)0, bw, lu11, reinterpret_cast<CT*>(panel_inv(work, 0)), ipiv,
                       flags, lbound);
  }
  for (int64_t k0 = 0; k0 < n; k0 += NB) {
    if (cancel && *cancel) return MW_EABORT;
    const int bw = (int)((n - k0) < NB ? (n - k0) : NB);
    const int64_t kn = k0 + bw;
    const int64_t rest = n - kn;
    const int bwn = (int)(rest < NB ? rest : NB);
    const int nrb = (int)((rest + LAW - 1) / LAW);
    // U12 blocks: every column right of the next panel (SYM: the right-hand
    // sides only; the L21 blocks store the rest)
    const int64_t cu = SYM ? std::max(kn + bwn, n) : kn + bwn;
    const int ncb = (int)((nc - cu + LAW - 1) / LAW);
    CT* inv = reinterpret_cast<CT*>(panel_inv(work, k0));
    int32_t* flag = flags + k0 / NB;
    if (!SYM) {
      const int nleaf = (int)((n - k0 + TROWS - 1) / TROWS);
      // interchanges from column k0 on: earlier multiplier columns keep their
      // order, as the per-panel forward solve expects
      hipLaunchKernelGGL(k_panel_pivot<T>, dim3(nleaf), dim3(256), 0, st, A, lda, n, k0, nc, k0,
                         bw, nleaf, tw, (const int32_t*)flag, lu11, inv, ipiv, info);
    }
    hipLaunchKernelGGL((k_panel_apply_la<T, SYM>), dim3(nrb + (ncb > 0 ? ncb : 0) + 1), dim3(256),
                       0, st, n, nc, A, lda, k0, bw, bwn, (const CT*)inv, (const CT*)lu11, nrb,
                       ncb > 0 ? ncb : 0);
    if (bwn > 0) {
      const int nnp = (int)((rest - bwn + AW - 1) / AW) + 1;
      const int64_t ncols = nc - kn - bwn;
      const int gxn = (int)((ncols + NB - 1) / NB);  // 0: no update left (no RHS)
      const int ntiles = SYM ? (int)SymTiles(n, nc, kn, bwn).total()
                             : gxn * (int)((rest + 63) / 64);
      // update blocks: what stays resident beside the factor's blocks
      const int G = std::max(1, std::min(ntiles, resident - nnp - 1));
      hipLaunchKernelGGL((k_lookahead<T, SYM>), dim3((unsigned)(nnp + 1 + (ntiles > 0 ? G : 0))),
                         dim3(256), 0, st, n, nc, A, lda, k0, bw, bwn, (const CT*)inv, lu11,
                         reinterpret_cast<CT*>(panel_inv(work, kn)), ipiv, flag + 1, lbound, nnp,
                         gxn > 0 ? gxn : 1, ntiles, G);
    }
    *herr = hipGetLastError();
    if (*herr != hipSuccess) return MW_EHIP;
  }
  return MW_OK;
}

// Factor the n x n block of A; columns [n, nc) are carried along (row
// interchanges, L^{-1} and trailing updates), so on return they hold
// L^{-1} P B for right-hand sides B stored there: the forward half of a solve.
//
// Two-level blocking: 32-column panels (threshold-pivot check, tournament
// fallback, apply) factor an nbo-column block, updating only that block; the
// block's U rows for the remaining columns follow by blocked substitution and
// the rest of the matrix takes one K = nbo update per block.
template <typename T>
int mw_getrf_t(int64_t n, int64_t nc, void* Av, int64_t lda, int32_t* ipiv, void* work,
               int32_t* info, const volatile int32_t* cancel, hipStream_t st, hipError_t* herr) {
  typedef typename cplx_t<T>::type CT;
  CT* A = reinterpret_cast<CT*>(Av);
  const int64_t npan = (n + NB - 1) / NB;
  const int64_t nodes = tour_nodes(n);
  char* wp = reinterpret_cast<char*>(work);
  CT* lu11 = reinterpret_cast<CT*>(wp);
  int32_t* flags = reinterpret_cast<int32_t*>(wp + (int64_t)NB * NB * 16 * (1 + 2 * npan));
  TourWs tw;
  tw.counter = flags + npan;
  tw.ncand = tw.counter + nodes;
  tw.cand = tw.ncand + nodes;
  const int64_t nzero = npan + nodes;
  hipLaunchKernelGGL(k_zero_i32, dim3((unsigned)((nzero + 255) / 256)), dim3(256), 0, st, flags,
                     nzero);
  hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, st, info, (int64_t)1);
  const double lbound = g_lbound;
  const int64_t nbo = outer_block(n);
  const bool one_level = nbo == NB;
  if (one_level) return getrf_lookahead<T>(n, nc, A, lda, ipiv, work, flags, tw, info, cancel, st,
                                           herr);
  for (int64_t K0 = 0; K0 < n; K0 += nbo) {
    const int64_t kend = (n - K0) < nbo ? n : K0 + nbo;
    for (int64_t k0 = K0; k0 < kend; k0 += NB) {
      if (cancel && *cancel) return MW_EABORT;
      const int bw = (int)((kend - k0) < NB ? (kend - k0) : NB);
      const int64_t rest = n - k0 - bw;         // rows below the panel
      const int64_t restc = kend - k0 - bw;     // block columns right of the panel
      // one level: the apply also forms U12 for every column to the right
      const int64_t capply = one_level ? nc : kend;
      const int nrb = (int)((rest + AW - 1) / AW);
      const int ncb = (int)((capply - k0 - bw + AW - 1) / AW);
      CT* inv = reinterpret_cast<CT*>(panel_inv(work, k0));
      int32_t* flag = flags + k0 / NB;
      hipLaunchKernelGGL(k_panel_nopiv<T>, dim3(nrb + 1), dim3(256), 0, st, n, (const CT*)A, lda,
                         k0, bw, lu11, inv, ipiv, flag, lbound);
      const int nleaf = (int)((n - k0 + TROWS - 1) / TROWS);
      hipLaunchKernelGGL(k_panel_pivot<T>, dim3(nleaf), dim3(256), 0, st, A, lda, n, K0, nc, k0,
                         bw, nleaf, tw, (const int32_t*)flag, lu11, inv, ipiv, info);
      hipLaunchKernelGGL(k_panel_apply<T>, dim3(nrb + ncb + 1), dim3(256), 0, st, n, capply, A,
                         lda, k0, bw, (const CT*)inv, (const CT*)lu11, nrb, ncb);
      if (rest > 0 && restc > 0) {
        *herr = launch_gemm_sub<T>(rest, restc, bw, A + (k0 + bw) * lda + k0, lda,
                                   A + k0 * lda + k0 + bw, lda, A + (k0 + bw) * lda + k0 + bw,
                                   lda, st);
        if (*herr != hipSuccess) return MW_EHIP;
      }
    }
    const int64_t ncr = nc - kend;  // columns right of the block
    if (ncr > 0) {
      CT* X = A + kend;  // U rows of the block, columns [kend, nc)
      const unsigned gx = (unsigned)((ncr + AW - 1) / AW);
      for (int64_t k0 = K0; k0 < kend && !one_level; k0 += NB) {
        const int bw = (int)((kend - k0) < NB ? (kend - k0) : NB);
        const CT* inv = reinterpret_cast<const CT*>(panel_inv(work, k0));
        hipLaunchKernelGGL(k_solve_apply<T>, dim3(gx), dim3(256), 0, st, ncr, inv, k0, bw, X,
                           lda, 1);
        const int64_t below = kend - k0 - bw;
        if (below > 0) {
          *herr = launch_gemm_sub<T>(below, ncr, bw, A + (k0 + bw) * lda + k0, lda, X + k0 * lda,
                                     lda, X + (k0 + bw) * lda, lda, st);
          if (*herr != hipSuccess) return MW_EHIP;
        }
      }
      if (n > kend) {
        *herr = launch_gemm_sub<T>(n - kend, ncr, kend - K0, A + kend * lda + K0, lda,
                                   A + K0 * lda + kend, lda, A + kend * lda + kend, lda, st);
        if (*herr != hipSuccess) return MW_EHIP;
      }
    }
    *herr = hipGetLastError();
    if (*herr != hipSuccess) return MW_EHIP;
  }
  return MW_OK;
}

void mw_set_pivot_bound(double lbound) { g_lbound = lbound; }
double mw_get_pivot_bound() { return g_lbound; }

template <typename T>
static int solve_fwd(int64_t n, int64_t nrhs, const typename cplx_t<T>::type* A, int64_t lda,
                     const int32_t* ipiv, const void* work, typename cplx_t<T>::type* X,
                     int64_t ldx, hipStream_t st, hipError_t* herr) {
  typedef typename cplx_t<T>::type CT;
  const unsigned gx = (unsigned)((nrhs + AW - 1) / AW);
  const int64_t nbo = outer_block(n);
  for (int64_t K0 = 0; K0 < n; K0 += nbo) {
    const int64_t kend = (n - K0) < nbo ? n : K0 + nbo;
    // a block's multiplier columns are stored in the block's final row order
    // (no ipiv: factors without row exchanges, as the symmetric ones)
    if (ipiv)
      hipLaunchKernelGGL(k_laswp_rows<T>, dim3((unsigned)((nrhs + 255) / 256)), dim3(256), 0, st,
                         X, ldx, nrhs, ipiv, K0, kend);
    // Rows inside the outer block take each panel's K = NB update; the rows
    // below it, one K = (kend - K0) update after the block, in the same row
    // order (the next block's interchanges come after it).
    for (int64_t k0 = K0; k0 < kend; k0 += NB) {
      const int bw = (int)((kend - k0) < NB ? (kend - k0) : NB);
      const CT* inv = reinterpret_cast<const CT*>(panel_inv(const_cast<void*>(work), k0));
      hipLaunchKernelGGL(k_solve_apply<T>, dim3(gx), dim3(256), 0, st, nrhs, inv, k0, bw, X, ldx,
                         1);
      const int64_t inner = kend - k0 - bw;
      if (inner > 0) {
        *herr = launch_gemm_sub<T>(inner, nrhs, bw, A + (k0 + bw) * lda + k0, lda, X + k0 * ldx,
                                   ldx, X + (k0 + bw) * ldx, ldx, st);
        if (*herr != hipSuccess) return MW_EHIP;
      }
    }
    if (n > kend) {
      *herr = launch_gemm_sub<T>(n - kend, nrhs, kend - K0, A + kend * lda + K0, lda,
                                 X + K0 * ldx, ldx, X + kend * ldx, ldx, st);
      if (*herr != hipSuccess) return MW_EHIP;
    }
  }
  return MW_OK;
}

template <typename T>
static int solve_bwd(int64_t n, int64_t nrhs, const typename cplx_t<T>::type* A, int64_t lda,
                     const void* work, typename cplx_t<T>::type* X, int64_t ldx, hipStream_t st,
                     hipError_t* herr) {
  typedef typename cplx_t<T>::type CT;
  if (nrhs <= 0 || n <= 0) return MW_OK;
  const int64_t npan = (n + NB - 1) / NB;
  const int64_t npair = npan / 2;
  const unsigned gx = (unsigned)((nrhs + NB - 1) / NB);
  const int64_t klone = (npan - 1) * NB;  // an odd last panel steps alone
  if (nrhs > 256) {
    // Many right-hand sides: each step's grid fills the chip, so the pair's
    // X' is formed once (not in every workgroup) and the rows above take a
    // plain K = 64 update.
    if (npan % 2 == 1) {
      const CT* inv = reinterpret_cast<const CT*>(panel_inv(const_cast<void*>(work), klone));
      hipLaunchKernelGGL(k_solve_apply<T>, dim3((unsigned)((nrhs + AW - 1) / AW)), dim3(256), 0,
                         st, nrhs, inv, klone, (int)(n - klone), X, ldx, 0);
      *herr = launch_gemm_sub<T>(klone, nrhs, n - klone, A + klone, lda, X + klone * ldx, ldx, X,
                                 ldx, st);
      if (*herr != hipSuccess) return MW_EHIP;
    }
    // Two levels, as the factorisation: rows inside the current outer block
    // take each pair's K = 64 update; the rows above the block, one K = nbo
    // update once the block is solved (nbo = NB: every pair updates all rows
    // above, as one level).
    const int64_t nbo = outer_block(n);
    const int64_t pairs_end = (npan % 2 == 1) ? klone : n;
    for (int64_t i = npair - 1; i >= 0; --i) {
      const int64_t kq = 2 * i * NB;
      const int bwp = (int)((n - kq - NB) < NB ? (n - kq - NB) : NB);
      hipLaunchKernelGGL(k_bwd2_pair<T>, dim3(gx), dim3(256), 0, st, kq, bwp, nrhs, A, lda, work,
                         X, ldx);
      const int64_t lo = nbo > NB ? (kq / nbo) * nbo : 0;
      if (kq > lo) {
        *herr = launch_gemm_sub<T>(kq - lo, nrhs, NB + bwp, A + lo * lda + kq, lda, X + kq * ldx,
                                   ldx, X + lo * ldx, ldx, st);
        if (*herr != hipSuccess) return MW_EHIP;
      }
      if (kq == lo && lo > 0) {
        const int64_t hi = (lo + nbo) < pairs_end ? lo + nbo : pairs_end;
        *herr = launch_gemm_sub<T>(lo, nrhs, hi - lo, A + lo, lda, X + lo * ldx, ldx, X, ldx, st);
        if (*herr != hipSuccess) return MW_EHIP;
      }
    }
    *herr = hipGetLastError();
    return *herr == hipSuccess ? MW_OK : MW_EHIP;
  }
  if (npan % 2 == 1 && klone > 0) {
    const CT* ui = reinterpret_cast<const CT*>(panel_inv(const_cast<void*>(work), klone)) + NB * NB;
    hipLaunchKernelGGL(k_bwd_fused<T>, dim3(gx, (unsigned)((klone + 63) / 64)), dim3(256), 0, st,
                       klone, (int)(n - klone), nrhs, A, lda, ui, X, ldx);
  }
  for (int64_t i = npair - 1; i >= 1; --i) {
    const int64_t kq = 2 * i * NB;
    const int bwp = (int)((n - kq - NB) < NB ? (n - kq - NB) : NB);
    hipLaunchKernelGGL(k_bwd2<T>, dim3(gx, (unsigned)((kq + 63) / 64)), dim3(256), 0, st, kq, bwp,
                       nrhs, A, lda, work, X, ldx);
  }
  if (npair > 0)
    hipLaunchKernelGGL(k_bwd2_final<T>, dim3(gx, (unsigned)npair), dim3(256), 0, st, n, nrhs, A,
                       lda, work, X, ldx);
  if (npan % 2 == 1) {
    const CT* inv = reinterpret_cast<const CT*>(panel_inv(const_cast<void*>(work), klone));
    hipLaunchKernelGGL(k_solve_apply<T>, dim3((unsigned)((nrhs + AW - 1) / AW)), dim3(256), 0, st,
                       nrhs, inv, klone, (int)(n - klone), X, ldx, 0);
  }
  *herr = hipGetLastError();
  return *herr == hipSuccess ? MW_OK : MW_EHIP;
}

template <typename T>
int mw_getrs_t(int64_t n, int64_t nrhs, const void* Av, int64_t lda, const int32_t* ipiv,
               const void* work, void* Xv, int64_t ldx, hipStream_t st, hipError_t* herr) {
  typedef typename cplx_t<T>::type CT;
  const CT* A = reinterpret_cast<const CT*>(Av);
  CT* X = reinterpret_cast<CT*>(Xv);
  int rc = solve_fwd<T>(n, nrhs, A, lda, ipiv, work, X, ldx, st, herr);
  if (rc) return rc;
  rc = solve_bwd<T>(n, nrhs, A, lda, work, X, ldx, st, herr);
  if (rc) return rc;
  *herr = hipGetLastError();
  return *herr == hipSuccess ? MW_OK : MW_EHIP;
}

// A is n x (n + nrhs): columns [n, n+nrhs) hold B on entry and X = A^{-1} B on
// return; the forward half rides inside the factorisation.
template <typename T>
int mw_gesv_t(int64_t n, int64_t nrhs, void* Av, int64_t lda, int32_t* ipiv, void* work,
              int32_t* info, const volatile int32_t* cancel, hipStream_t st, hipError_t* herr) {
  typedef typename cplx_t<T>::type CT;
  int rc = mw_getrf_t<T>(n, n + nrhs, Av, lda, ipiv, work, info, cancel, st, herr);
  if (rc) return rc;
  if (nrhs > 0) {
    CT* A = reinterpret_cast<CT*>(Av);
    rc = solve_bwd<T>(n, nrhs, A, lda, work, A + n, lda, st, herr);
    if (rc) return rc;
  }
  *herr = hipGetLastError();
  return *herr == hipSuccess ? MW_OK : MW_EHIP;
}

// The symmetric factorisation's workspace: mw_getrf_ws_bytes(n), then the
// diagonal-block inverses N (nblk x 512 x 512), V (the same) and a
// 512 x SYW staging block for the out-of-place block products (16-byte slots,
// so one size serves FP64 and FP32).
constexpr int64_t SYW = 1024;  // right-hand-side columns per block product
int64_t mw_sysv_ws_bytes(int64_t n) {
  const int64_t bz = sym_block(n);
  if (bz == 0) return mw_getrf_ws_bytes(n);
  const int64_t nblk = (n + bz - 1) / bz;
  return mw_getrf_ws_bytes(n) + 2 * nblk * SYB * SYB * 16 + SYB * SYW * 16;
}

template <typename T>
struct SymWs {
  typedef typename cplx_t<T>::type CT;
  CT *N, *V, *tmp;
  SymWs(void* work, int64_t n) {
    const int64_t bz = sym_block(n), nblk = (n + bz - 1) / bz;
    char* base = reinterpret_cast<char*>(work) + mw_getrf_ws_bytes(n);
    N = reinterpret_cast<CT*>(base);
    V = reinterpret_cast<CT*>(base + nblk * SYB * SYB * 16);
    tmp = reinterpret_cast<CT*>(base + 2 * nblk * SYB * SYB * 16);
  }
};

// N_b and V_b of every outer block (see k_blkinv_init): ten launches in all.
template <typename T>
static hipError_t sym_block_inverses(int64_t n, const typename cplx_t<T>::type* A, int64_t lda,
                                     void* work, hipStream_t st) {
  const int bz = (int)sym_block(n);
  const int64_t nblk = (n + bz - 1) / bz;
  SymWs<T> w(work, n);
  hipLaunchKernelGGL(k_blkinv_init<T>, dim3(bz, (unsigned)nblk), dim3(256), 0, st, n, bz,
                     (const void*)work, w.N, w.V);
  for (int s = NB; s < bz; s *= 2) {
    const int tiles = ((s + SBN - 1) / SBN) * ((s + SBM - 1) / SBM);
    const unsigned grid = (unsigned)(nblk * (bz / (2 * s)) * tiles);
    hipLaunchKernelGGL((k_blkinv_level<T, 1>), dim3(grid), dim3(256), 0, st, s, n, bz, A, lda,
                       w.N, w.V);
    hipLaunchKernelGGL((k_blkinv_level<T, 2>), dim3(grid), dim3(256), 0, st, s, n, bz, A, lda,
                       w.N, w.V);
  }
  hipLaunchKernelGGL(k_blkinv_vt<T>, dim3(bz / 64, bz / 64, (unsigned)nblk), dim3(256), 0, st,
                     n, bz, A, lda, (const typename cplx_t<T>::type*)w.N, w.V);
  return hipGetLastError();
}

// X <- Z^{-1} X with the symmetric factors (Z = L D L^T, U = D L^T) and the
// block inverses: per 512-row block one product with -N_b (forward) or -V_b
// (backward) on the block's rows, moved aside first, and one K = 512 update
// of the rows below (forward) or above (backward).  first_row (host, nrhs
// entries, or NULL): the first row where each right-hand side may be
// nonzero.  The forward pass then works, in block b, on the columns up to the
// last one that has started by the block's end: the others are still zero
// there, and L^{-1} keeps them so (a port's drive column starts at its feed
// row, so the forward pass does a third of the dense work on an array).
template <typename T>
static int sym_solve(int64_t n, int64_t nrhs, const typename cplx_t<T>::type* A, int64_t lda,
                     void* work, typename cplx_t<T>::type* X, int64_t ldx,
                     const int32_t* first_row, hipStream_t st, hipError_t* herr) {
  typedef typename cplx_t<T>::type CT;
  if (nrhs <= 0) return MW_OK;
  SymWs<T> w(work, n);
  const int64_t bz = sym_block(n), nblk = (n + bz - 1) / bz;
  auto diag = [&](const CT* M, bool lower, int64_t K0, int64_t bs, int64_t ncols) -> hipError_t {
    for (int64_t c0 = 0; c0 < ncols; c0 += SYW) {
      const int64_t cw = (ncols - c0) < SYW ? (ncols - c0) : SYW;
      const int64_t tot = bs * cw;
      hipLaunchKernelGGL(k_move_zero<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, bs,
                         cw, X + K0 * ldx + c0, ldx, w.tmp, cw);
      const unsigned grid = (unsigned)(((bs + 63) / 64) * ((cw + NB - 1) / NB));
      if (lower)
        hipLaunchKernelGGL((k_blk_trmm<T, true>), dim3(grid), dim3(256), 0, st, (int)bs, (int)cw, M,
                           (const CT*)w.tmp, X + K0 * ldx + c0, ldx);
      else
        hipLaunchKernelGGL((k_blk_trmm<T, false>), dim3(grid), dim3(256), 0, st, (int)bs, (int)cw,
                           M, (const CT*)w.tmp, X + K0 * ldx + c0, ldx);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  };
  for (int64_t b = 0; b < nblk; ++b) {
    const int64_t K0 = b * bz, kend = (n - K0) < bz ? n : K0 + bz, bs = kend - K0;
    int64_t act = nrhs;
    if (first_row) {
      act = 0;
      for (int64_t c = 0; c < nrhs; ++c)
        if (first_row[c] < kend) act = c + 1;
    }
    if (act == 0) continue;
    *herr = diag(w.N + b * SYB * SYB, true, K0, bs, act);
    if (*herr != hipSuccess) return MW_EHIP;
    if (n > kend) {
      *herr = launch_gemm_sub<T>(n - kend, act, bs, A + kend * lda + K0, lda, X + K0 * ldx, ldx,
                                 X + kend * ldx, ldx, st);
      if (*herr != hipSuccess) return MW_EHIP;
    }
  }
  for (int64_t b = nblk - 1; b >= 0; --b) {
    const int64_t K0 = b * bz, kend = (n - K0) < bz ? n : K0 + bz, bs = kend - K0;
    *herr = diag(w.V + b * SYB * SYB, false, K0, bs, nrhs);
    if (*herr != hipSuccess) return MW_EHIP;
    if (K0 > 0) {
      *herr = launch_gemm_sub<T>(K0, nrhs, bs, A + K0, lda, X + K0 * ldx, ldx, X, ldx, st);
      if (*herr != hipSuccess) return MW_EHIP;
    }
  }
  *herr = hipGetLastError();
  return *herr == hipSuccess ? MW_OK : MW_EHIP;
}

// Z = Z^T, factored without row exchanges.  Two levels only (n > 4096): per
// outer block the panel chain without its tournament fallback, the block's U
// rows right of it by symmetry (k_sym_u_rows), then the trailing update on
// the blocks on or below the block diagonal (the part the later blocks read;
// the rest of the upper triangle is overwritten by their U rows).  Then the
// 512 x 512 diagonal blocks' inverses (sym_block_inverses) and the
// right-hand sides (columns [n, n + nrhs) of A) by block products.  A panel
// whose threshold check fails is not pivoted: info[0] = -1 and the result is
// not valid (refactor the original matrix with mw_gesv_t).  One level:
// mw_gesv_t.  work: mw_sysv_ws_bytes(n).  first_row: see sym_solve.
template <typename T>
int mw_sysv_t(int64_t n, int64_t nrhs, void* Av, int64_t lda, int32_t* ipiv, void* work,
              int32_t* info, const volatile int32_t* cancel, const int32_t* first_row,
              hipStream_t st, hipError_t* herr) {
  typedef typename cplx_t<T>::type CT;
  const int64_t nbo = sym_block(n);
  CT* A = reinterpret_cast<CT*>(Av);
  const int64_t npan = (n + NB - 1) / NB;
  const int64_t nodes = tour_nodes(n);
  char* wp = reinterpret_cast<char*>(work);
  CT* lu11 = reinterpret_cast<CT*>(wp);
  int32_t* flags = reinterpret_cast<int32_t*>(wp + (int64_t)NB * NB * 16 * (1 + 2 * npan));
  hipLaunchKernelGGL(k_zero_i32, dim3((unsigned)((npan + nodes + 255) / 256)), dim3(256), 0, st,
                     flags, npan + nodes);
  hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, st, info, (int64_t)1);
  if (nbo == 0) {
    // one level: the look-ahead chain over [A | B] with the update on the
    // lower part and the panel-diagonal blocks, then the back substitution
    TourWs tw;
    tw.counter = flags + npan;
    tw.ncand = tw.counter + nodes;
    tw.cand = tw.ncand + nodes;
    int rc = getrf_lookahead<T, true>(n, n + nrhs, A, lda, ipiv, work, flags, tw, info, cancel,
                                      st, herr);
    if (rc) return rc;
    hipLaunchKernelGGL(k_sym_info, dim3(1), dim3(256), 0, st, (const int32_t*)flags, npan, info);
    if (nrhs > 0) {
      rc = solve_bwd<T>(n, nrhs, A, lda, work, A + n, lda, st, herr);
      if (rc) return rc;
    }
    *herr = hipGetLastError();
    return *herr == hipSuccess ? MW_OK : MW_EHIP;
  }
  const double lbound = g_lbound;
  int resident = 0;
  if ((*herr = lookahead_resident<T>(&resident)) != hipSuccess) return MW_EHIP;
  for (int64_t K0 = 0; K0 < n; K0 += nbo) {
    const int64_t kend = (n - K0) < nbo ? n : K0 + nbo;
    const int64_t bs = kend - K0;
    // Inside the block, the one-level look-ahead chain with the block's last
    // column as its right edge: the block's first panel factored alone, then
    // per panel k_panel_apply_la (L21, the block's U12, the next panel's
    // columns updated) and k_lookahead (the next panel's factor and threshold
    // check beside the rest of the block's K = 32 update, rows down to n).
    {
      const int bw0 = (int)(bs < NB ? bs : NB);
      const int nrb0 = (int)((n - K0 - bw0 + AW - 1) / AW);
      hipLaunchKernelGGL(k_panel_nopiv<T>, dim3(nrb0 + 1), dim3(256), 0, st, n, (const CT*)A, lda,
                         K0, bw0, lu11, reinterpret_cast<CT*>(panel_inv(work, K0)), ipiv,
                         flags + K0 / NB, lbound);
    }
    for (int64_t k0 = K0; k0 < kend; k0 += NB) {
      if (cancel && *cancel) return MW_EABORT;
      const int bw = (int)((kend - k0) < NB ? (kend - k0) : NB);
      const int64_t kn = k0 + bw, rest = n - kn, restc = kend - kn;
      const int bwn = (int)(restc < NB ? restc : NB);
      const int nrb = (int)((rest + LAW - 1) / LAW);
      const int ncb = (int)((restc - bwn + LAW - 1) / LAW);
      CT* inv = reinterpret_cast<CT*>(panel_inv(work, k0));
      hipLaunchKernelGGL(k_panel_apply_la<T>, dim3(nrb + ncb + 1), dim3(256), 0, st, n, kend, A,
                         lda, k0, bw, bwn, (const CT*)inv, (const CT*)lu11, nrb, ncb);
      if (bwn > 0) {
        const int nnp = (int)((rest - bwn + AW - 1) / AW) + 1;
        const int64_t ncols = restc - bwn;
        const int gxn = (int)((ncols + NB - 1) / NB);
        const int ntiles = gxn * (int)((rest + 63) / 64);
        // update blocks beside the factor's: what stays resident with them, and
        // at least half the device's worth when the factor's blocks alone fill
        // it (tall panels: they are then dispatched as those finish)
        const int G = std::max(1, std::min(ntiles, std::max(resident - nnp - 1, resident / 2)));
        hipLaunchKernelGGL(k_lookahead<T>, dim3((unsigned)(nnp + 1 + (ntiles > 0 ? G : 0))),
                           dim3(256), 0, st, n, kend, A, lda, k0, bw, bwn, (const CT*)inv, lu11,
                           reinterpret_cast<CT*>(panel_inv(work, kn)), ipiv, flags + kn / NB,
                           lbound, nnp, gxn > 0 ? gxn : 1, ntiles, G);
      }
    }
    if (n > kend) {
      hipLaunchKernelGGL(k_sym_u_rows<T>, dim3((unsigned)((bs + 63) / 64),
                                               (unsigned)((n - kend + 63) / 64)),
                         dim3(256), 0, st, K0, (int)bs, kend, n, A, lda);
      *herr = launch_gemm_stair<T>(n - kend, bs, A + kend * lda + K0, lda, A + K0 * lda + kend,
                                   lda, A + kend * lda + kend, lda, st);
      if (*herr != hipSuccess) return MW_EHIP;
    }
    *herr = hipGetLastError();
    if (*herr != hipSuccess) return MW_EHIP;
  }
  hipLaunchKernelGGL(k_sym_info, dim3(1), dim3(256), 0, st, (const int32_t*)flags, npan, info);
  *herr = sym_block_inverses<T>(n, A, lda, work, st);
  if (*herr != hipSuccess) return MW_EHIP;
  return sym_solve<T>(n, nrhs, A, lda, work, A + n, lda, first_row, st, herr);
}

// X <- Z^{-1} X with the factors of mw_sysv_t (its workspace); one level
// (n <= 4096: mw_gesv_t's pivoted factors): mw_getrs_t with ipiv.
template <typename T>
int mw_sytrs_t(int64_t n, int64_t nrhs, const void* Av, int64_t lda, const int32_t* ipiv,
               const void* work, void* Xv, int64_t ldx, hipStream_t st, hipError_t* herr) {
  typedef typename cplx_t<T>::type CT;
  if (sym_block(n) == 0) return mw_getrs_t<T>(n, nrhs, Av, lda, ipiv, work, Xv, ldx, st, herr);
  return sym_solve<T>(n, nrhs, reinterpret_cast<const CT*>(Av), lda, const_cast<void*>(work),
                      reinterpret_cast<CT*>(Xv), ldx, nullptr, st, herr);
}

template int mw_sytrs_t<double>(int64_t, int64_t, const void*, int64_t, const int32_t*,
                                const void*, void*, int64_t, hipStream_t, hipError_t*);
template int mw_sytrs_t<float>(int64_t, int64_t, const void*, int64_t, const int32_t*,
                               const void*, void*, int64_t, hipStream_t, hipError_t*);
template int mw_sysv_t<double>(int64_t, int64_t, void*, int64_t, int32_t*, void*, int32_t*,
                               const volatile int32_t*, const int32_t*, hipStream_t,
                               hipError_t*);
template int mw_sysv_t<float>(int64_t, int64_t, void*, int64_t, int32_t*, void*, int32_t*,
                               const volatile int32_t*, const int32_t*, hipStream_t,
                               hipError_t*);
template int mw_getrf_t<double>(int64_t, int64_t, void*, int64_t, int32_t*, void*, int32_t*,
                                const volatile int32_t*, hipStream_t, hipError_t*);
template int mw_getrf_t<float>(int64_t, int64_t, void*, int64_t, int32_t*, void*, int32_t*,
                               const volatile int32_t*, hipStream_t, hipError_t*);
template int mw_gesv_t<double>(int64_t, int64_t, void*, int64_t, int32_t*, void*, int32_t*,
                               const volatile int32_t*, hipStream_t, hipError_t*);
template int mw_gesv_t<float>(int64_t, int64_t, void*, int64_t, int32_t*, void*, int32_t*,
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
hipError_t mw_gemm_tn_add_z(int64_t m, int64_t n, int64_t k, const void* A, int64_t lda,
                            const void* B, int64_t ldb, void* C, int64_t ldc, hipStream_t st) {
  return launch_gemm_sub<double, true>(m, n, k, A, lda, B, ldb, C, ldc, st);
}
hipError_t mw_gemm_tn_add_c(int64_t m, int64_t n, int64_t k, const void* A, int64_t lda,
                            const void* B, int64_t ldb, void* C, int64_t ldc, hipStream_t st) {
  return launch_gemm_sub<float, true>(m, n, k, A, lda, B, ldb, C, ldc, st);
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

hipError_t mw_launch_add_sparse(int64_t nnz, const int32_t* rows, const int32_t* cols,
                                const void* vals, void* Z, int64_t ldz, hipStream_t st) {
  if (nnz <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_add_sparse, dim3((unsigned)((nnz + 255) / 256)), dim3(256), 0, st, nnz,
                     rows, cols, (const double2*)vals, (double2*)Z, ldz);
  return hipGetLastError();
}

hipError_t mw_launch_small_solve(int ns, int nt, void* S, int64_t lds, void* T, int64_t ldt,
                                 int32_t* info, hipStream_t st) {
  hipLaunchKernelGGL(k_small_solve, dim3(1), dim3(256), 0, st, ns, nt, (double2*)S, lds,
                     (double2*)T, ldt, info);
  return hipGetLastError();
}

hipError_t mw_launch_z2c(int64_t rows, int64_t cols, const void* src, int64_t lds, void* dst,
                         int64_t ldd, hipStream_t st) {
  if (rows <= 0 || cols <= 0) return hipSuccess;
  const int64_t tot = rows * cols;
  hipLaunchKernelGGL(k_z2c, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, rows, cols,
                     (const double2*)src, lds, (float2*)dst, ldd);
  return hipGetLastError();
}

hipError_t mw_launch_c2z(int64_t rows, int64_t cols, const void* src, int64_t lds, void* dst,
                         int64_t ldd, int accumulate, hipStream_t st) {
  if (rows <= 0 || cols <= 0) return hipSuccess;
  const int64_t tot = rows * cols;
  hipLaunchKernelGGL(k_c2z, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, rows, cols,
                     (const float2*)src, lds, (double2*)dst, ldd, accumulate);
  return hipGetLastError();
}

// scratch: 2 p column sums then the two maxima (16 (p + 1) bytes), cleared here
hipError_t mw_launch_readout_check(int64_t n, int64_t p, const void* X0, int64_t ldx,
                                   const void* C, int64_t ldc, const void* T, int64_t ldt, void* Y,
                                   int64_t ldy, double tol, void* scratch, double* stats,
                                   int32_t* gate, hipStream_t st) {
  double* acc = reinterpret_cast<double*>(scratch);
  unsigned long long* mx = reinterpret_cast<unsigned long long*>(acc + 2 * p);
  const int64_t words = 4 * (p + 1);
  hipLaunchKernelGGL(k_zero_i32, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, st,
                     reinterpret_cast<int32_t*>(scratch), words);
  const int64_t rpb = 256;
  hipLaunchKernelGGL(k_colsq2, dim3((unsigned)((p + 63) / 64), (unsigned)((n + rpb - 1) / rpb)),
                     dim3(256), 0, st, n, p, (const double2*)X0, ldx, (const double2*)C, ldc, rpb,
                     acc);
  hipLaunchKernelGGL(k_add_maxabs, dim3((unsigned)((p * p + 255) / 256)), dim3(256), 0, st, p,
                     (const double2*)T, ldt, (double2*)Y, ldy, mx);
  hipLaunchKernelGGL(k_readout_verdict, dim3(1), dim3(256), 0, st, p, (const double*)acc,
                     (const unsigned long long*)mx, tol, stats, gate);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Work lists of the batched factorisation.  A grid sized for the largest
// problem would carry, at every panel step, workgroups that only find their
// problem too small and exit (most of them, for uneven sizes); problem sizes
// are fixed per plan, so the (problem, block) pairs that have work are listed
// once on the host: per step, list 0 = threshold-check row blocks, 1 =
// tournament leaves, 2 = panel-apply blocks, 3 = trailing-update tiles
// (problem << 16 | tile row << 8 | tile column).  off[4 s + l] .. off[4 s + l
// + 1] delimits list l of step s.  hp is sorted by n, largest first; needs
// nprob <= 65536 and n <= 4096.
// ---------------------------------------------------------------------------
void mw_sweep_work(const mw_sweep_prob* hp, int nprob, int gemm_variant,
                   std::vector<uint32_t>& items, std::vector<int64_t>& off) {
  const int64_t nmax = hp[0].n;
  const int64_t tm = gemm_variant ? GBM : SBM, tn = gemm_variant ? GBN : SBN;
  items.clear();
  off.clear();
  for (int64_t k0 = 0; k0 < nmax; k0 += NB) {
    for (int list = 0; list < 4; ++list) {
      off.push_back((int64_t)items.size());
      for (int i = 0; i < nprob && hp[i].n > k0; ++i) {
        const int64_t n = hp[i].n, nc = n + hp[i].ncols;
        const int64_t bw = (n - k0) < NB ? (n - k0) : NB;
        const int64_t rest = n - k0 - bw;
        const uint32_t key = (uint32_t)i << 16;
        if (list == 0) {
          for (int64_t b = 0; b < (rest + AW - 1) / AW; ++b) items.push_back(key | (uint32_t)b);
        } else if (list == 1) {
          for (int64_t l = 0; l < (n - k0 + TROWS - 1) / TROWS; ++l)
            items.push_back(key | (uint32_t)l);
        } else if (list == 2) {
          const int64_t nblk = (rest + AW - 1) / AW + (nc - k0 - bw + AW - 1) / AW + 1;
          for (int64_t b = 0; b < nblk; ++b) items.push_back(key | (uint32_t)b);
        } else if (rest > 0) {
          const int64_t ncol = nc - k0 - bw;
          for (int64_t y = 0; y < (rest + tm - 1) / tm; ++y)
            for (int64_t x = 0; x < (ncol + tn - 1) / tn; ++x)
              items.push_back(key | (uint32_t)(y << 8) | (uint32_t)x);
        }
      }
    }
  }
  off.push_back((int64_t)items.size());
}

// ---------------------------------------------------------------------------
// Batched sweep driver.  `nact(k0)` = number of problems (sorted by n,
// largest first) with n > k0; `hp` is the host copy of the sorted table.
// ---------------------------------------------------------------------------
int mw_sweep_lu_run(const mw_sweep_prob* dP, const mw_ports* dports, const mw_sweep_prob* hp,
                    int nprob, int max_ports, int max_con, int64_t max_scatter,
                    int gemm_variant, double lbound, const uint32_t* items,
                    const int64_t* off, const volatile int32_t* cancel, hipStream_t st,
                    hipError_t* herr) {
  if (nprob <= 0) return MW_OK;
  const int64_t nmax = hp[0].n;
  int64_t maxcols = 0, maxrhs = 0, maxzero = 0;
  for (int i = 0; i < nprob; ++i) {
    maxcols = hp[i].ncols > maxcols ? hp[i].ncols : maxcols;
    maxrhs = hp[i].n * hp[i].ncols > maxrhs ? hp[i].n * hp[i].ncols : maxrhs;
  }
  maxzero = (nmax + NB - 1) / NB + tour_nodes(nmax);
  hipLaunchKernelGGL(kb_zero_ws, dim3((unsigned)((maxzero + 255) / 256), nprob), dim3(256), 0, st, dP);
  hipLaunchKernelGGL(kb_rhs, dim3((unsigned)((maxrhs + 255) / 256), nprob), dim3(256), 0, st, dP,
                     dports, 0);
  hipLaunchKernelGGL(kb_rhs, dim3((unsigned)((max_scatter + 255) / 256), nprob), dim3(256), 0, st,
                     dP, dports, 1);
  auto nact = [&](int64_t k0) {
    int lo = 0, hi = nprob;  // first index with n <= k0
    while (lo < hi) {
      const int mid = (lo + hi) / 2;
      if (hp[mid].n > k0) lo = mid + 1; else hi = mid;
    }
    return lo;
  };
  for (int64_t k0 = 0; k0 < nmax; k0 += NB) {
    if (cancel && *cancel) return MW_EABORT;
    const int na = nact(k0);
    // the step's work lists: check, pivot, apply, update (mw_sweep_work)
    const int64_t* o = off + 4 * (k0 / NB);
    auto count = [&](int l) { return (unsigned)(o[l + 1] - o[l]); };
    hipLaunchKernelGGL(kb_panel_a11, dim3(na), dim3(256), 0, st, dP, k0, lbound);
    if (count(0))
      hipLaunchKernelGGL(kb_panel_check, dim3(count(0)), dim3(256), 0, st, dP, k0, lbound,
                         items + o[0]);
    if (count(1))
      hipLaunchKernelGGL(kb_panel_pivot, dim3(count(1)), dim3(256), 0, st, dP, k0, items + o[1]);
    if (count(2))
      hipLaunchKernelGGL(kb_panel_apply, dim3(count(2)), dim3(256), 0, st, dP, k0, items + o[2]);
    if (count(3)) {
      if (gemm_variant)
        hipLaunchKernelGGL(kb_gemm<true>, dim3(count(3)), dim3(256), 0, st, dP, k0, 0, items + o[3]);
      else
        hipLaunchKernelGGL(kb_gemm<false>, dim3(count(3)), dim3(256), 0, st, dP, k0, 0,
                           items + o[3]);
    }
  }
  const int64_t nlast = ((nmax + NB - 1) / NB - 1) * NB;
  for (int64_t k0 = nlast; k0 >= 0; k0 -= NB) {
    const int na = nact(k0);
    hipLaunchKernelGGL(kb_bwd_apply, dim3((unsigned)((maxcols + AW - 1) / AW), na), dim3(256), 0, st,
                       dP, k0);
    if (k0 > 0) {
      if (gemm_variant)
        hipLaunchKernelGGL(kb_gemm<true>, dim3((unsigned)((maxcols + GBN - 1) / GBN),
                                               (unsigned)((k0 + GBM - 1) / GBM), na),
                           dim3(256), 0, st, dP, k0, 1, (const uint32_t*)nullptr);
      else
        hipLaunchKernelGGL(kb_gemm<false>, dim3((unsigned)((maxcols + SBN - 1) / SBN),
                                                (unsigned)((k0 + SBM - 1) / SBM), na),
                           dim3(256), 0, st, dP, k0, 1, (const uint32_t*)nullptr);
    }
  }
  if (max_con > 0) {
    hipLaunchKernelGGL(kb_schur_ax, dim3(max_con, nprob), dim3(64), 0, st, dP, dports);
    hipLaunchKernelGGL(kb_small_solve, dim3(nprob), dim3(256), 0, st, dP, dports);
    hipLaunchKernelGGL(kb_schur_update, dim3((unsigned)((nmax + 255) / 256), nprob), dim3(256), 0,
                       st, dP, dports);
  }
  hipLaunchKernelGGL(kb_readout, dim3(max_ports, nprob), dim3(64), 0, st, dP, dports);
  *herr = hipGetLastError();
  return *herr == hipSuccess ? MW_OK : MW_EHIP;
}

// ---------------------------------------------------------------------------
// Persistent batched solve (kb_lu_wg): the claim counter zeroed, then one
// launch of `per_cu` workgroups per compute unit (at most the occupancy
// limit, at most one per problem); `two`: 128-column outer blocks.
// `cancel`: a device-readable flag or NULL.
// ---------------------------------------------------------------------------
int mw_sweep_wg_run(const mw_sweep_prob* dP, const mw_ports* dports, int nprob, int two,
                    int per_cu, double lbound, int32_t* d_next, const volatile int32_t* cancel,
                    hipStream_t st, hipError_t* herr) {
  if (nprob <= 0) return MW_OK;
  int dev = 0, cus = 0, occ = 0;
  if ((*herr = hipGetDevice(&dev)) != hipSuccess) return MW_EHIP;
  if ((*herr = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) !=
      hipSuccess)
    return MW_EHIP;
  *herr = two ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kb_lu_wg<true>, 256, 0)
              : hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kb_lu_wg<false>, 256, 0);
  if (*herr != hipSuccess) return MW_EHIP;
  const int grid = std::max(1, std::min(nprob, cus * std::max(1, std::min(per_cu, occ))));
  hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, st, d_next, (int64_t)1);
  if (two)
    hipLaunchKernelGGL(kb_lu_wg<true>, dim3(grid), dim3(256), 0, st, dP, dports, nprob, d_next,
                       lbound, cancel);
  else
    hipLaunchKernelGGL(kb_lu_wg<false>, dim3(grid), dim3(256), 0, st, dP, dports, nprob, d_next,
                       lbound, cancel);
  *herr = hipGetLastError();
  return *herr == hipSuccess ? MW_OK : MW_EHIP;
}
