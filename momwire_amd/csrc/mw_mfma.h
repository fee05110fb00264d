// MFMA building blocks shared by the dense linear algebra and the readout:
// the 16 x 16 x 4 FP64 / FP32 instructions with their output layouts, and a
// complex product staged in LDS (A as reals, B as complex expanded to
// [[br, bi], [-bi, br]] when fragments are formed).
#pragma once

#include "mw_common.h"

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

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

// The MFMA core of the block products: acc = A_r B_exp with A_r (RA x 64
// reals, LDS rows of 65) and B (32 x CB complex, LDS rows of ldbs), expanded
// as above.  The (RA/16) x (2CB/16) output tiles of 16 x 16 are numbered
// row-major, TPW per wave: wave w holds tiles w TPW .. w TPW + TPW - 1, and
// acc[j][r] sits at row 16 tr + Mfma<T>::row(lane, r), real column
// 16 tc + lane % 16 of tile t = (tr, tc).
template <int RA, int CB>
struct TileGrid {
  static constexpr int TC = 2 * CB / 16;          // tile columns
  static constexpr int TPW = (RA / 16) * TC / 4;  // tiles per wave
};

// KR: real k extent (2 x the complex k); LDA: row stride of As in reals.
// The operands of k-step s + 1 are read from LDS before the MFMAs of step s
// issue, so an LDS read's latency hides behind the previous step's MFMAs.
template <typename T, int RA, int CB, bool ZERO = true, int KR = 64, int LDA = 65>
MW_DEV void lds_mfma(const T* As, const typename cplx_t<T>::type* Bs, int ldbs,
                     typename Mfma<T>::acc_t (&acc)[TileGrid<RA, CB>::TPW]) {
  typedef typename cplx_t<T>::type CT;
  constexpr int TC = TileGrid<RA, CB>::TC, TPW = TileGrid<RA, CB>::TPW;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (ZERO)
#pragma unroll
    for (int j = 0; j < TPW; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[j][r] = 0;
  T a[2][TPW];
  CT bv[2][TPW];
  auto fetch = [&](int buf, int ks) {
    const int kr = ks + (lane >> 4);
#pragma unroll
    for (int j = 0; j < TPW; ++j) {
      const int t = TPW * wave + j;
      const int tr = t / TC, tc = t % TC;
      a[buf][j] = As[(16 * tr + (lane & 15)) * LDA + kr];
      bv[buf][j] = Bs[(kr >> 1) * ldbs + ((16 * tc + (lane & 15)) >> 1)];
    }
  };
  fetch(0, 0);
#pragma unroll
  for (int ks = 0; ks < KR; ks += 4) {
    const int cur = (ks >> 2) & 1;
    if (ks + 4 < KR) fetch(cur ^ 1, ks + 4);
    const int kpar = (ks + (lane >> 4)) & 1;
#pragma unroll
    for (int j = 0; j < TPW; ++j) {
      const int cr = 16 * ((TPW * wave + j) % TC) + (lane & 15);
      const CT b = bv[cur][j];
      const T bb = (kpar == 0) ? ((cr & 1) ? b.y : b.x) : ((cr & 1) ? b.x : -b.y);
      acc[j] = Mfma<T>::mma(a[cur][j], bb, acc[j]);
    }
  }
}

}  // namespace
