// Learned Sommerfeld surrogate on gfx950: a SiLU MLP with 4 hidden layers of
// 256, BF16 weights and activations, FP32 accumulation on MFMA
// (v_mfma_f32_16x16x32_bf16), trained by the kernels below.
//
// The network maps the four inputs (rho/lambda, (z+z')/lambda, eps_r,
// sigma/(omega eps0)) of one Sommerfeld node to the four remainder surfaces
// (eight reals, divided by c1 k2^2 = -j omega mu/(4 pi)); it replaces the six
// lambda-integrals of the exact engine when a grid is filled.  No counterpart
// exists in the reference.
//
// Layout (all sample-major, n a multiple of 64):
//   H0 [n][64]  bf16 features (8 used)      Z_l [n][256] fp32 pre-activations
//   H_l [n][256] bf16 SiLU(Z_{l-1})          OUT [n][64] fp32 (8 used)
//   G_l [n][width_l] bf16 dLoss/dZ_l
// Parameters (fp32 master, then bf16 copies W [out][in] and W^T [in][out]):
//   per layer l: W_l [out_l][in_l], b_l [out_l];  in = 64,256,256,256,256,
//   out = 256,256,256,256,64.
//
// GEMMs.  C[m][n] = sum_k A[m][k] B[n][k] with k contiguous in both operands
// serves the forward pass (A = W, B = H) and the backward data pass (A = W^T,
// B = G): each lane feeds the MFMA straight from 16-byte loads.  The weight
// gradient dW[m][n] = sum_s G[s][m] H[s][n] is split over the samples and
// accumulated with FP32 atomics.
#include "mw_common.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int NL = 5;
constexpr int FIN[NL] = {64, 256, 256, 256, 256};
constexpr int FOUT[NL] = {256, 256, 256, 256, 64};
constexpr int NF_USED = 12;  // features in use
constexpr int NO_USED = 8;   // outputs in use

__host__ __device__ constexpr int64_t param_off(int l) {
  int64_t o = 0;
  for (int i = 0; i < l; ++i) o += (int64_t)FOUT[i] * FIN[i] + FOUT[i];
  return o;
}
__host__ __device__ constexpr int64_t w16_off(int l) {
  int64_t o = 0;
  for (int i = 0; i < l; ++i) o += (int64_t)FOUT[i] * FIN[i];
  return o;
}
constexpr int64_t N_PARAMS = param_off(NL);
constexpr int64_t N_W16 = w16_off(NL);

MW_DEV uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}
MW_DEV float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

MW_DEV float sigm(float z) { return 1.0f / (1.0f + expf(-z)); }

enum Epi { EPI_HIDDEN = 0, EPI_OUT = 1, EPI_BWD = 2 };

// C = A . B^T over a 64 x 64 tile per workgroup (4 waves of 32 x 32).
template <int EPI>
__global__ __launch_bounds__(256) void k_gemm_nt(const uint16_t* __restrict__ A, int K,
                                                  const uint16_t* __restrict__ B, int M,
                                                  const float* __restrict__ bias,
                                                  const float* __restrict__ zprev,
                                                  float* __restrict__ zout,
                                                  uint16_t* __restrict__ hout) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m0 = blockIdx.y * 64 + (wave >> 1) * 32;
  const int n0 = blockIdx.x * 64 + (wave & 1) * 32;
  const int r = lane & 15, g = lane >> 4;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 32) {
    bf16x8 a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const uint4 va = *reinterpret_cast<const uint4*>(A + (int64_t)(m0 + 16 * i + r) * K + k0 + 8 * g);
      a[i] = __builtin_bit_cast(bf16x8, va);
      const uint4 vb = *reinterpret_cast<const uint4*>(B + (int64_t)(n0 + 16 * i + r) * K + k0 + 8 * g);
      b[i] = __builtin_bit_cast(bf16x8, vb);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
  }
  // lane holds C[m0 + 16 i + 4 g + e][n0 + 16 j + r], e = 0..3
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int n = n0 + 16 * j + r;
      const int m = m0 + 16 * i + 4 * g;
      const int64_t o = (int64_t)n * M + m;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float c = acc[i][j][e];
        if (EPI == EPI_HIDDEN) {
          const float z = c + bias[m + e];
          zout[o + e] = z;
          hout[o + e] = f2bf(z * sigm(z));
        } else if (EPI == EPI_OUT) {
          zout[o + e] = c + bias[m + e];
        } else {
          const float z = zprev[o + e];
          const float s = sigm(z);
          hout[o + e] = f2bf(c * s * (1.0f + z * (1.0f - s)));
        }
      }
    }
}

// dW[m][n] += sum_{s in chunk} G[s][m] H[s][n]; G [ns][M], H [ns][N] (bf16).
__global__ __launch_bounds__(256) void k_gemm_wgrad(const uint16_t* __restrict__ G, int M,
                                                     const uint16_t* __restrict__ H, int N,
                                                     int chunk, float* __restrict__ dW) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m0 = blockIdx.y * 64 + (wave >> 1) * 32;
  const int n0 = blockIdx.x * 64 + (wave & 1) * 32;
  const int s_begin = blockIdx.z * chunk;
  const int r = lane & 15, g = lane >> 4;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = s_begin; k0 < s_begin + chunk; k0 += 32) {
    bf16x8 a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      uint16_t ta[8], tb[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int64_t s = k0 + 8 * g + e;
        ta[e] = G[s * M + m0 + 16 * i + r];
        tb[e] = H[s * N + n0 + 16 * i + r];
      }
      a[i] = __builtin_bit_cast(bf16x8, ta);
      b[i] = __builtin_bit_cast(bf16x8, tb);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e)
        atomicAdd(dW + (int64_t)(m0 + 16 * i + 4 * g + e) * N + n0 + 16 * j + r, acc[i][j][e]);
}

// db[m] += sum_s G[s][m] over a chunk of samples
__global__ __launch_bounds__(256) void k_bias_grad(const uint16_t* __restrict__ G, int M, int chunk,
                                                    float* __restrict__ db) {
  const int m = blockIdx.x * 256 + threadIdx.x;
  if (m >= M) return;
  const int64_t s0 = (int64_t)blockIdx.y * chunk;
  float acc = 0.f;
  for (int s = 0; s < chunk; ++s) acc += bf2f(G[(s0 + s) * M + m]);
  atomicAdd(db + m, acc);
}

// Features of the raw inputs x = (rho, h, eps_r, s) [n][4] (lambda = 1):
// scaled inputs, R1, theta, exp(-2 R1), sin(theta), and the ground's own
// scales: the beat (Re k1 - k2) R1 damped by Im k1 (cos and sin parts) and
// the interface layer exp(-|k1| R1), with k1 = 2 pi sqrt(eps_r - j s);
// 52 zero pads.
MW_DEV void features(const float* x, float* f) {
  const float rho = x[0], h = x[1], er = x[2], s = x[3];
  const float r1 = sqrtf(rho * rho + h * h);
  const float th = atan2f(h, rho);
  f[0] = 0.5f * rho - 1.0f;
  f[1] = 0.5f * h - 1.0f;
  f[2] = (er - 17.5f) * 0.08f;
  f[3] = logf(fmaxf(s, 1e-30f) * 0.0745355992f) * 0.321810032f;  // log(s/sqrt(180))/log(sqrt(500))
  f[4] = r1 * 0.353553391f - 1.0f;
  f[5] = th * 1.27323954f - 1.0f;
  f[6] = expf(-2.0f * r1);
  f[7] = sinf(th);
  // k1 / (2 pi) = sqrt(er - j s), branch with Im <= 0
  const float mag = sqrtf(sqrtf(er * er + s * s));
  const float half = 0.5f * atan2f(-s, er);
  const float k1r = 6.28318531f * mag * cosf(half), k1i = 6.28318531f * mag * sinf(half);
  const float damp = expf(k1i * r1);
  const float beat = (k1r - 6.28318531f) * r1;
  f[8] = cosf(beat) * damp;
  f[9] = sinf(beat) * damp;
  f[10] = expf(-6.28318531f * mag * r1);
  f[11] = logf(mag) * 0.5f - 1.0f;
}

__global__ __launch_bounds__(256) void k_features(const float* __restrict__ x, int64_t n, int64_t n_pad,
                                                  uint16_t* __restrict__ H0) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n_pad) return;
  float f[NF_USED];
  if (i < n) {
    features(x + i * 4, f);
  } else {
#pragma unroll
    for (int t = 0; t < NF_USED; ++t) f[t] = 0.f;
  }
  uint16_t* o = H0 + i * FIN[0];
#pragma unroll
  for (int t = 0; t < NF_USED; ++t) o[t] = f2bf(f[t]);
  for (int t = NF_USED; t < FIN[0]; ++t) o[t] = 0;
}

// loss += sum (out - t)^2 / (n 8); G4 = 2 (out - t) / (n 8) for the used outputs
__global__ __launch_bounds__(256) void k_loss(const float* __restrict__ out, const float* __restrict__ t,
                                              int64_t n, int64_t n_pad, uint16_t* __restrict__ G4,
                                              float* __restrict__ loss) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  float part = 0.f;
  if (i < n_pad) {
    const float sc = 1.0f / (float)(n * NO_USED);
    for (int o = 0; o < FOUT[NL - 1]; ++o) {
      float gval = 0.f;
      if (i < n && o < NO_USED) {
        const float d = out[i * FOUT[NL - 1] + o] - t[i * NO_USED + o];
        part += d * d * sc;
        gval = 2.0f * d * sc;
      }
      G4[i * FOUT[NL - 1] + o] = f2bf(gval);
    }
  }
  // wave reduction, one atomic per wave
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) part += __shfl_xor(part, off, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(loss, part);
}

// Adam on the fp32 master, then the bf16 copies W and W^T
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ P, const float* __restrict__ grad,
                                              float* __restrict__ m1, float* __restrict__ m2, float lr,
                                              float b1, float b2, float eps, float c1, float c2,
                                              uint16_t* __restrict__ w16, uint16_t* __restrict__ wt16,
                                              int update) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= N_PARAMS) return;
  float p = P[i];
  if (update) {
    const float gr = grad[i];
    const float a = b1 * m1[i] + (1.f - b1) * gr;
    const float v = b2 * m2[i] + (1.f - b2) * gr * gr;
    m1[i] = a;
    m2[i] = v;
    p -= lr * (a * c1) / (sqrtf(v * c2) + eps);
    P[i] = p;
  }
  // which layer / entry
  int l = 0;
#pragma unroll
  for (int t = 1; t < NL; ++t)
    if (i >= param_off(t)) l = t;
  const int64_t loc = i - param_off(l);
  const int64_t nw = (int64_t)FOUT[l] * FIN[l];
  if (loc < nw) {
    const int o = (int)(loc / FIN[l]), in = (int)(loc % FIN[l]);
    const uint16_t b = f2bf(p);
    w16[w16_off(l) + loc] = b;
    wt16[w16_off(l) + (int64_t)in * FOUT[l] + o] = b;
  }
}

__global__ __launch_bounds__(256) void k_zero_f32(float* __restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = 0.f;
}

// The surrogate's surfaces written into a normalised master table: for node
// idx[i], vals[idx[i]][s] = (out[i][2s] + j out[i][2s+1]) * c1 k2^2.
__global__ __launch_bounds__(256) void k_to_surfaces(const float* __restrict__ out, const int32_t* __restrict__ idx,
                                                     int64_t n, double sre, double sim,
                                                     double2* __restrict__ vals) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float* o = out + i * FOUT[NL - 1];
  const int64_t node = idx[i];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const double a = o[2 * s], b = o[2 * s + 1];
    vals[node * 4 + s] = make_double2(a * sre - b * sim, a * sim + b * sre);
  }
}

// Raw inputs of grid nodes (lambda = 1): (R1 cos th, R1 sin th, Re eps, -Im eps)
__global__ __launch_bounds__(256) void k_node_inputs(const double* __restrict__ R1, const double* __restrict__ th,
                                                     const int32_t* __restrict__ idx, int64_t n, double er,
                                                     double s, float* __restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int64_t node = idx[i];
  double sn, cs;
  sincos(th[node], &sn, &cs);
  x[i * 4 + 0] = (float)fmax(R1[node] * cs, 0.0);
  x[i * 4 + 1] = (float)fmax(R1[node] * sn, 0.0);
  x[i * 4 + 2] = (float)er;
  x[i * 4 + 3] = (float)s;
}

struct Ws {
  uint16_t* H[NL];   // layer inputs: H[0] features, H[l] = SiLU(Z[l-1])
  float* Z[NL - 1];
  float* OUT;
  uint16_t* G[NL];   // dLoss/dZ_l (G[4] = dLoss/dOUT)
  float* loss;
  float* x;          // scratch raw inputs for grid nodes
};

int64_t ws_layout(int64_t n, char* base, Ws* w) {
  int64_t off = 0;
  auto take = [&](int64_t bytes) {
    char* p = base ? base + off : nullptr;
    off += (bytes + 255) & ~int64_t(255);
    return p;
  };
  for (int l = 0; l < NL; ++l) w->H[l] = (uint16_t*)take(n * FIN[l] * 2);
  for (int l = 0; l < NL - 1; ++l) w->Z[l] = (float*)take(n * FOUT[l] * 4);
  w->OUT = (float*)take(n * FOUT[NL - 1] * 4);
  for (int l = 0; l < NL; ++l) w->G[l] = (uint16_t*)take(n * FOUT[l] * 2);
  w->loss = (float*)take(256);
  w->x = (float*)take(n * 16);
  return off;
}

hipError_t forward(const mw_mlp& net, int64_t np, const Ws& w, hipStream_t st) {
  for (int l = 0; l < NL; ++l) {
    const uint16_t* W = net.w16 + w16_off(l);
    const float* b = net.params + param_off(l) + (int64_t)FOUT[l] * FIN[l];
    dim3 grid((unsigned)(np / 64), FOUT[l] / 64);
    if (l < NL - 1)
      hipLaunchKernelGGL(k_gemm_nt<EPI_HIDDEN>, grid, dim3(256), 0, st, W, FIN[l], w.H[l], FOUT[l],
                         b, (const float*)nullptr, w.Z[l], w.H[l + 1]);
    else
      hipLaunchKernelGGL(k_gemm_nt<EPI_OUT>, grid, dim3(256), 0, st, W, FIN[l], w.H[l], FOUT[l], b,
                         (const float*)nullptr, w.OUT, (uint16_t*)nullptr);
  }
  return hipGetLastError();
}

}  // namespace

int64_t mw_mlp_n_params() { return N_PARAMS; }
int64_t mw_mlp_n_w16() { return N_W16; }
int64_t mw_mlp_ws_bytes(int64_t n) {
  Ws w;
  return ws_layout(n, nullptr, &w);
}

hipError_t mw_mlp_launch_sync(const mw_mlp& net, hipStream_t st) {
  hipLaunchKernelGGL(k_adam, dim3((unsigned)((N_PARAMS + 255) / 256)), dim3(256), 0, st, net.params,
                     (const float*)nullptr, (float*)nullptr, (float*)nullptr, 0.f, 0.f, 0.f, 0.f, 0.f,
                     0.f, net.w16, net.wt16, 0);
  return hipGetLastError();
}

hipError_t mw_mlp_launch_forward(const mw_mlp& net, const float* x, int64_t n, float* out, void* ws,
                                 hipStream_t st) {
  const int64_t np = (n + 63) / 64 * 64;
  Ws w;
  ws_layout(np, (char*)ws, &w);
  hipLaunchKernelGGL(k_features, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, x, n, np, w.H[0]);
  hipError_t e = forward(net, np, w, st);
  if (e != hipSuccess) return e;
  if (out)
    e = hipMemcpy2DAsync(out, NO_USED * sizeof(float), w.OUT, FOUT[NL - 1] * sizeof(float),
                         NO_USED * sizeof(float), n, hipMemcpyDeviceToDevice, st);
  return e;
}

hipError_t mw_mlp_launch_backward(const mw_mlp& net, const float* x, const float* t, int64_t n,
                                  float* loss_out, void* ws, hipStream_t st) {
  const int64_t np = (n + 63) / 64 * 64;
  Ws w;
  ws_layout(np, (char*)ws, &w);
  hipLaunchKernelGGL(k_features, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, x, n, np, w.H[0]);
  hipError_t e = forward(net, np, w, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_zero_f32, dim3((unsigned)((N_PARAMS + 255) / 256)), dim3(256), 0, st,
                     net.grads, N_PARAMS);
  hipLaunchKernelGGL(k_zero_f32, dim3(1), dim3(256), 0, st, w.loss, (int64_t)1);
  hipLaunchKernelGGL(k_loss, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, w.OUT, t, n, np,
                     w.G[NL - 1], w.loss);
  const int chunk = np % 256 == 0 ? 256 : 64;
  for (int l = NL - 1; l >= 0; --l) {
    float* gW = net.grads + param_off(l);
    float* gb = gW + (int64_t)FOUT[l] * FIN[l];
    hipLaunchKernelGGL(k_gemm_wgrad, dim3(FIN[l] / 64, FOUT[l] / 64, (unsigned)(np / chunk)), dim3(256),
                       0, st, w.G[l], FOUT[l], w.H[l], FIN[l], chunk, gW);
    hipLaunchKernelGGL(k_bias_grad, dim3((FOUT[l] + 255) / 256, (unsigned)(np / chunk)), dim3(256), 0,
                       st, w.G[l], FOUT[l], chunk, gb);
    if (l > 0) {
      // G[l-1] = (W_l^T G_l) * SiLU'(Z[l-1])
      const uint16_t* Wt = net.wt16 + w16_off(l);
      hipLaunchKernelGGL(k_gemm_nt<EPI_BWD>, dim3((unsigned)(np / 64), FIN[l] / 64), dim3(256), 0, st,
                         Wt, FOUT[l], w.G[l], FIN[l], (const float*)nullptr, w.Z[l - 1],
                         (float*)nullptr, w.G[l - 1]);
    }
  }
  if (loss_out)
    e = hipMemcpyAsync(loss_out, w.loss, sizeof(float), hipMemcpyDeviceToDevice, st);
  if (e != hipSuccess) return e;
  return hipGetLastError();
}

hipError_t mw_mlp_launch_adam(const mw_mlp& net, float lr, float b1, float b2, float eps, int step,
                              hipStream_t st) {
  const float c1 = 1.0f / (1.0f - powf(b1, (float)step));
  const float c2 = 1.0f / (1.0f - powf(b2, (float)step));
  hipLaunchKernelGGL(k_adam, dim3((unsigned)((N_PARAMS + 255) / 256)), dim3(256), 0, st, net.params,
                     net.grads, net.adam_m, net.adam_v, lr, b1, b2, eps, c1, c2, net.w16, net.wt16, 1);
  return hipGetLastError();
}

hipError_t mw_mlp_launch_grid(const mw_mlp& net, double eps_re, double eps_im, const double* R1,
                              const double* th, const int32_t* idx, int64_t n, double2* vals,
                              double sre, double sim, void* ws, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const int64_t np = (n + 63) / 64 * 64;
  Ws w;
  ws_layout(np, (char*)ws, &w);
  hipLaunchKernelGGL(k_node_inputs, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, R1, th, idx, n,
                     eps_re, -eps_im, w.x);
  hipError_t e = mw_mlp_launch_forward(net, w.x, n, nullptr, ws, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_to_surfaces, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, w.OUT, idx, n,
                     sre, sim, vals);
  return hipGetLastError();
}
