// coarse.hip -- kernels of the learned coarse-space correction (coarse.hpp).
//
// Network kernels work on m x 32 FP32 row-major blocks (one 128-byte row per
// DOF).  The dense layers and the weight gradients run on the matrix cores:
// v_mfma_f32_32x32x8bf16_1k, with FP32 values rounded to BF16 (nearest even)
// as they are loaded, and FP32 accumulation.  Operand layouts of the 32x32x8
// instruction (wave64, l = lane):
//   A (32 x 8): lane l holds A[l % 32][4 (l / 32) + e], e = 0..3;
//   B (8 x 32): lane l holds B[4 (l / 32) + e][l % 32];
//   D (32 x 32): lane l, value v holds D[8 (v / 4) + 4 (l / 32) + v % 4][l % 32]
// (the same D layout as v_mfma_f32_32x32x2f32).  Every reduction has a fixed
// order for a given size, so results are reproducible.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "coarse.hpp"
#include "engine.hpp"

namespace wa {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v16f __attribute__((ext_vector_type(16)));

constexpr int kGnnBlock = 256;  // four waves
constexpr int kXtyOut = kMat + kCoarse;

__device__ __forceinline__ bf16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (bf16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint32_t h) { return __uint_as_float((h & 0xFFFFu) << 16); }

static unsigned clamp_blocks(int64_t want, int64_t cap) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(want, cap));
}

// ---------------------------------------------------------------------------
// dense layer: one wave per 32-row tile
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kGnnBlock) k_gnn_layer(int64_t m, const float* __restrict__ A0,
                                                         const bf16_t* __restrict__ W0, const float* __restrict__ A1,
                                                         const float* __restrict__ s1, const bf16_t* __restrict__ W1,
                                                         const float* __restrict__ bias,
                                                         const float* __restrict__ mask, int relu,
                                                         float* __restrict__ out, float* __restrict__ out2,
                                                         const float* __restrict__ s2) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, hk = lane >> 5, col = lane & 31;
  v4s b0[4], b1[4];
#pragma unroll
  for (int kb = 0; kb < 4; ++kb)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int k = 8 * kb + 4 * hk + e;
      b0[kb][e] = (short)W0[k * kCoarse + col];
      b1[kb][e] = W1 ? (short)W1[k * kCoarse + col] : (short)0;
    }
  const float bj = bias ? bias[col] : 0.0f;
  const bool mm1 = A1 && W1;
  const int64_t tiles = (m + 31) / 32;
  for (int64_t t = (int64_t)blockIdx.x * 4 + wv; t < tiles; t += (int64_t)gridDim.x * 4) {
    const int64_t r0 = t * 32, ra = r0 + col;
    v16f acc = {};
    {
      VecT<float, 4> x[4];
#pragma unroll
      for (int kb = 0; kb < 4; ++kb) {
        if (ra < m) {
          x[kb] = vload<float, 4>(A0 + ra * kCoarse + 8 * kb + 4 * hk);
        } else {
#pragma unroll
          for (int e = 0; e < 4; ++e) x[kb].v[e] = 0.0f;
        }
      }
#pragma unroll
      for (int kb = 0; kb < 4; ++kb) {
        v4s a;
#pragma unroll
        for (int e = 0; e < 4; ++e) a[e] = (short)f2bf(x[kb].v[e]);
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b0[kb], acc, 0, 0, 0);
      }
    }
    if (mm1) {
      const float s = (ra < m && s1) ? s1[ra] : 1.0f;
      VecT<float, 4> x[4];
#pragma unroll
      for (int kb = 0; kb < 4; ++kb) {
        if (ra < m) {
          x[kb] = vload<float, 4>(A1 + ra * kCoarse + 8 * kb + 4 * hk);
        } else {
#pragma unroll
          for (int e = 0; e < 4; ++e) x[kb].v[e] = 0.0f;
        }
      }
#pragma unroll
      for (int kb = 0; kb < 4; ++kb) {
        v4s a;
#pragma unroll
        for (int e = 0; e < 4; ++e) a[e] = (short)f2bf(s * x[kb].v[e]);
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b1[kb], acc, 0, 0, 0);
      }
    }
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int64_t row = r0 + 8 * (v >> 2) + 4 * hk + (v & 3);
      if (row >= m) continue;
      const int64_t id = row * kCoarse + col;
      float val = acc[v] + bj;
      if (A1 && !W1) val += (s1 ? s1[row] : 1.0f) * A1[id];
      if (relu) val = val > 0.0f ? val : 0.0f;
      if (mask && !(mask[id] > 0.0f)) val = 0.0f;
      if (out) out[id] = val;
      if (out2) out2[id] = s2[row] * val;
    }
  }
}

void launch_gnn_layer(hipStream_t st, int64_t m, const float* A0, const bf16_t* W0, const float* A1,
                      const float* s1, const bf16_t* W1, const float* bias, const float* mask, int relu,
                      float* out, float* out2, const float* s2) {
  if (m <= 0) return;
  const int64_t tiles = (m + 31) / 32;
  k_gnn_layer<<<clamp_blocks((tiles + 3) / 4, 16384), kGnnBlock, 0, st>>>(m, A0, W0, A1, s1, W1, bias, mask, relu,
                                                                         out, out2, s2);
}

// ---------------------------------------------------------------------------
// X^T Y on the matrix cores: each wave takes 8-row groups (the k dimension)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kGnnBlock) k_gnn_xty(int64_t m, const float* __restrict__ X,
                                                       const float* __restrict__ sx, const float* __restrict__ Y,
                                                       float* __restrict__ part) {
  __shared__ float sh[4][kXtyOut];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, hk = lane >> 5, col = lane & 31;
  v16f acc = {};
  float cs = 0.0f;
  const int64_t groups = (m + 7) / 8, W = (int64_t)gridDim.x * 4;
  for (int64_t g = (int64_t)blockIdx.x * 4 + wv; g < groups; g += W) {
    float x[4], y[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int64_t row = g * 8 + 4 * hk + e;
      x[e] = 0.0f;
      y[e] = 0.0f;
      if (row < m) {
        x[e] = X[row * kCoarse + col];
        if (sx) x[e] *= sx[row];
        y[e] = Y[row * kCoarse + col];
      }
    }
    v4s a, b;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      a[e] = (short)f2bf(x[e]);
      b[e] = (short)f2bf(y[e]);
      cs += y[e];
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
  }
#pragma unroll
  for (int v = 0; v < 16; ++v) sh[wv][(8 * (v >> 2) + 4 * hk + (v & 3)) * kCoarse + col] = acc[v];
  const float other = __shfl_xor(cs, 32);
  if (hk == 0) sh[wv][kMat + col] = cs + other;
  __syncthreads();
  for (int i = threadIdx.x; i < kXtyOut; i += kGnnBlock)
    part[(int64_t)blockIdx.x * kXtyOut + i] = ((sh[0][i] + sh[1][i]) + sh[2][i]) + sh[3][i];
}

// Column sums of an nb x n partial matrix in two fixed-order stages: block j
// of the first sums rows [j c, (j + 1) c) (c = ceil(nb / kCprojMid)) of
// every column into mid; the second sums the kCprojMid rows of mid.
template <typename T>
__global__ void __launch_bounds__(1024) k_part_mid(const T* __restrict__ part, int nb, int n, T* __restrict__ mid) {
  const int chunk = (nb + kCprojMid - 1) / kCprojMid;
  const int b0 = blockIdx.x * chunk, b1 = min(nb, b0 + chunk);
  for (int i = threadIdx.x; i < n; i += 1024) {
    T s = 0;
    for (int b = b0; b < b1; ++b) s += part[(int64_t)b * n + i];
    mid[(int64_t)blockIdx.x * n + i] = s;
  }
}

__global__ void __launch_bounds__(1024) k_gnn_xty_fin(const float* __restrict__ mid, float* __restrict__ out,
                                                      float* __restrict__ out_cs) {
  for (int i = threadIdx.x; i < kXtyOut; i += 1024) {
    float s = 0.0f;
#pragma unroll 8
    for (int b = 0; b < kCprojMid; ++b) s += mid[(int64_t)b * kXtyOut + i];
    if (i < kMat) out[i] = s;
    else if (out_cs) out_cs[i - kMat] = s;
  }
}

void launch_gnn_xty(hipStream_t st, int64_t m, const float* X, const float* sx, const float* Y, float* part,
                    double* mid, float* out, float* out_cs) {
  const int64_t groups = (m + 7) / 8;
  const unsigned nb = clamp_blocks((groups + 63) / 64, kXtyBlocksMax);
  k_gnn_xty<<<nb, kGnnBlock, 0, st>>>(m, X, sx, Y, part);
  k_part_mid<float><<<kCprojMid, 1024, 0, st>>>(part, (int)nb, kXtyOut, reinterpret_cast<float*>(mid));
  k_gnn_xty_fin<<<1, 1024, 0, st>>>(reinterpret_cast<const float*>(mid), out, out_cs);
}

// ---------------------------------------------------------------------------
// X^T Y in FP64 (the coarse matrix and the projected residuals of the loss)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kGnnBlock) k_xty64(int64_t m, const float* __restrict__ X,
                                                     const float* __restrict__ Y, double* __restrict__ part) {
  constexpr int TR = 64;
  __shared__ float xs[TR][kCoarse + 1];
  __shared__ float ys[TR][kCoarse];
  const int a = threadIdx.x >> 3, b0 = (threadIdx.x & 7) * 4;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  const int64_t tiles = (m + TR - 1) / TR;
  for (int64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
    const int64_t r0 = t * TR;
    for (int i = threadIdx.x; i < TR * kCoarse; i += kGnnBlock) {
      const int r = i / kCoarse, c = i % kCoarse;
      const int64_t row = r0 + r;
      xs[r][c] = row < m ? X[row * kCoarse + c] : 0.0f;
      ys[r][c] = row < m ? Y[row * kCoarse + c] : 0.0f;
    }
    __syncthreads();
    for (int r = 0; r < TR; ++r) {
      const double x = xs[r][a];
#pragma unroll
      for (int c = 0; c < 4; ++c) acc[c] += x * (double)ys[r][b0 + c];
    }
    __syncthreads();
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) part[(int64_t)blockIdx.x * kMat + a * kCoarse + b0 + c] = acc[c];
}

__global__ void __launch_bounds__(1024) k_xty64_fin(const double* __restrict__ mid, double* __restrict__ out) {
  const int i = threadIdx.x;
  double s = 0.0;
#pragma unroll 8
  for (int b = 0; b < kCprojMid; ++b) s += mid[(int64_t)b * kMat + i];
  out[i] = s;
}

void launch_xty64(hipStream_t st, int64_t m, const float* X, const float* Y, double* part, double* mid,
                  double* out) {
  const int64_t tiles = (m + 63) / 64;
  const unsigned nb = clamp_blocks((tiles + 3) / 4, kXtyBlocksMax);
  k_xty64<<<nb, kGnnBlock, 0, st>>>(m, X, Y, part);
  k_part_mid<double><<<kCprojMid, 1024, 0, st>>>(part, (int)nb, kMat, mid);
  k_xty64_fin<<<1, kMat, 0, st>>>(mid, out);
}

// ---------------------------------------------------------------------------
// out = A0 M0 + A1 M1, accumulated in FP64 (FP32 inputs, FP64 M): the two
// terms of dL/dP cancel where P'GP is ill-conditioned.  8 threads per row,
// 4 outputs each.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kGnnBlock) k_tall_small(int64_t m, const float* __restrict__ A0,
                                                          const double* __restrict__ M0, const float* __restrict__ A1,
                                                          const double* __restrict__ M1, float* __restrict__ out) {
  __shared__ double m0[kMat], m1[kMat];
  for (int i = threadIdx.x; i < kMat; i += kGnnBlock) {
    m0[i] = M0[i];
    m1[i] = A1 ? M1[i] : 0.0;
  }
  __syncthreads();
  const int rl = threadIdx.x >> 3, j0 = (threadIdx.x & 7) * 4;
  for (int64_t r0 = (int64_t)blockIdx.x * 32; r0 < m; r0 += (int64_t)gridDim.x * 32) {
    const int64_t row = r0 + rl;
    if (row >= m) continue;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k4 = 0; k4 < 8; ++k4) {
      const VecT<float, 4> x = vload<float, 4>(A0 + row * kCoarse + 4 * k4);
#pragma unroll
      for (int kk = 0; kk < 4; ++kk)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[c] += (double)x.v[kk] * m0[(4 * k4 + kk) * kCoarse + j0 + c];
    }
    if (A1) {
#pragma unroll
      for (int k4 = 0; k4 < 8; ++k4) {
        const VecT<float, 4> x = vload<float, 4>(A1 + row * kCoarse + 4 * k4);
#pragma unroll
        for (int kk = 0; kk < 4; ++kk)
#pragma unroll
          for (int c = 0; c < 4; ++c) acc[c] += (double)x.v[kk] * m1[(4 * k4 + kk) * kCoarse + j0 + c];
      }
    }
    float o[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) o[c] = (float)acc[c];
    vstore<float, 4>(out + row * kCoarse + j0, o);
  }
}

void launch_tall_small(hipStream_t st, int64_t m, const float* A0, const double* M0, const float* A1,
                       const double* M1, float* out) {
  if (m <= 0) return;
  k_tall_small<<<clamp_blocks((m + 31) / 32, 16384), kGnnBlock, 0, st>>>(m, A0, M0, A1, M1, out);
}

// ---------------------------------------------------------------------------
// the loss's 32 x 32 solve (one workgroup, FP64)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(64) k_coarse_small(const double* __restrict__ Ac, const double* __restrict__ U,
                                                     double nu, double eps, double* __restrict__ V,
                                                     double* __restrict__ loss, double* __restrict__ W1,
                                                     double* __restrict__ W2) {
  constexpr int K = kCoarse;
  __shared__ double L[kMat], Vs[kMat];
  const int t = threadIdx.x;
  for (int i = t; i < kMat; i += 64) L[i] = Ac[i];
  __syncthreads();
  __shared__ double reg_s;
  if (t == 0) {
    double tr = 0.0;
    for (int i = 0; i < K; ++i) tr += L[i * (K + 1)];
    reg_s = eps * (tr > 0.0 ? tr : 0.0) / K + 1e-300;
  }
  __syncthreads();
  const double reg = reg_s;
  if (t < K) L[t * (K + 1)] += reg;
  __syncthreads();
  // lower Cholesky factor in place, column by column: the pivot, then the
  // entries below it in parallel (each sum in ascending k)
  for (int j = 0; j < K; ++j) {
    if (t == 0) {
      double d = L[j * (K + 1)];
      for (int k = 0; k < j; ++k) d -= L[j * K + k] * L[j * K + k];
      L[j * (K + 1)] = d > reg ? sqrt(d) : sqrt(reg);
    }
    __syncthreads();
    if (t > j && t < K) {
      double s = L[t * K + j];
      for (int k = 0; k < j; ++k) s -= L[t * K + k] * L[j * K + k];
      L[t * K + j] = s / L[j * (K + 1)];
    }
    __syncthreads();
  }
  if (t < K) {  // column t of U: L L^T v = u (forward, then backward substitution)
    for (int i = 0; i < K; ++i) {
      double s = U[i * K + t];
      for (int k = 0; k < i; ++k) s -= L[i * K + k] * Vs[k * K + t];
      Vs[i * K + t] = s / L[i * (K + 1)];
    }
    for (int i = K - 1; i >= 0; --i) {
      double s = Vs[i * K + t];
      for (int k = i + 1; k < K; ++k) s -= L[k * K + i] * Vs[k * K + t];
      Vs[i * K + t] = s / L[i * (K + 1)];
    }
  }
  __syncthreads();
  for (int i = t; i < kMat; i += 64) V[i] = Vs[i];
  if (t == 0) {
    double l = 0.0;
    for (int i = 0; i < kMat; ++i) l += U[i] * Vs[i];
    *loss = -l / nu;
  }
  const double f = 2.0 / nu;
  for (int i = t; i < kMat; i += 64) {
    const int r = i / K, c = i % K;
    W1[i] = -f * Vs[c * K + r];  // W1[s][a] = -(2/nu) V[a][s]
    double s = 0.0;
    for (int q = 0; q < K; ++q) s += Vs[r * K + q] * Vs[c * K + q];
    W2[i] = f * s;               // W2[b][a] = (2/nu) (V V^T)[b][a]
  }
}

void launch_coarse_small(hipStream_t st, const double* Ac, const double* U, double nu, double eps, double* V,
                         double* loss, double* W1, double* W2) {
  k_coarse_small<<<1, 64, 0, st>>>(Ac, U, nu, eps, V, loss, W1, W2);
}

// ---------------------------------------------------------------------------
// parameters
// ---------------------------------------------------------------------------
__global__ void k_gnn_adam(float* __restrict__ theta, const float* __restrict__ g, float* __restrict__ m1,
                           float* __restrict__ m2, double lr, double c1, double c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kGnnParams) return;
  const float gi = g[i];
  const float a = 0.9f * m1[i] + 0.1f * gi;
  const float b = 0.999f * m2[i] + 0.001f * gi * gi;
  m1[i] = a;
  m2[i] = b;
  const double mh = (double)a / c1, vh = (double)b / c2;
  theta[i] = (float)((double)theta[i] - lr * mh / (sqrt(vh) + 1e-8));
}

__global__ void k_gnn_pack(const float* __restrict__ theta, bf16_t* __restrict__ pack) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kGnnLayers * kLayerPack) return;
  const int l = i / kLayerPack, r = i % kLayerPack, w = r / kMat, e = r % kMat, k = e / kCoarse, j = e % kCoarse;
  const float* Ws = theta + (int64_t)l * kLayerParams;
  const float* Wa = Ws + kMat;
  const float v = w == 0 ? Ws[e] : w == 1 ? Wa[e] : w == 2 ? Ws[j * kCoarse + k] : Wa[j * kCoarse + k];
  pack[i] = f2bf(v);
}

void launch_gnn_adam(hipStream_t st, float* theta, const float* grad, float* m1, float* m2, double lr, double c1,
                     double c2) {
  k_gnn_adam<<<(kGnnParams + 255) / 256, 256, 0, st>>>(theta, grad, m1, m2, lr, c1, c2);
}
void launch_gnn_pack(hipStream_t st, const float* theta, bf16_t* pack) {
  k_gnn_pack<<<(kGnnLayers * kLayerPack + 255) / 256, 256, 0, st>>>(theta, pack);
}

// ---------------------------------------------------------------------------
// elementwise helpers
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kGnnBlock) k_rowscale(int64_t m, const float* __restrict__ s,
                                                        const float* __restrict__ u, float* __restrict__ v) {
  const int64_t tot = m * kCoarse / 4;
  for (int64_t i = (int64_t)blockIdx.x * kGnnBlock + threadIdx.x; i < tot; i += (int64_t)gridDim.x * kGnnBlock) {
    VecT<float, 4> x = vload<float, 4>(u + 4 * i);
    const float f = s[(4 * i) / kCoarse];
#pragma unroll
    for (int c = 0; c < 4; ++c) x.v[c] *= f;
    vstore<float, 4>(v + 4 * i, x);
  }
}
void launch_rowscale(hipStream_t st, int64_t m, const float* s, const float* u, float* v) {
  if (m <= 0) return;
  k_rowscale<<<clamp_blocks((m * kCoarse / 4 + kGnnBlock - 1) / kGnnBlock, 65536), kGnnBlock, 0, st>>>(m, s, u, v);
}

__global__ void k_to_bf16(int64_t count, const float* __restrict__ src, bf16_t* __restrict__ dst,
                          float* __restrict__ back) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
    const bf16_t h = f2bf(src[i]);
    dst[i] = h;
    if (back) back[i] = bf2f(h);
  }
}
void launch_to_bf16(hipStream_t st, int64_t count, const float* src, bf16_t* dst, float* back) {
  if (count <= 0) return;
  k_to_bf16<<<clamp_blocks((count + 255) / 256, 65536), 256, 0, st>>>(count, src, dst, back);
}

__device__ inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
__global__ void k_randn(int64_t count, uint64_t seed, float* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t h1 = splitmix64(seed ^ splitmix64(2 * (uint64_t)i));
    const uint64_t h2 = splitmix64(seed ^ splitmix64(2 * (uint64_t)i + 1));
    const double u1 = ((double)(h1 >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    const double u2 = ((double)(h2 >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    out[i] = (float)(sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
  }
}
void launch_randn(hipStream_t st, int64_t count, uint64_t seed, float* out) {
  if (count <= 0) return;
  k_randn<<<clamp_blocks((count + 255) / 256, 65536), 256, 0, st>>>(count, seed, out);
}

__global__ void k_sub(int64_t count, const float* __restrict__ a, const float* __restrict__ b,
                      float* __restrict__ c) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
    c[i] = a[i] - b[i];
}
void launch_sub(hipStream_t st, int64_t count, const float* a, const float* b, float* c) {
  if (count <= 0) return;
  k_sub<<<clamp_blocks((count + 255) / 256, 65536), 256, 0, st>>>(count, a, b, c);
}

// ---------------------------------------------------------------------------
// the coarse correction in the Krylov loop
// ---------------------------------------------------------------------------
// X^T R per block: E = 32 G threads own (coarse index a, NV columns from c0),
// the other 256 / E thread rows split the tile's rows; FP64 sums.  Tiles of
// TR rows are staged through LDS with 16-byte loads (X: 8 BF16 values each).
template <typename T, int NV>
__global__ void __launch_bounds__(kGnnBlock) k_cproj_part(int64_t m, const bf16_t* __restrict__ X,
                                                          const T* __restrict__ R, int kp,
                                                          double* __restrict__ part) {
  constexpr int TR = 512 / sizeof(T), VN = 16 / sizeof(T);
  __shared__ float xs[TR][kCoarse + 1];
  __shared__ __align__(16) T rs[TR * 32];
  __shared__ double red[kGnnBlock * NV];
  const int G = kp / NV, E = kCoarse * G, RL = kGnnBlock / E;
  const int e = threadIdx.x % E, rl = threadIdx.x / E, a = e / G, c0 = (e % G) * NV;
  double acc[NV];
#pragma unroll
  for (int c = 0; c < NV; ++c) acc[c] = 0.0;
  const int64_t tiles = (m + TR - 1) / TR, tot = m * kp;
  for (int64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
    const int64_t r0 = t * TR;
    for (int i = threadIdx.x; i < TR * 4; i += kGnnBlock) {
      const int r = i >> 2, q = i & 3;
      const int64_t row = r0 + r;
      VecT<uint32_t, 4> w;
      if (row < m) {
        w = vload<uint32_t, 4>(reinterpret_cast<const uint32_t*>(X + row * kCoarse) + 4 * q);
      } else {
#pragma unroll
        for (int h = 0; h < 4; ++h) w.v[h] = 0u;
      }
#pragma unroll
      for (int h = 0; h < 8; ++h) xs[r][8 * q + h] = bf2f(h & 1 ? (w.v[h >> 1] >> 16) : w.v[h >> 1]);
    }
    const int64_t base = r0 * kp;
    for (int i = threadIdx.x * VN; i < TR * kp; i += kGnnBlock * VN) {
      if (base + i + VN <= tot) {
        *reinterpret_cast<VecT<T, VN>*>(&rs[i]) = vload<T, VN>(R + base + i);
      } else {
#pragma unroll
        for (int j = 0; j < VN; ++j) rs[i + j] = base + i + j < tot ? R[base + i + j] : (T)0;
      }
    }
    __syncthreads();
    for (int r = rl; r < TR; r += RL) {
      const double x = xs[r][a];
#pragma unroll
      for (int c = 0; c < NV; ++c) acc[c] += x * (double)rs[r * kp + c0 + c];
    }
    __syncthreads();
  }
#pragma unroll
  for (int c = 0; c < NV; ++c) red[threadIdx.x * NV + c] = acc[c];
  __syncthreads();
  if (threadIdx.x < E) {
#pragma unroll
    for (int c = 0; c < NV; ++c) {
      double s = 0.0;
      for (int q = 0; q < RL; ++q) s += red[(q * E + threadIdx.x) * NV + c];
      part[(int64_t)blockIdx.x * kCoarse * kp + a * kp + c0 + c] = s;
    }
  }
}

template <typename T>
int launch_cproj_part(hipStream_t st, int64_t m, const bf16_t* X, const T* R, int kp, double* part) {
  const int64_t tiles = (m + 512 / sizeof(T) - 1) / (512 / sizeof(T));
  const unsigned nb = clamp_blocks((tiles + 1) / 2, kCprojBlocksMax);
  const int nv = kp >= 8 ? kp / 8 : 1;
  if (nv == 4) k_cproj_part<T, 4><<<nb, kGnnBlock, 0, st>>>(m, X, R, kp, part);
  else if (nv == 2) k_cproj_part<T, 2><<<nb, kGnnBlock, 0, st>>>(m, X, R, kp, part);
  else k_cproj_part<T, 1><<<nb, kGnnBlock, 0, st>>>(m, X, R, kp, part);
  return (int)nb;
}
template int launch_cproj_part<float>(hipStream_t, int64_t, const bf16_t*, const float*, int, double*);
template int launch_cproj_part<double>(hipStream_t, int64_t, const bf16_t*, const double*, int, double*);

// second stage: block j sums the partials [j c, (j + 1) c) of every entry
// (c = ceil(nb / kCprojMid)), in order
__global__ void __launch_bounds__(1024) k_cproj_mid(const double* __restrict__ part, int nb, int n,
                                                    double* __restrict__ mid) {
  const int chunk = (nb + kCprojMid - 1) / kCprojMid;
  const int b0 = blockIdx.x * chunk, b1 = min(nb, b0 + chunk);
  for (int i = threadIdx.x; i < n; i += 1024) {
    double s = 0.0;
    for (int b = b0; b < b1; ++b) s += part[(int64_t)b * n + i];
    mid[(int64_t)blockIdx.x * n + i] = s;
  }
}

// u = sum of the kCprojMid second-stage partials (in order); out = base + sign Ainv u
__global__ void __launch_bounds__(1024) k_cproj_fin(const double* __restrict__ mid, int kp,
                                                    const double* __restrict__ Ainv, const double* base, double sign,
                                                    double* out) {
  __shared__ double u[kCoarse * 32];
  const int n = kCoarse * kp;
  for (int i = threadIdx.x; i < n; i += 1024) {
    double s = 0.0;
#pragma unroll 8
    for (int b = 0; b < kCprojMid; ++b) s += mid[(int64_t)b * n + i];
    u[i] = s;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += 1024) {
    const int a = i / kp, c = i % kp;
    double s = 0.0;
    for (int b = 0; b < kCoarse; ++b) s += Ainv[a * kCoarse + b] * u[b * kp + c];
    out[i] = (base ? base[i] : 0.0) + sign * s;
  }
}

void launch_cproj_fin(hipStream_t st, const double* part, int nb, int kp, double* mid, const double* Ainv,
                      const double* base, double sign, double* out) {
  k_cproj_mid<<<kCprojMid, 1024, 0, st>>>(part, nb, kCoarse * kp, mid);
  k_cproj_fin<<<1, 1024, 0, st>>>(mid, kp, Ainv, base, sign, out);
}

template <typename T, int NV>
__global__ void __launch_bounds__(kGnnBlock) k_cexpand(int64_t m, const bf16_t* __restrict__ X,
                                                       const double* __restrict__ coef, double sign, const T* base,
                                                       T* out, int kp) {
  __shared__ T cf[kCoarse * 32];
  for (int i = threadIdx.x; i < kCoarse * kp; i += kGnnBlock) cf[i] = (T)(sign * coef[i]);
  __syncthreads();
  const int V = kp / NV, gpb = kGnnBlock / V, pt = threadIdx.x % V;
  for (int64_t i = (int64_t)blockIdx.x * gpb + threadIdx.x / V; i < m; i += (int64_t)gridDim.x * gpb) {
    const VecT<T, NV> b = vload<T, NV>(base + i * kp + pt * NV);
    T acc[NV];
#pragma unroll
    for (int c = 0; c < NV; ++c) acc[c] = b.v[c];
    const uint32_t* xr = reinterpret_cast<const uint32_t*>(X + i * kCoarse);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const VecT<uint32_t, 4> w = vload<uint32_t, 4>(xr + 4 * q);
#pragma unroll
      for (int h = 0; h < 8; ++h) {
        const int a = 8 * q + h;
        const T xv = (T)bf2f(h & 1 ? (w.v[h >> 1] >> 16) : w.v[h >> 1]);
#pragma unroll
        for (int c = 0; c < NV; ++c) acc[c] += xv * cf[a * kp + pt * NV + c];
      }
    }
    vstore<T, NV>(out + i * kp + pt * NV, acc);
  }
}

template <typename T>
void launch_cexpand(hipStream_t st, int64_t m, const bf16_t* X, const double* coef, double sign, const T* base,
                    T* out, int kp) {
  if (m <= 0) return;
  constexpr int NVM = 16 / sizeof(T);
  const int nv = std::min(NVM, kp);
  const int64_t rows_per_block = kGnnBlock / (kp / nv);
  const unsigned nb = clamp_blocks((m + rows_per_block - 1) / rows_per_block, 65536);
  if (nv == 4) k_cexpand<T, 4><<<nb, kGnnBlock, 0, st>>>(m, X, coef, sign, base, out, kp);
  else if (nv == 2) k_cexpand<T, 2><<<nb, kGnnBlock, 0, st>>>(m, X, coef, sign, base, out, kp);
  else k_cexpand<T, 1><<<nb, kGnnBlock, 0, st>>>(m, X, coef, sign, base, out, kp);
}
template void launch_cexpand<float>(hipStream_t, int64_t, const bf16_t*, const double*, double, const float*, float*,
                                    int);
template void launch_cexpand<double>(hipStream_t, int64_t, const bf16_t*, const double*, double, const double*,
                                     double*, int);

// ---------------------------------------------------------------------------
// flexible CG
// ---------------------------------------------------------------------------
template <typename T, int NV>
__global__ void __launch_bounds__(kBlock) k_dot2_part(const T* __restrict__ A, const T* __restrict__ B,
                                                      const T* __restrict__ C, int64_t m, int kp,
                                                      double* __restrict__ part) {
  __shared__ double sh[kBlock * NV];
  const int V = kp / NV, pt = threadIdx.x % V, gpb = kBlock / V;
  double ab[NV], ac[NV];
#pragma unroll
  for (int c = 0; c < NV; ++c) ab[c] = ac[c] = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * gpb + threadIdx.x / V; i < m; i += (int64_t)gridDim.x * gpb) {
    const int64_t id = i * kp + pt * NV;
    const VecT<T, NV> a = vload<T, NV>(A + id), b = vload<T, NV>(B + id), c = vload<T, NV>(C + id);
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      ab[q] += (double)a.v[q] * (double)b.v[q];
      ac[q] += (double)a.v[q] * (double)c.v[q];
    }
  }
  for (int pass = 0; pass < 2; ++pass) {
#pragma unroll
    for (int q = 0; q < NV; ++q) sh[threadIdx.x * NV + q] = pass == 0 ? ab[q] : ac[q];
    __syncthreads();
    if (threadIdx.x < kp) {
      double s = 0.0;
      for (int g = 0; g < gpb; ++g) s += sh[g * kp + threadIdx.x];
      part[((int64_t)blockIdx.x * 2 + pass) * kp + threadIdx.x] = s;
    }
    __syncthreads();
  }
}

template <typename T>
int launch_dot2_part(hipStream_t st, const T* A, const T* B, const T* C, int64_t m, int kp, double* part) {
  const int64_t want = (m * kp + kBlock - 1) / kBlock;
  const unsigned nb = clamp_blocks(want, kRedBlocks);
  constexpr int NVM = 16 / sizeof(T);
  const int nv = std::min(NVM, kp);
  if (nv == 4) k_dot2_part<T, 4><<<nb, kBlock, 0, st>>>(A, B, C, m, kp, part);
  else if (nv == 2) k_dot2_part<T, 2><<<nb, kBlock, 0, st>>>(A, B, C, m, kp, part);
  else k_dot2_part<T, 1><<<nb, kBlock, 0, st>>>(A, B, C, m, kp, part);
  return (int)nb;
}
template int launch_dot2_part<float>(hipStream_t, const float*, const float*, const float*, int64_t, int, double*);
template int launch_dot2_part<double>(hipStream_t, const double*, const double*, const double*, int64_t, int,
                                      double*);

__global__ void __launch_bounds__(1024) k_fcg_beta(ColState* __restrict__ s, const double* __restrict__ part, int nb,
                                                   const double* __restrict__ pap, double* __restrict__ rznew,
                                                   double* __restrict__ beta, int* __restrict__ mask,
                                                   int* __restrict__ n_active, int kp) {
  __shared__ double sh[1024];
  __shared__ double res[64];
  const int kp2 = 2 * kp, col = threadIdx.x % kp2, str = 1024 / kp2, s0 = threadIdx.x / kp2;
  double acc = 0.0;
  for (int b = s0; b < nb; b += str) acc += part[(int64_t)b * kp2 + col];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = str / 2; w >= 1; w >>= 1) {
    if (s0 < w) sh[threadIdx.x] += sh[threadIdx.x + w * kp2];
    __syncthreads();
  }
  if (s0 == 0) res[col] = sh[threadIdx.x];
  __syncthreads();
  const int c = threadIdx.x;
  if (c < kp) {
    const double v = res[c], zap = res[kp + c];
    rznew[c] = v;
    ColState& q = s[c];
    beta[c] = 0.0;
    mask[c] = 0;
    if (q.active) {
      const double thr = 4.930380657631324e-32 * fmax(fabs(q.rz0), 2.2250738585072014e-308);
      if (fabs(v) < thr) {
        q.active = 0;
        q.converged = q.rnorm / q.bnorm <= q.tol;
      } else {
        beta[c] = -zap / pap[c];
        q.beta = beta[c];
        q.rz = v;
        mask[c] = 1;
      }
    }
  }
  __syncthreads();
  if (c == 0) {
    int a = 0;
    for (int j = 0; j < kp; ++j) a += s[j].active;
    *n_active = a;
  }
}

void launch_fcg_beta(hipStream_t st, ColState* s, const double* part, int nb, const double* pap, double* rznew,
                     double* beta, int* mask, int* n_active, int kp) {
  k_fcg_beta<<<1, 1024, 0, st>>>(s, part, nb, pap, rznew, beta, mask, n_active, kp);
}

}  // namespace wa
