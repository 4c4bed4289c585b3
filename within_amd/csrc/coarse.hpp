// coarse.hpp -- the learned coarse-space correction: a graph network over the
// factor-level graph (the Gramian's sparsity) whose output columns span a
// coarse space P; the coarse correction enters one-level additive Schwarz as
// a balancing two-level preconditioner inside flexible CG.  The network's
// dense layers run on the matrix cores in BF16 (v_mfma_f32_32x32x8bf16_1k)
// for training and inference; its aggregation is the solver's own FP32
// operator apply.  No counterpart in the reference (it extends the CG loop of
// schwarz-precond/src/solve/cg.rs:45-133 and sits beside FePreconditioner,
// crates/within/src/operator/preconditioner.rs:47-154).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.hpp"

namespace wa {

constexpr int kCoarse = 32;      // coarse dimension = network width = input features
constexpr int kGnnLayers = 3;    // two hidden ReLU layers, one linear output layer
constexpr int kMat = kCoarse * kCoarse;
// fp32 parameters of layer l at l * kLayerParams: Ws (row k, column j), Wa, b
constexpr int kLayerParams = 2 * kMat + kCoarse;
constexpr int kGnnParams = kGnnLayers * kLayerParams;
// BF16 operands of layer l at l * kLayerPack: Ws, Wa, Ws^T, Wa^T
constexpr int kLayerPack = 4 * kMat;
using bf16_t = uint16_t;         // raw bfloat16 bits (round to nearest even)

// out = post(A0 W0 + (s1 o A1) W1 + bias); with W1 == nullptr the second term
// is added as it is: + s1 o A1.  post: ReLU when relu, then zero where
// mask <= 0 (mask == nullptr: none).  out2 = s2 o out (either may be null).
// Every matrix is m x 32 FP32 row-major; W0 / W1 are 32 x 32 BF16.
void launch_gnn_layer(hipStream_t st, int64_t m, const float* A0, const bf16_t* W0, const float* A1,
                      const float* s1, const bf16_t* W1, const float* bias, const float* mask, int relu,
                      float* out, float* out2, const float* s2);
// out[a * 32 + b] = sum_i (sx o X)[i, a] Y[i, b] (BF16 operands, FP32
// accumulation on the matrix cores); out_cs[b] = sum_i Y[i, b] (FP32) when
// out_cs is not null.  part: kXtyBlocksMax * (kMat + kCoarse) floats; mid:
// kCprojMid * kMat doubles (second-stage partials).
constexpr int kXtyBlocksMax = 512;
constexpr int kCprojMid = 32;  // second-stage partials of every two-stage reduction
void launch_gnn_xty(hipStream_t st, int64_t m, const float* X, const float* sx, const float* Y, float* part,
                    double* mid, float* out, float* out_cs);
// out[a * 32 + b] = sum_i X[i, a] Y[i, b] in FP64 (FP32 inputs).
// part: kXtyBlocksMax * kMat doubles; mid as above.
void launch_xty64(hipStream_t st, int64_t m, const float* X, const float* Y, double* part, double* mid,
                  double* out);
// out = A0 M0 + A1 M1, FP64 accumulation (M0, M1: 32 x 32 FP64)
void launch_tall_small(hipStream_t st, int64_t m, const float* A0, const double* M0, const float* A1,
                       const double* M1, float* out);
// The coarse solve of the training loss (one workgroup, FP64):
// V = (Ac + eps tr(Ac)/32 I)^-1 U, loss = -<U, V> / nu,
// W1 = -(2/nu) V^T, W2 = (2/nu) V V^T, so that dL/dP = R W1 + GP W2.
void launch_coarse_small(hipStream_t st, const double* Ac, const double* U, double nu, double eps, double* V,
                         double* loss, double* W1, double* W2);
// Adam step over the kGnnParams parameters (bias corrections c1, c2 from the
// host), then the BF16 operand pack of every layer
void launch_gnn_adam(hipStream_t st, float* theta, const float* grad, float* m1, float* m2, double lr, double c1,
                     double c2);
void launch_gnn_pack(hipStream_t st, const float* theta, bf16_t* pack);
// v = s o u (m x 32)
void launch_rowscale(hipStream_t st, int64_t m, const float* s, const float* u, float* v);
// dst = bf16(src) (count values); back != nullptr: back = float(dst)
void launch_to_bf16(hipStream_t st, int64_t count, const float* src, bf16_t* dst, float* back);
// N(0, 1) values from a counter-based generator (splitmix64, Box-Muller)
void launch_randn(hipStream_t st, int64_t count, uint64_t seed, float* out);
// c = a - b (count values)
void launch_sub(hipStream_t st, int64_t count, const float* a, const float* b, float* c);

// ---- the coarse correction inside the Krylov loop (T = float or double) ----
// part[blk] = X^T R over the block's rows (X: m x 32 BF16, R: m x kp), FP64;
// returns the block count
constexpr int kCprojBlocksMax = 1024;
template <typename T>
int launch_cproj_part(hipStream_t st, int64_t m, const bf16_t* X, const T* R, int kp, double* part);
// out = base + sign * Ainv u with u the sum of the nb partials (32 x kp),
// reduced in two fixed-order stages through mid (kCprojMid * 32 * kp
// doubles); base may be null
void launch_cproj_fin(hipStream_t st, const double* part, int nb, int kp, double* mid, const double* Ainv,
                      const double* base, double sign, double* out);
// out = base + sign * X coef (X: m x 32 BF16, coef: 32 x kp FP64); out may alias base
template <typename T>
void launch_cexpand(hipStream_t st, int64_t m, const bf16_t* X, const double* coef, double sign, const T* base,
                    T* out, int kp);
// flexible CG: partials of A.B and A.C in one pass (part: nb x 2 x kp), and
// beta = -<z, Ap> / <p, Ap> with the new <r, z> (Notay's FCG(1), which is
// the Polak-Ribiere form <z, r - r_old> / <z_old, r_old> of CG's beta)
template <typename T>
int launch_dot2_part(hipStream_t st, const T* A, const T* B, const T* C, int64_t m, int kp, double* part);
void launch_fcg_beta(hipStream_t st, ColState* s, const double* part, int nb, const double* pap, double* rznew,
                     double* beta, int* mask, int* n_active, int kp);

}  // namespace wa
