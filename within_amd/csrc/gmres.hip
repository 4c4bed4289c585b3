// gmres.hip -- the device side of right-preconditioned GMRES(restart)
// (schwarz-precond/src/solve/gmres.rs:201-530) for a block of kp columns, each
// with its own Arnoldi process:
//   * orthogonalisation of w against the basis V_0..V_j by two passes of block
//     (classical) Gram-Schmidt: one launch forms all j+1 dot products of a
//     pass, one launch subtracts them (the reference's modified Gram-Schmidt,
//     gmres.rs:201-220, gives the same Arnoldi relation; the second pass makes
//     classical Gram-Schmidt as stable);
//   * the Hessenberg column, the Givens rotations, the residual estimate and
//     the per-column stop decision in FP64 on the device (gmres.rs:223-260),
//     one thread per column, so an Arnoldi step needs no host round trip;
//   * the triangular solve and x += Z y (gmres.rs:499-530) on the device.
// Every sum has a fixed order.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "engine.hpp"
#include "gmres.hpp"
#include "kernels.hpp"

namespace wa {

constexpr int kGmBlock = 256;

// Lanes own NV consecutive columns of a row (16-byte accesses when the row is
// at least 16 bytes); V = kp / NV lanes per row, kGmBlock / V rows per block step.
constexpr int kGmGroup = 8;  // basis vectors per block of the dot-product grid

// part[(b * nv + i) * kp + c] = sum over block b's rows of W[r, c] * V_i[r, c]
// for the vectors i of this block's group (blockIdx.y), V_i = V + i * vstride
template <typename T, int NV>
__global__ void __launch_bounds__(kGmBlock) k_multi_dot(const T* __restrict__ W, const T* __restrict__ V,
                                                        int64_t vstride, int nv, int64_t m, int kp,
                                                        double* __restrict__ part) {
  __shared__ double sh[kGmBlock * NV];
  const int L = kp / NV, lane = threadIdx.x % L, g = threadIdx.x / L, G = kGmBlock / L;
  const int i0 = blockIdx.y * kGmGroup;
  double acc[kGmGroup][NV];
#pragma unroll
  for (int u = 0; u < kGmGroup; ++u)
#pragma unroll
    for (int t = 0; t < NV; ++t) acc[u][t] = 0.0;
  for (int64_t r = (int64_t)blockIdx.x * G + g; r < m; r += (int64_t)gridDim.x * G) {
    const int64_t e = r * kp + lane * NV;
    const VecT<T, NV> w = vload<T, NV>(W + e);
#pragma unroll
    for (int u = 0; u < kGmGroup; ++u)
      if (i0 + u < nv) {
        const VecT<T, NV> v = vload<T, NV>(V + (int64_t)(i0 + u) * vstride + e);
#pragma unroll
        for (int t = 0; t < NV; ++t) acc[u][t] += (double)w.v[t] * (double)v.v[t];
      }
  }
#pragma unroll
  for (int u = 0; u < kGmGroup; ++u) {
    if (i0 + u < nv) {  // uniform across the block
#pragma unroll
      for (int t = 0; t < NV; ++t) sh[threadIdx.x * NV + t] = acc[u][t];
      __syncthreads();
      for (int s = G / 2; s >= 1; s >>= 1) {
        if (g < s)
#pragma unroll
          for (int t = 0; t < NV; ++t) sh[threadIdx.x * NV + t] += sh[(threadIdx.x + s * L) * NV + t];
        __syncthreads();
      }
      if (g == 0)
#pragma unroll
        for (int t = 0; t < NV; ++t)
          part[((int64_t)blockIdx.x * nv + i0 + u) * kp + lane * NV + t] = sh[threadIdx.x * NV + t];
      __syncthreads();
    }
  }
}

// out[i * kp + c] (+)= sum over blocks of part, in block order
__global__ void k_multi_dot_total(const double* __restrict__ part, int nb, int nv, int kp, double* __restrict__ out,
                                  int accumulate) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nv * kp) return;
  const int i = j / kp, c = j % kp;
  double s = 0.0;
  for (int b = 0; b < nb; ++b) s += part[((int64_t)b * nv + i) * kp + c];
  out[j] = accumulate ? out[j] + s : s;
}

// W[r, c] += sign * sum_{i < nv, i < lim[c]} coef[i * kp + c] * V_i[r, c]
// (lim == nullptr: all nv), coefficients staged in LDS; with `part`, also the
// per-block sums of the new W^2
template <typename T, int NV>
__global__ void __launch_bounds__(kGmBlock) k_multi_axpy(T* __restrict__ W, const T* __restrict__ V, int64_t vstride,
                                                         int nv, const double* __restrict__ coef,
                                                         const int* __restrict__ lim, double sign, int64_t m, int kp,
                                                         double* __restrict__ part) {
  __shared__ double cf[kGmresMaxV * 32];
  __shared__ double sh[kGmBlock * NV];
  for (int j = threadIdx.x; j < nv * kp; j += kGmBlock) {
    const int i = j / kp, c = j % kp;
    cf[j] = (lim && i >= lim[c]) ? 0.0 : sign * coef[j];
  }
  __syncthreads();
  const int L = kp / NV, lane = threadIdx.x % L, g = threadIdx.x / L, G = kGmBlock / L;
  double acc2[NV];
#pragma unroll
  for (int t = 0; t < NV; ++t) acc2[t] = 0.0;
  for (int64_t r = (int64_t)blockIdx.x * G + g; r < m; r += (int64_t)gridDim.x * G) {
    const int64_t e = r * kp + lane * NV;
    const VecT<T, NV> w0 = vload<T, NV>(W + e);
    double w[NV];
#pragma unroll
    for (int t = 0; t < NV; ++t) w[t] = (double)w0.v[t];
#pragma unroll 4
    for (int i = 0; i < nv; ++i) {
      const VecT<T, NV> v = vload<T, NV>(V + (int64_t)i * vstride + e);
#pragma unroll
      for (int t = 0; t < NV; ++t) w[t] += cf[i * kp + lane * NV + t] * (double)v.v[t];
    }
    VecT<T, NV> o;
#pragma unroll
    for (int t = 0; t < NV; ++t) {
      o.v[t] = (T)w[t];
      acc2[t] += (double)o.v[t] * (double)o.v[t];
    }
    vstore<T, NV>(W + e, o);
  }
  if (!part) return;
#pragma unroll
  for (int t = 0; t < NV; ++t) sh[threadIdx.x * NV + t] = acc2[t];
  __syncthreads();
  for (int s = G / 2; s >= 1; s >>= 1) {
    if (g < s)
#pragma unroll
      for (int t = 0; t < NV; ++t) sh[threadIdx.x * NV + t] += sh[(threadIdx.x + s * L) * NV + t];
    __syncthreads();
  }
  if (g == 0)
#pragma unroll
    for (int t = 0; t < NV; ++t) part[(int64_t)blockIdx.x * kp + lane * NV + t] = sh[threadIdx.x * NV + t];
}

// column state reset at the start of a cycle: g = beta e_1, H / rotations zero
__global__ void k_gm_cycle_start(GmCol* __restrict__ s, double* __restrict__ H, const double* __restrict__ beta,
                                 const int* __restrict__ live, double* __restrict__ coef, int mr, int kp) {
  const int c = threadIdx.x;
  if (c >= kp) return;
  GmCol& q = s[c];
  double* h = H + (int64_t)c * mr * (mr + 1);
  for (int i = 0; i < mr * (mr + 1); ++i) h[i] = 0.0;
  for (int i = 0; i < kGmresMaxV; ++i) {
    q.cs[i] = 0.0;
    q.sn[i] = 0.0;
    q.g[i] = 0.0;
  }
  q.g[0] = live[c] ? beta[c] : 0.0;
  q.hmax = 0.0;
  q.inc = live[c];
  q.jc = 0;
  q.outcome = 3;  // restart unless decided otherwise
  q.resid = 0.0;
  coef[c] = live[c] ? 1.0 / beta[c] : 0.0;
}

// Hessenberg column j of every extending column: h_ij from the two
// Gram-Schmidt passes, h_{j+1,j} = ||w||; previous rotations, the new one
// (hypot), g update, residual estimate |g_{j+1}|, stop decision
// (gmres.rs:223-260); coef = 1 / h_{j+1,j} for the next basis vector
__global__ void k_gm_givens(GmCol* __restrict__ s, double* __restrict__ H, const double* __restrict__ h,
                            const double* __restrict__ h2, const double* __restrict__ part2, int nb, const double* __restrict__ abs_tol,
                            double* __restrict__ coef, int* __restrict__ n_inc, int j, int mr, int64_t total,
                            int64_t maxiter, int kp) {
  const int c = threadIdx.x;
  if (c < kp) {
    GmCol& q = s[c];
    coef[c] = 0.0;
    if (q.inc) {
      double w2 = 0.0;
      for (int b = 0; b < nb; ++b) w2 += part2[(int64_t)b * kp + c];
      double* hc = H + ((int64_t)c * mr + j) * (mr + 1);
      double lmax = 0.0;
      for (int i = 0; i <= j; ++i) {
        hc[i] = h[i * kp + c] + h2[i * kp + c];  // both Gram-Schmidt passes
        lmax = fmax(lmax, fabs(hc[i]));
      }
      const double hj1 = sqrt(w2);
      hc[j + 1] = hj1;
      q.hmax = fmax(fmax(q.hmax, lmax), hj1);
      for (int i = 0; i < j; ++i) {
        const double a = hc[i], bb = hc[i + 1];
        const double t = q.cs[i] * a + q.sn[i] * bb;
        hc[i + 1] = -q.sn[i] * a + q.cs[i] * bb;
        hc[i] = t;
      }
      const double a = hc[j], bb = hc[j + 1], rr = hypot(a, bb);
      const double cc = rr == 0.0 ? 1.0 : a / rr, ss = rr == 0.0 ? 0.0 : bb / rr;
      q.cs[j] = cc;
      q.sn[j] = ss;
      hc[j] = cc * a + ss * bb;
      hc[j + 1] = 0.0;
      const double t = cc * q.g[j] + ss * q.g[j + 1];
      q.g[j + 1] = -ss * q.g[j] + cc * q.g[j + 1];
      q.g[j] = t;
      q.jc = j + 1;
      q.resid = fabs(q.g[j + 1]);
      if (q.resid <= abs_tol[c]) {
        q.outcome = 0;
        q.inc = 0;
      } else if (total >= maxiter) {
        q.outcome = 1;
        q.inc = 0;
      } else if (hj1 > 1e-14 * q.hmax) {
        coef[c] = 1.0 / hj1;
      } else {
        q.outcome = 2;
        q.inc = 0;
      }
    }
  }
  __syncthreads();
  if (c == 0) {
    int a = 0;
    for (int i = 0; i < kp; ++i) a += s[i].inc;
    *n_inc = a;
  }
}

// y = R^-1 g per column (gmres.rs:499-519): y[i * kp + c], jc[c] entries
__global__ void k_gm_backsolve(const GmCol* __restrict__ s, const double* __restrict__ H, const int* __restrict__ live,
                               double* __restrict__ y, int* __restrict__ lim, int mr, int kp) {
  const int c = threadIdx.x;
  if (c >= kp) return;
  const GmCol& q = s[c];
  const int jc = live[c] ? q.jc : 0;
  lim[c] = jc;
  auto h = [&](int i, int jj) { return H[((int64_t)c * mr + jj) * (mr + 1) + i]; };
  for (int i = jc - 1; i >= 0; --i) {
    double sum = q.g[i];
    for (int k = i + 1; k < jc; ++k) sum -= h(i, k) * y[k * kp + c];
    const double dg = h(i, i);
    y[i * kp + c] = fabs(dg) < 1e-14 * (1.0 + fabs(sum)) ? 0.0 : sum / dg;
  }
}

template <typename T>
static inline int gm_nv(int kp) {
  const int full = 16 / (int)sizeof(T);
  return kp >= full ? full : kp;
}
static inline unsigned gm_blocks(int64_t m, int kp, int nvec) {
  const int64_t rows_per = kGmBlock / (kp / nvec);
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((m + rows_per - 1) / rows_per, kGmRedBlocks));
}

#define GM_VEC_SWITCH(T, kp, CALL)          \
  do {                                      \
    const int nvec_ = gm_nv<T>(kp);         \
    if (nvec_ == 16 / (int)sizeof(T)) {     \
      constexpr int NV = 16 / sizeof(T);    \
      CALL;                                 \
    } else if (nvec_ == 2) {                \
      constexpr int NV = 2;                 \
      CALL;                                 \
    } else {                                \
      constexpr int NV = 1;                 \
      CALL;                                 \
    }                                       \
  } while (0)

template <typename T>
void launch_multi_dot(hipStream_t st, const T* W, const T* V, int64_t vstride, int nv, int64_t m, int kp,
                      double* part, double* out, int accumulate) {
  const unsigned nb = gm_blocks(m, kp, gm_nv<T>(kp));
  const dim3 grid(nb, (unsigned)((nv + kGmGroup - 1) / kGmGroup));
  GM_VEC_SWITCH(T, kp, (k_multi_dot<T, NV><<<grid, kGmBlock, 0, st>>>(W, V, vstride, nv, m, kp, part)));
  k_multi_dot_total<<<(unsigned)((nv * kp + 255) / 256), 256, 0, st>>>(part, (int)nb, nv, kp, out, accumulate);
}
template <typename T>
int launch_multi_axpy(hipStream_t st, T* W, const T* V, int64_t vstride, int nv, const double* coef, const int* lim,
                      double sign, int64_t m, int kp, double* part) {
  const unsigned nb = gm_blocks(m, kp, gm_nv<T>(kp));
  GM_VEC_SWITCH(T, kp, (k_multi_axpy<T, NV><<<nb, kGmBlock, 0, st>>>(W, V, vstride, nv, coef, lim, sign, m, kp, part)));
  return (int)nb;
}
void launch_gm_cycle_start(hipStream_t st, GmCol* s, double* H, const double* beta, const int* live, double* coef,
                           int mr, int kp) {
  k_gm_cycle_start<<<1, 64, 0, st>>>(s, H, beta, live, coef, mr, kp);
}
void launch_gm_givens(hipStream_t st, GmCol* s, double* H, const double* h, const double* h2, const double* part2, int nb,
                      const double* abs_tol, double* coef, int* n_inc, int j, int mr, int64_t total, int64_t maxiter,
                      int kp) {
  k_gm_givens<<<1, 64, 0, st>>>(s, H, h, h2, part2, nb, abs_tol, coef, n_inc, j, mr, total, maxiter, kp);
}
void launch_gm_backsolve(hipStream_t st, const GmCol* s, const double* H, const int* live, double* y, int* lim, int mr,
                         int kp) {
  k_gm_backsolve<<<1, 64, 0, st>>>(s, H, live, y, lim, mr, kp);
}

template void launch_multi_dot<float>(hipStream_t, const float*, const float*, int64_t, int, int64_t, int, double*,
                                      double*, int);
template void launch_multi_dot<double>(hipStream_t, const double*, const double*, int64_t, int, int64_t, int, double*,
                                       double*, int);
template int launch_multi_axpy<float>(hipStream_t, float*, const float*, int64_t, int, const double*, const int*,
                                      double, int64_t, int, double*);
template int launch_multi_axpy<double>(hipStream_t, double*, const double*, int64_t, int, const double*, const int*,
                                       double, int64_t, int, double*);

}  // namespace wa
