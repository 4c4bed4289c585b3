// kernels.hip -- gfx950 kernels of the projection path.
//
// Layout: every block vector holds k right-hand sides as rows of kp values
// (kp = k rounded up to a power of two, <= 32): DOF blocks are m x kp and
// observation blocks n x kp, row-major, so one DOF or observation row is one
// contiguous gather of kp * sizeof(T) bytes.  A "group" of kp lanes owns one
// work item and each lane one column: gathers are coalesced across the group
// and no float atomics are used anywhere -- every sum has a fixed order, so
// repeated runs are bitwise identical.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <functional>

#include "engine.hpp"
#include "kernels.hpp"

namespace wa {

#define GROUP_SETUP(kp)                                              \
  const int lane = threadIdx.x & ((kp) - 1);                         \
  const int64_t grp = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / (kp);

static inline unsigned blocks_for(int64_t groups, int kp) {
  int64_t threads = groups * kp;
  return (unsigned)((threads + kBlock - 1) / kBlock);
}
static inline unsigned grid_stride_blocks(int64_t n, int kp) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(blocks_for(n, kp), 65536));
}

// ===========================================================================
// observation space (FP64): D'W V and V = Y - D X, row-major n x kp
// (rmatvec_wdt / matvec_d, domain.rs:390-418)
// ===========================================================================
// ===========================================================================
// layout conversion: column-major (ld) <-> row-major kp blocks, via LDS tiles
// ===========================================================================
// dst row i <- src row (map ? map[i] : i)
template <typename S, typename D>
__global__ void __launch_bounds__(kBlock) k_col_to_row(const S* __restrict__ src, int64_t ld, int64_t n,
                                                       int k, D* __restrict__ dst, int kp,
                                                       const uint32_t* __restrict__ map) {
  __shared__ double tile[32][65];
  const int64_t r0 = (int64_t)blockIdx.x * 64;
  for (int t = threadIdx.x; t < 64 * 32; t += kBlock) {
    int r = t & 63, c = t >> 6;
    int64_t i = r0 + r;
    double v = 0.0;
    if (c < k && i < n) v = (double)src[(map ? (int64_t)map[i] : i) + c * ld];
    tile[c][r] = v;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < 64 * kp; t += kBlock) {
    int c = t % kp, r = t / kp;
    int64_t i = r0 + r;
    if (i < n) dst[i * kp + c] = (D)tile[c][r];
  }
}

// dst row i <- src row (map ? map[i] : i); a map entry of UINT32_MAX gives zeros
template <typename S, typename D>
__global__ void __launch_bounds__(kBlock) k_row_to_col(const S* __restrict__ src, int kp, int64_t n, int k,
                                                       D* __restrict__ dst, int64_t ld,
                                                       const uint32_t* __restrict__ map) {
  __shared__ double tile[32][65];
  const int64_t r0 = (int64_t)blockIdx.x * 64;
  for (int t = threadIdx.x; t < 64 * kp; t += kBlock) {
    int c = t % kp, r = t / kp;
    int64_t i = r0 + r;
    double v = 0.0;
    if (i < n) {
      const uint32_t j = map ? map[i] : (uint32_t)i;
      if (j != UINT32_MAX) v = (double)src[(int64_t)j * kp + c];
    }
    tile[c][r] = v;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < 64 * 32; t += kBlock) {
    int r = t & 63, c = t >> 6;
    int64_t i = r0 + r;
    if (c < k && i < n) dst[i + c * ld] = (D)tile[c][r];
  }
}

void launch_col_to_row(hipStream_t st, const double* src, int64_t ld, int64_t n, int k, double* dst, int kp,
                       const uint32_t* map) {
  if (n > 0)
    k_col_to_row<double, double><<<(unsigned)((n + 63) / 64), kBlock, 0, st>>>(src, ld, n, k, dst, kp, map);
}
void launch_row_to_col(hipStream_t st, const double* src, int kp, int64_t n, int k, double* dst, int64_t ld,
                       const uint32_t* map) {
  if (n > 0)
    k_row_to_col<double, double><<<(unsigned)((n + 63) / 64), kBlock, 0, st>>>(src, kp, n, k, dst, ld, map);
}

// ===========================================================================
// block-vector ops with per-column scalars; every reduction is two-stage with
// a fixed grid (deterministic) and accumulates in FP64 (FP32 products are
// exact in FP64, so FP32 dots carry no accumulation error beyond FP64's)
// ===========================================================================
static inline unsigned red_blocks(int64_t m, int kp) {
  int64_t b = (m * kp + kBlock - 1) / kBlock;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(b, kRedBlocks));
}

template <int MODE, typename T>
__device__ inline double red_term(const T* A, const T* B, int64_t idx) {
  if (MODE == 0) return (double)A[idx] * (double)B[idx];
  return (double)A[idx] * (double)A[idx];
}

// Lanes own NV consecutive values of a kp-wide row (16-byte vectors when the
// row is at least 16 bytes); V = kp / NV lanes per row.
template <typename T>
static inline int vnv(int kp) {
  const int full = 16 / (int)sizeof(T);
  return kp >= full ? full : kp;
}
#define WA_VEC_SWITCH(T, kp, CALL)       \
  do {                                   \
    const int nv_ = vnv<T>(kp);          \
    if (nv_ == 16 / (int)sizeof(T)) {    \
      constexpr int NV = 16 / sizeof(T); \
      CALL;                              \
    } else if (nv_ == 2) {               \
      constexpr int NV = 2;              \
      CALL;                              \
    } else {                             \
      constexpr int NV = 1;              \
      CALL;                              \
    }                                    \
  } while (0)

// E[i] = Y[i] - sum_q X[off_q + code_q(i)]   (Y == nullptr: E = D X)
template <int NV>
__global__ void __launch_bounds__(kBlock) k_demean(int64_t n, int Q, const uint32_t* __restrict__ codes,
                                                   const uint32_t* __restrict__ offs,
                                                   const double* __restrict__ X,
                                                   const double* __restrict__ Y, double* __restrict__ E,
                                                   int kp) {
  const int V = kp / NV, pt = threadIdx.x % V;
  const int64_t gpb = kBlock / V;
  for (int64_t i = (int64_t)blockIdx.x * gpb + threadIdx.x / V; i < n; i += (int64_t)gridDim.x * gpb) {
    double s[NV];
#pragma unroll
    for (int c = 0; c < NV; ++c) s[c] = 0.0;
    for (int q = 0; q < Q; ++q) {
      const VecT<double, NV> x =
          vload<double, NV>(X + (int64_t)(offs[q] + codes[(int64_t)q * n + i]) * kp + pt * NV);
#pragma unroll
      for (int c = 0; c < NV; ++c) s[c] += x.v[c];
    }
    VecT<double, NV> e;
    if (Y) {
      const VecT<double, NV> y = vload<double, NV>(Y + i * kp + pt * NV);
#pragma unroll
      for (int c = 0; c < NV; ++c) e.v[c] = y.v[c] - s[c];
    } else {
#pragma unroll
      for (int c = 0; c < NV; ++c) e.v[c] = s[c];
    }
    vstore<double, NV>(E + i * kp + pt * NV, e);
  }
}

void launch_demean(hipStream_t st, int64_t n, int Q, const uint32_t* codes, const uint32_t* offs,
                   const double* X, const double* Y, double* E, int kp) {
  if (n <= 0) return;
  WA_VEC_SWITCH(double, kp, (k_demean<NV><<<grid_stride_blocks(n, kp / NV), kBlock, 0, st>>>(n, Q, codes, offs,
                                                                                               X, Y, E, kp)));
}

// per-block column sums of a thread's NV accumulators (sh: kBlock * NV)
template <int NV>
__device__ inline void block_col_sums(double (&acc)[NV], double* sh, int kp, int gpb, double* out) {
#pragma unroll
  for (int c = 0; c < NV; ++c) sh[threadIdx.x * NV + c] = acc[c];
  __syncthreads();
  if (threadIdx.x < kp) {
    double s = 0.0;
    for (int g = 0; g < gpb; ++g) s += sh[g * kp + threadIdx.x];
    out[(int64_t)blockIdx.x * kp + threadIdx.x] = s;
  }
}

// part[b * kp + c] = sum over this block's rows of A[:,c] * B[:,c]
template <typename T, int NV>
__global__ void __launch_bounds__(kBlock) k_dot_part(const T* __restrict__ A, const T* __restrict__ B,
                                                     int64_t m, int kp, double* __restrict__ part) {
  __shared__ double sh[kBlock * NV];
  const int V = kp / NV, pt = threadIdx.x % V, gpb = kBlock / V;
  double acc[NV];
#pragma unroll
  for (int c = 0; c < NV; ++c) acc[c] = 0.0;
#pragma unroll 2
  for (int64_t i = (int64_t)blockIdx.x * gpb + threadIdx.x / V; i < m; i += (int64_t)gridDim.x * gpb) {
    const VecT<T, NV> a = vload<T, NV>(A + i * kp + pt * NV), b = vload<T, NV>(B + i * kp + pt * NV);
#pragma unroll
    for (int c = 0; c < NV; ++c) acc[c] += (double)a.v[c] * (double)b.v[c];
  }
  block_col_sums<NV>(acc, sh, kp, gpb, part);
}

// out[c] = sum_b part[b * kp + c]: threads stride over the blocks, then a
// fixed-order tree over the workgroup
__global__ void __launch_bounds__(kBlock) k_dot_fin(const double* __restrict__ part, int nb, int kp,
                                                    double* __restrict__ out) {
  __shared__ double sh[kBlock];
  const int col = threadIdx.x % kp, str = kBlock / kp, s0 = threadIdx.x / kp;
  double acc = 0.0;
  for (int b = s0; b < nb; b += str) acc += part[(int64_t)b * kp + col];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = str / 2; w >= 1; w >>= 1) {
    if (s0 < w) sh[threadIdx.x] += sh[threadIdx.x + w * kp];
    __syncthreads();
  }
  if (s0 == 0) out[col] = sh[threadIdx.x];
}

// D[i + c ldd] = Y[i + c ldy] - sum_q X[off_q + code_q(i)] for c < k: the
// demeaning written straight into column-major output (same arithmetic as
// k_demean).  One lane per row, so the column-major reads and writes are
// coalesced along rows; the row's sums are formed 8 columns at a time.
__global__ void __launch_bounds__(kBlock) k_demean_cols(int64_t n, int Q, const uint32_t* __restrict__ codes,
                                                        const uint32_t* __restrict__ offs,
                                                        const double* __restrict__ X, const double* __restrict__ Y,
                                                        int64_t ldy, double* __restrict__ D, int64_t ldd, int k,
                                                        int kp) {
  constexpr int CB = 8, QMAX = 8;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    int64_t row[QMAX];
    for (int q = 0; q < Q; ++q) row[q] = (int64_t)(offs[q] + codes[(int64_t)q * n + i]) * kp;
    for (int c0 = 0; c0 < k; c0 += CB) {
      double s[CB];
#pragma unroll
      for (int c = 0; c < CB; ++c) s[c] = 0.0;
      const int w = min(CB, kp - c0);
      for (int q = 0; q < Q; ++q) {
        if (w == CB) {
#pragma unroll
          for (int h = 0; h < CB / 2; ++h) {
            const VecT<double, 2> x = vload<double, 2>(X + row[q] + c0 + 2 * h);
            s[2 * h] += x.v[0];
            s[2 * h + 1] += x.v[1];
          }
        } else {
          for (int c = 0; c < w; ++c) s[c] += X[row[q] + c0 + c];
        }
      }
#pragma unroll
      for (int c = 0; c < CB; ++c)
        if (c0 + c < k) D[i + (int64_t)(c0 + c) * ldd] = Y[i + (int64_t)(c0 + c) * ldy] - s[c];
    }
  }
}

void launch_demean_cols(hipStream_t st, int64_t n, int Q, const uint32_t* codes, const uint32_t* offs,
                        const double* X, const double* Y, int64_t ldy, double* D, int64_t ldd, int k, int kp) {
  if (n <= 0) return;
  k_demean_cols<<<grid_stride_blocks(n, 1), kBlock, 0, st>>>(n, Q, codes, offs, X, Y, ldy, D, ldd, k, kp);
}

template <typename T>
int launch_dot_part(hipStream_t st, const T* A, const T* B, int64_t m, int kp, double* part) {
  unsigned nb = red_blocks(m, kp);
  WA_VEC_SWITCH(T, kp, (k_dot_part<T, NV><<<nb, kBlock, 0, st>>>(A, B, m, kp, part)));
  return (int)nb;
}
template int launch_dot_part<float>(hipStream_t, const float*, const float*, int64_t, int, double*);
template int launch_dot_part<double>(hipStream_t, const double*, const double*, int64_t, int, double*);

template <typename T>
void launch_dot(hipStream_t st, const T* A, const T* B, int64_t m, int kp, double* part, double* out) {
  const int nb = launch_dot_part<T>(st, A, B, m, kp, part);
  k_dot_fin<<<1, kBlock, 0, st>>>(part, nb, kp, out);
}
template void launch_dot<float>(hipStream_t, const float*, const float*, int64_t, int, double*, double*);
template void launch_dot<double>(hipStream_t, const double*, const double*, int64_t, int, double*, double*);

// X += alpha P ; R -= alpha AP ; part = sum R^2   (alpha per column, FP64)
template <typename T, int NV, bool UPDX>
__global__ void __launch_bounds__(kBlock) k_cg_update(T* __restrict__ X, const T* __restrict__ P,
                                                      T* __restrict__ R, const T* __restrict__ AP,
                                                      const double* __restrict__ alpha, int64_t m, int kp,
                                                      double* __restrict__ part) {
  __shared__ double sh[kBlock * NV];
  const int V = kp / NV, pt = threadIdx.x % V, gpb = kBlock / V;
  T a[NV];
  double acc[NV];
#pragma unroll
  for (int c = 0; c < NV; ++c) {
    a[c] = (T)alpha[pt * NV + c];
    acc[c] = 0.0;
  }
#pragma unroll 2
  for (int64_t i = (int64_t)blockIdx.x * gpb + threadIdx.x / V; i < m; i += (int64_t)gridDim.x * gpb) {
    const int64_t id = i * kp + pt * NV;
    VecT<T, NV> r = vload<T, NV>(R + id);
    const VecT<T, NV> ap = vload<T, NV>(AP + id);
    if constexpr (UPDX) {
      VecT<T, NV> x = vload<T, NV>(X + id);
      const VecT<T, NV> p = vload<T, NV>(P + id);
#pragma unroll
      for (int c = 0; c < NV; ++c) x.v[c] += a[c] * p.v[c];
      vstore_nt<T, NV>(X + id, x);
    }
#pragma unroll
    for (int c = 0; c < NV; ++c) {
      r.v[c] = r.v[c] - a[c] * ap.v[c];
      acc[c] += (double)r.v[c] * (double)r.v[c];
    }
    vstore_nt<T, NV>(R + id, r);
  }
  block_col_sums<NV>(acc, sh, kp, gpb, part);
}

// X = nullptr: only R (and the partials); X is then updated by launch_cg_x
template <typename T>
int launch_cg_update(hipStream_t st, T* X, const T* P, T* R, const T* AP, const double* alpha, int64_t m, int kp,
                     double* part) {
  unsigned nb = red_blocks(m, kp);
  if (X)
    WA_VEC_SWITCH(T, kp, (k_cg_update<T, NV, true><<<nb, kBlock, 0, st>>>(X, P, R, AP, alpha, m, kp, part)));
  else
    WA_VEC_SWITCH(T, kp, (k_cg_update<T, NV, false><<<nb, kBlock, 0, st>>>(X, P, R, AP, alpha, m, kp, part)));
  return (int)nb;
}

// X += alpha P (alpha per column; the same arithmetic as k_cg_update's X)
template <typename T, int NV>
__global__ void __launch_bounds__(kBlock) k_cg_x(T* __restrict__ X, const T* __restrict__ P,
                                                 const double* __restrict__ alpha, int64_t m, int kp) {
  const int V = kp / NV, pt = threadIdx.x % V, gpb = kBlock / V;
  T a[NV];
#pragma unroll
  for (int c = 0; c < NV; ++c) a[c] = (T)alpha[pt * NV + c];
  for (int64_t i = (int64_t)blockIdx.x * gpb + threadIdx.x / V; i < m; i += (int64_t)gridDim.x * gpb) {
    const int64_t id = i * kp + pt * NV;
    VecT<T, NV> x = vload<T, NV>(X + id);
    const VecT<T, NV> p = vload<T, NV>(P + id);
#pragma unroll
    for (int c = 0; c < NV; ++c) x.v[c] += a[c] * p.v[c];
    vstore_nt<T, NV>(X + id, x);
  }
}
template <typename T>
void launch_cg_x(hipStream_t st, T* X, const T* P, const double* alpha, int64_t m, int kp) {
  WA_VEC_SWITCH(T, kp, (k_cg_x<T, NV><<<red_blocks(m, kp), kBlock, 0, st>>>(X, P, alpha, m, kp)));
}
template void launch_cg_x<float>(hipStream_t, float*, const float*, const double*, int64_t, int);
template void launch_cg_x<double>(hipStream_t, double*, const double*, const double*, int64_t, int);
template int launch_cg_update<float>(hipStream_t, float*, const float*, float*, const float*, const double*,
                                     int64_t, int, double*);
template int launch_cg_update<double>(hipStream_t, double*, const double*, double*, const double*, const double*,
                                      int64_t, int, double*);

// P = Z + beta P (beta per column; columns with mask 0 untouched)
template <typename T, int NV>
__global__ void __launch_bounds__(kBlock) k_xpby(T* __restrict__ P, const T* __restrict__ Z,
                                                 const double* __restrict__ beta, const int* __restrict__ mask,
                                                 int64_t m, int kp) {
  const int V = kp / NV, pt = threadIdx.x % V, gpb = kBlock / V;
  T b[NV];
  bool on[NV], any = false;
#pragma unroll
  for (int c = 0; c < NV; ++c) {
    b[c] = (T)beta[pt * NV + c];
    on[c] = mask[pt * NV + c] != 0;
    any = any || on[c];
  }
  if (!any) return;
  for (int64_t i = (int64_t)blockIdx.x * gpb + threadIdx.x / V; i < m; i += (int64_t)gridDim.x * gpb) {
    const int64_t id = i * kp + pt * NV;
    VecT<T, NV> p = vload<T, NV>(P + id);
    const VecT<T, NV> z = vload<T, NV>(Z + id);
#pragma unroll
    for (int c = 0; c < NV; ++c)
      if (on[c]) p.v[c] = z.v[c] + b[c] * p.v[c];
    vstore_nt<T, NV>(P + id, p);
  }
}
template <typename T>
void launch_xpby(hipStream_t st, T* P, const T* Z, const double* beta, const int* mask, int64_t m, int kp) {
  WA_VEC_SWITCH(T, kp, (k_xpby<T, NV><<<red_blocks(m, kp), kBlock, 0, st>>>(P, Z, beta, mask, m, kp)));
}
template void launch_xpby<float>(hipStream_t, float*, const float*, const double*, const int*, int64_t, int);
template void launch_xpby<double>(hipStream_t, double*, const double*, const double*, const int*, int64_t, int);

// dst = scale[c] * src  (type conversion; scale may be null = 1)
template <typename S, typename D>
__global__ void __launch_bounds__(kBlock) k_cast_scale(const S* __restrict__ src, D* __restrict__ dst,
                                                       const double* __restrict__ scale, int64_t m, int kp) {
  const int lane = threadIdx.x & (kp - 1);
  const double s = scale ? scale[lane] : 1.0;
  const int gpb = kBlock / kp;
  for (int64_t i = (int64_t)blockIdx.x * gpb + threadIdx.x / kp; i < m; i += (int64_t)gridDim.x * gpb)
    dst[i * kp + lane] = (D)((double)src[i * kp + lane] * s);
}
template <typename S, typename D>
void launch_cast_scale(hipStream_t st, const S* src, D* dst, const double* scale, int64_t m, int kp) {
  k_cast_scale<S, D><<<red_blocks(m, kp), kBlock, 0, st>>>(src, dst, scale, m, kp);
}
template void launch_cast_scale<double, float>(hipStream_t, const double*, float*, const double*, int64_t, int);
template void launch_cast_scale<double, double>(hipStream_t, const double*, double*, const double*, int64_t, int);

// X += scale[c] * D (FP64 accumulate of an FP32 or FP64 correction); mask 0 = skip
template <typename S>
__global__ void __launch_bounds__(kBlock) k_axpy_acc(double* __restrict__ X, const S* __restrict__ Dl,
                                                     const double* __restrict__ scale,
                                                     const int* __restrict__ mask, int64_t m, int kp) {
  const int lane = threadIdx.x & (kp - 1);
  if (mask && !mask[lane]) return;
  const double s = scale ? scale[lane] : 1.0;
  const int gpb = kBlock / kp;
  for (int64_t i = (int64_t)blockIdx.x * gpb + threadIdx.x / kp; i < m; i += (int64_t)gridDim.x * gpb)
    X[i * kp + lane] += s * (double)Dl[i * kp + lane];
}
template <typename S>
void launch_axpy_acc(hipStream_t st, double* X, const S* Dl, const double* scale, const int* mask, int64_t m,
                     int kp) {
  k_axpy_acc<S><<<red_blocks(m, kp), kBlock, 0, st>>>(X, Dl, scale, mask, m, kp);
}
template void launch_axpy_acc<float>(hipStream_t, double*, const float*, const double*, const int*, int64_t, int);
template void launch_axpy_acc<double>(hipStream_t, double*, const double*, const double*, const int*, int64_t,
                                      int);

// ===========================================================================
// CG bookkeeping on the device, one thread per column (solve/cg.rs:45-133)
// ===========================================================================
__global__ void k_cg_start(ColState* __restrict__ s, const double* __restrict__ bnorm2,
                           const double* __restrict__ rz, const double* __restrict__ tol, int k, int kp) {
  const int c = threadIdx.x;
  if (c >= kp) return;
  ColState& q = s[c];
  const double bn = sqrt(bnorm2[c]);
  q.bnorm = bn;
  q.rnorm = bn;
  q.tol = c < k ? tol[c] : 0.0;
  q.rz = rz[c];
  q.rz0 = rz[c];
  q.iters = 0;
  q.alpha = 0.0;
  q.beta = 0.0;
  q.nan_stop = 0;
  if (c >= k) { q.active = 0; q.converged = 1; return; }
  // zero-RHS shortcut: ||b|| < EPS -> x = 0, converged, 0 iterations
  if (bn < 2.220446049250313e-16) { q.active = 0; q.converged = 1; return; }
  q.active = 1;
  q.converged = 0;
}

// Column totals of per-block partials (part[b * kp + c], b < nb): threads
// stride over the blocks, then a fixed-order tree; the result is valid in
// threads [0, kp).  Launched with kScalarBlock threads (more loads in flight).
constexpr int kScalarBlock = 1024;
__device__ inline double col_total(const double* __restrict__ part, int nb, int kp, double* sh) {
  const int col = threadIdx.x % kp, str = kScalarBlock / kp, s0 = threadIdx.x / kp;
  double acc = 0.0;
  for (int b = s0; b < nb; b += str) acc += part[(int64_t)b * kp + col];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = str / 2; w >= 1; w >>= 1) {
    if (s0 < w) sh[threadIdx.x] += sh[threadIdx.x + w * kp];
    __syncthreads();
  }
  return sh[threadIdx.x];
}

// after A p: pap <= 0 -> stop (not converged); else alpha = rz / pap (a NaN
// pap or rz gives a NaN alpha, which carries into x as in cg.rs:77-86)
__global__ void __launch_bounds__(kScalarBlock) k_cg_alpha(ColState* __restrict__ s, const double* __restrict__ part,
                                                     int nb, double* __restrict__ pap, double* __restrict__ alpha,
                                                     int kp) {
  __shared__ double sh[kScalarBlock];
  const double v = col_total(part, nb, kp, sh);
  const int c = threadIdx.x;
  if (c >= kp) return;
  pap[c] = v;
  ColState& q = s[c];
  alpha[c] = 0.0;
  if (!q.active) return;
  // every active column is in its q.iters-th iteration (the count lives on
  // the device, so a replayed iteration graph keeps it)
  q.iters += 1;
  if (v <= 0.0) {
    q.active = 0;
    q.converged = 0;
    return;
  }
  alpha[c] = q.rz / v;
  q.alpha = alpha[c];
}

// after the update: ||r|| / ||b|| <= tol -> converged
__global__ void __launch_bounds__(kScalarBlock) k_cg_check(ColState* __restrict__ s, const double* __restrict__ part,
                                                     int nb, double* __restrict__ rr, int kp) {
  __shared__ double sh[kScalarBlock];
  const double v = col_total(part, nb, kp, sh);
  const int c = threadIdx.x;
  if (c >= kp) return;
  rr[c] = v;
  ColState& q = s[c];
  if (!q.active) return;
  q.rnorm = sqrt(v);
  if (q.rnorm / q.bnorm <= q.tol) {
    q.active = 0;
    q.converged = 1;
  } else if (isnan(q.rnorm)) {
    // a NaN residual has reached x and no later step can remove it: stop
    // (not converged) instead of running out the iteration budget; the host
    // reports the budget as spent, as the reference's loop does (cg.rs:75-125
    // keeps iterating on NaN and returns maxiter)
    q.active = 0;
    q.converged = 0;
    q.nan_stop = 1;
  }
}

// after z = M r: stagnation guard |rz_new| < EPS^2 max(|rz0|, MIN); beta
__global__ void __launch_bounds__(kScalarBlock) k_cg_beta(ColState* __restrict__ s, const double* __restrict__ part,
                                                    int nb, double* __restrict__ rznew, double* __restrict__ beta,
                                                    int* __restrict__ mask, int* __restrict__ n_active, int kp) {
  __shared__ double sh[kScalarBlock];
  const double v = col_total(part, nb, kp, sh);
  const int c = threadIdx.x;
  if (c < kp) {
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
        beta[c] = v / q.rz;
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

void launch_cg_start(hipStream_t st, ColState* s, const double* bnorm2, const double* rz, const double* tol, int k,
                     int kp) {
  k_cg_start<<<1, 64, 0, st>>>(s, bnorm2, rz, tol, k, kp);
}
void launch_cg_alpha(hipStream_t st, ColState* s, const double* part, int nb, double* pap, double* alpha, int kp) {
  k_cg_alpha<<<1, kScalarBlock, 0, st>>>(s, part, nb, pap, alpha, kp);
}
void launch_cg_check(hipStream_t st, ColState* s, const double* part, int nb, double* rr, int kp) {
  k_cg_check<<<1, kScalarBlock, 0, st>>>(s, part, nb, rr, kp);
}
void launch_cg_beta(hipStream_t st, ColState* s, const double* part, int nb, double* rznew, double* beta, int* mask,
                    int* n_active, int kp) {
  k_cg_beta<<<1, kScalarBlock, 0, st>>>(s, part, nb, rznew, beta, mask, n_active, kp);
}

// ===========================================================================
// LSMR (schwarz-precond/src/solve/lsmr.rs:122-316) over W^(1/2) D: the
// observation-space half of the modified Golub-Kahan step, the per-column
// scalar recurrences (one thread per column, FP64) and the fused DOF-space
// vector updates.  Columns advance in lockstep; a column that has stopped is
// left untouched by every vector kernel.
// ===========================================================================
template <int NV>
__global__ void __launch_bounds__(kBlock) k_lsmr_rhs(int64_t n, const double* __restrict__ sw,
                                                     const double* __restrict__ Y, double* __restrict__ U, int kp,
                                                     double* __restrict__ part) {
  __shared__ double sh[kBlock * NV];
  const int V = kp / NV, pt = threadIdx.x % V, gpb = kBlock / V;
  double acc[NV];
#pragma unroll
  for (int c = 0; c < NV; ++c) acc[c] = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * gpb + threadIdx.x / V; i < n; i += (int64_t)gridDim.x * gpb) {
    VecT<double, NV> u = vload<double, NV>(Y + i * kp + pt * NV);
    const double w = sw ? sw[i] : 1.0;
#pragma unroll
    for (int c = 0; c < NV; ++c) {
      u.v[c] = sw ? w * u.v[c] : u.v[c];
      acc[c] += u.v[c] * u.v[c];
    }
    vstore<double, NV>(U + i * kp + pt * NV, u);
  }
  block_col_sums<NV>(acc, sh, kp, gpb, part);
}

template <int NV>
__global__ void __launch_bounds__(kBlock) k_lsmr_fwd(int64_t n, int Q, const uint32_t* __restrict__ codes,
                                                     const double* __restrict__ sw, const double* __restrict__ X,
                                                     double* __restrict__ U, const LsmrCol* __restrict__ cs, int kp,
                                                     double* __restrict__ part) {
  __shared__ double sh[kBlock * NV];
  const int V = kp / NV, pt = threadIdx.x % V, gpb = kBlock / V;
  double acc[NV], sc[NV];
  bool act[NV], any = false;
#pragma unroll
  for (int c = 0; c < NV; ++c) {
    const LsmrCol& q = cs[pt * NV + c];
    act[c] = q.active != 0;
    sc[c] = -(q.alpha * q.beta_prev_inv);
    acc[c] = 0.0;
    any = any || act[c];
  }
  if (any) {
    for (int64_t i = (int64_t)blockIdx.x * gpb + threadIdx.x / V; i < n; i += (int64_t)gridDim.x * gpb) {
      double s[NV];
#pragma unroll
      for (int c = 0; c < NV; ++c) s[c] = 0.0;
      for (int q = 0; q < Q; ++q) {
        const VecT<double, NV> x = vload<double, NV>(X + (int64_t)codes[(int64_t)q * n + i] * kp + pt * NV);
#pragma unroll
        for (int c = 0; c < NV; ++c) s[c] += x.v[c];
      }
      const double w = sw ? sw[i] : 1.0;
      VecT<double, NV> u = vload<double, NV>(U + i * kp + pt * NV);
#pragma unroll
      for (int c = 0; c < NV; ++c)
        if (act[c]) {
          const double av = sw ? s[c] * w : s[c];
          const double val = av + sc[c] * u.v[c];
          u.v[c] = val;
          acc[c] += val * val;
        }
      vstore<double, NV>(U + i * kp + pt * NV, u);
    }
  }
  block_col_sums<NV>(acc, sh, kp, gpb, part);
}

static inline unsigned lsmr_blocks(int64_t n, int kp) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(blocks_for(n, kp), kLsmrBlocks));
}

int launch_lsmr_rhs(hipStream_t st, int64_t n, const double* sw, const double* Y, double* U, int kp, double* part) {
  const unsigned nb = lsmr_blocks(n, kp);
  WA_VEC_SWITCH(double, kp, (k_lsmr_rhs<NV><<<nb, kBlock, 0, st>>>(n, sw, Y, U, kp, part)));
  return (int)nb;
}

int launch_lsmr_fwd(hipStream_t st, int64_t n, int Q, const uint32_t* codes, const double* sw, const double* V,
                    double* U, const LsmrCol* cs, int kp, double* part) {
  const unsigned nb = lsmr_blocks(n, kp);
  WA_VEC_SWITCH(double, kp, (k_lsmr_fwd<NV><<<nb, kBlock, 0, st>>>(n, Q, codes, sw, V, U, cs, kp, part)));
  return (int)nb;
}

// beta_1 = ||b|| (lsmr.rs:188-201, bidiag.rs:551-560): a zero right-hand side
// stops at once (x = 0, converged); a non-finite one is flagged
__global__ void __launch_bounds__(kScalarBlock) k_lsmr_start(LsmrCol* __restrict__ cs, const double* __restrict__ part,
                                                       int nb, double tol, int k, int kp) {
  __shared__ double sh[kScalarBlock];
  const double v = col_total(part, nb, kp, sh);
  const int c = threadIdx.x;
  if (c >= kp) return;
  LsmrCol q = {};
  const double beta = sqrt(v);
  q.beta = beta;
  q.bnorm = beta;
  q.rel_tol = tol;
  q.abs_tol = tol * beta;
  q.c_bar = 1.0;
  q.prev_rho = q.prev_rho_bar = 1.0;
  if (c < k && !isfinite(v)) q.bad = 1;
  if (c < k && isfinite(v) && beta > 0.0) {
    q.active = 1;
    q.beta_inv = 1.0 / beta;
    q.beta_prev_inv = q.beta_inv;
  } else {
    q.converged = 1;  // zero right-hand side (or padding, or flagged)
  }
  cs[c] = q;
}

// alpha_1 = sqrt(<v, p~>) and the recurrence seeds (bidiag.rs:562-589,
// recurrence.rs:78-86, 240-246); A'b = 0 stops with x = 0 and |r| = ||b||
__global__ void __launch_bounds__(kScalarBlock) k_lsmr_alpha1(LsmrCol* __restrict__ cs,
                                                        const double* __restrict__ part, int nb, int kp,
                                                        int* __restrict__ n_active) {
  __shared__ double sh[kScalarBlock];
  const double v = col_total(part, nb, kp, sh);
  const int c = threadIdx.x;
  if (c < kp) {
    LsmrCol& q = cs[c];
    q.inv_alpha = 0.0;
    if (q.active) {
      const double alpha = v > 0.0 ? sqrt(v) : 0.0;
      q.alpha = alpha;
      q.inv_alpha = alpha > 0.0 ? 1.0 / alpha : 0.0;
      q.alpha_bar = alpha;
      q.phi_bar = q.bnorm;
      q.zeta_bar = alpha * q.bnorm;
      q.a_norm_sq = alpha * alpha;
      if (alpha == 0.0) {
        q.active = 0;
        q.converged = 1;
        q.stop = 1;
        q.resid = q.bnorm;
      }
    }
  }
  __syncthreads();
  if (c == 0) {
    int a = 0;
    for (int j = 0; j < kp; ++j) a += cs[j].active;
    *n_active = a;
  }
}

// beta_{k+1} = ||u|| (bidiag.rs:348-366); a zero beta zeroes p~ and so v
__global__ void __launch_bounds__(kScalarBlock) k_lsmr_beta(LsmrCol* __restrict__ cs, const double* __restrict__ part,
                                                      int nb, int kp) {
  __shared__ double sh[kScalarBlock];
  const double v = col_total(part, nb, kp, sh);
  const int c = threadIdx.x;
  if (c >= kp) return;
  LsmrCol& q = cs[c];
  if (!q.active) return;
  const double beta = sqrt(v);
  q.beta = beta;
  if (beta == 0.0) {
    q.beta_inv = 0.0;
    q.p_coeff = 0.0;
  } else {
    q.beta_inv = 1.0 / beta;
    q.p_coeff = beta / q.alpha;
  }
}

struct DevGivens {  // recurrence.rs:20-30
  double c, s, r;
  __device__ DevGivens(double a, double b) {
    r = hypot(a, b);
    if (r > 0.0) {
      c = a / r;
      s = b / r;
    } else {
      c = 1.0;
      s = 0.0;
    }
  }
};

// alpha_{k+1} = sqrt(<v, p~>), both rotations, the x / h / hbar coefficients
// and the stopping tests (recurrence.rs:89-138, 181-229, 263-275; lsmr.rs:225-251)
__global__ void __launch_bounds__(kScalarBlock) k_lsmr_step(LsmrCol* __restrict__ cs, const double* __restrict__ part,
                                                      int nb, int itn, int kp, int* __restrict__ n_active) {
  __shared__ double sh[kScalarBlock];
  const double v = col_total(part, nb, kp, sh);
  const int c = threadIdx.x;
  if (c < kp) {
    LsmrCol& q = cs[c];
    q.upd = q.active;
    if (q.active) {
      const double eps = 2.220446049250313e-16, tiny = 2.2250738585072014e-308;
      const double alpha = v > 0.0 ? sqrt(v) : 0.0, beta = q.beta;
      q.a_norm_sq += alpha * alpha + beta * beta;
      const DevGivens ph(q.alpha_bar, beta);
      const double theta_new = ph.s * alpha;
      const double alpha_bar_new = -ph.c * alpha, phi_bar_new = ph.s * q.phi_bar;
      const double theta_bar = q.s_bar * ph.r;
      const DevGivens pb(q.c_bar * ph.r, theta_new);
      const double zeta = pb.c * q.zeta_bar, zeta_bar_new = -pb.s * q.zeta_bar;
      q.alpha_bar = alpha_bar_new;
      q.phi_bar = phi_bar_new;
      q.c_bar = pb.c;
      q.s_bar = pb.s;
      q.zeta_bar = zeta_bar_new;
      const double rho = ph.r, rho_bar = pb.r;
      const double dx = rho * rho_bar, dh = q.prev_rho * q.prev_rho_bar;
      q.t_x = fabs(dx) > eps ? zeta / dx : 0.0;
      q.t_hbar = fabs(dh) > eps ? theta_bar * rho / dh : 0.0;
      q.t_h = fabs(rho) > eps ? theta_new / rho : 0.0;
      q.prev_rho = rho;
      q.prev_rho_bar = rho_bar;
      q.inv_alpha = alpha > 0.0 ? 1.0 / alpha : 0.0;
      q.alpha = alpha;
      q.beta_prev_inv = q.beta_inv;
      q.iters = itn;
      const double resid = fabs(q.phi_bar);
      q.resid = resid;
      int stop = -1;
      if (resid <= q.abs_tol) {
        stop = 2;
      } else {
        const double a_norm = fmax(sqrt(q.a_norm_sq), tiny);
        if (fabs(q.zeta_bar) / (a_norm * fmax(resid, tiny)) <= q.rel_tol) stop = 3;
      }
      if (stop < 0 && alpha == 0.0) stop = 4;
      if (stop >= 0) {
        q.active = 0;
        q.converged = 1;
        q.stop = stop;
      }
    }
  }
  __syncthreads();
  if (c == 0) {
    int a = 0;
    for (int j = 0; j < kp; ++j) a += cs[j].active;
    *n_active = a;
  }
}

void launch_lsmr_start(hipStream_t st, LsmrCol* cs, const double* part, int nb, double tol, int k, int kp) {
  k_lsmr_start<<<1, kScalarBlock, 0, st>>>(cs, part, nb, tol, k, kp);
}
void launch_lsmr_alpha1(hipStream_t st, LsmrCol* cs, const double* part, int nb, int kp, int* n_active) {
  k_lsmr_alpha1<<<1, kScalarBlock, 0, st>>>(cs, part, nb, kp, n_active);
}
void launch_lsmr_beta(hipStream_t st, LsmrCol* cs, const double* part, int nb, int kp) {
  k_lsmr_beta<<<1, kScalarBlock, 0, st>>>(cs, part, nb, kp);
}
void launch_lsmr_step(hipStream_t st, LsmrCol* cs, const double* part, int nb, int itn, int kp, int* n_active) {
  k_lsmr_step<<<1, kScalarBlock, 0, st>>>(cs, part, nb, itn, kp, n_active);
}

// DOF-space vector kernels: one lane per (row, column) value
__global__ void __launch_bounds__(kBlock) k_lsmr_ptilde(double* __restrict__ P, const double* __restrict__ ATU,
                                                        const LsmrCol* __restrict__ cs, int64_t m, int kp) {
  const int lane = threadIdx.x % kp;
  const LsmrCol& q = cs[lane];
  if (!q.active) return;
  const double bi = q.beta_inv, pc = -q.p_coeff;
  const int gpb = kBlock / kp;
  for (int64_t i = (int64_t)blockIdx.x * gpb + threadIdx.x / kp; i < m; i += (int64_t)gridDim.x * gpb) {
    const int64_t id = i * kp + lane;
    P[id] = bi * ATU[id] + pc * P[id];
  }
}

__global__ void __launch_bounds__(kBlock) k_lsmr_init_vec(double* __restrict__ V, double* __restrict__ H,
                                                          double* __restrict__ HB, double* __restrict__ X,
                                                          const LsmrCol* __restrict__ cs, int64_t m, int kp) {
  const int lane = threadIdx.x % kp;
  const double ia = cs[lane].inv_alpha;
  const int gpb = kBlock / kp;
  for (int64_t i = (int64_t)blockIdx.x * gpb + threadIdx.x / kp; i < m; i += (int64_t)gridDim.x * gpb) {
    const int64_t id = i * kp + lane;
    const double v = V[id] * ia;
    V[id] = v;
    H[id] = v;
    HB[id] = 0.0;
    X[id] = 0.0;
  }
}

__global__ void __launch_bounds__(kBlock) k_lsmr_update(double* __restrict__ V, double* __restrict__ H,
                                                        double* __restrict__ HB, double* __restrict__ X,
                                                        const LsmrCol* __restrict__ cs, int64_t m, int kp) {
  const int lane = threadIdx.x % kp;
  const LsmrCol& q = cs[lane];
  if (!q.upd) return;
  const double ia = q.inv_alpha, tx = q.t_x, thb = q.t_hbar, th = q.t_h;
  const bool scale = ia > 0.0;
  const int gpb = kBlock / kp;
  for (int64_t i = (int64_t)blockIdx.x * gpb + threadIdx.x / kp; i < m; i += (int64_t)gridDim.x * gpb) {
    const int64_t id = i * kp + lane;
    const double v = scale ? V[id] * ia : V[id];
    const double h_old = H[id];
    const double hb = h_old - thb * HB[id];
    HB[id] = hb;
    X[id] += tx * hb;
    H[id] = v - th * h_old;
    V[id] = v;
  }
}

__global__ void __launch_bounds__(kBlock) k_lsmr_reorth(double* __restrict__ V, double* __restrict__ P,
                                                        const double* __restrict__ Vj, const double* __restrict__ Pj,
                                                        const double* __restrict__ coef,
                                                        const LsmrCol* __restrict__ cs, int64_t m, int kp) {
  const int lane = threadIdx.x % kp;
  if (!cs[lane].active) return;
  const double c = -coef[lane];
  const int gpb = kBlock / kp;
  for (int64_t i = (int64_t)blockIdx.x * gpb + threadIdx.x / kp; i < m; i += (int64_t)gridDim.x * gpb) {
    const int64_t id = i * kp + lane;
    V[id] = c * Vj[id] + V[id];
    P[id] = c * Pj[id] + P[id];
  }
}

__global__ void __launch_bounds__(kBlock) k_lsmr_push(const double* __restrict__ V, const double* __restrict__ P,
                                                      double* __restrict__ Vs, double* __restrict__ Ps,
                                                      const LsmrCol* __restrict__ cs, int64_t m, int kp) {
  const int lane = threadIdx.x % kp;
  const double ia = cs[lane].inv_alpha;
  const int gpb = kBlock / kp;
  for (int64_t i = (int64_t)blockIdx.x * gpb + threadIdx.x / kp; i < m; i += (int64_t)gridDim.x * gpb) {
    const int64_t id = i * kp + lane;
    Vs[id] = V[id];
    Ps[id] = P[id] * ia;
  }
}

static inline unsigned vec_blocks(int64_t m, int kp) { return grid_stride_blocks(m, kp); }

void launch_lsmr_ptilde(hipStream_t st, double* P, const double* ATU, const LsmrCol* cs, int64_t m, int kp) {
  if (m > 0) k_lsmr_ptilde<<<vec_blocks(m, kp), kBlock, 0, st>>>(P, ATU, cs, m, kp);
}
void launch_lsmr_init_vec(hipStream_t st, double* V, double* H, double* HB, double* X, const LsmrCol* cs, int64_t m,
                          int kp) {
  if (m > 0) k_lsmr_init_vec<<<vec_blocks(m, kp), kBlock, 0, st>>>(V, H, HB, X, cs, m, kp);
}
void launch_lsmr_update(hipStream_t st, double* V, double* H, double* HB, double* X, const LsmrCol* cs, int64_t m,
                        int kp) {
  if (m > 0) k_lsmr_update<<<vec_blocks(m, kp), kBlock, 0, st>>>(V, H, HB, X, cs, m, kp);
}
void launch_lsmr_reorth(hipStream_t st, double* V, double* P, const double* Vj, const double* Pj, const double* c,
                        const LsmrCol* cs, int64_t m, int kp) {
  if (m > 0) k_lsmr_reorth<<<vec_blocks(m, kp), kBlock, 0, st>>>(V, P, Vj, Pj, c, cs, m, kp);
}
void launch_lsmr_push(hipStream_t st, const double* V, const double* P, double* Vs, double* Ps, const LsmrCol* cs,
                      int64_t m, int kp) {
  if (m > 0) k_lsmr_push<<<vec_blocks(m, kp), kBlock, 0, st>>>(V, P, Vs, Ps, cs, m, kp);
}

// Preconditioner elementwise kernels (same lane mapping as the vector ops).

// Rr[j] -= sum2[s] / n_red(s)
template <typename T, int NV>
__global__ void __launch_bounds__(kBlock) k_sub_mean(int64_t b0, int64_t n, const uint32_t* __restrict__ sub,
                                                     const uint32_t* __restrict__ ptr, const T* __restrict__ sums,
                                                     T* __restrict__ V_, int kp) {
  const int V = kp / NV, pt = threadIdx.x % V, gpb = kBlock / V;
  for (int64_t j = b0 + (int64_t)blockIdx.x * gpb + threadIdx.x / V; j < n; j += (int64_t)gridDim.x * gpb) {
    const uint32_t s = sub[j];
    const T cnt = (T)(ptr[s + 1] - ptr[s]);
    const VecT<T, NV> m = vload<T, NV>(sums + (int64_t)s * kp + pt * NV);
    VecT<T, NV> x = vload<T, NV>(V_ + j * kp + pt * NV);
#pragma unroll
    for (int c = 0; c < NV; ++c) x.v[c] -= m.v[c] / cnt;
    vstore<T, NV>(V_ + j * kp + pt * NV, x);
  }
}
template <typename T>
static void launch_sub_mean(hipStream_t st, int64_t b0, int64_t n, const uint32_t* sub, const uint32_t* ptr,
                            const T* sums, T* V, int kp) {
  if (n <= b0) return;
  WA_VEC_SWITCH(T, kp, (k_sub_mean<T, NV><<<grid_stride_blocks(n - b0, kp / NV), kBlock, 0, st>>>(b0, n, sub, ptr,
                                                                                                   sums, V, kp)));
}

// dst[idx[i]] = src[i]
template <typename T, int NV>
__global__ void __launch_bounds__(kBlock) k_scatter_rows(int64_t b0, int64_t n, const uint32_t* __restrict__ idx,
                                                         const T* __restrict__ src, T* __restrict__ dst, int kp) {
  const int V = kp / NV, pt = threadIdx.x % V, gpb = kBlock / V;
  for (int64_t i = b0 + (int64_t)blockIdx.x * gpb + threadIdx.x / V; i < n; i += (int64_t)gridDim.x * gpb)
    vstore<T, NV>(dst + (int64_t)idx[i] * kp + pt * NV, vload<T, NV>(src + i * kp + pt * NV));
}
template <typename T>
static void launch_scatter_rows(hipStream_t st, int64_t b0, int64_t n, const uint32_t* idx, const T* src, T* dst,
                                int kp) {
  if (n <= b0) return;
  WA_VEC_SWITCH(T, kp, (k_scatter_rows<T, NV><<<grid_stride_blocks(n - b0, kp / NV), kBlock, 0, st>>>(b0, n, idx, src,
                                                                                                       dst, kp)));
}

// Z[d] = sum_{l in dof d} scale_l (S[l] - mu3(sub(l)))   (0 for uncovered DOFs)
template <typename T, int NV>
__global__ void __launch_bounds__(kBlock) k_combine(int64_t m, const uint32_t* __restrict__ dptr,
                                                    const uint32_t* __restrict__ dloc,
                                                    const T* __restrict__ scale,
                                                    const uint32_t* __restrict__ loc_sub,
                                                    const uint32_t* __restrict__ loc_ptr,
                                                    const T* __restrict__ S, const T* __restrict__ sum3,
                                                    T* __restrict__ Z, int kp) {
  const int V = kp / NV, pt = threadIdx.x % V, gpb = kBlock / V;
  for (int64_t d = (int64_t)blockIdx.x * gpb + threadIdx.x / V; d < m; d += (int64_t)gridDim.x * gpb) {
    T acc[NV];
#pragma unroll
    for (int c = 0; c < NV; ++c) acc[c] = 0;
    for (uint32_t p = dptr[d]; p < dptr[d + 1]; ++p) {
      const uint32_t l = dloc[p];
      const uint32_t s = loc_sub[l];
      const T cnt = (T)(loc_ptr[s + 1] - loc_ptr[s]), sc = scale[l];
      const VecT<T, NV> mu = vload<T, NV>(sum3 + (int64_t)s * kp + pt * NV);
      const VecT<T, NV> x = vload<T, NV>(S + (int64_t)l * kp + pt * NV);
#pragma unroll
      for (int c = 0; c < NV; ++c) acc[c] += sc * (x.v[c] - mu.v[c] / cnt);
    }
    vstore<T, NV>(Z + d * kp + pt * NV, acc);
  }
}
template <typename T>
static void launch_combine(hipStream_t st, int64_t m, const uint32_t* dptr, const uint32_t* dloc, const T* scale,
                           const uint32_t* loc_sub, const uint32_t* loc_ptr, const T* S, const T* sum3, T* Z,
                           int kp) {
  if (m <= 0) return;
  WA_VEC_SWITCH(T, kp, (k_combine<T, NV><<<grid_stride_blocks(m, kp / NV), kBlock, 0, st>>>(
                          m, dptr, dloc, scale, loc_sub, loc_ptr, S, sum3, Z, kp)));
}

// ===========================================================================
// row kinds for the engine (engine.hpp)
// ===========================================================================
// implicit operator pass of factor q (GramianOperator::apply,
// operator/gramian.rs:84-89, as D'W D x):
//   Y[off + j] = diag_j X[off + j] + sum_{e in level j} w_e sum_{t<qm} X[other_t(e)]
template <typename T, int QM_, bool W>
struct OpKind {
  static constexpr int QM = QM_;
  static constexpr bool COEF = W;
  const T* X;
  T* Y;
  const uint32_t* row_ptr;
  const uint32_t* other;
  const T* ew;
  const T* diag;
  uint32_t off;
  int kp;
  __device__ uint32_t rbegin(uint32_t r) const { return row_ptr[r]; }
  __device__ uint32_t rend(uint32_t r) const { return row_ptr[r + 1]; }
  __device__ void load(uint32_t, uint32_t e, uint32_t* id, T& c) const {
#pragma unroll
    for (int t = 0; t < QM; ++t) id[t] = other[(int64_t)e * QM + t];
    if (W) c = ew[e];
  }
  template <int NV>
  struct Pre {
    T d;
    VecT<T, NV> x;
  };
  template <int NV>
  __device__ Pre<NV> pre(uint32_t r, int col) const {
    return {diag[r], vload<T, NV>(X + (int64_t)(off + r) * kp + col)};
  }
  template <int NV>
  __device__ void finish(uint32_t r, int col, const T (&acc)[NV], const Pre<NV>& q) const {
    VecT<T, NV> y;
#pragma unroll
    for (int c = 0; c < NV; ++c) y.v[c] = q.d * q.x.v[c] + acc[c];
    vstore<T, NV>(Y + (int64_t)(off + r) * kp + col, y);
  }
};

template <typename T, int QM, bool W>
static void run_op(hipStream_t st, const DevPass<T>& p, const T* X, T* Y, T* part, int kp) {
  OpKind<T, QM, W> k{X, Y, p.row_ptr, p.other, p.ew, p.diag, p.off, kp};
  run_rows<OpKind<T, QM, W>, T>(st, k, p.sched, part, kp);
}

template <typename T>
void launch_pull(hipStream_t st, const DevPass<T>& p, const T* X, T* Y, T* part, int kp) {
  const bool w = p.ew != nullptr;
  switch (p.qm) {
    case 0: w ? run_op<T, 0, true>(st, p, X, Y, part, kp) : run_op<T, 0, false>(st, p, X, Y, part, kp); break;
    case 1: w ? run_op<T, 1, true>(st, p, X, Y, part, kp) : run_op<T, 1, false>(st, p, X, Y, part, kp); break;
    case 2: w ? run_op<T, 2, true>(st, p, X, Y, part, kp) : run_op<T, 2, false>(st, p, X, Y, part, kp); break;
    case 3: w ? run_op<T, 3, true>(st, p, X, Y, part, kp) : run_op<T, 3, false>(st, p, X, Y, part, kp); break;
    case 4: w ? run_op<T, 4, true>(st, p, X, Y, part, kp) : run_op<T, 4, false>(st, p, X, Y, part, kp); break;
    case 5: w ? run_op<T, 5, true>(st, p, X, Y, part, kp) : run_op<T, 5, false>(st, p, X, Y, part, kp); break;
    case 6: w ? run_op<T, 6, true>(st, p, X, Y, part, kp) : run_op<T, 6, false>(st, p, X, Y, part, kp); break;
    case 7: w ? run_op<T, 7, true>(st, p, X, Y, part, kp) : run_op<T, 7, false>(st, p, X, Y, part, kp); break;
    default: break;  // more than 8 factors is rejected when the solver is built
  }
}
template void launch_pull<float>(hipStream_t, const DevPass<float>&, const float*, float*, float*, int);
template void launch_pull<double>(hipStream_t, const DevPass<double>&, const double*, double*, double*, int);

// observation-space D'W V for factor q (rmatvec_wdt, domain.rs:409-418)
template <bool W>
struct ObsKind {
  static constexpr int QM = 1;
  static constexpr bool COEF = W;
  const double* X;  // V, row-major n x kp
  double* out;
  const uint32_t* row_ptr;
  const uint32_t* obs;
  const double* w;
  uint32_t off;
  int kp;
  __device__ uint32_t rbegin(uint32_t r) const { return row_ptr[r]; }
  __device__ uint32_t rend(uint32_t r) const { return row_ptr[r + 1]; }
  __device__ void load(uint32_t, uint32_t e, uint32_t* id, double& c) const {
    const uint32_t i = obs[e];
    id[0] = i;
    if (W) c = w[i];
  }
  template <int NV>
  struct Pre {};
  template <int NV>
  __device__ Pre<NV> pre(uint32_t, int) const { return {}; }
  template <int NV>
  __device__ void finish(uint32_t r, int col, const double (&acc)[NV], const Pre<NV>&) const {
    vstore<double, NV>(out + (int64_t)(off + r) * kp + col, acc);
  }
};

void launch_scatter_obs(hipStream_t st, const DevPass<double>& p, const double* w, const double* V, double* out,
                        double* part, int kp) {
  if (w) {
    ObsKind<true> k{V, out, p.row_ptr, p.obs, w, p.off, kp};
    run_rows<ObsKind<true>, double>(st, k, p.sched, part, kp);
  } else {
    ObsKind<false> k{V, out, p.row_ptr, p.obs, nullptr, p.off, kp};
    run_rows<ObsKind<false>, double>(st, k, p.sched, part, kp);
  }
}

// explicit Gramian rows (gramian/explicit.rs:261-277)
template <typename T>
struct SpmmKind {
  static constexpr int QM = 1;
  static constexpr bool COEF = true;
  const T* X;
  T* Y;
  const uint32_t* ptr;
  const uint32_t* idx;
  const T* val;
  int kp;
  __device__ uint32_t rbegin(uint32_t r) const { return ptr[r]; }
  __device__ uint32_t rend(uint32_t r) const { return ptr[r + 1]; }
  __device__ void load(uint32_t, uint32_t e, uint32_t* id, T& c) const {
    id[0] = idx[e];
    c = val[e];
  }
  template <int NV>
  struct Pre {};
  template <int NV>
  __device__ Pre<NV> pre(uint32_t, int) const { return {}; }
  template <int NV>
  __device__ void finish(uint32_t r, int col, const T (&acc)[NV], const Pre<NV>&) const {
    VecT<T, NV> y;
#pragma unroll
    for (int c = 0; c < NV; ++c) y.v[c] = acc[c];
    vstore_nt<T, NV>(Y + (int64_t)r * kp + col, y);
  }
};

template <typename T>
void launch_spmm(hipStream_t st, const DevCsr<T>& g, const T* X, T* Y, T* part, int kp) {
  SpmmKind<T> k{X, Y, g.ptr, g.idx, g.val, kp};
  run_rows<SpmmKind<T>, T>(st, k, g.sched, part, kp);
}
template void launch_spmm<float>(hipStream_t, const DevCsr<float>&, const float*, float*, float*, int);
template void launch_spmm<double>(hipStream_t, const DevCsr<double>&, const double*, double*, double*, int);

// per-subdomain sums over contiguous ranges
template <typename T>
struct SegKind {
  static constexpr int QM = 1;
  static constexpr bool COEF = false;
  const T* X;
  T* out;
  const uint32_t* ptr;
  int kp;
  __device__ uint32_t rbegin(uint32_t r) const { return ptr[r]; }
  __device__ uint32_t rend(uint32_t r) const { return ptr[r + 1]; }
  __device__ void load(uint32_t, uint32_t e, uint32_t* id, T&) const { id[0] = e; }
  template <int NV>
  struct Pre {};
  template <int NV>
  __device__ Pre<NV> pre(uint32_t, int) const { return {}; }
  template <int NV>
  __device__ void finish(uint32_t r, int col, const T (&acc)[NV], const Pre<NV>&) const {
    vstore<T, NV>(out + (int64_t)r * kp + col, acc);
  }
};

// restriction fused with its per-subdomain sums (mu1): element l of row s
// (subdomain s) is B[l] = scale_l R[gidx_l], written as it is summed
template <typename T>
struct RestrictSegKind {
  static constexpr int QM = 1;
  static constexpr bool COEF = true;
  static constexpr bool EMIT = true;
  const T* X;  // R
  T* B;
  T* out;
  const uint32_t *ptr, *gidx;
  const T* scale;
  int kp;
  __device__ uint32_t rbegin(uint32_t r) const { return ptr[r]; }
  __device__ uint32_t rend(uint32_t r) const { return ptr[r + 1]; }
  __device__ void load(uint32_t, uint32_t e, uint32_t* id, T& c) const {
    id[0] = gidx[e];
    c = scale[e];
  }
  template <int NV>
  __device__ void emit(uint32_t e, int col, const VecT<T, NV>& v) const {
    vstore<T, NV>(B + (int64_t)e * kp + col, v);
  }
  template <int NV>
  struct Pre {};
  template <int NV>
  __device__ Pre<NV> pre(uint32_t, int) const { return {}; }
  template <int NV>
  __device__ void finish(uint32_t r, int col, const T (&acc)[NV], const Pre<NV>&) const {
    vstore<T, NV>(out + (int64_t)r * kp + col, acc);
  }
};

// Schur right-hand side rows (local_solver.rs:375-389):
//   Rr[j] = B[keep_j] - mu1 + sum_e (C[e,j]/D_e) B[e] - mu1 * rowsum_j
template <typename T>
struct KeepKind {
  static constexpr int QM = 1;
  static constexpr bool COEF = true;
  const T* X;  // B
  T* Rr;
  const uint32_t *ptr, *idx;
  const T *val, *rowsum;
  const uint32_t *red_loc, *red_sub, *loc_ptr;
  const T* sum1;
  int kp;
  __device__ uint32_t rbegin(uint32_t r) const { return ptr[r]; }
  __device__ uint32_t rend(uint32_t r) const { return ptr[r + 1]; }
  __device__ void load(uint32_t, uint32_t e, uint32_t* id, T& c) const {
    id[0] = idx[e];
    c = val[e];
  }
  template <int NV>
  struct Pre {
    T cnt, rs;
    VecT<T, NV> m, x;
  };
  template <int NV>
  __device__ Pre<NV> pre(uint32_t j, int col) const {
    const uint32_t s = red_sub[j];
    return {(T)(loc_ptr[s + 1] - loc_ptr[s]), rowsum[j], vload<T, NV>(sum1 + (int64_t)s * kp + col),
            vload<T, NV>(X + (int64_t)red_loc[j] * kp + col)};
  }
  template <int NV>
  __device__ void finish(uint32_t j, int col, const T (&acc)[NV], const Pre<NV>& q) const {
    VecT<T, NV> o;
#pragma unroll
    for (int c = 0; c < NV; ++c) {
      const T mu = q.m.v[c] / q.cnt;
      o.v[c] = q.x.v[c] - mu + acc[c] - mu * q.rs;
    }
    vstore<T, NV>(Rr + (int64_t)j * kp + col, o);
  }
};

// back-substitution rows (local_solver.rs:111-151, :400-447):
//   S[e] = (B[e] - mu1) / D_e + (1/D_e) sum_j C[e,j] x_keep[j]
template <typename T>
struct ElimKind {
  static constexpr int QM = 1;
  static constexpr bool SHARED_IDX = false;  // gathers hit a small, hot vector
  static constexpr bool COEF = true;
  const T* X;  // Rr (solved reduced system)
  T* S;
  const uint32_t *ptr, *idx;
  const T *val, *inv;
  const uint32_t *eloc, *loc_sub, *loc_ptr;
  const T *B, *sum1;
  int kp;
  __device__ uint32_t rbegin(uint32_t r) const { return ptr[r]; }
  __device__ uint32_t rend(uint32_t r) const { return ptr[r + 1]; }
  __device__ void load(uint32_t, uint32_t e, uint32_t* id, T& c) const {
    id[0] = idx[e];
    c = val[e];
  }
  template <int NV>
  struct Pre {
    uint32_t l;
    T cnt, iv;
    VecT<T, NV> m, b;
  };
  template <int NV>
  __device__ Pre<NV> pre(uint32_t i, int col) const {
    const uint32_t l = eloc[i];
    const uint32_t s = loc_sub[l];
    return {l, (T)(loc_ptr[s + 1] - loc_ptr[s]), inv[i], vload<T, NV>(sum1 + (int64_t)s * kp + col),
            vload<T, NV>(B + (int64_t)l * kp + col)};
  }
  template <int NV>
  __device__ void finish(uint32_t, int col, const T (&acc)[NV], const Pre<NV>& q) const {
    const uint32_t l = q.l;
    VecT<T, NV> o;
#pragma unroll
    for (int c = 0; c < NV; ++c) {
      const T mu = q.m.v[c] / q.cnt;
      o.v[c] = (q.b.v[c] - mu) * q.iv + q.iv * acc[c];
    }
    vstore<T, NV>(S + (int64_t)l * kp + col, o);
  }
};

// forward elimination level: z_v += sum_{a earlier} (w_av / d_a) z_a  (in place)
template <typename T>
struct FwdKind {
  static constexpr int QM = 1;
  static constexpr bool SHARED_IDX = false;  // gathers hit a small, hot vector
  static constexpr bool COEF = true;
  T* X;  // Z
  const uint32_t *rows, *ptr, *idx;
  const T* val;
  int kp;
  __device__ uint32_t rbegin(uint32_t r) const { return ptr[r]; }
  __device__ uint32_t rend(uint32_t r) const { return ptr[r + 1]; }
  __device__ void load(uint32_t, uint32_t e, uint32_t* id, T& c) const {
    id[0] = idx[e];
    c = val[e];
  }
  // a level's rows are not gathered by the same level: X[rows[r]] can be
  // read before the gathers
  template <int NV>
  struct Pre {
    T* p;
    VecT<T, NV> x;
  };
  template <int NV>
  __device__ Pre<NV> pre(uint32_t r, int col) const {
    T* p = X + (int64_t)rows[r] * kp + col;
    return {p, vload<T, NV>(p)};
  }
  template <int NV>
  __device__ void finish(uint32_t, int, const T (&acc)[NV], const Pre<NV>& q) const {
    VecT<T, NV> x = q.x;
#pragma unroll
    for (int c = 0; c < NV; ++c) x.v[c] += acc[c];
    vstore<T, NV>(q.p, x);
  }
};

// backward level: x_v = z_v / d_v + sum_{u later} (w_vu / d_v) x_u  (in place)
template <typename T>
struct BwdKind {
  static constexpr int QM = 1;
  static constexpr bool SHARED_IDX = false;  // gathers hit a small, hot vector
  static constexpr bool COEF = true;
  T* X;
  const uint32_t *rows, *ptr, *idx;
  const T *val, *inv_piv;
  int kp;
  __device__ uint32_t rbegin(uint32_t r) const { return ptr[r]; }
  __device__ uint32_t rend(uint32_t r) const { return ptr[r + 1]; }
  __device__ void load(uint32_t, uint32_t e, uint32_t* id, T& c) const {
    id[0] = idx[e];
    c = val[e];
  }
  template <int NV>
  struct Pre {
    T* p;
    T ip;
    VecT<T, NV> x;
  };
  template <int NV>
  __device__ Pre<NV> pre(uint32_t r, int col) const {
    const uint32_t v = rows[r];
    T* p = X + (int64_t)v * kp + col;
    return {p, inv_piv[v], vload<T, NV>(p)};
  }
  template <int NV>
  __device__ void finish(uint32_t, int, const T (&acc)[NV], const Pre<NV>& q) const {
    VecT<T, NV> x;
#pragma unroll
    for (int c = 0; c < NV; ++c) x.v[c] = q.ip * q.x.v[c] + acc[c];
    vstore<T, NV>(q.p, x);
  }
};

// dense cores: C[row] = sum_j Minv_c[row_local, j] z[core_j]  (rows of all cores)
template <typename T>
struct CoreKind {
  static constexpr int QM = 1;
  static constexpr bool SHARED_IDX = false;  // gathers hit a small, hot vector
  static constexpr bool COEF = true;
  const T* X;  // Z
  T* out;
  const uint32_t *row_ptr, *row_core, *cptr, *cidx;
  const uint64_t* moff;
  const T* mat;
  int kp;
  __device__ uint32_t rbegin(uint32_t r) const { return row_ptr[r]; }
  __device__ uint32_t rend(uint32_t r) const { return row_ptr[r + 1]; }
  __device__ void load(uint32_t r, uint32_t e, uint32_t* id, T& c) const {
    const uint32_t cc = row_core[r];
    const uint32_t b = cptr[cc], n = cptr[cc + 1] - b;
    const uint32_t j = e - row_ptr[r];
    id[0] = cidx[b + j];
    c = mat[moff[cc] + (uint64_t)(r - b) * n + j];
  }
  template <int NV>
  struct Pre {};
  template <int NV>
  __device__ Pre<NV> pre(uint32_t, int) const { return {}; }
  template <int NV>
  __device__ void finish(uint32_t r, int col, const T (&acc)[NV], const Pre<NV>&) const {
    vstore<T, NV>(out + (int64_t)r * kp + col, acc);
  }
};

// ---- dense cores ------------------------------------------------------------
// C[b + i] = sum_j M[i, j] Z[core_j] for a core of n rows (matrix padded to
// ld columns with zeros).  Z's rows are first gathered into Zc (ld x kp, rows
// beyond n zero).  kp >= 8: 128-row tiles through LDS, each thread holding a
// (kp/8) x 4 register block, the column range split in `split` parts whose
// partials are summed in part order; kp < 8: one wave per 4 rows with lanes
// along the row (a matrix-vector product).  Every sum has a fixed order.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_core_gather(const uint32_t* __restrict__ cidx, const T* __restrict__ X,
                                                        T* __restrict__ Zc, uint32_t n, uint32_t ld, int kp) {
  const int lane = threadIdx.x & (kp - 1);
  const int gpb = kBlock / kp;
  for (int64_t j = (int64_t)blockIdx.x * gpb + threadIdx.x / kp; j < ld; j += (int64_t)gridDim.x * gpb)
    Zc[j * kp + lane] = j < n ? X[(int64_t)cidx[j] * kp + lane] : T(0);
}

constexpr int kCoreRows = 128;  // rows per tile

template <typename T, int KP>
__global__ void __launch_bounds__(kBlock) k_core_tiled(const T* __restrict__ M, const T* __restrict__ Zc,
                                                       T* __restrict__ P, uint32_t n, uint32_t ld, uint32_t ks) {
  constexpr int NV = 16 / sizeof(T);  // values per 16-byte vector
  constexpr int TK = 128 / sizeof(T); // columns per step (128-byte row segments)
  constexpr int TR = kCoreRows, CG = KP / 4, RT = KP / 8;
  constexpr int VPR = TK / NV;                 // vectors per row segment
  constexpr int MV = TR * VPR / kBlock;        // M vectors per thread per step
  static_assert(TR * VPR % kBlock == 0, "tile shape");
  __shared__ T Ms[TK][TR + NV];
  __shared__ __align__(16) T Zs[TK * KP];
  const int tid = threadIdx.x;
  const int g = tid % CG, rg = tid / CG;
  const uint32_t r0 = blockIdx.x * TR;
  const uint32_t ka = blockIdx.y * ks, kb = min(ld, ka + ks);
  T acc[RT][4];
#pragma unroll
  for (int i = 0; i < RT; ++i)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[i][c] = 0;
  // the next M tile is loaded into registers while the current one is used
  VecT<T, NV> mr[MV];
  auto load_m = [&](uint32_t k0) {
#pragma unroll
    for (int u = 0; u < MV; ++u) {
      const int v = tid + u * kBlock;
      const int row = v / VPR, kv = (v % VPR) * NV;
      if (r0 + row < n) {
        mr[u] = vload<T, NV>(M + (size_t)(r0 + row) * ld + k0 + kv);
      } else {
#pragma unroll
        for (int c = 0; c < NV; ++c) mr[u].v[c] = 0;
      }
    }
  };
  if (ka < kb) load_m(ka);
  for (uint32_t k0 = ka; k0 < kb; k0 += TK) {
#pragma unroll
    for (int u = 0; u < MV; ++u) {
      const int v = tid + u * kBlock;
      const int row = v / VPR, kv = (v % VPR) * NV;
#pragma unroll
      for (int c = 0; c < NV; ++c) Ms[kv + c][row] = mr[u].v[c];
    }
    for (int v = tid; v < TK * KP / NV; v += kBlock)
      *reinterpret_cast<VecT<T, NV>*>(&Zs[v * NV]) = vload<T, NV>(Zc + (size_t)k0 * KP + v * NV);
    __syncthreads();
    if (k0 + TK < kb) load_m(k0 + TK);
#pragma unroll 4
    for (int kk = 0; kk < TK; ++kk) {
      T mv[RT], zv[4];
#pragma unroll
      for (int i = 0; i < RT; ++i) mv[i] = Ms[kk][rg * RT + i];
#pragma unroll
      for (int c = 0; c < 4; ++c) zv[c] = Zs[kk * KP + g * 4 + c];
#pragma unroll
      for (int i = 0; i < RT; ++i)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[i][c] += mv[i] * zv[c];
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < RT; ++i) {
    const uint32_t r = r0 + rg * RT + i;
    if (r < n)
#pragma unroll
      for (int c = 0; c < 4; ++c) P[((size_t)blockIdx.y * ld + r) * KP + g * 4 + c] = acc[i][c];
  }
}

// FP32, 32 columns: the same tiles on the matrix cores.  Each wave owns 32 of
// the tile's 128 rows and all 32 columns, accumulated by
// v_mfma_f32_32x32x2_f32 over pairs of k.  Operand layout (checked on gfx950):
// lane l supplies A[row l % 32][k l / 32] and B[k l / 32][column l % 32];
// result register r of lane l is D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].
// The core inverse is exactly symmetric, so the k-major tile Ms[k][row] =
// M[row][k] = M[k][row] is staged from rows k of M: 16-byte loads of
// contiguous row segments and 16-byte LDS stores; the row pitch (160 floats,
// 32 banks apart) keeps the two half-waves' operand reads on distinct banks.
typedef float f32x16 __attribute__((ext_vector_type(16)));
__global__ void __launch_bounds__(kBlock) k_core_mfma(const float* __restrict__ M, const float* __restrict__ Zc,
                                                      float* __restrict__ P, uint32_t n, uint32_t ld, uint32_t ks) {
  constexpr int KP = 32, NV = 4, TK = 32, TR = kCoreRows, VPK = TR / NV, MV = TK * VPK / kBlock;
  constexpr int MP = TR + 32;
  static_assert(TR == 4 * 32 && kBlock == 256, "four waves of 32 rows");
  __shared__ __align__(16) float Ms[TK][MP];
  __shared__ __align__(16) float Zs[TK * KP];
  const int tid = threadIdx.x, wv = tid / kWave, lane = tid & (kWave - 1);
  const uint32_t r0 = blockIdx.x * TR;
  const uint32_t ka = blockIdx.y * ks, kb = min(ld, ka + ks);
  f32x16 acc = {};
  VecT<float, NV> mr[MV];
  auto load_m = [&](uint32_t k0) {
#pragma unroll
    for (int u = 0; u < MV; ++u) {
      const int v = tid + u * kBlock;
      const int kk = v / VPK, rv = (v % VPK) * NV;
      if (k0 + kk < n && r0 + rv < ld) {
        mr[u] = vload<float, NV>(M + (size_t)(k0 + kk) * ld + r0 + rv);
      } else {
#pragma unroll
        for (int c = 0; c < NV; ++c) mr[u].v[c] = 0;
      }
    }
  };
  if (ka < kb) load_m(ka);
  const int ar = wv * 32 + lane % 32, hk = lane / 32, bc = lane % 32;
  for (uint32_t k0 = ka; k0 < kb; k0 += TK) {
#pragma unroll
    for (int u = 0; u < MV; ++u) {
      const int v = tid + u * kBlock;
      const int kk = v / VPK, rv = (v % VPK) * NV;
      *reinterpret_cast<VecT<float, NV>*>(&Ms[kk][rv]) = mr[u];
    }
    for (int v = tid; v < TK * KP / NV; v += kBlock)
      *reinterpret_cast<VecT<float, NV>*>(&Zs[v * NV]) = vload<float, NV>(Zc + (size_t)k0 * KP + v * NV);
    __syncthreads();
    if (k0 + TK < kb) load_m(k0 + TK);
#pragma unroll
    for (int kk = 0; kk < TK; kk += 2)
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(Ms[kk + hk][ar], Zs[(kk + hk) * KP + bc], acc, 0, 0, 0);
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const uint32_t row = r0 + wv * 32 + 8 * (r / 4) + 4 * hk + (r % 4);
    if (row < n) P[((size_t)blockIdx.y * ld + row) * KP + bc] = acc[r];
  }
}

// C = sum of the split partials, in split order; 16-byte vectors, all of a
// vector's partial loads in flight together (n * kp is a multiple of NV)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_core_sum(const T* __restrict__ P, T* __restrict__ C, uint32_t n,
                                                     uint32_t ld, int split, int kp) {
  constexpr int NV = 16 / sizeof(T), U = 8;
  const int64_t tot = (int64_t)n * kp / NV, stride = (int64_t)ld * kp;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < tot; i += (int64_t)gridDim.x * kBlock) {
    VecT<T, NV> a;
#pragma unroll
    for (int c = 0; c < NV; ++c) a.v[c] = 0;
    int s = 0;
    for (; s + U <= split; s += U) {
      VecT<T, NV> v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) v[u] = vload<T, NV>(P + (s + u) * stride + i * NV);
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int c = 0; c < NV; ++c) a.v[c] += v[u].v[c];
    }
    for (; s < split; ++s) {
      const VecT<T, NV> v = vload<T, NV>(P + s * stride + i * NV);
#pragma unroll
      for (int c = 0; c < NV; ++c) a.v[c] += v.v[c];
    }
    vstore<T, NV>(C + i * NV, a);
  }
}

template <typename T, int KP>
__global__ void __launch_bounds__(kBlock) k_core_gemv(const T* __restrict__ M, const T* __restrict__ Zc,
                                                      T* __restrict__ C, uint32_t n, uint32_t ld) {
  constexpr int NV = 16 / sizeof(T), R = 4;
  const int lane = threadIdx.x & (kWave - 1);
  const uint32_t r0 = (uint32_t)(((int64_t)blockIdx.x * kBlock + threadIdx.x) / kWave) * R;
  if (r0 >= n) return;
  T acc[R][KP];
#pragma unroll
  for (int i = 0; i < R; ++i)
#pragma unroll
    for (int c = 0; c < KP; ++c) acc[i][c] = 0;
  for (uint32_t k = lane * NV; k < ld; k += kWave * NV) {
    VecT<T, NV> mv[R];
#pragma unroll
    for (int i = 0; i < R; ++i) {
      if (r0 + i < n) {
        mv[i] = vload<T, NV>(M + (size_t)(r0 + i) * ld + k);
      } else {
#pragma unroll
        for (int u = 0; u < NV; ++u) mv[i].v[u] = 0;
      }
    }
#pragma unroll
    for (int u = 0; u < NV; ++u) {
      T zv[KP];
#pragma unroll
      for (int c = 0; c < KP; ++c) zv[c] = Zc[(size_t)(k + u) * KP + c];
#pragma unroll
      for (int i = 0; i < R; ++i)
#pragma unroll
        for (int c = 0; c < KP; ++c) acc[i][c] += mv[i].v[u] * zv[c];
    }
  }
#pragma unroll
  for (int i = 0; i < R; ++i)
#pragma unroll
    for (int c = 0; c < KP; ++c)
#pragma unroll
      for (int off = kWave / 2; off >= 1; off >>= 1) acc[i][c] += __shfl_xor(acc[i][c], off, kWave);
  if (lane == 0)
#pragma unroll
    for (int i = 0; i < R; ++i)
      if (r0 + i < n)
#pragma unroll
        for (int c = 0; c < KP; ++c) C[(size_t)(r0 + i) * KP + c] = acc[i][c];
}

// column split for a core: about eight workgroups per CU
static inline int core_split(uint32_t n, uint32_t ld, int tk) {
  const int tiles = (int)((n + kCoreRows - 1) / kCoreRows);
  int s = std::max(1, std::min(kCoreSplitMax, 2048 / std::max(tiles, 1)));
  return std::min<int>(s, (int)(ld / tk));
}

template <typename T>
static void run_dense_core(hipStream_t st, const DenseCore<T>& d, const uint32_t* core_idx, const T* X, T* C,
                           T* Zc, T* P, int kp) {
  k_core_gather<T><<<grid_stride_blocks(d.ld, kp), kBlock, 0, st>>>(core_idx + d.b, X, Zc, d.n, d.ld, kp);
  T* out = C + (size_t)d.b * kp;
  if (kp >= 8) {
    constexpr int TK = 128 / sizeof(T);
    const int split = core_split(d.n, d.ld, TK);
    const uint32_t ks = (uint32_t)(((d.ld / TK + split - 1) / split) * TK);
    const int nsplit = (int)((d.ld + ks - 1) / ks);
    dim3 grid((d.n + kCoreRows - 1) / kCoreRows, nsplit);
    switch (kp) {
      case 8: k_core_tiled<T, 8><<<grid, kBlock, 0, st>>>(d.M, Zc, P, d.n, d.ld, ks); break;
      case 16: k_core_tiled<T, 16><<<grid, kBlock, 0, st>>>(d.M, Zc, P, d.n, d.ld, ks); break;
      default:
        if constexpr (std::is_same<T, float>::value)
          k_core_mfma<<<grid, kBlock, 0, st>>>(d.M, Zc, P, d.n, d.ld, ks);
        else
          k_core_tiled<T, 32><<<grid, kBlock, 0, st>>>(d.M, Zc, P, d.n, d.ld, ks);
        break;
    }
    k_core_sum<T><<<grid_stride_blocks((int64_t)d.n * kp / (16 / sizeof(T)), 1), kBlock, 0, st>>>(P, out, d.n, d.ld,
                                                                                                 nsplit, kp);
  } else {
    const unsigned blocks = (unsigned)(((d.n + 3) / 4 * (int64_t)kWave + kBlock - 1) / kBlock);
    switch (kp) {
      case 1: k_core_gemv<T, 1><<<blocks, kBlock, 0, st>>>(d.M, Zc, out, d.n, d.ld); break;
      case 2: k_core_gemv<T, 2><<<blocks, kBlock, 0, st>>>(d.M, Zc, out, d.n, d.ld); break;
      default: k_core_gemv<T, 4><<<blocks, kBlock, 0, st>>>(d.M, Zc, out, d.n, d.ld); break;
    }
  }
}

template <typename T>
void launch_precond(hipStream_t st, const DevPrecond<T>& p, const T* R, T* Z, PrecondWork<T>& w, int kp,
                    int64_t m) {
  // 1. restrict (signed, partition-weighted) and per-subdomain sums -> mu1
  run_rows<RestrictSegKind<T>, T>(st, RestrictSegKind<T>{R, w.B, w.sum1, p.loc_ptr, p.loc_gidx, p.loc_scale, kp},
                                  p.s_loc, w.part, kp);
  if (p.n_red > 0) {
    // 2. Schur right-hand side, centred (mu2)
    run_rows<KeepKind<T>, T>(st,
                             KeepKind<T>{w.B, w.Rr, p.keep_ptr, p.keep_idx, p.keep_val, p.keep_rowsum, p.red_loc,
                                         p.red_sub, p.loc_ptr, w.sum1, kp},
                             p.s_keep, w.part, kp);
    run_rows<SegKind<T>, T>(st, SegKind<T>{w.Rr, w.sum2, p.red_ptr, kp}, p.s_red, w.part, kp);
    launch_sub_mean<T>(st, 0, p.n_red, p.red_sub, p.red_ptr, w.sum2, w.Rr, kp);
    // 3. reduced solve: forward levels, dense cores, backward levels
    FwdKind<T> fk{w.Rr, p.lvl_rows, p.fin_ptr, p.fin_idx, p.fin_val, kp};
    run_levels<FwdKind<T>, T>(st, fk, p.s_fwd, 0, (int)p.s_fwd.segs.size(), 1, w.part, kp);
    if (p.n_core_rows > 0) {
      run_rows<CoreKind<T>, T>(st,
                               CoreKind<T>{w.Rr, w.C, p.core_row_ptr, p.core_row_core, p.core_ptr, p.core_idx,
                                           p.core_mat_off, p.core_mat, kp},
                               p.s_core, w.part, kp);
      for (const DenseCore<T>& d : p.big) run_dense_core<T>(st, d, p.core_idx, w.Rr, w.C, w.Zc, w.P, kp);
      launch_scatter_rows<T>(st, 0, p.n_core_rows, p.core_idx, w.C, w.Rr, kp);
    }
    BwdKind<T> bk{w.Rr, p.lvl_rows, p.bout_ptr, p.bout_idx, p.bout_val, p.inv_pivot, kp};
    // the last level gathers the dense cores: no backward step
    run_levels<BwdKind<T>, T>(st, bk, p.s_bwd, 0, (int)p.s_bwd.segs.size() - 1, -1, w.part, kp);
    // 4. kept block copied into the local solution
    launch_scatter_rows<T>(st, 0, p.n_red, p.red_loc, w.Rr, w.S, kp);
  }
  // 5. back-substitution for the eliminated block
  if (p.n_elim > 0)
    run_rows<ElimKind<T>, T>(st,
                             ElimKind<T>{w.Rr, w.S, p.elim_ptr, p.elim_idx, p.elim_val, p.elim_inv, p.elim_loc,
                                         p.loc_sub, p.loc_ptr, w.B, w.sum1, kp},
                             p.s_elim, w.part, kp);
  // 6. centre (mu3), sign and partition weights, combine per DOF
  run_rows<SegKind<T>, T>(st, SegKind<T>{w.S, w.sum3, p.loc_ptr, kp}, p.s_loc, w.part, kp);
  launch_combine<T>(st, m, p.dof_ptr, p.dof_loc, p.loc_scale, p.loc_sub, p.loc_ptr, w.S, w.sum3, Z, kp);
}
template void launch_precond<float>(hipStream_t, const DevPrecond<float>&, const float*, float*,
                                    PrecondWork<float>&, int, int64_t);
template void launch_precond<double>(hipStream_t, const DevPrecond<double>&, const double*, double*,
                                     PrecondWork<double>&, int, int64_t);

}  // namespace wa

namespace wa {

// Y[:, c] += coef[c] * X[:, c]     (coef per column, FP64; sign applied)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_axpy_col(T* __restrict__ Y, const T* __restrict__ X,
                                                     const double* __restrict__ coef, double sign, int64_t m,
                                                     int kp) {
  const int lane = threadIdx.x & (kp - 1);
  const T a = (T)(sign * coef[lane]);
  const int gpb = kBlock / kp;
#pragma unroll 4
  for (int64_t i = (int64_t)blockIdx.x * gpb + threadIdx.x / kp; i < m; i += (int64_t)gridDim.x * gpb)
    Y[i * kp + lane] += a * X[i * kp + lane];
}
template <typename T>
void launch_axpy_col(hipStream_t st, T* Y, const T* X, const double* coef, double sign, int64_t m, int kp) {
  k_axpy_col<T><<<std::max<unsigned>(1, std::min<unsigned>((unsigned)((m * kp + kBlock - 1) / kBlock), kRedBlocks)),
                  kBlock, 0, st>>>(Y, X, coef, sign, m, kp);
}
template void launch_axpy_col<float>(hipStream_t, float*, const float*, const double*, double, int64_t, int);
template void launch_axpy_col<double>(hipStream_t, double*, const double*, const double*, double, int64_t, int);

// Y[:, c] = coef[c] * X[:, c]
template <typename T>
__global__ void __launch_bounds__(kBlock) k_scale_col(T* __restrict__ Y, const T* __restrict__ X,
                                                      const double* __restrict__ coef, int64_t m, int kp) {
  const int lane = threadIdx.x & (kp - 1);
  const T a = (T)coef[lane];
  const int gpb = kBlock / kp;
#pragma unroll 4
  for (int64_t i = (int64_t)blockIdx.x * gpb + threadIdx.x / kp; i < m; i += (int64_t)gridDim.x * gpb)
    Y[i * kp + lane] = a * X[i * kp + lane];
}
template <typename T>
void launch_scale_col(hipStream_t st, T* Y, const T* X, const double* coef, int64_t m, int kp) {
  k_scale_col<T><<<std::max<unsigned>(1, std::min<unsigned>((unsigned)((m * kp + kBlock - 1) / kBlock), kRedBlocks)),
                   kBlock, 0, st>>>(Y, X, coef, m, kp);
}
template void launch_scale_col<float>(hipStream_t, float*, const float*, const double*, int64_t, int);
template void launch_scale_col<double>(hipStream_t, double*, const double*, const double*, int64_t, int);

// R = B - R  (true residual from R = A x)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_rsub(T* __restrict__ R, const T* __restrict__ B, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    R[i] = B[i] - R[i];
}
template <typename T>
void launch_rsub(hipStream_t st, T* R, const T* B, int64_t n) {
  k_rsub<T><<<std::max<unsigned>(1, std::min<unsigned>((unsigned)((n + kBlock - 1) / kBlock), 65536)), kBlock, 0,
              st>>>(R, B, n);
}
template void launch_rsub<float>(hipStream_t, float*, const float*, int64_t);
template void launch_rsub<double>(hipStream_t, double*, const double*, int64_t);

}  // namespace wa

namespace wa {

// Local solves of one factor pair's subdomains, restricted to the pair's
// rows, in three stages: the centred Schur right-hand side, the reduced solve
// (forward levels, dense cores, backward levels) and the back-substitution of
// the eliminated block.  `part`, `Zc` and `P` are the split-row partials and
// dense-core scratch of the stream the stage runs on.  Each launch (or
// dependent group of launches) is handed to `emit`, which runs it at once or
// queues it for interleaved submission.
struct RunNow {
  template <class F>
  void operator()(F&& f) const { f(); }
};

template <typename T, class Emit>
static void pair_schur_rhs(Emit& emit, hipStream_t st, const DevPrecond<T>& p, const PairSched& ps,
                           PrecondWork<T>& w, T* part, int kp) {
  if (ps.red1 == ps.red0) return;
  emit([=, &p, &ps, &w] {
    run_rows<KeepKind<T>, T>(st,
                             KeepKind<T>{w.B, w.Rr, p.keep_ptr, p.keep_idx, p.keep_val, p.keep_rowsum, p.red_loc,
                                         p.red_sub, p.loc_ptr, w.sum1, kp},
                             ps.s_keep, part, kp);
  });
  emit([=, &p, &ps, &w] { run_rows<SegKind<T>, T>(st, SegKind<T>{w.Rr, w.sum2, p.red_ptr, kp}, ps.s_red, part, kp); });
  emit([=, &p, &ps, &w] { launch_sub_mean<T>(st, ps.red0, ps.red1, p.red_sub, p.red_ptr, w.sum2, w.Rr, kp); });
}

// clears the level-run counters of one apply; first checks that the previous
// apply's runs handed out every unit (a run whose workgroups all missed XCD 0
// would leave its rows untouched)
// (system-scope atomics throughout, like the runs that use the counters)
__global__ void k_flow_reset(uint32_t* ctr_f, int nf, uint32_t* ctr_b, int nb, uint32_t prev_f, uint32_t prev_b,
                             int check) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  auto get = [](uint32_t* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM); };
  auto put = [](uint32_t* p, uint32_t v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM); };
  if (check && (get(&ctr_f[0]) < prev_f || get(&ctr_b[0]) < prev_b))
    __hip_atomic_fetch_or(&ctr_f[1], 2u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  put(&ctr_f[0], 0);
  for (int l = 0; l < nf; ++l) put(&ctr_f[2 + l], 0);
  put(&ctr_b[0], 0);
  for (int l = 0; l < nb; ++l) put(&ctr_b[2 + l], 0);
}

template <typename T>
void flow_totals(const PairSched& ps, int kp, uint32_t& tot_f, uint32_t& tot_b) {
  tot_f = tot_b = 0;
  for (int t = ps.flow_t0; t < ps.flow_t1; ++t) {
    const Segment& gf = ps.s_fwd.segs[t];
    const Segment& gb = ps.s_bwd.segs[t];
    tot_f += flow_units<T>(FlowLevel{gf.r0, gf.r1, gf.i0, gf.i1}, kp);
    tot_b += flow_units<T>(FlowLevel{gb.r0, gb.r1, gb.i0, gb.i1}, kp);
  }
}
template void flow_totals<float>(const PairSched&, int, uint32_t&, uint32_t&);
template void flow_totals<double>(const PairSched&, int, uint32_t&, uint32_t&);

template <typename T, class Emit>
static void pair_reduced(Emit& emit, hipStream_t st, const DevPrecond<T>& p, const PairSched& ps,
                         PrecondWork<T>& w, T* part, T* Zc, T* P, int kp, const FlowCall* flow = nullptr) {
  if (ps.red1 == ps.red0) return;
  const bool fl = flow && ps.flow_t0 >= 0;
  const int nfl = fl ? ps.flow_t1 - ps.flow_t0 : 0;
  const FwdKind<T> fk{w.Rr, p.lvl_rows, p.fin_ptr, p.fin_idx, p.fin_val, kp};
  for (size_t t = 0; t < ps.s_fwd.segs.size(); ++t) {
    if (fl && (int)t >= ps.flow_t0 && (int)t < ps.flow_t1) {
      if ((int)t == ps.flow_t0) {
        const FlowCall fc = *flow;
        emit([=, &ps] {
          k_flow_reset<<<1, 64, 0, st>>>(fc.ctr_f, nfl, fc.ctr_b, nfl, fc.prev_f, fc.prev_b, fc.check);
          run_flow<FwdKind<T>, T>(st, fk, ps.s_fwd, ps.flow_fwd, nfl, part, fc.ctr_f, kp, fc.spin);
        });
      }
      continue;
    }
    const Segment& g = ps.s_fwd.segs[t];
    if (g.r1 > g.r0 || g.i1 > g.i0) emit([=, &ps] { run_seg<FwdKind<T>, T>(st, fk, ps.s_fwd, g, part, kp); });
  }
  if (ps.core1 > ps.core0) {
    emit([=, &p, &ps, &w] {
      run_rows<CoreKind<T>, T>(st,
                               CoreKind<T>{w.Rr, w.C, p.core_row_ptr, p.core_row_core, p.core_ptr, p.core_idx,
                                           p.core_mat_off, p.core_mat, kp},
                               ps.s_core, part, kp);
      for (size_t bi : ps.big) run_dense_core<T>(st, p.big[bi], p.core_idx, w.Rr, w.C, Zc, P, kp);
      launch_scatter_rows<T>(st, ps.core0, ps.core1, p.core_idx, w.C, w.Rr, kp);
    });
  }
  const BwdKind<T> bk{w.Rr, p.lvl_rows, p.bout_ptr, p.bout_idx, p.bout_val, p.inv_pivot, kp};
  // the last level gathers the dense cores: no backward step
  for (size_t step = 0; step + 1 < ps.s_bwd.segs.size(); ++step) {
    const int t = (int)ps.s_bwd.segs.size() - 2 - (int)step;
    if (fl && t >= ps.flow_t0 && t < ps.flow_t1) {
      if (t == ps.flow_t1 - 1) {
        const FlowCall fc = *flow;
        emit([=, &ps] { run_flow<BwdKind<T>, T>(st, bk, ps.s_bwd, ps.flow_bwd, nfl, part, fc.ctr_b, kp, fc.spin); });
      }
      continue;
    }
    const Segment& g = ps.s_bwd.segs[t];
    if (g.r1 > g.r0 || g.i1 > g.i0) emit([=, &ps] { run_seg<BwdKind<T>, T>(st, bk, ps.s_bwd, g, part, kp); });
  }
  emit([=, &p, &ps, &w] { launch_scatter_rows<T>(st, ps.red0, ps.red1, p.red_loc, w.Rr, w.S, kp); });
}

template <typename T, class Emit>
static void pair_elim(Emit& emit, hipStream_t st, const DevPrecond<T>& p, const PairSched& ps, PrecondWork<T>& w,
                      T* part, int kp) {
  emit([=, &p, &ps, &w] {
    run_rows<ElimKind<T>, T>(st,
                             ElimKind<T>{w.Rr, w.S, p.elim_ptr, p.elim_idx, p.elim_val, p.elim_inv, p.elim_loc,
                                         p.loc_sub, p.loc_ptr, w.B, w.sum1, kp},
                             ps.s_elim, part, kp);
  });
}

template <typename T>
static void pair_solve(hipStream_t st, const DevPrecond<T>& p, const PairSched& ps, PrecondWork<T>& w, T* part,
                       T* Zc, T* P, int kp) {
  RunNow now;
  pair_schur_rhs<T>(now, st, p, ps, w, part, kp);
  pair_reduced<T>(now, st, p, ps, w, part, Zc, P, kp);
  pair_elim<T>(now, st, p, ps, w, part, kp);
}

// Additive Schwarz with the factor pairs on two streams: the pair with the
// longest level chain (`lead`, latency bound: one short launch per
// elimination level) on `st`, the other pairs beside it on `st2`.  Each
// stream restricts, solves and sums its own pairs' subdomains; the combine
// waits for both.  Every row is computed by the same kernel code as in
// launch_precond and the pairs touch disjoint local rows, so the result is
// bitwise that of launch_precond.  The other pairs start once the lead's
// bandwidth-bound first stage (restriction, Schur right-hand side) is done,
// so that they run beside its level chain rather than slowing that stage.
// `side` (may be null) is more work for st2 after those pairs, still before
// the combine: independent work of the caller that can fill the chain's idle
// bandwidth.  The two streams' launches are submitted interleaved, in
// proportion to their counts, so that neither chain waits for the host to
// finish enqueueing the other.
template <typename T>
void launch_precond_overlap(hipStream_t st, hipStream_t st2, hipEvent_t fork, hipEvent_t join,
                            const DevPrecond<T>& p, const std::vector<PairSched>& pairs, size_t lead, const T* R,
                            T* Z, PrecondWork<T>& w, T* part2, T* Zc2, T* P2, int kp, int64_t m,
                            const SideWork* side, const FlowCall* flow) {
  std::vector<std::function<void()>> qa, qb;
  auto to_a = [&](auto&& f) { qa.emplace_back(std::forward<decltype(f)>(f)); };
  auto to_b = [&](auto&& f) { qb.emplace_back(std::forward<decltype(f)>(f)); };
  auto restrict_pair = [&](hipStream_t s, const PairSched& ps, T* part) {
    run_rows<RestrictSegKind<T>, T>(s, RestrictSegKind<T>{R, w.B, w.sum1, p.loc_ptr, p.loc_gidx, p.loc_scale, kp},
                                    ps.s_loc, part, kp);
  };
  auto centre_pair = [&](hipStream_t s, const PairSched& ps, T* part) {
    run_rows<SegKind<T>, T>(s, SegKind<T>{w.S, w.sum3, p.loc_ptr, kp}, ps.s_loc, part, kp);
  };
  const PairSched& a = pairs[lead];
  RunNow now;
  restrict_pair(st, a, w.part);
  pair_schur_rhs<T>(now, st, p, a, w, w.part, kp);
  (void)hipEventRecord(fork, st);
  (void)hipStreamWaitEvent(st2, fork, 0);
  pair_reduced<T>(to_a, st, p, a, w, w.part, w.Zc, w.P, kp, flow);
  pair_elim<T>(to_a, st, p, a, w, w.part, kp);
  to_a([&] { centre_pair(st, a, w.part); });
  for (size_t i = 0; i < pairs.size(); ++i) {
    if (i == lead) continue;
    const PairSched& b = pairs[i];
    to_b([&, i] { restrict_pair(st2, pairs[i], part2); });
    pair_schur_rhs<T>(to_b, st2, p, b, w, part2, kp);
    pair_reduced<T>(to_b, st2, p, b, w, part2, Zc2, P2, kp);
    pair_elim<T>(to_b, st2, p, b, w, part2, kp);
    to_b([&, i] { centre_pair(st2, pairs[i], part2); });
  }
  if (side) to_b([&] { side->run(st2, side->ctx); });
  const size_t per = std::max<size_t>(1, qa.size() / std::max<size_t>(1, qb.size()));
  size_t ia = 0, ib = 0;
  while (ia < qa.size() || ib < qb.size()) {
    for (size_t u = 0; u < per && ia < qa.size(); ++u) qa[ia++]();
    if (ib < qb.size()) {
      qb[ib++]();
      if (ib == qb.size()) (void)hipEventRecord(join, st2);
    }
  }
  if (qb.empty()) (void)hipEventRecord(join, st2);
  (void)hipStreamWaitEvent(st, join, 0);
  launch_combine<T>(st, m, p.dof_ptr, p.dof_loc, p.loc_scale, p.loc_sub, p.loc_ptr, w.S, w.sum3, Z, kp);
}
template void launch_precond_overlap<float>(hipStream_t, hipStream_t, hipEvent_t, hipEvent_t,
                                            const DevPrecond<float>&, const std::vector<PairSched>&, size_t,
                                            const float*, float*, PrecondWork<float>&, float*, float*, float*, int,
                                            int64_t, const SideWork*, const FlowCall*);
template void launch_precond_overlap<double>(hipStream_t, hipStream_t, hipEvent_t, hipEvent_t,
                                             const DevPrecond<double>&, const std::vector<PairSched>&, size_t,
                                             const double*, double*, PrecondWork<double>&, double*, double*, double*,
                                             int, int64_t, const SideWork*, const FlowCall*);

// prolongation of one pair (its subdomains cover distinct DOFs):
//   d = scale_l (S[l] - mu3); Z[g_l] += d; delta[g_l] = d
template <typename T>
__global__ void __launch_bounds__(kBlock) k_prolong_pair(int64_t b0, int64_t n, const uint32_t* __restrict__ gidx,
                                                         const T* __restrict__ scale,
                                                         const uint32_t* __restrict__ loc_sub,
                                                         const uint32_t* __restrict__ loc_ptr,
                                                         const T* __restrict__ S, const T* __restrict__ sum3,
                                                         T* __restrict__ Z, T* __restrict__ delta, int kp) {
  const int lane = threadIdx.x & (kp - 1);
  const int gpb = kBlock / kp;
  for (int64_t l = b0 + (int64_t)blockIdx.x * gpb + threadIdx.x / kp; l < n; l += (int64_t)gridDim.x * gpb) {
    const uint32_t s = loc_sub[l];
    const T mu = sum3[(int64_t)s * kp + lane] / (T)(loc_ptr[s + 1] - loc_ptr[s]);
    const T d = scale[l] * (S[l * kp + lane] - mu);
    const int64_t o = (int64_t)gidx[l] * kp + lane;
    Z[o] += d;
    delta[o] = d;
  }
}

// Multiplicative Schwarz, forward sweep (schwarz/multiplicative.rs:128-245):
// per factor pair (its components are mutually independent, so processing
// them together equals the sequential sweep), restrict the working residual,
// solve locally, add the correction to z and update r_work -= G delta with
// the explicit Gramian (operator/residual_update.rs:36-53).
template <typename T>
void launch_precond_mult(hipStream_t st, const DevPrecond<T>& p, const std::vector<PairSched>& pairs,
                         const DevCsr<T>& g, const T* R, T* Z, PrecondWork<T>& w, T* rwork, T* delta, T* gd,
                         const double* ones, int kp, int64_t m) {
  const size_t vbytes = sizeof(T) * (size_t)m * kp;
  (void)hipMemcpyAsync(rwork, R, vbytes, hipMemcpyDeviceToDevice, st);
  (void)hipMemsetAsync(Z, 0, vbytes, st);
  for (const PairSched& ps : pairs) {
    const int64_t nl = ps.loc1 - ps.loc0;
    run_rows<RestrictSegKind<T>, T>(
        st, RestrictSegKind<T>{rwork, w.B, w.sum1, p.loc_ptr, p.loc_gidx, p.loc_scale, kp}, ps.s_loc, w.part, kp);
    pair_solve<T>(st, p, ps, w, w.part, w.Zc, w.P, kp);
    run_rows<SegKind<T>, T>(st, SegKind<T>{w.S, w.sum3, p.loc_ptr, kp}, ps.s_loc, w.part, kp);
    (void)hipMemsetAsync(delta, 0, vbytes, st);
    k_prolong_pair<T><<<grid_stride_blocks(nl, kp), kBlock, 0, st>>>(ps.loc0, ps.loc1, p.loc_gidx, p.loc_scale,
                                                                     p.loc_sub, p.loc_ptr, w.S, w.sum3, Z, delta,
                                                                     kp);
    launch_spmm<T>(st, g, delta, gd, w.part, kp);
    launch_axpy_col<T>(st, rwork, gd, ones, -1.0, m, kp);
  }
}
template void launch_precond_mult<float>(hipStream_t, const DevPrecond<float>&, const std::vector<PairSched>&,
                                         const DevCsr<float>&, const float*, float*, PrecondWork<float>&, float*,
                                         float*, float*, const double*, int, int64_t);
template void launch_precond_mult<double>(hipStream_t, const DevPrecond<double>&, const std::vector<PairSched>&,
                                          const DevCsr<double>&, const double*, double*, PrecondWork<double>&,
                                          double*, double*, double*, const double*, int, int64_t);

template <typename T>
__global__ void k_pad_core(const double* __restrict__ src, uint32_t n, T* __restrict__ dst, uint32_t ld) {
  const int64_t tot = (int64_t)n * ld;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t / ld, j = t % ld;
    dst[t] = j < n ? (T)src[i * n + j] : (T)0;
  }
}
template <typename T>
void launch_pad_core(hipStream_t st, const double* src, uint32_t n, T* dst, uint32_t ld) {
  const int64_t tot = (int64_t)n * ld;
  if (tot <= 0) return;
  k_pad_core<T><<<(unsigned)std::min<int64_t>((tot + 255) / 256, 65536), 256, 0, st>>>(src, n, dst, ld);
}
template void launch_pad_core<float>(hipStream_t, const double*, uint32_t, float*, uint32_t);

__global__ void k_narrow(const double* __restrict__ src, float* __restrict__ dst, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = (float)src[i];
}
void launch_narrow(hipStream_t st, const double* src, float* dst, int64_t n) {
  if (n <= 0) return;
  k_narrow<<<(unsigned)std::min<int64_t>((n + 255) / 256, 65536), 256, 0, st>>>(src, dst, n);
}
template void launch_pad_core<double>(hipStream_t, const double*, uint32_t, double*, uint32_t);

}  // namespace wa
