// multi.hip -- block CG over many independent fixed-effects problems stacked
// block-diagonally into one design.  Each problem's DOFs are covered by row
// tiles that hold that problem only; every dot product is summed per tile and
// then per (problem, column) pair over that problem's tiles in a fixed order,
// so each pair keeps its own alpha, beta and convergence mask and follows
// exactly the CG recurrence of a solve of that problem alone (cg.rs:45-133
// for each right-hand side of Solver::solve_batch, solver.rs:242-277).  No
// float atomics: every sum has a fixed order.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.hpp"
#include "multi.hpp"

namespace wa {

constexpr int kSegBlock = 256;
constexpr int kPairBlock = 256;

// part[t * kp + c] = sum over the rows of tile t of A[i, c] * B[i, c]
// (B == nullptr: A[i, c]^2), accumulated in FP64 and reduced in a fixed tree
template <typename T>
__global__ void __launch_bounds__(kSegBlock) k_seg_dot(const T* __restrict__ A, const T* __restrict__ B,
                                                       const SegTile* __restrict__ tiles, int kp,
                                                       double* __restrict__ part) {
  __shared__ double sh[kSegBlock];
  const SegTile tl = tiles[blockIdx.x];
  const int c = threadIdx.x % kp, g = threadIdx.x / kp, G = kSegBlock / kp;
  double acc = 0.0;
  for (int64_t i = (int64_t)tl.r0 + g; i < tl.r1; i += G) {
    const double a = (double)A[i * kp + c];
    acc += B ? a * (double)B[i * kp + c] : a * a;
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = G / 2; w >= 1; w >>= 1) {
    if (g < w) sh[threadIdx.x] += sh[threadIdx.x + w * kp];
    __syncthreads();
  }
  if (g == 0) part[(int64_t)blockIdx.x * kp + c] = sh[c];
}

// total of pair (p, c) over problem p's tiles, in tile order
__device__ inline double pair_total(const double* __restrict__ part, const SegIndex& ix, int p, int c, int kp) {
  double s = 0.0;
  for (uint32_t j = ix.pt_ptr[p]; j < ix.pt_ptr[p + 1]; ++j) s += part[(int64_t)ix.pt_idx[j] * kp + c];
  return s;
}

__global__ void k_seg_totals(const double* __restrict__ part, SegIndex ix, int kp, double* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ix.P * kp) return;
  out[j] = pair_total(part, ix, j / kp, j % kp, kp);
}

__global__ void k_seg_start(ColState* __restrict__ s, const double* __restrict__ part_bb,
                            const double* __restrict__ part_rz, SegIndex ix, const double* __restrict__ tol,
                            const int* __restrict__ kcols, int kp) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ix.P * kp) return;
  const int p = j / kp, c = j % kp;
  ColState& q = s[j];
  const double bn = sqrt(pair_total(part_bb, ix, p, c, kp));
  const double rz = pair_total(part_rz, ix, p, c, kp);
  q.bnorm = bn;
  q.rnorm = bn;
  q.tol = tol[j];
  q.rz = rz;
  q.rz0 = rz;
  q.iters = 0;
  q.alpha = 0.0;
  q.beta = 0.0;
  q.nan_stop = 0;
  if (c >= kcols[p] || bn < 2.220446049250313e-16) {  // padding column or zero right-hand side
    q.active = 0;
    q.converged = 1;
    return;
  }
  q.active = 1;
  q.converged = 0;
}

__global__ void k_seg_alpha(ColState* __restrict__ s, const double* __restrict__ part, SegIndex ix,
                            double* __restrict__ alpha, int kp) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ix.P * kp) return;
  ColState& q = s[j];
  alpha[j] = 0.0;
  if (!q.active) return;
  q.iters += 1;  // the iteration count lives on the device (graph replays keep it)
  const double v = pair_total(part, ix, j / kp, j % kp, kp);
  if (v <= 0.0) {  // cg.rs:80-86: p'Ap <= 0 stops the column, not converged
    q.active = 0;
    q.converged = 0;
    return;
  }
  alpha[j] = q.rz / v;
  q.alpha = alpha[j];
}

__global__ void k_seg_check(ColState* __restrict__ s, const double* __restrict__ part, SegIndex ix, int kp) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ix.P * kp) return;
  ColState& q = s[j];
  if (!q.active) return;
  q.rnorm = sqrt(pair_total(part, ix, j / kp, j % kp, kp));
  if (q.rnorm / q.bnorm <= q.tol) {
    q.active = 0;
    q.converged = 1;
  } else if (isnan(q.rnorm)) {
    q.active = 0;
    q.converged = 0;
    q.nan_stop = 1;
  }
}

__global__ void k_seg_beta(ColState* __restrict__ s, const double* __restrict__ part, SegIndex ix,
                           double* __restrict__ beta, int* __restrict__ mask, int kp) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ix.P * kp) return;
  ColState& q = s[j];
  beta[j] = 0.0;
  mask[j] = 0;
  if (!q.active) return;
  const double v = pair_total(part, ix, j / kp, j % kp, kp);
  const double thr = 4.930380657631324e-32 * fmax(fabs(q.rz0), 2.2250738585072014e-308);
  if (fabs(v) < thr) {  // cg.rs:106-119 stagnation guard
    q.active = 0;
    q.converged = q.rnorm / q.bnorm <= q.tol;
  } else {
    beta[j] = v / q.rz;
    q.rz = v;
    mask[j] = 1;
  }
}

// number of active pairs (one workgroup, fixed tree)
__global__ void __launch_bounds__(1024) k_seg_count(const ColState* __restrict__ s, int npairs,
                                                    int* __restrict__ n_active) {
  __shared__ int sh[1024];
  int a = 0;
  for (int j = threadIdx.x; j < npairs; j += 1024) a += s[j].active;
  sh[threadIdx.x] = a;
  __syncthreads();
  for (int w = 512; w >= 1; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *n_active = sh[0];
}

// x += alpha p (x may be null), r -= alpha ap; part = per-tile sums of r^2
template <typename T>
__global__ void __launch_bounds__(kSegBlock) k_seg_update(T* __restrict__ x, const T* __restrict__ p,
                                                          T* __restrict__ r, const T* __restrict__ ap,
                                                          const double* __restrict__ alpha,
                                                          const SegTile* __restrict__ tiles, int kp,
                                                          double* __restrict__ part) {
  __shared__ double sh[kSegBlock];
  const SegTile tl = tiles[blockIdx.x];
  const int c = threadIdx.x % kp, g = threadIdx.x / kp, G = kSegBlock / kp;
  const double a = alpha[(int64_t)tl.prob * kp + c];
  const T at = (T)a;
  double acc = 0.0;
  for (int64_t i = (int64_t)tl.r0 + g; i < tl.r1; i += G) {
    const int64_t e = i * kp + c;
    if (x) x[e] += at * p[e];
    const T rv = r[e] - at * ap[e];
    r[e] = rv;
    acc += (double)rv * (double)rv;
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = G / 2; w >= 1; w >>= 1) {
    if (g < w) sh[threadIdx.x] += sh[threadIdx.x + w * kp];
    __syncthreads();
  }
  if (g == 0) part[(int64_t)blockIdx.x * kp + c] = sh[c];
}

// p = z + beta p where the pair's mask is set
template <typename T>
__global__ void __launch_bounds__(kSegBlock) k_seg_xpby(T* __restrict__ p, const T* __restrict__ z,
                                                        const double* __restrict__ beta,
                                                        const int* __restrict__ mask,
                                                        const SegTile* __restrict__ tiles, int kp) {
  const SegTile tl = tiles[blockIdx.x];
  const int c = threadIdx.x % kp, g = threadIdx.x / kp, G = kSegBlock / kp;
  const int64_t pc = (int64_t)tl.prob * kp + c;
  if (!mask[pc]) return;
  const T b = (T)beta[pc];
  for (int64_t i = (int64_t)tl.r0 + g; i < tl.r1; i += G) {
    const int64_t e = i * kp + c;
    p[e] = z[e] + b * p[e];
  }
}

// dst = scale[pair] * src (precision conversion with a per-pair scale)
template <typename S, typename D>
__global__ void __launch_bounds__(kSegBlock) k_seg_cast_scale(const S* __restrict__ src, D* __restrict__ dst,
                                                              const double* __restrict__ scale,
                                                              const SegTile* __restrict__ tiles, int kp) {
  const SegTile tl = tiles[blockIdx.x];
  const int c = threadIdx.x % kp, g = threadIdx.x / kp, G = kSegBlock / kp;
  const double sc = scale[(int64_t)tl.prob * kp + c];
  for (int64_t i = (int64_t)tl.r0 + g; i < tl.r1; i += G) dst[i * kp + c] = (D)(sc * (double)src[i * kp + c]);
}

// X += scale[pair] * Dl (scale == nullptr: 1) where the pair's mask is set
template <typename S>
__global__ void __launch_bounds__(kSegBlock) k_seg_axpy_acc(double* __restrict__ X, const S* __restrict__ Dl,
                                                            const double* __restrict__ scale,
                                                            const int* __restrict__ mask,
                                                            const SegTile* __restrict__ tiles, int kp) {
  const SegTile tl = tiles[blockIdx.x];
  const int c = threadIdx.x % kp, g = threadIdx.x / kp, G = kSegBlock / kp;
  const int64_t pc = (int64_t)tl.prob * kp + c;
  if (!mask[pc]) return;
  const double sc = scale ? scale[pc] : 1.0;
  for (int64_t i = (int64_t)tl.r0 + g; i < tl.r1; i += G) X[i * kp + c] += sc * (double)Dl[i * kp + c];
}

static inline unsigned pair_blocks(const SegIndex& ix, int kp) {
  return (unsigned)((ix.P * kp + kPairBlock - 1) / kPairBlock);
}

template <typename T>
void launch_seg_dot(hipStream_t st, const T* A, const T* B, const SegIndex& ix, int kp, double* part) {
  if (ix.n_tiles > 0) k_seg_dot<T><<<(unsigned)ix.n_tiles, kSegBlock, 0, st>>>(A, B, ix.tiles, kp, part);
}
void launch_seg_totals(hipStream_t st, const double* part, const SegIndex& ix, int kp, double* out) {
  k_seg_totals<<<pair_blocks(ix, kp), kPairBlock, 0, st>>>(part, ix, kp, out);
}
void launch_seg_start(hipStream_t st, ColState* s, const double* part_bb, const double* part_rz, const SegIndex& ix,
                      const double* tol, const int* kcols, int kp) {
  k_seg_start<<<pair_blocks(ix, kp), kPairBlock, 0, st>>>(s, part_bb, part_rz, ix, tol, kcols, kp);
}
void launch_seg_alpha(hipStream_t st, ColState* s, const double* part, const SegIndex& ix, double* alpha, int kp) {
  k_seg_alpha<<<pair_blocks(ix, kp), kPairBlock, 0, st>>>(s, part, ix, alpha, kp);
}
void launch_seg_check(hipStream_t st, ColState* s, const double* part, const SegIndex& ix, int kp) {
  k_seg_check<<<pair_blocks(ix, kp), kPairBlock, 0, st>>>(s, part, ix, kp);
}
void launch_seg_beta(hipStream_t st, ColState* s, const double* part, const SegIndex& ix, double* beta, int* mask,
                     int* n_active, int kp) {
  k_seg_beta<<<pair_blocks(ix, kp), kPairBlock, 0, st>>>(s, part, ix, beta, mask, kp);
  k_seg_count<<<1, 1024, 0, st>>>(s, ix.P * kp, n_active);
}
void launch_seg_count(hipStream_t st, const ColState* s, const SegIndex& ix, int kp, int* n_active) {
  k_seg_count<<<1, 1024, 0, st>>>(s, ix.P * kp, n_active);
}
template <typename T>
void launch_seg_update(hipStream_t st, T* x, const T* p, T* r, const T* ap, const double* alpha, const SegIndex& ix,
                       int kp, double* part) {
  if (ix.n_tiles > 0)
    k_seg_update<T><<<(unsigned)ix.n_tiles, kSegBlock, 0, st>>>(x, p, r, ap, alpha, ix.tiles, kp, part);
}
template <typename T>
void launch_seg_xpby(hipStream_t st, T* p, const T* z, const double* beta, const int* mask, const SegIndex& ix,
                     int kp) {
  if (ix.n_tiles > 0) k_seg_xpby<T><<<(unsigned)ix.n_tiles, kSegBlock, 0, st>>>(p, z, beta, mask, ix.tiles, kp);
}
template <typename S, typename D>
void launch_seg_cast_scale(hipStream_t st, const S* src, D* dst, const double* scale, const SegIndex& ix, int kp) {
  if (ix.n_tiles > 0)
    k_seg_cast_scale<S, D><<<(unsigned)ix.n_tiles, kSegBlock, 0, st>>>(src, dst, scale, ix.tiles, kp);
}
template <typename S>
void launch_seg_axpy_acc(hipStream_t st, double* X, const S* Dl, const double* scale, const int* mask,
                         const SegIndex& ix, int kp) {
  if (ix.n_tiles > 0)
    k_seg_axpy_acc<S><<<(unsigned)ix.n_tiles, kSegBlock, 0, st>>>(X, Dl, scale, mask, ix.tiles, kp);
}

template void launch_seg_dot<float>(hipStream_t, const float*, const float*, const SegIndex&, int, double*);
template void launch_seg_dot<double>(hipStream_t, const double*, const double*, const SegIndex&, int, double*);
template void launch_seg_update<float>(hipStream_t, float*, const float*, float*, const float*, const double*,
                                       const SegIndex&, int, double*);
template void launch_seg_update<double>(hipStream_t, double*, const double*, double*, const double*,
                                        const double*, const SegIndex&, int, double*);
template void launch_seg_xpby<float>(hipStream_t, float*, const float*, const double*, const int*, const SegIndex&,
                                     int);
template void launch_seg_xpby<double>(hipStream_t, double*, const double*, const double*, const int*,
                                      const SegIndex&, int);
template void launch_seg_cast_scale<double, float>(hipStream_t, const double*, float*, const double*,
                                                   const SegIndex&, int);
template void launch_seg_cast_scale<double, double>(hipStream_t, const double*, double*, const double*,
                                                    const SegIndex&, int);
template void launch_seg_axpy_acc<float>(hipStream_t, double*, const float*, const double*, const int*,
                                         const SegIndex&, int);
template void launch_seg_axpy_acc<double>(hipStream_t, double*, const double*, const double*, const int*,
                                          const SegIndex&, int);

}  // namespace wa
