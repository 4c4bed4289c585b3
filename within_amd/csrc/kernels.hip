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

#include "kernels.hpp"

namespace wa {

#define GROUP_SETUP(kp)                                              \
  const int lane = threadIdx.x & ((kp) - 1);                         \
  const int64_t grp = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / (kp);

static inline unsigned blocks_for(int64_t groups, int kp) {
  int64_t threads = groups * kp;
  return (unsigned)((threads + kBlock - 1) / kBlock);
}

// ===========================================================================
// implicit operator: one level-sorted pass per factor
//   Y[off_q + j] = diag_q[j] X[off_q + j] + sum_{e in level j} w_e sum_{r != q} X[other_r(e)]
// (GramianOperator::apply, operator/gramian.rs:84-89, as D'W D x)
// ===========================================================================
template <typename T, bool WEIGHTED>
__global__ void __launch_bounds__(kBlock) k_pull(const Item* __restrict__ items, int64_t n_items,
                                                 const uint32_t* __restrict__ row_ptr,
                                                 const uint32_t* __restrict__ other, int qm,
                                                 const T* __restrict__ ew, const T* __restrict__ diag,
                                                 uint32_t off, const T* __restrict__ X,
                                                 T* __restrict__ Y, T* __restrict__ part, int kp) {
  GROUP_SETUP(kp)
  if (grp >= n_items) return;
  const Item it = items[grp];
  if (it.slot < 0) {
    for (uint32_t r = it.row0; r < it.row1; ++r) {
      T acc = diag[r] * X[(int64_t)(off + r) * kp + lane];
      const uint32_t e1 = row_ptr[r + 1];
      for (uint32_t e = row_ptr[r]; e < e1; ++e) {
        T s = 0;
        for (int t = 0; t < qm; ++t) s += X[(int64_t)other[(int64_t)e * qm + t] * kp + lane];
        acc += WEIGHTED ? ew[e] * s : s;
      }
      Y[(int64_t)(off + r) * kp + lane] = acc;
    }
  } else {
    T acc = 0;
    for (uint32_t e = it.e0; e < it.e1; ++e) {
      T s = 0;
      for (int t = 0; t < qm; ++t) s += X[(int64_t)other[(int64_t)e * qm + t] * kp + lane];
      acc += WEIGHTED ? ew[e] * s : s;
    }
    part[(int64_t)it.slot * kp + lane] = acc;
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_pull_fix(const Fixup* __restrict__ fx, int64_t n_fix,
                                                     const T* __restrict__ diag, uint32_t off,
                                                     const T* __restrict__ X, T* __restrict__ Y,
                                                     const T* __restrict__ part, int kp) {
  GROUP_SETUP(kp)
  if (grp >= n_fix) return;
  const Fixup f = fx[grp];
  T acc = diag ? diag[f.row] * X[(int64_t)(off + f.row) * kp + lane] : T(0);
  for (uint32_t s = f.slot0; s < f.slot1; ++s) acc += part[(int64_t)s * kp + lane];
  Y[(int64_t)(off + f.row) * kp + lane] = acc;
}

template <typename T>
void launch_pull(hipStream_t st, const DevPass<T>& p, const T* X, T* Y, T* part, int kp) {
  const DevSched& sc = kp >= 8 ? p.wide : p.narrow;
  if (sc.n_items > 0) {
    if (p.ew)
      k_pull<T, true><<<blocks_for(sc.n_items, kp), kBlock, 0, st>>>(
          sc.items, sc.n_items, p.row_ptr, p.other, p.qm, p.ew, p.diag, p.off, X, Y, part, kp);
    else
      k_pull<T, false><<<blocks_for(sc.n_items, kp), kBlock, 0, st>>>(
          sc.items, sc.n_items, p.row_ptr, p.other, p.qm, nullptr, p.diag, p.off, X, Y, part, kp);
  }
  if (sc.n_fix > 0)
    k_pull_fix<T><<<blocks_for(sc.n_fix, kp), kBlock, 0, st>>>(sc.fix, sc.n_fix, p.diag, p.off, X, Y,
                                                               part, kp);
}
template void launch_pull<float>(hipStream_t, const DevPass<float>&, const float*, float*, float*, int);
template void launch_pull<double>(hipStream_t, const DevPass<double>&, const double*, double*, double*, int);

// ===========================================================================
// observation space (FP64): D'W V and V = Y - D X, row-major n x kp
// (rmatvec_wdt / matvec_d, domain.rs:390-418)
// ===========================================================================
__global__ void __launch_bounds__(kBlock) k_scatter_obs(const Item* __restrict__ items, int64_t n_items,
                                                        const uint32_t* __restrict__ row_ptr,
                                                        const uint32_t* __restrict__ obs,
                                                        const double* __restrict__ w,
                                                        const double* __restrict__ V, uint32_t off,
                                                        double* __restrict__ out,
                                                        double* __restrict__ part, int kp) {
  GROUP_SETUP(kp)
  if (grp >= n_items) return;
  const Item it = items[grp];
  if (it.slot < 0) {
    for (uint32_t r = it.row0; r < it.row1; ++r) {
      double acc = 0.0;
      for (uint32_t e = row_ptr[r]; e < row_ptr[r + 1]; ++e) {
        const uint32_t i = obs[e];
        const double v = V[(int64_t)i * kp + lane];
        acc += w ? w[i] * v : v;
      }
      out[(int64_t)(off + r) * kp + lane] = acc;
    }
  } else {
    double acc = 0.0;
    for (uint32_t e = it.e0; e < it.e1; ++e) {
      const uint32_t i = obs[e];
      const double v = V[(int64_t)i * kp + lane];
      acc += w ? w[i] * v : v;
    }
    part[(int64_t)it.slot * kp + lane] = acc;
  }
}

void launch_scatter_obs(hipStream_t st, const DevPass<double>& p, const uint32_t* obs, const double* w,
                        const double* V, double* out, double* part, int kp) {
  const DevSched& sc = kp >= 8 ? p.wide : p.narrow;
  if (sc.n_items > 0)
    k_scatter_obs<<<blocks_for(sc.n_items, kp), kBlock, 0, st>>>(sc.items, sc.n_items, p.row_ptr, obs, w,
                                                                 V, p.off, out, part, kp);
  if (sc.n_fix > 0)
    k_pull_fix<double><<<blocks_for(sc.n_fix, kp), kBlock, 0, st>>>(sc.fix, sc.n_fix, nullptr, p.off,
                                                                    nullptr, out, part, kp);
}

// E[i] = Y[i] - sum_q X[off_q + code_q(i)]   (Y == nullptr: E = D X)
__global__ void __launch_bounds__(kBlock) k_demean(int64_t n, int Q, const uint32_t* __restrict__ codes,
                                                   const uint32_t* __restrict__ offs,
                                                   const double* __restrict__ X,
                                                   const double* __restrict__ Y, double* __restrict__ E,
                                                   int kp) {
  const int lane = threadIdx.x & (kp - 1);
  const int64_t gpb = kBlock / kp;
  for (int64_t i = (int64_t)blockIdx.x * gpb + threadIdx.x / kp; i < n; i += (int64_t)gridDim.x * gpb) {
    double s = 0.0;
    for (int q = 0; q < Q; ++q) s += X[(int64_t)(offs[q] + codes[(int64_t)q * n + i]) * kp + lane];
    E[i * kp + lane] = Y ? Y[i * kp + lane] - s : s;
  }
}

void launch_demean(hipStream_t st, int64_t n, int Q, const uint32_t* codes, const uint32_t* offs,
                   const double* X, const double* Y, double* E, int kp) {
  int64_t groups = n;
  unsigned nb = (unsigned)std::min<int64_t>(blocks_for(groups, kp), 65536);
  k_demean<<<nb, kBlock, 0, st>>>(n, Q, codes, offs, X, Y, E, kp);
}

// ===========================================================================
// layout conversion: column-major (ld) <-> row-major kp blocks, via LDS tiles
// ===========================================================================
template <typename S, typename D>
__global__ void __launch_bounds__(kBlock) k_col_to_row(const S* __restrict__ src, int64_t ld, int64_t n,
                                                       int k, D* __restrict__ dst, int kp) {
  __shared__ double tile[32][65];
  const int64_t r0 = (int64_t)blockIdx.x * 64;
  for (int t = threadIdx.x; t < 64 * 32; t += kBlock) {
    int r = t & 63, c = t >> 6;
    int64_t i = r0 + r;
    tile[c][r] = (c < k && i < n) ? (double)src[i + c * ld] : 0.0;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < 64 * kp; t += kBlock) {
    int c = t % kp, r = t / kp;
    int64_t i = r0 + r;
    if (i < n) dst[i * kp + c] = (D)tile[c][r];
  }
}

template <typename S, typename D>
__global__ void __launch_bounds__(kBlock) k_row_to_col(const S* __restrict__ src, int kp, int64_t n, int k,
                                                       D* __restrict__ dst, int64_t ld) {
  __shared__ double tile[32][65];
  const int64_t r0 = (int64_t)blockIdx.x * 64;
  for (int t = threadIdx.x; t < 64 * kp; t += kBlock) {
    int c = t % kp, r = t / kp;
    int64_t i = r0 + r;
    tile[c][r] = i < n ? (double)src[i * kp + c] : 0.0;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < 64 * 32; t += kBlock) {
    int r = t & 63, c = t >> 6;
    int64_t i = r0 + r;
    if (c < k && i < n) dst[i + c * ld] = (D)tile[c][r];
  }
}

void launch_col_to_row(hipStream_t st, const double* src, int64_t ld, int64_t n, int k, double* dst, int kp) {
  if (n > 0) k_col_to_row<double, double><<<(unsigned)((n + 63) / 64), kBlock, 0, st>>>(src, ld, n, k, dst, kp);
}
void launch_row_to_col(hipStream_t st, const double* src, int kp, int64_t n, int k, double* dst, int64_t ld) {
  if (n > 0) k_row_to_col<double, double><<<(unsigned)((n + 63) / 64), kBlock, 0, st>>>(src, kp, n, k, dst, ld);
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

// part[b * kp + c] = sum over this block's rows of A[:,c] * B[:,c]
template <typename T>
__global__ void __launch_bounds__(kBlock) k_dot_part(const T* __restrict__ A, const T* __restrict__ B,
                                                     int64_t m, int kp, double* __restrict__ part) {
  __shared__ double sh[kBlock];
  const int lane = threadIdx.x & (kp - 1);
  const int gpb = kBlock / kp;
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * gpb + threadIdx.x / kp; i < m; i += (int64_t)gridDim.x * gpb)
    acc += (double)A[i * kp + lane] * (double)B[i * kp + lane];
  sh[threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.x < kp) {
    double s = 0.0;
    for (int g = 0; g < gpb; ++g) s += sh[g * kp + threadIdx.x];
    part[(int64_t)blockIdx.x * kp + threadIdx.x] = s;
  }
}

__global__ void k_dot_fin(const double* __restrict__ part, int nb, int kp, double* __restrict__ out) {
  const int c = threadIdx.x;
  if (c >= kp) return;
  double s = 0.0;
  for (int b = 0; b < nb; ++b) s += part[(int64_t)b * kp + c];
  out[c] = s;
}

template <typename T>
void launch_dot(hipStream_t st, const T* A, const T* B, int64_t m, int kp, double* part, double* out) {
  unsigned nb = red_blocks(m, kp);
  k_dot_part<T><<<nb, kBlock, 0, st>>>(A, B, m, kp, part);
  k_dot_fin<<<1, 64, 0, st>>>(part, (int)nb, kp, out);
}
template void launch_dot<float>(hipStream_t, const float*, const float*, int64_t, int, double*, double*);
template void launch_dot<double>(hipStream_t, const double*, const double*, int64_t, int, double*, double*);

// X += alpha P ; R -= alpha AP ; part = sum R^2   (alpha per column, FP64)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_cg_update(T* __restrict__ X, const T* __restrict__ P,
                                                      T* __restrict__ R, const T* __restrict__ AP,
                                                      const double* __restrict__ alpha, int64_t m, int kp,
                                                      double* __restrict__ part) {
  __shared__ double sh[kBlock];
  const int lane = threadIdx.x & (kp - 1);
  const int gpb = kBlock / kp;
  const T a = (T)alpha[lane];
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * gpb + threadIdx.x / kp; i < m; i += (int64_t)gridDim.x * gpb) {
    const int64_t id = i * kp + lane;
    X[id] += a * P[id];
    const T r = R[id] - a * AP[id];
    R[id] = r;
    acc += (double)r * (double)r;
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.x < kp) {
    double s = 0.0;
    for (int g = 0; g < gpb; ++g) s += sh[g * kp + threadIdx.x];
    part[(int64_t)blockIdx.x * kp + threadIdx.x] = s;
  }
}

template <typename T>
void launch_cg_update(hipStream_t st, T* X, const T* P, T* R, const T* AP, const double* alpha, int64_t m,
                      int kp, double* part, double* out) {
  unsigned nb = red_blocks(m, kp);
  k_cg_update<T><<<nb, kBlock, 0, st>>>(X, P, R, AP, alpha, m, kp, part);
  k_dot_fin<<<1, 64, 0, st>>>(part, (int)nb, kp, out);
}
template void launch_cg_update<float>(hipStream_t, float*, const float*, float*, const float*, const double*,
                                      int64_t, int, double*, double*);
template void launch_cg_update<double>(hipStream_t, double*, const double*, double*, const double*,
                                       const double*, int64_t, int, double*, double*);

// P = Z + beta P (beta per column; columns with mask 0 untouched)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_xpby(T* __restrict__ P, const T* __restrict__ Z,
                                                 const double* __restrict__ beta, const int* __restrict__ mask,
                                                 int64_t m, int kp) {
  const int lane = threadIdx.x & (kp - 1);
  if (!mask[lane]) return;
  const T b = (T)beta[lane];
  const int gpb = kBlock / kp;
  for (int64_t i = (int64_t)blockIdx.x * gpb + threadIdx.x / kp; i < m; i += (int64_t)gridDim.x * gpb) {
    const int64_t id = i * kp + lane;
    P[id] = Z[id] + b * P[id];
  }
}
template <typename T>
void launch_xpby(hipStream_t st, T* P, const T* Z, const double* beta, const int* mask, int64_t m, int kp) {
  k_xpby<T><<<red_blocks(m, kp), kBlock, 0, st>>>(P, Z, beta, mask, m, kp);
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
  if (c >= k) { q.active = 0; q.converged = 1; return; }
  // zero-RHS shortcut: ||b|| < EPS -> x = 0, converged, 0 iterations
  if (bn < 2.220446049250313e-16) { q.active = 0; q.converged = 1; return; }
  q.active = 1;
  q.converged = 0;
}

// after A p: pap <= 0 -> stop (not converged); else alpha = rz / pap
__global__ void k_cg_alpha(ColState* __restrict__ s, const double* __restrict__ pap, double* __restrict__ alpha,
                           int itn, int kp) {
  const int c = threadIdx.x;
  if (c >= kp) return;
  ColState& q = s[c];
  alpha[c] = 0.0;
  if (!q.active) return;
  if (!(pap[c] > 0.0)) {
    q.active = 0;
    q.converged = 0;
    q.iters = itn;
    return;
  }
  alpha[c] = q.rz / pap[c];
  q.alpha = alpha[c];
}

// after the update: ||r|| / ||b|| <= tol -> converged
__global__ void k_cg_check(ColState* __restrict__ s, const double* __restrict__ rr, int itn, int kp) {
  const int c = threadIdx.x;
  if (c >= kp) return;
  ColState& q = s[c];
  if (!q.active) return;
  q.rnorm = sqrt(rr[c]);
  q.iters = itn;
  if (q.rnorm / q.bnorm <= q.tol) {
    q.active = 0;
    q.converged = 1;
  }
}

// after z = M r: stagnation guard |rz_new| < EPS^2 max(|rz0|, MIN); beta
__global__ void k_cg_beta(ColState* __restrict__ s, const double* __restrict__ rznew, double* __restrict__ beta,
                          int* __restrict__ mask, int* __restrict__ n_active, int itn, int kp) {
  const int c = threadIdx.x;
  if (c < kp) {
    ColState& q = s[c];
    beta[c] = 0.0;
    mask[c] = 0;
    if (q.active) {
      const double thr = 4.930380657631324e-32 * fmax(fabs(q.rz0), 2.2250738585072014e-308);
      if (fabs(rznew[c]) < thr) {
        q.active = 0;
        q.converged = q.rnorm / q.bnorm <= q.tol;
        q.iters = itn;
      } else {
        beta[c] = rznew[c] / q.rz;
        q.rz = rznew[c];
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
void launch_cg_alpha(hipStream_t st, ColState* s, const double* pap, double* alpha, int itn, int kp) {
  k_cg_alpha<<<1, 64, 0, st>>>(s, pap, alpha, itn, kp);
}
void launch_cg_check(hipStream_t st, ColState* s, const double* rr, int itn, int kp) {
  k_cg_check<<<1, 64, 0, st>>>(s, rr, itn, kp);
}
void launch_cg_beta(hipStream_t st, ColState* s, const double* rznew, double* beta, int* mask, int* n_active,
                    int itn, int kp) {
  k_cg_beta<<<1, 64, 0, st>>>(s, rznew, beta, mask, n_active, itn, kp);
}

// ===========================================================================
// additive Schwarz: z = sum_i R_i' D_i A_i^+ D_i R_i r over all subdomains at
// once (schwarz-precond/src/local_solve.rs:182-248 per subdomain;
// operator/local_solver.rs:360-449 block elimination)
// ===========================================================================
// B[l] = scale_l * R[gidx_l]
template <typename T>
__global__ void __launch_bounds__(kBlock) k_restrict(const uint32_t* __restrict__ gidx,
                                                     const T* __restrict__ scale, const T* __restrict__ R,
                                                     T* __restrict__ B, int64_t n, int kp) {
  const int lane = threadIdx.x & (kp - 1);
  const int gpb = kBlock / kp;
  for (int64_t l = (int64_t)blockIdx.x * gpb + threadIdx.x / kp; l < n; l += (int64_t)gridDim.x * gpb)
    B[l * kp + lane] = scale[l] * R[(int64_t)gidx[l] * kp + lane];
}

// per-segment sums of V over rows [ptr[s], ptr[s+1])  -> out[s]
template <typename T>
__global__ void __launch_bounds__(kBlock) k_segsum(const Item* __restrict__ items, int64_t n_items,
                                                   const uint32_t* __restrict__ ptr, const T* __restrict__ V,
                                                   T* __restrict__ out, T* __restrict__ part, int kp) {
  GROUP_SETUP(kp)
  if (grp >= n_items) return;
  const Item it = items[grp];
  if (it.slot < 0) {
    for (uint32_t s = it.row0; s < it.row1; ++s) {
      T acc = 0;
      for (uint32_t e = ptr[s]; e < ptr[s + 1]; ++e) acc += V[(int64_t)e * kp + lane];
      out[(int64_t)s * kp + lane] = acc;
    }
  } else {
    T acc = 0;
    for (uint32_t e = it.e0; e < it.e1; ++e) acc += V[(int64_t)e * kp + lane];
    part[(int64_t)it.slot * kp + lane] = acc;
  }
}

template <typename T>
void launch_segsum(hipStream_t st, const DevSched& sc, const uint32_t* ptr, const T* V, T* out, T* part, int kp) {
  if (sc.n_items > 0)
    k_segsum<T><<<blocks_for(sc.n_items, kp), kBlock, 0, st>>>(sc.items, sc.n_items, ptr, V, out, part, kp);
  if (sc.n_fix > 0)
    k_pull_fix<T><<<blocks_for(sc.n_fix, kp), kBlock, 0, st>>>(sc.fix, sc.n_fix, nullptr, 0, nullptr, out, part,
                                                               kp);
}

// keep rows: Rr[j] = B[keep_j] - mu1 + sum_e val_e B[e] - mu1 * rowsum_j
template <typename T>
__global__ void __launch_bounds__(kBlock) k_keep_rows(int64_t n_red, const uint32_t* __restrict__ red_loc,
                                                      const uint32_t* __restrict__ red_sub,
                                                      const uint32_t* __restrict__ loc_ptr,
                                                      const uint32_t* __restrict__ kptr,
                                                      const uint32_t* __restrict__ kidx,
                                                      const T* __restrict__ kval, const T* __restrict__ krs,
                                                      const T* __restrict__ B, const T* __restrict__ sum1,
                                                      T* __restrict__ Rr, int kp) {
  GROUP_SETUP(kp)
  if (grp >= n_red) return;
  const uint32_t s = red_sub[grp];
  const T mu = sum1[(int64_t)s * kp + lane] / (T)(loc_ptr[s + 1] - loc_ptr[s]);
  T acc = 0;
  for (uint32_t e = kptr[grp]; e < kptr[grp + 1]; ++e) acc += kval[e] * B[(int64_t)kidx[e] * kp + lane];
  Rr[grp * kp + lane] = B[(int64_t)red_loc[grp] * kp + lane] - mu + acc - mu * krs[grp];
}

// Rr[j] -= sum2[s] / n_red(s)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_sub_mean(int64_t n, const uint32_t* __restrict__ sub,
                                                     const uint32_t* __restrict__ ptr, const T* __restrict__ sums,
                                                     T* __restrict__ V, int kp) {
  const int lane = threadIdx.x & (kp - 1);
  const int gpb = kBlock / kp;
  for (int64_t j = (int64_t)blockIdx.x * gpb + threadIdx.x / kp; j < n; j += (int64_t)gridDim.x * gpb) {
    const uint32_t s = sub[j];
    V[j * kp + lane] -= sums[(int64_t)s * kp + lane] / (T)(ptr[s + 1] - ptr[s]);
  }
}

// forward level: z_v = z_v + sum_{a earlier} c_av z_a  (in place)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_fwd_level(int64_t nrows, const uint32_t* __restrict__ rows,
                                                      const uint32_t* __restrict__ iptr,
                                                      const uint32_t* __restrict__ iidx,
                                                      const T* __restrict__ ival, T* __restrict__ Z, int kp) {
  GROUP_SETUP(kp)
  if (grp >= nrows) return;
  const uint32_t v = rows[grp];
  T acc = Z[(int64_t)v * kp + lane];
  for (uint32_t e = iptr[grp]; e < iptr[grp + 1]; ++e) acc += ival[e] * Z[(int64_t)iidx[e] * kp + lane];
  Z[(int64_t)v * kp + lane] = acc;
}

// backward level: x_v = z_v / d_v + sum_{u later} (w/d_v) x_u  (in place)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_bwd_level(int64_t nrows, const uint32_t* __restrict__ rows,
                                                      const uint32_t* __restrict__ optr,
                                                      const uint32_t* __restrict__ oidx,
                                                      const T* __restrict__ oval, const T* __restrict__ inv_piv,
                                                      T* __restrict__ Z, int kp) {
  GROUP_SETUP(kp)
  if (grp >= nrows) return;
  const uint32_t v = rows[grp];
  T acc = inv_piv[v] * Z[(int64_t)v * kp + lane];
  for (uint32_t e = optr[grp]; e < optr[grp + 1]; ++e) acc += oval[e] * Z[(int64_t)oidx[e] * kp + lane];
  Z[(int64_t)v * kp + lane] = acc;
}

// dense cores: out[cr] = sum_j Minv_c[row, j] z[core_j]
template <typename T>
__global__ void __launch_bounds__(kBlock) k_core(int64_t n_rows, const uint32_t* __restrict__ row_core,
                                                 const uint32_t* __restrict__ cptr,
                                                 const uint32_t* __restrict__ cidx,
                                                 const uint64_t* __restrict__ moff, const T* __restrict__ mat,
                                                 const T* __restrict__ Z, T* __restrict__ out, int kp) {
  GROUP_SETUP(kp)
  if (grp >= n_rows) return;
  const uint32_t c = row_core[grp];
  const uint32_t b = cptr[c], n = cptr[c + 1] - b;
  const uint32_t row = (uint32_t)grp - b;
  const T* M = mat + moff[c] + (uint64_t)row * n;
  T acc = 0;
  for (uint32_t j = 0; j < n; ++j) acc += M[j] * Z[(int64_t)cidx[b + j] * kp + lane];
  out[grp * kp + lane] = acc;
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_scatter_rows(int64_t n, const uint32_t* __restrict__ idx,
                                                         const T* __restrict__ src, T* __restrict__ dst, int kp) {
  const int lane = threadIdx.x & (kp - 1);
  const int gpb = kBlock / kp;
  for (int64_t i = (int64_t)blockIdx.x * gpb + threadIdx.x / kp; i < n; i += (int64_t)gridDim.x * gpb)
    dst[(int64_t)idx[i] * kp + lane] = src[i * kp + lane];
}

// elimination rows: S[e] = (B[e] - mu1) inv_e + inv_e * sum_j C[e,j] x_keep[j]
template <typename T>
__global__ void __launch_bounds__(kBlock) k_elim_rows(int64_t n_elim, const uint32_t* __restrict__ eloc,
                                                      const T* __restrict__ einv,
                                                      const uint32_t* __restrict__ eptr,
                                                      const uint32_t* __restrict__ eidx,
                                                      const T* __restrict__ eval,
                                                      const uint32_t* __restrict__ loc_sub,
                                                      const uint32_t* __restrict__ loc_ptr,
                                                      const T* __restrict__ B, const T* __restrict__ sum1,
                                                      const T* __restrict__ Rr, T* __restrict__ S, int kp) {
  GROUP_SETUP(kp)
  if (grp >= n_elim) return;
  const uint32_t l = eloc[grp];
  const uint32_t s = loc_sub[l];
  const T mu = sum1[(int64_t)s * kp + lane] / (T)(loc_ptr[s + 1] - loc_ptr[s]);
  T acc = 0;
  for (uint32_t e = eptr[grp]; e < eptr[grp + 1]; ++e) acc += eval[e] * Rr[(int64_t)eidx[e] * kp + lane];
  const T iv = einv[grp];
  S[(int64_t)l * kp + lane] = (B[(int64_t)l * kp + lane] - mu) * iv + iv * acc;
}

// Z[d] = sum_{l in dof d} scale_l (S[l] - mu3(sub(l)))   (0 for uncovered DOFs)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_combine(int64_t m, const uint32_t* __restrict__ dptr,
                                                    const uint32_t* __restrict__ dloc,
                                                    const T* __restrict__ scale,
                                                    const uint32_t* __restrict__ loc_sub,
                                                    const uint32_t* __restrict__ loc_ptr,
                                                    const T* __restrict__ S, const T* __restrict__ sum3,
                                                    T* __restrict__ Z, int kp) {
  const int lane = threadIdx.x & (kp - 1);
  const int gpb = kBlock / kp;
  for (int64_t d = (int64_t)blockIdx.x * gpb + threadIdx.x / kp; d < m; d += (int64_t)gridDim.x * gpb) {
    T acc = 0;
    for (uint32_t p = dptr[d]; p < dptr[d + 1]; ++p) {
      const uint32_t l = dloc[p];
      const uint32_t s = loc_sub[l];
      const T mu = sum3[(int64_t)s * kp + lane] / (T)(loc_ptr[s + 1] - loc_ptr[s]);
      acc += scale[l] * (S[(int64_t)l * kp + lane] - mu);
    }
    Z[d * kp + lane] = acc;
  }
}

static inline unsigned grid_stride_blocks(int64_t n, int kp) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(blocks_for(n, kp), 65536));
}

template <typename T>
void launch_precond(hipStream_t st, const DevPrecond<T>& p, const T* R, T* Z, PrecondWork<T>& w, int kp,
                    int64_t m) {
  // 1. restrict + per-subdomain sums (mu1)
  k_restrict<T><<<grid_stride_blocks(p.n_loc, kp), kBlock, 0, st>>>(p.loc_gidx, p.loc_scale, R, w.B, p.n_loc, kp);
  launch_segsum<T>(st, p.sched_loc, p.loc_ptr, w.B, w.sum1, w.part, kp);
  // 2. keep rows (Schur right-hand side) + sums (mu2), then centre
  if (p.n_red > 0) {
    k_keep_rows<T><<<blocks_for(p.n_red, kp), kBlock, 0, st>>>(p.n_red, p.red_loc, p.red_sub, p.loc_ptr, p.keep_ptr,
                                                               p.keep_idx, p.keep_val, p.keep_rowsum, w.B, w.sum1,
                                                               w.Rr, kp);
    launch_segsum<T>(st, p.sched_red, p.red_ptr, w.Rr, w.sum2, w.part, kp);
    k_sub_mean<T><<<grid_stride_blocks(p.n_red, kp), kBlock, 0, st>>>(p.n_red, p.red_sub, p.red_ptr, w.sum2, w.Rr,
                                                                      kp);
    // 3. reduced solve: forward levels, dense cores, backward levels
    for (int64_t t = 0; t < p.n_levels; ++t) {
      int64_t b = p.h_lvl_ptr[t], e = p.h_lvl_ptr[t + 1];
      if (e > b)
        k_fwd_level<T><<<blocks_for(e - b, kp), kBlock, 0, st>>>(e - b, p.lvl_rows + b, p.fin_ptr + b, p.fin_idx,
                                                                 p.fin_val, w.Rr, kp);
    }
    if (p.n_core_rows > 0) {
      k_core<T><<<blocks_for(p.n_core_rows, kp), kBlock, 0, st>>>(p.n_core_rows, p.core_row_core, p.core_ptr,
                                                                  p.core_idx, p.core_mat_off, p.core_mat, w.Rr,
                                                                  w.C, kp);
      k_scatter_rows<T><<<grid_stride_blocks(p.n_core_rows, kp), kBlock, 0, st>>>(p.n_core_rows, p.core_idx, w.C,
                                                                                  w.Rr, kp);
    }
    for (int64_t t = p.n_levels - 2; t >= 0; --t) {
      int64_t b = p.h_lvl_ptr[t], e = p.h_lvl_ptr[t + 1];
      if (e > b)
        k_bwd_level<T><<<blocks_for(e - b, kp), kBlock, 0, st>>>(e - b, p.lvl_rows + b, p.bout_ptr + b, p.bout_idx,
                                                                 p.bout_val, p.inv_pivot, w.Rr, kp);
    }
    // 4. back-substitution for the eliminated block; kept block copied over
    k_scatter_rows<T><<<grid_stride_blocks(p.n_red, kp), kBlock, 0, st>>>(p.n_red, p.red_loc, w.Rr, w.S, kp);
  }
  if (p.n_elim > 0)
    k_elim_rows<T><<<blocks_for(p.n_elim, kp), kBlock, 0, st>>>(p.n_elim, p.elim_loc, p.elim_inv, p.elim_ptr,
                                                                p.elim_idx, p.elim_val, p.loc_sub, p.loc_ptr, w.B,
                                                                w.sum1, w.Rr, w.S, kp);
  // 5. centre the local solutions (mu3), sign, partition weights, combine
  launch_segsum<T>(st, p.sched_loc, p.loc_ptr, w.S, w.sum3, w.part, kp);
  k_combine<T><<<grid_stride_blocks(m, kp), kBlock, 0, st>>>(m, p.dof_ptr, p.dof_loc, p.loc_scale, p.loc_sub,
                                                             p.loc_ptr, w.S, w.sum3, Z, kp);
}
template void launch_precond<float>(hipStream_t, const DevPrecond<float>&, const float*, float*,
                                    PrecondWork<float>&, int, int64_t);
template void launch_precond<double>(hipStream_t, const DevPrecond<double>&, const double*, double*,
                                     PrecondWork<double>&, int, int64_t);

// ===========================================================================
// explicit Gramian SpMM: Y[row] = sum_e val_e X[col_e]   (explicit.rs:261-277)
// ===========================================================================
template <typename T>
__global__ void __launch_bounds__(kBlock) k_spmm(const Item* __restrict__ items, int64_t n_items,
                                                 const uint32_t* __restrict__ ptr, const uint32_t* __restrict__ idx,
                                                 const T* __restrict__ val, const T* __restrict__ X,
                                                 T* __restrict__ Y, T* __restrict__ part, int kp) {
  GROUP_SETUP(kp)
  if (grp >= n_items) return;
  const Item it = items[grp];
  if (it.slot < 0) {
    for (uint32_t r = it.row0; r < it.row1; ++r) {
      T acc = 0;
      for (uint32_t e = ptr[r]; e < ptr[r + 1]; ++e) acc += val[e] * X[(int64_t)idx[e] * kp + lane];
      Y[(int64_t)r * kp + lane] = acc;
    }
  } else {
    T acc = 0;
    for (uint32_t e = it.e0; e < it.e1; ++e) acc += val[e] * X[(int64_t)idx[e] * kp + lane];
    part[(int64_t)it.slot * kp + lane] = acc;
  }
}

template <typename T>
void launch_spmm(hipStream_t st, const DevCsr<T>& g, const T* X, T* Y, T* part, int kp) {
  const DevSched& sc = kp >= 8 ? g.wide : g.narrow;
  if (sc.n_items > 0)
    k_spmm<T><<<blocks_for(sc.n_items, kp), kBlock, 0, st>>>(sc.items, sc.n_items, g.ptr, g.idx, g.val, X, Y, part,
                                                             kp);
  if (sc.n_fix > 0)
    k_pull_fix<T><<<blocks_for(sc.n_fix, kp), kBlock, 0, st>>>(sc.fix, sc.n_fix, nullptr, 0, nullptr, Y, part, kp);
}
template void launch_spmm<float>(hipStream_t, const DevCsr<float>&, const float*, float*, float*, int);
template void launch_spmm<double>(hipStream_t, const DevCsr<double>&, const double*, double*, double*, int);

}  // namespace wa
