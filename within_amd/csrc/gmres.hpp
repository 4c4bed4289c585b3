// gmres.hpp -- device side of block GMRES(restart): block Gram-Schmidt
// launches and the per-column Hessenberg / Givens state (FP64).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace wa {

constexpr int kGmresMaxV = 32;      // basis vectors per cycle on the device path (restart <= 31)
constexpr int kGmRedBlocks = 1024;  // fixed grid of the Gram-Schmidt reductions

// per-column Arnoldi state: rotations, rotated right-hand side g, the
// largest Hessenberg entry so far, extension flag, last column, outcome
// (0 converged, 1 maxiter, 2 breakdown, 3 restart), residual estimate
struct GmCol {
  double cs[kGmresMaxV], sn[kGmresMaxV], g[kGmresMaxV];
  double hmax, resid;
  int inc, jc, outcome, pad;
};

// out[i * kp + c] (accumulate: +=) = <W[:, c], V_i[:, c]> for i < nv, V_i = V + i * vstride
template <typename T>
void launch_multi_dot(hipStream_t st, const T* W, const T* V, int64_t vstride, int nv, int64_t m, int kp,
                      double* part, double* out, int accumulate);
// W[:, c] += sign * sum_{i < min(nv, lim[c])} coef[i * kp + c] V_i[:, c]; with
// part, the per-block sums of the new W^2 (returns the block count)
template <typename T>
int launch_multi_axpy(hipStream_t st, T* W, const T* V, int64_t vstride, int nv, const double* coef, const int* lim,
                      double sign, int64_t m, int kp, double* part);
void launch_gm_cycle_start(hipStream_t st, GmCol* s, double* H, const double* beta, const int* live, double* coef,
                           int mr, int kp);
void launch_gm_givens(hipStream_t st, GmCol* s, double* H, const double* h, const double* h2, const double* part2, int nb,
                      const double* abs_tol, double* coef, int* n_inc, int j, int mr, int64_t total, int64_t maxiter,
                      int kp);
void launch_gm_backsolve(hipStream_t st, const GmCol* s, const double* H, const int* live, double* y, int* lim, int mr,
                         int kp);

}  // namespace wa
