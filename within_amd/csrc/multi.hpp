// multi.hpp -- per-problem reductions for many fixed-effects problems stacked
// block-diagonally into one design (block CG with one recurrence per
// (problem, column) pair).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace wa {

// rows [r0, r1) of the DOF block, all of problem `prob`
struct SegTile {
  uint32_t r0, r1, prob, pad;
};
constexpr uint32_t kSegTileRows = 2048;

// device view: the tiles and, per problem, the list of its tiles (CSR)
struct SegIndex {
  const SegTile* tiles = nullptr;
  int64_t n_tiles = 0;
  const uint32_t* pt_ptr = nullptr;  // P + 1
  const uint32_t* pt_idx = nullptr;  // tile indices, ascending per problem
  int P = 0;
};

// per-tile column sums of A .* B (B == nullptr: A .* A) into part (n_tiles x kp)
template <typename T>
void launch_seg_dot(hipStream_t st, const T* A, const T* B, const SegIndex& ix, int kp, double* part);
// per-pair totals of per-tile partials: out[p * kp + c]
void launch_seg_totals(hipStream_t st, const double* part, const SegIndex& ix, int kp, double* out);
// CG scalar steps per (problem, column) pair; pairs are indexed p * kp + c
void launch_seg_start(hipStream_t st, ColState* s, const double* part_bb, const double* part_rz, const SegIndex& ix,
                      const double* tol, const int* kcols, int kp);
void launch_seg_alpha(hipStream_t st, ColState* s, const double* part, const SegIndex& ix, double* alpha, int kp);
void launch_seg_check(hipStream_t st, ColState* s, const double* part, const SegIndex& ix, int kp);
void launch_seg_beta(hipStream_t st, ColState* s, const double* part, const SegIndex& ix, double* beta, int* mask,
                     int* n_active, int kp);
void launch_seg_count(hipStream_t st, const ColState* s, const SegIndex& ix, int kp, int* n_active);
// x += alpha p (x may be null), r -= alpha ap, part = per-tile r'r
template <typename T>
void launch_seg_update(hipStream_t st, T* x, const T* p, T* r, const T* ap, const double* alpha, const SegIndex& ix,
                       int kp, double* part);
template <typename T>
void launch_seg_xpby(hipStream_t st, T* p, const T* z, const double* beta, const int* mask, const SegIndex& ix,
                     int kp);
template <typename S, typename D>
void launch_seg_cast_scale(hipStream_t st, const S* src, D* dst, const double* scale, const SegIndex& ix, int kp);
template <typename S>
void launch_seg_axpy_acc(hipStream_t st, double* X, const S* Dl, const double* scale, const int* mask,
                         const SegIndex& ix, int kp);

}  // namespace wa
