// engine.hpp -- the row-gather engine shared by every CSR-like kernel.
//
// A "row" is a sum over elements: out(row) = finish(row, sum_e coef(e) *
// sum_{t<QM} X[idx_t(e)]).  Rows are processed through a host-built schedule
// of work items: an item covers whole rows (finished in place) or one chunk of
// a long row (written to a partial slot and summed in slot order by the
// fix-up kernel), so long rows never serialise a group and the result is
// bitwise reproducible.
//
// Two lane mappings:
//   COLS  (kp >= 8): a group of kp lanes, lane = column.  Each lane loads the
//         indices of up to kp consecutive elements in one coalesced load and
//         the group walks them with __shfl broadcasts, so the row gathers of
//         one batch are independent and in flight together.
//   ELEMS (kp in {1,2,4}): a group of 8 lanes, lane = element slot; each lane
//         accumulates all kp columns of its elements, then an xor-butterfly
//         over the 8 lanes (fixed order) forms the row sum.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.hpp"

namespace wa {

constexpr int kElemGroup = 8;

// ---- COLS mode -------------------------------------------------------------
template <class K, typename T>
__device__ __forceinline__ T gather_cols(const K& k, uint32_t row, uint32_t e0, uint32_t e1, int lane, int kp) {
  T acc = 0;
  constexpr int QM = K::QM > 0 ? K::QM : 1;
  for (uint32_t b = e0; b < e1; b += (uint32_t)kp) {
    const uint32_t cnt = min((uint32_t)kp, e1 - b);
    uint32_t id[QM];
    T c = 0;
#pragma unroll
    for (int t = 0; t < QM; ++t) id[t] = 0;
    if ((uint32_t)lane < cnt) k.load(row, b + lane, id, c);
#pragma unroll 4
    for (uint32_t j = 0; j < cnt; ++j) {
      T s = 0;
#pragma unroll
      for (int t = 0; t < QM; ++t) {
        const uint32_t ix = (uint32_t)__shfl((int)id[t], (int)j, kp);
        s += k.X[(int64_t)ix * kp + lane];
      }
      if (K::COEF) {
        const T cj = __shfl(c, (int)j, kp);
        acc += cj * s;
      } else {
        acc += s;
      }
    }
  }
  return acc;
}

template <class K, typename T>
__global__ void __launch_bounds__(kBlock) k_rows_cols(K k, const Item* __restrict__ items, int64_t n_items,
                                                      T* __restrict__ part, int kp) {
  const int lane = threadIdx.x & (kp - 1);
  const int64_t g = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kp;
  if (g >= n_items) return;
  const Item it = items[g];
  if (it.slot < 0) {
    for (uint32_t r = it.row0; r < it.row1; ++r) {
      T acc = K::QM == 0 ? T(0) : gather_cols<K, T>(k, r, k.rbegin(r), k.rend(r), lane, kp);
      k.finish(r, lane, acc);
    }
  } else {
    T acc = K::QM == 0 ? T(0) : gather_cols<K, T>(k, it.row0, it.e0, it.e1, lane, kp);
    part[(int64_t)it.slot * kp + lane] = acc;
  }
}

// ---- ELEMS mode ------------------------------------------------------------
template <int KP, class K, typename T>
__device__ __forceinline__ void gather_elems(const K& k, uint32_t row, uint32_t e0, uint32_t e1, int sub,
                                             T (&acc)[KP]) {
  constexpr int QM = K::QM > 0 ? K::QM : 1;
#pragma unroll
  for (int c = 0; c < KP; ++c) acc[c] = 0;
  if (K::QM > 0) {
    for (uint32_t e = e0 + sub; e < e1; e += kElemGroup) {
      uint32_t id[QM];
      T cf = 1;
      k.load(row, e, id, cf);
#pragma unroll
      for (int c = 0; c < KP; ++c) {
        T s = 0;
#pragma unroll
        for (int t = 0; t < QM; ++t) s += k.X[(int64_t)id[t] * KP + c];
        acc[c] += K::COEF ? cf * s : s;
      }
    }
  }
#pragma unroll
  for (int off = kElemGroup / 2; off >= 1; off >>= 1)
#pragma unroll
    for (int c = 0; c < KP; ++c) acc[c] += __shfl_xor(acc[c], off, kElemGroup);
}

template <int KP, class K, typename T>
__global__ void __launch_bounds__(kBlock) k_rows_elems(K k, const Item* __restrict__ items, int64_t n_items,
                                                       T* __restrict__ part) {
  const int sub = threadIdx.x & (kElemGroup - 1);
  const int64_t g = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kElemGroup;
  if (g >= n_items) return;
  const Item it = items[g];
  T acc[KP];
  if (it.slot < 0) {
    for (uint32_t r = it.row0; r < it.row1; ++r) {
      gather_elems<KP, K, T>(k, r, k.rbegin(r), k.rend(r), sub, acc);
#pragma unroll
      for (int c = 0; c < KP; ++c)
        if (sub == c) k.finish(r, c, acc[c]);
    }
  } else {
    gather_elems<KP, K, T>(k, it.row0, it.e0, it.e1, sub, acc);
#pragma unroll
    for (int c = 0; c < KP; ++c)
      if (sub == c) part[(int64_t)it.slot * KP + c] = acc[c];
  }
}

// ---- fix-up: sum the partial slots of each split row in slot order ----------
template <class K, typename T>
__global__ void __launch_bounds__(kBlock) k_rows_fix(K k, const Fixup* __restrict__ fx, int64_t n_fix,
                                                     const T* __restrict__ part, int kp) {
  const int lane = threadIdx.x & (kp - 1);
  const int64_t g = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kp;
  if (g >= n_fix) return;
  const Fixup f = fx[g];
  T acc = 0;
  for (uint32_t s = f.slot0; s < f.slot1; ++s) acc += part[(int64_t)s * kp + lane];
  k.finish(f.row, lane, acc);
}

inline unsigned engine_blocks(int64_t groups, int lanes) {
  return (unsigned)((groups * lanes + kBlock - 1) / kBlock);
}

// Run items [i0, i1) and fix-ups [f0, f1) of a schedule.
template <class K, typename T>
void run_rows(hipStream_t st, const K& k, const DevSched& sc, int64_t i0, int64_t i1, int64_t f0, int64_t f1,
              T* part, int kp) {
  const int64_t ni = i1 - i0, nf = f1 - f0;
  if (ni > 0) {
    if (kp >= 8)
      k_rows_cols<K, T><<<engine_blocks(ni, kp), kBlock, 0, st>>>(k, sc.items + i0, ni, part, kp);
    else if (kp == 4)
      k_rows_elems<4, K, T><<<engine_blocks(ni, kElemGroup), kBlock, 0, st>>>(k, sc.items + i0, ni, part);
    else if (kp == 2)
      k_rows_elems<2, K, T><<<engine_blocks(ni, kElemGroup), kBlock, 0, st>>>(k, sc.items + i0, ni, part);
    else
      k_rows_elems<1, K, T><<<engine_blocks(ni, kElemGroup), kBlock, 0, st>>>(k, sc.items + i0, ni, part);
  }
  if (nf > 0) k_rows_fix<K, T><<<engine_blocks(nf, kp), kBlock, 0, st>>>(k, sc.fix + f0, nf, part, kp);
}

template <class K, typename T>
void run_rows(hipStream_t st, const K& k, const DevSched& sc, T* part, int kp) {
  run_rows<K, T>(st, k, sc, 0, sc.n_items, 0, sc.n_fix, part, kp);
}

}  // namespace wa
