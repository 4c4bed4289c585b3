// engine.hpp -- the row-gather engine shared by every CSR-like kernel.
//
// A "row" is a sum over elements: out(row) = finish(row, sum_e coef(e) *
// sum_{t<QM} X[idx_t(e)]).  Rows are processed through a host-built schedule
// of work items: an item covers whole rows (finished in place) or one chunk of
// a long row (written to a partial slot; a fix-up sums the slots of the row in
// slot order), so long rows never serialise a wave and every result is
// bitwise reproducible.
//
// Lane mapping (one wave64 per work item): a block-vector row holds kp values
// of type T; each lane owns NV consecutive values (a 16-byte vector when the
// row is at least 16 bytes), V = kp / NV lanes cover one row, and the wave
// works on S = 64 / V elements at once ("slots").  Every lane loads its
// element's indices itself (the V lanes of a slot read the same address, so
// the wave issues S distinct index loads), then its vector slices of the
// gathered rows: no cross-lane traffic until the row ends, where the S slot
// partials are combined by a fixed xor-butterfly.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.hpp"

namespace wa {

constexpr int kWave = 64;

template <typename T, int NV>
struct alignas(sizeof(T) * NV) VecT {
  T v[NV];
};

template <typename T, int NV>
__device__ __forceinline__ VecT<T, NV> vload(const T* p) {
  return *reinterpret_cast<const VecT<T, NV>*>(p);
}

// V lanes per row, NV values per lane (V * NV == kp)
template <class K, typename T, int V, int NV>
__device__ __forceinline__ void gather_vec(const K& k, uint32_t row, uint32_t e0, uint32_t e1, int slot, int part,
                                           T (&acc)[NV]) {
  constexpr int S = kWave / V;
  constexpr int QM = K::QM > 0 ? K::QM : 1;
  constexpr int KP = V * NV;
  constexpr int U = 2;  // elements per lane in flight together
#pragma unroll
  for (int c = 0; c < NV; ++c) acc[c] = 0;
  if (K::QM == 0) return;
  for (uint32_t e = e0 + slot; e < e1; e += U * S) {
    uint32_t id[U][QM];
    T cf[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      cf[u] = 1;
#pragma unroll
      for (int t = 0; t < QM; ++t) id[u][t] = 0;
      if (e + u * S < e1) k.load(row, e + u * S, id[u], cf[u]);
    }
    VecT<T, NV> g[U][QM];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int t = 0; t < QM; ++t) {
        if (e + u * S < e1) {
          g[u][t] = vload<T, NV>(k.X + (int64_t)id[u][t] * KP + part * NV);
        } else {
#pragma unroll
          for (int c = 0; c < NV; ++c) g[u][t].v[c] = 0;
        }
      }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int c = 0; c < NV; ++c) {
        T s = 0;
#pragma unroll
        for (int t = 0; t < QM; ++t) s += g[u][t].v[c];
        if (e + u * S < e1) acc[c] += K::COEF ? cf[u] * s : s;
      }
  }
  // combine the S slots (lanes V apart) in a fixed order
#pragma unroll
  for (int off = V; off < kWave; off <<= 1)
#pragma unroll
    for (int c = 0; c < NV; ++c) acc[c] += __shfl_xor(acc[c], off, kWave);
}

template <class K, typename T, int V, int NV>
__global__ void __launch_bounds__(kBlock) k_rows_vec(K k, const Item* __restrict__ items, int64_t n_items,
                                                     T* __restrict__ part) {
  const int lane = threadIdx.x & (kWave - 1);
  const int slot = lane / V, pt = lane % V;
  const int64_t g = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  if (g >= n_items) return;
  const Item it = items[g];
  T acc[NV];
  if (it.slot < 0) {
    for (uint32_t r = it.row0; r < it.row1; ++r) {
      gather_vec<K, T, V, NV>(k, r, k.rbegin(r), k.rend(r), slot, pt, acc);
      if (slot == 0)
#pragma unroll
        for (int c = 0; c < NV; ++c) k.finish(r, pt * NV + c, acc[c]);
    }
  } else {
    gather_vec<K, T, V, NV>(k, it.row0, it.e0, it.e1, slot, pt, acc);
    if (slot == 0)
#pragma unroll
      for (int c = 0; c < NV; ++c) part[(int64_t)it.slot * (V * NV) + pt * NV + c] = acc[c];
  }
}

// ---- fix-up: one workgroup per split row; threads stride over the slots,
// then a fixed-order tree over the workgroup ----------------------------------
template <class K, typename T>
__global__ void __launch_bounds__(kBlock) k_rows_fix(K k, const Fixup* __restrict__ fx, const T* __restrict__ part,
                                                     int kp) {
  __shared__ T sh[kBlock];
  const Fixup f = fx[blockIdx.x];
  const int col = threadIdx.x % kp, str = kBlock / kp, s0 = threadIdx.x / kp;
  T acc = 0;
  for (uint32_t s = f.slot0 + s0; s < f.slot1; s += str) acc += part[(int64_t)s * kp + col];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = str / 2; w >= 1; w >>= 1) {
    if (s0 < w) sh[threadIdx.x] += sh[threadIdx.x + w * kp];
    __syncthreads();
  }
  if (s0 == 0) k.finish(f.row, col, sh[threadIdx.x]);
}

inline unsigned engine_blocks(int64_t waves) { return (unsigned)((waves * kWave + kBlock - 1) / kBlock); }

template <class K, typename T, int V, int NV>
void launch_vec(hipStream_t st, const K& k, const Item* items, int64_t ni, T* part) {
  k_rows_vec<K, T, V, NV><<<engine_blocks(ni), kBlock, 0, st>>>(k, items, ni, part);
}

// Run items [i0, i1) and fix-ups [f0, f1) of a schedule.
template <class K, typename T>
void run_rows(hipStream_t st, const K& k, const DevSched& sc, int64_t i0, int64_t i1, int64_t f0, int64_t f1,
              T* part, int kp) {
  const int64_t ni = i1 - i0, nf = f1 - f0;
  if (ni > 0) {
    const Item* it = sc.items + i0;
    constexpr int NV16 = 16 / sizeof(T);  // values per 16-byte vector
    if (kp * (int)sizeof(T) >= 16) {
      switch (kp / NV16) {  // lanes per row
        case 1: launch_vec<K, T, 1, NV16>(st, k, it, ni, part); break;
        case 2: launch_vec<K, T, 2, NV16>(st, k, it, ni, part); break;
        case 4: launch_vec<K, T, 4, NV16>(st, k, it, ni, part); break;
        case 8: launch_vec<K, T, 8, NV16>(st, k, it, ni, part); break;
        default: launch_vec<K, T, 16, NV16>(st, k, it, ni, part); break;
      }
    } else if (kp == 2) {
      launch_vec<K, T, 1, 2>(st, k, it, ni, part);
    } else {
      launch_vec<K, T, 1, 1>(st, k, it, ni, part);
    }
  }
  if (nf > 0) k_rows_fix<K, T><<<(unsigned)nf, kBlock, 0, st>>>(k, sc.fix + f0, part, kp);
}

template <class K, typename T>
void run_rows(hipStream_t st, const K& k, const DevSched& sc, T* part, int kp) {
  run_rows<K, T>(st, k, sc, 0, sc.n_items, 0, sc.n_fix, part, kp);
}

// ---- many small levels in one workgroup -----------------------------------
// Levels t0..t1-1 (ascending when dir > 0, else descending) of a level
// schedule, processed by the 16 waves of one workgroup with a barrier between
// levels (rows of a level depend only on earlier levels).  The device-scope
// fences make each level's global writes visible to every wave before the
// next level reads them.
constexpr int kTailThreads = 1024;

template <class K, typename T, int V, int NV>
__global__ void __launch_bounds__(kTailThreads) k_levels_wg(K k, const Item* __restrict__ items,
                                                            const int64_t* __restrict__ item_ptr,
                                                            const Fixup* __restrict__ fx,
                                                            const int64_t* __restrict__ fix_ptr,
                                                            T* __restrict__ part, int t0, int t1, int dir) {
  constexpr int KP = V * NV;
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  constexpr int nw = kTailThreads / kWave;
  const int slot = lane / V, pt = lane % V;
  T acc[NV];
  for (int step = 0; step < t1 - t0; ++step) {
    const int t = dir > 0 ? t0 + step : t1 - 1 - step;
    for (int64_t g = item_ptr[t] + wave; g < item_ptr[t + 1]; g += nw) {
      const Item it = items[g];
      if (it.slot < 0) {
        for (uint32_t r = it.row0; r < it.row1; ++r) {
          gather_vec<K, T, V, NV>(k, r, k.rbegin(r), k.rend(r), slot, pt, acc);
          if (slot == 0)
#pragma unroll
            for (int c = 0; c < NV; ++c) k.finish(r, pt * NV + c, acc[c]);
        }
      } else {
        gather_vec<K, T, V, NV>(k, it.row0, it.e0, it.e1, slot, pt, acc);
        if (slot == 0)
#pragma unroll
          for (int c = 0; c < NV; ++c) part[(int64_t)it.slot * KP + pt * NV + c] = acc[c];
      }
    }
    __threadfence();
    __syncthreads();
    if (fix_ptr[t + 1] > fix_ptr[t]) {
      // one wave per split row: lanes stride over its slots, then an xor tree
      const int col = lane % KP;
      for (int64_t f = fix_ptr[t] + wave; f < fix_ptr[t + 1]; f += nw) {
        const Fixup fxr = fx[f];
        T a = 0;
        for (uint32_t s = fxr.slot0 + lane / KP; s < fxr.slot1; s += kWave / KP) a += part[(int64_t)s * KP + col];
#pragma unroll
        for (int off = KP; off < kWave; off <<= 1) a += __shfl_xor(a, off, kWave);
        if (lane < KP) k.finish(fxr.row, col, a);
      }
      __threadfence();
      __syncthreads();
    }
  }
}

template <class K, typename T>
void run_levels_wg(hipStream_t st, const K& k, const DevSched& sc, const int64_t* item_ptr, const int64_t* fix_ptr,
                   int t0, int t1, int dir, T* part, int kp) {
  if (t1 <= t0) return;
  constexpr int NV16 = 16 / sizeof(T);
#define WA_LEVELS_WG(V, NV) \
  k_levels_wg<K, T, V, NV><<<1, kTailThreads, 0, st>>>(k, sc.items, item_ptr, sc.fix, fix_ptr, part, t0, t1, dir)
  if (kp * (int)sizeof(T) >= 16) {
    switch (kp / NV16) {
      case 1: WA_LEVELS_WG(1, NV16); break;
      case 2: WA_LEVELS_WG(2, NV16); break;
      case 4: WA_LEVELS_WG(4, NV16); break;
      case 8: WA_LEVELS_WG(8, NV16); break;
      default: WA_LEVELS_WG(16, NV16); break;
    }
  } else if (kp == 2) {
    WA_LEVELS_WG(1, 2);
  } else {
    WA_LEVELS_WG(1, 1);
  }
#undef WA_LEVELS_WG
}

}  // namespace wa
