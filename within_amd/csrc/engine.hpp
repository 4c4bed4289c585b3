// engine.hpp -- the row-gather engine shared by every CSR-like kernel.
//
// A "row" is a sum over elements: out(row) = finish(row, sum_e coef(e) *
// sum_{t<QM} X[idx_t(e)]).  Rows are processed over a host-built schedule
// (plan.hpp Schedule): short rows (at most `short_max` elements) are summed by
// one lane group each, in element order; a longer row is a work item for a
// whole wave, and a row longer than one chunk is split into chunks written to
// partial slots that a fix-up sums in slot order.  Every result is therefore
// bitwise reproducible for a given schedule and column count.
//
// Lane mapping (wave64): a block-vector row holds kp values of type T; each
// lane owns NV consecutive values (a 16-byte vector when the row is at least
// 16 bytes), and V = kp / NV lanes form a lane group covering one row, so a
// wave holds S = 64 / V lane groups ("slots").
//  * short rows: slot s of wave w owns row r0 + w * S + s and walks its
//    elements with U gathers in flight;
//  * work items: the S slots stride over the item's elements, and the slot
//    partials are combined by a fixed xor-butterfly at the end of the row.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <type_traits>

#include "kernels.hpp"

namespace wa {

constexpr int kWave = 64;
constexpr int kFixBlock = 1024;

template <typename T, int NV>
struct alignas(sizeof(T) * NV) VecT {
  T v[NV];
};

template <typename T, int NV>
__device__ __forceinline__ VecT<T, NV> vload(const T* p) {
  return *reinterpret_cast<const VecT<T, NV>*>(p);
}

template <typename T, int NV>
__device__ __forceinline__ void vstore(T* p, const VecT<T, NV>& v) {
  *reinterpret_cast<VecT<T, NV>*>(p) = v;
}

// the same store with the non-temporal hint: for streams written once per
// pass and read back by another kernel (the Krylov vectors, the Gramian
// product), which then leave fewer lines in L2
template <typename T, int NV>
__device__ __forceinline__ void vstore_nt(T* p, const VecT<T, NV>& v) {
#pragma unroll
  for (int c = 0; c < NV; ++c) __builtin_nontemporal_store(v.v[c], p + c);
}

// A kind with EMIT = true also receives each element's term coef * X[id]
// (emit(e, col, v)), e.g. to write the restricted vector while summing it.
template <class K, class = void>
struct emits : std::false_type {};
template <class K>
struct emits<K, std::void_t<decltype(K::EMIT)>> : std::integral_constant<bool, K::EMIT> {};

template <typename T, int NV>
__device__ __forceinline__ void vstore(T* p, const T (&a)[NV]) {
  VecT<T, NV> v;
#pragma unroll
  for (int c = 0; c < NV; ++c) v.v[c] = a[c];
  vstore<T, NV>(p, v);
}

// Gathers are issued branch-free: an element past the end of the row loads
// the row's last element again (always in bounds) and its term is dropped by
// a select.  A guarded load would become a branch after which the compiler
// waits for every load in flight, leaving one gather in flight per lane.
//
// In a work item, the wave can also share its index loads: each lane loads
// one element's indices and coefficient, and the lanes that gather the
// element take them with a lane shuffle, so one index load feeds a batch of
// up to 8 gathers per lane.  That pays where the gathers miss the caches (the
// operator rows); where they hit a small, hot vector (the elimination and
// triangular-solve rows) the added shuffle latency and the rounds past the
// row's end cost more, so a kind opts out with SHARED_IDX = false.  The
// summation order is the same either way.
constexpr int kRowU = 4;   // short rows, one index per element: gathers in flight
constexpr int kRowUQ = 2;  // short rows, several indices per element
constexpr int kVecU = 4;   // work items without shared index loads: elements per lane in flight
template <class K, class = void>
struct shares_idx : std::true_type {};
template <class K>
struct shares_idx<K, std::void_t<decltype(K::SHARED_IDX)>> : std::integral_constant<bool, K::SHARED_IDX> {};
// gathers of one shared-index batch: at most 8 gathered vectors in flight per lane
constexpr int coop_batch(int v, int qm) { return v * qm <= 8 ? v : (8 / qm > 0 ? 8 / qm : 1); }

template <class K, typename T, int QM>
__device__ __forceinline__ void coop_take(const uint32_t (&mid)[QM], T mcf, int src, uint32_t (&id)[QM], T& cf) {
#pragma unroll
  for (int t = 0; t < QM; ++t) id[t] = (uint32_t)__shfl((int)mid[t], src, kWave);
  if constexpr (K::COEF) cf = __shfl(mcf, src, kWave);
  else cf = mcf;
}

// add the U terms of one batch in order; terms with ok[u] false are dropped
template <class K, typename T, int U, int QM, int NV>
__device__ __forceinline__ void add_terms(const K& k, const VecT<T, NV> (&g)[U][QM], const T (&cf)[U],
                                          const bool (&ok)[U], const uint32_t (&e)[U], int col, T (&acc)[NV]) {
#pragma unroll
  for (int u = 0; u < U; ++u) {
    VecT<T, NV> term;
#pragma unroll
    for (int c = 0; c < NV; ++c) {
      T s = 0;
#pragma unroll
      for (int t = 0; t < QM; ++t) s += g[u][t].v[c];
      term.v[c] = K::COEF ? cf[u] * s : s;
      acc[c] += ok[u] ? term.v[c] : T(0);
    }
    if constexpr (emits<K>::value)
      if (ok[u]) k.template emit<NV>(e[u], col, term);
  }
}

// one row by one lane group: elements in order
template <class K, typename T, int KP, int NV>
__device__ __forceinline__ void gather_row(const K& k, uint32_t row, uint32_t e0, uint32_t e1, int part,
                                           T (&acc)[NV]) {
  constexpr int QM = K::QM > 0 ? K::QM : 1;
#pragma unroll
  for (int c = 0; c < NV; ++c) acc[c] = 0;
  if (K::QM == 0) return;
  constexpr int U = QM == 1 ? kRowU : kRowUQ;  // gathers in flight
  for (uint32_t e = e0; e < e1; e += U) {
    uint32_t id[U][QM];
    T cf[U];
    bool ok[U];
    uint32_t ee[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      cf[u] = 1;
      ee[u] = e + u;
      ok[u] = ee[u] < e1;
      k.load(row, ok[u] ? ee[u] : e1 - 1, id[u], cf[u]);
    }
    VecT<T, NV> g[U][QM];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int t = 0; t < QM; ++t) g[u][t] = vload<T, NV>(k.X + (int64_t)id[u][t] * KP + part * NV);
    add_terms<K, T, U, QM, NV>(k, g, cf, ok, ee, part * NV, acc);
  }
}

// one row (or chunk) by a whole wave: V lanes per row, S slots over elements;
// slot s sums elements e0 + s, e0 + s + S, e0 + s + 2S, ... in that order
template <class K, typename T, int V, int NV>
__device__ __forceinline__ void gather_vec(const K& k, uint32_t row, uint32_t e0, uint32_t e1, int slot, int part,
                                           T (&acc)[NV]) {
  constexpr int S = kWave / V;
  constexpr int QM = K::QM > 0 ? K::QM : 1;
  constexpr int KP = V * NV;
#pragma unroll
  for (int c = 0; c < NV; ++c) acc[c] = 0;
  if (K::QM == 0) return;
  if constexpr (shares_idx<K>::value && V >= 4) {
    // lane l loads element base + l; in round u, slot s gathers element
    // base + u * S + s, whose indices lane u * S + s holds.  Control flow is
    // uniform over the wave (one item per wave).
    const int lane = threadIdx.x & (kWave - 1);
    constexpr int UB = coop_batch(V, QM);
    for (uint32_t base = e0; base < e1; base += kWave) {
      uint32_t mid[QM];
      T mcf = 1;
      k.load(row, min(base + lane, e1 - 1), mid, mcf);
#pragma unroll
      for (int u0 = 0; u0 < V; u0 += UB) {
        if (base + u0 * S >= e1) break;
        VecT<T, NV> g[UB][QM];
        T cf[UB];
        bool ok[UB];
        uint32_t e[UB];
#pragma unroll
        for (int u = 0; u < UB; ++u) {
          uint32_t id[QM];
          const int src = (u0 + u) * S + slot;
          coop_take<K, T, QM>(mid, mcf, src, id, cf[u]);
          e[u] = base + src;
          ok[u] = e[u] < e1;
#pragma unroll
          for (int t = 0; t < QM; ++t) g[u][t] = vload<T, NV>(k.X + (int64_t)id[t] * KP + part * NV);
        }
        add_terms<K, T, UB, QM, NV>(k, g, cf, ok, e, part * NV, acc);
      }
    }
  } else {
    constexpr int U = kVecU;
    for (uint32_t e = e0 + slot; e < e1; e += U * S) {
      uint32_t id[U][QM];
      T cf[U];
      bool ok[U];
      uint32_t ee[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        cf[u] = 1;
        ee[u] = e + u * S;
        ok[u] = ee[u] < e1;
        k.load(row, ok[u] ? ee[u] : e1 - 1, id[u], cf[u]);
      }
      VecT<T, NV> g[U][QM];
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int t = 0; t < QM; ++t) g[u][t] = vload<T, NV>(k.X + (int64_t)id[u][t] * KP + part * NV);
      add_terms<K, T, U, QM, NV>(k, g, cf, ok, ee, part * NV, acc);
    }
  }
  // combine the S slots (lanes V apart) in a fixed order
#pragma unroll
  for (int off = V; off < kWave; off <<= 1)
#pragma unroll
    for (int c = 0; c < NV; ++c) acc[c] += __shfl_xor(acc[c], off, kWave);
}

// Work of virtual wave g over rows [r0, r1): waves [0, n_sw) sum the short
// rows, S per wave; the remaining waves run one work item each.
template <class K, typename T, int V, int NV>
__device__ __forceinline__ void rows_wave(const K& k, int64_t g, uint32_t r0, uint32_t r1, uint32_t short_max,
                                          int64_t n_sw, const Item* __restrict__ items, int64_t n_items,
                                          T* __restrict__ part) {
  constexpr int S = kWave / V;
  constexpr int KP = V * NV;
  const int lane = threadIdx.x & (kWave - 1);
  const int slot = lane / V, pt = lane % V;
  T acc[NV];
  if (g < n_sw) {
    const int64_t r = (int64_t)r0 + g * S + slot;
    if (r >= (int64_t)r1) return;
    const uint32_t e0 = k.rbegin((uint32_t)r), e1 = k.rend((uint32_t)r);
    if (e1 - e0 > short_max) return;
    const auto q = k.template pre<NV>((uint32_t)r, pt * NV);  // finish inputs, loaded before the gathers
    gather_row<K, T, KP, NV>(k, (uint32_t)r, e0, e1, pt, acc);
    k.template finish<NV>((uint32_t)r, pt * NV, acc, q);
    return;
  }
  g -= n_sw;
  if (g >= n_items) return;
  const Item it = items[g];
  if (it.slot < 0) {
    // one whole row (or none: padding has row0 == row1); its element range
    // [e0, e1) is in the item, so the gathers need no row-pointer load
    if (it.row1 > it.row0) {
      const uint32_t r = it.row0;
      const auto q = k.template pre<NV>(r, pt * NV);
      gather_vec<K, T, V, NV>(k, r, it.e0, it.e1, slot, pt, acc);
      if (slot == 0) k.template finish<NV>(r, pt * NV, acc, q);
    }
  } else {
    gather_vec<K, T, V, NV>(k, it.row0, it.e0, it.e1, slot, pt, acc);
    if (slot == 0)
#pragma unroll
      for (int c = 0; c < NV; ++c) part[(int64_t)it.slot * KP + pt * NV + c] = acc[c];
  }
}

template <class K, typename T, int V, int NV>
__global__ void __launch_bounds__(kBlock) k_rows(K k, uint32_t r0, uint32_t r1, uint32_t short_max, int64_t n_sw,
                                                 const Item* __restrict__ items, int64_t n_items,
                                                 T* __restrict__ part) {
  const int64_t g = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  rows_wave<K, T, V, NV>(k, g, r0, r1, short_max, n_sw, items, n_items, part);
}

// ---- fix-up of one split row by a workgroup of NT threads: threads stride
// over the slots with eight independent sums (eight loads in flight: a row
// of a few thousand slots is latency bound), then a fixed-order tree -------
template <class K, typename T, int NT>
__device__ __forceinline__ void fix_row(const K& k, const Fixup& f, const T* __restrict__ part, int kp, T* sh) {
  constexpr int U = 8;
  const int col = threadIdx.x % kp, str = NT / kp, s0 = threadIdx.x / kp;
  T a[U];
#pragma unroll
  for (int u = 0; u < U; ++u) a[u] = 0;
  // two-stage fix-ups sum their first-stage blocks instead of the chunk slots
  const uint32_t b0 = f.chunk0 == kNoChunks ? f.slot0 : f.chunk0;
  const uint32_t b1 = f.chunk0 == kNoChunks ? f.slot1 : f.chunk0 + (f.slot1 - f.slot0 + kFixChunk - 1) / kFixChunk;
  uint32_t s = b0 + s0;
  for (; s + (U - 1) * str < b1; s += U * str)
#pragma unroll
    for (int u = 0; u < U; ++u) a[u] += part[(int64_t)(s + u * str) * kp + col];
  for (; s < b1; s += str) a[0] += part[(int64_t)s * kp + col];
  sh[threadIdx.x] = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
  __syncthreads();
  for (int w = str / 2; w >= 1; w >>= 1) {
    if (s0 < w) sh[threadIdx.x] += sh[threadIdx.x + w * kp];
    __syncthreads();
  }
  if (s0 == 0) {
    const T a[1] = {sh[threadIdx.x]};
    k.template finish<1>(f.row, col, a, k.template pre<1>(f.row, col));
  }
  __syncthreads();  // sh is reused by the next call
}

// First stage of the long fix-ups: block y of fix-up x sums its kFixChunk
// chunk slots (one column per lane group, a fixed tree) into partial slot
// chunk0 + y.  Fix-ups done in one stage, and blocks past a fix-up's end, exit.
template <typename T>
__global__ void __launch_bounds__(kFixChunk) k_fix_stage1(const Fixup* __restrict__ fx, T* __restrict__ part, int kp) {
  __shared__ T sh[kFixChunk];
  const Fixup f = fx[blockIdx.x];
  if (f.chunk0 == kNoChunks) return;
  const uint32_t s0 = f.slot0 + blockIdx.y * kFixChunk;
  if (s0 >= f.slot1) return;
  const uint32_t s1 = min(f.slot1, s0 + kFixChunk);
  const int col = threadIdx.x % kp, str = kFixChunk / kp, g = threadIdx.x / kp;
  T acc = 0;
  for (uint32_t s = s0 + g; s < s1; s += str) acc += part[(int64_t)s * kp + col];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = str / 2; w >= 1; w >>= 1) {
    if (g < w) sh[threadIdx.x] += sh[threadIdx.x + w * kp];
    __syncthreads();
  }
  if (g == 0) part[(int64_t)(f.chunk0 + blockIdx.y) * kp + col] = sh[col];
}

template <class K, typename T>
__global__ void __launch_bounds__(kFixBlock) k_rows_fix(K k, const Fixup* __restrict__ fx,
                                                        const T* __restrict__ part, int kp) {
  __shared__ T sh[kFixBlock];
  fix_row<K, T, kFixBlock>(k, fx[blockIdx.x], part, kp, sh);
}

inline unsigned engine_blocks(int64_t waves) { return (unsigned)((waves * kWave + kBlock - 1) / kBlock); }

template <class K, typename T, int V, int NV>
void launch_rows(hipStream_t st, const K& k, const DevSched& sc, const Segment& g, T* part) {
  constexpr int S = kWave / V;
  // short-row waves padded to whole XCD rounds, so that item i runs in
  // workgroup (n_sw + i) / 4 with the XCD (i / 4) mod 8 the schedule assumed
  static_assert(kBlock / kWave == (int)kWavesPerGroup, "schedules assume 4 waves per workgroup");
  constexpr int64_t kRound = (int64_t)kXcds * kWavesPerGroup;
  const int64_t raw = ((int64_t)g.r1 - g.r0 + S - 1) / S;
  const int64_t n_sw = (raw + kRound - 1) / kRound * kRound;
  const int64_t ni = g.i1 - g.i0;
  if (n_sw + ni == 0) return;
  k_rows<K, T, V, NV><<<engine_blocks(n_sw + ni), kBlock, 0, st>>>(k, g.r0, g.r1, sc.short_max, n_sw,
                                                                   sc.items + g.i0, ni, part);
}

// Run one segment of a schedule: short rows and items, then the fix-ups.
template <class K, typename T>
void run_seg(hipStream_t st, const K& k, const DevSched& sc, const Segment& g, T* part, int kp) {
  constexpr int NV16 = 16 / sizeof(T);  // values per 16-byte vector
  if (kp * (int)sizeof(T) >= 16) {
    switch (kp / NV16) {  // lanes per row
      case 1: launch_rows<K, T, 1, NV16>(st, k, sc, g, part); break;
      case 2: launch_rows<K, T, 2, NV16>(st, k, sc, g, part); break;
      case 4: launch_rows<K, T, 4, NV16>(st, k, sc, g, part); break;
      case 8: launch_rows<K, T, 8, NV16>(st, k, sc, g, part); break;
      default: launch_rows<K, T, 16, NV16>(st, k, sc, g, part); break;
    }
  } else if (kp == 2) {
    launch_rows<K, T, 1, 2>(st, k, sc, g, part);
  } else {
    launch_rows<K, T, 1, 1>(st, k, sc, g, part);
  }
  if (g.max_chunks > 0)
    k_fix_stage1<T><<<dim3((unsigned)(g.f1 - g.f0), g.max_chunks), kFixChunk, 0, st>>>(sc.fix + g.f0, part, kp);
  if (g.f1 > g.f0) k_rows_fix<K, T><<<(unsigned)(g.f1 - g.f0), kFixBlock, 0, st>>>(k, sc.fix + g.f0, part, kp);
}

// ---- a run of consecutive levels in one launch, on one XCD ------------------
// Levels are the segments of a level schedule, in processing order; each has
// ceil(rows / S) short-row units then one unit per item (no split rows).
// Waves take units in ticket order (ctr[0]); before a unit of level l a wave
// waits until ctr[2 + p] counts every unit of p, the nearest earlier level
// with units.
// A unit runs rows_wave exactly as a per-level launch would, so each row's
// sum is bitwise the same.  Only workgroups on XCD 0 take tickets: they share
// one L2, where a unit's stores are complete (vmcnt 0) before its count is
// published, and a waiting wave drops its L1 before reading them.  A wait
// longer than spin_limit polls sets ctr[1] and goes on (the caller reports
// the error); tickets are taken in order, so every unit waited for is already
// held by a running wave.  spin_limit is in ticks of the 100 MHz real-time
// counter.
constexpr uint32_t kXccIdReg = (3u << 11) | (0u << 6) | 20u;  // hwreg(HW_REG_XCC_ID, 0, 4)

template <class K, typename T, int V, int NV>
__global__ void __launch_bounds__(kBlock) k_level_flow(K k, const FlowLevel* __restrict__ lv, int nl,
                                                       uint32_t short_max, const Item* __restrict__ items,
                                                       T* __restrict__ part, uint32_t* __restrict__ ctr,
                                                       uint32_t spin_limit) {
  if ((__builtin_amdgcn_s_getreg(kXccIdReg) & 15u) != 0u) return;  // workgroup-uniform
  constexpr uint32_t S = kWave / V;
  const int lane = threadIdx.x & (kWave - 1);
  for (;;) {
    uint32_t w = 0;
    if (lane == 0) w = __hip_atomic_fetch_add(&ctr[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    w = __shfl(w, 0, kWave);
    int l = 0;
    uint32_t base = 0, nsw = 0, need = 0;
    int prev = -1;
    for (; l < nl; ++l) {
      nsw = (lv[l].r1 - lv[l].r0 + S - 1) / S;
      const uint32_t nu = nsw + (uint32_t)(lv[l].i1 - lv[l].i0);
      if (w < base + nu) break;
      base += nu;
      if (nu > 0) {  // the nearest earlier level that has units
        need = nu;
        prev = l;
      }
    }
    if (l == nl) return;
    if (prev >= 0) {
      if (lane == 0) {
        // every counter access is a system-scope atomic, so the polls and
        // the increments meet at one point; the wait is bounded by the
        // 100 MHz real-time counter, and gives up at once after an earlier
        // failed wait
        const uint64_t t_end = __builtin_amdgcn_s_memrealtime() + spin_limit;
        while (__hip_atomic_load(&ctr[2 + prev], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) < need) {
          if (__hip_atomic_load(&ctr[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0u) break;
          __builtin_amdgcn_s_sleep(2);
          if (__builtin_amdgcn_s_memrealtime() > t_end) {
            __hip_atomic_fetch_or(&ctr[1], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            break;
          }
        }
      }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    rows_wave<K, T, V, NV>(k, (int64_t)(w - base), lv[l].r0, lv[l].r1, short_max, (int64_t)nsw, items + lv[l].i0,
                           lv[l].i1 - lv[l].i0, part);
    // the unit's stores are complete at L2 before its count is published
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    if (lane == 0) __hip_atomic_fetch_add(&ctr[2 + l], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// units of a level run for kp columns of T (short-row units + items)
template <typename T>
inline uint32_t flow_units(const FlowLevel& f, int kp) {
  const int NV16 = 16 / (int)sizeof(T);
  const int V = kp * (int)sizeof(T) >= 16 ? kp / NV16 : 1;
  const uint32_t S = kWave / V;
  return (f.r1 - f.r0 + S - 1) / S + (uint32_t)(f.i1 - f.i0);
}

constexpr unsigned kFlowBlocks = 512;  // 64 workgroups on each XCD; those on XCD 0 work

template <class K, typename T>
void run_flow(hipStream_t st, const K& k, const DevSched& sc, const FlowLevel* lv, int nl, T* part, uint32_t* ctr,
              int kp, uint32_t spin_limit) {
  constexpr int NV16 = 16 / sizeof(T);
#define WA_FLOW(V_, NV_)                                                                                   \
  k_level_flow<K, T, V_, NV_><<<kFlowBlocks, kBlock, 0, st>>>(k, lv, nl, sc.short_max, sc.items, part, ctr, \
                                                              spin_limit)
  if (kp * (int)sizeof(T) >= 16) {
    switch (kp / NV16) {
      case 1: WA_FLOW(1, NV16); break;
      case 2: WA_FLOW(2, NV16); break;
      case 4: WA_FLOW(4, NV16); break;
      case 8: WA_FLOW(8, NV16); break;
      default: WA_FLOW(16, NV16); break;
    }
  } else if (kp == 2) {
    WA_FLOW(1, 2);
  } else {
    WA_FLOW(1, 1);
  }
#undef WA_FLOW
}

// Run every segment of a schedule (in order).
template <class K, typename T>
void run_rows(hipStream_t st, const K& k, const DevSched& sc, T* part, int kp) {
  for (const Segment& g : sc.segs) run_seg<K, T>(st, k, sc, g, part, kp);
}

// Levels t0..t1-1 of a level schedule, ascending when dir > 0 (one launch
// per level: a grid-wide barrier inside one cooperative launch costs more than
// a launch boundary here).
template <class K, typename T>
void run_levels(hipStream_t st, const K& k, const DevSched& sc, int t0, int t1, int dir, T* part, int kp) {
  for (int step = 0; step < t1 - t0; ++step)
    run_seg<K, T>(st, k, sc, sc.segs[dir > 0 ? t0 + step : t1 - 1 - step], part, kp);
}

}  // namespace wa
