// plan.cpp -- host-side setup (see plan.hpp).  Built once per design; its
// arrays are uploaded once and read by every GPU apply.
#include "plan.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <numeric>
#include <parallel/algorithm>
#include <thread>

namespace wa {

// phase timings on stderr when WITHIN_AMD_VERBOSE is set
static bool verbose() {
  static const bool v = std::getenv("WITHIN_AMD_VERBOSE") != nullptr;  // thread-safe initialisation
  return v;
}
struct PhaseTimer {
  const char* name;
  bool open = true;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  explicit PhaseTimer(const char* n) : name(n) {}
  void done() {
    if (open && verbose())
      std::fprintf(stderr, "[within_amd] %s: %.3f s\n", name,
                   std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    open = false;
  }
  ~PhaseTimer() { done(); }
};

enum { E_EMPTY = 1, E_OVERFLOW = 4, E_SINGULAR = 5, E_LOCAL = 6, E_INVALID = 9 };

// ---------------------------------------------------------------------------
// design -- WeightedDesign::from_store (crates/within/src/domain.rs:133-156)
// ---------------------------------------------------------------------------
Design make_design(const u32* codes, i64 n_obs, int Q, i64 ld, const double* weights) {
  if (n_obs <= 0) throw Error(E_EMPTY, "EmptyObservations: no observations");
  if (Q <= 0) throw Error(E_INVALID, "at least one factor is required");
  if (ld < n_obs) throw Error(E_INVALID, "leading dimension smaller than n_obs");
  if (n_obs >= (i64)UINT32_MAX) throw Error(E_OVERFLOW, "Overflow: n_obs exceeds u32 indexing");
  Design d;
  d.n_obs = n_obs;
  d.Q = Q;
  d.codes.resize(Q);
  // the factors side by side: each copied and its largest code found
  std::vector<u32> mx(Q, 0);
#pragma omp parallel for schedule(dynamic, 1) if (n_obs >= 65536)
  for (int q = 0; q < Q; ++q) {
    const u32* col = codes + (i64)q * ld;
    d.codes[q].assign(col, col + n_obs);
    u32 m = 0;
    for (i64 i = 0; i < n_obs; ++i) m = std::max(m, col[i]);
    mx[q] = m;
  }
  i64 off = 0;
  for (int q = 0; q < Q; ++q) {
    d.levels.push_back((i64)mx[q] + 1);
    d.offset.push_back(off);
    off += (i64)mx[q] + 1;
  }
  if (off >= (i64)UINT32_MAX) throw Error(E_OVERFLOW, "Overflow: n_dofs exceeds u32 indexing");
  d.n_dofs = off;
  if (weights) d.weights.assign(weights, weights + n_obs);
  return d;
}

// ---------------------------------------------------------------------------
// segmented-reduction work items
// ---------------------------------------------------------------------------
void append_schedule(Schedule& s, const std::vector<u32>& row_ptr, u32 r0, u32 r1, u32 chunk,
                     const std::vector<char>* skip) {
  Segment g{r0, r1, (i64)s.items.size(), 0, (i64)s.fixups.size(), 0, 0};
  std::vector<Item> whole;
  std::vector<u32> split;
  for (u32 r = r0; r < r1; ++r) {
    const u32 len = row_ptr[r + 1] - row_ptr[r];
    if (len <= s.short_max || (skip && (*skip)[r])) continue;
    // a whole-row item is one row and carries its element range (the
    // engine reads the range from the item, not from the row pointers)
    if (len <= chunk) whole.push_back({r, r + 1, row_ptr[r], row_ptr[r + 1], -1, 0});
    else split.push_back(r);
  }
  // split rows: each row's slots are consecutive (its fix-up sums them in
  // order); the chunk items are laid out XCD by XCD when several long rows
  // share the segment
  auto nchunks = [&](u32 r) { return (row_ptr[r + 1] - row_ptr[r] + chunk - 1) / chunk; };
  std::stable_sort(split.begin(), split.end(), [&](u32 a, u32 b) { return nchunks(a) > nchunks(b); });
  std::vector<u32> slot0(split.size());
  for (size_t t = 0; t < split.size(); ++t) {
    slot0[t] = (u32)s.n_slots;
    s.n_slots += nchunks(split[t]);
    s.fixups.push_back({split[t], slot0[t], (u32)s.n_slots, kNoChunks});
  }
  // long fix-ups: first-stage partial slots after the chunk slots
  for (size_t t = 0; t < split.size(); ++t) {
    Fixup& f = s.fixups[s.fixups.size() - split.size() + t];
    const u32 len = f.slot1 - f.slot0;
    if (len <= 2 * kFixChunk) continue;
    const u32 nb = (len + kFixChunk - 1) / kFixChunk;
    f.chunk0 = (u32)s.n_slots;
    s.n_slots += nb;
    g.max_chunks = std::max(g.max_chunks, nb);
  }
  auto chunk_item = [&](size_t t, u32 j) {
    const u32 r = split[t];
    const u32 e0 = row_ptr[r] + j * chunk, e1 = std::min(e0 + chunk, row_ptr[r + 1]);
    return Item{r, r + 1, e0, e1, (int32_t)(slot0[t] + j), 0};
  };
  if (split.size() >= kWavesPerGroup) {
    // chunk index j goes to workgroups ≡ j (mod kXcds); rows with more chunks
    // come first, so the rows still active at j are a prefix
    const Item none{0, 0, 0, 0, -1, 0};
    const u32 cmax = nchunks(split[0]);
    size_t act = split.size();
    for (u32 j0 = 0; j0 < cmax; j0 += kXcds) {
      while (act > 0 && nchunks(split[act - 1]) <= j0) --act;
      const size_t per = (act + kWavesPerGroup - 1) / kWavesPerGroup;  // workgroups per chunk index
      for (size_t w = 0; w < per * kXcds; ++w) {
        const u32 j = j0 + (u32)(w % kXcds);
        for (u32 v = 0; v < kWavesPerGroup; ++v) {
          const size_t t = (w / kXcds) * kWavesPerGroup + v;
          s.items.push_back(t < act && j < nchunks(split[t]) ? chunk_item(t, j) : none);
        }
      }
    }
  } else {
    for (size_t t = 0; t < split.size(); ++t)
      for (u32 j = 0; j < nchunks(split[t]); ++j) s.items.push_back(chunk_item(t, j));
  }
  s.items.insert(s.items.end(), whole.begin(), whole.end());
  g.i1 = (i64)s.items.size();
  g.f1 = (i64)s.fixups.size();
  s.segs.push_back(g);
}

Schedule make_schedule(const std::vector<u32>& row_ptr, u32 chunk, u32 short_max, const std::vector<char>* skip) {
  Schedule s;
  s.short_max = short_max;
  append_schedule(s, row_ptr, 0, (u32)row_ptr.size() - 1, chunk, skip);
  return s;
}

// stable counting sort of `idx` by key[idx]
static void counting_sort(std::vector<u32>& idx, const u32* key, i64 nkeys) {
  std::vector<u32> cnt(nkeys + 1, 0);
  for (u32 i : idx) cnt[key[i] + 1]++;
  for (i64 k = 0; k < nkeys; ++k) cnt[k + 1] += cnt[k];
  std::vector<u32> out(idx.size());
  for (u32 i : idx) out[cnt[key[i]]++] = i;
  idx.swap(out);
}

static FactorPass build_pass(const Design& d, int q) {
  FactorPass fp;
  fp.q = q;
  const i64 n = d.n_obs, L = d.levels[q];
  std::vector<u32> idx(n);
  std::iota(idx.begin(), idx.end(), 0u);
  // secondary key: the largest other factor (gather locality inside a level)
  int sec = -1;
  for (int r = 0; r < d.Q; ++r)
    if (r != q && (sec < 0 || d.levels[r] > d.levels[sec])) sec = r;
  if (sec >= 0) counting_sort(idx, d.codes[sec].data(), d.levels[sec]);
  counting_sort(idx, d.codes[q].data(), L);
  fp.row_ptr.assign(L + 1, 0);
  for (i64 i = 0; i < n; ++i) fp.row_ptr[d.codes[q][i] + 1]++;
  for (i64 j = 0; j < L; ++j) fp.row_ptr[j + 1] += fp.row_ptr[j];
  const int Qm = d.Q - 1;
  fp.other.resize((size_t)n * Qm);
  if (!d.weights.empty()) fp.w.resize(n);
#pragma omp parallel for schedule(static)
  for (i64 t = 0; t < n; ++t) {
    const u32 o = idx[t];
    int s = 0;
    for (int r = 0; r < d.Q; ++r)
      if (r != q) fp.other[(size_t)t * Qm + s++] = (u32)(d.offset[r] + d.codes[r][o]);
    if (!d.weights.empty()) fp.w[t] = d.weights[o];
  }
  fp.diag.assign(L, 0.0);
#pragma omp parallel for schedule(dynamic, 1024)
  for (i64 j = 0; j < L; ++j)
    for (u32 t = fp.row_ptr[j]; t < fp.row_ptr[j + 1]; ++t) fp.diag[j] += d.w(idx[t]);
  fp.obs = std::move(idx);
  return fp;
}

// ---------------------------------------------------------------------------
// CSR, cross-tabs, components, partition weights
//   gramian/cross_tab.rs:75-145, :381-501, :514-664; factor_pairs.rs:63-304
// ---------------------------------------------------------------------------
Csr Csr::transpose() const {
  Csr t;
  t.nrows = ncols;
  t.ncols = nrows;
  t.indptr.assign(ncols + 1, 0);
  for (u32 j : indices) t.indptr[j + 1]++;
  for (i64 j = 0; j < ncols; ++j) t.indptr[j + 1] += t.indptr[j];
  t.indices.resize(indices.size());
  t.data.resize(data.size());
  std::vector<u32> cur(t.indptr.begin(), t.indptr.end() - 1);
  for (i64 i = 0; i < nrows; ++i)
    for (u32 p = indptr[i]; p < indptr[i + 1]; ++p) {
      u32 at = cur[indices[p]]++;
      t.indices[at] = (u32)i;
      t.data[at] = data[p];
    }
  return t;
}

static CrossTab cross_tab(const Design& d, int q, int r, const std::vector<u32>& qmap,
                          const std::vector<u32>& rmap, i64 nq, i64 nr) {
  CrossTab ct;
  ct.diag_q.assign(nq, 0.0);
  ct.diag_r.assign(nr, 0.0);
  const u32* cq = d.codes[q].data();
  const u32* cr = d.codes[r].data();
  std::vector<u32> ptr(nq + 1, 0);
  for (i64 i = 0; i < d.n_obs; ++i) {
    u32 a = qmap[cq[i]], b = rmap[cr[i]];
    if (a == UINT32_MAX || b == UINT32_MAX) continue;
    double w = d.w(i);
    ct.diag_q[a] += w;
    ct.diag_r[b] += w;
    ptr[a + 1]++;
  }
  for (i64 j = 0; j < nq; ++j) ptr[j + 1] += ptr[j];
  std::vector<u32> col(ptr[nq]);
  std::vector<double> val(ptr[nq]);
  {
    std::vector<u32> cur(ptr.begin(), ptr.end() - 1);
    for (i64 i = 0; i < d.n_obs; ++i) {
      u32 a = qmap[cq[i]], b = rmap[cr[i]];
      if (a == UINT32_MAX || b == UINT32_MAX) continue;
      u32 at = cur[a]++;
      col[at] = b;
      val[at] = d.w(i);
    }
  }
  Csr& c = ct.c;
  c.nrows = nq;
  c.ncols = nr;
  c.indptr.assign(nq + 1, 0);
  std::vector<double> acc(nr, 0.0);
  std::vector<char> hit(nr, 0);
  std::vector<u32> cols;
  for (i64 row = 0; row < nq; ++row) {
    for (u32 p = ptr[row]; p < ptr[row + 1]; ++p) {
      if (!hit[col[p]]) { hit[col[p]] = 1; cols.push_back(col[p]); }
      acc[col[p]] += val[p];
    }
    std::sort(cols.begin(), cols.end());
    for (u32 k : cols) {
      if (acc[k] != 0.0) { c.indices.push_back(k); c.data.push_back(acc[k]); }
      acc[k] = 0.0;
      hit[k] = 0;
    }
    cols.clear();
    c.indptr[row + 1] = (u32)c.indices.size();
  }
  ct.ct = c.transpose();
  return ct;
}

std::vector<Subdomain> build_subdomains(const Design& d) {
  int Q = d.Q;
  std::vector<std::vector<char>> active(Q);
  for (int q = 0; q < Q; ++q) {
    active[q].assign(d.levels[q], 0);
    for (i64 i = 0; i < d.n_obs; ++i) active[q][d.codes[q][i]] = 1;
  }
  std::vector<std::pair<int, int>> pairs;
  for (int q = 0; q < Q; ++q)
    for (int r = q + 1; r < Q; ++r) pairs.push_back({q, r});
  std::vector<std::vector<Subdomain>> per(pairs.size());
#pragma omp parallel for schedule(dynamic, 1)
  for (size_t pi = 0; pi < pairs.size(); ++pi) {
    int q = pairs[pi].first, r = pairs[pi].second;
    std::vector<u32> qmap(d.levels[q], UINT32_MAX), rmap(d.levels[r], UINT32_MAX), l2g;
    i64 nq = 0, nr = 0;
    for (i64 j = 0; j < d.levels[q]; ++j)
      if (active[q][j]) { qmap[j] = (u32)nq++; l2g.push_back((u32)(d.offset[q] + j)); }
    for (i64 j = 0; j < d.levels[r]; ++j)
      if (active[r][j]) { rmap[j] = (u32)nr++; l2g.push_back((u32)(d.offset[r] + j)); }
    if (nq == 0 || nr == 0) continue;
    CrossTab full = cross_tab(d, q, r, qmap, rmap, nq, nr);
    // bipartite components by DFS (q nodes 0..nq, r nodes nq..)
    i64 n = nq + nr;
    std::vector<char> seen(n, 0);
    std::vector<i64> stack;
    std::vector<std::pair<std::vector<u32>, std::vector<u32>>> comps;
    for (i64 s0 = 0; s0 < n; ++s0) {
      if (seen[s0]) continue;
      seen[s0] = 1;
      stack.push_back(s0);
      std::vector<u32> qi, ri;
      while (!stack.empty()) {
        i64 v = stack.back();
        stack.pop_back();
        if (v < nq) {
          qi.push_back((u32)v);
          for (u32 p = full.c.indptr[v]; p < full.c.indptr[v + 1]; ++p) {
            i64 g = nq + full.c.indices[p];
            if (!seen[g]) { seen[g] = 1; stack.push_back(g); }
          }
        } else {
          ri.push_back((u32)(v - nq));
          for (u32 p = full.ct.indptr[v - nq]; p < full.ct.indptr[v - nq + 1]; ++p) {
            i64 g = full.ct.indices[p];
            if (!seen[g]) { seen[g] = 1; stack.push_back(g); }
          }
        }
      }
      std::sort(qi.begin(), qi.end());
      std::sort(ri.begin(), ri.end());
      comps.push_back({std::move(qi), std::move(ri)});
    }
    for (auto& cm : comps) {
      Subdomain s;
      s.q = q;
      s.r = r;
      if (comps.size() == 1) {
        s.ct = full;
      } else {
        std::vector<u32> rre(nr, UINT32_MAX);
        for (size_t k = 0; k < cm.second.size(); ++k) rre[cm.second[k]] = (u32)k;
        CrossTab& ct = s.ct;
        ct.c.nrows = cm.first.size();
        ct.c.ncols = cm.second.size();
        ct.c.indptr.assign(cm.first.size() + 1, 0);
        for (size_t k = 0; k < cm.first.size(); ++k) {
          u32 o = cm.first[k];
          ct.diag_q.push_back(full.diag_q[o]);
          for (u32 p = full.c.indptr[o]; p < full.c.indptr[o + 1]; ++p)
            if (rre[full.c.indices[p]] != UINT32_MAX) {
              ct.c.indices.push_back(rre[full.c.indices[p]]);
              ct.c.data.push_back(full.c.data[p]);
            }
          ct.c.indptr[k + 1] = (u32)ct.c.indices.size();
        }
        for (u32 o : cm.second) ct.diag_r.push_back(full.diag_r[o]);
        ct.ct = ct.c.transpose();
      }
      for (u32 i : cm.first) s.gidx.push_back(l2g[i]);
      for (u32 i : cm.second) s.gidx.push_back(l2g[nq + i]);
      per[pi].push_back(std::move(s));
    }
  }
  std::vector<Subdomain> subs;
  for (auto& v : per)
    for (auto& s : v) subs.push_back(std::move(s));
  std::vector<u32> cnt(d.n_dofs, 0);
  for (auto& s : subs)
    for (u32 g : s.gidx) cnt[g]++;
  for (auto& s : subs) {
    bool uniq = true;
    for (u32 g : s.gidx) uniq = uniq && cnt[g] <= 1;
    if (!uniq) {
      s.pou.resize(s.gidx.size());
      for (size_t k = 0; k < s.gidx.size(); ++k) s.pou[k] = 1.0 / std::sqrt((double)cnt[s.gidx[k]]);
    }
  }
  return subs;
}

// ---------------------------------------------------------------------------
// clique trees + approximate Cholesky (GKS 2023 clique-tree sampling,
// arXiv:2303.00709, as called at operator/schur_complement.rs:128-139)
//   neighbours split into `split` equal copies, stably sorted by (weight, id);
//   for i < len-1: rem = weight after i; j > i drawn with P ~ w_j;
//   edge (u_i, u_j) with weight w_i * rem / total (total = diag when split == 1,
//   else the weight sum); copies of one neighbour give no edge.
//   RNG: splitmix64 seeded with the star seed, u = (x >> 11) * 2^-53.
// ---------------------------------------------------------------------------
struct Edge { u32 lo, hi; double w; };

static inline uint64_t mix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

static void sample_clique_tree(const std::vector<std::pair<u32, double>>& nbrs, double diag,
                               uint64_t seed, int split, std::vector<Edge>& out) {
  if (split < 1) split = 1;
  std::vector<std::pair<u32, double>> e;
  e.reserve(nbrs.size() * split);
  for (const auto& p : nbrs)
    for (int c = 0; c < split; ++c) e.push_back({p.first, split == 1 ? p.second : p.second / split});
  std::stable_sort(e.begin(), e.end(), [](const std::pair<u32, double>& a, const std::pair<u32, double>& b) {
    return a.second != b.second ? a.second < b.second : a.first < b.first;
  });
  size_t len = e.size();
  if (len < 2) return;
  double total = diag;
  if (split > 1) {
    total = 0.0;
    for (const auto& p : e) total += p.second;
  }
  std::vector<double> suf(len + 1, 0.0);
  for (size_t i = len; i-- > 0;) suf[i] = suf[i + 1] + e[i].second;
  uint64_t st = seed;
  for (size_t i = 0; i + 1 < len; ++i) {
    double rem = suf[i + 1];
    double target = (double)(mix64(st) >> 11) * 0x1.0p-53 * rem;
    size_t lo = i + 1, hi = len - 1;
    while (lo < hi) {
      size_t mid = (lo + hi) / 2;
      if (suf[i + 1] - suf[mid + 1] > target) hi = mid; else lo = mid + 1;
    }
    double w = e[i].second * rem / total;
    u32 a = e[i].first, b = e[lo].first;
    if (a == b || !(w > 0.0)) continue;
    out.push_back({std::min(a, b), std::max(a, b), w});
  }
}

// y[j] -= c * x[j] and y[j] += c * x[j] over a range: the dense-core inner
// loops, built for AVX2 as well as the baseline ISA (chosen at load time).
// Either build does one multiply and one add or subtract per element (no
// contraction), so the results are the same bit for bit.
__attribute__((target_clones("avx2", "default"))) static void axpy_sub(double* __restrict y,
                                                                       const double* __restrict x, double c,
                                                                       i64 n) {
  for (i64 j = 0; j < n; ++j) y[j] -= c * x[j];
}
__attribute__((target_clones("avx2", "default"))) static void axpy_add(double* __restrict y,
                                                                       const double* __restrict x, double c,
                                                                       i64 n) {
  for (i64 j = 0; j < n; ++j) y[j] += c * x[j];
}

// Cholesky of the anchored minor (drop the last index) of a dense n x n
// Laplacian-like matrix, then its inverse.  Entry by entry the arithmetic is
// fixed:
//   l_ij = (a_ij - sum_{k<j} l_ik l_jk) / l_jj,  l_jj = sqrt(a_jj - sum_{k<j} l_jk^2),
//   U = L^-1:  U_ij = (delta_ij - sum_{k=j}^{i-1} l_ik U_kj) / l_ii,
//   inv_ij = sum_{k >= max(i,j)} U_ki U_kj,
// every sum taken term by term for k ascending from 0 (from j, from
// max(i,j)).  The loops below are blocked and vectorised across entries, which
// leaves each entry's sequence of operations unchanged (a term whose factor is
// an exact zero leaves a nonzero or +0 partial sum unchanged).  Empty result =
// minor not positive definite.
std::vector<double> anchored_inverse(const std::vector<double>& a, i64 n) {
  std::vector<double> inv(n * n, 0.0);
  const i64 m = n - 1;
  if (m <= 0) return inv;
  std::vector<double> l(m * m, 0.0);
  for (i64 i = 0; i < m; ++i)
    for (i64 j = 0; j <= i; ++j) l[i * m + j] = a[i * n + j];
  // 1. Cholesky, right-looking over panels of nb columns: the trailing update
  //    subtracts a panel's products from each entry in k order.
  constexpr i64 nb = 32, jt = 512;
  std::vector<double> pt;
  for (i64 k0 = 0; k0 < m; k0 += nb) {
    const i64 k1 = std::min(k0 + nb, m);
    for (i64 j = k0; j < k1; ++j) {
      double s = l[j * m + j];
      for (i64 k = k0; k < j; ++k) s -= l[j * m + k] * l[j * m + k];
      if (!(s > 0.0)) return {};
      const double ljj = std::sqrt(s);
      l[j * m + j] = ljj;
      for (i64 i = j + 1; i < k1; ++i) {
        double t = l[i * m + j];
        for (i64 k = k0; k < j; ++k) t -= l[i * m + k] * l[j * m + k];
        l[i * m + j] = t / ljj;
      }
    }
    if (k1 == m) break;
#pragma omp parallel for schedule(static) if (m - k1 > 256)
    for (i64 i = k1; i < m; ++i)
      for (i64 j = k0; j < k1; ++j) {
        double t = l[i * m + j];
        for (i64 k = k0; k < j; ++k) t -= l[i * m + k] * l[j * m + k];
        l[i * m + j] = t / l[j * m + j];
      }
    pt.assign((size_t)(k1 - k0) * m, 0.0);  // pt[kk][j] = l_j,(k0+kk)
    for (i64 j = k1; j < m; ++j)
      for (i64 kk = 0; kk < k1 - k0; ++kk) pt[kk * m + j] = l[j * m + k0 + kk];
#pragma omp parallel for schedule(dynamic, 16) if (m - k1 > 256)
    for (i64 i = k1; i < m; ++i) {
      double* __restrict li = &l[i * m];
      for (i64 j0 = k1; j0 <= i; j0 += jt) {
        const i64 j1 = std::min(j0 + jt, i + 1);
        for (i64 kk = 0; kk < k1 - k0; ++kk) {
          const double c = li[k0 + kk];
          const double* __restrict p = &pt[kk * m];
          axpy_sub(li + j0, p + j0, c, j1 - j0);
        }
      }
    }
  }
  // 2. U = L^-1 by forward substitution, columns in blocks of jb (one block
  //    per task, rows vectorised across the block's columns).
  constexpr i64 jb = 32;
  std::vector<double> u(m * m, 0.0);
  const i64 nblk = (m + jb - 1) / jb;
#pragma omp parallel if (m > 256)
  {
    std::vector<double> ub;  // ub[i][jj] = U_i,(j0+jj) for i >= j0
    std::vector<double> s(jb);
#pragma omp for schedule(dynamic, 1)
    for (i64 bi = 0; bi < nblk; ++bi) {
      const i64 j0 = bi * jb, w = std::min(jb, m - j0);
      ub.assign((size_t)(m - j0) * jb, 0.0);
      for (i64 i = j0; i < m; ++i) {
        for (i64 jj = 0; jj < jb; ++jj) s[jj] = (jj < w && i == j0 + jj) ? 1.0 : 0.0;
        const double* __restrict li = &l[i * m];
        double* __restrict sp = s.data();
        for (i64 k = j0; k < i; ++k) {
          const double c = li[k];
          const double* __restrict uk = &ub[(k - j0) * jb];
          axpy_sub(sp, uk, c, jb);
        }
        const double lii = li[i];
        double* ui = &ub[(i - j0) * jb];
        for (i64 jj = 0; jj < w; ++jj) ui[jj] = (i >= j0 + jj) ? s[jj] / lii : 0.0;
      }
      for (i64 i = j0; i < m; ++i)
        for (i64 jj = 0; jj < w && j0 + jj <= i; ++jj) u[i * m + j0 + jj] = ub[(i - j0) * jb + jj];
    }
  }
  // 3. inv_ij (j <= i) = sum_{k >= i} U_ki U_kj, rows in blocks of ib; mirrored.
  constexpr i64 ib = 32;
  const i64 nrb = (m + ib - 1) / ib;
#pragma omp parallel if (m > 256)
  {
    std::vector<double> acc;
#pragma omp for schedule(dynamic, 1)
    for (i64 bi = 0; bi < nrb; ++bi) {
      const i64 i0 = bi * ib, i1 = std::min(i0 + ib, m);
      acc.assign((size_t)(i1 - i0) * i1, 0.0);  // acc[i - i0][j], j <= i
      for (i64 k = i0; k < m; ++k) {
        const double* __restrict uk = &u[k * m];
        for (i64 i = i0; i < std::min(i1, k + 1); ++i) {
          const double c = uk[i];
          double* __restrict ai = &acc[(i - i0) * i1];
          axpy_add(ai, uk, c, i + 1);
        }
      }
      for (i64 i = i0; i < i1; ++i)
        for (i64 j = 0; j <= i; ++j) {
          const double v = acc[(i - i0) * i1 + j];
          inv[i * n + j] = v;
          inv[j * n + i] = v;
        }
    }
  }
  return inv;
}

static inline uint64_t round_key(uint64_t seed, uint64_t round, uint64_t v) {
  uint64_t s = seed ^ (round * 0xD1B54A32D192ED03ULL) ^ (v * 0x8CB92BA72F3D8DD7ULL);
  return mix64(s);
}

// Rounds of independent eliminations on a Laplacian (off-diagonals = -w):
// round t takes a greedy maximal independent set, visiting vertices in
// (degree, key(t, v), v) order; fill = clique tree of its neighbourhood (seed
// + v), merged after the round in (v ascending, emission) order.  Once at most
// `dense_core` vertices remain they are factored exactly (anchored at the
// last one); a core whose anchored minor is singular falls back to rounds.
static Factor factor_laplacian(const Csr& lap, uint64_t seed, int split, i64 dense_core, bool device,
                               bool report = false) {
  const auto t_start = std::chrono::steady_clock::now();
  auto secs = [&] { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count(); };
  using Nbr = std::pair<u32, double>;
  Factor f;
  i64 n = lap.nrows;
  f.n = n;
  // adjacency lists sorted by neighbour (a repeated column keeps its last value)
  std::vector<std::vector<Nbr>> adj(n);
#pragma omp parallel for schedule(dynamic, 256) if (n >= 4096)
  for (i64 i = 0; i < n; ++i) {
    std::vector<Nbr>& row = adj[i];
    for (u32 p = lap.indptr[i]; p < lap.indptr[i + 1]; ++p) {
      double w = -lap.data[p];
      if ((i64)lap.indices[p] != i && w > 0.0) row.push_back({lap.indices[p], w});
    }
    std::stable_sort(row.begin(), row.end(), [](const Nbr& x, const Nbr& y) { return x.first < y.first; });
    size_t o = 0;
    for (size_t k = 0; k < row.size(); ++k) {
      if (o > 0 && row[o - 1].first == row[k].first) row[o - 1] = row[k];
      else row[o++] = row[k];
    }
    row.resize(o);
  }
  f.pivot.assign(n, 0.0);
  std::vector<std::vector<Nbr>> cols(n);
  std::vector<char> live(n, 1);
  std::vector<uint64_t> key(n), taken(n, 0), dirty(n, 0);
  i64 left = n, cut = dense_core > 1 ? dense_core : 0;
  uint64_t round = 0;
  f.round_ptr.push_back(0);
  std::vector<u32> pick, order, touched;
  std::vector<std::vector<Edge>> fills;
  std::vector<i64> acnt(n, 0), abeg(n, 0);  // per touched vertex: its fill additions
  // a fill addition to a vertex, tagged with the edge's place in the round's
  // emission sequence (picked vertices in order, each one's edges in order)
  struct FillAdd {
    u32 nbr, seq;
    double w;
  };
  std::vector<FillAdd> adds;
  std::vector<u32> fill_at;  // per picked vertex: its first edge's place
  while (left > 0) {
    if (left <= cut) {
      std::vector<u32> core;
      for (i64 v = 0; v < n; ++v)
        if (live[v]) core.push_back((u32)v);
      i64 c = (i64)core.size();
      std::vector<i64> at(n, -1);
      for (i64 i = 0; i < c; ++i) at[core[i]] = i;
      // the core's Laplacian: its rows' neighbours (core numbering) and weights,
      // and its diagonal, the sum of the row's weights in adjacency order
      std::vector<u32> cptr(c + 1, 0);
      for (i64 i = 0; i < c; ++i) cptr[i + 1] = cptr[i] + (u32)adj[core[i]].size();
      std::vector<u32> ccol(cptr[c]);
      std::vector<double> cw(cptr[c]), cdiag(c);
#pragma omp parallel for schedule(static) if (c >= 1024)
      for (i64 i = 0; i < c; ++i) {
        double sum = 0.0;
        u32 o = cptr[i];
        for (const Nbr& kv : adj[core[i]]) {
          sum += kv.second;
          ccol[o] = (u32)at[kv.first];
          cw[o++] = kv.second;
        }
        cdiag[i] = sum;
      }
      const double t_rounds = secs();
      std::vector<double> inv;
      std::shared_ptr<DeviceDense> dinv;
      if (device && c >= kDeviceCore) {
        if (!device_anchored_inverse(c, cptr.data(), ccol.data(), cw.data(), cdiag.data(), dinv)) dinv.reset();
      } else {
        std::vector<double> a((size_t)c * c, 0.0);
        for (i64 i = 0; i < c; ++i) {
          for (u32 o = cptr[i]; o < cptr[i + 1]; ++o) a[i * c + ccol[o]] = -cw[o];
          a[i * c + i] = cdiag[i];
        }
        inv = anchored_inverse(a, c);
      }
      if (report)
        std::fprintf(stderr, "[within_amd]     approximate Cholesky: %llu rounds %.3f s, dense core %lld: %.3f s\n",
                     (unsigned long long)round, t_rounds, (long long)c, secs() - t_rounds);
      if (!inv.empty() || dinv) {
        f.core = std::move(core);
        f.core_inv = std::move(inv);
        f.core_dev = std::move(dinv);
        break;
      }
      cut = 0;
    }
    // greedy maximal independent set in (degree, key, id) order
    pick.clear();
    order.clear();
    for (i64 v = 0; v < n; ++v)
      if (live[v]) { key[v] = round_key(seed, round, (uint64_t)v); order.push_back((u32)v); }
    // a total order (ids break ties): the parallel sort gives std::sort's result
    __gnu_parallel::sort(order.begin(), order.end(), [&](u32 a, u32 b) {
      size_t da = adj[a].size(), db = adj[b].size();
      return da != db ? da < db : (key[a] != key[b] ? key[a] < key[b] : a < b);
    });
    for (u32 v : order) {
      if (taken[v] == round + 1) continue;
      pick.push_back(v);
      taken[v] = round + 1;
      for (const Nbr& kv : adj[v]) taken[kv.first] = round + 1;
    }
    std::sort(pick.begin(), pick.end());
    // eliminate the set: the picked vertices share no edge, so each sees the
    // round's starting adjacency; pivots and clique trees are independent
    const i64 np = (i64)pick.size();
    if ((i64)fills.size() < np) fills.resize(np);
#pragma omp parallel for schedule(dynamic, 64) if (np >= 1024)
    for (i64 t = 0; t < np; ++t) {
      const u32 v = pick[t];
      const std::vector<Nbr>& nb = adj[v];
      double dv = 0.0;
      for (const Nbr& q : nb) dv += q.second;
      f.pivot[v] = dv;
      fills[t].clear();
      if (nb.size() >= 2) sample_clique_tree(nb, dv, seed + (uint64_t)v, split, fills[t]);
    }
    // neighbours of the set lose their edge to it and gain the fill, added
    // per entry in (v ascending, emission) order: the additions are scattered
    // in parallel and each vertex's are then sorted by (neighbour, place in the
    // emission sequence), which is the sequential order of the additions
    const bool par = np >= 1024;
    touched.clear();
#pragma omp parallel if (par)
    {
      std::vector<u32> mine;
#pragma omp for schedule(dynamic, 64) nowait
      for (i64 t = 0; t < np; ++t)
        for (const Nbr& q : adj[pick[t]])
          if (__atomic_exchange_n(&dirty[q.first], round + 1, __ATOMIC_RELAXED) != round + 1) mine.push_back(q.first);
#pragma omp critical
      touched.insert(touched.end(), mine.begin(), mine.end());
    }
    const i64 nt = (i64)touched.size();
    for (i64 k = 0; k < nt; ++k) acnt[touched[k]] = 0;
#pragma omp parallel for schedule(dynamic, 64) if (par)
    for (i64 t = 0; t < np; ++t)
      for (const Edge& e : fills[t]) {
        __atomic_fetch_add(&acnt[e.lo], 1, __ATOMIC_RELAXED);
        __atomic_fetch_add(&acnt[e.hi], 1, __ATOMIC_RELAXED);
      }
    i64 run = 0;
    for (u32 u : touched) {  // acnt becomes the write cursor of u's range
      abeg[u] = run;
      run += acnt[u];
      acnt[u] = abeg[u];
    }
    adds.resize(run);
    fill_at.resize(np + 1);
    fill_at[0] = 0;
    for (i64 t = 0; t < np; ++t) fill_at[t + 1] = fill_at[t] + (u32)fills[t].size();
#pragma omp parallel for schedule(dynamic, 64) if (par)
    for (i64 t = 0; t < np; ++t) {
      u32 sq = fill_at[t];
      for (const Edge& e : fills[t]) {
        adds[__atomic_fetch_add(&acnt[e.lo], 1, __ATOMIC_RELAXED)] = {e.hi, sq, e.w};
        adds[__atomic_fetch_add(&acnt[e.hi], 1, __ATOMIC_RELAXED)] = {e.lo, sq, e.w};
        ++sq;
      }
    }
    for (i64 t = 0; t < np; ++t) {
      const u32 v = pick[t];
      cols[v] = std::move(adj[v]);
      adj[v] = std::vector<Nbr>();
      live[v] = 0;
      f.order.push_back(v);
    }
#pragma omp parallel if (nt >= 1024)
    {
      std::vector<Nbr> merged;
      auto before = [](const FillAdd& x, const FillAdd& y) { return x.nbr != y.nbr ? x.nbr < y.nbr : x.seq < y.seq; };
#pragma omp for schedule(dynamic, 64)
      for (i64 k = 0; k < nt; ++k) {
        const u32 u = touched[k];
        FillAdd* a0 = adds.data() + abeg[u];
        FillAdd* a1 = adds.data() + acnt[u];
        if (a1 - a0 <= 16) {
          for (FillAdd* x = a0 + 1; x < a1; ++x)
            for (FillAdd* y = x; y > a0 && before(*y, *(y - 1)); --y) std::swap(*y, *(y - 1));
        } else {
          std::sort(a0, a1, before);
        }
        const std::vector<Nbr>& old = adj[u];
        merged.clear();
        size_t i = 0;
        const FillAdd* p = a0;
        while (i < old.size() || p < a1) {
          if (i < old.size() && !live[old[i].first]) { ++i; continue; }
          if (p == a1 || (i < old.size() && old[i].first < p->nbr)) {
            merged.push_back(old[i++]);
            continue;
          }
          const u32 j = p->nbr;
          double acc = 0.0;
          if (i < old.size() && old[i].first == j) acc = old[i++].second;
          for (; p < a1 && p->nbr == j; ++p) acc += p->w;
          merged.push_back({j, acc});
        }
        adj[u].assign(merged.begin(), merged.end());
      }
    }
    left -= np;
    f.round_ptr.push_back((u32)f.order.size());
    ++round;
  }
  f.col_ptr.assign(n + 1, 0);
  for (i64 v = 0; v < n; ++v) f.col_ptr[v + 1] = f.col_ptr[v] + (u32)cols[v].size();
  f.col_idx.reserve(f.col_ptr[n]);
  f.col_w.reserve(f.col_ptr[n]);
  for (i64 v = 0; v < n; ++v)
    for (const auto& p : cols[v]) { f.col_idx.push_back(p.first); f.col_w.push_back(p.second); }
  return f;
}

// ---------------------------------------------------------------------------
// Schur complements (operator/schur_complement.rs:170-216, :331-447, :534-586,
// :697-711) and the local-solver build (operator/schwarz.rs:207-303)
// ---------------------------------------------------------------------------
struct ElimView {
  bool eq;
  i64 n_keep, n_elim;
  const std::vector<double>* dkeep;
  const std::vector<double>* delim;
  const Csr* k2e;
  const Csr* e2k;
};

static ElimView elim_view(const CrossTab& ct, std::vector<double>& inv) {
  ElimView e;
  e.eq = ct.n_q() >= ct.n_r();
  e.n_keep = e.eq ? ct.n_r() : ct.n_q();
  e.n_elim = e.eq ? ct.n_q() : ct.n_r();
  e.delim = e.eq ? &ct.diag_q : &ct.diag_r;
  e.dkeep = e.eq ? &ct.diag_r : &ct.diag_q;
  e.k2e = e.eq ? &ct.ct : &ct.c;
  e.e2k = e.eq ? &ct.c : &ct.ct;
  inv.clear();
  for (i64 i = 0; i < e.n_elim; ++i) {
    double v = (*e.delim)[i];
    if (!(v > 0.0))
      throw Error(E_SINGULAR, "SingularDiagonal: eliminated block " + std::string(e.eq ? "q" : "r") +
                                  " has a zero diagonal at index " + std::to_string(i));
    inv.push_back(1.0 / v);
  }
  return e;
}

static Csr schur_exact(const ElimView& e, const std::vector<double>& inv) {
  Csr S;
  S.nrows = S.ncols = e.n_keep;
  S.indptr.assign(e.n_keep + 1, 0);
  std::vector<double> acc(e.n_keep, 0.0);
  std::vector<char> hit(e.n_keep, 0);
  std::vector<u32> cols;
  for (i64 i = 0; i < e.n_keep; ++i) {
    acc[i] = (*e.dkeep)[i];
    hit[i] = 1;
    cols.push_back((u32)i);
    for (u32 p = e.k2e->indptr[i]; p < e.k2e->indptr[i + 1]; ++p) {
      u32 k = e.k2e->indices[p];
      double sc = e.k2e->data[p] * inv[k];
      for (u32 q = e.e2k->indptr[k]; q < e.e2k->indptr[k + 1]; ++q) {
        u32 j = e.e2k->indices[q];
        if (!hit[j]) { hit[j] = 1; cols.push_back(j); }
        acc[j] -= sc * e.e2k->data[q];
      }
    }
    std::sort(cols.begin(), cols.end());
    for (u32 j : cols) {
      if (acc[j] != 0.0 || (i64)j == i) { S.indices.push_back(j); S.data.push_back(acc[j]); }
      acc[j] = 0.0;
      hit[j] = 0;
    }
    cols.clear();
    S.indptr[i + 1] = (u32)S.indices.size();
  }
  return S;
}

static Csr schur_sampled(const ElimView& e, uint64_t seed, int split) {
  i64 n = e.n_keep;
  // clique trees of every eliminated star, in star order (blocks of stars
  // sampled in parallel and concatenated in block order)
  const i64 nblk = std::max<i64>(1, std::min<i64>(256, e.n_elim / 4096));
  std::vector<std::vector<Edge>> part(nblk);
#pragma omp parallel for schedule(dynamic, 1) if (nblk > 1)
  for (i64 bk = 0; bk < nblk; ++bk) {
    std::vector<std::pair<u32, double>> nb;
    const i64 k0 = e.n_elim * bk / nblk, k1 = e.n_elim * (bk + 1) / nblk;
    for (i64 k = k0; k < k1; ++k) {
      u32 b = e.e2k->indptr[k], en = e.e2k->indptr[k + 1];
      if (en - b <= 1) continue;
      nb.clear();
      for (u32 p = b; p < en; ++p) nb.push_back({e.e2k->indices[p], e.e2k->data[p]});
      sample_clique_tree(nb, (*e.delim)[k], seed + (uint64_t)k, split, part[bk]);
    }
  }
  // a stable sort by (lo, hi): a stable counting sort by lo, then each lo
  // bucket stably by hi; duplicates then summed in emission order
  std::vector<i64> lo_ptr(n + 1, 0);
  for (const auto& pe : part)
    for (const Edge& x : pe) lo_ptr[x.lo + 1]++;
  for (i64 i = 0; i < n; ++i) lo_ptr[i + 1] += lo_ptr[i];
  std::vector<Edge> edges(lo_ptr[n]);
  {
    std::vector<i64> cur(lo_ptr.begin(), lo_ptr.end() - 1);
    for (const auto& pe : part)
      for (const Edge& x : pe) edges[cur[x.lo]++] = x;
  }
  part.clear();
  std::vector<i64> m_cnt(n, 0);
#pragma omp parallel for schedule(dynamic, 64) if (n >= 256)
  for (i64 i = 0; i < n; ++i) {
    Edge* a = edges.data() + lo_ptr[i];
    Edge* b = edges.data() + lo_ptr[i + 1];
    std::stable_sort(a, b, [](const Edge& x, const Edge& y) { return x.hi < y.hi; });
    i64 o = 0;
    for (Edge* x = a; x < b; ++x) {
      if (o > 0 && a[o - 1].hi == x->hi) a[o - 1].w += x->w;
      else a[o++] = *x;
    }
    m_cnt[i] = o;
  }
  std::vector<Edge> m;
  m.reserve(edges.size());
  for (i64 i = 0; i < n; ++i) m.insert(m.end(), edges.begin() + lo_ptr[i], edges.begin() + lo_ptr[i] + m_cnt[i]);
  edges.clear();
  std::vector<u32> lo_cnt(n, 0), up_cnt(n, 0);
  std::vector<double> dg(n, 0.0);
  for (const Edge& x : m) {
    up_cnt[x.lo]++;
    lo_cnt[x.hi]++;
    dg[x.lo] += x.w;
    dg[x.hi] += x.w;
  }
  Csr S;
  S.nrows = S.ncols = n;
  S.indptr.assign(n + 1, 0);
  for (i64 i = 0; i < n; ++i) S.indptr[i + 1] = S.indptr[i] + lo_cnt[i] + 1 + up_cnt[i];
  S.indices.resize(S.indptr[n]);
  S.data.resize(S.indptr[n]);
  std::vector<u32> lc(n), uc(n);
  for (i64 i = 0; i < n; ++i) {
    lc[i] = S.indptr[i];
    u32 dp = S.indptr[i] + lo_cnt[i];
    S.indices[dp] = (u32)i;
    S.data[dp] = dg[i];
    uc[i] = dp + 1;
  }
  for (const Edge& x : m) {
    u32 p = uc[x.lo]++;
    S.indices[p] = x.hi;
    S.data[p] = -x.w;
    p = lc[x.hi]++;
    S.indices[p] = x.lo;
    S.data[p] = -x.w;
  }
  return S;
}

void build_local_solver(Subdomain& s, const LocalConfig& cfg) {
  ElimView e = elim_view(s.ct, s.inv_diag_elim);
  s.eliminate_q = e.eq;
  if (cfg.dense_threshold > 0 && e.n_keep <= cfg.dense_threshold) {
    // dense fast path: exact Schur, anchored at the last kept index
    Csr S = schur_exact(e, s.inv_diag_elim);
    i64 c = e.n_keep;
    std::vector<double> a(c * c, 0.0);
    for (i64 i = 0; i < c; ++i)
      for (u32 p = S.indptr[i]; p < S.indptr[i + 1]; ++p) a[i * c + S.indices[p]] = S.data[p];
    std::vector<double> inv = anchored_inverse(a, c);
    if (!inv.empty()) {
      s.dense = true;
      s.f = Factor();
      s.f.n = c;
      s.f.round_ptr.push_back(0);
      s.f.pivot.assign(c, 0.0);
      s.f.col_ptr.assign(c + 1, 0);
      for (i64 i = 0; i < c; ++i) s.f.core.push_back((u32)i);
      s.f.core_inv = std::move(inv);
      return;
    }
  }
  const bool big = e.n_keep >= 50000;
  auto t0 = std::chrono::steady_clock::now();
  Csr S = cfg.approx_schur ? schur_sampled(e, cfg.schur_seed, cfg.schur_split) : schur_exact(e, s.inv_diag_elim);
  if (big && verbose())
    std::fprintf(stderr, "[within_amd]     schur complement (n=%lld nnz=%zu): %.3f s\n", (long long)e.n_keep,
                 S.indices.size(), std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
  s.f = factor_laplacian(S, cfg.ac_seed, cfg.ac_split, cfg.dense_core, cfg.device, big && verbose());
}

// ---------------------------------------------------------------------------
// explicit Gramian -- gramian/explicit.rs:25-108, :307-409 (sorted rows, diag
// kept, zero off-diagonals dropped)
// ---------------------------------------------------------------------------
Csr build_gramian(const Design& d, const std::vector<FactorPass>& passes) {
  // Row a (level j of factor q) = its diagonal (sum of weights) and, per other
  // factor r, the weight sums of the levels of r that share observations with
  // it.  Read off the level-sorted passes: the elements of a row are ordered by
  // the code of q's largest other factor, so that factor's sums are runs; the
  // remaining factors' sums go through a dense per-thread accumulator.
  Csr g;
  g.nrows = g.ncols = d.n_dofs;
  g.indptr.assign(d.n_dofs + 1, 0);
  const int Q = d.Q, Qm = Q - 1;
  i64 max_levels = 0;
  for (int q = 0; q < Q; ++q) max_levels = std::max(max_levels, d.levels[q]);
  std::vector<std::vector<std::pair<u32, double>>> rows(d.n_dofs);
  for (int q = 0; q < Q; ++q) {
    const FactorPass& fp = passes[q];
    int sec = -1;
    for (int r = 0; r < Q; ++r)
      if (r != q && (sec < 0 || d.levels[r] > d.levels[sec])) sec = r;
    const int ts = sec < 0 ? -1 : (sec < q ? sec : sec - 1);
    const i64 nl = d.levels[q];
    const int chunk = (int)std::max<i64>(1, std::min<i64>(64, nl / 256));
#pragma omp parallel
    {
      std::vector<double> acc(Qm > 1 ? max_levels : 0, 0.0);
      std::vector<char> hit(Qm > 1 ? max_levels : 0, 0);
      std::vector<u32> touched;
      std::vector<std::pair<u32, double>> blk;
#pragma omp for schedule(dynamic, chunk)
      for (i64 j = 0; j < nl; ++j) {
        const u32 e0 = fp.row_ptr[j], e1 = fp.row_ptr[j + 1];
        if (e0 == e1) continue;
        const u32 a = (u32)(d.offset[q] + j);
        auto& row = rows[a];
        // blocks in factor order; the diagonal sits in q's own block
        for (int r = 0, t = 0; r < Q; ++r) {
          if (r == q) {
            row.push_back({a, fp.diag[j]});
            continue;
          }
          const int tt = t++;
          blk.clear();
          if (tt == ts) {
            for (u32 e = e0; e < e1;) {
              const u32 c = fp.other[(size_t)e * Qm + tt];
              double sum = 0.0;
              for (; e < e1 && fp.other[(size_t)e * Qm + tt] == c; ++e) sum += fp.w.empty() ? 1.0 : fp.w[e];
              if (sum != 0.0) blk.push_back({c, sum});
            }
          } else {
            touched.clear();
            for (u32 e = e0; e < e1; ++e) {
              const u32 c = fp.other[(size_t)e * Qm + tt];
              const u32 l = c - (u32)d.offset[r];
              if (!hit[l]) { hit[l] = 1; touched.push_back(l); }
              acc[l] += fp.w.empty() ? 1.0 : fp.w[e];
            }
            std::sort(touched.begin(), touched.end());
            for (u32 l : touched) {
              if (acc[l] != 0.0) blk.push_back({(u32)(d.offset[r] + l), acc[l]});
              acc[l] = 0.0;
              hit[l] = 0;
            }
          }
          row.insert(row.end(), blk.begin(), blk.end());
        }
      }
    }
  }
  for (i64 a = 0; a < d.n_dofs; ++a) g.indptr[a + 1] = g.indptr[a] + (u32)rows[a].size();
  g.indices.resize(g.indptr[d.n_dofs]);
  g.data.resize(g.indptr[d.n_dofs]);
#pragma omp parallel for schedule(static)
  for (i64 a = 0; a < d.n_dofs; ++a) {
    u32 p = g.indptr[a];
    for (const auto& kv : rows[a]) { g.indices[p] = kv.first; g.data[p] = kv.second; ++p; }
  }
  return g;
}

// ---------------------------------------------------------------------------
// flatten subdomains into concatenated device arrays
// ---------------------------------------------------------------------------
static void flatten(Plan& p) {
  PrecondFlat& f = p.pre;
  const auto& subs = p.subs;
  f.n_sub = (i64)subs.size();
  f.loc_ptr.assign(1, 0);
  f.red_ptr.assign(1, 0);
  for (const auto& s : subs) {
    f.loc_ptr.push_back(f.loc_ptr.back() + (u32)s.n_local());
    f.red_ptr.push_back(f.red_ptr.back() + (u32)s.n_keep());
  }
  f.n_loc = f.loc_ptr.back();
  f.n_red = f.red_ptr.back();
  f.loc_gidx.resize(f.n_loc);
  f.loc_scale.resize(f.n_loc);
  f.loc_sub.resize(f.n_loc);
  f.red_sub.resize(f.n_red);
  f.red_loc.resize(f.n_red);
  f.inv_pivot.assign(f.n_red, 0.0);
  i64 max_rounds = 0;
  for (const auto& s : subs) max_rounds = std::max<i64>(max_rounds, (i64)s.f.round_ptr.size() - 1);
  // keep rows and elimination rows, subdomain by subdomain: row and nonzero
  // offsets first, then every subdomain's rows filled in parallel
  std::vector<i64> er0(f.n_sub + 1, 0), ez0(f.n_sub + 1, 0), kz0(f.n_sub + 1, 0);
  for (i64 si = 0; si < f.n_sub; ++si) {
    const Subdomain& s = subs[si];
    const Csr& e2k = s.eliminate_q ? s.ct.c : s.ct.ct;
    const Csr& k2e = s.eliminate_q ? s.ct.ct : s.ct.c;
    er0[si + 1] = er0[si] + e2k.nrows;
    ez0[si + 1] = ez0[si] + (i64)e2k.indices.size();
    kz0[si + 1] = kz0[si] + (i64)k2e.indices.size();
    f.c_nnz += (i64)s.ct.c.indices.size();
    if (s.dense) f.dense_subs++;
  }
  f.elim_loc.resize(er0[f.n_sub]);
  f.elim_inv.resize(er0[f.n_sub]);
  f.elim_ptr.resize(er0[f.n_sub] + 1);
  f.elim_ptr[0] = 0;
  f.elim_idx.resize(ez0[f.n_sub]);
  f.elim_val.resize(ez0[f.n_sub]);
  f.keep_ptr.resize(f.n_red + 1);
  f.keep_ptr[0] = 0;
  f.keep_idx.resize(kz0[f.n_sub]);
  f.keep_val.resize(kz0[f.n_sub]);
  f.keep_rowsum.resize(f.n_red);
#pragma omp parallel
  for (i64 si = 0; si < f.n_sub; ++si) {
    const Subdomain& s = subs[si];
    const u32 L0 = f.loc_ptr[si], R0 = f.red_ptr[si];
    const i64 nq = s.ct.n_q(), nl = s.n_local();
    const Csr& e2k = s.eliminate_q ? s.ct.c : s.ct.ct;   // elim rows -> keep cols
    const Csr& k2e = s.eliminate_q ? s.ct.ct : s.ct.c;   // keep rows -> elim cols
    const u32 elim_base = s.eliminate_q ? L0 : L0 + (u32)nq;
    const u32 keep_base = s.eliminate_q ? L0 + (u32)nq : L0;
#pragma omp for schedule(static)
    for (i64 k = 0; k < nl; ++k) {
      f.loc_gidx[L0 + k] = s.gidx[k];
      double w = s.pou.empty() ? 1.0 : s.pou[k];
      f.loc_scale[L0 + k] = k >= nq ? -w : w;
      f.loc_sub[L0 + k] = (u32)si;
    }
#pragma omp for schedule(dynamic, 1024)
    for (i64 e = 0; e < e2k.nrows; ++e) {
      const i64 row = er0[si] + e;
      f.elim_loc[row] = elim_base + (u32)e;
      f.elim_inv[row] = s.inv_diag_elim[e];
      i64 o = ez0[si] + e2k.indptr[e];
      for (u32 q = e2k.indptr[e]; q < e2k.indptr[e + 1]; ++q, ++o) {
        f.elim_idx[o] = R0 + e2k.indices[q];
        f.elim_val[o] = e2k.data[q];
      }
      f.elim_ptr[row + 1] = (u32)(ez0[si] + e2k.indptr[e + 1]);
    }
#pragma omp for schedule(dynamic, 1024)
    for (i64 j = 0; j < k2e.nrows; ++j) {
      f.red_sub[R0 + j] = (u32)si;
      f.red_loc[R0 + j] = keep_base + (u32)j;
      double rs = 0.0;
      i64 o = kz0[si] + k2e.indptr[j];
      for (u32 q = k2e.indptr[j]; q < k2e.indptr[j + 1]; ++q, ++o) {
        double v = k2e.data[q] * s.inv_diag_elim[k2e.indices[q]];
        f.keep_idx[o] = elim_base + k2e.indices[q];
        f.keep_val[o] = v;
        rs += v;
      }
      f.keep_ptr[R0 + j + 1] = (u32)(kz0[si] + k2e.indptr[j + 1]);
      f.keep_rowsum[R0 + j] = rs;
    }
  }
  // factor levels: round t of every subdomain, then the core gather level.
  // In-edges of each reduced row as CSR, in (subdomain, elimination order,
  // column) order; then the level rows in order, their forward (in-edge) and
  // backward (out-edge) lists sized, offset and filled in parallel.
  std::vector<u32> in_ptr(f.n_red + 1, 0);
  for (i64 si = 0; si < f.n_sub; ++si) {
    const Factor& F = subs[si].f;
    const u32 R0 = f.red_ptr[si];
    for (u32 k = 0; k < F.round_ptr.back(); ++k) {
      const u32 a = F.order[k];
      if (!(F.pivot[a] > 0.0)) continue;
      for (u32 p = F.col_ptr[a]; p < F.col_ptr[a + 1]; ++p) in_ptr[R0 + F.col_idx[p] + 1]++;
    }
  }
  for (i64 r = 0; r < f.n_red; ++r) in_ptr[r + 1] += in_ptr[r];
  std::vector<u32> in_src(in_ptr[f.n_red]);
  std::vector<double> in_w(in_ptr[f.n_red]);
  {
    std::vector<u32> cur(in_ptr.begin(), in_ptr.end() - 1);
    for (i64 si = 0; si < f.n_sub; ++si) {
      const Factor& F = subs[si].f;
      const u32 R0 = f.red_ptr[si];
      for (u32 k = 0; k < F.round_ptr.back(); ++k) {
        const u32 a = F.order[k];
        if (!(F.pivot[a] > 0.0)) continue;
        for (u32 p = F.col_ptr[a]; p < F.col_ptr[a + 1]; ++p) {
          const u32 at = cur[R0 + F.col_idx[p]]++;
          in_src[at] = R0 + a;
          in_w[at] = F.col_w[p] / F.pivot[a];
        }
        f.inv_pivot[R0 + a] = 1.0 / F.pivot[a];
      }
      f.ac_nnz += (i64)F.col_idx.size();
    }
  }
  // the level rows: (level, subdomain, elimination order); per row its
  // subdomain and whether it pulls its later neighbours on the way back
  // (eliminated rows do; core rows get their values from the dense solve)
  std::vector<u32> row_sub;
  std::vector<char> row_out;
  f.lvl_ptr.assign(1, 0);
  for (i64 t = 0; t <= max_rounds; ++t) {
    for (i64 si = 0; si < f.n_sub; ++si) {
      const Factor& F = subs[si].f;
      const u32 R0 = f.red_ptr[si];
      auto add = [&](u32 v) {
        f.lvl_rows.push_back(R0 + v);
        row_sub.push_back((u32)si);
        row_out.push_back(t < max_rounds && F.pivot[v] > 0.0 ? 1 : 0);
      };
      if (t < (i64)F.round_ptr.size() - 1) {
        for (u32 k = F.round_ptr[t]; k < F.round_ptr[t + 1]; ++k) add(F.order[k]);
      } else if (t == max_rounds) {
        for (u32 v : F.core) add(v);  // last level: core vertices gather their in-edges
      }
    }
    f.lvl_ptr.push_back((u32)f.lvl_rows.size());
  }
  const i64 nrows = (i64)f.lvl_rows.size();
  f.fin_ptr.assign(nrows + 1, 0);
  f.bout_ptr.assign(nrows + 1, 0);
  for (i64 k = 0; k < nrows; ++k) {
    const u32 g = f.lvl_rows[k];
    const Factor& F = subs[row_sub[k]].f;
    const u32 v = g - f.red_ptr[row_sub[k]];
    f.fin_ptr[k + 1] = f.fin_ptr[k] + (in_ptr[g + 1] - in_ptr[g]);
    f.bout_ptr[k + 1] = f.bout_ptr[k] + (row_out[k] ? F.col_ptr[v + 1] - F.col_ptr[v] : 0);
  }
  f.fin_idx.resize(f.fin_ptr[nrows]);
  f.fin_val.resize(f.fin_ptr[nrows]);
  f.bout_idx.resize(f.bout_ptr[nrows]);
  f.bout_val.resize(f.bout_ptr[nrows]);
#pragma omp parallel for schedule(dynamic, 1024)
  for (i64 k = 0; k < nrows; ++k) {
    const u32 g = f.lvl_rows[k];
    const u32 si = row_sub[k], R0 = f.red_ptr[si], v = g - R0;
    const Factor& F = subs[si].f;
    u32 o = f.fin_ptr[k];
    for (u32 e = in_ptr[g]; e < in_ptr[g + 1]; ++e, ++o) {
      f.fin_idx[o] = in_src[e];
      f.fin_val[o] = in_w[e];
    }
    if (!row_out[k]) continue;
    o = f.bout_ptr[k];
    for (u32 p = F.col_ptr[v]; p < F.col_ptr[v + 1]; ++p, ++o) {
      f.bout_idx[o] = R0 + F.col_idx[p];
      f.bout_val[o] = F.col_w[p] / F.pivot[v];
    }
  }
  // dense cores
  f.core_ptr.assign(1, 0);
  for (i64 si = 0; si < f.n_sub; ++si) {
    const Factor& F = subs[si].f;
    if (F.core.empty()) continue;
    u32 R0 = f.red_ptr[si];
    f.core_mat_off.push_back(f.core_mat.size());
    for (u32 v : F.core) {
      f.core_idx.push_back(R0 + v);
      f.core_row_core.push_back((u32)(f.core_ptr.size() - 1));
    }
    f.core_ptr.push_back((u32)f.core_idx.size());
    f.core_mat.insert(f.core_mat.end(), F.core_inv.begin(), F.core_inv.end());
    f.core_dev.push_back(F.core_dev);
    f.max_core = std::max<i64>(f.max_core, (i64)F.core.size());
  }
  // DOF -> local entries
  f.dof_ptr.assign(p.d.n_dofs + 1, 0);
  for (u32 g : f.loc_gidx) f.dof_ptr[g + 1]++;
  for (i64 j = 0; j < p.d.n_dofs; ++j) f.dof_ptr[j + 1] += f.dof_ptr[j];
  f.dof_loc.resize(f.n_loc);
  {
    std::vector<u32> cur(f.dof_ptr.begin(), f.dof_ptr.end() - 1);
    for (i64 l = 0; l < f.n_loc; ++l) f.dof_loc[cur[f.loc_gidx[l]]++] = (u32)l;
  }
}

DeviceLayout make_device_layout(const Plan& p) {
  if (p.d.Q == 0) {
    // a preconditioner on its own (no design): the user numbering as is
    DeviceLayout L;
    L.m = p.d.n_dofs;
    L.int_of.resize(L.m);
    std::iota(L.int_of.begin(), L.int_of.end(), 0u);
    L.ext_of = L.int_of;
    L.gram = p.gram;
    L.pre = p.pre;
    return L;
  }
  DeviceLayout L = layout_operator(p);
  layout_precond(p, L);
  return L;
}

DeviceLayout layout_operator(const Plan& p) {
  const Design& d = p.d;
  const int Q = d.Q;
  const i64 n = d.n_obs;
  DeviceLayout L;
  PhaseTimer t_levels("layout: levels");
  // active levels: the non-empty rows of each factor's pass
  // factor with the most levels, ordered by the factor with the next-most
  int big = 0, ref = -1;
  for (int q = 1; q < Q; ++q)
    if (d.levels[q] > d.levels[big]) big = q;
  for (int q = 0; q < Q; ++q)
    if (q != big && (ref < 0 || d.levels[q] > d.levels[ref])) ref = q;
  L.int_of.assign(d.n_dofs, UINT32_MAX);
  for (int q = 0; q < Q; ++q) {
    const std::vector<u32>& rp = p.passes[q].row_ptr;
    std::vector<u32> lv;
    for (i64 l = 0; l < d.levels[q]; ++l)
      if (rp[l + 1] > rp[l]) lv.push_back((u32)l);
    if (q == big && ref >= 0) {
      // the smallest `ref` code a level meets is its first element's: a
      // pass's rows are sorted by the code of the factor with the most
      // levels among the others, which is `ref` for `big`
      const FactorPass& fp = p.passes[q];
      const int ts = ref < q ? ref : ref - 1;
      std::vector<u32> key(d.levels[q], UINT32_MAX);
#pragma omp parallel for schedule(static)
      for (i64 l = 0; l < d.levels[q]; ++l)
        if (rp[l + 1] > rp[l]) key[l] = fp.other[(size_t)rp[l] * (Q - 1) + ts];
      std::stable_sort(lv.begin(), lv.end(), [&](u32 a, u32 b) { return key[a] < key[b]; });
    }
    L.off.push_back((i64)L.ext_of.size());
    L.levels.push_back((i64)lv.size());
    for (u32 l : lv) {
      L.int_of[d.offset[q] + l] = (u32)L.ext_of.size();
      L.ext_of.push_back((u32)(d.offset[q] + l));
    }
  }
  L.m = (i64)L.ext_of.size();
  // observations -> internal DOFs
  L.dof_of_obs.resize((size_t)Q * n);
  for (int q = 0; q < Q; ++q) {
    const u32* c = d.codes[q].data();
    const i64 o = d.offset[q];
    u32* out = L.dof_of_obs.data() + (size_t)q * n;
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < n; ++i) out[i] = L.int_of[o + c[i]];
  }
  t_levels.done();
  PhaseTimer t_passes("layout: passes");
  // factor passes: active rows in internal order; elements re-sorted by the
  // internal index of the largest other factor (gather locality)
  const int Qm = Q - 1;
  L.passes.resize(Q);
  for (int q = 0; q < Q; ++q) {
    const FactorPass& fp = p.passes[q];
    FactorPass& np = L.passes[q];
    np.q = q;
    int sec = -1;
    for (int r = 0; r < Q; ++r)
      if (r != q && (sec < 0 || d.levels[r] > d.levels[sec])) sec = r;
    const int ts = sec < 0 ? 0 : (sec < q ? sec : sec - 1);
    const i64 nr = L.levels[q];
    np.row_ptr.assign(nr + 1, 0);
    np.diag.resize(nr);
    for (i64 r = 0; r < nr; ++r) {
      const u32 l = L.ext_of[L.off[q] + r] - (u32)d.offset[q];
      np.row_ptr[r + 1] = np.row_ptr[r] + (fp.row_ptr[l + 1] - fp.row_ptr[l]);
      np.diag[r] = fp.diag[l];
    }
    np.obs.resize(fp.obs.size());
    np.other.resize(fp.other.size());
    if (!fp.w.empty()) np.w.resize(fp.w.size());
    // rows in parallel, each a stable sort by (internal index of `sec`, element)
    const int chunk = (int)std::max<i64>(1, std::min<i64>(256, nr / 512));
#pragma omp parallel
    {
      std::vector<std::pair<u32, u32>> idx;
#pragma omp for schedule(dynamic, chunk)
      for (i64 r = 0; r < nr; ++r) {
        const u32 l = L.ext_of[L.off[q] + r] - (u32)d.offset[q];
        idx.clear();
        for (u32 e = fp.row_ptr[l]; e < fp.row_ptr[l + 1]; ++e)
          idx.push_back({Qm > 0 ? L.int_of[fp.other[(size_t)e * Qm + ts]] : 0u, e});
        if (Qm > 0) std::sort(idx.begin(), idx.end());
        size_t o = np.row_ptr[r];
        for (const auto& ke : idx) {
          const u32 e = ke.second;
          np.obs[o] = fp.obs[e];
          for (int t = 0; t < Qm; ++t) np.other[o * Qm + t] = L.int_of[fp.other[(size_t)e * Qm + t]];
          if (!fp.w.empty()) np.w[o] = fp.w[e];
          ++o;
        }
      }
    }
  }
  t_passes.done();
  PhaseTimer t_gram("layout: gramian");
  // explicit Gramian: rows and columns renumbered, columns ascending
  if (p.has_gram) {
    const Csr& g = p.gram;
    Csr& ng = L.gram;
    ng.nrows = ng.ncols = L.m;
    ng.indptr.assign(L.m + 1, 0);
    for (i64 r = 0; r < L.m; ++r) {
      const u32 u = L.ext_of[r];
      ng.indptr[r + 1] = ng.indptr[r] + (g.indptr[u + 1] - g.indptr[u]);
    }
    ng.indices.resize(g.indices.size());
    ng.data.resize(g.data.size());
#pragma omp parallel
    {
      std::vector<std::pair<u32, double>> row;
#pragma omp for schedule(dynamic, 1024)
      for (i64 r = 0; r < L.m; ++r) {
        const u32 u = L.ext_of[r];
        row.clear();
        for (u32 k = g.indptr[u]; k < g.indptr[u + 1]; ++k) row.push_back({L.int_of[g.indices[k]], g.data[k]});
        std::sort(row.begin(), row.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
        size_t o = ng.indptr[r];
        for (const auto& e : row) {
          ng.indices[o] = e.first;
          ng.data[o++] = e.second;
        }
      }
    }
  }
  t_gram.done();
  return L;
}

// a deep copy of a flattened preconditioner, its arrays copied side by side
static void copy_flat(const PrecondFlat& f, PrecondFlat& c) {
  static_assert(sizeof(PrecondFlat) == 7 * sizeof(i64) + 33 * sizeof(std::vector<u32>),
                "copy_flat copies every member of PrecondFlat");
  c.n_loc = f.n_loc, c.n_red = f.n_red, c.n_sub = f.n_sub;
  c.max_core = f.max_core, c.ac_nnz = f.ac_nnz, c.c_nnz = f.c_nnz, c.dense_subs = f.dense_subs;
#pragma omp parallel
#pragma omp single
  {
#pragma omp task
    c.loc_gidx = f.loc_gidx;
#pragma omp task
    c.loc_scale = f.loc_scale;
#pragma omp task
    c.loc_ptr = f.loc_ptr;
#pragma omp task
    c.red_ptr = f.red_ptr;
#pragma omp task
    c.loc_sub = f.loc_sub;
#pragma omp task
    c.red_sub = f.red_sub;
#pragma omp task
    c.red_loc = f.red_loc;
#pragma omp task
    c.elim_loc = f.elim_loc;
#pragma omp task
    c.elim_inv = f.elim_inv;
#pragma omp task
    c.elim_ptr = f.elim_ptr;
#pragma omp task
    c.elim_idx = f.elim_idx;
#pragma omp task
    c.elim_val = f.elim_val;
#pragma omp task
    c.keep_ptr = f.keep_ptr;
#pragma omp task
    c.keep_idx = f.keep_idx;
#pragma omp task
    c.keep_val = f.keep_val;
#pragma omp task
    c.keep_rowsum = f.keep_rowsum;
#pragma omp task
    c.lvl_ptr = f.lvl_ptr;
#pragma omp task
    c.lvl_rows = f.lvl_rows;
#pragma omp task
    c.fin_ptr = f.fin_ptr;
#pragma omp task
    c.fin_idx = f.fin_idx;
#pragma omp task
    c.fin_val = f.fin_val;
#pragma omp task
    c.bout_ptr = f.bout_ptr;
#pragma omp task
    c.bout_idx = f.bout_idx;
#pragma omp task
    c.bout_val = f.bout_val;
#pragma omp task
    c.inv_pivot = f.inv_pivot;
#pragma omp task
    c.core_ptr = f.core_ptr;
#pragma omp task
    c.core_idx = f.core_idx;
#pragma omp task
    c.core_mat_off = f.core_mat_off;
#pragma omp task
    c.core_mat = f.core_mat;
#pragma omp task
    c.core_row_core = f.core_row_core;
#pragma omp task
    c.core_dev = f.core_dev;
#pragma omp task
    c.dof_ptr = f.dof_ptr;
#pragma omp task
    c.dof_loc = f.dof_loc;
  }
}

void layout_precond(const Plan& p, DeviceLayout& L) {
  PhaseTimer t_pre("layout: preconditioner");
  // preconditioner: DOF indices renumbered; each subdomain's eliminated block
  // (its elimination rows) re-laid out in internal DOF order
  if (!p.meta.empty()) {
    const PrecondFlat& f = p.pre;
    // (bytes built for another design with the same DOF count may name
    // levels this design never observes)
    for (u32 g : f.loc_gidx)
      if (L.int_of[g] == UINT32_MAX)
        throw Error(E_INVALID, "the preconditioner does not match the design (it covers an unobserved level)");
    PrecondFlat nf;
    copy_flat(f, nf);
    const i64 ne = (i64)f.elim_loc.size();
    std::vector<u32> lp(f.n_loc);
    std::iota(lp.begin(), lp.end(), 0u);
    std::vector<u32> order(ne);
    std::iota(order.begin(), order.end(), 0u);
    // runs of elimination rows of one subdomain, each ordered by internal DOF
    std::vector<i64> run(1, 0);
    for (i64 t = 1; t < ne; ++t)
      if (f.loc_sub[f.elim_loc[t]] != f.loc_sub[f.elim_loc[t - 1]]) run.push_back(t);
    if (ne > 0) run.push_back(ne);
    const i64 nruns = (i64)run.size() - 1;
#pragma omp parallel
    {
      std::vector<std::pair<u32, u32>> kv;
#pragma omp for schedule(dynamic, 1)
      for (i64 q = 0; q < nruns; ++q) {
        const i64 a = run[q], b = run[q + 1];
        u32 base = f.elim_loc[a];
        kv.clear();
        for (i64 t = a; t < b; ++t) {
          base = std::min(base, f.elim_loc[t]);
          kv.push_back({L.int_of[f.loc_gidx[f.elim_loc[t]]], (u32)t});
        }
        std::sort(kv.begin(), kv.end());  // = a stable sort by internal DOF
        for (i64 t = a; t < b; ++t) {
          order[t] = kv[t - a].second;
          lp[f.elim_loc[order[t]]] = base + (u32)(t - a);
        }
      }
    }
#pragma omp parallel for schedule(static)
    for (i64 l = 0; l < f.n_loc; ++l) {
      nf.loc_gidx[lp[l]] = L.int_of[f.loc_gidx[l]];
      nf.loc_scale[lp[l]] = f.loc_scale[l];
      nf.loc_sub[lp[l]] = f.loc_sub[l];
    }
    nf.elim_ptr.assign(ne + 1, 0);
    for (i64 t = 0; t < ne; ++t) {
      const u32 i = order[t];
      nf.elim_ptr[t + 1] = nf.elim_ptr[t] + (f.elim_ptr[i + 1] - f.elim_ptr[i]);
    }
#pragma omp parallel for schedule(dynamic, 4096)
    for (i64 t = 0; t < ne; ++t) {
      const u32 i = order[t];
      nf.elim_loc[t] = lp[f.elim_loc[i]];
      nf.elim_inv[t] = f.elim_inv[i];
      u32 o = nf.elim_ptr[t];
      for (u32 k = f.elim_ptr[i]; k < f.elim_ptr[i + 1]; ++k, ++o) {
        nf.elim_idx[o] = f.elim_idx[k];
        nf.elim_val[o] = f.elim_val[k];
      }
    }
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < f.keep_idx.size(); ++k) nf.keep_idx[k] = lp[f.keep_idx[k]];
#pragma omp parallel for schedule(static)
    for (i64 j = 0; j < f.n_red; ++j) nf.red_loc[j] = lp[f.red_loc[j]];
    nf.dof_ptr.assign(L.m + 1, 0);
    for (u32 g : nf.loc_gidx) nf.dof_ptr[g + 1]++;
    for (i64 j = 0; j < L.m; ++j) nf.dof_ptr[j + 1] += nf.dof_ptr[j];
    nf.dof_loc.resize(nf.n_loc);
    std::vector<u32> cur(nf.dof_ptr.begin(), nf.dof_ptr.end() - 1);
    for (i64 l = 0; l < nf.n_loc; ++l) nf.dof_loc[cur[nf.loc_gidx[l]]++] = (u32)l;
    L.pre = std::move(nf);
  }
}

void build_plan(Plan& p) {
  const Design& d = p.d;
  // the subdomains (cross-tabulations and components) read only d: built on a
  // thread of their own while the operator passes and the Gramian are built
  const bool want_subs = (p.precond == 1 || p.precond == 2) && !p.pre_given;
  struct SubsJoiner {
    std::thread t;
    ~SubsJoiner() {
      if (t.joinable()) t.join();
    }
  } subs_thread;
  std::string subs_err;
  int subs_code = 0;
  if (want_subs) {
    subs_thread.t = std::thread([&p, &subs_err, &subs_code] {
      try {
        PhaseTimer t("subdomains");
        p.subs = build_subdomains(p.d);
      } catch (const Error& e) {
        subs_err = e.what();
        subs_code = e.code;
      } catch (const std::exception& e) {
        subs_err = e.what();
        subs_code = E_INVALID;
      }
    });
  }
  p.passes.resize(d.Q);
  {
    PhaseTimer t("operator passes");
    // the passes side by side, each on a host thread with its own team
    std::vector<std::thread> th;
    std::vector<std::string> perr(d.Q);
    for (int q = 1; q < d.Q; ++q)
      th.emplace_back([&p, &d, &perr, q] {
        try {
          p.passes[q] = build_pass(d, q);
        } catch (const std::exception& e) {
          perr[q] = e.what();
        }
      });
    try {
      if (d.Q > 0) p.passes[0] = build_pass(d, 0);
    } catch (const std::exception& e) {
      perr[0] = e.what();
    }
    for (std::thread& x : th) x.join();
    for (int q = 0; q < d.Q; ++q)
      if (!perr[q].empty()) throw Error(E_INVALID, perr[q]);
  }
  const bool want_gram = p.explicit_op || p.precond == 2;
  auto build_gram = [&p] {
    PhaseTimer t("explicit gramian");
    p.gram = build_gramian(p.d, p.passes);
    p.has_gram = true;
  };
  // the Gramian and the operator's device layout read only d, passes and gram:
  // when the caller lays the plan out early they run on a thread of their own
  // while the preconditioner is built below (which writes subs, pre and
  // meta); joined on every way out of this function
  if (want_gram && !(p.layout_early && d.Q > 0)) build_gram();
  struct Joiner {
    std::thread t;
    ~Joiner() {
      if (t.joinable()) t.join();
    }
  } lay;
  std::string lay_err;
  int lay_code = 0;
  if (p.layout_early && d.Q > 0) {
    p.lay_op.reset();
    lay.t = std::thread([&p, &lay_err, &lay_code, want_gram, &build_gram] {
      try {
        if (want_gram) build_gram();
        p.lay_op = std::make_shared<DeviceLayout>(layout_operator(p));
      } catch (const Error& e) {
        lay_err = e.what();
        lay_code = e.code;
      } catch (const std::exception& e) {
        lay_err = e.what();
        lay_code = E_INVALID;
      }
    });
  }
  if (want_subs) {
    subs_thread.t.join();
    if (subs_code) throw Error(subs_code, subs_err);
    PhaseTimer t("local solvers");
    std::vector<std::string> err(p.subs.size());
    std::vector<int> code(p.subs.size(), 0);
    // huge subdomains first, each on a host thread of its own with its own
    // thread team (their Schur sampling, rounds and dense cores are parallel
    // inside; while one subdomain's dense core is factored on the GPU the
    // others keep the host busy), then the large ones side by side (one
    // thread each: a stack of many problems has hundreds of them), then the
    // rest.  Each build's arithmetic does not depend on its thread count.
    std::vector<size_t> serial, large, rest;
    for (size_t i = 0; i < p.subs.size(); ++i) {
      const Subdomain& sd = p.subs[i];
      const i64 keep = std::min(sd.ct.n_q(), sd.ct.n_r());  // the smaller block is kept
      if (sd.n_local() >= ((i64)1 << 20) || keep >= 16384) serial.push_back(i);
      else if (keep >= 1024 || sd.n_local() >= 65536) large.push_back(i);
      else rest.push_back(i);
    }
    // a few large subdomains keep their inner parallelism
    if (large.size() < 4) {
      serial.insert(serial.end(), large.begin(), large.end());
      large.clear();
    }
    auto build_one = [&](size_t i) {
      try {
        auto t0 = std::chrono::steady_clock::now();
        build_local_solver(p.subs[i], p.cfg);
        double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (verbose() && dt > 0.05)
          std::fprintf(stderr, "[within_amd]   subdomain %zu (n_q=%lld n_r=%lld): %.3f s, rounds %zu core %zu\n", i,
                       (long long)p.subs[i].ct.n_q(), (long long)p.subs[i].ct.n_r(), dt,
                       p.subs[i].f.round_ptr.size() - 1, p.subs[i].f.core.size());
      } catch (const Error& e) {
        err[i] = e.what();
        code[i] = e.code;
      }
    };
    {
      PhaseTimer ts("local solvers, large subdomains");
      std::vector<std::thread> th;
      for (size_t t = 1; t < serial.size(); ++t) th.emplace_back(build_one, serial[t]);
      if (!serial.empty()) build_one(serial[0]);
      for (std::thread& x : th) x.join();
    }
    {
      PhaseTimer tl("local solvers, large subdomains side by side");
      // largest first (dynamic schedule)
      std::sort(large.begin(), large.end(),
                [&](size_t a, size_t b) { return p.subs[a].n_local() > p.subs[b].n_local(); });
#pragma omp parallel for schedule(dynamic, 1)
      for (size_t t = 0; t < large.size(); ++t) build_one(large[t]);
    }
    {
      PhaseTimer tr("local solvers, the rest");
#pragma omp parallel for schedule(dynamic, 1)
      for (size_t t = 0; t < rest.size(); ++t) build_one(rest[t]);
    }
    for (size_t i = 0; i < p.subs.size(); ++i)
      if (code[i]) throw Error(code[i], err[i]);
    PhaseTimer tf("flatten");
    flatten(p);
    p.meta.clear();
    for (const Subdomain& sd : p.subs)
      p.meta.push_back({sd.q, sd.r, sd.eliminate_q ? sd.ct.n_q() : sd.ct.n_r(), (i64)sd.f.core.size(), sd.ct.n_q(),
                        sd.ct.n_r(), (i64)sd.ct.c.indices.size()});
  }
  if (lay.t.joinable()) lay.t.join();
  if (lay_code) throw Error(lay_code, lay_err);
}

// ---------------------------------------------------------------------------
// preconditioner bytes
// ---------------------------------------------------------------------------
namespace {
constexpr uint32_t kMagic = 0x43504157u;  // "WAPC"
constexpr uint32_t kVersion = 1;

struct Writer {
  std::vector<uint8_t> b;
  template <typename T>
  void put(const T& v) {
    const uint8_t* p = reinterpret_cast<const uint8_t*>(&v);
    b.insert(b.end(), p, p + sizeof(T));
  }
  template <typename T>
  void vec(const std::vector<T>& v) {
    put<uint64_t>(v.size());
    const uint8_t* p = reinterpret_cast<const uint8_t*>(v.data());
    b.insert(b.end(), p, p + sizeof(T) * v.size());
  }
};

struct Reader {
  const uint8_t* p;
  i64 left;
  void need(i64 n) {
    if (n < 0 || n > left) throw Error(E_INVALID, "preconditioner bytes are truncated or malformed");
  }
  template <typename T>
  T get() {
    need((i64)sizeof(T));
    T v;
    std::memcpy(&v, p, sizeof(T));
    p += sizeof(T);
    left -= (i64)sizeof(T);
    return v;
  }
  template <typename T>
  void vec(std::vector<T>& v) {
    const uint64_t n = get<uint64_t>();
    if (n > (uint64_t)left / sizeof(T)) throw Error(E_INVALID, "preconditioner bytes are truncated or malformed");
    v.resize(n);
    if (n) std::memcpy(v.data(), p, sizeof(T) * n);  // an empty vector's data() may be null
    p += sizeof(T) * n;
    left -= (i64)(sizeof(T) * n);
  }
};

template <class IO>
void flat_fields(IO& io, PrecondFlat& f) {
  io.vec(f.loc_gidx); io.vec(f.loc_scale); io.vec(f.loc_ptr); io.vec(f.red_ptr); io.vec(f.loc_sub);
  io.vec(f.red_sub); io.vec(f.red_loc); io.vec(f.elim_loc); io.vec(f.elim_inv); io.vec(f.elim_ptr);
  io.vec(f.elim_idx); io.vec(f.elim_val); io.vec(f.keep_ptr); io.vec(f.keep_idx); io.vec(f.keep_val);
  io.vec(f.keep_rowsum); io.vec(f.lvl_ptr); io.vec(f.lvl_rows); io.vec(f.fin_ptr); io.vec(f.fin_idx);
  io.vec(f.fin_val); io.vec(f.bout_ptr); io.vec(f.bout_idx); io.vec(f.bout_val); io.vec(f.inv_pivot);
  io.vec(f.core_ptr); io.vec(f.core_idx); io.vec(f.core_mat_off); io.vec(f.core_mat); io.vec(f.core_row_core);
  io.vec(f.dof_ptr); io.vec(f.dof_loc);
}
}  // namespace

void materialize_cores(PrecondFlat& f) {
  bool any = false;
  for (const auto& d : f.core_dev) any = any || d != nullptr;
  if (!any) return;
  const size_t n_cores = f.core_mat_off.size();
  std::vector<double> mat;
  std::vector<uint64_t> off(n_cores);
  for (size_t c = 0; c < n_cores; ++c) {
    off[c] = mat.size();
    const DeviceDense* d = device_core(f, c);
    if (d) {
      const std::vector<double> h = d->download();
      mat.insert(mat.end(), h.begin(), h.end());
    } else {
      const uint64_t nc = f.core_ptr[c + 1] - f.core_ptr[c];
      mat.insert(mat.end(), f.core_mat.begin() + (ptrdiff_t)f.core_mat_off[c],
                 f.core_mat.begin() + (ptrdiff_t)(f.core_mat_off[c] + nc * nc));
    }
  }
  f.core_mat = std::move(mat);
  f.core_mat_off = std::move(off);
  f.core_dev.clear();
}

std::vector<uint8_t> serialize_precond(const Plan& p) {
  if (p.precond != 1 && p.precond != 2) throw Error(E_INVALID, "solver has no preconditioner");
  Writer w;
  w.put(kMagic);
  w.put(kVersion);
  w.put<int32_t>(p.precond);
  w.put<i64>(p.d.n_dofs);
  w.put<uint64_t>(p.cfg.ac_seed); w.put<int32_t>(p.cfg.ac_split); w.put<int32_t>(p.cfg.approx_schur ? 1 : 0);
  w.put<uint64_t>(p.cfg.schur_seed); w.put<int32_t>(p.cfg.schur_split); w.put<int32_t>(p.cfg.dense_threshold);
  w.put<i64>(p.cfg.dense_core);
  w.vec(p.meta);
  PrecondFlat host;  // a copy with the device cores brought to the host, when there are any
  bool dev = false;
  for (const auto& d : p.pre.core_dev) dev = dev || d != nullptr;
  if (dev) {
    host = p.pre;
    materialize_cores(host);
  }
  const PrecondFlat& f = dev ? host : p.pre;
  w.put<i64>(f.n_loc); w.put<i64>(f.n_red); w.put<i64>(f.n_sub);
  w.put<i64>(f.max_core); w.put<i64>(f.ac_nnz); w.put<i64>(f.c_nnz); w.put<i64>(f.dense_subs);
  flat_fields(w, const_cast<PrecondFlat&>(f));
  if (p.precond == 2) {
    w.put<i64>(p.gram.nrows);
    w.vec(p.gram.indptr); w.vec(p.gram.indices); w.vec(p.gram.data);
  }
  return std::move(w.b);
}

i64 deserialize_precond(const uint8_t* buf, i64 len, Plan& p) {
  if (!buf || len <= 0) throw Error(E_INVALID, "empty preconditioner bytes");
  Reader r{buf, len};
  if (r.get<uint32_t>() != kMagic) throw Error(E_INVALID, "not a within_amd preconditioner");
  if (r.get<uint32_t>() != kVersion) throw Error(E_INVALID, "unsupported preconditioner format version");
  p.precond = r.get<int32_t>();
  if (p.precond != 1 && p.precond != 2) throw Error(E_INVALID, "bad preconditioner kind");
  const i64 n_dofs = r.get<i64>();
  p.cfg.ac_seed = r.get<uint64_t>(); p.cfg.ac_split = r.get<int32_t>(); p.cfg.approx_schur = r.get<int32_t>() != 0;
  p.cfg.schur_seed = r.get<uint64_t>(); p.cfg.schur_split = r.get<int32_t>(); p.cfg.dense_threshold = r.get<int32_t>();
  p.cfg.dense_core = r.get<i64>();
  r.vec(p.meta);
  PrecondFlat& f = p.pre;
  f.n_loc = r.get<i64>(); f.n_red = r.get<i64>(); f.n_sub = r.get<i64>();
  f.max_core = r.get<i64>(); f.ac_nnz = r.get<i64>(); f.c_nnz = r.get<i64>(); f.dense_subs = r.get<i64>();
  flat_fields(r, f);
  // consistency of the sizes every kernel relies on
  bool ok = n_dofs > 0 && (i64)f.loc_ptr.size() == f.n_sub + 1 && (i64)f.red_ptr.size() == f.n_sub + 1 &&
            (i64)p.meta.size() == f.n_sub && (i64)f.loc_gidx.size() == f.n_loc && (i64)f.loc_scale.size() == f.n_loc &&
            (i64)f.loc_sub.size() == f.n_loc && (i64)f.red_sub.size() == f.n_red && (i64)f.red_loc.size() == f.n_red &&
            (i64)f.dof_ptr.size() == n_dofs + 1 && (i64)f.dof_loc.size() == f.n_loc &&
            f.loc_ptr.back() == (u32)f.n_loc && f.red_ptr.back() == (u32)f.n_red;
  auto below = [&](const std::vector<u32>& v, i64 n) {
    for (u32 x : v) ok = ok && (i64)x < n;
  };
  auto csr = [&](const std::vector<u32>& ptr, i64 rows, size_t nnz) {
    ok = ok && (i64)ptr.size() == rows + 1 && ptr[0] == 0 && ptr.back() == nnz;
    for (size_t i = 0; ok && i + 1 < ptr.size(); ++i) ok = ptr[i] <= ptr[i + 1];
  };
  if (ok) {
    const i64 n_elim = (i64)f.elim_loc.size(), n_lv = (i64)f.lvl_rows.size();
    const i64 n_cores = f.core_ptr.empty() ? 0 : (i64)f.core_ptr.size() - 1;
    below(f.loc_gidx, n_dofs); below(f.loc_sub, f.n_sub); below(f.red_sub, f.n_sub); below(f.red_loc, f.n_loc);
    below(f.elim_loc, f.n_loc); below(f.elim_idx, f.n_red); below(f.keep_idx, f.n_loc); below(f.lvl_rows, f.n_red);
    below(f.fin_idx, f.n_red); below(f.bout_idx, f.n_red); below(f.core_idx, f.n_red); below(f.dof_loc, f.n_loc);
    below(f.core_row_core, n_cores);
    csr(f.loc_ptr, f.n_sub, (size_t)f.n_loc); csr(f.red_ptr, f.n_sub, (size_t)f.n_red);
    csr(f.elim_ptr, n_elim, f.elim_idx.size()); csr(f.keep_ptr, f.n_red, f.keep_idx.size());
    csr(f.fin_ptr, n_lv, f.fin_idx.size()); csr(f.bout_ptr, n_lv, f.bout_idx.size());
    csr(f.dof_ptr, n_dofs, (size_t)f.n_loc);
    ok = ok && f.elim_inv.size() == f.elim_loc.size() && f.elim_val.size() == f.elim_idx.size() &&
         f.keep_val.size() == f.keep_idx.size() && (i64)f.keep_rowsum.size() == f.n_red &&
         f.fin_val.size() == f.fin_idx.size() && f.bout_val.size() == f.bout_idx.size() &&
         (i64)f.inv_pivot.size() == f.n_red && (i64)f.core_mat_off.size() == n_cores &&
         f.core_row_core.size() == f.core_idx.size() && !f.lvl_ptr.empty() && f.lvl_ptr.back() == (u32)n_lv;
    for (size_t i = 0; ok && i + 1 < f.lvl_ptr.size(); ++i) ok = f.lvl_ptr[i] <= f.lvl_ptr[i + 1];
    if (n_cores > 0) csr(f.core_ptr, n_cores, f.core_idx.size());
    i64 max_core = 0;
    for (i64 c = 0; ok && c < n_cores; ++c) {
      const uint64_t nc = f.core_ptr[c + 1] - f.core_ptr[c];
      ok = f.core_mat_off[c] <= f.core_mat.size() && nc * nc <= f.core_mat.size() - f.core_mat_off[c];
      for (u32 i = f.core_ptr[c]; ok && i < f.core_ptr[c + 1]; ++i) ok = f.core_row_core[i] == (u32)c;
      max_core = std::max<i64>(max_core, (i64)nc);
    }
    ok = ok && max_core == f.max_core;
    // local / reduced entries belong to the subdomain whose range holds them
    for (i64 s = 0; ok && s < f.n_sub; ++s) {
      for (u32 l = f.loc_ptr[s]; ok && l < f.loc_ptr[s + 1]; ++l) ok = f.loc_sub[l] == (u32)s;
      for (u32 j = f.red_ptr[s]; ok && j < f.red_ptr[s + 1]; ++j)
        ok = f.red_sub[j] == (u32)s && f.red_loc[j] >= f.loc_ptr[s] && f.red_loc[j] < f.loc_ptr[s + 1];
    }
    // elimination rows: grouped by subdomain in order, meta[s].n_elim of them,
    // their local entries a contiguous block of that subdomain, each once
    if (ok) {
      std::vector<char> seen(f.n_loc, 0);
      i64 a = 0;
      for (i64 s = 0; ok && s < f.n_sub; ++s) {
        const i64 b = a + p.meta[s].n_elim;
        ok = b <= n_elim;
        u32 lo = UINT32_MAX, hi = 0;
        for (i64 t = a; ok && t < b; ++t) {
          const u32 l = f.elim_loc[t];
          ok = f.loc_sub[l] == (u32)s && !seen[l];
          seen[l] = 1;
          lo = std::min(lo, l);
          hi = std::max(hi, l);
        }
        ok = ok && (a == b || (i64)(hi - lo) + 1 == b - a);
        a = b;
      }
      ok = ok && a == n_elim;
    }
  }
  {
    i64 se = 0, sc = 0;
    for (size_t s = 0; s < p.meta.size(); ++s) {
      const SubMeta& m = p.meta[s];
      ok = ok && m.q >= 0 && m.r >= 0 && m.n_elim >= 0 && m.n_core >= 0 && m.n_q >= 0 && m.n_r >= 0 &&
           m.c_nnz >= 0 && (m.n_elim == m.n_q || m.n_elim == m.n_r) &&
           (f.loc_ptr.size() != (size_t)f.n_sub + 1 || m.n_q + m.n_r == (i64)f.loc_ptr[s + 1] - f.loc_ptr[s]);
      se += m.n_elim;
      sc += m.n_core;
    }
    ok = ok && se == (i64)f.elim_loc.size() && sc == (i64)f.core_idx.size();
  }
  if (!ok) throw Error(E_INVALID, "preconditioner bytes are inconsistent");
  if (p.precond == 2) {
    Csr& g = p.gram;
    g.nrows = g.ncols = r.get<i64>();
    r.vec(g.indptr); r.vec(g.indices); r.vec(g.data);
    bool gok = g.nrows == n_dofs && (i64)g.indptr.size() == n_dofs + 1 && g.indices.size() == g.data.size() &&
               g.indptr[0] == 0 && g.indptr.back() == g.indices.size();
    for (i64 i = 0; gok && i < n_dofs; ++i) gok = g.indptr[i] <= g.indptr[i + 1];
    if (!gok) throw Error(E_INVALID, "preconditioner bytes are inconsistent");
    for (u32 j : g.indices)
      if ((i64)j >= n_dofs) throw Error(E_INVALID, "preconditioner bytes are inconsistent");
    p.has_gram = true;
  }
  if (r.left != 0) throw Error(E_INVALID, "trailing bytes after the preconditioner");
  p.pre_given = true;
  return n_dofs;
}

}  // namespace wa
