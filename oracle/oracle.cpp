// oracle.cpp -- FP64 CPU restatement of within's projection path.
// TEST INFRASTRUCTURE ONLY (see oracle.hpp for scope and parity status).
#include "oracle.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <numeric>

namespace oracle {

static const double EPS = std::numeric_limits<double>::epsilon();

// ===========================================================================
// Design ops -- crates/within/src/domain.rs:133-156 (levels = max+1),
// :312-345 (gather_add), :359-418 (scatter_add / rmatvec_wdt), :423-432 (diag)
// ===========================================================================
Design make_design(const u32* codes, i64 n_obs, int n_factors, i64 ld, const double* weights) {
  if (n_obs <= 0) throw WithinError(ERR_EMPTY_OBS, "no observations");
  if (n_factors <= 0) throw WithinError(ERR_INVALID, "no factors");
  Design d;
  d.n_obs = n_obs;
  d.n_factors = n_factors;
  d.codes.resize(n_factors);
  i64 off = 0;
  for (int q = 0; q < n_factors; ++q) {
    d.codes[q].assign(codes + q * ld, codes + q * ld + n_obs);
    u32 mx = 0;
    for (i64 i = 0; i < n_obs; ++i) mx = std::max(mx, d.codes[q][i]);
    d.n_levels.push_back((i64)mx + 1);
    d.offset.push_back(off);
    off += (i64)mx + 1;
  }
  d.n_dofs = off;
  if (weights) d.weights.assign(weights, weights + n_obs);
  return d;
}

void matvec_d(const Design& d, const double* x, double* y) {
  for (i64 i = 0; i < d.n_obs; ++i) y[i] = 0.0;
  for (int q = 0; q < d.n_factors; ++q) {
    const u32* c = d.codes[q].data();
    const double* xs = x + d.offset[q];
    for (i64 i = 0; i < d.n_obs; ++i) y[i] += xs[c[i]];
  }
}

void rmatvec_wdt(const Design& d, const double* r, double* x) {
  for (i64 j = 0; j < d.n_dofs; ++j) x[j] = 0.0;
#pragma omp parallel for schedule(static) if (d.n_obs > 100000)
  for (int q = 0; q < d.n_factors; ++q) {
    const u32* c = d.codes[q].data();
    double* xs = x + d.offset[q];
    if (d.weights.empty())
      for (i64 i = 0; i < d.n_obs; ++i) xs[c[i]] += r[i];
    else
      for (i64 i = 0; i < d.n_obs; ++i) xs[c[i]] += d.weights[i] * r[i];
  }
}

std::vector<double> gramian_diagonal(const Design& d) {
  std::vector<double> g(d.n_dofs, 0.0);
  for (int q = 0; q < d.n_factors; ++q)
    for (i64 i = 0; i < d.n_obs; ++i) g[d.offset[q] + d.codes[q][i]] += d.w(i);
  return g;
}

// ===========================================================================
// CSR helpers -- operator/csr_block.rs:31-62 (transpose), :106-135 (spmv)
// ===========================================================================
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
      u32 pos = cur[indices[p]]++;
      t.indices[pos] = (u32)i;
      t.data[pos] = data[p];
    }
  return t;
}

void Csr::spmv_assign_add(const double* x, const double* base, double* y) const {
  for (i64 i = 0; i < nrows; ++i) {
    double acc = base[i];
    for (u32 p = indptr[i]; p < indptr[i + 1]; ++p) acc += data[p] * x[indices[p]];
    y[i] = acc;
  }
}

void Csr::matvec(const double* x, double* y) const {
  for (i64 i = 0; i < nrows; ++i) {
    double acc = 0.0;
    for (u32 p = indptr[i]; p < indptr[i + 1]; ++p) acc += data[p] * x[indices[p]];
    y[i] = acc;
  }
}

// ===========================================================================
// Cross-tabs, components, subdomains, PoU
//   gramian/cross_tab.rs:75-145 (active levels, compact maps),
//   :514-664 (accumulation; zero sums dropped, observation-order sums),
//   :381-444 (DFS components, sorted), :450-501 (component extraction);
//   domain/factor_pairs.rs:63-80, :178-205, :272-304 (subdomains, 1/sqrt(c))
// ===========================================================================
static CrossTab accumulate_cross(const Design& d, int q, int r, const std::vector<u32>& qmap,
                                 const std::vector<u32>& rmap, i64 nq, i64 nr) {
  CrossTab ct;
  ct.diag_q.assign(nq, 0.0);
  ct.diag_r.assign(nr, 0.0);
  const u32* cq = d.codes[q].data();
  const u32* cr = d.codes[r].data();
  // bucket observations by compact q-row, preserving observation order
  std::vector<u32> cnt(nq + 1, 0);
  for (i64 i = 0; i < d.n_obs; ++i) {
    u32 a = qmap[cq[i]], b = rmap[cr[i]];
    if (a == UINT32_MAX || b == UINT32_MAX) continue;
    double w = d.w(i);
    ct.diag_q[a] += w;
    ct.diag_r[b] += w;
    cnt[a + 1]++;
  }
  for (i64 j = 0; j < nq; ++j) cnt[j + 1] += cnt[j];
  std::vector<u32> bcol(cnt[nq]);
  std::vector<double> bval(cnt[nq]);
  std::vector<u32> cur(cnt.begin(), cnt.end() - 1);
  for (i64 i = 0; i < d.n_obs; ++i) {
    u32 a = qmap[cq[i]], b = rmap[cr[i]];
    if (a == UINT32_MAX || b == UINT32_MAX) continue;
    u32 p = cur[a]++;
    bcol[p] = b;
    bval[p] = d.w(i);
  }
  Csr& c = ct.c;
  c.nrows = nq;
  c.ncols = nr;
  c.indptr.assign(nq + 1, 0);
  std::vector<double> work(nr, 0.0);
  std::vector<char> seen(nr, 0);
  std::vector<u32> touched;
  for (i64 row = 0; row < nq; ++row) {
    for (u32 p = cnt[row]; p < cnt[row + 1]; ++p) {
      u32 col = bcol[p];
      if (!seen[col]) { seen[col] = 1; touched.push_back(col); }
      work[col] += bval[p];
    }
    std::sort(touched.begin(), touched.end());
    for (u32 col : touched) {
      if (work[col] != 0.0) { c.indices.push_back(col); c.data.push_back(work[col]); }
      work[col] = 0.0;
      seen[col] = 0;
    }
    touched.clear();
    c.indptr[row + 1] = (u32)c.indices.size();
  }
  ct.ct = c.transpose();
  return ct;
}

struct Component { std::vector<u32> qi, ri; };

static std::vector<Component> bipartite_components(const CrossTab& ct) {
  i64 nq = ct.n_q(), nr = ct.n_r(), n = nq + nr;
  std::vector<char> vis(n, 0);
  std::vector<Component> comps;
  std::vector<i64> stack;
  for (i64 s = 0; s < n; ++s) {
    if (vis[s]) continue;
    vis[s] = 1;
    stack.push_back(s);
    Component comp;
    while (!stack.empty()) {
      i64 node = stack.back();
      stack.pop_back();
      if (node < nq) {
        comp.qi.push_back((u32)node);
        for (u32 p = ct.c.indptr[node]; p < ct.c.indptr[node + 1]; ++p) {
          i64 g = nq + ct.c.indices[p];
          if (!vis[g]) { vis[g] = 1; stack.push_back(g); }
        }
      } else {
        i64 ri = node - nq;
        comp.ri.push_back((u32)ri);
        for (u32 p = ct.ct.indptr[ri]; p < ct.ct.indptr[ri + 1]; ++p) {
          i64 g = ct.ct.indices[p];
          if (!vis[g]) { vis[g] = 1; stack.push_back(g); }
        }
      }
    }
    std::sort(comp.qi.begin(), comp.qi.end());
    std::sort(comp.ri.begin(), comp.ri.end());
    comps.push_back(std::move(comp));
  }
  return comps;
}

static CrossTab extract_component(const CrossTab& full, const Component& comp) {
  CrossTab ct;
  i64 nq = comp.qi.size(), nr = comp.ri.size();
  std::vector<u32> rremap(full.n_r(), UINT32_MAX);
  for (size_t k = 0; k < comp.ri.size(); ++k) rremap[comp.ri[k]] = (u32)k;
  for (u32 i : comp.qi) ct.diag_q.push_back(full.diag_q[i]);
  for (u32 i : comp.ri) ct.diag_r.push_back(full.diag_r[i]);
  ct.c.nrows = nq;
  ct.c.ncols = nr;
  ct.c.indptr.assign(nq + 1, 0);
  for (i64 k = 0; k < nq; ++k) {
    u32 old = comp.qi[k];
    for (u32 p = full.c.indptr[old]; p < full.c.indptr[old + 1]; ++p) {
      u32 nc = rremap[full.c.indices[p]];
      if (nc != UINT32_MAX) { ct.c.indices.push_back(nc); ct.c.data.push_back(full.c.data[p]); }
    }
    ct.c.indptr[k + 1] = (u32)ct.c.indices.size();
  }
  ct.ct = ct.c.transpose();
  return ct;
}

std::vector<Subdomain> build_local_domains(const Design& d) {
  int Q = d.n_factors;
  std::vector<std::vector<char>> active(Q);
  for (int q = 0; q < Q; ++q) {
    active[q].assign(d.n_levels[q], 0);
    for (i64 i = 0; i < d.n_obs; ++i) active[q][d.codes[q][i]] = 1;
  }
  std::vector<std::pair<int, int>> pairs;
  for (int q = 0; q < Q; ++q)
    for (int r = q + 1; r < Q; ++r) pairs.push_back({q, r});
  std::vector<std::vector<Subdomain>> per_pair(pairs.size());
#pragma omp parallel for schedule(dynamic, 1)
  for (size_t pi = 0; pi < pairs.size(); ++pi) {
    int q = pairs[pi].first, r = pairs[pi].second;
    std::vector<u32> qmap(d.n_levels[q], UINT32_MAX), rmap(d.n_levels[r], UINT32_MAX);
    std::vector<u32> l2g;
    i64 nq = 0, nr = 0;
    for (i64 j = 0; j < d.n_levels[q]; ++j)
      if (active[q][j]) { qmap[j] = (u32)nq++; l2g.push_back((u32)(d.offset[q] + j)); }
    for (i64 j = 0; j < d.n_levels[r]; ++j)
      if (active[r][j]) { rmap[j] = (u32)nr++; l2g.push_back((u32)(d.offset[r] + j)); }
    if (nq == 0 || nr == 0) continue;
    CrossTab full = accumulate_cross(d, q, r, qmap, rmap, nq, nr);
    std::vector<Component> comps = bipartite_components(full);
    for (const Component& comp : comps) {
      Subdomain s;
      s.q = q;
      s.r = r;
      s.ct = comps.size() == 1 ? full : extract_component(full, comp);
      for (u32 i : comp.qi) s.gidx.push_back(l2g[i]);
      for (u32 i : comp.ri) s.gidx.push_back(l2g[nq + i]);
      per_pair[pi].push_back(std::move(s));
    }
  }
  std::vector<Subdomain> subs;
  for (auto& v : per_pair)
    for (auto& s : v) subs.push_back(std::move(s));
  // partition of unity: 1/sqrt(count) where a DOF is shared
  std::vector<u32> counts(d.n_dofs, 0);
  for (auto& s : subs)
    for (u32 g : s.gidx) counts[g]++;
  for (auto& s : subs) {
    bool uniq = true;
    for (u32 g : s.gidx) uniq = uniq && counts[g] <= 1;
    if (!uniq) {
      s.pou.resize(s.gidx.size());
      for (size_t k = 0; k < s.gidx.size(); ++k) s.pou[k] = 1.0 / std::sqrt((double)counts[s.gidx[k]]);
    }
  }
  return subs;
}

// ===========================================================================
// Clique-tree sampling.  Restates GKS 2023 (arXiv:2303.00709) clique-tree
// sampling as used at operator/schur_complement.rs:128-139 (seed = seed +
// star index) through the absent approx-chol 0.1.0 crate.  Parity unpinned vs
// the crate; pinned by unbiasedness / exactness checks in tests.
//
//   * neighbours expanded into `split` copies of weight w/split,
//     stably sorted by (weight ascending, neighbour id ascending);
//   * total = diag (split == 1) or sum of weights (split > 1);
//   * for i = 0..len-2: rem = sum of weights after i; draw j > i with
//     probability w_j / rem; emit edge (u_i, u_j) of weight w_i * rem / total;
//     self-loops (copies of one neighbour) are dropped.
// RNG: splitmix64 stream seeded with the star seed; u = (next >> 11) * 2^-53.
// ===========================================================================
static inline uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
static inline double uniform01(uint64_t& s) { return (double)(splitmix64(s) >> 11) * 0x1.0p-53; }

void clique_tree_sample(std::vector<std::pair<u32, double>>& nbrs, double diag, uint64_t seed,
                        int split, std::vector<Edge>& out) {
  if (split < 1) split = 1;
  std::vector<std::pair<u32, double>> e;
  e.reserve(nbrs.size() * split);
  for (auto& p : nbrs)
    for (int c = 0; c < split; ++c) e.push_back({p.first, split == 1 ? p.second : p.second / split});
  std::stable_sort(e.begin(), e.end(), [](const std::pair<u32, double>& a, const std::pair<u32, double>& b) {
    if (a.second != b.second) return a.second < b.second;
    return a.first < b.first;
  });
  size_t len = e.size();
  if (len < 2) return;
  double total = diag;
  if (split > 1) {
    total = 0.0;
    for (auto& p : e) total += p.second;
  }
  // suffix sums S[i] = sum_{l >= i} w_l
  std::vector<double> S(len + 1, 0.0);
  for (size_t i = len; i-- > 0;) S[i] = S[i + 1] + e[i].second;
  uint64_t st = seed;
  for (size_t i = 0; i + 1 < len; ++i) {
    double rem = S[i + 1];
    double target = uniform01(st) * rem;
    // smallest j in (i, len) with S[i+1] - S[j+1] > target
    size_t lo = i + 1, hi = len - 1;
    while (lo < hi) {
      size_t mid = (lo + hi) / 2;
      if (S[i + 1] - S[mid + 1] > target) hi = mid; else lo = mid + 1;
    }
    size_t j = lo;
    double w = e[i].second * rem / total;
    u32 a = e[i].first, b = e[j].first;
    if (a == b || !(w > 0.0)) continue;
    out.push_back({std::min(a, b), std::max(a, b), w});
  }
}

// ===========================================================================
// Approximate Cholesky in rounds of independent eliminations (the build's
// GPU-schedulable restatement of GKS 2023 AC; see header for parity status).
//   round t: key(v) = (degree, hash(seed, t, v), v); v is eliminated when its
//   key is smaller than every remaining neighbour's key.  Eliminated columns:
//   pivot d_v = sum of incident weights; fill = clique tree of the
//   neighbourhood (seed = seed + v); fill edges merged in (v ascending,
//   emission order).  Isolated vertices get pivot 0 (anchored to 0).
// ===========================================================================
static inline uint64_t mis_hash(uint64_t seed, uint64_t round, uint64_t v) {
  uint64_t s = seed ^ (round * 0xD1B54A32D192ED03ULL) ^ (v * 0x8CB92BA72F3D8DD7ULL);
  return splitmix64(s);
}

// Textbook LL' of an SPD row-major matrix; false when a pivot is not positive
// (local_solver.rs:229-243 delegates this to faer's llt).
static bool cholesky(std::vector<double>& a, i64 m) {
  for (i64 j = 0; j < m; ++j) {
    double s = a[j * m + j];
    for (i64 k = 0; k < j; ++k) s -= a[j * m + k] * a[j * m + k];
    if (!(s > 0.0)) return false;
    double ljj = std::sqrt(s);
    a[j * m + j] = ljj;
    for (i64 i = j + 1; i < m; ++i) {
      double t = a[i * m + j];
      for (i64 k = 0; k < j; ++k) t -= a[i * m + k] * a[j * m + k];
      a[i * m + j] = t / ljj;
    }
    for (i64 k = j + 1; k < m; ++k) a[j * m + k] = 0.0;
  }
  return true;
}

// Laplacian core inverse: L = chol(minor); T = L^-1 stored transposed (row j =
// column j of L^-1); inv[i][j] = sum_{k >= max(i,j)} T[i][k] T[j][k].  Every
// sum runs over k ascending.
std::vector<double> anchored_inverse(const std::vector<double>& a, i64 n) {
  std::vector<double> inv(n * n, 0.0);
  i64 m = n - 1;
  if (m <= 0) return inv;
  std::vector<double> l(m * m);
  for (i64 i = 0; i < m; ++i)
    for (i64 j = 0; j < m; ++j) l[i * m + j] = a[i * n + j];
  // right-looking column Cholesky, rows of the trailing column in parallel
  for (i64 j = 0; j < m; ++j) {
    double s = l[j * m + j];
    for (i64 k = 0; k < j; ++k) s -= l[j * m + k] * l[j * m + k];
    if (!(s > 0.0)) return {};
    double ljj = std::sqrt(s);
    l[j * m + j] = ljj;
#pragma omp parallel for schedule(static) if (m - j > 512)
    for (i64 i = j + 1; i < m; ++i) {
      double t = l[i * m + j];
      for (i64 k = 0; k < j; ++k) t -= l[i * m + k] * l[j * m + k];
      l[i * m + j] = t / ljj;
    }
  }
  std::vector<double> T(m * m, 0.0);
#pragma omp parallel for schedule(dynamic, 8) if (m > 256)
  for (i64 j = 0; j < m; ++j)
    for (i64 i = j; i < m; ++i) {
      double s = i == j ? 1.0 : 0.0;
      for (i64 k = j; k < i; ++k) s -= l[i * m + k] * T[j * m + k];
      T[j * m + i] = s / l[i * m + i];
    }
#pragma omp parallel for schedule(dynamic, 8) if (m > 256)
  for (i64 i = 0; i < m; ++i)
    for (i64 j = 0; j < m; ++j) {
      double s = 0.0;
      for (i64 k = std::max(i, j); k < m; ++k) s += T[i * m + k] * T[j * m + k];
      inv[i * n + j] = s;
    }
  return inv;
}

ACFactor ac_factor(const Csr& lap, uint64_t seed, int split, i64 dense_core) {
  ACFactor f;
  i64 n = lap.nrows;
  f.n = n;
  std::vector<std::map<u32, double>> adj(n);
  for (i64 i = 0; i < n; ++i)
    for (u32 p = lap.indptr[i]; p < lap.indptr[i + 1]; ++p) {
      u32 j = lap.indices[p];
      double w = -lap.data[p];
      if ((i64)j != i && w > 0.0) adj[i][j] = w;
    }
  f.pivot.assign(n, 0.0);
  std::vector<std::vector<std::pair<u32, double>>> cols(n);
  std::vector<char> alive(n, 1);
  i64 remaining = n;
  f.round_ptr.push_back(0);
  uint64_t round = 0;
  std::vector<u32> sel;
  std::vector<uint64_t> hkey(n);
  std::vector<uint64_t> blocked(n, 0);
  i64 cut = dense_core > 1 ? dense_core : 0;
  while (remaining > 0) {
    if (remaining <= cut) {
      // exact dense factor of the remaining (core) Laplacian, anchored at its last vertex
      std::vector<u32> core;
      for (i64 v = 0; v < n; ++v)
        if (alive[v]) core.push_back((u32)v);
      i64 c = (i64)core.size();
      std::vector<i64> pos(n, -1);
      for (i64 i = 0; i < c; ++i) pos[core[i]] = i;
      std::vector<double> a(c * c, 0.0);
      for (i64 i = 0; i < c; ++i) {
        double dsum = 0.0;
        for (auto& kv : adj[core[i]]) {
          dsum += kv.second;
          a[i * c + pos[kv.first]] = -kv.second;
        }
        a[i * c + i] = dsum;
      }
      std::vector<double> inv = anchored_inverse(a, c);
      if (!inv.empty()) {
        f.core = core;
        f.core_inv = std::move(inv);
        break;
      }
      cut = 0;  // disconnected core: finish with plain rounds
    }
    // greedy maximal independent set over alive vertices in (degree, hash, id) order
    sel.clear();
    std::vector<u32> cand;
    for (i64 v = 0; v < n; ++v)
      if (alive[v]) { hkey[v] = mis_hash(seed, round, (uint64_t)v); cand.push_back((u32)v); }
    std::sort(cand.begin(), cand.end(), [&](u32 a, u32 b) {
      size_t da = adj[a].size(), db = adj[b].size();
      if (da != db) return da < db;
      if (hkey[a] != hkey[b]) return hkey[a] < hkey[b];
      return a < b;
    });
    for (u32 v : cand) {
      if (blocked[v] == round + 1) continue;
      sel.push_back(v);
      blocked[v] = round + 1;
      for (auto& kv : adj[v]) blocked[kv.first] = round + 1;
    }
    std::sort(sel.begin(), sel.end());
    std::vector<Edge> fill;
    std::vector<std::pair<u32, double>> nb;
    for (u32 v : sel) {
      nb.assign(adj[v].begin(), adj[v].end());
      double d = 0.0;
      for (auto& p : nb) d += p.second;
      f.pivot[v] = d;
      cols[v] = nb;
      if (nb.size() >= 2) clique_tree_sample(nb, d, seed + (uint64_t)v, split, fill);
      for (auto& p : cols[v]) adj[p.first].erase(v);
      adj[v].clear();
      alive[v] = 0;
      f.order.push_back(v);
    }
    for (const Edge& e : fill) {
      adj[e.lo][e.hi] += e.w;
      adj[e.hi][e.lo] += e.w;
    }
    remaining -= (i64)sel.size();
    f.round_ptr.push_back((u32)f.order.size());
    ++round;
  }
  f.col_ptr.assign(n + 1, 0);
  for (i64 v = 0; v < n; ++v) f.col_ptr[v + 1] = f.col_ptr[v] + (u32)cols[v].size();
  for (i64 v = 0; v < n; ++v)
    for (auto& p : cols[v]) { f.col_idx.push_back(p.first); f.col_w.push_back(p.second); }
  return f;
}

void ACFactor::solve_in_place(double* x) const {
  size_t R = round_ptr.size() - 1;
  for (size_t t = 0; t < R; ++t)
    for (u32 k = round_ptr[t]; k < round_ptr[t + 1]; ++k) {
      u32 v = order[k];
      if (pivot[v] <= 0.0) continue;
      double s = x[v] / pivot[v];
      for (u32 p = col_ptr[v]; p < col_ptr[v + 1]; ++p) x[col_idx[p]] += col_w[p] * s;
    }
  for (u32 k = 0; k < round_ptr[R]; ++k) {
    u32 v = order[k];
    x[v] = pivot[v] > 0.0 ? x[v] / pivot[v] : 0.0;
  }
  if (!core.empty()) {
    size_t c = core.size();
    std::vector<double> z(c), y(c, 0.0);
    for (size_t i = 0; i < c; ++i) z[i] = x[core[i]];
    for (size_t i = 0; i < c; ++i) {
      double s = 0.0;
      for (size_t j = 0; j < c; ++j) s += core_inv[i * c + j] * z[j];
      y[i] = s;
    }
    for (size_t i = 0; i < c; ++i) x[core[i]] = y[i];
  }
  for (size_t t = R; t-- > 0;)
    for (u32 k = round_ptr[t]; k < round_ptr[t + 1]; ++k) {
      u32 v = order[k];
      if (pivot[v] <= 0.0) continue;
      double acc = 0.0;
      for (u32 p = col_ptr[v]; p < col_ptr[v + 1]; ++p) acc += col_w[p] * x[col_idx[p]];
      x[v] += acc / pivot[v];
    }
}

// ===========================================================================
// Schur complements -- operator/schur_complement.rs:170-216 (eliminate the
// larger block, ties eliminate q, zero diag -> SingularDiagonal), :331-447
// (exact row workspace), :490-527 (anchored minor drops the LAST keep index),
// :534-586 (Laplacian CSR from sorted edges), :697-711 (approximate).
// ===========================================================================
struct Elim {
  bool eliminate_q;
  i64 n_keep, n_elim;
  std::vector<double> inv_diag;
  const std::vector<double>* diag_keep;
  const std::vector<double>* diag_elim;
  const Csr* keep_to_elim;
  const Csr* elim_to_keep;
};

static Elim make_elim(const CrossTab& ct) {
  Elim e;
  e.eliminate_q = ct.n_q() >= ct.n_r();
  e.n_keep = e.eliminate_q ? ct.n_r() : ct.n_q();
  e.n_elim = e.eliminate_q ? ct.n_q() : ct.n_r();
  e.diag_elim = e.eliminate_q ? &ct.diag_q : &ct.diag_r;
  e.diag_keep = e.eliminate_q ? &ct.diag_r : &ct.diag_q;
  e.keep_to_elim = e.eliminate_q ? &ct.ct : &ct.c;
  e.elim_to_keep = e.eliminate_q ? &ct.c : &ct.ct;
  for (i64 i = 0; i < e.n_elim; ++i) {
    double dd = (*e.diag_elim)[i];
    if (!(dd > 0.0))
      throw WithinError(ERR_SINGULAR_DIAG, std::string("singular eliminated diagonal at index ") +
                                               std::to_string(i));
    e.inv_diag.push_back(1.0 / dd);
  }
  return e;
}

Csr exact_schur(const CrossTab& ct, bool& eliminate_q, std::vector<double>& inv_diag) {
  Elim e = make_elim(ct);
  Csr S;
  S.nrows = S.ncols = e.n_keep;
  S.indptr.assign(e.n_keep + 1, 0);
  std::vector<double> work(e.n_keep, 0.0);
  std::vector<char> seen(e.n_keep, 0);
  std::vector<u32> touched;
  for (i64 i = 0; i < e.n_keep; ++i) {
    work[i] = (*e.diag_keep)[i];
    seen[i] = 1;
    touched.push_back((u32)i);
    const Csr& A = *e.keep_to_elim;
    const Csr& B = *e.elim_to_keep;
    for (u32 p = A.indptr[i]; p < A.indptr[i + 1]; ++p) {
      u32 k = A.indices[p];
      double scale = A.data[p] * e.inv_diag[k];
      for (u32 q = B.indptr[k]; q < B.indptr[k + 1]; ++q) {
        u32 j = B.indices[q];
        if (!seen[j]) { seen[j] = 1; touched.push_back(j); }
        work[j] -= scale * B.data[q];
      }
    }
    std::sort(touched.begin(), touched.end());
    for (u32 j : touched) {
      if (work[j] != 0.0 || (i64)j == i) { S.indices.push_back(j); S.data.push_back(work[j]); }
      work[j] = 0.0;
      seen[j] = 0;
    }
    touched.clear();
    S.indptr[i + 1] = (u32)S.indices.size();
  }
  eliminate_q = e.eliminate_q;
  inv_diag = e.inv_diag;
  return S;
}

static Csr laplacian_from_edges(std::vector<Edge>& edges, i64 n) {
  std::stable_sort(edges.begin(), edges.end(), [](const Edge& a, const Edge& b) {
    return a.lo != b.lo ? a.lo < b.lo : a.hi < b.hi;
  });
  std::vector<Edge> m;
  for (const Edge& e : edges) {
    if (!m.empty() && m.back().lo == e.lo && m.back().hi == e.hi) m.back().w += e.w;
    else m.push_back(e);
  }
  std::vector<u32> lc(n, 0), uc(n, 0);
  std::vector<double> diag(n, 0.0);
  for (const Edge& e : m) {
    uc[e.lo]++;
    lc[e.hi]++;
    diag[e.lo] += e.w;
    diag[e.hi] += e.w;
  }
  Csr S;
  S.nrows = S.ncols = n;
  S.indptr.assign(n + 1, 0);
  for (i64 i = 0; i < n; ++i) S.indptr[i + 1] = S.indptr[i] + lc[i] + 1 + uc[i];
  S.indices.resize(S.indptr[n]);
  S.data.resize(S.indptr[n]);
  std::vector<u32> lcur(n), ucur(n);
  for (i64 i = 0; i < n; ++i) {
    lcur[i] = S.indptr[i];
    u32 dpos = S.indptr[i] + lc[i];
    S.indices[dpos] = (u32)i;
    S.data[dpos] = diag[i];
    ucur[i] = dpos + 1;
  }
  for (const Edge& e : m) {
    u32 p = ucur[e.lo]++;
    S.indices[p] = e.hi;
    S.data[p] = -e.w;
    p = lcur[e.hi]++;
    S.indices[p] = e.lo;
    S.data[p] = -e.w;
  }
  return S;
}

Csr approx_schur(const CrossTab& ct, uint64_t seed, int split, bool& eliminate_q,
                 std::vector<double>& inv_diag) {
  Elim e = make_elim(ct);
  std::vector<Edge> edges;
  std::vector<std::pair<u32, double>> nb;
  const Csr& B = *e.elim_to_keep;
  for (i64 k = 0; k < e.n_elim; ++k) {
    u32 deg = B.indptr[k + 1] - B.indptr[k];
    if (deg <= 1) continue;
    nb.clear();
    for (u32 p = B.indptr[k]; p < B.indptr[k + 1]; ++p) nb.push_back({B.indices[p], B.data[p]});
    clique_tree_sample(nb, (*e.diag_elim)[k], seed + (uint64_t)k, split, edges);
  }
  eliminate_q = e.eliminate_q;
  inv_diag = e.inv_diag;
  return laplacian_from_edges(edges, e.n_keep);
}

// anchored minor: rows/cols 0..n_keep-2 of the exact Schur complement
static std::vector<double> anchored_minor(const Elim& e) {
  i64 m = e.n_keep - 1;
  if (m <= 0) return {};
  std::vector<double> a(m * m, 0.0);
  for (i64 i = 0; i < m; ++i) a[i * m + i] = (*e.diag_keep)[i];
  const Csr& A = *e.keep_to_elim;
  const Csr& B = *e.elim_to_keep;
  for (i64 i = 0; i < m; ++i)
    for (u32 p = A.indptr[i]; p < A.indptr[i + 1]; ++p) {
      u32 k = A.indices[p];
      double scale = A.data[p] * e.inv_diag[k];
      for (u32 q = B.indptr[k]; q < B.indptr[k + 1]; ++q) {
        u32 j = B.indices[q];
        if ((i64)j < m) a[i * m + j] -= scale * B.data[q];
      }
    }
  return a;
}

void DenseAnchored::solve_in_place(double* x) const {
  if (n == 0) return;
  if (n == 1) { x[0] = 0.0; return; }
  i64 m = n - 1;
  for (i64 i = 0; i < m; ++i) {
    double s = x[i];
    for (i64 j = 0; j < i; ++j) s -= l[i * m + j] * x[j];
    x[i] = s / l[i * m + i];
  }
  for (i64 i = m; i-- > 0;) {
    double s = x[i];
    for (i64 j = i + 1; j < m; ++j) s -= l[j * m + i] * x[j];
    x[i] = s / l[i * m + i];
  }
  x[m] = 0.0;
}

// operator/schwarz.rs:207-303 -- dense fast path (always exact Schur) when
// n_keep <= dense_threshold; falls back to the sparse path if LL' fails.
BlockElim build_block_elim(const CrossTab& ct, const LocalSolverConfig& cfg) {
  BlockElim b;
  b.ct = ct;
  i64 n_keep = std::min(ct.n_q(), ct.n_r());
  if (cfg.dense_threshold > 0 && n_keep <= cfg.dense_threshold) {
    Elim e = make_elim(ct);
    std::vector<double> a = anchored_minor(e);
    i64 m = std::max<i64>(n_keep - 1, 0);
    if (m == 0 || cholesky(a, m)) {
      b.dense = true;
      b.eliminate_q = e.eliminate_q;
      b.inv_diag_elim = e.inv_diag;
      b.dn.n = n_keep;
      b.dn.l = std::move(a);
      return b;
    }
  }
  if (cfg.approx_schur)
    b.schur = approx_schur(ct, cfg.schur_seed, cfg.schur_split, b.eliminate_q, b.inv_diag_elim);
  else
    b.schur = exact_schur(ct, b.eliminate_q, b.inv_diag_elim);
  b.ac = ac_factor(b.schur, cfg.ac_seed, cfg.ac_split, cfg.dense_core);
  return b;
}

static void subtract_mean(double* v, i64 n) {
  if (n == 0) return;
  double s = 0.0;
  for (i64 i = 0; i < n; ++i) s += v[i];
  double mean = s / (double)n;
  for (i64 i = 0; i < n; ++i) v[i] -= mean;
}

// operator/local_solver.rs:360-449
void BlockElim::solve(double* rhs, double* sol) const {
  i64 nq = ct.n_q(), nr = ct.n_r(), n = nq + nr;
  for (i64 i = nq; i < n; ++i) rhs[i] = -rhs[i];
  subtract_mean(rhs, n);
  i64 nred = n_reduced();
  if (eliminate_q) {
    for (i64 i = 0; i < nq; ++i) rhs[i] *= inv_diag_elim[i];
    ct.ct.spmv_assign_add(rhs, rhs + nq, rhs + n);
    subtract_mean(rhs + n, nred);
    std::memcpy(sol + nq, rhs + n, sizeof(double) * nred);
    if (dense) dn.solve_in_place(sol + nq); else ac.solve_in_place(sol + nq);
    for (i64 i = 0; i < nq; ++i) {
      double s = 0.0;
      for (u32 p = ct.c.indptr[i]; p < ct.c.indptr[i + 1]; ++p) s += ct.c.data[p] * sol[nq + ct.c.indices[p]];
      sol[i] = rhs[i] + inv_diag_elim[i] * s;
    }
  } else {
    for (i64 i = 0; i < nr; ++i) rhs[nq + i] *= inv_diag_elim[i];
    ct.c.spmv_assign_add(rhs + nq, rhs, rhs + n);
    subtract_mean(rhs + n, nred);
    std::memcpy(sol, rhs + n, sizeof(double) * nred);
    if (dense) dn.solve_in_place(sol); else ac.solve_in_place(sol);
    for (i64 i = 0; i < nr; ++i) {
      double s = 0.0;
      for (u32 p = ct.ct.indptr[i]; p < ct.ct.indptr[i + 1]; ++p) s += ct.ct.data[p] * sol[ct.ct.indices[p]];
      sol[nq + i] = rhs[nq + i] + inv_diag_elim[i] * s;
    }
  }
  subtract_mean(sol, n);
  for (i64 i = nq; i < n; ++i) sol[i] = -sol[i];
}

// ===========================================================================
// Explicit Gramian -- gramian/explicit.rs:25-108, :307-409 (observation scan,
// per-row column sort, zero off-diagonals dropped)
// ===========================================================================
Csr build_gramian(const Design& d) {
  i64 m = d.n_dofs;
  int Q = d.n_factors;
  std::vector<std::map<u32, double>> rows(m);
  for (i64 i = 0; i < d.n_obs; ++i) {
    double w = d.w(i);
    for (int q = 0; q < Q; ++q) {
      u32 a = (u32)(d.offset[q] + d.codes[q][i]);
      rows[a][a] += w;
      for (int r = 0; r < Q; ++r) {
        if (r == q) continue;
        u32 b = (u32)(d.offset[r] + d.codes[r][i]);
        rows[a][b] += w;
      }
    }
  }
  Csr g;
  g.nrows = g.ncols = m;
  g.indptr.assign(m + 1, 0);
  for (i64 a = 0; a < m; ++a) {
    for (auto& kv : rows[a])
      if (kv.second != 0.0 || (i64)kv.first == a) { g.indices.push_back(kv.first); g.data.push_back(kv.second); }
    g.indptr[a + 1] = (u32)g.indices.size();
  }
  return g;
}

// ===========================================================================
// Schwarz -- schwarz-precond/src/domain.rs:190-243 (restrict / prolongate),
// local_solve.rs:182-248; additive/preconditioner.rs:164-214;
// multiplicative.rs:128-245 with residual_update.rs:36-53 (r -= G delta)
// ===========================================================================
Schwarz build_schwarz(const Design& d, const LocalSolverConfig& cfg, bool multiplicative,
                      const Csr* gramian) {
  Schwarz s;
  s.n_dofs = d.n_dofs;
  s.multiplicative = multiplicative;
  s.gramian = gramian;
  s.subs = build_local_domains(d);
  s.solvers.resize(s.subs.size());
  std::vector<std::string> errs(s.subs.size());
  std::vector<int> codes(s.subs.size(), 0);
#pragma omp parallel for schedule(dynamic, 1)
  for (size_t i = 0; i < s.subs.size(); ++i) {
    try {
      s.solvers[i] = build_block_elim(s.subs[i].ct, cfg);
    } catch (const WithinError& e) {
      errs[i] = e.what();
      codes[i] = e.code;
    }
  }
  for (size_t i = 0; i < s.subs.size(); ++i)
    if (codes[i]) throw WithinError(codes[i], errs[i]);
  return s;
}

// local[k] = w_k * global[gidx[k]]  (schwarz-precond/src/domain.rs:190-214,
// SubdomainCore::restrict_weighted; w = 1 when the weights are uniform)
void restrict_weighted(const u32* gidx, const double* w, i64 nl, const double* global, double* local) {
  for (i64 k = 0; k < nl; ++k) local[k] = global[gidx[k]] * (w ? w[k] : 1.0);
}

// out[gidx[k]] += w_k * local[k]  (domain.rs:216-243, prolongate_weighted_add)
void prolongate_weighted_add(const u32* gidx, const double* w, i64 nl, const double* local, double* out) {
  for (i64 k = 0; k < nl; ++k) out[gidx[k]] += local[k] * (w ? w[k] : 1.0);
}

// r_work -= G[:, gidx] delta, through the rows of the symmetric G
// (operator/residual_update.rs:36-53, SparseGramianUpdater::update; the same
// vector as OperatorResidualUpdater, multiplicative.rs:55-103)
void residual_update(const Csr& g, const u32* gidx, const double* delta, i64 nl, double* r_work) {
  for (i64 k = 0; k < nl; ++k) {
    const double c = delta[k];
    if (c == 0.0) continue;
    const u32 row = gidx[k];
    for (u32 p = g.indptr[row]; p < g.indptr[row + 1]; ++p) r_work[g.indices[p]] -= c * g.data[p];
  }
}

// Additive: z = sum_i R_i' W_i A_i^+ W_i R_i r (additive/preconditioner.rs:
// 164-214, local_solve.rs:182-248).  Multiplicative: the forward sweep of
// multiplicative.rs:128-245 -- each subdomain restricts the current r_work,
// and its weighted correction is added to z and taken off r_work.
void Schwarz::apply(const double* r, double* z) const {
  for (i64 j = 0; j < n_dofs; ++j) z[j] = 0.0;
  std::vector<double> rw;
  if (multiplicative) rw.assign(r, r + n_dofs);
  const double* src = multiplicative ? rw.data() : r;
  for (size_t i = 0; i < subs.size(); ++i) {
    const Subdomain& sd = subs[i];
    const BlockElim& le = solvers[i];
    const i64 nl = sd.gidx.size();
    const double* w = sd.pou.empty() ? nullptr : sd.pou.data();
    std::vector<double> rs(nl + le.n_reduced() + 1, 0.0), zs(nl + le.n_reduced() + 1, 0.0);
    restrict_weighted(sd.gidx.data(), w, nl, src, rs.data());
    le.solve(rs.data(), zs.data());
    prolongate_weighted_add(sd.gidx.data(), w, nl, zs.data(), z);
    if (multiplicative) {
      std::vector<double> delta(nl);
      for (i64 k = 0; k < nl; ++k) delta[k] = zs[k] * (w ? w[k] : 1.0);
      residual_update(*gramian, sd.gidx.data(), delta.data(), nl, rw.data());
    }
  }
}

// ===========================================================================
// Krylov -- solve/cg.rs:45-133 and solve/gmres.rs:201-530, solve.rs:24-38
// ===========================================================================
static double dot(const double* a, const double* b, i64 n) {
  double s = 0.0;
  for (i64 i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}
static double vnorm(const double* a, i64 n) {
  double s = 0.0;
  for (i64 i = 0; i < n; ++i) s += a[i] * a[i];
  return std::sqrt(s);
}

KrylovResult pcg(const LinOp& A, const double* b, i64 n, const LinOp* M, double tol, i64 maxiter) {
  KrylovResult res;
  double bn = vnorm(b, n);
  if (bn < EPS) { res.x.assign(n, 0.0); res.converged = true; res.iterations = 0; res.residual = 0; return res; }
  std::vector<double> x(n, 0.0), r(b, b + n), z(n), ap(n), p;
  if (M) M->apply(r.data(), z.data()); else z = r;
  p = z;
  double rz = dot(r.data(), z.data(), n), rz0 = rz, rn = bn;
  for (i64 it = 1; it <= maxiter; ++it) {
    A.apply(p.data(), ap.data());
    double pap = dot(p.data(), ap.data(), n);
    if (pap <= 0.0) { res.x = x; res.converged = false; res.iterations = it; res.residual = rn; return res; }
    double alpha = rz / pap, rr = 0.0;
    for (i64 i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * ap[i];
      rr += r[i] * r[i];
    }
    rn = std::sqrt(rr);
    if (rn / bn <= tol) { res.x = x; res.converged = true; res.iterations = it; res.residual = rn; return res; }
    if (M) M->apply(r.data(), z.data()); else z = r;
    double rzn = dot(r.data(), z.data(), n);
    double thr = EPS * EPS * std::max(std::fabs(rz0), std::numeric_limits<double>::min());
    if (std::fabs(rzn) < thr) { res.x = x; res.converged = rn / bn <= tol; res.iterations = it; res.residual = rn; return res; }
    double beta = rzn / rz;
    for (i64 i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    rz = rzn;
  }
  res.x = x; res.converged = false; res.iterations = maxiter; res.residual = rn;
  return res;
}

KrylovResult pgmres(const LinOp& A, const double* b, i64 n, const LinOp* M, double tol,
                    i64 maxiter, i64 restart) {
  KrylovResult res;
  double bn = vnorm(b, n);
  if (bn < EPS) { res.x.assign(n, 0.0); res.converged = true; res.iterations = 0; res.residual = 0; return res; }
  double abs_tol = tol * bn;
  i64 m = restart;
  std::vector<double> x(n, 0.0), r(n), w(n);
  std::vector<std::vector<double>> V, Z;
  std::vector<double> H((m + 1) * m), cs(m), sn(m), g(m + 1);
  auto h = [&](i64 i, i64 j) -> double& { return H[j * (m + 1) + i]; };
  i64 total = 0;
  auto residual = [&]() {
    A.apply(x.data(), r.data());
    for (i64 i = 0; i < n; ++i) r[i] = b[i] - r[i];
    return vnorm(r.data(), n);
  };
  while (true) {
    double beta = residual();
    if (beta <= abs_tol) { res.x = x; res.converged = true; res.iterations = total; res.residual = beta; return res; }
    if (total >= maxiter) { res.x = x; res.converged = false; res.iterations = total; res.residual = beta; return res; }
    V.clear(); Z.clear();
    std::fill(H.begin(), H.end(), 0.0);
    std::fill(cs.begin(), cs.end(), 0.0);
    std::fill(sn.begin(), sn.end(), 0.0);
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    for (i64 i = 0; i < n; ++i) r[i] *= 1.0 / beta;
    V.push_back(r);
    i64 cycle = std::min(maxiter - total, m);
    i64 j = 0;
    int outcome = 3;  // 0 conv, 1 maxiter, 2 breakdown, 3 restart
    double resid = 0.0, hmax = 0.0;
    while (j < cycle) {
      Z.push_back(std::vector<double>(n));
      if (M) M->apply(V[j].data(), Z[j].data()); else Z[j] = V[j];
      A.apply(Z[j].data(), w.data());
      double lmax = 0.0;
      for (i64 i = 0; i <= j; ++i) {
        double hij = dot(w.data(), V[i].data(), n);
        h(i, j) = hij;
        lmax = std::max(lmax, std::fabs(hij));
        for (i64 k = 0; k < n; ++k) w[k] -= hij * V[i][k];
      }
      double hj1 = vnorm(w.data(), n);
      h(j + 1, j) = hj1;
      hmax = std::max(std::max(hmax, lmax), hj1);
      for (i64 i = 0; i < j; ++i) {
        double a = h(i, j), bb = h(i + 1, j);
        double t = cs[i] * a + sn[i] * bb;
        h(i + 1, j) = -sn[i] * a + cs[i] * bb;
        h(i, j) = t;
      }
      double a = h(j, j), bb = h(j + 1, j), rr = std::hypot(a, bb);
      double c = rr == 0.0 ? 1.0 : a / rr, s = rr == 0.0 ? 0.0 : bb / rr;
      cs[j] = c; sn[j] = s;
      h(j, j) = c * a + s * bb;
      h(j + 1, j) = 0.0;
      double t = c * g[j] + s * g[j + 1];
      g[j + 1] = -s * g[j] + c * g[j + 1];
      g[j] = t;
      ++j; ++total;
      resid = std::fabs(g[j]);
      if (resid <= abs_tol) { outcome = 0; break; }
      if (total >= maxiter) { outcome = 1; break; }
      if (hj1 > 1e-14 * hmax) {
        for (i64 k = 0; k < n; ++k) w[k] *= 1.0 / hj1;
        V.push_back(w);
      } else { outcome = 2; break; }
    }
    std::vector<double> y(j, 0.0);
    for (i64 i = j; i-- > 0;) {
      double sum = g[i];
      for (i64 k = i + 1; k < j; ++k) sum -= h(i, k) * y[k];
      double dg = h(i, i);
      y[i] = std::fabs(dg) < 1e-14 * (1.0 + std::fabs(sum)) ? 0.0 : sum / dg;
    }
    for (i64 k = 0; k < j; ++k)
      for (i64 i = 0; i < n; ++i) x[i] += y[k] * Z[k][i];
    if (outcome == 0) { res.x = x; res.converged = true; res.iterations = total; res.residual = resid; return res; }
    if (outcome == 1) { res.x = x; res.converged = false; res.iterations = total; res.residual = resid; return res; }
    if (outcome == 2) {
      double rr = residual();
      res.x = x; res.converged = rr <= abs_tol; res.iterations = total; res.residual = rr;
      return res;
    }
  }
}

// ===========================================================================
// LSMR -- schwarz-precond/src/solve/lsmr.rs:122-316 (driver),
// lsmr/bidiag.rs:86-590 (Golub-Kahan / modified Golub-Kahan, windowed MGS),
// lsmr/recurrence.rs:15-285 (the two Givens chains, the x/h/hbar recurrence,
// the two stopping tests)
// ===========================================================================
namespace {
struct Givens {  // recurrence.rs:20-30
  double c, s, r;
  Givens(double a, double b) {
    r = std::hypot(a, b);
    if (r > 0.0) { c = a / r; s = b / r; } else { c = 1.0; s = 0.0; }
  }
};
struct Rot { double rho = 1.0, rho_bar = 1.0, theta_new = 0.0, theta_bar = 0.0, zeta = 0.0; };

// ring of past v (and, preconditioned, p~ = M v) vectors (bidiag.rs:86-262)
struct Ring {
  i64 n = 0, cap = 0, next = 0, count = 0;
  std::vector<double> v, p;
  Ring(i64 rows, i64 cols, i64 local_size, bool paired) : n(cols) {
    cap = std::max<i64>(0, std::min(local_size, std::min(rows, cols)));
    v.assign(cap * n, 0.0);
    if (paired) p.assign(cap * n, 0.0);
  }
  // modified Gram-Schmidt, oldest first; the coefficient is <y, v_j> (plain)
  // or <y, p_j> (M-weighted, p~ updated in lockstep)
  void reorth(double* y, double* pt) const {
    const i64 start = count < cap ? 0 : next;
    for (i64 i = 0; i < count; ++i) {
      const i64 j = (start + i) % cap;
      const double* vj = v.data() + j * n;
      const double* pj = pt ? p.data() + j * n : vj;
      const double c = dot(y, pj, n);
      for (i64 t = 0; t < n; ++t) y[t] = -c * vj[t] + y[t];
      if (pt)
        for (i64 t = 0; t < n; ++t) pt[t] = -c * pj[t] + pt[t];
    }
  }
  void push(const double* y, const double* pt, double inv_alpha) {
    std::copy(y, y + n, v.begin() + next * n);
    if (pt)
      for (i64 t = 0; t < n; ++t) p[next * n + t] = pt[t] * inv_alpha;
    next = (next + 1) % cap;
    if (count < cap) ++count;
  }
};

struct Bidiag {  // GolubKahan (M == nullptr) or ModifiedGolubKahan
  const RectOp& A;
  const LinOp* M;
  i64 m, n;
  std::vector<double> u, v, pt, av, atu;
  Ring ring;
  double alpha = 0.0, beta_prev_inv = 1.0;
  Bidiag(const RectOp& a, const LinOp* mm, i64 local_size)
      : A(a), M(mm), m(a.rows()), n(a.cols()), u(m, 0.0), v(n, 0.0), pt(mm ? n : 0, 0.0), av(m), atu(n),
        ring(a.rows(), a.cols(), local_size, mm != nullptr) {}

  // bidiag.rs:476-505 / 551-589; returns (alpha_1, beta_1)
  std::pair<double, double> init(const double* b) {
    const double beta = std::sqrt(dot(b, b, m));
    if (beta > 0.0) {
      const double inv = 1.0 / beta;
      for (i64 i = 0; i < m; ++i) u[i] = b[i] * inv;
    }
    if (!M) {
      A.adjoint(u.data(), v.data());
      alpha = std::sqrt(dot(v.data(), v.data(), n));
      if (alpha > 0.0) scale(v, 1.0 / alpha);
      if (ring.cap) ring.push(v.data(), nullptr, 0.0);
    } else {
      A.adjoint(u.data(), pt.data());
      M->apply(pt.data(), v.data());
      const double vp = dot(v.data(), pt.data(), n);
      alpha = vp > 0.0 ? std::sqrt(vp) : 0.0;
      if (alpha > 0.0) scale(v, 1.0 / alpha);
      if (ring.cap) ring.push(v.data(), pt.data(), alpha > 0.0 ? 1.0 / alpha : 0.0);
      beta_prev_inv = 1.0;
    }
    return {alpha, beta};
  }
  static void scale(std::vector<double>& y, double s) {
    for (double& t : y) t *= s;
  }
  // y = x + s y, returns ||y||^2 (bidiag.rs:19-38)
  static double axpy_sq(std::vector<double>& y, const std::vector<double>& x, double s) {
    double acc = 0.0;
    for (size_t i = 0; i < y.size(); ++i) {
      const double val = x[i] + s * y[i];
      y[i] = val;
      acc += val * val;
    }
    return acc;
  }
  // one step (bidiag.rs:286-340 / 342-430): returns (alpha_{k+1}, beta_{k+1})
  std::pair<double, double> step() {
    if (!M) {
      A.apply(v.data(), av.data());
      const double beta = std::sqrt(axpy_sq(u, av, -alpha));
      if (beta == 0.0) {
        std::fill(v.begin(), v.end(), 0.0);
        alpha = 0.0;
        return {0.0, beta};
      }
      scale(u, 1.0 / beta);
      A.adjoint(u.data(), atu.data());
      double a2 = axpy_sq(v, atu, -beta);
      if (ring.cap) {
        ring.reorth(v.data(), nullptr);
        a2 = dot(v.data(), v.data(), n);
      }
      const double a = std::sqrt(a2);
      if (a > 0.0) scale(v, 1.0 / a);
      if (ring.cap) ring.push(v.data(), nullptr, 0.0);
      alpha = a;
      return {a, beta};
    }
    const double s = -(alpha * beta_prev_inv);
    A.apply(v.data(), av.data());
    const double beta = std::sqrt(axpy_sq(u, av, s));
    if (beta == 0.0) {
      std::fill(v.begin(), v.end(), 0.0);
      std::fill(pt.begin(), pt.end(), 0.0);
      alpha = 0.0;
      beta_prev_inv = 0.0;
      return {0.0, beta};
    }
    const double beta_inv = beta > 0.0 ? 1.0 / beta : 0.0;
    A.adjoint(u.data(), atu.data());
    const double p_coeff = beta / alpha;
    for (i64 i = 0; i < n; ++i) pt[i] = beta_inv * atu[i] + (-p_coeff) * pt[i];
    M->apply(pt.data(), v.data());
    if (ring.cap) ring.reorth(v.data(), pt.data());
    const double vp = dot(v.data(), pt.data(), n);
    const double a = vp > 0.0 ? std::sqrt(vp) : 0.0;
    if (a > 0.0) scale(v, 1.0 / a);
    if (ring.cap) ring.push(v.data(), pt.data(), a > 0.0 ? 1.0 / a : 0.0);
    alpha = a;
    beta_prev_inv = beta_inv;
    return {a, beta};
  }
};
}  // namespace

LsmrResult lsmr(const RectOp& A, const double* b, const LinOp* M, double tol, i64 maxiter, i64 local_size) {
  const i64 m = A.rows(), n = A.cols();
  // lsmr.rs:254-283 input validation
  if (!std::isfinite(tol) || tol < 0.0) throw WithinError(ERR_SOLVE, "lsmr: tolerance must be finite and nonnegative");
  for (i64 i = 0; i < m; ++i)
    if (!std::isfinite(b[i])) throw WithinError(ERR_SOLVE, "lsmr: rhs entries must be finite");
  LsmrResult res;
  const double b_norm = vnorm(b, m);
  if (b_norm == 0.0) {  // lsmr.rs:193-201
    res.x.assign(n, 0.0); res.converged = true; res.iterations = 0; res.residual_norm = 0.0;
    res.stop = LSMR_ZERO_RHS;
    return res;
  }
  Bidiag bd(A, M, local_size);
  const auto s1 = bd.init(b);
  if (s1.first == 0.0) {  // lsmr.rs:214-222
    res.x.assign(n, 0.0); res.converged = true; res.iterations = 0; res.residual_norm = b_norm;
    res.stop = LSMR_ATB_ZERO;
    return res;
  }
  // recurrence state (recurrence.rs:78-86), solution state (:168-174),
  // convergence state (:240-246)
  double alpha_bar = s1.first, phi_bar = s1.second, c_bar = 1.0, s_bar = 0.0, zeta_bar = s1.first * s1.second;
  std::vector<double> x(n, 0.0), h(bd.v), hb(n, 0.0);
  const double abs_tol = tol * b_norm, rel_tol = tol;
  double a_norm_sq = s1.first * s1.first;
  Rot prev;
  const double eps = std::numeric_limits<double>::epsilon(), tiny = std::numeric_limits<double>::min();
  for (i64 itn = 1; itn <= maxiter; ++itn) {
    const auto st = bd.step();
    const double alpha = st.first, beta = st.second;
    a_norm_sq += alpha * alpha + beta * beta;
    // P-hat then P-bar (recurrence.rs:89-138)
    const Givens ph(alpha_bar, beta);
    Rot cur;
    cur.theta_new = ph.s * alpha;
    const double alpha_bar_new = -ph.c * alpha, phi_bar_new = ph.s * phi_bar;
    cur.theta_bar = s_bar * ph.r;
    const Givens pb(c_bar * ph.r, cur.theta_new);
    cur.zeta = pb.c * zeta_bar;
    const double zeta_bar_new = -pb.s * zeta_bar;
    alpha_bar = alpha_bar_new; phi_bar = phi_bar_new; c_bar = pb.c; s_bar = pb.s; zeta_bar = zeta_bar_new;
    cur.rho = ph.r;
    cur.rho_bar = pb.r;
    // x / h / hbar (recurrence.rs:181-229)
    const double dx = cur.rho * cur.rho_bar, dh = prev.rho * prev.rho_bar;
    const double t_x = std::fabs(dx) > eps ? cur.zeta / dx : 0.0;
    const double t_hbar = std::fabs(dh) > eps ? cur.theta_bar * cur.rho / dh : 0.0;
    const double t_h = std::fabs(cur.rho) > eps ? cur.theta_new / cur.rho : 0.0;
    for (i64 i = 0; i < n; ++i) {
      const double h_old = h[i];
      const double hbi = h_old - t_hbar * hb[i];
      hb[i] = hbi;
      x[i] += t_x * hbi;
      h[i] = bd.v[i] - t_h * h_old;
    }
    // stopping tests (recurrence.rs:263-275); convergence wins over breakdown
    const double resid = std::fabs(phi_bar);
    int stop = -1;
    if (resid <= abs_tol) stop = LSMR_RESIDUAL_TOL;
    else {
      const double a_norm = std::max(std::sqrt(a_norm_sq), tiny);
      if (std::fabs(zeta_bar) / (a_norm * std::max(resid, tiny)) <= rel_tol) stop = LSMR_NORMAL_TOL;
    }
    if (stop < 0 && alpha == 0.0) stop = LSMR_BREAKDOWN;
    if (stop >= 0) {
      res.x = std::move(x); res.converged = true; res.iterations = itn; res.residual_norm = resid; res.stop = stop;
      return res;
    }
    prev = cur;
  }
  res.x = std::move(x); res.converged = false; res.iterations = maxiter; res.residual_norm = std::fabs(phi_bar);
  res.stop = LSMR_MAXITER;
  return res;
}

// ===========================================================================
// Solver -- crates/within/src/solver.rs:98-132 (build), :158-239 (solve),
// :296-347 (iterative refinement), :353-398 (dispatch)
// ===========================================================================
struct ImplicitOp : LinOp {
  const Design& d;
  explicit ImplicitOp(const Design& dd) : d(dd) {}
  void apply(const double* x, double* y) const override {
    std::vector<double> t(d.n_obs);
    matvec_d(d, x, t.data());
    if (!d.weights.empty())
      for (i64 i = 0; i < d.n_obs; ++i) t[i] *= d.weights[i];
    // G x = D' (W (D x)); rmatvec_dt on the weighted intermediate
    for (i64 j = 0; j < d.n_dofs; ++j) y[j] = 0.0;
    for (int q = 0; q < d.n_factors; ++q) {
      const u32* c = d.codes[q].data();
      double* ys = y + d.offset[q];
      for (i64 i = 0; i < d.n_obs; ++i) ys[c[i]] += t[i];
    }
  }
};
struct CsrOp : LinOp {
  const Csr& g;
  explicit CsrOp(const Csr& gg) : g(gg) {}
  void apply(const double* x, double* y) const override { g.matvec(x, y); }
};
struct SchwarzOp : LinOp {
  const Schwarz& s;
  explicit SchwarzOp(const Schwarz& ss) : s(ss) {}
  void apply(const double* x, double* y) const override { s.apply(x, y); }
};

Solver::Solver(Design d, const Params& p) : design(std::move(d)), params(p) {
  auto t0 = std::chrono::steady_clock::now();
  bool need_g = p.explicit_op || p.precond == 2;
  if (need_g) { gramian = build_gramian(design); has_gramian = true; }
  if (p.precond == 1 || p.precond == 2) {
    schwarz = build_schwarz(design, p.local, p.precond == 2, has_gramian ? &gramian : nullptr);
    has_precond = true;
  }
  time_setup = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// W^(1/2) D and its adjoint (operator.rs:99-181)
struct WeightedDesignOp : RectOp {
  const Design& d;
  std::vector<double> sw;  // sqrt of the weights (empty = unit)
  explicit WeightedDesignOp(const Design& dd) : d(dd) {
    if (!d.weights.empty()) {
      sw.resize(d.n_obs);
      for (i64 i = 0; i < d.n_obs; ++i) sw[i] = std::sqrt(d.weights[i]);
    }
  }
  i64 rows() const override { return d.n_obs; }
  i64 cols() const override { return d.n_dofs; }
  void apply(const double* x, double* y) const override {
    matvec_d(d, x, y);
    if (!sw.empty())
      for (i64 i = 0; i < d.n_obs; ++i) y[i] *= sw[i];
  }
  void adjoint(const double* u, double* x) const override {
    std::vector<double> t(u, u + d.n_obs);
    if (!sw.empty())
      for (i64 i = 0; i < d.n_obs; ++i) t[i] = sw[i] * u[i];
    for (i64 j = 0; j < d.n_dofs; ++j) x[j] = 0.0;
    for (int q = 0; q < d.n_factors; ++q) {
      const u32* c = d.codes[q].data();
      double* xs = x + d.offset[q];
      for (i64 i = 0; i < d.n_obs; ++i) xs[c[i]] += t[i];
    }
  }
};

SolveOut Solver::solve(const double* y) const {
  const Design& d = design;
  i64 m = d.n_dofs, n = d.n_obs;
  if (params.krylov == 2) {  // solver.rs:162-207
    WeightedDesignOp rect(d);
    std::vector<double> b(y, y + n);
    if (!rect.sw.empty())
      for (i64 i = 0; i < n; ++i) b[i] = rect.sw[i] * y[i];
    SchwarzOp sop(schwarz);
    LsmrResult r = lsmr(rect, b.data(), has_precond ? &sop : nullptr, params.tol, params.maxiter,
                        params.local_size);
    SolveOut out;
    out.x = std::move(r.x);
    out.converged = r.converged;
    out.iterations = r.iterations;
    out.demeaned.assign(n, 0.0);
    matvec_d(d, out.x.data(), out.demeaned.data());
    for (i64 i = 0; i < n; ++i) out.demeaned[i] = y[i] - out.demeaned[i];
    std::vector<double> rhs(m), res(m);
    rmatvec_wdt(d, y, rhs.data());
    rmatvec_wdt(d, out.demeaned.data(), res.data());
    out.residual = vnorm(res.data(), m) / std::max(vnorm(rhs.data(), m), 1e-15);
    return out;
  }
  ImplicitOp iop(d);
  CsrOp cop(gramian);
  const LinOp& A = has_gramian && params.explicit_op ? (const LinOp&)cop : (const LinOp&)iop;
  SchwarzOp sop(schwarz);
  const LinOp* M = has_precond ? &sop : nullptr;
  auto krylov = [&](const double* rhs, double tol) {
    if (params.krylov == 1) return pgmres(A, rhs, m, M, tol, params.maxiter, params.restart);
    return pcg(A, rhs, m, M, tol, params.maxiter);
  };
  std::vector<double> rhs(m);
  rmatvec_wdt(d, y, rhs.data());
  double rhs_norm = std::max(vnorm(rhs.data(), m), 1e-15);
  double abs_tol = params.tol * rhs_norm;
  KrylovResult k0 = krylov(rhs.data(), params.tol);
  SolveOut out;
  out.x = k0.x;
  out.iterations = k0.iterations;
  out.converged = k0.converged;
  out.demeaned.assign(n, 0.0);
  std::vector<double> corr(m);
  for (i64 it = 0; it < params.max_refinements; ++it) {
    matvec_d(d, out.x.data(), out.demeaned.data());
    for (i64 i = 0; i < n; ++i) out.demeaned[i] = y[i] - out.demeaned[i];
    rmatvec_wdt(d, out.demeaned.data(), corr.data());
    double cn = vnorm(corr.data(), m);
    if (cn <= abs_tol) break;
    double ctol = std::min(abs_tol / cn, 1.0);
    KrylovResult kc = krylov(corr.data(), ctol);
    for (i64 j = 0; j < m; ++j) out.x[j] += kc.x[j];
    out.iterations += kc.iterations;
    out.converged = kc.converged;
  }
  matvec_d(d, out.x.data(), out.demeaned.data());
  for (i64 i = 0; i < n; ++i) out.demeaned[i] = y[i] - out.demeaned[i];
  rmatvec_wdt(d, out.demeaned.data(), corr.data());
  out.residual = vnorm(corr.data(), m) / rhs_norm;
  return out;
}

}  // namespace oracle
