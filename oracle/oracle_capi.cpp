// oracle_capi.cpp -- ctypes-facing C ABI of the FP64 CPU oracle.
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py, never by within_amd.
#include <omp.h>

#include <chrono>
#include <cstring>
#include <memory>

#include "oracle.hpp"

using namespace oracle;

namespace {
thread_local std::string g_err;
int fail(const std::exception& e) {
  g_err = e.what();
  if (auto* we = dynamic_cast<const WithinError*>(&e)) return we->code;
  return ERR_INVALID;
}
}  // namespace

extern "C" {

struct OracleParams {
  int krylov;            // 0 CG, 1 GMRES
  int64_t restart;
  int explicit_op;
  double tol;
  int64_t maxiter;
  int64_t max_refinements;
  int precond;           // 0 none, 1 additive, 2 multiplicative
  uint64_t ac_seed;
  int ac_split;
  int approx_schur;
  uint64_t schur_seed;
  int schur_split;
  int dense_threshold;
  int64_t dense_core;
  int64_t local_size;    // LSMR window (0 = None)
};

const char* oracle_last_error(void) { return g_err.c_str(); }

// LSMR on a dense row-major rows x cols matrix, optionally with the diagonal
// preconditioner M^-1 = diag(minv) (the operators of lsmr/tests.rs).
// info = [converged, iterations, residual_norm, stop].
int oracle_lsmr_dense(int64_t rows, int64_t cols, const double* a, const double* b, const double* minv,
                      double tol, int64_t maxiter, int64_t local_size, double* x, double* info) {
  struct Dense : RectOp {
    int64_t r, c;
    const double* a;
    i64 rows() const override { return r; }
    i64 cols() const override { return c; }
    void apply(const double* x, double* y) const override {
      for (int64_t i = 0; i < r; ++i) {
        double s = 0.0;
        for (int64_t j = 0; j < c; ++j) s += a[i * c + j] * x[j];
        y[i] = s;
      }
    }
    void adjoint(const double* u, double* x) const override {
      for (int64_t j = 0; j < c; ++j) {
        double s = 0.0;
        for (int64_t i = 0; i < r; ++i) s += a[i * c + j] * u[i];
        x[j] = s;
      }
    }
  };
  struct Diag : LinOp {
    int64_t n;
    const double* d;
    void apply(const double* x, double* y) const override {
      for (int64_t i = 0; i < n; ++i) y[i] = d[i] * x[i];
    }
  };
  try {
    Dense A;
    A.r = rows; A.c = cols; A.a = a;
    Diag M;
    M.n = cols; M.d = minv;
    LsmrResult r = lsmr(A, b, minv ? &M : nullptr, tol, maxiter, local_size);
    std::copy(r.x.begin(), r.x.end(), x);
    info[0] = r.converged; info[1] = (double)r.iterations; info[2] = r.residual_norm; info[3] = r.stop;
    return 0;
  } catch (const std::exception& e) {
    return fail(e);
  }
}

int oracle_solver_new(const uint32_t* codes, int64_t n_obs, int n_factors, int64_t ld,
                      const double* weights, const OracleParams* p, void** out) {
  try {
    Params pr;
    pr.krylov = p->krylov;
    pr.restart = p->restart;
    pr.explicit_op = p->explicit_op;
    pr.tol = p->tol;
    pr.maxiter = p->maxiter;
    pr.max_refinements = p->max_refinements;
    pr.precond = p->precond;
    pr.local.ac_seed = p->ac_seed;
    pr.local.ac_split = p->ac_split;
    pr.local.approx_schur = p->approx_schur != 0;
    pr.local.schur_seed = p->schur_seed;
    pr.local.schur_split = p->schur_split;
    pr.local.dense_threshold = p->dense_threshold;
    pr.local.dense_core = p->dense_core;
    pr.local_size = p->local_size;
    Design d = make_design(codes, n_obs, n_factors, ld, weights);
    *out = new Solver(std::move(d), pr);
    return 0;
  } catch (const std::exception& e) {
    *out = nullptr;
    return fail(e);
  }
}

void oracle_solver_free(void* s) { delete (Solver*)s; }
int64_t oracle_n_dofs(void* s) { return ((Solver*)s)->design.n_dofs; }
double oracle_time_setup(void* s) { return ((Solver*)s)->time_setup; }

// Y column-major (ldy >= n_obs), X (n_dofs x k) and demeaned (n_obs x k)
// column-major.  RHS are solved in parallel with `threads` OpenMP threads.
int oracle_solve_batch(void* sv, const double* Y, int k, int64_t ldy, double* X, double* dem,
                       int* converged, int64_t* iterations, double* residual, int threads) {
  Solver* s = (Solver*)sv;
  int64_t n = s->design.n_obs, m = s->design.n_dofs;
  int rc = 0;
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
  for (int j = 0; j < k; ++j) {
    try {
      SolveOut o = s->solve(Y + (int64_t)j * ldy);
      std::memcpy(X + (int64_t)j * m, o.x.data(), sizeof(double) * m);
      if (dem) std::memcpy(dem + (int64_t)j * n, o.demeaned.data(), sizeof(double) * n);
      converged[j] = o.converged;
      iterations[j] = o.iterations;
      residual[j] = o.residual;
    } catch (const std::exception& e) {
#pragma omp critical
      rc = fail(e);
    }
  }
  return rc;
}

int oracle_matvec_d(void* sv, const double* x, double* y) {
  matvec_d(((Solver*)sv)->design, x, y);
  return 0;
}
int oracle_rmatvec_wdt(void* sv, const double* r, double* x) {
  rmatvec_wdt(((Solver*)sv)->design, r, x);
  return 0;
}
int oracle_gramian_diag(void* sv, double* out) {
  auto g = gramian_diagonal(((Solver*)sv)->design);
  std::memcpy(out, g.data(), sizeof(double) * g.size());
  return 0;
}
// G x (implicit: D'(W(Dx)))
int oracle_gram_apply(void* sv, const double* x, double* y) {
  Solver* s = (Solver*)sv;
  const Design& d = s->design;
  std::vector<double> t(d.n_obs);
  matvec_d(d, x, t.data());
  rmatvec_wdt(d, t.data(), y);
  return 0;
}
int oracle_precond_apply(void* sv, const double* r, double* z) {
  Solver* s = (Solver*)sv;
  if (!s->has_precond) { g_err = "no preconditioner"; return ERR_INVALID; }
  s->schwarz.apply(r, z);
  return 0;
}

// ---- explicit Gramian export -------------------------------------------
int64_t oracle_gramian_nnz(void* sv) {
  Solver* s = (Solver*)sv;
  if (!s->has_gramian) { s->gramian = build_gramian(s->design); s->has_gramian = true; }
  return (int64_t)s->gramian.indices.size();
}
int oracle_gramian_export(void* sv, uint32_t* indptr, uint32_t* indices, double* data) {
  Solver* s = (Solver*)sv;
  const Csr& g = s->gramian;
  std::memcpy(indptr, g.indptr.data(), sizeof(uint32_t) * g.indptr.size());
  std::memcpy(indices, g.indices.data(), sizeof(uint32_t) * g.indices.size());
  std::memcpy(data, g.data.data(), sizeof(double) * g.data.size());
  return 0;
}

// ---- subdomain / local-solver introspection ------------------------------
int64_t oracle_n_subdomains(void* sv) { return (int64_t)((Solver*)sv)->schwarz.subs.size(); }

// info: [q, r, n_q, n_r, eliminate_q, dense, nnz_c, ac_rounds, ac_nnz, schur_nnz, has_pou,
//        core_size]
int oracle_subdomain_info(void* sv, int64_t i, int64_t* info) {
  Solver* s = (Solver*)sv;
  const Subdomain& sd = s->schwarz.subs[i];
  const BlockElim& b = s->schwarz.solvers[i];
  info[0] = sd.q; info[1] = sd.r;
  info[2] = sd.ct.n_q(); info[3] = sd.ct.n_r();
  info[4] = b.eliminate_q; info[5] = b.dense;
  info[6] = (int64_t)sd.ct.c.indices.size();
  info[7] = b.dense ? 0 : (int64_t)b.ac.round_ptr.size() - 1;
  info[8] = b.dense ? 0 : (int64_t)b.ac.col_idx.size();
  info[9] = b.dense ? 0 : (int64_t)b.schur.indices.size();
  info[10] = !sd.pou.empty();
  info[11] = b.dense ? 0 : (int64_t)b.ac.core.size();
  return 0;
}
int oracle_subdomain_gidx(void* sv, int64_t i, uint32_t* gidx, double* pou) {
  Solver* s = (Solver*)sv;
  const Subdomain& sd = s->schwarz.subs[i];
  std::memcpy(gidx, sd.gidx.data(), sizeof(uint32_t) * sd.gidx.size());
  for (size_t k = 0; k < sd.gidx.size(); ++k) pou[k] = sd.pou.empty() ? 1.0 : sd.pou[k];
  return 0;
}
// Export the factored reduced Laplacian's AC factor: pivots (n), col_ptr (n+1),
// col_idx/col_w (nnz), round_ptr (rounds+1), order (n).
int oracle_subdomain_ac(void* sv, int64_t i, double* pivot, uint32_t* col_ptr, uint32_t* col_idx,
                        double* col_w, uint32_t* round_ptr, uint32_t* order, uint32_t* core,
                        double* core_inv) {
  Solver* s = (Solver*)sv;
  const ACFactor& f = s->schwarz.solvers[i].ac;
  std::memcpy(pivot, f.pivot.data(), sizeof(double) * f.pivot.size());
  std::memcpy(col_ptr, f.col_ptr.data(), sizeof(uint32_t) * f.col_ptr.size());
  std::memcpy(col_idx, f.col_idx.data(), sizeof(uint32_t) * f.col_idx.size());
  std::memcpy(col_w, f.col_w.data(), sizeof(double) * f.col_w.size());
  std::memcpy(round_ptr, f.round_ptr.data(), sizeof(uint32_t) * f.round_ptr.size());
  std::memcpy(order, f.order.data(), sizeof(uint32_t) * f.order.size());
  std::memcpy(core, f.core.data(), sizeof(uint32_t) * f.core.size());
  std::memcpy(core_inv, f.core_inv.data(), sizeof(double) * f.core_inv.size());
  return 0;
}
// Local solve of subdomain i on a local rhs (n_local) -> sol (n_local).
int oracle_local_solve(void* sv, int64_t i, const double* rhs, double* sol) {
  Solver* s = (Solver*)sv;
  const BlockElim& b = s->schwarz.solvers[i];
  int64_t nl = b.n_local();
  std::vector<double> r(nl + b.n_reduced() + 1, 0.0), z(nl + b.n_reduced() + 1, 0.0);
  std::memcpy(r.data(), rhs, sizeof(double) * nl);
  b.solve(r.data(), z.data());
  std::memcpy(sol, z.data(), sizeof(double) * nl);
  return 0;
}
// Reduced Schur Laplacian of subdomain i (sparse path only).
int64_t oracle_subdomain_schur_nnz(void* sv, int64_t i) {
  return (int64_t)((Solver*)sv)->schwarz.solvers[i].schur.indices.size();
}
int oracle_subdomain_schur(void* sv, int64_t i, uint32_t* indptr, uint32_t* indices, double* data) {
  const Csr& S = ((Solver*)sv)->schwarz.solvers[i].schur;
  std::memcpy(indptr, S.indptr.data(), sizeof(uint32_t) * S.indptr.size());
  std::memcpy(indices, S.indices.data(), sizeof(uint32_t) * S.indices.size());
  std::memcpy(data, S.data.data(), sizeof(double) * S.data.size());
  return 0;
}

// ---- standalone building blocks (for math checks) -------------------------
// Clique tree of one star; returns the number of edges written (<= cap).
int64_t oracle_clique_tree(const uint32_t* idx, const double* w, int64_t deg, double diag,
                           uint64_t seed, int split, uint32_t* lo, uint32_t* hi, double* wo,
                           int64_t cap) {
  std::vector<std::pair<u32, double>> nb;
  for (int64_t i = 0; i < deg; ++i) nb.push_back({idx[i], w[i]});
  std::vector<Edge> out;
  clique_tree_sample(nb, diag, seed, split, out);
  int64_t c = std::min<int64_t>(cap, (int64_t)out.size());
  for (int64_t i = 0; i < c; ++i) { lo[i] = out[i].lo; hi[i] = out[i].hi; wo[i] = out[i].w; }
  return (int64_t)out.size();
}

// AC-factor a Laplacian (CSR) and solve L x = b in place; returns #rounds.
int64_t oracle_ac_solve(int64_t n, const uint32_t* indptr, const uint32_t* indices,
                        const double* data, uint64_t seed, int split, int64_t dense_core,
                        double* x) {
  Csr L;
  L.nrows = L.ncols = n;
  L.indptr.assign(indptr, indptr + n + 1);
  L.indices.assign(indices, indices + indptr[n]);
  L.data.assign(data, data + indptr[n]);
  ACFactor f = ac_factor(L, seed, split, dense_core);
  f.solve_in_place(x);
  return (int64_t)f.round_ptr.size() - 1;
}

// Exact or approximate Schur complement of a cross-tab given by dense C (nq x nr
// row-major) and diagonals.  Writes dense S (n_keep^2), returns eliminate_q or
// a negative error code.
int oracle_schur_dense(const double* cdense, int64_t nq, int64_t nr, const double* dq,
                       const double* dr, int approx, uint64_t seed, int split, double* S,
                       double* inv_diag) {
  try {
    CrossTab ct;
    ct.c.nrows = nq;
    ct.c.ncols = nr;
    ct.c.indptr.assign(nq + 1, 0);
    for (int64_t i = 0; i < nq; ++i) {
      for (int64_t j = 0; j < nr; ++j)
        if (cdense[i * nr + j] != 0.0) { ct.c.indices.push_back((u32)j); ct.c.data.push_back(cdense[i * nr + j]); }
      ct.c.indptr[i + 1] = (u32)ct.c.indices.size();
    }
    ct.ct = ct.c.transpose();
    ct.diag_q.assign(dq, dq + nq);
    ct.diag_r.assign(dr, dr + nr);
    bool eq;
    std::vector<double> inv;
    Csr s = approx ? approx_schur(ct, seed, split, eq, inv) : exact_schur(ct, eq, inv);
    int64_t nk = s.nrows;
    for (int64_t i = 0; i < nk * nk; ++i) S[i] = 0.0;
    for (int64_t i = 0; i < nk; ++i)
      for (u32 p = s.indptr[i]; p < s.indptr[i + 1]; ++p) S[i * nk + s.indices[p]] = s.data[p];
    std::memcpy(inv_diag, inv.data(), sizeof(double) * inv.size());
    return eq ? 1 : 0;
  } catch (const std::exception& e) {
    return -fail(e);
  }
}

// Restriction / prolongation / residual update of the Schwarz apply on their
// own (pinned by the reference's local_solve.rs:313-376 and
// multiplicative.rs:392-430 values); w may be NULL (uniform weights).
void oracle_restrict(const uint32_t* gidx, const double* w, int64_t nl, const double* global, double* local) {
  restrict_weighted(gidx, w, nl, global, local);
}
void oracle_prolongate_add(const uint32_t* gidx, const double* w, int64_t nl, const double* local, double* out) {
  prolongate_weighted_add(gidx, w, nl, local, out);
}
void oracle_residual_update(int64_t n, const uint32_t* indptr, const uint32_t* indices, const double* data,
                            const uint32_t* gidx, const double* delta, int64_t nl, double* r_work) {
  Csr g;
  g.nrows = g.ncols = n;
  g.indptr.assign(indptr, indptr + n + 1);
  g.indices.assign(indices, indices + indptr[n]);
  g.data.assign(data, data + indptr[n]);
  residual_update(g, gidx, delta, nl, r_work);
}

// Timed CPU baseline: solve k RHS (columns of Y) with `threads` threads; returns
// wall seconds of the solve phase (setup excluded; the solver is built already).
double oracle_time_solve_batch(void* sv, const double* Y, int k, int64_t ldy, int threads,
                               int* converged) {
  Solver* s = (Solver*)sv;
  int64_t m = s->design.n_dofs;
  std::vector<double> X((size_t)m * k);
  std::vector<int64_t> it(k);
  std::vector<double> res(k);
  auto t0 = std::chrono::steady_clock::now();
  oracle_solve_batch(sv, Y, k, ldy, X.data(), nullptr, converged, it.data(), res.data(), threads);
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // extern "C"
