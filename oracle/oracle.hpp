// oracle.hpp -- FP64 CPU restatement of within's D'WDx = D'Wz projection path.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in within_amd/ may include, link or call
// this code; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg use it, as the checker and as the timed CPU baseline.
//
// Every routine cites the reference file:line whose behaviour it restates
// (paths relative to the py-econometrics/within tree).
//
// Parity status of each part (what pins it):
//   design ops (matvec_d / rmatvec_wdt / diag)   -- pinned: domain.rs:491-545 known answers
//   cross-tabs, components, PoU weights          -- pinned: docs AKM example G, invariants
//   exact Schur, anchored dense Cholesky         -- pinned: dense formula S = Dk - C' De^-1 C,
//                                                   numpy Cholesky
//   clique-tree sampling, approximate Cholesky   -- the reference's approx-chol 0.1.0 crate is
//                                                   not in the tree; this is a restatement of
//                                                   GKS 2023 (arXiv:2303.00709).  Values
//                                                   "parity unpinned" against the reference;
//                                                   pinned by math: unbiasedness, exactness on
//                                                   degree<=2 eliminations, Laplacian invariants.
//   block-elimination local solve, Schwarz       -- pinned: local solve == pseudo-inverse on
//                                                   exact factors, PoU invariants; restrict /
//                                                   prolongate / residual update on the
//                                                   reference's values (local_solve.rs:313-376,
//                                                   multiplicative.rs:392-430, operator/tests.rs:
//                                                   195-346); additive sum and multiplicative
//                                                   sweep on overlapping subdomains == dense
//                                                   sum_i R_i' W_i G_i^+ W_i R_i formulas
//   CG / GMRES / refinement                      -- pinned: cg.rs SpdMatrix3 case, closed forms,
//                                                   dense lstsq projection of small problems
//   LSMR (plain and preconditioned, windowed)    -- pinned: the reference's lsmr/tests.rs
//                                                   operators and answers, numpy lstsq, the
//                                                   CG solve of the same design
#pragma once
#include <cstdint>
#include <vector>
#include <string>
#include <stdexcept>

namespace oracle {

using u32 = uint32_t;
using i64 = int64_t;

struct WithinError : std::runtime_error {
  int code;
  WithinError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
// Error codes mirror WithinError variants (crates/within/src/error.rs:25-59).
enum ErrCode {
  ERR_EMPTY_OBS = 1, ERR_OBS_MISMATCH = 2, ERR_WEIGHT_MISMATCH = 3,
  ERR_OVERFLOW = 4, ERR_SINGULAR_DIAG = 5, ERR_LOCAL_BUILD = 6,
  ERR_PRECOND_BUILD = 7, ERR_SOLVE = 8, ERR_INVALID = 9
};

// ---------------------------------------------------------------------------
// Design (crates/within/src/domain.rs:133-156, observation.rs)
// ---------------------------------------------------------------------------
struct Design {
  i64 n_obs = 0;
  int n_factors = 0;
  std::vector<std::vector<u32>> codes;  // factor-major: codes[q][i]
  std::vector<double> weights;          // empty = unit weights
  std::vector<i64> n_levels, offset;
  i64 n_dofs = 0;
  double w(i64 i) const { return weights.empty() ? 1.0 : weights[i]; }
};
Design make_design(const u32* codes, i64 n_obs, int n_factors, i64 ld, const double* weights);
void matvec_d(const Design& d, const double* x, double* y);
void rmatvec_wdt(const Design& d, const double* r, double* x);
std::vector<double> gramian_diagonal(const Design& d);

// ---------------------------------------------------------------------------
// CSR blocks (operator/csr_block.rs, schwarz-precond/src/sparse_matrix.rs)
// ---------------------------------------------------------------------------
struct Csr {
  i64 nrows = 0, ncols = 0;
  std::vector<u32> indptr, indices;
  std::vector<double> data;
  Csr transpose() const;
  // y = base + A x  (csr_block.rs:106-135)
  void spmv_assign_add(const double* x, const double* base, double* y) const;
  void matvec(const double* x, double* y) const;
};

struct CrossTab {  // gramian/cross_tab.rs:189-199
  Csr c, ct;
  std::vector<double> diag_q, diag_r;
  i64 n_q() const { return c.nrows; }
  i64 n_r() const { return c.ncols; }
};

struct Subdomain {
  int q = 0, r = 0;
  std::vector<u32> gidx;      // q-levels then r-levels (compact-active, ascending)
  std::vector<double> pou;    // empty = uniform (all 1)
  CrossTab ct;
};

std::vector<Subdomain> build_local_domains(const Design& d);

// ---------------------------------------------------------------------------
// Approximate Cholesky / clique trees (GKS 2023 restatement; see header)
// ---------------------------------------------------------------------------
struct Edge { u32 lo, hi; double w; };
// Emit the sampled clique tree of a star into `out` (u<v pairs, self-loops dropped).
void clique_tree_sample(std::vector<std::pair<u32, double>>& nbrs, double diag,
                        uint64_t seed, int split, std::vector<Edge>& out);

struct ACFactor {
  i64 n = 0;
  std::vector<u32> round_ptr;   // rounds of independent eliminations (before the core)
  std::vector<u32> order;       // vertices, round-major, ascending inside a round
  std::vector<double> pivot;    // per vertex; 0 = anchor (solution pinned to 0)
  std::vector<u32> col_ptr;     // per vertex (indexed by vertex id)
  std::vector<u32> col_idx;     // later neighbours at elimination time
  std::vector<double> col_w;    // their edge weights  (L[u,v] = -w / pivot[v])
  std::vector<u32> core;        // dense-core vertices (ascending); last one is the anchor
  std::vector<double> core_inv; // |core|^2 row-major: inverse of the anchored core minor,
                                // zero anchor row/column
  void solve_in_place(double* x) const;  // x <- (L D L')^+ x (anchored)
};
// Laplacian given by (sorted) symmetric CSR; off-diagonal entries are -w.
// Rounds run while more than `dense_core` vertices remain; the rest is
// factored exactly (anchored dense inverse).  dense_core <= 1 = pure AC.
ACFactor ac_factor(const Csr& lap, uint64_t seed, int split, i64 dense_core);
// Inverse of the anchored minor (drop last index) of a dense Laplacian-like
// matrix a (n x n row-major): returns n x n with zero last row/col, or empty
// when the minor is not positive definite.
std::vector<double> anchored_inverse(const std::vector<double>& a, i64 n);

struct DenseAnchored {  // operator/local_solver.rs:207-300
  i64 n = 0;
  std::vector<double> l;  // (n-1)^2 row-major lower factor
  void solve_in_place(double* x) const;
};

struct LocalSolverConfig {  // crates/within/src/config.rs:83-118
  uint64_t ac_seed = 0;
  int ac_split = 1;          // split_merge (None -> 1)
  bool approx_schur = true;  // Some(ApproxSchurConfig) vs None (exact)
  uint64_t schur_seed = 0;
  int schur_split = 1;
  int dense_threshold = 24;
  i64 dense_core = 16384;    // AC stops and factors exactly below this size
};

struct BlockElim {  // operator/local_solver.rs:306-468
  CrossTab ct;
  std::vector<double> inv_diag_elim;
  bool eliminate_q = true;
  bool dense = false;
  ACFactor ac;
  DenseAnchored dn;
  Csr schur;  // the reduced Laplacian that was factored (sparse path)
  i64 n_local() const { return ct.n_q() + ct.n_r(); }
  i64 n_reduced() const { return eliminate_q ? ct.n_r() : ct.n_q(); }
  void solve(double* rhs, double* sol) const;  // rhs clobbered; buffers >= n_local + n_reduced
};
BlockElim build_block_elim(const CrossTab& ct, const LocalSolverConfig& cfg);
Csr exact_schur(const CrossTab& ct, bool& eliminate_q, std::vector<double>& inv_diag);
Csr approx_schur(const CrossTab& ct, uint64_t seed, int split, bool& eliminate_q,
                 std::vector<double>& inv_diag);

// ---------------------------------------------------------------------------
// Explicit Gramian (gramian/explicit.rs)
// ---------------------------------------------------------------------------
Csr build_gramian(const Design& d);

// ---------------------------------------------------------------------------
// Schwarz preconditioners (schwarz-precond/src/schwarz/**)
// ---------------------------------------------------------------------------
void restrict_weighted(const u32* gidx, const double* w, i64 nl, const double* global, double* local);
void prolongate_weighted_add(const u32* gidx, const double* w, i64 nl, const double* local, double* out);
void residual_update(const Csr& g, const u32* gidx, const double* delta, i64 nl, double* r_work);

struct Schwarz {
  i64 n_dofs = 0;
  bool multiplicative = false;
  std::vector<Subdomain> subs;
  std::vector<BlockElim> solvers;
  const Csr* gramian = nullptr;  // multiplicative residual updates
  void apply(const double* r, double* z) const;
};
Schwarz build_schwarz(const Design& d, const LocalSolverConfig& cfg, bool multiplicative,
                      const Csr* gramian);

// ---------------------------------------------------------------------------
// Krylov + solver (solve/cg.rs, solve/gmres.rs, crates/within/src/solver.rs)
// ---------------------------------------------------------------------------
struct KrylovResult { std::vector<double> x; bool converged; i64 iterations; double residual; };
struct LinOp {
  virtual ~LinOp() {}
  virtual void apply(const double* x, double* y) const = 0;
};
KrylovResult pcg(const LinOp& A, const double* b, i64 n, const LinOp* M, double tol, i64 maxiter);
KrylovResult pgmres(const LinOp& A, const double* b, i64 n, const LinOp* M, double tol,
                    i64 maxiter, i64 restart);

// LSMR (schwarz-precond/src/solve/lsmr.rs:122-316): min ||b - A x|| with an
// optional preconditioner M ~ A'A (modified Golub-Kahan, one M^-1 per step),
// windowed reorthogonalisation over the last `local_size` v vectors.
struct RectOp {
  virtual ~RectOp() {}
  virtual i64 rows() const = 0;
  virtual i64 cols() const = 0;
  virtual void apply(const double* x, double* y) const = 0;    // y = A x
  virtual void adjoint(const double* u, double* x) const = 0;  // x = A' u
};
enum LsmrStop { LSMR_ZERO_RHS = 0, LSMR_ATB_ZERO = 1, LSMR_RESIDUAL_TOL = 2, LSMR_NORMAL_TOL = 3,
                LSMR_BREAKDOWN = 4, LSMR_MAXITER = 5 };
struct LsmrResult { std::vector<double> x; bool converged; i64 iterations; double residual_norm; int stop; };
LsmrResult lsmr(const RectOp& A, const double* b, const LinOp* M, double tol, i64 maxiter, i64 local_size);

struct Params {
  int krylov = 0;  // 0 = CG, 1 = GMRES, 2 = LSMR
  i64 restart = 30;
  int explicit_op = 0;
  double tol = 1e-8;
  i64 maxiter = 1000;
  i64 max_refinements = 2;
  int precond = 1;  // 0 = none, 1 = additive, 2 = multiplicative
  i64 local_size = 0;  // LSMR reorthogonalisation window (0 = None)
  LocalSolverConfig local;
};

struct SolveOut {
  std::vector<double> x, demeaned;
  bool converged = false;
  i64 iterations = 0;
  double residual = 0;
};

struct Solver {
  Design design;
  Params params;
  Csr gramian;
  bool has_gramian = false;
  Schwarz schwarz;
  bool has_precond = false;
  double time_setup = 0;
  Solver(Design d, const Params& p);
  SolveOut solve(const double* y) const;
};

}  // namespace oracle
