// kernels.hpp -- device-side views of the uploaded plan and kernel launchers.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <vector>

#include "plan.hpp"

namespace wa {

constexpr int kBlock = 256;       // threads per workgroup (4 wave64)
constexpr int kRedBlocks = 512;   // fixed reduction grid (deterministic sums)
constexpr uint32_t kChunk = 256;  // elements per work item (long rows are split)
// The explicit Gramian's rows: on the 20M AKM Gramian (k = 32 FP32) one SpMM
// took 0.394 / 0.383 / 0.371 / 0.367 / 0.402 ms with chunks of 128 / 256 /
// 1024 / 2048 / 4096 elements; longer chunks mean fewer slots to fix up.
constexpr uint32_t kGramChunk = 1024;
constexpr uint32_t kShortRow = 16;   // rows this short are summed by one lane group
// Elimination levels: rows of 9-16 elements run faster as whole-wave items
// (preconditioner 52.8 against 53.2 ms per 32-RHS block on the 20M AKM panel;
// a limit of 4 gave 53.3 ms).
constexpr uint32_t kShortLevel = 8;
constexpr uint32_t kShortElim = 32;  // the same for elimination rows (their gathers hit L1)
constexpr uint32_t kShortPass = 8;   // the same for factor passes (several gathers per element)

struct DevSched {
  const Item* items = nullptr;
  int64_t n_items = 0;
  const Fixup* fix = nullptr;
  int64_t n_fix = 0;
  int64_t n_slots = 0;
  uint32_t short_max = 0;
  std::vector<Segment> segs;  // row ranges (one per level for level schedules)
};

// one level-sorted factor pass (plan.hpp FactorPass) in precision T
template <typename T>
struct DevPass {
  const uint32_t* row_ptr = nullptr;
  const uint32_t* other = nullptr;  // element-major (qm per element)
  const uint32_t* obs = nullptr;
  const T* ew = nullptr;            // element weights (null = unit)
  const T* diag = nullptr;
  int qm = 0;
  uint32_t off = 0;
  int64_t levels = 0, n_elem = 0;
  DevSched sched;
};

template <typename T>
struct DevCsr {
  const uint32_t* ptr = nullptr;
  const uint32_t* idx = nullptr;
  const T* val = nullptr;
  int64_t nrows = 0, nnz = 0;
  DevSched sched;
};

// A dense core applied by the dense kernels: core rows [b, b + n), matrix
// n x n row-major with leading dimension ld (a multiple of 32, zero padded).
constexpr uint32_t kDenseCore = 512;  // cores at least this large
constexpr int kCoreSplitMax = 32;     // split of the column range (partials)
template <typename T>
struct DenseCore {
  uint32_t b = 0, n = 0, ld = 0;
  const T* M = nullptr;
};

template <typename T>
struct DevPrecond {
  int64_t n_loc = 0, n_red = 0, n_sub = 0, n_elim = 0, n_levels = 0, n_core_rows = 0;
  const uint32_t *loc_gidx, *loc_ptr, *red_ptr, *loc_sub, *red_sub, *red_loc;
  const T* loc_scale;
  const uint32_t *elim_loc, *elim_ptr, *elim_idx;
  const T *elim_inv, *elim_val;
  const uint32_t *keep_ptr, *keep_idx;
  const T *keep_val, *keep_rowsum;
  const uint32_t *lvl_rows, *fin_ptr, *fin_idx, *bout_ptr, *bout_idx;
  const T *fin_val, *bout_val, *inv_pivot;
  const uint32_t *core_ptr, *core_idx, *core_row_core, *core_row_ptr;
  const uint64_t* core_mat_off;
  const T* core_mat;
  const uint32_t *dof_ptr, *dof_loc;
  DevSched s_loc, s_red, s_keep, s_elim, s_fwd, s_bwd, s_core;
  std::vector<DenseCore<T>> big;  // cores applied by the dense kernels (not in s_core)
};

// one level of a level run: its rows and items (a segment of the schedule)
struct FlowLevel {
  uint32_t r0, r1;
  int64_t i0, i1;
};
// subranges of one factor pair (multiplicative sweep, two-stream apply)
struct PairSched {
  int64_t loc0 = 0, loc1 = 0, red0 = 0, red1 = 0, core0 = 0, core1 = 0;
  DevSched s_loc, s_red, s_keep, s_elim, s_fwd, s_bwd, s_core;
  std::vector<size_t> big;  // indices into DevPrecond::big
  // levels [flow_t0, flow_t1) run as one launch per direction (flow_t0 < 0: none);
  // flow_fwd / flow_bwd list their segments in processing order (device)
  int flow_t0 = -1, flow_t1 = -1;
  const FlowLevel *flow_fwd = nullptr, *flow_bwd = nullptr;
};
// one apply's use of the level runs: counters of the forward and backward
// runs (error bits in ctr_f[1]); the previous apply's ticket totals, checked
// when `check` is set before the counters are cleared
struct FlowCall {
  uint32_t *ctr_f = nullptr, *ctr_b = nullptr;
  uint32_t spin = 0, prev_f = 0, prev_b = 0;
  int check = 0;
};

// scratch for one preconditioner apply (kp columns)
template <typename T>
struct PrecondWork {
  T *B = nullptr, *S = nullptr, *Rr = nullptr, *C = nullptr;
  T *sum1 = nullptr, *sum2 = nullptr, *sum3 = nullptr, *part = nullptr;
  T *Zc = nullptr, *P = nullptr;  // dense cores: gathered right-hand side, column-split partials
};

// per-column Krylov state (FP64 scalars)
struct ColState {
  double bnorm, rnorm, tol, rz, rz0, alpha, beta, pad;
  int active, converged, iters, nan_stop;  // nan_stop: stopped on a NaN residual
};

// per-column LSMR state (schwarz-precond/src/solve/lsmr/bidiag.rs,
// recurrence.rs): bidiagonalisation scalars, the two Givens chains, the
// x / h / hbar coefficients of the current step and the stopping state
struct LsmrCol {
  double alpha, beta, beta_inv, beta_prev_inv, p_coeff;
  double alpha_bar, phi_bar, c_bar, s_bar, zeta_bar, a_norm_sq, prev_rho, prev_rho_bar;
  double abs_tol, rel_tol, bnorm, resid;
  double inv_alpha, t_x, t_hbar, t_h;
  int active, upd, converged, iters, stop, bad;
};
constexpr int kLsmrBlocks = 2048;  // fixed grid of the observation-space LSMR reductions

// U = sw o Y (sw == nullptr: unit), part = column sums of U^2 (kLsmrBlocks blocks)
int launch_lsmr_rhs(hipStream_t st, int64_t n, const double* sw, const double* Y, double* U, int kp, double* part);
// U = sw o (D V) + s_c U with s_c = -alpha_c / beta_c (active columns), part = sums of U^2
int launch_lsmr_fwd(hipStream_t st, int64_t n, int Q, const uint32_t* codes, const double* sw, const double* V,
                    double* U, const LsmrCol* cs, int kp, double* part);
void launch_lsmr_start(hipStream_t st, LsmrCol* cs, const double* part, int nb, double tol, int k, int kp);
void launch_lsmr_alpha1(hipStream_t st, LsmrCol* cs, const double* part, int nb, int kp, int* n_active);
void launch_lsmr_beta(hipStream_t st, LsmrCol* cs, const double* part, int nb, int kp);
void launch_lsmr_step(hipStream_t st, LsmrCol* cs, const double* part, int nb, int itn, int kp, int* n_active);
// P = beta_inv ATU - p_coeff P (active columns)
void launch_lsmr_ptilde(hipStream_t st, double* P, const double* ATU, const LsmrCol* cs, int64_t m, int kp);
// V *= inv_alpha; H = V; HB = 0; X = 0 (every column)
void launch_lsmr_init_vec(hipStream_t st, double* V, double* H, double* HB, double* X, const LsmrCol* cs, int64_t m,
                          int kp);
// V *= inv_alpha; HB = H - t_hbar HB; X += t_x HB; H = V - t_h H (columns updated this step)
void launch_lsmr_update(hipStream_t st, double* V, double* H, double* HB, double* X, const LsmrCol* cs, int64_t m,
                        int kp);
// windowed Gram-Schmidt: V -= c Vj, P -= c Pj (active columns); ring push:
// Vs = V, Ps = inv_alpha P
void launch_lsmr_reorth(hipStream_t st, double* V, double* P, const double* Vj, const double* Pj, const double* c,
                        const LsmrCol* cs, int64_t m, int kp);
void launch_lsmr_push(hipStream_t st, const double* V, const double* P, double* Vs, double* Ps, const LsmrCol* cs,
                      int64_t m, int kp);

template <typename T>
void launch_pull(hipStream_t st, const DevPass<T>& p, const T* X, T* Y, T* part, int kp);
void launch_scatter_obs(hipStream_t st, const DevPass<double>& p, const double* w, const double* V, double* out,
                        double* part, int kp);
void launch_demean(hipStream_t st, int64_t n, int Q, const uint32_t* codes, const uint32_t* offs,
                   const double* X, const double* Y, double* E, int kp);
// the same demeaning from column-major Y straight into column-major D (k columns)
void launch_demean_cols(hipStream_t st, int64_t n, int Q, const uint32_t* codes, const uint32_t* offs,
                        const double* X, const double* Y, int64_t ldy, double* D, int64_t ldd, int k, int kp);
// column-major (ld) <-> row-major kp blocks; with `map`, row i of the
// destination comes from row map[i] of the source (UINT32_MAX: zeros)
void launch_col_to_row(hipStream_t st, const double* src, int64_t ld, int64_t n, int k, double* dst, int kp,
                       const uint32_t* map = nullptr);
void launch_row_to_col(hipStream_t st, const double* src, int kp, int64_t n, int k, double* dst, int64_t ld,
                       const uint32_t* map = nullptr);
template <typename T>
void launch_dot(hipStream_t st, const T* A, const T* B, int64_t m, int kp, double* part, double* out);
// per-block partials only; returns their count (finished by the CG scalar kernels)
template <typename T>
int launch_dot_part(hipStream_t st, const T* A, const T* B, int64_t m, int kp, double* part);
template <typename T>
int launch_cg_update(hipStream_t st, T* X, const T* P, T* R, const T* AP, const double* alpha, int64_t m, int kp,
                     double* part);
template <typename T>
void launch_xpby(hipStream_t st, T* P, const T* Z, const double* beta, const int* mask, int64_t m, int kp);
template <typename S, typename D>
void launch_cast_scale(hipStream_t st, const S* src, D* dst, const double* scale, int64_t m, int kp);
template <typename S>
void launch_axpy_acc(hipStream_t st, double* X, const S* Dl, const double* scale, const int* mask, int64_t m,
                     int kp);
void launch_cg_start(hipStream_t st, ColState* s, const double* bnorm2, const double* rz, const double* tol, int k,
                     int kp);
// the CG scalar steps take the per-block partials of their dot product; the
// iteration count of each column is kept in its ColState (device side)
void launch_cg_alpha(hipStream_t st, ColState* s, const double* part, int nb, double* pap, double* alpha, int kp);
void launch_cg_check(hipStream_t st, ColState* s, const double* part, int nb, double* rr, int kp);
void launch_cg_beta(hipStream_t st, ColState* s, const double* part, int nb, double* rznew, double* beta, int* mask,
                    int* n_active, int kp);
template <typename T>
void launch_precond(hipStream_t st, const DevPrecond<T>& p, const T* R, T* Z, PrecondWork<T>& w, int kp,
                    int64_t m);
template <typename T>
void launch_spmm(hipStream_t st, const DevCsr<T>& g, const T* X, T* Y, T* part, int kp);
template <typename T>
void launch_precond_mult(hipStream_t st, const DevPrecond<T>& p, const std::vector<PairSched>& pairs,
                         const DevCsr<T>& g, const T* R, T* Z, PrecondWork<T>& w, T* rwork, T* delta, T* gd,
                         const double* ones, int kp, int64_t m);
// work enqueued on a given stream by a callback (side work of an apply)
struct SideWork {
  void (*run)(hipStream_t, void*) = nullptr;
  void* ctx = nullptr;
};
// the additive apply with the pairs' local solves split over two streams
template <typename T>
void launch_precond_overlap(hipStream_t st, hipStream_t st2, hipEvent_t fork, hipEvent_t join,
                            const DevPrecond<T>& p, const std::vector<PairSched>& pairs, size_t lead, const T* R,
                            T* Z, PrecondWork<T>& w, T* part2, T* Zc2, T* P2, int kp, int64_t m,
                            const SideWork* side, const FlowCall* flow);
// ticket totals of the lead pair's level runs for kp columns of T
template <typename T>
void flow_totals(const PairSched& ps, int kp, uint32_t& tot_f, uint32_t& tot_b);
template <typename T>
void launch_cg_x(hipStream_t st, T* X, const T* P, const double* alpha, int64_t m, int kp);
template <typename T>
void launch_axpy_col(hipStream_t st, T* Y, const T* X, const double* coef, double sign, int64_t m, int kp);
template <typename T>
void launch_scale_col(hipStream_t st, T* Y, const T* X, const double* coef, int64_t m, int kp);
template <typename T>
void launch_rsub(hipStream_t st, T* R, const T* B, int64_t n);
// dst[i * ld + j] = (T) src[i * n + j] for j < n, 0 for n <= j < ld (i < n)
template <typename T>
void launch_pad_core(hipStream_t st, const double* src, uint32_t n, T* dst, uint32_t ld);
// dst[i] = (float) src[i], i < n (the host cast's rounding: to nearest even)
void launch_narrow(hipStream_t st, const double* src, float* dst, int64_t n);

}  // namespace wa
