/*
 * within_amd.h -- C ABI of the MI355X-native projection path of
 * py-econometrics/within (D'WDx = D'Wz and the demeaned residual z - Dx).
 *
 * Plain pointers and sizes only.  Host-pointer entry points copy in and out;
 * *_device entry points take pointers to gfx950 device memory.  Every call is
 * synchronous on return.  A solver handle may be shared between host threads:
 * its calls are serialised (each uses the handle's stream and work buffers).
 * Status codes mirror the reference's WithinError
 * variants (crates/within/src/error.rs:25-59); non-convergence is reported in
 * the per-RHS stats, never as an error (solver.rs:158-239).
 *
 * Each declaration cites the reference interface it replaces.
 */
#ifndef WITHIN_AMD_H
#define WITHIN_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct wa_design wa_design;
typedef struct wa_solver wa_solver;
typedef struct wa_multi wa_multi;

typedef enum {
  WA_OK = 0,
  WA_ERR_EMPTY_OBSERVATIONS = 1,   /* WithinError::EmptyObservations            */
  WA_ERR_OBS_COUNT_MISMATCH = 2,   /* WithinError::ObservationCountMismatch     */
  WA_ERR_WEIGHT_COUNT_MISMATCH = 3,/* WithinError::WeightCountMismatch          */
  WA_ERR_OVERFLOW = 4,             /* WithinError::Overflow                     */
  WA_ERR_SINGULAR_DIAGONAL = 5,    /* WithinError::SingularDiagonal             */
  WA_ERR_LOCAL_SOLVER_BUILD = 6,   /* WithinError::LocalSolverBuild             */
  WA_ERR_PRECONDITIONER_BUILD = 7, /* WithinError::PreconditionerBuild          */
  WA_ERR_ITERATIVE_SOLVE = 8,      /* WithinError::IterativeSolve               */
  WA_ERR_INVALID_ARGUMENT = 9,     /* bad sizes / pointers / config values      */
  WA_ERR_DEVICE = 10,              /* HIP runtime failure, or no gfx950 device  */
  WA_ERR_UNSUPPORTED = 11          /* configurations this library does not build */
} wa_status;

/* LocalSolverConfig + ApproxCholConfig + ApproxSchurConfig
 * (crates/within/src/config.rs:40-149).  Rust LocalSolverConfig::default():
 * ac_split 1, approx_schur 1 (seed 0, split 1), dense_threshold 24;
 * solver_default(): ac_split 2. */
typedef struct {
  uint64_t ac_seed;
  int32_t ac_split;        /* split_merge: None -> 1                              */
  int32_t approx_schur;    /* 0 = exact Schur complement, 1 = clique-tree sampled  */
  uint64_t schur_seed;
  int32_t schur_split;
  int32_t dense_threshold; /* n_keep <= threshold -> dense anchored Cholesky      */
  int64_t dense_core;      /* approximate Cholesky hands its last <= dense_core
                              vertices to an exact dense factor (0 = default 16384,
                              1 = pure approximate Cholesky)                     */
} wa_local_config;

/* SolverParams + Preconditioner (config.rs:157-237). */
typedef struct {
  int32_t krylov;          /* 0 = Cg, 1 = Gmres{restart}, 2 = Lsmr{local_size}     */
  int32_t restart;         /* GMRES restart (default 30)                           */
  int32_t operator_repr;   /* 0 = Implicit, 1 = Explicit                           */
  int32_t preconditioner;  /* 0 = none, 1 = Additive, 2 = Multiplicative           */
  double tol;              /* default 1e-8                                         */
  int64_t maxiter;         /* default 1000                                         */
  int64_t max_refinements; /* default 2                                            */
  int32_t precision;       /* 0 = FP64 throughout (reference semantics),
                              1 = FP32 inner Krylov + FP64 iterative refinement:
                              FP32 solves to max(tol', 1e-5) are repeated from
                              the FP64 residual until it meets tol, at most
                              max_refinements + 8 times (a column whose FP64
                              residual falls by less than 10 % over one FP32
                              solve is finished by FP64 refinement instead);
                              converged is that final residual test.  With the
                              explicit Gramian the
                              residual is b - Gx in FP64 (b = D'Wy), the same
                              vector as D'W(y - Dx) up to ~eps (|b| + |G||x|)  */
  int32_t block;           /* RHS per block-vector solve: 0 = auto (32), 1 = one
                              RHS per solve                                        */
  int32_t host_only;       /* 1 = build the host-side setup only (no GPU); apply
                              and solve calls then fail with WA_ERR_DEVICE        */
  int32_t local_size;      /* Lsmr reorthogonalisation window (0 = None)           */
  wa_local_config local;
} wa_params;

/* Per-RHS result fields of SolveResult / BatchSolveResult
 * (orchestrate.rs:37-134). */
typedef struct {
  int32_t converged;
  int64_t iterations;      /* all Krylov iterations incl. refinement corrections  */
  double residual;         /* ||D'W(y - Dx)|| / max(||D'Wy||, 1e-15)              */
  double time_solve;       /* seconds (the block's solve time, shared by its RHS)  */
} wa_rhs_stats;

/* Message of the last failing call on this thread (never NULL). */
const char* wa_last_error(void);

/* Fills wa_params with SolverParams::default() and Rust
 * LocalSolverConfig::default() (config.rs:96-104, :227-237), additive Schwarz. */
void wa_params_default(wa_params* p);

/* ---- design (WeightedDesign::from_store, domain.rs:133-156) ----------------
 * codes: column q of the (n_obs x n_factors) uint32 category matrix starts at
 * codes + q * ld (F-order, observation.rs:236-291).  weights: NULL = unit. */
wa_status wa_design_create(const uint32_t* codes, int64_t n_obs, int32_t n_factors, int64_t ld,
                           const double* weights, wa_design** out);
void wa_design_destroy(wa_design* d);
int64_t wa_design_n_dofs(const wa_design* d);
int64_t wa_design_n_obs(const wa_design* d);
int32_t wa_design_n_factors(const wa_design* d);
/* level count of factor q (max code + 1) */
int64_t wa_design_n_levels(const wa_design* d, int32_t q);

/* ---- solver (Solver::from_design, solver.rs:98-132) ----------------------- */
wa_status wa_solver_create(const wa_design* d, const wa_params* p, wa_solver** out);
void wa_solver_destroy(wa_solver* s);
/* Change the per-solve settings (Krylov method, tolerances, precision, block
 * size) of a built solver; preconditioner settings are fixed at build.  No
 * reference counterpart: it lets one built preconditioner serve several
 * SolverParams (the reference rebuilds a Solver per SolverParams). */
wa_status wa_solver_set_params(wa_solver* s, const wa_params* p);

/* Solver::solve_batch (solver.rs:242-277; orchestrate.rs:172-184).
 * Y: n_obs x k column-major (ldy >= n_obs).  X: n_dofs x k (ldx >= n_dofs),
 * demeaned: n_obs x k (ldd >= n_obs), both column-major; demeaned may be NULL.
 * stats: k entries. */
wa_status wa_solve_batch(wa_solver* s, const double* Y, int32_t k, int64_t ldy, double* X,
                         int64_t ldx, double* demeaned, int64_t ldd, wa_rhs_stats* stats);
/* Same, with Y / X / demeaned in device memory. */
wa_status wa_solve_batch_device(wa_solver* s, const double* Y, int32_t k, int64_t ldy, double* X,
                                int64_t ldx, double* demeaned, int64_t ldd, wa_rhs_stats* stats);

/* ---- Level A: Operator drop-ins (schwarz-precond/src/lib.rs:106-131) ------
 * Column-major blocks of k vectors (any k >= 0) with leading dimension ld.
 * The operators run on blocks of up to 32 columns in the solver's own work
 * vectors.  The *_device variants take gfx950 device pointers and allocate
 * nothing; the host-pointer variants stage through a device buffer owned by
 * the solver, allocated on the first call that needs a larger one. */
/* GramianOperator::apply (operator/gramian.rs:84-89): Y = D'W D X */
wa_status wa_gram_apply(wa_solver* s, const double* X, double* Y, int32_t k, int64_t ld);
/* Gramian (explicit CSR) Operator::apply (gramian/explicit.rs:261-277) */
wa_status wa_gram_apply_explicit(wa_solver* s, const double* X, double* Y, int32_t k, int64_t ld);
/* FePreconditioner::apply (operator/preconditioner.rs:108-154) */
wa_status wa_precond_apply(wa_solver* s, const double* R, double* Z, int32_t k, int64_t ld);
/* WeightedDesign::matvec_d (domain.rs:390-395): Y (n_obs x k) = D X */
wa_status wa_matvec_d(wa_solver* s, const double* X, int64_t ldx, double* Y, int64_t ldy, int32_t k);
/* WeightedDesign::rmatvec_wdt (domain.rs:409-418): X (n_dofs x k) = D'W R */
wa_status wa_rmatvec_wdt(wa_solver* s, const double* R, int64_t ldr, double* X, int64_t ldx,
                         int32_t k);
/* The same five operators on device pointers (no allocation, no staging). */
wa_status wa_gram_apply_device(wa_solver* s, const double* X, double* Y, int32_t k, int64_t ld);
wa_status wa_gram_apply_explicit_device(wa_solver* s, const double* X, double* Y, int32_t k, int64_t ld);
wa_status wa_precond_apply_device(wa_solver* s, const double* R, double* Z, int32_t k, int64_t ld);
wa_status wa_matvec_d_device(wa_solver* s, const double* X, int64_t ldx, double* Y, int64_t ldy, int32_t k);
wa_status wa_rmatvec_wdt_device(wa_solver* s, const double* R, int64_t ldr, double* X, int64_t ldx,
                                int32_t k);
/* The implicit operator Y = D'W D X on device pointers with no allocation
 * and without the solver's work vectors.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  Its scratch (the
 * row-major copies of up to 32 columns of X and of Y, and the factor passes'
 * partial sums) is a 64 MiB device array of the library that each GPU holds
 * until the process ends; column blocks are narrowed to fit it, and a design
 * needing more than 64 MiB per column (about 4M DOFs) is rejected with
 * WA_ERR_INVALID_ARGUMENT.  Nothing in it has to start at zero.  Every call
 * on one GPU shares the array: calls that overlap on one GPU (from several
 * host threads on different solvers) race.  The solver's device must be
 * current. */
wa_status wa_gram_apply_noalloc(wa_solver* s, const double* X, double* Y, int32_t k, int64_t ld);

/* ---- many independent problems --------------------------------------------
 * Solver::solve_batch (solver.rs:242-277) for each of n_problems independent
 * designs, run together: the problems with the same number of factors are
 * stacked block-diagonally into one design whose preconditioner, operator and
 * vector kernels serve all of them in each launch, while every (problem,
 * column) pair keeps its own CG scalars, tolerance and convergence test, as
 * a separate solve would.  CG only (KrylovMethod::Cg), additive Schwarz or no
 * preconditioner; the reference has no multi-problem entry point (each
 * problem is its own Solver).
 * wa_multi_solve: problem i has k[i] right-hand sides Y[i] (n_obs_i x k[i],
 * column-major, ld ldy[i]) and receives X[i] (n_dofs_i x k[i], ld ldx[i]),
 * demeaned[i] (may be NULL, or demeaned itself NULL) and stats[i] (k[i]
 * entries).  The pointers may be host or device memory.
 * wa_multi_info: wa_solver_info of stack `group` (groups in order of first
 * appearance of each factor count), with entry 15 = its problem count. */
wa_status wa_multi_create(const wa_design* const* designs, int32_t n_problems, const wa_params* p,
                          wa_multi** out);
void wa_multi_destroy(wa_multi* mp);
wa_status wa_multi_info(const wa_multi* mp, int32_t group, int64_t* info16);
wa_status wa_multi_solve(wa_multi* mp, const double* const* Y, const int32_t* k, const int64_t* ldy,
                         double* const* X, const int64_t* ldx, double* const* demeaned, const int64_t* ldd,
                         wa_rhs_stats* const* stats);

/* ---- serialized preconditioners (pickling of FePreconditioner) -----------
 * Replaces the postcard bytes of FePreconditioner.__reduce__ / FePreconditioner
 * (data) (crates/within-py/src/lib.rs:892-913) and Solver(categories,
 * preconditioner=p) (lib.rs:925-982) with this library's own format (not
 * byte-compatible with the reference's).
 * wa_solver_precond_bytes: buf == NULL -> *len = size; else copies (needs
 * *len >= size) and sets *len.  wa_solver_create_with_precond: a solver for the
 * design whose preconditioner (kind, local config, factors) comes from the
 * bytes; no subdomain setup runs.  wa_precond_create: the preconditioner on
 * its own (no design): only wa_precond_apply, wa_solver_info,
 * wa_solver_precond_bytes and wa_solver_params work on it. */
wa_status wa_solver_precond_bytes(const wa_solver* s, uint8_t* buf, int64_t* len);
wa_status wa_solver_create_with_precond(const wa_design* d, const wa_params* p, const uint8_t* bytes,
                                        int64_t len, wa_solver** out);
wa_status wa_precond_create(const uint8_t* bytes, int64_t len, int32_t host_only, wa_solver** out);
/* The effective parameters of a built solver (the preconditioner kind and
 * local config come from the bytes when it was built from them). */
wa_status wa_solver_params(const wa_solver* s, wa_params* out);

/* Scheduling metrics of the built Schwarz preconditioner
 * (schwarz/additive/planning.rs:140-212, operator/preconditioner.rs:64-79;
 * AdditiveSchwarzDiagnostics / subdomain_inner_parallel_work in
 * within-py/src/lib.rs:197-258, 872-876): work[i] (n_subdomains entries, may
 * be NULL) = 2 nnz(C) + n_local when the larger block of subdomain i exceeds
 * 10000 rows, else 0 (local_solver.rs:348-356); out4 = total and max of work,
 * total scatter entries (sum of local sizes), n_dofs. */
wa_status wa_precond_work(const wa_solver* s, int64_t* work, int64_t* out4);

/* ---- setup introspection (host data; no GPU needed) ------------------------
 * info[0..15]: n_subdomains, n_local_total, n_reduced_total, ac_levels,
 * ac_nnz, core_total, n_cores, c_nnz_total, gramian_nnz, n_dofs, n_obs,
 * operator_passes, dense_subdomains, max_core, device_dofs (active levels
 * the device solves over; 0 for a host-only solver), apply_streams (low byte:
 * 2 when the additive apply runs the factor pairs' local solves on two
 * streams, else 1, 0 for a host-only solver; bits 8 and up: the number of
 * trailing elimination levels run as one launch per direction). */
wa_status wa_solver_info(const wa_solver* s, int64_t* info16);
/* Subdomain i: meta[0..7] = q, r, n_q, n_r, eliminate_q, dense, rounds, core
 * size; gidx (n_q + n_r), pou (n_q + n_r) may be NULL. */
wa_status wa_solver_subdomain(const wa_solver* s, int64_t i, int64_t* meta8, uint32_t* gidx,
                              double* pou);
/* The reduced-system factor of subdomain i in the oracle's layout: pivot (n),
 * col_ptr (n+1), col_idx / col_w (nnz), round_ptr (rounds+1), order (eliminated
 * count), core (c), core_inv (c*c).  nnz_out receives the column nnz. */
wa_status wa_solver_subdomain_factor(const wa_solver* s, int64_t i, int64_t* nnz_out,
                                     double* pivot, uint32_t* col_ptr, uint32_t* col_idx,
                                     double* col_w, uint32_t* round_ptr, uint32_t* order,
                                     uint32_t* core, double* core_inv);

/* ---- autotuner ---------------------------------------------------------------
 * Picks the shape of the operator's segmented-reduction schedules (elements
 * per work item: 256..4096; longest row summed by one lane group: 4..32) for
 * the explicit Gramian SpMM and for the implicit factor passes, by timing each
 * candidate (`reps` FP32 launches of k columns, HIP events) and keeping the
 * fastest; the reference fixes the equivalent constants (domain.rs:191,199;
 * csr_block.rs:4-6).  out8: Gramian chunk, short-row limit, microseconds with
 * the default and the chosen schedule, then the same four for the factor
 * passes (-1 where the solver has no such operator). */
wa_status wa_solver_autotune(wa_solver* s, int32_t k, int32_t reps, int64_t* out8);

/* ---- learned coarse-space correction -----------------------------------------
 * No counterpart in the reference: it extends the CG loop
 * (schwarz-precond/src/solve/cg.rs:45-133) beside FePreconditioner
 * (crates/within/src/operator/preconditioner.rs:47-154).  A three-layer graph
 * network over the factor-level graph (the Gramian's sparsity; BF16 matrix
 * cores for its dense layers, the solver's FP32 operator for aggregation)
 * outputs 32 coarse vectors P.  It is trained on the device (Adam) on error /
 * residual pairs of the plain solver's own CG iterations, to maximise the
 * share of their G-energy that the coarse space captures.  When enabled, CG
 * (additive Schwarz only) runs in its flexible form with the balancing
 * two-level preconditioner z = Q r + (I - Q G) M (I - G Q) r,
 * Q = P (P'GP)^+ P'.  Plain Schwarz-PCG is the fallback: by default the
 * correction is enabled only when it cuts both the iterations and the solve
 * time of a validation block; the FP64 residual of the refinement loop judges
 * every solve. */
typedef struct {
  int32_t steps;          /* Adam steps (default 200; 0 = prepare and evaluate only) */
  int32_t sample_iters;   /* plain CG iterations behind the error samples (default 8) */
  uint64_t seed;
  double lr;              /* Adam step size (default 3e-3) */
  int32_t force;          /* 0: enable when it cuts validation iterations; 1: enable; -1: off */
  int32_t reserved;
  double ridge;           /* shift of P'GP in the loss, relative to its mean diagonal (default 1e-4) */
} wa_coarse_params;
typedef struct {
  int32_t steps, enabled;
  double loss_first, loss_last;       /* minus the captured share of the samples' G-energy */
  double iters_plain, iters_coarse;   /* mean FP32 CG iterations of the validation block (tol 1e-6) */
  double ms_plain, ms_coarse;         /* its solve times */
  double seconds;                     /* wall time of wa_coarse_train */
  int32_t kept;                       /* coarse vectors kept (G-orthonormal, at most 32) */
  int32_t reserved;
} wa_coarse_report;
void wa_coarse_params_default(wa_coarse_params* p);
/* Builds the training set, trains, validates and enables or not (see force). */
wa_status wa_coarse_train(wa_solver* s, const wa_coarse_params* p, wa_coarse_report* out);
/* Switches a trained correction on or off (later solves). */
wa_status wa_coarse_enable(wa_solver* s, int32_t on);
/* Loss and gradient (kGnnParams = 6240 floats) at the given parameters on the
 * current training set; the parameters replace the network's. */
wa_status wa_coarse_eval(wa_solver* s, const float* theta, double* loss, float* grad);
/* Copies internal data for checking (internal DOF numbering, m rows):
 * 0 features (m x 32 f32), 1 error samples E (m x 32 f32), 2 residuals R = G E,
 * 3 parameters (6240 f32), 4 internal -> user DOF (m u32), 5 nu = sum e'r (1 f64),
 * 6 coarse space P (m x 32, f32 values of the BF16 copy), 7 GP (the same),
 * 8 (P'GP)^+ (32 x 32 f64), 9 D^-1/2 (m f32), and of the last evaluation
 * (m x 32 f32): 10 dL/dP, 11 the network output P, 12 G P.  *len = element
 * count; buf may be
 * NULL to query it. */
wa_status wa_coarse_export(const wa_solver* s, int32_t what, void* buf, int64_t cap, int64_t* len);
/* The two-level preconditioner apply (FP64) on k user-numbered columns. */
wa_status wa_coarse_precond_apply(wa_solver* s, const double* R, double* Z, int32_t k, int64_t ld);

/* ---- kernel timing (HIP events on the solver's stream) ---------------------
 * Kernel classes: 0 operator, 1 preconditioner, 2 vector ops, 3 observation
 * passes, 4 explicit operator, 5 coarse correction.  Returns launches and summed milliseconds
 * since the last reset.  Timing adds event records: off unless enabled. */
void wa_timing_enable(wa_solver* s, int32_t on);
void wa_timing_reset(wa_solver* s);
wa_status wa_timing_get(wa_solver* s, int32_t kclass, int64_t* launches, double* ms);

#ifdef __cplusplus
}
#endif
#endif /* WITHIN_AMD_H */
