"""Python surface of ``within._within`` on the MI355X-native path.

Mirrors the PyO3 bridge of the reference (``crates/within-py/src/lib.rs``
:62-1074, stubs in ``python/within/_within.pyi``): the same config classes,
defaults, result types and validation, backed by ``libwithin_amd.so``.
Front-door defaults follow the reference exactly:

* ``preconditioner=None`` means additive Schwarz with
  ``LocalSolverConfig::solver_default()`` (AC split 2) (lib.rs:546-552);
* ``SchurComplement()`` with no arguments means an exact Schur complement and
  AC split 1 (lib.rs:499-514);
* CG or LSMR with a multiplicative preconditioner raises ``ValueError``
  (lib.rs:639-656).

Extension: ``CG`` / ``GMRES`` accept ``precision="mixed"`` (FP32 inner
Krylov + FP64 iterative refinement) and ``block`` (right-hand sides per
block-vector solve) in addition to the reference fields.
"""

from __future__ import annotations

import ctypes as C
import time
import warnings
from enum import IntEnum

import numpy as np

from . import _native as N


class OperatorRepr(IntEnum):
    Implicit = 0
    Explicit = 1


class Preconditioner(IntEnum):
    Additive = 0
    Multiplicative = 1
    Off = 2


class ReductionStrategy(IntEnum):
    """Accepted for API compatibility; the GPU combine is always the
    deterministic gather over each DOF's subdomain slots."""
    Auto = 0
    AtomicScatter = 1
    ParallelReduction = 2


_PRECISIONS = {"fp64": 0, "mixed": 1}


class CG:
    def __init__(self, tol=1e-8, maxiter=1000, operator=OperatorRepr.Implicit, max_refinements=2,
                 precision="fp64", block=None):
        self.tol = float(tol)
        self.maxiter = int(maxiter)
        self.operator = OperatorRepr(operator)
        self.max_refinements = int(max_refinements)
        if precision not in _PRECISIONS:
            raise ValueError("precision must be 'fp64' or 'mixed'")
        self.precision = precision
        self.block = block

    def __repr__(self):
        return (f"CG(tol={self.tol}, maxiter={self.maxiter}, operator={self.operator.name}, "
                f"max_refinements={self.max_refinements}, precision={self.precision!r})")


class GMRES:
    def __init__(self, tol=1e-8, maxiter=1000, restart=30, operator=OperatorRepr.Implicit, max_refinements=2,
                 precision="fp64", block=None):
        self.tol = float(tol)
        self.maxiter = int(maxiter)
        self.restart = int(restart)
        self.operator = OperatorRepr(operator)
        self.max_refinements = int(max_refinements)
        if precision not in _PRECISIONS:
            raise ValueError("precision must be 'fp64' or 'mixed'")
        self.precision = precision
        self.block = block

    def __repr__(self):
        return (f"GMRES(tol={self.tol}, maxiter={self.maxiter}, restart={self.restart}, "
                f"operator={self.operator.name}, max_refinements={self.max_refinements})")


class LSMR:
    """Modified LSMR on W^(1/2) D with the Schwarz preconditioner as M ~ D'WD
    (crates/within-py/src/lib.rs:398-421; schwarz-precond/src/solve/lsmr.rs).
    ``local_size``: windowed reorthogonalisation over that many past vectors
    (None = off).  Runs in FP64 on the GPU with no refinement step."""

    def __init__(self, tol=1e-8, maxiter=1000, local_size=None):
        self.tol = float(tol)
        self.maxiter = int(maxiter)
        if local_size is not None and int(local_size) < 0:
            raise ValueError("local_size must be >= 0")
        self.local_size = None if local_size is None else int(local_size)

    def __repr__(self):
        return f"LSMR(tol={self.tol}, maxiter={self.maxiter}, local_size={self.local_size})"


class ApproxCholConfig:
    def __init__(self, seed=0, split=1):
        if int(split) < 1:
            raise ValueError("split must be >= 1")
        self.seed = int(seed)
        self.split = int(split)


class ApproxSchurConfig:
    def __init__(self, seed=0, split=1):
        if int(split) < 1:
            raise ValueError("split must be >= 1")
        self.seed = int(seed)
        self.split = int(split)


class SchurComplement:
    def __init__(self, approx_chol=None, approx_schur=None, dense_threshold=24):
        if approx_chol is not None and not isinstance(approx_chol, ApproxCholConfig):
            raise TypeError("approx_chol must be ApproxCholConfig or None")
        if approx_schur is not None and not isinstance(approx_schur, ApproxSchurConfig):
            raise TypeError("approx_schur must be ApproxSchurConfig or None")
        self.approx_chol = approx_chol
        self.approx_schur = approx_schur
        self.dense_threshold = int(dense_threshold)


class AdditiveSchwarz:
    def __init__(self, local_solver=None, reduction=ReductionStrategy.Auto):
        if local_solver is not None and not isinstance(local_solver, SchurComplement):
            raise TypeError("local_solver must be SchurComplement or None")
        self.local_solver = local_solver
        self.reduction = ReductionStrategy(reduction)


class MultiplicativeSchwarz:
    def __init__(self, local_solver=None):
        if local_solver is not None and not isinstance(local_solver, SchurComplement):
            raise TypeError("local_solver must be SchurComplement or None")
        self.local_solver = local_solver


class SolveResult:
    __slots__ = ("x", "demeaned", "converged", "iterations", "residual", "time_total", "time_setup",
                 "time_solve")

    def __init__(self, **kw):
        for k, v in kw.items():
            setattr(self, k, v)

    def __repr__(self):
        return (f"SolveResult(converged={self.converged}, iterations={self.iterations}, "
                f"residual={self.residual:.3e}, n_dofs={len(self.x)})")


class BatchSolveResult:
    __slots__ = ("x", "demeaned", "converged", "iterations", "residual", "time_solve", "time_total")

    def __init__(self, **kw):
        for k, v in kw.items():
            setattr(self, k, v)

    def __repr__(self):
        return f"BatchSolveResult(k={len(self.converged)}, converged={all(self.converged)})"


class AdditiveSchwarzDiagnostics:
    """Build-time scheduling metrics (schwarz/additive/planning.rs:140-212;
    within-py/src/lib.rs:197-258).  The device combine is a fixed-order
    per-DOF gather (no atomics), so the resolved strategy is always
    ``ParallelReduction``."""

    def __init__(self, work4, reduction):
        total, mx, scatter, n_dofs = (int(v) for v in work4)
        self.reduction_strategy = reduction
        self.resolved_reduction_strategy = ReductionStrategy.ParallelReduction
        self.total_inner_parallel_work = total
        self.max_inner_parallel_work = mx
        self.total_scatter_dofs = scatter
        self.outer_parallel_capacity = 0.0 if mx == 0 else total / mx
        self.scatter_overlap = scatter / max(n_dofs, 1)

    def __repr__(self):
        return ("AdditiveSchwarzDiagnostics("
                f"reduction_strategy={self.reduction_strategy.name}, "
                f"resolved_reduction_strategy={self.resolved_reduction_strategy.name}, "
                f"total_inner_parallel_work={self.total_inner_parallel_work}, "
                f"max_inner_parallel_work={self.max_inner_parallel_work}, "
                f"total_scatter_dofs={self.total_scatter_dofs}, "
                f"outer_parallel_capacity={self.outer_parallel_capacity:.3f}, "
                f"scatter_overlap={self.scatter_overlap:.3f})")


# ---------------------------------------------------------------------------
# config extraction (crates/within-py/src/lib.rs:496-671)
# ---------------------------------------------------------------------------
def _local_from(sc, default_split):
    """-> (ac_seed, ac_split, approx_schur, schur_seed, schur_split, dense_threshold)."""
    if sc is None:
        return (0, default_split, 1, 0, 1, 24)
    if not isinstance(sc, SchurComplement):
        raise TypeError("local_solver must be SchurComplement or None")
    ac = sc.approx_chol or ApproxCholConfig()
    if sc.approx_schur is None:
        return (ac.seed, ac.split, 0, 0, 1, sc.dense_threshold)
    return (ac.seed, ac.split, 1, sc.approx_schur.seed, sc.approx_schur.split, sc.dense_threshold)


def _precond_config(preconditioner):
    """-> (kind, local tuple, reduction); kind 0 none, 1 additive, 2 multiplicative."""
    if preconditioner is None:
        return 1, _local_from(None, 2), ReductionStrategy.Auto
    if isinstance(preconditioner, (Preconditioner, int)) and not isinstance(preconditioner, bool):
        p = Preconditioner(int(preconditioner))
        if p == Preconditioner.Off:
            return 0, _local_from(None, 2), ReductionStrategy.Auto
        if p == Preconditioner.Additive:
            return 1, _local_from(None, 2), ReductionStrategy.Auto
        return 2, _local_from(None, 2), ReductionStrategy.Auto
    if isinstance(preconditioner, AdditiveSchwarz):
        return 1, _local_from(preconditioner.local_solver, 2), preconditioner.reduction
    if isinstance(preconditioner, MultiplicativeSchwarz):
        return 2, _local_from(preconditioner.local_solver, 2), ReductionStrategy.Auto
    raise TypeError("preconditioner must be Preconditioner.Additive, Preconditioner.Multiplicative, "
                    "Preconditioner.Off, AdditiveSchwarz(...), MultiplicativeSchwarz(...), or None")


def _params(config, preconditioner):
    L = N.load()
    p = N.Params()
    L.wa_params_default(C.byref(p))
    if config is None:
        config = CG()
    if not isinstance(config, (CG, GMRES, LSMR)):
        raise TypeError("config must be CG, GMRES, or LSMR")
    kind, loc, reduction = _precond_config(preconditioner)
    if isinstance(config, (CG, LSMR)) and kind == 2:
        method = "CG" if isinstance(config, CG) else "LSMR"
        raise ValueError(f"{method} requires a symmetric preconditioner; use 'additive' or switch to GMRES")
    p.preconditioner = kind
    p.tol = config.tol
    p.maxiter = config.maxiter
    if isinstance(config, LSMR):  # lib.rs:621-631: implicit operator, no refinement
        p.krylov = 2
        p.local_size = config.local_size or 0
        p.operator_repr = 0
        p.max_refinements = 0
    else:
        p.krylov = 1 if isinstance(config, GMRES) else 0
        p.restart = config.restart if isinstance(config, GMRES) else 30
        p.operator_repr = int(config.operator)
        p.max_refinements = config.max_refinements
        p.precision = _PRECISIONS[config.precision]
        p.block = 0 if config.block is None else int(config.block)
    (p.local.ac_seed, p.local.ac_split, p.local.approx_schur, p.local.schur_seed, p.local.schur_split,
     p.local.dense_threshold) = loc
    return p, reduction


def _categories(categories):
    a = np.asarray(categories)
    if a.ndim != 2:
        raise ValueError("categories must be a 2-D array of shape (n_obs, n_factors)")
    if a.dtype != np.uint32:
        raise TypeError("categories must have dtype uint32")
    # within-py/src/lib.rs:732-744: warn unless consecutive observations of a
    # factor are adjacent (element stride 1 along axis 0; 0 for empty arrays)
    if a.strides[0] // a.itemsize not in (0, 1):
        warnings.warn("categories array is not F-contiguous (column-major). Use np.asfortranarray(categories) "
                      "for faster solves.", UserWarning, stacklevel=3)
    return np.asfortranarray(a)


def _vector(v, n, name):
    a = np.ascontiguousarray(np.asarray(v, dtype=np.float64))
    if a.ndim != 1 or a.shape[0] != n:
        raise ValueError(f"{name} must have shape ({n},), got {a.shape}")
    return a


class _Design:
    def __init__(self, categories, weights):
        L = N.load()
        cats = _categories(categories)
        self.n_obs, self.n_factors = cats.shape
        w = None if weights is None else _vector(weights, self.n_obs, "weights")
        h = C.c_void_p()
        N.check(L.wa_design_create(N.ptr(cats), self.n_obs, self.n_factors, max(self.n_obs, 1), N.ptr(w),
                                   C.byref(h)))
        self._destroy = L.wa_design_destroy
        self.h = h
        self.n_dofs = L.wa_design_n_dofs(h)

    def __del__(self):
        # the destroy function is held by the instance: at interpreter exit
        # the module's globals may already be gone
        if getattr(self, "h", None) and getattr(self, "_destroy", None):
            self._destroy(self.h)
            self.h = None


class FePreconditioner:
    """A built Schwarz preconditioner (device resident)
    (crates/within-py/src/lib.rs:820-913).

    Obtained from ``Solver.preconditioner()`` or rebuilt from its bytes with
    ``FePreconditioner(data)``; pickling stores the bytes (this library's own
    format, not the reference's postcard encoding).  ``apply`` runs the GPU
    additive / multiplicative Schwarz apply.  ``Solver(categories,
    preconditioner=p)`` uploads these factors for a design with the same DOFs
    instead of building new ones (solver.rs:135-155).
    """

    def __init__(self, data=None, *, _solver=None, _host_only=False):
        self._bytes = None
        self._own = None
        if _solver is not None:
            self._solver = _solver
            self._h = _solver._h
            return
        if data is None:
            raise TypeError("FePreconditioner(data) needs the bytes of a serialised preconditioner")
        b = bytes(data)
        h = C.c_void_p()
        rc = N.load().wa_precond_create(_byte_buffer(b), len(b), 1 if _host_only else 0, C.byref(h))
        if rc != 0:
            raise ValueError(f"failed to deserialize preconditioner: {N.last_error()}")
        self._destroy = N.load().wa_solver_destroy
        self._own = h
        self._h = h
        self._solver = None
        self._bytes = b

    def __del__(self):
        if getattr(self, "_own", None) and getattr(self, "_destroy", None):
            self._destroy(self._own)
            self._own = None

    def _info(self):
        out = np.zeros(16, dtype=np.int64)
        N.check(N.load().wa_solver_info(self._h, N.ptr(out)))
        return out

    def _kind(self):
        p = N.Params()
        N.check(N.load().wa_solver_params(self._h, C.byref(p)))
        return int(p.preconditioner)

    def to_bytes(self):
        """The serialised preconditioner (kind, local config, factors)."""
        if self._bytes is None:
            self._bytes = _precond_bytes(self._h)
        return self._bytes

    def apply(self, x):
        n = self.nrows
        x = _vector(x, n, "x")
        y = np.zeros(n)
        N.check(N.load().wa_precond_apply(self._h, N.ptr(x), N.ptr(y), 1, n))
        return y

    @property
    def nrows(self):
        return int(self._info()[9])

    @property
    def ncols(self):
        return self.nrows

    @property
    def n_subdomains(self):
        return int(self._info()[0])

    def _work(self):
        w = np.zeros(max(self.n_subdomains, 1), dtype=np.int64)
        out = np.zeros(4, dtype=np.int64)
        N.check(N.load().wa_precond_work(self._h, N.ptr(w), N.ptr(out)))
        return w[: self.n_subdomains], out

    @property
    def subdomain_inner_parallel_work(self):
        """Estimated nested-parallel work per subdomain (preconditioner.rs:64-79)."""
        return [int(v) for v in self._work()[0]]

    def additive_schwarz_diagnostics(self):
        if self._kind() != 1:
            return None
        red = self._solver._reduction if self._solver is not None else ReductionStrategy.Auto
        return AdditiveSchwarzDiagnostics(self._work()[1], red)

    def __reduce__(self):
        return (FePreconditioner, (self.to_bytes(),))

    def __repr__(self):
        variant = "Additive" if self._kind() == 1 else "Multiplicative"
        return f"FePreconditioner({variant}, n={self.nrows})"


def _byte_buffer(b):
    return (C.c_uint8 * max(len(b), 1)).from_buffer_copy(b if b else b"\0")


def _precond_bytes(h):
    L = N.load()
    n = C.c_int64(0)
    N.check(L.wa_solver_precond_bytes(h, None, C.byref(n)))
    buf = (C.c_uint8 * max(n.value, 1))()
    N.check(L.wa_solver_precond_bytes(h, buf, C.byref(n)))
    return bytes(buf)[: n.value]


class Solver:
    """Persistent solver: preconditioner built and uploaded once
    (crates/within/src/solver.rs:82-447)."""

    def __init__(self, categories, config=None, weights=None, preconditioner=None, *, _host_only=False,
                 _dense_core=0):
        t0 = time.perf_counter()
        L = N.load()
        given = preconditioner if isinstance(preconditioner, FePreconditioner) else None
        self._precond_arg = None if given is not None else preconditioner
        self._design = _Design(categories, weights)
        p, self._reduction = _params(config, self._precond_arg)
        p.host_only = 1 if _host_only else 0
        p.local.dense_core = int(_dense_core)
        h = C.c_void_p()
        if given is None:
            N.check(L.wa_solver_create(self._design.h, C.byref(p), C.byref(h)))
        else:
            if given._solver is not None:
                self._reduction = given._solver._reduction
            b = given.to_bytes()
            N.check(L.wa_solver_create_with_precond(self._design.h, C.byref(p), _byte_buffer(b), len(b),
                                                    C.byref(h)))
            N.check(L.wa_solver_params(h, C.byref(p)))
        self._destroy = L.wa_solver_destroy
        self._h = h
        self._kind = int(p.preconditioner)
        self._params = p
        self.time_setup = time.perf_counter() - t0

    def __del__(self):
        if getattr(self, "_h", None) and getattr(self, "_destroy", None):
            self._destroy(self._h)
            self._h = None

    def set_config(self, config):
        """Switch the Krylov settings (CG/GMRES, tol, maxiter, refinements,
        precision, block) of this built solver; the preconditioner stays."""
        p, _ = _params(config, self._precond_arg)
        p.host_only = self._params.host_only
        p.preconditioner = self._params.preconditioner
        p.local = self._params.local
        N.check(N.load().wa_solver_set_params(self._h, C.byref(p)))
        self._params = p

    @property
    def n_dofs(self):
        return self._design.n_dofs

    @property
    def n_obs(self):
        return self._design.n_obs

    def info(self):
        out = np.zeros(16, dtype=np.int64)
        N.check(N.load().wa_solver_info(self._h, N.ptr(out)))
        return out

    def preconditioner(self):
        if self._kind == 0:
            return None
        return FePreconditioner(_solver=self)

    def autotune(self, k=32, reps=5):
        """Time candidate schedule shapes of the operator (explicit Gramian and
        implicit factor passes) for k right-hand sides and keep the fastest;
        returns the choices and the timings (microseconds per launch)."""
        out = np.zeros(8, dtype=np.int64)
        N.check(N.load().wa_solver_autotune(self._h, int(k), int(reps), N.ptr(out)))
        keys = ("gram_chunk", "gram_short", "gram_us_default", "gram_us_best", "pass_chunk", "pass_short",
                "pass_us_default", "pass_us_best")
        return dict(zip(keys, (int(v) for v in out)))

    def train_coarse(self, steps=200, sample_iters=8, seed=0, lr=3e-3, force=None, ridge=1e-4):
        """Train the learned coarse-space correction on this solver's own CG
        iterations and keep it for later solves when it cuts iterations.

        A three-layer graph network over the factor-level graph outputs 32
        coarse vectors; with it, CG runs in its flexible form with the
        balancing two-level preconditioner.  It is kept only when a validation
        block converges in fewer iterations and less time; ``force=True`` /
        ``False`` overrides that decision.  Returns the training report
        (losses, mean validation iterations and solve times with and without
        the correction, whether it is enabled).  No counterpart in the
        reference."""
        p = N.CoarseParams()
        L = N.load()
        L.wa_coarse_params_default(C.byref(p))
        p.steps, p.sample_iters, p.seed, p.lr = int(steps), int(sample_iters), int(seed), float(lr)
        p.ridge = float(ridge)
        p.force = 0 if force is None else (1 if force else -1)
        r = N.CoarseReport()
        N.check(L.wa_coarse_train(self._h, C.byref(p), C.byref(r)))
        return {k: getattr(r, k) for k, _ in N.CoarseReport._fields_ if k != "reserved"}

    def enable_coarse(self, on=True):
        """Switch a trained coarse correction on or off for later solves."""
        N.check(N.load().wa_coarse_enable(self._h, 1 if on else 0))

    def _solve_cols(self, Y):
        n, m = self.n_obs, self.n_dofs
        Y = np.asfortranarray(np.asarray(Y, dtype=np.float64))
        if Y.ndim != 2 or Y.shape[0] != n:
            raise ValueError(f"Y must have shape ({n}, k), got {Y.shape}")
        k = Y.shape[1]
        X = np.zeros((m, k), order="F")
        D = np.zeros((n, k), order="F")
        stats = (N.RhsStats * max(k, 1))()
        t0 = time.perf_counter()
        N.check(N.load().wa_solve_batch(self._h, N.ptr(Y), k, n, N.ptr(X), m, N.ptr(D), n, stats))
        return X, D, stats, time.perf_counter() - t0

    def solve(self, y):
        t0 = time.perf_counter()
        y = _vector(y, self.n_obs, "y")
        X, D, st, _ = self._solve_cols(y[:, None])
        return SolveResult(x=X[:, 0], demeaned=D[:, 0], converged=bool(st[0].converged),
                           iterations=int(st[0].iterations), residual=float(st[0].residual),
                           time_total=time.perf_counter() - t0, time_setup=0.0, time_solve=float(st[0].time_solve))

    def solve_batch(self, Y):
        t0 = time.perf_counter()
        X, D, st, _ = self._solve_cols(Y)
        k = X.shape[1]
        return BatchSolveResult(x=X, demeaned=D, converged=[bool(st[i].converged) for i in range(k)],
                                iterations=[int(st[i].iterations) for i in range(k)],
                                residual=[float(st[i].residual) for i in range(k)],
                                time_solve=[float(st[i].time_solve) for i in range(k)],
                                time_total=time.perf_counter() - t0)


def solve(categories, y, config=None, weights=None, preconditioner=None):
    """One-shot solve (crates/within-py/src/lib.rs:750-775)."""
    t0 = time.perf_counter()
    s = Solver(categories, config=config, weights=weights, preconditioner=preconditioner)
    r = s.solve(y)
    r.time_setup = s.time_setup
    r.time_total = time.perf_counter() - t0
    return r


def solve_batch(categories, Y, config=None, weights=None, preconditioner=None):
    """One-shot batch solve (crates/within-py/src/lib.rs:781-815)."""
    t0 = time.perf_counter()
    s = Solver(categories, config=config, weights=weights, preconditioner=preconditioner)
    r = s.solve_batch(Y)
    r.time_total = time.perf_counter() - t0
    return r


class ProblemBatch:
    """Many independent fixed-effects problems solved together on the GPU.

    Each problem is what ``Solver(categories_i, config, weights_i,
    preconditioner)`` would solve (crates/within/src/solver.rs:82-447); the
    problems with the same number of factors are stacked block-diagonally and
    share every operator and preconditioner launch, while each right-hand side
    of each problem keeps its own CG recurrence, tolerance and convergence
    flag.  CG only, with additive Schwarz or no preconditioner.  The
    reference has no such entry point: there each problem is its own Solver.
    """

    def __init__(self, problems, config=None, weights=None, preconditioner=None):
        t0 = time.perf_counter()
        L = N.load()
        if config is None:
            config = CG()
        if not isinstance(config, CG):
            raise TypeError("a ProblemBatch is solved with CG")
        problems = list(problems)
        if not problems:
            raise ValueError("a ProblemBatch needs at least one problem")
        if weights is None:
            weights = [None] * len(problems)
        if len(weights) != len(problems):
            raise ValueError("weights must have one entry per problem")
        self._designs = [_Design(c, w) for c, w in zip(problems, weights)]
        p, _ = _params(config, preconditioner)
        arr = (C.c_void_p * len(self._designs))(*[d.h.value for d in self._designs])
        h = C.c_void_p()
        N.check(L.wa_multi_create(arr, len(self._designs), C.byref(p), C.byref(h)))
        self._destroy = L.wa_multi_destroy
        self._h = h
        self.time_setup = time.perf_counter() - t0

    def __del__(self):
        if getattr(self, "_h", None) and getattr(self, "_destroy", None):
            self._destroy(self._h)
            self._h = None

    def __len__(self):
        return len(self._designs)

    def info(self, group=0):
        out = np.zeros(16, dtype=np.int64)
        N.check(N.load().wa_multi_info(self._h, int(group), N.ptr(out)))
        return out

    def solve(self, Ys):
        """``Ys[i]``: (n_obs_i,) or (n_obs_i, k_i); returns one
        BatchSolveResult per problem."""
        t0 = time.perf_counter()
        P = len(self._designs)
        Ys = list(Ys)
        if len(Ys) != P:
            raise ValueError(f"expected {P} right-hand-side arrays, got {len(Ys)}")
        ys, xs, ds, sts = [], [], [], []
        for d, Y in zip(self._designs, Ys):
            Y = np.asfortranarray(np.asarray(Y, dtype=np.float64))
            if Y.ndim == 1:
                Y = Y[:, None]
            if Y.ndim != 2 or Y.shape[0] != d.n_obs:
                raise ValueError(f"Y must have shape ({d.n_obs}, k), got {Y.shape}")
            k = Y.shape[1]
            ys.append(Y)
            xs.append(np.zeros((d.n_dofs, k), order="F"))
            ds.append(np.zeros((d.n_obs, k), order="F"))
            sts.append((N.RhsStats * max(k, 1))())
        PP = C.c_void_p * P
        I64 = C.c_int64 * P
        N.check(N.load().wa_multi_solve(
            self._h, PP(*[y.ctypes.data for y in ys]), (C.c_int32 * P)(*[y.shape[1] for y in ys]),
            I64(*[d.n_obs for d in self._designs]), PP(*[x.ctypes.data for x in xs]),
            I64(*[d.n_dofs for d in self._designs]), PP(*[x.ctypes.data for x in ds]),
            I64(*[d.n_obs for d in self._designs]), PP(*[C.addressof(st) for st in sts])))
        total = time.perf_counter() - t0
        out = []
        for y, x, dm, st in zip(ys, xs, ds, sts):
            k = y.shape[1]
            out.append(BatchSolveResult(x=x, demeaned=dm, converged=[bool(st[i].converged) for i in range(k)],
                                        iterations=[int(st[i].iterations) for i in range(k)],
                                        residual=[float(st[i].residual) for i in range(k)],
                                        time_solve=[float(st[i].time_solve) for i in range(k)],
                                        time_total=total))
        return out


def solve_many(problems, Ys, config=None, weights=None, preconditioner=None):
    """One-shot solve of many independent problems (see ``ProblemBatch``)."""
    return ProblemBatch(problems, config=config, weights=weights, preconditioner=preconditioner).solve(Ys)
