"""GPU runs of the two workload configurations the other test files do not
cover:

* the 5M-observation AKM panel (worker 500k Zipf 1.1, firm 50k with mobility
  0.05, year 20; y = worker and firm effects + N(0,1)): the mixed-precision
  block solve (FP32 inner Krylov, FP64 refinement) against the FP64 solve,
  RHS by RHS within 1e-8 relative, with converged flags;
* many uneven independent problems (n log-uniform over 10k..300k, 2 or 3
  uniform factors with n/20, n/200 and 25 levels, 1-4 RHS each), one solver
  per problem as a caller with many regressions would: the smaller ones
  against the oracle (demeaned within 1e-8), every one through properties
  that fix the projection whatever its size (converged flag, the reported
  residual recomputed as ||D'W(y - Dx)|| / ||D'W y|| of the returned vector,
  and demeaned = y - D x with the returned x).
"""
import numpy as np
import pytest

import within_amd as W
from within_amd import _native as N
from within_amd import synth

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def test_five_million_panel_mixed_against_fp64():
    p = synth.akm_panel(5_000_000, 500_000, 50_000, 20, zipf_s=1.1, mobility=0.05, seed=2)
    codes = p["codes"]
    Y = synth.akm_response(p, k=4, beta_cols=3)
    prec = W.AdditiveSchwarz(W.SchurComplement(W.ApproxCholConfig(0, 1), W.ApproxSchurConfig(0, 1)))
    s = W.Solver(codes, config=W.CG(precision="mixed", tol=1e-9, operator=W.OperatorRepr.Explicit),
                 preconditioner=prec)
    mixed = s.solve_batch(Y)
    s.set_config(W.CG(tol=1e-11, max_refinements=4, operator=W.OperatorRepr.Explicit))
    ref = s.solve_batch(Y)
    assert all(mixed.converged) and all(ref.converged)
    for j in range(Y.shape[1]):
        assert rel(mixed.demeaned[:, j], ref.demeaned[:, j]) <= 1e-8, j
        assert mixed.residual[j] <= 1.01e-9


PROBLEMS = synth.many_problems(n_problems=10, n_min=10_000, n_max=300_000, seed=4)


@pytest.mark.parametrize("i", range(len(PROBLEMS)))
def test_many_uneven_problems(i, oracle_mod):
    codes, Y = PROBLEMS[i]
    n = codes.shape[0]
    r = W.solve_batch(codes, Y, config=W.CG(tol=1e-10, max_refinements=3))
    assert all(r.converged)
    L = N.load()
    s = W.Solver(codes, preconditioner=W.Preconditioner.Off)
    m = s.n_dofs
    k = Y.shape[1]
    # demeaned = y - D x
    D = np.zeros((n, k), order="F")
    N.check(L.wa_matvec_d(s._h, N.ptr(np.asfortranarray(r.x)), m, N.ptr(D), n, k))
    assert rel(Y - D, r.demeaned) <= 1e-12
    # the reported residual is that of the returned vector
    for j in range(k):
        res = np.zeros(m)
        b = np.zeros(m)
        N.check(L.wa_rmatvec_wdt(s._h, N.ptr(np.ascontiguousarray(r.demeaned[:, j])), n, N.ptr(res), m, 1))
        N.check(L.wa_rmatvec_wdt(s._h, N.ptr(np.ascontiguousarray(Y[:, j])), n, N.ptr(b), m, 1))
        want = np.linalg.norm(res) / max(np.linalg.norm(b), 1e-15)
        assert want <= 1.01e-10
        assert abs(r.residual[j] - want) <= 1e-12 + 1e-3 * want
    if n <= 60_000:
        ro = oracle_mod.OracleSolver(codes, None, ac_split=2, tol=1e-10, max_refinements=3).solve_batch(Y)
        assert all(ro["converged"])
        for j in range(k):
            assert rel(r.demeaned[:, j], ro["demeaned"][:, j]) <= 1e-8, j
