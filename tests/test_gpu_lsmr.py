"""GPU LSMR (solver.rs:162-207; schwarz-precond/src/solve/lsmr.rs) against the
oracle's LSMR restatement on the same seeded inputs.

Bars: demeaned residuals within 1e-8 relative of the oracle's LSMR per RHS
(FP64 both; summation orders differ), matching converged flags, iteration
counts within a few steps (the stopping tests compare rounded scalars), the
reported residual equal to ||D'W(y - Dx)|| / ||D'W y|| to FP64 rounding.
"""
import numpy as np
import pytest

import within_amd as W
from within_amd import _native as N
from within_amd import synth

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def problems():
    out = []
    c, _, _ = synth.uniform_problem(3000, [60, 25], seed=1)
    out.append((c, None, "2fe-uniform"))
    c, w, _ = synth.uniform_problem(4000, [80, 30, 7], seed=2, weights=True)
    out.append((c, w, "3fe-weighted"))
    p = synth.akm_panel(20000, 2000, 300, 12, seed=3)
    out.append((p["codes"], None, "akm-3fe"))
    p = synth.akm_panel(15000, 1500, 200, 10, n_industries=5, seed=4, weights="lognormal")
    out.append((p["codes"], p["weights"], "akm-4fe-weighted"))
    return out


PROBS = problems()


def check_against(r, ro, Y, o, its_slack):
    k = Y.shape[1]
    assert list(r.converged) == [bool(c) for c in ro["converged"]]
    for j in range(k):
        assert rel(r.demeaned[:, j], ro["demeaned"][:, j]) <= 1e-8, (j, rel(r.demeaned[:, j], ro["demeaned"][:, j]))
        assert abs(r.iterations[j] - ro["iterations"][j]) <= its_slack, (r.iterations[j], ro["iterations"][j])
        res = np.linalg.norm(o.rmatvec_wdt(r.demeaned[:, j])) / max(np.linalg.norm(o.rmatvec_wdt(Y[:, j])), 1e-15)
        assert abs(r.residual[j] - res) <= 1e-12 + 1e-6 * res


@pytest.mark.parametrize("idx", range(len(PROBS)), ids=[p[2] for p in PROBS])
def test_preconditioned_lsmr_matches_oracle(idx, oracle_mod):
    codes, w, _ = PROBS[idx]
    n = codes.shape[0]
    Y = np.asfortranarray(np.random.default_rng(20 + idx).standard_normal((n, 4)))
    Y[:, 0] += np.cos(codes[:, 0].astype(float))
    o = oracle_mod.OracleSolver(codes, w, krylov=2, precond=1, ac_split=2, tol=1e-10, maxiter=1000)
    ro = o.solve_batch(Y)
    r = W.Solver(codes, weights=w, config=W.LSMR(tol=1e-10)).solve_batch(Y)
    assert all(r.converged)
    check_against(r, ro, Y, o, its_slack=3)


def test_unpreconditioned_lsmr_matches_oracle(oracle_mod):
    codes, w, _ = PROBS[1]
    n = codes.shape[0]
    Y = np.asfortranarray(np.random.default_rng(5).standard_normal((n, 3)))
    o = oracle_mod.OracleSolver(codes, w, krylov=2, precond=0, tol=1e-9, maxiter=3000)
    ro = o.solve_batch(Y)
    r = W.Solver(codes, weights=w, config=W.LSMR(tol=1e-9, maxiter=3000),
                 preconditioner=W.Preconditioner.Off).solve_batch(Y)
    assert all(r.converged)
    # M = I on the device vs plain Golub-Kahan in the oracle: the same
    # recurrences in exact arithmetic (lsmr/tests.rs:270-309)
    check_against(r, ro, Y, o, its_slack=max(3, int(0.02 * max(ro["iterations"]))))


@pytest.mark.parametrize("window", [1, 5])
def test_windowed_lsmr_matches_oracle(window, oracle_mod):
    codes, w, _ = PROBS[3]
    n = codes.shape[0]
    Y = np.asfortranarray(np.random.default_rng(8).standard_normal((n, 3)))
    o = oracle_mod.OracleSolver(codes, w, krylov=2, precond=1, ac_split=2, tol=1e-10, local_size=window)
    ro = o.solve_batch(Y)
    r = W.Solver(codes, weights=w, config=W.LSMR(tol=1e-10, local_size=window)).solve_batch(Y)
    assert all(r.converged)
    check_against(r, ro, Y, o, its_slack=3)


def test_lsmr_agrees_with_cg_on_the_projection(oracle_mod):
    """The demeaned residual is unique, so LSMR and CG reach the same one."""
    codes, w, _ = PROBS[2]
    n = codes.shape[0]
    Y = np.asfortranarray(np.random.default_rng(2).standard_normal((n, 5)))
    a = W.Solver(codes, weights=w, config=W.LSMR(tol=1e-11)).solve_batch(Y)
    b = W.Solver(codes, weights=w, config=W.CG(tol=1e-11, max_refinements=3)).solve_batch(Y)
    for j in range(5):
        assert rel(a.demeaned[:, j], b.demeaned[:, j]) <= 1e-8


def test_edge_columns_and_errors():
    codes, w, _ = PROBS[0]
    n = codes.shape[0]
    Y = np.asfortranarray(np.random.default_rng(1).standard_normal((n, 3)))
    Y[:, 1] = 0.0
    s = W.Solver(codes, config=W.LSMR(tol=1e-10))
    r = s.solve_batch(Y)
    assert r.converged[1] and r.iterations[1] == 0 and not np.any(r.demeaned[:, 1])
    assert r.converged[0] and r.converged[2]
    # one RHS alone equals its column of the batch
    r1 = s.solve(Y[:, 2])
    assert rel(r1.demeaned, r.demeaned[:, 2]) <= 1e-12 and r1.iterations == r.iterations[2]
    # maxiter exhaustion
    s.set_config(W.LSMR(tol=1e-14, maxiter=2))
    r2 = s.solve(Y[:, 0])
    assert not r2.converged and r2.iterations == 2
    # non-finite right-hand side (lsmr.rs:267-274)
    bad = Y[:, 0].copy()
    bad[7] = np.nan
    with pytest.raises(N.NativeError):
        s.solve(bad)
    # a negative or NaN tolerance is an input error for LSMR (lsmr.rs:262-266)
    for bad_tol in (-1.0, float("nan")):
        s.set_config(W.LSMR(tol=bad_tol))
        with pytest.raises(N.NativeError):
            s.solve(Y[:, 0])
    # the solver keeps working after the errors
    s.set_config(W.LSMR(tol=1e-10))
    assert s.solve(Y[:, 0]).converged
