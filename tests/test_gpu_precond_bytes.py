"""GPU: preconditioners rebuilt from their serialised bytes
(crates/within-py/src/lib.rs:892-913 pickling, 925-982 Solver(...,
preconditioner=p); crates/within/src/solver.rs:135-155).

Bars: a standalone preconditioner loaded from bytes (user DOF numbering on the
device) applies within 1e-12 relative of the solver it came from (internal
numbering; same FP64 factors, different summation order) and within 1e-10 of
the oracle; solvers built from the bytes reach the oracle's demeaned
residuals within 1e-8 relative per RHS, with converged flags.
"""
import pickle

import numpy as np
import pytest

import within_amd as W
from within_amd import _native as N
from within_amd import synth

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def akm3():
    p = synth.akm_panel(20000, 2000, 300, 12, seed=3)
    return p["codes"], None


def akm4w():
    p = synth.akm_panel(15000, 1500, 200, 10, n_industries=5, seed=4, weights="lognormal")
    return p["codes"], p["weights"]


def apply_cols(h, R):
    m, k = R.shape
    Z = np.zeros((m, k), order="F")
    N.check(N.load().wa_precond_apply(h, N.ptr(R), N.ptr(Z), k, m))
    return Z


@pytest.mark.parametrize("make", [akm3, akm4w], ids=["akm-3fe", "akm-4fe-weighted"])
@pytest.mark.parametrize("kind", ["additive", "multiplicative"])
def test_unpickled_preconditioner_applies_like_the_original(make, kind, oracle_mod):
    codes, w = make()
    if kind == "additive":
        s = W.Solver(codes, weights=w)
        o = oracle_mod.OracleSolver(codes, w, ac_split=2)
    else:
        s = W.Solver(codes, weights=w, config=W.GMRES(), preconditioner=W.MultiplicativeSchwarz())
        o = oracle_mod.OracleSolver(codes, w, krylov=1, precond=2, ac_split=2)
    p = s.preconditioner()
    q = pickle.loads(pickle.dumps(p))
    assert isinstance(q, W.FePreconditioner) and q.to_bytes() == p.to_bytes()
    assert (q.nrows, q.n_subdomains) == (p.nrows, p.n_subdomains)
    m = s.n_dofs
    R = np.asfortranarray(np.random.default_rng(5).standard_normal((m, 4)))
    Zs, Zq = apply_cols(s._h, R), apply_cols(q._h, R)
    assert rel(Zq, Zs) <= 1e-12, rel(Zq, Zs)
    ref = np.column_stack([o.precond_apply(R[:, j]) for j in range(R.shape[1])])
    assert rel(Zq, ref) <= 1e-10, rel(Zq, ref)
    # the single-vector method agrees with the batched call
    assert rel(q.apply(R[:, 0]), Zq[:, 0]) <= 1e-14


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_solver_from_bytes_matches_oracle(precision, oracle_mod):
    codes, w = akm4w()
    n = codes.shape[0]
    Y = np.asfortranarray(np.random.default_rng(9).standard_normal((n, 4)))
    o = oracle_mod.OracleSolver(codes, w, ac_split=2, tol=1e-10, max_refinements=3)
    ro = o.solve_batch(Y)
    cfg = W.CG(tol=1e-10, max_refinements=3, precision=precision)
    b = W.Solver(codes, weights=w, config=cfg).preconditioner().to_bytes()
    s = W.Solver(codes, weights=w, config=cfg, preconditioner=W.FePreconditioner(b))
    r = s.solve_batch(Y)
    assert all(r.converged)
    for j in range(Y.shape[1]):
        assert rel(r.demeaned[:, j], ro["demeaned"][:, j]) <= 1e-8, j
    # one-shot solve with a preconditioner object
    r1 = W.solve(codes, Y[:, 1], config=cfg, weights=w, preconditioner=W.FePreconditioner(b))
    assert r1.converged and rel(r1.demeaned, ro["demeaned"][:, 1]) <= 1e-8


def test_multiplicative_solver_from_bytes_matches_oracle(oracle_mod):
    codes, w = akm3()
    n = codes.shape[0]
    Y = np.asfortranarray(np.random.default_rng(2).standard_normal((n, 3)))
    o = oracle_mod.OracleSolver(codes, w, ac_split=2, tol=1e-10, max_refinements=3)
    ro = o.solve_batch(Y)
    cfg = W.GMRES(tol=1e-10, max_refinements=3)
    p = W.Solver(codes, weights=w, config=cfg, preconditioner=W.MultiplicativeSchwarz()).preconditioner()
    s = W.Solver(codes, weights=w, config=cfg, preconditioner=pickle.loads(pickle.dumps(p)))
    r = s.solve_batch(Y)
    assert all(r.converged)
    for j in range(Y.shape[1]):
        assert rel(r.demeaned[:, j], ro["demeaned"][:, j]) <= 1e-8, j
