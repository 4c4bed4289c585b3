"""GPU: the Python surface behaves as the reference's own Python tests expect
(tests/test_solve.py, test_edge_cases.py, test_properties.py,
test_preconditioner.py of py-econometrics/within), run through the gfx950
path.  Numerical bars: the demeaned residual is the unique least-squares
residual, so two routes to it agree to 1e-8 relative; properties the
reference states with their own tolerances keep those tolerances."""
import pickle

import numpy as np
import pytest

import within_amd as W

pytestmark = pytest.mark.gpu


def cats(*cols):
    return np.asfortranarray(np.column_stack(cols).astype(np.uint32))


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


@pytest.fixture(scope="module")
def problem():  # tests/test_solve.py fixture: 2 x 50 levels, 10k observations
    rng = np.random.RandomState(42)
    n = 10000
    fe1, fe2 = rng.randint(0, 50, n), rng.randint(0, 50, n)
    y = rng.randn(n) + 0.5 * fe1 - 0.3 * fe2
    return cats(fe1, fe2), y


def dense_design(c):
    n = c.shape[0]
    cols = []
    for q in range(c.shape[1]):
        L = int(c[:, q].max()) + 1
        D = np.zeros((n, L))
        D[np.arange(n), c[:, q]] = 1.0
        cols.append(D)
    return np.hstack(cols)


# ---- test_solve.py ---------------------------------------------------------
def test_configs_agree_on_the_projection(problem):
    c, y = problem
    assert W.solve(c, y).converged  # default configuration
    t = 1e-11
    ref = W.solve(c, y, W.CG(tol=t))
    runs = [W.solve(c, y, W.CG(tol=t, operator=W.OperatorRepr.Explicit)),
            W.solve(c, y, W.GMRES(tol=t), preconditioner=W.Preconditioner.Multiplicative),
            W.solve(c, y, W.CG(tol=t), preconditioner=W.Preconditioner.Off),
            W.solve(c, y, W.GMRES(tol=t), preconditioner=W.Preconditioner.Off),
            W.solve(c, y, W.CG(tol=t), preconditioner=W.AdditiveSchwarz(W.SchurComplement(W.ApproxCholConfig(3, 2)))),
            W.solve(c, y, W.GMRES(tol=t), preconditioner=W.MultiplicativeSchwarz(W.SchurComplement())),
            W.solve(c, y, W.LSMR(tol=t))]
    for r in runs:
        assert r.converged and r.x.shape == (100,) and r.demeaned.shape == (10000,)
        assert rel(r.demeaned, ref.demeaned) <= 1e-8


def test_recovers_fixed_effects_and_orthogonality():  # test_solve.py:150-188
    rng = np.random.default_rng(0)
    n = 5000
    f1, f2 = rng.integers(0, 40, n), rng.integers(0, 30, n)
    a, b = rng.standard_normal(40), rng.standard_normal(30)
    y = a[f1] + b[f2]
    r = W.solve(cats(f1, f2), y, W.CG(tol=1e-12))
    # x recovers (a, b) up to one constant shifted between the factors
    shift = r.x[0] - a[0]
    assert np.abs(r.x[:40] - (a + shift)).max() < 1e-6
    assert np.abs(r.x[40:] - (b - shift)).max() < 1e-6
    e = rng.standard_normal(n)
    r2 = W.solve(cats(f1, f2), y + e)
    D = dense_design(cats(f1, f2))
    assert np.abs(D.T @ r2.demeaned).max() < 1e-4


def test_result_fields_and_solver_reuse(problem):
    c, y = problem
    s = W.Solver(c)
    assert (s.n_dofs, s.n_obs) == (100, 10000)
    r1, r2 = s.solve(y), s.solve(2 * y)
    assert r1.converged and r1.iterations > 0 and r1.residual >= 0
    assert r1.time_total >= 0 and r1.time_solve >= 0
    assert rel(r2.demeaned, 2 * r1.demeaned) <= 1e-8
    one = W.solve(c, y)
    assert rel(r1.demeaned, one.demeaned) <= 1e-10
    Y = np.column_stack([y, 2 * y, y])
    b = s.solve_batch(Y)
    assert b.x.shape == (100, 3) and b.demeaned.shape == (10000, 3)
    assert len(b.converged) == len(b.iterations) == len(b.residual) == len(b.time_solve) == 3
    assert np.array_equal(b.demeaned[:, 0], b.demeaned[:, 2])  # identical columns
    assert rel(b.demeaned[:, 0], r1.demeaned) <= 1e-10
    bb = W.solve_batch(c, Y[:, :1])
    assert bb.x.shape == (100, 1) and rel(bb.demeaned[:, 0], r1.demeaned) <= 1e-10
    assert W.Solver(c, preconditioner=W.Preconditioner.Off).preconditioner() is None


def test_weighted_and_unit_weights(problem):
    c, y = problem
    w = np.random.default_rng(1).uniform(0.5, 2.0, len(y))
    r = W.solve(c, y, weights=w)
    assert r.converged
    D = dense_design(c)
    assert np.abs(D.T @ (w * r.demeaned)).max() < 1e-4
    a, b = W.solve(c, y, weights=np.ones(len(y))), W.solve(c, y)
    assert rel(a.demeaned, b.demeaned) <= 1e-10


def test_preconditioner_object_api(problem):  # test_preconditioner.py
    c, y = problem
    s = W.Solver(c)
    p = s.preconditioner()
    assert (p.nrows, p.ncols) == (s.n_dofs, s.n_dofs)
    assert repr(p).startswith("FePreconditioner(Additive")
    x = np.random.default_rng(2).standard_normal(s.n_dofs)
    y1 = p.apply(x)
    assert y1.shape == (s.n_dofs,) and np.all(np.isfinite(y1))
    assert np.array_equal(p.apply(x), y1)  # deterministic
    with pytest.raises(ValueError):
        p.apply(np.ones(s.n_dofs + 1))
    q = W.FePreconditioner(p.to_bytes())
    assert np.array_equal(q.apply(x), W.FePreconditioner(p.to_bytes()).apply(x))
    assert rel(q.apply(x), y1) <= 1e-12
    assert rel(pickle.loads(pickle.dumps(p)).apply(x), y1) <= 1e-12
    with pytest.raises(ValueError):
        W.FePreconditioner(b"not a preconditioner")
    m = W.Solver(c, W.GMRES(), preconditioner=W.Preconditioner.Multiplicative).preconditioner()
    assert repr(m).startswith("FePreconditioner(Multiplicative")
    r = W.Solver(c, preconditioner=q).solve(y)
    assert r.converged and rel(r.demeaned, s.solve(y).demeaned) <= 1e-8


# ---- test_edge_cases.py -----------------------------------------------------
def test_nan_and_inf_in_y_propagate():
    c = cats(np.array([0, 1, 0]), np.array([0, 0, 1]))
    r = W.solve(c, np.array([1.0, np.nan, 3.0]))
    assert np.any(np.isnan(r.x)) and not r.converged
    r = W.solve(c, np.array([1.0, np.inf, 3.0]))
    assert np.any(~np.isfinite(r.x))
    for cfg in (W.CG(precision="mixed"), W.GMRES()):
        r = W.solve(c, np.array([1.0, np.nan, 3.0]), cfg)
        assert np.any(np.isnan(r.x)) or not r.converged


def test_nan_residual_spends_the_iteration_budget():
    # cg.rs:75-125 keeps iterating on a NaN residual until maxiter, and every
    # refinement correction adds its own budget (solver.rs:296-347)
    c = cats(np.array([0, 1, 0, 2]), np.array([0, 0, 1, 1]))
    r = W.solve(c, np.array([1.0, np.nan, 3.0, 2.0]), W.CG(maxiter=40, max_refinements=2))
    assert not r.converged and r.iterations == 3 * 40
    r = W.solve(c, np.array([1.0, np.nan, 3.0, 2.0]), W.CG(maxiter=40, max_refinements=2, precision="mixed"))
    assert not r.converged and r.iterations == 3 * 40


def test_weight_edge_cases_do_not_crash():
    c = cats(np.array([0, 1, 0, 1, 2]), np.array([0, 0, 1, 1, 0]))
    y = np.ones(5)
    for w in (np.zeros(5), np.array([1.0, -1.0, 1.0, 1.0, 1.0]), np.array([1.0, np.nan, 1.0, 1.0, 1.0])):
        try:
            r = W.solve(c, y, weights=w)
            assert not r.converged or np.any(~np.isfinite(r.x))
        except ValueError:
            pass  # raising is acceptable (test_edge_cases.py:39-105)
    for w in (np.full(5, 1e-300), np.full(5, 1e300), np.array([1.0, np.inf, 1.0, 1.0, 1.0])):
        try:
            r = W.solve(c, y, weights=w)
            assert isinstance(r.converged, bool)
        except ValueError:
            pass


def test_degenerate_inputs():
    c = cats(np.array([0, 1, 0, 1, 2]), np.array([0, 0, 1, 1, 0]))
    r = W.solve(c, np.zeros(5))
    assert r.converged and np.abs(r.x).max() <= 1e-10
    rng = np.random.RandomState(42)
    c2 = cats(rng.randint(0, 10, 100), rng.randint(0, 10, 100))
    r = W.solve(c2, np.full(100, 5.0))
    assert r.converged and np.abs(r.demeaned).max() <= 1e-4
    r = W.solve(np.zeros((5, 2), dtype=np.uint32, order="F"), np.arange(1.0, 6.0),
                preconditioner=W.Preconditioner.Off)
    assert np.all(np.isfinite(r.x))
    for cc, yy in ((np.array([[0, 0]], dtype=np.uint32, order="F"), np.array([5.0])),
                   (np.array([[0, 0], [1, 1]], dtype=np.uint32, order="F"), np.array([1.0, 2.0]))):
        r = W.solve(cc, yy, preconditioner=W.Preconditioner.Off)
        assert np.all(np.isfinite(r.x))
    r = W.solve(cats(np.array([0, 0, 1, 1]), np.array([0, 1, 0, 1])), np.array([1.0, 2.0, 3.0, 4.0]))
    assert r.converged and np.all(np.isfinite(r.x))


def test_iteration_limits_and_tolerances():
    rng = np.random.default_rng(42)
    c = cats(rng.integers(0, 50, 1000), rng.integers(0, 50, 1000))
    y = rng.standard_normal(1000)
    for cfg in (W.CG(maxiter=1, tol=1e-15), W.CG(maxiter=2, tol=1e-15), W.GMRES(maxiter=1, tol=1e-15)):
        r = W.solve(c, y, cfg)
        assert not r.converged and np.all(np.isfinite(r.x))
    rng = np.random.default_rng(0)
    c2 = cats(rng.integers(0, 20, 200), rng.integers(0, 20, 200))
    r = W.solve(c2, rng.standard_normal(200), W.CG(tol=1.0))
    assert r.converged and r.iterations <= 5
    c3 = cats(np.array([0, 1, 0, 1, 2]), np.array([0, 0, 1, 1, 0]))
    y3 = np.arange(1.0, 6.0)
    assert np.all(np.isfinite(W.solve(c3, y3, W.CG(tol=1e-14)).x))
    assert np.all(np.isfinite(W.solve(c3, y3, W.CG(tol=0.0)).x))
    r = W.solve(np.asfortranarray(c3[:3]), y3[:3], W.CG(tol=float("nan")))
    assert isinstance(r.converged, bool)


def test_non_contiguous_inputs_match():
    c = cats(np.array([0, 1, 0, 1, 2]), np.array([0, 0, 1, 1, 0]))
    y = np.array([1.0, 2.0, 3.0, 4.0, 5.0])
    yf = np.zeros(10)
    yf[::2] = y
    a, b = W.solve(c, y), W.solve(c, yf[::2])
    assert np.abs(a.x - b.x).max() <= 1e-12
    w = np.array([0.5, 1.0, 1.5, 0.8, 1.2])
    wf = np.zeros(10)
    wf[::2] = w
    a, b = W.solve(c, y, weights=w), W.solve(c, y, weights=wf[::2])
    assert np.abs(a.x - b.x).max() <= 1e-12


# ---- test_properties.py ------------------------------------------------------
def test_properties_on_random_panels():
    rng = np.random.default_rng(7)
    for trial in range(4):
        n = int(rng.integers(200, 3000))
        c = cats(rng.integers(0, int(rng.integers(5, 60)), n), rng.integers(0, int(rng.integers(5, 40)), n),
                 rng.integers(0, int(rng.integers(2, 9)), n))
        y = rng.standard_normal(n)
        s = W.Solver(c)
        r = s.solve(y)
        assert r.converged and r.residual >= 0
        assert np.array_equal(s.solve(y).x, r.x)  # determinism
        D = dense_design(c)
        assert np.abs(r.demeaned - (y - D @ r.x)).max() <= 1e-10 * max(1.0, np.abs(y).max())
        e = W.solve(c, y, W.CG(operator=W.OperatorRepr.Explicit))
        assert rel(e.demeaned, r.demeaned) <= 1e-7
        assert r.x.shape == (s.n_dofs,) and r.demeaned.shape == (s.n_obs,)
        for red in (W.ReductionStrategy.AtomicScatter, W.ReductionStrategy.ParallelReduction):
            rr = W.solve(c, y, preconditioner=W.AdditiveSchwarz(reduction=red))
            assert rr.converged and rel(rr.demeaned, r.demeaned) <= 1e-8


def test_one_solver_shared_by_threads(problem):
    """The reference's Solver is shared across threads (solver.rs:242-277);
    concurrent calls on one handle give the sequential results."""
    import threading

    c, y = problem
    s = W.Solver(c, W.CG(tol=1e-10))
    rng = np.random.default_rng(5)
    ys = [y + rng.standard_normal(len(y)) for _ in range(6)]
    seq = [s.solve(v).demeaned for v in ys]
    out = [None] * len(ys)

    def work(i):
        out[i] = s.solve(ys[i]).demeaned

    ts = [threading.Thread(target=work, args=(i,)) for i in range(len(ys))]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    for a, b in zip(out, seq):
        assert np.array_equal(a, b)


def test_single_factor_matches_the_reference_semantics(oracle_mod):
    """With one factor the Schwarz preconditioner has no subdomains and its
    apply is zero (the additive sum over no subdomains), so CG stops at once,
    unconverged, with x = 0, as the oracle (and the reference) do; with the
    preconditioner off the solve converges to the group means."""
    rng = np.random.default_rng(0)
    c1 = np.asfortranarray(rng.integers(0, 50, (1000, 1)).astype(np.uint32))
    y = rng.standard_normal(1000)
    r = W.solve(c1, y)
    ro = oracle_mod.OracleSolver(c1, None, ac_split=2).solve_batch(y[:, None])
    assert r.converged == bool(ro["converged"][0]) is False
    assert r.iterations == int(ro["iterations"][0])
    assert np.array_equal(r.demeaned, y) and not np.any(r.x)
    p = W.Solver(c1).preconditioner()
    assert p.n_subdomains == 0 and not np.any(p.apply(np.ones(50)))
    r0 = W.solve(c1, y, preconditioner=W.Preconditioner.Off)
    ro0 = oracle_mod.OracleSolver(c1, None, precond=0).solve_batch(y[:, None])
    assert r0.converged and bool(ro0["converged"][0])
    assert rel(r0.demeaned, ro0["demeaned"][:, 0]) <= 1e-10
    means = np.bincount(c1[:, 0], weights=y, minlength=50) / np.maximum(np.bincount(c1[:, 0], minlength=50), 1)
    r1 = W.solve(c1, y, W.CG(tol=1e-13), preconditioner=W.Preconditioner.Off)
    assert np.abs(r1.demeaned - (y - means[c1[:, 0]])).max() <= 1e-12


@pytest.mark.parametrize("operator", [W.OperatorRepr.Explicit, W.OperatorRepr.Implicit])
def test_stalled_mixed_columns_finish_in_fp64(oracle_mod, monkeypatch, operator):
    """A column whose FP64 residual stops falling under FP32 corrections is
    finished by FP64 refinement from where it stands (SURVEY's FP32 -> FP64
    retry).  A stall ratio of 1e-300 forces the hand-over after the first FP32
    correction; the result must still be the oracle's projection."""
    from within_amd import synth

    p = synth.akm_panel(40000, 4000, 400, 10, seed=14, weights="lognormal")
    codes, w = p["codes"], p["weights"]
    Y = synth.akm_response(p, k=3)
    ro = oracle_mod.OracleSolver(codes, w, tol=1e-10).solve_batch(Y)
    cfg = W.CG(tol=1e-10, precision="mixed", operator=operator)
    plain = W.Solver(codes, weights=w, config=cfg).solve_batch(Y)
    monkeypatch.setenv("WITHIN_AMD_STALL_RATIO", "1e-300")
    r = W.Solver(codes, weights=w, config=cfg).solve_batch(Y)
    for c in range(Y.shape[1]):
        assert r.converged[c] and rel(r.demeaned[:, c], ro["demeaned"][:, c]) <= 1e-8
        # the FP64 corrections are counted with the FP32 ones
        assert r.iterations[c] > 0
    # without the forced stall, the mixed path never hands over on this problem
    monkeypatch.delenv("WITHIN_AMD_STALL_RATIO")
    again = W.Solver(codes, weights=w, config=cfg).solve_batch(Y)
    np.testing.assert_array_equal(again.demeaned, plain.demeaned)
    assert again.iterations == plain.iterations
