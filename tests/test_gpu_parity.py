"""GPU parity: the gfx950 path through the C ABI against the FP64 CPU oracle.

Bars (written per test): design ops and operator applies to 1e-12 relative
(FP64, different but fixed summation orders); preconditioner applies to 1e-10;
end-to-end demeaned residuals and Dx within 1e-8 relative of the oracle's own
solve on the same seeded inputs, with matching converged flags.
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
    rng = np.random.default_rng(11)
    # (codes, weights, name)
    c, _, _ = synth.uniform_problem(3000, [60, 25], seed=1)
    out.append((c, None, "2fe-uniform"))
    c, w, _ = synth.uniform_problem(4000, [80, 30, 7], seed=2, weights=True)
    out.append((c, w, "3fe-weighted"))
    p = synth.akm_panel(20000, 2000, 300, 12, seed=3)
    out.append((p["codes"], None, "akm-3fe"))
    p = synth.akm_panel(15000, 1500, 200, 10, n_industries=5, seed=4, weights="lognormal")
    out.append((p["codes"], p["weights"], "akm-4fe-weighted"))
    c = np.asfortranarray(np.column_stack([rng.integers(0, 40, 500), rng.integers(0, 3, 500)]).astype(np.uint32))
    out.append((c, None, "dense-path"))
    return out


PROBS = problems()


def test_design_ops_known_answers():
    # crates/within/src/domain.rs:491-545
    cats = np.asfortranarray(np.array([[0, 0], [1, 1], [2, 2], [0, 3], [1, 0]], dtype=np.uint32))
    s = W.Solver(cats, preconditioner=W.Preconditioner.Off)
    L = N.load()
    x = np.array([1.0, 2.0, 3.0, 10.0, 20.0, 30.0, 40.0])
    y = np.zeros(5)
    N.check(L.wa_matvec_d(s._h, N.ptr(x), 7, N.ptr(y), 5, 1))
    assert y.tolist() == [11.0, 22.0, 33.0, 41.0, 12.0]
    r = np.array([1.0, 2.0, 3.0, 4.0, 5.0])
    out = np.zeros(7)
    N.check(L.wa_rmatvec_wdt(s._h, N.ptr(r), 5, N.ptr(out), 7, 1))
    assert out.tolist() == [5.0, 7.0, 3.0, 6.0, 2.0, 3.0, 4.0]
    cats2 = np.asfortranarray(np.array([[0, 0], [1, 0], [0, 1], [1, 1]], dtype=np.uint32))
    s2 = W.Solver(cats2, weights=np.array([1.0, 2.0, 3.0, 4.0]), preconditioner=W.Preconditioner.Off)
    out = np.zeros(4)
    N.check(L.wa_rmatvec_wdt(s2._h, N.ptr(np.ones(4)), 4, N.ptr(out), 4, 1))
    assert out.tolist() == [4.0, 6.0, 3.0, 7.0]


@pytest.mark.parametrize("idx", range(len(PROBS)), ids=[p[2] for p in PROBS])
def test_operator_and_design_ops_match_oracle(idx, oracle_mod):
    codes, w, _ = PROBS[idx]
    o = oracle_mod.OracleSolver(codes, w, precond=0)
    s = W.Solver(codes, weights=w, preconditioner=W.Preconditioner.Off)
    L = N.load()
    m, n = s.n_dofs, s.n_obs
    rng = np.random.default_rng(idx)
    for k in (1, 3, 8, 32):
        X = np.asfortranarray(rng.standard_normal((m, k)))
        Y = np.zeros((m, k), order="F")
        N.check(L.wa_gram_apply(s._h, N.ptr(X), N.ptr(Y), k, m))
        ref = np.column_stack([o.gram_apply(X[:, j]) for j in range(k)])
        assert rel(Y, ref) <= 1e-12, k
        D = np.zeros((n, k), order="F")
        N.check(L.wa_matvec_d(s._h, N.ptr(X), m, N.ptr(D), n, k))
        refd = np.column_stack([o.matvec_d(X[:, j]) for j in range(k)])
        assert rel(D, refd) <= 1e-14, k
        R = np.asfortranarray(rng.standard_normal((n, k)))
        B = np.zeros((m, k), order="F")
        N.check(L.wa_rmatvec_wdt(s._h, N.ptr(R), n, N.ptr(B), m, k))
        refb = np.column_stack([o.rmatvec_wdt(R[:, j]) for j in range(k)])
        assert rel(B, refb) <= 1e-12, k


@pytest.mark.parametrize("idx", range(len(PROBS)), ids=[p[2] for p in PROBS])
@pytest.mark.parametrize("cfg", ["rust-default", "python-default", "exact-schur", "pure-ac"])
def test_preconditioner_apply_matches_oracle(idx, cfg, oracle_mod):
    codes, w, _ = PROBS[idx]
    if cfg == "rust-default":
        prec = W.AdditiveSchwarz(W.SchurComplement(W.ApproxCholConfig(0, 1), W.ApproxSchurConfig(0, 1)))
        okw = dict(ac_split=1)
    elif cfg == "python-default":
        prec, okw = None, dict(ac_split=2)
    elif cfg == "exact-schur":
        prec, okw = W.AdditiveSchwarz(W.SchurComplement()), dict(ac_split=1, approx_schur=False)
    else:
        prec = W.AdditiveSchwarz(W.SchurComplement(W.ApproxCholConfig(3, 1), W.ApproxSchurConfig(5, 2), 0))
        okw = dict(ac_seed=3, schur_seed=5, schur_split=2, dense_threshold=0)
    o = oracle_mod.OracleSolver(codes, w, **okw)
    s = W.Solver(codes, weights=w, preconditioner=prec)
    L = N.load()
    m = s.n_dofs
    rng = np.random.default_rng(100 + idx)
    for k in (1, 4, 32):
        R = np.asfortranarray(rng.standard_normal((m, k)))
        Z = np.zeros((m, k), order="F")
        N.check(L.wa_precond_apply(s._h, N.ptr(R), N.ptr(Z), k, m))
        ref = np.column_stack([o.precond_apply(R[:, j]) for j in range(k)])
        assert rel(Z, ref) <= 1e-10, (cfg, k, rel(Z, ref))


@pytest.mark.parametrize("idx", range(len(PROBS)), ids=[p[2] for p in PROBS])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_solve_batch_matches_oracle(idx, precision, oracle_mod):
    codes, w, _ = PROBS[idx]
    n = codes.shape[0]
    rng = np.random.default_rng(7 + idx)
    Y = np.asfortranarray(rng.standard_normal((n, 5)))
    Y[:, 0] += np.sin(codes[:, 0].astype(float))  # some fixed-effect signal
    o = oracle_mod.OracleSolver(codes, w, ac_split=2, tol=1e-10, max_refinements=3)
    ro = o.solve_batch(Y)
    s = W.Solver(codes, config=W.CG(tol=1e-10, max_refinements=3, precision=precision), weights=w)
    r = s.solve_batch(Y)
    assert all(r.converged) and all(ro["converged"])
    # demeaned residuals and Dx are unique (x is not): 1e-8 relative per RHS
    for j in range(Y.shape[1]):
        assert rel(r.demeaned[:, j], ro["demeaned"][:, j]) <= 1e-8, (j, rel(r.demeaned[:, j], ro["demeaned"][:, j]))
        dx, dxo = Y[:, j] - r.demeaned[:, j], Y[:, j] - ro["demeaned"][:, j]
        assert rel(dx, dxo) <= 1e-8
        assert r.residual[j] <= 1e-10 * 1.0001


def test_default_tolerance_flags_and_iterations(oracle_mod):
    codes, w, _ = PROBS[2]
    Y = np.asfortranarray(np.random.default_rng(5).standard_normal((codes.shape[0], 3)))
    ro = oracle_mod.OracleSolver(codes, w, ac_split=2).solve_batch(Y)
    r = W.solve_batch(codes, Y)
    assert r.converged == [bool(c) for c in ro["converged"]]
    for j in range(3):
        assert r.residual[j] <= 1e-8
        assert rel(r.demeaned[:, j], ro["demeaned"][:, j]) <= 1e-6


def test_readme_fwl_known_answer():
    fe, Y, _ = synth.readme_fwl()
    r = W.solve_batch(fe, Y)
    beta = np.linalg.lstsq(r.demeaned[:, 1:], r.demeaned[:, 0], rcond=None)[0]
    assert np.round(beta, 4).tolist() == [0.9982, -2.006, 0.5005]


def test_repeated_solve_is_bitwise_deterministic():
    codes, w, _ = PROBS[3]
    y = np.random.default_rng(1).standard_normal(codes.shape[0])
    s = W.Solver(codes, weights=w)
    a, b = s.solve(y), s.solve(y)
    assert np.array_equal(a.x, b.x) and np.array_equal(a.demeaned, b.demeaned)
    sm = W.Solver(codes, weights=w, config=W.CG(precision="mixed"))
    a, b = sm.solve(y), sm.solve(y)
    assert np.array_equal(a.x, b.x)


def test_zero_rhs_gives_zero_solution():
    codes, _, _ = PROBS[0]
    r = W.solve(codes, np.zeros(codes.shape[0]))
    assert r.converged and r.iterations == 0 and np.all(r.x == 0.0)


def test_batch_equals_single():
    codes, w, _ = PROBS[1]
    Y = np.asfortranarray(np.random.default_rng(3).standard_normal((codes.shape[0], 3)))
    s = W.Solver(codes, weights=w, config=W.CG(tol=1e-10))
    rb = s.solve_batch(Y)
    for j in range(3):
        r1 = s.solve(Y[:, j])
        assert rel(r1.demeaned, rb.demeaned[:, j]) <= 1e-9


# ---- GMRES, multiplicative Schwarz, explicit operator ---------------------
@pytest.mark.parametrize("idx", [0, 2, 3], ids=["2fe-uniform", "akm-3fe", "akm-4fe-weighted"])
def test_explicit_operator_matches_oracle(idx, oracle_mod):
    codes, w, _ = PROBS[idx]
    o = oracle_mod.OracleSolver(codes, w, precond=0)
    ip, ix, dt = o.explicit_gramian()
    m = o.n_dofs
    G = np.zeros((m, m))
    for r in range(m):
        G[r, ix[ip[r]:ip[r + 1]]] = dt[ip[r]:ip[r + 1]]
    s = W.Solver(codes, weights=w, config=W.CG(operator=W.OperatorRepr.Explicit), preconditioner=W.Preconditioner.Off)
    X = np.asfortranarray(np.random.default_rng(idx).standard_normal((m, 5)))
    Y = np.zeros((m, 5), order="F")
    N.check(N.load().wa_gram_apply_explicit(s._h, N.ptr(X), N.ptr(Y), 5, m))
    assert rel(Y, G @ X) <= 1e-12


@pytest.mark.parametrize("idx", [0, 2, 3], ids=["2fe-uniform", "akm-3fe", "akm-4fe-weighted"])
def test_multiplicative_apply_matches_oracle(idx, oracle_mod):
    codes, w, _ = PROBS[idx]
    o = oracle_mod.OracleSolver(codes, w, krylov=1, precond=2, ac_split=2)
    s = W.Solver(codes, weights=w, config=W.GMRES(), preconditioner=W.MultiplicativeSchwarz())
    m = s.n_dofs
    R = np.asfortranarray(np.random.default_rng(50 + idx).standard_normal((m, 3)))
    Z = np.zeros((m, 3), order="F")
    N.check(N.load().wa_precond_apply(s._h, N.ptr(R), N.ptr(Z), 3, m))
    ref = np.column_stack([o.precond_apply(R[:, j]) for j in range(3)])
    assert rel(Z, ref) <= 1e-10, rel(Z, ref)


@pytest.mark.parametrize("idx", [1, 2, 3], ids=["3fe-weighted", "akm-3fe", "akm-4fe-weighted"])
@pytest.mark.parametrize("kind", ["gmres-additive", "gmres-multiplicative", "gmres-mixed", "cg-explicit",
                                  "cg-mixed-explicit"])
def test_other_solver_paths_match_oracle(idx, kind, oracle_mod):
    codes, w, _ = PROBS[idx]
    n = codes.shape[0]
    Y = np.asfortranarray(np.random.default_rng(idx).standard_normal((n, 3)))
    o = oracle_mod.OracleSolver(codes, w, ac_split=2, tol=1e-10, max_refinements=3)
    ro = o.solve_batch(Y)
    if kind == "gmres-additive":
        cfg, prec = W.GMRES(tol=1e-10, max_refinements=3), None
    elif kind == "gmres-multiplicative":
        cfg, prec = W.GMRES(tol=1e-10, max_refinements=3), W.MultiplicativeSchwarz()
    elif kind == "gmres-mixed":
        cfg, prec = W.GMRES(tol=1e-10, max_refinements=3, precision="mixed"), None
    elif kind == "cg-explicit":
        cfg, prec = W.CG(tol=1e-10, max_refinements=3, operator=W.OperatorRepr.Explicit), None
    else:  # residuals b - Gx from the FP64 Gramian, demeaned written from Y
        cfg = W.CG(tol=1e-10, max_refinements=3, operator=W.OperatorRepr.Explicit, precision="mixed")
        prec = None
    r = W.solve_batch(codes, Y, config=cfg, weights=w, preconditioner=prec)
    assert all(r.converged)
    for j in range(3):
        assert rel(r.demeaned[:, j], ro["demeaned"][:, j]) <= 1e-8
        # the reported residual is the observation-space one, to FP64 rounding
        res = o.rmatvec_wdt(r.demeaned[:, j])
        b = o.rmatvec_wdt(Y[:, j])
        assert abs(r.residual[j] - np.linalg.norm(res) / np.linalg.norm(b)) <= 1e-12 + 1e-3 * r.residual[j]


def test_gmres_iterations_track_oracle(oracle_mod):
    # same operator and preconditioner, same GMRES recurrence: iteration counts agree
    codes, w, _ = PROBS[2]
    y = np.random.default_rng(9).standard_normal(codes.shape[0])
    ro = oracle_mod.OracleSolver(codes, w, krylov=1, precond=2, ac_split=2).solve(y)
    r = W.solve(codes, y, config=W.GMRES(), weights=w, preconditioner=W.MultiplicativeSchwarz())
    assert r.converged and ro["converged"]
    assert abs(r.iterations - ro["iterations"]) <= 1


# A factor pair whose reduced system ends in a dense core of at least 512
# vertices (the tiled and matrix-vector dense-core kernels): 971 vertices with
# the default core size, 538 after one round of eliminations at dense_core=600.
BIG = synth.akm_panel(60000, 6000, 1000, 10, seed=8)["codes"]
RUST_DEFAULT = W.AdditiveSchwarz(W.SchurComplement(W.ApproxCholConfig(0, 1), W.ApproxSchurConfig(0, 1)))


@pytest.mark.parametrize("dense_core", [0, 600])
def test_dense_core_apply_matches_oracle(dense_core, oracle_mod):
    okw = dict(ac_split=1)
    if dense_core:
        okw["dense_core"] = dense_core
    o = oracle_mod.OracleSolver(BIG, None, **okw)
    s = W.Solver(BIG, preconditioner=RUST_DEFAULT, _dense_core=dense_core)
    assert s.info()[13] >= 512
    m = s.n_dofs
    rng = np.random.default_rng(123 + dense_core)
    for k in (1, 2, 3, 4, 8, 16, 32):  # every kernel width
        R = np.asfortranarray(rng.standard_normal((m, k)))
        Z = np.zeros((m, k), order="F")
        N.check(N.load().wa_precond_apply(s._h, N.ptr(R), N.ptr(Z), k, m))
        ref = np.column_stack([o.precond_apply(R[:, j]) for j in range(k)])
        assert rel(Z, ref) <= 1e-10, (k, rel(Z, ref))


@pytest.mark.parametrize("k", [4, 32])
def test_dense_core_mixed_solve_matches_oracle(k, oracle_mod):
    n = BIG.shape[0]
    Y = np.asfortranarray(np.random.default_rng(9 + k).standard_normal((n, k)))
    o = oracle_mod.OracleSolver(BIG, None, ac_split=1, tol=1e-10, max_refinements=3)
    ro = o.solve_batch(Y)
    s = W.Solver(BIG, config=W.CG(tol=1e-10, max_refinements=3, precision="mixed"), preconditioner=RUST_DEFAULT)
    r = s.solve_batch(Y)
    assert all(r.converged) and all(ro["converged"])
    for j in range(k):
        assert rel(r.demeaned[:, j], ro["demeaned"][:, j]) <= 1e-8, (j, rel(r.demeaned[:, j], ro["demeaned"][:, j]))


def test_dense_core_multiplicative_apply_matches_oracle(oracle_mod):
    o = oracle_mod.OracleSolver(BIG, None, krylov=1, precond=2, ac_split=2)
    s = W.Solver(BIG, config=W.GMRES(), preconditioner=W.MultiplicativeSchwarz())
    m = s.n_dofs
    R = np.asfortranarray(np.random.default_rng(77).standard_normal((m, 3)))
    Z = np.zeros((m, 3), order="F")
    N.check(N.load().wa_precond_apply(s._h, N.ptr(R), N.ptr(Z), 3, m))
    ref = np.column_stack([o.precond_apply(R[:, j]) for j in range(3)])
    assert rel(Z, ref) <= 1e-10, rel(Z, ref)


def test_two_stream_apply_is_bitwise_the_one_stream_apply(oracle_mod, monkeypatch):
    """With a 64-vertex dense core the worker-firm pair keeps 11 elimination
    levels, so the additive apply runs the other pairs' local solves on a
    second stream beside that level chain.  Same kernels on disjoint rows: the
    result equals the one-stream apply bit for bit, for applies and for whole
    mixed-precision solves, and the oracle to 1e-10."""
    o = oracle_mod.OracleSolver(BIG, None, ac_split=1, dense_core=64)
    cfg = W.CG(tol=1e-10, max_refinements=3, precision="mixed")
    two = W.Solver(BIG, config=cfg, preconditioner=RUST_DEFAULT, _dense_core=64)
    monkeypatch.setenv("WITHIN_AMD_PRECOND_STREAMS", "1")
    one = W.Solver(BIG, config=cfg, preconditioner=RUST_DEFAULT, _dense_core=64)
    assert two.info()[15] & 255 == 2 and one.info()[15] == 1
    m = two.n_dofs
    rng = np.random.default_rng(31)
    for k in (1, 4, 32):
        R = np.asfortranarray(rng.standard_normal((m, k)))
        Z1, Z2 = np.zeros((m, k), order="F"), np.zeros((m, k), order="F")
        N.check(N.load().wa_precond_apply(one._h, N.ptr(R), N.ptr(Z1), k, m))
        N.check(N.load().wa_precond_apply(two._h, N.ptr(R), N.ptr(Z2), k, m))
        assert np.array_equal(Z1, Z2), k
        ref = np.column_stack([o.precond_apply(R[:, j]) for j in range(k)])
        assert rel(Z2, ref) <= 1e-10, (k, rel(Z2, ref))
    Y = np.asfortranarray(rng.standard_normal((BIG.shape[0], 8)))
    a, b = one.solve_batch(Y), two.solve_batch(Y)
    assert all(a.converged) and list(a.iterations) == list(b.iterations)
    assert np.array_equal(a.demeaned, b.demeaned) and np.array_equal(a.x, b.x)


@pytest.mark.parametrize("config", [W.CG(tol=1e-10, max_refinements=3), W.GMRES(tol=1e-9),
                                    W.GMRES(tol=1e-9, precision="mixed"), W.LSMR(tol=1e-9)],
                         ids=["cg-fp64", "gmres-fp64", "gmres-mixed", "lsmr"])
def test_two_stream_solves_are_bitwise_the_one_stream_solves(config, monkeypatch):
    """Every Krylov method that applies the additive preconditioner gives the
    same bits with the two-stream apply (the x update of CG on the second
    stream included)."""
    two = W.Solver(BIG, config=config, preconditioner=RUST_DEFAULT, _dense_core=64)
    monkeypatch.setenv("WITHIN_AMD_PRECOND_STREAMS", "1")
    one = W.Solver(BIG, config=config, preconditioner=RUST_DEFAULT, _dense_core=64)
    assert two.info()[15] & 255 == 2 and one.info()[15] == 1
    Y = np.asfortranarray(np.random.default_rng(41).standard_normal((BIG.shape[0], 5)))
    a, b = one.solve_batch(Y), two.solve_batch(Y)
    assert all(a.converged) and list(a.iterations) == list(b.iterations)
    assert np.array_equal(a.demeaned, b.demeaned) and np.array_equal(a.x, b.x)


@pytest.mark.skip(reason="level runs are switched off: both GPU runs of this test stalled in the first run's "
                         "waits for the previous level's count; the cause is not found")
def test_level_runs_are_bitwise_the_per_level_launches(oracle_mod, monkeypatch):
    """The lead pair's trailing elimination levels as one launch per direction
    (waves ordered by tickets, one XCD) give the same bits as one launch per
    level, over many repeated applies (a missed wait would show up as a
    changed or varying result), and the oracle's apply to 1e-10."""
    flow = W.Solver(BIG, config=W.CG(tol=1e-10, max_refinements=3, precision="mixed"), preconditioner=RUST_DEFAULT,
                    _dense_core=64)
    monkeypatch.setenv("WITHIN_AMD_LEVEL_FLOW", "0")
    plain = W.Solver(BIG, config=W.CG(tol=1e-10, max_refinements=3, precision="mixed"), preconditioner=RUST_DEFAULT,
                     _dense_core=64)
    assert flow.info()[15] & 255 == 2 and flow.info()[15] >> 8 >= 4
    assert plain.info()[15] == 2
    o = oracle_mod.OracleSolver(BIG, None, ac_split=1, dense_core=64)
    m = flow.n_dofs
    rng = np.random.default_rng(51)
    L = N.load()
    for k in (1, 4, 32):
        R = np.asfortranarray(rng.standard_normal((m, k)))
        Zp = np.zeros((m, k), order="F")
        N.check(L.wa_precond_apply(plain._h, N.ptr(R), N.ptr(Zp), k, m))
        for rep in range(10):
            Zf = np.zeros((m, k), order="F")
            N.check(L.wa_precond_apply(flow._h, N.ptr(R), N.ptr(Zf), k, m))
            assert np.array_equal(Zf, Zp), (k, rep)
        ref = np.column_stack([o.precond_apply(R[:, j]) for j in range(k)])
        assert rel(Zf, ref) <= 1e-10
    Y = np.asfortranarray(rng.standard_normal((BIG.shape[0], 32)))
    a, b = plain.solve_batch(Y), flow.solve_batch(Y)
    assert all(a.converged) and list(a.iterations) == list(b.iterations)
    assert np.array_equal(a.demeaned, b.demeaned) and np.array_equal(a.x, b.x)
    c = flow.solve_batch(Y)
    assert np.array_equal(c.x, b.x)
