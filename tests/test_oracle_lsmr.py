"""The oracle's LSMR restatement pinned against the reference's own LSMR tests
(crates/schwarz-precond/src/solve/lsmr/tests.rs: the same operators, right-hand
sides, tolerances and expected answers), numpy's least squares, and the CG
solve of the same fixed-effects design (solver.rs:162-207)."""
import numpy as np
import pytest

from within_amd import synth

OVER = np.array([[1.0, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0]])  # tests.rs:8-31
DIAG_MINV = np.array([0.5, 0.5, 1.0])                            # tests.rs:33-53


def normal_residual(a, x, b):
    return float(np.linalg.norm(a.T @ (b - a @ x)))


def vandermonde(rows, cols):  # tests.rs:384-398
    x = np.arange(rows) / max(rows - 1, 1)
    return np.column_stack([x ** j for j in range(cols)])


def smooth_rhs(rows):
    return np.log1p(np.arange(rows) / (rows - 1))


@pytest.fixture(scope="module")
def L(oracle_mod):
    return oracle_mod.lsmr_dense


def test_overdetermined_plain_and_preconditioned(L):  # tests.rs:64-93
    b = np.array([1.0, 2.0, 3.0, 3.0])
    for minv in (None, DIAG_MINV):
        r = L(OVER, b, minv, tol=1e-10, maxiter=100)
        assert r["converged"]
        assert np.linalg.norm(r["x"] - [1, 2, 3]) < 1e-6


def test_inconsistent_system(L):  # tests.rs:95-108
    b = np.array([1.0, 2.0, 3.0, 0.0])
    r = L(OVER, b, tol=1e-10)
    assert r["converged"] and normal_residual(OVER, r["x"], b) < 1e-6


def test_under_rank_deficient_and_degenerate(L):  # tests.rs:110-195
    r = L(np.array([[1.0, 0, 0], [0, 1, 0]]), np.array([1.0, 2.0]), tol=1e-12)
    assert r["converged"] and np.abs(r["x"] - [1, 2, 0]).max() < 1e-10
    a = np.array([[1.0, 1.0], [2.0, 2.0]])
    r = L(a, np.array([3.0, 6.0]), tol=1e-12)
    assert r["converged"] and abs(r["x"].sum() - 3.0) < 1e-10
    assert normal_residual(a, r["x"], np.array([3.0, 6.0])) < 1e-10
    a = np.array([[1.0, 0.0], [0.0, 0.0]])
    r = L(a, np.array([2.0, 3.0]), tol=1e-12)
    assert r["converged"] and abs(r["x"][0] - 2.0) < 1e-10 and abs(r["x"][1]) < 1e-10


def test_mid_stream_beta_zero_breakdown(L):  # tests.rs:196-225
    r = L(np.array([[1.0, 0.0], [0.0, 0.0]]), np.array([5.0, 0.0]), tol=1e-12)
    assert r["converged"] and abs(r["x"][0] - 5.0) < 1e-12 and abs(r["x"][1]) < 1e-12


def test_maxiter_exhaustion(L):  # tests.rs:227-237
    r = L(OVER, np.array([1.0, 2.0, 3.0, 3.0]), tol=1e-15, maxiter=1)
    assert not r["converged"] and r["iterations"] == 1 and r["stop"] == "maxiter"


def test_plain_matches_identity_preconditioner(L):  # tests.rs:270-309
    b = np.array([1.0, 2.0, 3.0, 3.0])
    a = L(OVER, b, tol=1e-12)
    i = L(OVER, b, np.ones(3), tol=1e-12)
    assert a["converged"] and i["converged"] and abs(a["iterations"] - i["iterations"]) <= 1
    assert np.linalg.norm(a["x"] - i["x"]) < 1e-10
    assert abs(a["residual_norm"] - i["residual_norm"]) < 1e-10


def test_windowed_plain_matches_identity(L):  # tests.rs:311-367
    A, b = vandermonde(30, 12), smooth_rhs(30)
    a = L(A, b, tol=1e-12, maxiter=50, local_size=10)
    i = L(A, b, np.ones(12), tol=1e-12, maxiter=50, local_size=10)
    assert a["converged"] and i["converged"] and abs(a["iterations"] - i["iterations"]) <= 1
    assert np.linalg.norm(a["x"] - i["x"]) < 1e-7
    assert abs(a["residual_norm"] - i["residual_norm"]) < 1e-7


def test_window_zero_is_no_window(L):  # tests.rs:424-440
    b = np.array([1.0, 2.0, 3.0, 3.0])
    a, z = L(OVER, b, tol=1e-10), L(OVER, b, tol=1e-10, local_size=0)
    assert a["iterations"] == z["iterations"] and np.array_equal(a["x"], z["x"]) and a["converged"]


def test_window_recovers_ill_conditioned(L):  # tests.rs:442-531
    A, b = vandermonde(30, 12), smooth_rhs(30)
    r0 = L(A, b, tol=1e-9, maxiter=30)
    r10 = L(A, b, tol=1e-9, maxiter=30, local_size=10)
    assert r10["converged"]
    assert (not r0["converged"]) or r10["iterations"] <= r0["iterations"]
    assert normal_residual(A, r10["x"], b) < 1e-6
    d = (A * A).sum(axis=0)
    rp = L(A, b, np.where(d > 0, 1.0 / d, 1.0), tol=1e-9, maxiter=30, local_size=10)
    assert rp["converged"] and normal_residual(A, rp["x"], b) < 1e-6


@pytest.mark.parametrize("window", [1, 12, 13])
def test_window_boundary_sizes(L, window):  # tests.rs:652-677
    A, b = vandermonde(30, 12), smooth_rhs(30)
    r = L(A, b, tol=1e-9, maxiter=200, local_size=window)
    assert r["converged"] and normal_residual(A, r["x"], b) < 1e-6


def test_early_exits(L):  # tests.rs:580-650
    r = L(np.array([[1.0], [0.0]]), np.array([0.0, 1.0]), tol=1e-12)
    assert r["converged"] and r["iterations"] == 0 and r["x"].tolist() == [0.0]
    assert r["stop"] == "atb_zero" and abs(r["residual_norm"] - 1.0) < 1e-15
    r = L(OVER, np.zeros(4), tol=1e-12)
    assert r["converged"] and r["iterations"] == 0 and r["stop"] == "zero_rhs" and r["residual_norm"] == 0.0


def test_rejects_invalid_inputs(L, oracle_mod):  # tests.rs:679-700
    with pytest.raises(oracle_mod.OracleError):
        L(OVER, np.array([1.0, 2.0, 3.0, 4.0]), tol=float("nan"))
    with pytest.raises(oracle_mod.OracleError):
        L(OVER, np.array([1.0, np.inf, 3.0, 4.0]), tol=1e-10)


@pytest.mark.parametrize("precond", [False, True])
def test_random_least_squares_match_numpy(L, precond):
    rng = np.random.default_rng(1)
    A = rng.standard_normal((60, 9)) * rng.uniform(0.2, 5.0, 9)
    b = rng.standard_normal(60)
    minv = 1.0 / (A * A).sum(axis=0) if precond else None
    r = L(A, b, minv, tol=1e-14, maxiter=200)
    ref = np.linalg.lstsq(A, b, rcond=None)[0]
    assert np.linalg.norm(r["x"] - ref) <= 1e-9 * np.linalg.norm(ref)


@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("precond", [0, 1])
def test_fixed_effects_lsmr_matches_cg(oracle_mod, weighted, precond):
    """The demeaned residual is the unique least-squares residual, so LSMR and
    CG solves of the same design agree on it (solver.rs:162-207)."""
    p = synth.akm_panel(6000, 600, 80, 8, seed=12, weights="lognormal" if weighted else None)
    codes, w = p["codes"], p.get("weights") if weighted else None
    Y = np.asfortranarray(np.random.default_rng(3).standard_normal((codes.shape[0], 2)))
    ls = oracle_mod.OracleSolver(codes, w, krylov=2, precond=precond, ac_split=2, tol=1e-12, maxiter=5000)
    cg = oracle_mod.OracleSolver(codes, w, krylov=0, precond=1, ac_split=2, tol=1e-12, max_refinements=3)
    rl, rc = ls.solve_batch(Y), cg.solve_batch(Y)
    assert all(rl["converged"])
    for j in range(2):
        d = np.linalg.norm(rl["demeaned"][:, j] - rc["demeaned"][:, j]) / np.linalg.norm(rc["demeaned"][:, j])
        assert d <= 1e-8, d
        assert rl["residual"][j] <= 1e-8
