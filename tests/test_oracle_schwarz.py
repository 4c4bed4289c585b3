"""Pin the oracle's Schwarz preconditioners independently (CPU only).

The GPU additive and multiplicative applies are compared with the oracle's
``Schwarz::apply``; this file pins that oracle code itself:

* the reference's own exact values for restriction / prolongation
  (schwarz-precond/src/local_solve.rs:313-376) and for the residual updater
  (schwarz/multiplicative.rs:392-430), and the sparse updater equal to the
  naive ``r - G y_accum`` updater (crates/within/src/operator/tests.rs:195-346);
* dense formulas on small 3- and 4-factor designs whose factor-pair
  subdomains overlap (every DOF sits in Q-1 subdomains, PoU weights
  1/sqrt(Q-1)), with exact local solves (exact Schur complement and a dense
  core covering the whole reduced system), so that each local solve is the
  pseudo-inverse of the subdomain's block of G:
    additive        z = sum_i R_i' W_i G_i^+ W_i R_i r
                    (additive/preconditioner.rs:164-214, local_solve.rs:182-248)
    multiplicative  the forward sweep of multiplicative.rs:128-245: subdomain i
                    restricts the current r_work, z += R_i' W_i z_i and
                    r_work -= G R_i' W_i z_i (residual_update.rs:36-53);
* one subdomain covering a connected 2-factor design: the preconditioner is
  G^+ itself.

A dropped PoU weight, a missing sign flip of the r-block, or a flipped or
missing residual update changes these vectors by O(1) and fails the checks.
"""

import numpy as np
import pytest


def dense_D(cats):
    cats = np.asarray(cats)
    n, Q = cats.shape
    levels = cats.max(axis=0) + 1
    off = np.concatenate([[0], np.cumsum(levels)[:-1]])
    D = np.zeros((n, levels.sum()))
    for q in range(Q):
        D[np.arange(n), off[q] + cats[:, q]] = 1.0
    return D


def gramian(cats, w=None):
    D = dense_D(cats)
    wv = np.ones(len(D)) if w is None else np.asarray(w, float)
    return D.T @ (D * wv[:, None])


# --- reference values: restriction / prolongation (local_solve.rs:313-376) ---

def test_restrict_prolongate_uniform_values(oracle_mod):
    g = [1, 3, 4]
    glob = np.array([10.0, 20, 30, 40, 50, 60])
    loc = oracle_mod.restrict(g, None, glob)
    assert loc.tolist() == [20.0, 40.0, 50.0]
    out = oracle_mod.prolongate_add(g, None, loc, np.zeros(6))
    assert out.tolist() == [0.0, 20.0, 0.0, 40.0, 50.0, 0.0]


def test_restrict_prolongate_weighted_values(oracle_mod):
    g, w = [0, 1, 2], [1.0, 0.5, 0.25]
    loc = oracle_mod.restrict(g, w, np.array([4.0, 8.0, 16.0]))
    assert np.allclose(loc, [4.0, 4.0, 4.0], atol=1e-14)
    out = oracle_mod.prolongate_add(g, w, loc, np.zeros(3))
    assert np.allclose(out, [4.0, 2.0, 1.0], atol=1e-14)


def test_identity_local_solve_round_trip(oracle_mod):
    # local_solve.rs:360-376: R' (I) R r with uniform weights
    g = [1, 2]
    r = np.array([10.0, 20, 30, 40])
    out = oracle_mod.prolongate_add(g, None, oracle_mod.restrict(g, None, r), np.zeros(4))
    assert out.tolist() == [0.0, 20.0, 30.0, 0.0]


# --- reference values: residual updater (multiplicative.rs:392-430) ---------

def csr_of(a):
    a = np.asarray(a, float)
    ip, ix, dt = [0], [], []
    for row in a:
        nz = np.nonzero(row)[0]
        ix += nz.tolist()
        dt += row[nz].tolist()
        ip.append(len(ix))
    return np.array(ip), np.array(ix), np.array(dt)


def test_residual_updater_diag_values(oracle_mod):
    ip, ix, dt = csr_of(np.diag([2.0, 3.0, 1.0, 4.0]))
    r = np.array([10.0, 12.0, 5.0, 8.0])
    oracle_mod.residual_update(ip, ix, dt, [0, 2], [1.0, 2.0], r)
    assert np.allclose(r, [8.0, 12.0, 3.0, 8.0], atol=1e-12)
    oracle_mod.residual_update(ip, ix, dt, [1, 3], [0.5, 1.0], r)
    assert np.allclose(r, [8.0, 10.5, 3.0, 4.0], atol=1e-12)


def test_residual_updater_reset_values(oracle_mod):
    ip, ix, dt = csr_of(np.eye(2))
    r = np.array([5.0, 3.0])
    oracle_mod.residual_update(ip, ix, dt, [0], [2.0], r)
    assert abs(r[0] - 3.0) < 1e-12
    r = np.array([10.0, 7.0])
    oracle_mod.residual_update(ip, ix, dt, [1], [1.0], r)
    assert abs(r[0] - 10.0) < 1e-12 and abs(r[1] - 6.0) < 1e-12


@pytest.mark.parametrize("weighted", [False, True])
def test_sparse_updater_equals_naive_updater(oracle_mod, weighted):
    # operator/tests.rs:195-346: the oracle's explicit Gramian, two sparse
    # updates against r_original - G y_accum
    if weighted:
        cats, w = [[0, 0], [1, 0], [0, 1], [1, 1]], [1.0, 2.0, 3.0, 4.0]
        steps = [([0, 2], [0.5, -0.3])]
        r0 = np.array([5.0, 3.0, 7.0, 1.0])
    else:
        cats, w = [[0, 0], [1, 1], [2, 2], [0, 3], [1, 0]], None
        steps = [([0, 1, 3, 4], [0.5, -0.3, 0.2, 0.1]), ([2, 5, 6], [1.0, -0.5, 0.8]), ([0], [1.0]),
                 ([0, 1, 3], [0.0, 0.0, 0.0])]
        r0 = np.array([10.0, 20, 30, 40, 50, 60, 70])
    o = oracle_mod.OracleSolver(np.asarray(cats, np.uint32), None if w is None else np.array(w), precond=0,
                                explicit_op=True)
    ip, ix, dt = o.explicit_gramian()
    G = gramian(cats, w)
    r = r0.copy()
    y = np.zeros(len(r0))
    for gi, c in steps:
        oracle_mod.residual_update(ip, ix, dt, gi, c, r)
        y[gi] += c
        assert np.allclose(r, r0 - G @ y, rtol=0, atol=1e-12)
    zero = r.copy()
    oracle_mod.residual_update(ip, ix, dt, [0, 1], [0.0, 0.0], zero)
    assert np.array_equal(zero, r)


# --- dense formulas on overlapping subdomains ---------------------------------

def overlapping_designs():
    rng = np.random.default_rng(31)
    out = []
    # 3 factors: worker-like, firm-like, year-like; every DOF in 2 subdomains
    n = 160
    cats = np.column_stack([rng.integers(0, 30, n), rng.integers(0, 8, n), rng.integers(0, 5, n)])
    out.append(("3fe", cats, None))
    out.append(("3fe-weighted", cats, rng.uniform(0.3, 2.5, n)))
    # 4 factors: every DOF in 3 subdomains
    n = 200
    cats = np.column_stack([rng.integers(0, 26, n), rng.integers(0, 9, n), rng.integers(0, 4, n),
                            rng.integers(0, 6, n)])
    out.append(("4fe", cats, None))
    # large enough keep block to take the sparse path (n_keep > dense_threshold)
    n = 600
    cats = np.column_stack([rng.integers(0, 70, n), rng.integers(0, 40, n), rng.integers(0, 3, n)])
    out.append(("3fe-sparse-path", cats, rng.uniform(0.5, 2.0, n)))
    return out


def exact_oracle(oracle_mod, cats, w, precond):
    return oracle_mod.OracleSolver(np.asarray(cats, np.uint32), w, precond=precond, krylov=1 if precond == 2 else 0,
                                   approx_schur=False, dense_core=4096)


def subdomain_terms(o, G):
    terms = []
    for i in range(o.n_subdomains):
        sd = o.subdomain(i)
        g = sd["gidx"].astype(int)
        assert sd["ac_rounds"] == 0  # exact local solve: no sampled rounds
        terms.append((g, sd["pou"], np.linalg.pinv(G[np.ix_(g, g)], rcond=1e-12)))
    return terms


@pytest.mark.parametrize("case", range(4))
def test_additive_apply_equals_dense_sum(oracle_mod, case):
    name, cats, w = overlapping_designs()[case]
    o = exact_oracle(oracle_mod, cats, w, 1)
    G = gramian(cats, w)
    terms = subdomain_terms(o, G)
    counts = np.zeros(len(G))
    for g, _, _ in terms:
        counts[g] += 1
    Q = cats.shape[1]
    assert np.all(counts[counts > 0] == Q - 1), name  # every DOF overlaps Q-1 subdomains
    assert any(not np.allclose(p, 1.0) for _, p, _ in terms)  # the PoU weights are not trivial
    rng = np.random.default_rng(case)
    for _ in range(3):
        r = rng.standard_normal(len(G))
        z_ref = np.zeros(len(G))
        for g, p, Ginv in terms:
            z_ref[g] += p * (Ginv @ (p * r[g]))
        z = o.precond_apply(r)
        assert np.linalg.norm(z - z_ref) <= 1e-12 * np.linalg.norm(z_ref), name


@pytest.mark.parametrize("case", range(4))
def test_multiplicative_apply_equals_dense_sweep(oracle_mod, case):
    name, cats, w = overlapping_designs()[case]
    o = exact_oracle(oracle_mod, cats, w, 2)
    G = gramian(cats, w)
    terms = subdomain_terms(o, G)
    rng = np.random.default_rng(100 + case)
    for _ in range(3):
        r = rng.standard_normal(len(G))
        rw = r.copy()
        z_ref = np.zeros(len(G))
        for g, p, Ginv in terms:
            corr = np.zeros(len(G))
            corr[g] = p * (Ginv @ (p * rw[g]))
            z_ref += corr
            rw -= G @ corr
        z = o.precond_apply(r)
        assert np.linalg.norm(z - z_ref) <= 1e-11 * np.linalg.norm(z_ref), name
        # the sweep is not the additive sum on overlapping subdomains
        z_add = sum(np.bincount(g, p * (Ginv @ (p * r[g])), len(G)) for g, p, Ginv in terms)
        assert np.linalg.norm(z - z_add) > 1e-3 * np.linalg.norm(z_add)


@pytest.mark.parametrize("precond", [1, 2])
def test_single_subdomain_is_the_pseudo_inverse(oracle_mod, precond):
    # a connected 2-factor design is one subdomain with uniform weights: the
    # exact local solve makes the preconditioner G^+ on range(G)
    rng = np.random.default_rng(5)
    n = 300
    cats = np.column_stack([rng.integers(0, 40, n), rng.integers(0, 30, n)])
    w = rng.uniform(0.5, 1.5, n)
    o = exact_oracle(oracle_mod, cats, w, precond)
    assert o.n_subdomains == 1
    assert np.all(o.subdomain(0)["pou"] == 1.0)
    G = gramian(cats, w)
    r = G @ rng.standard_normal(len(G))
    z = o.precond_apply(r)
    z_ref = np.linalg.pinv(G, rcond=1e-12) @ r
    assert np.linalg.norm(z - z_ref) <= 1e-10 * np.linalg.norm(z_ref)
    assert np.allclose(G @ z, r, atol=1e-10 * np.linalg.norm(r))
