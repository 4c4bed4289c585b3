"""Pins of the oracle's learned coarse-space network (oracle/gnn.py), CPU only.

The reference has no learned component, so the oracle is pinned by
mathematics: its gradient against central finite differences (every
operation in FP64), the loss against its meaning, and the two-level apply
against dense linear algebra.  Plausible mistakes planted in a copy of the
gradient code each fail the finite-difference check.
"""

import inspect

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import gnn


def _graph(seed=0, n=3000, levels=(80, 25, 6)):
    """A small 3-factor Gramian G = D'D (dense-free, scipy CSR) and its diagonal."""
    rng = np.random.default_rng(seed)
    cols = [rng.integers(0, L, n) for L in levels]
    off = np.concatenate([[0], np.cumsum(levels)])
    rows = np.repeat(np.arange(n), len(levels))
    idx = np.concatenate([c[:, None] + o for c, o in zip(cols, off[:-1])], axis=1).ravel()
    D = sp.csr_matrix((np.ones(len(idx)), (rows, idx)), shape=(n, off[-1]))
    G = (D.T @ D).tocsr()
    factor = np.repeat(np.arange(len(levels)), levels)
    return G, np.asarray(G.diagonal()), factor, len(levels)


def _problem(seed=0, scale=1.0):
    G, d, factor, Q = _graph(seed)
    m = G.shape[0]
    rng = np.random.default_rng(seed + 1)
    X0 = gnn.features(d, factor, np.arange(m), seed)
    dinv = 1.0 / np.sqrt(d)
    E = rng.standard_normal((m, gnn.K))
    R = G @ E
    nu = float(np.sum(E * R))
    a = np.sqrt(6.0 / 96.0)
    theta = scale * rng.uniform(-a, a, gnn.N_PARAMS)
    theta[[l * gnn.LAYER_PARAMS + 2 * gnn.MAT + j for l in range(gnn.LAYERS) for j in range(gnn.K)]] = (
        0.1 * rng.standard_normal(gnn.LAYERS * gnn.K))
    return G, X0, dinv, Q, E, R, nu, theta


def _fd_errors(fn, theta, args, coords, h=1e-5):
    loss, g = fn(theta, *args)
    errs = []
    for i in coords:
        tp = theta.copy()
        tm = theta.copy()
        tp[i] += h
        tm[i] -= h
        fd = (fn(tp, *args, grad=False)[0] - fn(tm, *args, grad=False)[0]) / (2 * h)
        errs.append(abs(fd - g[i]) / max(abs(fd), abs(g[i]), 1e-12))
    rng = np.random.default_rng(7)
    u = rng.standard_normal(len(theta))
    u /= np.linalg.norm(u)
    fd = (fn(theta + h * u, *args, grad=False)[0] - fn(theta - h * u, *args, grad=False)[0]) / (2 * h)
    errs.append(abs(fd - g @ u) / max(abs(fd), 1e-12))
    return np.array(errs), loss


def _exact(fn):
    def f(theta, G, X0, dinv, Q, E, R, nu, grad=True):
        return fn(theta, X0, G, dinv, Q, E, R, nu, eps=0.0, mimic=False, grad=grad)
    return f


def _coords(theta):
    # a few weights of every matrix and bias of every layer
    rng = np.random.default_rng(3)
    out = []
    for l in range(gnn.LAYERS):
        o = l * gnn.LAYER_PARAMS
        out += list(o + rng.choice(gnn.MAT, 4, replace=False))
        out += list(o + gnn.MAT + rng.choice(gnn.MAT, 4, replace=False))
        out += list(o + 2 * gnn.MAT + rng.choice(gnn.K, 2, replace=False))
    return out


def test_gradient_matches_finite_differences():
    G, X0, dinv, Q, E, R, nu, theta = _problem()
    errs, loss = _fd_errors(_exact(gnn.loss_grad), theta, (G, X0, dinv, Q, E, R, nu), _coords(theta))
    assert -1.0 <= loss <= 0.0
    assert errs.max() < 1e-5, errs


# (description, text in loss_grad, planted replacement)
PLANTS = [
    ("drop the GP V V' term of dL/dP", "(R - GP @ V) @ V.T", "R @ V.T"),
    ("Wa instead of Wa' in the aggregation branch", "r(dz) @ r(Wa.T)", "r(dz) @ r(Wa)"),
    ("Ws instead of Ws' in the direct branch", "r(dz) @ r(Ws.T)", "r(dz) @ r(Ws)"),
    ("no ReLU mask", "dz = dh * (H[l] > 0.0)", "dz = dh"),
    ("unscaled aggregation input of dWa", "r(c[:, None] * T[l]).T", "r(T[l]).T"),
    ("no D^-1/2 before the transposed aggregation", "ys = dinv[:, None] * (", "ys = ("),
    ("wrong sign of dL/dP", "return loss, -(2.0 / nu)", "return loss, (2.0 / nu)"),
]


@pytest.mark.parametrize("desc,old,new", PLANTS, ids=[p[0] for p in PLANTS])
def test_planted_gradient_mistakes_fail_the_check(desc, old, new):
    src = inspect.getsource(gnn)
    assert src.count(old) == 1
    ns = {"__name__": "planted_gnn"}
    exec(compile(src.replace(old, new), "<planted>", "exec"), ns)
    G, X0, dinv, Q, E, R, nu, theta = _problem()
    errs, _ = _fd_errors(_exact(ns["loss_grad"]), theta, (G, X0, dinv, Q, E, R, nu), _coords(theta))
    assert errs.max() > 1e-3, (desc, errs.max())


def test_loss_is_minus_the_captured_energy_share():
    G, X0, dinv, Q, E, R, nu, theta = _problem()
    # span(P) holds every sample: the projection removes all of their G-energy
    Ac = E.T @ (G @ E)
    V, loss = gnn.coarse_solve(Ac, E.T @ R, nu, 0.0)
    assert abs(loss + 1.0) < 1e-10
    np.testing.assert_allclose(V, np.eye(gnn.K), atol=1e-8)
    # P G-orthogonal to the samples: nothing is captured
    rng = np.random.default_rng(5)
    P = rng.standard_normal(E.shape)
    P -= E @ np.linalg.solve(E.T @ R, R.T @ P)  # P' G E = 0
    _, loss0 = gnn.coarse_solve(P.T @ (G @ P), P.T @ R, nu, 0.0)
    assert abs(loss0) < 1e-10
    # a single coarse vector equal to one sample captures that sample's share
    P1 = np.zeros_like(E)
    P1[:, 0] = E[:, 0]
    P1[:, 1:] = P[:, 1:]
    Ac1 = P1.T @ (G @ P1)
    _, l1 = gnn.coarse_solve(Ac1, P1.T @ R, nu, 0.0)
    share = float(E[:, 0] @ R[:, 0]) / nu
    assert l1 <= -share + 1e-10


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -9, -2.5, 3.0e38], dtype=np.float32)
    got = gnn.bf16(x)
    want = [1.0, 1.0, 1.0 + 2.0 ** -6, 1.0, -2.5]
    np.testing.assert_allclose(got[:5], want, rtol=0, atol=0)
    # 3e38 = 1.7632 * 2^127 rounds to 226/128 * 2^127
    assert got[5] == 226.0 / 128.0 * 2.0 ** 127
    # every bf16 value is kept exactly
    y = np.array([1.5, -0.0078125, 65280.0], dtype=np.float32)
    np.testing.assert_array_equal(gnn.bf16(y), y.astype(np.float64))


def test_two_level_apply_against_dense_algebra():
    G, X0, dinv, Q, E, R, nu, theta = _problem()
    Gd = G.toarray()
    m = Gd.shape[0]
    rng = np.random.default_rng(9)
    P = rng.standard_normal((m, gnn.K))
    GP = Gd @ P
    Ainv = gnn.sym_pinv(P.T @ GP)
    r = Gd @ rng.standard_normal(m)  # a consistent right-hand side
    Gplus = np.linalg.pinv(Gd)
    # with the exact one-level solve, the two-level apply is still a solve
    z = gnn.two_level_apply(P, GP, Ainv, lambda v: Gplus @ v, r)
    np.testing.assert_allclose(Gd @ z, r, rtol=0, atol=1e-8 * np.linalg.norm(r))
    # with any one-level M, P'G z = P'r (the coarse residual vanishes)
    Mdiag = 1.0 / np.diag(Gd)
    z2 = gnn.two_level_apply(P, GP, Ainv, lambda v: Mdiag * v, r)
    np.testing.assert_allclose(P.T @ (Gd @ z2), P.T @ r, rtol=1e-9, atol=1e-9 * np.linalg.norm(P.T @ r))
    # and it is symmetric: <s, z(r)> = <r, z(s)>
    s = Gd @ rng.standard_normal(m)
    zs = gnn.two_level_apply(P, GP, Ainv, lambda v: Mdiag * v, s)
    assert abs(s @ z2 - r @ zs) < 1e-9 * abs(s @ z2)
    # the pseudo-inverse drops null directions
    Pn = np.concatenate([P[:, :31], P[:, :1]], axis=1)
    An = gnn.sym_pinv(Pn.T @ (Gd @ Pn))
    np.testing.assert_allclose(An @ (Pn.T @ Gd @ Pn) @ An, An, atol=1e-8 * np.abs(An).max())


def test_features_definition():
    G, d, factor, Q = _graph()
    F = gnn.features(d, factor, np.arange(len(d)), 11)
    z = np.log(d)
    z = (z - z.mean()) / z.std()
    np.testing.assert_allclose(F[:, 1], z, rtol=1e-6, atol=1e-6)
    assert np.all(F[:, 0] == 1.0)
    np.testing.assert_array_equal(F[:, 2:6].sum(axis=1), 1.0)
    np.testing.assert_array_equal(np.argmax(F[:, 2:6], axis=1), factor & 3)
    r = F[:, 7:].ravel()
    assert abs(r.mean()) < 0.02 and abs(r.std() - 1.0) < 0.02
    # keyed by the user DOF: a permuted numbering permutes the rows
    perm = np.random.default_rng(0).permutation(len(d))
    F2 = gnn.features(d[perm], factor[perm], perm, 11)
    np.testing.assert_allclose(F2[:, 7:], F[perm, 7:], rtol=0, atol=0)
