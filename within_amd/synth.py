"""Seeded synthetic fixed-effects problems (panel layouts) for tests and benchmarks.

The AKM-style generator follows the semantics of the reference's benchmark
helpers (``benchmarks/_problems.py:417-428`` Zipf sizes, ``:431-539``
mobility), written for vectorised generation at tens of millions of rows:

* workers: observation counts drawn with P(rank r) proportional to r^-s over
  ``n_workers`` ranks; worker ids are a random relabelling of the ranks and
  rows come in panel order (all rows of a worker contiguous);
* firms: each worker starts at a uniform random firm and switches to a new
  uniform random firm with probability ``mobility`` before each later row;
* years: uniform in ``[0, n_years)``;
* optional fourth factor ``industry x year`` where each firm has a uniform
  random industry.
"""

from __future__ import annotations

import numpy as np


def readme_problem(n=100_000, seed=1):
    """Config-1 / README inputs: ``np.random.seed(1)`` stream (README.md:23-31, 44-54)."""
    rs = np.random.RandomState(seed)
    fe = np.asfortranarray(np.column_stack([
        rs.randint(0, 500, n).astype(np.uint32),
        rs.randint(0, 200, n).astype(np.uint32),
    ]))
    y = rs.randn(n)
    return fe, y, rs


def readme_fwl(n=100_000, seed=1):
    """README FWL example: returns (fe, Y=[y, X], beta_true).  Known answer
    beta_hat = [0.9982, -2.006, 0.5005] (README.md:44-54)."""
    fe, _y, rs = readme_problem(n, seed)
    beta = np.array([1.0, -2.0, 0.5])
    X = rs.randn(n, 3)
    y = X @ beta + rs.randn(n)
    return fe, np.asfortranarray(np.column_stack([y, X])), beta


def zipf_counts(rng, n_obs, n_levels, s):
    """Draw n_obs level ranks with P(r) ~ (r+1)^-s; returns int64 ranks."""
    w = np.arange(1, n_levels + 1, dtype=np.float64) ** (-s)
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    return np.searchsorted(cdf, rng.random(n_obs), side="right").clip(0, n_levels - 1)


def akm_panel(n_obs, n_workers, n_firms, n_years, *, zipf_s=1.1, mobility=0.05, n_industries=0,
              seed=0, weights=None):
    """AKM panel codes (F-order uint32, shape (n_obs, Q)) plus effect draws.

    Returns dict with ``codes``, ``worker_fe``, ``firm_fe`` and optional ``weights``.
    ``weights='lognormal'`` draws log-normal(0, 0.5) observation weights.
    """
    rng = np.random.default_rng(seed)
    ranks = zipf_counts(rng, n_obs, n_workers, zipf_s)
    relabel = rng.permutation(n_workers).astype(np.uint32)
    worker = np.sort(relabel[ranks])  # panel order: rows of a worker contiguous
    first = np.ones(n_obs, dtype=bool)
    first[1:] = worker[1:] != worker[:-1]
    switch = first | (rng.random(n_obs) < mobility)
    spell = np.cumsum(switch) - 1
    spell_firm = rng.integers(0, n_firms, int(spell[-1]) + 1, dtype=np.uint32)
    firm = spell_firm[spell]
    year = rng.integers(0, n_years, n_obs, dtype=np.uint32)
    cols = [worker, firm, year]
    if n_industries:
        ind = rng.integers(0, n_industries, n_firms, dtype=np.uint32)
        cols.append(ind[firm] * np.uint32(n_years) + year)
    out = {
        "codes": np.asfortranarray(np.column_stack(cols).astype(np.uint32)),
        "worker_fe": rng.standard_normal(n_workers),
        "firm_fe": rng.standard_normal(n_firms),
    }
    if weights == "lognormal":
        out["weights"] = rng.lognormal(0.0, 0.5, n_obs)
    out["rng"] = rng
    return out


def akm_response(panel, k=1, beta_cols=0):
    """y = effects + N(0,1) (k == 1) or Y = [y, X] with X ~ N(0,1) and
    y = X beta + effects + N(0,1) (beta_cols = k - 1 regressors)."""
    rng = panel["rng"]
    c = panel["codes"]
    n = c.shape[0]
    eff = panel["worker_fe"][c[:, 0]] + panel["firm_fe"][c[:, 1]]
    if beta_cols <= 0:
        Y = np.empty((n, k), order="F")
        for j in range(k):
            Y[:, j] = eff + rng.standard_normal(n)
        return Y
    X = rng.standard_normal((n, beta_cols))
    beta = rng.standard_normal(beta_cols)
    y = X @ beta + eff + rng.standard_normal(n)
    return np.asfortranarray(np.column_stack([y, X]))


def uniform_problem(n_obs, levels, seed=0, weights=False):
    """Uniform codes for the given per-factor level counts (config 4 shapes)."""
    rng = np.random.default_rng(seed)
    codes = np.asfortranarray(np.column_stack(
        [rng.integers(0, L, n_obs, dtype=np.uint32) for L in levels]).astype(np.uint32))
    w = rng.lognormal(0.0, 0.5, n_obs) if weights else None
    return codes, w, rng


def many_problems(n_problems=256, n_min=10_000, n_max=2_000_000, seed=0):
    """Config 4: n ~ logU[n_min, n_max], 2 or 3 factors with ~n/20, n/200 and 25
    levels, uniform codes, 1-4 RHS of N(0,1)."""
    rng = np.random.default_rng(seed)
    probs = []
    for _ in range(n_problems):
        n = int(np.exp(rng.uniform(np.log(n_min), np.log(n_max))))
        q = int(rng.integers(2, 4))
        levels = [max(2, n // 20), max(2, n // 200), 25][:q]
        codes = np.asfortranarray(np.column_stack(
            [rng.integers(0, L, n, dtype=np.uint32) for L in levels]).astype(np.uint32))
        k = int(rng.integers(1, 5))
        Y = np.asfortranarray(rng.standard_normal((n, k)))
        probs.append((codes, Y))
    return probs
