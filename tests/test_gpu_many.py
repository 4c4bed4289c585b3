"""Many independent problems solved together (ProblemBatch / wa_multi_*).

Bars:
* every problem's demeaned residuals within 1e-8 (relative) of the FP64
  oracle's own solve of that problem alone (tol 1e-10), with converged flags;
* the batch equals a Solver per problem (FP64 CG, tol 1e-10) to 1e-8, and its
  per-RHS iteration counts are those of the separate solves up to one
  iteration (the recurrences are the same; only summation orders differ);
* a zero right-hand side converges in 0 iterations with x = 0;
* the 256-problem workload (n log-uniform in [10k, 2M], 2 or 3 factors with
  n/20, n/200 and 25 levels, 1-4 RHS): every RHS converged, its reported
  residual recomputed from the returned demeaned vector (D'W(y - Dx), numpy),
  demeaned = y - D x, and the oracle on every problem of at most 200k
  observations.
"""

import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import within_amd as W
from within_amd import synth

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def small_problems():
    probs = synth.many_problems(n_problems=14, n_min=2_000, n_max=60_000, seed=7)
    rng = np.random.default_rng(3)
    out = []
    for i, (codes, Y) in enumerate(probs):
        w = rng.uniform(0.5, 2.0, codes.shape[0]) if i % 5 == 2 else None
        out.append((codes, Y, w))
    # an AKM-shaped problem (Zipf workers, mobile firms) and a zero right-hand side
    p = synth.akm_panel(30_000, 3_000, 400, 10, seed=9)
    Y = np.asfortranarray(np.column_stack([rng.standard_normal(30_000), np.zeros(30_000)]))
    out.append((p["codes"], Y, None))
    return out


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_matches_oracle_and_separate_solves(precision, oracle_mod):
    probs = small_problems()
    cfg = W.CG(tol=1e-10, max_refinements=3, precision=precision)
    batch = W.ProblemBatch([c for c, _, _ in probs], config=cfg, weights=[w for _, _, w in probs])
    assert batch.info(0)[15] + batch.info(1)[15] == len(probs)  # two stacks: 2- and 3-factor problems
    res = batch.solve([Y for _, Y, _ in probs])
    # the per-problem CG iterations replayed from a captured graph equal the
    # same iterations enqueued launch by launch, bit for bit
    os.environ["WITHIN_AMD_GRAPHS"] = "0"
    try:
        eager = batch.solve([Y for _, Y, _ in probs])
    finally:
        del os.environ["WITHIN_AMD_GRAPHS"]
    for r, e in zip(res, eager):
        assert np.array_equal(r.demeaned, e.demeaned) and r.iterations == e.iterations
    for i, ((codes, Y, w), r) in enumerate(zip(probs, res)):
        ro = oracle_mod.OracleSolver(codes, w, ac_split=2, tol=1e-10, max_refinements=3).solve_batch(Y)
        sep = W.Solver(codes, config=W.CG(tol=1e-10, max_refinements=3), weights=w).solve_batch(Y)
        for j in range(Y.shape[1]):
            assert r.converged[j], (i, j)
            if not np.any(Y[:, j]):
                assert r.iterations[j] == 0 and not np.any(r.x[:, j])
                continue
            assert rel(r.demeaned[:, j], ro["demeaned"][:, j]) <= 1e-8, (i, j)
            assert rel(r.demeaned[:, j], sep.demeaned[:, j]) <= 1e-8, (i, j)
            assert r.residual[j] <= 1.01e-10
            if precision == "fp64":
                assert abs(r.iterations[j] - sep.iterations[j]) <= 1, (i, j, r.iterations[j], sep.iterations[j])


def dx(codes, x):
    levels = codes.max(axis=0).astype(np.int64) + 1
    off = np.concatenate([[0], np.cumsum(levels)[:-1]])
    out = np.zeros((codes.shape[0], x.shape[1]))
    for q in range(codes.shape[1]):
        out += x[off[q] + codes[:, q].astype(np.int64)]
    return out


def dtw(codes, r):
    levels = codes.max(axis=0).astype(np.int64) + 1
    return np.concatenate([np.bincount(codes[:, q], r, levels[q]) for q in range(codes.shape[1])])


def test_many_uneven_problems_256(oracle_mod):
    probs = synth.many_problems(n_problems=256, n_min=10_000, n_max=2_000_000, seed=0)
    tol = 1e-9
    batch = W.ProblemBatch([c for c, _ in probs], config=W.CG(tol=tol, precision="mixed",
                                                              operator=W.OperatorRepr.Explicit))
    print(f"256 problems, {sum(c.shape[0] for c, _ in probs)} observations: setup {batch.time_setup:.1f}s",
          flush=True)
    res = batch.solve([Y for _, Y in probs])
    print(f"solve {res[0].time_total:.2f}s", flush=True)
    small = []
    for i, ((codes, Y), r) in enumerate(zip(probs, res)):
        assert all(r.converged), i
        assert rel(Y - dx(codes, r.x), r.demeaned) <= 1e-12, i
        for j in range(Y.shape[1]):
            want = np.linalg.norm(dtw(codes, r.demeaned[:, j])) / np.linalg.norm(dtw(codes, Y[:, j]))
            assert want <= 1.01 * tol, (i, j, want)
            assert abs(r.residual[j] - want) <= 1e-13 + 1e-3 * want, (i, j)
        if codes.shape[0] <= 200_000:
            small.append(i)

    def check(i):
        codes, Y = probs[i]
        ro = oracle_mod.OracleSolver(codes, None, ac_split=2, tol=1e-10).solve_batch(Y, threads=2)
        return [rel(res[i].demeaned[:, j], ro["demeaned"][:, j]) for j in range(Y.shape[1])] + [
            float(all(ro["converged"]))]

    with ThreadPoolExecutor(max_workers=max(1, min(8, len(os.sched_getaffinity(0)) // 2))) as ex:
        outs = list(ex.map(check, small))
    print(f"oracle checked {len(small)} problems of <= 200k observations", flush=True)
    for i, o in zip(small, outs):
        assert o[-1] == 1.0, i
        assert max(o[:-1]) <= 1e-8, (i, o)
