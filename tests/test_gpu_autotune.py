"""The operator-schedule autotuner (wa_solver_autotune / Solver.autotune).

It times candidate schedule shapes for the explicit Gramian SpMM and the
implicit factor passes and keeps the fastest.  The chosen shapes must be
among the candidates, never slower than the default by more than timing
noise, and the solver's results afterwards must still match the oracle
(1e-8 on the demeaned residuals) with the explicit and the implicit operator.
"""

import numpy as np
import pytest

import within_amd as W
from within_amd import synth

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_autotune_keeps_results(oracle_mod):
    p = synth.akm_panel(300_000, 30_000, 3_000, 15, seed=12)
    codes = p["codes"]
    Y = np.asfortranarray(np.random.default_rng(0).standard_normal((codes.shape[0], 8)))
    s = W.Solver(codes, config=W.CG(tol=1e-10, operator=W.OperatorRepr.Explicit))
    before = s.solve_batch(Y)
    t = s.autotune(k=8, reps=3)
    print(t)
    assert t["gram_chunk"] in (256, 512, 1024, 2048, 4096) and t["gram_short"] in (4, 8, 16, 32)
    assert t["pass_chunk"] in (256, 512, 1024, 2048, 4096) and t["pass_short"] in (4, 8, 16, 32)
    assert t["gram_us_best"] <= t["gram_us_default"] and t["pass_us_best"] <= t["pass_us_default"]
    after = s.solve_batch(Y)
    ro = oracle_mod.OracleSolver(codes, None, ac_split=2, tol=1e-10).solve_batch(Y)
    for j in range(8):
        assert after.converged[j]
        assert rel(after.demeaned[:, j], before.demeaned[:, j]) <= 1e-8
        assert rel(after.demeaned[:, j], ro["demeaned"][:, j]) <= 1e-8
    s.set_config(W.CG(tol=1e-10, operator=W.OperatorRepr.Implicit, precision="mixed"))
    imp = s.solve_batch(Y)
    for j in range(8):
        assert imp.converged[j] and rel(imp.demeaned[:, j], ro["demeaned"][:, j]) <= 1e-8
