"""CG iterations replayed from captured HIP graphs.

The block CG captures one iteration (operator, dot products, scalar steps,
preconditioner -- on two streams where the additive apply overlaps factor
pairs -- and the direction update) and replays it, several iterations per host
check.  Replays must give bitwise the same results and iteration counts as the
same iterations enqueued one launch at a time (WITHIN_AMD_GRAPHS=0), on every
replay, with eager solves in between.  The two-stream apply inside a graph is
checked at full size (tests/test_gpu_fullsize.py), where the additive apply
overlaps the factor pairs.
"""

import os

import numpy as np
import pytest

import within_amd as W
from within_amd import synth

pytestmark = pytest.mark.gpu


def solve(s, Y, graphs):
    old = os.environ.get("WITHIN_AMD_GRAPHS")
    os.environ["WITHIN_AMD_GRAPHS"] = "1" if graphs else "0"
    try:
        return s.solve_batch(Y)
    finally:
        if old is None:
            del os.environ["WITHIN_AMD_GRAPHS"]
        else:
            os.environ["WITHIN_AMD_GRAPHS"] = old


CASES = [
    ("3fe-fp64", dict(n=40_000, workers=4_000, firms=500, years=12), W.CG(tol=1e-10)),
    ("3fe-mixed-explicit", dict(n=40_000, workers=4_000, firms=500, years=12),
     W.CG(tol=1e-9, precision="mixed", operator=W.OperatorRepr.Explicit)),
    ("akm-2m", dict(n=2_000_000, workers=200_000, firms=20_000, years=20),
     W.CG(tol=1e-9, precision="mixed", operator=W.OperatorRepr.Explicit)),
    ("no-preconditioner", dict(n=20_000, workers=2_000, firms=300, years=8), W.CG(tol=1e-9)),
]


@pytest.mark.parametrize("case", range(len(CASES)), ids=[c[0] for c in CASES])
def test_graph_replay_is_bitwise_the_eager_loop(case):
    name, shape, cfg = CASES[case]
    p = synth.akm_panel(shape["n"], shape["workers"], shape["firms"], shape["years"], seed=case)
    prec = W.Preconditioner.Off if name == "no-preconditioner" else W.AdditiveSchwarz(
        W.SchurComplement(W.ApproxCholConfig(0, 1), W.ApproxSchurConfig(0, 1)))
    s = W.Solver(p["codes"], config=cfg, preconditioner=prec)
    Y = np.asfortranarray(np.random.default_rng(case).standard_normal((shape["n"], 5)))
    ref = solve(s, Y, graphs=False)
    assert all(ref.converged)
    for rep in range(3):
        g = solve(s, Y, graphs=True)
        assert np.array_equal(g.demeaned, ref.demeaned), (name, rep)
        assert np.array_equal(g.x, ref.x), (name, rep)
        assert g.iterations == ref.iterations, (name, rep)
        assert g.converged == ref.converged
        e = solve(s, Y, graphs=False)  # eager work between the replays
        assert np.array_equal(e.demeaned, ref.demeaned)
