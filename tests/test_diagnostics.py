"""Preconditioner scheduling metrics (AdditiveSchwarzDiagnostics,
subdomain_inner_parallel_work; schwarz/additive/planning.rs:140-212,
operator/local_solver.rs:348-356, within-py/src/lib.rs:197-258, 872-876),
recomputed here from the design itself: a subdomain's work is 2 nnz(C) +
n_local when its larger block exceeds 10000 rows, with nnz(C) the number of
distinct (q-level, r-level) pairs observed inside the subdomain."""
import ctypes as C

import numpy as np
import pytest

import within_amd as W
from within_amd import _native as N
from within_amd import synth


@pytest.fixture(scope="module")
def big():
    p = synth.akm_panel(150000, 24000, 300, 6, seed=9)
    return p["codes"]


def expected_work(s, codes):
    L = N.load()
    out = []
    scatter = 0
    for i in range(int(s.info()[0])):
        meta = np.zeros(8, dtype=np.int64)
        N.check(L.wa_solver_subdomain(s._h, i, N.ptr(meta), None, None))
        q, r, n_q, n_r = (int(v) for v in meta[:4])
        g = np.zeros(n_q + n_r, dtype=np.uint32)
        N.check(L.wa_solver_subdomain(s._h, i, N.ptr(meta), N.ptr(g), None))
        off_q = int(codes[:, :q].max(axis=0).sum() + q) if q else 0
        lv_q = g[:n_q].astype(np.int64) - off_q
        inside = np.isin(codes[:, q], lv_q)
        nnz = len(np.unique(codes[inside, q].astype(np.int64) * (1 << 32) + codes[inside, r]))
        out.append(2 * nnz + n_q + n_r if max(n_q, n_r) > 10000 else 0)
        scatter += n_q + n_r
    return out, scatter


def test_inner_work_and_diagnostics_match_the_design(big):
    s = W.Solver(big, _host_only=True)
    p = s.preconditioner()
    work, scatter = expected_work(s, big)
    assert p.subdomain_inner_parallel_work == work
    assert max(work) > 0  # the worker-firm block is past the 10000-row threshold
    d = p.additive_schwarz_diagnostics()
    assert d.total_inner_parallel_work == sum(work)
    assert d.max_inner_parallel_work == max(work)
    assert d.total_scatter_dofs == scatter
    assert d.outer_parallel_capacity == pytest.approx(sum(work) / max(work))
    assert d.scatter_overlap == pytest.approx(scatter / s.n_dofs)
    assert d.reduction_strategy == W.ReductionStrategy.Auto
    assert d.resolved_reduction_strategy == W.ReductionStrategy.ParallelReduction
    assert "AdditiveSchwarzDiagnostics(reduction_strategy=Auto" in repr(d)
    # configured strategy is reported back; the bytes keep the metrics
    s2 = W.Solver(big, preconditioner=W.AdditiveSchwarz(reduction=W.ReductionStrategy.AtomicScatter),
                  _host_only=True)
    assert s2.preconditioner().additive_schwarz_diagnostics().reduction_strategy == \
        W.ReductionStrategy.AtomicScatter
    q = W.FePreconditioner(p.to_bytes(), _host_only=True)
    assert q.subdomain_inner_parallel_work == work
    assert q.additive_schwarz_diagnostics().total_inner_parallel_work == sum(work)


def test_small_problem_has_no_inner_work():
    codes = synth.readme_problem(3000)[0]
    p = W.Solver(codes, _host_only=True).preconditioner()
    assert p.subdomain_inner_parallel_work == [0] * p.n_subdomains
    d = p.additive_schwarz_diagnostics()
    assert d.max_inner_parallel_work == 0 and d.outer_parallel_capacity == 0.0


def test_multiplicative_has_no_additive_diagnostics(big):
    s = W.Solver(np.asfortranarray(big[:20000]), config=W.GMRES(), preconditioner=W.MultiplicativeSchwarz(), _host_only=True)
    assert s.preconditioner().additive_schwarz_diagnostics() is None
