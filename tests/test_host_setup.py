"""CPU checks of the native library: every C-ABI symbol is exported, the
host-side setup (subdomains, partition weights, Schur complements and the
reduced-system factors) reproduces the oracle's exactly, and the Python
surface validates configs the way within._within does.  No GPU calls.
"""

import ctypes as C
import os
import re
import warnings

import numpy as np
import pytest

import within_amd as W
from within_amd import _native as N
from within_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_every_declared_symbol_is_exported():
    hdr = open(os.path.join(ROOT, "include", "within_amd.h")).read()
    declared = set(re.findall(r"\b(wa_[a-z0-9_]+)\s*\(", hdr))
    assert len(declared) >= 20
    lib = C.CDLL(N.LIB_PATH)
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing
    assert declared == set(N.EXPORTS)


def test_params_default_matches_reference_defaults():
    p = N.Params()
    N.load().wa_params_default(C.byref(p))
    # SolverParams::default() (config.rs:227-237), LocalSolverConfig::default() (:96-104)
    assert (p.krylov, p.operator_repr, p.tol, p.maxiter, p.max_refinements) == (0, 0, 1e-8, 1000, 2)
    assert (p.local.ac_seed, p.local.ac_split, p.local.approx_schur, p.local.schur_split,
            p.local.dense_threshold) == (0, 1, 1, 1, 24)


def host_solver(codes, w=None, **kw):
    return W.Solver(codes, weights=w, _host_only=True, **kw)


CASES = [
    ("readme", lambda: (synth.readme_problem(20000)[0], None)),
    ("akm", lambda: (synth.akm_panel(30000, 3000, 400, 15, seed=3)["codes"], None)),
    ("akm-weighted-4fe", lambda: (lambda p: (p["codes"], p["weights"]))(
        synth.akm_panel(20000, 2000, 300, 10, n_industries=4, seed=5, weights="lognormal"))),
]
CONFIGS = [
    ("rust-default", dict(ac_split=1), W.AdditiveSchwarz(W.SchurComplement(W.ApproxCholConfig(0, 1),
                                                                            W.ApproxSchurConfig(0, 1)))),
    ("python-default", dict(ac_split=2), None),
    ("exact-schur", dict(ac_split=1, approx_schur=False), W.AdditiveSchwarz(W.SchurComplement())),
    ("seeded-split", dict(ac_seed=7, ac_split=3, schur_seed=11, schur_split=2),
     W.AdditiveSchwarz(W.SchurComplement(W.ApproxCholConfig(7, 3), W.ApproxSchurConfig(11, 2)))),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("cfg", CONFIGS, ids=[c[0] for c in CONFIGS])
def test_host_setup_reproduces_oracle(case, cfg, oracle_mod):
    codes, w = case[1]()
    okw, prec = cfg[1], cfg[2]
    o = oracle_mod.OracleSolver(codes, w, **okw)
    s = host_solver(codes, w, preconditioner=prec)
    L = N.load()
    assert int(s.info()[0]) == o.n_subdomains
    for i in range(o.n_subdomains):
        od = o.subdomain(i)
        meta = np.zeros(8, dtype=np.int64)
        g = np.zeros(od["n_q"] + od["n_r"], dtype=np.uint32)
        pw = np.zeros(len(g))
        N.check(L.wa_solver_subdomain(s._h, i, N.ptr(meta), N.ptr(g), N.ptr(pw)))
        assert meta[:6].tolist() == [od["q"], od["r"], od["n_q"], od["n_r"], od["eliminate_q"], od["dense"]]
        assert np.array_equal(g, od["gidx"]) and np.array_equal(pw, od["pou"])
        if od["dense"]:
            continue
        oa = o.subdomain_ac(i)
        n = len(oa["pivot"])
        nnz = C.c_int64()
        arrs = dict(pivot=np.zeros(n), col_ptr=np.zeros(n + 1, np.uint32), col_idx=np.zeros(len(oa["col_idx"]), np.uint32),
                    col_w=np.zeros(len(oa["col_w"])), round_ptr=np.zeros(len(oa["round_ptr"]), np.uint32),
                    order=np.zeros(len(oa["order"]), np.uint32), core=np.zeros(len(oa["core"]), np.uint32),
                    core_inv=np.zeros(oa["core_inv"].size))
        N.check(L.wa_solver_subdomain_factor(s._h, i, C.byref(nnz), *[N.ptr(a) for a in arrs.values()]))
        assert nnz.value == len(oa["col_idx"])
        for k, a in arrs.items():
            assert np.array_equal(a, np.ravel(oa[k])), (i, k)


def test_device_calls_fail_loudly_on_host_only_solver():
    codes = synth.readme_problem(2000)[0]
    s = host_solver(codes)
    with pytest.raises(N.NativeError) as e:
        s.solve(np.ones(2000))
    assert e.value.code == 10  # WA_ERR_DEVICE


def test_config_validation_mirrors_within_py():
    codes = synth.readme_problem(500)[0]
    # CG / LSMR with multiplicative Schwarz (within-py/src/lib.rs:639-656)
    with pytest.raises(ValueError, match="symmetric"):
        host_solver(codes, config=W.CG(), preconditioner=W.MultiplicativeSchwarz())
    with pytest.raises(ValueError, match="LSMR requires a symmetric"):
        host_solver(codes, config=W.LSMR(), preconditioner=W.MultiplicativeSchwarz())
    with pytest.raises(ValueError):
        W.LSMR(local_size=-1)
    s = host_solver(codes, config=W.LSMR(local_size=5))
    p = N.Params()
    N.check(N.load().wa_solver_params(s._h, C.byref(p)))
    assert (p.krylov, p.local_size, p.operator_repr, p.max_refinements) == (2, 5, 0, 0)
    with pytest.raises(TypeError):
        host_solver(codes, config="cg")
    with pytest.raises(TypeError):
        host_solver(codes, preconditioner="additive")
    with pytest.raises(TypeError):
        W.AdditiveSchwarz(local_solver=3)
    with pytest.raises(ValueError):
        W.ApproxCholConfig(split=0)
    # GMRES + multiplicative is accepted and builds the explicit Gramian
    s = host_solver(codes, config=W.GMRES(), preconditioner=W.MultiplicativeSchwarz())
    assert s.info()[8] > 0


def test_input_validation():
    with pytest.raises(ValueError):
        host_solver(np.zeros((0, 2), dtype=np.uint32))
    with pytest.raises(TypeError):
        host_solver(np.zeros((10, 2), dtype=np.int64))
    with pytest.raises(ValueError):
        host_solver(np.zeros(10, dtype=np.uint32))
    with pytest.raises(ValueError):
        host_solver(np.zeros((10, 2), dtype=np.uint32), w=np.ones(9))
    c = np.ascontiguousarray(np.random.default_rng(0).integers(0, 5, (50, 2)).astype(np.uint32))
    with warnings.catch_warnings(record=True) as rec:
        warnings.simplefilter("always")
        host_solver(c)
    assert any("F-contiguous" in str(r.message) for r in rec)
    # the reference's test is the element stride along the observations
    # (within-py/src/lib.rs:732-744): F-order and single-factor C-order
    # arrays pass silently, a column slice of an F-order array too; a column
    # slice of a C-order array (stride 2) warns
    with pytest.warns(UserWarning, match="F-contiguous"):
        host_solver(c[:, :1])
    for ok in (np.asfortranarray(c), np.ascontiguousarray(c[:, :1]), np.asfortranarray(np.column_stack([c, c]))[:, ::2]):
        with warnings.catch_warnings():
            warnings.simplefilter("error")
            host_solver(ok)


def test_empty_observations_error_code():
    L = N.load()
    h = C.c_void_p()
    rc = L.wa_design_create(None, 0, 2, 1, None, C.byref(h))
    assert rc == 1  # EmptyObservations


def test_zero_weights_with_preconditioner_errors():
    # tests/edge_cases.rs:83-101 -> SingularDiagonal at build
    cats = np.asfortranarray(np.array([[0, 0], [1, 0], [0, 1], [1, 1], [2, 0]], dtype=np.uint32))
    with pytest.raises(N.NativeError) as e:
        host_solver(cats, w=np.zeros(5))
    assert e.value.code == 5


def test_level_schedule_invariants():
    # every reduced vertex is in exactly one forward level; in-edges point to
    # strictly earlier levels, out-edges to later levels or the core
    codes = synth.akm_panel(20000, 2000, 300, 12, seed=3)["codes"]
    s = host_solver(codes)
    info = s.info()
    assert info[3] >= 0 and info[5] > 0


@pytest.mark.parametrize("dense_core", [0, 600])
def test_large_dense_core_is_bitwise_the_oracles(dense_core, oracle_mod):
    # the blocked, vectorised host factorisation of a ~1000-vertex core keeps
    # every entry's operation sequence: bitwise equal to the oracle's loops
    codes = synth.akm_panel(60000, 6000, 1000, 10, seed=8)["codes"]
    okw = dict(ac_split=1)
    if dense_core:
        okw["dense_core"] = dense_core
    o = oracle_mod.OracleSolver(codes, None, **okw)
    prec = W.AdditiveSchwarz(W.SchurComplement(W.ApproxCholConfig(0, 1), W.ApproxSchurConfig(0, 1)))
    s = W.Solver(codes, preconditioner=prec, _host_only=True, _dense_core=dense_core)
    L = N.load()
    big = 0
    for i in range(o.n_subdomains):
        if o.subdomain(i)["dense"]:
            continue
        oa = o.subdomain_ac(i)
        n = len(oa["pivot"])
        nnz = C.c_int64()
        arrs = dict(pivot=np.zeros(n), col_ptr=np.zeros(n + 1, np.uint32), col_idx=np.zeros(len(oa["col_idx"]), np.uint32),
                    col_w=np.zeros(len(oa["col_w"])), round_ptr=np.zeros(len(oa["round_ptr"]), np.uint32),
                    order=np.zeros(len(oa["order"]), np.uint32), core=np.zeros(len(oa["core"]), np.uint32),
                    core_inv=np.zeros(oa["core_inv"].size))
        N.check(L.wa_solver_subdomain_factor(s._h, i, C.byref(nnz), *[N.ptr(a) for a in arrs.values()]))
        for k, a in arrs.items():
            assert np.array_equal(a, np.ravel(oa[k])), (i, k)
        big = max(big, len(oa["core"]))
    assert big >= 512


@pytest.mark.parametrize("dense_core", [1, 1500])
def test_parallel_elimination_rounds_are_bitwise_the_oracles(dense_core, oracle_mod):
    # a reduced system large enough that the rounds eliminate and merge in
    # parallel (>= 1024 picks per round): the same per-entry sums in the same
    # order as the oracle's sequential map-based rounds
    codes = synth.akm_panel(150000, 15000, 6000, 8, seed=21)["codes"]
    o = oracle_mod.OracleSolver(codes, None, ac_split=1, dense_core=dense_core)
    prec = W.AdditiveSchwarz(W.SchurComplement(W.ApproxCholConfig(0, 1), W.ApproxSchurConfig(0, 1)))
    s = W.Solver(codes, preconditioner=prec, _host_only=True, _dense_core=dense_core)
    L = N.load()
    most = 0
    for i in range(o.n_subdomains):
        if o.subdomain(i)["dense"]:
            continue
        oa = o.subdomain_ac(i)
        n = len(oa["pivot"])
        nnz = C.c_int64()
        arrs = dict(pivot=np.zeros(n), col_ptr=np.zeros(n + 1, np.uint32), col_idx=np.zeros(len(oa["col_idx"]), np.uint32),
                    col_w=np.zeros(len(oa["col_w"])), round_ptr=np.zeros(len(oa["round_ptr"]), np.uint32),
                    order=np.zeros(len(oa["order"]), np.uint32), core=np.zeros(len(oa["core"]), np.uint32),
                    core_inv=np.zeros(oa["core_inv"].size))
        N.check(L.wa_solver_subdomain_factor(s._h, i, C.byref(nnz), *[N.ptr(a) for a in arrs.values()]))
        for k, a in arrs.items():
            assert np.array_equal(a, np.ravel(oa[k])), (i, k)
        rp = np.ravel(oa["round_ptr"]).astype(np.int64)
        if len(rp) > 1:
            most = max(most, int(np.diff(rp).max()))
    assert most >= 1024
