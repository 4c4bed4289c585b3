"""Serialised preconditioner on the host: bytes round trips, solvers built
from bytes, and malformed input (crates/within-py/src/lib.rs:892-913,
925-982; crates/within/src/solver.rs:135-155).  No GPU: every solver here is
host-only, so these check the format and its validation, not the apply (the
GPU tests compare applies and solves)."""
import ctypes as C
import pickle

import numpy as np
import pytest

import within_amd as W
from within_amd import _native as N
from within_amd import synth

SEEDED = W.AdditiveSchwarz(W.SchurComplement(W.ApproxCholConfig(7, 3), W.ApproxSchurConfig(11, 2)))


def host_solver(codes, w=None, **kw):
    return W.Solver(codes, weights=w, _host_only=True, **kw)


def effective(s):
    p = N.Params()
    N.check(N.load().wa_solver_params(s._h, C.byref(p)))
    loc = p.local
    return (p.preconditioner, loc.ac_seed, loc.ac_split, loc.approx_schur, loc.schur_seed, loc.schur_split,
            loc.dense_threshold, loc.dense_core)


@pytest.fixture(scope="module")
def akm():
    p = synth.akm_panel(20000, 2000, 300, 10, n_industries=4, seed=5, weights="lognormal")
    return p["codes"], p["weights"]


@pytest.mark.parametrize("prec,config", [(SEEDED, None), (None, None),
                                         (W.MultiplicativeSchwarz(), W.GMRES()),
                                         (W.AdditiveSchwarz(W.SchurComplement()), None)],
                         ids=["additive-seeded", "additive-default", "multiplicative", "exact-schur"])
def test_bytes_round_trip_and_solver_from_bytes(akm, prec, config):
    codes, w = akm
    s = host_solver(codes, w, preconditioner=prec, config=config)
    p = s.preconditioner()
    b = p.to_bytes()
    assert isinstance(b, bytes) and len(b) > 64
    # the standalone object re-serialises to the same bytes
    p2 = W.FePreconditioner(b, _host_only=True)
    assert p2.to_bytes() == b
    assert (p2.nrows, p2.ncols, p2.n_subdomains) == (s.n_dofs, s.n_dofs, int(s.info()[0]))
    assert repr(p2).startswith("FePreconditioner(Multiplicative" if prec is not None and
                               isinstance(prec, W.MultiplicativeSchwarz) else "FePreconditioner(Additive")
    # a solver built from the bytes carries the same preconditioner: same
    # bytes out, same setup counts, same effective kind and local config
    s2 = host_solver(codes, w, preconditioner=p2, config=config)
    assert s2.preconditioner().to_bytes() == b
    assert s2.info()[:8].tolist() == s.info()[:8].tolist()
    assert s2.info()[12:14].tolist() == s.info()[12:14].tolist()
    assert effective(s2)[:7] == effective(s)[:7]
    # and a solver built from the original object does too
    s3 = host_solver(codes, w, preconditioner=p, config=config)
    assert s3.preconditioner().to_bytes() == b


def test_precond_overrides_and_fixes_the_preconditioner_settings(akm):
    codes, w = akm
    b = host_solver(codes, w, preconditioner=SEEDED).preconditioner().to_bytes()
    p = W.FePreconditioner(b, _host_only=True)
    # the solver takes kind and local config from the bytes, not from the default
    s = host_solver(codes, w, preconditioner=p)
    e = effective(s)
    assert e[:7] == (1, 7, 3, 1, 11, 2, 24)
    # the Krylov settings can change, the preconditioner stays
    s.set_config(W.CG(tol=1e-6, maxiter=50))
    s.set_config(W.GMRES(restart=20))
    assert effective(s)[:7] == e[:7]


def test_multiplicative_bytes_need_gmres(akm):
    codes, w = akm
    b = host_solver(codes, w, config=W.GMRES(), preconditioner=W.MultiplicativeSchwarz()).preconditioner().to_bytes()
    p = W.FePreconditioner(b, _host_only=True)
    with pytest.raises(ValueError, match="symmetric"):
        host_solver(codes, w, preconditioner=p)
    s = host_solver(codes, w, preconditioner=p, config=W.GMRES())
    assert s.info()[8] > 0  # the Gramian the multiplicative sweeps use is rebuilt from the design


def test_bytes_for_another_design_are_rejected(akm):
    codes, w = akm
    b = host_solver(codes, w).preconditioner().to_bytes()
    other = synth.readme_problem(3000)[0]
    with pytest.raises(ValueError, match="different number of DOFs"):
        host_solver(other, preconditioner=W.FePreconditioner(b, _host_only=True))
    # same DOF count, but a level the preconditioner covers is never observed
    moved = np.array(codes, order="F")
    moved[moved[:, 0] == 5, 0] = 6
    assert moved[:, 0].max() == codes[:, 0].max()
    with pytest.raises(ValueError, match="does not match the design"):
        host_solver(moved, w, preconditioner=W.FePreconditioner(b, _host_only=True))


def test_pickle_stores_the_bytes(akm):
    codes, w = akm
    p = host_solver(codes, w, preconditioner=SEEDED).preconditioner()
    fn, args = p.__reduce__()
    assert fn is W.FePreconditioner and args == (p.to_bytes(),)
    # pickle.dumps goes through __reduce__ (loading builds a device handle,
    # which the GPU tests cover)
    assert pickle.dumps(p) == pickle.dumps(p)


def test_malformed_bytes_raise_value_error(akm):
    codes, w = akm
    b = host_solver(codes, w, preconditioner=SEEDED).preconditioner().to_bytes()
    bad = [b"", b"\0" * 16, b[: len(b) // 2], b[:-1], b + b"\0", b"XXXX" + b[4:],
           b[:4] + bytes([b[4] ^ 0xFF]) + b[5:]]
    for data in bad:
        with pytest.raises(ValueError, match="failed to deserialize"):
            W.FePreconditioner(data, _host_only=True)
    with pytest.raises(TypeError):
        W.FePreconditioner()


def test_corrupted_bytes_never_crash(akm):
    """Random byte flips either fail validation or give a self-consistent
    preconditioner (one whose bytes round-trip); none may crash the
    process."""
    codes, w = akm
    b = host_solver(codes, w, preconditioner=SEEDED).preconditioner().to_bytes()
    rng = np.random.default_rng(0)
    rejected = 0
    for _ in range(300):
        a = bytearray(b)
        for pos in rng.integers(0, len(a), size=int(rng.integers(1, 4))):
            a[pos] ^= int(rng.integers(1, 256))
        try:
            p = W.FePreconditioner(bytes(a), _host_only=True)
        except ValueError:
            rejected += 1
            continue
        assert p.to_bytes() == bytes(a)
    assert rejected > 0


def test_single_factor_preconditioner_has_no_subdomains():
    """One factor: no factor pairs, so the Schwarz preconditioner has no
    subdomains (factor_pairs.rs:63-304); it still exists, serialises and
    reports zero scheduling work."""
    c1 = np.asfortranarray(np.random.default_rng(0).integers(0, 50, (1000, 1)).astype(np.uint32))
    s = host_solver(c1)
    p = s.preconditioner()
    assert p is not None and p.n_subdomains == 0 and p.subdomain_inner_parallel_work == []
    q = W.FePreconditioner(p.to_bytes(), _host_only=True)
    assert q.to_bytes() == p.to_bytes() and q.nrows == 50
    d = q.additive_schwarz_diagnostics()
    assert (d.total_inner_parallel_work, d.total_scatter_dofs, d.outer_parallel_capacity) == (0, 0, 0.0)
    assert host_solver(c1, preconditioner=q).info()[0] == 0
