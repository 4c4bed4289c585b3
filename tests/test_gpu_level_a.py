"""Level A operator drop-ins on device pointers and on more than 32 columns.

The ``*_device`` entry points take gfx950 device pointers and allocate
nothing; the host-pointer entry points stage through the solver's own device
buffer.  Both split k > 32 columns into blocks of 32.  Checked against the
FP64 oracle (1e-12 for the operators, 1e-10 for the preconditioner, as the
single-block tests) and the device variants against the host variants bit
for bit, with leading dimensions larger than the row counts.
"""

import ctypes as C

import numpy as np
import pytest

import within_amd as W
from within_amd import _native as N
from within_amd import synth

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def dptr(t):
    return C.c_void_p(t.data_ptr())


def padded(rng, rows, k, pad):
    """Column-major rows x k block inside a (rows + pad) x k buffer."""
    buf = np.asfortranarray(rng.standard_normal((rows + pad, k)))
    return buf


@pytest.mark.parametrize("k", [1, 40, 70])
def test_level_a_blocks_and_device_variants(k, oracle_mod):
    import torch

    p = synth.akm_panel(12000, 1200, 150, 9, n_industries=4, seed=8, weights="lognormal")
    codes, w = p["codes"], p["weights"]
    o = oracle_mod.OracleSolver(codes, w, explicit_op=True)
    s = W.Solver(codes, weights=w, config=W.CG(operator=W.OperatorRepr.Explicit),
                 preconditioner=W.AdditiveSchwarz(W.SchurComplement(W.ApproxCholConfig(0, 1),
                                                                    W.ApproxSchurConfig(0, 1))))
    L = N.load()
    m, n = s.n_dofs, s.n_obs
    rng = np.random.default_rng(k)
    pad = 5
    ldm, ldn = m + pad, n + pad
    X = padded(rng, m, k, pad)
    R = padded(rng, n, k, pad)
    dX, dR = torch.from_numpy(X.T.copy()).cuda(), torch.from_numpy(R.T.copy()).cuda()  # rows = columns

    def host_dof(fn, src):
        out = np.full((ldm, k), 7.0, order="F")
        N.check(getattr(L, fn)(s._h, N.ptr(src), N.ptr(out), k, ldm))
        return out

    def dev_dof(fn, src):
        out = torch.full((k, ldm), 7.0, dtype=torch.float64, device="cuda")
        N.check(getattr(L, fn)(s._h, dptr(src), dptr(out), k, ldm))
        return out.cpu().numpy().T

    for fn, ref_fn, bar in (("wa_gram_apply", o.gram_apply, 1e-12), ("wa_gram_apply_explicit", o.gram_apply, 1e-12),
                            ("wa_precond_apply", o.precond_apply, 1e-10)):
        Yh = host_dof(fn, X)
        Yd = dev_dof(fn + "_device", dX)
        ref = np.column_stack([ref_fn(X[:m, j]) for j in range(k)])
        assert rel(Yh[:m], ref) <= bar, (fn, rel(Yh[:m], ref))
        assert np.array_equal(Yh[:m], Yd[:m]), fn
        assert np.all(Yh[m:] == 7.0) and np.all(Yd[m:] == 7.0), fn  # padding rows untouched

    Dh = np.full((ldn, k), 7.0, order="F")
    N.check(L.wa_matvec_d(s._h, N.ptr(X), ldm, N.ptr(Dh), ldn, k))
    Dd = torch.full((k, ldn), 7.0, dtype=torch.float64, device="cuda")
    N.check(L.wa_matvec_d_device(s._h, dptr(dX), ldm, dptr(Dd), ldn, k))
    refd = np.column_stack([o.matvec_d(X[:m, j]) for j in range(k)])
    assert rel(Dh[:n], refd) <= 1e-14
    assert np.array_equal(Dh[:n], Dd.cpu().numpy().T[:n]) and np.all(Dh[n:] == 7.0)

    Bh = np.full((ldm, k), 7.0, order="F")
    N.check(L.wa_rmatvec_wdt(s._h, N.ptr(R), ldn, N.ptr(Bh), ldm, k))
    Bd = torch.full((k, ldm), 7.0, dtype=torch.float64, device="cuda")
    N.check(L.wa_rmatvec_wdt_device(s._h, dptr(dR), ldn, dptr(Bd), ldm, k))
    refb = np.column_stack([o.rmatvec_wdt(R[:n, j]) for j in range(k)])
    assert rel(Bh[:m], refb) <= 1e-12
    assert np.array_equal(Bh[:m], Bd.cpu().numpy().T[:m]) and np.all(Bh[m:] == 7.0)


def test_level_a_rejects_bad_arguments():
    codes, _, _ = synth.uniform_problem(500, [20, 10], seed=3)
    s = W.Solver(codes)
    L = N.load()
    m = s.n_dofs
    X = np.zeros((m, 2), order="F")
    Y = np.zeros((m, 2), order="F")
    assert L.wa_gram_apply(s._h, N.ptr(X), N.ptr(Y), 2, m - 1) == 9  # leading dimension below the rows
    assert L.wa_gram_apply(s._h, N.ptr(X), N.ptr(Y), -1, m) == 9
    assert L.wa_gram_apply(s._h, N.ptr(X), N.ptr(Y), 0, m) == 0  # nothing to do
    assert L.wa_gram_apply_device(s._h, None, N.ptr(Y), 2, m) == 9
