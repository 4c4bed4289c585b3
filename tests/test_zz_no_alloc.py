"""The allocation-free implicit operator apply (wa_gram_apply_noalloc).

Its scratch is a 64 MiB device array of the library, shared by every call on
one GPU, instead of the solver's work vectors.  Two calls in a row -- the
second on a smaller problem and with fewer columns -- must equal the existing
entry point (wa_gram_apply_device, same kernels) bit for bit, and the oracle
to 1e-12.  A design needing more scratch per column than the array holds is
rejected with an error.
"""

import ctypes as C

import numpy as np
import pytest
import torch

import within_amd as W
from within_amd import _native as N
from within_amd import synth

pytestmark = pytest.mark.gpu


def dptr(t):
    return C.c_void_p(t.data_ptr())


def apply_both(s, X, k, ld):
    L = N.load()
    dX = torch.from_numpy(np.ascontiguousarray(X.T)).cuda()
    a = torch.full((k, ld), 3.0, dtype=torch.float64, device="cuda")
    b = torch.full((k, ld), 3.0, dtype=torch.float64, device="cuda")
    N.check(L.wa_gram_apply_noalloc(s._h, dptr(dX), dptr(a), k, ld))
    N.check(L.wa_gram_apply_device(s._h, dptr(dX), dptr(b), k, ld))
    return a.cpu().numpy().T, b.cpu().numpy().T


def test_noalloc_gram_apply_matches_device_entry_point(oracle_mod):
    big = synth.akm_panel(60000, 6000, 700, 12, n_industries=3, seed=21, weights="lognormal")
    small = synth.akm_panel(5000, 500, 60, 6, seed=22)
    rng = np.random.default_rng(0)
    for p, k, w in ((big, 40, big["weights"]), (small, 3, None)):
        s = W.Solver(p["codes"], weights=w, preconditioner=W.Preconditioner.Off)
        m = s.n_dofs
        ld = m + 3
        X = np.asfortranarray(rng.standard_normal((ld, k)))
        a, b = apply_both(s, X, k, ld)
        assert np.array_equal(a[:m], b[:m]), (m, k)
        assert np.all(a[m:] == 3.0)
        o = oracle_mod.OracleSolver(p["codes"], w, precond=0)
        ref = np.column_stack([o.gram_apply(X[:m, j]) for j in range(min(k, 4))])
        err = np.linalg.norm(a[:m, :ref.shape[1]] - ref) / np.linalg.norm(ref)
        assert err <= 1e-12, err


def test_noalloc_rejects_designs_beyond_its_scratch():
    n = 4_400_000  # 4.4M worker levels: 2 x 8 x 4.4M bytes per column exceed 64 MiB
    codes = np.asfortranarray(np.column_stack([np.arange(n), np.arange(n) % 10]).astype(np.uint32))
    s = W.Solver(codes, preconditioner=W.Preconditioner.Off)
    L = N.load()
    X = torch.zeros((1, s.n_dofs), dtype=torch.float64, device="cuda")
    Y = torch.zeros((1, s.n_dofs), dtype=torch.float64, device="cuda")
    rc = L.wa_gram_apply_noalloc(s._h, dptr(X), dptr(Y), 1, s.n_dofs)
    assert rc == 9
    assert "scratch" in N.last_error()
