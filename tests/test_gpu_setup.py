"""GPU-side setup: dense cores of at least 2048 vertices are inverted on the
GPU with the host factorisation's per-entry arithmetic, so the factor a
device solver uploads is the host-only solver's (and the oracle's) bit for
bit (operator/local_solver.rs:207-300)."""
import ctypes as C

import numpy as np
import pytest

import within_amd as W
from within_amd import _native as N
from within_amd import synth

pytestmark = pytest.mark.gpu

PREC = W.AdditiveSchwarz(W.SchurComplement(W.ApproxCholConfig(0, 1), W.ApproxSchurConfig(0, 1)))


def factor(s, i):
    L = N.load()
    meta = np.zeros(8, dtype=np.int64)
    N.check(L.wa_solver_subdomain(s._h, i, N.ptr(meta), None, None))
    nnz = C.c_int64()
    N.check(L.wa_solver_subdomain_factor(s._h, i, C.byref(nnz), None, None, None, None, None, None, None, None))
    return meta, nnz.value


def arrays(s, i, n, nnz, n_rounds, n_core):
    L = N.load()
    a = dict(pivot=np.zeros(n), col_ptr=np.zeros(n + 1, np.uint32), col_idx=np.zeros(nnz, np.uint32),
             col_w=np.zeros(nnz), round_ptr=np.zeros(n_rounds + 1, np.uint32), order=np.zeros(n - n_core, np.uint32),
             core=np.zeros(n_core, np.uint32), core_inv=np.zeros(n_core * n_core))
    N.check(L.wa_solver_subdomain_factor(s._h, i, C.byref(C.c_int64()), *[N.ptr(v) for v in a.values()]))
    return a


@pytest.mark.parametrize("n,workers,firms,dense_core", [(150000, 15000, 3000, 0), (300000, 30000, 12000, 6000)],
                         ids=["core-only", "rounds-then-core"])
def test_device_dense_core_is_bitwise_the_host(n, workers, firms, dense_core):
    codes = synth.akm_panel(n, workers, firms, 8, seed=23)["codes"]
    h = W.Solver(codes, preconditioner=PREC, _host_only=True, _dense_core=dense_core)
    d = W.Solver(codes, preconditioner=PREC, _dense_core=dense_core)
    biggest = 0
    for i in range(int(h.info()[0])):
        mh, nh = factor(h, i)
        md, nd = factor(d, i)
        assert mh.tolist() == md.tolist() and nh == nd
        if mh[5]:  # dense subdomain: its factor is the dense inverse of the reduced system
            continue
        n = int(mh[3] if mh[4] else mh[2])
        ah = arrays(h, i, n, nh, int(mh[6]), int(mh[7]))
        ad = arrays(d, i, n, nd, int(md[6]), int(md[7]))
        for k in ah:
            assert np.array_equal(ah[k], ad[k]), (i, k)
        biggest = max(biggest, int(mh[7]))
    assert biggest >= 2048
    # and the device solver solves with it
    Y = np.asfortranarray(np.random.default_rng(0).standard_normal((codes.shape[0], 2)))
    r = d.solve_batch(Y)
    assert all(r.converged)
