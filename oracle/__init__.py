"""ctypes front end of the FP64 CPU oracle.

TEST INFRASTRUCTURE ONLY.  The oracle is a CPU restatement of within's
projection path (D'WDx = D'Wz) used as the checker in tests/, by
``__graft_entry__.smoke()`` and as the timed CPU baseline in ``bench.py``.
``within_amd`` never imports it.

Parity status per function is listed in ``oracle/oracle.hpp``; in short, the
approximate-Cholesky / clique-tree values restate GKS 2023 because the
reference's ``approx-chol 0.1.0`` crate is absent, so they are pinned by
mathematical checks (unbiasedness, exactness on degree <= 2 eliminations,
Laplacian invariants) rather than by reference vectors.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


def build() -> None:
    """Compile the oracle with its Makefile (g++, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


class _Params(C.Structure):
    _fields_ = [
        ("krylov", C.c_int),
        ("restart", C.c_int64),
        ("explicit_op", C.c_int),
        ("tol", C.c_double),
        ("maxiter", C.c_int64),
        ("max_refinements", C.c_int64),
        ("precond", C.c_int),
        ("ac_seed", C.c_uint64),
        ("ac_split", C.c_int),
        ("approx_schur", C.c_int),
        ("schur_seed", C.c_uint64),
        ("schur_split", C.c_int),
        ("dense_threshold", C.c_int),
        ("dense_core", C.c_int64),
        ("local_size", C.c_int64),
    ]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P = C.c_void_p
        L.oracle_last_error.restype = C.c_char_p
        L.oracle_solver_new.argtypes = [P, C.c_int64, C.c_int, C.c_int64, P, C.POINTER(_Params), C.POINTER(C.c_void_p)]
        L.oracle_solver_free.argtypes = [P]
        L.oracle_n_dofs.argtypes = [P]
        L.oracle_n_dofs.restype = C.c_int64
        L.oracle_time_setup.argtypes = [P]
        L.oracle_time_setup.restype = C.c_double
        L.oracle_solve_batch.argtypes = [P, P, C.c_int, C.c_int64, P, P, P, P, P, C.c_int]
        for nm in ("oracle_matvec_d", "oracle_rmatvec_wdt", "oracle_gram_apply", "oracle_precond_apply"):
            getattr(L, nm).argtypes = [P, P, P]
        L.oracle_gramian_diag.argtypes = [P, P]
        L.oracle_gramian_nnz.argtypes = [P]
        L.oracle_gramian_nnz.restype = C.c_int64
        L.oracle_gramian_export.argtypes = [P, P, P, P]
        L.oracle_n_subdomains.argtypes = [P]
        L.oracle_n_subdomains.restype = C.c_int64
        L.oracle_subdomain_info.argtypes = [P, C.c_int64, P]
        L.oracle_subdomain_gidx.argtypes = [P, C.c_int64, P, P]
        L.oracle_subdomain_ac.argtypes = [P, C.c_int64, P, P, P, P, P, P, P, P]
        L.oracle_local_solve.argtypes = [P, C.c_int64, P, P]
        L.oracle_subdomain_schur_nnz.argtypes = [P, C.c_int64]
        L.oracle_subdomain_schur_nnz.restype = C.c_int64
        L.oracle_subdomain_schur.argtypes = [P, C.c_int64, P, P, P]
        L.oracle_clique_tree.argtypes = [P, P, C.c_int64, C.c_double, C.c_uint64, C.c_int, P, P, P, C.c_int64]
        L.oracle_clique_tree.restype = C.c_int64
        L.oracle_ac_solve.argtypes = [C.c_int64, P, P, P, C.c_uint64, C.c_int, C.c_int64, P]
        L.oracle_ac_solve.restype = C.c_int64
        L.oracle_schur_dense.argtypes = [P, C.c_int64, C.c_int64, P, P, C.c_int, C.c_uint64, C.c_int, P, P]
        L.oracle_time_solve_batch.argtypes = [P, P, C.c_int, C.c_int64, C.c_int, P]
        L.oracle_time_solve_batch.restype = C.c_double
        L.oracle_restrict.argtypes = [P, P, C.c_int64, P, P]
        L.oracle_prolongate_add.argtypes = [P, P, C.c_int64, P, P]
        L.oracle_residual_update.argtypes = [C.c_int64, P, P, P, P, P, C.c_int64, P]
        L.oracle_lsmr_dense.argtypes = [C.c_int64, C.c_int64, P, P, P, C.c_double, C.c_int64, C.c_int64, P, P]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class OracleError(ValueError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


class OracleSolver:
    """FP64 CPU solver restating within's ``Solver`` (solver.rs:82-447).

    precond: 0 none, 1 additive Schwarz, 2 multiplicative Schwarz.
    krylov:  0 CG, 1 GMRES(restart), 2 LSMR (window ``local_size``, 0 = None).
    Local solver defaults follow ``LocalSolverConfig::default()`` (config.rs:96-104).
    """

    def __init__(self, categories, weights=None, *, krylov=0, restart=30, explicit_op=False,
                 tol=1e-8, maxiter=1000, max_refinements=2, precond=1, ac_seed=0, ac_split=1,
                 approx_schur=True, schur_seed=0, schur_split=1, dense_threshold=24,
                 dense_core=16384, local_size=0):
        L = lib()
        cats = np.asfortranarray(np.asarray(categories, dtype=np.uint32))
        if cats.ndim != 2:
            raise ValueError("categories must be 2-D")
        self.n_obs, self.n_factors = cats.shape
        self._cats = cats
        self._w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        p = _Params(krylov, restart, int(explicit_op), tol, maxiter, max_refinements, precond,
                    ac_seed, ac_split, int(approx_schur), schur_seed, schur_split, dense_threshold,
                    dense_core, int(local_size or 0))
        h = C.c_void_p()
        rc = L.oracle_solver_new(_ptr(cats), self.n_obs, self.n_factors, max(self.n_obs, 1),
                                 _ptr(self._w), C.byref(p), C.byref(h))
        if rc != 0:
            raise OracleError(rc, L.oracle_last_error().decode())
        self._h = h
        self.n_dofs = L.oracle_n_dofs(h)
        self.time_setup = L.oracle_time_setup(h)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().oracle_solver_free(self._h)
            self._h = None

    def solve_batch(self, Y, threads=0):
        Y = np.asfortranarray(np.asarray(Y, dtype=np.float64))
        if Y.ndim == 1:
            Y = Y[:, None]
        k = Y.shape[1]
        X = np.zeros((self.n_dofs, k), order="F")
        D = np.zeros((self.n_obs, k), order="F")
        conv = np.zeros(k, dtype=np.int32)
        its = np.zeros(k, dtype=np.int64)
        res = np.zeros(k)
        rc = lib().oracle_solve_batch(self._h, _ptr(Y), k, self.n_obs, _ptr(X), _ptr(D), _ptr(conv),
                                      _ptr(its), _ptr(res), threads)
        if rc != 0:
            raise OracleError(rc, lib().oracle_last_error().decode())
        return {"x": X, "demeaned": D, "converged": conv.astype(bool), "iterations": its, "residual": res}

    def solve(self, y):
        r = self.solve_batch(np.asarray(y)[:, None])
        return {"x": r["x"][:, 0], "demeaned": r["demeaned"][:, 0], "converged": bool(r["converged"][0]),
                "iterations": int(r["iterations"][0]), "residual": float(r["residual"][0])}

    def time_solve_batch(self, Y, threads=0):
        Y = np.asfortranarray(np.asarray(Y, dtype=np.float64))
        k = Y.shape[1]
        conv = np.zeros(k, dtype=np.int32)
        t = lib().oracle_time_solve_batch(self._h, _ptr(Y), k, self.n_obs, threads, _ptr(conv))
        return t, conv.astype(bool)

    def _vec_op(self, name, x, n_out):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(n_out)
        rc = getattr(lib(), name)(self._h, _ptr(x), _ptr(y))
        if rc != 0:
            raise OracleError(rc, lib().oracle_last_error().decode())
        return y

    def matvec_d(self, x):
        return self._vec_op("oracle_matvec_d", x, self.n_obs)

    def rmatvec_wdt(self, r):
        return self._vec_op("oracle_rmatvec_wdt", r, self.n_dofs)

    def gram_apply(self, x):
        return self._vec_op("oracle_gram_apply", x, self.n_dofs)

    def precond_apply(self, r):
        return self._vec_op("oracle_precond_apply", r, self.n_dofs)

    def gramian_diagonal(self):
        out = np.zeros(self.n_dofs)
        lib().oracle_gramian_diag(self._h, _ptr(out))
        return out

    def explicit_gramian(self):
        L = lib()
        nnz = L.oracle_gramian_nnz(self._h)
        indptr = np.zeros(self.n_dofs + 1, dtype=np.uint32)
        indices = np.zeros(nnz, dtype=np.uint32)
        data = np.zeros(nnz)
        L.oracle_gramian_export(self._h, _ptr(indptr), _ptr(indices), _ptr(data))
        return indptr, indices, data

    @property
    def n_subdomains(self):
        return lib().oracle_n_subdomains(self._h)

    def subdomain(self, i):
        info = np.zeros(12, dtype=np.int64)
        lib().oracle_subdomain_info(self._h, i, _ptr(info))
        keys = ["q", "r", "n_q", "n_r", "eliminate_q", "dense", "nnz_c", "ac_rounds", "ac_nnz",
                "schur_nnz", "has_pou", "core"]
        d = dict(zip(keys, (int(v) for v in info)))
        nl = d["n_q"] + d["n_r"]
        g = np.zeros(nl, dtype=np.uint32)
        w = np.zeros(nl)
        lib().oracle_subdomain_gidx(self._h, i, _ptr(g), _ptr(w))
        d["gidx"], d["pou"] = g, w
        return d

    def subdomain_ac(self, i):
        d = self.subdomain(i)
        if d["dense"]:
            return None
        n = d["n_r"] if d["eliminate_q"] else d["n_q"]
        piv = np.zeros(n)
        cp = np.zeros(n + 1, dtype=np.uint32)
        ci = np.zeros(d["ac_nnz"], dtype=np.uint32)
        cw = np.zeros(d["ac_nnz"])
        rp = np.zeros(d["ac_rounds"] + 1, dtype=np.uint32)
        order = np.zeros(n, dtype=np.uint32)
        core = np.zeros(d["core"], dtype=np.uint32)
        core_inv = np.zeros(d["core"] ** 2)
        lib().oracle_subdomain_ac(self._h, i, _ptr(piv), _ptr(cp), _ptr(ci), _ptr(cw), _ptr(rp), _ptr(order),
                                  _ptr(core), _ptr(core_inv))
        return {"pivot": piv, "col_ptr": cp, "col_idx": ci, "col_w": cw, "round_ptr": rp, "order": order,
                "core": core, "core_inv": core_inv.reshape(d["core"], d["core"])}

    def subdomain_schur(self, i):
        d = self.subdomain(i)
        n = d["n_r"] if d["eliminate_q"] else d["n_q"]
        nnz = lib().oracle_subdomain_schur_nnz(self._h, i)
        ip = np.zeros(n + 1, dtype=np.uint32)
        ix = np.zeros(nnz, dtype=np.uint32)
        dt = np.zeros(nnz)
        lib().oracle_subdomain_schur(self._h, i, _ptr(ip), _ptr(ix), _ptr(dt))
        return ip, ix, dt

    def local_solve(self, i, rhs):
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        out = np.zeros_like(rhs)
        lib().oracle_local_solve(self._h, i, _ptr(rhs), _ptr(out))
        return out


def clique_tree(idx, w, diag, seed, split=1):
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    w = np.ascontiguousarray(w, dtype=np.float64)
    cap = max(1, len(idx) * max(split, 1))
    lo = np.zeros(cap, dtype=np.uint32)
    hi = np.zeros(cap, dtype=np.uint32)
    wo = np.zeros(cap)
    c = lib().oracle_clique_tree(_ptr(idx), _ptr(w), len(idx), diag, seed, split, _ptr(lo), _ptr(hi), _ptr(wo), cap)
    return lo[:c], hi[:c], wo[:c]


def ac_solve(indptr, indices, data, b, seed=0, split=1, dense_core=0):
    """AC-factor the Laplacian CSR and return (x, rounds) with L x ~= b."""
    indptr = np.ascontiguousarray(indptr, dtype=np.uint32)
    indices = np.ascontiguousarray(indices, dtype=np.uint32)
    data = np.ascontiguousarray(data, dtype=np.float64)
    x = np.array(b, dtype=np.float64)
    rounds = lib().oracle_ac_solve(len(indptr) - 1, _ptr(indptr), _ptr(indices), _ptr(data), seed, split,
                                   dense_core, _ptr(x))
    return x, rounds


def schur_dense(c_dense, diag_q, diag_r, approx=False, seed=0, split=1):
    c = np.ascontiguousarray(c_dense, dtype=np.float64)
    nq, nr = c.shape
    nk = min(nq, nr)
    S = np.zeros((nk, nk))
    inv = np.zeros(max(nq, nr))
    dq = np.ascontiguousarray(diag_q, dtype=np.float64)
    dr = np.ascontiguousarray(diag_r, dtype=np.float64)
    rc = lib().oracle_schur_dense(_ptr(c), nq, nr, _ptr(dq), _ptr(dr), int(approx), seed, split, _ptr(S), _ptr(inv))
    if rc < 0:
        raise OracleError(-rc, lib().oracle_last_error().decode())
    return S, bool(rc), inv


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def restrict(gidx, w, global_vec):
    """local[k] = w[k] * global[gidx[k]] (w None = uniform)."""
    g, wv, gv = _u32(gidx), _f64(w), _f64(global_vec)
    out = np.zeros(len(g))
    lib().oracle_restrict(_ptr(g), _ptr(wv), len(g), _ptr(gv), _ptr(out))
    return out


def prolongate_add(gidx, w, local, out):
    """out[gidx[k]] += w[k] * local[k], in place on a float64 array."""
    g, wv, lv = _u32(gidx), _f64(w), _f64(local)
    assert out.dtype == np.float64 and out.flags.c_contiguous
    lib().oracle_prolongate_add(_ptr(g), _ptr(wv), len(g), _ptr(lv), _ptr(out))
    return out


def residual_update(indptr, indices, data, gidx, delta, r_work):
    """r_work -= G[:, gidx] delta through the rows of the symmetric CSR G, in place."""
    ip, ix, dt = _u32(indptr), _u32(indices), _f64(data)
    g, dv = _u32(gidx), _f64(delta)
    assert r_work.dtype == np.float64 and r_work.flags.c_contiguous
    lib().oracle_residual_update(len(ip) - 1, _ptr(ip), _ptr(ix), _ptr(dt), _ptr(g), _ptr(dv), len(g), _ptr(r_work))
    return r_work


LSMR_STOPS = ("zero_rhs", "atb_zero", "residual_tol", "normal_tol", "breakdown", "maxiter")


def lsmr_dense(a, b, minv=None, tol=1e-10, maxiter=100, local_size=None):
    """LSMR (lsmr.rs:122-316) on a dense matrix; ``minv`` is a diagonal M^-1."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    rows, cols = a.shape
    mv = None if minv is None else np.ascontiguousarray(minv, dtype=np.float64)
    x = np.zeros(cols)
    info = np.zeros(4)
    L = lib()
    rc = L.oracle_lsmr_dense(rows, cols, _ptr(a), _ptr(b), _ptr(mv), tol, maxiter, int(local_size or 0),
                             _ptr(x), _ptr(info))
    if rc != 0:
        raise OracleError(rc, L.oracle_last_error().decode())
    return dict(x=x, converged=bool(info[0]), iterations=int(info[1]), residual_norm=float(info[2]),
                stop=LSMR_STOPS[int(info[3])])
