"""ctypes binding of ``libwithin_amd.so`` (the C ABI in ``include/within_amd.h``).

There is no CPU fallback: if the library is missing the import fails loudly.
"""

from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libwithin_amd.so")

WA_OK = 0
STATUS_NAMES = {
    1: "EmptyObservations", 2: "ObservationCountMismatch", 3: "WeightCountMismatch", 4: "Overflow",
    5: "SingularDiagonal", 6: "LocalSolverBuild", 7: "PreconditionerBuild", 8: "IterativeSolve",
    9: "InvalidArgument", 10: "Device", 11: "Unsupported",
}


class LocalConfig(C.Structure):
    _fields_ = [
        ("ac_seed", C.c_uint64),
        ("ac_split", C.c_int32),
        ("approx_schur", C.c_int32),
        ("schur_seed", C.c_uint64),
        ("schur_split", C.c_int32),
        ("dense_threshold", C.c_int32),
        ("dense_core", C.c_int64),
    ]


class Params(C.Structure):
    _fields_ = [
        ("krylov", C.c_int32),
        ("restart", C.c_int32),
        ("operator_repr", C.c_int32),
        ("preconditioner", C.c_int32),
        ("tol", C.c_double),
        ("maxiter", C.c_int64),
        ("max_refinements", C.c_int64),
        ("precision", C.c_int32),
        ("block", C.c_int32),
        ("host_only", C.c_int32),
        ("local_size", C.c_int32),
        ("local", LocalConfig),
    ]


class RhsStats(C.Structure):
    _fields_ = [
        ("converged", C.c_int32),
        ("iterations", C.c_int64),
        ("residual", C.c_double),
        ("time_solve", C.c_double),
    ]


class CoarseParams(C.Structure):
    _fields_ = [
        ("steps", C.c_int32),
        ("sample_iters", C.c_int32),
        ("seed", C.c_uint64),
        ("lr", C.c_double),
        ("force", C.c_int32),
        ("reserved", C.c_int32),
        ("ridge", C.c_double),
    ]


class CoarseReport(C.Structure):
    _fields_ = [
        ("steps", C.c_int32),
        ("enabled", C.c_int32),
        ("loss_first", C.c_double),
        ("loss_last", C.c_double),
        ("iters_plain", C.c_double),
        ("iters_coarse", C.c_double),
        ("ms_plain", C.c_double),
        ("ms_coarse", C.c_double),
        ("seconds", C.c_double),
        ("kept", C.c_int32),
        ("reserved", C.c_int32),
    ]


GNN_PARAMS = 3 * (2 * 32 * 32 + 32)  # kGnnParams (coarse.hpp)

EXPORTS = [
    "wa_last_error", "wa_params_default", "wa_design_create", "wa_design_destroy", "wa_design_n_dofs",
    "wa_design_n_obs", "wa_design_n_factors", "wa_design_n_levels", "wa_solver_create", "wa_solver_destroy", "wa_solver_set_params",
    "wa_solve_batch", "wa_solve_batch_device", "wa_gram_apply", "wa_gram_apply_explicit", "wa_precond_apply",
    "wa_matvec_d", "wa_rmatvec_wdt", "wa_solver_info", "wa_solver_subdomain", "wa_solver_subdomain_factor",
    "wa_timing_enable", "wa_timing_reset", "wa_timing_get", "wa_solver_precond_bytes",
    "wa_solver_create_with_precond", "wa_precond_create", "wa_solver_params", "wa_precond_work",
    "wa_gram_apply_device", "wa_gram_apply_explicit_device", "wa_precond_apply_device", "wa_matvec_d_device",
    "wa_rmatvec_wdt_device", "wa_gram_apply_noalloc", "wa_multi_create", "wa_multi_destroy", "wa_multi_info",
    "wa_multi_solve", "wa_solver_autotune", "wa_coarse_params_default", "wa_coarse_train", "wa_coarse_enable",
    "wa_coarse_eval", "wa_coarse_export", "wa_coarse_precond_apply",
]

_lib = None


def load():
    """Load the native library (built in-tree by ``within_amd.build``)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"within_amd native library not found at {LIB_PATH}; build it with "
            "`python -m within_amd.build` (hipcc, gfx950)")
    try:
        # PyTorch ships its own HIP runtime.  When this library initialises the
        # device before torch is imported, torch later finds no GPU; importing
        # torch first lets both share the device (torch is optional here).
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    i64, i32 = C.c_int64, C.c_int32
    L.wa_last_error.restype = C.c_char_p
    L.wa_params_default.argtypes = [C.POINTER(Params)]
    L.wa_params_default.restype = None
    L.wa_design_create.argtypes = [P, i64, i32, i64, P, C.POINTER(P)]
    L.wa_design_destroy.argtypes = [P]
    L.wa_design_destroy.restype = None
    for nm in ("wa_design_n_dofs", "wa_design_n_obs"):
        getattr(L, nm).argtypes = [P]
        getattr(L, nm).restype = i64
    L.wa_design_n_factors.argtypes = [P]
    L.wa_design_n_factors.restype = i32
    L.wa_design_n_levels.argtypes = [P, i32]
    L.wa_design_n_levels.restype = i64
    L.wa_solver_create.argtypes = [P, C.POINTER(Params), C.POINTER(P)]
    L.wa_solver_set_params.argtypes = [P, C.POINTER(Params)]
    L.wa_solver_destroy.argtypes = [P]
    L.wa_solver_destroy.restype = None
    for nm in ("wa_solve_batch", "wa_solve_batch_device"):
        getattr(L, nm).argtypes = [P, P, i32, i64, P, i64, P, i64, P]
    for nm in ("wa_gram_apply", "wa_gram_apply_explicit", "wa_precond_apply", "wa_gram_apply_device",
               "wa_gram_apply_explicit_device", "wa_precond_apply_device", "wa_gram_apply_noalloc"):
        getattr(L, nm).argtypes = [P, P, P, i32, i64]
    for nm in ("wa_matvec_d", "wa_rmatvec_wdt", "wa_matvec_d_device", "wa_rmatvec_wdt_device"):
        getattr(L, nm).argtypes = [P, P, i64, P, i64, i32]
    L.wa_solver_info.argtypes = [P, P]
    L.wa_solver_subdomain.argtypes = [P, i64, P, P, P]
    L.wa_solver_subdomain_factor.argtypes = [P, i64, P, P, P, P, P, P, P, P, P]
    L.wa_timing_enable.argtypes = [P, i32]
    L.wa_timing_enable.restype = None
    L.wa_timing_reset.argtypes = [P]
    L.wa_timing_reset.restype = None
    L.wa_timing_get.argtypes = [P, i32, P, P]
    L.wa_solver_precond_bytes.argtypes = [P, P, C.POINTER(i64)]
    L.wa_solver_create_with_precond.argtypes = [P, C.POINTER(Params), P, i64, C.POINTER(P)]
    L.wa_precond_create.argtypes = [P, i64, i32, C.POINTER(P)]
    L.wa_solver_params.argtypes = [P, C.POINTER(Params)]
    L.wa_precond_work.argtypes = [P, P, P]
    L.wa_multi_create.argtypes = [P, i32, C.POINTER(Params), C.POINTER(P)]
    L.wa_multi_destroy.argtypes = [P]
    L.wa_multi_destroy.restype = None
    L.wa_multi_info.argtypes = [P, i32, P]
    L.wa_multi_solve.argtypes = [P, P, P, P, P, P, P, P, P]
    L.wa_solver_autotune.argtypes = [P, i32, i32, P]
    L.wa_coarse_params_default.argtypes = [C.POINTER(CoarseParams)]
    L.wa_coarse_params_default.restype = None
    L.wa_coarse_train.argtypes = [P, C.POINTER(CoarseParams), C.POINTER(CoarseReport)]
    L.wa_coarse_enable.argtypes = [P, i32]
    L.wa_coarse_eval.argtypes = [P, P, P, P]
    L.wa_coarse_export.argtypes = [P, i32, P, i64, C.POINTER(i64)]
    L.wa_coarse_precond_apply.argtypes = [P, P, P, i32, i64]
    _lib = L
    return L


def last_error() -> str:
    return load().wa_last_error().decode(errors="replace")


class NativeError(ValueError):
    """A failing C-ABI call; ``code`` is the wa_status value."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


def check(rc: int) -> None:
    if rc != WA_OK:
        raise NativeError(rc, last_error())


def ptr(a):
    return None if a is None else C.c_void_p(a.ctypes.data)
