"""within_amd -- MI355X-native projection path of py-econometrics/within.

Solves the fixed-effects normal equations D'WDx = D'Wy and returns the
demeaned residual y - Dx with hand-written gfx950 kernels behind the
reference's API (``solve`` / ``solve_batch`` / ``Solver`` and the config
classes of ``within._within``).  See ``include/within_amd.h`` for the C ABI.
"""

from ._within import (  # noqa: F401
    AdditiveSchwarz,
    AdditiveSchwarzDiagnostics,
    ApproxCholConfig,
    ApproxSchurConfig,
    BatchSolveResult,
    CG,
    FePreconditioner,
    GMRES,
    LSMR,
    MultiplicativeSchwarz,
    OperatorRepr,
    Preconditioner,
    ProblemBatch,
    ReductionStrategy,
    SchurComplement,
    SolveResult,
    Solver,
    solve,
    solve_batch,
    solve_many,
)

__all__ = [
    "OperatorRepr", "Preconditioner", "SolveResult", "BatchSolveResult", "CG", "GMRES", "LSMR",
    "FePreconditioner", "Solver", "solve", "solve_batch", "AdditiveSchwarz", "AdditiveSchwarzDiagnostics",
    "MultiplicativeSchwarz", "ReductionStrategy", "ApproxCholConfig", "ApproxSchurConfig", "SchurComplement",
    "ProblemBatch", "solve_many",
]
