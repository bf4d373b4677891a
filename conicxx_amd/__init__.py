"""MI355X-native conic interior-point solver (gfx950 HIP kernels behind a C-ABI).

The public Python surface mirrors the reference's Solver / Settings /
ConeSpec / Solution interface; see conicxx_amd.solver.
"""
from .solver import (ConeSpec, FactorPrecision, Info, LinearSolverBackend, Settings, Solution,  # noqa: F401
                     Solver, Status, analyze_pattern, equilibrate, toString)
from .batch import BatchSolver  # noqa: F401,E402
