"""coolpdlp_hip -- MI355X-native PDLP/PDHG engine with CoolPDLP.jl's API.

Exports mirror src/CoolPDLP.jl:53-63: the matrix plug-ins, MILP and its counters,
PrimalDualSolution, preprocess / initialize / solve / solve! (solve_inplace) / PDHG / PDLP /
Algorithm, is_feasible, objective_value.
All computation runs in libcoolpdlp_hip.so (hand-written HIP kernels for gfx950).
"""
from ._lib import ArgumentError, DimensionMismatch, HIPError
from .algorithms import (DUAL_INFEASIBLE, ITERATION_LIMIT, NUMERICAL_ERROR, OPTIMAL,
                         PRIMAL_INFEASIBLE, STILL_RUNNING, TIME_LIMIT, Algorithm,
                         ConvergenceStats, Engine, KKTErrors, LoopState, PDHG, PDLP,
                         TerminationStatus, absolute, initialize, preprocess, relative, solve,
                         solve_inplace)
from .batch import BatchSolver
from .matrices import GPUSparseMatrixCSR, GPUSparseMatrixELL
from .problems import (MILP, PrimalDualSolution, is_feasible, nbcons, nbcons_eq, nbcons_ineq,
                       nbvar, nbvar_cont, nbvar_int, objective_value)
from . import moi
from .moi import Optimizer
from .qps import QPSData, read_mps, write_mps
from . import permutation

__all__ = [
    "GPUSparseMatrixCSR", "GPUSparseMatrixELL", "MILP", "nbvar", "nbvar_int", "nbvar_cont",
    "nbcons", "nbcons_eq", "nbcons_ineq", "PrimalDualSolution", "preprocess", "initialize",
    "solve", "solve_inplace", "PDHG", "PDLP",
    "Algorithm", "is_feasible", "objective_value", "TerminationStatus", "OPTIMAL",
    "TIME_LIMIT", "ITERATION_LIMIT", "STILL_RUNNING", "PRIMAL_INFEASIBLE", "DUAL_INFEASIBLE",
    "NUMERICAL_ERROR", "KKTErrors", "ConvergenceStats",
    "Engine", "BatchSolver", "LoopState", "relative", "absolute", "DimensionMismatch", "ArgumentError", "HIPError",
    "Optimizer", "moi", "QPSData", "read_mps", "write_mps",
]


def version() -> str:
    from ._lib import lib
    return lib.cpdlp_version().decode()
