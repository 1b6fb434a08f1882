"""MathOptInterface-style optimizer over the engine, mirroring src/moi.jl.

`Optimizer` keeps the reference optimizer's state and contract (src/moi.jl:16-246):

* options: `RawOptimizerAttribute`s are `Algorithm` keyword arguments, plus "algorithm"
  (PDLP / PDHG), "float_type", "int_type" and "matrix_type"; `Silent` maps to
  `show_progress`; `TimeLimitSec` to `time_limit`;
* a small model API stands in for `MOI.copy_to`: variables with bounds (a variable in
  GreaterThan / LessThan / EqualTo / Interval), affine constraints in the same four sets,
  and an affine objective with a sense and a constant;
* `optimize()` builds the MILP (maximisation negates c), solves it, and derives the
  reduced costs, the objective and dual objective values and the MOI statuses exactly as
  src/moi.jl:165-246 does (`finish`, which is plain numpy).

One departure: the reference pops "algorithm", "float_type", "int_type" and
"matrix_type" from the option dictionary on each optimize! (src/moi.jl:194-201), so a
second optimize! falls back to their defaults; here they persist.
"""
from __future__ import annotations

import enum
import warnings
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp

from ._lib import ArgumentError
from .problems import MILP


class TerminationStatusCode(enum.Enum):
    """The MOI termination statuses the reference optimizer reports."""
    OPTIMIZE_NOT_CALLED = 0
    OPTIMAL = 1
    TIME_LIMIT = 2
    ITERATION_LIMIT = 3
    OTHER_ERROR = 4
    # only with the opt-in "detect_infeasibility" attribute
    INFEASIBLE = 5
    DUAL_INFEASIBLE = 6
    NUMERICAL_ERROR = 7


class ResultStatusCode(enum.Enum):
    """The MOI result statuses the reference optimizer reports."""
    NO_SOLUTION = 0
    FEASIBLE_POINT = 1
    UNKNOWN_RESULT_STATUS = 2


class ObjectiveSense(enum.Enum):
    MIN_SENSE = 0
    MAX_SENSE = 1
    FEASIBILITY_SENSE = 2


# ------------------------------------------------------------------ sets and functions
@dataclass(frozen=True)
class EqualTo:
    value: float


@dataclass(frozen=True)
class GreaterThan:
    lower: float


@dataclass(frozen=True)
class LessThan:
    upper: float


@dataclass(frozen=True)
class Interval:
    lower: float
    upper: float


SUPPORTED_SETS = (EqualTo, GreaterThan, LessThan, Interval)


def _bounds(s) -> tuple[float, float]:
    if isinstance(s, EqualTo):
        return float(s.value), float(s.value)
    if isinstance(s, GreaterThan):
        return float(s.lower), np.inf
    if isinstance(s, LessThan):
        return -np.inf, float(s.upper)
    if isinstance(s, Interval):
        return float(s.lower), float(s.upper)
    raise ArgumentError(f"unsupported set {s!r}")


def _set_of(lo: float, up: float):
    """The set whose finite sides are [lo, up] (None when both are infinite)."""
    if lo == up:
        return EqualTo(lo)
    if np.isfinite(lo) and np.isfinite(up):
        return Interval(lo, up)
    if np.isfinite(lo):
        return GreaterThan(lo)
    if np.isfinite(up):
        return LessThan(up)
    return None


@dataclass(frozen=True)
class VariableIndex:
    value: int   # 0-based


@dataclass
class ScalarAffineFunction:
    """sum(coef * x[var]) + constant; terms is a list of (coefficient, VariableIndex)."""
    terms: list = field(default_factory=list)
    constant: float = 0.0


@dataclass(frozen=True)
class ConstraintIndex:
    kind: str        # "variable" (a bound on a variable) or "affine" (a row of A)
    set_type: type
    value: int       # the variable, or the row


# ------------------------------------------------------------------ the helpers of moi.jl
def positive_part(a):
    return np.maximum(a, 0.0)


def negative_part(a):
    return -np.minimum(a, 0.0)


def safeprod_left(left, right):
    """src/utils/linalg.jl:17: `right` where `left` is infinite, else left * right."""
    left = np.asarray(left, dtype=np.float64)
    right = np.asarray(right, dtype=np.float64)
    with np.errstate(invalid="ignore"):
        return np.where(np.isinf(left), right, left * right)


def proj_multiplier(lam, l, u):
    """src/utils/linalg.jl:26-42: the multiplier of l <= x <= u, either bound infinite."""
    lam = np.asarray(lam, dtype=np.float64)
    lmin = np.asarray(l) == -np.inf
    umax = np.asarray(u) == np.inf
    return np.where(lmin, np.where(umax, 0.0, -negative_part(lam)),
                    np.where(umax, positive_part(lam), lam))


_STATUS = {
    "OPTIMAL": (TerminationStatusCode.OPTIMAL, ResultStatusCode.FEASIBLE_POINT,
                ResultStatusCode.FEASIBLE_POINT),
    "TIME_LIMIT": (TerminationStatusCode.TIME_LIMIT, ResultStatusCode.UNKNOWN_RESULT_STATUS,
                   ResultStatusCode.UNKNOWN_RESULT_STATUS),
    "ITERATION_LIMIT": (TerminationStatusCode.ITERATION_LIMIT, ResultStatusCode.UNKNOWN_RESULT_STATUS,
                        ResultStatusCode.UNKNOWN_RESULT_STATUS),
    "STILL_RUNNING": (TerminationStatusCode.OTHER_ERROR, ResultStatusCode.NO_SOLUTION,
                      ResultStatusCode.NO_SOLUTION),
    # the certificates themselves are not returned: no solution is reported
    "PRIMAL_INFEASIBLE": (TerminationStatusCode.INFEASIBLE, ResultStatusCode.NO_SOLUTION,
                          ResultStatusCode.NO_SOLUTION),
    "DUAL_INFEASIBLE": (TerminationStatusCode.DUAL_INFEASIBLE, ResultStatusCode.NO_SOLUTION,
                        ResultStatusCode.NO_SOLUTION),
    "NUMERICAL_ERROR": (TerminationStatusCode.NUMERICAL_ERROR, ResultStatusCode.NO_SOLUTION,
                        ResultStatusCode.NO_SOLUTION),
}


class Optimizer:
    """CoolPDLP.Optimizer (src/moi.jl:16-246): the solver behind a JuMP-like model."""

    def __init__(self):
        self.options: dict = {}
        self.silent = False
        self._clear_model()
        self._clear_results()

    # ---- state (MOI.is_empty / MOI.empty!, src/moi.jl:41-67)
    def _clear_results(self):
        self.x = np.zeros(0)
        self.y = np.zeros(0)
        self.z = np.zeros(0)
        self.obj_value = np.nan
        self.dual_obj_value = np.nan
        self.termination_status = TerminationStatusCode.OPTIMIZE_NOT_CALLED
        self.primal_status = ResultStatusCode.UNKNOWN_RESULT_STATUS
        self.dual_status = ResultStatusCode.UNKNOWN_RESULT_STATUS
        self.solve_time = 0.0

    def _clear_model(self):
        self._lv: list = []
        self._uv: list = []
        self._bound_set: dict = {}   # variable -> set type of its bound constraint
        self._rows: list = []        # (terms, lower, upper, set type); constants moved into the bounds
        self._obj = ScalarAffineFunction()
        self._sense = ObjectiveSense.MIN_SENSE

    def is_empty(self) -> bool:
        return (self.x.size == 0 and self.y.size == 0 and self.z.size == 0
                and np.isnan(self.obj_value) and np.isnan(self.dual_obj_value)
                and self.termination_status == TerminationStatusCode.OPTIMIZE_NOT_CALLED
                and self.primal_status == ResultStatusCode.UNKNOWN_RESULT_STATUS
                and self.dual_status == ResultStatusCode.UNKNOWN_RESULT_STATUS
                and self.solve_time == 0.0 and not self._lv and not self._rows)

    def empty(self):
        self._clear_model()
        self._clear_results()

    # ---- attributes (src/moi.jl:79-106)
    solver_name = "CoolPDLP"

    def set_silent(self, value: bool):
        self.silent = bool(value)

    @property
    def time_limit_sec(self):
        return self.options.get("time_limit")

    @time_limit_sec.setter
    def time_limit_sec(self, value):
        if value is None:
            self.options.pop("time_limit", None)
        else:
            self.options["time_limit"] = float(value)

    def set_raw(self, name: str, value):
        """MOI.RawOptimizerAttribute(name) = value"""
        self.options[str(name)] = value

    def get_raw(self, name: str):
        return self.options[str(name)]

    @property
    def result_count(self) -> int:
        return 0 if self.termination_status == TerminationStatusCode.OPTIMIZE_NOT_CALLED else 1

    @property
    def raw_status_string(self) -> str:
        return self.termination_status.name

    # ---- model
    def add_variable(self) -> VariableIndex:
        self._lv.append(-np.inf)
        self._uv.append(np.inf)
        return VariableIndex(len(self._lv) - 1)

    def add_variables(self, k: int) -> list:
        return [self.add_variable() for _ in range(k)]

    def add_constraint(self, func, s) -> ConstraintIndex:
        if not isinstance(s, SUPPORTED_SETS):
            raise ArgumentError(f"unsupported set {s!r}")
        lo, up = _bounds(s)
        if isinstance(func, VariableIndex):
            j = func.value
            if j in self._bound_set:
                raise ArgumentError(f"variable {j} already has a {self._bound_set[j].__name__} bound")
            self._bound_set[j] = type(s)
            self._lv[j], self._uv[j] = lo, up
            return ConstraintIndex("variable", type(s), j)
        if isinstance(func, ScalarAffineFunction):
            terms = [(float(cf), v.value) for cf, v in func.terms]
            # the function's constant moves into the set (MOI.Utilities' rule)
            k = float(func.constant)
            self._rows.append((terms, lo - k, up - k, type(s)))
            return ConstraintIndex("affine", type(s), len(self._rows) - 1)
        raise ArgumentError(f"unsupported function {func!r}")

    def read_from_file(self, source):
        """Load an MPS model (a path or an open text file), as the tutorial's
        JuMP.read_from_file(path; format = MOI.FileFormats.FORMAT_MPS) followed by
        set_optimizer does (test/tutorial.jl:87-93): the model is emptied, then one variable
        per column with its bounds (a bound constraint of the set its finite sides make),
        one affine constraint per row, and the objective with its constant and sense.
        Integer markers are read and dropped: the optimizer solves the continuous
        relaxation.  Returns the variables in column order."""
        from .qps import read_mps
        q = read_mps(source)
        self.empty()
        xs = self.add_variables(len(q.c))
        for j, v in enumerate(xs):
            s = _set_of(float(q.lvar[j]), float(q.uvar[j]))
            if s is not None:
                self.add_constraint(v, s)
        rows = [[] for _ in range(len(q.lcon))]
        for i, j, a in zip(q.arows, q.acols, q.avals):
            rows[int(i)].append((float(a), xs[int(j)]))
        for i, terms in enumerate(rows):
            s = _set_of(float(q.lcon[i]), float(q.ucon[i]))
            if s is None:   # a free row constrains nothing; kept so rows keep their indices
                s = Interval(-np.inf, np.inf)
            self.add_constraint(ScalarAffineFunction(terms, 0.0), s)
        obj = ScalarAffineFunction([(float(q.c[j]), xs[j]) for j in range(len(q.c)) if q.c[j] != 0.0],
                                   float(q.c0))
        self.set_objective(obj, ObjectiveSense.MAX_SENSE if q.objsense.upper().startswith("MAX")
                           else ObjectiveSense.MIN_SENSE)
        return xs

    def set_objective(self, func, sense=ObjectiveSense.MIN_SENSE):
        if isinstance(func, VariableIndex):
            func = ScalarAffineFunction([(1.0, func)], 0.0)
        if not isinstance(func, ScalarAffineFunction):
            raise ArgumentError(f"unsupported objective {func!r}")
        self._obj = func
        self._sense = sense

    def _problem(self):
        n, m = len(self._lv), len(self._rows)
        ri, ci, vv = [], [], []
        for i, (terms, *_rest) in enumerate(self._rows):
            for cf, j in terms:
                ri.append(i)
                ci.append(j)
                vv.append(cf)
        A = sp.csc_matrix((vv, (ri, ci)), shape=(m, n))   # duplicates summed
        A.sum_duplicates()
        c = np.zeros(n)
        for cf, v in self._obj.terms:
            c[v.value] += float(cf)
        max_sense = self._sense == ObjectiveSense.MAX_SENSE
        if max_sense:
            c = -c
        lv, uv = np.array(self._lv, dtype=np.float64), np.array(self._uv, dtype=np.float64)
        lc = np.array([r[1] for r in self._rows], dtype=np.float64)
        uc = np.array([r[2] for r in self._rows], dtype=np.float64)
        return MILP(c=c, lv=lv, uv=uv, A=A, lc=lc, uc=uc), max_sense, float(self._obj.constant)

    # ---- optimize! (src/moi.jl:165-246)
    def optimize(self):
        from .algorithms import PDLP, solve
        self._clear_results()
        milp, max_sense, constant = self._problem()
        opts = dict(self.options)
        algorithm = opts.pop("algorithm", PDLP)
        float_type = opts.pop("float_type", "float64")
        if float_type not in ("float64", "Float64", np.float64):
            warnings.warn(f"Got mismatched float type: solving in {float_type} but returning "
                          "the solution in float64.")
        int_type = opts.pop("int_type", None)
        matrix_type = opts.pop("matrix_type", None)
        algo_opts = {"show_progress": int(not self.silent)}
        algo_opts.update(opts)
        algo = algorithm(float_type, int_type, matrix_type, **algo_opts)
        sol, stats = solve(milp, algo)
        self.finish(milp, sol.x, sol.y, stats.termination_status.name, stats.time_elapsed,
                    max_sense, constant)

    def finish(self, milp: MILP, x, y, status: str, solve_time: float, max_sense: bool = False,
               constant: float = 0.0):
        """The results of a solve, as src/moi.jl:210-244 derives them: reduced costs
        z = proj_multiplier(c - A'y, lv, uv); the objective c'x; the dual objective
        lc'|y|+ - uc'|y|- + lv'|z|+ - uv'|z|- (each product `safeprod_left`); both sign-
        flipped for maximisation, plus the objective constant; the statuses."""
        self.x = np.asarray(x, dtype=np.float64).copy()
        self.y = np.asarray(y, dtype=np.float64).copy()
        self.z = proj_multiplier(milp.c - milp.At @ self.y, milp.lv, milp.uv)
        raw_obj = float(np.dot(self.x, milp.c))
        raw_dual = float(np.sum(safeprod_left(milp.lc, positive_part(self.y)))
                         - np.sum(safeprod_left(milp.uc, negative_part(self.y)))
                         + np.sum(safeprod_left(milp.lv, positive_part(self.z)))
                         - np.sum(safeprod_left(milp.uv, negative_part(self.z))))
        self.obj_value = (-raw_obj if max_sense else raw_obj) + constant
        self.dual_obj_value = (-raw_dual if max_sense else raw_dual) + constant
        self.solve_time = float(solve_time)
        if status not in _STATUS:
            raise ArgumentError(f"unknown termination status {status!r}")
        self.termination_status, self.primal_status, self.dual_status = _STATUS[status]

    # ---- results (src/moi.jl:99-154)
    def _check_result(self, result_index: int = 1):
        if result_index < 1 or result_index > self.result_count:
            raise ArgumentError(f"result index {result_index} out of bounds (ResultCount "
                                f"{self.result_count})")

    def get_primal_status(self, result_index: int = 1) -> ResultStatusCode:
        return ResultStatusCode.NO_SOLUTION if result_index > self.result_count else self.primal_status

    def get_dual_status(self, result_index: int = 1) -> ResultStatusCode:
        return ResultStatusCode.NO_SOLUTION if result_index > self.result_count else self.dual_status

    def objective_value(self, result_index: int = 1) -> float:
        self._check_result(result_index)
        return self.obj_value

    def dual_objective_value(self, result_index: int = 1) -> float:
        self._check_result(result_index)
        return self.dual_obj_value

    def variable_primal(self, vi: VariableIndex, result_index: int = 1) -> float:
        self._check_result(result_index)
        return float(self.x[vi.value])

    def constraint_dual(self, ci: ConstraintIndex, result_index: int = 1) -> float:
        self._check_result(result_index)
        if ci.kind == "affine":
            return float(self.y[ci.value])
        zj = float(self.z[ci.value])
        if ci.set_type is GreaterThan:
            return float(positive_part(zj))
        if ci.set_type is LessThan:
            return float(-negative_part(zj))
        return zj
