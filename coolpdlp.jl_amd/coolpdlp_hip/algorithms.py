"""Algorithm configs, results and solve(), mirroring src/algorithms/*.jl and the
components' public structs (KKTErrors, ConvergenceStats, TerminationStatus).

Every solve runs on the HIP engine through the C ABI; nothing here computes the
method itself.
"""
from __future__ import annotations

import copy
import ctypes as C
import enum
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import ArgumentError, check, lib
from .problems import MILP, PrimalDualSolution, require_unpreconditioned


class TerminationStatus(enum.IntEnum):
    """src/components/termination.jl:13 (same order), plus three statuses that only the
    opt-in detect_infeasibility=True produces (the reference has no such check)."""
    OPTIMAL = 0
    TIME_LIMIT = 1
    ITERATION_LIMIT = 2
    STILL_RUNNING = 3
    PRIMAL_INFEASIBLE = 4     # a dual ray (Farkas certificate) was found
    DUAL_INFEASIBLE = 5       # a primal ray was found
    NUMERICAL_ERROR = 6       # the KKT errors became NaN or infinite


OPTIMAL = TerminationStatus.OPTIMAL
TIME_LIMIT = TerminationStatus.TIME_LIMIT
ITERATION_LIMIT = TerminationStatus.ITERATION_LIMIT
STILL_RUNNING = TerminationStatus.STILL_RUNNING
PRIMAL_INFEASIBLE = TerminationStatus.PRIMAL_INFEASIBLE
DUAL_INFEASIBLE = TerminationStatus.DUAL_INFEASIBLE
NUMERICAL_ERROR = TerminationStatus.NUMERICAL_ERROR


@dataclass
class KKTErrors:
    """src/components/errors.jl:8-21"""
    primal: float = math.nan
    primal_scale: float = math.nan
    dual: float = math.nan
    dual_scale: float = math.nan
    gap: float = math.nan
    gap_scale: float = math.nan

    @classmethod
    def from_c(cls, e):
        return cls(e.primal, e.primal_scale, e.dual, e.dual_scale, e.gap, e.gap_scale)

    def relative(self) -> float:
        """errors.jl:55-58"""
        return max(self.primal / self.primal_scale, self.dual / self.dual_scale,
                   self.gap / self.gap_scale)

    def absolute(self, omega: float) -> float:
        """errors.jl:60-63"""
        return math.sqrt(omega ** 2 * self.primal ** 2 + omega ** -2 * self.dual ** 2
                         + self.gap ** 2)

    def isapprox(self, other: "KKTErrors", rtol=1e-8, atol=0.0) -> bool:
        """errors.jl:33-42"""
        return all(math.isclose(getattr(self, f), getattr(other, f), rel_tol=rtol, abs_tol=atol)
                   for f in ("primal", "dual", "gap", "primal_scale", "dual_scale", "gap_scale"))

    def __str__(self):
        return (f"KKT relative errors: primal {self.primal / self.primal_scale:.3e}, "
                f"dual {self.dual / self.dual_scale:.3e}, gap {self.gap / self.gap_scale:.3e}")


relative = KKTErrors.relative
absolute = KKTErrors.absolute


@dataclass
class ConvergenceStats:
    """src/components/termination.jl:44-56, plus the engine's counters."""
    err: KKTErrors
    time_elapsed: float
    kkt_passes: int
    termination_status: TerminationStatus
    error_history: list = field(default_factory=list)
    eta: float = math.nan
    omega: float = math.nan
    powm_iters: int = 0
    restarts: int = 0
    fp32_passes: int = 0
    setup_seconds: float = 0.0
    loop_seconds: float = 0.0
    step_kernel_ms: float = 0.0
    step_kernel_launches: int = 0
    format_A: int = 0
    format_At: int = 0
    line_tiles: int = 0   # line-grouped short tiles in use: bit 0 A, bit 1 A'


_FLOAT_TYPES = {"float64": _lib.PREC_FP64, "Float64": _lib.PREC_FP64, np.float64: _lib.PREC_FP64,
                "float32": _lib.PREC_FP32, "Float32": _lib.PREC_FP32, np.float32: _lib.PREC_FP32,
                "mixed": _lib.PREC_MIXED}


@dataclass
class LoopState:
    """The PDLP loop's state between two checks (PDLPState, src/algorithms/pdlp.jl:19-43,
    less its scratch and per-check statistics): the five primal vectors `x` (rows sol,
    sol_last, sol_avg, sol_avg_last, sol_restart; n entries each) and dual vectors `y` (m
    each), scaled, in the LP's original order; the scaling they belong to (`D2`, `D1`:
    the unscaled point is x * D2, y * D1); and the scalars.  Engine.get_state and
    BatchSolver.take_state produce it, Engine.resume continues from it."""
    x: np.ndarray
    y: np.ndarray
    D2: np.ndarray
    D1: np.ndarray
    eta: float
    omega: float
    eta_sum: float
    kkt_passes: int
    inner: int
    total: int
    restarts: int
    fp32_passes: int

    @classmethod
    def from_c(cls, s, x, y, D2, D1):
        return cls(x, y, D2, D1, s.eta, s.omega, s.eta_sum, s.kkt_passes, s.inner, s.total,
                   s.restarts, s.fp32_passes)

    def to_c(self):
        s = _lib.CState()
        for f, _ in s._fields_:
            setattr(s, f, getattr(self, f))
        return s

    _SCALARS = ("eta", "omega", "eta_sum", "kkt_passes", "inner", "total", "restarts", "fp32_passes")

    def save(self, path):
        """Write the state to an .npz file (arrays and scalars only, no pickled objects)."""
        np.savez(path, x=self.x, y=self.y, D2=self.D2, D1=self.D1,
                 **{f: np.asarray(getattr(self, f)) for f in self._SCALARS})

    @classmethod
    def load(cls, path) -> "LoopState":
        """Read a state written by `save`."""
        with np.load(path, allow_pickle=False) as z:
            kw = {f: z[f].item() for f in cls._SCALARS}
            return cls(z["x"].copy(), z["y"].copy(), z["D2"].copy(), z["D1"].copy(), **kw)

    @classmethod
    def buffers(cls, m: int, n: int):
        k = _lib.STATE_VECTORS
        return (np.zeros((k, max(n, 1))), np.zeros((k, max(m, 1))), np.zeros(max(n, 1)),
                np.zeros(max(m, 1)))


class Algorithm:
    """Algorithm{A}(T, Ti, M; backend, kwargs...), src/algorithms/common.jl:53-116.

    Positional: float type T ("float64" | "float32" | "mixed"), index type Ti
    (accepted; the engine indexes with int32), matrix type M (GPUSparseMatrixCSR,
    GPUSparseMatrixELL or "auto").  Keyword arguments are the reference's, with the same
    defaults, plus engine options (device, powm_tol, powm_maxiter, powm_seed,
    polish_reltol, use_graph, long_row)."""

    name = "PDLP"
    _algo = _lib.ALGO_PDLP

    def __init__(self, float_type="float64", int_type=None, matrix_type=None, *, backend="hip",
                 **kwargs):
        if backend not in ("hip", "rocm", "gpu"):
            raise ArgumentError(f"unsupported backend {backend!r}: this engine runs on HIP")
        self.params = _lib.default_params(self._algo)
        if float_type not in _FLOAT_TYPES:
            raise ArgumentError(f"unsupported float type {float_type!r}")
        self.params.precision = _FLOAT_TYPES[float_type]
        self.float_type = float_type
        self.int_type = int_type
        from .matrices import format_of
        self.params.format = format_of(matrix_type)
        self.matrix_type = matrix_type
        names = {f for f, _ in _lib.CParams._fields_}
        for k, v in kwargs.items():
            if k not in names or k in ("algorithm", "precision", "format"):
                raise ArgumentError(f"unknown keyword argument {k!r}")
            setattr(self.params, k, v)

    def __repr__(self):
        p = self.params
        return (f"{self.name} algorithm: float_type={self.float_type}, check_every={p.check_every}, "
                f"termination_reltol={p.termination_reltol}, max_kkt_passes={p.max_kkt_passes}, "
                f"time_limit={p.time_limit}, ruiz_iter={p.ruiz_iter}, "
                f"chambolle_pock_alpha={p.chambolle_pock_alpha}")


class PDLP(Algorithm):
    """PDLP(args...; kwargs...), src/algorithms/pdlp.jl:6-11"""
    name = "PDLP"
    _algo = _lib.ALGO_PDLP


class PDHG(Algorithm):
    """PDHG(args...; kwargs...), src/algorithms/pdhg.jl:6-16 (ruiz_iter=0, no restarts)"""
    name = "PDHG"
    _algo = _lib.ALGO_PDHG


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class Engine:
    """One LP resident on the GPU (a cpdlp_handle): build once, solve, query."""

    def __init__(self, milp: MILP, algo: Algorithm):
        require_unpreconditioned(milp)
        self.milp = milp
        self.algo = algo
        A = milp.A
        self._colptr = np.ascontiguousarray(A.indptr, dtype=np.int64)
        self._rowval = np.ascontiguousarray(A.indices, dtype=np.int64)
        self._nzval = np.ascontiguousarray(A.data, dtype=np.float64)
        self._vecs = [np.ascontiguousarray(v, dtype=np.float64)
                      for v in (milp.c, milp.lv, milp.uv, milp.lc, milp.uc)]
        lp = _lib.CLP()
        lp.m, lp.n = A.shape
        lp.index_base, lp.index_bits = 0, 64
        lp.colptr, lp.rowval, lp.nzval = (_ptr(self._colptr), _ptr(self._rowval),
                                          _ptr(self._nzval))
        lp.c, lp.lv, lp.uv, lp.lc, lp.uc = [_ptr(v) for v in self._vecs]
        h = C.c_void_p()
        check(lib.cpdlp_create(C.byref(lp), C.byref(algo.params), C.byref(h)))
        self.h = h
        self.m, self.n = A.shape

    def close(self):
        if getattr(self, "h", None):
            lib.cpdlp_destroy(self.h)
            self.h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def prepare(self):
        check(lib.cpdlp_prepare(self.h))

    def solve(self, sol: PrimalDualSolution | None = None):
        x0, y0 = self._start(sol)
        check(lib.cpdlp_solve(self.h, _ptr(x0), _ptr(y0)))
        return self.solution(), self.stats()

    def _start(self, sol):
        x0 = y0 = None
        if sol is not None:
            x0 = np.ascontiguousarray(sol.x, dtype=np.float64)
            y0 = np.ascontiguousarray(sol.y, dtype=np.float64)
            if len(x0) != self.n or len(y0) != self.m:
                raise _lib.DimensionMismatch("starting point has the wrong size")
        return x0, y0

    def initialize(self, sol: PrimalDualSolution | None = None):
        """initialize(milp, sol, algo), common.jl:176: everything before the loop."""
        x0, y0 = self._start(sol)
        check(lib.cpdlp_initialize(self.h, _ptr(x0), _ptr(y0)))

    def run_checks(self, max_checks: int = 1) -> TerminationStatus:
        """solve!'s loop for at most `max_checks` chunks of check_every steps
        (pdlp.jl:69-94); returns the termination status after the last check."""
        st = C.c_int32()
        check(lib.cpdlp_run_checks(self.h, int(max_checks), C.byref(st)))
        return TerminationStatus(st.value)

    def get_state(self) -> LoopState:
        """The loop's state after the last check (cpdlp_get_state): a checkpoint that
        `resume` continues, here or on another engine of the same LP."""
        x, y, D2, D1 = LoopState.buffers(self.m, self.n)
        s = _lib.CState()
        check(lib.cpdlp_get_state(self.h, C.byref(s), _ptr(x), _ptr(y), _ptr(D2), _ptr(D1)))
        return LoopState.from_c(s, x[:, :self.n], y[:, :self.m], D2[:self.n], D1[:self.m])

    def resume(self, state: LoopState):
        """initialize, then continue from `state` (cpdlp_resume): run_checks carries on
        that run.  A state taken from an engine with the same scaling resumes bit for bit."""
        k = _lib.STATE_VECTORS
        x = np.ascontiguousarray(state.x, dtype=np.float64)
        y = np.ascontiguousarray(state.y, dtype=np.float64)
        D2 = np.ascontiguousarray(state.D2, dtype=np.float64)
        D1 = np.ascontiguousarray(state.D1, dtype=np.float64)
        if x.shape != (k, self.n) or y.shape != (k, self.m) or D2.shape != (self.n,) or D1.shape != (self.m,):
            raise _lib.DimensionMismatch("state does not match the LP's size")
        s = state.to_c()
        check(lib.cpdlp_resume(self.h, C.byref(s), _ptr(x), _ptr(y), _ptr(D2), _ptr(D1)))

    def solution_device(self, x_dev, y_dev):
        """Unscaled solution into torch device tensors (float64, cuda): no host copy."""
        for t, k in ((x_dev, self.n), (y_dev, self.m)):
            if t is not None and (t.dtype.itemsize != 8 or not t.is_cuda or t.numel() < k
                                  or not t.is_contiguous()):
                raise _lib.ArgumentError("solution_device needs contiguous float64 cuda tensors")
        import torch
        torch.cuda.current_stream().synchronize()
        check(lib.cpdlp_get_solution_device(
            self.h, C.c_void_p(x_dev.data_ptr()) if x_dev is not None else None,
            C.c_void_p(y_dev.data_ptr()) if y_dev is not None else None))

    def solution(self) -> PrimalDualSolution:
        x = np.zeros(self.n)
        y = np.zeros(self.m)
        check(lib.cpdlp_get_solution(self.h, _ptr(x), _ptr(y)))
        return PrimalDualSolution(x, y)

    def stats(self) -> ConvergenceStats:
        s = _lib.CStats()
        check(lib.cpdlp_get_stats(self.h, C.byref(s)))
        cnt = C.c_int64()
        check(lib.cpdlp_get_history(self.h, None, None, 0, C.byref(cnt)))
        k = cnt.value
        passes = np.zeros(max(k, 1), np.int64)
        errs = (_lib.CKKT * max(k, 1))()
        check(lib.cpdlp_get_history(self.h, _ptr(passes), errs, k, C.byref(cnt)))
        hist = [(int(passes[i]), KKTErrors.from_c(errs[i])) for i in range(k)]
        return ConvergenceStats(
            err=KKTErrors.from_c(s.err), time_elapsed=s.time_elapsed, kkt_passes=s.kkt_passes,
            termination_status=TerminationStatus(s.termination_status), error_history=hist,
            eta=s.eta, omega=s.omega, powm_iters=s.powm_iters, restarts=s.restarts,
            fp32_passes=s.fp32_passes, setup_seconds=s.setup_seconds,
            loop_seconds=s.loop_seconds, step_kernel_ms=s.step_kernel_ms,
            step_kernel_launches=s.step_kernel_launches, format_A=s.format_A,
            format_At=s.format_At, line_tiles=s.line_tiles)

    def kkt_errors(self, sol: PrimalDualSolution) -> KKTErrors:
        """kkt_errors! on the preconditioned problem for an unscaled point, errors.jl:65-109"""
        x = np.ascontiguousarray(sol.x, dtype=np.float64)
        y = np.ascontiguousarray(sol.y, dtype=np.float64)
        e = _lib.CKKT()
        check(lib.cpdlp_kkt_errors(self.h, _ptr(x), _ptr(y), C.byref(e)))
        return KKTErrors.from_c(e)

    def scaling(self):
        """(D1, D2, scaled A values, scaled At values) in CSC(A) layout"""
        nnz = len(self._nzval)
        D1, D2 = np.zeros(max(self.m, 1)), np.zeros(max(self.n, 1))
        a, at = np.zeros(max(nnz, 1)), np.zeros(max(nnz, 1))
        check(lib.cpdlp_get_scaling(self.h, _ptr(D1), _ptr(D2), _ptr(a), _ptr(at)))
        return D1[: self.m], D2[: self.n], a[:nnz], at[:nnz]

    def csr(self):
        nnz = len(self._nzval)
        rp = np.zeros(self.m + 1, np.int64)
        col = np.zeros(max(nnz, 1), np.int32)
        check(lib.cpdlp_get_csr(self.h, _ptr(rp), _ptr(col)))
        return rp, col[:nnz]

    def profile_steps(self, reps=20):
        """(ms per primal-step launch, ms per dual-step launch), HIP-event timed."""
        a, b = C.c_double(), C.c_double()
        check(lib.cpdlp_profile_steps(self.h, int(reps), C.byref(a), C.byref(b)))
        return a.value, b.value

    def step_size(self):
        eta, omega, it = C.c_double(), C.c_double(), C.c_int64()
        check(lib.cpdlp_get_step_size(self.h, C.byref(eta), C.byref(omega), C.byref(it)))
        return eta.value, omega.value, it.value


def preprocess(milp: MILP, sol: PrimalDualSolution | None, algo: Algorithm):
    """preprocess(milp_init, sol_init, algo), src/algorithms/common.jl:147-168.

    Returns (MILP, PrimalDualSolution): the problem and start point after the
    preconditioning (Ruiz + Chambolle-Pock, src/components/preconditioning.jl:95-155),
    whose D1 and D2 the engine computes on the device, applied as precondition
    (preconditioning.jl:37-91): A~ = (D1 a) D2, c~ = D2 c, lv~ = lv / D2, uv~ = uv / D2,
    lc~ = D1 lc, uc~ = D1 uc, x~ = x / D2, y~ = y / D1, in the caller's order; the MILP
    carries D1 and D2 (milp.jl:73).  No start point: zeros.  The scaling is computed in
    FP64 in every precision, so a mixed-precision algorithm is scaled by its FP64 form."""
    run = algo
    if algo.params.precision == _lib.PREC_MIXED:
        run = copy.copy(algo)
        run.params = type(algo.params).from_buffer_copy(algo.params)
        run.params.precision = _lib.PREC_FP64
    with Engine(milp, run) as eng:
        eng.initialize(sol)
        st = eng.get_state()
    D1, D2 = st.D1.copy(), st.D2.copy()
    A = milp.A.tocsc(copy=True)
    rows = A.indices
    cols = np.repeat(np.arange(A.shape[1]), np.diff(A.indptr))
    A.data = (D1[rows] * A.data) * D2[cols]
    scaled = MILP(c=D2 * milp.c, lv=milp.lv / D2, uv=milp.uv / D2, A=A, lc=D1 * milp.lc,
                  uc=D1 * milp.uc, D1=D1 * milp.D1, D2=milp.D2 * D2, int_var=milp.int_var,
                  var_names=milp.var_names, dataset=milp.dataset, name=milp.name, path=milp.path)
    if sol is None:
        x0, y0 = np.zeros(milp.A.shape[1]), np.zeros(milp.A.shape[0])
    else:
        x0, y0 = np.asarray(sol.x, np.float64), np.asarray(sol.y, np.float64)
    return scaled, PrimalDualSolution(x0 / D2, y0 / D1)


def initialize(milp: MILP, sol: PrimalDualSolution | None, algo: Algorithm) -> Engine:
    """initialize(milp, sol, algo), src/algorithms/common.jl:170-176: the state of a solve
    of `milp` from `sol` (an Engine, initialised; preprocessing included), which
    solve_inplace runs to termination and Engine.run_checks runs chunk by chunk.  Close it
    (or use it in a `with` block) to release its device memory.  In mixed precision the
    loop runs the FP32 phase, then the FP64 one, as solve does."""
    eng = Engine(milp, algo)
    try:
        eng.initialize(sol)
    except BaseException:
        eng.close()
        raise
    return eng


def solve_inplace(state: Engine, milp: MILP | None = None, algo: Algorithm | None = None):
    """solve!(state, milp, algo), src/algorithms/common.jl:211-216 (pdlp.jl:69-94): runs an
    initialised state's loop to termination; returns (PrimalDualSolution, ConvergenceStats)
    as solve does.  `milp` and `algo` are the state's own (accepted for the reference's
    signature)."""
    while state.run_checks(100) == STILL_RUNNING:
        pass
    return state.solution(), state.stats()


def solve(milp: MILP, *args):
    """solve(milp, algo) / solve(milp, sol, algo), src/algorithms/common.jl:186-209.

    Returns (PrimalDualSolution, ConvergenceStats); the solution is unscaled."""
    if len(args) == 1:
        sol, algo = None, args[0]
    elif len(args) == 2:
        sol, algo = args
    else:
        raise TypeError("solve(milp, algo) or solve(milp, sol, algo)")
    with Engine(milp, algo) as eng:
        return eng.solve(sol)
