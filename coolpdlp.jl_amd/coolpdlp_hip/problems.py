"""MILP container and PrimalDualSolution, mirroring src/problems/{milp,solution}.jl."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp

from ._lib import ArgumentError, DimensionMismatch


def _vec(v, name):
    a = np.asarray(v)
    if a.ndim != 1:
        raise DimensionMismatch(f"{name} must be a vector")
    if not np.issubdtype(a.dtype, np.floating):
        a = a.astype(np.float64)
    return a


class MILP:
    """min c'x  s.t.  lv <= x <= uv,  lc <= A x <= uc   ("cuPDLPx form").

    Mirrors MILP(; c, lv, uv, A, At, lc, uc, D1, D2, int_var, var_names, dataset, name,
    path), src/problems/milp.jl:55-115, including its checks: inconsistent sizes raise
    DimensionMismatch (:71-76); a non-numeric element type raises ArgumentError
    (:83-90).  A is any scipy.sparse matrix (kept as CSC with sorted rows, the
    reference's SparseMatrixCSC) or a dense array.
    """

    def __init__(self, *, c, lv, uv, A, lc, uc, At=None, D1=None, D2=None, int_var=None,
                 var_names=None, dataset="", name="", path=""):
        self.c = _vec(c, "c")
        self.lv = _vec(lv, "lv")
        self.uv = _vec(uv, "uv")
        self.lc = _vec(lc, "lc")
        self.uc = _vec(uc, "uc")
        if sp.issparse(A):
            A = A.tocsc()
        else:
            A = np.asarray(A)
            if A.ndim != 2:
                raise DimensionMismatch("A must be a matrix")
            A = sp.csc_matrix(A)
        if not np.issubdtype(A.dtype, np.number):
            raise ArgumentError("Abstract type parameter")
        A = A.astype(np.float64) if not np.issubdtype(A.dtype, np.floating) else A
        A.sort_indices()
        self.A = A
        m, n = A.shape
        self.At = A.T.tocsc() if At is None else sp.csc_matrix(At)
        self.D1 = np.ones(m) if D1 is None else _vec(D1, "D1")
        self.D2 = np.ones(n) if D2 is None else _vec(D2, "D2")
        self.int_var = np.zeros(n, bool) if int_var is None else np.asarray(int_var, bool)
        self.var_names = [str(j + 1) for j in range(n)] if var_names is None else list(var_names)
        if not (n == len(self.c) == len(self.lv) == len(self.uv) == len(self.D2)
                == len(self.int_var) == len(self.var_names)):
            raise DimensionMismatch("Variable size not consistent")
        if not (m == len(self.lc) == len(self.uc) == len(self.D1)):
            raise DimensionMismatch("Constraint size not consistent")
        if self.At.shape != (n, m):
            raise DimensionMismatch("At has the wrong shape")
        self.dataset, self.path = str(dataset), str(path)
        if not name and path:
            import os
            name = os.path.splitext(os.path.basename(path))[0]
        self.name = str(name)

    @classmethod
    def from_generated(cls, g, **kw):
        A = sp.csc_matrix((g.nzval, g.rowval, g.colptr), shape=(g.m, g.n))
        return cls(c=g.c, lv=g.lv, uv=g.uv, A=A, lc=g.lc, uc=g.uc, name=g.name, **kw)

    @classmethod
    def from_qps(cls, qps, **kw):
        """MILP(qps::QPSData; kwargs...), src/problems/milp.jl:122-137, for a QPSData from
        coolpdlp_hip.qps.read_mps."""
        from .qps import milp_from_qps
        return milp_from_qps(qps, **kw)

    def __repr__(self):
        return (f"MILP instance {self.name} from dataset {self.dataset}: {nbvar(self)} variables "
                f"({nbvar_int(self)} integer), {nbcons(self)} constraints "
                f"({nbcons_eq(self)} equalities), {self.A.nnz} nonzeros")


def require_unpreconditioned(lp) -> None:
    """The engine scales on the device and reports errors and solutions through its own
    D1, D2 only.  The reference composes those with the D1, D2 a MILP already carries
    (precondition, src/components/preconditioning.jl:73; kkt_errors!, errors.jl:77-84;
    get_solution, common.jl:233-235), so a preprocessed MILP would be measured and
    unscaled differently here: it is refused (ArgumentError).  Generated LPs carry none."""
    for name in ("D1", "D2"):
        d = getattr(lp, name, None)
        if d is not None and not np.all(np.asarray(d) == 1.0):
            raise ArgumentError("the engine takes an unpreconditioned MILP (D1 = D2 = I); "
                                "pass the problem given to preprocess, not its result")


def nbvar(milp: MILP) -> int:
    """milp.jl:164"""
    return len(milp.c)


def nbvar_int(milp: MILP) -> int:
    """milp.jl:171"""
    return int(np.sum(milp.int_var))


def nbvar_cont(milp: MILP) -> int:
    """milp.jl:178"""
    return nbvar(milp) - nbvar_int(milp)


def nbcons(milp: MILP) -> int:
    """milp.jl:185"""
    return milp.A.shape[0]


def nbcons_eq(milp: MILP) -> int:
    """milp.jl:192"""
    return int(np.sum(milp.lc == milp.uc))


def nbcons_ineq(milp: MILP) -> int:
    """milp.jl:199"""
    return nbcons(milp) - nbcons_eq(milp)


@dataclass
class PrimalDualSolution:
    """src/problems/solution.jl:49-54 (host arrays; the reference keeps them on its
    backend until Array(...) is called)."""
    x: np.ndarray
    y: np.ndarray

    def copy(self):
        return PrimalDualSolution(self.x.copy(), self.y.copy())


def objective_value(x, milp: MILP) -> float:
    """dot(x, milp.c), solution.jl:40"""
    return float(np.dot(np.asarray(x, dtype=np.float64), milp.c))


def is_feasible(x, milp: MILP, *, cons_tol=1e-6, int_tol=1e-5, verbose=True) -> bool:
    """solution.jl:12-33: bound, constraint and integrality violations vs tolerances."""
    import warnings
    x = np.asarray(x, dtype=np.float64)
    bounds_err = max(np.max(x - milp.uv, initial=-np.inf), np.max(milp.lv - x, initial=-np.inf))
    Ax = milp.A @ x
    cons_err = max(np.max(Ax - milp.uc, initial=-np.inf), np.max(milp.lc - Ax, initial=-np.inf))
    xint = x[milp.int_var]
    int_err = np.max(np.abs(xint - np.round(xint)), initial=-np.inf)
    if bounds_err > cons_tol:
        if verbose:
            warnings.warn(f"Variable bounds not satisfied: bounds_err={bounds_err}")
        return False
    if cons_err > cons_tol:
        if verbose:
            warnings.warn(f"Constraints not satisfied: cons_err={cons_err}")
        return False
    if int_err > int_tol:
        if verbose:
            warnings.warn(f"Integrality not satisfied: int_err={int_err}")
        return False
    return True
