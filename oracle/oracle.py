"""ctypes wrapper of the CPU oracle (libcpdlp_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker / CPU baseline, never by the
product package.

Parity pinning of each oracle function (see tests/test_oracle_*.py):
  csc_to_csr / ell      -- scipy.sparse transpose + dense products (bit-exact)
  spmv_csr              -- dense alpha*A*b + beta*c (test/utils/matrices.jl:26)
  scaling               -- Ruiz -> unit inf-norms (test/components/preconditioning.jl:36-41),
                           composition/involution, MILP objective & y'Ax invariance
  kkt_errors            -- closed forms of test/components/errors.jl:27-34 and
                           invariance under preconditioning (:36-38)
  spectral_norm         -- numpy 2-norm (test/utils/linalg.jl:52-60), stopping rule
  solve                 -- closed-form optima c'x0, scipy HiGHS, the reference's
                           known answers (205.0, test/moi.jl:44-52; OPTIMAL on
                           test/components/termination.jl:2-9)
  restart_decision      -- hand-worked known answers at every boundary of should_restart
                           and the candidate choice (tests/test_oracle_restart.py)
  primal_weight_update  -- its power form (dy/dx)^t w^(1-t), t in {0, 0.3, 0.5, 1}, and
                           the zero_tol guards; at the first restart of a solve against
                           distances of independently computed iterates
  averaging / restart   -- the recorded check errors against iterates computed one pass
                           at a time with restarts disabled (mean of the iterates since
                           the start, the iterate one step earlier, F10) and against a
                           warm-started solve from the restart candidate
  powm_start            -- parity unpinned (the reference draws StableRNG(0) randn)
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcpdlp_oracle.so")

STATUS_NAMES = ["OPTIMAL", "TIME_LIMIT", "ITERATION_LIMIT", "STILL_RUNNING", "PRIMAL_INFEASIBLE",
                "DUAL_INFEASIBLE", "NUMERICAL_ERROR"]


class Params(C.Structure):
    _fields_ = [
        ("algorithm", C.c_int32), ("ruiz_iter", C.c_int32),
        ("chambolle_pock_alpha", C.c_double), ("invnorm_scaling", C.c_double),
        ("primal_weight_damping", C.c_double), ("zero_tol", C.c_double),
        ("sufficient_decay", C.c_double), ("necessary_decay", C.c_double),
        ("artificial_decay", C.c_double), ("check_every", C.c_int32),
        ("record_error_history", C.c_int32), ("termination_reltol", C.c_double),
        ("max_kkt_passes", C.c_int64), ("time_limit", C.c_double),
        ("powm_tol", C.c_double), ("powm_maxiter", C.c_int64), ("powm_seed", C.c_uint64),
        ("detect_infeasibility", C.c_int32), ("ray_candidates", C.c_int32),
        ("infeasibility_tol", C.c_double),
    ]


class KKT(C.Structure):
    _fields_ = [(f, C.c_double) for f in
                ("primal", "primal_scale", "dual", "dual_scale", "gap", "gap_scale")]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}

    def relative(self):
        return max(self.primal / self.primal_scale, self.dual / self.dual_scale,
                   self.gap / self.gap_scale)


class Hist(C.Structure):
    _fields_ = [("kkt_passes", C.c_int64), ("err", KKT)]


class Stats(C.Structure):
    _fields_ = [("err", KKT), ("time_elapsed", C.c_double), ("kkt_passes", C.c_int64),
                ("termination_status", C.c_int32), ("pad", C.c_int32),
                ("eta", C.c_double), ("omega", C.c_double), ("powm_iters", C.c_int64),
                ("restarts", C.c_int64), ("n_history", C.c_int64), ("setup_seconds", C.c_double)]


class Trace(C.Structure):
    _fields_ = [("cap", C.c_int64), ("count", C.c_int64), ("passes", C.POINTER(C.c_int64)),
                ("err", C.POINTER(KKT)), ("from_avg", C.POINTER(C.c_int32)),
                ("restarted", C.POINTER(C.c_int32)), ("omega", C.POINTER(C.c_double)),
                ("omega_after", C.POINTER(C.c_double)), ("inner", C.POINTER(C.c_int64)),
                ("total", C.POINTER(C.c_int64)), ("x_avg", C.POINTER(C.c_double)),
                ("y_avg", C.POINTER(C.c_double))]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        d, i64, i32 = P(C.c_double), P(C.c_int64), P(C.c_int32)
        L.orc_default_params.argtypes = [P(Params), C.c_int32]
        L.orc_set_threads.argtypes = [C.c_int32]
        L.orc_csc_to_csr.argtypes = [C.c_int64, C.c_int64, i64, i32, d, i64, i32, d]
        L.orc_spmv_csr.argtypes = [C.c_int64, i64, i32, d, C.c_double, d, C.c_double, d]
        L.orc_ell_width.argtypes = [C.c_int64, i64]
        L.orc_ell_width.restype = C.c_int64
        L.orc_csr_to_ell.argtypes = [C.c_int64, C.c_int64, i64, i32, d, i32, d]
        L.orc_scaling.argtypes = [C.c_int64, C.c_int64, i64, i32, d, C.c_int32, C.c_double,
                                  d, d, d, d]
        L.orc_kkt_errors.argtypes = [C.c_int64, C.c_int64, i64, i32, d, i64, i32, d,
                                     d, d, d, d, d, d, d, d, d, P(KKT)]
        L.orc_powm_start.argtypes = [C.c_int64, C.c_uint64, d]
        L.orc_spectral_norm.argtypes = [C.c_int64, C.c_int64, i64, i32, d, i64, i32, d,
                                        C.c_double, C.c_int64, C.c_uint64, i64]
        L.orc_spectral_norm.restype = C.c_double
        L.orc_solve.argtypes = [C.c_int64, C.c_int64, i64, i32, d, d, d, d, d, d, d, d,
                                P(Params), d, d, P(Stats), P(Hist), C.c_int64]
        L.orc_solve.restype = C.c_int
        L.orc_ray_test.argtypes = [C.c_int64, C.c_int64, i64, i32, d, i64, i32, d,
                                   d, d, d, d, d, d, d, d, d, C.c_double]
        L.orc_ray_test.restype = C.c_int32
        L.orc_restart_decision.argtypes = [P(KKT)] * 5 + [C.c_double, C.c_int64, C.c_int64,
                                                         P(Params), i32]
        L.orc_restart_decision.restype = C.c_int32
        L.orc_primal_weight_update.argtypes = [C.c_double] * 5
        L.orc_primal_weight_update.restype = C.c_double
        L.orc_solve_traced.argtypes = L.orc_solve.argtypes + [P(Trace)]
        L.orc_solve_traced.restype = C.c_int
        _lib = L
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def set_threads(n: int):
    lib().orc_set_threads(int(n))


def default_params(algorithm="PDLP", **kw) -> Params:
    p = Params()
    lib().orc_default_params(C.byref(p), 0 if algorithm == "PDLP" else 1)
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def csc_to_csr(m, n, colptr, rowval, nzval):
    colptr, rowval, nzval = _i64(colptr), _i32(rowval), _f64(nzval)
    nnz = int(colptr[-1])
    rowptr = np.zeros(m + 1, np.int64)
    col = np.zeros(max(nnz, 1), np.int32)
    val = np.zeros(max(nnz, 1), np.float64)
    lib().orc_csc_to_csr(m, n, _p(colptr, C.c_int64), _p(rowval, C.c_int32), _p(nzval, C.c_double),
                         _p(rowptr, C.c_int64), _p(col, C.c_int32), _p(val, C.c_double))
    return rowptr, col[:nnz], val[:nnz]


def spmv_csr(rowptr, colval, val, alpha, b, beta, c):
    rowptr, colval, val, b = _i64(rowptr), _i32(colval), _f64(val), _f64(b)
    out = _f64(c).copy()
    lib().orc_spmv_csr(len(rowptr) - 1, _p(rowptr, C.c_int64), _p(colval, C.c_int32),
                       _p(val, C.c_double), float(alpha), _p(b, C.c_double), float(beta),
                       _p(out, C.c_double))
    return out


def csr_to_ell(rowptr, colval, val):
    rowptr, colval, val = _i64(rowptr), _i32(colval), _f64(val)
    m = len(rowptr) - 1
    d = int(lib().orc_ell_width(m, _p(rowptr, C.c_int64)))
    ec = np.zeros(max(m * d, 1), np.int32)
    ev = np.zeros(max(m * d, 1), np.float64)
    lib().orc_csr_to_ell(m, d, _p(rowptr, C.c_int64), _p(colval, C.c_int32), _p(val, C.c_double),
                         _p(ec, C.c_int32), _p(ev, C.c_double))
    return d, ec[: m * d].reshape(d, m), ev[: m * d].reshape(d, m)


def scaling(m, n, colptr, rowval, nzval, ruiz_iter=10, cp_alpha=1.0):
    colptr, rowval, nzval = _i64(colptr), _i32(rowval), _f64(nzval)
    nnz = int(colptr[-1])
    D1 = np.zeros(max(m, 1))
    D2 = np.zeros(max(n, 1))
    a = np.zeros(max(nnz, 1))
    at = np.zeros(max(nnz, 1))
    lib().orc_scaling(m, n, _p(colptr, C.c_int64), _p(rowval, C.c_int32), _p(nzval, C.c_double),
                      int(ruiz_iter), float(cp_alpha), _p(D1, C.c_double), _p(D2, C.c_double),
                      _p(a, C.c_double), _p(at, C.c_double))
    return D1[:m], D2[:n], a[:nnz], at[:nnz]


def kkt_errors(m, n, A_csr, At_csr, c, lv, uv, lc, uc, D1, D2, x, y) -> KKT:
    ar, ac, av = _i64(A_csr[0]), _i32(A_csr[1]), _f64(A_csr[2])
    tr, tc, tv = _i64(At_csr[0]), _i32(At_csr[1]), _f64(At_csr[2])
    vs = [_f64(v) for v in (c, lv, uv, lc, uc, D1, D2, x, y)]
    out = KKT()
    lib().orc_kkt_errors(m, n, _p(ar, C.c_int64), _p(ac, C.c_int32), _p(av, C.c_double),
                         _p(tr, C.c_int64), _p(tc, C.c_int32), _p(tv, C.c_double),
                         *[_p(v, C.c_double) for v in vs], C.byref(out))
    return out


def ray_test(m, n, colptr, rowval, nzval, c, lv, uv, lc, uc, D1, D2, dx, dy, tol=1e-8) -> str:
    """The opt-in infeasibility test on one ray (dx, dy) of an LP given in scaled form
    (CSC A_s and scaled vectors) with its scaling D1, D2."""
    A_csr = csc_to_csr(m, n, colptr, rowval, nzval)
    colptr, rowval, nzval = _i64(colptr), _i32(rowval), _f64(nzval)
    vs = [_f64(v) for v in (c, lv, uv, lc, uc, D1, D2, dx, dy)]
    st = lib().orc_ray_test(m, n, _p(A_csr[0], C.c_int64), _p(A_csr[1], C.c_int32),
                            _p(A_csr[2], C.c_double), _p(colptr, C.c_int64), _p(rowval, C.c_int32),
                            _p(nzval, C.c_double), *[_p(v, C.c_double) for v in vs], float(tol))
    return STATUS_NAMES[st]


def powm_start(n, seed=0):
    x = np.zeros(max(n, 1))
    lib().orc_powm_start(n, seed, _p(x, C.c_double))
    return x[:n]


def spectral_norm(m, n, A_csr, At_csr, tol=None, maxiter=None, seed=0):
    ar, ac, av = _i64(A_csr[0]), _i32(A_csr[1]), _f64(A_csr[2])
    tr, tc, tv = _i64(At_csr[0]), _i32(At_csr[1]), _f64(At_csr[2])
    if tol is None:
        n3 = (n * n * n) & ((1 << 64) - 1)
        if n3 >= 1 << 63:
            n3 -= 1 << 64
        tol = np.finfo(np.float64).eps * float(n3)
    it = np.zeros(1, np.int64)
    s = lib().orc_spectral_norm(m, n, _p(ar, C.c_int64), _p(ac, C.c_int32), _p(av, C.c_double),
                                _p(tr, C.c_int64), _p(tc, C.c_int32), _p(tv, C.c_double),
                                float(tol), int(n if maxiter is None else maxiter), int(seed),
                                _p(it, C.c_int64))
    return s, int(it[0])


@dataclass
class OracleResult:
    x: np.ndarray
    y: np.ndarray
    stats: Stats
    history: list

    @property
    def status(self) -> str:
        return STATUS_NAMES[self.stats.termination_status]


def solve(lp, params: Params | None = None, x0=None, y0=None, hist_cap=100_000, **kw) -> OracleResult:
    """Solve a GeneratedLP-like object (m, n, colptr, rowval, nzval, c, lv, uv, lc, uc)."""
    if params is None:
        params = default_params(**kw)
    m, n = lp.m, lp.n
    colptr, rowval, nzval = _i64(lp.colptr), _i32(lp.rowval), _f64(lp.nzval)
    vs = [_f64(v) for v in (lp.c, lp.lv, lp.uv, lp.lc, lp.uc)]
    xo = np.zeros(max(n, 1))
    yo = np.zeros(max(m, 1))
    st = Stats()
    hist = (Hist * max(hist_cap, 1))()
    x0p = _p(_f64(x0), C.c_double) if x0 is not None else None
    y0p = _p(_f64(y0), C.c_double) if y0 is not None else None
    lib().orc_solve(m, n, _p(colptr, C.c_int64), _p(rowval, C.c_int32), _p(nzval, C.c_double),
                    *[_p(v, C.c_double) for v in vs], x0p, y0p, C.byref(params),
                    _p(xo, C.c_double), _p(yo, C.c_double), C.byref(st), hist, hist_cap)
    h = [(hist[i].kkt_passes, hist[i].err.as_dict()) for i in range(min(st.n_history, hist_cap))]
    return OracleResult(xo[:n], yo[:m], st, h)


def kkt(**fields) -> KKT:
    """A KKTErrors value; unset fields are 0 (scales 1)."""
    e = KKT()
    for f, _ in KKT._fields_:
        setattr(e, f, fields.get(f, 1.0 if f.endswith("_scale") else 0.0))
    return e


def restart_decision(e_sol, e_avg, e_last, e_avg_last, e_restart, omega, inner, total,
                     params: Params | None = None):
    """(restart?, candidate is the average?) of restart_check! + should_restart
    (pdlp.jl:149-169, restart.jl:40-57) on five KKT errors."""
    params = params or default_params()
    fa = C.c_int32()
    r = lib().orc_restart_decision(C.byref(e_sol), C.byref(e_avg), C.byref(e_last),
                                   C.byref(e_avg_last), C.byref(e_restart), float(omega),
                                   int(inner), int(total), C.byref(params), C.byref(fa))
    return bool(r), bool(fa.value)


def primal_weight_update(omega, dx, dy, damping=0.5, zero_tol=1e-8) -> float:
    """primal_weight_update!, step_size.jl:58-75, on the distances dx, dy."""
    return lib().orc_primal_weight_update(float(omega), float(dx), float(dy), float(damping),
                                          float(zero_tol))


@dataclass
class TraceResult:
    result: OracleResult
    passes: np.ndarray       # kkt_passes at each check
    err: list                # per check: dict name -> KKT (sol, avg, last, avg_last, restart)
    from_avg: np.ndarray
    restarted: np.ndarray
    omega: np.ndarray
    omega_after: np.ndarray
    inner: np.ndarray
    total: np.ndarray
    x_avg: np.ndarray        # final average, unscaled
    y_avg: np.ndarray


def solve_traced(lp, params: Params, cap=4096, x0=None, y0=None) -> TraceResult:
    """solve() with a per-check record of the restart check's inputs and decisions."""
    m, n = lp.m, lp.n
    colptr, rowval, nzval = _i64(lp.colptr), _i32(lp.rowval), _f64(lp.nzval)
    vs = [_f64(v) for v in (lp.c, lp.lv, lp.uv, lp.lc, lp.uc)]
    xo, yo = np.zeros(max(n, 1)), np.zeros(max(m, 1))
    st = Stats()
    hist = (Hist * 16)()
    arr = {"passes": np.zeros(cap, np.int64), "from_avg": np.zeros(cap, np.int32),
           "restarted": np.zeros(cap, np.int32), "omega": np.zeros(cap), "omega_after": np.zeros(cap),
           "inner": np.zeros(cap, np.int64), "total": np.zeros(cap, np.int64)}
    errs = (KKT * (5 * cap))()
    xa, ya = np.zeros(max(n, 1)), np.zeros(max(m, 1))
    tr = Trace()
    tr.cap, tr.count = cap, 0
    tr.err = errs
    ctype = {np.dtype(np.int64): C.c_int64, np.dtype(np.int32): C.c_int32, np.dtype(np.float64): C.c_double}
    for k, a in arr.items():
        setattr(tr, k, _p(a, ctype[a.dtype]))
    tr.x_avg, tr.y_avg = _p(xa, C.c_double), _p(ya, C.c_double)
    x0p = _p(_f64(x0), C.c_double) if x0 is not None else None
    y0p = _p(_f64(y0), C.c_double) if y0 is not None else None
    lib().orc_solve_traced(m, n, _p(colptr, C.c_int64), _p(rowval, C.c_int32), _p(nzval, C.c_double),
                           *[_p(v, C.c_double) for v in vs], x0p, y0p, C.byref(params),
                           _p(xo, C.c_double), _p(yo, C.c_double), C.byref(st), hist, 16, C.byref(tr))
    k = min(tr.count, cap)
    names = ("sol", "avg", "last", "avg_last", "restart")
    err = [{nm: errs[5 * i + q] for q, nm in enumerate(names)} for i in range(k)]
    res = OracleResult(xo[:n], yo[:m], st, [])
    return TraceResult(res, arr["passes"][:k], err, arr["from_avg"][:k].astype(bool),
                       arr["restarted"][:k].astype(bool), arr["omega"][:k], arr["omega_after"][:k],
                       arr["inner"][:k], arr["total"][:k], xa[:n], ya[:m])
