"""Batched and served solves: many independent LPs on one GPU (cpdlp_batch_* C ABI).

The reference solves one LP per `solve` call (src/algorithms/common.jl:186-209); a
BatchSolver runs that same PDLP algorithm on many LPs at once, one workgroup per LP,
with every per-LP decision (scaling, step size, restarts, termination) on the device.
LPs can be added while others run (continuous batching): `run` advances all admitted
LPs by a bounded number of checks, `poll` returns the finished ones.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib
from .algorithms import Algorithm, ConvergenceStats, KKTErrors, PDLP, TerminationStatus
from .problems import PrimalDualSolution


def footprint(m: int, n: int, nnz: int) -> int:
    """HBM bytes an LP of this size takes in a batch arena."""
    return int(lib.cpdlp_batch_footprint(m, n, nnz))


class BatchSolver:
    """A batch arena of `hbm_bytes` on GPU `algo.params.device` solving with `algo`
    (default: PDLP in mixed precision -- FP32 bulk, FP64 checks and polish)."""

    def __init__(self, algo: Algorithm | None = None, hbm_bytes: int = 1 << 30):
        self.algo = algo if algo is not None else PDLP("mixed")
        h = C.c_void_p()
        check(lib.cpdlp_batch_create(C.byref(self.algo.params), int(hbm_bytes), C.byref(h)))
        self.h = h
        self.shapes = {}

    def close(self):
        if getattr(self, "h", None):
            lib.cpdlp_batch_destroy(self.h)
            self.h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def add(self, lp, sol: PrimalDualSolution | None = None) -> int:
        """Admit an LP (a MILP or a generated LP with CSC arrays); returns its id.
        `sol` is an optional unscaled start, as in solve(milp, sol, algo)
        (src/algorithms/common.jl:186-209).  Raises MemoryError when the arena is full
        (retry after release)."""
        if hasattr(lp, "A"):   # MILP
            A = lp.A
            m, n = A.shape
            colptr, rowval, nzval = A.indptr, A.indices, A.data
        else:
            m, n = lp.m, lp.n
            colptr, rowval, nzval = lp.colptr, lp.rowval, lp.nzval
        arrs = [np.ascontiguousarray(colptr, dtype=np.int64), np.ascontiguousarray(rowval, dtype=np.int64),
                np.ascontiguousarray(nzval, dtype=np.float64)]
        vecs = [np.ascontiguousarray(v, dtype=np.float64) for v in (lp.c, lp.lv, lp.uv, lp.lc, lp.uc)]
        c_lp = _lib.CLP()
        c_lp.m, c_lp.n, c_lp.index_base, c_lp.index_bits = m, n, 0, 64
        c_lp.colptr, c_lp.rowval, c_lp.nzval = [a.ctypes.data_as(C.c_void_p) for a in arrs]
        c_lp.c, c_lp.lv, c_lp.uv, c_lp.lc, c_lp.uc = [v.ctypes.data_as(C.c_void_p) for v in vecs]
        i = C.c_int64()
        if sol is None:
            check(lib.cpdlp_batch_add(self.h, C.byref(c_lp), C.byref(i)))
        else:
            x0 = np.ascontiguousarray(sol.x, dtype=np.float64)
            y0 = np.ascontiguousarray(sol.y, dtype=np.float64)
            if x0.shape != (n,) or y0.shape != (m,):
                raise _lib.DimensionMismatch("start point does not match the LP's size")
            check(lib.cpdlp_batch_add_start(self.h, C.byref(c_lp), x0.ctypes.data_as(C.c_void_p),
                                            y0.ctypes.data_as(C.c_void_p), C.byref(i)))
        self.shapes[i.value] = (m, n)
        return i.value

    def run(self, max_checks: int = 10) -> int:
        """Advance every unfinished LP by up to max_checks checks; returns how many
        are still running."""
        r = C.c_int64()
        check(lib.cpdlp_batch_run(self.h, int(max_checks), C.byref(r)))
        return r.value

    def poll(self) -> list:
        ids = np.zeros(max(len(self.shapes), 1), np.int64)
        cnt = C.c_int64()
        check(lib.cpdlp_batch_poll(self.h, ids.ctypes.data_as(C.c_void_p), len(ids), C.byref(cnt)))
        return ids[: cnt.value].tolist()

    def get(self, i: int):
        m, n = self.shapes[i]
        x, y = np.zeros(max(n, 1)), np.zeros(max(m, 1))
        s = _lib.CStats()
        check(lib.cpdlp_batch_get(self.h, int(i), x.ctypes.data_as(C.c_void_p),
                                  y.ctypes.data_as(C.c_void_p), C.byref(s)))
        st = ConvergenceStats(
            err=KKTErrors.from_c(s.err), time_elapsed=s.time_elapsed, kkt_passes=s.kkt_passes,
            termination_status=TerminationStatus(s.termination_status), eta=s.eta, omega=s.omega,
            powm_iters=s.powm_iters, restarts=s.restarts, fp32_passes=s.fp32_passes)
        return PrimalDualSolution(x[:n], y[:m]), st

    def release(self, i: int):
        check(lib.cpdlp_batch_release(self.h, int(i)))
        self.shapes.pop(i, None)

    def memory(self):
        used, size = C.c_int64(), C.c_int64()
        check(lib.cpdlp_batch_memory(self.h, C.byref(used), C.byref(size)))
        return used.value, size.value

    def solve_all(self, lps, checks_per_round: int = 20, sols=None):
        """Solve a list of LPs together (optionally from start points `sols`, one per LP or
        None); returns [(solution, stats)] in input order."""
        sols = [None] * len(lps) if sols is None else list(sols)
        if len(sols) != len(lps):
            raise _lib.DimensionMismatch("one start point (or None) per LP")
        ids = [self.add(lp, s) for lp, s in zip(lps, sols)]
        while self.run(checks_per_round) > 0:
            pass
        out = [self.get(i) for i in ids]
        for i in ids:
            self.release(i)
        return out
