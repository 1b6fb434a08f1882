"""Batched and served solves: many independent LPs on one GPU (cpdlp_batch_* C ABI).

The reference solves one LP per `solve` call (src/algorithms/common.jl:186-209); a
BatchSolver runs that same PDLP algorithm on many LPs at once, one workgroup per LP,
with every per-LP decision (scaling, step size, restarts, termination) on the device.
LPs can be added while others run (continuous batching): `run` advances all admitted
LPs by a bounded number of checks, `poll` returns the finished ones.
"""
from __future__ import annotations

import concurrent.futures as cf
import ctypes as C
import copy
import dataclasses

import numpy as np

from . import _lib
from ._lib import check, lib
from .algorithms import Algorithm, ConvergenceStats, KKTErrors, LoopState, PDLP, TerminationStatus
from .problems import PrimalDualSolution, require_unpreconditioned


def footprint(m: int, n: int, nnz: int) -> int:
    """HBM bytes an LP of this size takes in a batch arena."""
    return int(lib.cpdlp_batch_footprint(m, n, nnz))


class BatchSolver:
    """A batch arena of `hbm_bytes` on GPU `algo.params.device` solving with `algo`
    (default: PDLP in mixed precision -- FP32 bulk, FP64 checks and polish).

    Tail continuation (`tail` > 0, off by default): a batch LP runs on one workgroup, so
    the last few unfinished LPs leave the device nearly idle.  Once at most `tail` LPs are
    unfinished, every one of them that has run at least `tail_min_passes` passes leaves
    the batch with its loop state (take_state) and continues on a plain FP64 engine (CSR,
    Engine.resume, many CUs per LP) on one of `tail_workers` threads; poll and get report
    it like any other.  A continued LP's result comes from a different engine than the
    batch's, so it is not the batch-alone result bit for bit; its pass count includes the
    batch's passes, and its time limit counts again from the hand-over."""

    def __init__(self, algo: Algorithm | None = None, hbm_bytes: int = 1 << 30, *, tail: int = 0,
                 tail_min_passes: int = 2000, tail_workers: int = 8):
        self.algo = algo if algo is not None else PDLP("mixed")
        h = C.c_void_p()
        check(lib.cpdlp_batch_create(C.byref(self.algo.params), int(hbm_bytes), C.byref(h)))
        self.h = h
        self.shapes = {}
        self.tail, self.tail_min_passes = int(tail), int(tail_min_passes)
        self._lps = {}           # id -> LP, kept for a continuation
        self._cont = {}          # id -> future of (solution, stats) on a plain engine
        self._cont_reported = set()
        self._pool = cf.ThreadPoolExecutor(max(1, int(tail_workers))) if self.tail > 0 else None

    def close(self):
        if getattr(self, "_pool", None):
            self._pool.shutdown(wait=True)   # the plain engines finish first
            self._pool = None
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
        require_unpreconditioned(lp)
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
        if self.tail > 0:
            self._lps[i.value] = lp
        return i.value

    def run(self, max_checks: int = 10) -> int:
        """Advance every unfinished LP by up to max_checks checks; returns how many
        are still running (continued ones included)."""
        r = C.c_int64()
        check(lib.cpdlp_batch_run(self.h, int(max_checks), C.byref(r)))
        running = r.value
        if self.tail > 0:
            if 0 < running + self._cont_running() <= self.tail:
                running -= self._continue_stragglers(self.tail_min_passes)
            pending = [f for f in self._cont.values() if not f.done()]
            if running == 0 and pending:   # only continued LPs left: wait for one
                cf.wait(pending, timeout=0.05, return_when=cf.FIRST_COMPLETED)
            running += self._cont_running()
        return running

    def _cont_running(self) -> int:
        return sum(1 for f in self._cont.values() if not f.done())

    def _stats_only(self, i: int) -> ConvergenceStats:
        s = _lib.CStats()
        check(lib.cpdlp_batch_get(self.h, int(i), None, None, C.byref(s)))
        return ConvergenceStats(
            err=KKTErrors.from_c(s.err), time_elapsed=s.time_elapsed, kkt_passes=s.kkt_passes,
            termination_status=TerminationStatus(s.termination_status), eta=s.eta, omega=s.omega,
            powm_iters=s.powm_iters, restarts=s.restarts, fp32_passes=s.fp32_passes)

    def _continue_stragglers(self, min_passes: int) -> int:
        """Hand the unfinished LPs past min_passes passes to plain engines; returns how many."""
        k = 0
        for i in list(self.shapes):
            if i in self._cont:
                continue
            st0 = self._stats_only(i)
            if st0.termination_status != TerminationStatus.STILL_RUNNING or st0.kkt_passes < min_passes:
                continue
            shape = self.shapes[i]
            state = self.take_state(i)
            self.shapes[i] = shape
            self._cont[i] = self._pool.submit(self._continue, self._lps.pop(i), state, st0)
            k += 1
        return k

    def _continue(self, lp, state: LoopState, st0: ConvergenceStats):
        from .algorithms import Engine
        from .problems import MILP
        milp = lp if hasattr(lp, "A") else MILP.from_generated(lp)
        algo = copy.copy(self.algo)
        algo.params = _lib.CParams.from_buffer_copy(self.algo.params)
        algo.params.precision = _lib.PREC_FP64
        algo.params.format = _lib.FMT_CSR
        with Engine(milp, algo) as e:
            e.resume(state)
            while e.run_checks(50) == TerminationStatus.STILL_RUNNING:
                pass
            sol, st = e.solution(), e.stats()
        return sol, dataclasses.replace(st, time_elapsed=st0.time_elapsed + st.time_elapsed,
                                        powm_iters=st0.powm_iters)

    def poll(self) -> list:
        ids = np.zeros(max(len(self.shapes), 1), np.int64)
        cnt = C.c_int64()
        check(lib.cpdlp_batch_poll(self.h, ids.ctypes.data_as(C.c_void_p), len(ids), C.byref(cnt)))
        out = ids[: cnt.value].tolist()
        for i, f in self._cont.items():
            if f.done() and i not in self._cont_reported:
                self._cont_reported.add(i)
                out.append(i)
        return out

    def continued(self, i: int) -> bool:
        """Whether LP `i` was handed over to a plain engine (tail continuation)."""
        return i in self._cont

    def get(self, i: int):
        if i in self._cont:
            return self._cont[i].result()   # re-raises a failure of the plain engine
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
        if i in self._cont:
            self._cont.pop(i).result()
            self._cont_reported.discard(i)
        else:
            check(lib.cpdlp_batch_release(self.h, int(i)))
            self._lps.pop(i, None)
        self.shapes.pop(i, None)

    def take_state(self, i: int) -> LoopState:
        """Take a running instance out of the batch with its loop state
        (cpdlp_batch_take_state), for Engine.resume on a plain engine of the same LP; the
        id is then unknown to the batch."""
        if i not in self.shapes:
            raise _lib.ArgumentError("unknown batch instance id")
        m, n = self.shapes[i]
        x, y, D2, D1 = LoopState.buffers(m, n)
        s = _lib.CState()
        check(lib.cpdlp_batch_take_state(self.h, int(i), C.byref(s), x.ctypes.data_as(C.c_void_p),
                                         y.ctypes.data_as(C.c_void_p), D2.ctypes.data_as(C.c_void_p),
                                         D1.ctypes.data_as(C.c_void_p)))
        self.shapes.pop(i, None)
        return LoopState.from_c(s, x[:, :n], y[:, :m], D2[:n], D1[:m])

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
