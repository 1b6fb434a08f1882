"""Serving: continuous batching of LP solve requests on one GPU.

Small requests (m + n <= batch_limit) are admitted into the batch engine's HBM arena as
room frees up (footprint accounting, cpdlp_batch_footprint) and advanced in rounds of
PDHG checks; every round, finished requests are harvested and their HBM released, and
waiting requests are admitted -- so requests arriving while others are on the GPU join
the running batch (their work overlaps on the device).  Rounds are work-budgeted by the
batch engine: a large instance runs fewer checks per round than small ones, so a small
request is not held back by a larger neighbour.  Larger requests run on the plain
engine (one HIP stream each) in worker threads, in lanes by size with their own queues,
under an HBM budget: one workgroup per LP is the right shape for many small LPs, while
a medium or large LP solves much faster on the whole GPU.  Iterates, scalings and
restart state stay on the device between rounds.

The first-come-first-served baseline (`fcfs_latencies`) is the one-solve-at-a-time
queue: request k starts when both it has arrived and request k-1 has finished, and
takes its measured plain-engine service time.
"""
from __future__ import annotations

import threading
import time
from collections import deque
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from ._lib import ArgumentError, DimensionMismatch
from .algorithms import Engine, PDLP
from .batch import BatchSolver, footprint
from .problems import MILP, require_unpreconditioned


def plain_footprint(m: int, n: int, nnz: int) -> int:
    """Estimated HBM bytes of one plain-engine solve (matrices in two layouts, FP64 and
    FP32 values, vectors and iterates)."""
    return int(40 * nnz + 250 * (m + n))


@dataclass
class Request:
    rid: int
    lp: object
    t_arrival: float
    route: str = "batch"
    t_done: float = float("nan")
    t_start: float = float("nan")      # admitted to the batch arena / taken by a plain worker
    status: str = ""
    objective: float = float("nan")
    kkt_passes: int = 0
    x: np.ndarray | None = None
    y: np.ndarray | None = None
    sol: object = None                 # optional start point (PrimalDualSolution)


@dataclass
class Lane:
    """A plain-engine route: its own queue (shortest job first) and worker threads, for
    requests of at most `max_nnz` nonzeros, so that a long solve in one lane never holds
    back the requests of another."""
    max_nnz: float
    workers: int
    waiting: deque = field(default_factory=deque)
    running: int = 0
    pool: ThreadPoolExecutor | None = None


@dataclass
class Server:
    reltol: float = 1e-6
    precision: str = "mixed"
    batch_bytes: int = 16 << 30
    plain_bytes: int = 64 << 30
    checks_per_round: int = 20         # checks of the smallest instances per round
    batch_streams: int = 1             # + one stream per plain worker: 4 hardware queues
    batch_limit: int = 8192            # m + n limit of the batch route
    # when > 0: a request for a plain lane that already has `overflow_queue` x workers
    # waiting goes to the batch engine instead if it fits there (m + n <= 65536).  Off by
    # default: medium LPs in the batch lengthen the rounds of the small ones.
    overflow_queue: int = 0
    # plain lanes (max nonzeros, workers): medium requests, then everything larger
    lanes: tuple = ((1_000_000, 3), (float("inf"), 1))
    time_limit: float = 600.0
    keep_solutions: bool = False
    # opt-in: infeasible / unbounded / NaN requests stop at the check that proves it
    # (PRIMAL_INFEASIBLE, DUAL_INFEASIBLE, NUMERICAL_ERROR) instead of at a limit
    detect_infeasibility: bool = False
    # power-method tolerance of every request (NaN: the reference's eps(T)*n^3 rule, which
    # stops after one Rayleigh quotient for n >~ 3e5 and then diverges)
    powm_tol: float = float("nan")
    done: list = field(default_factory=list)

    def __post_init__(self):
        algo = PDLP(self.precision, termination_reltol=self.reltol, time_limit=self.time_limit,
                    batch_streams=self.batch_streams, detect_infeasibility=int(self.detect_infeasibility),
                    powm_tol=self.powm_tol)
        self.batch = BatchSolver(algo, hbm_bytes=self.batch_bytes)
        self.waiting: deque = deque()
        self.in_batch: dict = {}
        self.plain_lanes = [Lane(float(mx), int(w)) for mx, w in self.lanes]
        for ln in self.plain_lanes:
            ln.pool = ThreadPoolExecutor(ln.workers)
        self.plain_used = 0
        self.lock = threading.Lock()
        self.next_id = 0
        self.rounds = 0
        self.round_seconds = []            # wall time of each batch round (run + harvest)

    def close(self):
        for ln in self.plain_lanes:
            ln.pool.shutdown(wait=True)
        self.batch.close()

    # -------------------------------------------------------------- submission
    def submit(self, lp, t_arrival: float | None = None, sol=None) -> int:
        """Queue an LP; `sol` is an optional start point, as in solve(milp, sol, algo)
        (e.g. the solution of an earlier request for a nearby LP, kept with
        keep_solutions)."""
        require_unpreconditioned(lp)
        t = time.perf_counter() if t_arrival is None else t_arrival
        r = Request(self.next_id, lp, t, sol=sol)
        self.next_id += 1
        k = route(lp.m, lp.n, lp.nnz, self.batch_limit, self.plain_lanes, self.overflow_queue)
        if k < 0:
            self.waiting.append(r)
        else:
            r.route = "plain"
            self.plain_lanes[k].waiting.append(r)
        return r.rid

    def pending(self) -> int:
        with self.lock:
            return (len(self.waiting) + len(self.in_batch)
                    + sum(len(ln.waiting) + ln.running for ln in self.plain_lanes))

    # -------------------------------------------------------------- plain route
    def _plain_job(self, r: Request, fp: int, lane: Lane):
        try:
            milp = MILP.from_generated(r.lp)
            # fixed CSR format: a request's result must not depend on autotuner timing
            algo = PDLP(self.precision, None, "csr", termination_reltol=self.reltol,
                        time_limit=self.time_limit, detect_infeasibility=int(self.detect_infeasibility),
                        powm_tol=self.powm_tol)
            with Engine(milp, algo) as eng:
                sol, st = eng.solve(r.sol)
            r.status = st.termination_status.name
            r.kkt_passes = st.kkt_passes
            r.objective = float(sol.x @ r.lp.c)
            if self.keep_solutions:
                r.x, r.y = sol.x, sol.y
        except Exception as e:   # noqa: BLE001 -- a failed request is reported, not fatal
            r.status = f"ERROR: {e}"
        r.t_done = time.perf_counter()
        with self.lock:
            lane.running -= 1
            self.plain_used -= fp
            self.done.append(r)

    def _admit_plain(self):
        for lane in self.plain_lanes:
            # shortest job first (by nonzeros) within the lane
            if len(lane.waiting) > 1:
                lane.waiting = deque(sorted(lane.waiting, key=lambda q: q.lp.nnz))
            while lane.waiting and lane.running < lane.workers:
                r = lane.waiting[0]
                fp = plain_footprint(r.lp.m, r.lp.n, r.lp.nnz)
                with self.lock:
                    busy = sum(ln.running for ln in self.plain_lanes)
                    if busy > 0 and self.plain_used + fp > self.plain_bytes:
                        break
                    lane.running += 1
                    self.plain_used += fp
                lane.waiting.popleft()
                r.t_start = time.perf_counter()
                lane.pool.submit(self._plain_job, r, fp, lane)

    # -------------------------------------------------------------- batch route
    def _admit_batch(self):
        while self.waiting:
            r = self.waiting[0]
            try:
                bid = self.batch.add(r.lp, r.sol)
            except MemoryError:
                return   # no room until a running request finishes
            except (ArgumentError, DimensionMismatch) as e:   # never admissible: reported, not retried
                self.waiting.popleft()
                r.status = f"ERROR: {e}"
                r.t_done = time.perf_counter()
                with self.lock:
                    self.done.append(r)
                continue
            self.waiting.popleft()
            r.t_start = time.perf_counter()
            self.in_batch[bid] = r

    def pump(self):
        """One scheduling round: admit, advance the batch, harvest."""
        self._admit_plain()
        self._admit_batch()
        if self.in_batch:
            t0 = time.perf_counter()
            self.batch.run(self.checks_per_round)
            now = time.perf_counter()
            self.rounds += 1
            self.round_seconds.append(now - t0)
            for bid in self.batch.poll():
                r = self.in_batch.pop(bid)
                sol, st = self.batch.get(bid)
                self.batch.release(bid)
                r.t_done = now
                r.status = st.termination_status.name
                r.kkt_passes = st.kkt_passes
                r.objective = float(sol.x @ r.lp.c)
                if self.keep_solutions:
                    r.x, r.y = sol.x, sol.y
                with self.lock:
                    self.done.append(r)
        else:
            time.sleep(0.0005)

    def drain(self, deadline: float):
        while self.pending() > 0 and time.perf_counter() < deadline:
            self.pump()


class NativeServer:
    """The same continuous batching as `Server`, run by the library itself
    (cpdlp_server_*): requests are copied at submit; small ones (m + n <= batch_limit)
    go to the batch engine, medium ones (m + n <= batch2_limit) to a second batch engine
    with rounds of its own, both advanced by `pump` (or the library's threads after
    `start`); larger ones to plain engines on the library's worker threads, in lanes of
    (max nonzeros, workers), shortest job first, under an HBM budget.  batch2_limit <=
    batch_limit (the default) turns the second batch route off.  This is the interface a
    C or Julia caller uses."""

    def __init__(self, algo=None, *, batch_bytes: int = 16 << 30, plain_bytes: int = 64 << 30,
                 batch_limit: int = 8192, checks_per_round: int = 20,
                 lanes=((1_000_000, 3), (float("inf"), 1)),
                 batch2_limit: int = 0, batch2_bytes: int = 2 << 30, batch2_checks: int = 0):
        import ctypes as C
        from . import _lib
        self._C, self._lib = C, _lib
        # one stream per batch route beside the plain workers' streams keeps the total
        # within the device's 4 hardware queues
        self.algo = algo if algo is not None else PDLP("mixed", batch_streams=1)
        cfg = _lib.CServerConfig()
        _lib.check(_lib.lib.cpdlp_server_default_config(C.byref(cfg)))
        cfg.batch_bytes, cfg.plain_bytes = int(batch_bytes), int(plain_bytes)
        cfg.batch_limit, cfg.checks_per_round = int(batch_limit), int(checks_per_round)
        cfg.batch2_limit, cfg.batch2_bytes = int(batch2_limit), int(batch2_bytes)
        cfg.batch2_checks = int(batch2_checks)
        if not 1 <= len(lanes) <= _lib.SERVER_MAX_LANES:
            raise _lib.ArgumentError(f"1 to {_lib.SERVER_MAX_LANES} lanes")
        cfg.n_lanes = len(lanes)
        for k, (mx, w) in enumerate(lanes):
            cfg.lane_max_nnz[k], cfg.lane_workers[k] = float(mx), int(w)
        h = C.c_void_p()
        _lib.check(_lib.lib.cpdlp_server_create(C.byref(self.algo.params), C.byref(cfg), C.byref(h)))
        self.h = h
        self.shapes = {}
        self._held = {}          # arrays lent to the library, by request id

    def close(self):
        if getattr(self, "h", None):
            self._lib.lib.cpdlp_server_destroy(self.h)   # joins the workers first
            self.h = None
            self._held.clear()

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, lp, sol=None, borrow: bool = True) -> int:
        """Queue an LP (a MILP or a generated LP); `sol` is an optional unscaled start.
        With `borrow` (the default) the library reads the arrays in place
        (cpdlp_server_submit_borrowed): they are held here until `get` and must not be
        modified meanwhile.  Otherwise the library copies them at submit."""
        C, _lib = self._C, self._lib
        require_unpreconditioned(lp)
        if hasattr(lp, "A"):
            A = lp.A
            m, n = A.shape
            colptr, rowval, nzval = A.indptr, A.indices, A.data
        else:
            m, n = lp.m, lp.n
            colptr, rowval, nzval = lp.colptr, lp.rowval, lp.nzval
        bits = 32 if int(colptr[-1]) < 2**31 else 64   # int32 row indices pass without a copy
        it = np.int32 if bits == 32 else np.int64
        arrs = [np.ascontiguousarray(colptr, dtype=it), np.ascontiguousarray(rowval, dtype=it),
                np.ascontiguousarray(nzval, dtype=np.float64)]
        vecs = [np.ascontiguousarray(v, dtype=np.float64) for v in (lp.c, lp.lv, lp.uv, lp.lc, lp.uc)]
        c_lp = _lib.CLP()
        c_lp.m, c_lp.n, c_lp.index_base, c_lp.index_bits = m, n, 0, bits
        c_lp.colptr, c_lp.rowval, c_lp.nzval = [a.ctypes.data_as(C.c_void_p) for a in arrs]
        c_lp.c, c_lp.lv, c_lp.uv, c_lp.lc, c_lp.uc = [v.ctypes.data_as(C.c_void_p) for v in vecs]
        x0 = y0 = None
        if sol is not None:
            x0 = np.ascontiguousarray(sol.x, dtype=np.float64)
            y0 = np.ascontiguousarray(sol.y, dtype=np.float64)
            if x0.shape != (n,) or y0.shape != (m,):
                raise _lib.DimensionMismatch("start point does not match the LP's size")
        i = C.c_int64()
        fn = _lib.lib.cpdlp_server_submit_borrowed if borrow else _lib.lib.cpdlp_server_submit
        _lib.check(fn(self.h, C.byref(c_lp), None if x0 is None else x0.ctypes.data_as(C.c_void_p),
                      None if y0 is None else y0.ctypes.data_as(C.c_void_p), C.byref(i)))
        self.shapes[i.value] = (m, n)
        if borrow:
            self._held[i.value] = (arrs, vecs, x0, y0)
        return i.value

    def start(self):
        """Let the library pump on its own thread (cpdlp_server_start); `pump` then only
        reports the pending count."""
        self._lib.check(self._lib.lib.cpdlp_server_start(self.h))

    def pump(self) -> int:
        """One scheduling round; returns how many requests are not finished."""
        k = self._C.c_int64()
        self._lib.check(self._lib.lib.cpdlp_server_pump(self.h, self._C.byref(k)))
        return k.value

    def poll(self) -> list:
        C = self._C
        ids = np.zeros(max(len(self.shapes), 1), np.int64)
        cnt = C.c_int64()
        self._lib.check(self._lib.lib.cpdlp_server_poll(self.h, ids.ctypes.data_as(C.c_void_p), len(ids),
                                                        C.byref(cnt)))
        return ids[: cnt.value].tolist()

    def get(self, i: int):
        """(PrimalDualSolution, ConvergenceStats, seconds from submit to finish, route:
        -1 batch or the plain lane); frees the request."""
        from .algorithms import ConvergenceStats, KKTErrors, TerminationStatus
        from .problems import PrimalDualSolution
        C, _lib = self._C, self._lib
        if i not in self.shapes:
            raise _lib.ArgumentError("unknown request id")
        m, n = self.shapes[i]
        x, y = np.zeros(max(n, 1)), np.zeros(max(m, 1))
        s, secs, rt = _lib.CStats(), C.c_double(), C.c_int32()
        rc = _lib.lib.cpdlp_server_get(self.h, int(i), x.ctypes.data_as(C.c_void_p),
                                       y.ctypes.data_as(C.c_void_p), C.byref(s), C.byref(secs), C.byref(rt))
        running = (rc == _lib.ERR_STATE
                   and _lib.lib.cpdlp_last_error().decode(errors="replace") == "request still running")
        if not running:
            # the request is finished (solved or failed) and freed by the library, which
            # no longer reads a borrowed request's arrays; while it still runs (ERR_STATE)
            # they stay held and the id stays known
            self.shapes.pop(i, None)
            self._held.pop(i, None)
        _lib.check(rc)
        st = ConvergenceStats(
            err=KKTErrors.from_c(s.err), time_elapsed=s.time_elapsed, kkt_passes=s.kkt_passes,
            termination_status=TerminationStatus(s.termination_status), eta=s.eta, omega=s.omega,
            powm_iters=s.powm_iters, restarts=s.restarts, fp32_passes=s.fp32_passes)
        return PrimalDualSolution(x[:n], y[:m]), st, secs.value, rt.value

    def drain(self, deadline: float):
        """Pump until every request has finished or the deadline (perf_counter) passes."""
        while self.pump() > 0 and time.perf_counter() < deadline:
            time.sleep(0.0005)


def route(m: int, n: int, nnz: int, batch_limit: int, lanes, overflow_queue: int = 0) -> int:
    """Where a request goes: -1 for the batch engine, else the index of its plain lane
    (the first whose max_nnz admits it, the last lane otherwise).  With overflow_queue > 0
    a request whose lane already has overflow_queue x workers waiting goes to the batch
    engine instead when it fits there (m + n <= 65536)."""
    if m + n <= batch_limit:
        return -1
    k = next((i for i, ln in enumerate(lanes) if nnz <= ln.max_nnz), len(lanes) - 1)
    ln = lanes[k]
    if overflow_queue > 0 and m + n <= 65536 and len(ln.waiting) >= overflow_queue * ln.workers:
        return -1
    return k


def fcfs_latencies(arrivals: np.ndarray, service: np.ndarray) -> np.ndarray:
    """Latencies of a first-come-first-served single server: start_k = max(arrival_k,
    finish_{k-1}), finish_k = start_k + service_k."""
    lat = np.empty(len(arrivals))
    free = -np.inf
    for k, (a, s) in enumerate(zip(arrivals, service)):
        start = max(a, free)
        free = start + s
        lat[k] = free - a
    return lat


def poisson_arrivals(qps: float, duration: float, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / qps)
        if t >= duration:
            return np.array(out)
        out.append(t)
