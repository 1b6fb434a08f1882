"""Seeded synthetic LP generators for the benchmark workloads and tests.

Every generated LP comes with a known optimal primal-dual point (x0, y0, s0), so
its optimal objective is the closed form ``c'x0 = b'y0``.  Sign convention is the
reference's: for a row ``lc <= a'x <= uc`` the multiplier lies in
``proj_multiplier(y, lc, uc)`` (src/utils/linalg.jl:26-42), so a ``<=`` row
(lc = -inf) has ``y <= 0``; a variable with lv = 0, uv = inf has reduced cost
``r = c - A'y >= 0``.

Construction (shared by all workloads):
  * x0 ~ U(0,1) on a random 60% support of the columns, 0 elsewhere; b = A x0.
  * Rows: a random half are equalities (lc = uc = b), the rest are ``<=`` rows
    (lc = -inf, uc = b), so every row is active at x0.
  * y0 ~ N(0,1), with y0 <= 0 forced on ``<=`` rows.
  * s0 ~ U(0,1) off the support, 0 on it;  c = A'y0 + s0.
  * Bounds 0 <= x <= inf (optionally 0 <= x <= ub on a fraction of columns; x0 <= 1
    keeps those upper bounds inactive).
Then (x0, y0, r = s0) satisfies every KKT condition of the reference's
``kkt_errors!`` (src/components/errors.jl:65-109) exactly in real arithmetic.

Matrices are returned in CSC form (colptr int64, rowval int32, nzval float64,
0-based, rows sorted within each column), the reference's input format
(SparseMatrixCSC, src/problems/milp.jl:20-115).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp


@dataclass
class GeneratedLP:
    m: int
    n: int
    colptr: np.ndarray
    rowval: np.ndarray
    nzval: np.ndarray
    c: np.ndarray
    lv: np.ndarray
    uv: np.ndarray
    lc: np.ndarray
    uc: np.ndarray
    x0: np.ndarray | None = None
    y0: np.ndarray | None = None
    objective: float = float("nan")
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def nnz(self) -> int:
        return int(self.colptr[-1])

    def csc(self) -> sp.csc_matrix:
        return sp.csc_matrix((self.nzval, self.rowval, self.colptr), shape=(self.m, self.n))


def _dedupe(major: np.ndarray, minor: np.ndarray, minor_size: int, rng) -> np.ndarray:
    """Redraw ``minor`` entries that repeat within the same ``major`` group."""
    minor = minor.astype(np.int64, copy=True)
    for _ in range(64):
        key = major.astype(np.int64) * minor_size + minor
        order = np.argsort(key, kind="stable")
        ks = key[order]
        dup = np.zeros(len(ks), dtype=bool)
        dup[1:] = ks[1:] == ks[:-1]
        if not dup.any():
            return minor
        idx = order[dup]
        minor[idx] = rng.integers(0, minor_size, size=len(idx))
    raise RuntimeError("could not draw distinct indices")


def _csc_from_pairs(m: int, n: int, rows: np.ndarray, cols: np.ndarray, vals: np.ndarray):
    key = cols.astype(np.int64) * m + rows.astype(np.int64)
    order = np.argsort(key, kind="stable")
    rows_s = rows[order].astype(np.int32)
    cols_s = cols[order]
    vals_s = vals[order].astype(np.float64)
    counts = np.bincount(cols_s, minlength=n)
    colptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=colptr[1:])
    return colptr, rows_s, vals_s


def _finish(m, n, colptr, rowval, nzval, rng, eq_frac=0.5, box_frac=0.0, box_ub=10.0,
            support_frac=0.6, name="", meta=None) -> GeneratedLP:
    A = sp.csc_matrix((nzval, rowval, colptr), shape=(m, n))
    x0 = np.zeros(n)
    n_sup = int(round(support_frac * n))
    sup = rng.permutation(n)[:n_sup]
    x0[sup] = rng.random(n_sup)
    b = A @ x0
    is_eq = np.zeros(m, dtype=bool)
    is_eq[rng.permutation(m)[: int(round(eq_frac * m))]] = True
    lc = np.where(is_eq, b, -np.inf)
    uc = b.copy()
    y0 = rng.standard_normal(m)
    y0 = np.where(is_eq, y0, -np.abs(y0))
    s0 = rng.random(n)
    s0[sup] = 0.0
    c = A.T @ y0 + s0
    lv = np.zeros(n)
    uv = np.full(n, np.inf)
    if box_frac > 0:
        boxed = rng.permutation(n)[: int(round(box_frac * n))]
        uv[boxed] = box_ub
    obj = float(c @ x0)
    return GeneratedLP(m, n, colptr, rowval, nzval, c, lv, uv, lc, uc, x0, y0, obj, name,
                       dict(meta or {}, dual_objective=float(b @ y0)))


def column_lp(m: int, n: int, nnz_per_col, seed: int, eq_frac=0.5, box_frac=0.0,
              name="column_lp") -> GeneratedLP:
    """Random LP with a fixed (or per-column random) count of nonzeros per column at
    uniform random distinct rows, N(0,1) values.  ``nnz_per_col`` is an int or a
    (lo, hi) inclusive range drawn per column."""
    rng = np.random.default_rng(seed)
    if isinstance(nnz_per_col, tuple):
        k = rng.integers(nnz_per_col[0], nnz_per_col[1] + 1, size=n)
    else:
        k = np.full(n, int(nnz_per_col))
    k = np.minimum(k, m)
    cols = np.repeat(np.arange(n, dtype=np.int64), k)
    rows = rng.integers(0, m, size=len(cols))
    rows = _dedupe(cols, rows, m, rng)
    vals = rng.standard_normal(len(cols))
    colptr, rowval, nzval = _csc_from_pairs(m, n, rows, cols, vals)
    return _finish(m, n, colptr, rowval, nzval, rng, eq_frac=eq_frac, box_frac=box_frac,
                   name=name, meta=dict(seed=seed))


def small_dense_lp(m=2_000, n=5_000, seed=0) -> GeneratedLP:
    """The reference's CPU-runnable case: m=2,000 (half =, half <=), n=5,000, 8 nonzeros
    per column, 0 <= x < inf."""
    return column_lp(m, n, 8, seed, name=f"plain_lp_m{m}_n{n}")


def zipf_row_lengths(m: int, target_nnz: int, exponent: float, cap: int, rng) -> np.ndarray:
    """Row lengths: Z ~ Zipf(exponent) truncated to [1, cap], then scaled by one common
    factor f so that sum(clamp(round(f*Z), 1, cap)) is as close to ``target_nnz`` as
    bisection on f gets (Zipf(1.5) capped at 5,000 has mean ~54, so f ~ 0.37 for a
    mean of 20)."""
    k = np.arange(1, cap + 1, dtype=np.float64)
    p = k ** (-exponent)
    cdf = np.cumsum(p / p.sum())
    z = np.searchsorted(cdf, rng.random(m), side="right") + 1
    z = np.minimum(z, cap).astype(np.float64)
    lo, hi = 1e-6, float(cap)
    for _ in range(80):
        f = 0.5 * (lo + hi)
        s = np.clip(np.rint(f * z), 1, cap).sum()
        if s < target_nnz:
            lo = f
        else:
            hi = f
    return np.clip(np.rint(hi * z), 1, cap).astype(np.int64)


def power_law_lp(m=1_000_000, n=2_000_000, target_nnz=20_000_000, exponent=1.5, cap=5_000,
                 box_frac=0.3, seed=1) -> GeneratedLP:
    """Single large LP: power-law row lengths (Zipf exponent 1.5, capped at 5,000 per
    row, rescaled to ~20M nonzeros), uniform random distinct columns per row,
    N(0,1) values, box bounds 0 <= x <= 10 on 30% of the columns."""
    rng = np.random.default_rng(seed)
    L = zipf_row_lengths(m, target_nnz, exponent, cap, rng)
    rows = np.repeat(np.arange(m, dtype=np.int64), L)
    cols = rng.integers(0, n, size=len(rows))
    cols = _dedupe(rows, cols, n, rng)
    vals = rng.standard_normal(len(rows))
    colptr, rowval, nzval = _csc_from_pairs(m, n, rows, cols, vals)
    return _finish(m, n, colptr, rowval, nzval, rng, box_frac=box_frac,
                   name=f"power_law_lp_m{m}_n{n}",
                   meta=dict(seed=seed, max_row=int(L.max()), mean_row=float(L.mean())))


def batch_specs(count=4096, m_lo=50, m_hi=5_000, seed=3):
    """Shapes of the uneven batch: m log-uniform in [m_lo, m_hi], n = round(m*U(1.5,3)),
    per-instance seeds seed*1_000_003 + i."""
    rng = np.random.default_rng(seed)
    ms = np.rint(np.exp(rng.uniform(math.log(m_lo), math.log(m_hi), size=count))).astype(np.int64)
    ns = np.rint(ms * rng.uniform(1.5, 3.0, size=count)).astype(np.int64)
    return [(int(ms[i]), int(ns[i]), seed * 1_000_003 + i) for i in range(count)]


def batch_lps(count=4096, m_lo=50, m_hi=5_000, seed=3):
    """The 4,096 uneven LPs: 6-12 nonzeros per column, same feasible construction."""
    return [column_lp(m, n, (6, 12), s, name=f"batch_{i}")
            for i, (m, n, s) in enumerate(batch_specs(count, m_lo, m_hi, seed))]


def request_sizes(count: int, seed=5):
    """Serving request mix: 70% m log-uniform in [50, 2k], 25% in [2k, 50k],
    5% in [50k, 500k]; n = 2m; 8 nonzeros per column."""
    rng = np.random.default_rng(seed)
    u = rng.random(count)
    lo = np.where(u < 0.70, 50, np.where(u < 0.95, 2_000, 50_000))
    hi = np.where(u < 0.70, 2_000, np.where(u < 0.95, 50_000, 500_000))
    ms = np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi)))).astype(np.int64)
    return [(int(mm), 2 * int(mm), seed * 7_919 + i) for i, mm in enumerate(ms)]


def request_lp(m: int, n: int, seed: int) -> GeneratedLP:
    return column_lp(m, n, 8, seed, name=f"req_m{m}")


def random_milp_and_sol(m: int, n: int, p: float, seed: int):
    """Mirror of the reference's test generator (src/utils/test.jl:1-50): A = sprandn,
    each bound 25% equal (lower == upper), 25% infinite, 50% finite gap; x, y random
    and projected onto the bounds / multiplier sets.  Returns (GeneratedLP, x, y)."""
    rng = np.random.default_rng(seed)
    c = rng.random(n)
    A = sp.random(m, n, density=p, format="csc", random_state=rng,
                  data_rvs=lambda k: rng.standard_normal(k))
    A.sort_indices()

    def bounds(z, lower):
        r = rng.random(len(z))
        gap = rng.random(len(z))
        inf = -np.inf if lower else np.inf
        shifted = z - gap if lower else z + gap
        return np.where(r < 0.25, z, np.where(r < 0.5, inf, shifted))

    luv = rng.standard_normal(n)
    luc = rng.standard_normal(m)
    lv, uv = bounds(luv, True), bounds(luv, False)
    lc, uc = bounds(luc, True), bounds(luc, False)
    x = np.minimum(uv, np.maximum(lv, rng.standard_normal(n)))
    yr = rng.standard_normal(m)
    lmin, umax = np.isneginf(lc), np.isposinf(uc)
    y = np.where(lmin, np.where(umax, 0.0, np.minimum(yr, 0.0)), np.where(umax, np.maximum(yr, 0.0), yr))
    lp = GeneratedLP(m, n, A.indptr.astype(np.int64), A.indices.astype(np.int32),
                     A.data.astype(np.float64), c, lv, uv, lc, uc, name="random_milp")
    return lp, x, y


def min_cost_flow_lp(nodes=500_000, out_degree=8, src_frac=0.01, seed=2,
                     check_feasible=True) -> GeneratedLP:
    """Min-cost flow: random directed graph, each node with ``out_degree`` arcs to
    uniform random heads (no self loops); A = node-arc incidence (+1 at the tail,
    -1 at the head), A x = supply, 0 <= x <= capacity.  Capacities integer U{1..100},
    costs integer U{1..10}; 1% of nodes are sources and 1% sinks with magnitudes
    U(10, 1000), the sinks rescaled so supplies sum to zero.  Each source's supply and
    each sink's demand is then replaced by 90% of what a max-flow from a super source
    to a super sink routes through it, so the LP is feasible by construction (a scaled
    feasible flow).  No closed-form optimum: the objective is checked against an FP64
    solve and, at small scale, HiGHS."""
    rng = np.random.default_rng(seed)
    E = nodes * out_degree
    tails = np.repeat(np.arange(nodes, dtype=np.int64), out_degree)
    heads = rng.integers(0, nodes - 1, size=E)
    heads = heads + (heads >= tails)                      # skip self loops
    cap = rng.integers(1, 101, size=E).astype(np.float64)
    cost = rng.integers(1, 11, size=E).astype(np.float64)
    perm = rng.permutation(nodes)
    k = max(1, int(round(src_frac * nodes)))
    sources, sinks = perm[:k], perm[k:2 * k]
    supply = np.zeros(nodes)
    supply[sources] = rng.uniform(10, 1000, size=k)
    dem = rng.uniform(10, 1000, size=k)
    supply[sinks] = -dem * (supply[sources].sum() / dem.sum())
    scale = 1.0
    if check_feasible:
        # route the supplies with a max-flow from a super source to a super sink (in
        # units of 1/100, integer capacities) and keep, for every source and sink, what
        # the flow actually carries through it, times 0.9: a scaled feasible flow, so
        # the LP is feasible by construction
        from scipy.sparse.csgraph import maximum_flow
        S, T = nodes, nodes + 1
        unit = 100
        r = np.concatenate([tails, np.full(k, S), sinks])
        cc = np.concatenate([heads, sources, np.full(k, T)])
        w = np.concatenate([(cap * unit).astype(np.int64),
                            np.ceil(supply[sources] * unit).astype(np.int64),
                            np.ceil(-supply[sinks] * unit).astype(np.int64)])
        Gm = sp.csr_matrix((w.astype(np.int32), (r, cc)), shape=(nodes + 2, nodes + 2))
        Gm.sum_duplicates()
        res = maximum_flow(Gm, S, T)
        F = res.flow.tocsr()
        routed_src = np.asarray(F[S, sources].todense()).ravel() / unit
        routed_snk = np.asarray(F[sinks, T].todense()).ravel() / unit
        new = np.zeros(nodes)
        new[sources] = 0.9 * routed_src
        new[sinks] = -0.9 * routed_snk
        total = supply[sources].sum()
        scale = float(new[sources].sum() / total) if total > 0 else 1.0
        supply = new
    cols = np.concatenate([np.arange(E), np.arange(E)])
    rows = np.concatenate([tails, heads])
    vals = np.concatenate([np.ones(E), -np.ones(E)])
    colptr, rowval, nzval = _csc_from_pairs(nodes, E, rows, cols, vals)
    lc = supply.copy()
    uc = supply.copy()
    return GeneratedLP(nodes, E, colptr, rowval, nzval, cost, np.zeros(E), cap, lc, uc,
                       name=f"min_cost_flow_{nodes}x{E}",
                       meta=dict(seed=seed, supply_scale=scale))
