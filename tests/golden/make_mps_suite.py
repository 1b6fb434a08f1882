"""Writes tests/golden/mps/: small LPs in free MPS with the features of the Netlib set
that the reference's integration tests solve (test/algorithms/all.jl:12-62) -- ranged
rows, free / fixed / upper-only variables, bad scaling, degeneracy and redundant rows,
staircase (multi-period) structure, a transportation problem -- plus highs.json with
each instance's optimal objective from SciPy's HiGHS.  Netlib itself cannot be fetched
here, so these stand in for it.

Every instance is feasible and bounded by construction: a point x inside the variable
bounds fixes the row bounds around A x, and c = A'y + z with the multiplier signs of a
dual-feasible point (y >= 0 on rows with only a lower bound, y <= 0 on rows with only an
upper bound; z >= 0 on variables with only a lower bound, z <= 0 with only an upper
bound, z = 0 on free variables).

Run from the repository root:  python tests/golden/make_mps_suite.py
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np
import scipy
import scipy.sparse as sp
from scipy.optimize import linprog

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "coolpdlp.jl_amd"))
from coolpdlp_hip.qps import write_mps  # noqa: E402

INF = np.inf


def _random_matrix(rng, m, n, per_col):
    rows, cols = [], []
    for j in range(n):
        k = min(m, per_col)
        rows.append(rng.choice(m, k, replace=False))
        cols.append(np.full(k, j))
    rows, cols = np.concatenate(rows), np.concatenate(cols)
    A = sp.csc_matrix((rng.standard_normal(len(rows)), (rows, cols)), shape=(m, n))
    # no empty rows
    empty = np.flatnonzero(np.diff(A.tocsr().indptr) == 0)
    if len(empty):
        A = A.tolil()
        for i in empty:
            A[i, rng.integers(n)] = rng.standard_normal()
        A = A.tocsc()
    return A


def _assemble(rng, A, row_kinds, var_kinds, tight=0.3):
    """Bounds around a feasible point and a dual-feasible objective (module docstring)."""
    m, n = A.shape
    lv, uv, x = np.zeros(n), np.full(n, INF), np.zeros(n)
    z = np.zeros(n)
    for j, k in enumerate(var_kinds):
        at_bound = rng.random() < tight
        if k == "lo":                      # [l, inf)
            lv[j] = 0.0 if rng.random() < 0.7 else -rng.uniform(0, 3)
            x[j] = lv[j] if at_bound else lv[j] + rng.uniform(0, 2)
            z[j] = abs(rng.standard_normal())
        elif k == "up":                    # (-inf, u]
            lv[j], uv[j] = -INF, rng.uniform(-2, 2)
            x[j] = uv[j] if at_bound else uv[j] - rng.uniform(0, 2)
            z[j] = -abs(rng.standard_normal())
        elif k == "box":
            lv[j] = -rng.uniform(0, 2) if rng.random() < 0.3 else 0.0
            uv[j] = lv[j] + rng.uniform(0.5, 5)
            x[j] = rng.choice([lv[j], uv[j]]) if at_bound else rng.uniform(lv[j], uv[j])
            z[j] = rng.standard_normal()
        elif k == "free":
            lv[j] = -INF
            x[j] = rng.standard_normal()
        elif k == "fixed":
            lv[j] = uv[j] = x[j] = rng.uniform(-1, 2)
            z[j] = rng.standard_normal()
    r = A @ x
    lc, uc, y = np.empty(m), np.empty(m), np.zeros(m)
    for i, k in enumerate(row_kinds):
        slack = 0.0 if rng.random() < tight else rng.uniform(0, 1)
        if k == "E":
            lc[i] = uc[i] = r[i]
            y[i] = rng.standard_normal()
        elif k == "L":
            lc[i], uc[i] = -INF, r[i] + slack
            y[i] = -abs(rng.standard_normal())
        elif k == "G":
            lc[i], uc[i] = r[i] - slack, INF
            y[i] = abs(rng.standard_normal())
        else:                              # ranged
            lc[i], uc[i] = r[i] - slack, r[i] + rng.uniform(0.1, 1)
            y[i] = rng.standard_normal()
    c = A.T @ y + z
    return dict(A=A, c=c, lv=lv, uv=uv, lc=lc, uc=uc)


def ranges(rng):
    m, n = 60, 100
    A = _random_matrix(rng, m, n, 4)
    rows = rng.choice(["E", "L", "G", "R"], m, p=[0.25, 0.25, 0.25, 0.25])
    return _assemble(rng, A, rows, rng.choice(["lo", "box"], n, p=[0.7, 0.3]))


def free_fixed(rng):
    m, n = 50, 90
    A = _random_matrix(rng, m, n, 4)
    rows = rng.choice(["E", "L", "G", "R"], m, p=[0.4, 0.2, 0.2, 0.2])
    var = rng.choice(["lo", "up", "box", "free", "fixed"], n, p=[0.4, 0.15, 0.2, 0.15, 0.1])
    return _assemble(rng, A, rows, var)


def badscale(rng):
    m, n = 80, 150
    A = _random_matrix(rng, m, n, 5)
    rows = rng.choice(["E", "L", "G", "R"], m, p=[0.3, 0.3, 0.2, 0.2])
    d = _assemble(rng, A, rows, rng.choice(["lo", "box", "up"], n, p=[0.6, 0.3, 0.1]))
    rs = 10.0 ** rng.uniform(-3, 3, m)      # row scales: the same constraints, rescaled
    cs = 10.0 ** rng.uniform(-2, 2, n)      # column scales: x = cs * x'
    d["A"] = sp.csc_matrix(sp.diags(rs) @ d["A"] @ sp.diags(cs))
    d["lc"], d["uc"] = d["lc"] * rs, d["uc"] * rs
    d["c"] = d["c"] * cs
    d["lv"], d["uv"] = d["lv"] / cs, d["uv"] / cs
    return d


def degenerate(rng):
    m0, n = 90, 160
    A0 = _random_matrix(rng, m0, n, 4)
    # 30 redundant rows: scaled copies of existing inequality rows
    src = rng.choice(m0, 30, replace=False)
    scale = rng.uniform(0.5, 2.0, 30)
    A = sp.vstack([A0, sp.diags(scale) @ A0.tocsr()[src]]).tocsc()
    rows0 = rng.choice(["E", "L", "G"], m0, p=[0.2, 0.4, 0.4])
    d = _assemble(rng, A0, rows0, rng.choice(["lo", "box"], n, p=[0.8, 0.2]), tight=0.7)
    lc2, uc2 = d["lc"][src] * scale, d["uc"][src] * scale
    d["A"], d["lc"], d["uc"] = A, np.concatenate([d["lc"], lc2]), np.concatenate([d["uc"], uc2])
    # the copies get no multiplier: c stays dual feasible
    return d


def staircase(rng):
    T, mr, nc = 6, 15, 25
    blocks = []
    for t in range(T):
        row = [None] * T
        row[t] = _random_matrix(rng, mr, nc, 3)
        if t > 0:
            row[t - 1] = sp.random(mr, nc, density=0.08, random_state=rng,
                                   data_rvs=rng.standard_normal)
        blocks.append(row)
    A = sp.bmat(blocks, format="csc")
    m, n = A.shape
    rows = rng.choice(["E", "L", "G"], m, p=[0.5, 0.25, 0.25])
    return _assemble(rng, A, rows, rng.choice(["lo", "box"], n, p=[0.8, 0.2]))


def transport(rng):
    ns, nd = 12, 18
    supply = rng.integers(20, 80, ns).astype(float)
    demand = rng.multinomial(int(supply.sum()), np.ones(nd) / nd).astype(float)
    n = ns * nd
    rows, cols = [], []
    for i in range(ns):
        for j in range(nd):
            k = i * nd + j
            rows += [i, ns + j]
            cols += [k, k]
    A = sp.csc_matrix((np.ones(len(rows)), (rows, cols)), shape=(ns + nd, n))
    b = np.concatenate([supply, demand])     # balanced: one row is redundant
    return dict(A=A, c=rng.integers(1, 20, n).astype(float), lv=np.zeros(n), uv=np.full(n, INF),
                lc=b, uc=b.copy())


INSTANCES = [("netlike_ranges", ranges, 101), ("netlike_free_fixed", free_fixed, 102),
             ("netlike_badscale", badscale, 103), ("netlike_degenerate", degenerate, 104),
             ("netlike_staircase", staircase, 105), ("netlike_transport", transport, 106)]


def _highs(d):
    A = sp.csr_matrix(d["A"])
    lc, uc = d["lc"], d["uc"]
    eq = lc == uc
    ub_rows, ub_rhs = [], []
    up = ~eq & np.isfinite(uc)
    lo = ~eq & np.isfinite(lc)
    if up.any():
        ub_rows.append(A[up])
        ub_rhs.append(uc[up])
    if lo.any():
        ub_rows.append(-A[lo])
        ub_rhs.append(-lc[lo])
    h = linprog(d["c"], A_ub=sp.vstack(ub_rows) if ub_rows else None,
                b_ub=np.concatenate(ub_rhs) if ub_rhs else None,
                A_eq=A[eq] if eq.any() else None, b_eq=lc[eq] if eq.any() else None,
                bounds=list(zip(d["lv"], d["uv"])), method="highs")
    return h


def main():
    out = os.path.join(HERE, "mps")
    os.makedirs(out, exist_ok=True)
    record = {"solver": f"scipy {scipy.__version__} linprog(method='highs')", "instances": {}}
    for name, fn, seed in INSTANCES:
        d = fn(np.random.default_rng(seed))
        h = _highs(d)
        if h.status != 0:
            raise SystemExit(f"{name}: HiGHS status {h.status} ({h.message})")
        write_mps(os.path.join(out, name + ".mps"), name=name, **d)
        m, n = d["A"].shape
        record["instances"][name] = {"m": m, "n": n, "nnz": int(d["A"].nnz),
                                     "objective": float(h.fun)}
        print(f"{name}: m={m} n={n} nnz={d['A'].nnz} objective={h.fun:.10g}")
    with open(os.path.join(out, "highs.json"), "w") as f:
        json.dump(record, f, indent=1, sort_keys=True)
        f.write("\n")


if __name__ == "__main__":
    main()
