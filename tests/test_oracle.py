"""Pins the CPU oracle against the reference's own tests, known answers, closed forms
and independent libraries (scipy, HiGHS).  CPU only.

Each check is chosen so that a plausible slip (dropped term, wrong sign, transposed
operand, swapped bound) fails it.  Reference test files are cited per test.
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.optimize import linprog

import oracle as O
from coolpdlp_hip import generators as G

INF = np.inf


def _lp(A, c, lv, uv, lc, uc):
    A = sp.csc_matrix(np.asarray(A, float))
    return G.GeneratedLP(A.shape[0], A.shape[1], A.indptr.astype(np.int64),
                         A.indices.astype(np.int32), A.data.astype(float),
                         *[np.asarray(v, float) for v in (c, lv, uv, lc, uc)])


def _kkt_plain(lp, x, y):
    """KKT errors of (x, y) on lp with D1 = D2 = 1 (no scaling)."""
    A_csr = O.csc_to_csr(lp.m, lp.n, lp.colptr, lp.rowval, lp.nzval)
    return O.kkt_errors(lp.m, lp.n, A_csr, (lp.colptr, lp.rowval, lp.nzval), lp.c, lp.lv, lp.uv,
                        lp.lc, lp.uc, np.ones(lp.m), np.ones(lp.n), x, y)


# ------------------------------------------------------------------ matrices
def test_csc_to_csr_is_scipy_transpose():
    rng = np.random.default_rng(0)
    for (m, n, p) in [(1, 1, 1.0), (10, 30, 0.2), (300, 100, 0.05), (5, 0, 0.1), (0, 4, 0.1)]:
        A = sp.random(m, n, density=p, format="csc", random_state=rng)
        rp, col, val = O.csc_to_csr(m, n, A.indptr, A.indices, A.data)
        R = A.tocsr()
        R.sort_indices()
        assert np.array_equal(rp, R.indptr) and np.array_equal(col, R.indices)
        assert np.array_equal(val, R.data)


def test_spmv_matches_dense():
    # test/utils/matrices.jl:10-34: mul!(c, A, b, α, β) ≈ αAb + βc
    rng = np.random.default_rng(1)
    for m in (10, 20, 30):
        for n in (10, 20, 30):
            for p in (0.01, 0.1, 0.2, 0.3):
                A = sp.random(m, n, density=p, format="csc", random_state=rng)
                rp, col, val = O.csc_to_csr(m, n, A.indptr, A.indices, A.data)
                b, c = rng.random(n), rng.random(m)
                out = O.spmv_csr(rp, col, val, 0.3, b, 0.7, c)
                np.testing.assert_allclose(out, 0.3 * (A.toarray() @ b) + 0.7 * c, rtol=1e-14)
                # beta == 0 never reads c
                out0 = O.spmv_csr(rp, col, val, 1.0, b, 0.0, np.full(m, np.nan))
                np.testing.assert_allclose(out0, A.toarray() @ b, rtol=1e-14, atol=1e-15)


def test_ell_layout():
    # mat_ell.jl:54-70: d = max row length, left-packed, column-major m x d
    A = sp.csc_matrix(np.array([[1.0, 0, 2], [0, 0, 0], [0, 3, 4], [5, 0, 0]]))
    rp, col, val = O.csc_to_csr(4, 3, A.indptr, A.indices, A.data)
    d, ec, ev = O.csr_to_ell(rp, col, val)
    assert d == 2
    assert ec.tolist() == [[0, -1, 1, 0], [2, -1, 2, -1]]
    assert ev.tolist() == [[1, 0, 3, 5], [2, 0, 4, 0]]


# ------------------------------------------------------------------ linalg helpers
def test_combine_known_answers():
    # test/utils/linalg.jl:33-39: combine(1,2)=2, (3,3)=3, (-Inf,2)=2, (3,Inf)=3, (-Inf,Inf)=0
    # primal_scale = 1 + ||combine(lc, uc)|| with D1 = 1 and any x
    lc = [1, 3, -INF, 3, -INF]
    uc = [2, 3, 2, INF, INF]
    lp = _lp(np.zeros((5, 1)), [0.0], [0.0], [1.0], lc, uc)
    e = _kkt_plain(lp, np.zeros(1), np.zeros(5))
    assert e.primal_scale == 1 + math.sqrt(4 + 9 + 4 + 9 + 0)


def test_proj_multiplier_cases():
    # test/utils/linalg.jl:24-31 via the dual residual c - A'y - r with A'y = 0:
    # (-Inf, Inf) -> r = 0; (-Inf, 3) -> r = -neg(c); (-3, Inf) -> r = pos(c); (-3, 3) -> r = c
    rng = np.random.default_rng(2)
    for _ in range(20):
        c = rng.standard_normal(4)
        lp = _lp(np.zeros((1, 4)), c, [-INF, -INF, -3, -3], [INF, 3, INF, 3], [0.0], [0.0])
        e = _kkt_plain(lp, np.zeros(4), np.zeros(1))
        expected = math.sqrt(c[0] ** 2 + max(c[1], 0) ** 2 + min(c[2], 0) ** 2 + 0.0)
        assert abs(e.dual - expected) <= 1e-15 * (1 + expected)


def test_kkt_closed_forms():
    # test/components/errors.jl:27-34 (with safe() for infinite bounds, :7-15)
    lp, x, y = G.random_milp_and_sol(100, 200, 0.4, seed=3)
    A = lp.csc().toarray()
    e = _kkt_plain(lp, x, y)
    Ax, Aty = A @ x, A.T @ y
    r = np.where(np.isneginf(lp.lv), np.where(np.isposinf(lp.uv), 0.0, np.minimum(lp.c - Aty, 0)),
                 np.where(np.isposinf(lp.uv), np.maximum(lp.c - Aty, 0), lp.c - Aty))

    def safe(v):
        return np.where(np.isposinf(v), np.finfo(float).max, np.where(np.isneginf(v), -np.finfo(float).max, v))

    def p(z, l, u):
        return np.dot(np.maximum(z, 0), safe(u)) - np.dot(-np.minimum(z, 0), safe(l))

    comb = np.maximum(0, np.maximum(np.where(np.isfinite(lp.lc), abs(lp.lc), 0),
                                    np.where(np.isfinite(lp.uc), abs(lp.uc), 0)))
    np.testing.assert_allclose(e.primal, np.linalg.norm(Ax - np.clip(Ax, lp.lc, lp.uc)), rtol=1e-12)
    np.testing.assert_allclose(e.dual, np.linalg.norm(lp.c - Aty - r), rtol=1e-12)
    np.testing.assert_allclose(e.primal_scale, 1 + np.linalg.norm(comb), rtol=1e-12)
    np.testing.assert_allclose(e.dual_scale, 1 + np.linalg.norm(lp.c), rtol=1e-12)
    # y and r lie in their multiplier sets, so p() never multiplies an infinite bound
    gap = abs(np.dot(lp.c, x) + p(-y, lp.lc, lp.uc) + p(-r, lp.lv, lp.uv))
    np.testing.assert_allclose(e.gap, gap, rtol=1e-10)
    gs = 1 + abs(np.dot(lp.c, x)) + abs(p(-y, lp.lc, lp.uc) + p(-r, lp.lv, lp.uv))
    np.testing.assert_allclose(e.gap_scale, gs, rtol=1e-12)


def test_kkt_invariant_under_preconditioning():
    # test/components/errors.jl:36-38 and docs/src/preconditioning.md
    lp, x, y = G.random_milp_and_sol(100, 200, 0.4, seed=4)
    e = _kkt_plain(lp, x, y)
    D1, D2, a, at = O.scaling(lp.m, lp.n, lp.colptr, lp.rowval, lp.nzval, 10, 1.0)
    assert not np.allclose(D1, 1) and not np.allclose(D2, 1)
    A_csr = O.csc_to_csr(lp.m, lp.n, lp.colptr, lp.rowval, a)
    ep = O.kkt_errors(lp.m, lp.n, A_csr, (lp.colptr, lp.rowval, at), D2 * lp.c, lp.lv / D2,
                      lp.uv / D2, D1 * lp.lc, D1 * lp.uc, D1, D2, x / D2, y / D1)
    for f in ("primal", "primal_scale", "dual", "dual_scale", "gap", "gap_scale"):
        assert math.isclose(getattr(e, f), getattr(ep, f), rel_tol=1e-9, abs_tol=1e-12), f


def test_generator_point_is_optimal():
    # (x0, y0) of every generated LP is a KKT point: all three errors vanish
    for g in (G.column_lp(200, 500, 6, seed=1), G.column_lp(100, 300, (6, 12), seed=2),
              G.power_law_lp(2000, 4000, 40_000, seed=3)):
        e = _kkt_plain(g, g.x0, g.y0)
        assert e.primal / e.primal_scale < 1e-13
        assert e.dual / e.dual_scale < 1e-13
        assert e.gap / e.gap_scale < 1e-13
        assert math.isclose(g.objective, g.meta["dual_objective"], rel_tol=1e-10, abs_tol=1e-10)


# ------------------------------------------------------------------ scaling
def test_ruiz_converges_to_unit_norms():
    # test/components/preconditioning.jl:36-41 (Ruiz alone, 10^4 iterations, rtol 1e-2)
    rng = np.random.default_rng(42)
    A = sp.random(10, 20, density=0.4, format="csc", random_state=rng)
    D1, D2, a, at = O.scaling(10, 20, A.indptr, A.indices, A.data, 10_000, float("nan"))
    As = sp.csc_matrix((a, A.indices, A.indptr), shape=A.shape)
    Ats = sp.csc_matrix((at, A.indices, A.indptr), shape=A.shape)
    cols = abs(As).max(axis=0).toarray().ravel()
    rows = abs(Ats).max(axis=1).toarray().ravel()
    assert np.allclose(cols[cols > 0], 1, rtol=1e-2) and np.allclose(rows[rows > 0], 1, rtol=1e-2)
    # the scaled values are D1 A D2
    np.testing.assert_allclose(As.toarray(), D1[:, None] * A.toarray() * D2[None, :], rtol=1e-14)


def test_scaling_one_pass_closed_form():
    # one Ruiz pass + CP(alpha=1) by hand: d1 = 1/sqrt(max|row|), d2 = 1/sqrt(max|col|), then
    # 1-norms of the once-scaled matrix
    A = np.array([[4.0, 0, -1], [0, 9, 0], [2, 0, 16]])
    S = sp.csc_matrix(A)
    D1, D2, a, at = O.scaling(3, 3, S.indptr, S.indices, S.data, 1, 1.0)
    r1 = 1 / np.sqrt(abs(A).max(1))
    c1 = 1 / np.sqrt(abs(A).max(0))
    A1 = r1[:, None] * A * c1[None, :]
    r2 = 1 / np.sqrt(abs(A1).sum(1))
    c2 = 1 / np.sqrt(abs(A1).sum(0))
    np.testing.assert_allclose(D1, r1 * r2, rtol=1e-15)
    np.testing.assert_allclose(D2, c1 * c2, rtol=1e-15)


def test_scaling_preserves_objective_and_bilinear_form():
    # test/components/preconditioning.jl:44-75
    rng = np.random.default_rng(5)
    lp, _, _ = G.random_milp_and_sol(10, 20, 0.4, seed=5)
    x, y = rng.standard_normal(20), rng.standard_normal(10)
    D1, D2, a, at = O.scaling(lp.m, lp.n, lp.colptr, lp.rowval, lp.nzval, 10, 1.0)
    A = lp.csc().toarray()
    Ap = sp.csc_matrix((a, lp.rowval, lp.colptr), shape=A.shape).toarray()
    assert math.isclose(lp.c @ x, (D2 * lp.c) @ (x / D2), rel_tol=1e-12)
    assert math.isclose(y @ A @ x, (y / D1) @ Ap @ (x / D2), rel_tol=1e-12)
    np.testing.assert_allclose(np.clip(x, lp.lv, lp.uv), D2 * np.clip(x / D2, lp.lv / D2, lp.uv / D2),
                               rtol=1e-14)


# ------------------------------------------------------------------ spectral norm
def test_spectral_norm_vs_numpy():
    # test/utils/linalg.jl:52-60 (rtol 0.1 there; converged power iteration is far tighter)
    rng = np.random.default_rng(42)
    for _ in range(10):
        A = rng.standard_normal((10, 20))
        S = sp.csc_matrix(A)
        A_csr = O.csc_to_csr(10, 20, S.indptr, S.indices, S.data)
        s, it = O.spectral_norm(10, 20, A_csr, (S.indptr, S.indices, S.data), tol=1e-7, maxiter=10_000)
        assert math.isclose(s, np.linalg.norm(A, 2), rel_tol=1e-3)


def test_spectral_norm_one_rayleigh_quotient_rule():
    # powm! stops after one iteration when tol*|θ| exceeds the first residual (large n):
    # then λ = x0ᵀAᵀA x0 for the normalised start vector
    g = G.column_lp(50, 120, 4, seed=6)
    S = g.csc()
    A_csr = O.csc_to_csr(g.m, g.n, g.colptr, g.rowval, g.nzval)
    s, it = O.spectral_norm(g.m, g.n, A_csr, (g.colptr, g.rowval, g.nzval), tol=1e6)
    assert it == 1
    x0 = O.powm_start(g.n)
    x0 /= np.linalg.norm(x0)
    assert math.isclose(s, np.linalg.norm(S @ x0), rel_tol=1e-12)
    # default rule: tol = eps*n^3, maxiter = n -> stops at residual <= tol*|θ|
    s2, it2 = O.spectral_norm(g.m, g.n, A_csr, (g.colptr, g.rowval, g.nzval))
    assert 1 <= it2 <= g.n + 1
    assert s2 <= np.linalg.norm(S.toarray(), 2) * (1 + 1e-12)


def test_powm_start_is_deterministic_and_centered():
    a, b = O.powm_start(100_000, 0), O.powm_start(100_000, 0)
    assert np.array_equal(a, b)
    assert abs(a.mean()) < 0.02 and abs(a.std() - 1.0) < 0.02   # Irwin-Hall(12) - 6
    assert not np.array_equal(a, O.powm_start(100_000, 1))


# ------------------------------------------------------------------ full solves
def test_moi_two_variable_lp_205():
    # test/moi.jl:38-53: objective 205.0, atol 1e-2
    g = _lp([[6, 8], [7, 12]], [12, 20], [0, 0], [INF, 3], [100, 120], [INF, INF])
    r = O.solve(g)
    assert r.status == "OPTIMAL"
    assert abs(r.x @ g.c - 205.0) <= 1e-2


def test_termination_check_not_skipped():
    # test/components/termination.jl:1-10
    g = _lp([[1, 1]], [1, 1], [0, 0], [INF, INF], [1], [INF])
    r = O.solve(g)
    assert r.status == "OPTIMAL" and r.stats.kkt_passes == 100
    assert abs(r.x.sum() - 1) < 1e-3


@pytest.mark.parametrize("algo,reltol", [("PDLP", 1e-6), ("PDHG", 1e-4)])
def test_solve_vs_closed_form_and_highs(algo, reltol):
    g = G.column_lp(150, 400, 5, seed=7)
    r = O.solve(g, O.default_params(algo, termination_reltol=reltol, max_kkt_passes=10**6))
    assert r.status == "OPTIMAL"
    assert r.stats.err.relative() <= reltol
    h = linprog(g.c, bounds=list(zip(g.lv, np.where(np.isinf(g.uv), None, g.uv))),
                method="highs", **_highs_rows(g))
    assert h.status == 0
    assert abs(h.fun - g.objective) <= 1e-7 * (1 + abs(g.objective))   # closed form is right
    assert abs(r.x @ g.c - g.objective) <= 20 * reltol * (1 + abs(g.objective))


def _highs_rows(g):
    A = g.csc().tocsr()
    eq = np.isfinite(g.lc) & (g.lc == g.uc)
    le = ~eq
    return dict(A_eq=A[eq], b_eq=g.uc[eq], A_ub=A[le], b_ub=g.uc[le])


def test_pdlp_reference_sized_case_closed_form():
    # the reference's CPU-runnable workload: m=2,000 n=5,000, 8 nnz per column, reltol 1e-4
    g = G.small_dense_lp()
    r = O.solve(g)
    assert r.status == "OPTIMAL" and r.stats.kkt_passes % 100 == 0
    assert abs(r.x @ g.c - g.objective) <= 1e-3 * (1 + abs(g.objective))


def test_termination_limits_and_history():
    g = G.column_lp(100, 250, 4, seed=8)
    r = O.solve(g, O.default_params(max_kkt_passes=1, termination_reltol=1e-14))
    assert r.status == "ITERATION_LIMIT" and r.stats.kkt_passes == 100
    # one check + the duplicate entry of the outer termination check (pdlp.jl:88-90)
    assert r.stats.n_history == 2 and r.history[0] == r.history[1]
    r = O.solve(g, O.default_params(time_limit=0.0, termination_reltol=1e-14))
    assert r.status == "TIME_LIMIT"
    r = O.solve(g, O.default_params("PDHG", max_kkt_passes=1, termination_reltol=1e-14))
    assert r.status == "ITERATION_LIMIT" and r.stats.n_history == 1   # PDHG: no duplicate


def test_first_check_restarts():
    # F9: inner == total at the first restart check, so long_inner_loop is true
    g = G.column_lp(100, 250, 4, seed=9)
    r = O.solve(g, O.default_params(max_kkt_passes=100, termination_reltol=1e-14))
    assert r.stats.restarts == 0   # terminated by the iteration limit before the restart check
    r = O.solve(g, O.default_params(max_kkt_passes=200, termination_reltol=1e-14))
    assert r.stats.restarts == 1


def test_warm_start_at_optimum_terminates_at_first_check():
    g = G.column_lp(100, 250, 4, seed=10)
    r = O.solve(g, x0=g.x0, y0=g.y0)
    assert r.status == "OPTIMAL" and r.stats.kkt_passes == 100


def test_planted_sign_error_is_caught():
    """A wrong multiplier sign on <= rows (y0 >= 0, the configs' literal wording) makes
    (x0, y0) dual-infeasible in the reference's convention: the KKT check sees it."""
    g = G.column_lp(100, 250, 4, seed=11)
    bad_y = np.where(np.isneginf(g.lc), -g.y0, g.y0)
    e = _kkt_plain(g, g.x0, bad_y)
    assert e.dual / e.dual_scale > 1e-3 or e.gap / e.gap_scale > 1e-3
