"""GPU parity of the plain (one LP) path, through the C ABI, against the oracle.

Bars: index maps and scaling bit-exact; SpMV bit-exact on rows that the reference's
thread-per-row order covers (rows <= 64 nonzeros), 1e-13 relative otherwise; solve
results: same termination status, objective within 1e-6 relative of the oracle's and
of the closed form c'x0, kkt_passes within two checks.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
from coolpdlp_hip import generators as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cp():
    import coolpdlp_hip
    return coolpdlp_hip


def _lp(A, c, lv, uv, lc, uc):
    A = sp.csc_matrix(np.asarray(A, float))
    g = G.GeneratedLP(A.shape[0], A.shape[1], A.indptr.astype(np.int64), A.indices.astype(np.int32),
                      A.data.astype(float), *[np.asarray(v, float) for v in (c, lv, uv, lc, uc)])
    return g


# ------------------------------------------------------------------ matrices
def _rand_mats(seed=0):
    rng = np.random.default_rng(seed)
    mats = []
    for m in (10, 20, 30):
        for n in (10, 20, 30):
            for p in (0.01, 0.1, 0.2, 0.3):   # test/utils/matrices.jl:10-15
                mats.append(sp.random(m, n, density=p, format="csc", random_state=rng))
    # long rows (wavefront path) mixed with short ones, and an empty row / column
    A = sp.random(400, 3000, density=0.002, format="lil", random_state=rng)
    for i in (0, 7, 399):
        A[i, rng.choice(3000, 700, replace=False)] = rng.standard_normal(700)
    mats.append(sp.csc_matrix(A))
    return mats


@pytest.mark.parametrize("fmt", ["csr", "ell", "auto"])
def test_mul_matches_dense(cp, fmt):
    rng = np.random.default_rng(1)
    for A in _rand_mats():
        cls = cp.GPUSparseMatrixELL if fmt == "ell" else cp.GPUSparseMatrixCSR
        kw = {} if fmt != "auto" else {"format": 0}
        M = cls(A, **kw)
        assert M.shape == A.shape and M.nnz == A.nnz
        b = rng.random(A.shape[1])
        c = rng.random(A.shape[0])
        alpha, beta = 0.37, 0.81
        out = M.mul(c.copy(), b, alpha, beta)
        ref = alpha * (A @ b) + beta * c
        np.testing.assert_allclose(out, ref, rtol=1e-13, atol=1e-14)
        # beta == 0 must not read c (NaN there)
        out0 = M.mul(np.full(A.shape[0], np.nan), b, 1.0, 0.0)
        assert np.all(np.isfinite(out0))
        # bit-exact vs the oracle's thread-per-row sum on short rows
        rp, col, val = O.csc_to_csr(A.shape[0], A.shape[1], A.indptr, A.indices, A.data)
        seq = O.spmv_csr(rp, col, val, 1.0, b, 0.0, np.zeros(A.shape[0]))
        short = np.diff(rp) <= 64
        assert np.array_equal(out0[short], seq[short])
        np.testing.assert_allclose(out0, seq, rtol=1e-13, atol=1e-14)


def test_mul_transpose_and_torch(cp):
    import torch
    rng = np.random.default_rng(2)
    A = sp.random(500, 800, density=0.01, format="csc", random_state=rng)
    Mt = cp.GPUSparseMatrixCSR(A, transpose=True)
    assert Mt.shape == (800, 500)
    y = torch.tensor(rng.standard_normal(500), device="cuda:0")
    out = torch.empty(800, dtype=torch.float64, device="cuda:0")
    Mt.mul(out, y)
    torch.cuda.synchronize()
    np.testing.assert_allclose(out.cpu().numpy(), A.T @ y.cpu().numpy(), rtol=1e-13, atol=1e-14)
    with pytest.raises(cp.ArgumentError):
        Mt.mul(np.zeros(800), y)   # mixed backends, utils/device.jl:6-12


def test_ell_layout_matches_oracle(cp):
    g = G.column_lp(60, 90, 3, seed=4)
    A = g.csc()
    M = cp.GPUSparseMatrixELL(A)
    col, val = M.ell()
    rp, c, v = O.csc_to_csr(g.m, g.n, g.colptr, g.rowval, g.nzval)
    d, ec, ev = O.csr_to_ell(rp, c, v)
    assert M.ell_width == d
    assert np.array_equal(col, ec) and np.array_equal(val, ev)


# ------------------------------------------------------------------ preprocessing
def test_csr_and_scaling_bit_exact(cp):
    lp, _, _ = G.random_milp_and_sol(120, 260, 0.05, seed=7)
    milp = cp.MILP.from_generated(lp)
    with cp.Engine(milp, cp.PDLP()) as eng:
        rp, col = eng.csr()
        orp, ocol, _ = O.csc_to_csr(lp.m, lp.n, lp.colptr, lp.rowval, lp.nzval)
        assert np.array_equal(rp, orp) and np.array_equal(col, ocol)
        D1, D2, a, at = eng.scaling()
    oD1, oD2, oa, oat = O.scaling(lp.m, lp.n, lp.colptr, lp.rowval, lp.nzval, 10, 1.0)
    assert np.array_equal(D1, oD1) and np.array_equal(D2, oD2)
    assert np.array_equal(a, oa) and np.array_equal(at, oat)


def test_step_size_matches_oracle(cp):
    g = G.column_lp(200, 500, 5, seed=3)
    milp = cp.MILP.from_generated(g)
    with cp.Engine(milp, cp.PDLP()) as eng:
        eta, omega, iters = eng.step_size()
    D1, D2, a, at = O.scaling(g.m, g.n, g.colptr, g.rowval, g.nzval, 10, 1.0)
    A_csr = O.csc_to_csr(g.m, g.n, g.colptr, g.rowval, a)
    At_csr = (g.colptr, g.rowval, at)
    s, it = O.spectral_norm(g.m, g.n, A_csr, At_csr)
    assert abs(iters - it) <= 2
    assert abs(eta - 0.9 / s) <= 1e-9 * eta


def test_step_size_vs_exact_spectral_norm(cp):
    """test/utils/linalg.jl "Spectral norm": dense N(0,1) 10 x 20 matrices; the engine's
    power method gives the operator 2-norm of the scaled matrix (the step size is
    invnorm_scaling / norm, 0.9 by default): within the reference's 1e-1 with its
    tolerance 1e-7 and the default cap of n iterations, within 1e-6 with the cap lifted."""
    rng = np.random.default_rng(42)
    for k in range(10):
        A = sp.csc_matrix(rng.standard_normal((10, 20)))
        g = _lp(A.toarray(), np.ones(20), np.zeros(20), np.full(20, np.inf), np.zeros(10), np.full(10, np.inf))
        milp = cp.MILP.from_generated(g)
        with cp.Engine(milp, cp.PDLP(powm_tol=1e-7)) as eng:
            eta, _, it = eng.step_size()
            D1, D2, a, _ = eng.scaling()
        with cp.Engine(milp, cp.PDLP(powm_tol=1e-7, powm_maxiter=10_000)) as eng:
            eta_c, _, it_c = eng.step_size()
        As = sp.csc_matrix((a, A.indices, A.indptr), shape=A.shape).toarray()
        assert np.array_equal(As, D1[:, None] * A.toarray() * D2[None, :])
        exact = np.linalg.norm(As, 2)
        assert it <= 21 and abs(0.9 / eta - exact) <= 1e-1 * exact, (k, 0.9 / eta, exact, it)
        assert abs(0.9 / eta_c - exact) <= 1e-6 * exact, (k, 0.9 / eta_c, exact, it_c)


# ------------------------------------------------------------------ solves
def test_moi_two_variable_lp(cp):
    # test/moi.jl:44-52: min 12x + 20y, 6x + 8y >= 100, 7x + 12y >= 120, x >= 0, 0 <= y <= 3
    g = _lp([[6, 8], [7, 12]], [12, 20], [0, 0], [np.inf, 3], [100, 120], [np.inf, np.inf])
    milp = cp.MILP.from_generated(g)
    sol, st = cp.solve(milp, cp.PDLP())
    assert st.termination_status == cp.OPTIMAL
    assert abs(cp.objective_value(sol.x, milp) - 205.0) <= 1e-2
    assert cp.is_feasible(sol.x, milp, cons_tol=1e-2)


def test_termination_not_skipped(cp):
    # test/components/termination.jl:1-10
    g = _lp([[1, 1]], [1, 1], [0, 0], [np.inf, np.inf], [1], [np.inf])
    sol, st = cp.solve(cp.MILP.from_generated(g), cp.PDLP())
    assert st.termination_status == cp.OPTIMAL
    assert abs(sol.x.sum() - 1.0) < 1e-3


def test_mul_on_torch_default_stream_is_ordered(cp):
    """mul! on device tensors runs on the caller's stream (torch's current one, the null
    stream by default): b written by torch just before and c read by torch just after
    need no synchronisation of their own."""
    import torch
    rng = np.random.default_rng(7)
    m, n = 300_000, 200_000
    A = sp.random(m, n, density=10 / n, format="csc", random_state=rng)
    M = cp.GPUSparseMatrixCSR(A)
    for seed in range(3):
        torch.manual_seed(seed)
        b = torch.rand(n, dtype=torch.float64, device="cuda:0")
        c = torch.full((m,), float("nan"), dtype=torch.float64, device="cuda:0")
        M.mul(c, b)
        out = c.cpu().numpy()
        np.testing.assert_allclose(out, A @ b.cpu().numpy(), rtol=1e-12, atol=1e-13)
    M.close()


@pytest.mark.parametrize("algo,use_graph,fmt", [
    ("PDLP", 0, "csr"), ("PDLP", 1, "csr"), ("PDLP", 1, "ell"), ("PDLP", 1, "xslice"),
    ("PDLP", 1, "xslice_all"), ("PDLP", 1, "auto"), ("PDHG", 0, "csr"), ("PDHG", 1, "auto")])
def test_solve_matches_oracle(cp, algo, use_graph, fmt):
    g = G.column_lp(400, 1000, 6, seed=21)
    milp = cp.MILP.from_generated(g)
    A = cp.PDLP if algo == "PDLP" else cp.PDHG
    reltol = 1e-6 if algo == "PDLP" else 1e-4
    sol, st = cp.solve(milp, A("float64", None, fmt, termination_reltol=reltol, use_graph=use_graph))
    code = {"ell": 2, "xslice": 3, "xslice_all": 4}
    if fmt in code:
        assert (st.format_A, st.format_At) == (code[fmt], code[fmt])
    ref = O.solve(g, O.default_params(algo, termination_reltol=reltol))
    assert st.termination_status.name == ref.status == "OPTIMAL"
    assert abs(st.kkt_passes - ref.stats.kkt_passes) <= 200
    assert st.err.relative() <= reltol
    obj = cp.objective_value(sol.x, milp)
    obj_ref = float(ref.x @ g.c)
    assert abs(obj - obj_ref) <= 10 * reltol * (1 + abs(obj_ref))
    assert abs(obj - g.objective) <= 10 * reltol * (1 + abs(g.objective))
    assert abs(st.eta - ref.stats.eta) <= 1e-9 * ref.stats.eta
    # the history is recorded every check, with the duplicate final entry of PDLP
    assert len(st.error_history) == ref.stats.n_history


@pytest.mark.parametrize("fmt", ["csr", "auto"])
def test_small_dense_lp_at_its_stated_size(cp, fmt):
    """The reference's CPU-runnable case at its stated size: m = 2,000 (1,000 =, 1,000 <=),
    n = 5,000, 8 nonzeros per column, 0 <= x, FP64 PDLP to relative KKT 1e-4 with the
    reference's power-method rule, against the oracle on the same seeded LP: same status,
    step size, pass count within two checks, early checks to reduction-order rounding,
    objective against the oracle and the closed form c'x0."""
    g = G.small_dense_lp()
    assert (g.m, g.n, g.nnz) == (2_000, 5_000, 40_000)
    milp = cp.MILP.from_generated(g)
    sol, st = cp.solve(milp, cp.PDLP("float64", None, fmt, termination_reltol=1e-4))
    ref = O.solve(g, O.default_params(termination_reltol=1e-4))
    assert st.termination_status.name == ref.status == "OPTIMAL"
    assert abs(st.eta - ref.stats.eta) <= 1e-9 * ref.stats.eta
    assert abs(st.powm_iters - ref.stats.powm_iters) <= 2
    assert abs(st.kkt_passes - ref.stats.kkt_passes) <= 200
    assert st.err.relative() <= 1e-4
    for (p1, e1), (p2, e2) in list(zip(st.error_history, ref.history))[:3]:
        assert p1 == p2
        for f in ("primal", "dual", "gap"):
            assert abs(getattr(e1, f) - e2[f]) <= 1e-8 * (1 + abs(e2[f])), (p1, f)
    obj, obj_ref = cp.objective_value(sol.x, milp), float(ref.x @ g.c)
    assert abs(obj - obj_ref) <= 1e-3 * (1 + abs(obj_ref))
    assert abs(obj - g.objective) <= 1e-3 * (1 + abs(g.objective))


def test_trajectory_matches_oracle_exactly_early(cp):
    """Short rows + FP64: the first checks' KKT errors agree to reduction-order
    rounding (sums differ only in order)."""
    g = G.column_lp(300, 700, 6, seed=5)
    milp = cp.MILP.from_generated(g)
    _, st = cp.solve(milp, cp.PDLP("float64", None, "csr", max_kkt_passes=500, termination_reltol=1e-12))
    ref = O.solve(g, O.default_params(max_kkt_passes=500, termination_reltol=1e-12))
    assert st.termination_status == cp.ITERATION_LIMIT and ref.status == "ITERATION_LIMIT"
    assert st.kkt_passes == ref.stats.kkt_passes == 500
    for (p1, e1), (p2, e2) in zip(st.error_history, ref.history):
        assert p1 == p2
        for f in ("primal", "dual", "gap"):
            assert abs(getattr(e1, f) - e2[f]) <= 1e-9 * (1 + abs(e2[f])), (p1, f)


def test_kkt_errors_match_oracle(cp):
    lp, x, y = G.random_milp_and_sol(80, 150, 0.1, seed=9)
    milp = cp.MILP.from_generated(lp)
    with cp.Engine(milp, cp.PDLP()) as eng:
        e = eng.kkt_errors(cp.PrimalDualSolution(x, y))
    D1, D2, a, at = O.scaling(lp.m, lp.n, lp.colptr, lp.rowval, lp.nzval, 10, 1.0)
    A_csr = O.csc_to_csr(lp.m, lp.n, lp.colptr, lp.rowval, a)
    e_ref = O.kkt_errors(lp.m, lp.n, A_csr, (lp.colptr, lp.rowval, at), D2 * lp.c, lp.lv / D2,
                         lp.uv / D2, D1 * lp.lc, D1 * lp.uc, D1, D2, x / D2, y / D1)
    for f in ("primal", "primal_scale", "dual", "dual_scale", "gap", "gap_scale"):
        assert abs(getattr(e, f) - getattr(e_ref, f)) <= 1e-12 * (1 + abs(getattr(e_ref, f))), f


def test_limits(cp):
    g = G.column_lp(200, 400, 5, seed=8)
    milp = cp.MILP.from_generated(g)
    _, st = cp.solve(milp, cp.PDLP(max_kkt_passes=1, termination_reltol=1e-12))
    assert st.termination_status == cp.ITERATION_LIMIT and st.kkt_passes == 100
    _, st = cp.solve(milp, cp.PDLP(time_limit=0.0, termination_reltol=1e-12))
    assert st.termination_status == cp.TIME_LIMIT and st.kkt_passes == 100


def test_warm_start_and_empty_constraints(cp):
    g = G.column_lp(200, 400, 5, seed=8)
    milp = cp.MILP.from_generated(g)
    sol, st = cp.solve(milp, cp.PrimalDualSolution(g.x0, g.y0), cp.PDLP())
    assert st.termination_status == cp.OPTIMAL and st.kkt_passes == 100
    # no constraints and c == 0: early exit, x = proj_box(0, lv, uv) (common.jl:194-198)
    e = cp.MILP(c=np.zeros(3), lv=np.array([1.0, -2.0, -np.inf]), uv=np.array([2.0, -1.0, np.inf]),
                A=sp.csc_matrix((0, 3)), lc=np.zeros(0), uc=np.zeros(0))
    sol, st = cp.solve(e, cp.PDLP())
    assert st.termination_status == cp.OPTIMAL
    assert np.array_equal(sol.x, [1.0, -1.0, 0.0])


@pytest.mark.parametrize("algo,fmt,prec", [("PDLP", "csr", "float64"), ("PDLP", "auto", "float64"),
                                           ("PDLP", "ell", "float64"), ("PDHG", "csr", "float64"),
                                           ("PDLP", "csr", "float32"), ("PDLP", "auto", "mixed")])
def test_degenerate_shapes_match_oracle(cp, algo, fmt, prec):
    """Ragged and degenerate inputs through the whole loop, against the oracle on the same
    LP: no constraints with a nonzero cost (A has no rows: the step size has no operator
    norm to follow, every variable goes to the bound its cost points at), an empty row and
    an empty column, a single variable with a single free-sided constraint, and a free
    row with a fixed variable.  Same status and pass count; the solution within 1e-6 (FP32
    and mixed: the same status, the solution within 1e-5 at a 1e-5 tolerance)."""
    cases = [
        _lp(np.zeros((0, 3)), [1.0, -2.0, 0.5], [0, -1, -3], [2, 4, 5], [], []),
        _lp([[1.0, 0, 1.0], [0, 0, 0], [2.0, 0, -1.0]], [1.0, 1.0, 1.0], [0, 0, 0], [10, 10, 10],
            [1, -1, 0], [1, 1, 5]),
        _lp([[2.0]], [3.0], [-np.inf], [np.inf], [4.0], [np.inf]),
        # a free row (both bounds infinite) and a fixed variable (lv = uv)
        _lp([[1.0, 1.0, 0.0], [1.0, -1.0, 2.0], [0.0, 1.0, 1.0]], [1.0, 2.0, -1.0], [0, 1.5, -5], [4, 1.5, 5],
            [-np.inf, 0.5, 1.0], [np.inf, 3.0, 1.0]),
    ]
    A_ = cp.PDLP if algo == "PDLP" else cp.PDHG
    reltol = 1e-6 if prec == "float64" else 1e-5
    for k, g in enumerate(cases):
        ref = O.solve(g, O.default_params(algo, termination_reltol=reltol))
        sol, st = cp.solve(cp.MILP.from_generated(g), A_(prec, None, fmt, termination_reltol=reltol))
        assert st.termination_status.name == ref.status, (k, st.termination_status, ref.status)
        if prec == "float64":
            assert abs(st.kkt_passes - ref.stats.kkt_passes) <= 100, (k, st.kkt_passes, ref.stats.kkt_passes)
        np.testing.assert_allclose(sol.x, ref.x, rtol=reltol, atol=reltol, err_msg=str(k))


@pytest.mark.parametrize("fmt", ["csr", "auto"])
def test_random_shapes_match_oracle(cp, fmt):
    """Twelve feasible LPs of random shape (m 10-300, n/m 1.2-3, 2-8 nonzeros per column,
    0-60% boxed columns) to 1e-6, and six of the reference's own random problems
    (src/utils/test.jl: random bounds, a quarter infinite and a quarter equal, which do not
    converge) over their first checks: the GPU run follows the oracle's FP64 run of the
    same LP -- the same status and pass count within one check, the objective within 1e-6,
    and every relative KKT error of the first five checks within 1e-9."""
    rng = np.random.default_rng(7)
    lps = []
    for k in range(12):
        m = int(rng.integers(10, 300))
        n, kk = int(m * rng.uniform(1.2, 3.0)), int(rng.integers(2, 9))
        lps.append((G.column_lp(m, n, kk, seed=500 + k, box_frac=float(rng.uniform(0, 0.6))), 60_000))
    for k in range(6):
        m, n = int(rng.integers(5, 200)), int(rng.integers(5, 400))
        lps.append((G.random_milp_and_sol(m, n, float(rng.uniform(0.03, 0.3)), seed=100 + k)[0], 600))
    for k, (g, cap) in enumerate(lps):
        ref = O.solve(g, O.default_params("PDLP", termination_reltol=1e-6, max_kkt_passes=cap))
        sol, st = cp.solve(cp.MILP.from_generated(g),
                           cp.PDLP("float64", None, fmt, termination_reltol=1e-6, max_kkt_passes=cap))
        assert st.termination_status.name == ref.status, (k, st.termination_status, ref.status)
        assert abs(st.kkt_passes - ref.stats.kkt_passes) <= 100, (k, st.kkt_passes, ref.stats.kkt_passes)
        if ref.status == "OPTIMAL":
            obj, obj_ref = float(sol.x @ g.c), float(ref.x @ g.c)
            assert abs(obj - obj_ref) <= 1e-6 * (1 + abs(obj_ref)), (k, obj, obj_ref)
        for (p1, d1), (p2, e2) in list(zip(ref.history, st.error_history))[:5]:
            e1 = max(d1["primal"] / d1["primal_scale"], d1["dual"] / d1["dual_scale"],
                     d1["gap"] / d1["gap_scale"])
            assert p1 == p2 and abs(e1 - e2.relative()) <= 1e-9 * e1, (k, p1, e1, e2.relative())


def test_float32_and_mixed(cp):
    g = G.column_lp(300, 700, 6, seed=13)
    milp = cp.MILP.from_generated(g)
    sol, st = cp.solve(milp, cp.PDLP("mixed", termination_reltol=1e-7))
    assert st.termination_status == cp.OPTIMAL
    assert st.fp32_passes > 0 and st.kkt_passes > st.fp32_passes
    assert abs(cp.objective_value(sol.x, milp) - g.objective) <= 1e-5 * (1 + abs(g.objective))
    # the eps(Float32)*n^3 power-method rule stops after one Rayleigh quotient here
    # (eta*||A|| > 1 diverges, as it would in the reference): converge it instead
    sol, st = cp.solve(milp, cp.PDLP("float32", termination_reltol=1e-4, powm_tol=1e-6))
    assert st.termination_status == cp.OPTIMAL
    assert abs(cp.objective_value(sol.x, milp) - g.objective) <= 1e-2 * (1 + abs(g.objective))


@pytest.mark.parametrize("fmt", ["csr", "xslice", "xslice_all", "auto"])
def test_power_law_lp_formats(cp, fmt):
    """Long rows (wavefront path) + XCD slices on a small power-law LP: same status and
    objective as the oracle; the slice format's row sums differ only in rounding."""
    g = G.power_law_lp(3000, 6000, 60_000, seed=2)
    milp = cp.MILP.from_generated(g)
    kw = dict(termination_reltol=1e-5, powm_tol=1e-6, max_kkt_passes=400_000)
    sol, st = cp.solve(milp, cp.PDLP("float64", None, fmt, **kw))
    ref = O.solve(g, O.default_params(termination_reltol=1e-5, powm_tol=1e-6, max_kkt_passes=400_000))
    assert st.termination_status == cp.OPTIMAL and ref.status == "OPTIMAL"
    obj = cp.objective_value(sol.x, milp)
    assert abs(obj - g.objective) <= 1e-4 * (1 + abs(g.objective))
    assert abs(obj - float(ref.x @ g.c)) <= 1e-4 * (1 + abs(g.objective))


@pytest.mark.parametrize("powm_tol", [1e-6, 1e-4])
def test_min_cost_flow_mixed_vs_highs(cp, powm_tol):
    """Mixed precision (FP32 bulk, FP64 checks + polish) on a small min-cost flow:
    objective within 1e-5 relative of HiGHS's optimum, flow conservation to 1e-5.  With
    powm_tol 1e-4 the power method runs on the FP32 matrix."""
    from scipy.optimize import linprog
    g = G.min_cost_flow_lp(nodes=2000, out_degree=8, seed=5)
    A = g.csc()
    h = linprog(g.c, A_eq=A, b_eq=g.lc, bounds=list(zip(g.lv, g.uv)), method="highs")
    assert h.status == 0
    milp = cp.MILP.from_generated(g)
    sol, st = cp.solve(milp, cp.PDLP("mixed", termination_reltol=1e-6, powm_tol=powm_tol,
                                     max_kkt_passes=2_000_000))
    assert st.termination_status == cp.OPTIMAL
    assert st.fp32_passes > 0
    obj = float(sol.x @ g.c)
    assert abs(obj - h.fun) <= 1e-5 * (1 + abs(h.fun)), (obj, h.fun)
    assert np.max(np.abs(A @ sol.x - g.lc)) <= 1e-5 * (1 + np.max(np.abs(g.lc)))


def test_min_cost_flow_mixed_vs_oracle_fp64_at_1e8(cp):
    """The min-cost-flow workload's stated check at reduced size (5,000 nodes, 40,000
    arcs): the mixed-precision solve at relative KKT 1e-6 against the oracle's FP64 solve
    of the same LP at 1e-8.  Stated errors: objective within 1e-5 relative of the oracle's,
    flow conservation |Ax - b| within 1e-5 of (1 + max |b|), bounds within 1e-9."""
    g = G.min_cost_flow_lp(nodes=5000, out_degree=8, seed=6)
    A = g.csc()
    ref = O.solve(g, O.default_params(termination_reltol=1e-8, powm_tol=1e-6, max_kkt_passes=2_000_000,
                                      time_limit=600))
    assert ref.status == "OPTIMAL"
    milp = cp.MILP.from_generated(g)
    sol, st = cp.solve(milp, cp.PDLP("mixed", termination_reltol=1e-6, powm_tol=1e-6,
                                     max_kkt_passes=2_000_000))
    assert st.termination_status == cp.OPTIMAL and st.fp32_passes > 0
    obj, obj_ref = float(sol.x @ g.c), float(ref.x @ g.c)
    assert abs(obj - obj_ref) <= 1e-5 * (1 + abs(obj_ref)), (obj, obj_ref)
    assert np.max(np.abs(A @ sol.x - g.lc)) <= 1e-5 * (1 + np.max(np.abs(g.lc)))
    assert np.all(sol.x >= g.lv - 1e-9) and np.all(sol.x <= g.uv + 1e-9)


@pytest.mark.parametrize("fmt,code,slices", [("xslice", 3, None), ("xslice_all", 4, None),
                                             ("xslice", 3, "2,4"), ("xslice_all", 4, "4,2")])
def test_xslice_long_rows_track_csr(cp, fmt, code, slices):
    """XCD column slices of the long rows (xslice) or of every row (xslice_all),
    including row parts longer than one tile (rows of 15k-30k nonzeros): the KKT history
    tracks the CSR run's to rounding over the first checks, and the format is reported
    per side.  Also with 2 and 4 slices per side (CPDLP_SLICES="<A>,<A'>"), where the
    combine adds fewer slice sums than its 8 registers hold."""
    import os
    g = G.power_law_lp(2000, 60000, 300_000, cap=40_000, seed=4)
    milp = cp.MILP.from_generated(g)
    kw = dict(termination_reltol=1e-12, max_kkt_passes=1500, powm_tol=1e-6)
    _, s_csr = cp.solve(milp, cp.PDLP("float64", None, "csr", **kw))
    if slices:
        os.environ["CPDLP_SLICES"] = slices
    try:
        _, s_xs = cp.solve(milp, cp.PDLP("float64", None, fmt, **kw))
    finally:
        os.environ.pop("CPDLP_SLICES", None)
    assert (s_xs.format_A, s_xs.format_At) == (code, code)
    assert len(s_csr.error_history) == len(s_xs.error_history) >= 10
    for (p1, e1), (p2, e2) in list(zip(s_csr.error_history, s_xs.error_history))[:8]:
        assert p1 == p2
        assert abs(e1.relative() - e2.relative()) <= 1e-7 * e1.relative(), (p1, e1, e2)


@pytest.mark.parametrize("algo,prec", [("PDLP", "float64"), ("PDHG", "float64"), ("PDLP", "float32")])
def test_chunked_loop_equals_solve(cp, algo, prec):
    """initialize + run_checks one chunk at a time (the reference's solve! yielding per
    chunk) reproduces one solve exactly: same iterates, passes and history."""
    import torch
    g = G.column_lp(300, 700, 6, seed=9)
    milp = cp.MILP.from_generated(g)
    A = cp.PDLP if algo == "PDLP" else cp.PDHG
    algo_ = A(prec, termination_reltol=1e-5, max_kkt_passes=40_000, powm_tol=1e-6)
    with cp.Engine(milp, algo_) as eng:
        sol, st = eng.solve()
        eng.initialize()
        status, calls = cp.STILL_RUNNING, 0
        while status == cp.STILL_RUNNING:
            status = eng.run_checks(1)
            calls += 1
        st2 = eng.stats()
        sol2 = eng.solution()
        xd = torch.empty(g.n, dtype=torch.float64, device="cuda")
        yd = torch.empty(g.m, dtype=torch.float64, device="cuda")
        eng.solution_device(xd, yd)
    assert status == st.termination_status == st2.termination_status
    assert st2.kkt_passes == st.kkt_passes
    assert calls == st.kkt_passes // 100
    assert np.array_equal(sol.x, sol2.x) and np.array_equal(sol.y, sol2.y)
    assert [p for p, _ in st.error_history] == [p for p, _ in st2.error_history]
    assert np.array_equal(xd.cpu().numpy(), sol2.x) and np.array_equal(yd.cpu().numpy(), sol2.y)


def test_preprocess_initialize_solve_inplace(cp):
    """The reference's free functions (common.jl:147-216): preprocess gives the scaled
    problem and start point with D1, D2 equal to the oracle's scaling bit for bit (and
    the scaled matrix equal to the oracle's (D1 a) D2); initialize + solve_inplace equals
    solve bit for bit."""
    lp = G.column_lp(300, 700, 6, seed=9)
    rng = np.random.default_rng(3)
    x0, y0 = rng.standard_normal(lp.n), rng.standard_normal(lp.m)
    milp = cp.MILP.from_generated(lp)
    algo = cp.PDLP(termination_reltol=1e-6, max_kkt_passes=40_000)
    start = cp.PrimalDualSolution(x0, y0)
    scaled, ssol = cp.preprocess(milp, start, algo)
    oD1, oD2, oa, _ = O.scaling(lp.m, lp.n, lp.colptr, lp.rowval, lp.nzval, 10, 1.0)
    assert np.array_equal(scaled.D1, oD1) and np.array_equal(scaled.D2, oD2)
    assert np.array_equal(scaled.A.indptr, lp.colptr) and np.array_equal(scaled.A.data, oa)
    assert np.array_equal(scaled.c, oD2 * lp.c) and np.array_equal(scaled.uc, oD1 * lp.uc)
    assert np.array_equal(scaled.lv, lp.lv / oD2) and np.array_equal(ssol.y, y0 / oD1)
    assert np.array_equal(ssol.x, x0 / oD2)
    sol, st = cp.solve(milp, start, algo)
    with cp.initialize(milp, start, algo) as state:
        first = state.get_state()
        assert np.array_equal(first.x[0], ssol.x) and np.array_equal(first.y[0], ssol.y)
        sol2, st2 = cp.solve_inplace(state, milp, algo)
    assert st.termination_status == st2.termination_status == cp.OPTIMAL
    assert st2.kkt_passes == st.kkt_passes
    assert np.array_equal(sol.x, sol2.x) and np.array_equal(sol.y, sol2.y)
    assert [p for p, _ in st.error_history] == [p for p, _ in st2.error_history]
    # mixed precision: the same (FP64) scaling
    scaled_m, _ = cp.preprocess(milp, start, cp.PDLP("mixed", termination_reltol=1e-6))
    assert np.array_equal(scaled_m.D1, scaled.D1) and np.array_equal(scaled_m.D2, scaled.D2)


@pytest.mark.parametrize("reltol,polish_reltol", [(1e-7, 1e-4), (1e-9, 1e-12)])
def test_chunked_loop_mixed_equals_solve(cp, reltol, polish_reltol):
    """Mixed precision through initialize + run_checks one chunk at a time: the FP32 phase,
    the hand-over and the FP64 phase give solve's bits (the same history, passes, FP32
    passes and solution), whether the FP32 phase ends at its own tolerance (1e-4) or by
    the early switch (1e-9, which FP32 iterates cannot reach)."""
    g = G.column_lp(300, 700, 6, seed=9)
    milp = cp.MILP.from_generated(g)
    algo = cp.PDLP("mixed", termination_reltol=reltol, polish_reltol=polish_reltol, max_kkt_passes=60_000)
    with cp.Engine(milp, algo) as eng:
        sol, st = eng.solve()
        eng.initialize()
        status = cp.STILL_RUNNING
        while status == cp.STILL_RUNNING:
            status = eng.run_checks(1)
        sol2, st2 = eng.solution(), eng.stats()
    assert st.termination_status == st2.termination_status == cp.OPTIMAL
    assert 0 < st.fp32_passes < st.kkt_passes
    assert (st2.kkt_passes, st2.fp32_passes, st2.restarts) == (st.kkt_passes, st.fp32_passes, st.restarts)
    assert np.array_equal(sol.x, sol2.x) and np.array_equal(sol.y, sol2.y)
    assert [(p, e) for p, e in st.error_history] == [(p, e) for p, e in st2.error_history]


# ------------------------------------------------------------------ full size
@pytest.fixture(scope="module")
def large_lp():
    return G.power_law_lp()   # m=1,000,000, n=2,000,000, nnz=20,000,027 (the benchmark LP)


def test_full_size_csr_and_scaling_bit_exact(cp, large_lp):
    """At the benchmark LP's real shape: CSR(A) construction and the Ruiz + Chambolle-Pock
    scaling (D1, D2 and both scaled value arrays) equal the oracle's bit for bit."""
    g = large_lp
    milp = cp.MILP.from_generated(g)
    with cp.Engine(milp, cp.PDLP("float64", None, "csr", termination_reltol=1e-6)) as eng:
        rp, col = eng.csr()
        orp, ocol, _ = O.csc_to_csr(g.m, g.n, g.colptr, g.rowval, g.nzval)
        assert np.array_equal(rp, orp) and np.array_equal(col, ocol)
        del orp, ocol, rp, col
        D1, D2, a, at = eng.scaling()
    oD1, oD2, oa, oat = O.scaling(g.m, g.n, g.colptr, g.rowval, g.nzval, 10, 1.0)
    assert np.array_equal(D1, oD1) and np.array_equal(D2, oD2)
    assert np.array_equal(a, oa) and np.array_equal(at, oat)


@pytest.mark.parametrize("fmt,code", [("xslice", 3), ("xslice_all", 4)])
def test_full_size_xslice_tracks_csr(cp, large_lp, fmt, code):
    """At the benchmark LP's real shape: the XCD-slice formats (steps, KKT check and power
    method) follow the CSR run's KKT history to rounding over the first checks."""
    milp = cp.MILP.from_generated(large_lp)
    kw = dict(termination_reltol=1e-12, max_kkt_passes=500, powm_tol=1e-4)
    _, s_csr = cp.solve(milp, cp.PDLP("float64", None, "csr", **kw))
    _, s_xs = cp.solve(milp, cp.PDLP("float64", None, fmt, **kw))
    assert s_xs.format_A == code and s_csr.format_A == 1
    assert abs(s_xs.eta - s_csr.eta) <= 1e-10 * s_csr.eta
    assert len(s_csr.error_history) == len(s_xs.error_history) >= 5
    for (p1, e1), (p2, e2) in zip(s_csr.error_history, s_xs.error_history):
        assert p1 == p2
        assert abs(e1.relative() - e2.relative()) <= 1e-8 * e1.relative(), (p1, e1, e2)


@pytest.mark.parametrize("precision,fmt,which", [("float64", "csr", "column"), ("mixed", "csr", "column"),
                                                 ("float64", "xslice", "power_law"),
                                                 ("float64", "xslice", "wide")])
def test_restart_error_cache_keeps_the_bits(cp, precision, fmt, which):
    """The check reuses the restart point's KKT errors between restarts (4 candidates
    evaluated instead of 5) and gathers the 4 candidates interleaved (one request per
    entry); with the candidates as separate vectors (CPDLP_KKT_SEPARATE) and n beyond 1M
    sliced columns it runs its rows side as two passes of 2 candidates ("wide": 4 x 8 B x
    n / 8 slices > 4 MiB per XCD).  In every form the whole error history, restarts, pass
    count and solution are the same bits as evaluating all 5 as separate vectors in one
    pass at every check (CPDLP_KKT_FULL_CHECK)."""
    import os
    g = (G.column_lp(400, 1000, 6, seed=23, box_frac=0.3) if which == "column"
         else G.power_law_lp(2000, 60000, 300_000, cap=40_000, seed=4) if which == "power_law"
         else G.power_law_lp(20_000, 1_100_000, 3_000_000, cap=40_000, seed=6))
    milp = cp.MILP.from_generated(g)
    kw = dict(termination_reltol=1e-7, powm_tol=1e-6, max_kkt_passes=20_000 if which != "wide" else 4_000)
    runs = []
    for env in (None, "CPDLP_KKT_SEPARATE", "CPDLP_KKT_FULL_CHECK"):
        if env:
            os.environ[env] = "1"
        try:
            runs.append(cp.solve(milp, cp.PDLP(precision, None, fmt, **kw)))
        finally:
            if env:
                os.environ.pop(env, None)
    s2, t2 = runs[-1]
    for s1, t1 in runs[:-1]:
        assert t1.restarts >= 2 and t1.restarts == t2.restarts
        assert t1.kkt_passes == t2.kkt_passes and t1.termination_status == t2.termination_status
        assert len(t1.error_history) == len(t2.error_history)
        for (p1, e1), (p2, e2) in zip(t1.error_history, t2.error_history):
            assert p1 == p2 and e1 == e2, (p1, e1, e2)
        assert np.array_equal(s1.x, s2.x) and np.array_equal(s1.y, s2.y)


@pytest.mark.parametrize("order", ["density", "given"])
def test_row_orders_reach_the_same_optimum(cp, order):
    """CPDLP_ORDER lays the short rows and columns out in density order (increasing
    stored-entry count, the order sort_rows_columns gives, src/components/permutation.jl:
    16-46) or in the caller's order instead of first touch: the same LP reaches the same
    optimum (objective to its closed form and to the first-touch solve's) through every
    format; an unknown order is refused."""
    import os
    g = G.power_law_lp(2000, 60000, 300_000, cap=40_000, seed=4)
    milp = cp.MILP.from_generated(g)
    kw = dict(termination_reltol=1e-8, powm_tol=1e-6, max_kkt_passes=40_000)
    runs = {}
    for o in ("first_touch", order):
        os.environ["CPDLP_ORDER"] = o
        try:
            for fmt in ("csr", "auto"):
                runs[o, fmt] = cp.solve(milp, cp.PDLP("float64", None, fmt, **kw))
        finally:
            os.environ.pop("CPDLP_ORDER", None)
    ref = float(runs["first_touch", "csr"][0].x @ g.c)
    for (o, fmt), (sol, st) in runs.items():
        assert st.termination_status == cp.OPTIMAL, (o, fmt, st)
        obj = float(sol.x @ g.c)
        assert abs(obj - g.objective) <= 1e-6 * (1 + abs(g.objective)), (o, fmt)
        assert abs(obj - ref) <= 1e-6 * (1 + abs(ref)), (o, fmt)
    os.environ["CPDLP_ORDER"] = "random"
    try:
        with pytest.raises(cp.ArgumentError, match="CPDLP_ORDER"):
            cp.solve(milp, cp.PDLP("float64", **kw))
    finally:
        os.environ.pop("CPDLP_ORDER", None)


@pytest.mark.parametrize("precision,fmt,which", [("float64", "csr", "column"), ("mixed", "csr", "column"),
                                                 ("float64", "csr", "power_law"),
                                                 ("float64", "auto", "wide")])
def test_line_grouped_tiles_keep_the_bits(cp, precision, fmt, which):
    """The steps on line-grouped tiles (CPDLP_LINE_TILES=1: 512-row, 4,096-entry tiles whose
    entries are stored in column order, so lanes of one load instruction share 128-byte
    lines of the gathered vector) give the row-ordered tiles' bits (=0): each product lands
    in its own LDS slot and every row is summed in the same order.  The whole error
    history, restarts, pass count and solution are identical, in FP64 and mixed, with and
    without long rows (wavefront rows in the same launch) and under the tuner's choice
    (unset: built where the request estimate gains >= 10%, kept where it times faster);
    the stats report the sides that used them."""
    import os
    g = (G.column_lp(400, 1000, 6, seed=23, box_frac=0.3) if which == "column"
         else G.power_law_lp(2000, 60000, 300_000, cap=40_000, seed=4) if which == "power_law"
         else G.power_law_lp(20_000, 1_100_000, 3_000_000, cap=40_000, seed=6))
    milp = cp.MILP.from_generated(g)
    kw = dict(termination_reltol=1e-7, powm_tol=1e-6, max_kkt_passes=20_000 if which != "wide" else 4_000)
    runs = {}
    for mode in ("0", "1", None):
        if mode is not None:
            os.environ["CPDLP_LINE_TILES"] = mode
        try:
            runs[mode] = cp.solve(milp, cp.PDLP(precision, None, fmt, **kw))
        finally:
            os.environ.pop("CPDLP_LINE_TILES", None)
    s0, t0 = runs["0"]
    assert t0.line_tiles == 0
    t1 = runs["1"][1]   # forced: every CSR side runs its steps on them (CSR = format code 1)
    assert t1.line_tiles == (t1.format_A == 1) + 2 * (t1.format_At == 1) and t1.format_At == 1, t1
    assert t0.restarts >= 1
    for mode in ("1", None):
        s1, t1 = runs[mode]
        assert t1.kkt_passes == t0.kkt_passes and t1.restarts == t0.restarts, mode
        assert t1.termination_status == t0.termination_status
        assert len(t1.error_history) == len(t0.error_history)
        for (p1, e1), (p0, e0) in zip(t1.error_history, t0.error_history):
            assert p1 == p0 and e1 == e0, (mode, p1, e1, e0)
        assert np.array_equal(s1.x, s0.x) and np.array_equal(s1.y, s0.y), mode
