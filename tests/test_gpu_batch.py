"""GPU tests of the batched engine (many LPs, one workgroup each), through the C ABI.

Bars: every instance reaches OPTIMAL at its tolerance with the objective within
20 x reltol (relative) of the closed form c'x0 and of the oracle's FP64 solve; an
instance's result is bit-identical whether it is solved alone or in a mix, and
whatever the round length; the FP64 batch agrees with the plain engine's FP64 solve.
"""
import numpy as np
import pytest

import oracle as O
from coolpdlp_hip import generators as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cp():
    import coolpdlp_hip
    return coolpdlp_hip


def _mix(count, seed):
    return [G.column_lp(m, n, (6, 12), s, name=f"b{i}")
            for i, (m, n, s) in enumerate(G.batch_specs(count, 50, 600, seed))]


@pytest.mark.parametrize("precision,reltol", [("mixed", 1e-6), ("float64", 1e-6)])
def test_batch_matches_closed_form_and_oracle(cp, precision, reltol):
    lps = _mix(24, seed=7)
    with cp.BatchSolver(cp.PDLP(precision, termination_reltol=reltol), hbm_bytes=256 << 20) as B:
        res = B.solve_all(lps, checks_per_round=7)
    for g, (sol, st) in zip(lps, res):
        assert st.termination_status == cp.OPTIMAL, (g.name, st)
        assert st.err.relative() <= reltol
        obj = float(sol.x @ g.c)
        assert abs(obj - g.objective) <= 20 * reltol * (1 + abs(g.objective)), (g.name, obj, g.objective)
        if precision == "mixed":
            assert st.fp32_passes > 0
    for g, (sol, st) in list(zip(lps, res))[:4]:
        ref = O.solve(g, O.default_params(termination_reltol=reltol))
        assert ref.status == "OPTIMAL"
        assert abs(float(sol.x @ g.c) - float(ref.x @ g.c)) <= 20 * reltol * (1 + abs(g.objective))


def _lp(A, c, lv, uv, lc, uc):
    import scipy.sparse as sp
    A = sp.csc_matrix(np.asarray(A, float))
    return G.GeneratedLP(A.shape[0], A.shape[1], A.indptr.astype(np.int64), A.indices.astype(np.int32),
                         A.data.astype(float), *[np.asarray(v, float) for v in (c, lv, uv, lc, uc)])


@pytest.mark.parametrize("precision", ["float64", "mixed"])
def test_batch_degenerate_shapes(cp, precision):
    """Ragged and degenerate instances in one batch with an ordinary LP, against the
    oracle's FP64 solve: no constraints with a nonzero cost, no constraints and c = 0
    (the reference's early exit: x = proj_box(0), common.jl:194-198), an empty row and an
    empty column, a single variable with one constraint, and a free row with a fixed
    variable.  Same status; the solution
    within 1e-6 (1e-5 mixed); the ordinary LP keeps its bits from a batch of its own."""
    cases = [
        _lp(np.zeros((0, 3)), [1.0, -2.0, 0.5], [0, -1, -3], [2, 4, 5], [], []),
        _lp(np.zeros((0, 3)), [0.0, 0.0, 0.0], [1, -2, -np.inf], [2, -1, np.inf], [], []),
        _lp([[1.0, 0, 1.0], [0, 0, 0], [2.0, 0, -1.0]], [1.0, 1.0, 1.0], [0, 0, 0], [10, 10, 10],
            [1, -1, 0], [1, 1, 5]),
        _lp([[2.0]], [3.0], [-np.inf], [np.inf], [4.0], [np.inf]),
        # a free row (both bounds infinite) and a fixed variable (lv = uv)
        _lp([[1.0, 1.0, 0.0], [1.0, -1.0, 2.0], [0.0, 1.0, 1.0]], [1.0, 2.0, -1.0], [0, 1.5, -5], [4, 1.5, 5],
            [-np.inf, 0.5, 1.0], [np.inf, 3.0, 1.0]),
    ]
    ordinary = G.column_lp(120, 300, 6, seed=5)
    algo = cp.PDLP(precision, termination_reltol=1e-6)
    with cp.BatchSolver(algo, hbm_bytes=64 << 20) as B:
        res = B.solve_all(cases + [ordinary])
    with cp.BatchSolver(algo, hbm_bytes=64 << 20) as B:
        (sol_alone, st_alone), = B.solve_all([ordinary])
    tol = 1e-6 if precision == "float64" else 1e-5
    for k, (g, (sol, st)) in enumerate(zip(cases, res)):
        ref = O.solve(g, O.default_params(termination_reltol=1e-6))
        assert st.termination_status.name == ref.status == "OPTIMAL", (k, st.termination_status)
        np.testing.assert_allclose(sol.x, ref.x, rtol=tol, atol=tol, err_msg=str(k))
    assert np.array_equal(res[1][0].x, [1.0, -1.0, 0.0])
    sol_o, st_o = res[-1]
    assert st_o.termination_status == cp.OPTIMAL and st_o.kkt_passes == st_alone.kkt_passes
    assert np.array_equal(sol_o.x, sol_alone.x) and np.array_equal(sol_o.y, sol_alone.y)


def test_batch_fp64_agrees_with_plain_engine(cp):
    g = G.column_lp(200, 450, 6, seed=31)
    milp = cp.MILP.from_generated(g)
    sol_p, st_p = cp.solve(milp, cp.PDLP(termination_reltol=1e-7))
    with cp.BatchSolver(cp.PDLP("float64", termination_reltol=1e-7), hbm_bytes=64 << 20) as B:
        (sol_b, st_b), = B.solve_all([g])
    assert st_p.termination_status == st_b.termination_status == cp.OPTIMAL
    assert abs(st_p.eta - st_b.eta) <= 1e-9 * st_p.eta
    assert abs(st_b.kkt_passes - st_p.kkt_passes) <= 300
    assert abs(float(sol_b.x @ g.c) - float(sol_p.x @ g.c)) <= 1e-5 * (1 + abs(g.objective))


def test_batch_warm_start(cp):
    """cpdlp_batch_add_start: a start point as sol in solve(milp, sol, algo)
    (src/algorithms/common.jl:186-209).  The re-solve of an LP after a change of cost
    from the old optimum agrees with the plain engine's warm start; a cold instance in the
    same batch keeps its bits; a start at the optimum stops within the first checks."""
    import dataclasses
    g = G.column_lp(200, 450, 6, seed=33)
    milp = cp.MILP.from_generated(g)
    rng = np.random.default_rng(1)
    g_old = dataclasses.replace(g, c=g.c * (1 + 0.01 * rng.standard_normal(g.n)))
    start, _ = cp.solve(cp.MILP.from_generated(g_old), cp.PDLP(termination_reltol=1e-8))
    sol_p, st_p = cp.solve(milp, start, cp.PDLP(termination_reltol=1e-7))
    algo = cp.PDLP("float64", termination_reltol=1e-7)
    with cp.BatchSolver(algo, hbm_bytes=64 << 20) as B:
        (sol_w, st_w), (sol_c, st_c) = B.solve_all([g, g], sols=[start, None])
    with cp.BatchSolver(algo, hbm_bytes=64 << 20) as B:
        (sol_a, st_a), = B.solve_all([g])
    assert st_p.termination_status == st_w.termination_status == cp.OPTIMAL
    assert abs(st_w.kkt_passes - st_p.kkt_passes) <= 300
    assert abs(float(sol_w.x @ g.c) - float(sol_p.x @ g.c)) <= 1e-5 * (1 + abs(g.objective))
    assert abs(float(sol_w.x @ g.c) - g.objective) <= 1e-5 * (1 + abs(g.objective))
    # the cold instance beside the warm one is the cold instance alone, bit for bit
    assert np.array_equal(sol_c.x, sol_a.x) and np.array_equal(sol_c.y, sol_a.y)
    assert st_c.kkt_passes == st_a.kkt_passes
    # from an optimum at 1e-9, a re-solve at 1e-6 stops within the first checks, in FP64
    # and in mixed precision
    best, _ = cp.solve(milp, cp.PDLP(termination_reltol=1e-9))
    for prec in ("float64", "mixed"):
        with cp.BatchSolver(cp.PDLP(prec, termination_reltol=1e-6), hbm_bytes=64 << 20) as B:
            (sol_o, st_o), = B.solve_all([g], sols=[best])
        assert st_o.termination_status == cp.OPTIMAL and st_o.kkt_passes <= 300, (prec, st_o.kkt_passes)
        assert st_o.kkt_passes < st_a.kkt_passes
    with cp.BatchSolver(algo, hbm_bytes=64 << 20) as B:
        with pytest.raises(cp.DimensionMismatch):
            B.add(g, cp.PrimalDualSolution(np.zeros(g.n - 1), np.zeros(g.m)))


def test_alone_equals_in_mix_and_round_length_free(cp):
    lps = _mix(12, seed=9)
    target = lps[5]
    algo = cp.PDLP("mixed", termination_reltol=1e-6)
    with cp.BatchSolver(algo, hbm_bytes=128 << 20) as B:
        (alone, st_alone), = B.solve_all([target], checks_per_round=3)
    with cp.BatchSolver(algo, hbm_bytes=128 << 20) as B:
        mix = B.solve_all(lps, checks_per_round=11)
    sol_mix, st_mix = mix[5]
    assert st_alone.kkt_passes == st_mix.kkt_passes
    assert np.array_equal(alone.x, sol_mix.x) and np.array_equal(alone.y, sol_mix.y)


def test_continuous_admission_and_release(cp):
    """Instances admitted while others run finish with the same results; the arena
    refuses what does not fit and accepts it after a release."""
    lps = _mix(6, seed=11)
    algo = cp.PDLP("mixed", termination_reltol=1e-6)
    with cp.BatchSolver(algo, hbm_bytes=64 << 20) as B:
        ref = B.solve_all(lps, checks_per_round=5)
    with cp.BatchSolver(algo, hbm_bytes=64 << 20) as B:
        ids = {B.add(lps[0]): 0, B.add(lps[1]): 1}
        B.run(2)
        ids[B.add(lps[2])] = 2
        B.run(1)
        for k in (3, 4, 5):
            ids[B.add(lps[k])] = k
        got = {}
        while len(got) < 6:
            B.run(4)
            for i in B.poll():
                got[ids[i]] = B.get(i)
                B.release(i)
        for k in range(6):
            assert np.array_equal(got[k][0].x, ref[k][0].x)
            assert got[k][1].kkt_passes == ref[k][1].kkt_passes
    g = lps[0]
    fp = max(cp.batch.footprint(h.m, h.n, h.nnz) for h in lps[:2])
    with cp.BatchSolver(algo, hbm_bytes=fp + 1024) as B:
        i0 = B.add(g)
        with pytest.raises(MemoryError):
            B.add(lps[1])
        while B.run(10) > 0:
            pass
        B.get(i0)
        B.release(i0)
        B.add(lps[1])   # fits again


def test_batch_limits(cp):
    lps = _mix(3, seed=13)
    with cp.BatchSolver(cp.PDLP("mixed", termination_reltol=1e-12, max_kkt_passes=300),
                        hbm_bytes=64 << 20) as B:
        res = B.solve_all(lps)
    for sol, st in res:
        assert st.termination_status == cp.ITERATION_LIMIT and st.kkt_passes == 300


def test_batch_long_rows_and_columns(cp):
    """Rows and columns longer than one tile (2,048 nonzeros) take the whole-block path of
    the batch engine: same status and objective as the plain engine and the closed form."""
    import scipy.sparse as sp
    rng = np.random.default_rng(17)
    m, n = 600, 4000
    A = sp.random(m, n, density=0.002, format="lil", random_state=rng)
    A[0, :] = rng.standard_normal(n)            # one dense row (4,000 nonzeros)
    A[:, 1] = rng.standard_normal((m, 1))       # one dense column (600 nonzeros)
    A[5, :2500] = rng.standard_normal(2500)     # a row just over one tile
    A = sp.csc_matrix(A)
    x0 = np.where(rng.random(n) < 0.6, rng.random(n), 0.0)
    y0 = rng.standard_normal(m)
    s0 = np.where(x0 > 0, 0.0, rng.random(n))
    b = A @ x0
    g = G.GeneratedLP(m, n, A.indptr.astype(np.int64), A.indices.astype(np.int32), A.data.copy(),
                      A.T @ y0 + s0, np.zeros(n), np.full(n, np.inf), b, b)
    obj = float(g.c @ x0)
    milp = cp.MILP.from_generated(g)
    sol_p, st_p = cp.solve(milp, cp.PDLP("float64", termination_reltol=1e-6, max_kkt_passes=400_000))
    with cp.BatchSolver(cp.PDLP("float64", termination_reltol=1e-6, max_kkt_passes=400_000),
                        hbm_bytes=64 << 20) as B:
        (sol_b, st_b), = B.solve_all([g])
    assert st_p.termination_status == st_b.termination_status == cp.OPTIMAL
    assert abs(st_b.eta - st_p.eta) <= 1e-9 * st_p.eta
    for sol in (sol_p, sol_b):
        assert abs(float(g.c @ sol.x) - obj) <= 1e-4 * (1 + abs(obj))


def test_wide_workgroups_bit_identical(cp, monkeypatch):
    """Large instances get workgroups of 4 x 256 threads summing four tiles side by side.
    Forcing every instance onto them (CPDLP_BATCH_WIDE_NNZ=1) gives the same bits as the
    256-thread kernels: same passes, same iterates, in FP32 and FP64, for tiles of short
    rows and for rows longer than a tile (the whole-sub-block path)."""
    import scipy.sparse as sp
    lps = _mix(10, seed=21)
    rng = np.random.default_rng(5)
    m, n = 400, 3000
    A = sp.random(m, n, density=0.003, format="lil", random_state=rng)
    A[0, :] = rng.standard_normal(n)            # a row of 3,000 nonzeros
    A[:, 2] = rng.standard_normal((m, 1))
    A = sp.csc_matrix(A)
    x0 = np.where(rng.random(n) < 0.6, rng.random(n), 0.0)
    y0 = rng.standard_normal(m)
    s0 = np.where(x0 > 0, 0.0, rng.random(n))
    b = A @ x0
    lps.append(G.GeneratedLP(m, n, A.indptr.astype(np.int64), A.indices.astype(np.int32),
                             A.data.copy(), A.T @ y0 + s0, np.zeros(n), np.full(n, np.inf), b, b))
    algo = cp.PDLP("mixed", termination_reltol=1e-6, max_kkt_passes=200_000)
    with cp.BatchSolver(algo, hbm_bytes=256 << 20) as B:
        narrow = B.solve_all(lps, checks_per_round=6)
    monkeypatch.setenv("CPDLP_BATCH_WIDE_NNZ", "1")
    with cp.BatchSolver(algo, hbm_bytes=256 << 20) as B:
        wide = B.solve_all(lps, checks_per_round=6)
    for (s1, t1), (s2, t2) in zip(narrow, wide):
        assert t1.termination_status == t2.termination_status == cp.OPTIMAL
        assert t1.kkt_passes == t2.kkt_passes and t1.fp32_passes == t2.fp32_passes
        assert np.array_equal(s1.x, s2.x) and np.array_equal(s1.y, s2.y)


def _upper_range():
    """c3's upper range: m up to 5,000, n = m x U(1.5, 3), 6-12 nonzeros per column."""
    shapes = [(5000, 14000, 91), (5000, 9000, 92), (3100, 7500, 93), (2400, 4000, 94), (700, 1500, 95),
              (60, 150, 96)]
    return [G.column_lp(m, n, (6, 12), sd, name=f"u{m}_{n}") for m, n, sd in shapes]


def _solve_batch(cp, lps, precision, env=None):
    import os
    old = {k: os.environ.get(k) for k in (env or {})}
    try:
        os.environ.update(env or {})
        with cp.BatchSolver(cp.PDLP(precision, termination_reltol=1e-6), hbm_bytes=512 << 20) as B:
            return B.solve_all(lps, checks_per_round=10)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


@pytest.mark.parametrize("precision", ["float64", "mixed"])
def test_batch_upper_range_gathers_from_global_memory(cp, precision):
    """Instances up to m = 5,000 (the largest FP64 ones exceed the LDS budget) solved
    with the default LDS classes and with every instance gathering from global memory
    (CPDLP_BATCH_LDS_BYTES=0: the narrow and the wide k_batch_steps<V, false, ...>).  The
    two runs give the same bits; the results match the closed form c'x0 and the oracle's
    FP64 solve."""
    lps = _upper_range()
    res_lds = _solve_batch(cp, lps, precision)
    res_glb = _solve_batch(cp, lps, precision, {"CPDLP_BATCH_LDS_BYTES": "0"})
    for g, (sa, ta), (sb, tb) in zip(lps, res_lds, res_glb):
        assert ta.termination_status == tb.termination_status == cp.OPTIMAL, (g.name, ta, tb)
        assert ta.kkt_passes == tb.kkt_passes, g.name
        assert np.array_equal(sa.x, sb.x) and np.array_equal(sa.y, sb.y), g.name
        obj = float(sb.x @ g.c)
        assert abs(obj - g.objective) <= 20e-6 * (1 + abs(g.objective)), (g.name, obj, g.objective)
    for g, (sol, st) in list(zip(lps, res_glb))[:2]:
        ref = O.solve(g, O.default_params(termination_reltol=1e-6))
        assert ref.status == "OPTIMAL"
        assert abs(float(sol.x @ g.c) - float(ref.x @ g.c)) <= 20e-6 * (1 + abs(g.objective)), g.name


def test_batch_add_beyond_the_whole_arena(cp):
    """An LP larger than the whole arena is refused with CPDLP_ERR_ARGUMENT (it can never
    fit); one that does not fit now gives MemoryError; neither changes the arena."""
    from coolpdlp_hip.batch import footprint
    small = G.column_lp(60, 150, 6, seed=41)
    big = G.column_lp(1500, 3500, 8, seed=42)
    fp = footprint(small.m, small.n, small.nnz)
    with cp.BatchSolver(cp.PDLP("float64", termination_reltol=1e-6), hbm_bytes=2 * fp) as B:
        with pytest.raises(cp.ArgumentError, match="whole batch arena"):
            B.add(big)
        assert B.memory()[0] == 0
        i1, i2 = B.add(small), B.add(small)
        used = B.memory()[0]
        with pytest.raises(MemoryError):
            B.add(small)
        assert B.memory()[0] == used
        while B.run(10):
            pass
        for i in (i1, i2):
            sol, st = B.get(i)
            assert st.termination_status == cp.OPTIMAL
            assert abs(float(sol.x @ small.c) - small.objective) <= 20e-6 * (1 + abs(small.objective))
            B.release(i)
        assert B.memory()[0] == 0

