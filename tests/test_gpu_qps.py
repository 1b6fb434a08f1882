"""GPU tests of MPS ingestion end to end: the reference's MOI test LP (test/moi.jl:44-52,
objective 205.0) read from fixed-format MPS, turned into a MILP as MILP(qps::QPSData)
does (src/problems/milp.jl:122-137) and solved by the HIP engine, as the tutorial does
with a Netlib file (test/tutorial.jl:14-48); and the reference's Netlib integration
tests (test/algorithms/all.jl:20-73) on the Netlib-style MPS fixtures: PDLP and PDHG
against HiGHS's objective, and the CPU-GPU coherence check against the oracle."""
import io

import numpy as np
import pytest

import oracle as O
from coolpdlp_hip import (MILP, OPTIMAL, PDHG, PDLP, is_feasible, objective_value, read_mps,
                          solve)

from test_qps import LP205_FIXED, generated, row_scale, suite

INF = np.inf

pytestmark = pytest.mark.gpu


def test_mps_lp205_solves_on_gpu():
    milp = MILP.from_qps(read_mps(io.StringIO(LP205_FIXED)), dataset="hand", name="lp205")
    sol, stats = solve(milp, PDLP(termination_reltol=1e-8, time_limit=30.0))
    assert stats.termination_status == OPTIMAL
    assert abs(objective_value(sol.x, milp) - 205.0) <= 1e-4
    assert np.allclose(sol.x, [15.0, 1.25], atol=1e-4)
    assert is_feasible(sol.x, milp, cons_tol=1e-6, verbose=False)


@pytest.mark.parametrize("fmt", ["csr", "ell", "auto"])
def test_suite_pdlp_vs_highs(fmt):
    """PDLP at reltol 1e-5: OPTIMAL, feasible to 1e-2, objective within 1e-2 of HiGHS
    (test/algorithms/all.jl:54-62, 20-41)."""
    for name, milp, obj in suite():
        sol, st = solve(milp, PDLP("float64", None, fmt, termination_reltol=1e-5,
                                   max_kkt_passes=10 ** 7, time_limit=60.0))
        assert st.termination_status == OPTIMAL, name
        assert is_feasible(sol.x, milp, cons_tol=1e-2, int_tol=INF, verbose=False), name
        assert abs(objective_value(sol.x, milp) - obj) <= 1e-2 * abs(obj), name


def test_suite_pdhg_vs_highs():
    """PDHG at reltol 1e-1, objective rtol 1e-1 (test/algorithms/all.jl:43-52); cons_tol
    1e-1 scaled by 1 + the largest row bound, as in the oracle's test of the same suite."""
    for name, milp, obj in suite():
        sol, st = solve(milp, PDHG("float64", None, "auto", termination_reltol=1e-1,
                                   max_kkt_passes=10 ** 7, time_limit=60.0))
        assert st.termination_status == OPTIMAL, name
        assert is_feasible(sol.x, milp, cons_tol=1e-1 * row_scale(milp), int_tol=INF, verbose=False), name
        assert abs(objective_value(sol.x, milp) - obj) <= 1e-1 * abs(obj), name


def test_suite_cpu_gpu_coherence():
    """test/algorithms/all.jl:64-73: the fourth-smallest instance, reltol 1e-3, a check
    every pass; the GPU engine and the oracle stop at the same pass with the same errors
    up to reduction-order rounding."""
    name, milp, _ = suite()[3]
    _, st = solve(milp, PDLP("float64", None, "csr", termination_reltol=1e-3, check_every=1))
    ref = O.solve(generated(milp), O.default_params(termination_reltol=1e-3, check_every=1))
    assert st.termination_status == OPTIMAL and ref.status == "OPTIMAL"
    assert st.kkt_passes == ref.stats.kkt_passes, name
    for f in ("primal", "primal_scale", "dual", "dual_scale", "gap", "gap_scale"):
        a, b = getattr(st.err, f), getattr(ref.stats.err, f)
        assert abs(a - b) <= 1e-8 * (1 + abs(b)), (name, f, a, b)


def test_tutorial_workflow():
    """test/tutorial.jl end to end on the fourth-smallest fixture (the tutorial takes
    Netlib's fourth instance): MILP(qps) solved by PDLP at 1e-6 with a 10 s time limit, the
    FP32 solve, and the Optimizer reading the same MPS file (matrix_type CSR); each x
    feasible to 1e-3 and each objective within 1e-3 of HiGHS's; the relative error falls
    from the first check to the last (tutorial.jl:122-131)."""
    from coolpdlp_hip import GPUSparseMatrixCSR, Optimizer, relative
    name, milp, obj = suite()[3]
    sol, stats = solve(milp, PDLP(termination_reltol=1e-6, time_limit=10.0))
    sol32, stats32 = solve(milp, PDLP("float32", None, "csr", termination_reltol=1e-6, time_limit=10.0))
    opt = Optimizer()
    xs = opt.read_from_file(milp.path)
    opt.set_silent(True)
    opt.set_raw("termination_reltol", 1e-6)
    opt.set_raw("matrix_type", GPUSparseMatrixCSR)
    opt.optimize()
    x_jump = np.array([opt.variable_primal(v) for v in xs])
    assert relative(stats.error_history[-1][1]) < relative(stats.error_history[0][1])
    for x in (sol.x, sol32.x, x_jump):
        assert is_feasible(x, milp, cons_tol=1e-3, int_tol=INF, verbose=False), name
        assert abs(objective_value(x, milp) - obj) <= 1e-3 * abs(obj), (name, objective_value(x, milp), obj)
