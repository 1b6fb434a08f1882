"""CPU tests of the MOI-style optimizer (coolpdlp_hip.moi, mirroring src/moi.jl): model
building, options, status mapping, and the post-processing of a solve -- reduced costs,
objective and dual objective, constraint duals -- fed with the oracle's solution of the
reference's own MOI test LP (test/moi.jl:44-52, objective 205.0).  The LP's exact
primal and dual solutions are worked by hand: x = 15, y = 1.25; row duals 0.25 and 1.5."""
import numpy as np
import pytest

import oracle as O
from coolpdlp_hip import generators as G
from coolpdlp_hip import moi
from coolpdlp_hip._lib import ArgumentError

INF = np.inf


def _model(max_sense=False, constant=0.0):
    """x >= 0, 0 <= y <= 3; min 12x + 20y (+ constant); 6x + 8y >= 100; 7x + 12y >= 120."""
    opt = moi.Optimizer()
    x, y = opt.add_variables(2)
    cx = opt.add_constraint(x, moi.GreaterThan(0.0))
    cy = opt.add_constraint(y, moi.Interval(0.0, 3.0))
    sign = -1.0 if max_sense else 1.0
    opt.set_objective(moi.ScalarAffineFunction([(sign * 12.0, x), (sign * 20.0, y)], sign * constant),
                      moi.ObjectiveSense.MAX_SENSE if max_sense else moi.ObjectiveSense.MIN_SENSE)
    c1 = opt.add_constraint(moi.ScalarAffineFunction([(6.0, x), (8.0, y)]), moi.GreaterThan(100.0))
    c2 = opt.add_constraint(moi.ScalarAffineFunction([(7.0, x), (12.0, y)]), moi.GreaterThan(120.0))
    return opt, (x, y), (cx, cy, c1, c2)


def _oracle_solution(milp):
    A = milp.A.tocsc()
    g = G.GeneratedLP(A.shape[0], A.shape[1], A.indptr.astype(np.int64), A.indices.astype(np.int32),
                      A.data.copy(), milp.c, milp.lv, milp.uv, milp.lc, milp.uc)
    r = O.solve(g, O.default_params(termination_reltol=1e-8))
    assert r.status == "OPTIMAL"
    return r.x, r.y


def test_model_to_milp():
    opt, _, _ = _model()
    milp, max_sense, constant = opt._problem()
    assert not max_sense and constant == 0.0
    assert np.array_equal(milp.A.toarray(), [[6, 8], [7, 12]])
    assert np.array_equal(milp.c, [12, 20])
    assert np.array_equal(milp.lv, [0, 0]) and np.array_equal(milp.uv, [INF, 3])
    assert np.array_equal(milp.lc, [100, 120]) and np.array_equal(milp.uc, [INF, INF])
    # maximisation negates c (src/moi.jl:181-183); the constant is kept aside
    opt, _, _ = _model(max_sense=True, constant=7.0)
    milp, max_sense, constant = opt._problem()
    assert max_sense and constant == -7.0 and np.array_equal(milp.c, [12, 20])


def test_results_of_the_205_lp():
    opt, (x, y), (cx, cy, c1, c2) = _model()
    milp, max_sense, constant = opt._problem()
    xs, ys = _oracle_solution(milp)
    opt.finish(milp, xs, ys, "OPTIMAL", 0.5, max_sense, constant)
    assert opt.termination_status == moi.TerminationStatusCode.OPTIMAL
    assert opt.get_primal_status() == moi.ResultStatusCode.FEASIBLE_POINT
    assert opt.get_dual_status() == moi.ResultStatusCode.FEASIBLE_POINT
    assert opt.result_count == 1 and opt.raw_status_string == "OPTIMAL" and opt.solve_time == 0.5
    assert abs(opt.objective_value() - 205.0) <= 1e-4
    assert abs(opt.dual_objective_value() - 205.0) <= 1e-4        # strong duality
    assert abs(opt.variable_primal(x) - 15.0) <= 1e-4 and abs(opt.variable_primal(y) - 1.25) <= 1e-4
    assert abs(opt.constraint_dual(c1) - 0.25) <= 1e-4 and abs(opt.constraint_dual(c2) - 1.5) <= 1e-4
    # both variables are strictly inside their bounds: zero reduced costs
    assert abs(opt.constraint_dual(cx)) <= 1e-4 and abs(opt.constraint_dual(cy)) <= 1e-4
    assert opt.get_primal_status(2) == moi.ResultStatusCode.NO_SOLUTION
    with pytest.raises(ArgumentError):
        opt.objective_value(2)


def test_maximisation_and_constant():
    opt, _, _ = _model(max_sense=True, constant=7.0)
    milp, max_sense, constant = opt._problem()
    xs, ys = _oracle_solution(milp)
    opt.finish(milp, xs, ys, "OPTIMAL", 0.0, max_sense, constant)
    assert abs(opt.objective_value() - (-212.0)) <= 1e-4            # max -(12x + 20y) - 7
    assert abs(opt.dual_objective_value() - (-212.0)) <= 1e-4


def test_reduced_costs_and_bound_duals():
    """z = proj_multiplier(c - A'y, lv, uv) and the sign rules of the bound duals
    (src/moi.jl:128-153): GreaterThan gives |z|+, LessThan -|z|-, EqualTo/Interval z."""
    opt = moi.Optimizer()
    v = opt.add_variables(4)
    cis = [opt.add_constraint(v[0], moi.GreaterThan(1.0)), opt.add_constraint(v[1], moi.LessThan(2.0)),
           opt.add_constraint(v[2], moi.EqualTo(0.5)), opt.add_constraint(v[3], moi.Interval(-1.0, 1.0))]
    opt.set_objective(moi.ScalarAffineFunction([(3.0, v[0]), (-2.0, v[1]), (1.0, v[2]), (-4.0, v[3])]))
    opt.add_constraint(moi.ScalarAffineFunction([(1.0, v[0]), (1.0, v[1]), (1.0, v[2]), (1.0, v[3])], 1.0),
                       moi.LessThan(10.0))
    milp, max_sense, constant = opt._problem()
    assert milp.uc[0] == 9.0                      # the function's constant moves into the set
    y = np.array([-0.5])
    opt.finish(milp, np.zeros(4), y, "ITERATION_LIMIT", 1.0, max_sense, constant)
    r = milp.c - milp.A.T @ y                     # 3.5, -1.5, 1.5, -3.5
    assert np.allclose(opt.z, [3.5, -1.5, 1.5, -3.5])
    assert np.allclose(r, opt.z)
    assert opt.constraint_dual(cis[0]) == 3.5       # GreaterThan: |z|+
    assert opt.constraint_dual(cis[1]) == -1.5      # LessThan: -|z|-
    assert opt.constraint_dual(cis[2]) == 1.5       # EqualTo: z
    assert opt.constraint_dual(cis[3]) == -3.5      # Interval: z
    assert opt.termination_status == moi.TerminationStatusCode.ITERATION_LIMIT
    assert opt.get_primal_status() == moi.ResultStatusCode.UNKNOWN_RESULT_STATUS
    # dual objective lc'|y|+ - uc'|y|- + lv'|z|+ - uv'|z|-, where an infinite bound times a
    # part contributes the part itself (safeprod_left).  |y|+ = 0, |y|- = 0.5;
    # |z|+ = (3.5, 0, 1.5, 0), |z|- = (0, 1.5, 0, 3.5); lv = (1, -inf, 0.5, -1),
    # uv = (inf, 2, 0.5, 1):  0 - 9*0.5 + (3.5 + 0 + 0.75 + 0) - (0 + 3 + 0 + 3.5) = -6.75
    assert opt.dual_objective_value() == pytest.approx(-6.75)
    assert opt.objective_value() == 0.0


def test_status_mapping_and_state():
    opt, _, _ = _model()
    assert opt.is_empty() is False                # it has a model
    milp, _, _ = opt._problem()
    for status, want in [("TIME_LIMIT", (moi.TerminationStatusCode.TIME_LIMIT,
                                         moi.ResultStatusCode.UNKNOWN_RESULT_STATUS)),
                         ("ITERATION_LIMIT", (moi.TerminationStatusCode.ITERATION_LIMIT,
                                              moi.ResultStatusCode.UNKNOWN_RESULT_STATUS)),
                         ("STILL_RUNNING", (moi.TerminationStatusCode.OTHER_ERROR,
                                            moi.ResultStatusCode.NO_SOLUTION))]:
        opt.finish(milp, np.zeros(2), np.zeros(2), status, 0.0)
        assert (opt.termination_status, opt.primal_status) == want
        assert opt.dual_status == want[1]
    opt.empty()
    assert opt.is_empty() and opt.result_count == 0
    assert opt.termination_status == moi.TerminationStatusCode.OPTIMIZE_NOT_CALLED


def test_options_and_errors():
    opt = moi.Optimizer()
    assert opt.solver_name == "CoolPDLP" and opt.time_limit_sec is None
    opt.time_limit_sec = 12
    assert opt.get_raw("time_limit") == 12.0
    opt.time_limit_sec = None
    assert "time_limit" not in opt.options
    opt.set_raw("termination_reltol", 1e-5)
    assert opt.get_raw("termination_reltol") == 1e-5
    x = opt.add_variable()
    opt.add_constraint(x, moi.GreaterThan(0.0))
    with pytest.raises(ArgumentError):
        opt.add_constraint(x, moi.LessThan(1.0))  # a second bound on the same variable
    with pytest.raises(ArgumentError):
        opt.add_constraint(x, "not a set")


def test_read_from_file_builds_the_mps_problem():
    """Optimizer.read_from_file (the tutorial's JuMP.read_from_file of an MPS file): the
    model it builds is the problem MILP(qps) holds, for every Netlib-style fixture (ranged,
    free, fixed, upper-only variables, equality and ranged rows), with the objective
    constant and the sense kept; a second read replaces the first model."""
    import os

    import scipy.sparse as sp

    from coolpdlp_hip import MILP, read_mps
    d = os.path.join(os.path.dirname(__file__), "golden", "mps")
    opt = moi.Optimizer()
    for f in sorted(os.listdir(d)):
        if not f.endswith(".mps"):
            continue
        path = os.path.join(d, f)
        xs = opt.read_from_file(path)
        q = read_mps(path)
        ref = MILP.from_qps(q)
        milp, max_sense, constant = opt._problem()
        assert len(xs) == ref.A.shape[1] and not max_sense and constant == q.c0, f
        for name in ("c", "lv", "uv", "lc", "uc"):
            assert np.array_equal(getattr(milp, name), getattr(ref, name)), (f, name)
        assert (sp.csc_matrix(milp.A) != sp.csc_matrix(ref.A)).nnz == 0, f
    # a maximisation with a constant and a free row
    import io
    text = """NAME T
OBJSENSE
    MAX
ROWS
 N obj
 L r1
 N free
COLUMNS
    x1 obj 1.0 r1 1.0
    x1 free 2.0
    x2 obj 2.0 r1 1.0
RHS
    rhs obj -3.5 r1 4.0
BOUNDS
 UP bnd x1 3.0
 FR bnd x2
ENDATA
"""
    xs = opt.read_from_file(io.StringIO(text))
    milp, max_sense, constant = opt._problem()
    assert max_sense and len(xs) == 2
    assert np.array_equal(milp.c, [-1.0, -2.0])
    assert np.array_equal(milp.lv, [0.0, -np.inf]) and np.array_equal(milp.uv, [3.0, np.inf])
    assert milp.A.shape == (1, 2) and np.array_equal(milp.uc, [4.0]) and milp.lc[0] == -np.inf
    assert constant == read_mps(io.StringIO(text)).c0
