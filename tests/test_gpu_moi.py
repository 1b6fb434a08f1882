"""GPU tests of the MOI-style optimizer end to end (coolpdlp_hip.moi.Optimizer.optimize):
the reference's own MOI cases, test/moi.jl:38-84 -- the 2-variable LP with objective
205.0 (atol 1e-2), solved in FP64, in FP32 (with the mismatched-float warning) and with
the ELL matrix type, and with PDHG through the "algorithm" raw attribute."""

import pytest

from coolpdlp_hip import moi

pytestmark = pytest.mark.gpu


def _build(opt):
    x, y = opt.add_variables(2)
    opt.add_constraint(x, moi.GreaterThan(0.0))
    opt.add_constraint(y, moi.Interval(0.0, 3.0))
    opt.set_objective(moi.ScalarAffineFunction([(12.0, x), (20.0, y)]))
    c1 = opt.add_constraint(moi.ScalarAffineFunction([(6.0, x), (8.0, y)]), moi.GreaterThan(100.0))
    c2 = opt.add_constraint(moi.ScalarAffineFunction([(7.0, x), (12.0, y)]), moi.GreaterThan(120.0))
    return x, y, c1, c2


@pytest.mark.parametrize("attrs", [{}, {"matrix_type": "ell"}, {"termination_reltol": 1e-8}])
def test_optimizer_205(attrs):
    opt = moi.Optimizer()
    opt.set_silent(True)
    for k, v in attrs.items():
        opt.set_raw(k, v)
    x, y, c1, c2 = _build(opt)
    opt.optimize()
    assert opt.termination_status == moi.TerminationStatusCode.OPTIMAL
    assert opt.get_primal_status() == moi.ResultStatusCode.FEASIBLE_POINT
    assert abs(opt.objective_value() - 205.0) <= 1e-2
    assert abs(opt.dual_objective_value() - 205.0) <= 1e-1
    assert abs(opt.variable_primal(x) - 15.0) <= 1e-2 and abs(opt.variable_primal(y) - 1.25) <= 1e-2
    assert abs(opt.constraint_dual(c1) - 0.25) <= 1e-2 and abs(opt.constraint_dual(c2) - 1.5) <= 1e-2
    assert opt.solve_time > 0
    assert opt.options == attrs          # the options persist across optimize()


def test_optimizer_float32_warns():
    opt = moi.Optimizer()
    opt.set_silent(True)
    opt.set_raw("float_type", "float32")
    x, y, _, _ = _build(opt)
    with pytest.warns(UserWarning, match="mismatched float type"):
        opt.optimize()
    assert opt.termination_status == moi.TerminationStatusCode.OPTIMAL
    assert abs(opt.objective_value() - 205.0) <= 1e-2
    assert isinstance(opt.variable_primal(x), float)


def test_optimizer_pdhg_and_limits():
    from coolpdlp_hip import PDHG
    opt = moi.Optimizer()
    opt.set_silent(True)
    opt.set_raw("algorithm", PDHG)
    opt.set_raw("termination_reltol", 1e-4)
    _build(opt)
    opt.optimize()
    assert opt.termination_status == moi.TerminationStatusCode.OPTIMAL
    assert abs(opt.objective_value() - 205.0) <= 1e-1
    opt2 = moi.Optimizer()
    opt2.set_silent(True)
    opt2.set_raw("max_kkt_passes", 100)
    opt2.set_raw("termination_reltol", 1e-12)
    _build(opt2)
    opt2.optimize()
    assert opt2.termination_status == moi.TerminationStatusCode.ITERATION_LIMIT
    assert opt2.get_primal_status() == moi.ResultStatusCode.UNKNOWN_RESULT_STATUS
