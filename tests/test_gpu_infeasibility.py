"""GPU tests of the opt-in infeasibility and NaN detection (detect_infeasibility; no
counterpart in the reference): the engine's verdicts equal the oracle's (and so
SciPy-HiGHS's classification) on the infeasible and unbounded LPs of
tests/test_infeasibility.py, in FP64 and mixed precision, PDLP and PDHG, at about the
same pass; feasible LPs keep their trajectory bit for bit; NaN stops the solve; the MOI
front end maps the statuses; the batch engine and both serving routes give the same
verdicts."""
import time

import numpy as np
import pytest

import oracle as O
from coolpdlp_hip import moi

from test_infeasibility import WANT, cases
from test_qps import generated, suite

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cp():
    import coolpdlp_hip
    return coolpdlp_hip


@pytest.mark.parametrize("algo,prec,fmt", [("PDLP", "float64", "csr"), ("PDLP", "float64", "auto"),
                                           ("PDLP", "mixed", "csr"), ("PDHG", "float64", "csr")])
def test_verdicts_match_oracle(cp, algo, prec, fmt):
    A = cp.PDLP if algo == "PDLP" else cp.PDHG
    for name, milp, hs in cases():
        _, st = cp.solve(milp, A(prec, None, fmt, termination_reltol=1e-6, max_kkt_passes=200_000,
                                 time_limit=60.0, detect_infeasibility=True))
        ref = O.solve(generated(milp), O.default_params(algo, termination_reltol=1e-6, max_kkt_passes=200_000,
                                                        time_limit=60.0, detect_infeasibility=1))
        assert ref.status == WANT[hs]
        assert st.termination_status.name == WANT[hs], (name, st.termination_status, st.kkt_passes)
        if prec == "float64":
            tol = max(300, 0.05 * ref.stats.kkt_passes)
            assert abs(st.kkt_passes - ref.stats.kkt_passes) <= tol, (name, st.kkt_passes, ref.stats.kkt_passes)


def test_feasible_trajectories_unchanged(cp):
    for name, milp, _ in suite():
        args = dict(termination_reltol=1e-6, max_kkt_passes=10 ** 6, time_limit=60.0)
        sol0, st0 = cp.solve(milp, cp.PDLP("float64", None, "csr", **args))
        sol1, st1 = cp.solve(milp, cp.PDLP("float64", None, "csr", detect_infeasibility=True, **args))
        assert st0.termination_status == st1.termination_status == cp.OPTIMAL, name
        assert st0.kkt_passes == st1.kkt_passes, name
        assert np.array_equal(sol0.x, sol1.x) and np.array_equal(sol0.y, sol1.y), name


def test_nan_and_front_ends(cp):
    from coolpdlp_hip import generators as G
    milp = cp.MILP.from_generated(G.column_lp(30, 60, 4, seed=2))
    milp.c[5] = np.nan
    _, st = cp.solve(milp, cp.PDLP(detect_infeasibility=True))
    assert st.termination_status == cp.NUMERICAL_ERROR
    # MOI: x1 + x2 <= 1, x1 + x2 >= 3, x >= 0
    for detect, want in [(True, moi.TerminationStatusCode.INFEASIBLE),
                         (False, moi.TerminationStatusCode.ITERATION_LIMIT)]:
        opt = moi.Optimizer()
        opt.set_silent(True)
        opt.set_raw("max_kkt_passes", 5_000)
        if detect:
            opt.set_raw("detect_infeasibility", True)
        x, y = opt.add_variables(2)
        opt.add_constraint(x, moi.GreaterThan(0.0))
        opt.add_constraint(y, moi.GreaterThan(0.0))
        opt.set_objective(moi.ScalarAffineFunction([(1.0, x), (1.0, y)]))
        opt.add_constraint(moi.ScalarAffineFunction([(1.0, x), (1.0, y)]), moi.LessThan(1.0))
        opt.add_constraint(moi.ScalarAffineFunction([(1.0, x), (1.0, y)]), moi.GreaterThan(3.0))
        opt.optimize()
        assert opt.termination_status == want


@pytest.mark.parametrize("prec", ["float64", "mixed"])
def test_batch_and_serving_verdicts(cp, prec):
    """The batch engine's verdicts equal the plain engine's on the same LPs, mixed in one
    batch with feasible ones, which keep their bits (detection only reads the iterates)."""
    from coolpdlp_hip import generators as G
    from coolpdlp_hip.serving import Server
    bad = [(name, milp, WANT[hs]) for name, milp, hs in cases()]
    good = [G.column_lp(m, 2 * m, (4, 8), 40 + m) for m in (80, 200, 350)]
    args = dict(termination_reltol=1e-6, max_kkt_passes=200_000, time_limit=60.0)
    with cp.BatchSolver(cp.PDLP(prec, **args), hbm_bytes=64 << 20) as B:
        ref = B.solve_all(good)
    with cp.BatchSolver(cp.PDLP(prec, detect_infeasibility=True, **args), hbm_bytes=64 << 20) as B:
        res = B.solve_all([milp for _, milp, _ in bad] + good)
    for (name, milp, want), (_, st) in zip(bad, res):
        assert st.termination_status.name == want, (name, prec, st.termination_status, st.kkt_passes)
    for g, (sol, st), (sol0, st0) in zip(good, res[len(bad):], ref):
        assert st.termination_status == st0.termination_status == cp.OPTIMAL
        assert st.kkt_passes == st0.kkt_passes and np.array_equal(sol.x, sol0.x), g.name
    # serving: an infeasible request on each route stops with its verdict
    for limit in (8000, 0):
        srv = Server(reltol=1e-6, precision=prec, batch_bytes=64 << 20, plain_bytes=1 << 30,
                     batch_limit=limit, detect_infeasibility=True)
        milp = bad[3][1]
        g = G.GeneratedLP(milp.A.shape[0], milp.A.shape[1], milp.A.indptr.astype(np.int64),
                          milp.A.indices.astype(np.int32), milp.A.data.copy(), milp.c, milp.lv,
                          milp.uv, milp.lc, milp.uc)
        srv.submit(g)
        srv.drain(time.perf_counter() + 120)
        srv.close()
        r, = srv.done
        assert r.status == bad[3][2] and r.route == ("batch" if limit else "plain"), (r.status, r.route)
