"""Opt-in infeasibility and NaN detection (detect_infeasibility; no counterpart in the
reference, whose TerminationStatus has no infeasible values and whose MOI tests skip
infeasible models, test/moi.jl:31).  At every check that is not OPTIMAL, three candidate
rays -- the iterate minus the restart point, the iterate, the iterate minus the previous
one -- are tested as Farkas certificates (a dual ray: primal infeasible) and as primal
rays (dual infeasible); NaN or infinite KKT errors stop the solve.

CPU tests on the oracle: every verdict agrees with SciPy-HiGHS's classification of the
same LP (status 2 infeasible, 3 unbounded); feasible LPs keep their trajectory bit for
bit (the test reads the iterates, it never changes them); a planted sign error in the
Farkas objective is caught.  The GPU engine is compared with the oracle in
tests/test_gpu_infeasibility.py."""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.optimize import linprog

import oracle as O
from coolpdlp_hip import MILP, generators as G

from test_qps import generated, suite

INF = np.inf


def _lp(A, c, lv, uv, lc, uc):
    return MILP(c=c, lv=lv, uv=uv, A=sp.csc_matrix(np.asarray(A, float)), lc=lc, uc=uc)


def cases():
    """(name, MILP, HiGHS status): small infeasible and unbounded LPs."""
    out = [("tiny_infeasible", _lp([[1, 1], [1, 1]], [1, 1], [0, 0], [INF, INF], [-INF, 3], [1, INF])),
           ("tiny_unbounded", _lp([[1, -1]], [-1, 0], [0, 0], [INF, INF], [-INF], [1])),
           ("box_infeasible", _lp([[1, 1]], [1, 2], [0, 0], [1, 1], [3], [INF]))]
    m = MILP.from_generated(G.column_lp(60, 120, 5, seed=3))
    A = m.A.toarray()
    # a copy of row 0 that contradicts it
    bound = m.lc[0] - 5.0 if np.isfinite(m.lc[0]) else m.uc[0] - 5.0
    out.append(("generated_infeasible", _lp(np.vstack([A, A[0]]), m.c, m.lv, m.uv,
                                            np.append(m.lc, -INF), np.append(m.uc, bound))))
    # a variable of cost -1 that no row holds back
    out.append(("generated_unbounded", _lp(np.hstack([A, np.zeros((A.shape[0], 1))]), np.append(m.c, -1.0),
                                           np.append(m.lv, 0.0), np.append(m.uv, INF), m.lc, m.uc)))
    return [(name, milp, _highs_status(milp)) for name, milp in out]


def _highs_status(milp):
    A = sp.csr_matrix(milp.A)
    eq = milp.lc == milp.uc
    up, lo = ~eq & np.isfinite(milp.uc), ~eq & np.isfinite(milp.lc)
    h = linprog(milp.c, A_ub=sp.vstack([A[up], -A[lo]]), b_ub=np.concatenate([milp.uc[up], -milp.lc[lo]]),
                A_eq=A[eq] if eq.any() else None, b_eq=milp.lc[eq] if eq.any() else None,
                bounds=list(zip(milp.lv, milp.uv)), method="highs")
    return h.status


WANT = {2: "PRIMAL_INFEASIBLE", 3: "DUAL_INFEASIBLE"}


def _solve(milp, algo="PDLP", **kw):
    kw.setdefault("termination_reltol", 1e-6)
    kw.setdefault("max_kkt_passes", 200_000)
    kw.setdefault("time_limit", 60.0)
    return O.solve(generated(milp), O.default_params(algo, **kw))


@pytest.mark.parametrize("algo", ["PDLP", "PDHG"])
def test_verdicts_match_highs(algo):
    for name, milp, hs in cases():
        assert hs in WANT, (name, hs)
        r = _solve(milp, algo, detect_infeasibility=1)
        assert r.status == WANT[hs], (name, algo, r.status, r.stats.kkt_passes)
        # without detection the reference's behaviour: the loop runs to a limit
        if name.startswith("tiny"):
            r0 = _solve(milp, algo, max_kkt_passes=5_000)
            assert r0.status == "ITERATION_LIMIT"


def test_feasible_trajectories_unchanged():
    for name, milp, _ in suite():
        r0 = _solve(milp)
        r1 = _solve(milp, detect_infeasibility=1)
        assert r0.status == r1.status == "OPTIMAL", name
        assert r0.stats.kkt_passes == r1.stats.kkt_passes
        assert np.array_equal(r0.x, r1.x) and np.array_equal(r0.y, r1.y)
        assert r0.history == r1.history


def test_nan_stops_with_numerical_error():
    milp = MILP.from_generated(G.column_lp(30, 60, 4, seed=2))
    milp.c[5] = np.nan
    assert _solve(milp, detect_infeasibility=1).status == "NUMERICAL_ERROR"
    assert _solve(milp, max_kkt_passes=1_000).status == "ITERATION_LIMIT"


def test_each_candidate_alone():
    """PDLP detects through any of the three candidates; PDHG, which never restarts,
    through the difference of consecutive iterates."""
    name, milp, hs = cases()[0]
    for mask in (1, 2, 4):
        assert _solve(milp, detect_infeasibility=1, ray_candidates=mask).status == WANT[hs], mask
    assert _solve(milp, "PDHG", detect_infeasibility=1, ray_candidates=4).status == WANT[hs]
    assert _solve(milp, "PDHG", detect_infeasibility=1, ray_candidates=3).status == "ITERATION_LIMIT"


def _ray(milp, dx, dy, D1=None, D2=None):
    """orc_ray_test on the LP scaled by D1, D2 and the ray given in original units."""
    m, n = milp.A.shape
    D1 = np.ones(m) if D1 is None else D1
    D2 = np.ones(n) if D2 is None else D2
    As = sp.csc_matrix(sp.diags(D1) @ milp.A @ sp.diags(D2))
    return O.ray_test(m, n, As.indptr, As.indices, As.data, D2 * milp.c, milp.lv / D2, milp.uv / D2,
                      D1 * milp.lc, D1 * milp.uc, D1, D2, np.asarray(dx, float) / D2,
                      np.asarray(dy, float) / D1)


def test_hand_worked_certificates_and_planted_sign_errors():
    """x1 + x2 <= 1 and x1 + x2 >= 3 (x >= 0): e = (-1, 1) is a Farkas ray (signs: e1 <= 0 on
    the <= row, e2 >= 0 on the >= row; r = -A'e = 0; objective 3 - 1 = 2 > 0).  The
    flipped ray has the wrong signs and fails; so does a ray with the objective's sign
    planted wrong (e = (-1, 0.2): objective 0.6 - 1 < 0).  min -x1 s.t. x1 - x2 <= 1
    (x >= 0): d = (1, 1) is a primal ray (A d = 0, d >= 0, c'd = -1); (1, 0) violates
    the row's cone (A d = 1 > 0).  The verdicts hold under any positive scaling."""
    inf_lp = cases()[0][1]
    unb_lp = cases()[1][1]
    rng = np.random.default_rng(0)
    for D in (None, "random"):
        k1 = {} if D is None else dict(D1=rng.uniform(0.1, 10, 2), D2=rng.uniform(0.1, 10, 2))
        k2 = {} if D is None else dict(D1=rng.uniform(0.1, 10, 1), D2=rng.uniform(0.1, 10, 2))
        assert _ray(inf_lp, [0, 0], [-1, 1], **k1) == "PRIMAL_INFEASIBLE"
        assert _ray(inf_lp, [0, 0], [1, -1], **k1) == "STILL_RUNNING"
        assert _ray(inf_lp, [0, 0], [-1, 0.2], **k1) == "STILL_RUNNING"
        assert _ray(unb_lp, [1, 1], [0], **k2) == "DUAL_INFEASIBLE"
        assert _ray(unb_lp, [1, 0], [0], **k2) == "STILL_RUNNING"
        assert _ray(unb_lp, [-1, -1], [0], **k2) == "STILL_RUNNING"
