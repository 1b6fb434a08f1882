"""Pins the oracle's restart machinery: should_restart and the candidate choice
(src/components/restart.jl:40-57, src/algorithms/pdlp.jl:138-170), the averaging
(pdlp.jl:126-136), the primal-weight update (src/components/step_size.jl:58-75) and the
restart itself (pdlp.jl:172-196).  CPU only.

The decision rule is pinned on hand-built KKT errors whose answers are worked out from
the reference's lines (boundaries of <= and >, ties, the weighting of absolute()).  The
solve-level parts are pinned against iterates computed independently: the oracle run
one pass at a time with restarts disabled, whose trajectory does not depend on where the
checks fall.  Each test is chosen so that a plausible slip (< for <=, the previous check's
candidate in place of the one one step earlier, swapped averaging weights, a wrong
damping exponent) fails it.
"""
import math

import numpy as np
import pytest

import oracle as O
from coolpdlp_hip import generators as G

K = O.kkt


def _gap(v):
    """A KKT error whose absolute() is v for every omega (only the gap term)."""
    return K(gap=v)


def _decide(sol, avg, last, avg_last, restart, omega=1.0, inner=1, total=100, **kw):
    return O.restart_decision(sol, avg, last, avg_last, restart, omega, inner, total,
                              O.default_params(**kw) if kw else None)


# ------------------------------------------------------------ should_restart
# (restart?, candidate is the average?) for hand-built errors; omega = 1 unless given,
# inner = 1 of total = 100 so that long_inner_loop (inner >= 0.36 total) is off.
CASES = [
    # sufficient decay, boundary: 0.2 <= 0.2 * 1.0 (a strict < would not restart)
    ("sufficient at its boundary", (_gap(0.2), _gap(0.9), _gap(0.5), _gap(0.5), _gap(1.0)), {}, (True, False)),
    ("just above sufficient, local progress", (_gap(0.2000001), _gap(0.9), _gap(0.5), _gap(0.5), _gap(1.0)), {},
     (False, False)),
    # necessary decay and no local progress: 0.5 <= 0.8 and 0.5 > min(0.4, 0.45)
    ("necessary + no local progress", (_gap(0.5), _gap(0.9), _gap(0.4), _gap(0.45), _gap(1.0)), {}, (True, False)),
    # no_local_progress is strict: 0.5 > 0.5 is false
    ("no local progress is strict", (_gap(0.5), _gap(0.9), _gap(0.5), _gap(0.6), _gap(1.0)), {}, (False, False)),
    # necessary boundary: 0.8 <= 0.8 * 1.0
    ("necessary at its boundary", (_gap(0.8), _gap(0.9), _gap(0.7), _gap(0.75), _gap(1.0)), {}, (True, False)),
    ("just above necessary", (_gap(0.8000001), _gap(0.9), _gap(0.7), _gap(0.75), _gap(1.0)), {}, (False, False)),
    # the last candidate is the better of (last, avg_last): either one can carry the test
    ("last candidate = sol_last", (_gap(0.5), _gap(0.9), _gap(0.3), _gap(0.6), _gap(1.0)), {}, (True, False)),
    ("last candidate = avg_last", (_gap(0.5), _gap(0.9), _gap(0.6), _gap(0.3), _gap(1.0)), {}, (True, False)),
    # the candidate is the average when it is better, and on ties
    ("candidate = average", (_gap(0.9), _gap(0.5), _gap(0.4), _gap(0.45), _gap(1.0)), {}, (True, True)),
    ("tie goes to the average", (_gap(0.5), _gap(0.5), _gap(0.5), _gap(0.6), _gap(1.0)), {}, (False, True)),
    ("average judged, not the iterate", (_gap(0.95), _gap(0.9), _gap(0.5), _gap(0.5), _gap(1.0)), {},
     (False, True)),
    # long inner loop: 36 >= 0.36 * 100 restarts without any decay; 35 does not
    ("long inner loop at its boundary", (_gap(0.9), _gap(0.95), _gap(0.1), _gap(0.1), _gap(1.0)),
     {"inner": 36, "total": 100}, (True, False)),
    ("inner loop one short", (_gap(0.9), _gap(0.95), _gap(0.1), _gap(0.1), _gap(1.0)),
     {"inner": 35, "total": 100}, (False, False)),
    # F9: the first check has inner == total
    ("first check", (_gap(0.99), _gap(0.995), _gap(0.1), _gap(0.1), _gap(1.0)), {"inner": 100, "total": 100},
     (True, False)),
    # absolute(err, w) = sqrt(w^2 primal^2 + w^-2 dual^2 + gap^2): with w = 2 the candidate
    # (primal 0.3) weighs 0.6 and the restart point (dual 2.0) weighs 1.0, so only
    # "necessary" holds; swapping the weights would make it 0.15 vs 4.0, a sufficient decay
    ("weights: primal * w", (K(primal=0.3), K(primal=0.4), _gap(0.7), _gap(0.7), K(dual=2.0)), {"omega": 2.0},
     (False, False)),
    ("weights: dual / w", (K(dual=0.3), K(dual=0.4), _gap(0.7), _gap(0.7), K(primal=2.0)), {"omega": 0.5},
     (False, False)),
    # the gap is not weighted: 0.2 against w * 0.5 = 1.0 is a sufficient decay
    ("gap unweighted", (_gap(0.2), _gap(0.9), _gap(0.5), _gap(0.5), K(primal=0.5)), {"omega": 2.0}, (True, False)),
]


@pytest.mark.parametrize("name,errs,kw,expected", CASES, ids=[c[0] for c in CASES])
def test_should_restart_known_answers(name, errs, kw, expected):
    assert _decide(*errs, **kw) == expected


def test_restart_parameters_are_used():
    errs = (_gap(0.3), _gap(0.9), _gap(0.5), _gap(0.5), _gap(1.0))
    assert _decide(*errs) == (False, False)                             # 0.3 > 0.2
    assert _decide(*errs, sufficient_decay=0.3) == (True, False)        # 0.3 <= 0.3
    errs = (_gap(0.5), _gap(0.9), _gap(0.4), _gap(0.4), _gap(1.0))
    assert _decide(*errs) == (True, False)
    assert _decide(*errs, necessary_decay=0.45) == (False, False)       # 0.5 > 0.45
    errs = (_gap(0.9), _gap(0.95), _gap(0.1), _gap(0.1), _gap(1.0))
    assert _decide(*errs, inner=20, total=100, artificial_decay=0.2) == (True, False)
    assert _decide(*errs, inner=20, total=100, artificial_decay=0.21) == (False, False)


# ------------------------------------------------------------ primal weight update
def test_primal_weight_update_power_form():
    # exp(t log(dy/dx) + (1 - t) log w) = (dy/dx)^t * w^(1-t)
    assert math.isclose(O.primal_weight_update(4.0, 1.0, 16.0, 0.5), 8.0, rel_tol=1e-14)
    assert math.isclose(O.primal_weight_update(2.0, 2.0, 8.0, 0.3), 2.0 ** 1.3, rel_tol=1e-14)
    assert math.isclose(O.primal_weight_update(3.0, 5.0, 2.0, 1.0), 0.4, rel_tol=1e-14)
    assert math.isclose(O.primal_weight_update(3.0, 5.0, 2.0, 0.0), 3.0, rel_tol=1e-14)
    # distances at or below zero_tol keep w (strict >)
    assert O.primal_weight_update(3.0, 1e-8, 2.0, 0.5, 1e-8) == 3.0
    assert O.primal_weight_update(3.0, 2.0, 1e-8, 0.5, 1e-8) == 3.0
    assert O.primal_weight_update(3.0, 2e-8, 2.0, 0.5, 1e-8) != 3.0


# ------------------------------------------------------------ solve-level pins
def _lp():
    return G.column_lp(30, 70, 4, seed=21, box_frac=0.3)


def _no_restart(**kw):
    return O.default_params(sufficient_decay=0.0, necessary_decay=0.0, artificial_decay=math.inf,
                            termination_reltol=0.0, **kw)


def _iterates(g, count):
    """Unscaled iterates after 1..count passes with restarts disabled, each from its own
    run of the oracle (one pass per check, stopped by the pass limit)."""
    xs, ys = [], []
    for k in range(1, count + 1):
        r = O.solve(g, _no_restart(check_every=1, max_kkt_passes=k))
        assert r.stats.kkt_passes == k and r.stats.restarts == 0 and r.status == "ITERATION_LIMIT"
        xs.append(r.x)
        ys.append(r.y)
    return np.array(xs), np.array(ys)


class _Scaled:
    """The preconditioned LP of the oracle (Ruiz 10 + Chambolle-Pock) and the KKT
    errors of an unscaled point on it."""

    def __init__(self, g):
        self.g = g
        self.D1, self.D2, a, at = O.scaling(g.m, g.n, g.colptr, g.rowval, g.nzval, 10, 1.0)
        self.A = O.csc_to_csr(g.m, g.n, g.colptr, g.rowval, a)
        self.At = (g.colptr, g.rowval, at)

    def kkt(self, x, y):
        g, D1, D2 = self.g, self.D1, self.D2
        return O.kkt_errors(g.m, g.n, self.A, self.At, D2 * g.c, g.lv / D2, g.uv / D2, D1 * g.lc,
                            D1 * g.uc, D1, D2, x / D2, y / D1)


def _close(e, f, rtol=1e-10):
    return all(math.isclose(getattr(e, k), getattr(f, k), rel_tol=rtol, abs_tol=1e-13)
               for k in ("primal", "primal_scale", "dual", "dual_scale", "gap", "gap_scale"))


def test_check_errors_are_the_reference_iterates():
    """At a check after t passes: err(sol) is the iterate after t passes, err(avg) the mean
    of iterates 1..t, err(last) and err(avg_last) the same one PDHG step earlier (t - 1),
    not at the previous check (F10), and err(restart) the start point."""
    g, ce, checks = _lp(), 3, 4
    tr = O.solve_traced(g, _no_restart(check_every=ce, max_kkt_passes=ce * (checks + 1)))
    assert len(tr.passes) == checks + 1 and not tr.restarted.any()
    xs, ys = _iterates(g, ce * checks)
    S = _Scaled(g)
    e0 = S.kkt(np.zeros(g.n), np.zeros(g.m))
    for i in range(checks):
        t = ce * (i + 1)
        assert tr.passes[i] == t
        e = tr.err[i]
        assert _close(e["sol"], S.kkt(xs[t - 1], ys[t - 1]))
        assert _close(e["avg"], S.kkt(xs[:t].mean(0), ys[:t].mean(0)))
        assert _close(e["last"], S.kkt(xs[t - 2], ys[t - 2]))
        assert _close(e["avg_last"], S.kkt(xs[:t - 1].mean(0), ys[:t - 1].mean(0)))
        assert _close(e["restart"], e0)
        if i > 0:   # the previous check's iterate is a different point
            prev = tr.err[i - 1]["sol"]
            assert not _close(e["last"], prev, rtol=1e-6)


def test_average_is_the_mean_since_start_without_restarts():
    """With eta fixed and no restart, update_average! gives the arithmetic mean of the
    iterates after passes 1..t (the start point has weight 0 once eta_sum = 0)."""
    g, t = _lp(), 17
    tr = O.solve_traced(g, _no_restart(check_every=t, max_kkt_passes=t))
    xs, ys = _iterates(g, t)
    np.testing.assert_allclose(tr.x_avg, xs.mean(0), rtol=1e-11, atol=1e-13)
    np.testing.assert_allclose(tr.y_avg, ys.mean(0), rtol=1e-11, atol=1e-13)
    # the weights favour no iterate: the first and the last ones move the mean equally
    assert not np.allclose(tr.x_avg, xs[-1]) and not np.allclose(tr.x_avg, xs[0])


@pytest.mark.parametrize("damping", [0.5, 0.3, 1.0])
def test_first_restart_primal_weight(damping):
    """F9: the first check restarts.  The new weight is (dy/dx)^t * w0^(1-t), with dx, dy
    the scaled distances of the candidate from the start point (0 here), and w0 =
    ||c~|| / ||combine(lc~, uc~)||."""
    g, ce = _lp(), 5
    tr = O.solve_traced(g, O.default_params(check_every=ce, max_kkt_passes=2 * ce, termination_reltol=0.0,
                                            primal_weight_damping=damping))
    assert tr.restarted[0] and tr.inner[0] == tr.total[0] == ce
    S = _Scaled(g)
    comb = np.maximum.reduce([np.zeros(g.m), np.where(np.isfinite(g.lc), np.abs(S.D1 * g.lc), 0.0),
                              np.where(np.isfinite(g.uc), np.abs(S.D1 * g.uc), 0.0)])
    w0 = np.linalg.norm(S.D2 * g.c) / np.linalg.norm(comb)
    assert math.isclose(tr.omega[0], w0, rel_tol=1e-12)
    xs, ys = _iterates(g, ce)
    if tr.from_avg[0]:
        xc, yc = xs.mean(0), ys.mean(0)
    else:
        xc, yc = xs[-1], ys[-1]
    dx, dy = np.linalg.norm(xc / S.D2), np.linalg.norm(yc / S.D1)
    expected = (dy / dx) ** damping * w0 ** (1.0 - damping)
    assert math.isclose(tr.omega_after[0], expected, rel_tol=1e-10)
    assert tr.omega[1] == tr.omega_after[0]


def test_restart_moves_the_restart_point_and_counters():
    """After a restart, the restart point is the candidate (its errors, bit for bit, at the
    next check), inner restarts from 0 and total keeps counting."""
    g, ce = G.column_lp(60, 150, 5, seed=22), 10
    tr = O.solve_traced(g, O.default_params(check_every=ce, termination_reltol=1e-9, max_kkt_passes=3000))
    n = len(tr.passes) - 1          # the last check terminated and has no restart record
    assert tr.restarted[:n].sum() >= 3 and (~tr.restarted[:n]).sum() >= 3
    inner = 0
    for i in range(n):
        inner += ce
        assert tr.total[i] == ce * (i + 1) and tr.inner[i] == inner
        nxt = tr.err[i + 1]["restart"] if i + 1 < n else None
        if tr.restarted[i]:
            inner = 0
            cand = tr.err[i]["avg" if tr.from_avg[i] else "sol"]
            if nxt is not None:
                assert all(getattr(nxt, f) == getattr(cand, f) for f, _ in O.KKT._fields_)
        else:
            assert tr.omega_after[i] == tr.omega[i]
            if nxt is not None:
                assert all(getattr(nxt, f) == getattr(tr.err[i]["restart"], f) for f, _ in O.KKT._fields_)
        # the recorded decision is the rule applied to the recorded errors
        e = tr.err[i]
        assert O.restart_decision(e["sol"], e["avg"], e["last"], e["avg_last"], e["restart"], tr.omega[i],
                                  tr.inner[i], tr.total[i]) == (bool(tr.restarted[i]), bool(tr.from_avg[i]))


def _iterates_from(g, x0, y0, count):
    """Unscaled iterates after 1..count passes from the start (x0, y0), restarts disabled."""
    xs, ys = [], []
    for k in range(1, count + 1):
        r = O.solve(g, _no_restart(check_every=1, max_kkt_passes=k), x0=x0, y0=y0)
        assert r.stats.kkt_passes == k and r.stats.restarts == 0
        xs.append(r.x)
        ys.append(r.y)
    return np.array(xs), np.array(ys)


@pytest.mark.parametrize("seed", [21, 23, 24])
def test_after_a_restart_the_loop_starts_over_from_the_candidate(seed):
    """restart! (pdlp.jl:172-196): sol <- candidate, avg <- 0 with eta_sum <- 0, restart
    point <- sol.  With damping 0 the primal weight never changes, so the state after
    the first restart (F9) is a fresh solve warm-started at the candidate: the next
    check's iterate, average, last iterate and restart point are those of that solve."""
    g, ce = G.column_lp(30, 70, 4, seed=seed, box_frac=0.3), 4
    tr = O.solve_traced(g, O.default_params(check_every=ce, max_kkt_passes=3 * ce, termination_reltol=0.0,
                                            primal_weight_damping=0.0))
    assert tr.restarted[0] and math.isclose(tr.omega_after[0], tr.omega[0], rel_tol=1e-14)   # exp(log w)
    xs, ys = _iterates(g, ce)
    xc, yc = (xs.mean(0), ys.mean(0)) if tr.from_avg[0] else (xs[-1], ys[-1])
    zx, zy = _iterates_from(g, xc, yc, ce)
    S = _Scaled(g)
    e = tr.err[1]
    assert _close(e["restart"], S.kkt(xc, yc), rtol=1e-9)
    assert _close(e["sol"], S.kkt(zx[-1], zy[-1]), rtol=1e-9)
    assert _close(e["last"], S.kkt(zx[-2], zy[-2]), rtol=1e-9)
    assert _close(e["avg"], S.kkt(zx.mean(0), zy.mean(0)), rtol=1e-9)
    assert _close(e["avg_last"], S.kkt(zx[:-1].mean(0), zy[:-1].mean(0)), rtol=1e-9)


def test_restart_from_average_is_exercised():
    """The warm-start test above covers a restart from the average for one of its seeds."""
    ce = 4
    kinds = set()
    for seed in (21, 23, 24):
        g = G.column_lp(30, 70, 4, seed=seed, box_frac=0.3)
        tr = O.solve_traced(g, O.default_params(check_every=ce, max_kkt_passes=3 * ce, termination_reltol=0.0,
                                                primal_weight_damping=0.0))
        kinds.add(bool(tr.from_avg[0]))
    assert kinds == {True, False}
