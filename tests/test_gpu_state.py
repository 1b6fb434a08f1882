"""GPU tests of the loop state (cpdlp_get_state, cpdlp_resume, cpdlp_batch_take_state):
a run stopped after some checks and resumed on a fresh engine of the same LP gives the
same bits as one that never stopped (FP64 and FP32, PDLP and PDHG); a batch instance
taken out of its batch mid-run continues on a plain engine to the LP's optimum while the
instances left in the batch keep their bits; misuse is refused."""
import numpy as np
import pytest

from coolpdlp_hip import generators as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cp():
    import coolpdlp_hip
    return coolpdlp_hip


def _run_to_end(eng, cp):
    while eng.run_checks(50) == cp.STILL_RUNNING:
        pass
    return eng.solution(), eng.stats()


@pytest.mark.parametrize("algo,precision,fmt,stop", [("PDLP", "float64", "csr", 3), ("PDLP", "float64", "auto", 11),
                                                     ("PDLP", "float32", "csr", 2), ("PDHG", "float64", "csr", 5)])
def test_resume_is_bit_exact(cp, algo, precision, fmt, stop):
    g = G.column_lp(400, 1000, 6, seed=31, box_frac=0.3)
    milp = cp.MILP.from_generated(g)
    make = cp.PDLP if algo == "PDLP" else cp.PDHG
    params = dict(termination_reltol=1e-7 if precision == "float64" else 1e-5, powm_tol=1e-6,
                  max_kkt_passes=30_000)
    with cp.Engine(milp, make(precision, None, fmt, **params)) as a:
        a.initialize()
        assert a.run_checks(stop) == cp.STILL_RUNNING
        state = a.get_state()
        assert state.kkt_passes == stop * 100 and state.x.shape == (5, g.n) and state.y.shape == (5, g.m)
        sol_a, st_a = _run_to_end(a, cp)
    with cp.Engine(milp, make(precision, None, fmt, **params)) as b:
        b.resume(state)
        mid = b.get_state()   # nothing run yet: the state as given
        for f in ("x", "y", "D2", "D1"):
            assert np.array_equal(getattr(mid, f), getattr(state, f)), f
        assert (mid.eta, mid.omega, mid.eta_sum, mid.kkt_passes, mid.restarts) == \
            (state.eta, state.omega, state.eta_sum, state.kkt_passes, state.restarts)
        sol_b, st_b = _run_to_end(b, cp)
    assert st_b.termination_status == st_a.termination_status == cp.OPTIMAL
    assert st_b.kkt_passes == st_a.kkt_passes and st_b.restarts == st_a.restarts
    assert np.array_equal(sol_a.x, sol_b.x) and np.array_equal(sol_a.y, sol_b.y)
    tail = [(p, e) for p, e in st_a.error_history if p > state.kkt_passes]
    assert len(tail) == len(st_b.error_history)
    for (p1, e1), (p2, e2) in zip(tail, st_b.error_history):
        assert p1 == p2 and e1 == e2, (p1, e1, e2)


@pytest.mark.parametrize("precision", ["float64", "mixed"])
def test_batch_instance_continues_on_a_plain_engine(cp, precision):
    """Take the largest instance out of a batch after two rounds (in mixed precision it is
    still in its FP32 phase), resume it on an FP64 plain engine: it reaches the optimum;
    the instances left in the batch keep the bits they have when solved without it."""
    lps = [G.column_lp(m, int(2.2 * m), (6, 12), 500 + m) for m in (120, 700, 1900)]
    algo = cp.PDLP(precision, termination_reltol=1e-6)
    with cp.BatchSolver(algo, hbm_bytes=256 << 20) as B:
        alone = B.solve_all(lps[:2])
    with cp.BatchSolver(algo, hbm_bytes=256 << 20) as B:
        ids = [B.add(g) for g in lps]
        for _ in range(200):   # until the instance is set up and has run 3 checks
            assert B.run(1) > 0
            try:
                if B.get(ids[2])[1].kkt_passes >= 300:
                    break
            except cp.HIPError:   # not set up yet
                pass
        state = B.take_state(ids[2])
        assert 300 <= state.kkt_passes < 10_000 and state.x.shape == (5, lps[2].n)
        assert state.fp32_passes == (state.kkt_passes if precision == "mixed" else 0)
        with pytest.raises(cp.ArgumentError):
            B.take_state(ids[2])   # no longer in the batch (unknown id)
        while B.run(20) > 0:
            pass
        rest = [B.get(i) for i in ids[:2]]
    for (s1, t1), (s2, t2) in zip(alone, rest):
        assert np.array_equal(s1.x, s2.x) and np.array_equal(s1.y, s2.y) and t1.kkt_passes == t2.kkt_passes
    g = lps[2]
    with cp.Engine(cp.MILP.from_generated(g), cp.PDLP("float64", None, "csr", termination_reltol=1e-6)) as e:
        e.resume(state)
        sol, st = _run_to_end(e, cp)
    assert st.termination_status == cp.OPTIMAL and st.kkt_passes > state.kkt_passes
    assert abs(float(sol.x @ g.c) - g.objective) <= 2e-5 * (1 + abs(g.objective))


def test_state_misuse_is_refused(cp):
    g = G.column_lp(200, 500, 6, seed=7)
    milp = cp.MILP.from_generated(g)
    with cp.Engine(milp, cp.PDLP("float64", termination_reltol=1e-6)) as e:
        with pytest.raises(cp.HIPError, match="no state"):
            e.get_state()                        # nothing run yet
        e.initialize()
        e.run_checks(1)
        st = e.get_state()
        with pytest.raises(cp.DimensionMismatch):
            e.resume(cp.LoopState(st.x[:, :-1], st.y, st.D2, st.D1, st.eta, st.omega, st.eta_sum,
                                  st.kkt_passes, st.inner, st.total, st.restarts, st.fp32_passes))
        bad = cp.LoopState(st.x, st.y, st.D2, st.D1, -1.0, st.omega, st.eta_sum, st.kkt_passes, st.inner,
                           st.total, st.restarts, st.fp32_passes)
        with pytest.raises(cp.ArgumentError, match="eta"):
            e.resume(bad)
    with cp.Engine(milp, cp.PDLP("mixed", termination_reltol=1e-6)) as e:
        with pytest.raises(cp.ArgumentError, match="mixed"):
            e.resume(st)
    with cp.BatchSolver(cp.PDLP("float64", termination_reltol=1e-6), hbm_bytes=64 << 20) as B:
        i = B.add(g)
        with pytest.raises(cp.HIPError, match="not set up"):
            B.take_state(i)                      # admitted, not run yet
        while B.run(20) > 0:
            pass
        with pytest.raises(cp.HIPError, match="finished"):
            B.take_state(i)


def test_batch_tail_continuation(cp):
    """BatchSolver(tail=...): once at most `tail` LPs are unfinished, those past
    tail_min_passes continue on plain FP64 engines; every LP reaches its optimum, the
    continued ones are reported by poll and get like the rest, and the LPs that finish in
    the batch keep their batch-alone bits."""
    lps = [G.column_lp(m, int(2.2 * m), (6, 12), 900 + m) for m in (60, 90, 150, 400, 1200, 2500)]
    algo = cp.PDLP("mixed", termination_reltol=1e-6)
    with cp.BatchSolver(algo, hbm_bytes=256 << 20) as B:
        alone = B.solve_all(lps)
    with cp.BatchSolver(algo, hbm_bytes=256 << 20, tail=3, tail_min_passes=300, tail_workers=2) as B:
        ids = [B.add(g) for g in lps]
        seen = []
        for _ in range(100_000):
            running = B.run(5)
            seen += B.poll()
            if running == 0:
                break
        seen += B.poll()
        assert sorted(seen) == sorted(ids)
        res = [B.get(i) for i in ids]
        cont = [B.continued(i) for i in ids]
        for i in ids:
            B.release(i)
    assert any(cont)
    for g, (sol, st), (sa, ta), c in zip(lps, res, alone, cont):
        assert st.termination_status == cp.OPTIMAL, (g.name, c, st)
        assert abs(float(sol.x @ g.c) - g.objective) <= 2e-5 * (1 + abs(g.objective)), (g.name, c)
        if c:
            assert st.kkt_passes >= 300
        else:
            assert np.array_equal(sol.x, sa.x) and np.array_equal(sol.y, sa.y) and st.kkt_passes == ta.kkt_passes
