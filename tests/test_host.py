"""Host-side logic that needs no GPU: the FCFS queue baseline, Poisson arrivals, the
batch arena footprint, generators' shapes."""
import numpy as np

from coolpdlp_hip import generators as G


def test_fcfs_queue():
    from coolpdlp_hip.serving import fcfs_latencies
    lat = fcfs_latencies(np.array([0.0, 0.5, 0.6, 5.0]), np.array([1.0, 1.0, 1.0, 0.5]))
    assert np.allclose(lat, [1.0, 1.5, 2.4, 0.5])


def test_poisson_arrivals_rate():
    from coolpdlp_hip.serving import poisson_arrivals
    a = poisson_arrivals(1000.0, 20.0, seed=1)
    assert abs(len(a) / 20.0 - 1000.0) < 50 and np.all(np.diff(a) > 0) and a[-1] < 20.0


def test_batch_footprint_grows_with_size():
    from coolpdlp_hip.batch import footprint
    a, b = footprint(100, 200, 1000), footprint(200, 400, 2000)
    assert 0 < a < b and a % 256 == 0


def test_batch_specs_shapes():
    specs = G.batch_specs(4096)
    ms = np.array([s[0] for s in specs])
    ns = np.array([s[1] for s in specs])
    assert len(specs) == 4096 and ms.min() >= 50 and ms.max() <= 5000
    r = ns / ms
    assert r.min() >= 1.5 - 1e-3 and r.max() <= 3.0 + 1e-3
    assert len({s[2] for s in specs}) == 4096   # per-instance seeds


def test_min_cost_flow_small_is_balanced_and_feasible():
    from scipy.optimize import linprog
    g = G.min_cost_flow_lp(nodes=400, out_degree=8, seed=3)
    assert abs(g.lc.sum()) < 1e-8 and np.array_equal(g.lc, g.uc)
    A = g.csc()
    assert np.allclose(np.asarray(A.sum(axis=0)).ravel(), 0)   # +1 / -1 per arc
    h = linprog(g.c, A_eq=A, b_eq=g.lc, bounds=list(zip(g.lv, g.uv)), method="highs")
    assert h.status == 0


def test_serving_routes_by_size():
    from collections import deque
    from coolpdlp_hip.serving import Lane, route
    lanes = [Lane(1_000_000, 2), Lane(float("inf"), 1)]
    assert route(1000, 2000, 16_000, 8192, lanes) == -1            # small: batch engine
    assert route(3000, 6000, 48_000, 8192, lanes) == 0             # medium: first lane
    assert route(200_000, 400_000, 3_200_000, 8192, lanes) == 1    # large: last lane
    # overflow: only when asked for, only when the lane is backed up, only if it fits
    lanes[0].waiting = deque(range(4))
    assert route(3000, 6000, 48_000, 8192, lanes) == 0
    assert route(3000, 6000, 48_000, 8192, lanes, overflow_queue=2) == -1
    assert route(30_000, 60_000, 480_000, 8192, lanes, overflow_queue=2) == 0


def test_preprocessed_milp_is_refused_before_any_device_call():
    """A MILP that already carries D1, D2 (the result of preprocess) is refused with
    ArgumentError by every route: the reference would compose its scaling with the new
    one (src/components/preconditioning.jl:73) and measure errors through both
    (errors.jl:77-84), which the engine does not do.  The check runs on the host."""
    import pytest
    import scipy.sparse as sp
    import coolpdlp_hip as cp
    from coolpdlp_hip.problems import require_unpreconditioned
    A = sp.csc_matrix(np.array([[1.0, 2.0], [0.0, 1.0]]))
    kw = dict(c=[1.0, 1.0], lv=[0.0, 0.0], uv=[1.0, 1.0], A=A, lc=[0.0, 0.0], uc=[1.0, 1.0])
    require_unpreconditioned(cp.MILP(**kw))
    require_unpreconditioned(G.column_lp(20, 40, 3, seed=1))
    for d in (dict(D1=[1.0, 2.0]), dict(D2=[0.5, 1.0])):
        milp = cp.MILP(**kw, **d)
        with pytest.raises(cp.ArgumentError):
            cp.Engine(milp, cp.PDLP())
        with pytest.raises(cp.ArgumentError):
            cp.solve(milp, cp.PDLP())
        with pytest.raises(cp.ArgumentError):
            cp.preprocess(milp, None, cp.PDLP())
