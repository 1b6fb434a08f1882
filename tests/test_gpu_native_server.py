"""GPU tests of the library's own server (cpdlp_server_*, through coolpdlp_hip.serving.
NativeServer): requests submitted while others run give the same bits as when solved
alone -- batch-routed ones (either batch route) as the batch engine alone, plain-routed
ones as a CSR-format plain solve alone; a start point is honoured on every route; the
opt-in infeasibility detection reaches every route; errors are reported per request and
per call."""
import time

import numpy as np
import pytest

from coolpdlp_hip import generators as G

from test_infeasibility import WANT, cases

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cp():
    import coolpdlp_hip
    return coolpdlp_hip


def _drain(srv, want, deadline_s=180):
    """Pump until `want` requests have finished; returns {id: (sol, st, seconds, route)}."""
    out = {}
    t_end = time.perf_counter() + deadline_s
    while len(out) < want and time.perf_counter() < t_end:
        srv.pump()
        for i in srv.poll():
            out[i] = srv.get(i)
        time.sleep(0.0005)
    assert len(out) == want, f"{len(out)} of {want} finished"
    return out


@pytest.mark.parametrize("borrow,driver,batch2", [(True, False, 0), (False, False, 12000), (True, True, 0),
                                                  (True, True, 12000)])
def test_native_server_alone_equals_in_mix(cp, borrow, driver, batch2):
    """Copied (cpdlp_server_submit) and borrowed (cpdlp_server_submit_borrowed) requests,
    pumped by the caller or by the library's own threads (cpdlp_server_start), with the
    second batch route off or taking 8000 < m + n <= 12000."""
    from coolpdlp_hip.serving import NativeServer
    small = [G.request_lp(m, 2 * m, 700 + m) for m in (60, 150, 420, 900, 1500, 2600)]
    big = [G.request_lp(3000, 6000, 78), G.request_lp(5000, 10000, 79), G.request_lp(3500, 7000, 80)]
    algo = cp.PDLP("mixed", termination_reltol=1e-6)
    with NativeServer(algo, batch_bytes=256 << 20, plain_bytes=4 << 30, batch_limit=8000,
                      checks_per_round=2, batch2_limit=batch2, batch2_bytes=256 << 20,
                      batch2_checks=3) as srv:
        if driver:
            srv.start()
        ids = {srv.submit(small[0], borrow=borrow): small[0], srv.submit(big[0], borrow=borrow): big[0]}
        for g in small[1:] + big[1:]:
            for _ in range(3):
                srv.pump()
            ids[srv.submit(g, borrow=borrow)] = g   # arrives while the others are on the GPU
        res = _drain(srv, len(ids))
    for i, g in ids.items():
        sol, st, secs, route = res[i]
        assert st.termination_status == cp.OPTIMAL, (g.name, st)
        assert secs > 0
        assert abs(float(sol.x @ g.c) - g.objective) <= 2e-5 * (1 + abs(g.objective))
        if g.m + g.n <= 8000 or g.m + g.n <= batch2:
            assert route == (-1 if g.m + g.n <= 8000 else -2)
            with cp.BatchSolver(algo, hbm_bytes=64 << 20) as B:
                (alone, st_a), = B.solve_all([g])
        else:
            assert route == 0                # <= 1e6 nonzeros: the first lane
            alone, st_a = cp.solve(cp.MILP.from_generated(g), cp.PDLP("mixed", None, "csr",
                                                                      termination_reltol=1e-6))
        assert np.array_equal(alone.x, sol.x) and np.array_equal(alone.y, sol.y), g.name
        assert st_a.kkt_passes == st.kkt_passes


@pytest.mark.parametrize("batch_limit,batch2,want", [(8000, 0, -1), (0, 0, 0), (0, 8000, -2)])
def test_native_server_warm_start_and_detection(cp, batch_limit, batch2, want):
    from coolpdlp_hip.serving import NativeServer
    g = G.request_lp(400, 800, 91)
    algo = cp.PDLP("mixed", termination_reltol=1e-6, detect_infeasibility=True)
    with NativeServer(algo, batch_bytes=64 << 20, plain_bytes=1 << 30, batch_limit=batch_limit,
                      batch2_limit=batch2, batch2_bytes=64 << 20) as srv:
        first = srv.submit(g)
        (sol, st, _, route), = _drain(srv, 1).values()
        assert st.termination_status == cp.OPTIMAL and route == want
        srv.submit(g, sol)
        (_, st2, _, _), = _drain(srv, 1).values()
        assert st2.termination_status == cp.OPTIMAL and st2.kkt_passes <= st.kkt_passes // 2
        bad = {srv.submit(milp): (name, WANT[hs]) for name, milp, hs in cases()}
        res = _drain(srv, len(bad))
        for i, (name, want) in bad.items():
            assert res[i][1].termination_status.name == want, (name, res[i][1].termination_status)
        with pytest.raises(cp.ArgumentError):
            srv.get(first)                   # already handed over and freed


def test_native_server_reports_request_errors(cp):
    """Errors are per call: a get of an id the library does not know and a start point of
    the wrong size are refused, and the server keeps serving."""
    from coolpdlp_hip.serving import NativeServer
    with NativeServer(cp.PDLP("mixed", termination_reltol=1e-6, max_kkt_passes=2000),
                      batch_bytes=16 << 20, plain_bytes=1 << 30, batch2_bytes=16 << 20) as srv:
        g = G.request_lp(100, 200, 5)
        i = srv.submit(g)
        srv.shapes[10**9] = (1, 1)
        with pytest.raises(cp.ArgumentError):
            srv.get(10**9)
        (_, st, _, _), = _drain(srv, 1).values()
        assert st.termination_status == cp.OPTIMAL
        with pytest.raises(cp.DimensionMismatch):
            srv.submit(g, cp.PrimalDualSolution(np.zeros(g.n + 1), np.zeros(g.m)))
        assert i not in srv.shapes


def test_get_on_a_running_borrowed_request_keeps_its_arrays(cp):
    """cpdlp_server_get on a request that has not finished fails with "request still
    running" and leaves the request queued: the borrowed arrays stay held (the library
    still reads them at admission) and the id stays valid for a later get."""
    import gc
    from coolpdlp_hip.serving import NativeServer
    g = G.request_lp(3000, 6000, 81)          # plain route (m + n > batch_limit)
    algo = cp.PDLP("mixed", termination_reltol=1e-6)
    with NativeServer(algo, batch_bytes=64 << 20, plain_bytes=4 << 30, batch_limit=2000,
                      batch2_limit=0) as srv:
        i = srv.submit(G.request_lp(3000, 6000, 81), borrow=True)   # no other reference to the LP
        with pytest.raises(Exception, match="request still running"):
            srv.get(i)                      # not even admitted: nothing pumped yet
        assert i in srv._held and i in srv.shapes
        gc.collect()
        res = _drain(srv, 1)
    sol, st, _, route = res[i]
    assert route >= 0 and st.termination_status == cp.OPTIMAL
    assert abs(float(sol.x @ g.c) - g.objective) <= 2e-5 * (1 + abs(g.objective))
    assert not srv._held and not srv.shapes


def test_request_larger_than_the_batch_arena_fails_alone(cp):
    """A batch-route request whose footprint exceeds the whole arena can never be
    admitted: it is reported as failed (CPDLP_ERR_ARGUMENT) instead of blocking the
    requests behind it, which complete."""
    from coolpdlp_hip.batch import footprint
    from coolpdlp_hip.serving import NativeServer
    big = G.request_lp(1800, 3600, 82)
    small = [G.request_lp(m, 2 * m, 83 + m) for m in (60, 120)]
    arena = footprint(small[1].m, small[1].n, small[1].nnz) * 3
    assert footprint(big.m, big.n, big.nnz) > arena
    algo = cp.PDLP("mixed", termination_reltol=1e-6)
    with NativeServer(algo, batch_bytes=arena, plain_bytes=1 << 30, batch_limit=8000,
                      checks_per_round=5, batch2_limit=0) as srv:
        ib = srv.submit(big)
        ids = [srv.submit(g) for g in small]
        done, err = {}, None
        t_end = time.perf_counter() + 120
        while len(done) + (err is not None) < 3 and time.perf_counter() < t_end:
            srv.pump()
            for i in srv.poll():
                try:
                    done[i] = srv.get(i)
                except cp.ArgumentError as e:
                    assert i == ib
                    err = e
            time.sleep(0.0005)
    assert err is not None and "arena" in str(err)
    for i, g in zip(ids, small):
        sol, st, _, route = done[i]
        assert route == -1 and st.termination_status == cp.OPTIMAL
        assert abs(float(sol.x @ g.c) - g.objective) <= 2e-5 * (1 + abs(g.objective))


@pytest.mark.parametrize("batch_limit", [0, 8000])
def test_native_server_degenerate_requests(cp, batch_limit):
    """Degenerate requests on each route (batch_limit 0: every request to a plain lane;
    8000: every request to the batch route): no constraints with a nonzero cost and with
    c = 0 (the early exit), an empty row and column, a free row with a fixed variable.
    Each ends OPTIMAL with the oracle's solution to 1e-6."""
    import scipy.sparse as sp

    import oracle as O
    from coolpdlp_hip.serving import NativeServer

    def lp(A, c, lv, uv, lc, uc):
        A = sp.csc_matrix(np.asarray(A, float))
        return G.GeneratedLP(A.shape[0], A.shape[1], A.indptr.astype(np.int64), A.indices.astype(np.int32),
                             A.data.astype(float), *[np.asarray(v, float) for v in (c, lv, uv, lc, uc)])

    cases = [lp(np.zeros((0, 3)), [1.0, -2.0, 0.5], [0, -1, -3], [2, 4, 5], [], []),
             lp(np.zeros((0, 3)), [0.0, 0.0, 0.0], [1, -2, -np.inf], [2, -1, np.inf], [], []),
             lp([[1.0, 0, 1.0], [0, 0, 0], [2.0, 0, -1.0]], [1.0, 1.0, 1.0], [0, 0, 0], [10, 10, 10],
                [1, -1, 0], [1, 1, 5]),
             lp([[1.0, 1.0, 0.0], [1.0, -1.0, 2.0], [0.0, 1.0, 1.0]], [1.0, 2.0, -1.0], [0, 1.5, -5],
                [4, 1.5, 5], [-np.inf, 0.5, 1.0], [np.inf, 3.0, 1.0])]
    algo = cp.PDLP("float64", termination_reltol=1e-6)
    with NativeServer(algo, batch_bytes=64 << 20, plain_bytes=1 << 30, batch_limit=batch_limit) as srv:
        ids = {srv.submit(g): g for g in cases}
        res = _drain(srv, len(ids))
    for i, g in ids.items():
        sol, st, _, route = res[i]
        assert route == (-1 if batch_limit else 0)
        ref = O.solve(g, O.default_params(termination_reltol=1e-6))
        assert st.termination_status == cp.OPTIMAL and ref.status == "OPTIMAL"
        np.testing.assert_allclose(sol.x, ref.x, rtol=1e-6, atol=1e-6)
