"""GPU tests of continuous-batching serving: requests that arrive while others run give
the same result as when solved alone (bit-identical for batch-routed requests, and for
plain-routed ones, which use a fixed CSR format); every result is OPTIMAL at 1e-6 and
matches the closed-form optimum."""
import time

import numpy as np
import pytest

from coolpdlp_hip import generators as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cp():
    import coolpdlp_hip
    return coolpdlp_hip


def test_serving_alone_equals_in_mix(cp):
    from coolpdlp_hip.serving import Server
    lps = [G.request_lp(m, 2 * m, 500 + m) for m in (60, 150, 420, 900, 1500, 2600)]
    big = G.request_lp(3000, 6000, 77)
    srv = Server(reltol=1e-6, batch_bytes=256 << 20, plain_bytes=4 << 30, checks_per_round=2,
                 batch_limit=8000, keep_solutions=True)
    t0 = time.perf_counter()
    srv.submit(lps[0])
    srv.submit(big)                      # plain route (m + n > batch_limit)
    for k in range(1, len(lps)):
        for _ in range(3):
            srv.pump()
        srv.submit(lps[k])               # arrives while the others are on the GPU
    srv.drain(time.perf_counter() + 300)
    srv.close()
    assert len(srv.done) == len(lps) + 1
    by_lp = {id(r.lp): r for r in srv.done}
    algo = cp.PDLP("mixed", termination_reltol=1e-6)
    for g in lps:
        r = by_lp[id(g)]
        assert r.status == "OPTIMAL" and r.route == "batch"
        with cp.BatchSolver(algo, hbm_bytes=64 << 20) as B:
            (alone, st), = B.solve_all([g])
        assert np.array_equal(alone.x, r.x)
        assert abs(r.objective - g.objective) <= 2e-5 * (1 + abs(g.objective))
    r = by_lp[id(big)]
    assert r.status == "OPTIMAL" and r.route == "plain"
    sol, st = cp.solve(cp.MILP.from_generated(big), cp.PDLP("mixed", None, "csr", termination_reltol=1e-6))
    assert np.array_equal(sol.x, r.x)
    assert abs(r.objective - big.objective) <= 2e-5 * (1 + abs(big.objective))



@pytest.mark.parametrize("batch_limit", [8000, 0])
def test_serving_warm_start_resubmission(cp, batch_limit):
    """A request submitted with the solution of an earlier one as its start (the re-solve
    of a nearby LP) takes at most half the passes of the cold solve, on the batch route (batch_limit 8000)
    and on the plain route (batch_limit 0)."""
    from coolpdlp_hip.serving import Server
    g = G.request_lp(400, 800, 91)
    srv = Server(reltol=1e-6, batch_bytes=64 << 20, plain_bytes=2 << 30, batch_limit=batch_limit,
                 keep_solutions=True)
    srv.submit(g)
    srv.drain(time.perf_counter() + 120)
    first, = srv.done
    assert first.status == "OPTIMAL"
    srv.submit(g, sol=cp.PrimalDualSolution(first.x, first.y))
    srv.drain(time.perf_counter() + 120)
    srv.close()
    second = srv.done[1]
    assert second.route == ("batch" if batch_limit else "plain")
    assert second.status == "OPTIMAL"
    assert second.kkt_passes <= first.kkt_passes // 2
    assert abs(second.objective - g.objective) <= 2e-5 * (1 + abs(g.objective))
