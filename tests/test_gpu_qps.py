"""GPU test of MPS ingestion end to end: the reference's MOI test LP (test/moi.jl:44-52,
objective 205.0) read from fixed-format MPS, turned into a MILP as MILP(qps::QPSData)
does (src/problems/milp.jl:122-137) and solved by the HIP engine, as the tutorial does
with a Netlib file (test/tutorial.jl:14-48)."""
import io

import numpy as np
import pytest

from coolpdlp_hip import MILP, OPTIMAL, PDLP, is_feasible, objective_value, read_mps, solve

from test_qps import LP205_FIXED

pytestmark = pytest.mark.gpu


def test_mps_lp205_solves_on_gpu():
    milp = MILP.from_qps(read_mps(io.StringIO(LP205_FIXED)), dataset="hand", name="lp205")
    sol, stats = solve(milp, PDLP(termination_reltol=1e-8, time_limit=30.0))
    assert stats.termination_status == OPTIMAL
    assert abs(objective_value(sol.x, milp) - 205.0) <= 1e-4
    assert np.allclose(sol.x, [15.0, 1.25], atol=1e-4)
    assert is_feasible(sol.x, milp, cons_tol=1e-6, verbose=False)
