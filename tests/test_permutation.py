"""Row and column density permutations (coolpdlp_hip.permutation, mirroring
src/components/permutation.jl): the reference's own test (test/components/permutation.jl:
5-15) on the same sizes and densities, stability of the order on ties, the permuted MILP's
fields, and -- on the GPU -- that the permuted MILP solves to the same optimum."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
from coolpdlp_hip import MILP, generators as G
from coolpdlp_hip._lib import DimensionMismatch
from coolpdlp_hip.permutation import (increasing_column_order, permute_rows_columns,
                                      sort_rows_columns)

from test_qps import generated


def test_sort_columns_reference_cases():
    """test/components/permutation.jl:5-15 (over m x n rather than n x n)."""
    rng = np.random.default_rng(0)
    for m in (10, 20, 30):
        for n in (10, 20, 30):
            for p in (0.01, 0.1, 0.3):
                A = sp.random(m, n, density=p, format="csc", random_state=rng)
                perm_col = increasing_column_order(A)
                perm_row = increasing_column_order(A.T.tocsc())
                B = permute_rows_columns(A, perm_col=perm_col, perm_row=perm_row)
                assert np.array_equal(A.toarray()[:, perm_col][perm_row, :], B.toarray())
                assert np.all(np.diff(np.diff(B.indptr)) >= 0)             # columns
                assert np.all(np.diff(np.diff(B.tocsr().indptr)) >= 0)     # rows


def test_ties_keep_their_order():
    # column counts 2, 1, 2, 1, 0 -> stable order 4, 1, 3, 0, 2
    A = sp.csc_matrix(np.array([[1, 0, 1, 0, 0], [1, 1, 1, 1, 0]], float))
    assert increasing_column_order(A).tolist() == [4, 1, 3, 0, 2]
    with pytest.raises(DimensionMismatch):
        permute_rows_columns(A, perm_col=[0, 1, 2, 3, 3], perm_row=[0, 1])
    with pytest.raises(DimensionMismatch):
        permute_rows_columns(A, perm_col=[0, 1, 2, 3, 4], perm_row=[1])


def _milp():
    g = G.column_lp(60, 120, (1, 8), seed=4, box_frac=0.3)
    milp = MILP.from_generated(g)
    milp.int_var[::7] = True
    milp.D1 = np.arange(1.0, 61.0)
    milp.D2 = np.arange(1.0, 121.0)
    return milp


def test_sort_rows_columns_fields():
    milp = _milp()
    s = sort_rows_columns(milp)
    pv = increasing_column_order(milp.A)
    pc = increasing_column_order(milp.At)
    assert np.array_equal(s.A.toarray(), milp.A.toarray()[pc][:, pv])
    assert np.array_equal(s.At.toarray(), s.A.toarray().T)
    for f in ("c", "lv", "uv", "D2", "int_var"):
        assert np.array_equal(getattr(s, f), getattr(milp, f)[pv]), f
    for f in ("lc", "uc", "D1"):
        assert np.array_equal(getattr(s, f), getattr(milp, f)[pc]), f
    assert s.var_names == [milp.var_names[j] for j in pv]
    assert np.all(np.diff(np.diff(s.A.indptr)) >= 0) and np.all(np.diff(np.diff(s.At.indptr)) >= 0)
    # the oracle reaches the same optimum on both orders
    milp.D1, milp.D2 = np.ones(60), np.ones(120)
    s = sort_rows_columns(milp)
    r0 = O.solve(generated(milp), O.default_params(termination_reltol=1e-8))
    r1 = O.solve(generated(s), O.default_params(termination_reltol=1e-8))
    assert r0.status == r1.status == "OPTIMAL"
    assert abs(milp.c @ r0.x - s.c @ r1.x) <= 1e-6 * (1 + abs(milp.c @ r0.x))


@pytest.mark.gpu
def test_sorted_milp_solves_to_the_same_optimum_on_gpu():
    from coolpdlp_hip import OPTIMAL, PDLP, objective_value, solve
    milp = _milp()
    milp.D1, milp.D2 = np.ones(60), np.ones(120)
    s = sort_rows_columns(milp)
    sol0, st0 = solve(milp, PDLP(termination_reltol=1e-8))
    sol1, st1 = solve(s, PDLP(termination_reltol=1e-8))
    assert st0.termination_status == st1.termination_status == OPTIMAL
    o0, o1 = objective_value(sol0.x, milp), objective_value(sol1.x, s)
    assert abs(o0 - o1) <= 1e-6 * (1 + abs(o0))
    # the sorted solution, put back in the original order, is optimal for the original
    x = np.empty_like(sol1.x)
    x[increasing_column_order(milp.A)] = sol1.x
    assert abs(objective_value(x, milp) - o0) <= 1e-6 * (1 + abs(o0))
