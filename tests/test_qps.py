"""CPU tests of MPS ingestion (coolpdlp_hip.qps): the fields of QPSReader.jl's QPSData
that MILP(qps::QPSData) reads (src/problems/milp.jl:122-137), on hand-written files whose
expected bounds follow from the MPS format's definition (worked by hand in the comments),
a write -> read round trip that must be exact, and the reference's MOI test LP
(test/moi.jl:44-52, objective 205.0) read from fixed-format MPS and solved by the oracle."""
import io
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
from coolpdlp_hip import MILP, generators as G, is_feasible, nbcons, nbvar, nbvar_int, objective_value
from coolpdlp_hip._lib import ArgumentError
from coolpdlp_hip.qps import VTYPE_BINARY, VTYPE_CONTINUOUS, VTYPE_INTEGER, read_mps, write_mps

INF = np.inf
SUITE_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mps")


def suite():
    """(name, MILP, HiGHS objective) for the Netlib-style MPS fixtures, smallest first by
    variable count, as test/algorithms/all.jl:12-15 sorts Netlib."""
    with open(os.path.join(SUITE_DIR, "highs.json")) as f:
        rec = json.load(f)["instances"]
    out = []
    for name in sorted(rec):
        path = os.path.join(SUITE_DIR, name + ".mps")
        milp = MILP.from_qps(read_mps(path), dataset="netlike", path=path)
        out.append((name, milp, rec[name]["objective"]))
    return sorted(out, key=lambda t: nbvar(t[1]))


def generated(milp):
    """The MILP as the oracle takes it."""
    A = milp.A
    return G.GeneratedLP(A.shape[0], A.shape[1], A.indptr.astype(np.int64), A.indices.astype(np.int32),
                         A.data.copy(), milp.c, milp.lv, milp.uv, milp.lc, milp.uc)


def row_scale(milp):
    """1 + the largest finite row bound: the relative KKT error bounds the residual by
    about this factor times the tolerance."""
    b = np.concatenate([milp.lc, milp.uc])
    b = b[np.isfinite(b)]
    return 1.0 + (np.max(np.abs(b)) if len(b) else 0.0)

LP205_FIXED = """\
NAME          LP205
ROWS
 N  COST
 G  C1
 G  C2
COLUMNS
    X         COST          12.0   C1             6.0
    X         C2             7.0
    Y         COST          20.0   C1             8.0
    Y         C2            12.0
RHS
    RHS       C1           100.0   C2           120.0
BOUNDS
 UP BND       Y              3.0
ENDATA
"""

# Every bound type and every row type with and without a range.  Expected, by the
# format's rules:
#   a UP 4          -> [0, 4]           b UP -2        -> [-inf, -2] (lower still default 0)
#   c LO -1         -> [-1, inf]        d FX 2.5       -> [2.5, 2.5]
#   e FR            -> [-inf, inf]      f MI           -> [-inf, inf]
#   g PL            -> [0, inf]         h BV           -> [0, 1], binary
#   i LI 2          -> [2, inf], int    j UI 7         -> [0, 7], int
#   k LO -5, UP -1  -> [-5, -1]         m1 (INTORG)    -> [0, inf], int
#   rows: e_pos E 3 R 2 -> [3, 5]; e_neg E 3 R -2 -> [1, 3]; e_zero E 3 -> [3, 3];
#         l_row L 10 R -4 -> [6, 10]; g_row G -1 R 6 -> [-1, 5]; l_plain L 8 -> [-inf, 8];
#         g_none G (no RHS) -> [0, inf]
ALL_TYPES_FREE = """\
* comment line
NAME all_types
OBJSENSE
    MAX
ROWS
 N obj
 E e_pos
 E e_neg
 E e_zero
 L l_row
 G g_row
 L l_plain
 N spare
 G g_none
COLUMNS
 a obj 1 e_pos 1
 a e_neg 2 spare 9
 b e_zero 3
 c l_row 4
 d g_row 5
 e l_plain 6
 f g_none 7
 g obj -1
 h e_pos 1
 i e_neg 1
 j e_zero 1
 k l_row 1
 MARKER1 'MARKER' 'INTORG'
 m1 obj 2 g_row 1
 MARKER2 'MARKER' 'INTEND'
RHS
 rhs e_pos 3 e_neg 3
 rhs e_zero 3 l_row 10
 rhs g_row -1 l_plain 8
 rhs obj -3.5 spare 4
RANGES
 rng e_pos 2 e_neg -2
 rng l_row -4 g_row 6
BOUNDS
 UP bnd a 4
 UP bnd b -2
 LO bnd c -1
 FX bnd d 2.5
 FR bnd e
 MI bnd f
 PL bnd g
 BV bnd h
 LI bnd i 2
 UI bnd j 7
 LO bnd k -5
 UP bnd k -1
ENDATA
"""


def test_lp205_fields_and_solution():
    q = read_mps(io.StringIO(LP205_FIXED))
    assert q.name == "LP205" and q.objname == "COST" and q.objsense == "MIN"
    assert q.varnames == ["X", "Y"] and q.connames == ["C1", "C2"]
    milp = MILP.from_qps(q, dataset="hand", path="lp205.mps")
    assert (nbvar(milp), nbcons(milp), nbvar_int(milp)) == (2, 2, 0)
    assert np.array_equal(milp.A.toarray(), [[6, 8], [7, 12]])
    assert np.array_equal(milp.At.toarray(), [[6, 7], [8, 12]])
    assert np.array_equal(milp.c, [12, 20])
    assert np.array_equal(milp.lv, [0, 0]) and np.array_equal(milp.uv, [INF, 3])
    assert np.array_equal(milp.lc, [100, 120]) and np.array_equal(milp.uc, [INF, INF])
    # the MILP's name comes from the keywords (here derived from the path), not from the
    # file's NAME card, as in the reference
    assert milp.var_names == ["X", "Y"] and milp.name == "lp205" and milp.dataset == "hand"
    r = O.solve(generated(milp), O.default_params(termination_reltol=1e-8))
    assert r.status == "OPTIMAL"
    assert abs(float(milp.c @ r.x) - 205.0) <= 1e-4
    assert np.allclose(r.x, [15.0, 1.25], atol=1e-4)


def test_every_bound_and_row_type():
    q = read_mps(io.StringIO(ALL_TYPES_FREE))
    assert q.objsense == "MAX" and q.c0 == 3.5
    assert q.varnames == list("abcdefghijk") + ["m1"]
    assert q.connames == ["e_pos", "e_neg", "e_zero", "l_row", "g_row", "l_plain", "g_none"]
    assert np.array_equal(q.lvar, [0, -INF, -1, 2.5, -INF, -INF, 0, 0, 2, 0, -5, 0])
    assert np.array_equal(q.uvar, [4, -2, INF, 2.5, INF, INF, INF, 1, INF, 7, -1, INF])
    assert np.array_equal(q.lcon, [3, 1, 3, 6, -1, -INF, 0])
    assert np.array_equal(q.ucon, [5, 3, 3, 10, 5, 8, INF])
    want_types = [VTYPE_CONTINUOUS] * 7 + [VTYPE_BINARY, VTYPE_INTEGER, VTYPE_INTEGER,
                                           VTYPE_CONTINUOUS, VTYPE_INTEGER]
    assert q.vartypes.tolist() == want_types
    # c keeps the file's signs (the sense is recorded, not applied); the spare N row's
    # entry (a, 9) is dropped
    assert np.array_equal(q.c, [1, 0, 0, 0, 0, 0, -1, 0, 0, 0, 0, 2])
    milp = MILP.from_qps(q)
    assert np.array_equal(milp.c, q.c)
    assert milp.int_var.tolist() == [False] * 7 + [True, True, True, False, True]
    dense = np.zeros((7, 12))
    for i, j, v in [(0, 0, 1), (1, 0, 2), (2, 1, 3), (3, 2, 4), (4, 3, 5), (5, 4, 6), (6, 5, 7),
                    (0, 7, 1), (1, 8, 1), (2, 9, 1), (3, 10, 1), (4, 11, 1)]:
        dense[i, j] = v
    assert np.array_equal(milp.A.toarray(), dense)


def test_duplicate_entries_are_summed():
    """sparse(arows, acols, avals) in the reference sums repeated (row, column) pairs."""
    text = "NAME dup\nROWS\n N o\n L r\nCOLUMNS\n x r 1.5\n x r 2.0\nRHS\n r 1\nENDATA\n"
    q = read_mps(io.StringIO(text))
    assert len(q.avals) == 2
    assert MILP.from_qps(q).A.toarray().tolist() == [[3.5]]


def test_write_read_round_trip_is_exact():
    g = G.column_lp(40, 60, 4, seed=11, eq_frac=0.3, box_frac=0.4)
    lv, uv, lc, uc = g.lv.copy(), g.uv.copy(), g.lc.copy(), g.uc.copy()
    lv[:5], uv[:5] = -INF, INF                # free
    lv[5:8] = -INF                            # upper only
    uv[8:10] = lv[8:10] = 1.25                # fixed
    lv[10:12], uv[10:12] = -2.0, 3.0          # box with a negative lower bound
    lv[12], uv[12] = 0.0, -1.0                # a negative UP over the default lower bound
    lc[:3], uc[:3] = -1.0, 2.0                # ranged rows
    lc[3:5], uc[3:5] = -INF, 7.0
    src = MILP.from_generated(G.GeneratedLP(g.m, g.n, g.colptr, g.rowval, g.nzval, g.c, lv, uv, lc, uc))
    buf = io.StringIO()
    write_mps(buf, c=src.c, lv=src.lv, uv=src.uv, A=src.A, lc=src.lc, uc=src.uc, name="rt")
    buf.seek(0)
    back = MILP.from_qps(read_mps(buf))
    for f in ("c", "lv", "uv", "lc", "uc"):
        assert np.array_equal(getattr(back, f), getattr(src, f)), f
    assert (back.A != src.A).nnz == 0 and back.A.nnz == src.A.nnz


def test_refusals():
    quad = "NAME q\nROWS\n N o\nCOLUMNS\n x o 1\nQUADOBJ\n x x 1\nENDATA\n"
    with pytest.raises(ArgumentError):
        read_mps(io.StringIO(quad))
    unknown = "NAME u\nROWS\n N o\n L r\nCOLUMNS\n x r 1\nBOUNDS\n UP b y 1\nENDATA\n"
    with pytest.raises(ArgumentError):
        read_mps(io.StringIO(unknown))
    unknown_row = "NAME u\nROWS\n N o\n L r\nCOLUMNS\n x s 1\nENDATA\n"
    with pytest.raises(ArgumentError):
        read_mps(io.StringIO(unknown_row))
    bad_row = "NAME u\nROWS\n N o\n Q r\nENDATA\n"
    with pytest.raises(ArgumentError):
        read_mps(io.StringIO(bad_row))
    with pytest.raises(ArgumentError):
        write_mps(io.StringIO(), c=[1.0], lv=[0.0], uv=[1.0], A=np.ones((1, 1)), lc=[-INF], uc=[INF])


def test_suite_reads_back_the_lps_highs_solved():
    """HiGHS on the LPs read from the fixtures reproduces the objectives it found on the
    in-memory LPs before they were written (tests/golden/make_mps_suite.py)."""
    from scipy.optimize import linprog
    import scipy.sparse as sp
    cases = suite()
    assert [n for n, _, _ in cases][:3] == ["netlike_free_fixed", "netlike_ranges", "netlike_badscale"]
    for name, milp, obj in cases:
        A = sp.csr_matrix(milp.A)
        eq = milp.lc == milp.uc
        up, lo = ~eq & np.isfinite(milp.uc), ~eq & np.isfinite(milp.lc)
        h = linprog(milp.c, A_ub=sp.vstack([A[up], -A[lo]]), b_ub=np.concatenate([milp.uc[up], -milp.lc[lo]]),
                    A_eq=A[eq], b_eq=milp.lc[eq], bounds=list(zip(milp.lv, milp.uv)), method="highs")
        assert h.status == 0, name
        assert abs(h.fun - obj) <= 1e-9 * (1 + abs(obj)), name


@pytest.mark.parametrize("algo", ["PDLP", "PDHG"])
def test_suite_oracle_vs_highs(algo):
    """test_optimizer, test/algorithms/all.jl:20-62, on the oracle: PDLP at reltol 1e-5 with
    objective rtol 1e-2 and cons_tol 1e-2; PDHG at reltol 1e-1 with objective rtol 1e-1.
    PDHG's cons_tol 1e-1 is scaled by 1 + the largest row bound: these right-hand sides
    reach 80, where a relative KKT error of 1e-1 allows residuals above 0.1."""
    for name, milp, obj in suite():
        reltol = 1e-5 if algo == "PDLP" else 1e-1
        r = O.solve(generated(milp), O.default_params(algo, termination_reltol=reltol,
                                                      max_kkt_passes=10 ** 7, time_limit=1e9))
        assert r.status == "OPTIMAL", name
        if algo == "PDLP":
            assert is_feasible(r.x, milp, cons_tol=1e-2, int_tol=INF, verbose=False), name
            assert abs(objective_value(r.x, milp) - obj) <= 1e-2 * abs(obj), name
        else:
            assert is_feasible(r.x, milp, cons_tol=1e-1 * row_scale(milp), int_tol=INF, verbose=False), name
            assert abs(objective_value(r.x, milp) - obj) <= 1e-1 * abs(obj), name


def test_round_trip_property():
    """write_mps -> read_mps returns the same LP for random shapes, sparsity, bound kinds
    (free, lower-only, upper-only, boxed, fixed) and row kinds (=, <=, >=, ranged with
    integer-valued bounds, so that lc + (uc - lc) is exact)."""
    from hypothesis import given, settings, strategies as st

    @settings(max_examples=60, deadline=None)
    @given(st.integers(1, 12), st.integers(1, 12), st.integers(0, 2**31 - 1))
    def check(m, n, seed):
        rng = np.random.default_rng(seed)
        A = sp.random(m, n, density=rng.uniform(0.1, 0.8), format="csc", random_state=rng,
                      data_rvs=lambda k: rng.standard_normal(k) * 10.0 ** rng.integers(-3, 4, k))
        kinds_v = rng.integers(0, 5, n)
        lv = np.where(kinds_v == 0, -INF, np.where(kinds_v == 2, -INF, rng.uniform(-5, 5, n)))
        uv = np.where(kinds_v <= 1, INF, lv + rng.uniform(0, 5, n))
        uv = np.where(kinds_v == 2, rng.uniform(-5, 5, n), uv)
        uv = np.where(kinds_v == 4, lv, uv)
        kinds_r = rng.integers(0, 4, m)
        b = rng.integers(-20, 20, m).astype(float)
        lc = np.where(kinds_r == 1, -INF, b)
        uc = np.where(kinds_r == 2, INF, np.where(kinds_r == 3, b + rng.integers(1, 9, m), b))
        c = rng.standard_normal(n) * (rng.random(n) < 0.7)
        buf = io.StringIO()
        write_mps(buf, c=c, lv=lv, uv=uv, A=A, lc=lc, uc=uc)
        buf.seek(0)
        back = MILP.from_qps(read_mps(buf))
        for name, want in (("c", c), ("lv", lv), ("uv", uv), ("lc", lc), ("uc", uc)):
            assert np.array_equal(getattr(back, name), want), name
        assert back.A.shape == (m, n) and (back.A != A).nnz == 0

    check()
