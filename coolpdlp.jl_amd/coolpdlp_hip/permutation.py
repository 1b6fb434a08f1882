"""Row and column density permutations, src/components/permutation.jl:1-46.

Permutations are 0-based index arrays (the reference's are 1-based); `perm[k]` is the
original index placed at position k, as with Julia's `sortperm`.  The sorts are stable,
as `sortperm` is, so equal counts keep their original order.

The engine does not need these to run fast: its own layout builder bins rows by length
and orders rows and columns by first touch on the device side.  They are the
reference's host-side transforms, for callers that permute a MILP themselves.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from ._lib import DimensionMismatch


def increasing_column_order(A) -> np.ndarray:
    """permutation.jl:1-4: the columns by increasing stored-entry count (stable)."""
    A = sp.csc_matrix(A)
    return np.argsort(np.diff(A.indptr), kind="stable")


def _check_perm(p, size, what):
    p = np.asarray(p, dtype=np.int64)
    if p.shape != (size,) or not np.array_equal(np.sort(p), np.arange(size)):
        raise DimensionMismatch(f"{what} is not a permutation of 0..{size - 1}")
    return p


def permute_rows_columns(A, *, perm_col, perm_row):
    """permutation.jl:6-14: the matrix B with B[k, l] = A[perm_row[k], perm_col[l]],
    built from the triplets through the inverse permutations, as the reference does."""
    A = sp.coo_matrix(A)
    m, n = A.shape
    perm_row = _check_perm(perm_row, m, "perm_row")
    perm_col = _check_perm(perm_col, n, "perm_col")
    inv_row = np.empty(m, np.int64)
    inv_row[perm_row] = np.arange(m)
    inv_col = np.empty(n, np.int64)
    inv_col[perm_col] = np.arange(n)
    B = sp.csc_matrix((A.data, (inv_row[A.row], inv_col[A.col])), shape=(m, n))
    B.sort_indices()
    return B


def sort_rows_columns(milp):
    """sort_rows_columns(milp), permutation.jl:16-46: a new MILP whose variables are in
    increasing column density of A and whose constraints are in increasing row density
    (the column density of A'), with every per-variable and per-constraint field
    permuted to match.  Returns the MILP; the permutations are `increasing_column_order`
    of A and of A'."""
    from .problems import MILP
    perm_var = increasing_column_order(milp.A)
    perm_cons = increasing_column_order(milp.At)
    return MILP(
        c=milp.c[perm_var], lv=milp.lv[perm_var], uv=milp.uv[perm_var],
        A=permute_rows_columns(milp.A, perm_col=perm_var, perm_row=perm_cons),
        At=permute_rows_columns(milp.At, perm_col=perm_cons, perm_row=perm_var),
        lc=milp.lc[perm_cons], uc=milp.uc[perm_cons],
        D1=milp.D1[perm_cons], D2=milp.D2[perm_var],
        int_var=milp.int_var[perm_var], var_names=[milp.var_names[j] for j in perm_var],
        dataset=milp.dataset, name=milp.name, path=milp.path)
