"""MPS ingestion: the fields of QPSReader.jl's QPSData (QPSReader 0.2.1, the reference's
dependency, Project.toml:20,36; not vendored), which the reference turns into a MILP
(src/problems/milp.jl:117-137).  This is a restatement of the MPS format (free and fixed
column layout), written from the format's definition, for linear and mixed-integer
programs:

  NAME, OBJSENSE (MIN / MAX, on the same or the next line), ROWS (the first N row is the
  objective, later N rows are free rows whose entries, RHS and RANGES are ignored; E / L /
  G; an entry on an undeclared row is an error), COLUMNS (with MARKER INTORG / INTEND blocks for integer
  variables), RHS (a value on the objective row is minus the objective constant), RANGES
  (E: [rhs, rhs+|R|] for R > 0, [rhs-|R|, rhs] for R < 0; L: [rhs-|R|, rhs];
  G: [rhs, rhs+|R|]), BOUNDS (UP, LO, FX, FR, MI, PL, BV, LI, UI; an UP bound below 0 on
  a variable whose lower bound is still the default 0 makes that lower bound -inf, the
  usual MPS convention), ENDATA.  Default bounds are [0, +inf]; integer variables keep
  them (BV makes [0, 1]).  Quadratic sections (QUADOBJ / QMATRIX) are refused: the
  engine solves LPs.

`read_mps` returns a `QPSData`; `milp_from_qps` builds the MILP as the reference's
`MILP(qps::QPSData)` does (A from the triplets, duplicates summed, int_var from the
variable types, var_names).
"""
from __future__ import annotations

import io
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp

from ._lib import ArgumentError

VTYPE_CONTINUOUS, VTYPE_INTEGER, VTYPE_BINARY = 0, 1, 2


@dataclass
class QPSData:
    """The fields of QPSReader.jl's QPSData that a MILP uses."""
    name: str = ""
    objsense: str = "MIN"
    objname: str = ""
    c0: float = 0.0
    c: np.ndarray = field(default_factory=lambda: np.zeros(0))
    lvar: np.ndarray = field(default_factory=lambda: np.zeros(0))
    uvar: np.ndarray = field(default_factory=lambda: np.zeros(0))
    lcon: np.ndarray = field(default_factory=lambda: np.zeros(0))
    ucon: np.ndarray = field(default_factory=lambda: np.zeros(0))
    arows: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    acols: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    avals: np.ndarray = field(default_factory=lambda: np.zeros(0))
    vartypes: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int8))
    varnames: list = field(default_factory=list)
    connames: list = field(default_factory=list)


_SECTIONS = {"NAME", "OBJSENSE", "OBJSENCE", "ROWS", "COLUMNS", "RHS", "RANGES", "BOUNDS",
             "ENDATA", "QUADOBJ", "QMATRIX", "QSECTION", "QCMATRIX", "SOS"}


def read_mps(source) -> QPSData:
    """Parse an MPS file: `source` is a path or an open text file (io.StringIO for text
    held in memory)."""
    if hasattr(source, "read"):
        text = source.read()
    else:
        with open(source) as f:
            text = f.read()
    q = QPSData()
    row_index: dict = {}
    free_rows: set = set()             # N rows after the objective: ignored
    row_type: list = []
    col_index: dict = {}
    c_list: list = []
    arows: list = []
    acols: list = []
    avals: list = []
    vtypes: list = []
    rhs: dict = {}
    ranges: dict = {}
    bounds: list = []
    section = None
    integer_block = False
    for raw in io.StringIO(text):
        line = raw.rstrip("\n")
        if not line.strip() or line.lstrip().startswith("*"):
            continue
        head = line.split()[0].upper()
        if not line[0].isspace() and head in _SECTIONS:
            section = "OBJSENSE" if head == "OBJSENCE" else head
            rest = line.split()[1:]
            if section == "NAME":
                q.name = " ".join(rest)
            elif section == "OBJSENSE" and rest:
                q.objsense = rest[0].upper()
            elif section in ("QUADOBJ", "QMATRIX", "QSECTION", "QCMATRIX"):
                raise ArgumentError(f"quadratic section {head} is not supported (LP engine)")
            elif section == "SOS":
                raise ArgumentError("SOS constraints are not supported")
            if section == "ENDATA":
                break
            continue
        f = line.split()   # free MPS; fixed MPS without spaces in names splits alike
        if section == "OBJSENSE":
            q.objsense = f[0].upper()
        elif section == "ROWS":
            t, nm = f[0].upper(), f[1]
            if t == "N":
                if not q.objname:
                    q.objname = nm
                else:
                    free_rows.add(nm)   # further N rows are free rows: ignored
                continue
            if t not in ("E", "L", "G"):
                raise ArgumentError(f"unknown row type {t!r}")
            if nm in row_index:
                raise ArgumentError(f"duplicate row {nm!r}")
            row_index[nm] = len(row_type)
            row_type.append(t)
            q.connames.append(nm)
        elif section == "COLUMNS":
            if len(f) >= 3 and f[1].strip("'\"").upper() == "MARKER":
                m = f[2].strip("'\"").upper()
                if m == "INTORG":
                    integer_block = True
                elif m == "INTEND":
                    integer_block = False
                continue
            cn = f[0]
            if cn not in col_index:
                col_index[cn] = len(c_list)
                c_list.append(0.0)
                vtypes.append(VTYPE_INTEGER if integer_block else VTYPE_CONTINUOUS)
                q.varnames.append(cn)
            j = col_index[cn]
            for k in range(1, len(f) - 1, 2):
                rn, val = f[k], float(f[k + 1])
                if rn == q.objname:
                    c_list[j] += val
                elif rn in row_index:
                    arows.append(row_index[rn])
                    acols.append(j)
                    avals.append(val)
                elif rn not in free_rows:
                    raise ArgumentError(f"entry on unknown row {rn!r}")
        elif section in ("RHS", "RANGES"):
            # (set name,) row, value (, row, value): an odd field count carries a set name
            pairs = f[1:] if len(f) % 2 == 1 else f
            for k in range(0, len(pairs) - 1, 2):
                rn, val = pairs[k], float(pairs[k + 1])
                if section == "RHS":
                    if rn == q.objname:
                        q.c0 = -val
                    elif rn in row_index:
                        rhs[row_index[rn]] = val
                    elif rn not in free_rows:
                        raise ArgumentError(f"RHS on unknown row {rn!r}")
                elif rn in row_index:
                    ranges[row_index[rn]] = val
                elif rn not in free_rows:
                    raise ArgumentError(f"RANGES on unknown row {rn!r}")
        elif section == "BOUNDS":
            # type (, set name), column (, value); FR / MI / PL carry no value, BV may
            t = f[0].upper()
            if t in ("FR", "MI", "PL", "BV"):
                cn = f[2] if len(f) >= 3 and f[2] in col_index else f[1]
                val = None
            elif len(f) >= 4:
                cn, val = f[2], float(f[3])
            elif len(f) == 3:
                cn, val = f[1], float(f[2])
            else:
                raise ArgumentError(f"bound line {line!r} has no value")
            if cn not in col_index:
                raise ArgumentError(f"bound on unknown column {cn!r}")
            bounds.append((t, col_index[cn], val))
        elif section is None:
            raise ArgumentError(f"data before any section: {line!r}")
    n, m = len(c_list), len(row_type)
    q.c = np.array(c_list, dtype=np.float64)
    q.vartypes = np.array(vtypes, dtype=np.int8)
    lvar = np.zeros(n)
    uvar = np.full(n, np.inf)
    lower_set = np.zeros(n, bool)
    for t, j, val in bounds:
        if t == "UP":
            uvar[j] = val
            if val < 0 and not lower_set[j] and lvar[j] == 0.0:
                lvar[j] = -np.inf
        elif t == "LO":
            lvar[j] = val
            lower_set[j] = True
        elif t == "FX":
            lvar[j] = uvar[j] = val
            lower_set[j] = True
        elif t == "FR":
            lvar[j], uvar[j] = -np.inf, np.inf
            lower_set[j] = True
        elif t == "MI":
            lvar[j] = -np.inf
            lower_set[j] = True
        elif t == "PL":
            uvar[j] = np.inf
        elif t == "BV":
            lvar[j], uvar[j] = 0.0, 1.0
            lower_set[j] = True
            q.vartypes[j] = VTYPE_BINARY
        elif t == "LI":
            lvar[j] = val
            lower_set[j] = True
            q.vartypes[j] = VTYPE_INTEGER
        elif t == "UI":
            uvar[j] = val
            q.vartypes[j] = VTYPE_INTEGER
        else:
            raise ArgumentError(f"unknown bound type {t!r}")
    q.lvar, q.uvar = lvar, uvar
    lcon = np.empty(m)
    ucon = np.empty(m)
    for i, t in enumerate(row_type):
        b = rhs.get(i, 0.0)
        r = ranges.get(i)
        if t == "E":
            lcon[i] = ucon[i] = b
            if r is not None:
                if r > 0:
                    ucon[i] = b + abs(r)
                elif r < 0:
                    lcon[i] = b - abs(r)
        elif t == "L":
            lcon[i], ucon[i] = -np.inf, b
            if r is not None:
                lcon[i] = b - abs(r)
        else:   # G
            lcon[i], ucon[i] = b, np.inf
            if r is not None:
                ucon[i] = b + abs(r)
    q.lcon, q.ucon = lcon, ucon
    q.arows = np.array(arows, dtype=np.int64)
    q.acols = np.array(acols, dtype=np.int64)
    q.avals = np.array(avals, dtype=np.float64)
    return q


def milp_from_qps(q: QPSData, **kw):
    """MILP(qps::QPSData; kwargs...), src/problems/milp.jl:122-137: A from the triplets
    (duplicates summed, as `sparse` does), c, bounds and var_names as read, int_var for
    integer and binary variables.  Like the reference, the objective sense and the
    objective constant stay in the QPSData and are not applied to the MILP."""
    from .problems import MILP
    m, n = len(q.lcon), len(q.lvar)
    A = sp.csc_matrix((q.avals, (q.arows, q.acols)), shape=(m, n))
    A.sum_duplicates()
    int_var = (q.vartypes == VTYPE_BINARY) | (q.vartypes == VTYPE_INTEGER)
    return MILP(c=q.c, lv=q.lvar, uv=q.uvar, A=A, lc=q.lcon, uc=q.ucon, int_var=int_var,
                var_names=list(q.varnames), **kw)


def _num(v) -> str:
    """Shortest decimal that reads back as the same double."""
    return repr(float(v))


def write_mps(path_or_file, *, c, lv, uv, A, lc, uc, name="LP", varnames=None, connames=None):
    """Write an LP in free MPS (min c'x, lc <= Ax <= uc, lv <= x <= uv).  A row with two
    different finite bounds becomes a G row with a range; a row with no finite bound is
    refused.  A ranged row reads back as [lc, lc + (uc - lc)], which can differ from uc
    in the last bit."""
    A = sp.csc_matrix(A)
    m, n = A.shape
    varnames = varnames or [f"x{j + 1}" for j in range(n)]
    connames = connames or [f"r{i + 1}" for i in range(m)]
    lines = [f"NAME {name}", "ROWS", " N obj"]
    kinds = []
    for i in range(m):
        lo, up = lc[i], uc[i]
        if lo == up:
            kinds.append("E")
        elif np.isfinite(lo):
            kinds.append("G")
        elif np.isfinite(up):
            kinds.append("L")
        else:
            raise ArgumentError(f"row {i} has no finite bound")
        lines.append(f" {kinds[-1]} {connames[i]}")
    lines.append("COLUMNS")
    for j in range(n):
        if c[j] != 0.0:
            lines.append(f" {varnames[j]} obj {_num(c[j])}")
        for k in range(A.indptr[j], A.indptr[j + 1]):
            lines.append(f" {varnames[j]} {connames[A.indices[k]]} {_num(A.data[k])}")
        if c[j] == 0.0 and A.indptr[j] == A.indptr[j + 1]:
            lines.append(f" {varnames[j]} obj 0.0")
    lines.append("RHS")
    for i in range(m):
        b = lc[i] if kinds[i] in ("E", "G") else uc[i]
        if b != 0.0:
            lines.append(f" rhs {connames[i]} {_num(b)}")
    rng = [(i, uc[i] - lc[i]) for i in range(m) if kinds[i] == "G" and np.isfinite(uc[i])]
    if rng:
        lines.append("RANGES")
        for i, r in rng:
            lines.append(f" rng {connames[i]} {_num(r)}")
    lines.append("BOUNDS")
    for j in range(n):
        lo, up = lv[j], uv[j]
        if lo == up:
            lines.append(f" FX bnd {varnames[j]} {_num(lo)}")
            continue
        if lo == -np.inf and up == np.inf:
            lines.append(f" FR bnd {varnames[j]}")
            continue
        if lo == -np.inf:
            lines.append(f" MI bnd {varnames[j]}")
        elif lo != 0.0 or up < 0.0:   # a negative UP alone would make the lower bound -inf
            lines.append(f" LO bnd {varnames[j]} {_num(lo)}")
        if up != np.inf:
            lines.append(f" UP bnd {varnames[j]} {_num(up)}")
    lines.append("ENDATA")
    text = "\n".join(lines) + "\n"
    if hasattr(path_or_file, "write"):
        path_or_file.write(text)
    else:
        with open(path_or_file, "w") as f:
            f.write(text)
