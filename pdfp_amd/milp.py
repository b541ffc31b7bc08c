"""The MILP data contract and the host-side feasibility checks.

Mirrors src/milp.jl of the reference:

- ``MILP``                       src/milp.jl:15-26
- ``read_mps``                   src/milp.jl:42-94
- ``max_violation``              src/milp.jl:102-114
- ``is_feasible``                src/milp.jl:122-127
- ``is_integral`` (``_is_integral``) src/milp.jl:129-130
- ``write_solution``             src/milp.jl:138-145

The matrix is a ``scipy.sparse.csc_matrix`` with float64 values and int64
indices — the same compressed-sparse-column layout as Julia's
``SparseMatrixCSC{Float64,Int}``, 0-based.  Row senses are ``"le"`` / ``"eq"``
(the reference's ``:le`` / ``:eq``); ``int_idx`` is 0-based.
"""
from __future__ import annotations

import gzip
import math
import os
import re
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp

MPS_INFINITY = 1.0e20


@dataclass
class MILP:
    """min c'x + c0  s.t.  A x ≤ b (rows "le"), A x = b (rows "eq"),
    l ≤ x ≤ u, x_j integer for j in int_idx  (src/milp.jl:1-26)."""

    c: np.ndarray
    c0: float
    A: sp.csc_matrix
    b: np.ndarray
    row_sense: np.ndarray          # array of "le" / "eq"
    l: np.ndarray
    u: np.ndarray
    int_idx: np.ndarray            # 0-based, ascending
    var_names: list = field(default_factory=list)
    name: str = ""

    def __post_init__(self):
        self.c = np.asarray(self.c, dtype=np.float64)
        self.b = np.asarray(self.b, dtype=np.float64)
        self.l = np.asarray(self.l, dtype=np.float64)
        self.u = np.asarray(self.u, dtype=np.float64)
        self.int_idx = np.asarray(self.int_idx, dtype=np.int64)
        self.row_sense = np.asarray(self.row_sense, dtype="<U2")
        A = sp.csc_matrix(self.A, dtype=np.float64)
        A.sort_indices()
        A.indices = A.indices.astype(np.int64)
        A.indptr = A.indptr.astype(np.int64)
        self.A = A
        if not self.var_names:
            self.var_names = [f"x{j + 1}" for j in range(len(self.c))]
        if A.shape != (len(self.b), len(self.c)):
            raise ValueError(f"A is {A.shape}, expected {(len(self.b), len(self.c))}")

    @property
    def is_eq(self) -> np.ndarray:
        return self.row_sense == "eq"

    @property
    def is_int(self) -> np.ndarray:
        v = np.zeros(len(self.c), dtype=bool)
        v[self.int_idx] = True
        return v


def num_variables(milp: MILP) -> int:
    return len(milp.c)


def num_constraints(milp: MILP) -> int:
    return len(milp.b)


def num_continuous(milp: MILP) -> int:
    return len(milp.c) - len(milp.int_idx)


def _is_infinite(v: float) -> bool:
    return (not math.isfinite(v)) or abs(v) >= MPS_INFINITY


def max_violation(milp: MILP, x) -> float:
    """Largest violation of A x ≤ b / A x = b at x  (src/milp.jl:102-114);
    only the rows count, not the column bounds or integrality."""
    x = np.asarray(x, dtype=np.float64)
    Ax = _csc_matvec(milp.A, x)
    d = Ax - milp.b
    eq = milp.row_sense == "eq"
    v = 0.0
    if len(d):
        le_max = np.max(np.where(eq, -np.inf, d)) if (~eq).any() else -np.inf
        eq_max = np.max(np.abs(d[eq])) if eq.any() else -np.inf
        v = max(0.0, le_max, eq_max)
        if np.isnan(d).any():
            return float("nan")
    return float(v)


class DeviceViolationCheck:
    """max_violation(milp, x) on the device (pdfp_checker_*): the original A
    stays in HBM; each call uploads x and reads back one scalar, bit-identical
    to the reference's max_violation (src/milp.jl:102-114; a NaN activity
    never raises the maximum, as there)."""

    def __init__(self, milp: MILP, device: int = 0):
        import ctypes as C

        from . import _lib

        self._L = _lib
        self._lib = lib = _lib.load()
        A = sp.csc_matrix(milp.A, dtype=np.float64)
        self.n = A.shape[1]
        colptr = np.ascontiguousarray(A.indptr, dtype=np.int64)
        rowval = np.ascontiguousarray(A.indices, dtype=np.int64)
        nzval = np.ascontiguousarray(A.data, dtype=np.float64)
        b = np.ascontiguousarray(milp.b, dtype=np.float64)
        is_le = np.ascontiguousarray(milp.row_sense == "le", dtype=np.uint8)
        ctx = C.c_void_p()
        st = lib.pdfp_checker_create(C.byref(ctx), A.shape[0], self.n, _lib.i64ptr(colptr),
                                     _lib.i64ptr(rowval), _lib.dptr(nzval), 0, _lib.dptr(b),
                                     _lib.u8ptr(is_le), device)
        _lib.check(st, lib.pdfp_checker_last_error, None)
        self._ctx = ctx
        self._out = C.c_double()
        self._C = C

    def __del__(self):
        ctx = getattr(self, "_ctx", None)
        if ctx:
            self._lib.pdfp_checker_destroy(ctx)
            self._ctx = None

    def __call__(self, x) -> float:
        xs = np.ascontiguousarray(x, dtype=np.float64)
        if len(xs) != self.n:
            raise ValueError("x has the wrong length")
        st = self._lib.pdfp_checker_max_violation(self._ctx, self._L.dptr(xs),
                                                  self._C.byref(self._out))
        self._L.check(st, self._lib.pdfp_checker_last_error, self._ctx)
        return float(self._out.value)


def _csc_matvec(A: sp.csc_matrix, x: np.ndarray) -> np.ndarray:
    """y = A x in the column order of SparseArrays' CSC product: for each
    column j, y[row] += a * x[j], accumulated in that order (np.bincount
    accumulates sequentially in input order)."""
    m = A.shape[0]
    if A.nnz == 0:
        return np.zeros(m)
    cols = np.repeat(np.arange(A.shape[1]), np.diff(A.indptr))
    return np.bincount(A.indices, weights=A.data * x[cols], minlength=m)


def is_integral(x, idx, tol: float = 1e-6) -> bool:
    """_is_integral(x, idx, tol)  src/milp.jl:129-130 (Julia's round is
    ties-to-even, as is np.round)."""
    x = np.asarray(x, dtype=np.float64)
    if len(idx) == 0:
        return True
    v = x[idx]
    return bool(np.all(np.abs(v - np.round(v)) <= tol))


def is_feasible(milp: MILP, x, tol: float = 1e-6) -> bool:
    """Integrality, bounds and every constraint to the absolute tolerance
    (src/milp.jl:122-127)."""
    x = np.asarray(x, dtype=np.float64)
    if len(x) != num_variables(milp):
        return False
    if not is_integral(x, milp.int_idx, tol):
        return False
    if not np.all((milp.l - tol <= x) & (x <= milp.u + tol)):
        return False
    return max_violation(milp, x) <= tol


def write_solution(path: str, milp: MILP, x) -> str:
    """One ``name<TAB>value`` line per variable (src/milp.jl:138-145)."""
    with open(path, "w") as io:
        for name, v in zip(milp.var_names, x):
            io.write(f"{name}\t{float(v)!r}\n")
    return path


# ---- MPS reader ------------------------------------------------------------


def read_mps(path: str) -> MILP:
    """Read a free-format MPS file, optionally gzipped (src/milp.jl:42-94).

    Conventions, as the reference's reader applies them: a maximisation
    objective is negated; binaries (BV) get [0, 1]; variables between
    MARKER INTORG / INTEND and LI / UI bounds are integer; default bounds are
    [0, ∞); ``UP`` with a negative value on a variable whose lower bound is
    still the default 0 makes the lower bound -∞; |v| ≥ 1e20 is infinite.
    Each row maps to canonical rows: a row with equal finite bounds is one
    "eq" row; otherwise a finite upper bound gives ``a x ≤ hi`` and a finite
    lower bound gives ``-a x ≤ -lo`` (ranged rows give both).
    """
    opener = gzip.open if path.endswith(".gz") else open
    with opener(path, "rt", encoding="latin-1") as f:
        lines = f.read().splitlines()

    name = None
    objsense = "min"
    obj_row = None
    row_names: list[str] = []
    row_type: dict[str, str] = {}
    row_index: dict[str, int] = {}
    col_names: list[str] = []
    col_index: dict[str, int] = {}
    col_int: list[bool] = []
    entries_r: list[int] = []
    entries_c: list[int] = []
    entries_v: list[float] = []
    cost: dict[int, float] = {}
    rhs: dict[str, float] = {}
    ranges: dict[str, float] = {}
    lower: dict[int, float] = {}
    upper: dict[int, float] = {}
    binary: set = set()
    integer_marker = False
    section = None
    obj_constant = 0.0

    for raw in lines:
        if not raw.strip() or raw.startswith("*"):
            continue
        if not raw[0].isspace():
            tok = raw.split()
            head = tok[0].upper()
            if head == "NAME":
                name = tok[1] if len(tok) > 1 else None
                section = None
            elif head in ("OBJSENSE", "OBJSENCE"):
                section = "OBJSENSE"
                if len(tok) > 1:
                    objsense = "max" if tok[1].upper().startswith("MAX") else "min"
            elif head in ("ROWS", "COLUMNS", "RHS", "RANGES", "BOUNDS"):
                section = head
            elif head == "ENDATA":
                break
            else:
                section = head
            continue
        tok = raw.split()
        if section == "OBJSENSE":
            objsense = "max" if tok[0].upper().startswith("MAX") else "min"
        elif section == "ROWS":
            t, rname = tok[0].upper(), tok[1]
            if t == "N":
                if obj_row is None:
                    obj_row = rname
                row_type[rname] = "N"
            else:
                row_type[rname] = t
                row_index[rname] = len(row_names)
                row_names.append(rname)
        elif section == "COLUMNS":
            if len(tok) >= 3 and tok[1].upper().strip("'\"") == "MARKER":
                mk = tok[2].upper().strip("'\"")
                if mk == "INTORG":
                    integer_marker = True
                elif mk == "INTEND":
                    integer_marker = False
                continue
            cname = tok[0]
            if cname not in col_index:
                col_index[cname] = len(col_names)
                col_names.append(cname)
                col_int.append(integer_marker)
            j = col_index[cname]
            for k in range(1, len(tok) - 1, 2):
                rname, val = tok[k], float(tok[k + 1])
                if rname == obj_row:
                    cost[j] = cost.get(j, 0.0) + val
                elif row_type.get(rname) == "N":
                    continue
                else:
                    entries_r.append(row_index[rname])
                    entries_c.append(j)
                    entries_v.append(val)
        elif section == "RHS":
            pairs = tok[1:] if len(tok) % 2 == 1 else tok
            for k in range(0, len(pairs) - 1, 2):
                rname, val = pairs[k], float(pairs[k + 1])
                if rname == obj_row:
                    obj_constant = -val
                else:
                    rhs[rname] = val
        elif section == "RANGES":
            pairs = tok[1:] if len(tok) % 2 == 1 else tok
            for k in range(0, len(pairs) - 1, 2):
                ranges[pairs[k]] = float(pairs[k + 1])
        elif section == "BOUNDS":
            bt = tok[0].upper()
            if bt in ("FR", "MI", "PL", "BV") and len(tok) == 3:
                cname = tok[2]
                val = None
            elif bt in ("FR", "MI", "PL", "BV") and len(tok) == 2:
                cname = tok[1]
                val = None
            else:
                cname = tok[2]
                val = float(tok[3])
            j = col_index[cname]
            if bt == "UP":
                if val < 0 and lower.get(j, 0.0) == 0.0 and j not in lower:
                    lower[j] = -math.inf
                upper[j] = val
            elif bt == "LO":
                lower[j] = val
            elif bt == "FX":
                lower[j] = val
                upper[j] = val
            elif bt == "FR":
                lower[j] = -math.inf
                upper[j] = math.inf
            elif bt == "MI":
                lower[j] = -math.inf
            elif bt == "PL":
                upper[j] = math.inf
            elif bt == "BV":
                binary.add(j)
                lower[j] = 0.0
                upper[j] = 1.0
            elif bt == "LI":
                col_int[j] = True
                lower[j] = val
            elif bt == "UI":
                col_int[j] = True
                upper[j] = val
            elif bt == "SC":
                raise ValueError("semi-continuous bounds are not supported")

    n = len(col_names)
    sign = -1.0 if objsense == "max" else 1.0
    c = np.zeros(n)
    for j, v in cost.items():
        c[j] = v
    c = sign * c
    c0 = sign * obj_constant

    def fin(v):
        return (-math.inf if v < 0 else math.inf) if _is_infinite(v) else v

    l = np.array([fin(lower.get(j, 0.0)) for j in range(n)], dtype=np.float64)
    u = np.array([fin(upper.get(j, math.inf)) for j in range(n)], dtype=np.float64)
    int_idx = []
    for j in range(n):
        if j in binary:
            int_idx.append(j)
            l[j] = max(l[j], 0.0)
            u[j] = min(u[j], 1.0)
        elif col_int[j]:
            int_idx.append(j)

    # Original row bounds [lo, hi] from the type, the RHS and the ranges.
    b_list: list[float] = []
    sense: list[str] = []
    canonical: list[list[tuple[int, float]]] = [[] for _ in row_names]
    for i, rname in enumerate(row_names):
        t = row_type[rname]
        r = rhs.get(rname, 0.0)
        if t == "E":
            lo = hi = r
            if rname in ranges:
                R = ranges[rname]
                if R > 0:
                    hi = r + abs(R)
                elif R < 0:
                    lo = r - abs(R)
        elif t == "L":
            lo, hi = -math.inf, r
            if rname in ranges:
                lo = r - abs(ranges[rname])
        elif t == "G":
            lo, hi = r, math.inf
            if rname in ranges:
                hi = r + abs(ranges[rname])
        else:
            raise ValueError(f"unknown row type {t}")
        if not _is_infinite(lo) and not _is_infinite(hi) and lo == hi:
            b_list.append(hi); sense.append("eq"); canonical[i].append((len(b_list) - 1, 1.0))
            continue
        if not _is_infinite(hi):
            b_list.append(hi); sense.append("le"); canonical[i].append((len(b_list) - 1, 1.0))
        if not _is_infinite(lo):
            b_list.append(-lo); sense.append("le"); canonical[i].append((len(b_list) - 1, -1.0))

    I, J, V = [], [], []
    for r, cj, v in zip(entries_r, entries_c, entries_v):
        for row, s in canonical[r]:
            I.append(row); J.append(cj); V.append(s * v)
    # sparse(I, J, V): duplicates are summed
    A = sp.coo_matrix((np.array(V, dtype=np.float64), (np.array(I, dtype=np.int64),
                       np.array(J, dtype=np.int64))), shape=(len(b_list), n)).tocsc()
    A.sum_duplicates()
    if name is None:
        name = re.sub(r"\.mps(\.gz)?$", "", os.path.basename(path))
    return MILP(c=c, c0=c0, A=A, b=np.array(b_list, dtype=np.float64),
                row_sense=np.array(sense, dtype="<U2"), l=l, u=u,
                int_idx=np.array(int_idx, dtype=np.int64), var_names=list(col_names),
                name=name)
